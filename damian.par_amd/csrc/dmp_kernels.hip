// dmp_kernels.hip -- gfx950 kernels of the explicit elastic-wave step.
//
// One launch advances every tile by k time steps of run_explicit
// (explicit.f90:125-153).  A workgroup owns a tile of consecutive active
// nodes and keeps, in LDS, u at the tile and at the rings of nodes around it
// that its next k updates read.  Per sub-step it updates the rings that are
// still exact, so after k sub-steps the tile's own nodes hold u_{t0+k} and
// nothing outside the workgroup is read inside the launch.  Per node:
// lumped mass / Lysmer damping / K_e u gathered from the node's elements
// (explicit.f90:129-142), hanging-node reaction transfer (sum_reactanchors,
// P2), plane-wave forces (wave_reactions, model.f90:745-800, FP64 phase and
// time, evaluated in the workgroup for its k steps), point sources, division
// by e_mfull (:148); receivers are sampled every step from LDS
// (visu_field, visualization.f90:156-181).  u_j lives in U[j mod 10]; hanging
// nodes are never stored, their value is the mean of their anchors
// (model.f90:612-624).
//
// The element stiffness is K_e = lambda_e K_lambda + mu_e K_mu with size-
// independent 8x8 tables of the square bilinear plane-strain element
// (formulation.f90:276-313); the tables are derived below from the shape
// functions and fold to immediates.  Rows of K annihilate rigid translations,
// so K_e u_e is evaluated on differences u_j - u_n (exact in FP64 up to
// rounding, and what keeps the FP32 mode accurate).
#include "dmp_internal.h"

#include <mutex>

namespace dmp {

// ---------------------------------------------------------------------------
// Element tables (square bilinear quad, corners BL, BR, TR, TL, x right,
// z down; DOF 2a = x, 2a+1 = z of corner a).
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int cx_(int a) { return (a == 1 || a == 2) ? 1 : 0; }
__host__ __device__ constexpr int cz_(int a) { return a >= 2 ? 1 : 0; }
// integral over [0,1] of phi_a * phi_b (phi = s or 1-s)
__host__ __device__ constexpr double ipp_(int ca, int cb) { return ca == cb ? 1.0 / 3.0 : 1.0 / 6.0; }
__host__ __device__ constexpr double sg_(int c) { return c ? 1.0 : -1.0; }

// K = lambda * KL + mu * KM  (plane strain, D = [[l+2m, l, 0],[l, l+2m, 0],[0,0,m]])
__host__ __device__ constexpr double KL(int r, int c)
{
    const int a = r >> 1, p = r & 1, b = c >> 1, q = c & 1;
    const double xx = sg_(cx_(a)) * sg_(cx_(b)) * ipp_(cz_(a), cz_(b));
    const double zz = sg_(cz_(a)) * sg_(cz_(b)) * ipp_(cx_(a), cx_(b));
    const double xz = sg_(cx_(a)) * sg_(cz_(b)) * 0.25;
    const double zx = sg_(cz_(a)) * sg_(cx_(b)) * 0.25;
    return p == 0 ? (q == 0 ? xx : xz) : (q == 0 ? zx : zz);
}
__host__ __device__ constexpr double KM(int r, int c)
{
    const int a = r >> 1, p = r & 1, b = c >> 1, q = c & 1;
    const double xx = sg_(cx_(a)) * sg_(cx_(b)) * ipp_(cz_(a), cz_(b));
    const double zz = sg_(cz_(a)) * sg_(cz_(b)) * ipp_(cx_(a), cx_(b));
    const double xz = sg_(cx_(a)) * sg_(cz_(b)) * 0.25;
    const double zx = sg_(cz_(a)) * sg_(cx_(b)) * 0.25;
    return p == 0 ? (q == 0 ? 2.0 * xx + zz : zx) : (q == 0 ? xz : 2.0 * zz + xx);
}
// Consistent-mass pattern scaled like m_mmat (formulation.f90:256-268):
// 4 * integral of N_a N_b over the unit square, same component only.
__host__ __device__ constexpr double MH(int r, int c)
{
    const int a = r >> 1, p = r & 1, b = c >> 1, q = c & 1;
    return p == q ? 4.0 * ipp_(cx_(a), cx_(b)) * ipp_(cz_(a), cz_(b)) : 0.0;
}

static_assert(KM(0, 0) == 2.0 / 3.0 + 1.0 / 3.0, "K_mu(1,1)");
static_assert(KL(0, 1) == 0.25 && KM(0, 1) == 0.25, "K(1,2) = (mu+lambda)/4");

// Neighbour order: the corner of the square across x from corner k (same z),
// across z (same x), and the diagonal one (t = 0, 1, 2), and the rows of
// K_lambda / K_mu at corner k for them in terms of q = s / 4, s = +1 for even
// k and -1 for odd k (p: row x or z of the corner, c: column x or z of the
// neighbour).  Checked against KL / KM for every corner and neighbour: the
// rows of the four corners differ only in the sign of q.
__host__ __device__ constexpr int nbr_corner(int k, int t)
{
    return t == 0 ? (k ^ 1) : t == 1 ? (3 - k) : ((k + 2) & 3);
}
__host__ __device__ constexpr double nbr_kl(int t, int p, int c, double q)
{
    return t == 0 ? (p == 0 ? (c == 0 ? -1.0 / 3.0 : q) : (c == 0 ? -q : 1.0 / 6.0))
         : t == 1 ? (p == 0 ? (c == 0 ? 1.0 / 6.0 : -q) : (c == 0 ? q : -1.0 / 3.0))
                  : (p == 0 ? (c == 0 ? -1.0 / 6.0 : -q) : (c == 0 ? -q : -1.0 / 6.0));
}
__host__ __device__ constexpr double nbr_km(int t, int p, int c, double q)
{
    return t == 0 ? (p == 0 ? (c == 0 ? -0.5 : -q) : (c == 0 ? q : 0.0))
         : t == 1 ? (p == 0 ? (c == 0 ? 0.0 : q) : (c == 0 ? -q : -0.5))
                  : (p == 0 ? (c == 0 ? -0.5 : -q) : (c == 0 ? -q : -0.5));
}
__host__ __device__ constexpr bool nbr_tables_match()
{
    for (int k = 0; k < 4; ++k)
        for (int t = 0; t < 3; ++t)
            for (int p = 0; p < 2; ++p)
                for (int c = 0; c < 2; ++c) {
                    const double q = (k & 1) ? -0.25 : 0.25;
                    const int j = nbr_corner(k, t);
                    if (KL(2 * k + p, 2 * j + c) != nbr_kl(t, p, c, q)) return false;
                    if (KM(2 * k + p, 2 * j + c) != nbr_km(t, p, c, q)) return false;
                }
    return true;
}
static_assert(nbr_tables_match(), "neighbour-order rows differ from K_lambda / K_mu");

// ---------------------------------------------------------------------------
// Precision policies.
// ---------------------------------------------------------------------------
template <class R> struct V2 { R x, y; };

template <class R>
__device__ __forceinline__ V2<R> operator+(V2<R> a, V2<R> b) { return {a.x + b.x, a.y + b.y}; }

struct PF64 {
    using real = double;
    using node_t = double2;
    static constexpr int kOneWaves = 6;   // one-step kernel: waves per SIMD it is compiled for
    __device__ static node_t ld(const void* U, int i) { return static_cast<const double2*>(U)[i]; }
    __device__ static void st(void* U, int i, node_t v) { static_cast<double2*>(U)[i] = v; }
    __device__ static V2<real> sub(node_t a, node_t b) { return {a.x - b.x, a.y - b.y}; }
    __device__ static double2 full(node_t a) { return a; }
    __device__ static real coef(const void* c, int n) { return static_cast<const double*>(c)[n]; }
    __device__ static V2<real> lm(const void* t, int m)
    {
        double2 v = static_cast<const double2*>(t)[m];
        return {v.x, v.y};
    }
    // u_{i} = 2 u_{i-1} - u_{i-2} + corr
    __device__ static node_t advance(node_t u, node_t up, V2<real> c)
    {
        return make_double2((2.0 * u.x - up.x) + c.x, (2.0 * u.y - up.y) + c.y);
    }
};

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Compensated FP32: node_t = (hi.x, hi.z, lo.x, lo.z), value = hi + lo.
struct PF32C {
    using real = float;
    using node_t = float4;
    static constexpr int kOneWaves = 7;
    __device__ static node_t ld(const void* U, int i) { return static_cast<const float4*>(U)[i]; }
    __device__ static void st(void* U, int i, node_t v) { static_cast<float4*>(U)[i] = v; }
    __device__ static V2<real> sub(node_t a, node_t b)
    {
        return {(a.x - b.x) + (a.z - b.z), (a.y - b.y) + (a.w - b.w)};
    }
    __device__ static double2 full(node_t a)
    {
        return make_double2((double)a.x + (double)a.z, (double)a.y + (double)a.w);
    }
    __device__ static real coef(const void* c, int n) { return static_cast<const float*>(c)[n]; }
    __device__ static V2<real> lm(const void* t, int m)
    {
        float2 v = static_cast<const float2*>(t)[m];
        return {v.x, v.y};
    }
    __device__ static void adv1(float uh, float ul, float ph, float pl, float c, float& rh, float& rl)
    {
        float s, e;
        two_sum(2.0f * uh, -ph, s, e);             // 2*uh exact
        const float lo = ((2.0f * ul - pl) + e) + c;
        two_sum(s, lo, rh, rl);
    }
    __device__ static node_t advance(node_t u, node_t up, V2<real> c)
    {
        node_t r;
        adv1(u.x, u.z, up.x, up.z, c.x, r.x, r.z);
        adv1(u.y, u.w, up.y, up.w, c.y, r.y, r.w);
        return r;
    }
};

// ---------------------------------------------------------------------------
// Node update (LDS-tiled).
// ---------------------------------------------------------------------------
// Mean of two stored values: the hanging-node value (u_a1 + u_a2)/2 of
// set_dispshanging (model.f90:622).
__device__ __forceinline__ double2 avg2(double2 a, double2 b)
{
    return make_double2((a.x + b.x) * 0.5, (a.y + b.y) * 0.5);
}
__device__ __forceinline__ float4 avg2(float4 a, float4 b)
{
    float4 r;
    float s, e;
    two_sum(a.x, b.x, s, e);
    two_sum(0.5f * s, 0.5f * ((a.z + b.z) + e), r.x, r.z);
    two_sum(a.y, b.y, s, e);
    two_sum(0.5f * s, 0.5f * ((a.w + b.w) + e), r.y, r.w);
    return r;
}

// f -= rows (2K, 2K+1) of K_e (lam, mu) applied to u_j - uref over the three
// other corners of the element, whose LDS slots are packed in w.
template <class P, int K>
__device__ __forceinline__ void slot_rows(const typename P::node_t* val, uint32_t w,
                                          V2<typename P::real> lm, typename P::node_t uref,
                                          V2<typename P::real>& f)
{
    using R = typename P::real;
    R xl = 0, xm = 0, zl = 0, zm = 0;
    int q = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j == K) continue;
        const int slot = (w >> (10 * q)) & 1023;
        ++q;
        const V2<R> d = P::sub(val[slot], uref);
        // two chained fused multiply-adds per accumulator
        xl = fma(R(KL(2 * K, 2 * j + 1)), d.y, fma(R(KL(2 * K, 2 * j)), d.x, xl));
        xm = fma(R(KM(2 * K, 2 * j + 1)), d.y, fma(R(KM(2 * K, 2 * j)), d.x, xm));
        zl = fma(R(KL(2 * K + 1, 2 * j + 1)), d.y, fma(R(KL(2 * K + 1, 2 * j)), d.x, zl));
        zm = fma(R(KM(2 * K + 1, 2 * j + 1)), d.y, fma(R(KM(2 * K + 1, 2 * j)), d.x, zm));
    }
    f.x -= lm.x * xl + lm.y * xm;
    f.y -= lm.x * zl + lm.y * zm;
}

// slot_rows on corner values already in registers (c[q] = value of the q-th
// corner other than K, in descriptor order).
template <class P, int K>
__device__ __forceinline__ void slot_rows_v(const typename P::node_t (&c)[3], V2<typename P::real> lm,
                                            typename P::node_t uref, V2<typename P::real>& f)
{
    using R = typename P::real;
    R xl = 0, xm = 0, zl = 0, zm = 0;
    int q = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j == K) continue;
        const V2<R> d = P::sub(c[q], uref);
        ++q;
        // two chained fused multiply-adds per accumulator
        xl = fma(R(KL(2 * K, 2 * j + 1)), d.y, fma(R(KL(2 * K, 2 * j)), d.x, xl));
        xm = fma(R(KM(2 * K, 2 * j + 1)), d.y, fma(R(KM(2 * K, 2 * j)), d.x, xm));
        zl = fma(R(KL(2 * K + 1, 2 * j + 1)), d.y, fma(R(KL(2 * K + 1, 2 * j)), d.x, zl));
        zm = fma(R(KM(2 * K + 1, 2 * j + 1)), d.y, fma(R(KM(2 * K + 1, 2 * j)), d.x, zm));
    }
    f.x -= lm.x * xl + lm.y * xm;
    f.y -= lm.x * zl + lm.y * zm;
}

// Element rows of one node (or hanging node) from its descriptor: per
// quadrant k one word holding the LDS slots of the element's three other
// corners (10 bits each, ascending corner order), 0xFFFFFFFF for an empty
// quadrant.  A tile with at most four materials keeps their (lambda, mu) in a
// 4-entry table, indexed by bits 30-31 of the word.  A tile with more
// (heterogeneous media, ELEM) holds an element's (lambda, mu) in LDS at the
// slot of the element's corner 0 (BL, each node being corner 0 of at most
// one element): the node itself for k = 0, the word's first slot otherwise.
template <class P>
struct Rows {
    using R = typename P::real;
    uint32_t w[4];

    __device__ void load(uint4 d0)
    {
        w[0] = d0.x;
        w[1] = d0.y;
        w[2] = d0.z;
        w[3] = d0.w;
    }
    __device__ bool present(int k) const { return w[k] != 0xFFFFFFFFu; }
    __device__ V2<R> mat(int k, const V2<R>* lmt, const V2<R>* em, int self, bool elem) const
    {
        if (!elem) return lmt[(w[k] >> 30) & 3];
        const int slot = (k == 0 || !present(k)) ? self : (int)(w[k] & 1023);
        return em[slot];
    }
    // Two elements at a time: the six corner reads are issued before any of
    // them is used (an absent element reads the node's own slot, whose
    // differences are exactly 0, with a zero material, so it adds +0 as
    // skipping it would).
    template <int K0>
    __device__ void pair_rows(const typename P::node_t* val, typename P::node_t uref, int self,
                              const V2<R>* lmt, const V2<R>* em, bool elem, V2<R>& f) const
    {
        using node_t = typename P::node_t;
        const uint32_t sw = (uint32_t)self * ((1u << 20) | (1u << 10) | 1u);
        const bool p0 = present(K0), p1 = present(K0 + 1);
        const uint32_t w0 = p0 ? w[K0] : sw, w1 = p1 ? w[K0 + 1] : sw;
        const V2<R> m0 = mat(K0, lmt, em, self, elem), m1 = mat(K0 + 1, lmt, em, self, elem);
        node_t c0[3], c1[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            c0[q] = val[(w0 >> (10 * q)) & 1023];
            c1[q] = val[(w1 >> (10 * q)) & 1023];
        }
        __builtin_amdgcn_sched_barrier(0);
        slot_rows_v<P, K0>(c0, p0 ? m0 : V2<R>{0, 0}, uref, f);
        slot_rows_v<P, K0 + 1>(c1, p1 ? m1 : V2<R>{0, 0}, uref, f);
    }
    __device__ V2<R> apply_pairs(const typename P::node_t* val, typename P::node_t uref, int self,
                                 const V2<R>* lmt, const V2<R>* em, bool elem) const
    {
        V2<R> f{0, 0};
        pair_rows<0>(val, uref, self, lmt, em, elem, f);
        pair_rows<2>(val, uref, self, lmt, em, elem, f);
        return f;
    }
    // Rows (2K, 2K+1) of corner K's element from a word whose slots are in
    // neighbour order, K a runtime value: the rows of the square element's
    // four corners differ only by the sign s = +1 (K even) or -1 (K odd) of
    // their off-diagonal terms, q = s / 4 (nbr_kl / nbr_km above, checked
    // against K_lambda / K_mu).  The element's (lambda, mu) of an ELEM tile
    // sits at its corner 0: the node itself (K = 0), the x- (K = 1), diagonal
    // (K = 2) or z-neighbour (K = 3).
    // coefficient of table M (0: K_lambda, 1: K_mu), neighbour T, row PP,
    // column CC: a constant or +-q (nbr_kl / nbr_km)
    template <int M, int T, int PP, int CC>
    __device__ static R kc(R q)
    {
        constexpr double c0 = M ? nbr_km(T, PP, CC, 0.0) : nbr_kl(T, PP, CC, 0.0);
        constexpr double c1 = (M ? nbr_km(T, PP, CC, 1.0) : nbr_kl(T, PP, CC, 1.0)) - c0;
        static_assert(c0 == 0.0 || c1 == 0.0, "a coefficient is a constant or +-q");
        return c1 == 0.0 ? R(c0) : (c1 > 0.0 ? q : -q);
    }
    // acc + coefficient * d, the term left out where the coefficient is 0
    template <int M, int T, int PP, int CC>
    __device__ static R term(R q, R d, R acc)
    {
        constexpr double c0 = M ? nbr_km(T, PP, CC, 0.0) : nbr_kl(T, PP, CC, 0.0);
        constexpr double c1 = (M ? nbr_km(T, PP, CC, 1.0) : nbr_kl(T, PP, CC, 1.0)) - c0;
        if constexpr (c0 == 0.0 && c1 == 0.0) return acc;
        else return fma(kc<M, T, PP, CC>(q), d, acc);
    }
    // row PP of table M over the three neighbours, in neighbour order, column
    // x then z of each
    template <int M, int PP>
    __device__ static R row(R q, V2<R> dx, V2<R> dz, V2<R> dd)
    {
        static_assert((M ? nbr_km(0, PP, 0, 1.0) : nbr_kl(0, PP, 0, 1.0)) != 0.0, "first term present");
        R acc = kc<M, 0, PP, 0>(q) * dx.x;
        acc = term<M, 0, PP, 1>(q, dx.y, acc);
        acc = term<M, 1, PP, 0>(q, dz.x, acc);
        acc = term<M, 1, PP, 1>(q, dz.y, acc);
        acc = term<M, 2, PP, 0>(q, dd.x, acc);
        return term<M, 2, PP, 1>(q, dd.y, acc);
    }
    __device__ void quad_rows(uint32_t wk, int K, R q, const typename P::node_t* val, typename P::node_t uref,
                              int self, const V2<R>* lmt, const V2<R>* em, bool elem, V2<R>& f) const
    {
        const int sx = wk & 1023, sz = (wk >> 10) & 1023, sd = (wk >> 20) & 1023;
        const V2<R> dx = P::sub(val[sx], uref), dz = P::sub(val[sz], uref), dd = P::sub(val[sd], uref);
        const int ms = K == 0 ? self : K == 1 ? sx : K == 2 ? sd : sz;
        const V2<R> lm = elem ? em[ms] : lmt[(wk >> 30) & 3];
        const R xl = row<0, 0>(q, dx, dz, dd), xm = row<1, 0>(q, dx, dz, dd);
        const R zl = row<0, 1>(q, dx, dz, dd), zm = row<1, 1>(q, dx, dz, dd);
        f.x -= lm.x * xl + lm.y * xm;
        f.y -= lm.x * zl + lm.y * zm;
    }
    // A hanging node's row, one half per lane of a pair: side 0 takes
    // quadrants 0 and 2, side 1 quadrants 1 and 3, so that the node's two
    // elements (adjacent quadrants) fall one to each lane whichever edge it
    // hangs on, and both lanes run the same instructions.  The descriptor
    // holds its slots in neighbour order (quad_rows).
    __device__ V2<R> half_row(int side, const typename P::node_t* val, typename P::node_t uref, int self,
                              const V2<R>* lmt, const V2<R>* em, bool elem) const
    {
        const R q = side ? R(-0.25) : R(0.25);
        const uint32_t wa = side ? w[1] : w[0], wb = side ? w[3] : w[2];
        const bool pa = wa != 0xFFFFFFFFu, pb = wb != 0xFFFFFFFFu;
        V2<R> f{0, 0};
        if (pa || pb) quad_rows(pa ? wa : wb, side + (pa ? 0 : 2), q, val, uref, self, lmt, em, elem, f);
        if (pa && pb) quad_rows(wb, side + 2, q, val, uref, self, lmt, em, elem, f);
        return f;
    }
    __device__ V2<R> apply(const typename P::node_t* val, typename P::node_t uref, int self,
                           const V2<R>* lmt, const V2<R>* em, bool elem) const
    {
        V2<R> f{0, 0};
        if (present(0)) slot_rows<P, 0>(val, w[0], mat(0, lmt, em, self, elem), uref, f);
        if (present(1)) slot_rows<P, 1>(val, w[1], mat(1, lmt, em, self, elem), uref, f);
        if (present(2)) slot_rows<P, 2>(val, w[2], mat(2, lmt, em, self, elem), uref, f);
        if (present(3)) slot_rows<P, 3>(val, w[3], mat(3, lmt, em, self, elem), uref, f);
        return f;
    }
};

// sigma-Ricker of formulation.f90:53-59 at time j*dt - start.
__device__ __forceinline__ double point_ricker(const InstParams& P, int j)
{
    const double t = (double)j * P.dt - P.p_start;
    return exp(-(t * t) / (4 * P.p_sigma)) * ((t * t) - 2 * P.p_sigma) / P.p_den * P.p_amp;
}

__device__ __forceinline__ double2 set_z(double2 u, double z) { return make_double2(u.x, z); }
__device__ __forceinline__ float4 set_z(float4 u, double z)
{
    const float hi = (float)z;
    return make_float4(u.x, hi, u.z, (float)(z - (double)hi));
}

// ---------------------------------------------------------------------------
// Plane-wave free field and forces (FP64).
// ---------------------------------------------------------------------------
// Incident + two reflected Ricker trains of pwaves.f90:148-217 at a point with
// precomputed delays v[w]: displacement (column 1) and acceleration (column 3)
// at step j; layer nodes use time j*dt - t0 (model.f90:761), receivers
// (dt*k)*io_interval - t0 with k = j / io_interval (visualization.f90:329).
__device__ double4 ff_value(const FFT& F, int interval, int j)
{
    double t;
    if (F.recv_time) {
        const int I = max(interval, 1);
        t = (F.dt * (double)(j / I)) * (double)I - F.t0;
    } else {
        t = (double)j * F.dt - F.t0;
    }
    const double fp = F.freq * 3.14159265358979323846;
    double ux = 0.0, uz = 0.0, ax = 0.0, az = 0.0;
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        const double tv = t - F.v[w];
        double s = fp * tv;
        s = s * s;
        const double ex = exp(s);
        const double d = (1.0 - 2.0 * s) / ex;
        const double a = -2.0 * (4.0 * (s * s) - 12.0 * s + 3.0) * (fp * fp) / ex;
        ux = ux + d * F.UX[w] * F.A[w];
        uz = uz + d * F.UZ[w] * F.A[w];
        ax = ax + a * F.UX[w] * F.A[w];
        az = az + a * F.UZ[w] * F.A[w];
    }
    return make_double4(ux, uz, ax, az);
}

// Rows of one seismic element at its corner k (multdrm, formulation.f90:378-393,
// applied to K_e u0 and to m_e * M a0).
template <int K>
__device__ void drm_rows(const DrmElem& E, const double4 f[4], double& fx, double& fz)
{
    const bool ek = (E.ext >> K) & 1;
    double kx = 0, kz = 0, mx = 0, mz = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const bool ej = (E.ext >> j) & 1;
        if (ek == ej) continue;
        const double sg = ek ? 1.0 : -1.0;
        const double k00 = E.lam * KL(2 * K, 2 * j) + E.mu * KM(2 * K, 2 * j);
        const double k01 = E.lam * KL(2 * K, 2 * j + 1) + E.mu * KM(2 * K, 2 * j + 1);
        const double k10 = E.lam * KL(2 * K + 1, 2 * j) + E.mu * KM(2 * K + 1, 2 * j);
        const double k11 = E.lam * KL(2 * K + 1, 2 * j + 1) + E.mu * KM(2 * K + 1, 2 * j + 1);
        kx += sg * (k00 * f[j].x + k01 * f[j].y);
        kz += sg * (k10 * f[j].x + k11 * f[j].y);
        mx += sg * (MH(2 * K, 2 * j) * f[j].z);
        mz += sg * (MH(2 * K + 1, 2 * j + 1) * f[j].w);
    }
    fx = kx + mx * E.mass;
    fz = kz + mz * E.mass;
}

// ---------------------------------------------------------------------------
// The step kernel.
// ---------------------------------------------------------------------------
// Receiver samples of step j = t0 + s (visu_field, visualization.f90:156-181)
// from the tile's values at that step; the free field is added in the
// exterior region.
// Lane of receiver work (sampling, free-field loads): several steps per
// launch put it on the last wave of the workgroup, which forms no
// hanging-node values or rows (one lane each, from lane 0 up; tiles hold a
// few dozen, at most a few hundred), so that sampling overlaps those phases;
// in the one-step kernel every wave has core nodes to update and the lane is
// the thread index.
template <bool ONE>
__device__ __forceinline__ int recv_lane()
{
    const int t = (int)threadIdx.x, nt = (int)blockDim.x;
    return ONE || nt == 64 ? t : (t + 64) % nt;
}

// First lane of a small list copied into LDS during the load phase, the lane
// order rotated by `off` lanes: in a multi-step launch wave 0 has the longest
// chain of dependent loads (its entry loads, then the lists whose lanes
// start at 0), so the short lists start on later waves, whose chains are
// shorter; the rotation is a bijection on the workgroup's lanes.  The
// one-step kernel keeps lane q for word q.
template <bool ONE>
__device__ __forceinline__ int copy_lane(int off)
{
    const int t = (int)threadIdx.x, nt = (int)blockDim.x;
    return ONE ? t : (t + nt - off % nt) % nt;
}

// Receiver value accumulation shared by the tile sampler (from LDS) and the
// HBM sampler, so that both give the same bits: corner value a, or for a
// hanging corner the mean of its anchors a and b (set_dispshanging), times the
// bilinear weight, summed over the corners in order (matmul(nmat, dofs),
// formulation.f90:325-330).
__device__ __forceinline__ void recv_accum(double2 a, double2 b, bool hang, double w, double& fx,
                                           double& fz)
{
    double2 u = a;
    if (hang) u = make_double2((a.x + b.x) / 2.0, (a.y + b.y) / 2.0);
    fx = fx + w * u.x;
    fz = fz + w * u.y;
}

__device__ __forceinline__ double* recv_slot(const DevView& V, int j, int r)
{
    const int q = j / V.recv_interval;
    return V.recv_buf + 2 * ((size_t)(q % V.recv_cap) * V.nrecv + r);
}

// Samples the tile's receivers at step t0 + s from the LDS values into
// rff[s][q], in place of that receiver's free field (read once, here, by the
// same thread).  The samples go out to the ring in flush_receivers at the end
// of the launch: a barrier waits for every global store its wave has issued
// (stores count in vmcnt), so nothing is stored while barriers remain.
template <class P>
__device__ void sample_receivers(const DevView& V, const typename P::node_t* cur, double2* rff,
                                 const RecvT* rcv, int n_recv, int t0, int s, int lane)
{
    const int I = V.recv_interval;
    const int j = t0 + s;
    if (I <= 0 || (j % I) != 0) return;
    for (int q = lane; q < n_recv; q += blockDim.x) {
        const RecvT& Rc = rcv[q];
        double fx = 0.0, fz = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool hang = (Rc.hang >> c) & 1;
            const double2 a = P::full(cur[Rc.s[c].x]);
            const double2 b = hang ? P::full(cur[Rc.s[c].y]) : a;
            recv_accum(a, b, hang, Rc.w[c], fx, fz);
        }
        if (Rc.ff >= 0) {
            const double2 g = rff[s * n_recv + q];
            fx = fx + g.x;
            fz = fz + g.y;
        }
        rff[s * n_recv + q] = make_double2(fx, fz);
    }
}

// Stores the samples of rows 0 .. rows-1 (steps t0 .. t0 + rows - 1) that
// sample_receivers left in rff; the same lane mapping, so each thread stores
// what it computed and no barrier is needed.
__device__ void flush_receivers(const DevView& V, const double2* rff, const RecvT* rcv, int n_recv, int t0,
                                int rows, int lane)
{
    const int I = V.recv_interval;
    if (I <= 0) return;
    for (int q = lane; q < n_recv; q += blockDim.x) {
        const int r = rcv[q].r;
        for (int s = 0; s < rows; ++s) {
            const int j = t0 + s;
            if (j % I != 0) continue;
            const double2 v = rff[s * n_recv + q];
            double* out = recv_slot(V, j, r);
            out[0] = v.x;
            out[1] = v.y;
        }
    }
}

// Plane-wave contribution of one (node, seismic element) pair from the free
// field at its corners.
__device__ double2 drm_pair(const DrmElem& E, const double4* fs)
{
    const double4 g[4] = {fs[E.c[0]], fs[E.c[1]], fs[E.c[2]], fs[E.c[3]]};
    double px = 0.0, pz = 0.0;
    switch (E.k) {
    case 0: drm_rows<0>(E, g, px, pz); break;
    case 1: drm_rows<1>(E, g, px, pz); break;
    case 2: drm_rows<2>(E, g, px, pz); break;
    default: drm_rows<3>(E, g, px, pz); break;
    }
    return make_double2(px, pz);
}

// Several steps per launch: the values of a tile's entries into LDS from its
// fixed-stride entry list (slot q for list word q), two list words per lane
// at a time so that their loads overlap.
template <class P>
__device__ __forceinline__ void load_entries(const DevView& V, int tile, const void* U0, const void* Um,
                                             typename P::node_t* buf, int S)
{
    using node_t = typename P::node_t;
    constexpr int kNode = kPrevFlag - 1;
    const int NT = blockDim.x, cap = V.elist_cap;
    const int* L = V.elist + (size_t)tile * cap;
    for (int q0 = threadIdx.x; q0 < cap; q0 += 2 * NT) {
        const int q1 = q0 + NT;
        const int w0 = L[q0];
        const int w1 = q1 < cap ? L[q1] : -1;
        node_t a0{}, b0{}, a1{}, b1{};
        if (w0 >= 0) {
            a0 = P::ld(U0, w0 & kNode);
            if (w0 & kPrevFlag) b0 = P::ld(Um, w0 & kNode);
        }
        if (w1 >= 0) {
            a1 = P::ld(U0, w1 & kNode);
            if (w1 & kPrevFlag) b1 = P::ld(Um, w1 & kNode);
        }
        if (w0 >= 0) {
            buf[q0] = a0;
            if (w0 & kPrevFlag) buf[S + q0] = b0;
        }
        if (w1 >= 0) {
            buf[q1] = a1;
            if (w1 & kPrevFlag) buf[S + q1] = b1;
        }
    }
}

// ONE: the layout has depth 1 (one step per launch): two LDS buffers, no
// rotation, side records read straight into registers.
template <class P, bool ONE>
__device__ void tile_steps(const DevView& V, int tile, int t0, int k, int end)
{
    using R = typename P::real;
    using node_t = typename P::node_t;
    using RowsT = Rows<P>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int NT = blockDim.x;
    const int tid = threadIdx.x;
    const TileInfo T = V.tiles[tile];
    // one-step layouts cut tiles of exactly `block` nodes (the last one
    // shorter): the core range follows from the tile index, so the core loads
    // need not wait for the tile record
    const int n0 = ONE ? tile * V.block : T.n0;
    const int cnt = ONE ? min(V.block, V.na - n0) : T.cnt;
    const int S = V.lds_slots;
    const int D = ONE ? 1 : V.depth;

    // LDS: three value buffers, forces [D][n_force], receivers' free field
    // [D+1][n_recv], side records, receivers, material table, hanging-node
    // rows / masses / anchors
    node_t* buf = reinterpret_cast<node_t*>(smem);
    double2* frc = reinterpret_cast<double2*>(buf + (D > 1 ? 3 : 2) * S);
    double2* rff = frc + D * V.lds_force;
    SideT* sides = reinterpret_cast<SideT*>(rff + D * V.lds_recv);
    RecvT* rcv = reinterpret_cast<RecvT*>(sides + V.lds_side);
    V2<R>* lmt = reinterpret_cast<V2<R>*>(rcv + V.lds_recv);
    V2<R>* ft = lmt + V.lds_mat;   // [2 * n_vrow]: transfer terms of a hanging node to its anchors
    int2* va = reinterpret_cast<int2*>(ft + 2 * V.lds_vrows);
    V2<R>* em = reinterpret_cast<V2<R>*>(va + V.lds_virt);   // ELEM tiles: (lambda, mu) per slot
    const bool elem = T.mat_mode != 0;

    node_t* cur = buf;           // u_{t0+s-1}
    node_t* prv = buf + S;       // u_{t0+s-2}
    node_t* nxt = buf + 2 * S;   // u_{t0+s} (a third buffer only when D > 1)
    const void* U0 = V.U[t0 % kStateSlots];
    const void* Um = V.U[(t0 + kStateSlots - 1) % kStateSlots];

    // ---- phase 0: every global load of the launch
    // material table: 4 per tile at a fixed stride (no dependence on the
    // tile record)
    if (tid < 4) lmt[tid] = P::lm(V.tile_lm, 4 * tile + tid);
    // T.lev_end[j] as scalars and register selects (a dynamic index would put
    // the tile record on the scratch stack)
    static_assert(kMaxDepth == 8, "ring levels read as nine scalars");
    const int le0 = T.lev_end[0], le1 = T.lev_end[1], le2 = T.lev_end[2], le3 = T.lev_end[3],
              le4 = T.lev_end[4], le5 = T.lev_end[5], le6 = T.lev_end[6], le7 = T.lev_end[7],
              le8 = T.lev_end[8];
    auto lev_end = [=](int j) {
        return j <= 0 ? le0 : j == 1 ? le1 : j == 2 ? le2 : j == 3 ? le3 : j == 4 ? le4
             : j == 5 ? le5 : j == 6 ? le6 : j == 7 ? le7 : le8;
    };
    const int upd_end = lev_end(D - 1);   // entries ever updated
    // values of every entry (the outer ring is read, never updated), and the
    // lane's entry, updated while it is inside the exact rings
    const int e = tid;
    RowsT rows{};
    R cf = 0;
    if (ONE) {
        // one step: only the core is updated (one entry per lane, the
        // workgroup being at least as wide as the tile); its values,
        // descriptor and coefficient are issued first, the ring entries
        // after the tile record
        node_t cu{}, cp{};
        uint4 d0{};
        if (e < cnt) {
            cu = P::ld(U0, n0 + e);
            cp = P::ld(Um, n0 + e);
            d0 = V.core_desc[n0 + e];
            cf = P::coef(V.core_coef, n0 + e);
        }
        for (int q = cnt + tid; q < T.n_ent; q += NT) {
            const int gid = V.ring_gid[T.ring_off + q - cnt];
            buf[q] = P::ld(U0, gid);
            if (q < T.prv_end) buf[S + q] = P::ld(Um, gid);
            if (elem) em[q] = P::lm(V.node_em, gid);
        }
        if (elem && e < cnt) em[e] = P::lm(V.node_em, n0 + e);
        if (e < cnt) {
            buf[e] = cu;
            buf[S + e] = cp;
            rows.load(d0);
        }
    } else {
        load_entries<P>(V, tile, U0, Um, buf, S);
        if (elem) {
            const int* L = V.elist + (size_t)tile * V.elist_cap;
            for (int q = tid; q < T.n_ent; q += NT) em[q] = P::lm(V.node_em, L[q] & (kPrevFlag - 1));
        }
    }
    if (elem)
        for (int q = tid; q < T.n_virt; q += NT) em[T.n_ent + q] = P::lm(V.virt_em, T.virt_off + q);
    if (!ONE && e < upd_end) {
        uint4 d0{};
        if (e < cnt) {
            d0 = V.core_desc[n0 + e];
            cf = P::coef(V.core_coef, n0 + e);
        } else {
            const int r = T.ring_off + e - cnt;
            d0 = V.ring_desc[r];
            cf = P::coef(V.ring_coef, r);
        }
        rows.load(d0);
    }
    {
        // side records and receivers, copied as 16-byte words
        constexpr int SW = sizeof(SideT) / 16, RW = sizeof(RecvT) / 16;
        if (!ONE) {
            const uint4* src = reinterpret_cast<const uint4*>(V.side + T.side_off);
            uint4* dst = reinterpret_cast<uint4*>(sides);
            for (int q = tid; q < SW * T.n_side; q += NT) dst[q] = src[q];
        }
        const uint4* rsrc = reinterpret_cast<const uint4*>(V.recvs + T.recv_off);
        uint4* rdst = reinterpret_cast<uint4*>(rcv);
        for (int q = copy_lane<ONE>(8 * 64); q < RW * T.n_recv; q += NT) rdst[q] = rsrc[q];
    }
    // anchor slots of the hanging nodes: in a multi-step launch lane q keeps
    // entry q in registers for its sub-steps, loaded without a wait (the LDS
    // copy serves lists longer than the workgroup, and the one-step kernel,
    // which saves the registers)
    int2 my_va{0, 0};
    if (!ONE && tid < T.n_virt) my_va = V.virt[T.virt_off + tid];
    for (int q = copy_lane<ONE>(7 * 64); q < T.n_virt; q += NT) va[q] = V.virt[T.virt_off + q];
    // two lanes per hanging-node row: lane 2 r + side takes half of row r's
    // element terms and the transfer term of its anchor `side`
    const int vr = tid >> 1, vside = tid & 1;
    const bool has_vr = vr < T.n_vrow;
    RowsT vrow{};
    int v_slot = 0, v_force = -1;
    int2 v_anc{0, 0};
    R v_a1m = 0, v_a2c0 = 0, v_a2c1 = 0;   // a1 * M_v, a2 * C_v of the lane's hanging node
    if (has_vr) {
        const VRowT& r = V.vrows[T.vrow_off + vr];
        vrow.load(make_uint4(r.desc[0], r.desc[1], r.desc[2], r.desc[3]));
        v_slot = r.v_slot;
        v_force = r.force;
        if (!ONE) v_anc = make_int2(r.anc[0], r.anc[1]);
        v_a1m = R(r.a1m);
        v_a2c0 = R(r.a2c[0]);
        v_a2c1 = R(r.a2c[1]);
    }
    // forces of steps t0+1 .. t0+k and the receivers' free field of the steps
    // the launch samples, t0 .. t0+max(k,1)-1, from the tables
    for (int q = copy_lane<ONE>(3 * 64); q < k * T.n_force; q += NT) {
        const int s1 = q / T.n_force, f = q - s1 * T.n_force;
        const int slot = V.force_slots[T.force_off + f];
        frc[q] = V.ftab[(size_t)((t0 + s1 + 1) % V.chunk) * V.nforce + slot];
    }
    const int rlane = recv_lane<ONE>();
    // end (several-step layouts, k < depth): the launch also samples step
    // t0 + k, ring 1 being still exact after its k sub-steps
    const int srows = max(k, 1) + (!ONE && end ? 1 : 0);
    for (int q = rlane; q < srows * T.n_recv; q += NT) {
        const int s = q / T.n_recv, r = q - s * T.n_recv;
        const int ff = V.recvs[T.recv_off + r].ff;
        if (ff >= 0) {
            const double4 g = V.fftab[(size_t)((t0 + s) % V.chunk) * V.nff + ff];
            rff[q] = make_double2(g.x, g.y);
        }
    }
    // one step: the lane's side record goes to registers, issued before the
    // barrier
    const bool side = cf < R(0);
    const int sidx = side ? (int)(-cf) - 1 : 0;
    SideT sreg{};
    if (ONE && side) sreg = V.side[T.side_off + sidx];
    // the load phase's barrier; several steps per launch take it at the top
    // of the first sub-step instead, so that the loop-invariant work the
    // compiler places before the loop (descriptor decode) overlaps the other
    // waves' loads
    if (ONE) __syncthreads();

    const int n_recv = T.n_recv;
    // receivers at t0+s-1 are sampled during sub-step s, from cur: on the
    // last wave while the first ones form the hanging-node rows, when the
    // tile has them (that wave is idle there), else before the hanging-node
    // means (a sample takes longer than the means phase)
    const bool recv_in_rows = !ONE && T.n_vrow > 0;

    void* Uk = V.U[(t0 + k) % kStateSlots];
    void* Uk1 = V.U[(t0 + k - 1 + kStateSlots) % kStateSlots];
    // several steps per launch: the record's words (force index, transfers) in
    // registers, so that the force and transfer reads of a sub-step do not
    // wait for the record
    uint32_t m_word = 0, m_tr0 = 0, m_tr1 = 0, m_tr2 = 0, m_tr3 = 0;
    // FP32C also keeps the record's inverse mass and damping there, converted
    // once (FP64 reads them from LDS: its registers are nearly all in use)
    constexpr bool kSideRegs = !ONE && sizeof(R) == sizeof(float);
    R m_invm0 = 0, m_invm1 = 0, m_c20 = 0, m_c21 = 0;
    for (int s = 1; s <= (ONE ? min(k, 1) : k); ++s) {
        if (!ONE && s == 1) {
            __syncthreads();
            if (side) {
                const SideT& q = sides[sidx];
                if (kSideRegs) {
                    m_invm0 = R(q.invm[0]);
                    m_invm1 = R(q.invm[1]);
                    m_c20 = R(q.c2[0]);
                    m_c21 = R(q.c2[1]);
                }
                m_word = q.word;
                m_tr0 = q.tr[0];
                m_tr1 = q.tr[1];
                m_tr2 = q.tr[2];
                m_tr3 = q.tr[3];
            }
        }
        if (!recv_in_rows) sample_receivers<P>(V, cur, rff, rcv, n_recv, t0, s - 1, rlane);
        // hanging-node values at t0+s-1 (means of anchors)
        const int vb = T.n_ent;
        if (ONE) {
            for (int q = tid; q < T.n_virt; q += NT) cur[vb + q] = avg2(cur[va[q].x], cur[va[q].y]);
        } else {
            if (tid < T.n_virt) cur[vb + tid] = avg2(cur[my_va.x], cur[my_va.y]);
            for (int q = tid + NT; q < T.n_virt; q += NT) cur[vb + q] = avg2(cur[va[q].x], cur[va[q].y]);
        }
        __syncthreads();
        const double2* fs = frc + (s - 1) * T.n_force;
        if (T.n_vrow) {
            // A hanging node's reaction, split between its anchors (P2): the
            // term each anchor adds, computed once here.  With v = (a + p) / 2:
            //   g = -(K u)_v + F_v + a1 M_v (2 (u_v - u_a) - (u_{i-2})_v + (u_{i-2})_a)
            //       + a2 C_v ((u_{i-2})_v - (u_{i-2})_a - 2 (u_a - (u_{i-2})_a)),
            // and (u_{i-2})_v - (u_{i-2})_a = ((u_{i-2})_p - (u_{i-2})_a) / 2.
            if (has_vr) {
                const node_t uv = cur[v_slot];
                // -(K u)_v: this lane's half and its partner's (a + b is b + a
                // in floating point, so both lanes hold the same bits)
                const V2<R> gh = vrow.half_row(vside, cur, uv, v_slot, lmt, em, elem);
                V2<R> gv = {gh.x + __shfl_xor(gh.x, 1), gh.y + __shfl_xor(gh.y, 1)};
                if (v_force >= 0) {
                    const double2 F = fs[v_force];
                    gv.x += R(F.x);
                    gv.y += R(F.y);
                }
                const int2 an = ONE ? va[v_slot - vb] : v_anc;
                const int a = vside ? an.y : an.x, p = vside ? an.x : an.y;
                const node_t ua = cur[a], upa = prv[a];
                const V2<R> vel = P::sub(ua, upa);
                const V2<R> du = P::sub(uv, ua);
                const V2<R> dpp = P::sub(prv[p], upa);
                const V2<R> dp = {R(0.5) * dpp.x, R(0.5) * dpp.y};
                V2<R> g = gv;
                g.x += v_a1m * (R(2) * du.x - dp.x) + v_a2c0 * (dp.x - R(2) * vel.x);
                g.y += v_a1m * (R(2) * du.y - dp.y) + v_a2c1 * (dp.y - R(2) * vel.y);
                ft[tid] = {R(0.5) * g.x, R(0.5) * g.y};   // ft[2 * row + side]
            }
            if (recv_in_rows) sample_receivers<P>(V, cur, rff, rcv, n_recv, t0, s - 1, rlane);
            __syncthreads();
        }
        if (e < lev_end(D - s)) {
            const node_t un = cur[e], upn = prv[e];
            V2<R> f = ONE ? rows.apply(cur, un, e, lmt, em, elem)
                          : rows.apply_pairs(cur, un, e, lmt, em, elem);   // -K u
            V2<R> corr;
            int point = 0;
            double pz = 0.0;
            if (!side) {
                corr = {f.x * cf, f.y * cf};                                   // interior node
            } else {
                // the record as scalars (a dynamic index into it would put it
                // on the scratch stack)
                const SideT* qp = ONE ? &sreg : &sides[sidx];
                R invm0, invm1, c20, c21;
                uint32_t word, tr0, tr1, tr2, tr3;
                if (ONE) {
                    invm0 = R(sreg.invm[0]);
                    invm1 = R(sreg.invm[1]);
                    c20 = R(sreg.c2[0]);
                    c21 = R(sreg.c2[1]);
                    word = sreg.word;
                    tr0 = sreg.tr[0];
                    tr1 = sreg.tr[1];
                    tr2 = sreg.tr[2];
                    tr3 = sreg.tr[3];
                } else {
                    invm0 = kSideRegs ? m_invm0 : R(qp->invm[0]);
                    invm1 = kSideRegs ? m_invm1 : R(qp->invm[1]);
                    c20 = kSideRegs ? m_c20 : R(qp->c2[0]);
                    c21 = kSideRegs ? m_c21 : R(qp->c2[1]);
                    word = m_word;
                    tr0 = m_tr0;
                    tr1 = m_tr1;
                    tr2 = m_tr2;
                    tr3 = m_tr3;
                }
                const V2<R> vel = P::sub(un, upn);
                point = (word >> 3) & 3;
                if (word & 32) {
                    const double2 F = fs[(word >> 6) & 1023];
                    if (point == 2) {
                        pz = F.y;                                              // point_displacements
                    } else {
                        f.x += R(F.x);
                        f.y += R(F.y);
                    }
                }
                f.x -= R(c20) * vel.x;
                f.y -= R(c21) * vel.y;
                const int ntr = word & 7;
                // the hanging nodes' reactions, in the anchor's transfer order
                // (terms computed by the hanging-node rows)
#pragma unroll
                for (int t = 0; t < kMaxTransfers; ++t) {
                    if (t >= ntr) break;
                    const uint32_t w = t == 0 ? tr0 : t == 1 ? tr1 : t == 2 ? tr2 : tr3;
                    const V2<R> h = ft[w & 2047];
                    f.x += h.x;
                    f.y += h.y;
                }
                corr = {f.x * R(invm0), f.y * R(invm1)};
            }
            node_t unew = P::advance(un, upn, corr);
            if (point == 2) unew = set_z(unew, pz);
            if (s < k || (!ONE && end)) nxt[e] = unew;
            // the core's u_{t0+k} is stored here only when no barrier follows
            // (a barrier would wait for the store); u_{t0+k-1} after the loop
            if (s == k && e < cnt && (ONE || !end)) P::st(Uk, n0 + e, unew);
        }
        if (s == k) {
            if (!ONE && end) {
                __syncthreads();
                sample_receivers<P>(V, nxt, rff, rcv, n_recv, t0, k, rlane);
            }
            break;
        }
        __syncthreads();
        node_t* t = prv;
        prv = cur;
        cur = nxt;
        nxt = t;
    }
    // after the last sub-step: cur holds u_{t0+k-1} and, with end, nxt holds
    // u_{t0+k} (no lane writes either after its last barrier)
    if (e < cnt) {
        if (!ONE && end) P::st(Uk, n0 + e, nxt[e]);
        if (k > 1) P::st(Uk1, n0 + e, cur[e]);
    }
    flush_receivers(V, rff, rcv, n_recv, t0, srows, rlane);
}

// Advances every tile from step t0 = t0_arg (+ the graph base) by k steps and
// samples the receivers at steps t0 .. t0 + k - 1 (and t0 + k when end is set,
// which needs k < depth).  The one-step kernel is bandwidth-bound, so it is
// compiled for as many waves per SIMD as its registers allow without spills
// (FP32C 7 with 72 VGPRs; FP64 6); the deeper one is latency-bound and keeps
// the registers a 1024-thread workgroup leaves (4 waves per SIMD).
template <class P, bool ONE>
__global__ __launch_bounds__(kMaxThreads) __attribute__((amdgpu_waves_per_eu(ONE ? P::kOneWaves : 4)))
void step_kernel(DevView V, int t0_arg, int use_base, int k, int end)
{
    const int t0 = t0_arg + (use_base ? V.ctrl->base : 0);
    const int b = blockIdx.x;
    const int tile = V.xcd_remap ? (b & 7) * V.tiles_per_xcd + (b >> 3) : b;
    if (tile >= V.node_tiles) return;   // whole block exits together
    tile_steps<P, ONE>(V, tile, t0, k, end);
}

// Receivers at step j (j % recv_interval == 0) straight from the state level
// in HBM: the last step of a run call, which no step launch of the call
// starts from.
template <class P>
__global__ void recv_kernel(DevView V, int j)
{
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= V.nrecv) return;
    const RecvG g = V.rg[r];
    const void* U = V.U[j % kStateSlots];
    double fx = 0.0, fz = 0.0;
    if (g.found) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool hang = (g.hang >> c) & 1;
            const double2 a = P::full(P::ld(U, g.a[c]));
            const double2 b = hang ? P::full(P::ld(U, g.b[c])) : a;
            recv_accum(a, b, hang, g.w[c], fx, fz);
        }
        if (g.ff >= 0) {
            const double4 f = V.fftab[(size_t)(j % V.chunk) * V.nff + g.ff];
            fx = fx + f.x;
            fz = fz + f.y;
        }
    }
    double* out = recv_slot(V, j, r);
    out[0] = fx;
    out[1] = fz;
}

// ---------------------------------------------------------------------------
// Source tables.  The plane-wave and point-source terms depend on time and
// geometry only, so they are evaluated ahead of the steps that use them, for
// a window of `chunk` steps: step j in row j mod chunk.
// ---------------------------------------------------------------------------
// Free field (u0x, u0z, a0x, a0z) of steps [j0, j0 + n) at every point.
__global__ void ff_table_kernel(const FFT* ff, int nff, int interval, double4* tab, int chunk, int j0,
                                int n)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)n * nff) return;
    const int jj = (int)(q / nff), p = (int)(q - (long)jj * nff);
    const int j = j0 + jj;
    tab[(size_t)(j % chunk) * nff + p] = ff_value(ff[p], interval, j);
}

// Forces of steps [j0, j0 + n) at every forced node: plane-wave pairs summed in
// the reference's element order, plus the point source.
__global__ void force_table_kernel(const ForceT* forces, int nforce, const DrmElem* pairs,
                                   const double4* fftab, int nff, const InstParams* inst,
                                   double2* tab, int chunk, int j0, int n)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)n * nforce) return;
    const int jj = (int)(q / nforce), f = (int)(q - (long)jj * nforce);
    const int j = j0 + jj;
    const ForceT F = forces[f];
    const double4* fs = fftab + (size_t)(j % chunk) * nff;
    double fx = 0.0, fz = 0.0;
    for (int p = F.first; p < F.first + F.count; ++p) {
        const double2 g = drm_pair(pairs[p], fs);
        fx += g.x;
        fz += g.y;
    }
    if (F.point) fz += point_ricker(inst[F.inst], j);
    tab[(size_t)(j % chunk) * nforce + f] = make_double2(fx, fz);
}

__global__ void advance_kernel(Ctrl* c, int g)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) c->base += g;
}

__global__ void set_base_kernel(Ctrl* c, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) c->base = v;
}

// ---------------------------------------------------------------------------
// Receiver-sample scratch of the allocation-free run call: a file-scope
// device array, one copy per GPU for the life of the process.  Host code
// never passes the array's name to a kernel (on the host it names a
// stand-in, not device memory): it uses the address hipGetSymbolAddress
// gives on each device, taken once per device at set-up.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

size_t scratch_bytes() { return kScratchBytes; }

// Device address of `scratch` on the current device (which must be
// `device`), looked up on the first call for that device.
hipError_t scratch_address(int device, double** out)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = static_cast<double*>(p);
    }
    *out = addr[device];
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
static int node_grid(const DevView& V)
{
    if (V.na == 0) return 0;
    return V.xcd_remap ? 8 * V.tiles_per_xcd : V.node_tiles;
}

size_t step_lds_bytes(const DevView& V, int precision)
{
    const size_t rb = precision == 0 ? sizeof(double) : sizeof(float);
    return (size_t)(V.depth > 1 ? 3 : 2) * V.lds_slots * 16 +
           (size_t)V.depth * V.lds_force * sizeof(double2) +
           (size_t)V.depth * V.lds_recv * sizeof(double2) +
           (size_t)V.lds_side * sizeof(SideT) + (size_t)V.lds_recv * sizeof(RecvT) +
           (size_t)V.lds_mat * 2 * rb + (size_t)V.lds_vrows * 4 * rb +
           (size_t)V.lds_virt * sizeof(int2) + (V.has_elem ? (size_t)V.lds_slots * 2 * rb : 0);
}

template <class P>
static void launch_one(const DevView& V, int grid, size_t lds, hipStream_t st, int t0, int use_base, int k,
                       int end)
{
    if (V.depth == 1)
        hipLaunchKernelGGL((step_kernel<P, true>), dim3(grid), dim3(V.threads), lds, st, V, t0, use_base, k, 0);
    else
        hipLaunchKernelGGL((step_kernel<P, false>), dim3(grid), dim3(V.threads), lds, st, V, t0, use_base, k,
                           end);
}

hipError_t launch_step(const DevView& V, int precision, int t0, int use_base, int k, hipStream_t st, int end)
{
    const int grid = node_grid(V);
    if (grid == 0) return hipSuccess;
    if (end && k >= V.depth) return hipErrorInvalidValue;   // ring 1 would not be exact at t0 + k
    const size_t lds = step_lds_bytes(V, precision);
    if (precision == 0)
        launch_one<PF64>(V, grid, lds, st, t0, use_base, k, end);
    else
        launch_one<PF32C>(V, grid, lds, st, t0, use_base, k, end);
    return hipGetLastError();
}

hipError_t launch_recv(const DevView& V, int precision, int j, hipStream_t st)
{
    if (V.nrecv <= 0) return hipSuccess;
    const int grid = (V.nrecv + 255) / 256;
    if (precision == 0)
        hipLaunchKernelGGL(recv_kernel<PF64>, dim3(grid), dim3(256), 0, st, V, j);
    else
        hipLaunchKernelGGL(recv_kernel<PF32C>, dim3(grid), dim3(256), 0, st, V, j);
    return hipGetLastError();
}

// Lets every step kernel take up to *max_lds bytes of dynamic LDS, less what
// the kernel declares statically, and lowers *max_lds to the smallest such
// budget (the most any layout may then need).
hipError_t configure_kernels(size_t* max_lds)
{
    const void* fns[4] = {reinterpret_cast<const void*>(step_kernel<PF64, true>),
                          reinterpret_cast<const void*>(step_kernel<PF32C, true>),
                          reinterpret_cast<const void*>(step_kernel<PF64, false>),
                          reinterpret_cast<const void*>(step_kernel<PF32C, false>)};
    size_t budget = *max_lds;
    for (const void* f : fns) {
        hipFuncAttributes a{};
        hipError_t e = hipFuncGetAttributes(&a, f);
        if (e != hipSuccess) return e;
        const size_t dyn = *max_lds > a.sharedSizeBytes ? *max_lds - a.sharedSizeBytes : 0;
        e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
        budget = std::min(budget, dyn);
    }
    *max_lds = budget;
    return hipSuccess;
}

hipError_t launch_tables(const DevView& V, int j0, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (V.nff > 0) {
        const long total = (long)n * V.nff;
        hipLaunchKernelGGL(ff_table_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                           V.ffp, V.nff, V.recv_interval, const_cast<double4*>(V.fftab), V.chunk, j0, n);
    }
    if (V.nforce > 0) {
        const long total = (long)n * V.nforce;
        hipLaunchKernelGGL(force_table_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                           V.forces, V.nforce, V.pairs, V.fftab, V.nff, V.inst,
                           const_cast<double2*>(V.ftab), V.chunk, j0, n);
    }
    return hipGetLastError();
}

hipError_t launch_advance(Ctrl* c, int g, hipStream_t st)
{
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, st, c, g);
    return hipGetLastError();
}

hipError_t launch_set_base(Ctrl* c, int v, hipStream_t st)
{
    hipLaunchKernelGGL(set_base_kernel, dim3(1), dim3(64), 0, st, c, v);
    return hipGetLastError();
}

} // namespace dmp
