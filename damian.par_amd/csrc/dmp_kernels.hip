// dmp_kernels.hip -- gfx950 kernels of the explicit elastic-wave step.
//
// One launch = one time step i of run_explicit (explicit.f90:125-153) for
// every instance of a layout:
//   * node lanes (the bulk): u_i at one active node, gathered from the node's
//     elements: lumped mass / Lysmer damping / K_e u (explicit.f90:129-142),
//     hanging-node reaction transfer (sum_reactanchors, P2), plane-wave force
//     (computed by the previous launch), division by e_mfull (:148);
//   * auxiliary lanes (a few leading workgroups): plane-wave (DRM) forces for
//     step i+1 (wave_reactions, model.f90:745-800, FP64 phase/time) and the
//     receiver sample of step i-1 (visu_field, visualization.f90:156-181).
// State rotates through three buffers (u_i in U[i % 3]); hanging nodes are
// never stored, their value is the mean of their anchors (model.f90:612-624).
//
// The element stiffness is K_e = lambda_e K_lambda + mu_e K_mu with size-
// independent 8x8 tables of the square bilinear plane-strain element
// (formulation.f90:276-313); the tables are derived below from the shape
// functions and fold to immediates.  Rows of K annihilate rigid translations,
// so K_e u_e is evaluated on differences u_j - u_n (exact in FP64 up to
// rounding, and what keeps the FP32 mode accurate).
#include "dmp_internal.h"

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

// ---------------------------------------------------------------------------
// Precision policies.
// ---------------------------------------------------------------------------
template <class R> struct V2 { R x, y; };

template <class R>
__device__ __forceinline__ V2<R> operator+(V2<R> a, V2<R> b) { return {a.x + b.x, a.y + b.y}; }

struct PF64 {
    using real = double;
    using node_t = double2;
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
        xl += R(KL(2 * K, 2 * j)) * d.x + R(KL(2 * K, 2 * j + 1)) * d.y;
        xm += R(KM(2 * K, 2 * j)) * d.x + R(KM(2 * K, 2 * j + 1)) * d.y;
        zl += R(KL(2 * K + 1, 2 * j)) * d.x + R(KL(2 * K + 1, 2 * j + 1)) * d.y;
        zm += R(KM(2 * K + 1, 2 * j)) * d.x + R(KM(2 * K + 1, 2 * j + 1)) * d.y;
    }
    f.x -= lm.x * xl + lm.y * xm;
    f.y -= lm.x * zl + lm.y * zm;
}

// Element rows of one node (or hanging node) from its descriptor.
template <class P, bool COMPACT>
struct Rows {
    using R = typename P::real;
    uint32_t w[4];
    V2<R> lm[4];

    // full format: material words resolved with dependent loads
    __device__ void load(const DevView& V, uint4 d0, uint4 d1, const V2<R>* pal)
    {
        if (COMPACT) {
            const uint32_t ww[4] = {d0.x, d0.y, d0.z, d0.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                w[k] = ww[k];
                // register selects (a dynamic index would put the palette on
                // the scratch stack)
                const uint32_t z = (ww[k] >> 30) & 3;
                const V2<R> a = (z & 1) ? pal[1] : pal[0];
                const V2<R> b = (z & 1) ? pal[3] : pal[2];
                lm[k] = (z & 2) ? b : a;
            }
        } else {
            const uint32_t ww[4] = {d0.x, d0.z, d1.x, d1.z};
            const uint32_t mm[4] = {d0.y, d0.w, d1.y, d1.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                w[k] = ww[k];
                lm[k] = (ww[k] & (1u << 30)) ? P::lm(V.mat_lm, (int)mm[k]) : V2<R>{0, 0};
            }
        }
    }
    __device__ bool present(int k) const
    {
        return COMPACT ? w[k] != 0xFFFFFFFFu : (w[k] & (1u << 30)) != 0;
    }
    __device__ V2<R> apply(const typename P::node_t* val, typename P::node_t uref) const
    {
        V2<R> f{0, 0};
        if (present(0)) slot_rows<P, 0>(val, w[0], lm[0], uref, f);
        if (present(1)) slot_rows<P, 1>(val, w[1], lm[1], uref, f);
        if (present(2)) slot_rows<P, 2>(val, w[2], lm[2], uref, f);
        if (present(3)) slot_rows<P, 3>(val, w[3], lm[3], uref, f);
        return f;
    }
};

// sigma-Ricker of formulation.f90:53-59 at time t - start.
__device__ __forceinline__ double point_ricker(const InstParams& P, int i)
{
    const double t = (double)i * P.dt - P.p_start;
    return exp(-(t * t) / (4 * P.p_sigma)) * ((t * t) - 2 * P.p_sigma) / P.p_den * P.p_amp;
}

__device__ __forceinline__ double2 set_z(double2 u, double z) { return make_double2(u.x, z); }
__device__ __forceinline__ float4 set_z(float4 u, double z)
{
    const float hi = (float)z;
    return make_float4(u.x, hi, u.z, (float)(z - (double)hi));
}

template <class P, bool COMPACT>
__device__ void tile_update(const DevView& V, int tile, int i)
{
    using R = typename P::real;
    using node_t = typename P::node_t;
    using RowsT = Rows<P, COMPACT>;
    extern __shared__ __align__(16) unsigned char smem[];
    node_t* val = reinterpret_cast<node_t*>(smem);
    V2<R>* fv = reinterpret_cast<V2<R>*>(smem + (size_t)V.lds_val_slots * sizeof(node_t));

    const TileInfo T = V.tiles[tile];
    const int B = V.block;
    const int n0 = tile * B;
    const int cnt = min(B, V.na - n0);
    const int tid = threadIdx.x;
    const void* Uc = V.U[(i + 2) % 3];   // u_{i-1}
    const void* Up = V.U[(i + 1) % 3];   // u_{i-2}
    void* Uo = V.U[i % 3];

    V2<R> pal[4];
    if (COMPACT) {
#pragma unroll
        for (int k = 0; k < 4; ++k) pal[k] = P::lm(V.palette, 4 * tile + k);
    }
    node_t un{}, upn{};
    RowsT rows{};
    R cf = 0;
    const int n = n0 + tid;
    if (tid < cnt) {
        un = P::ld(Uc, n);
        upn = P::ld(Up, n);
        if (COMPACT)
            rows.load(V, V.desc[n], uint4{}, pal);
        else
            rows.load(V, V.desc[2 * n], V.desc[2 * n + 1], pal);
        cf = P::coef(V.coef, n);
        val[tid] = un;
    }
    for (int k = tid; k < T.n_extra; k += B) val[cnt + k] = P::ld(Uc, V.extra_gid[T.extra_off + k]);
    __syncthreads();
    if (T.n_virt) {
        const int vb = cnt + T.n_extra;
        for (int k = tid; k < T.n_virt; k += B) {
            const int2 a = V.virt_anch[T.virt_off + k];
            val[vb + k] = avg2(val[a.x], val[a.y]);
        }
        __syncthreads();
    }
    if (T.n_vrow) {
        for (int k = tid; k < T.n_vrow; k += B) {
            const VRow r = V.vrows[T.vrow_off + k];
            RowsT vr;
            vr.load(V, make_uint4(r.desc[0], r.desc[1], r.desc[2], r.desc[3]),
                    make_uint4(r.desc[4], r.desc[5], r.desc[6], r.desc[7]), pal);
            V2<R> g = vr.apply(val, val[r.v_lds]);                     // -(K u)_v
            if (r.drm_slot >= 0) {
                const double2 F = V.drm_force[i & 1][r.drm_slot];
                g.x += R(F.x);
                g.y += R(F.y);
            }
            fv[k] = g;
        }
        __syncthreads();
    }
    if (tid >= cnt) return;

    V2<R> f = rows.apply(val, un);                                    // -K u, own elements
    V2<R> corr;
    int point = 0;
    double pz = 0.0;
    if (cf >= R(0)) {
        corr = {f.x * cf, f.y * cf};                                   // interior node
    } else {
        const SideRec s = V.side[(int)(-cf) - 1];
        if (s.point) {
            point = s.point;
            pz = point_ricker(V.inst[s.inst], i);
            if (point == 1) f.y += R(pz);                              // point_reactions
        }
        const V2<R> vel = P::sub(un, upn);
        if (s.drm_slot >= 0) {
            const double2 F = V.drm_force[i & 1][s.drm_slot];
            f.x += R(F.x);
            f.y += R(F.y);
        }
        f.x -= R(s.c2[0]) * vel.x;
        f.y -= R(s.c2[1]) * vel.y;
        for (int t = s.tr_first; t < s.tr_first + s.tr_count; ++t) {
            const TransferT r = V.tr[t];
            const node_t upp = P::ld(Up, r.partner_gid);
            V2<R> g = fv[r.vrow];
            // u_v - u_n, and (u_{i-2})_v - (u_{i-2})_n = ((up_p - up_n) + 0)/2
            const V2<R> du = P::sub(val[r.v_lds], un);
            const V2<R> dpp = P::sub(upp, upn);
            const V2<R> dp = {R(0.5) * dpp.x, R(0.5) * dpp.y};
            const R a1m = R(r.a1m);
            g.x += a1m * (R(2) * du.x - dp.x) + R(r.a2c[0]) * (dp.x - R(2) * vel.x);
            g.y += a1m * (R(2) * du.y - dp.y) + R(r.a2c[1]) * (dp.y - R(2) * vel.y);
            f.x += R(r.w) * g.x;
            f.y += R(r.w) * g.y;
        }
        corr = {f.x * R(s.invm[0]), f.y * R(s.invm[1])};
    }
    node_t unew = P::advance(un, upn, corr);
    if (point == 2) unew = set_z(unew, pz);                            // point_displacements
    P::st(Uo, n, unew);
}

// ---------------------------------------------------------------------------
// Auxiliary lanes: plane-wave forces and receivers (FP64).
// ---------------------------------------------------------------------------
// Incident + two reflected Ricker trains of pwaves.f90:148-217 at a point with
// precomputed delays v[w]: displacement (column 1) and acceleration (column 3).
__device__ void free_field(const InstParams& P, double t, const double v[3], double& ux,
                           double& uz, double& ax, double& az)
{
    const double fp = P.freq * 3.14159265358979323846;
    ux = uz = ax = az = 0.0;
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        const double tv = t - v[w];
        double s = fp * tv;
        s = s * s;
        const double ex = exp(s);
        const double d = (1.0 - 2.0 * s) / ex;
        const double a = -2.0 * (4.0 * (s * s) - 12.0 * s + 3.0) * (fp * fp) / ex;
        ux = ux + d * P.UX[w] * P.A[w];
        uz = uz + d * P.UZ[w] * P.A[w];
        ax = ax + a * P.UX[w] * P.A[w];
        az = az + a * P.UZ[w] * P.A[w];
    }
}

// Free field at layer node p for step j (time j*dt - t0, model.f90:761).
__device__ void ff_node(const DevView& V, int p, int j)
{
    const FFNode N = V.ff[p];
    const InstParams& P = V.inst[N.inst];
    const double t = (double)j * P.dt - P.t0;
    double4 r;
    free_field(P, t, N.v, r.x, r.y, r.z, r.w);
    V.ffval[((j % 3) + 3) % 3][p] = r;
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
    fx += kx + mx * E.mass;
    fz += kz + mz * E.mass;
}

// Plane-wave force of slot s for step j from the free field of step j.
__device__ void drm_slot_force(const DevView& V, int s, int j)
{
    const DrmSlot S = V.drm[s];
    const double4* F = V.ffval[((j % 3) + 3) % 3];
    double fx = 0.0, fz = 0.0;
    for (int q = S.first; q < S.first + S.count; ++q) {
        const DrmElem E = V.drm_el[q];
        const double4 f[4] = {F[E.c[0]], F[E.c[1]], F[E.c[2]], F[E.c[3]]};
        switch (E.k) {
        case 0: drm_rows<0>(E, f, fx, fz); break;
        case 1: drm_rows<1>(E, f, fx, fz); break;
        case 2: drm_rows<2>(E, f, fx, fz); break;
        default: drm_rows<3>(E, f, fx, fz); break;
        }
    }
    V.drm_force[j & 1][s] = make_double2(fx, fz);
}

template <class P>
__device__ double2 node_value(const DevView& V, const void* U, int c)
{
    if (c < V.na) return P::full(P::ld(U, c));
    const int2 a = V.vanch[c - V.na];
    const double2 x = P::full(P::ld(U, a.x)), y = P::full(P::ld(U, a.y));
    return make_double2((x.x + y.x) / 2.0, (x.y + y.y) / 2.0);
}

template <class P>
__device__ void sample_receiver(const DevView& V, int r, int j, int force)
{
    const int I = V.recv_interval;
    if (I <= 0 || j < 0 || (j % I) != 0) return;
    const Ctrl c = *V.ctrl;
    if (!force && j <= c.block_start) return;
    const int slot = j / I - c.slot_base - 1 + c.slot_offset;
    const RecvRec R = V.recv[r];
    double fx = 0.0, fz = 0.0;
    if (R.c[0] >= 0) {
        const void* U = V.U[j % 3];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double2 u = node_value<P>(V, U, R.c[k]);
            fx = fx + R.w[k] * u.x;
            fz = fz + R.w[k] * u.y;
        }
        if (R.ff) {
            const InstParams& Pp = V.inst[R.inst];
            const double time = (Pp.dt * (double)(j / I)) * (double)I;
            double ux, uz, ax, az;
            free_field(Pp, time - Pp.t0, R.v, ux, uz, ax, az);
            fx = fx + ux;
            fz = fz + uz;
        }
    }
    double* out = V.recv_buf + 2 * ((size_t)slot * V.nrecv + r);
    out[0] = fx;
    out[1] = fz;
}

// mode bits: 1 node lanes; 2 plane-wave pipeline (free field of step i+2 at
// the layer nodes, forces of step i+1 from the free field computed by the
// previous launch); 4 receiver sample of step i-1; 8 force the sample even
// when i-1 <= block_start.
template <class P>
__global__ __launch_bounds__(kMaxBlock) void step_kernel(DevView V, int step_arg, int use_base,
                                                         int mode)
{
    const int i = step_arg + (use_base ? V.ctrl->base : 0);
    int b = blockIdx.x;
    if (b < V.naux_blocks) {
        int t = b * blockDim.x + threadIdx.x;
        if (t < V.nff) {
            if (mode & 2) ff_node(V, t, i + 2);
            return;
        }
        t -= V.nff;
        if (t < V.ndrm) {
            if (mode & 2) drm_slot_force(V, t, i + 1);
            return;
        }
        t -= V.ndrm;
        if (t < V.nrecv && (mode & 4)) sample_receiver<P>(V, t, i - 1, mode & 8);
        return;
    }
    if (!(mode & 1)) return;
    b -= V.naux_blocks;
    const int tile = V.xcd_remap ? (b & 7) * V.tiles_per_xcd + (b >> 3) : b;
    if (tile >= V.node_tiles) return;   // whole block exits together
    if (V.compact)
        tile_update<P, true>(V, tile, i);
    else
        tile_update<P, false>(V, tile, i);
}

__global__ void advance_kernel(Ctrl* c, int g)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) c->base += g;
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
    const size_t vb = precision == 0 ? sizeof(double2) : sizeof(float4);
    const size_t fb = precision == 0 ? sizeof(double2) : sizeof(float2);
    return (size_t)V.lds_val_slots * vb + (size_t)V.lds_vrows * fb;
}

hipError_t launch_step(const DevView& V, int precision, int step_arg, int use_base, int mode,
                       hipStream_t st)
{
    const int grid = V.naux_blocks + ((mode & 1) ? node_grid(V) : 0);
    if (grid == 0) return hipSuccess;
    const size_t lds = (mode & 1) ? step_lds_bytes(V, precision) : 0;
    if (precision == 0)
        hipLaunchKernelGGL(step_kernel<PF64>, dim3(grid), dim3(V.block), lds, st, V, step_arg,
                           use_base, mode);
    else
        hipLaunchKernelGGL(step_kernel<PF32C>, dim3(grid), dim3(V.block), lds, st, V, step_arg,
                           use_base, mode);
    return hipGetLastError();
}

hipError_t launch_advance(Ctrl* c, int g, hipStream_t st)
{
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, st, c, g);
    return hipGetLastError();
}

} // namespace dmp
