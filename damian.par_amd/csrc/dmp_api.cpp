// dmp_api.cpp -- C-ABI of the explicit elastic-wave step library (host side).
//
// Implements the entry points declared in include/damian_step.h: input
// validation and copies, the FP64 precompute of init_explicit
// (explicit.f90:79-109), the device layout (space-filling-curve node order,
// per-node element slots, hanging nodes as virtual nodes), the run loop with
// hipGraph-captured step blocks, state and mesh export, launch tuning,
// batching of many instances into one layout, and the snapshot writers.
#include "damian_step.h"
#include "dmp_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace dmp {
hipError_t launch_step(const DevView& V, int precision, int t0, int use_base, int k, hipStream_t st,
                       int end = 0);
hipError_t launch_recv(const DevView& V, int precision, int j, hipStream_t st);
size_t step_lds_bytes(const DevView& V, int precision);
hipError_t launch_tables(const DevView& V, int j0, int n, hipStream_t st);
hipError_t configure_kernels(size_t* max_lds);
hipError_t launch_advance(Ctrl* c, int g, hipStream_t st);
hipError_t launch_set_base(Ctrl* c, int v, hipStream_t st);
size_t scratch_bytes();
hipError_t scratch_address(int device, double** out);
}

using namespace dmp;

namespace {

// Extra work of a node that anchors hanging nodes, in units of a plain node's,
// when tiles are cut (build_tiles).  Measured on the canyon quadtree (3 steps
// per launch, one tile per compute unit), per step: 2 -> 3.94 us, 4 -> 3.72,
// 5 -> 3.67, 6 -> 3.68, 8 -> 3.93 (FP32C; FP64 alike); the valley of config 3
// is within 1 % for 2 and 5.
constexpr double kAnchorWeight = 5.0;

// Autotuning: timed runs per launch shape (after a warm-up), until there are
// two and kTuneMinSeconds of them, at most kTuneMaxRuns.
constexpr int kTuneMaxRuns = 15;
constexpr double kTuneMinSeconds = 2e-3;

const double kAngZero = (double)0.0001f;   // f_wangzero is a default-real literal (pwaves.f90:8)

thread_local std::string g_lib_error;

struct Error {
    int code;
    std::string msg;
};

Error fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return Error{code, buf};
}

#define HIPCHECK(expr)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw fail(DMP_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

// --------------------------------------------------------------------------
// Free-field constants (pwaves.f90:16-98).
// --------------------------------------------------------------------------
int init_plane_wave(int kind, double mag, double alpha, double beta, double angle, double freq,
                    InstParams& P)
{
    double sine = 0, cose = 0, sinf_ = 0, cosf_ = 0, a = 0, a1 = 0, b = 0, b1 = 0;
    double ui[6];
    P.kind = kind;
    P.freq = freq;
    if (kind == 1) {
        const double e = angle;
        a = mag;
        sine = std::sin(e);
        sinf_ = sine * (beta / alpha);
        const double f = std::asin(sinf_);
        cose = std::cos(e);
        cosf_ = std::cos(f);
        if (cose < kAngZero) return DMP_EWAVE;
        if (std::fabs(sine) < kAngZero) {
            a1 = -a;
            b1 = 0.0;
            cose = 1;
            sine = 0;
        } else {
            const double sin2e = std::sin(2 * e), sin2f = std::sin(2 * f), cos2f = std::cos(2 * f);
            const double r = alpha / beta;
            const double den = sin2e * sin2f + (r * r) * (cos2f * cos2f);
            a1 = (sin2e * sin2f - (r * r) * (cos2f * cos2f)) / den * a;
            b1 = -(2 * r * sin2e * cos2f) / den * a;
        }
        ui[0] = sine; ui[1] = -cose; ui[2] = sine; ui[3] = cose; ui[4] = -cosf_; ui[5] = sinf_;
        const double S[3] = {sine, sine, sinf_}, C[3] = {-cose, cose, cosf_}, V[3] = {alpha, alpha, beta};
        const double A[3] = {a, a1, b1};
        for (int w = 0; w < 3; ++w) {
            P.S[w] = S[w]; P.C[w] = C[w]; P.V[w] = V[w]; P.A[w] = A[w];
            P.UX[w] = ui[2 * w]; P.UZ[w] = ui[2 * w + 1];
        }
    } else if (kind == 2) {
        const double f = angle;
        b = mag;
        sinf_ = std::sin(f);
        const double fcr = std::asin(beta / alpha);
        sine = sinf_ * (alpha / beta);
        const double e = std::asin(sine);
        cose = std::cos(e);
        cosf_ = std::cos(f);
        if (std::fabs(sinf_) < kAngZero) {
            b1 = -b;
            a1 = 0.0;
            cosf_ = 1;
            sinf_ = 0;
        } else if ((fcr - f) > kAngZero) {
            const double sin2e = std::sin(2 * e), sin2f = std::sin(2 * f), cos2f = std::cos(2 * f);
            const double r = alpha / beta;
            const double den = sin2e * sin2f + (r * r) * (cos2f * cos2f);
            b1 = (sin2e * sin2f - (r * r) * (cos2f * cos2f)) / den * b;
            a1 = r * std::sin(4.0 * f) / den * b;
        } else {
            return DMP_EWAVE;
        }
        ui[0] = cosf_; ui[1] = sinf_; ui[2] = sine; ui[3] = cose; ui[4] = -cosf_; ui[5] = sinf_;
        const double S[3] = {sinf_, sine, sinf_}, C[3] = {-cosf_, cose, cosf_}, V[3] = {beta, alpha, beta};
        const double A[3] = {b, a1, b1};
        for (int w = 0; w < 3; ++w) {
            P.S[w] = S[w]; P.C[w] = C[w]; P.V[w] = V[w]; P.A[w] = A[w];
            P.UX[w] = ui[2 * w]; P.UZ[w] = ui[2 * w + 1];
        }
    } else {
        for (int w = 0; w < 3; ++w) {
            P.S[w] = 0; P.C[w] = 0; P.V[w] = 1; P.A[w] = 0; P.UX[w] = 0; P.UZ[w] = 0;
        }
    }
    return DMP_OK;
}

// --------------------------------------------------------------------------
// Host-side layout of one instance (local ids), later packed into a device
// layout with other instances.
// --------------------------------------------------------------------------
// Range of DrmElem pairs received by one node.
struct PairRange {
    int first = 0, count = 0;
};

// Side node before tiling: transfers are a range of TransferRec.
struct HostSide {
    double invm[2];
    double c2[2];
    int tr_first, tr_count;
    int drm;                 // PairRange index (local) or global pair first (packed), -1 none
    int drm_count;           // packed only
    int point, inst;
};

// Hanging node v anchored by a side node: weight w of its reactions
// (sum_reactanchors P2: w = 1/2), its mass and damping, its plane-wave pairs.
struct TransferRec {
    double w;
    double a1m;              // a1 * M_v
    double a2c[2];           // a2 * C_v
    int v;                   // virtual index
    int drm;                 // as HostSide::drm
    int drm_count;
};

// Receiver: corners (active, or hanging) and bilinear weights at the clamped
// point (shapefuncs, formulation.f90:334-363).
struct RecvRec {
    int c[4];            // local: >= 0 active, < 0 -(v+1); packed: active id or na + v
    double w[4];
    int ffi;             // free-field point when the free field is added, else -1
    int inst;
};

struct LocalLayout {
    int na = 0, nv = 0, ne = 0, nnodes = 0;
    std::vector<int> node_int;                 // orig node -> active id, or -(v+1)
    std::vector<int> act_orig, virt_orig;      // local id -> orig node
    std::vector<int> elem_orig;                // internal elem -> orig elem
    std::vector<std::array<int, 4>> node_elems, vnode_elems;   // local elem ids / -1
    std::vector<std::array<int, 2>> vanch;     // local active ids
    std::vector<std::array<int, 4>> ecorner;   // >=0 active, <0 -(v+1)
    std::vector<std::array<int, 4>> ehang;     // virtual local id per edge or -1
    std::vector<std::array<int, 2>> seis;      // (internal elem, kind) in input order
    std::vector<double> lm;                    // (lambda, mu) per material entry
    std::vector<int> emat;
    std::vector<double> invm;                  // per active node (regular) or 0
    std::vector<int> side_of;                  // per active node: side index or -1
    std::vector<HostSide> side;
    std::vector<TransferRec> tr;               // v local, drm local
    std::vector<FFNode> ff;                    // free-field points (layer nodes, receivers)
    std::vector<PairRange> drm;                // pairs received by a node
    std::vector<DrmElem> drm_el;               // c[] local free-field point ids
    std::vector<RecvRec> recv;                 // corners encoded like ecorner
    InstParams ip{};
};

struct Instance;

// Packed (multi-instance) host arrays the tile layout is derived from.
struct PackedHost {
    std::vector<int4> node_elems, vnode_elems, elem_corners;
    std::vector<int> elem_mat;
    std::vector<int2> vanch;
    std::vector<int> node_side;      // HostSide index, or -1 for a plain node
    std::vector<double> invm;        // plain nodes: 1 / e_mfull
    std::vector<HostSide> side;      // drm = global pair first
    std::vector<TransferRec> tr;     // v global virtual index, drm = global pair first
    std::vector<int> vtr;            // per hanging node: one of its TransferRec, or -1
    std::vector<DrmElem> pairs;      // c = global free-field point
    std::vector<FFNode> ff;
    std::vector<RecvRec> recv;       // c: active id or na + virtual index; c[0] = -1: not found
    std::vector<double> lm;          // (lambda, mu) per material entry
    std::vector<InstParams> inst;
    std::vector<ForceT> forces;      // forced nodes (global pair ranges)
    std::vector<int> node_force;     // per active node: force row or -1
    std::vector<int> vforce;         // per hanging node: force row or -1
    std::vector<int> dep_ptr, dep;   // per active node: the other active nodes its update reads
    // (lambda, mu) of the element whose corner 0 (BL) a node is, per active
    // and per hanging node (zero where a node is no element's corner 0)
    std::vector<double> node_em, vnode_em;
};

// Device layout of one or more instances plus the run loop state.
struct Runtime {
    int precision = 0, device = 0;
    hipStream_t stream = nullptr;
    DevView V{};
    std::vector<void*> allocs;
    // nodes per tile, XCD remap, steps per captured graph, steps per launch
    int tuning[4] = {256, 1, 60, 4};
    int init_tuning[4] = {256, 1, 60, 4};   // the shape dmp_init settled on
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0, graph_k = 0;
    int step = 0;
    bool first_run = true;
    // value of the device step base once the enqueued work has run
    // (INT32_MIN: unknown, set before the next graph launch)
    int dev_base = 0;
    size_t lds_limit = 65536;      // dynamic LDS bytes one workgroup may use (the device's, at set-up)
    int table_lo = -1;             // source tables hold steps [table_lo, table_lo + chunk)
    std::vector<int> inst_node_base, inst_vnode_base, inst_elem_base, inst_recv_base;
    std::vector<int> inst_na, inst_nrecv;
    PackedHost host;
    int tiles_block = 0, tiles_depth = 0;   // shape the tile arrays were built for
    std::vector<void*> tile_allocs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;          // device time of the last run (events on `stream`)
    double* scratch = nullptr;     // this device's receiver-sample scratch (allocation-free runs)
    // receiver samples: the kernels write them straight into a ring of
    // mapped pinned host memory (V.recv_buf is its device address); the host
    // moves finished chunks into the caller's buffer while later launches run
    std::vector<hipEvent_t> free_ev;    // events not marking a pending chunk
    double* hstage = nullptr;           // pinned, hstage_cap doubles
    size_t hstage_cap = 0;
    // node numbering of the caller (host copies, for state downloads)
    std::vector<int> h_act_orig, h_virt_orig;
    std::vector<int> recv_missing;   // receivers in no element: sampled as zero
    int* d_act_orig = nullptr;
    int* d_virt_orig = nullptr;
    int* d_elem_orig = nullptr;
    int4* d_ehang = nullptr;
    int4* d_elem_corners = nullptr;
    int2* d_vanch = nullptr;
    int2* d_seis = nullptr;
    std::vector<int> inst_seis_base, inst_nseis;

    ~Runtime() { release(); }

    void release()
    {
        // launches still in flight (a run that failed part way) may write the
        // receiver staging and read the layout: let them finish first
        if (stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
        for (void* p : tile_allocs) (void)hipFree(p);
        tile_allocs.clear();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        ev0 = ev1 = nullptr;
        for (hipEvent_t e : free_ev) (void)hipEventDestroy(e);
        free_ev.clear();
        if (hstage) (void)hipHostFree(hstage);
        hstage = nullptr;
        hstage_cap = 0;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    template <class T> T* dalloc(size_t n)
    {
        if (n == 0) n = 1;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, n * sizeof(T));
        if (e != hipSuccess)
            throw fail(DMP_ENOMEM, "hipMalloc(%zu bytes) failed: %s", n * sizeof(T), hipGetErrorString(e));
        allocs.push_back(p);
        return static_cast<T*>(p);
    }
    template <class T> T* upload(const std::vector<T>& v)
    {
        T* d = dalloc<T>(v.size());
        if (!v.empty()) HIPCHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return d;
    }
    // allocations owned by the current tile layout (freed on rebuild)
    template <class T> T* upload_tile(const std::vector<T>& v)
    {
        T* d = upload(v);
        allocs.pop_back();
        tile_allocs.push_back(d);
        return d;
    }

    void apply_tuning();
    void apply_launch_shape()
    {
        V.xcd_remap = tuning[1];
        V.tiles_per_xcd = (V.node_tiles + 7) / 8;
        drop_graph();
    }
    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        graph_steps = 0;
        graph_k = 0;
    }
};

struct Instance {
    int precision = 0, device = 0;
    std::string err;
    bool have_mesh = false, have_bounds = false, have_time = false;
    int nelems = 0, nnodes = 0;
    std::vector<int> eptr, eind;
    std::vector<uint8_t> ediv;
    std::vector<double> nodes;
    std::vector<int> ematprop;
    int nummat = 0;
    std::vector<double> matconst, eprops;
    std::vector<int> absorbing, hanging, seismic;
    int wave_kind = 0;
    double origin[2] = {0, 0}, t0 = 0, angle = 0, amp = 0, freq = 0, vp = 0, vs = 0;
    double box[3] = {0, 0, 0};
    int point_kind = 0;
    double p_amp = 0, p_sigma = 0, p_start = 0;
    std::vector<int> point_nodes;   // 1-based
    double dt = 0;
    int nsteps = 0, io_interval = 1;
    std::vector<int> recv_elem;
    std::vector<double> recv_xy;
    std::vector<uint8_t> recv_ff;
    // derived
    std::vector<double> mfull;
    std::unique_ptr<LocalLayout> L;
    std::unique_ptr<Runtime> rt;
};

struct Batch {
    std::string err;
    std::vector<Instance*> inst;
    std::unique_ptr<Runtime> rt;
};

// Wall time of a setup phase, printed to stderr when DMP_TIMING is set.
struct PhaseTimer {
    const char* name;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    bool done = false;
    explicit PhaseTimer(const char* n) : name(n) {}
    void stop()
    {
        if (done) return;
        done = true;
        if (std::getenv("DMP_TIMING"))
            std::fprintf(stderr, "dmp setup %-14s %8.1f ms\n", name,
                         std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    ~PhaseTimer() { stop(); }
};

// Host threads for set-up work: DMP_THREADS, else OMP_NUM_THREADS, else the
// hardware threads, at most 16.
int setup_threads()
{
    for (const char* var : {"DMP_THREADS", "OMP_NUM_THREADS"}) {
        const char* v = std::getenv(var);
        if (v && std::atoi(v) > 0) return std::min(std::atoi(v), 16);
    }
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(hw, 16u));
}

// f(i) for i in [0, n) on up to nthreads threads taking indices in order;
// the first exception thrown (lowest index) is rethrown after all threads end.
template <class F> void parallel_items(int n, int nthreads, F&& f)
{
    nthreads = std::max(1, std::min(nthreads, n));
    if (nthreads == 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::exception_ptr> errs(n);
    auto work = [&]() {
        for (int i = next++; i < n; i = next++) {
            try {
                f(i);
            } catch (...) {
                errs[i] = std::current_exception();
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    for (auto& e : errs)
        if (e) std::rethrow_exception(e);
}

uint64_t spread_bits(uint32_t v)
{
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

// --------------------------------------------------------------------------
// init_explicit precompute + layout of one instance.
// --------------------------------------------------------------------------
void build_local(Instance& I)
{
    auto Lp = std::make_unique<LocalLayout>();
    LocalLayout& L = *Lp;
    const int ne = I.nelems, nn = I.nnodes;
    L.ne = ne;
    L.nnodes = nn;
    const double dt = I.dt;
    const double a1 = 1.0 / (dt * dt), a2 = 1.0 / (2.0 * dt);

    // ---- element rows: 4 corners, hanging entries, geometry (model.f90:112-174)
    std::vector<std::array<int, 4>> corner(ne), hang(ne);
    std::vector<double> h(ne);
    for (int e = 0; e < ne; ++e) {
        const int b = I.eptr[e] - 1, en = I.eptr[e + 1] - 1;
        if (b < 0 || en <= b || en > (int)I.eind.size())
            throw fail(DMP_EINVAL, "element %d: bad eptr range", e + 1);
        int nc = 0;
        hang[e] = {-1, -1, -1, -1};
        for (int j = b; j < en; ++j) {
            const int node = I.eind[j] - 1;
            if (node < 0 || node >= nn) throw fail(DMP_EINVAL, "element %d: node id %d out of range", e + 1, node + 1);
            if (!I.ediv[j]) {
                if (nc >= 4) throw fail(DMP_EINVAL, "element %d has more than 4 non-hanging nodes", e + 1);
                corner[e][nc++] = node;
            } else {
                if (nc == 0) throw fail(DMP_EUNSUPPORTED, "element %d: row starts with a hanging node", e + 1);
                if (hang[e][nc - 1] >= 0) throw fail(DMP_EUNSUPPORTED, "element %d: two hanging nodes on one edge", e + 1);
                hang[e][nc - 1] = node;
            }
        }
        if (nc != 4) throw fail(DMP_EINVAL, "element %d has %d non-hanging nodes (need 4)", e + 1, nc);
        const double* p0 = &I.nodes[2 * corner[e][0]];
        const double* p1 = &I.nodes[2 * corner[e][1]];
        const double* p2 = &I.nodes[2 * corner[e][2]];
        const double* p3 = &I.nodes[2 * corner[e][3]];
        h[e] = p2[1] - p0[1];  // elength (model.f90:112-118)
        const double tol = 1e-6 * std::max(1.0, std::fabs(h[e]));
        const double hx = p1[0] - p0[0];
        if (!(h[e] > 0) || std::fabs(hx - h[e]) > tol || std::fabs(p1[1] - p0[1]) > tol ||
            std::fabs(p2[0] - p1[0]) > tol || std::fabs(p3[0] - p0[0]) > tol ||
            std::fabs(p3[1] - p2[1]) > tol)
            throw fail(DMP_EUNSUPPORTED,
                       "element %d is not an axis-aligned square with corners BL, BR, TR, TL", e + 1);
    }

    // ---- hanging nodes (m_hanging order defines virtual ids)
    const int nh = (int)I.hanging.size() / 3;
    std::vector<int> hang_idx(nn, -1);
    for (int i = 0; i < nh; ++i) {
        const int hn = I.hanging[3 * i] - 1;
        if (hn < 0 || hn >= nn) throw fail(DMP_EINVAL, "hanging entry %d out of range", i + 1);
        if (hang_idx[hn] >= 0) throw fail(DMP_EUNSUPPORTED, "node %d listed twice in m_hanging", hn + 1);
        hang_idx[hn] = i;
    }
    for (int i = 0; i < nh; ++i)
        for (int k = 1; k <= 2; ++k) {
            const int an = I.hanging[3 * i + k] - 1;
            if (an < 0 || an >= nn) throw fail(DMP_EINVAL, "hanging anchor %d out of range", i + 1);
            if (hang_idx[an] >= 0)
                throw fail(DMP_EUNSUPPORTED, "hanging node %d has a hanging anchor %d", I.hanging[3 * i], an + 1);
        }
    {
        int flagged = 0;
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < 4; ++k)
                if (hang[e][k] >= 0) {
                    ++flagged;
                    if (hang_idx[hang[e][k]] < 0)
                        throw fail(DMP_EINVAL, "flagged node %d missing from m_hanging", hang[e][k] + 1);
                }
        if (flagged != nh) throw fail(DMP_EINVAL, "m_hanging has %d entries, ediv flags %d", nh, flagged);
    }

    // ---- node order: active nodes along a Morton curve
    double hmin = 1e300, xmin = 1e300, zmin = 1e300;
    for (int e = 0; e < ne; ++e) hmin = std::min(hmin, h[e]);
    for (int n = 0; n < nn; ++n) {
        xmin = std::min(xmin, I.nodes[2 * n]);
        zmin = std::min(zmin, I.nodes[2 * n + 1]);
    }
    if (ne == 0) hmin = 1.0;
    auto key_of = [&](double x, double z) {
        const double qx = std::llround((x - xmin) / hmin * 2.0);
        const double qz = std::llround((z - zmin) / hmin * 2.0);
        return (spread_bits((uint32_t)qx) << 1) | spread_bits((uint32_t)qz);
    };
    std::vector<std::pair<uint64_t, int>> ord;
    ord.reserve(nn);
    for (int n = 0; n < nn; ++n)
        if (hang_idx[n] < 0) ord.push_back({key_of(I.nodes[2 * n], I.nodes[2 * n + 1]), n});
    std::sort(ord.begin(), ord.end());
    L.na = (int)ord.size();
    L.nv = nh;
    L.node_int.assign(nn, 0);
    L.act_orig.resize(L.na);
    for (int a = 0; a < L.na; ++a) {
        L.act_orig[a] = ord[a].second;
        L.node_int[ord[a].second] = a;
    }
    L.virt_orig.resize(nh);
    L.vanch.resize(nh);
    for (int i = 0; i < nh; ++i) {
        const int hn = I.hanging[3 * i] - 1;
        L.virt_orig[i] = hn;
        L.node_int[hn] = -(i + 1);
    }
    for (int i = 0; i < nh; ++i)
        L.vanch[i] = {L.node_int[I.hanging[3 * i + 1] - 1], L.node_int[I.hanging[3 * i + 2] - 1]};

    // ---- element order: Morton key of the BL corner
    std::vector<std::pair<uint64_t, int>> eord(ne);
    for (int e = 0; e < ne; ++e)
        eord[e] = {key_of(I.nodes[2 * corner[e][0]], I.nodes[2 * corner[e][0] + 1]), e};
    std::sort(eord.begin(), eord.end());
    L.elem_orig.resize(ne);
    std::vector<int> elem_int(ne);
    for (int i = 0; i < ne; ++i) {
        L.elem_orig[i] = eord[i].second;
        elem_int[eord[i].second] = i;
    }
    L.ecorner.resize(ne);
    L.ehang.resize(ne);
    L.node_elems.assign(L.na, {-1, -1, -1, -1});
    L.vnode_elems.assign(nh, {-1, -1, -1, -1});
    for (int i = 0; i < ne; ++i) {
        const int e = L.elem_orig[i];
        for (int k = 0; k < 4; ++k) {
            const int node = corner[e][k];
            const int id = L.node_int[node];
            L.ecorner[i][k] = id;
            auto& slot = id >= 0 ? L.node_elems[id][k] : L.vnode_elems[-id - 1][k];
            if (slot >= 0)
                throw fail(DMP_EUNSUPPORTED, "node %d is corner %d of two elements", node + 1, k + 1);
            slot = i;
            L.ehang[i][k] = hang[e][k] >= 0 ? hang_idx[hang[e][k]] : -1;
        }
    }

    // ---- materials
    const bool per_elem = !I.eprops.empty();
    std::vector<std::array<double, 5>> emc(ne);
    for (int e = 0; e < ne; ++e) {
        if (per_elem) {
            for (int q = 0; q < 5; ++q) emc[e][q] = I.eprops[5 * e + q];
        } else {
            const int m = I.ematprop[e] - 1;
            if (m < 0 || m >= I.nummat) throw fail(DMP_EINVAL, "element %d: material %d out of range", e + 1, m + 1);
            for (int q = 0; q < 5; ++q) emc[e][q] = I.matconst[5 * m + q];
        }
    }
    L.emat.resize(ne);
    if (per_elem) {
        L.lm.resize(2 * (size_t)ne);
        for (int i = 0; i < ne; ++i) {
            const int e = L.elem_orig[i];
            L.lm[2 * i] = emc[e][1];
            L.lm[2 * i + 1] = emc[e][0];
            L.emat[i] = i;
        }
    } else {
        L.lm.resize(2 * (size_t)I.nummat);
        for (int m = 0; m < I.nummat; ++m) {
            L.lm[2 * m] = I.matconst[5 * m + 1];
            L.lm[2 * m + 1] = I.matconst[5 * m + 0];
        }
        for (int i = 0; i < ne; ++i) L.emat[i] = I.ematprop[L.elem_orig[i]] - 1;
    }

    // ---- lumped mass, Lysmer damping, effective mass in reference order
    //      (getmassof/getdampingof model.f90:564-604; explicit.f90:96-106)
    std::vector<double> M(nn, 0.0), C(2 * (size_t)nn, 0.0), damp(8 * (size_t)ne, 0.0), mass(ne);
    for (int e = 0; e < ne; ++e) {
        const double sc = h[e] / 2.0;
        mass[e] = (sc * sc) * emc[e][2];
        for (int k = 0; k < 4; ++k) {
            const int ab = I.absorbing[corner[e][k]];
            if (ab == 1 || ab == 2) {
                double r0 = emc[e][3] * emc[e][2] * h[e], r1 = emc[e][4] * emc[e][2] * h[e];
                if (ab == 1) std::swap(r0, r1);
                damp[8 * e + 2 * k] = r0;
                damp[8 * e + 2 * k + 1] = r1;
            }
        }
    }
    I.mfull.assign(2 * (size_t)nn, 0.0);
    for (int e = 0; e < ne; ++e)
        for (int j = 0; j < 8; ++j) {
            const int d = 2 * corner[e][j / 2] + (j % 2);
            I.mfull[d] = I.mfull[d] + a2 * damp[8 * e + j] + a1 * mass[e];
            C[d] += damp[8 * e + j];
            if ((j % 2) == 0) M[corner[e][j / 2]] += mass[e];
        }
    for (int i = 0; i < nh; ++i) {   // sum_reactanchors, P2 (communicator.f90:1232)
        const int hn = I.hanging[3 * i] - 1, x1 = I.hanging[3 * i + 1] - 1, x2 = I.hanging[3 * i + 2] - 1;
        for (int c = 0; c < 2; ++c) {
            const double half = I.mfull[2 * hn + c] / 2.0;
            I.mfull[2 * x1 + c] = I.mfull[2 * x1 + c] + half;
            I.mfull[2 * x2 + c] = I.mfull[2 * x2 + c] + half;
        }
    }

    // ---- instance constants (initplanewaves, pwaves.f90:16-98)
    L.ip = InstParams{};
    L.ip.dt = dt;
    L.ip.t0 = I.t0;
    L.ip.ox = I.origin[0];
    L.ip.oz = I.origin[1];
    L.ip.recv_interval = I.io_interval;
    if (init_plane_wave(I.wave_kind, I.amp, I.vp, I.vs, I.angle, I.freq, L.ip) != DMP_OK)
        throw fail(DMP_EWAVE, I.wave_kind == 1 ? "grazing P incidence not implemented"
                                               : "post-critical SV incidence not implemented");
    L.ip.point_kind = I.point_kind;
    L.ip.p_amp = I.p_amp;
    L.ip.p_sigma = I.p_sigma;
    L.ip.p_start = I.p_start;
    L.ip.p_den = 8 * std::sqrt(3.14159265358979323846) * std::pow(I.p_sigma, 2.5);
    // delays of the three trains at a point (x, z) relative to m_worigin
    auto delays = [&](double x, double z, double* v) {
        for (int w = 0; w < 3; ++w) v[w] = (x * L.ip.S[w] + z * L.ip.C[w]) / L.ip.V[w];
    };

    // ---- plane-wave slots: every corner of a seismic element
    const int ns = (int)I.seismic.size() / 2;
    std::vector<int> ff_of(nn, -1);
    std::vector<int> drm_of(nn, -1);
    std::vector<std::pair<int, DrmElem>> pairs;
    L.seis.resize(ns);
    for (int s = 0; s < ns; ++s) {
        const int e = I.seismic[2 * s] - 1, kind = I.seismic[2 * s + 1];
        if (e < 0 || e >= ne) throw fail(DMP_EINVAL, "m_seismic entry %d: element out of range", s + 1);
        int ext;
        switch (kind) {
        case 2: ext = 1 | 8; break;        // SVERL: BL, TL
        case 4: ext = 2 | 4; break;        // SVERR: BR, TR
        case 1: ext = 4 | 8; break;        // SHORI: TR, TL
        case 6: ext = 1 | 4 | 8; break;    // CORNL
        case 5: ext = 2 | 4 | 8; break;    // CORNR
        default: throw fail(DMP_EINVAL, "m_seismic entry %d: unknown kind %d", s + 1, kind);
        }
        L.seis[s] = {elem_int[e], kind};
        if (I.wave_kind == 0) continue;
        for (int j = 0; j < 4; ++j) {
            const int node = corner[e][j];
            if (ff_of[node] < 0) {
                ff_of[node] = (int)L.ff.size();
                FFNode F{};
                delays(I.nodes[2 * node] - I.origin[0], I.nodes[2 * node + 1] - I.origin[1], F.v);
                F.inst = 0;
                F.recv_time = 0;
                L.ff.push_back(F);
            }
        }
        for (int k = 0; k < 4; ++k) {
            const int node = corner[e][k];
            if (drm_of[node] < 0) drm_of[node] = 0;  // mark; numbered below
            DrmElem E{};
            for (int j = 0; j < 4; ++j) E.c[j] = ff_of[corner[e][j]];
            E.lam = emc[e][1];
            E.mu = emc[e][0];
            E.mass = mass[e];
            E.k = k;
            E.ext = ext;
            pairs.push_back({node, E});
        }
    }
    // number slots in internal node order (active first, then virtual)
    {
        int next = 0;
        for (int a = 0; a < L.na; ++a)
            if (drm_of[L.act_orig[a]] >= 0) drm_of[L.act_orig[a]] = next++;
        for (int v = 0; v < nh; ++v)
            if (drm_of[L.virt_orig[v]] >= 0) drm_of[L.virt_orig[v]] = next++;
        L.drm.assign(next, PairRange{});
        std::stable_sort(pairs.begin(), pairs.end(), [&](const auto& x, const auto& y) {
            return drm_of[x.first] < drm_of[y.first];
        });
        L.drm_el.resize(pairs.size());
        for (size_t q = 0; q < pairs.size(); ++q) {
            const int s = drm_of[pairs[q].first];
            if (L.drm[s].count == 0) L.drm[s].first = (int)q;
            L.drm[s].count++;
            L.drm_el[q] = pairs[q].second;
        }
    }

    // ---- per active node: plain reciprocal or side record
    std::vector<std::vector<int>> anchored(L.na);
    for (int i = 0; i < nh; ++i) {
        anchored[L.vanch[i][0]].push_back(i);
        anchored[L.vanch[i][1]].push_back(i);
    }
    std::vector<char> is_point(nn, 0);
    for (int q : I.point_nodes) {
        if (q < 1 || q > nn) throw fail(DMP_EINVAL, "point node %d out of range", q);
        if (hang_idx[q - 1] >= 0) throw fail(DMP_EUNSUPPORTED, "point node %d is a hanging node", q);
        is_point[q - 1] = 1;
    }
    L.invm.assign(L.na, 0.0);
    L.side_of.assign(L.na, -1);
    for (int a = 0; a < L.na; ++a) {
        const int n = L.act_orig[a];
        const double m0 = I.mfull[2 * n], m1 = I.mfull[2 * n + 1];
        if (!(m0 > 0) || !(m1 > 0))
            throw fail(DMP_EINVAL, "node %d has no mass (not a corner of any element)", n + 1);
        const bool plain = C[2 * n] == 0.0 && C[2 * n + 1] == 0.0 && anchored[a].empty() &&
                           drm_of[n] < 0 && m0 == m1 && !is_point[n];
        if (plain) {
            L.invm[a] = 1.0 / m0;
            continue;
        }
        HostSide s{};
        s.invm[0] = 1.0 / m0;
        s.invm[1] = 1.0 / m1;
        s.c2[0] = 2.0 * a2 * C[2 * n];
        s.c2[1] = 2.0 * a2 * C[2 * n + 1];
        s.drm = drm_of[n];
        s.point = is_point[n] ? I.point_kind : 0;
        s.inst = 0;
        s.tr_first = (int)L.tr.size();
        s.tr_count = (int)anchored[a].size();
        if (s.tr_count > kMaxTransfers)
            throw fail(DMP_EUNSUPPORTED, "node %d anchors %d hanging nodes (max %d)", n + 1,
                       s.tr_count, kMaxTransfers);
        for (int v : anchored[a]) {
            const int hn = L.virt_orig[v];
            TransferRec r{};
            r.w = 0.5;
            r.a1m = a1 * M[hn];
            r.a2c[0] = a2 * C[2 * hn];
            r.a2c[1] = a2 * C[2 * hn + 1];
            r.v = v;
            r.drm = drm_of[hn];
            L.tr.push_back(r);
        }
        L.side_of[a] = (int)L.side.size();
        L.side.push_back(s);
    }

    // ---- receivers (visu_field: shapeparams/shapefuncs, formulation.f90:334-363)
    const int nr = (int)I.recv_elem.size();
    L.recv.resize(nr);
    for (int r = 0; r < nr; ++r) {
        RecvRec R{};
        const int e = I.recv_elem[r] - 1;
        R.ffi = -1;
        R.inst = 0;
        if (e < 0 || e >= ne) {
            R.c[0] = R.c[1] = R.c[2] = R.c[3] = -1;
            R.c[0] = INT32_MIN;
        } else {
            if (I.recv_ff[r] && I.wave_kind != 0) {
                FFNode F{};
                delays(I.recv_xy[2 * r] - I.origin[0], I.recv_xy[2 * r + 1] - I.origin[1], F.v);
                F.inst = 0;
                F.recv_time = 1;
                R.ffi = (int)L.ff.size();
                L.ff.push_back(F);
            }
            const double* n1 = &I.nodes[2 * corner[e][0]];
            const double* n3 = &I.nodes[2 * corner[e][2]];
            double par[2] = {(I.recv_xy[2 * r] - n1[0]) / (n3[0] - n1[0]),
                             (I.recv_xy[2 * r + 1] - n1[1]) / (n3[1] - n1[1])};
            for (int c = 0; c < 2; ++c) {
                par[c] = 2.0 * par[c] - 1.0;
                par[c] = std::max(-1.0, par[c]);
                par[c] = std::min(1.0, par[c]);
            }
            const double rr = par[0], ss = par[1];
            R.w[0] = (1 - rr) * (1 - ss) / 4;
            R.w[1] = (1 + rr) * (1 - ss) / 4;
            R.w[2] = (1 + rr) * (1 + ss) / 4;
            R.w[3] = (1 - rr) * (1 + ss) / 4;
            for (int k = 0; k < 4; ++k) R.c[k] = L.node_int[corner[e][k]];
        }
        L.recv[r] = R;
    }

    I.L = std::move(Lp);
}

// --------------------------------------------------------------------------
// Tiles of B core nodes and their rings for launches of up to K steps
// (R.tuning[0], R.tuning[3]).  Tiles are independent once their core ranges
// are cut: they are built in chunks on host threads, each chunk into lists
// of its own, and the lists are joined in tile order (the same layout as a
// serial build).
// --------------------------------------------------------------------------
// A chunk's share of the tile lists; TileInfo offsets are chunk-local until
// the join.
struct TilePart {
    std::vector<TileInfo> tiles;
    std::vector<int> ring_gid;
    std::vector<uint32_t> ring_full;   // 8 words per ring entry (full format)
    std::vector<double> ring_coef;
    std::vector<SideT> sides;
    std::vector<VRowT> vrows;
    std::vector<int2> virt;
    std::vector<int> tile_mats;
    std::vector<int> force_slots;
    std::vector<RecvT> recvs;
    std::vector<int> virt_ids;         // hanging node of each virt entry
    std::vector<uint32_t> vrow_full;   // 8 words per hanging-node row (full format)
    int max_slots = 2, max_vrows = 0, max_virt = 0, max_force = 0;
    int max_side = 0, max_recv = 0, max_mat = 0, max_ent = 0;
};

// Per-thread marks: entry i holds t + 1 once tile t has visited it, so they
// never need clearing.  The arrays are calloc'ed: untouched pages stay the
// shared zero page, so a worker pays only for the nodes its tiles reach.
struct MarkArray {
    int* p = nullptr;
    ~MarkArray() { std::free(p); }
    void zeroed(size_t count)
    {
        std::free(p);
        p = static_cast<int*>(std::calloc(std::max<size_t>(count, 1), sizeof(int)));
        if (!p) throw std::bad_alloc();
    }
    int& operator[](size_t i) { return p[i]; }
};

struct TileScratch {
    MarkArray stamp, slot, vstamp, vslot, vrstamp, vridx, fstamp, flocal, matstamp, matlocal;
    std::vector<int> ent, vl, vlist;
    bool ready = false;
    void init(size_t na, size_t nv, size_t nf, size_t nm)
    {
        if (ready) return;
        for (MarkArray* a : {&stamp, &slot}) a->zeroed(na);
        for (MarkArray* a : {&vstamp, &vslot, &vrstamp, &vridx}) a->zeroed(nv);
        for (MarkArray* a : {&fstamp, &flocal}) a->zeroed(nf);
        for (MarkArray* a : {&matstamp, &matlocal}) a->zeroed(nm);
        ready = true;
    }
};

void build_tiles(Runtime& R)
{
    PhaseTimer pt("tiles");
    const PackedHost& H = R.host;
    DevView& V = R.V;
    for (void* p : R.tile_allocs) (void)hipFree(p);
    R.tile_allocs.clear();
    R.tiles_block = 0;   // no valid layout until this build completes
    R.tiles_depth = 0;
    const int NA = V.na, NV = V.nv;
    const int B = R.tuning[0], K = R.tuning[3];
    const int ntiles = (NA + B - 1) / B;
    const int NR = (int)H.recv.size();
    const int NFF = (int)H.ff.size();
    const int NM = (int)(H.lm.size() / 2);
    auto quad = [](int4 q) { return std::array<int, 4>{q.x, q.y, q.z, q.w}; };

    // Tiles: ceil(NA / B) contiguous ranges of the Morton order.  With several
    // steps per launch and at most two tiles per compute unit, the slowest
    // tile sets the launch time, so tiles are cut at equal work: a node that
    // anchors hanging nodes counts 1 + kAnchorWeight, since its tile also
    // evaluates the hanging nodes' rows and transfer terms every sub-step.
    // Otherwise (one step per launch, or many tiles per compute unit, where
    // bandwidth bounds the step) tiles keep B nodes: a block is sized for the
    // largest tile.
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, R.device) != hipSuccess || ncu < 1)
        ncu = 256;
    std::vector<int> tstart(ntiles + 1, 0);
    if (K == 1 || ntiles > 2 * ncu) {
        for (int t = 1; t <= ntiles; ++t) tstart[t] = std::min(t * B, NA);
    } else {
        std::vector<double> cum(NA + 1, 0.0);
        for (int i = 0; i < NA; ++i) {
            const int si = H.node_side[i];
            const bool anchor = si >= 0 && H.side[si].tr_count > 0;
            cum[i + 1] = cum[i] + 1.0 + (anchor ? kAnchorWeight : 0.0);
        }
        for (int t = 1; t < ntiles; ++t) {
            const double goal = cum[NA] * t / ntiles;
            int i = (int)(std::lower_bound(cum.begin(), cum.end(), goal) - cum.begin());
            i = std::max(i, tstart[t - 1] + 1);   // every tile keeps a node
            i = std::min(i, NA - (ntiles - t));
            tstart[t] = i;
        }
        tstart[ntiles] = NA;
    }
    auto tile_of = [&](int node) {
        return (int)(std::upper_bound(tstart.begin(), tstart.end(), node) - tstart.begin()) - 1;
    };

    // receivers: sampled by the tile that owns the first stored node of corner 0
    std::vector<std::vector<int>> tile_recv(std::max(ntiles, 1));
    for (int r = 0; r < NR; ++r) {
        const int c = H.recv[r].c[0];
        const int owner = c < 0 ? 0 : tile_of(c < NA ? c : H.vanch[c - NA].x);
        tile_recv[owner].push_back(r);
    }

    std::vector<uint32_t> core_full(8 * (size_t)NA, 0);   // per node, written by its tile
    std::vector<double> core_coef(NA, 0.0);

    // one tile into part P (offsets local to P)
    auto build_one = [&](int t, TileScratch& S, TilePart& P) {
        const int n0 = tstart[t], cnt = tstart[t + 1] - tstart[t];
        std::vector<int>& ent = S.ent;
        std::vector<int>& vl = S.vl;
        std::vector<int>& vlist = S.vlist;
        MarkArray& stamp = S.stamp;
        MarkArray& slot = S.slot;
        const int mark = t + 1;
        TileInfo T{};
        T.n0 = n0;
        T.cnt = cnt;
        ent.clear();
        vl.clear();
        vlist.clear();
        auto add_ent = [&](int g) {
            if (stamp[g] == mark) return;
            stamp[g] = mark;
            slot[g] = (int)ent.size();
            ent.push_back(g);
        };
        // ring 0: the core; ring 1: what the core reads, and the stored nodes
        // of the tile's receivers (sampled at sub-steps 0 .. k-1, while ring 1
        // is still exact); ring j+1: what ring j reads
        for (int i = 0; i < cnt; ++i) add_ent(n0 + i);
        T.lev_end[0] = (int)ent.size();
        size_t lo = 0;
        for (int j = 1; j <= kMaxDepth; ++j) {
            if (j <= K) {
                const size_t hi = ent.size();
                for (size_t q = lo; q < hi; ++q) {
                    const int g = ent[q];
                    for (int d = H.dep_ptr[g]; d < H.dep_ptr[g + 1]; ++d) add_ent(H.dep[d]);
                }
                if (j == 1)
                    for (int r : tile_recv[t])
                        for (int c : H.recv[r].c) {
                            if (c < 0) continue;
                            if (c < NA) {
                                add_ent(c);
                            } else {
                                add_ent(H.vanch[c - NA].x);
                                add_ent(H.vanch[c - NA].y);
                            }
                        }
                lo = hi;
            }
            T.lev_end[j] = (int)ent.size();
        }
        const int N = (int)ent.size();
        const int upd_end = T.lev_end[K - 1];
        // u_{t0-1} is read for the updated entries and for the other anchors of
        // their hanging nodes: those of the outer ring go first in it
        {
            std::vector<int> first, rest;
            for (int q = upd_end; q < N; ++q) {
                const int g = ent[q];
                bool partner = false;
                for (int d = H.dep_ptr[g]; d < H.dep_ptr[g + 1] && !partner; ++d) {
                    const int n = H.dep[d];   // dep is symmetric for anchors: n reads g
                    if (stamp[n] != mark || slot[n] >= upd_end || H.node_side[n] < 0) continue;
                    const HostSide& hs = H.side[H.node_side[n]];
                    for (int x = hs.tr_first; x < hs.tr_first + hs.tr_count; ++x) {
                        const int2 a = H.vanch[H.tr[x].v];
                        if ((a.x == n && a.y == g) || (a.y == n && a.x == g)) partner = true;
                    }
                }
                (partner ? first : rest).push_back(g);
            }
            int q = upd_end;
            for (int g : first) {
                ent[q] = g;
                slot[g] = q++;
            }
            T.prv_end = q;
            for (int g : rest) {
                ent[q] = g;
                slot[g] = q++;
            }
        }
        if (upd_end > kMaxThreads)
            throw fail(DMP_EUNSUPPORTED, "tile %d updates %d nodes for %d steps per launch (max %d): "
                       "use fewer steps per launch or a smaller block", t, upd_end, K, kMaxThreads);
        // hanging nodes the updated entries read, and the hanging-node rows
        // their anchors take
        auto need_v = [&](int v) {
            if (S.vstamp[v] == mark) return;
            S.vstamp[v] = mark;
            S.vslot[v] = N + (int)vl.size();
            vl.push_back(v);
        };
        auto elem_virts = [&](int e) {
            if (e < 0) return;
            for (int c : quad(H.elem_corners[e]))
                if (c >= NA) need_v(c - NA);
        };
        for (int q = 0; q < upd_end; ++q) {
            const int g = ent[q];
            for (int e : quad(H.node_elems[g])) elem_virts(e);
            const int si = H.node_side[g];
            if (si < 0) continue;
            const HostSide& hs = H.side[si];
            for (int x = hs.tr_first; x < hs.tr_first + hs.tr_count; ++x) {
                const int v = H.tr[x].v;
                need_v(v);
                if (S.vrstamp[v] != mark) {
                    S.vrstamp[v] = mark;
                    S.vridx[v] = (int)vlist.size();
                    vlist.push_back(v);
                    for (int e : quad(H.vnode_elems[v])) elem_virts(e);
                }
            }
        }
        const int nslots = N + (int)vl.size();
        if (nslots > kMaxSlots)
            throw fail(DMP_EUNSUPPORTED, "tile %d needs %d LDS slots (max %d): use a smaller block "
                       "or fewer steps per launch", t, nslots, kMaxSlots);
        // the value buffers alone must fit the LDS budget (checked in full at the end)
        if ((size_t)(K > 1 ? 3 : 2) * ((nslots + 1) & ~1) * 16 > R.lds_limit)
            throw fail(DMP_EUNSUPPORTED, "tile %d needs more than %zu B of LDS for %d steps per launch",
                       t, R.lds_limit, K);
        if (2 * (int)vlist.size() > kMaxThreads)
            throw fail(DMP_EUNSUPPORTED, "tile %d needs %d hanging-node rows (max %d): use a smaller block",
                       t, (int)vlist.size(), kMaxThreads / 2);
        auto fslot = [&](int id) {
            if (id < NA) {
                if (stamp[id] != mark) throw fail(DMP_EINVAL, "internal: node %d outside tile %d", id, t);
                return slot[id];
            }
            if (S.vstamp[id - NA] != mark) throw fail(DMP_EINVAL, "internal: hanging node outside tile %d", t);
            return S.vslot[id - NA];
        };
        // material table of the rows this tile evaluates
        T.mat_off = (int)P.tile_mats.size();
        auto mat_loc = [&](int m) {
            if (S.matstamp[m] != mark) {
                S.matstamp[m] = mark;
                S.matlocal[m] = (int)P.tile_mats.size() - T.mat_off;
                P.tile_mats.push_back(m);
            }
            return S.matlocal[m];
        };
        // full format: w[2k] = slots | present bit, w[2k+1] = material table index
        auto pack = [&](int4 q4, uint32_t* w) {
            const auto q = quad(q4);
            for (int k = 0; k < 4; ++k) {
                w[2 * k] = 0;
                w[2 * k + 1] = 0;
                if (q[k] < 0) continue;
                const auto c = quad(H.elem_corners[q[k]]);
                uint32_t word = 1u << 30;
                int sh = 0;
                for (int j = 0; j < 4; ++j)
                    if (j != k) {
                        word |= (uint32_t)fslot(c[j]) << sh;
                        sh += 10;
                    }
                w[2 * k] = word;
                w[2 * k + 1] = (uint32_t)mat_loc(H.elem_mat[q[k]]);
            }
        };
        // forced nodes: rows of the force table, listed once per tile
        T.force_off = (int)P.force_slots.size();
        T.side_off = (int)P.sides.size();
        auto add_force = [&](int fs) {
            if (S.fstamp[fs] != mark) {
                S.fstamp[fs] = mark;
                S.flocal[fs] = (int)P.force_slots.size() - T.force_off;
                P.force_slots.push_back(fs);
                if (S.flocal[fs] > 1023) throw fail(DMP_EUNSUPPORTED, "tile %d has more than 1024 forced nodes", t);
            }
            return S.flocal[fs];
        };
        T.vrow_off = (int)P.vrows.size();
        T.n_vrow = (int)vlist.size();
        for (int v : vlist) {
            const TransferRec& x = H.tr[H.vtr[v]];
            VRowT r{};
            r.a1m = x.a1m;
            r.a2c[0] = x.a2c[0];
            r.a2c[1] = x.a2c[1];
            r.v_slot = S.vslot[v];
            r.anc[0] = fslot(H.vanch[v].x);
            r.anc[1] = fslot(H.vanch[v].y);
            r.force = H.vforce[v] >= 0 ? add_force(H.vforce[v]) : -1;
            uint32_t full[8];
            pack(H.vnode_elems[v], full);
            P.vrow_full.insert(P.vrow_full.end(), full, full + 8);
            P.vrows.push_back(r);
        }
        // coefficient of an updated entry: 1/mfull, or -(side record + 1)
        auto coef_of = [&](int g) {
            const int si = H.node_side[g];
            if (si < 0) return H.invm[g];
            const HostSide& hs = H.side[si];
            SideT d{};
            d.invm[0] = hs.invm[0];
            d.invm[1] = hs.invm[1];
            d.c2[0] = hs.c2[0];
            d.c2[1] = hs.c2[1];
            uint32_t word = (uint32_t)hs.tr_count | ((uint32_t)hs.point << 3);
            if (H.node_force[g] >= 0) {
                const int fi = add_force(H.node_force[g]);
                word |= 32u | ((uint32_t)fi << 6);
            }
            d.word = word;
            for (int x = 0; x < hs.tr_count; ++x) {
                // the hanging node's row, and which of its two anchors g is
                const TransferRec& tr = H.tr[hs.tr_first + x];
                const int2 a = H.vanch[tr.v];
                d.tr[x] = 2 * (uint32_t)S.vridx[tr.v] + (a.x == g ? 0u : 1u);
            }
            P.sides.push_back(d);
            return -(double)(P.sides.size() - T.side_off);   // -(tile-local index + 1)
        };
        for (int i = 0; i < cnt; ++i) {
            pack(H.node_elems[n0 + i], &core_full[8 * (size_t)(n0 + i)]);
            core_coef[n0 + i] = coef_of(n0 + i);
        }
        T.n_ent = N;
        T.ring_off = (int)P.ring_gid.size();
        for (int q = cnt; q < N; ++q) {
            const int g = ent[q];
            P.ring_gid.push_back(g);
            uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            double c = 0.0;
            if (q < upd_end) {
                pack(H.node_elems[g], w);
                c = coef_of(g);
            }
            P.ring_full.insert(P.ring_full.end(), w, w + 8);
            P.ring_coef.push_back(c);
        }
        T.n_force = (int)P.force_slots.size() - T.force_off;
        T.n_side = (int)P.sides.size() - T.side_off;
        T.n_mat = (int)P.tile_mats.size() - T.mat_off;
        T.mat_mode = T.n_mat > 4 ? 1 : 0;
        // receivers
        T.recv_off = (int)P.recvs.size();
        T.n_recv = (int)tile_recv[t].size();
        for (int r : tile_recv[t]) {
            const RecvRec& x = H.recv[r];
            RecvT y{};
            y.r = r;
            y.ff = -1;
            if (x.c[0] >= 0) {
                for (int k = 0; k < 4; ++k) {
                    const int c = x.c[k];
                    if (c < NA) {
                        y.s[k] = make_int2(fslot(c), fslot(c));
                    } else {
                        y.hang |= 1 << k;
                        y.s[k] = make_int2(fslot(H.vanch[c - NA].x), fslot(H.vanch[c - NA].y));
                    }
                    y.w[k] = x.w[k];
                }
                y.ff = x.ffi;
            }
            P.recvs.push_back(y);
        }
        T.virt_off = (int)P.virt.size();
        T.n_virt = (int)vl.size();
        for (int v : vl) {
            P.virt.push_back(make_int2(fslot(H.vanch[v].x), fslot(H.vanch[v].y)));
            P.virt_ids.push_back(v);
        }
        P.tiles.push_back(T);
        P.max_slots = std::max(P.max_slots, nslots);
        P.max_vrows = std::max(P.max_vrows, T.n_vrow);
        P.max_virt = std::max(P.max_virt, T.n_virt);
        P.max_force = std::max(P.max_force, T.n_force);
        P.max_side = std::max(P.max_side, T.n_side);
        P.max_recv = std::max(P.max_recv, T.n_recv);
        P.max_mat = std::max(P.max_mat, T.n_mat);
        P.max_ent = std::max(P.max_ent, upd_end);
    };

    // chunks of tiles on worker threads (each with its own marks), joined in order
    const int nthreads = std::min(setup_threads(), std::max(1, ntiles / 64));
    const int nchunk = std::max(1, std::min(ntiles, 4 * nthreads));
    std::vector<TilePart> parts(nchunk);
    {
        std::vector<TileScratch> scratch(nthreads);
        std::atomic<int> next{0};
        std::vector<std::exception_ptr> errs(nchunk);
        auto work = [&](int w) {
            TileScratch& S = scratch[w];
            for (int c = next++; c < nchunk; c = next++) {
                try {
                    S.init((size_t)NA, (size_t)NV, H.forces.size(), (size_t)NM);
                    const int t0 = (int)((long)ntiles * c / nchunk), t1 = (int)((long)ntiles * (c + 1) / nchunk);
                    for (int t = t0; t < t1; ++t) build_one(t, S, parts[c]);
                } catch (...) {
                    errs[c] = std::current_exception();
                }
            }
        };
        std::vector<std::thread> pool;
        for (int w = 1; w < nthreads; ++w) pool.emplace_back(work, w);
        work(0);
        for (auto& th : pool) th.join();
        for (auto& e : errs)
            if (e) std::rethrow_exception(e);
    }
    std::vector<TileInfo> tiles;
    std::vector<int> ring_gid;
    std::vector<uint32_t> ring_full;
    std::vector<double> ring_coef;
    std::vector<SideT> sides;
    std::vector<VRowT> vrows;
    std::vector<int2> virt;
    std::vector<int> tile_mats;
    std::vector<int> force_slots;
    std::vector<RecvT> recvs;
    std::vector<int> tile_virt_ids;
    std::vector<uint32_t> vrow_full;
    int max_slots = 2, max_vrows = 0, max_virt = 0, max_force = 0;
    int max_side = 0, max_recv = 0, max_mat = 0, max_ent = 0;
    {
        size_t nr = 0, ns = 0, nvr = 0, nvi = 0, nm = 0, nf = 0, nrc = 0;
        for (const TilePart& P : parts) {
            nr += P.ring_gid.size();
            ns += P.sides.size();
            nvr += P.vrows.size();
            nvi += P.virt.size();
            nm += P.tile_mats.size();
            nf += P.force_slots.size();
            nrc += P.recvs.size();
        }
        if (nr > (size_t)INT32_MAX / 8) throw fail(DMP_EUNSUPPORTED, "ring lists too long");
        tiles.reserve(ntiles);
        ring_gid.reserve(nr);
        ring_full.reserve(8 * nr);
        ring_coef.reserve(nr);
        sides.reserve(ns);
        vrows.reserve(nvr);
        vrow_full.reserve(8 * nvr);
        virt.reserve(nvi);
        tile_virt_ids.reserve(nvi);
        tile_mats.reserve(nm);
        force_slots.reserve(nf);
        recvs.reserve(nrc);
    }
    for (TilePart& P : parts) {
        const int o_ring = (int)ring_gid.size(), o_side = (int)sides.size(), o_vrow = (int)vrows.size(),
                  o_virt = (int)virt.size(), o_mat = (int)tile_mats.size(), o_force = (int)force_slots.size(),
                  o_recv = (int)recvs.size();
        for (TileInfo T : P.tiles) {
            T.ring_off += o_ring;
            T.side_off += o_side;
            T.vrow_off += o_vrow;
            T.virt_off += o_virt;
            T.mat_off += o_mat;
            T.force_off += o_force;
            T.recv_off += o_recv;
            tiles.push_back(T);
        }
        ring_gid.insert(ring_gid.end(), P.ring_gid.begin(), P.ring_gid.end());
        ring_full.insert(ring_full.end(), P.ring_full.begin(), P.ring_full.end());
        ring_coef.insert(ring_coef.end(), P.ring_coef.begin(), P.ring_coef.end());
        sides.insert(sides.end(), P.sides.begin(), P.sides.end());
        vrows.insert(vrows.end(), P.vrows.begin(), P.vrows.end());
        vrow_full.insert(vrow_full.end(), P.vrow_full.begin(), P.vrow_full.end());
        virt.insert(virt.end(), P.virt.begin(), P.virt.end());
        tile_virt_ids.insert(tile_virt_ids.end(), P.virt_ids.begin(), P.virt_ids.end());
        tile_mats.insert(tile_mats.end(), P.tile_mats.begin(), P.tile_mats.end());
        force_slots.insert(force_slots.end(), P.force_slots.begin(), P.force_slots.end());
        recvs.insert(recvs.end(), P.recvs.begin(), P.recvs.end());
        max_slots = std::max(max_slots, P.max_slots);
        max_vrows = std::max(max_vrows, P.max_vrows);
        max_virt = std::max(max_virt, P.max_virt);
        max_force = std::max(max_force, P.max_force);
        max_side = std::max(max_side, P.max_side);
        max_recv = std::max(max_recv, P.max_recv);
        max_mat = std::max(max_mat, P.max_mat);
        max_ent = std::max(max_ent, P.max_ent);
        TilePart().tiles.swap(P.tiles);   // free the part as it is joined
        P = TilePart();
    }
    if ((int)tiles.size() != ntiles) throw fail(DMP_EINVAL, "internal: %zu tiles built of %d", tiles.size(), ntiles);

    // descriptors, per tile: a present element's word holds its three slots
    // and, in a tile with at most 4 materials, the table index in bits 30-31
    auto to_word = [](const uint32_t* full, int k, int mode) {
        if (!(full[2 * k] & (1u << 30))) return 0xFFFFFFFFu;
        return (full[2 * k] & 0x3FFFFFFFu) | (mode ? 0u : full[2 * k + 1] << 30);
    };
    // hanging-node rows: the three slots of quadrant k in (x-neighbour,
    // z-neighbour, diagonal) order instead of ascending corner order (corner
    // k's neighbour across x shares its z, across z its x; the step kernel
    // evaluates these rows with the quadrant as a runtime value)
    auto canonical_word = [](uint32_t word, int k) {
        if (word == 0xFFFFFFFFu) return word;
        static constexpr int kPos[4][3] = {{0, 2, 1}, {0, 1, 2}, {2, 1, 0}, {2, 0, 1}};
        uint32_t out = word & 0xC0000000u;
        for (int t = 0; t < 3; ++t) out |= ((word >> (10 * kPos[k][t])) & 1023u) << (10 * t);
        return out;
    };
    auto emit = [&](uint4* out, const uint32_t* full, int mode) {
        *out = make_uint4(to_word(full, 0, mode), to_word(full, 1, mode), to_word(full, 2, mode),
                          to_word(full, 3, mode));
    };
    std::vector<uint4> core_desc(NA), ring_desc(ring_gid.size());
    bool has_elem = false;
    for (const TileInfo& T : tiles) has_elem = has_elem || T.mat_mode;
    {
        const int nblk = std::max(1, std::min(ntiles, 16 * setup_threads()));
        parallel_items(nblk, setup_threads(), [&](int b) {
            const int t0 = (int)((long)ntiles * b / nblk), t1 = (int)((long)ntiles * (b + 1) / nblk);
            for (int t = t0; t < t1; ++t) {
                const TileInfo& T = tiles[t];
                for (int e = 0; e < T.cnt; ++e)
                    emit(&core_desc[T.n0 + e], &core_full[8 * (size_t)(T.n0 + e)], T.mat_mode);
                for (int e = T.cnt; e < T.n_ent; ++e) {
                    const size_t r = (size_t)T.ring_off + e - T.cnt;
                    emit(&ring_desc[r], &ring_full[8 * r], T.mat_mode);
                }
                for (int q = 0; q < T.n_vrow; ++q) {
                    VRowT& r = vrows[T.vrow_off + q];
                    uint4 w;
                    emit(&w, &vrow_full[8 * (size_t)(T.vrow_off + q)], T.mat_mode);
                    r.desc[0] = canonical_word(w.x, 0);
                    r.desc[1] = canonical_word(w.y, 1);
                    r.desc[2] = canonical_word(w.z, 2);
                    r.desc[3] = canonical_word(w.w, 3);
                }
            }
        });
    }
    // (lambda, mu) per entry of the tiles' hanging-node lists, for ELEM tiles
    std::vector<double> virt_em64;
    if (has_elem) {
        virt_em64.assign(2 * virt.size(), 0.0);
        for (const TileInfo& T : tiles) {
            if (!T.mat_mode) continue;
            for (int q = 0; q < T.n_virt; ++q) {
                // the hanging node at slot n_ent + q is the one listed q-th
                const int v = tile_virt_ids[T.virt_off + q];
                virt_em64[2 * (size_t)(T.virt_off + q)] = H.vnode_em[2 * (size_t)v];
                virt_em64[2 * (size_t)(T.virt_off + q) + 1] = H.vnode_em[2 * (size_t)v + 1];
            }
        }
    }
    V.has_elem = has_elem ? 1 : 0;
    V.block = B;
    V.threads = std::max(64, (max_ent + 63) / 64 * 64);
    V.depth = K;
    V.node_tiles = ntiles;
    V.lds_slots = (max_slots + 1) & ~1;
    V.lds_vrows = (max_vrows + 1) & ~1;
    V.lds_virt = max_virt;
    V.lds_force = max_force;
    V.lds_side = K == 1 ? 0 : max_side;   // one step per launch: records stay in registers
    V.lds_recv = max_recv;
    V.lds_mat = 4;   // every tile loads its 4-entry table
    V.threads = std::max(V.threads, (2 * max_vrows + 63) / 64 * 64);   // two lanes per hanging-node row
    if (V.threads > kMaxThreads)
        throw fail(DMP_EUNSUPPORTED, "tiles need %d threads (max %d)", V.threads, kMaxThreads);
    // every index the kernel follows stays inside its array (checked before
    // anything is launched)
    {
        auto bad = [&](int t, const char* what) {
            return fail(DMP_EINVAL, "internal: tile %d: %s out of range", t, what);
        };
        auto check_tile = [&](int t) {
            const TileInfo& T = tiles[t];
            const int cnt = T.cnt;
            if (T.n0 != tstart[t] || cnt < 1 || T.n0 + cnt != tstart[t + 1]) throw bad(t, "core range");
            // the one-step kernel derives the core range from the tile index
            if (K == 1 && (T.n0 != t * B || cnt != std::min(B, NA - t * B) || cnt > V.threads))
                throw bad(t, "one-step core range");
            const int nslots = T.n_ent + T.n_virt;
            if (T.n_ent < cnt || T.lev_end[K - 1] > V.threads || nslots > V.lds_slots || nslots > kMaxSlots)
                throw bad(t, "entries");
            for (int j = 0; j < kMaxDepth; ++j)
                if (T.lev_end[j] > T.lev_end[j + 1] || T.lev_end[j] < cnt) throw bad(t, "ring levels");
            if (T.lev_end[kMaxDepth] != T.n_ent) throw bad(t, "ring levels");
            if (T.prv_end < T.lev_end[K - 1] || T.prv_end > T.n_ent) throw bad(t, "previous-step range");
            if (T.ring_off < 0 || T.ring_off + (T.n_ent - cnt) > (int)ring_gid.size()) throw bad(t, "ring list");
            if (2 * T.n_vrow > V.threads || T.n_vrow > V.lds_vrows || T.n_virt > V.lds_virt ||
                T.n_force > V.lds_force || (K > 1 && T.n_side > V.lds_side) || T.n_recv > V.lds_recv ||
                (T.mat_mode == 0 && T.n_mat > V.lds_mat))
                throw bad(t, "LDS list");
            if (T.virt_off + T.n_virt > (int)virt.size() || T.vrow_off + T.n_vrow > (int)vrows.size() ||
                T.force_off + T.n_force > (int)force_slots.size() || T.side_off + T.n_side > (int)sides.size() ||
                T.recv_off + T.n_recv > (int)recvs.size() || T.mat_off + T.n_mat > (int)tile_mats.size())
                throw bad(t, "list range");
            auto desc_ok = [&](const uint32_t* w) {
                for (int k = 0; k < 4; ++k) {
                    if (!(w[2 * k] & (1u << 30))) continue;
                    for (int q = 0; q < 3; ++q)
                        if ((int)((w[2 * k] >> (10 * q)) & 1023) >= nslots) return false;
                    if ((int)w[2 * k + 1] >= T.n_mat) return false;
                }
                return true;
            };
            const int upd_end = T.lev_end[K - 1];
            for (int e = 0; e < upd_end; ++e) {
                const uint32_t* w = e < cnt ? &core_full[8 * (size_t)(T.n0 + e)]
                                            : &ring_full[8 * (size_t)(T.ring_off + e - cnt)];
                if (!desc_ok(w)) throw bad(t, "descriptor");
                const double c = e < cnt ? core_coef[T.n0 + e] : ring_coef[T.ring_off + e - cnt];
                if (c < 0 && (int)(-c) - 1 >= T.n_side) throw bad(t, "side index");
            }
            for (int e = cnt; e < T.n_ent; ++e)
                if (ring_gid[T.ring_off + e - cnt] < 0 || ring_gid[T.ring_off + e - cnt] >= NA)
                    throw bad(t, "ring node");
            for (int q = 0; q < T.n_virt; ++q) {
                const int2 a = virt[T.virt_off + q];
                if (a.x < 0 || a.x >= T.n_ent || a.y < 0 || a.y >= T.n_ent) throw bad(t, "anchor slot");
            }
            auto word_ok = [&](const uint32_t* w) {
                for (int k = 0; k < 4; ++k) {
                    if (w[k] == 0xFFFFFFFFu) continue;
                    for (int q = 0; q < 3; ++q)
                        if ((int)((w[k] >> (10 * q)) & 1023) >= nslots) return false;
                    if (T.mat_mode ? (w[k] >> 30) != 0 : (int)(w[k] >> 30) >= T.n_mat) return false;
                }
                return true;
            };
            if (T.mat_mode != (T.n_mat > 4 ? 1 : 0)) throw bad(t, "material mode");
            for (int e = 0; e < upd_end; ++e) {
                const uint4 d = e < cnt ? core_desc[T.n0 + e] : ring_desc[(size_t)T.ring_off + e - cnt];
                const uint32_t w[4] = {d.x, d.y, d.z, d.w};
                if (!word_ok(w)) throw bad(t, "descriptor word");
            }
            for (int q = 0; q < T.n_vrow; ++q) {
                const VRowT& r = vrows[T.vrow_off + q];
                if (r.v_slot < T.n_ent || r.v_slot >= nslots || r.force >= T.n_force || !word_ok(r.desc) ||
                    !desc_ok(&vrow_full[8 * (size_t)(T.vrow_off + q)]))
                    throw bad(t, "hanging-node row");
                // the row computes both anchors' transfer terms from u_{i-2}
                const int2 a = virt[T.virt_off + r.v_slot - T.n_ent];
                if (a.x != r.anc[0] || a.y != r.anc[1] || a.x >= T.prv_end || a.y >= T.prv_end)
                    throw bad(t, "hanging-node anchors");
            }
            for (int q = 0; q < T.n_side; ++q) {
                const SideT& d = sides[T.side_off + q];
                const int ntr = d.word & 7;
                if (ntr > kMaxTransfers) throw bad(t, "transfer count");
                if ((d.word & 32) && (int)((d.word >> 6) & 1023) >= T.n_force) throw bad(t, "force index");
                for (int x = 0; x < ntr; ++x) {
                    if ((int)d.tr[x] >= 2 * T.n_vrow) throw bad(t, "transfer");
                }
            }
            for (int q = 0; q < T.n_force; ++q)
                if (force_slots[T.force_off + q] < 0 || force_slots[T.force_off + q] >= (int)H.forces.size())
                    throw bad(t, "force row");
            for (int q = 0; q < T.n_recv; ++q) {
                const RecvT& y = recvs[T.recv_off + q];
                if (y.r < 0 || y.r >= NR || y.ff >= NFF) throw bad(t, "receiver");
                for (int c = 0; c < 4; ++c)
                    if (y.s[c].x < 0 || y.s[c].x >= T.n_ent || y.s[c].y < 0 || y.s[c].y >= T.n_ent)
                        throw bad(t, "receiver slot");
            }
            for (int q = 0; q < T.n_mat; ++q)
                if (tile_mats[T.mat_off + q] < 0 || tile_mats[T.mat_off + q] >= NM) throw bad(t, "material");
        };
        const int nblk = std::max(1, std::min(ntiles, 16 * setup_threads()));
        parallel_items(nblk, setup_threads(), [&](int b) {
            const int t0 = (int)((long)ntiles * b / nblk), t1 = (int)((long)ntiles * (b + 1) / nblk);
            for (int t = t0; t < t1; ++t) check_tile(t);
        });
    }
    const size_t lds = step_lds_bytes(V, R.precision);
    if (std::getenv("DMP_LAYOUT_STATS")) {
        double ent = 0, ring = 0;
        std::vector<int> hist(17, 0);
        for (const TileInfo& T : tiles) {
            ent += T.n_ent;
            hist[std::min(16, T.n_ent / 64)]++;
        }
        ring = (double)ring_gid.size();
        int elem_tiles = 0;
        for (const TileInfo& T : tiles) elem_tiles += T.mat_mode;
        std::fprintf(stderr, "layout B=%d K=%d tiles=%d threads=%d mean_ent=%.1f ring/core=%.3f lds=%zu "
                     "slots=%d vrows=%d virt=%d force=%d side=%d recv=%d elem_tiles=%d hist=",
                     B, K, ntiles, V.threads, ent / std::max(ntiles, 1), ring / std::max(NA, 1), lds,
                     V.lds_slots, V.lds_vrows, V.lds_virt, V.lds_force, V.lds_side,
                     V.lds_recv, elem_tiles);
        for (int h : hist) std::fprintf(stderr, "%d,", h);
        int with_vrow = 0;
        for (const TileInfo& T : tiles) with_vrow += T.n_vrow > 0;
        std::fprintf(stderr, " tiles with hanging-node rows %d\n", with_vrow);
    }
    if (lds > R.lds_limit)
        throw fail(DMP_EUNSUPPORTED, "tiles of %d nodes for %d steps per launch need %zu B of LDS "
                   "(limit %zu)", B, K, lds, R.lds_limit);
    V.tiles = R.upload_tile(tiles);
    V.core_desc = R.upload_tile(core_desc);
    V.ring_gid = R.upload_tile(ring_gid);
    V.elist = nullptr;
    V.elist_cap = 0;
    if (K > 1) {
        // entry lists at a fixed stride per tile (padding costs traffic, so
        // the bandwidth-bound one-step layouts keep the ring list instead)
        int cap = 1;
        for (const TileInfo& T : tiles) cap = std::max(cap, T.n_ent);
        if (NA >= kPrevFlag) throw fail(DMP_EUNSUPPORTED, "more than %d active nodes", kPrevFlag - 1);
        std::vector<int> elist((size_t)ntiles * cap, -1);
        for (int t = 0; t < ntiles; ++t) {
            const TileInfo& T = tiles[t];
            for (int e = 0; e < T.n_ent; ++e) {
                const int g = e < T.cnt ? T.n0 + e : ring_gid[T.ring_off + e - T.cnt];
                elist[(size_t)t * cap + e] = g | (e < T.prv_end ? kPrevFlag : 0);
            }
        }
        V.elist = R.upload_tile(elist);
        V.elist_cap = cap;
    }
    V.ring_desc = R.upload_tile(ring_desc);
    if (R.precision == DMP_FP64) {
        V.core_coef = R.upload_tile(core_coef);
        V.ring_coef = R.upload_tile(ring_coef);
    } else {
        std::vector<float> c32(core_coef.begin(), core_coef.end()), r32(ring_coef.begin(), ring_coef.end());
        V.core_coef = R.upload_tile(c32);
        V.ring_coef = R.upload_tile(r32);
    }
    V.side = R.upload_tile(sides);
    V.vrows = R.upload_tile(vrows);
    V.virt = R.upload_tile(virt);
    {
        // (lambda, mu) of each tile's 4-entry table (the kernel indexes
        // 4 * tile); ELEM tiles leave theirs zero
        std::vector<double> t64(8 * (size_t)ntiles, 0.0);
        for (int t = 0; t < ntiles; ++t) {
            const TileInfo& T = tiles[t];
            if (T.mat_mode) continue;
            for (int q = 0; q < T.n_mat; ++q) {
                const int m = tile_mats[T.mat_off + q];
                t64[8 * (size_t)t + 2 * q] = H.lm[2 * m];
                t64[8 * (size_t)t + 2 * q + 1] = H.lm[2 * m + 1];
            }
        }
        if (R.precision == DMP_FP64) {
            V.tile_lm = R.upload_tile(t64);
            V.virt_em = R.upload_tile(virt_em64);
        } else {
            std::vector<float> t32(t64.begin(), t64.end());
            V.tile_lm = R.upload_tile(t32);
            std::vector<float> v32(virt_em64.begin(), virt_em64.end());
            V.virt_em = R.upload_tile(v32);
        }
    }
    V.force_slots = R.upload_tile(force_slots);
    V.recvs = R.upload_tile(recvs);
    R.tiles_block = B;
    R.tiles_depth = K;
}

void prepare_graph(Runtime& R);

void Runtime::apply_tuning()
{
    if (tiles_block != tuning[0] || tiles_depth != tuning[3]) build_tiles(*this);
    apply_launch_shape();
    // the receiver ring is part of the captured view: a layout built before
    // the ring exists gets its graph when the ring is allocated
    if (V.nrecv <= 0 || V.recv_interval <= 0 || V.recv_cap > 0) prepare_graph(*this);
}

// Switches to tuning t; a shape the layout cannot use leaves the previous one
// in place and reports why.
void set_tuning(Runtime& R, const int* t)
{
    int old[4];
    for (int k = 0; k < 4; ++k) {
        old[k] = R.tuning[k];
        R.tuning[k] = t[k];
    }
    try {
        R.apply_tuning();
    } catch (const Error&) {
        for (int k = 0; k < 4; ++k) R.tuning[k] = old[k];
        R.apply_tuning();
        throw;
    }
}

// --------------------------------------------------------------------------
// Pack local layouts into a device layout.
// --------------------------------------------------------------------------
void build_runtime(Runtime& R, const std::vector<Instance*>& insts)
{
    const int ninst = (int)insts.size();
    R.precision = insts[0]->precision;
    R.device = insts[0]->device;
    HIPCHECK(hipSetDevice(R.device));
    HIPCHECK(hipStreamCreateWithFlags(&R.stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreate(&R.ev0));
    HIPCHECK(hipEventCreate(&R.ev1));
    HIPCHECK(scratch_address(R.device, &R.scratch));
    PhaseTimer pt_pack("pack");
    int NA = 0, NV = 0, NE = 0, NM = 0, NS = 0, NT = 0, NDE = 0, NR = 0, NSE = 0, NF = 0;
    for (Instance* I : insts) {
        const LocalLayout& L = *I->L;
        R.inst_node_base.push_back(NA);
        R.inst_vnode_base.push_back(NV);
        R.inst_elem_base.push_back(NE);
        R.inst_recv_base.push_back(NR);
        R.inst_seis_base.push_back(NSE);
        R.inst_na.push_back(L.na);
        R.inst_nrecv.push_back((int)L.recv.size());
        R.inst_nseis.push_back((int)L.seis.size());
        NA += L.na;
        NV += L.nv;
        NE += L.ne;
        NM += (int)L.lm.size() / 2;
        NS += (int)L.side.size();
        NT += (int)L.tr.size();
        NDE += (int)L.drm_el.size();
        NR += (int)L.recv.size();
        NSE += (int)L.seis.size();
        NF += (int)L.ff.size();
    }
    auto enc = [&](int inst, int id) {
        return id >= 0 ? R.inst_node_base[inst] + id : NA + R.inst_vnode_base[inst] + (-id - 1);
    };
    std::vector<int4> node_elems(NA), vnode_elems(NV), elem_corners(NE), ehang(NE);
    std::vector<int2> vanch(NV), seis(NSE);
    std::vector<int> elem_mat(NE), act_orig(NA), virt_orig(NV), elem_orig(NE);
    std::vector<double> lm64(2 * (size_t)NM);
    std::vector<float> lm32(2 * (size_t)NM);
    std::vector<double> invm(NA, 0.0);
    std::vector<int> node_side(NA, -1);
    std::vector<HostSide> side(NS);
    std::vector<TransferRec> tr(NT);
    std::vector<DrmElem> drm_el(NDE);
    std::vector<FFNode> ffn(NF);
    std::vector<RecvRec> recv(NR);
    std::vector<InstParams> ip(ninst);
    int mb = 0, sb = 0, tb = 0, deb = 0, fb = 0;
    for (int q = 0; q < ninst; ++q) {
        const LocalLayout& L = *insts[q]->L;
        const int nb = R.inst_node_base[q], vb = R.inst_vnode_base[q], eb = R.inst_elem_base[q];
        auto e4 = [&](const std::array<int, 4>& a) {
            return make_int4(a[0] >= 0 ? eb + a[0] : -1, a[1] >= 0 ? eb + a[1] : -1,
                             a[2] >= 0 ? eb + a[2] : -1, a[3] >= 0 ? eb + a[3] : -1);
        };
        for (int a = 0; a < L.na; ++a) {
            node_elems[nb + a] = e4(L.node_elems[a]);
            act_orig[nb + a] = L.act_orig[a];
            invm[nb + a] = L.invm[a];
            node_side[nb + a] = L.side_of[a] < 0 ? -1 : sb + L.side_of[a];
        }
        for (int v = 0; v < L.nv; ++v) {
            vnode_elems[vb + v] = e4(L.vnode_elems[v]);
            vanch[vb + v] = make_int2(nb + L.vanch[v][0], nb + L.vanch[v][1]);
            virt_orig[vb + v] = L.virt_orig[v];
        }
        for (int e = 0; e < L.ne; ++e) {
            const auto& c = L.ecorner[e];
            elem_corners[eb + e] = make_int4(enc(q, c[0]), enc(q, c[1]), enc(q, c[2]), enc(q, c[3]));
            const auto& hh = L.ehang[e];
            ehang[eb + e] = make_int4(hh[0] >= 0 ? vb + hh[0] : -1, hh[1] >= 0 ? vb + hh[1] : -1,
                                      hh[2] >= 0 ? vb + hh[2] : -1, hh[3] >= 0 ? vb + hh[3] : -1);
            elem_mat[eb + e] = mb + L.emat[e];
            elem_orig[eb + e] = L.elem_orig[e];
        }
        for (size_t s = 0; s < L.seis.size(); ++s)
            seis[R.inst_seis_base[q] + s] = make_int2(eb + L.seis[s][0], L.seis[s][1]);
        for (size_t m = 0; m < L.lm.size() / 2; ++m) {
            lm64[2 * (mb + m)] = L.lm[2 * m];
            lm64[2 * (mb + m) + 1] = L.lm[2 * m + 1];
            lm32[2 * (mb + m)] = (float)L.lm[2 * m];
            lm32[2 * (mb + m) + 1] = (float)L.lm[2 * m + 1];
        }
        // plane-wave ranges: local PairRange index -> global (first, count)
        auto pair_range = [&](int d, int& first, int& count) {
            if (d < 0) {
                first = 0;
                count = 0;
            } else {
                first = deb + L.drm[d].first;
                count = L.drm[d].count;
            }
        };
        for (size_t s = 0; s < L.side.size(); ++s) {
            HostSide x = L.side[s];
            x.tr_first += tb;
            pair_range(x.drm, x.drm, x.drm_count);
            x.inst = q;
            side[sb + s] = x;
        }
        for (size_t t = 0; t < L.tr.size(); ++t) {
            TransferRec x = L.tr[t];
            x.v += vb;
            pair_range(x.drm, x.drm, x.drm_count);
            tr[tb + t] = x;
        }
        for (size_t d = 0; d < L.drm_el.size(); ++d) {
            DrmElem x = L.drm_el[d];
            for (int j = 0; j < 4; ++j) x.c[j] += fb;
            drm_el[deb + d] = x;
        }
        for (size_t f = 0; f < L.ff.size(); ++f) {
            FFNode x = L.ff[f];
            x.inst = q;
            ffn[fb + f] = x;
        }
        for (size_t r = 0; r < L.recv.size(); ++r) {
            RecvRec x = L.recv[r];
            if (x.c[0] != INT32_MIN) {
                for (int k = 0; k < 4; ++k) x.c[k] = enc(q, x.c[k]);
            } else {
                x.c[0] = -1;
            }
            if (x.ffi >= 0) x.ffi += fb;
            x.inst = q;
            recv[R.inst_recv_base[q] + r] = x;
        }
        ip[q] = L.ip;
        mb += (int)L.lm.size() / 2;
        sb += (int)L.side.size();
        tb += (int)L.tr.size();
        deb += (int)L.drm_el.size();
        fb += (int)L.ff.size();
    }
    int interval = insts[0]->io_interval;
    for (Instance* I : insts)
        if (I->io_interval != interval)
            throw fail(DMP_EINVAL, "batched instances must share io_interval");

    DevView& V = R.V;
    V = DevView{};
    V.na = NA;
    V.nv = NV;
    V.ne = NE;
    V.nrecv = NR;
    V.recv_interval = interval;
    V.mat_lm = R.precision == DMP_FP64 ? (const void*)R.upload(lm64) : (const void*)R.upload(lm32);
    V.inst = R.upload(ip);
    PackedHost& H = R.host;
    H.node_elems = std::move(node_elems);
    H.vnode_elems = std::move(vnode_elems);
    H.elem_corners = elem_corners;
    H.elem_mat = std::move(elem_mat);
    H.vanch = vanch;
    H.node_side = std::move(node_side);
    H.invm = std::move(invm);
    H.side = std::move(side);
    H.tr = std::move(tr);
    H.pairs = std::move(drm_el);
    H.ff = std::move(ffn);
    H.recv = std::move(recv);
    H.lm = lm64;
    H.inst = ip;
    H.vtr.assign(NV, -1);
    for (size_t x = 0; x < H.tr.size(); ++x)
        if (H.vtr[H.tr[x].v] < 0) H.vtr[H.tr[x].v] = (int)x;
    // each node is corner 0 of at most one element (build_local checks every
    // corner position): that element's (lambda, mu) is kept at the node
    H.node_em.assign(2 * (size_t)NA, 0.0);
    H.vnode_em.assign(2 * (size_t)NV, 0.0);
    for (int e = 0; e < NE; ++e) {
        const int c0 = H.elem_corners[e].x, m = H.elem_mat[e];
        double* dst = c0 < NA ? &H.node_em[2 * (size_t)c0] : &H.vnode_em[2 * (size_t)(c0 - NA)];
        dst[0] = H.lm[2 * (size_t)m];
        dst[1] = H.lm[2 * (size_t)m + 1];
    }
    if (R.precision == DMP_FP64) {
        V.node_em = R.upload(H.node_em);
    } else {
        std::vector<float> f(H.node_em.begin(), H.node_em.end());
        V.node_em = R.upload(f);
    }
    pt_pack.stop();
    // dependencies: the active nodes an update at node n reads -- corners of
    // its elements and, for each hanging node it anchors, the hanging node's
    // anchors and the corners of its elements (hanging corners as anchors);
    // node ranges on host threads, joined in node order
    {
        PhaseTimer pt_dep("dependencies");
        H.dep_ptr.assign((size_t)NA + 1, 0);
        H.dep.clear();
        const int nblk = std::max(1, std::min(256, NA / 4096 + 1));
        std::vector<std::vector<int>> part(nblk);
        parallel_items(nblk, setup_threads(), [&](int b) {
            const int n0 = (int)((long)NA * b / nblk), n1 = (int)((long)NA * (b + 1) / nblk);
            std::vector<int> tmp;
            std::vector<int>& out = part[b];
            auto add = [&](int c) {
                if (c < NA) {
                    tmp.push_back(c);
                } else {
                    tmp.push_back(H.vanch[c - NA].x);
                    tmp.push_back(H.vanch[c - NA].y);
                }
            };
            auto add_elem = [&](int e) {
                if (e < 0) return;
                const int4 c = H.elem_corners[e];
                add(c.x);
                add(c.y);
                add(c.z);
                add(c.w);
            };
            for (int n = n0; n < n1; ++n) {
                tmp.clear();
                const int4 q = H.node_elems[n];
                add_elem(q.x);
                add_elem(q.y);
                add_elem(q.z);
                add_elem(q.w);
                const int si = H.node_side[n];
                if (si >= 0) {
                    const HostSide& hs = H.side[si];
                    for (int x = hs.tr_first; x < hs.tr_first + hs.tr_count; ++x) {
                        const int v = H.tr[x].v;
                        tmp.push_back(H.vanch[v].x);
                        tmp.push_back(H.vanch[v].y);
                        const int4 w = H.vnode_elems[v];
                        add_elem(w.x);
                        add_elem(w.y);
                        add_elem(w.z);
                        add_elem(w.w);
                    }
                }
                std::sort(tmp.begin(), tmp.end());
                tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
                int cnt = 0;
                for (int d : tmp)
                    if (d != n) {
                        out.push_back(d);
                        ++cnt;
                    }
                H.dep_ptr[n + 1] = cnt;
            }
        });
        size_t total = 0;
        for (const auto& v : part) total += v.size();
        if (total > (size_t)INT32_MAX) throw fail(DMP_EUNSUPPORTED, "dependency lists too long");
        for (int n = 0; n < NA; ++n) H.dep_ptr[n + 1] += H.dep_ptr[n];
        H.dep.reserve(total);
        for (auto& v : part) {
            H.dep.insert(H.dep.end(), v.begin(), v.end());
            std::vector<int>().swap(v);
        }
    }
    // forced nodes (plane-wave pairs and/or point source) and the source tables
    {
        PhaseTimer pt_src("sources");
        H.node_force.assign(NA, -1);
        H.vforce.assign(NV, -1);
        H.forces.clear();
        for (int n = 0; n < NA; ++n) {
            const int si = H.node_side[n];
            if (si < 0) continue;
            const HostSide& hs = H.side[si];
            if (hs.drm_count == 0 && !hs.point) continue;
            H.node_force[n] = (int)H.forces.size();
            H.forces.push_back(ForceT{hs.drm_count ? hs.drm : 0, hs.drm_count, hs.point, hs.inst});
        }
        for (int v = 0; v < NV; ++v) {
            if (H.vtr[v] < 0) continue;
            const TransferRec& x = H.tr[H.vtr[v]];
            if (x.drm_count == 0) continue;
            H.vforce[v] = (int)H.forces.size();
            H.forces.push_back(ForceT{x.drm, x.drm_count, 0, 0});
        }
        std::vector<FFT> fft(H.ff.size());
        for (size_t f = 0; f < H.ff.size(); ++f) {
            const FFNode& x = H.ff[f];
            const InstParams& P = ip[x.inst];
            FFT& y = fft[f];
            for (int w = 0; w < 3; ++w) {
                y.v[w] = x.v[w];
                y.A[w] = P.A[w];
                y.UX[w] = P.UX[w];
                y.UZ[w] = P.UZ[w];
            }
            y.dt = P.dt;
            y.t0 = P.t0;
            y.freq = P.freq;
            y.recv_time = x.recv_time;
        }
        V.nff = (int)fft.size();
        V.nforce = (int)H.forces.size();
        // window of steps held: a multiple of 60 (whole captured graphs) plus
        // one, at most about 1 GiB
        const size_t per_step = (size_t)V.nff * sizeof(double4) + (size_t)V.nforce * sizeof(double2);
        int chunk = 960;
        while (chunk > 60 && per_step * (size_t)(chunk + 1) > ((size_t)1 << 30)) chunk -= 60;
        V.chunk = chunk + 1;
        V.ffp = R.upload(fft);
        V.forces = R.upload(H.forces);
        V.pairs = R.upload(H.pairs);
        V.fftab = R.dalloc<double4>((size_t)std::max(V.nff, 1) * V.chunk);
        V.ftab = R.dalloc<double2>((size_t)std::max(V.nforce, 1) * V.chunk);
        R.table_lo = -1;
    }
    PhaseTimer pt_state("state+uploads");
    for (int k = 0; k < kStateSlots; ++k) {
        V.U[k] = R.dalloc<double2>(NA);   // 16 B per node in both precisions
        HIPCHECK(hipMemset(V.U[k], 0, (size_t)std::max(NA, 1) * 16));
    }
    // problem_displacements(0, e_u), problem_displacements(-dt, e_up)
    // (explicit.f90:93-94): prescribed z of POINTD nodes in u_0 and u_-1
    for (int q = 0; q < ninst; ++q) {
        const Instance& I = *insts[q];
        if (I.point_kind != 2) continue;
        const LocalLayout& L = *I.L;
        const double den = L.ip.p_den;
        auto rk = [&](double t) {
            t = t - I.p_start;
            return std::exp(-(t * t) / (4 * I.p_sigma)) * ((t * t) - 2 * I.p_sigma) / den * I.p_amp;
        };
        const double z0 = rk(0.0), zm = rk(-I.dt);
        for (int nd1 : I.point_nodes) {
            const int g = R.inst_node_base[q] + L.node_int[nd1 - 1];
            for (int k : {0, kStateSlots - 1}) {
                const double z = k == 0 ? z0 : zm;
                char* base = static_cast<char*>(V.U[k]) + (size_t)g * 16;
                if (R.precision == DMP_FP64) {
                    HIPCHECK(hipMemcpy(base + 8, &z, 8, hipMemcpyHostToDevice));
                } else {
                    const float hi = (float)z, lo = (float)(z - (double)hi);
                    HIPCHECK(hipMemcpy(base + 4, &hi, 4, hipMemcpyHostToDevice));
                    HIPCHECK(hipMemcpy(base + 12, &lo, 4, hipMemcpyHostToDevice));
                }
            }
        }
    }
    V.ctrl = R.dalloc<Ctrl>(1);
    HIPCHECK(hipMemset(V.ctrl, 0, sizeof(Ctrl)));
    R.dev_base = 0;
    {
        // receivers in global node ids (sampled from HBM at the end of a run call)
        std::vector<RecvG> rg(H.recv.size());
        R.recv_missing.clear();
        for (size_t r = 0; r < H.recv.size(); ++r) {
            const RecvRec& x = H.recv[r];
            RecvG& g = rg[r];
            g = RecvG{};
            g.ff = -1;
            if (x.c[0] < 0) {
                R.recv_missing.push_back((int)r);
                continue;
            }
            g.found = 1;
            g.ff = x.ffi;
            for (int k = 0; k < 4; ++k) {
                const int c = x.c[k];
                if (c < 0 || c >= NA + NV) throw fail(DMP_EINVAL, "internal: receiver %zu corner out of range", r);
                if (c < NA) {
                    g.a[k] = g.b[k] = c;
                } else {
                    g.a[k] = H.vanch[c - NA].x;
                    g.b[k] = H.vanch[c - NA].y;
                    g.hang |= 1 << k;
                }
                g.w[k] = x.w[k];
            }
            if (g.ff >= V.nff) throw fail(DMP_EINVAL, "internal: receiver %zu free-field point out of range", r);
        }
        V.rg = R.upload(rg);
    }
    R.h_act_orig = act_orig;
    R.h_virt_orig = virt_orig;
    R.d_act_orig = R.upload(act_orig);
    R.d_virt_orig = R.upload(virt_orig);
    R.d_elem_orig = R.upload(elem_orig);
    R.d_ehang = R.upload(ehang);
    R.d_elem_corners = R.upload(elem_corners);
    R.d_vanch = R.upload(vanch);
    R.d_seis = R.upload(seis);
    V.recv_cap = 0;
    V.recv_buf = nullptr;
    R.step = 0;
    R.first_run = true;
    pt_state.stop();
    int lds_max = 0;
    HIPCHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, R.device));
    // dynamic LDS per workgroup: what the device allows (160 KiB on gfx950);
    // a layout takes only what its largest tile needs
    R.lds_limit = (size_t)lds_max;
    HIPCHECK(configure_kernels(&R.lds_limit));
    HIPCHECK(hipDeviceSynchronize());
    // default shape: several steps per launch where launches dominate (small
    // meshes), one where HBM does (large ones); fewer steps or smaller tiles
    // where a tile of the default shape does not fit
    const bool small = NA < 200000;
    const int shapes[][2] = {{256, small ? 3 : 1}, {256, 2}, {128, 3}, {128, 2}, {256, 1}, {128, 1}, {64, 1}};
    bool ok = false;
    for (const auto& sh : shapes) {
        R.tuning[0] = sh[0];
        R.tuning[3] = sh[1];
        try {
            R.apply_tuning();
            ok = true;
            break;
        } catch (const Error& e) {
            if (e.code != DMP_EUNSUPPORTED) throw;
            if (std::getenv("DMP_TIMING")) std::fprintf(stderr, "dmp setup shape %d x %d: %s\n", sh[0], sh[1], e.msg.c_str());
        }
    }
    if (!ok) throw fail(DMP_EUNSUPPORTED, "no tile shape fits this mesh");
    for (int k = 0; k < 4; ++k) R.init_tuning[k] = R.tuning[k];
}

// g steps per graph as g/k launches of k steps, then the base advance.
void capture_graph(Runtime& R, int g, int k)
{
    R.drop_graph();
    hipStream_t cs;
    HIPCHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    HIPCHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    for (int m = 0; m < g / k; ++m) HIPCHECK(launch_step(R.V, R.precision, m * k, 1, k, cs));
    HIPCHECK(launch_advance(R.V.ctrl, g, cs));
    HIPCHECK(hipStreamEndCapture(cs, &R.graph));
    HIPCHECK(hipGraphInstantiate(&R.gexec, R.graph, nullptr, nullptr, 0));
    HIPCHECK(hipStreamDestroy(cs));
    R.graph_steps = g;
    R.graph_k = k;
}

// Steps per graph launch: whole launches, inside one source-table window,
// with the samples of one graph within a quarter of the receiver ring; 0 for
// eager launches.
int graph_length(const Runtime& R)
{
    const DevView& V = R.V;
    const int K = R.tuning[3];
    int G = std::min(R.tuning[2], V.chunk - 1);
    if (V.nrecv > 0 && V.recv_interval > 0)
        G = (int)std::min<long>(G, (V.recv_cap / 4 - 1) * (long)V.recv_interval);
    return std::max(0, G / K) * K;
}

// Instantiates the graph of the current launch shape when the shape is set,
// so that no run call pays for it.
void prepare_graph(Runtime& R)
{
    const int G = graph_length(R);
    if (G > 0 && (R.graph_steps != G || R.graph_k != R.tuning[3] || !R.gexec)) capture_graph(R, G, R.tuning[3]);
}

// The source tables cover steps [a, b] (b - a < chunk); otherwise they are
// evaluated for [a, a + chunk), in stream order after the launches before.
void ensure_tables(Runtime& R, int a, int b)
{
    if (R.V.nff == 0 && R.V.nforce == 0) return;
    if (R.table_lo >= 0 && a >= R.table_lo && b < R.table_lo + R.V.chunk) return;
    HIPCHECK(launch_tables(R.V, a, R.V.chunk, R.stream));
    R.table_lo = a;
}

// Receiver sample ring: sample q (step q * io_interval) lives in slot
// q mod recv_cap of pinned, device-mapped host memory, allocated once at
// set-up: nsteps / io_interval + 2 slots, at most 256 MiB (at least 64
// slots); run calls longer than the ring stream their samples out through it.
void alloc_ring(Runtime& R, int nsteps)
{
    const int I = R.V.recv_interval;
    if (R.V.nrecv <= 0 || I <= 0) return;
    constexpr size_t kRingBytes = size_t(256) << 20;
    const size_t row = (size_t)2 * R.V.nrecv;
    const size_t want = (size_t)std::max(nsteps, 0) / I + 2;
    const size_t slots = std::max<size_t>(64, std::min(want, kRingBytes / (row * sizeof(double))));
    if (slots > (size_t)INT32_MAX) throw fail(DMP_EINVAL, "receiver ring too large");
    void* p = nullptr;
    const hipError_t e = hipHostMalloc(&p, slots * row * sizeof(double), hipHostMallocMapped);
    if (e != hipSuccess)
        throw fail(DMP_ENOMEM, "hipHostMalloc(%zu bytes) failed: %s", slots * row * sizeof(double),
                   hipGetErrorString(e));
    std::memset(p, 0, slots * row * sizeof(double));
    R.hstage = static_cast<double*>(p);
    R.hstage_cap = slots * row;
    void* dp = nullptr;
    HIPCHECK(hipHostGetDevicePointer(&dp, R.hstage, 0));
    R.V.recv_buf = static_cast<double*>(dp);
    R.V.recv_cap = (int)slots;
    R.drop_graph();   // kernels capture the view by value
    prepare_graph(R);
}

// Waits for an event by polling it: the end of a run call is on the
// caller's critical path, and a blocking wait adds its wake-up latency to
// every short call.
void wait_event(hipEvent_t e)
{
    hipError_t st;
    while ((st = hipEventQuery(e)) == hipErrorNotReady) {
    }
    HIPCHECK(st);
}

hipEvent_t take_event(Runtime& R)
{
    if (!R.free_ev.empty()) {
        hipEvent_t e = R.free_ev.back();
        R.free_ev.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
}

// Advances n steps; returns the number of receiver samples copied to out.
//
// A launch from step t0 samples the receivers at t0 .. t0 + k - 1, so the
// last step of a call is sampled by a receivers-only kernel reading that
// state level from HBM (and sampled again, with the same bits, by the first
// launch of the next call).  The call returns the samples of steps j in
// (s, s + n] with j % io_interval == 0, and step s itself on the first call.
int run_steps(Runtime& R, int n, double* out)
{
    if (n < 0) throw fail(DMP_EINVAL, "negative step count");
    DevView& V = R.V;
    const int s = R.step;
    const int I = V.recv_interval;
    const bool rec = V.nrecv > 0 && I > 0;
    if (rec && V.recv_cap <= 0) throw fail(DMP_ESTATE, "receiver ring not allocated");
    const long q_lo = rec ? (R.first_run ? (long)s / I : (long)s / I + 1) : 0;
    const long q_hi = rec ? (long)(s + n) / I : -1;   // inclusive
    const long nsamp = rec ? std::max(0L, q_hi - q_lo + 1) : 0;
    const bool stream_out = out && nsamp > 0;
    const size_t row = (size_t)2 * V.nrecv;
    const long cap = V.recv_cap;
    const int K = R.tuning[3];
    const int G = graph_length(R);

    // samples [q_lo, copied) are in out; a mark's event completes once the
    // samples below its `upto` are written
    struct Mark {
        hipEvent_t ev;
        long upto;
    };
    std::vector<Mark> marks;
    size_t mhead = 0;
    long copied = q_lo, written = q_lo;
    constexpr size_t kFlushBytes = size_t(2) << 20;
    const long flush_samples =
        rec ? std::max<long>(1, std::min<long>((long)(kFlushBytes / (row * sizeof(double))), cap / 4)) : 1;
    auto copy_out = [&](long upto) {
        upto = std::min(upto, q_hi + 1);
        while (copied < upto) {
            const long slot = copied % cap;
            const long cnt = std::min(upto - copied, cap - slot);
            std::memcpy(out + (size_t)(copied - q_lo) * row, R.hstage + (size_t)slot * row,
                        (size_t)cnt * row * sizeof(double));
            copied += cnt;
        }
    };
    auto mark = [&]() {
        Mark m{take_event(R), written};
        HIPCHECK(hipEventRecord(m.ev, R.stream));
        marks.push_back(m);
    };
    auto retire = [&]() {   // oldest mark: wait, copy its samples out
        Mark m = marks[mhead++];
        HIPCHECK(hipEventSynchronize(m.ev));
        R.free_ev.push_back(m.ev);
        copy_out(m.upto);
    };
    // before enqueueing work that writes samples up to qmax: the ring slot
    // of every sample it writes must hold one already copied out
    auto before = [&](long qmax) {
        if (!stream_out) return;
        while (qmax - copied >= cap) {
            if (mhead == marks.size() || marks.back().upto < written) mark();
            retire();
        }
    };
    auto after = [&](long upto) {   // the enqueued work writes samples < upto
        if (!stream_out) return;
        written = std::max(written, std::min(upto, q_hi + 1));
        const long last = mhead < marks.size() ? marks.back().upto : copied;
        if (written - last >= flush_samples) mark();
    };
    auto last_sample = [&](int t0, int k) { return rec ? (long)(t0 + k - 1) / I : 0L; };

    HIPCHECK(hipEventRecord(R.ev0, R.stream));
    int done = 0;
    if (G > 0 && n >= G) {
        if (R.graph_steps != G || R.graph_k != K || !R.gexec) capture_graph(R, G, K);
        if (R.dev_base != s) HIPCHECK(launch_set_base(V.ctrl, s, R.stream));
        R.dev_base = INT32_MIN;
        while (n - done >= G) {
            ensure_tables(R, s + done, s + done + G);
            before(last_sample(s + done, G));
            HIPCHECK(hipGraphLaunch(R.gexec, R.stream));
            done += G;
            after(last_sample(s + done, 0) + 1);
        }
        R.dev_base = s + done;
    }
    // the call's last step is a sample: a last launch of fewer steps than
    // the layout's depth samples it itself, otherwise a receivers-only kernel
    const bool end_sample = rec && (s + n) % I == 0 && (n > 0 || R.first_run);
    bool end_done = false;
    while (done < n) {
        const int k = std::min(K, n - done);
        const int end = end_sample && done + k == n && k < V.depth ? 1 : 0;
        ensure_tables(R, s + done, s + done + k + end);
        before(last_sample(s + done, k + end));
        HIPCHECK(launch_step(V, R.precision, s + done, 0, k, R.stream, end));
        done += k;
        after(last_sample(s + done, end) + 1);
        end_done = end_done || end;
    }
    if (end_sample && !end_done) {
        ensure_tables(R, s + n, s + n);
        before(q_hi);
        HIPCHECK(launch_recv(V, R.precision, s + n, R.stream));
        after(q_hi + 1);
    }
    HIPCHECK(hipEventRecord(R.ev1, R.stream));
    while (mhead < marks.size()) retire();   // copies overlap the launches still running
    wait_event(R.ev1);
    if (stream_out) copy_out(q_hi + 1);
    HIPCHECK(hipEventElapsedTime(&R.last_ms, R.ev0, R.ev1));
    R.step = s + n;
    R.first_run = false;
    return (int)nsamp;
}

// run_steps without any allocation: eager launches only (no graph is
// instantiated, no event created), the receiver samples written to this
// device's scratch array instead of the pinned ring and copied to out at the
// end.  The samples of one call (and the one its first launch writes again)
// must fit the scratch.  Calls on one device share the scratch: calls that
// overlap in time race.  Nothing in it needs to start at zero: every sample
// copied out is written by the call.
int run_steps_noalloc(Runtime& R, int n, double* out)
{
    if (n < 0) throw fail(DMP_EINVAL, "negative step count");
    DevView V = R.V;   // the view with the scratch in place of the ring
    const int s = R.step;
    const int I = V.recv_interval;
    const bool rec = V.nrecv > 0 && I > 0;
    const long q_lo = rec ? (R.first_run ? (long)s / I : (long)s / I + 1) : 0;
    const long q_hi = rec ? (long)(s + n) / I : -1;
    const long nsamp = rec ? std::max(0L, q_hi - q_lo + 1) : 0;
    const size_t row = (size_t)2 * std::max(V.nrecv, 0);
    if (rec) {
        const long slots = (long)(scratch_bytes() / (row * sizeof(double)));
        // samples q_lo - 1 .. q_hi are written during the call
        if (nsamp + 1 > slots)
            throw fail(DMP_EINVAL, "dmp_run_noalloc: %ld receiver samples of %d receivers exceed the "
                       "%zu-byte scratch (at most %ld samples per call)", nsamp, V.nrecv,
                       scratch_bytes(), slots - 1);
        V.recv_buf = R.scratch;
        V.recv_cap = (int)slots;
    }
    const int K = R.tuning[3];
    const bool end_sample = rec && (s + n) % I == 0 && (n > 0 || R.first_run);
    bool end_done = false;
    HIPCHECK(hipEventRecord(R.ev0, R.stream));
    for (int done = 0; done < n;) {
        const int k = std::min(K, n - done);
        const int end = end_sample && done + k == n && k < V.depth ? 1 : 0;
        ensure_tables(R, s + done, s + done + k + end);
        HIPCHECK(launch_step(V, R.precision, s + done, 0, k, R.stream, end));
        done += k;
        end_done = end_done || end;
    }
    if (end_sample && !end_done) {
        ensure_tables(R, s + n, s + n);
        HIPCHECK(launch_recv(V, R.precision, s + n, R.stream));
    }
    if (out && nsamp > 0) {
        const long cap = V.recv_cap;
        for (long q = q_lo; q <= q_hi;) {
            const long slot = q % cap;
            const long cnt = std::min(q_hi + 1 - q, cap - slot);
            HIPCHECK(hipMemcpyAsync(out + (size_t)(q - q_lo) * row, V.recv_buf + (size_t)slot * row,
                                    (size_t)cnt * row * sizeof(double), hipMemcpyDeviceToHost, R.stream));
            q += cnt;
        }
    }
    HIPCHECK(hipEventRecord(R.ev1, R.stream));
    HIPCHECK(hipEventSynchronize(R.ev1));
    HIPCHECK(hipEventElapsedTime(&R.last_ms, R.ev0, R.ev1));
    // no kernel writes the samples of receivers in no element; the scratch
    // may hold another call's samples there
    if (out)
        for (long q = 0; q < nsamp; ++q)
            for (const int r : R.recv_missing) {
                out[(size_t)q * row + 2 * (size_t)r] = 0.0;
                out[(size_t)q * row + 2 * (size_t)r + 1] = 0.0;
            }
    R.step = s + n;
    R.first_run = false;
    return (int)nsamp;
}

// Active-node state of instance q at step j, in the caller's node numbering
// (hanging nodes projected), as doubles: one device-to-host copy of the
// instance's nodes.
void fetch_state(Runtime& R, int q, int nnodes, int j, double* u)
{
    const int na = R.inst_na[q], nb = R.inst_node_base[q];
    const void* U = R.V.U[((j % kStateSlots) + kStateSlots) % kStateSlots];
    std::vector<double2> all(na);
    if (R.precision == DMP_FP64) {
        HIPCHECK(hipMemcpy(all.data(), static_cast<const double2*>(U) + nb, sizeof(double2) * na,
                           hipMemcpyDeviceToHost));
    } else {
        std::vector<float4> f(na);
        HIPCHECK(hipMemcpy(f.data(), static_cast<const float4*>(U) + nb, sizeof(float4) * na,
                           hipMemcpyDeviceToHost));
        for (int a = 0; a < na; ++a)
            all[a] = make_double2((double)f[a].x + (double)f[a].z, (double)f[a].y + (double)f[a].w);
    }
    std::fill(u, u + 2 * (size_t)nnodes, 0.0);
    for (int a = 0; a < na; ++a) {
        const int n = R.h_act_orig[nb + a];
        u[2 * n] = all[a].x;
        u[2 * n + 1] = all[a].y;
    }
    const int vb = R.inst_vnode_base[q];
    const int nvq = (q + 1 < (int)R.inst_vnode_base.size() ? R.inst_vnode_base[q + 1] : R.V.nv) - vb;
    for (int v = 0; v < nvq; ++v) {
        const int2 an = R.host.vanch[vb + v];
        const int n = R.h_virt_orig[vb + v];
        const double2 x = all[an.x - nb], y = all[an.y - nb];
        u[2 * n] = (x.x + y.x) / 2.0;
        u[2 * n + 1] = (x.y + y.y) / 2.0;
    }
}

template <class H> H* as(void* p) { return static_cast<H*>(p); }

int finish(std::string* err, const Error& e)
{
    if (err) *err = e.msg;
    g_lib_error = e.msg;
    return e.code;
}

template <class F> int guarded(std::string* err, F&& f)
{
    try {
        f();
        return DMP_OK;
    } catch (const Error& e) {
        return finish(err, e);
    } catch (const std::bad_alloc&) {
        return finish(err, Error{DMP_ENOMEM, "host allocation failed"});
    } catch (const std::exception& ex) {
        return finish(err, Error{DMP_EINVAL, ex.what()});
    }
}

void check_tuning(const int* t)
{
    if (t[0] < kMinBlock || t[0] > kMaxBlock)
        throw fail(DMP_EINVAL, "nodes per tile must be in [%d, %d] (got %d)", kMinBlock, kMaxBlock, t[0]);
    if (t[1] != 0 && t[1] != 1) throw fail(DMP_EINVAL, "xcd remap must be 0 or 1");
    if (t[2] < 0 || t[2] > 4096) throw fail(DMP_EINVAL, "graph steps must be in [0, 4096]");
    if (t[3] < 1 || t[3] > kMaxDepth)
        throw fail(DMP_EINVAL, "steps per launch must be in [1, %d] (got %d)", kMaxDepth, t[3]);
}

// Runs every candidate shape for nsteps steps from a saved state, once to
// warm up and twice timed; keeps the shape with the shortest timed run.  The
// state (the six state buffers, step) is restored.
void autotune(Runtime& R, int nsteps, int* out)
{
    if (nsteps <= 0) throw fail(DMP_EINVAL, "autotune needs a positive step count");
    struct Scratch {
        std::vector<void*> p;
        ~Scratch() { for (void* q : p) (void)hipFree(q); }
        void* get(size_t n)
        {
            void* q = nullptr;
            if (hipMalloc(&q, n) != hipSuccess) throw fail(DMP_ENOMEM, "autotune scratch allocation failed");
            p.push_back(q);
            return q;
        }
    } scratch;
    const size_t ub = (size_t)std::max(R.V.na, 1) * 16;
    void* save[kStateSlots];
    for (int k = 0; k < kStateSlots; ++k) {
        save[k] = scratch.get(ub);
        HIPCHECK(hipMemcpyAsync(save[k], R.V.U[k], ub, hipMemcpyDeviceToDevice, R.stream));
    }
    const int step0 = R.step;
    const bool fr0 = R.first_run;
    auto restore = [&]() {
        for (int k = 0; k < kStateSlots; ++k)
            HIPCHECK(hipMemcpyAsync(R.V.U[k], save[k], ub, hipMemcpyDeviceToDevice, R.stream));
        HIPCHECK(hipStreamSynchronize(R.stream));
        R.step = step0;
    };
    int best[4] = {R.tuning[0], R.tuning[1], R.tuning[2], R.tuning[3]};
    double best_t = 1e300;
    // tile sizes: 64, 128, 256, and those giving one or two tiles per compute
    // unit (small meshes are latency-bound: one tile per CU, no CU twice)
    std::vector<int> cand_block = {64, 128, 256};
    {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, R.device) != hipSuccess)
            ncu = 256;
        for (int m = 1; m <= 2; ++m) {
            const int b = (R.V.na + m * ncu - 1) / (m * ncu);
            if (b >= kMinBlock && b <= kMaxBlock &&
                std::find(cand_block.begin(), cand_block.end(), b) == cand_block.end())
                cand_block.push_back(b);
        }
    }
    // many tiles per compute unit (bandwidth-bound): deep launches only add
    // ring work, so one and two steps per launch are tried
    std::vector<int> cand_depth = {1, 2, 3, 4, 6, 8};
    {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, R.device) != hipSuccess || ncu < 1)
            ncu = 256;
        if (R.V.na > 2 * 256 * ncu) cand_depth = {1, 2};
    }
    for (int b : cand_block)
        for (int K : cand_depth) {
            const int cand_graph[2] = {0, std::max(K, (std::min(nsteps, 60) / K) * K)};
            for (int x = 0; x <= 1; ++x)
                for (int g : cand_graph) {
                    R.tuning[0] = b;
                    R.tuning[1] = x;
                    R.tuning[2] = g;
                    R.tuning[3] = K;
                    try {
                        R.apply_tuning();
                    } catch (const Error& e) {
                        if (e.code != DMP_EUNSUPPORTED) throw;
                        continue;   // shape the layout cannot use
                    }
                    // a warm-up, then timed runs from the saved state until
                    // there are two and 2 ms of them (at most 15: short
                    // calls vary by a few percent from run to run, more than
                    // what separates the best shapes); a shape counts the
                    // lower median of its runs
                    std::vector<double> ts;
                    double timed = 0.0;
                    for (int pass = 0; pass <= kTuneMaxRuns; ++pass) {
                        restore();
                        R.first_run = false;
                        auto t0 = std::chrono::steady_clock::now();
                        run_steps(R, nsteps, nullptr);
                        auto t1 = std::chrono::steady_clock::now();
                        const double dt = std::chrono::duration<double>(t1 - t0).count();
                        if (pass == 0) continue;
                        ts.push_back(dt);
                        timed += dt;
                        if (ts.size() >= 2 && timed >= kTuneMinSeconds) break;
                    }
                    std::sort(ts.begin(), ts.end());
                    const double t = ts[(ts.size() - 1) / 2];
                    if (t < best_t) {
                        best_t = t;
                        best[0] = b;
                        best[1] = x;
                        best[2] = g;
                        best[3] = K;
                    }
                }
        }
    for (int k = 0; k < 4; ++k) R.tuning[k] = best[k];
    R.apply_tuning();
    restore();
    R.first_run = fr0;
    if (out)
        for (int k = 0; k < 4; ++k) out[k] = best[k];
}

} // namespace

// ==========================================================================
// C-ABI
// ==========================================================================
extern "C" {

int dmp_create(void** h, const int* precision, const int* device)
{
    if (!h || !precision || !device) return finish(nullptr, Error{DMP_EINVAL, "null argument"});
    *h = nullptr;
    return guarded(nullptr, [&] {
        if (*precision != DMP_FP64 && *precision != DMP_FP32C)
            throw fail(DMP_EINVAL, "precision must be DMP_FP64 (0) or DMP_FP32C (1)");
        int ndev = 0;
        hipError_t e = hipGetDeviceCount(&ndev);
        if (e != hipSuccess || ndev <= 0)
            throw fail(DMP_EHIP, "no HIP device available (%s)", hipGetErrorString(e));
        if (*device < 0 || *device >= ndev) throw fail(DMP_EINVAL, "device %d out of range", *device);
        auto* I = new Instance();
        I->precision = *precision;
        I->device = *device;
        *h = I;
    });
}

int dmp_destroy(void** h)
{
    if (!h) return DMP_EINVAL;
    delete as<Instance>(*h);
    *h = nullptr;
    return DMP_OK;
}

int dmp_last_error(void* h, char* buf, const int* buflen)
{
    if (!buf || !buflen || *buflen < 0) return DMP_EINVAL;
    const std::string& s = h ? as<Instance>(h)->err : g_lib_error;
    const int n = std::min<int>(*buflen, (int)s.size());
    memcpy(buf, s.data(), n);
    for (int i = n; i < *buflen; ++i) buf[i] = ' ';
    return DMP_OK;
}

int dmp_set_mesh(void* h, const int* nelems, const int* eptr, const int* eind, const bool* ediv,
                 const int* nnodes, const double* nodes, const int* ematprop, const int* nummat,
                 const double* matconst, const double* elem_props)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (I->L) throw fail(DMP_ESTATE, "mesh already initialised");
        if (!nelems || !eptr || !eind || !ediv || !nnodes || !nodes || !nummat)
            throw fail(DMP_EINVAL, "null argument");
        if (*nelems < 0 || *nnodes < 0) throw fail(DMP_EINVAL, "negative sizes");
        if (!elem_props && (!ematprop || !matconst || *nummat <= 0))
            throw fail(DMP_EINVAL, "materials missing");
        const int ne = *nelems, nn = *nnodes;
        if (eptr[0] != 1) throw fail(DMP_EINVAL, "eptr(1) must be 1");
        for (int e = 0; e < ne; ++e)
            if (eptr[e + 1] < eptr[e]) throw fail(DMP_EINVAL, "eptr not monotone at %d", e + 1);
        const int len = eptr[ne] - 1;
        I->nelems = ne;
        I->nnodes = nn;
        I->eptr.assign(eptr, eptr + ne + 1);
        I->eind.assign(eind, eind + len);
        I->ediv.resize(len);
        for (int j = 0; j < len; ++j) I->ediv[j] = ediv[j] ? 1 : 0;
        I->nodes.assign(nodes, nodes + 2 * (size_t)nn);
        I->nummat = *nummat;
        if (elem_props) {
            I->eprops.assign(elem_props, elem_props + 5 * (size_t)ne);
            I->ematprop.assign(ne, 1);
        } else {
            I->eprops.clear();
            I->ematprop.assign(ematprop, ematprop + ne);
            I->matconst.assign(matconst, matconst + 5 * (size_t)*nummat);
        }
        I->have_mesh = true;
    });
}

int dmp_set_boundaries(void* h, const int* absorbing, const int* nhang, const int* hanging,
                       const int* nseis, const int* seismic)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (!I->have_mesh) throw fail(DMP_ESTATE, "dmp_set_mesh must come first");
        if (I->L) throw fail(DMP_ESTATE, "already initialised");
        if (!absorbing || !nhang || !nseis) throw fail(DMP_EINVAL, "null argument");
        if (*nhang < 0 || *nseis < 0) throw fail(DMP_EINVAL, "negative sizes");
        if ((*nhang > 0 && !hanging) || (*nseis > 0 && !seismic)) throw fail(DMP_EINVAL, "null array");
        I->absorbing.assign(absorbing, absorbing + I->nnodes);
        for (int v : I->absorbing)
            if (v < 0 || v > 2) throw fail(DMP_EINVAL, "absorbing flag %d not in {0,1,2}", v);
        I->hanging.assign(hanging, hanging + 3 * (size_t)*nhang);
        I->seismic.assign(seismic, seismic + 2 * (size_t)*nseis);
        I->have_bounds = true;
    });
}

int dmp_set_wave(void* h, const char* prob, const double* origin, const double* t0,
                 const double* angle, const double* amp, const double* freq, const double* vp,
                 const double* vs, const double* box)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (I->L) throw fail(DMP_ESTATE, "already initialised");
        if (!prob || !origin || !t0 || !angle || !amp || !freq || !vp || !vs || !box)
            throw fail(DMP_EINVAL, "null argument");
        std::string p(prob, 6);
        if (I->point_kind) throw fail(DMP_EINVAL, "a problem has either a plane wave or a point source");
        if (p == "WAVEP ") I->wave_kind = 1;
        else if (p == "WAVESV") I->wave_kind = 2;
        else throw fail(DMP_EUNSUPPORTED, "problem '%s' has no plane wave", p.c_str());
        if (!(*vp > 0) || !(*vs > 0)) throw fail(DMP_EINVAL, "wave velocities must be positive");
        I->origin[0] = origin[0];
        I->origin[1] = origin[1];
        I->t0 = *t0;
        I->angle = *angle;
        I->amp = *amp;
        I->freq = *freq;
        I->vp = *vp;
        I->vs = *vs;
        for (int k = 0; k < 3; ++k) I->box[k] = box[k];
        InstParams tmp{};
        if (init_plane_wave(I->wave_kind, I->amp, I->vp, I->vs, I->angle, I->freq, tmp) != DMP_OK)
            throw fail(DMP_EWAVE, I->wave_kind == 1 ? "grazing P incidence not implemented"
                                                    : "post-critical SV incidence not implemented");
    });
}

int dmp_set_point_source(void* h, const char* prob, const double* amp, const double* sigma,
                         const double* start, const int* npoints, const int* pointnodes)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (I->L) throw fail(DMP_ESTATE, "already initialised");
        if (!prob || !amp || !sigma || !start || !npoints || (*npoints > 0 && !pointnodes))
            throw fail(DMP_EINVAL, "null argument");
        std::string p(prob, 6);
        if (p == "POINTN") I->point_kind = 1;
        else if (p == "POINTD") I->point_kind = 2;
        else throw fail(DMP_EUNSUPPORTED, "problem '%s' is not a point source", p.c_str());
        if (I->wave_kind) throw fail(DMP_EINVAL, "a problem has either a plane wave or a point source");
        if (!(*sigma > 0)) throw fail(DMP_EINVAL, "sigma must be positive");
        I->p_amp = *amp;
        I->p_sigma = *sigma;
        I->p_start = *start;
        I->point_nodes.assign(pointnodes, pointnodes + std::max(*npoints, 0));
    });
}

int dmp_set_time(void* h, const double* dt, const int* nsteps, const int* io_interval)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (I->L) throw fail(DMP_ESTATE, "already initialised");
        if (!dt || !nsteps || !io_interval) throw fail(DMP_EINVAL, "null argument");
        if (!(*dt > 0) || !std::isfinite(*dt)) throw fail(DMP_EINVAL, "dt must be positive");
        if (*io_interval < 0 || *nsteps < 0) throw fail(DMP_EINVAL, "negative step counts");
        I->dt = *dt;
        I->nsteps = *nsteps;
        I->io_interval = *io_interval;
        I->have_time = true;
    });
}

int dmp_set_receivers(void* h, const int* nrecv, const int* recv_elem, const double* recv_xy,
                      const bool* recv_add_freefield)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (I->L) throw fail(DMP_ESTATE, "already initialised");
        if (!nrecv || *nrecv < 0) throw fail(DMP_EINVAL, "bad receiver count");
        const int n = *nrecv;
        if (n > 0 && (!recv_elem || !recv_xy || !recv_add_freefield)) throw fail(DMP_EINVAL, "null array");
        I->recv_elem.assign(recv_elem, recv_elem + n);
        I->recv_xy.assign(recv_xy, recv_xy + 2 * (size_t)n);
        I->recv_ff.resize(n);
        for (int r = 0; r < n; ++r) I->recv_ff[r] = recv_add_freefield[r] ? 1 : 0;
    });
}

int dmp_init(void* h)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (!I->have_mesh || !I->have_time) throw fail(DMP_ESTATE, "mesh and time must be set before dmp_init");
        if (!I->have_bounds) {
            I->absorbing.assign(I->nnodes, 0);
            I->hanging.clear();
            I->seismic.clear();
            I->have_bounds = true;
        }
        if (I->rt) throw fail(DMP_ESTATE, "already initialised");
        const bool built_here = !I->L;   // a batch may have built the local layout already
        if (built_here) {
            PhaseTimer pt("local layout");
            build_local(*I);
        }
        auto rt = std::make_unique<Runtime>();
        try {
            PhaseTimer pt("device layout");
            build_runtime(*rt, {I});
        } catch (...) {
            if (built_here) I->L.reset();
            throw;
        }
        alloc_ring(*rt, I->nsteps);
        I->rt = std::move(rt);
    });
}

int dmp_run(void* h, const int* nsteps_block, double* recv_out, int* nsaved)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (!I->rt) throw fail(DMP_ESTATE, "dmp_init must come first");
        if (!nsteps_block) throw fail(DMP_EINVAL, "null argument");
        HIPCHECK(hipSetDevice(I->device));
        const int ns = run_steps(*I->rt, *nsteps_block, recv_out);
        if (nsaved) *nsaved = ns;
    });
}

int dmp_run_noalloc(void* h, const int* nsteps_block, double* recv_out, int* nsaved)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (!I->rt) throw fail(DMP_ESTATE, "dmp_init must come first");
        if (!nsteps_block) throw fail(DMP_EINVAL, "null argument");
        HIPCHECK(hipSetDevice(I->device));
        const int ns = run_steps_noalloc(*I->rt, *nsteps_block, recv_out);
        if (nsaved) *nsaved = ns;
    });
}

int dmp_get_state(void* h, double* u, double* up, double* mfull)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (!I->rt) throw fail(DMP_ESTATE, "dmp_init must come first");
        HIPCHECK(hipSetDevice(I->device));
        Runtime& R = *I->rt;
        if (u) fetch_state(R, 0, I->nnodes, R.step, u);
        if (up) fetch_state(R, 0, I->nnodes, R.step - 1, up);
        if (mfull) memcpy(mfull, I->mfull.data(), I->mfull.size() * sizeof(double));
    });
}

int dmp_last_run_ms(void* h, float* ms)
{
    if (!h || !ms) return DMP_EINVAL;
    Instance* I = as<Instance>(h);
    if (!I->rt) return DMP_ESTATE;
    *ms = I->rt->last_ms;
    return DMP_OK;
}

int dmp_batch_last_run_ms(void* b, float* ms)
{
    if (!b || !ms) return DMP_EINVAL;
    *ms = as<Batch>(b)->rt->last_ms;
    return DMP_OK;
}

int dmp_get_step(void* h, int* step)
{
    if (!h || !step) return DMP_EINVAL;
    Instance* I = as<Instance>(h);
    *step = I->rt ? I->rt->step : 0;
    return DMP_OK;
}

int dmp_export_mesh(void* h, int* eptr, int* eind, bool* ediv, int* hanging, int* seismic)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (!I->rt) throw fail(DMP_ESTATE, "dmp_init must come first");
        Runtime& R = *I->rt;
        HIPCHECK(hipSetDevice(I->device));
        const int NA = R.V.na, NV = R.V.nv, NE = R.V.ne;
        std::vector<int> act(NA), virt(NV), eorig(NE);
        std::vector<int4> corners(NE), eh(NE);
        std::vector<int2> vanch(NV), seis(I->seismic.size() / 2);
        HIPCHECK(hipMemcpy(act.data(), R.d_act_orig, sizeof(int) * NA, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(eorig.data(), R.d_elem_orig, sizeof(int) * NE, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(corners.data(), R.d_elem_corners, sizeof(int4) * NE, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(eh.data(), R.d_ehang, sizeof(int4) * NE, hipMemcpyDeviceToHost));
        if (NV) {
            HIPCHECK(hipMemcpy(virt.data(), R.d_virt_orig, sizeof(int) * NV, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(vanch.data(), R.d_vanch, sizeof(int2) * NV, hipMemcpyDeviceToHost));
        }
        if (!seis.empty())
            HIPCHECK(hipMemcpy(seis.data(), R.d_seis, sizeof(int2) * seis.size(), hipMemcpyDeviceToHost));
        auto orig_node = [&](int id) { return id < NA ? act[id] : virt[id - NA]; };
        // rows in the caller's element order
        std::vector<int> inv(NE);
        for (int i = 0; i < NE; ++i) inv[eorig[i]] = i;
        int pos = 0;
        if (eptr) eptr[0] = 1;
        for (int e = 0; e < NE; ++e) {
            const int i = inv[e];
            const int c[4] = {corners[i].x, corners[i].y, corners[i].z, corners[i].w};
            const int hh[4] = {eh[i].x, eh[i].y, eh[i].z, eh[i].w};
            for (int k = 0; k < 4; ++k) {
                if (eind) eind[pos] = orig_node(c[k]) + 1;
                if (ediv) ediv[pos] = false;
                ++pos;
                if (hh[k] >= 0) {
                    if (eind) eind[pos] = virt[hh[k]] + 1;
                    if (ediv) ediv[pos] = true;
                    ++pos;
                }
            }
            if (eptr) eptr[e + 1] = pos + 1;
        }
        if (hanging)
            for (int v = 0; v < NV; ++v) {
                hanging[3 * v] = virt[v] + 1;
                hanging[3 * v + 1] = act[vanch[v].x] + 1;
                hanging[3 * v + 2] = act[vanch[v].y] + 1;
            }
        if (seismic)
            for (size_t s = 0; s < seis.size(); ++s) {
                seismic[2 * s] = eorig[seis[s].x] + 1;
                seismic[2 * s + 1] = seis[s].y;
            }
    });
}

int dmp_set_tuning(void* h, const int* tuning)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (!I->rt) throw fail(DMP_ESTATE, "dmp_init must come first");
        const int* t = tuning ? tuning : I->rt->init_tuning;
        check_tuning(t);
        set_tuning(*I->rt, t);
    });
}

int dmp_get_tuning(void* h, int* tuning)
{
    if (!h || !tuning) return DMP_EINVAL;
    Instance* I = as<Instance>(h);
    if (!I->rt) return DMP_ESTATE;
    for (int k = 0; k < 4; ++k) tuning[k] = I->rt->tuning[k];
    return DMP_OK;
}

int dmp_autotune(void* h, const int* nsteps, int* tuning_out)
{
    if (!h) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Instance* I = as<Instance>(h);
    return guarded(&I->err, [&] {
        if (!I->rt) throw fail(DMP_ESTATE, "dmp_init must come first");
        if (!nsteps) throw fail(DMP_EINVAL, "null argument");
        HIPCHECK(hipSetDevice(I->device));
        autotune(*I->rt, *nsteps, tuning_out);
    });
}

int dmp_batch_create(void** b, const int* n, void* const* handles)
{
    if (!b || !n || !handles) return finish(nullptr, Error{DMP_EINVAL, "null argument"});
    *b = nullptr;
    return guarded(nullptr, [&] {
        if (*n <= 0) throw fail(DMP_EINVAL, "empty batch");
        auto B = std::make_unique<Batch>();
        std::vector<Instance*> todo;
        for (int i = 0; i < *n; ++i) {
            Instance* I = as<Instance>(handles[i]);
            if (!I) throw fail(DMP_EINVAL, "instance %d is a null handle", i);
            if (I->precision != as<Instance>(handles[0])->precision ||
                I->device != as<Instance>(handles[0])->device)
                throw fail(DMP_EINVAL, "batched instances must share precision and device");
            if (!I->L) {
                if (!I->have_mesh || !I->have_time)
                    throw fail(DMP_ESTATE, "instance %d: mesh and time must be set", i);
                if (!I->have_bounds) {
                    I->absorbing.assign(I->nnodes, 0);
                    I->hanging.clear();
                    I->seismic.clear();
                    I->have_bounds = true;
                }
                todo.push_back(I);
            }
            B->inst.push_back(I);
        }
        {
            // instances given set up but not initialised: their local layouts
            // (effective mass, Lysmer, plane-wave pairs, node order) on host
            // threads; no per-instance device layout is made
            PhaseTimer pt("local layouts");
            try {
                parallel_items((int)todo.size(), setup_threads(), [&](int i) { build_local(*todo[i]); });
            } catch (...) {
                for (Instance* I : todo) I->L.reset();
                throw;
            }
        }
        B->rt = std::make_unique<Runtime>();
        {
            PhaseTimer pt("batch layout");
            build_runtime(*B->rt, B->inst);
        }
        int nsteps = 0;
        for (const Instance* I : B->inst) nsteps = std::max(nsteps, I->nsteps);
        alloc_ring(*B->rt, nsteps);
        *b = B.release();
    });
}

int dmp_batch_destroy(void** b)
{
    if (!b) return DMP_EINVAL;
    delete as<Batch>(*b);
    *b = nullptr;
    return DMP_OK;
}

int dmp_batch_run(void* b, const int* nsteps, double* recv_out, int* nsaved)
{
    if (!b) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Batch* B = as<Batch>(b);
    return guarded(&B->err, [&] {
        if (!nsteps) throw fail(DMP_EINVAL, "null argument");
        HIPCHECK(hipSetDevice(B->rt->device));
        const int ns = run_steps(*B->rt, *nsteps, recv_out);
        if (nsaved) *nsaved = ns;
    });
}

int dmp_batch_get_state(void* b, const int* instance, double* u, double* up)
{
    if (!b) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Batch* B = as<Batch>(b);
    return guarded(&B->err, [&] {
        if (!instance || *instance < 0 || *instance >= (int)B->inst.size())
            throw fail(DMP_EINVAL, "instance out of range");
        HIPCHECK(hipSetDevice(B->rt->device));
        Runtime& R = *B->rt;
        const int q = *instance;
        if (u) fetch_state(R, q, B->inst[q]->nnodes, R.step, u);
        if (up) fetch_state(R, q, B->inst[q]->nnodes, R.step - 1, up);
    });
}

int dmp_batch_set_tuning(void* b, const int* tuning)
{
    if (!b) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Batch* B = as<Batch>(b);
    return guarded(&B->err, [&] {
        const int* t = tuning ? tuning : B->rt->init_tuning;
        check_tuning(t);
        set_tuning(*B->rt, t);
    });
}

int dmp_batch_autotune(void* b, const int* nsteps, int* tuning_out)
{
    if (!b) return finish(nullptr, Error{DMP_EINVAL, "null handle"});
    Batch* B = as<Batch>(b);
    return guarded(&B->err, [&] {
        if (!nsteps) throw fail(DMP_EINVAL, "null argument");
        HIPCHECK(hipSetDevice(B->rt->device));
        autotune(*B->rt, *nsteps, tuning_out);
    });
}

int dmp_write_step(const int* nfname, const char* fname, const char* mode, const int* step,
                   const int* ndofs, const double* u)
{
    if (!nfname || !fname || !mode || !step || !ndofs || (*ndofs > 0 && !u)) return DMP_EINVAL;
    std::string name(fname, (size_t)std::max(*nfname, 0));
    FILE* f = fopen(name.c_str(), mode[0] == 'n' ? "wb" : "ab");
    if (!f) return finish(nullptr, Error{DMP_EINVAL, "cannot open " + name});
    bool ok = fwrite(step, sizeof(int), 1, f) == 1 && fwrite(ndofs, sizeof(int), 1, f) == 1 &&
              fwrite(u, sizeof(double), (size_t)*ndofs, f) == (size_t)*ndofs;
    ok = (fclose(f) == 0) && ok;
    return ok ? DMP_OK : finish(nullptr, Error{DMP_EINVAL, "short write to " + name});
}

int dmp_write_mass(const int* nfname, const char* fname, const int* ndofs, const double* mass)
{
    if (!nfname || !fname || !ndofs || (*ndofs > 0 && !mass)) return DMP_EINVAL;
    std::string name(fname, (size_t)std::max(*nfname, 0));
    FILE* f = fopen(name.c_str(), "wb");
    if (!f) return finish(nullptr, Error{DMP_EINVAL, "cannot open " + name});
    bool ok = fwrite(ndofs, sizeof(int), 1, f) == 1 &&
              fwrite(mass, sizeof(double), (size_t)*ndofs, f) == (size_t)*ndofs;
    ok = (fclose(f) == 0) && ok;
    return ok ? DMP_OK : finish(nullptr, Error{DMP_EINVAL, "short write to " + name});
}

int dmp_write_mesh(const int* nfname, const char* fname, const int* nelems, const int* eptr,
                   const int* eind, const bool* ediv, const int* nnodes, const double* nodes)
{
    if (!nfname || !fname || !nelems || !eptr || !nnodes || *nelems < 0 || *nnodes < 0 ||
        (*nnodes > 0 && !nodes))
        return DMP_EINVAL;
    const int ne = *nelems;
    const int nent = eptr[ne] - 1;   // entries the mesh holds (1-based pointers)
    if (nent < 0 || (nent > 0 && (!eind || !ediv))) return DMP_EINVAL;
    std::string name(fname, (size_t)std::max(*nfname, 0));
    FILE* f = fopen(name.c_str(), "wb");
    if (!f) return finish(nullptr, Error{DMP_EINVAL, "cannot open " + name});
    // eptr[nelems] entries of eind and ediv, one more than the mesh holds
    // (io.cpp:354, 357); the extra entry is written as 0 / false instead of
    // being read past the caller's arrays
    const int zero = 0;
    const bool no = false;
    bool ok = fwrite(nelems, sizeof(int), 1, f) == 1 &&
              fwrite(eptr, sizeof(int), (size_t)ne + 1, f) == (size_t)ne + 1 &&
              fwrite(eind, sizeof(int), (size_t)nent, f) == (size_t)nent &&
              fwrite(&zero, sizeof(int), 1, f) == 1 &&
              fwrite(ediv, sizeof(bool), (size_t)nent, f) == (size_t)nent &&
              fwrite(&no, sizeof(bool), 1, f) == 1 && fwrite(nnodes, sizeof(int), 1, f) == 1 &&
              fwrite(nodes, sizeof(double), 2 * (size_t)*nnodes, f) == 2 * (size_t)*nnodes;
    ok = (fclose(f) == 0) && ok;
    return ok ? DMP_OK : finish(nullptr, Error{DMP_EINVAL, "short write to " + name});
}

int dmp_write_tparams(const int* nfname, const char* fname, const double* deltat,
                      const int* iointerval)
{
    if (!nfname || !fname || !deltat || !iointerval) return DMP_EINVAL;
    std::string name(fname, (size_t)std::max(*nfname, 0));
    FILE* f = fopen(name.c_str(), "wb");
    if (!f) return finish(nullptr, Error{DMP_EINVAL, "cannot open " + name});
    bool ok = fwrite(deltat, sizeof(double), 1, f) == 1 && fwrite(iointerval, sizeof(int), 1, f) == 1;
    ok = (fclose(f) == 0) && ok;
    return ok ? DMP_OK : finish(nullptr, Error{DMP_EINVAL, "short write to " + name});
}

int dmp_write_rank(const int* nfname, const char* fname, const int* myrank)
{
    if (!nfname || !fname || !myrank) return DMP_EINVAL;
    std::string name(fname, (size_t)std::max(*nfname, 0));
    // [int count][int visualize cursor][count x int rank]: a new file starts
    // with two zeros; each call appends one rank and bumps the count
    FILE* f = fopen(name.c_str(), "rb+");
    if (!f) {
        f = fopen(name.c_str(), "wb+");
        if (!f) return finish(nullptr, Error{DMP_EINVAL, "cannot open " + name});
        const int zero[2] = {0, 0};
        if (fwrite(zero, sizeof(int), 2, f) != 2) {
            fclose(f);
            return finish(nullptr, Error{DMP_EINVAL, "short write to " + name});
        }
    }
    int count = 0;
    bool ok = fseek(f, 0, SEEK_SET) == 0 && fread(&count, sizeof(int), 1, f) == 1 && count >= 0 &&
              fseek(f, (long)(sizeof(int) * (2 + (size_t)count)), SEEK_SET) == 0 &&
              fwrite(myrank, sizeof(int), 1, f) == 1;
    ++count;
    ok = ok && fseek(f, 0, SEEK_SET) == 0 && fwrite(&count, sizeof(int), 1, f) == 1;
    ok = (fclose(f) == 0) && ok;
    return ok ? DMP_OK : finish(nullptr, Error{DMP_EINVAL, "cannot update " + name});
}

} // extern "C"
