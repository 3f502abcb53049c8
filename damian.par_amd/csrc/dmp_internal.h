// dmp_internal.h -- device layout of the explicit step library (host + device).
//
// One "layout" holds one or more site-response instances packed end to end:
//   active nodes   [0, na)        : state-carrying nodes, space-filling-curve
//                                   ordered per instance;
//   virtual nodes  [na, na + nv)  : hanging nodes; never stored, their value is
//                                   the mean of two active anchors
//                                   (set_dispshanging, model.f90:612-624);
//   elements       [0, ne)        : 4 corners (active or virtual ids).
// The per-node update reproduces run_explicit (explicit.f90:125-153) as a
// gather over the node's (at most 4) elements, so no scatter-add and no
// atomics are needed; hanging-node force transfer (sum_reactanchors, P2) is a
// gather over the anchored virtual nodes.
//
// Time blocking: one launch advances every tile by k <= depth steps.  A tile
// owns a run of consecutive active nodes (its core) and holds in LDS the "rings"
// around them: ring j+1 = the nodes ring j's updates read.  Sub-step s of a
// launch updates the entries of rings 0 .. depth-s, so after k sub-steps the
// core is exact, while no other tile's state is needed inside the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace dmp {

constexpr int kMinBlock = 32;       // core nodes per tile
constexpr int kMaxBlock = 1024;
constexpr int kMaxThreads = 1024;   // threads per workgroup
constexpr int kMaxDepth = 8;       // steps per launch
// u_j lives in U[j mod kStateSlots].  A launch of k <= kMaxDepth steps from
// t0 reads U[t0] and U[t0-1] (any workgroup, at any time during the launch)
// and writes U[t0+k] and U[t0+k-1]; the four are distinct slots only when
// kStateSlots >= kMaxDepth + 2.
constexpr int kStateSlots = 10;
static_assert(kStateSlots >= kMaxDepth + 2, "a launch must not write a state level another workgroup reads");
constexpr int kMaxSlots = 1023;    // LDS slots addressable by a 10-bit descriptor field

// Per instance constants of the plane-wave and point sources.
struct InstParams {
    double dt;
    double t0;           // m_wttoorigin
    double ox, oz;       // m_worigin
    double freq;         // f_wfreq
    // three wave trains (incident, reflected P, reflected S):
    //   delay v_w = (x*S[w] + z*C[w]) / V[w], amplitude A[w], direction (UX, UZ)
    double S[3], C[3], V[3], A[3], UX[3], UZ[3];
    int kind;            // 0 none, 1 P, 2 SV
    int recv_interval;
    // point source (model.f90:809-843): 0 none, 1 POINTN force, 2 POINTD
    // prescribed z displacement; amp * ricker((t - start), sigma)
    int point_kind, pad;
    double p_amp, p_sigma, p_start, p_den;   // p_den = 8 sqrt(pi) sigma^2.5
};

// Point where the free field is evaluated (plane-wave layer node or receiver):
// delays of the three wave trains, v_w = (x S_w + z C_w) / V_w
// (pwaves.f90:153-155, 189-191), time independent.  Layer nodes use the time
// of wave_reactions, j*dt (model.f90:761); receivers the time of the visualize
// pass, (dt*k)*io_interval (visualization.f90:329).
struct FFNode {
    double v[3];
    int inst;
    int recv_time;
};

// A tile's copy of a free-field point with the constants of its instance, so
// that one load has everything the evaluation needs.
struct FFT {
    double v[3];
    double dt, t0, freq;
    double A[3], UX[3], UZ[3];
    int recv_time, pad;
};

// One seismic element seen from one of its corner nodes: that node's two rows
// of multdrm(K_e, u0) + m_e multdrm(M, a0) (model.f90:763-799).
struct DrmElem {
    int c[4];            // free-field point of each corner
    double lam, mu, mass;
    int k;               // corner of this element at the receiving node
    int ext;             // bit j set: corner j is exterior (model.f90:778-793)
};

// ---- tile layout ----------------------------------------------------------
// LDS slots of a tile: entries [0, n_ent) (core first, then rings in order),
// hanging nodes after them.  A node's element rows are described per corner
// position k (the node is corner k of that element) by one word holding the
// LDS slots of the other three corners (10 bits each, ascending corner
// order), 0xFFFFFFFF for an empty position: 16 B per node.  A tile with at
// most 4 materials (mat_mode 0) indexes its 4-entry (lambda, mu) table with
// bits 30-31; a tile with more (mat_mode 1) reads an element's (lambda, mu)
// from LDS at the slot of the element's corner 0, loaded per slot from
// node_em / virt_em.
constexpr int kPrevFlag = 1 << 30;   // entry-list word: u_{t0-1} is read as well

struct TileInfo {
    int n0, cnt;                 // core: active nodes n0 .. n0 + cnt - 1 (entries 0 .. cnt - 1)
    int n_ent;                   // entries; those >= cnt are ring entries
    int ring_off;                // ring entries in ring_gid / ring_desc / ring_coef
    int lev_end[kMaxDepth + 1];  // entries in rings <= j
    int virt_off, n_virt;        // hanging nodes: anchor slots in virt
    int vrow_off, n_vrow;        // hanging nodes whose reactions the tile's anchors take
    int force_off, n_force;      // forced nodes: rows of the force table in force_slots
    int recv_off, n_recv;        // receivers sampled by the tile
    int side_off, n_side;        // SideT of the tile's entries (coef = -(index - side_off + 1))
    int mat_off, n_mat;          // materials of the rows the tile evaluates (host bookkeeping)
    int prv_end;                 // entries whose u_{t0-1} is read (updated ones, then anchors)
    int mat_mode;                // 0: 4-entry material table; 1: (lambda, mu) per slot
    int pad[3];
};

// Node whose update needs more than the plain interior formula: absorbing
// dashpots, hanging-node anchors, plane-wave forces or a point source.
// word: bits 0-2 transfer count, 3-4 point kind (1 force, 2 prescribed z),
// 5 forced, 6-15 tile-local force index.  tr[t] = 2 * vrow + (0 if the node
// is the hanging node's first anchor, 1 if its second): the index of the
// transfer term that the hanging-node row computes for this anchor.
struct SideT {
    double invm[2];      // 1 / e_mfull per DOF
    double c2[2];        // 2 * a2 * C_n (Lysmer damping of the node itself)
    uint32_t word;
    uint32_t tr[4];
    uint32_t pad[3];
};

constexpr int kMaxTransfers = 4;   // a node anchors at most one hanging node per incident edge

// Hanging node whose reactions anchors of the tile take (sum_reactanchors,
// P2 weight 1/2): its rows, mass and damping.
struct VRowT {
    double a1m;          // a1 * M_v
    double a2c[2];       // a2 * C_v
    int v_slot;          // LDS slot of the hanging node
    int force;           // tile-local force index or -1
    uint32_t desc[4];    // same layout as a node descriptor
    int anc[2];          // LDS slots of its two anchors (as in the tile's virt list)
    int pad[2];
};
static_assert(sizeof(VRowT) % 16 == 0, "hanging-node rows are read as whole records");

// Node (active or hanging) with an external force: plane-wave pairs
// [first, first + count) and/or a point source.
struct ForceT {
    int first, count;
    int point;           // 1: force on z, 2: prescribed z displacement
    int inst;
};

// Receiver sampled by a tile: bilinear weights at the clamped point
// (shapefuncs), each corner as two LDS slots (the same slot twice for a
// stored node, the two anchors for a hanging node).
struct RecvT {
    int r;               // receiver index in the layout
    int ff;              // free-field point, or -1
    int hang;            // bit k: corner k is a hanging node
    int pad;
    int2 s[4];
    double w[4];
};

static_assert(sizeof(SideT) % 16 == 0 && sizeof(RecvT) % 16 == 0 && sizeof(FFT) % 16 == 0,
              "records copied to LDS as 16-byte words");

// Receiver in global node ids, for sampling straight from a state level in
// HBM (the last step of a run call): corner k is the stored node a[k], or for
// a hanging corner (bit k of hang) the mean of a[k] and b[k].
struct RecvG {
    int a[4], b[4];
    double w[4];
    int ff;              // free-field point, or -1
    int hang;
    int found;           // 0: receiver outside the mesh, zeros
    int pad;
};

// Device control block: the step base of graph-launched kernels, set by a
// one-thread kernel before a run's first graph launch and advanced by one at
// the end of every graph.
struct Ctrl {
    int base;
    int pad[3];
};

// All device pointers and sizes the step kernel needs.
struct DevView {
    int na, nv, ne;
    int nrecv;
    int block;           // core nodes per tile on average (tiles are cut at equal work)
    int threads;         // threads per workgroup (one LDS entry each)
    int depth;           // rings per tile = most steps one launch may take
    int node_tiles;      // ceil(na / block)
    int xcd_remap;       // 1: consecutive tiles stay on one XCD
    int tiles_per_xcd;
    int has_elem;        // some tile reads (lambda, mu) per slot (mat_mode 1)
    const void* mat_lm;       // (lambda, mu) per material entry, double2 or float2
    const void* node_em;      // per active node: (lambda, mu) of the element whose corner 0 it is
    const void* virt_em;      // per entry of the tiles' hanging-node lists: the same
    const TileInfo* tiles;
    const uint4* core_desc;   // per active node
    const void* core_coef;    // per active node: 1/mfull, or -(tile side + 1)
    const int* ring_gid;      // ring entries (one-step layouts)
    // several steps per launch: every entry of a tile, at a fixed stride (so
    // the loads need not wait for the tile record): node | kPrevFlag when
    // u_{t0-1} is read too, -1 past the tile's last entry
    const int* elist;
    int elist_cap;
    const uint4* ring_desc;   // per ring entry
    const void* ring_coef;
    const SideT* side;
    const VRowT* vrows;
    const int2* virt;         // hanging nodes: LDS slots of the two anchors
    const int* force_slots;
    const RecvT* recvs;
    // source tables: step j in row j mod chunk
    int chunk, nff, nforce;
    const FFT* ffp;           // free-field points
    const ForceT* forces;     // forced nodes
    const DrmElem* pairs;     // plane-wave pairs, corners are free-field points
    const double4* fftab;     // [chunk][nff]
    const double2* ftab;      // [chunk][nforce]
    const void* tile_lm;      // per tile 4 (lambda, mu), double2 or float2
    // LDS sizing (maxima over tiles)
    int lds_slots, lds_vrows, lds_virt, lds_force, lds_side, lds_recv, lds_mat;
    const InstParams* inst;
    // receiver samples: sample q (the step q * recv_interval) in slot
    // q mod recv_cap of [recv_cap][nrecv][2]
    double* recv_buf;
    int recv_cap;
    int recv_interval;
    const RecvG* rg;          // receivers in global node ids
    void* U[kStateSlots];     // u_j in U[j mod kStateSlots], 16 B per active node
    Ctrl* ctrl;
};

} // namespace dmp
