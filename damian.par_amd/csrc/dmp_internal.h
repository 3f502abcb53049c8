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
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace dmp {

constexpr int kMaxBlock = 256;

// Per instance constants read by the auxiliary (plane wave + receiver) lanes.
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

// Node whose update needs more than the plain interior formula: absorbing
// dashpots, hanging-node anchors, or plane-wave forces.
struct SideRec {
    double invm[2];      // 1 / e_mfull per DOF
    double c2[2];        // 2 * a2 * C_n (Lysmer damping of the node itself)
    int tr_first, tr_count;
    int drm_slot;        // -1: no plane-wave force at this node
    int point;           // 1: point force on z, 2: prescribed z displacement
    int inst, pad;
};

// Hanging node v anchored by the side node: weight w of its reactions
// (sum_reactanchors P2: w = 1/2), its mass and damping.
struct TransferRec {
    double w;
    double a1m;          // a1 * M_v
    double a2c[2];       // a2 * C_v
    int v;               // virtual index (0-based within [na, na+nv) minus na)
    int drm_slot;        // -1 or plane-wave slot of v
};

// Tiles: block b of the node part updates active nodes [t*B, t*B+B) of tile
// t.  The tile's values live in LDS slots: owned nodes [0, cnt), halo nodes
// [cnt, cnt+n_extra), hanging (virtual) nodes after them.  A node's element
// rows are described per corner slot k (the node is corner k of that
// element) by one word holding the LDS slots of the other three corners (10
// bits each, ascending corner order).  Compact descriptors (16 B per node):
// bits 30-31 select one of 4 (lambda, mu) of the tile's palette, and the
// word 0xFFFFFFFF marks an empty slot.  Full descriptors (32 B per node): bit
// 30 marks a present element and a second word holds the material index.
struct TileInfo {
    int extra_off, n_extra;  // halo nodes: global ids in extra_gid
    int virt_off, n_virt;    // hanging nodes: (anchor slot, anchor slot) in virt_anch
    int vrow_off, n_vrow;    // hanging nodes whose rows the tile's anchors need
    int pad0, pad1;
};

// Rows of a hanging node gathered by its anchors, computed once per tile.
struct VRow {
    int v_lds;           // LDS slot of the hanging node
    int drm_slot;        // plane-wave force at the hanging node, -1 none
    int pad0, pad1;
    uint32_t desc[8];    // same layout as a node descriptor
};

// Per transfer of a side node, tile-local view.
struct TransferT {
    double w, a1m, a2c[2];
    int vrow;            // index into the tile's VRow list
    int v_lds;           // LDS slot of the hanging node
    int partner_lds;     // LDS slot of the other anchor
    int partner_gid;     // global id of the other anchor (for u_{i-2})
};

// Node of the plane-wave layer: delays of the three wave trains at the node,
// v_w = (x S_w + z C_w) / V_w (pwaves.f90:153-155, 189-191), time independent.
struct FFNode {
    double v[3];
    int inst, pad;
};

// A node that receives plane-wave (DRM) forces and its seismic elements.
struct DrmSlot {
    int first, count;    // into DrmElem
    int inst, pad;
};

struct DrmElem {
    int c[4];            // layer-node (FFNode) index of each corner
    double lam, mu, mass;
    int k;               // corner of this element at the slot node
    int ext;             // bit j set: corner j is exterior (model.f90:778-793)
};

struct RecvRec {
    int c[4];            // corner ids (active or virtual); c[0] < 0 -> not found
    double w[4];         // bilinear weights (shapefuncs at the clamped point)
    double v[3];         // free-field delays at the point
    int ff;              // add the free field (inside_soil)
    int inst;
};

// Device control block (device memory; written by host memcpy outside graphs
// and by the step-advance kernel inside them).
struct Ctrl {
    int base;            // step index base for graph-launched kernels
    int block_start;     // s: step reached before the current run call
    int slot_base;       // s / interval
    int slot_offset;     // 1 when the step-0 sample is slot 0 of this run
};

template <class T> struct Vec2 { T x, y; };

// All device pointers and sizes the step kernel needs.
struct DevView {
    int na, nv, ne;
    int nside, ndrm, nrecv;
    int block;           // threads per block used for the node part
    int naux_blocks;     // leading blocks that run plane-wave / receiver lanes
    int node_tiles;      // ceil(na / block)
    int xcd_remap;       // 1: consecutive tiles stay on one XCD
    int tiles_per_xcd;
    const int2* vanch;        // nv  : anchors (active ids)
    const int4* elem_corners; // ne  : corners (active id, or na + virtual index)
    const void* mat_lm;       // (lambda, mu) per material entry, double2 or float2
    const void* coef;         // per active node: 1/mfull, or -(side+1)
    const SideRec* side;
    const TransferT* tr;      // tile-local transfer records
    // tiles
    const TileInfo* tiles;
    const int* extra_gid;
    const int2* virt_anch;
    const VRow* vrows;
    const uint4* desc;        // per active node: 1 (compact) or 2 (full) uint4
    int compact;              // 1: 16-byte descriptors with a per-tile palette
    const void* palette;      // compact: 4 (lambda, mu) per tile, double2 or float2
    int lds_val_slots;        // max LDS value slots over tiles
    int lds_vrows;            // max hanging rows over tiles
    int nff;                  // plane-wave layer nodes
    const FFNode* ff;
    double4* ffval[3];        // (u0x, u0z, a0x, a0z) at step j in ffval[j % 3]
    const DrmSlot* drm;
    const DrmElem* drm_el;
    double2* drm_force[2];    // double buffered by step parity
    const RecvRec* recv;
    double* recv_buf;         // [slot][nrecv][2]
    int recv_interval;
    const InstParams* inst;
    void* U[3];               // rotating state buffers: u_i in U[i % 3]
    Ctrl* ctrl;
};

} // namespace dmp
