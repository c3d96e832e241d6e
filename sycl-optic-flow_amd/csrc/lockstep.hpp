// lockstep.hpp -- device tables of the batched TV-L1 solver.
//
// A batch is solved scale by scale, coarse to fine, in lockstep: every
// launch covers the same scale (and warping) of all pairs of the batch.
// Iteration launches walk a work list of (pair, tile) entries; each pair
// keeps its own control block (TvCtl) and stopping decision, taken by the
// workgroup that finishes the pair's last tile, exactly as in a one-pair
// launch.  The workgroup that finishes a launch last rebuilds the list
// from the pairs that are still iterating, so later launches skip the
// converged ones.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "tvl1_kernels.hpp"

namespace amdflow {

constexpr int kLsMaxPairs = 512;   // pairs per lockstep group

// One pair at one scale s of a lockstep solve.
struct LsPair {
    IterArgs a;            // iteration arguments of the pair at scale s
    const float* I0s;      // pyramid images at scale s
    const float* I1s;
    float4* wsrc;          // (I1, I1x, I1y, 0) of scale s
    float* p0;             // dual buffer 0 (4 planes of a.plane floats)
    float* ufine;          // buffer 0 of scale s-1 (null at s = 0)
    float* dst;            // the pair's output flow (planar, used at s = 0)
    int nxf, nyf;          // size of scale s-1
    float fx, fy;          // zoom-in factors nxf/nx, nyf/ny
    float invz;            // 1/zfactor
    float* ucoarse;        // buffer 0 of the flow at scale s when s is the pair's
                           // coarsest (zero initial flow), else null
    unsigned pair_id;      // index of the pair in the batch
    unsigned gx;           // iteration tiles per row at the scale's launch shape
    unsigned ntiles;       // iteration tiles
    unsigned bx;           // 64x4 pixel blocks per row at scale s
    unsigned bxf;          // the same at scale s-1
};

// The tables of one scale, passed by value to every lockstep kernel.
struct LsPhase {
    const LsPair* pairs;        // npairs entries (the pairs that have scale s)
    const unsigned* blk_base;   // [npairs+1] prefix sums of pixel blocks at scale s
    const unsigned* fblk_base;  // [npairs+1] the same at scale s-1
    const unsigned* tile_base;  // [npairs+1] prefix sums of iteration tiles
    int* iters;                 // iteration counts [pair_id][kMaxScales][warps]
    int npairs;
    unsigned nblk;              // pixel blocks at scale s, all pairs
    unsigned nfblk;             // the same at scale s-1
    unsigned ntiles;            // iteration tiles, all pairs
    int level;                  // s
    int warps;                  // warpings per scale
    int warp;                   // the running warping (warp set-up and iterations)
};

// Device control of a lockstep group.
struct LsCtl {
    unsigned counter;    // workgroups finished in the running iteration launch
    unsigned launches;   // iteration launches completed in the running warping
    unsigned nlist;      // pairs still iterating
    unsigned ntiles;     // their tiles
    unsigned list_pair[kLsMaxPairs];   // phase-local index of each listed pair
    unsigned list_base[kLsMaxPairs];   // first work-list entry of each listed pair
};

// Phase 1 (frames to pyramids) of one pair, for the batched preparation.
struct LsPrep {
    const uint8_t* s0;     // the u8 frames
    const uint8_t* s1;
    float* A0;             // frame 0: float copy, then blur pass buffers
    float* A1;
    float* B0;             // frame 1: blur pass buffers (the handle's warping scratch)
    float* B1;
    float* I0p;            // the pyramids
    float* I1p;
    TvCtl* ctl;
    unsigned n0;           // W*H
    unsigned nmm;          // workgroups of the min/max pass
    int ns;
    int nx[kMaxScales], ny[kMaxScales];
    unsigned off[kMaxScales];
};

// One step of the batched preparation: the participating pairs and the
// prefix sums of their work items.  step 0 = conversion / normalisation /
// pre-smoothing at scale 0; step s >= 1 builds scale s from scale s-1.
struct LsPrepStep {
    const LsPrep* preps;
    const unsigned* part;   // [npart] batch indices of the participating pairs
    const unsigned* base;   // [npart+1] prefix sums of work items
    int npart;
    unsigned nitems;
    int s;
};

// Batched preparation launches (batch_lockstep.hip).  load: u8 -> float and
// per-workgroup min/max keys (items = nmm per pair); final: keys combined
// and control blocks reset (one workgroup per pair, npairs of them);
// normalize: items = 256-pixel chunks; blur rows / cols and zoom-out: items
// = 64x4 blocks of both frames.
int launch_lp_load(const LsPrepStep& st, hipStream_t s);
int launch_lp_final(const LsPrep* preps, int npairs, hipStream_t s);
int launch_lp_normalize(const LsPrepStep& st, hipStream_t s);
int launch_lp_blur(const LsPrepStep& st, const GaussTaps& g, hipStream_t s);
int launch_lp_zoom_out(const LsPrepStep& st, float zfactor, hipStream_t s);

// Iteration launch over the work list (tvl1_iter.hip): `grid` workgroups
// take the listed tiles round robin.  host_status receives
// (epoch << 32) | (all stopped << 31) | completed launches.
int launch_ls_iter(int K, int R, const LsPhase& ph, LsCtl* mc, unsigned grid,
                   unsigned long long* host_status, unsigned epoch, hipStream_t s);
// Workgroups of k_ls_iter<K,R> that fit on the device at once.
int ls_iter_resident(int K, int R, unsigned* blocks);
// Scale start: packed gradient of I1 and dual buffer 0 cleared, all pairs.
int launch_ls_level_begin(const LsPhase& ph, hipStream_t s);
// Warping start: warp set-up of all pairs, their control blocks reset and
// the work list set to every pair.
int launch_ls_warp_setup(const LsPhase& ph, LsCtl* mc, hipStream_t s);
// Scale end: current flow of scale s carried to buffer 0 of scale s-1.
int launch_ls_zoom_in(const LsPhase& ph, hipStream_t s);
// Finest scale's current flow into each pair's output.
int launch_ls_copy_flow(const LsPhase& ph, hipStream_t s);

}  // namespace amdflow
