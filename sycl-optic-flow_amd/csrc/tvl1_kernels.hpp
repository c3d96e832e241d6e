// tvl1_kernels.hpp -- device-side data structures of the TV-L1 solver.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace amdflow {

constexpr int kMaxScales = 32;
constexpr int kMinMaxBlocks = 1024;   // grid of the frame conversion + min/max kernel
// Workgroup groups of the two-level grid barrier: workgroup b belongs to
// group b % 8, the XCD the dispatcher places it on.
constexpr int kBarGroups = 8;
struct alignas(256) BarWord {
    unsigned v;
};
struct alignas(256) SeqWord {
    unsigned long long v;
};

// Device control block of one solve.  Written by the last workgroup of each
// iteration launch (and by the set-up kernels), read by every workgroup at
// the start of the next launch: kernel boundaries order the accesses.
struct TvCtl {
    int cur[kMaxScales];        // which of the two u/p buffers is current, per scale
    int n;                      // iterations completed in the current warping
    int stop;                   // warping finished: further launches are no-ops
    int redo;                   // >0: re-run this many iterations from `cur`
    unsigned counter;           // workgroups finished in the running launch
    unsigned launches;          // iteration launches that did work (stats)
    unsigned wlaunches;         // the same, counted within the current warping
    unsigned long long maxkey[2];   // (|x| bits << 32) | (~index), per image
    unsigned long long minkey[2];   // (|x| bits << 32) | index, per image
    unsigned long long mmpart[4][kMinMaxBlocks];   // per-workgroup max0, max1, min0, min1
    // grid barrier of the persistent warping kernel: the arrival counter and
    // the generation word that waiting workgroups poll sit on cache lines of
    // their own, away from each other and from the fields above
    alignas(256) unsigned bar_count;
    alignas(256) unsigned bar_gen;
    BarWord grp_count[kBarGroups];   // two-level barrier: per-group arrivals
    BarWord grp_gen[kBarGroups];     // and per-group release epochs
    // neighbour-synchronised warping kernel, by chunk parity: workgroups
    // whose error partials of the chunk are published, and the stopping
    // decision of the chunk, (seq << 24) | ((chunk + 1) << 4) | keep
    BarWord chunk_count[2];
    SeqWord decision[2];
    alignas(256) int error;         // kBarTimeout or kBarNotResident, else 0
};

// ctl->error values.  kBarTimeout: a grid barrier wait ran past its poll
// bound.  kBarNotResident: the first grid barrier of a persistent launch did
// not see every workgroup arrive within its wall-clock bound, so the grid may
// not be co-resident; the host then repeats the solve with host-paced
// launches.
constexpr int kBarTimeout = 1;
constexpr int kBarNotResident = 2;

struct IterArgs {
    const float* wx;
    const float* wy;
    const float* rhoc;
    float* u[2];        // u[b]: u1 plane, u2 plane at +plane
    float* p[2];        // p[b]: p11, p12, p21, p22 planes, stride `plane`
    int nx, ny;
    int plane;          // nx*ny
    float l_t, theta, taut;
    double eps2;        // (float)(epsilon*epsilon) widened
    int max_iter;
    int level;
    TvCtl* ctl;
    int* iters_slot;    // receives the warping's iteration count
    double* partials;   // [K][nblocks] per-workgroup error partials
    unsigned long long* host_status;  // host-mapped progress word (may be null)
    unsigned epoch;
    double* err_log;    // if set: per-iteration error sums, indexed by iteration
    // persistent kernel: wall-clock ticks that the first grid barrier waits
    // for all workgroups before it reports kBarNotResident (0: no bound)
    unsigned long long resident_ticks;
    int force_not_resident;   // test hook: the first barrier reports kBarNotResident
    // neighbour-synchronised warping kernel: per tile, (seq << 20) | chunks
    // completed.  `seq` numbers the warpings of the process from 1 (never
    // reused, below 2^40): it tags a warping's flags and decisions, so words
    // left by earlier warpings, of any handle whose memory a buffer reuses,
    // always compare below the current warping's
    unsigned long long* tile_flags;
    unsigned long long seq;
};

// Launch modes of a warping (LaunchRule::persistent, of_tvl1_set_launch_rule).
constexpr int kHostPaced = 0;
constexpr int kPersistBarrier = 1;     // one launch, grid barrier between chunks
constexpr int kPersistNeighbour = 2;   // one launch, neighbour flags between chunks
// Chunks of one neighbour-synchronised warping are counted in 20 bits.
constexpr int kNbrMaxChunks = (1 << 20) - 2;
constexpr unsigned long long kNbrMaxSeq = 1ull << 40;
// The next warping number of the process (tvl1_iter.hip).
unsigned long long next_warp_seq();

// Launch helpers implemented in tvl1_iter.hip.
int launch_iter(int K, int R, const IterArgs& a, hipStream_t s);
bool iter_config_ok(int K, int R);
void iter_grid(int K, int R, int nx, int ny, unsigned* gx, unsigned* gy);
// Whole warping in one launch of a grid that fits the device at once;
// `nblocks` = 0 picks the largest such grid not above the tile count.  Its
// first grid barrier reports kBarNotResident if the grid did not fully start.
// `mode`: kPersistBarrier or kPersistNeighbour.
int launch_persist(int K, int R, const IterArgs& a, int nblocks, hipStream_t s,
                   int mode = kPersistBarrier);
// Tiles of launch shape (K, R) at nx x ny (the size of a tile-flag array).
unsigned iter_tiles(int K, int R, int nx, int ny);
// Process-wide: true = warpings whose launch rule is persistent run
// host-paced instead (after a residency failure, with AMDFLOW_HOST_PACED=1,
// or when set); false (default) = persistent launches.
bool persist_host_paced();
void set_persist_host_paced(bool on);
// Test hook: persistent launches report kBarNotResident at their first
// barrier while set.
void set_persist_force_not_resident(bool on);

}  // namespace amdflow
