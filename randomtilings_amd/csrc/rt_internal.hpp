// rt_internal.hpp — declarations shared by the kernel and host translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/rt_hip.h"
#include "rt_device.hpp"

namespace rt {

// Context accessors and the thread-local error message (rt_last_error) for the other units.
int ctx_device(rt_ctx* c);
hipStream_t ctx_stream(rt_ctx* c);
int set_error(int code, const std::string& msg);

// Internal model id of the coupled domino pair (the two chains of coupling from the past).
constexpr int kDominoPair = 4;

struct StepwiseArgs {
    RuleParams rp;
    void* cls;  // NC contiguous class arrays of N rows of pitch P elements
    int N, W, P;
    uint4* gens;
    const uint4* params;  // nullptr => uniform parameters m1/m2/tm
    uint32_t m1, m2, tm;
    const uint8_t* sched_host;  // colour class per step
    int64_t nsteps;
};

// One chain of a batched launch: the blocks [tile0, tile0 + ntiles) of the grid run its tiles.
struct BatchEntry {
    RuleParams rp;
    void* cls[2];
    uint4* gens[2];
    const uint8_t* sched;  // the chain's colour schedule for the whole batched run
    int N, W, P, tiles_c;
    int buf0, g0;  // copies holding the state / generators when the run starts
    int tile0, ntiles;
};

struct PersistArgs {
    RuleParams rp;
    void* cls[2];  // two copies of the NC contiguous class arrays
    const uint4* gen_in;
    uint4* gen_out;
    const uint4* params;
    uint32_t m1, m2, tm;
    const uint8_t* sched;  // device, colour class per step
    int64_t nsteps;
    int N, W, P;
    int tiles_c;
    int buf;  // which copy holds the current state
    uint32_t* flags;  // one kLine-word line per tile: the last block epoch it published
    uint32_t epoch0;  // block epochs completed before this launch
    int* err;
    uint32_t spin_limit;
    // Skipped tiles (packed lozenge kernel): tiles whose cells can never change are not run.
    // live lists the tiles that run, one workgroup each, and live_map[t] is 1 for those;
    // ntiles counts every tile of the grid.  live == nullptr: every tile runs.
    const int32_t* live;
    const uint8_t* live_map;
    int ntiles;
    // batched runs (blocked path): every block looks up its chain; the steps done so far in
    // the run are read from step_ctr (a device counter, so one captured round replays as is)
    const BatchEntry* batch;
    int nbatch;
    int block_h;
    const int64_t* step_ctr;
    int64_t step_off;  // steps this launch lies past the counter (launches of one captured round)
};

// A value every lane of the workgroup holds alike (derived from blockIdx and memory every lane
// reads at the same address), moved to a scalar register.  The compiler cannot prove such
// values uniform when they are loaded through a pointer it cannot prove unwritten, and then
// computes with them per lane and turns branches on them into exec-mask sequences.
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
template <class T>
__device__ __forceinline__ T* uniform(T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

// Resolve a batched block: the chain owning blockIdx.x, with the copies and schedule offset
// of the current block of steps.  Returns the chain's rule tables where they lie, in the entry
// in global memory: copying them into the caller's PersistArgs (which the kernel has to copy to
// change) would make every lane hold that copy in scratch memory, the tables being indexed
// per lane.
__device__ __forceinline__ const RuleParams& resolve_batch(PersistArgs& a, int& tile, int& ntiles) {
    const int b = blockIdx.x;
    int lo = 0, hi = a.nbatch - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.batch[mid].tile0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const BatchEntry& e = a.batch[lo];
    const int64_t s0 = *a.step_ctr + a.step_off;
    const int par = (int)((s0 / a.block_h) & 1);
    a.cls[0] = uniform(e.cls[0]);
    a.cls[1] = uniform(e.cls[1]);
    a.buf = uniform(e.buf0 ^ par);
    a.gen_in = uniform(e.gens[e.g0 ^ par]);
    a.gen_out = uniform(e.gens[e.g0 ^ par ^ 1]);
    a.sched = uniform(e.sched) + s0;
    a.N = uniform(e.N);
    a.W = uniform(e.W);
    a.P = uniform(e.P);
    a.tiles_c = uniform(e.tiles_c);
    tile = uniform(b - e.tile0);
    ntiles = uniform(e.ntiles);
    return e.rp;
}

hipError_t launch_init_gens(int64_t G, uint32_t seed, const uint4* params, uint32_t m1, uint32_t m2, uint32_t tm,
                            uint4* gens, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, int n, uint32_t v, hipStream_t st);
hipError_t launch_pair_differs(const uint16_t* a, int64_t n, uint32_t* count, hipStream_t st);
hipError_t launch_stepwise(int model, const StepwiseArgs& a, hipStream_t st);

// Height functions (rt_heights.hip).  heights: initial values (INT_MAX = unset), replaced by
// the closure; fixed: vertices that may not move.  With `tiling`, also reads the tiling off.
int height_closure_device(int device, hipStream_t s, int model, const std::vector<int>& d, int dir,
                          std::vector<int32_t>& heights, const std::vector<uint8_t>& fixed, std::vector<int32_t>* tiling,
                          int* launches, std::string& err);
bool boundary_heights_host(int model, const std::vector<int>& d, std::vector<int32_t>& h, std::vector<uint8_t>& fixed);

int persist_config_count(int model);
bool persist_config(int model, int idx, int* tr, int* tc, int* h, int* nt);
bool persist_config_packed(int model, int idx);
int persist_capacity(int model, int idx, bool uni, int device);
bool persist_uses_scratch(int model, int idx, bool uni);
int persist_waves_per_cu(int model, int idx, bool uni, bool batch, bool one);
// coop: launch cooperatively, so that the runtime refuses a grid whose workgroups cannot all
// be resident at once (hipErrorCooperativeLaunchTooLarge) instead of queuing tiles that would
// wait on tiles that never start.
hipError_t launch_persist(int model, int idx, bool uni, const PersistArgs& a, int nblocks, hipStream_t st,
                          bool coop = false);
hipError_t launch_add_i64(int64_t* p, int64_t v, hipStream_t st);

// Domino height observables (rt_observe.hip): chains points at nchains device pointers to
// class arrays (two classes of N rows, pitch P bytes); out: 4 uint64 per chain (count, sum of
// heights, sum of squares, sum of |256 h - ref256|) over the masked vertices, added to.
hipError_t launch_domino_observe(const void* chains, int nchains, int N, int P, const uint8_t* faces,
                                 const int32_t* hmin, const uint8_t* mask, const int32_t* ref256, int32_t* heights,
                                 unsigned long long* out, hipStream_t st);

// The allocation-free entry point's device memory (rt_noalloc.hip): a file-scope array of
// kNoAllocScratchBytes per device.  noalloc_setup takes its device address once per device
// (with that device current); noalloc_scratch returns it, or nullptr with the reason.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
void noalloc_setup(int device);
void* noalloc_scratch(int device, std::string& why);

}  // namespace rt
