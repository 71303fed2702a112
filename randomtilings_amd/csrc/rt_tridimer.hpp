// rt_tridimer.hpp — host/device interface of the triangle-dimer chain kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rt {

// Halo of the one-step tile kernel: the longest chain of neighbour-reading update kernels in
// one reference step (butterfly moves: four butterfly updates, lozenges, the other
// butterflies and up-triangles, then down-triangles).
constexpr int kHalo = 6;
// Largest move lattice (M x M sites) the single-workgroup kernel holds in LDS.
constexpr int kSmallMax = 64;

struct TdArgs {
    const uint4* cells_in;  // M*M cells: eight 16-bit move states each
    uint4* cells_out;
    const uint4* gen_in;  // TinyMT states, generator id = site index
    uint4* gen_out;
    const uint4* params;  // per-generator (mat1, mat2, tmat, 0), or nullptr
    uint32_t m1, m2, tm;  // uniform parameters when params == nullptr
    int M;
    const uint8_t* sched;  // device move codes (single-workgroup kernel)
    int64_t nsteps;
};

// One triangle-dimer chain of a batched launch (blocks [tile0, tile0 + ntiles)).
struct TdBatchEntry {
    uint4* cells[2];
    uint4* gens[2];
    const uint8_t* sched;  // move codes of the whole batched run
    int M;
    int buf0, g0;
    int tile0, ntiles;
};

// Move code of one reference step: type | r1 << 2 | r2 << 4 | p1 << 5 | p2 << 7.
inline uint8_t td_move(int type, int r1, int r2, int p1, int p2) {
    return (uint8_t)(type | r1 << 2 | r2 << 4 | p1 << 5 | p2 << 7);
}

int td_config_count();
bool td_config(int idx, int* tr, int* tc, int* nt);
int td_config_base(int idx);   // the one-step configuration with the same tiles
int td_config_steps(int idx);  // steps per launch: 1, or 2 for the two-step configurations
hipError_t td_launch_derive(const int32_t* t, int N, uint4* cells, hipStream_t st);
hipError_t td_launch_merge(const uint4* cells, int N, int32_t* t, hipStream_t st);
hipError_t td_launch_small(const TdArgs& a, bool uni, hipStream_t st);
hipError_t td_launch_step(int cfg, const TdArgs& a, bool uni, uint32_t move, hipStream_t st);
// Two steps in one launch on the tiles of one-step configuration cfg; moves = first | second << 8.
hipError_t td_launch_step2(int cfg, const TdArgs& a, bool uni, uint32_t moves, hipStream_t st);
// cfg: one-step configuration (tiles); the launch's first step is *step_ctr + off; two: it runs
// two steps; spl: steps per launch of the run's earlier launches (fixes the copy parity)
hipError_t td_launch_step_batch(int cfg, const TdArgs& a, bool uni, const TdBatchEntry* batch, int nbatch, int blocks,
                                const int64_t* step_ctr, int64_t off, bool two, int spl, hipStream_t st);

}  // namespace rt
