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

// Steps whose rotation draws one draw launch makes: every site's generator advances that many
// times and the outcome of each draw (tempered output < 2^31) is kept as one bit per site and
// step.  The tile kernels only read those bits, so a halo site's number is never drawn twice.
constexpr int kDrawSteps = 32;

// Words (64 sites each) of one row of a rotation-bit plane of an M x M move lattice.
__host__ __device__ inline int td_bit_words(int M) { return (M + 63) / 64; }
// Bytes of the rotation bits of a chain: kDrawSteps planes of M rows.
inline size_t td_bit_bytes(int M) { return sizeof(uint64_t) * (size_t)kDrawSteps * M * td_bit_words(M); }

struct TdArgs {
    const uint2* cells_in;  // M*M cells: the eight move states of a site in 64 bits
    uint2* cells_out;
    const uint4* gen_in;  // TinyMT states, generator id = site index
    uint4* gen_out;
    const uint4* params;  // per-generator (mat1, mat2, tmat, 0), or nullptr
    uint32_t m1, m2, tm;  // uniform parameters when params == nullptr
    int M;
    const uint8_t* sched;  // device move codes (single-workgroup kernel)
    int64_t nsteps;
    uint64_t* rbits;  // rotation bits: plane k, row gi, word gj / 64 (bit gj % 64)
    int plane;              // plane of the launch's first step
};

// One triangle-dimer chain of a batched launch (step blocks [tile0, tile0 + ntiles), draw
// blocks [dblk0, dblk0 + ndblk)).
struct TdBatchEntry {
    uint2* cells[2];
    uint4* gens[2];
    uint64_t* rbits;
    const uint8_t* sched;  // move codes of the whole batched run
    int M;
    int buf0, g0;
    int tile0, ntiles;
    int dblk0, ndblk;
};

// Move code of one reference step: type | r1 << 2 | r2 << 4 | p1 << 5 | p2 << 7.
inline uint8_t td_move(int type, int r1, int r2, int p1, int p2) {
    return (uint8_t)(type | r1 << 2 | r2 << 4 | p1 << 5 | p2 << 7);
}

int td_config_count();
bool td_config(int idx, int* tr, int* tc, int* nt);
int td_config_base(int idx);   // the one-step configuration with the same tiles
int td_config_steps(int idx);  // steps per launch: 1, or 2 for the two-step configurations
hipError_t td_launch_derive(const int32_t* t, int N, uint2* cells, hipStream_t st);
hipError_t td_launch_merge(const uint2* cells, int N, int32_t* t, hipStream_t st);
// Draw the rotation bits of k <= kDrawSteps steps: gen_in -> gen_out advanced k draws.
hipError_t td_launch_draw(const TdArgs& a, bool uni, int k, hipStream_t st);
// Batched: every chain's draws for the k steps from *step_ctr + off (a multiple of kDrawSteps
// from the run's start); blocks = the summed ndblk.
hipError_t td_launch_draw_batch(const TdArgs& a, bool uni, const TdBatchEntry* batch, int nbatch, int blocks,
                                const int64_t* step_ctr, int64_t off, int k, hipStream_t st);
// Draw blocks of one chain.
int td_draw_blocks(int M);
hipError_t td_launch_small(const TdArgs& a, bool uni, hipStream_t st);
// One step on the tiles of configuration cfg, its rotation bits in plane a.plane.
hipError_t td_launch_step(int cfg, const TdArgs& a, uint32_t move, hipStream_t st);
// Two steps in one launch on the tiles of one-step configuration cfg (planes a.plane and
// a.plane + 1); moves = first | second << 8.
hipError_t td_launch_step2(int cfg, const TdArgs& a, uint32_t moves, hipStream_t st);
// cfg: one-step configuration (tiles); the launch's first step is *step_ctr + off; two: it runs
// two steps; spl: steps per launch of the run's earlier launches (fixes the copy parity)
hipError_t td_launch_step_batch(int cfg, const TdArgs& a, const TdBatchEntry* batch, int nbatch, int blocks,
                                const int64_t* step_ctr, int64_t off, bool two, int spl, hipStream_t st);

}  // namespace rt
