// rt_device.hpp — device building blocks shared by every sweep kernel (gfx950).
//
//  * TinyMT32 held in VGPRs: one 127-bit state per site; the parameter triple either
//    per generator (a table entry) or uniform (kernel arguments, i.e. SGPRs).
//  * epoch flags for the persistent kernels' neighbour synchronisation (every wait ends:
//    it gives up after a fixed number of polls and raises an error flag the host reports).
//  * XCD-aware block -> tile remapping.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rt {

// Chain memory is always device (global) memory, but a pointer the compiler meets as a
// generic pointer (a field of a table in memory, as in batched launches) makes it emit flat
// loads and stores, which also count against the LDS wait counter: every wait on an LDS access
// then waits for the flat loads too.  gload/gstore access through a global-address-space
// pointer (the HIP vector types through the compiler's own vector types, whose copies may bind
// to any address space).
#define RT_GLOBAL __attribute__((address_space(1)))
typedef unsigned int rt_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int rt_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint4 gload(const uint4* p) {
    const rt_u32x4 v = *(const RT_GLOBAL rt_u32x4*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 gload(const uint2* p) {
    const rt_u32x2 v = *(const RT_GLOBAL rt_u32x2*)p;
    return make_uint2(v.x, v.y);
}
template <class T>
__device__ __forceinline__ T gload(const T* p) {
    return *(const RT_GLOBAL T*)p;
}
__device__ __forceinline__ void gstore(uint4* p, uint4 v) { *(RT_GLOBAL rt_u32x4*)p = rt_u32x4{v.x, v.y, v.z, v.w}; }
__device__ __forceinline__ void gstore(uint2* p, uint2 v) { *(RT_GLOBAL rt_u32x2*)p = rt_u32x2{v.x, v.y}; }
template <class T>
__device__ __forceinline__ void gstore(T* p, T v) {
    *(RT_GLOBAL T*)p = v;
}

// Rule thresholds on k = u >> 9 (the 23-bit mantissa of the reference's rd), computed on
// the host from the reference's float/double predicates.  A transition X -> Y fires iff
// lo <= k < hi.  Indexed [colour class][row parity].
struct FlipRule {
    uint32_t ab_lo[3][2], ab_hi[3][2];  // A -> B (domino 3 -> 12, lozenge 25 -> 38)
    uint32_t ba_lo[3][2], ba_hi[3][2];  // B -> A
};

// Rectangle-triangle rule (recttrianglekernel.cl:25-118) as integer thresholds.
struct RectTriRule {
    uint32_t gate;        // proceed iff k < gate  (rd < 0.9f)
    uint32_t band[3];     // rd/0.9f < .25/.5/.75  iff  k < band[b]
    uint32_t cyc_acc[5][4];  // 5-cycle state s, band b: accept iff k < cyc_acc[s][b]
    uint32_t tog_acc[12];    // paired toggles: accept iff k < tog_acc[t]
    // The 17 states that can move, found by one hash probe instead of 17 comparisons:
    // slot h = (state * kRtHash mod 2^32) >> 27 holds the state (hkey, 0xffffffff when empty)
    // and its index (hval: 0-4 the 5-cycle, 5-16 the toggles); the targets per index and band.
    uint32_t hkey[32];
    uint32_t hval[32];
    int32_t cyc_nt[5][4];
    int32_t tog_nt[12];
};
// A multiplier (24-bit, for the 24 x 24-bit multiply) under which the 17 movable
// rect-triangle states fall in distinct slots of 32.
constexpr uint32_t kRtHash = 0xc9c201u;

struct RuleParams {
    FlipRule fr;
    RectTriRule rt;
};

// TinyMT32 step + temper (tinymt32.clh:39-55, 62-73), returning k = u >> 9.
__device__ __forceinline__ uint32_t tm_draw(uint32_t& s0, uint32_t& s1, uint32_t& s2, uint32_t& s3, uint32_t mat1,
                                            uint32_t mat2, uint32_t tmat) {
    uint32_t x = (s0 & 0x7fffffffu) ^ s1 ^ s2;
    uint32_t y = s3;
    x ^= x << 1;
    y ^= (y >> 1) ^ x;
    const uint32_t m = 0u - (y & 1u);
    const uint32_t n1 = s2 ^ (m & mat1);
    const uint32_t n2 = x ^ (y << 10) ^ (m & mat2);
    s0 = s1;
    s1 = n1;
    s2 = n2;
    s3 = y;
    const uint32_t t1 = s0 + (s2 >> 8);
    const uint32_t t0 = s3 ^ t1 ^ ((0u - (t1 & 1u)) & tmat);
    return t0 >> 9;
}

// gfx950 three-input bitwise op.  The truth table is indexed by (a << 2) | (b << 1) | c.
template <class F>
constexpr uint32_t tt3(F f) {
    uint32_t t = 0;
    for (int i = 0; i < 8; i++)
        if (f((i >> 2) & 1, (i >> 1) & 1, i & 1)) t |= 1u << i;
    return t;
}
#define RT_BITOP3(a, b, c, expr) \
    __builtin_amdgcn_bitop3_b32((a), (b), (c), tt3([](int A, int B, int C) { return (expr); }))

// The same TinyMT32 step as tm_draw, returning the full tempered output t0 (k = t0 >> 9),
// written with three-input bitwise ops: 16 VALU operations per draw.
__device__ __forceinline__ uint32_t tm_draw_raw(uint32_t& s0, uint32_t& s1, uint32_t& s2, uint32_t& s3, uint32_t mat1,
                                                uint32_t mat2, uint32_t tmat) {
    const uint32_t a = s0 & 0x7fffffffu;
    const uint32_t x1 = RT_BITOP3(a, s1, s2, A ^ B ^ C);
    const uint32_t x = x1 ^ (x1 << 1);
    const uint32_t y = RT_BITOP3(s3, s3 >> 1, x, A ^ B ^ C);
    const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int)y, 0, 1);  // 0 or ~0 from bit 0
    const uint32_t n1 = RT_BITOP3(s2, m, mat1, A ^ (B & C));
    const uint32_t z = RT_BITOP3(y << 10, m, mat2, A ^ (B & C));
    const uint32_t n2 = z ^ x;
    s0 = s1;
    s1 = n1;
    s2 = n2;
    s3 = y;
    const uint32_t t1 = s0 + (n2 >> 8);
    const uint32_t m2 = (uint32_t)__builtin_amdgcn_sbfe((int)t1, 0, 1);
    return RT_BITOP3(t1, m2, tmat, A ^ (B & C)) ^ y;
}

// tinymt32_init (tinymt32.clh:182-203)
__device__ __forceinline__ void tm_init(uint32_t seed, uint32_t mat1, uint32_t mat2, uint32_t tmat, uint32_t& s0,
                                        uint32_t& s1, uint32_t& s2, uint32_t& s3) {
    uint32_t st[4] = {seed, mat1, mat2, tmat};
#pragma unroll
    for (int i = 1; i < 8; i++)
        st[i & 3] ^= (uint32_t)i + 1812433253u * (st[(i - 1) & 3] ^ (st[(i - 1) & 3] >> 30));
    s0 = st[0];
    s1 = st[1];
    s2 = st[2];
    s3 = st[3];
    if ((s0 & 0x7fffffffu) == 0 && s1 == 0 && s2 == 0 && s3 == 0) {
        s0 = 'T';
        s1 = 'I';
        s2 = 'N';
        s3 = 'Y';
    }
#pragma unroll
    for (int i = 0; i < 8; i++) (void)tm_draw(s0, s1, s2, s3, mat1, mat2, tmat);
}

// blockIdx -> tile index such that the blocks an XCD receives (round-robin dispatch:
// block b runs on XCD b % 8) form one contiguous run of tiles, so halo exchanges between
// neighbouring tiles mostly stay inside one XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    constexpr int X = 8;
    const int x = b % X, r = b / X;
    const int base = nb / X, extra = nb % X;
    const int off = x * base + (x < extra ? x : extra);
    return off + r;
}

// Flag lines of the persistent kernels' neighbour synchronisation: one 128-byte line per
// tile holding the last block epoch the tile published.  Epochs increase monotonically
// across launches (the host tracks them) and are compared modulo 2^32.  Every wait on a
// flag is bounded: after a fixed number of polls the waiter raises an error flag and stops
// waiting, and a raised error flag ends every later wait at once, so every wave reaches
// the kernel's end.
constexpr int kLine = 32;  // words per flag line

__device__ __forceinline__ bool epoch_reached(uint32_t v, uint32_t epoch) { return (int32_t)(v - epoch) >= 0; }

}  // namespace rt
