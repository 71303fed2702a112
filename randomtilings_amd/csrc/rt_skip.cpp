// rt_skip.cpp — tiles of a lozenge chain whose cells can never change.
//
// A lozenge domain is embedded in the N x N vertex array with zero codes outside it
// (LozengeTiler.cpp:31-37 splits the array into three colour classes of N x N/3).  The
// reference rotates every position of every class (lozengekernel.cl:11-12), but a zero cell
// never flips (only codes 25 and 38 do, lozengekernel.cl:21-24) and the update rewrites a
// cell's edge bits from its neighbours' (lozengekernel.cl:36-54), so a cell whose whole
// neighbourhood is zero stays zero as long as no change reaches it.  In a tiling that is a
// fixed point of the update (every edge bit agrees with the neighbour it mirrors, which holds
// for every valid tiling), changes start only at flips, and a flip at a code-25/38 vertex
// touches only its six neighbours, which in a valid tiling are vertices of the tiled region.
// A tile of the persistent sweep whose cells, halo and a margin are all zero therefore never
// changes and need not run: the tiling after any number of steps is the same, bit for bit.
// The sweep kernel re-checks this while it runs (the halo words inside a skipped tile must
// stay zero) and raises the error flag if a change ever gets that far, so an array that is
// consistent but not a valid tiling cannot go wrong silently.
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <vector>

namespace rt {

// Whether the three class arrays (each N rows of pitch P, W used columns) are a fixed point
// of the update of every class: applying lozengekernel.cl:36-54 for class c = 0, 1, 2 to the
// current state leaves every updatable cell (rows 1..N-2, columns 1..W-2) as it is.
bool lozenge_consistent(const uint8_t* cls, int N, int W, int P) {
    const size_t CS = (size_t)N * P;
    for (int c = 0; c < 3; c++) {
        const uint8_t* s = cls + c * CS;
        const uint8_t* d2 = cls + (c == 2 ? 0 : c + 1) * CS;
        const uint8_t* d3 = cls + (c == 0 ? 2 : c - 1) * CS;
        for (int i = 1; i <= N - 2; i++) {
            const int r = i % 3;
            const int p2 = r == (2 * c + 1) % 3, p4 = r == (2 * c + 2) % 3, p8 = r == (2 * c) % 3;
            for (int j = 1; j <= W - 2; j++) {
                const size_t x = (size_t)i * P + j;
                const int n2 = (d2[x] & ~38) | ((s[x - P + p2] & 16) >> 3) | ((s[x - p4] & 8) >> 1) |
                               ((s[x + P] & 1) << 5);
                const int n3 = (d3[x] & ~25) | ((s[x - P] & 32) >> 5) | ((s[x + p8] & 4) << 1) |
                               ((s[x + P - p2] & 2) << 3);
                if (n2 != d2[x] || n3 != d3[x]) return false;
            }
        }
    }
    return true;
}

// 2-D prefix counts of the positions (N rows, P columns) holding a nonzero cell in any class.
void lozenge_nonzero_counts(const uint8_t* cls, int N, int P, std::vector<int32_t>& ps) {
    const size_t CS = (size_t)N * P;
    ps.assign((size_t)(N + 1) * (P + 1), 0);
    for (int i = 0; i < N; i++) {
        int32_t row = 0;
        for (int j = 0; j < P; j++) {
            const size_t x = (size_t)i * P + j;
            row += (cls[x] | cls[CS + x] | cls[2 * CS + x]) != 0;
            ps[(size_t)(i + 1) * (P + 1) + j + 1] = ps[(size_t)i * (P + 1) + j + 1] + row;
        }
    }
}

// Tiles of TR x TC positions (rows x class columns, halo H) that run: a tile is skipped when
// every cell of every class within its halo plus a margin of two is zero.  map gets one byte
// per tile, row-major (1 = runs); returns the number that run.
int lozenge_live_tiles(const std::vector<int32_t>& ps, int N, int W, int P, int TR, int TC, int H,
                       std::vector<uint8_t>& map) {
    auto count = [&](int r0, int r1, int c0, int c1) {
        return ps[(size_t)r1 * (P + 1) + c1] - ps[(size_t)r0 * (P + 1) + c1] - ps[(size_t)r1 * (P + 1) + c0] +
               ps[(size_t)r0 * (P + 1) + c0];
    };
    constexpr int kMargin = 2;
    const int trn = (N + TR - 1) / TR, tcn = (W + TC - 1) / TC;
    map.assign((size_t)trn * tcn, 1);
    int live = 0;
    for (int r = 0; r < trn; r++)
        for (int c = 0; c < tcn; c++) {
            const int r0 = std::max(0, r * TR - H - kMargin), r1 = std::min(N, (r + 1) * TR + H + kMargin);
            const int c0 = std::max(0, c * TC - H - kMargin), c1 = std::min(P, (c + 1) * TC + H + kMargin);
            const bool run = count(r0, r1, c0, c1) != 0;
            map[(size_t)r * tcn + c] = run;
            live += run;
        }
    return live;
}

}  // namespace rt
