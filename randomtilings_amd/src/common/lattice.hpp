// lattice.hpp — vertex lattices of the domino and lozenge models: neighbour stencils, the
// height step along an uncrossed edge, which faces border an edge, and the classification of
// vertices against a domain.  Shared by the tiler classes (host statics) and the device
// height-function path of librt_hip.so.
//
// Domino (DominoTiler.cpp:202-232 of the reference): an M x M domain of unit squares, an
// N x N vertex array (N = M + 1); neighbour nb = down, up, right, left carries state bit
// 2, 1, 8, 4; crossing a black-on-left edge raises the height by one.
// Lozenge (LozengeTiler.cpp:12-16): an M x 2M domain of triangles, N x N vertices;
// neighbours (0,+1), (-1,+1), (-1,0), (0,-1), (+1,-1), (+1,0) carry bits 8, 2, 1, 4, 16, 32
// and height steps -1, +1, -1, +1, -1, +1.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <thread>
#include <vector>

namespace rtlat {

// f(i) for every row i in [0, n), the rows split over up to 16 host threads (row loops over
// vertex arrays of large domains; every row writes only its own entries).
template <class F>
void for_rows(int n, F f) {
    const int t = std::max(1, std::min({16, (int)std::thread::hardware_concurrency(), n / 64}));
    if (t <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> pool;
    for (int k = 0; k < t; ++k)
        pool.emplace_back([&, k] {
            for (int i = k * n / t; i < (k + 1) * n / t; ++i) f(i);
        });
    for (auto& th : pool) th.join();
}

inline int side_of(size_t n) { return (int)std::lround(std::sqrt((double)n)); }

struct DominoGeom {
    static constexpr int DEG = 4;
    static constexpr int di[4] = {1, -1, 0, 0};
    static constexpr int dj[4] = {0, 0, 1, -1};
    static constexpr int bit[4] = {2, 1, 8, 4};
    const std::vector<int>& d;
    int N, M;  // vertices per side, faces per side
    DominoGeom(const std::vector<int>& dd, int n) : d(dd), N(n), M(n - 1) {}
    bool face(int fi, int fj) const { return fi >= 0 && fj >= 0 && fi < M && fj < M && d[(size_t)fi * M + fj] != 0; }
    // the two squares beside the edge from vertex (i,j) in direction nb
    void faces(int i, int j, int nb, bool& a, bool& b) const {
        switch (nb) {
            case 0: a = face(i, j - 1); b = face(i, j); break;          // down
            case 1: a = face(i - 1, j - 1); b = face(i - 1, j); break;  // up
            case 2: a = face(i - 1, j); b = face(i, j); break;          // right
            default: a = face(i - 1, j - 1); b = face(i, j - 1); break; // left
        }
    }
    // From a vertex with i+j even the horizontal steps go up and the vertical ones down; the
    // other way round when i+j is odd.
    int delta(int i, int j, int nb) const { return (nb > 1 ? 1 : -1) * (((i + j) & 1) == 0 ? 1 : -1); }
};

struct LozGeom {
    static constexpr int DEG = 6;
    static constexpr int di[6] = {0, -1, -1, 0, 1, 1};
    static constexpr int dj[6] = {1, 1, 0, -1, -1, 0};
    static constexpr int bit[6] = {8, 2, 1, 4, 16, 32};
    static constexpr int dh[6] = {-1, 1, -1, 1, -1, 1};
    // the two triangles beside edge nb, as (row offset, column offset from 2j) in the M x 2M domain
    static constexpr int face_off[6][2][2] = {{{0, 0}, {-1, 1}},   {{-1, 1}, {-1, 0}}, {{-1, 0}, {-1, -1}},
                                              {{-1, -1}, {0, -2}}, {{0, -2}, {0, -1}}, {{0, -1}, {0, 0}}};
    const std::vector<int>& d;
    int N, M;
    LozGeom(const std::vector<int>& dd, int n) : d(dd), N(n), M(n - 1) {}
    // Flat index into the M x 2M domain, the way the reference addresses it: a column offset
    // of -1/-2 at j = 0 reaches the end of the previous row.  Out-of-array reads count as empty.
    bool face_flat(long r, long c) const {
        const long k = r * 2 * M + c;
        return k >= 0 && k < (long)d.size() && d[(size_t)k] != 0;
    }
    void faces(int i, int j, int nb, bool& a, bool& b) const {
        a = face_flat(i + face_off[nb][0][0], 2L * j + face_off[nb][0][1]);
        b = face_flat(i + face_off[nb][1][0], 2L * j + face_off[nb][1][1]);
    }
    int delta(int, int, int nb) const { return dh[nb]; }
};

// Vertex classes: 1 = all adjacent faces in the domain, 2 = some, 0 = none.
inline std::vector<int> domino_vertices(const std::vector<int>& d) {
    const int M = side_of(d.size());
    const int N = M + 1;
    auto at = [&](int fi, int fj) {
        fi = std::min(std::max(fi, 0), M - 1);
        fj = std::min(std::max(fj, 0), M - 1);
        return d[(size_t)fi * M + fj] != 0;
    };
    std::vector<int> v((size_t)N * N, 0);
    for_rows(N, [&](int i) {
        for (int j = 0; j < N; ++j) {
            const bool a = at(i - 1, j - 1), b = at(i - 1, j), c = at(i, j - 1), e = at(i, j);
            const bool border = (i == 0 || j == 0 || i == M || j == M);
            if (!border && a && b && c && e) v[(size_t)i * N + j] = 1;
            else if (a || b || c || e) v[(size_t)i * N + j] = 2;
        }
    });
    return v;
}

inline std::vector<int> lozenge_vertices(const std::vector<int>& d) {
    const int M = (int)std::lround(std::sqrt(d.size() / 2.0));
    const int N = M + 1;
    std::vector<int> v((size_t)N * N, 0);
    auto f = [&](int r, int c) { return d[(size_t)r * 2 * M + c] != 0; };
    for_rows(M, [&](int i) {
        if (i < 1) return;
        for (int j = 1; j < M; ++j) {
            // the six triangles around vertex (i,j)
            const bool a = f(i, 2 * j), b = f(i, 2 * j - 1), c = f(i, 2 * j - 2);
            const bool e = f(i - 1, 2 * j), g = f(i - 1, 2 * j + 1), h = f(i - 1, 2 * j - 1);
            if (a && b && c && e && g && h) v[(size_t)i * N + j] = 1;
            else if (a || b || c || e || g || h) v[(size_t)i * N + j] = 2;
        }
    });
    return v;
}

}  // namespace rtlat
