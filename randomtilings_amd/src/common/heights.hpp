// heights.hpp — height functions on planar lattices: boundary walk, Thurston's max/min
// construction and tileability checks, shared by the domino and lozenge classes.
//
// A Geom describes the vertex lattice (an N x N array of vertices, DEG neighbours each):
//   di[nb], dj[nb]            neighbour offsets
//   delta(i, j, nb)           the height change along the edge when no tile crosses it
//                              (+1 or -1; a crossed edge changes the height by -K*delta)
//   faces(i, j, nb, &a, &b)   whether each of the two faces beside the edge is in the domain
//
// The maximal height function is the min-plus closure of the boundary heights over the
// directed "+1" edges inside the domain (every other constraint h(v) <= h(u) + K follows by
// walking around a face that is inside the domain); the minimal one is the max-plus closure
// over the "-1" edges.  The domain is tileable iff the boundary walk closes consistently and
// the closure lowers (raises) no boundary height.  Unit weights make the closure a
// bucket-ordered breadth-first search, linear in the number of vertices.
#pragma once

#include <algorithm>
#include <climits>
#include <cstdint>
#include <deque>
#include <vector>

namespace rtapi {

constexpr int kInf = INT_MAX;

// First vertex in row-major order whose class is 2 (boundary), or -1.
inline int first_boundary(const std::vector<int>& vertices) {
    for (size_t k = 0; k < vertices.size(); k++)
        if (vertices[k] == 2) return (int)k;
    return -1;
}

// Heights along the boundary: breadth-first from the first boundary vertex (height 0) over
// boundary edges (exactly one adjacent face in the domain) between boundary vertices.
// Returns false if the boundary loop does not close consistently.
template <class Geom>
bool boundary_heights(const Geom& g, const std::vector<int>& vertices, std::vector<int>& hf) {
    const int N = g.N;
    hf.assign((size_t)N * N, kInf);
    const int s = first_boundary(vertices);
    if (s < 0) return true;
    hf[s] = 0;
    std::deque<int> q{s};
    while (!q.empty()) {
        const int u = q.front();
        q.pop_front();
        const int i0 = u / N, j0 = u % N;
        for (int nb = 0; nb < Geom::DEG; nb++) {
            const int i = i0 + g.di[nb], j = j0 + g.dj[nb];
            if (i < 0 || j < 0 || i >= N || j >= N) continue;
            const int v = i * N + j;
            if (vertices[v] != 2) continue;
            bool fa, fb;
            g.faces(i0, j0, nb, fa, fb);
            if (fa == fb) continue;  // not a boundary edge
            const int h = hf[u] + g.delta(i0, j0, nb);
            if (hf[v] == kInf) {
                hf[v] = h;
                q.push_back(v);
            } else if (hf[v] != h) {
                return false;
            }
        }
    }
    return true;
}

// Extend boundary heights into the domain.  dir = +1 builds the maximal height function
// (closure over +1 edges, lowest heights first), dir = -1 the minimal one.  Returns false if
// a boundary height would have to move (the domain has no tiling).
template <class Geom>
bool thurston_closure(const Geom& g, const std::vector<int>& vertices, std::vector<int>& hf, int dir) {
    const int N = g.N;
    // bucket queue keyed by dir*h, increasing
    long lo = LONG_MAX;
    for (size_t k = 0; k < hf.size(); k++)
        if (hf[k] != kInf && vertices[k] > 0) lo = std::min(lo, (long)dir * hf[k]);
    if (lo == LONG_MAX) return true;
    std::vector<std::vector<int>> buckets(1);
    auto push = [&](int v, long key) {
        const size_t b = (size_t)(key - lo);
        if (b >= buckets.size()) buckets.resize(b + 1);
        buckets[b].push_back(v);
    };
    for (size_t k = 0; k < hf.size(); k++)
        if (hf[k] != kInf && vertices[k] > 0) push((int)k, (long)dir * hf[k]);
    for (size_t b = 0; b < buckets.size(); b++) {
        for (size_t x = 0; x < buckets[b].size(); x++) {
            const int u = buckets[b][x];
            const int i0 = u / N, j0 = u % N;
            for (int nb = 0; nb < Geom::DEG; nb++) {
                if (g.delta(i0, j0, nb) != dir) continue;
                const int i = i0 + g.di[nb], j = j0 + g.dj[nb];
                if (i < 0 || j < 0 || i >= N || j >= N) continue;
                const int v = i * N + j;
                if (vertices[v] == 0) continue;
                bool fa, fb;
                g.faces(i0, j0, nb, fa, fb);
                if (!fa && !fb) continue;  // edge outside the domain
                const int h = hf[u] + dir;
                if (hf[v] == kInf) {
                    hf[v] = h;
                    push(v, (long)dir * h);
                } else if ((long)dir * h < (long)dir * hf[v]) {
                    return false;  // only boundary vertices are preset
                }
            }
        }
        std::vector<int>().swap(buckets[b]);
    }
    return true;
}

// Height function of a tiling: breadth-first from the first boundary vertex (height 0)
// over edges inside the domain that no tile crosses (crossed(i, j, nb) true when it is).
template <class Geom, class Crossed>
std::vector<int> tiling_heights(const Geom& g, const std::vector<int>& vertices, Crossed crossed) {
    const int N = g.N;
    std::vector<int> hf((size_t)N * N, kInf);
    const int s = first_boundary(vertices);
    if (s < 0) return hf;
    hf[s] = 0;
    std::deque<int> q{s};
    while (!q.empty()) {
        const int u = q.front();
        q.pop_front();
        const int i0 = u / N, j0 = u % N;
        for (int nb = 0; nb < Geom::DEG; nb++) {
            if (crossed(i0, j0, nb)) continue;
            const int i = i0 + g.di[nb], j = j0 + g.dj[nb];
            if (i < 0 || j < 0 || i >= N || j >= N) continue;
            const int v = i * N + j;
            if (vertices[v] == 0 || hf[v] != kInf) continue;
            bool fa, fb;
            g.faces(i0, j0, nb, fa, fb);
            if (!fa && !fb) continue;
            hf[v] = hf[u] + g.delta(i0, j0, nb);
            q.push_back(v);
        }
    }
    return hf;
}

}  // namespace rtapi
