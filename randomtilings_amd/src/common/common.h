#ifndef SRC_COMMON_UTIL_H_
#define SRC_COMMON_UTIL_H_

// common.h — shared types and matrix utilities of the RandomTilings C++ API
// (the reference's src/common/common.h:40-85), on top of the HIP library.
//
// The standard headers below are part of the interface: callers of the reference get
// <random>, <chrono>, <iostream> ... through this header (e.g. Examples/
// MinimalLozengeEx.cpp:12,172 use std::chrono and std::random_device without including
// them).

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <fstream>
#include <iostream>
#include <queue>
#include <random>
#include <sstream>
#include <stack>
#include <string>
#include <time.h>
#include <vector>

#include "cl_compat.h"

#define infty INT_MAX
#define sqrt3 1.7320508075688772935274

typedef std::vector<int> domain;
typedef std::vector<int> heightfunc;
typedef std::vector<int> tiling;

// Prints the visible HIP devices (the reference prints its OpenCL platforms).
void PrintOpenCLInfo();

// Saves an MxN matrix to the disk with name filename (whitespace-separated decimals).
void SaveMatrix(const std::vector<double> &tiling, int M, int N, std::string filename);
void SaveMatrix(const std::vector<int> &tiling, int M, int N, std::string filename);

// Prints an MxN matrix to the console (entries equal to infty print as blanks).
void PrintMatrix(const std::vector<int> &mat, int M, int N);

// Adds/removes a zero padding around an MxN matrix v.
void PadMatrix(std::vector<int> &v, int M, int N);
void UnPadMatrix(std::vector<int> &v, int M, int N);

// The above functions for square matrices.
void PadMatrix(std::vector<int> &v);
void UnPadMatrix(std::vector<int> &v);
void PrintMatrix(const std::vector<int> &mat);
void SaveMatrix(const std::vector<int> &mat, std::string filename);
void SaveMatrix(const std::vector<double> &mat, std::string filename);

// Loads a non-negative integer matrix (every maximal run of decimal digits is one entry).
std::vector<int> LoadMatrix(std::string filename);

// Heap orderings on (vertex, height) pairs, kept for source compatibility.
struct incheightPriority {
    bool operator()(std::pair<std::pair<int, int>, int> a, std::pair<std::pair<int, int>, int> b) const {
        return a.second < b.second;
    }
};

struct decheightPriority {
    bool operator()(std::pair<std::pair<int, int>, int> a, std::pair<std::pair<int, int>, int> b) const {
        return a.second > b.second;
    }
};

namespace rtapi {
// Side of a square array stored as a flat vector (the reference's `int N = sqrt(t.size())`).
inline int side(size_t n) { return (int)std::lround(std::sqrt((double)n)); }
// Throws std::runtime_error carrying rt_last_error() when rc is not RT_OK.
void check(int rc, const char* what);
}  // namespace rtapi

#endif /* SRC_COMMON_UTIL_H_ */
