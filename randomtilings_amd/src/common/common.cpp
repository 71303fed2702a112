// common.cpp — matrix utilities and the cl:: compatibility objects
// (reference: src/common/common.cpp:4-161 for the matrix functions' behaviour).
#include "common.h"

#include <stdexcept>

#include "rt_hip.h"

namespace rtapi {
void check(int rc, const char* what) {
    if (rc != RT_OK) throw std::runtime_error(std::string(what) + ": " + rt_last_error());
}
}  // namespace rtapi

// ---- cl:: shim ------------------------------------------------------------------------

cl_int cl::Device::getInfo(cl_device_info what, std::string* out) const {
    if (what != CL_DEVICE_NAME || !out) return -30;  // CL_INVALID_VALUE
    char buf[256];
    if (rt_device_name(id_, buf, sizeof(buf)) != RT_OK) return -33;  // CL_INVALID_DEVICE
    *out = buf;
    return CL_SUCCESS;
}

cl::Context::Context(cl_device_type type) : device_(0) { (void)type; }

cl_int cl::Context::getInfo(cl_context_info what, std::vector<Device>* out) const {
    if (what != CL_CONTEXT_DEVICES || !out) return -30;
    out->clear();
    const int n = rt_device_count();
    for (int d = 0; d < n; d++) out->push_back(Device(d));
    if (out->empty()) out->push_back(Device(0));  // keep devices[0] addressable, as callers assume
    return CL_SUCCESS;
}

void PrintOpenCLInfo() {
    const int n = rt_device_count();
    std::cout << "Number of HIP devices: " << n << std::endl;
    for (int d = 0; d < n; d++) {
        char buf[256];
        if (rt_device_name(d, buf, sizeof(buf)) == RT_OK) std::cout << "\tDevice " << d << ": " << buf << std::endl;
    }
}

// ---- matrices -------------------------------------------------------------------------

template <class V>
static void save_rows(const std::vector<V>& m, int M, int N, const std::string& filename) {
    std::ofstream f(filename.c_str());
    for (int i = 0; i < M; ++i) {
        for (int j = 0; j < N; ++j) f << m[(size_t)i * N + j] << " ";
        f << "\n";
    }
}

void SaveMatrix(const std::vector<int>& m, int M, int N, std::string filename) { save_rows(m, M, N, filename); }
void SaveMatrix(const std::vector<double>& m, int M, int N, std::string filename) { save_rows(m, M, N, filename); }

// Square shorthands: the side is the integer part of sqrt(size), as in the reference, so a
// non-square vector is written as its leading floor(sqrt(n))^2 entries.
void SaveMatrix(const std::vector<int>& m, std::string filename) {
    const int s = (int)std::sqrt((double)m.size());
    SaveMatrix(m, s, s, filename);
}
void SaveMatrix(const std::vector<double>& m, std::string filename) {
    const int s = (int)std::sqrt((double)m.size());
    SaveMatrix(m, s, s, filename);
}

void PrintMatrix(const std::vector<int>& m, int M, int N) {
    for (int i = 0; i < M; ++i) {
        for (int j = 0; j < N; ++j) {
            const int v = m[(size_t)i * N + j];
            if (v < infty) std::cout << v << " ";
            else std::cout << "  ";
        }
        std::cout << "\n";
    }
}

void PrintMatrix(const std::vector<int>& m) {
    const int s = (int)std::sqrt((double)m.size());
    PrintMatrix(m, s, s);
}

void PadMatrix(std::vector<int>& v, int M, int N) {
    std::vector<int> out((size_t)(M + 2) * (N + 2), 0);
    for (int i = 0; i < M; i++)
        for (int j = 0; j < N; j++) out[(size_t)(i + 1) * (N + 2) + j + 1] = v[(size_t)i * N + j];
    v.swap(out);
}

void UnPadMatrix(std::vector<int>& v, int M, int N) {
    std::vector<int> out((size_t)(M - 2) * (N - 2));
    for (int i = 0; i < M - 2; i++)
        for (int j = 0; j < N - 2; j++) out[(size_t)i * (N - 2) + j] = v[(size_t)(i + 1) * N + j + 1];
    v.swap(out);
}

void PadMatrix(std::vector<int>& v) {
    const int s = (int)std::sqrt((double)v.size());
    PadMatrix(v, s, s);
}

void UnPadMatrix(std::vector<int>& v) {
    const int s = (int)std::sqrt((double)v.size());
    UnPadMatrix(v, s, s);
}

// Every maximal run of ASCII digits becomes one entry; everything else separates entries
// (so a '-' sign is dropped, exactly as the reference's reader does).
std::vector<int> LoadMatrix(std::string filename) {
    std::ifstream is(filename.c_str());
    if (!is.is_open()) {
        std::cout << "Cannot open file: " << filename << std::endl;
        exit(1);
    }
    std::vector<int> mat;
    bool in_num = false;
    int val = 0;
    char c;
    while (is.get(c)) {
        if (c >= '0' && c <= '9') {
            val = in_num ? 10 * val + (c - '0') : (c - '0');
            in_num = true;
        } else if (in_num) {
            mat.push_back(val);
            in_num = false;
        }
    }
    if (in_num) mat.push_back(val);
    return mat;
}

// ---- rtapi helpers ------------------------------------------------------------------------

#include "rtapi.hpp"

namespace rtapi {

std::shared_ptr<rt_ctx> make_ctx(int device) {
    rt_ctx* c = nullptr;
    check(rt_create(device, &c), "creating the HIP context");
    return std::shared_ptr<rt_ctx>(c, [](rt_ctx* p) { rt_destroy(p); });
}

std::string basename(const std::string& path) {
    const size_t k = path.find_last_of("/\\");
    return k == std::string::npos ? path : path.substr(k + 1);
}

void load_default_params(rt_ctx* ctx, int64_t n) {
    std::vector<uint32_t> m((size_t)n * 3);
    for (int64_t i = 0; i < n; i++) {
        m[3 * i + 0] = 0x8f7011eeu;
        m[3 * i + 1] = 0xfc78ff1fu;
        m[3 * i + 2] = 0x3793fdffu;
    }
    check(rt_load_params(ctx, m.data(), n), "loading the default TinyMT parameters");
}

}  // namespace rtapi
