// rt_capi.cpp — the C ABI (include/rt_hip.h): contexts, TinyMT parameter tables, chains,
// the host colour schedule and the reference-compatible Walk entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <random>
#include <tuple>
#include <string>
#include <vector>

#include "rt_internal.hpp"
#include "rt_tridimer.hpp"

namespace rt {
bool domino_rule(int rule, const double* p12, FlipRule& fr);
bool lozenge_rule(int rule, double q, FlipRule& fr);
bool recttri_rule(const float* weights16, RectTriRule& r);
bool lozenge_consistent(const uint8_t* cls, int N, int W, int P);
void lozenge_nonzero_counts(const uint8_t* cls, int N, int P, std::vector<int32_t>& ps);
int lozenge_live_tiles(const std::vector<int32_t>& ps, int N, int W, int P, int TR, int TC, int H,
                       std::vector<uint8_t>& map);
}  // namespace rt

using namespace rt;

static thread_local std::string g_err;

// exchange flags per chain: one 128-byte line per tile of the persistent grid
static constexpr int kMaxBlocks = 4096;
static constexpr size_t kBarBytes = (size_t)kMaxBlocks * kLine * sizeof(uint32_t);

static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail(RT_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint4* params = nullptr;  // (mat1, mat2, tmat, 0) per generator id
    int64_t nparams = 0;
    bool uniform = false;  // every entry identical -> pass as kernel arguments
    uint32_t um1 = 0, um2 = 0, utm = 0;
    int default_path = RT_PATH_AUTO;
    int default_cfg = -1;
    hipEvent_t marks[16] = {};
    int live_chains = 0;   // chains created on this context and not yet destroyed
    bool closing = false;  // rt_destroy was called while chains were live
};

struct rt_chain {
    rt_ctx* ctx = nullptr;
    int model = 0, rule = 0;
    int N = 0, W = 0, NC = 0, esize = 0;
    int P = 0;  // row pitch in elements: rows start 16-byte aligned
    int64_t G = 0;  // generators used by Rotate
    RuleParams rp{};
    void* cls[2] = {nullptr, nullptr};
    int cur = 0;
    uint4* gens[2] = {nullptr, nullptr};
    int gcur = 0;
    uint8_t* dsched = nullptr;
    int64_t dsched_cap = 0;
    uint32_t* misc = nullptr;  // [0] unused, [1] error flag
    uint32_t* bar = nullptr;   // barrier lines: kMaxBlocks arrival lines + 8 release lines
    uint32_t epoch = 0;        // barrier epochs completed once queued launches finish
    bool seeded = false;
    std::mt19937 mt;                      // domino schedule (DominoTiler.cpp:31,81)
    std::minstd_rand0 lcg;                // lozenge / rect-tri schedule (LozengeTiler.cpp:50-52)
    std::uniform_int_distribution<int> d3{0, 2};
    std::uniform_int_distribution<int> d2{0, 1};  // triangle-dimer schedule (TriangleDimerTiler.cpp:211-213)
    int32_t* tsave = nullptr;  // triangle dimers: the caller's tiling; merge writes the move states into it
    uint64_t* tbits = nullptr;  // triangle dimers: rotation bits of one draw block (td_bit_bytes)
    int path = RT_PATH_AUTO;
    int cfg = -1;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_sched = nullptr;  // the last schedule upload has read hsched
    bool sched_pending = false;
    bool timed = false;
    std::vector<uint8_t> hsched;
    int64_t last_launches = 0;
    int32_t last_cfg = -1;
    bool counted = false;  // counted in ctx->live_chains (chains handed out by rt_chain_create)
    // Lozenge chains: tiles whose cells never change are not run by the packed persistent
    // sweep (rt_skip.cpp).  consistent: the uploaded tiling is a fixed point of the update;
    // runs[cfg]: which tiles of that configuration run (built at creation, uploaded on use).
    struct Runs {
        std::vector<uint8_t> map;  // one byte per tile, row-major; empty: every tile runs
        std::vector<int32_t> list;
        int32_t* d_list = nullptr;
        uint8_t* d_map = nullptr;
    };
    bool skip = true;
    bool consistent = false;
    std::vector<Runs> runs;
    int32_t last_tiles = 0, last_running = 0;
    int64_t coop_refused = 0;  // resident launches the runtime refused (continued blocked)
    int coop_error = 0;        // the hipError_t of the last refusal
    // Device memory borrowed from the caller's scratch (rt_lozenge_walk_noalloc): nothing of it
    // is allocated or freed here, and there are no events.
    bool borrowed = false;
};

// The tiles a persistent run of configuration cfg skips, or nullptr when every tile runs.
static rt_chain::Runs* skip_runs(rt_chain* ch, int cfg) {
    if (!ch->skip || !ch->consistent || cfg < 0 || cfg >= (int)ch->runs.size()) return nullptr;
    rt_chain::Runs& r = ch->runs[cfg];
    return r.map.empty() ? nullptr : &r;
}

static void build_runs(rt_chain* ch, const std::vector<uint8_t>& h) {
    ch->runs.clear();
    if (ch->model != RT_LOZENGE) return;
    ch->consistent = lozenge_consistent(h.data(), ch->N, ch->W, ch->P);
    if (!ch->consistent) return;
    std::vector<int32_t> ps;
    lozenge_nonzero_counts(h.data(), ch->N, ch->P, ps);
    const int n = persist_config_count(ch->model);
    ch->runs.resize(n);
    for (int i = 0; i < n; i++) {
        int tr, tc, hh, nt;
        if (!persist_config_packed(ch->model, i) || !persist_config(ch->model, i, &tr, &tc, &hh, &nt)) continue;
        std::vector<uint8_t> map;
        const int running = lozenge_live_tiles(ps, ch->N, ch->W, ch->P, tr, tc, hh, map);
        if (running == (int)map.size() || running == 0) continue;
        rt_chain::Runs& r = ch->runs[i];
        r.map = map;
        for (int t = 0; t < (int)map.size(); t++)
            if (map[t]) r.list.push_back(t);
    }
}

static hipError_t upload_runs(rt_chain::Runs& r) {
    if (r.d_list) return hipSuccess;
    hipError_t e = hipMalloc(&r.d_list, r.list.size() * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&r.d_map, r.map.size());
    if (e == hipSuccess) e = hipMemcpy(r.d_list, r.list.data(), r.list.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(r.d_map, r.map.data(), r.map.size(), hipMemcpyHostToDevice);
    return e;
}

static void free_runs(rt_chain* ch) {
    for (auto& r : ch->runs) {
        if (r.d_list) hipFree(r.d_list);
        if (r.d_map) hipFree(r.d_map);
        r.d_list = nullptr;
        r.d_map = nullptr;
    }
}

// Tile configurations chosen by rt_chain_autotune, keyed by (model, side, uniform params).
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, bool>, int> g_tuned;

// Host colour schedule of the planar models, drawn exactly as the reference's Walk loop
// draws it.
static void draw_schedule(rt_chain* ch, int64_t steps, uint8_t* out) {
    for (int64_t s = 0; s < steps; s++) {
        if (ch->model == RT_DOMINO || ch->model == kDominoPair) {
            const int r = ch->mt() % 2;  // DominoTiler.cpp:35-36: r==1 rotates B (class 0)
            out[s] = (r == 1) ? 0 : 1;
        } else {
            out[s] = (uint8_t)ch->d3(ch->lcg);  // LozengeTiler.cpp:50-56, RectTriangleTiler.cpp:47-50
        }
    }
}

// ---- triangle dimers (TriangleDimerTiler::Walk, TriangleDimerTiler.cpp:86-483) ----------
// The chain state is M x M 64-bit cells of the eight move states (M = N/2), derived on the
// device from the uploaded tiling; the host draws each step's move exactly as Walk does.  The
// tiled kernels read every step's Rotate outcomes from rotation bits drawn kDrawSteps steps at
// a time.

// Device buffers of a triangle-dimer chain: two copies of the cells and of the generators, and
// the rotation bits.
static hipError_t td_alloc(rt_chain* ch) {
    hipError_t e = hipSuccess;
    for (int b = 0; b < 2 && e == hipSuccess; b++) e = hipMalloc(&ch->cls[b], sizeof(uint2) * (size_t)ch->G);
    for (int b = 0; b < 2 && e == hipSuccess; b++) e = hipMalloc(&ch->gens[b], sizeof(uint4) * (size_t)ch->G);
    if (e == hipSuccess) e = hipMalloc(&ch->tbits, td_bit_bytes(ch->W));
    return e;
}

static int td_create(rt_chain* ch, const int32_t* tiling) {
    const int N = ch->N;
    // TriangleDimerTiler.h:61-63: N even, three rows/columns of zero padding
    if (N < 8 || N % 2) return fail(RT_ERR_ARG, "triangle-dimer tiling side %d must be even and >= 8", N);
    for (int64_t i = 0; i < (int64_t)N * N; i++)
        if (tiling[i] < 0 || tiling[i] > 15)
            return fail(RT_ERR_ARG, "triangle-dimer lozenge state %d out of the 4-bit range", tiling[i]);
    const int M = N / 2;
    ch->NC = 8;
    ch->W = M;
    ch->P = M;
    ch->esize = 2;
    ch->G = (int64_t)M * M;
    hipError_t e = hipSetDevice(ch->ctx->device);
    if (e == hipSuccess) e = td_alloc(ch);
    if (e == hipSuccess) e = hipMalloc(&ch->tsave, sizeof(int32_t) * (size_t)N * N);
    if (e == hipSuccess) e = hipMalloc(&ch->misc, 64);
    if (e == hipSuccess) e = hipEventCreate(&ch->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ch->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ch->ev_sched, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMemcpy(ch->tsave, tiling, sizeof(int32_t) * (size_t)N * N, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(ch->misc, 0, 64);
    if (e == hipSuccess) e = td_launch_derive(ch->tsave, N, (uint2*)ch->cls[0], ch->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ch->ctx->stream);
    if (e != hipSuccess) return fail(RT_ERR_HIP, "triangle-dimer chain setup: %s", hipGetErrorString(e));
    ch->cur = 0;
    return RT_OK;
}

// TriangleDimerTiler.cpp:217-222, :291-292: type, r1, r2, then p1, p2 for butterflies
static void draw_td_moves(rt_chain* ch, int64_t steps, uint8_t* out) {
    for (int64_t s = 0; s < steps; s++) {
        const int type = ch->d3(ch->lcg), r1 = ch->d3(ch->lcg), r2 = ch->d2(ch->lcg);
        int p1 = 0, p2 = 0;
        if (type == 2) {
            p1 = ch->d3(ch->lcg);
            p2 = ch->d2(ch->lcg);
        }
        out[s] = td_move(type, r1, r2, p1, p2);
    }
}

static int td_run(rt_chain* ch, int64_t steps) {
    rt_ctx* c = ch->ctx;
    draw_td_moves(ch, steps, ch->hsched.data());
    TdArgs a{};
    a.params = c->uniform ? nullptr : c->params;
    a.m1 = c->um1;
    a.m2 = c->um2;
    a.tm = c->utm;
    a.M = ch->W;
    const bool small = ch->W <= kSmallMax && ch->path != RT_PATH_STEPWISE && ch->path != RT_PATH_BLOCKED;
    if (ch->path == RT_PATH_PERSISTENT && !small)
        return fail(RT_ERR_ARG, "the single-workgroup triangle-dimer kernel holds at most %d x %d sites", kSmallMax,
                    kSmallMax);
    HIPCHK(hipEventRecord(ch->ev0, c->stream));
    if (small) {
        if (ch->dsched_cap < steps) {
            if (ch->borrowed)
                return fail(RT_ERR_ARG, "rt_chain_run: %lld steps exceed the schedule buffer (%lld)", (long long)steps,
                            (long long)ch->dsched_cap);
            if (ch->dsched) {
                HIPCHK(hipStreamSynchronize(c->stream));
                HIPCHK(hipFree(ch->dsched));
                ch->dsched = nullptr;
            }
            HIPCHK(hipMalloc(&ch->dsched, (size_t)steps + 16));
            ch->dsched_cap = steps;
        }
        HIPCHK(hipMemcpyAsync(ch->dsched, ch->hsched.data(), (size_t)steps, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipEventRecord(ch->ev_sched, c->stream));
        ch->sched_pending = true;
        const int64_t chunk = 1 << 16;  // keeps single launches short
        for (int64_t s0 = 0; s0 < steps; s0 += chunk) {
            a.cells_in = (const uint2*)ch->cls[ch->cur];
            a.cells_out = (uint2*)ch->cls[ch->cur ^ 1];
            a.gen_in = ch->gens[ch->gcur];
            a.gen_out = ch->gens[ch->gcur ^ 1];
            a.sched = ch->dsched + s0;
            a.nsteps = std::min<int64_t>(chunk, steps - s0);
            HIPCHK(td_launch_small(a, c->uniform, c->stream));
            ch->cur ^= 1;
            ch->gcur ^= 1;
            ch->last_launches++;
        }
        ch->last_cfg = -1;
    } else {
        int cfg = ch->cfg;
        if (cfg < 0) {
            cfg = 0;
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tuned.find(std::make_tuple(ch->model, ch->N, c->uniform));
            if (it != g_tuned.end()) cfg = it->second;
        }
        // blocks of kDrawSteps steps: one draw launch, then the steps; two-step configurations
        // advance pairs of steps per launch, an odd last step on the one-step kernel of the same
        // tiles
        const int base = td_config_base(cfg);
        a.rbits = ch->tbits;
        for (int64_t s0 = 0; s0 < steps; s0 += kDrawSteps) {
            const int k = (int)std::min<int64_t>(kDrawSteps, steps - s0);
            a.gen_in = ch->gens[ch->gcur];
            a.gen_out = ch->gens[ch->gcur ^ 1];
            HIPCHK(td_launch_draw(a, c->uniform, k, c->stream));
            ch->gcur ^= 1;
            ch->last_launches++;
            for (int p = 0; p < k; p++) {
                const bool pair = td_config_steps(cfg) == 2 && p + 1 < k;
                a.cells_in = (const uint2*)ch->cls[ch->cur];
                a.cells_out = (uint2*)ch->cls[ch->cur ^ 1];
                a.plane = p;
                if (pair) {
                    HIPCHK(td_launch_step2(base, a, (uint32_t)ch->hsched[s0 + p] | ((uint32_t)ch->hsched[s0 + p + 1] << 8),
                                           c->stream));
                    p++;
                } else {
                    HIPCHK(td_launch_step(base, a, ch->hsched[s0 + p], c->stream));
                }
                ch->cur ^= 1;
                ch->last_launches++;
            }
        }
        ch->last_cfg = cfg;
    }
    if (ch->ev1) HIPCHK(hipEventRecord(ch->ev1, c->stream));
    ch->timed = ch->ev1 != nullptr;
    return RT_OK;
}

static int td_read(rt_chain* ch, int32_t* tiling) {
    HIPCHK(hipSetDevice(ch->ctx->device));
    HIPCHK(td_launch_merge((const uint2*)ch->cls[ch->cur], ch->N, ch->tsave, ch->ctx->stream));
    HIPCHK(hipStreamSynchronize(ch->ctx->stream));
    HIPCHK(hipMemcpy(tiling, ch->tsave, sizeof(int32_t) * (size_t)ch->N * ch->N, hipMemcpyDeviceToHost));
    return RT_OK;
}

// Measure every tile configuration (one and two steps per launch) on a scratch copy of the chain (its own state is
// not advanced); chains that fit one workgroup always run the single-workgroup kernel.
static int td_autotune(rt_chain* ch, int64_t probe_steps, int32_t* best) {
    rt_ctx* c = ch->ctx;
    if (ch->W <= kSmallMax) {
        if (best) *best = -1;
        return RT_OK;
    }
    rt_chain* t = new rt_chain();
    t->ctx = c;
    t->model = ch->model;
    t->N = ch->N;
    t->W = ch->W;
    t->P = ch->P;
    t->NC = ch->NC;
    t->esize = ch->esize;
    t->G = ch->G;
    const size_t cbytes = sizeof(uint2) * (size_t)ch->G;
    hipError_t e = td_alloc(t);
    if (e == hipSuccess) e = hipMalloc(&t->misc, 64);
    if (e == hipSuccess) e = hipEventCreate(&t->ev0);
    if (e == hipSuccess) e = hipEventCreate(&t->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&t->ev_sched, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMemsetAsync(t->misc, 0, 64, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(t->cls[0], ch->cls[ch->cur], cbytes, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) {
        rt_chain_destroy(t);
        return fail(RT_ERR_HIP, "autotune clone: %s", hipGetErrorString(e));
    }
    int rc = rt_chain_seed(t, RT_SEED_WALK, 12345);
    int chosen = -1;
    double best_ms = 1e30;
    for (int i = 0; i < td_config_count() && rc == RT_OK; i++) {
        t->path = RT_PATH_STEPWISE;
        t->cfg = i;
        rc = rt_chain_run(t, probe_steps / 4 + 1);
        if (rc == RT_OK) rc = rt_chain_run(t, probe_steps);
        double ms = 0;
        if (rc == RT_OK) rc = rt_chain_elapsed_ms(t, &ms);
        if (rc == RT_OK) rc = rt_chain_sync(t);
        if (rc == RT_OK && ms < best_ms) {
            best_ms = ms;
            chosen = i;
        }
    }
    rt_chain_destroy(t);
    if (rc != RT_OK) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[std::make_tuple(ch->model, ch->N, c->uniform)] = chosen;
    }
    ch->cfg = chosen;
    if (best) *best = chosen;
    return RT_OK;
}

namespace rt {
int ctx_device(rt_ctx* c) { return c->device; }
hipStream_t ctx_stream(rt_ctx* c) { return c->stream; }
int set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
}  // namespace rt

extern "C" {

const char* rt_last_error(void) { return g_err.c_str(); }

const char* rt_version(void) { return "randomtilings_amd 0.1 (gfx950)"; }

int rt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int rt_device_name(int device, char* buf, int32_t cap) {
    if (!buf || cap <= 0) return fail(RT_ERR_ARG, "rt_device_name: bad buffer");
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, device));
    snprintf(buf, (size_t)cap, "%s (%s)", p.name, p.gcnArchName);
    return RT_OK;
}

int rt_create(int device, rt_ctx** out) {
    if (!out) return fail(RT_ERR_ARG, "rt_create: null out");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(RT_ERR_ARG, "rt_create: device %d of %d", device, n);
    HIPCHK(hipSetDevice(device));
    noalloc_setup(device);  // the allocation-free entry point's scratch address on this device
    rt_ctx* c = new rt_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(RT_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = c;
    return RT_OK;
}

static void ctx_free(rt_ctx* c) {
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (hipEvent_t& e : c->marks)
        if (e) hipEventDestroy(e);
    if (c->params) hipFree(c->params);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

// A context outlives its chains: destroying it while chains are live defers the release to
// the destruction of the last one.
int rt_destroy(rt_ctx* c) {
    if (!c) return RT_OK;
    if (c->live_chains > 0) {
        c->closing = true;
        return RT_OK;
    }
    ctx_free(c);
    return RT_OK;
}

int rt_ctx_mark(rt_ctx* c, int32_t slot) {
    if (!c || slot < 0 || slot >= 16) return fail(RT_ERR_ARG, "rt_ctx_mark: bad slot");
    HIPCHK(hipSetDevice(c->device));
    if (!c->marks[slot]) HIPCHK(hipEventCreate(&c->marks[slot]));
    HIPCHK(hipEventRecord(c->marks[slot], c->stream));
    return RT_OK;
}

int rt_ctx_elapsed_ms(rt_ctx* c, int32_t a, int32_t b, double* ms) {
    if (!c || !ms || a < 0 || b < 0 || a >= 16 || b >= 16 || !c->marks[a] || !c->marks[b])
        return fail(RT_ERR_ARG, "rt_ctx_elapsed_ms: unrecorded slot");
    HIPCHK(hipEventSynchronize(c->marks[b]));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, c->marks[a], c->marks[b]));
    *ms = f;
    return RT_OK;
}

int rt_ctx_sync(rt_ctx* c) {
    if (!c) return fail(RT_ERR_ARG, "rt_ctx_sync: null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_ctx_set_path(rt_ctx* c, int path, int32_t cfg) {
    if (!c || path < RT_PATH_AUTO || path > RT_PATH_BLOCKED) return fail(RT_ERR_ARG, "rt_ctx_set_path: bad args");
    c->default_path = path;
    c->default_cfg = cfg;
    return RT_OK;
}

int rt_load_params(rt_ctx* c, const uint32_t* mat3, int64_t n) {
    if (!c || !mat3 || n <= 0) return fail(RT_ERR_ARG, "rt_load_params: bad args");
    HIPCHK(hipSetDevice(c->device));
    std::vector<uint4> h((size_t)n);
    bool uni = true;
    for (int64_t i = 0; i < n; i++) {
        h[i] = make_uint4(mat3[3 * i], mat3[3 * i + 1], mat3[3 * i + 2], 0u);
        if (h[i].x != h[0].x || h[i].y != h[0].y || h[i].z != h[0].z) uni = false;
    }
    if (c->params) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(c->params));
        c->params = nullptr;
        c->nparams = 0;
    }
    HIPCHK(hipMalloc(&c->params, sizeof(uint4) * (size_t)n));
    HIPCHK(hipMemcpy(c->params, h.data(), sizeof(uint4) * (size_t)n, hipMemcpyHostToDevice));
    c->nparams = n;
    c->uniform = uni;
    c->um1 = h[0].x;
    c->um2 = h[0].y;
    c->utm = h[0].z;
    return RT_OK;
}

// file_reader.cpp:51-70,114-166: entry i is the i-th line that does not start with '#'; its
// mat1, mat2 and tmat are the hexadecimal numbers (strtoul base 16) after the 3rd, 4th and 5th
// comma.  An empty line where an entry is expected (also the end of the file), a missing comma
// or a number strtoul reports out of range (errno) is a format error, as in the reference.
// Read sequentially (the reference reopens the file and seeks for every entry).
int rt_parse_params_file(const char* path, int64_t n, uint32_t* mat3) {
    if (!path || n <= 0 || !mat3) return fail(RT_ERR_ARG, "rt_parse_params_file: bad args");
    std::ifstream f(path);
    if (!f) return fail(RT_ERR_IO, "parameter file not found: %s", path);
    std::string line;
    for (int64_t got = 0; got < n; got++) {
        bool have = false;
        while ((have = (bool)std::getline(f, line)) && !line.empty() && line[0] == '#') {
        }
        if (!have || line.empty())
            return fail(RT_ERR_IO, "file format error at entry %lld of %s (%s)", (long long)got, path,
                        have ? "empty line" : "end of file");
        errno = 0;
        for (int k = 0; k < 3; k++) {
            int commas = 0;
            size_t pos = 0;
            for (; pos < line.size(); pos++)
                if (line[pos] == ',' && ++commas == 3 + k) break;
            if (pos >= line.size()) return fail(RT_ERR_IO, "comma not found in line: %s", line.c_str());
            mat3[3 * got + k] = (uint32_t)strtoul(line.c_str() + pos + 1, nullptr, 16);
        }
        if (errno != 0) return fail(RT_ERR_IO, "file format error: %s", line.c_str());
    }
    return RT_OK;
}

int rt_load_params_file(rt_ctx* c, const char* path, int64_t n) {
    if (!c || !path || n <= 0) return fail(RT_ERR_ARG, "rt_load_params_file: bad args");
    std::vector<uint32_t> m((size_t)n * 3);
    const int rc = rt_parse_params_file(path, n, m.data());
    if (rc != RT_OK) return rc;
    return rt_load_params(c, m.data(), n);
}

int64_t rt_params_count(rt_ctx* c) { return c ? c->nparams : 0; }

// Rule thresholds for inspection by tests: FlipRule (24 words) then RectTriRule (36).
int32_t rt_rule_thresholds(int model, int rule, const double* rparams, int32_t nrparams, uint32_t* out, int32_t cap) {
    RuleParams rp{};
    if (model == RT_DOMINO) {
        if (rule == RT_DOMINO_WEIGHTED && (!rparams || nrparams < 4)) return fail(RT_ERR_ARG, "p12[4] required");
        if (!domino_rule(rule, rparams, rp.fr)) return fail(RT_ERR_ARG, "bad domino rule");
    } else if (model == RT_LOZENGE) {
        const double q = (rparams && nrparams > 0) ? rparams[0] : 0.99;
        if (!lozenge_rule(rule, q, rp.fr)) return fail(RT_ERR_ARG, "bad lozenge rule");
    } else if (model == RT_RECTTRI) {
        float w[16];
        const float* wp = nullptr;
        if (rparams && nrparams >= 16) {
            for (int k = 0; k < 16; k++) w[k] = (float)rparams[k];
            wp = w;
        }
        recttri_rule(wp, rp.rt);
    } else {
        return fail(RT_ERR_ARG, "bad model");
    }
    const int32_t nw = (int32_t)((sizeof(FlipRule) + sizeof(RectTriRule)) / 4);
    if (out && cap >= nw) {
        std::memcpy(out, &rp.fr, sizeof(FlipRule));
        std::memcpy(out + sizeof(FlipRule) / 4, &rp.rt, sizeof(RectTriRule));
    }
    return nw;
}

// ---- chains --------------------------------------------------------------------------

static int mod3(int x) { return ((x % 3) + 3) % 3; }

static void split(const rt_chain* ch, const int32_t* t, std::vector<uint8_t>& out) {
    const int N = ch->N, W = ch->P;
    const size_t CS = (size_t)N * W;
    out.assign(CS * ch->NC * ch->esize, 0);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            int c, col;
            if (ch->model == RT_DOMINO || ch->model == kDominoPair) {
                c = (i + j) % 2;  // DominoTiler.cpp:20-21
                col = j / 2;
            } else if (ch->model == RT_LOZENGE) {
                c = mod3(j - i);  // LozengeTiler.cpp:31-37
                col = j / 3;
            } else {
                c = mod3(i - j);  // RectTriangleTiler.cpp:26-32
                col = j / 3;
            }
            const size_t at = c * CS + (size_t)i * W + col;
            if (ch->esize == 1) out[at] = (uint8_t)t[(size_t)i * N + j];
            else if (ch->esize == 2) {
                const uint16_t v = (uint16_t)t[(size_t)i * N + j];
                std::memcpy(&out[at * 2], &v, 2);
            } else std::memcpy(&out[at * 4], &t[(size_t)i * N + j], 4);
        }
}

static void merge(const rt_chain* ch, const std::vector<uint8_t>& in, int32_t* t) {
    const int N = ch->N, W = ch->P;
    const size_t CS = (size_t)N * W;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            int c, col;
            if (ch->model == RT_DOMINO || ch->model == kDominoPair) {
                c = (i + j) % 2;
                col = j / 2;
            } else if (ch->model == RT_LOZENGE) {
                c = mod3(j - i);
                col = j / 3;
            } else {
                c = mod3(i - j);
                col = j / 3;
            }
            const size_t at = c * CS + (size_t)i * W + col;
            if (ch->esize == 1) t[(size_t)i * N + j] = in[at];
            else if (ch->esize == 2) {
                uint16_t v;
                std::memcpy(&v, &in[at * 2], 2);
                t[(size_t)i * N + j] = v;
            } else std::memcpy(&t[(size_t)i * N + j], &in[at * 4], 4);
        }
}

int rt_chain_destroy(rt_chain* ch) {
    if (!ch) return RT_OK;
    hipSetDevice(ch->ctx->device);
    hipStreamSynchronize(ch->ctx->stream);
    for (int b = 0; b < 2; b++) {
        if (ch->cls[b]) hipFree(ch->cls[b]);
        if (ch->gens[b]) hipFree(ch->gens[b]);
    }
    free_runs(ch);
    if (ch->dsched) hipFree(ch->dsched);
    if (ch->tsave) hipFree(ch->tsave);
    if (ch->tbits) hipFree(ch->tbits);
    if (ch->misc) hipFree(ch->misc);
    if (ch->bar) hipFree(ch->bar);
    if (ch->ev0) hipEventDestroy(ch->ev0);
    if (ch->ev1) hipEventDestroy(ch->ev1);
    if (ch->ev_sched) hipEventDestroy(ch->ev_sched);
    rt_ctx* c = ch->ctx;
    const bool counted = ch->counted;
    delete ch;
    if (counted && --c->live_chains == 0 && c->closing) ctx_free(c);
    return RT_OK;
}

int rt_chain_create(rt_ctx* ctx, int model, int rule, const double* rparams, int32_t nrparams, const int32_t* tiling,
                    int32_t N, rt_chain** out) {
    if (!ctx || !tiling || !out) return fail(RT_ERR_ARG, "rt_chain_create: null argument");
    rt_chain* ch = new rt_chain();
    ch->ctx = ctx;
    ch->model = model;
    ch->rule = rule;
    ch->N = N;
    ch->path = ctx->default_path;
    ch->cfg = ctx->default_cfg;
    if (model == RT_DOMINO || model == kDominoPair) {
        // DominoTiler.h:34-37: 2M x 2M, twice zero-padded.  The reference only warns on odd N
        // (DominoTiler.cpp:12) and then indexes past its arrays; odd N is rejected here.
        if (N < 6 || N % 2) {
            delete ch;
            return fail(RT_ERR_ARG, "domino tiling side %d must be even and >= 6", N);
        }
        ch->NC = 2;
        ch->W = N / 2;
        ch->esize = model == kDominoPair ? 2 : 1;
        ch->G = (int64_t)(N - 2) * (ch->W - 2);
        if (rule == RT_DOMINO_WEIGHTED && (!rparams || nrparams < 4)) {
            delete ch;
            return fail(RT_ERR_ARG, "weighted domino rule needs p12[4]");
        }
        if (!domino_rule(rule, rparams, ch->rp.fr)) {
            delete ch;
            return fail(RT_ERR_ARG, "invalid domino rule %d or probabilities", rule);
        }
    } else if (model == RT_LOZENGE || model == RT_RECTTRI) {
        // LozengeTiler.h:26-28, RectTriangleTiler.h:26-27: N divisible by 3, triple padding
        if (N < 6 || N % 3) {
            delete ch;
            return fail(RT_ERR_ARG, "tiling side %d must be a multiple of 3 and >= 6", N);
        }
        ch->NC = 3;
        ch->W = N / 3;
        ch->G = (int64_t)N * ch->W;
        if (model == RT_LOZENGE) {
            ch->esize = 1;
            const double q = (rparams && nrparams > 0) ? rparams[0] : 0.99;
            if (!lozenge_rule(rule, q, ch->rp.fr)) {
                delete ch;
                return fail(RT_ERR_ARG, "invalid lozenge rule %d or q", rule);
            }
        } else {
            ch->esize = 4;
            float w[16];
            const float* wp = nullptr;
            if (rparams && nrparams >= 16) {
                for (int k = 0; k < 16; k++) w[k] = (float)rparams[k];
                wp = w;
            }
            recttri_rule(wp, ch->rp.rt);
        }
    } else if (model == RT_TRIDIMER) {
        const int rc = td_create(ch, tiling);
        if (rc != RT_OK) {
            rt_chain_destroy(ch);
            return rc;
        }
        ch->counted = true;
        ctx->live_chains++;
        *out = ch;
        return RT_OK;
    } else {
        delete ch;
        return fail(RT_ERR_ARG, "rt_chain_create: model %d has no planar chain", model);
    }
    if (model == RT_LOZENGE) {
        for (int i = 0; i < N * N; i++)
            if (tiling[i] < 0 || tiling[i] > 255) {
                delete ch;
                return fail(RT_ERR_ARG, "lozenge vertex state %d out of the 8-bit range", tiling[i]);
            }
    } else if (model == RT_DOMINO) {
        for (int i = 0; i < N * N; i++)
            if (tiling[i] < 0 || tiling[i] > 255) {
                delete ch;
                return fail(RT_ERR_ARG, "domino vertex state %d out of the 8-bit range", tiling[i]);
            }
    } else if (model == kDominoPair) {
        for (int i = 0; i < N * N; i++)
            if (tiling[i] < 0 || tiling[i] > 0xffff) {
                delete ch;
                return fail(RT_ERR_ARG, "coupled domino state %d out of range", tiling[i]);
            }
    }
    ch->P = (int)((((size_t)ch->W * ch->esize + 15) / 16) * 16 / ch->esize);
    hipError_t e = hipSetDevice(ctx->device);
    const size_t bytes = (size_t)N * ch->P * ch->NC * ch->esize;
    for (int b = 0; b < 2 && e == hipSuccess; b++) e = hipMalloc(&ch->cls[b], bytes);
    for (int b = 0; b < 2 && e == hipSuccess; b++) e = hipMalloc(&ch->gens[b], sizeof(uint4) * (size_t)ch->G);
    if (e == hipSuccess) e = hipMalloc(&ch->misc, 64);
    if (e == hipSuccess) e = hipMalloc(&ch->bar, kBarBytes);
    if (e == hipSuccess) e = hipEventCreate(&ch->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ch->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ch->ev_sched, hipEventDisableTiming);
    if (e != hipSuccess) {
        rt_chain_destroy(ch);
        return fail(RT_ERR_HIP, "chain allocation: %s", hipGetErrorString(e));
    }
    std::vector<uint8_t> h;
    split(ch, tiling, h);
    build_runs(ch, h);
    e = hipMemcpy(ch->cls[0], h.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ch->cls[1], h.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(ch->misc, 0, 64);
    if (e == hipSuccess) e = hipMemset(ch->bar, 0, kBarBytes);
    if (e != hipSuccess) {
        rt_chain_destroy(ch);
        return fail(RT_ERR_HIP, "chain upload: %s", hipGetErrorString(e));
    }
    ch->counted = true;
    ctx->live_chains++;
    *out = ch;
    return RT_OK;
}

int rt_chain_set_path(rt_chain* ch, int path, int32_t cfg) {
    if (!ch || path < RT_PATH_AUTO || path > RT_PATH_BLOCKED) return fail(RT_ERR_ARG, "rt_chain_set_path: bad args");
    const int ncfg = ch->model == RT_TRIDIMER ? td_config_count() : persist_config_count(ch->model);
    if (cfg >= ncfg) return fail(RT_ERR_ARG, "tile config %d out of range", cfg);
    ch->path = path;
    ch->cfg = cfg;
    return RT_OK;
}

int rt_chain_seed(rt_chain* ch, int mode, int64_t seed) {
    if (!ch) return fail(RT_ERR_ARG, "rt_chain_seed: null chain");
    rt_ctx* c = ch->ctx;
    if (!c->params) return fail(RT_ERR_PARAMS, "no TinyMT parameter table loaded (LoadTinyMT)");
    if (c->nparams < ch->G)
        return fail(RT_ERR_PARAMS, "TinyMT parameter table holds %lld sets; this chain needs %lld",
                    (long long)c->nparams, (long long)ch->G);
    uint32_t tmseed;
    const long lseed = (long)seed;
    if (ch->model == RT_DOMINO || ch->model == kDominoPair) {
        ch->mt.seed(lseed);
        // Walk: TinyMT seeded with the first mt19937 output (DominoTiler.cpp:31-32);
        // epoch: with the epoch seed itself (:81-82).  Both pass through a 32-bit kernel arg.
        tmseed = (mode == RT_SEED_EPOCH) ? (uint32_t)(uint64_t)seed : (uint32_t)ch->mt();
    } else {
        // LozengeTiler.cpp:47-52, RectTriangleTiler.cpp:41-45
        ch->lcg.seed(lseed);
        ch->d3.reset();
        ch->d2.reset();
        tmseed = (uint32_t)(uint64_t)seed;
    }
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(launch_init_gens(ch->G, tmseed, c->uniform ? nullptr : c->params, c->um1, c->um2, c->utm,
                            ch->gens[ch->gcur], c->stream));
    ch->seeded = true;
    return RT_OK;
}


// The parameter table covers the chain (what seeding and autotuning need).
static bool c_has_params(const rt_ctx* c, const rt_chain* ch) { return c->params && c->nparams >= ch->G; }

// Whether the packed sweep can run this chain's rule: every lozenge and rectangle-triangle
// rule (lozenge intervals are disjoint; the rectangle-triangle row-wave body reads the same
// rule tables as the generic one); for dominoes and coupled domino pairs the rules that split k at one threshold
// 0 < t < 2^23 per class and row parity (3 -> 12 below t, 12 -> 3 from t on: the CFTP and
// weighted heat-bath rules), and the toggle rule of dominokernel.cl (both 3 -> 12 and 12 -> 3
// below one threshold t, the same for every class and row parity).
static bool packed_rule_ok(const rt_chain* ch) {
    if (ch->model == RT_LOZENGE || ch->model == RT_RECTTRI) return true;
    if (ch->model != RT_DOMINO && ch->model != kDominoPair) return false;
    const FlipRule& f = ch->rp.fr;
    bool split = true, toggle = true;
    const uint32_t t0 = f.ab_hi[0][0];
    for (int c = 0; c < 2; c++)
        for (int p = 0; p < 2; p++) {
            const uint32_t t = f.ab_hi[c][p];
            if (f.ab_lo[c][p] != 0 || f.ba_lo[c][p] != t || f.ba_hi[c][p] != (1u << 23) || t == 0 || t >= (1u << 23))
                split = false;
            if (f.ab_lo[c][p] != 0 || f.ba_lo[c][p] != 0 || t != t0 || f.ba_hi[c][p] != t0 || t0 == 0 ||
                t0 >= (1u << 23))
                toggle = false;
        }
    return split || toggle;
}

// Tile configuration for the LDS-resident paths.  *resident is true when the whole tile grid
// fits on the device at once (one launch runs many blocks of h steps, tiles exchanging
// borders through flags); otherwise the chain runs "blocked": one launch per block of h
// steps, the launch boundary being the exchange.
// Resident grids are launched cooperatively unless RT_COOP_LAUNCH=0 is set in the environment
// (read once): then as plain launches, for tools that do not survive cooperative launches
// (rocprofv3 ends before writing its results).  Co-residency then rests on the one-workgroup-
// per-CU capacity alone, and every neighbour wait stays bounded.
static bool coop_launches() {
    static const bool on = [] {
        const char* v = std::getenv("RT_COOP_LAUNCH");
        return !(v && std::strcmp(v, "0") == 0);
    }();
    return on;
}

static int choose_persist_cfg(rt_chain* ch, int* nblocks, int* tiles_c, bool* resident, bool no_packed = false,
                              bool blocked = false) {
    const int n = persist_config_count(ch->model);
    const bool uni = ch->ctx->uniform;
    auto grid = [&](int idx, bool* res) {
        int tr, tc, h, nt;
        if (!persist_config(ch->model, idx, &tr, &tc, &h, &nt)) return false;
        if (persist_config_packed(ch->model, idx) && (no_packed || !packed_rule_ok(ch))) return false;
        const int trn = (ch->N + tr - 1) / tr, tcn = (ch->W + tc - 1) / tc;
        const int cap = persist_capacity(ch->model, idx, uni, ch->ctx->device);
        if (cap <= 0) return false;
        const rt_chain::Runs* r = skip_runs(ch, idx);
        const int running = r ? (int)r->list.size() : trn * tcn;  // workgroups: skipped tiles do not run
        *res = running <= cap && trn * tcn <= kMaxBlocks && !persist_uses_scratch(ch->model, idx, uni);
        *nblocks = running;
        *tiles_c = tcn;
        return true;
    };
    if (ch->cfg >= 0) return grid(ch->cfg, resident) ? ch->cfg : -1;
    if (!blocked) {  // tuned choices come from resident single-chain probes
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(std::make_tuple(ch->model, ch->N, uni));
        if (it != g_tuned.end() && grid(it->second, resident)) return it->second;
    }
    // Default.  A resident grid runs every tile at once, one per CU, so a step costs what one
    // tile costs: the configuration with the least work per tile (tile plus halo, the packed
    // sweep counted at half per position), the exchange every h steps added as 1/h, wins.  A grid
    // run blocked (too many tiles, or a batch) streams every tile through HBM per block: there
    // the fewest redundant halo positions per owned one win, plus 4/h for the launch, load and
    // store of every block (batches of 32 lozenge hexagons: halo 8 at 2.3e11 flip attempts/s,
    // every halo-4 tile at 1.3-1.5e11), divided by the fraction of 16 waves per CU the kernel
    // keeps in flight (fewer leave the streaming latency exposed: the rect-triangle 48 x 48
    // tiles, 2 waves per SIMD, ran batches at 0.65x the 32 x 32 ones).
    int best = -1;
    double best_cost = 1e30;
    for (int i = 0; i < n; i++) {
        bool res = false;
        if (!grid(i, &res)) continue;
        int tr, tc, h, nt;
        persist_config(ch->model, i, &tr, &tc, &h, &nt);
        const double ext = (double)(tr + 2 * h) * (tc + 2 * h);
        const double redund = ext / ((double)tr * tc);
        double cost;
        if (res && !blocked) {
            cost = ext * (1.0 + 1.0 / h) * (persist_config_packed(ch->model, i) ? 0.5 : 1.0);
        } else {
            const int waves = persist_waves_per_cu(ch->model, i, uni, blocked, true);
            const double occ = std::max(std::min(1.0, waves / 16.0), 1.0 / 64);
            cost = 1e9 + (redund + 4.0 / h) / occ;
        }
        if (cost < best_cost) {
            best_cost = cost;
            best = i;
        }
    }
    if (best >= 0) grid(best, resident);
    return best;
}

int rt_chain_autotune(rt_chain* ch, int64_t probe_steps, int32_t* best) {
    if (!ch || probe_steps <= 0) return fail(RT_ERR_ARG, "rt_chain_autotune: bad arguments");
    rt_ctx* c = ch->ctx;
    HIPCHK(hipSetDevice(c->device));
    if (!c->params || c->nparams < ch->G) return fail(RT_ERR_PARAMS, "autotune needs the parameter table loaded");
    if (ch->model == RT_TRIDIMER) return td_autotune(ch, probe_steps, best);
    // scratch clone: same shape and rule, the chain's current state, throwaway generators
    rt_chain* t = new rt_chain();
    t->ctx = c;
    t->model = ch->model;
    t->rule = ch->rule;
    t->N = ch->N;
    t->W = ch->W;
    t->P = ch->P;
    t->NC = ch->NC;
    t->esize = ch->esize;
    t->G = ch->G;
    t->rp = ch->rp;
    t->skip = ch->skip;
    t->consistent = ch->consistent;
    t->runs = ch->runs;
    for (auto& r : t->runs) r.d_list = nullptr, r.d_map = nullptr;  // the clone uploads its own
    const size_t bytes = (size_t)ch->N * ch->P * ch->NC * ch->esize;
    hipError_t e = hipSuccess;
    for (int b = 0; b < 2 && e == hipSuccess; b++) e = hipMalloc(&t->cls[b], bytes);
    for (int b = 0; b < 2 && e == hipSuccess; b++) e = hipMalloc(&t->gens[b], sizeof(uint4) * (size_t)t->G);
    if (e == hipSuccess) e = hipMalloc(&t->misc, 64);
    if (e == hipSuccess) e = hipMalloc(&t->bar, kBarBytes);
    if (e == hipSuccess) e = hipEventCreate(&t->ev0);
    if (e == hipSuccess) e = hipEventCreate(&t->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&t->ev_sched, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMemsetAsync(t->misc, 0, 64, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(t->bar, 0, kBarBytes, c->stream);
    for (int b = 0; b < 2 && e == hipSuccess; b++)
        e = hipMemcpyAsync(t->cls[b], ch->cls[ch->cur], bytes, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) {
        rt_chain_destroy(t);
        return fail(RT_ERR_HIP, "autotune clone: %s", hipGetErrorString(e));
    }
    int rc = rt_chain_seed(t, RT_SEED_WALK, 12345);
    int chosen = -1;
    double best_ms = 1e30;
    std::string last_fault;
    const int n = persist_config_count(ch->model);
    for (int i = 0; i < n && rc == RT_OK; i++) {
        t->path = RT_PATH_PERSISTENT;
        t->cfg = i;
        int nb = 0, tcn = 0;
        bool res = false;
        if (choose_persist_cfg(t, &nb, &tcn, &res) != i) continue;  // does not fit
        rc = rt_chain_run(t, probe_steps / 4 + 1);  // warm-up
        double ms = 1e30;
        for (int rep = 0; rep < 2 && rc == RT_OK; rep++) {  // the faster of two probes
            double m = 0;
            rc = rt_chain_run(t, probe_steps);
            if (rc == RT_OK) rc = rt_chain_elapsed_ms(t, &m);
            if (rc == RT_OK) rc = rt_chain_sync(t);
            if (rc == RT_OK && m < ms) ms = m;
        }
        if (rc == RT_ERR_DEVICE) {
            // This configuration raised the device error flag (e.g. a bounded wait expired):
            // it is not used.  The flag is sticky, so clear it, the exchange flags and the
            // block epoch, and restore the chain's state before probing the next one.
            last_fault = g_err;
            hipError_t er = hipStreamSynchronize(c->stream);
            if (er == hipSuccess) er = hipMemsetAsync(t->misc, 0, 64, c->stream);
            if (er == hipSuccess) er = hipMemsetAsync(t->bar, 0, kBarBytes, c->stream);
            for (int b = 0; b < 2 && er == hipSuccess; b++)
                er = hipMemcpyAsync(t->cls[b], ch->cls[ch->cur], bytes, hipMemcpyDeviceToDevice, c->stream);
            if (er == hipSuccess) er = hipStreamSynchronize(c->stream);
            t->epoch = 0;
            t->cur = 0;
            rc = er == hipSuccess ? rt_chain_seed(t, RT_SEED_WALK, 12345)
                                  : fail(RT_ERR_HIP, "autotune reset: %s", hipGetErrorString(er));
            continue;
        }
        if (rc == RT_OK && ms < best_ms) {
            best_ms = ms;
            chosen = i;
        }
    }
    rt_chain_destroy(t);
    if (rc != RT_OK) return rc;
    if (chosen < 0)
        return fail(last_fault.empty() ? RT_ERR_ARG : RT_ERR_DEVICE, "no persistent tile configuration %s%s",
                    last_fault.empty() ? "fits this chain" : "ran without a device error: ", last_fault.c_str());
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[std::make_tuple(ch->model, ch->N, c->uniform)] = chosen;
    }
    ch->path = RT_PATH_PERSISTENT;
    ch->cfg = chosen;
    if (best) *best = chosen;
    return RT_OK;
}

int rt_chain_run(rt_chain* ch, int64_t steps) {
    if (!ch) return fail(RT_ERR_ARG, "rt_chain_run: null chain");
    if (!ch->seeded) return fail(RT_ERR_ARG, "rt_chain_run: chain not seeded");
    if (steps < 0) return fail(RT_ERR_ARG, "rt_chain_run: negative steps");
    rt_ctx* c = ch->ctx;
    HIPCHK(hipSetDevice(c->device));
    if (ch->sched_pending) {  // the previous upload may still be reading hsched
        if (ch->ev_sched) HIPCHK(hipEventSynchronize(ch->ev_sched));
        else HIPCHK(hipStreamSynchronize(c->stream));
        ch->sched_pending = false;
    }
    ch->hsched.resize((size_t)std::max<int64_t>(steps, 1));
    if (ch->model == RT_TRIDIMER) {
        ch->timed = false;
        ch->last_launches = 0;
        return steps == 0 ? RT_OK : td_run(ch, steps);
    }
    draw_schedule(ch, steps, ch->hsched.data());
    ch->timed = false;
    ch->last_launches = 0;
    if (steps == 0) return RT_OK;

    int nblocks = 0, tiles_c = 0, cfg = -1;
    bool resident = false;
    if (ch->path != RT_PATH_STEPWISE) cfg = choose_persist_cfg(ch, &nblocks, &tiles_c, &resident);
    if (ch->path == RT_PATH_BLOCKED) resident = false;
    if ((ch->path == RT_PATH_PERSISTENT || ch->path == RT_PATH_BLOCKED) && cfg < 0)
        return fail(RT_ERR_ARG, "no persistent tile configuration fits this chain on the device");

    if (cfg < 0) {
        if (ch->ev0) HIPCHK(hipEventRecord(ch->ev0, c->stream));
        StepwiseArgs a{};
        a.rp = ch->rp;
        a.cls = ch->cls[ch->cur];
        a.N = ch->N;
        a.W = ch->W;
        a.P = ch->P;
        a.gens = ch->gens[ch->gcur];
        a.params = c->uniform ? nullptr : c->params;
        a.m1 = c->um1;
        a.m2 = c->um2;
        a.tm = c->utm;
        a.sched_host = ch->hsched.data();
        a.nsteps = steps;
        HIPCHK(launch_stepwise(ch->model, a, c->stream));
        ch->last_launches = 2 * steps;
        ch->last_cfg = -1;
    } else {
        if (ch->dsched_cap < steps) {
            if (ch->borrowed)
                return fail(RT_ERR_ARG, "rt_chain_run: %lld steps exceed the schedule buffer (%lld)", (long long)steps,
                            (long long)ch->dsched_cap);
            if (ch->dsched) {
                HIPCHK(hipStreamSynchronize(c->stream));
                HIPCHK(hipFree(ch->dsched));
                ch->dsched = nullptr;
            }
            // 16 bytes of slack: kernels fetch a block's colours as one aligned 8-byte word
            HIPCHK(hipMalloc(&ch->dsched, (size_t)steps + 16));
            ch->dsched_cap = steps;
        }
        HIPCHK(hipMemcpyAsync(ch->dsched, ch->hsched.data(), (size_t)steps, hipMemcpyHostToDevice, c->stream));
        if (ch->ev_sched) HIPCHK(hipEventRecord(ch->ev_sched, c->stream));
        ch->sched_pending = true;
        int tr, tc, h, nt;
        persist_config(ch->model, cfg, &tr, &tc, &h, &nt);
        rt_chain::Runs* runs = skip_runs(ch, cfg);
        if (runs) HIPCHK(upload_runs(*runs));
        ch->last_tiles = ((ch->N + tr - 1) / tr) * tiles_c;
        ch->last_running = nblocks;
        if (ch->ev0) HIPCHK(hipEventRecord(ch->ev0, c->stream));
        // one launch per chunk keeps single kernels short; chunk is a multiple of h.  A grid
        // that is not resident runs one block of h steps per launch.  Resident grids launch
        // cooperatively: the runtime then guarantees that every workgroup is resident at once
        // (the neighbour waits rely on it), or refuses the launch, in which case the chain
        // continues on the blocked path (the same kernel, one block per launch, no waits).
        int64_t chunk = resident ? (int64_t)h * 2048 : (int64_t)h;
        for (int64_t s0 = 0; s0 < steps; s0 += chunk) {
            const int64_t ns = std::min<int64_t>(chunk, steps - s0);
            PersistArgs a{};
            a.rp = ch->rp;
            a.cls[0] = ch->cls[0];
            a.cls[1] = ch->cls[1];
            a.gen_in = ch->gens[ch->gcur];
            a.gen_out = ch->gens[ch->gcur ^ 1];
            a.params = c->uniform ? nullptr : c->params;
            a.m1 = c->um1;
            a.m2 = c->um2;
            a.tm = c->utm;
            a.sched = ch->dsched + s0;
            a.nsteps = ns;
            a.N = ch->N;
            a.W = ch->W;
            a.P = ch->P;
            a.tiles_c = tiles_c;
            a.buf = ch->cur;
            a.flags = ch->bar;
            a.epoch0 = ch->epoch;
            a.err = (int*)(ch->misc + 1);
            a.spin_limit = 1u << 22;
            a.live = runs ? runs->d_list : nullptr;
            a.live_map = runs ? runs->d_map : nullptr;
            a.ntiles = ch->last_tiles;
            const hipError_t el =
                launch_persist(ch->model, cfg, c->uniform, a, nblocks, c->stream, resident && coop_launches());
            if (el != hipSuccess && resident) {
                // Refused before anything ran: the grid cannot be resident
                // (hipErrorCooperativeLaunchTooLarge), or the runtime cannot give this process
                // a cooperative queue now (hipErrorIllegalState was seen while another process
                // held one).  Redo this chunk, and the rest of the run, on the blocked path.
                (void)hipGetLastError();
                ch->coop_error = (int)el;
                resident = false;
                chunk = h;
                s0 -= chunk;
                ch->coop_refused++;
                continue;
            }
            HIPCHK(el);
            ch->gcur ^= 1;
            const int64_t nblk = (ns + h - 1) / h;  // each block of h steps publishes once
            ch->epoch += (uint32_t)(nblk - 1);
            if (nblk & 1) ch->cur ^= 1;
            ch->last_launches++;
        }
        ch->last_cfg = cfg;
    }
    if (ch->ev1) HIPCHK(hipEventRecord(ch->ev1, c->stream));
    ch->timed = ch->ev1 != nullptr;
    return RT_OK;
}

int rt_chain_sync(rt_chain* ch) {
    if (!ch) return fail(RT_ERR_ARG, "rt_chain_sync: null chain");
    HIPCHK(hipSetDevice(ch->ctx->device));
    HIPCHK(hipStreamSynchronize(ch->ctx->stream));
    uint32_t m[2];
    HIPCHK(hipMemcpy(m, ch->misc, sizeof(m), hipMemcpyDeviceToHost));
    if (m[1] == 2)
        return fail(RT_ERR_DEVICE, "a skipped tile's cells would have changed: the tiling is not closed under the "
                                   "chain (rt_chain_set_skip(ch, 0) runs every tile)");
    if (m[1] != 0) return fail(RT_ERR_DEVICE, "device error flag raised (grid barrier wait expired)");
    return RT_OK;
}

int rt_chain_set_skip(rt_chain* ch, int enable) {
    if (!ch) return fail(RT_ERR_ARG, "rt_chain_set_skip: null chain");
    ch->skip = enable != 0;
    return RT_OK;
}

int rt_chain_last_grid(rt_chain* ch, int32_t* tiles, int32_t* running) {
    if (!ch) return fail(RT_ERR_ARG, "rt_chain_last_grid: null chain");
    if (tiles) *tiles = ch->last_cfg >= 0 ? ch->last_tiles : 0;
    if (running) *running = ch->last_cfg >= 0 ? ch->last_running : 0;
    return RT_OK;
}

int rt_lozenge_skip_map(const int32_t* tiling, int32_t N, int32_t tr, int32_t tc, int32_t h, uint8_t* map,
                        int32_t* running, int32_t* consistent) {
    if (!tiling || N < 3 || tr <= 0 || tc <= 0 || h < 0) return fail(RT_ERR_ARG, "rt_lozenge_skip_map: bad arguments");
    for (int64_t i = 0; i < (int64_t)N * N; i++)
        if (tiling[i] < 0 || tiling[i] > 63) return fail(RT_ERR_ARG, "lozenge vertex code %d out of range", tiling[i]);
    rt_chain ch;
    ch.model = RT_LOZENGE;
    ch.N = N;
    ch.W = N / 3;
    ch.NC = 3;
    ch.esize = 1;
    ch.P = (int)(((size_t)ch.W + 15) / 16 * 16);
    std::vector<uint8_t> cls;
    split(&ch, tiling, cls);
    const bool ok = lozenge_consistent(cls.data(), N, ch.W, ch.P);
    std::vector<int32_t> ps;
    lozenge_nonzero_counts(cls.data(), N, ch.P, ps);
    std::vector<uint8_t> m;
    int run = lozenge_live_tiles(ps, N, ch.W, ch.P, tr, tc, h, m);
    if (!ok) {
        std::fill(m.begin(), m.end(), 1);
        run = (int)m.size();
    }
    if (map) std::memcpy(map, m.data(), m.size());
    if (running) *running = run;
    if (consistent) *consistent = ok;
    return RT_OK;
}

int rt_chain_elapsed_ms(rt_chain* ch, double* ms) {
    if (!ch || !ms) return fail(RT_ERR_ARG, "rt_chain_elapsed_ms: null argument");
    if (!ch->timed) {
        *ms = 0.0;
        return RT_OK;
    }
    HIPCHK(hipEventSynchronize(ch->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, ch->ev0, ch->ev1));
    *ms = f;
    return RT_OK;
}

int rt_chain_read(rt_chain* ch, int32_t* tiling) {
    if (!ch || !tiling) return fail(RT_ERR_ARG, "rt_chain_read: null argument");
    int rc = rt_chain_sync(ch);
    if (rc) return rc;
    if (ch->model == RT_TRIDIMER) return td_read(ch, tiling);
    std::vector<uint8_t> h((size_t)ch->N * ch->P * ch->NC * ch->esize);
    HIPCHK(hipMemcpy(h.data(), ch->cls[ch->cur], h.size(), hipMemcpyDeviceToHost));
    merge(ch, h, tiling);
    return RT_OK;
}

int64_t rt_chain_attempts_per_step(rt_chain* ch) { return ch ? ch->G : 0; }

int64_t rt_chain_state_bytes(rt_chain* ch) {
    if (ch && ch->model == RT_TRIDIMER) return ch->G * (int64_t)sizeof(uint2);
    return ch ? (int64_t)ch->N * ch->W * ch->NC * ch->esize : 0;
}

int32_t rt_persist_capacity(int model, int32_t idx, int device, int uniform_params) {
    if (model == RT_TRIDIMER || idx < 0 || idx >= persist_config_count(model)) return -1;
    if (hipSetDevice(device) != hipSuccess) return -1;
    const bool uni = uniform_params != 0;
    if (persist_uses_scratch(model, idx, uni)) return 0;
    return persist_capacity(model, idx, uni, device);
}

int32_t rt_tile_config_count(int model) {
    return model == RT_TRIDIMER ? td_config_count() : persist_config_count(model);
}

int rt_tile_config(int model, int32_t idx, int32_t* rows, int32_t* cols, int32_t* halo, int32_t* threads) {
    if (!rows || !cols || !halo || !threads) return fail(RT_ERR_ARG, "null argument");
    int tr, tc, h = kHalo, nt;
    const bool ok = model == RT_TRIDIMER ? td_config(idx, &tr, &tc, &nt) : persist_config(model, idx, &tr, &tc, &h, &nt);
    if (!ok) return fail(RT_ERR_ARG, "config %d out of range", idx);
    if (model == RT_TRIDIMER) h = kHalo * td_config_steps(idx);  // a halo of kHalo per step of a launch
    *rows = tr;
    *cols = tc;
    *halo = h;
    *threads = nt;
    return RT_OK;
}

int32_t rt_tile_config_packed(int model, int32_t idx) {
    return model != RT_TRIDIMER && persist_config_packed(model, idx) ? 1 : 0;
}

int64_t rt_chain_last_launches(rt_chain* ch) { return ch ? ch->last_launches : 0; }

int32_t rt_chain_last_config(rt_chain* ch) { return ch ? ch->last_cfg : -1; }

int rt_chain_coop_refusals(rt_chain* ch, int64_t* count, int32_t* last_error) {
    if (!ch) return fail(RT_ERR_ARG, "rt_chain_coop_refusals: null chain");
    if (count) *count = ch->coop_refused;
    if (last_error) *last_error = ch->coop_error;
    return RT_OK;
}

int32_t rt_chain_tile_configs(rt_chain* ch) { return ch ? rt_tile_config_count(ch->model) : 0; }

int rt_chain_tile_config_info(rt_chain* ch, int32_t idx, int32_t* rows, int32_t* cols, int32_t* halo,
                              int32_t* threads) {
    if (!ch) return fail(RT_ERR_ARG, "null argument");
    return rt_tile_config(ch->model, idx, rows, cols, halo, threads);
}

// ---- batches ---------------------------------------------------------------------------

}  // extern "C"

struct rt_batch {
    rt_ctx* ctx = nullptr;
    std::vector<rt_chain*> chains;
    uint8_t* dsched = nullptr;
    size_t dsched_cap = 0;
    BatchEntry* dent = nullptr;
    size_t dent_cap = 0;
    TdBatchEntry* dtd = nullptr;  // triangle-dimer chains
    size_t dtd_cap = 0;
    int64_t* ctr = nullptr;  // [0] step counter of the current group, [1] the kernels' error word
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_up = nullptr;
    bool up_pending = false;
    std::vector<uint8_t> hsched;
    std::vector<BatchEntry> hent;
    std::vector<rt_chain*> hent_chain;  // the chain of each entry
    int64_t launches = 0;
    int graphs = 0;          // graphs instantiated by the last run (0: the kept one was replayed)
    int64_t replays = 0;     // graph replays of the last run
    hipGraphExec_t exec = nullptr;  // one round of every group, kept across runs
    std::vector<int64_t> graph_key;  // what the kept graph was captured for
    int64_t graph_launches = 0;      // kernel launches in one replay
    int cfg_override[5] = {-1, -1, -1, -1, -1};  // per model: the configuration rt_batch_autotune chose
    bool timed = false;
};

extern "C" {

int rt_batch_create(rt_ctx* ctx, rt_batch** out) {
    if (!ctx || !out) return fail(RT_ERR_ARG, "rt_batch_create: null argument");
    rt_batch* b = new rt_batch();
    b->ctx = ctx;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipMalloc(&b->ctr, 64);
    if (e == hipSuccess) e = hipMemset(b->ctr, 0, 64);  // step counter and the kernels' error word
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_up, hipEventDisableTiming);
    if (e != hipSuccess) {
        rt_batch_destroy(b);
        return fail(RT_ERR_HIP, "rt_batch_create: %s", hipGetErrorString(e));
    }
    *out = b;
    return RT_OK;
}

int rt_batch_destroy(rt_batch* b) {
    if (!b) return RT_OK;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    if (b->dsched) hipFree(b->dsched);
    if (b->dent) hipFree(b->dent);
    if (b->dtd) hipFree(b->dtd);
    if (b->ctr) hipFree(b->ctr);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->ev_up) hipEventDestroy(b->ev_up);
    delete b;
    return RT_OK;
}

int rt_batch_add(rt_batch* b, rt_chain* ch) {
    if (!b || !ch) return fail(RT_ERR_ARG, "rt_batch_add: null argument");
    if (ch->ctx != b->ctx) return fail(RT_ERR_ARG, "rt_batch_add: the chain belongs to another context");
    if (std::find(b->chains.begin(), b->chains.end(), ch) != b->chains.end())
        return fail(RT_ERR_ARG, "rt_batch_add: chain already in the batch");
    b->chains.push_back(ch);
    return RT_OK;
}

// Steps per captured round of a batched run: a multiple of every planar block height (4, 8) and
// of the triangle-dimer steps per launch (1, 2), so every group's launches of one round cover
// exactly these steps and one step-counter update serves them all.
constexpr int64_t kRoundSteps = 32;
static_assert(kRoundSteps % kDrawSteps == 0, "a round holds whole triangle-dimer draw blocks");

int rt_batch_run(rt_batch* b, int64_t steps, int use_graph) {
    if (!b || steps < 0) return fail(RT_ERR_ARG, "rt_batch_run: bad arguments");
    rt_ctx* c = b->ctx;
    for (rt_chain* ch : b->chains)
        if (!ch->seeded) return fail(RT_ERR_ARG, "rt_batch_run: every chain must be seeded");
    HIPCHK(hipSetDevice(c->device));
    if (b->up_pending) {
        HIPCHK(hipEventSynchronize(b->ev_up));
        b->up_pending = false;
    }
    b->launches = 0;
    b->graphs = 0;
    b->replays = 0;
    b->timed = false;
    if (steps == 0 || b->chains.empty()) return RT_OK;
    // planar chains are grouped by model and run together; triangle-dimer chains run their
    // own launches
    const int models[] = {RT_DOMINO, RT_LOZENGE, RT_RECTTRI, kDominoPair};
    // per-chain schedule segments: 16-byte aligned with 16 bytes of slack, as a chain's own
    // schedule buffer (kernels fetch a block's colours as 8-byte words)
    const size_t seg = (((size_t)steps + 16) + 15) / 16 * 16;
    std::vector<rt_chain*> planar, tdc;
    for (rt_chain* ch : b->chains) (ch->model == RT_TRIDIMER ? tdc : planar).push_back(ch);
    b->hsched.assign(seg * (planar.size() + tdc.size()), 0);
    for (size_t k = 0; k < planar.size(); k++) draw_schedule(planar[k], steps, b->hsched.data() + k * seg);
    for (size_t k = 0; k < tdc.size(); k++) draw_td_moves(tdc[k], steps, b->hsched.data() + (planar.size() + k) * seg);
    if (b->dsched_cap < b->hsched.size()) {
        if (b->dsched) {
            HIPCHK(hipStreamSynchronize(c->stream));
            HIPCHK(hipFree(b->dsched));
            b->dsched = nullptr;
        }
        HIPCHK(hipMalloc(&b->dsched, b->hsched.size()));
        b->dsched_cap = b->hsched.size();
    }
    struct Group {
        int model, cfg, h;
        size_t first, count;
        int blocks;
    };
    std::vector<Group> groups;
    b->hent.clear();
    b->hent_chain.clear();
    for (int m : models) {
        Group g{m, -1, 0, b->hent.size(), 0, 0};
        rt_chain* largest = nullptr;
        for (rt_chain* ch : planar)
            if (ch->model == m && (!largest || ch->G > largest->G)) largest = ch;
        if (!largest) continue;
        int nb = 0, tcn = 0;
        bool res = false;
        // one configuration serves the whole group: packed only if every chain's rule allows it
        bool all_packed_ok = true;
        for (rt_chain* ch : planar)
            if (ch->model == m && !packed_rule_ok(ch)) all_packed_ok = false;
        g.cfg = choose_persist_cfg(largest, &nb, &tcn, &res, !all_packed_ok, true);
        if (b->cfg_override[m] >= 0) {
            g.cfg = b->cfg_override[m];
            if (persist_config_packed(m, g.cfg) && !all_packed_ok)
                return fail(RT_ERR_ARG, "rt_batch_run: configuration %d of model %d is packed and a chain's rule "
                            "cannot use it", g.cfg, m);
        }
        if (g.cfg < 0) return fail(RT_ERR_ARG, "rt_batch_run: no tile configuration for model %d", m);
        int tr, tc, h, nt;
        persist_config(m, g.cfg, &tr, &tc, &h, &nt);
        g.h = h;
        for (size_t k = 0; k < planar.size(); k++) {
            rt_chain* ch = planar[k];
            if (ch->model != m) continue;
            BatchEntry e{};
            e.rp = ch->rp;
            e.cls[0] = ch->cls[0];
            e.cls[1] = ch->cls[1];
            e.gens[0] = ch->gens[0];
            e.gens[1] = ch->gens[1];
            e.sched = b->dsched + k * seg;
            e.N = ch->N;
            e.W = ch->W;
            e.P = ch->P;
            e.tiles_c = (ch->W + tc - 1) / tc;
            e.buf0 = ch->cur;
            e.g0 = ch->gcur;
            e.tile0 = g.blocks;
            e.ntiles = ((ch->N + tr - 1) / tr) * e.tiles_c;
            g.blocks += e.ntiles;
            b->hent.push_back(e);
            b->hent_chain.push_back(ch);
            g.count++;
        }
        groups.push_back(g);
    }
    if (b->dent_cap < b->hent.size()) {
        if (b->dent) {
            HIPCHK(hipStreamSynchronize(c->stream));
            HIPCHK(hipFree(b->dent));
            b->dent = nullptr;
        }
        HIPCHK(hipMalloc(&b->dent, sizeof(BatchEntry) * std::max<size_t>(b->hent.size(), 1)));
        b->dent_cap = b->hent.size();
    }
    // triangle-dimer entries: one tiled launch per step covers every chain
    // configuration: the largest chain's own if set, else default 0 (32 x 32 tiles, one step
    // per launch; tuned single-chain choices are not used, large tiles being slower in batches)
    int td_cfg = 0, td_blocks = 0, td_dblocks = 0;
    std::vector<TdBatchEntry> htd;
    if (!tdc.empty()) {
        rt_chain* largest = tdc[0];
        for (rt_chain* ch : tdc)
            if (ch->G > largest->G) largest = ch;
        if (largest->cfg >= 0) td_cfg = largest->cfg;
        if (b->cfg_override[RT_TRIDIMER] >= 0) td_cfg = b->cfg_override[RT_TRIDIMER];
        int tr, tc, nt;
        td_config(td_cfg, &tr, &tc, &nt);
        for (size_t k = 0; k < tdc.size(); k++) {
            rt_chain* ch = tdc[k];
            TdBatchEntry e{};
            e.cells[0] = (uint2*)ch->cls[0];
            e.cells[1] = (uint2*)ch->cls[1];
            e.gens[0] = ch->gens[0];
            e.gens[1] = ch->gens[1];
            e.rbits = ch->tbits;
            e.sched = b->dsched + (planar.size() + k) * seg;
            e.M = ch->W;
            e.buf0 = ch->cur;
            e.g0 = ch->gcur;
            e.tile0 = td_blocks;
            e.ntiles = ((ch->W + tr - 1) / tr) * ((ch->W + tc - 1) / tc);
            td_blocks += e.ntiles;
            e.dblk0 = td_dblocks;
            e.ndblk = td_draw_blocks(ch->W);
            td_dblocks += e.ndblk;
            htd.push_back(e);
        }
        if (b->dtd_cap < htd.size()) {
            if (b->dtd) {
                HIPCHK(hipStreamSynchronize(c->stream));
                HIPCHK(hipFree(b->dtd));
                b->dtd = nullptr;
            }
            HIPCHK(hipMalloc(&b->dtd, sizeof(TdBatchEntry) * htd.size()));
            b->dtd_cap = htd.size();
        }
    }
    HIPCHK(hipMemcpyAsync(b->dsched, b->hsched.data(), b->hsched.size(), hipMemcpyHostToDevice, c->stream));
    if (!b->hent.empty())
        HIPCHK(hipMemcpyAsync(b->dent, b->hent.data(), sizeof(BatchEntry) * b->hent.size(), hipMemcpyHostToDevice,
                              c->stream));
    if (!htd.empty())
        HIPCHK(hipMemcpyAsync(b->dtd, htd.data(), sizeof(TdBatchEntry) * htd.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(b->ev_up, c->stream));
    b->up_pending = true;
    HIPCHK(hipEventRecord(b->ev0, c->stream));
    // Every group advances in rounds of kRoundSteps steps (each block height h and each
    // triangle-dimer launch length divides it): one round of every group plus one update of the
    // shared step counter is captured once as a HIP graph and replayed steps / kRoundSteps
    // times; the graph is kept and replayed by later runs of the same batch as long as its
    // chains, tile configurations, device tables and parameters are unchanged.  The steps left
    // over run as direct launches at their offsets from the counter.
    auto planar_args = [&](const Group& g) {
        PersistArgs a{};
        a.params = c->uniform ? nullptr : c->params;
        a.m1 = c->um1;
        a.m2 = c->um2;
        a.tm = c->utm;
        a.batch = b->dent + g.first;
        a.nbatch = (int)g.count;
        a.block_h = g.h;
        a.step_ctr = b->ctr;
        a.err = (int*)(b->ctr + 1);
        a.spin_limit = 1;
        a.nsteps = g.h;
        return a;
    };
    TdArgs ta{};
    ta.params = c->uniform ? nullptr : c->params;
    ta.m1 = c->um1;
    ta.m2 = c->um2;
    ta.tm = c->utm;
    const int td_base = tdc.empty() ? 0 : td_config_base(td_cfg), td_spl = tdc.empty() ? 1 : td_config_steps(td_cfg);
    // launches of one group covering `n` steps from offset `off` (planar: blocks of h and a
    // partial last block; triangle dimers: per kDrawSteps steps (off and n multiples of it but
    // for the run's last block) one draw launch, then pairs of steps when the configuration
    // runs two, then single steps)
    auto run_span = [&](int64_t off, int64_t n, int64_t* nl) -> hipError_t {
        for (const Group& g : groups) {
            PersistArgs a = planar_args(g);
            for (int64_t j = 0; j * g.h < n; j++) {
                a.step_off = off + j * g.h;
                a.nsteps = std::min<int64_t>(g.h, n - j * g.h);
                const hipError_t e = launch_persist(g.model, g.cfg, c->uniform, a, g.blocks, c->stream);
                if (e != hipSuccess) return e;
                ++*nl;
            }
        }
        for (int64_t d0 = 0; !htd.empty() && d0 < n; d0 += kDrawSteps) {
            const int64_t k = std::min<int64_t>(kDrawSteps, n - d0);
            hipError_t e = td_launch_draw_batch(ta, c->uniform, b->dtd, (int)htd.size(), td_dblocks, b->ctr, off + d0,
                                                (int)k, c->stream);
            if (e != hipSuccess) return e;
            ++*nl;
            const int64_t pairs = td_spl == 2 ? k / 2 : 0;
            for (int64_t j = 0; j < k - pairs; j++) {  // pairs first, then single steps
                const bool two = j < pairs;
                const int64_t o = two ? 2 * j : 2 * pairs + (j - pairs);
                e = td_launch_step_batch(td_base, ta, b->dtd, (int)htd.size(), td_blocks, b->ctr, off + d0 + o, two,
                                         td_spl, c->stream);
                if (e != hipSuccess) return e;
                ++*nl;
            }
        }
        return hipSuccess;
    };
    // the step counter restarts; the error word (b->ctr[1]) keeps what any run raised until
    // rt_batch_sync reports it
    HIPCHK(launch_fill_u32((uint32_t*)b->ctr, 2, 0u, c->stream));
    bool round_ok = true;  // every block height divides the round
    for (const Group& g : groups) round_ok = round_ok && kRoundSteps % g.h == 0;
    const int64_t replays = use_graph && round_ok ? steps / kRoundSteps : 0;
    if (replays > 0) {
        std::vector<int64_t> key{(int64_t)(intptr_t)b->dent, (int64_t)(intptr_t)b->dtd, (int64_t)c->uniform,
                                 (int64_t)(intptr_t)c->params, (int64_t)c->um1, (int64_t)c->um2, (int64_t)c->utm,
                                 (int64_t)td_cfg, (int64_t)td_blocks, (int64_t)td_dblocks, (int64_t)htd.size()};
        for (const Group& g : groups)
            for (int64_t v : {(int64_t)g.model, (int64_t)g.cfg, (int64_t)g.first, (int64_t)g.count, (int64_t)g.blocks})
                key.push_back(v);
        if (!b->exec || key != b->graph_key) {
            if (b->exec) HIPCHK(hipGraphExecDestroy(b->exec));
            b->exec = nullptr;
            hipGraph_t graph = nullptr;
            int64_t nl = 0;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t e1 = run_span(0, kRoundSteps, &nl);
            const hipError_t e2 = launch_add_i64(b->ctr, kRoundSteps, c->stream);
            const hipError_t e3 = hipStreamEndCapture(c->stream, &graph);
            if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
                if (graph) hipGraphDestroy(graph);
                return fail(RT_ERR_HIP, "rt_batch_run: graph capture failed");
            }
            const hipError_t e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            if (e != hipSuccess) {
                b->exec = nullptr;
                return fail(RT_ERR_HIP, "rt_batch_run: graph instantiation: %s", hipGetErrorString(e));
            }
            b->graph_key = key;
            b->graph_launches = nl;
            b->graphs++;
        }
        for (int64_t r = 0; r < replays; r++) HIPCHK(hipGraphLaunch(b->exec, c->stream));
        b->launches += replays * b->graph_launches;
        b->replays = replays;
    }
    {
        int64_t nl = 0;
        HIPCHK(run_span(0, steps - replays * kRoundSteps, &nl));
        b->launches += nl;
    }
    // copy flips: a planar chain flips its copies once per block, a triangle-dimer chain once
    // per launch
    for (const Group& g : groups) {
        const int64_t rounds = (steps + g.h - 1) / g.h;
        for (size_t k = g.first; k < g.first + g.count; k++) {
            rt_chain* ch = b->hent_chain[k];
            if (rounds & 1) {
                ch->cur ^= 1;
                ch->gcur ^= 1;
            }
            ch->last_launches = rounds;
            ch->last_cfg = g.cfg;
            ch->timed = false;
        }
    }
    if (!htd.empty()) {
        // cells flip once per step launch, generators once per draw launch (every kDrawSteps
        // steps, the run's last block possibly shorter); pairs never straddle a draw block
        int64_t launches = 0;
        for (int64_t d0 = 0; d0 < steps; d0 += kDrawSteps) {
            const int64_t k = std::min<int64_t>(kDrawSteps, steps - d0);
            launches += k - (td_spl == 2 ? k / 2 : 0);
        }
        const int64_t draws = (steps + kDrawSteps - 1) / kDrawSteps;
        for (rt_chain* ch : tdc) {
            if (launches & 1) ch->cur ^= 1;
            if (draws & 1) ch->gcur ^= 1;
            ch->last_launches = launches + draws;
            ch->last_cfg = td_cfg;
            ch->timed = false;
        }
    }
    HIPCHK(hipEventRecord(b->ev1, c->stream));
    b->timed = true;
    return RT_OK;
}

int rt_batch_sync(rt_batch* b) {
    if (!b) return fail(RT_ERR_ARG, "rt_batch_sync: null batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    int64_t err = 0;
    HIPCHK(hipMemcpy(&err, b->ctr + 1, sizeof(err), hipMemcpyDeviceToHost));
    if (err != 0) {
        HIPCHK(hipMemset(b->ctr + 1, 0, sizeof(err)));
        return fail(RT_ERR_DEVICE, "rt_batch_sync: a batched launch raised the device error flag (%d)", (int)err);
    }
    return RT_OK;
}

// A chain's state for the batch autotuner to put back after a probe: its current copies of the
// state and generators, and the host schedule RNGs.
struct ChainSnap {
    void* cls = nullptr;
    void* gens = nullptr;
    size_t cbytes = 0, gbytes = 0;
    int cur = 0, gcur = 0;
    uint32_t epoch = 0;
    std::mt19937 mt;
    std::minstd_rand0 lcg;
    std::uniform_int_distribution<int> d3, d2;
};

static size_t state_bytes(const rt_chain* ch) {
    return ch->model == RT_TRIDIMER ? sizeof(uint2) * (size_t)ch->G : (size_t)ch->N * ch->P * ch->NC * ch->esize;
}

int rt_batch_autotune(rt_batch* b, int64_t probe_steps, int32_t* chosen) {
    if (!b || probe_steps <= 0) return fail(RT_ERR_ARG, "rt_batch_autotune: bad arguments");
    rt_ctx* c = b->ctx;
    HIPCHK(hipSetDevice(c->device));
    for (rt_chain* ch : b->chains)
        if (!ch->seeded) return fail(RT_ERR_ARG, "rt_batch_autotune: every chain must be seeded");
    std::vector<ChainSnap> snap(b->chains.size());
    int rc = RT_OK;
    hipError_t e = hipStreamSynchronize(c->stream);
    for (size_t k = 0; k < b->chains.size() && e == hipSuccess; k++) {
        rt_chain* ch = b->chains[k];
        ChainSnap& sn = snap[k];
        sn.cbytes = state_bytes(ch);
        sn.gbytes = sizeof(uint4) * (size_t)ch->G;
        e = hipMalloc(&sn.cls, sn.cbytes);
        if (e == hipSuccess) e = hipMalloc(&sn.gens, sn.gbytes);
        if (e == hipSuccess) e = hipMemcpyAsync(sn.cls, ch->cls[ch->cur], sn.cbytes, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(sn.gens, ch->gens[ch->gcur], sn.gbytes, hipMemcpyDeviceToDevice, c->stream);
        sn.cur = ch->cur;
        sn.gcur = ch->gcur;
        sn.epoch = ch->epoch;
        sn.mt = ch->mt;
        sn.lcg = ch->lcg;
        sn.d3 = ch->d3;
        sn.d2 = ch->d2;
    }
    auto restore = [&](const std::vector<size_t>& which) -> hipError_t {
        hipError_t er = hipSuccess;
        for (size_t k : which) {
            rt_chain* ch = b->chains[k];
            const ChainSnap& sn = snap[k];
            ch->cur = sn.cur;
            ch->gcur = sn.gcur;
            ch->epoch = sn.epoch;
            ch->mt = sn.mt;
            ch->lcg = sn.lcg;
            ch->d3 = sn.d3;
            ch->d2 = sn.d2;
            if (er == hipSuccess)
                er = hipMemcpyAsync(ch->cls[ch->cur], sn.cls, sn.cbytes, hipMemcpyDeviceToDevice, c->stream);
            if (er == hipSuccess)
                er = hipMemcpyAsync(ch->gens[ch->gcur], sn.gens, sn.gbytes, hipMemcpyDeviceToDevice, c->stream);
        }
        return er;
    };
    if (e != hipSuccess) rc = fail(RT_ERR_HIP, "rt_batch_autotune: snapshot: %s", hipGetErrorString(e));
    const int models[] = {RT_DOMINO, RT_LOZENGE, RT_RECTTRI, RT_TRIDIMER};
    for (int m : models) {
        if (rc != RT_OK) break;
        std::vector<size_t> which;
        for (size_t k = 0; k < b->chains.size(); k++)
            if (b->chains[k]->model == m) which.push_back(k);
        if (which.empty()) continue;
        bool all_packed_ok = true;
        for (size_t k : which) all_packed_ok = all_packed_ok && packed_rule_ok(b->chains[k]);
        const int ncfg = m == RT_TRIDIMER ? td_config_count() : persist_config_count(m);
        int best = -1;
        double best_ms = 1e30;
        for (int i = 0; i < ncfg && rc == RT_OK; i++) {
            if (m != RT_TRIDIMER) {
                int tr, tc, h, nt;
                if (!persist_config(m, i, &tr, &tc, &h, &nt) || persist_waves_per_cu(m, i, c->uniform, true, true) <= 0) continue;
                if (persist_config_packed(m, i) && !all_packed_ok) continue;
            }
            // the group alone, on configuration i, in a batch of its own
            rt_batch* sub = nullptr;
            rc = rt_batch_create(c, &sub);
            if (rc != RT_OK) break;
            for (size_t k : which) rt_batch_add(sub, b->chains[k]);
            sub->cfg_override[m] = i;
            double ms = 1e30;
            rc = rt_batch_run(sub, probe_steps / 4 + 1, 1);  // warm-up (captures the graph)
            if (rc == RT_OK) rc = rt_batch_run(sub, probe_steps, 1);
            if (rc == RT_OK) rc = rt_batch_elapsed_ms(sub, &ms);
            if (rc == RT_OK) rc = rt_batch_sync(sub);
            rt_batch_destroy(sub);
            const hipError_t er = restore(which);
            if (rc == RT_OK && er != hipSuccess) rc = fail(RT_ERR_HIP, "rt_batch_autotune: restore: %s", hipGetErrorString(er));
            if (rc == RT_OK && ms < best_ms) {
                best_ms = ms;
                best = i;
            }
        }
        if (rc == RT_OK) b->cfg_override[m] = best;
    }
    (void)hipStreamSynchronize(c->stream);
    for (ChainSnap& sn : snap) {
        if (sn.cls) hipFree(sn.cls);
        if (sn.gens) hipFree(sn.gens);
    }
    if (rc != RT_OK) return rc;
    if (chosen)
        for (int m = 0; m < 4; m++) chosen[m] = b->cfg_override[m];
    return RT_OK;
}

int rt_batch_set_config(rt_batch* b, int model, int32_t idx) {
    if (!b) return fail(RT_ERR_ARG, "rt_batch_set_config: null batch");
    if (model != RT_DOMINO && model != RT_LOZENGE && model != RT_RECTTRI && model != RT_TRIDIMER)
        return fail(RT_ERR_ARG, "rt_batch_set_config: unknown model %d", model);
    const int n = model == RT_TRIDIMER ? td_config_count() : persist_config_count(model);
    if (idx < -1 || idx >= n)
        return fail(RT_ERR_ARG, "rt_batch_set_config: configuration %d of model %d (%d compiled)", idx, model, n);
    b->cfg_override[model] = idx;
    return RT_OK;
}

int rt_batch_elapsed_ms(rt_batch* b, double* ms) {
    if (!b || !ms) return fail(RT_ERR_ARG, "rt_batch_elapsed_ms: null argument");
    *ms = 0.0;
    if (!b->timed) return RT_OK;
    HIPCHK(hipEventSynchronize(b->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, b->ev0, b->ev1));
    *ms = f;
    return RT_OK;
}

int64_t rt_batch_last_launches(rt_batch* b) { return b ? b->launches : 0; }
int32_t rt_batch_last_graphs(rt_batch* b) { return b ? b->graphs : 0; }

// ---- Walk entry points ----------------------------------------------------------------

static int walk(rt_ctx* ctx, int model, int rule, const double* rp, int nrp, int32_t* t, int32_t N,
                const int64_t* steps, const int64_t* seeds, int K, int mode) {
    rt_chain* ch = nullptr;
    int rc = rt_chain_create(ctx, model, rule, rp, nrp, t, N, &ch);
    if (rc) return rc;
    for (int k = 0; k < K && rc == RT_OK; k++) {
        rc = rt_chain_seed(ch, mode, seeds[k]);
        if (rc == RT_OK) rc = rt_chain_run(ch, steps[k]);
    }
    if (rc == RT_OK) rc = rt_chain_read(ch, t);
    rt_chain_destroy(ch);
    return rc;
}

int rt_domino_walk(rt_ctx* ctx, int rule, const double* rparams, int32_t* tiling, int32_t N, int64_t steps,
                   int64_t seed) {
    return walk(ctx, RT_DOMINO, rule, rparams, rparams ? 4 : 0, tiling, N, &steps, &seed, 1, RT_SEED_WALK);
}

int rt_domino_cftp(rt_ctx* ctx, int rule, const double* rparams, const int32_t* tmax, const int32_t* tmin,
                   int32_t N, int64_t seed, int mode, int64_t first_steps, int32_t max_restarts, int32_t* out,
                   int64_t* total_steps, int32_t* restarts) {
    return rt_domino_cftp_ex(ctx, rule, rparams, tmax, tmin, N, seed, mode, first_steps, max_restarts, out,
                             total_steps, restarts, nullptr, nullptr);
}

int rt_domino_cftp_ex(rt_ctx* ctx, int rule, const double* rparams, const int32_t* tmax, const int32_t* tmin,
                      int32_t N, int64_t seed, int mode, int64_t first_steps, int32_t max_restarts, int32_t* out,
                      int64_t* total_steps, int32_t* restarts, rt_cftp_progress progress, void* user) {
    if (!ctx || !tmax || !tmin || !out) return fail(RT_ERR_ARG, "rt_domino_cftp: null argument");
    if (rule != RT_DOMINO_CFTP && rule != RT_DOMINO_WEIGHTED)
        return fail(RT_ERR_ARG, "coupling from the past needs a monotone rule (RT_DOMINO_CFTP or RT_DOMINO_WEIGHTED)");
    if (mode != RT_CFTP_REFERENCE && mode != RT_CFTP_DOUBLING) return fail(RT_ERR_ARG, "unknown CFTP schedule %d", mode);
    if (first_steps <= 0 || max_restarts < 0) return fail(RT_ERR_ARG, "rt_domino_cftp: bad step bounds");
    if (N < 6 || N % 2) return fail(RT_ERR_ARG, "domino tiling side %d must be even and >= 6", N);
    const size_t NN = (size_t)N * N;
    std::vector<int32_t> pair(NN);
    for (size_t i = 0; i < NN; i++) {
        if (tmax[i] < 0 || tmax[i] > 255 || tmin[i] < 0 || tmin[i] > 255)
            return fail(RT_ERR_ARG, "domino vertex state out of the 8-bit range");
        pair[i] = tmax[i] | (tmin[i] << 8);
    }
    rt_chain* ch = nullptr;
    int rc = rt_chain_create(ctx, kDominoPair, rule, rparams, rparams ? 4 : 0, pair.data(), N, &ch);
    if (rc) return rc;
    // Coalescence takes many launches of the same shape: unless the caller fixed the path,
    // time the tile configurations once per (side, parameter mode) on a scratch copy (the
    // sample does not depend on the configuration) and keep the fastest.
    if (ch->path == RT_PATH_AUTO && ch->cfg < 0 && c_has_params(ctx, ch)) {
        bool tuned;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            tuned = g_tuned.count(std::make_tuple((int)kDominoPair, N, ctx->uniform)) != 0;
        }
        if (!tuned) {
            rc = rt_chain_autotune(ch, 1024, nullptr);
            if (rc) {
                rt_chain_destroy(ch);
                return rc;
            }
            ch->path = RT_PATH_AUTO;  // the tuned choice is found again through the tuning table
            ch->cfg = -1;
        }
    }
    const size_t bytes = (size_t)N * ch->P * ch->NC * ch->esize;
    void* init = nullptr;
    uint32_t* count = nullptr;
    hipError_t e = hipMalloc(&init, bytes);
    if (e == hipSuccess) e = hipMalloc(&count, 64);
    if (e == hipSuccess) e = hipMemcpy(init, ch->cls[ch->cur], bytes, hipMemcpyDeviceToDevice);
    if (e != hipSuccess) {
        if (init) hipFree(init);
        if (count) hipFree(count);
        rt_chain_destroy(ch);
        return fail(RT_ERR_HIP, "rt_domino_cftp setup: %s", hipGetErrorString(e));
    }
    // epochs, oldest first (AztecDiamondCftpEx.cpp:47-48, :67)
    std::mt19937 mt((std::mt19937::result_type)seed);
    std::vector<std::pair<int64_t, int64_t>> epochs{{first_steps, (int64_t)mt()}};
    rc = RT_ERR_NOT_COALESCED;
    int r = 0;
    for (; r <= max_restarts; r++) {
        int st = RT_OK;
        if (hipMemcpyAsync(ch->cls[ch->cur], init, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
            st = fail(RT_ERR_HIP, "rt_domino_cftp: reset failed");
        int64_t total = 0;
        for (size_t k = 0; k < epochs.size() && st == RT_OK; k++) {
            st = rt_chain_seed(ch, RT_SEED_EPOCH, epochs[k].second);
            if (st == RT_OK) st = rt_chain_run(ch, epochs[k].first);
            total += epochs[k].first;
            if (progress && st == RT_OK && k + 1 < epochs.size()) {  // a replay's epochs can take minutes
                st = rt_chain_sync(ch);
                if (st == RT_OK) progress(user, r, total, -1);
            }
        }
        uint32_t differ = 1;
        if (st == RT_OK && launch_pair_differs((const uint16_t*)ch->cls[ch->cur], (int64_t)(bytes / 2), count,
                                               ctx->stream) != hipSuccess)
            st = fail(RT_ERR_HIP, "rt_domino_cftp: coupling test launch failed");
        if (st == RT_OK) st = rt_chain_sync(ch);
        if (st == RT_OK && hipMemcpy(&differ, count, sizeof(differ), hipMemcpyDeviceToHost) != hipSuccess)
            st = fail(RT_ERR_HIP, "rt_domino_cftp: coupling test read failed");
        if (st != RT_OK) {
            rc = st;
            break;
        }
        if (progress) progress(user, r, total, differ == 0 ? 1 : 0);
        if (differ == 0) {
            std::vector<int32_t> both(NN);
            rc = rt_chain_read(ch, both.data());
            if (rc == RT_OK) {
                for (size_t i = 0; i < NN; i++) out[i] = both[i] & 0xff;
                if (total_steps) *total_steps = total;
                if (restarts) *restarts = r;
            }
            break;
        }
        const int64_t len = (mode == RT_CFTP_REFERENCE) ? 2 * epochs.back().first : total;
        epochs.insert(epochs.begin(), {len, (int64_t)mt()});
    }
    if (rc == RT_ERR_NOT_COALESCED) fail(rc, "no coalescence after %d restarts", max_restarts);
    hipFree(init);
    hipFree(count);
    rt_chain_destroy(ch);
    return rc;
}

// ---- observables on the device -----------------------------------------------------------

struct rt_observer {
    rt_ctx* ctx = nullptr;
    int N = 0;
    uint8_t* faces = nullptr;  // M x M squares of the domain
    int32_t* hmin = nullptr;   // N x N, INT32_MAX outside
    uint8_t* mask = nullptr;   // N x N vertices the statistics cover
    int32_t* ref = nullptr;    // N x N reference field x 256, or null
    const uint8_t** dchains = nullptr;
    unsigned long long* dout = nullptr;
    int32_t* dheights = nullptr;
    int cap = 0, hcap = 0;
};

extern "C" {

int rt_observer_create(rt_ctx* ctx, const int32_t* domain, int32_t M, const int32_t* hmin, const uint8_t* mask,
                       const int32_t* ref256, rt_observer** out) {
    if (!ctx || !domain || !hmin || !out || M < 2) return fail(RT_ERR_ARG, "rt_observer_create: bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    const int N = M + 1;
    const size_t NN = (size_t)N * N, MM = (size_t)M * M;
    std::vector<uint8_t> f(MM), m(NN);
    for (size_t k = 0; k < MM; k++) f[k] = domain[k] != 0;
    for (size_t k = 0; k < NN; k++) m[k] = mask ? (mask[k] != 0) : (hmin[k] != INT32_MAX);
    rt_observer* o = new rt_observer();
    o->ctx = ctx;
    o->N = N;
    hipError_t e = hipMalloc(&o->faces, MM);
    if (e == hipSuccess) e = hipMalloc(&o->hmin, NN * 4);
    if (e == hipSuccess) e = hipMalloc(&o->mask, NN);
    if (e == hipSuccess && ref256) e = hipMalloc(&o->ref, NN * 4);
    if (e == hipSuccess) e = hipMemcpy(o->faces, f.data(), MM, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(o->hmin, hmin, NN * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(o->mask, m.data(), NN, hipMemcpyHostToDevice);
    if (e == hipSuccess && ref256) e = hipMemcpy(o->ref, ref256, NN * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        rt_observer_destroy(o);
        return fail(RT_ERR_HIP, "rt_observer_create: %s", hipGetErrorString(e));
    }
    *out = o;
    return RT_OK;
}

int rt_observer_destroy(rt_observer* o) {
    if (!o) return RT_OK;
    hipSetDevice(o->ctx->device);
    hipStreamSynchronize(o->ctx->stream);
    for (void* p : {(void*)o->faces, (void*)o->hmin, (void*)o->mask, (void*)o->ref, (void*)o->dchains,
                    (void*)o->dout, (void*)o->dheights})
        if (p) hipFree(p);
    delete o;
    return RT_OK;
}

int rt_observe(rt_observer* o, rt_chain* const* chains, int32_t n, int64_t* out, int32_t* heights) {
    if (!o || !chains || n <= 0 || !out) return fail(RT_ERR_ARG, "rt_observe: bad arguments");
    rt_ctx* c = o->ctx;
    std::vector<const uint8_t*> ptrs((size_t)n);
    int P = -1;
    for (int k = 0; k < n; k++) {
        const rt_chain* ch = chains[k];
        if (!ch || ch->model != RT_DOMINO || ch->N != o->N || ch->ctx != c)
            return fail(RT_ERR_ARG, "rt_observe: chain %d is not a domino chain of side %d on this context", k, o->N);
        if (P >= 0 && ch->P != P) return fail(RT_ERR_ARG, "rt_observe: chains of different pitch");
        P = ch->P;
        ptrs[k] = (const uint8_t*)ch->cls[ch->cur];
    }
    HIPCHK(hipSetDevice(c->device));
    if (o->cap < n) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (o->dchains) HIPCHK(hipFree(o->dchains));
        if (o->dout) HIPCHK(hipFree(o->dout));
        o->dchains = nullptr;
        o->dout = nullptr;
        HIPCHK(hipMalloc(&o->dchains, sizeof(void*) * n));
        HIPCHK(hipMalloc(&o->dout, sizeof(unsigned long long) * 4 * n));
        o->cap = n;
    }
    const size_t NN = (size_t)o->N * o->N;
    if (heights && o->hcap < n) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (o->dheights) HIPCHK(hipFree(o->dheights));
        o->dheights = nullptr;
        HIPCHK(hipMalloc(&o->dheights, NN * 4 * n));
        o->hcap = n;
    }
    HIPCHK(hipMemcpyAsync(o->dchains, ptrs.data(), sizeof(void*) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_fill_u32((uint32_t*)o->dout, 8 * n, 0u, c->stream));
    HIPCHK(launch_domino_observe(o->dchains, n, o->N, P, o->faces, o->hmin, o->mask, o->ref,
                                 heights ? o->dheights : nullptr, o->dout, c->stream));
    HIPCHK(hipMemcpyAsync(out, o->dout, sizeof(int64_t) * 4 * n, hipMemcpyDeviceToHost, c->stream));
    if (heights) HIPCHK(hipMemcpyAsync(heights, o->dheights, NN * 4 * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return RT_OK;
}

}  // extern "C"

// ---- height functions on the device -----------------------------------------------------

static int domain_of(int model, const int32_t* domain, int32_t M, std::vector<int>& d) {
    if (model != RT_DOMINO && model != RT_LOZENGE)
        return fail(RT_ERR_ARG, "height functions are defined for RT_DOMINO and RT_LOZENGE domains");
    if (!domain || M < 2) return fail(RT_ERR_ARG, "bad domain");
    const size_t n = (size_t)M * (model == RT_LOZENGE ? 2 * M : M);
    d.assign(domain, domain + n);
    return RT_OK;
}

int rt_height_closure(rt_ctx* ctx, int model, const int32_t* domain, int32_t M, int dir, int32_t* heights,
                      const uint8_t* fixed, int32_t* launches) {
    if (!ctx || !heights || (dir != 1 && dir != -1)) return fail(RT_ERR_ARG, "rt_height_closure: bad arguments");
    std::vector<int> d;
    int rc = domain_of(model, domain, M, d);
    if (rc) return rc;
    const size_t n = (size_t)(M + 1) * (M + 1);
    std::vector<int32_t> h(heights, heights + n);
    std::vector<uint8_t> fx(n, 0);
    if (fixed)
        for (size_t k = 0; k < n; k++) fx[k] = fixed[k] != 0;
    std::string err;
    int nl = 0;
    rc = height_closure_device(ctx->device, ctx->stream, model, d, dir, h, fx, nullptr, &nl, err);
    if (launches) *launches = nl;
    if (rc == RT_ERR_UNTILEABLE) return fail(rc, "the closure moves a fixed vertex: no admissible height function");
    if (rc) return fail(rc, "rt_height_closure: %s", err.c_str());
    std::memcpy(heights, h.data(), n * sizeof(int32_t));
    return RT_OK;
}

int rt_extremal_tiling(rt_ctx* ctx, int model, const int32_t* domain, int32_t M, int dir, int32_t* tiling,
                       int32_t* heights) {
    if (!ctx || !tiling || (dir != 1 && dir != -1)) return fail(RT_ERR_ARG, "rt_extremal_tiling: bad arguments");
    std::vector<int> d;
    int rc = domain_of(model, domain, M, d);
    if (rc) return rc;
    const size_t n = (size_t)(M + 1) * (M + 1);
    std::memset(tiling, 0, n * sizeof(int32_t));
    std::vector<int32_t> h;
    std::vector<uint8_t> fx;
    if (!boundary_heights_host(model, d, h, fx))
        return fail(RT_ERR_UNTILEABLE, "the region is not tileable (its boundary does not close)");
    std::vector<int32_t> t;
    std::string err;
    int nl = 0;
    rc = height_closure_device(ctx->device, ctx->stream, model, d, dir, h, fx, &t, &nl, err);
    if (rc == RT_ERR_UNTILEABLE) return fail(rc, "the region is not tileable (the closure moves a boundary height)");
    if (rc) return fail(rc, "rt_extremal_tiling: %s", err.c_str());
    std::memcpy(tiling, t.data(), n * sizeof(int32_t));
    if (heights) std::memcpy(heights, h.data(), n * sizeof(int32_t));
    return RT_OK;
}

int rt_domino_walk_epochs(rt_ctx* ctx, int rule, const double* rparams, int32_t* tiling, int32_t N,
                          const int64_t* steps, const int64_t* seeds, int32_t K) {
    if (K < 0 || (K > 0 && (!steps || !seeds))) return fail(RT_ERR_ARG, "rt_domino_walk_epochs: bad epochs");
    return walk(ctx, RT_DOMINO, rule, rparams, rparams ? 4 : 0, tiling, N, steps, seeds, K, RT_SEED_EPOCH);
}

int rt_lozenge_walk(rt_ctx* ctx, int rule, double q, int32_t* tiling, int32_t N, int64_t steps, int64_t seed) {
    return walk(ctx, RT_LOZENGE, rule, &q, 1, tiling, N, &steps, &seed, 1, RT_SEED_WALK);
}

// rt_lozenge_walk without allocating: the chain's state copies, generators, exchange flags,
// error word and colour schedule are carved out of this device's file-scope scratch array
// (rt_noalloc.hip), zeroed where they must start at zero by a fill kernel, and the chain runs
// on the same kernels as rt_lozenge_walk.  Tile skipping is off (it would upload per-chain
// tile lists); the result does not depend on it.  Calls on one device share the array:
// overlapping calls race.
int rt_lozenge_walk_noalloc(rt_ctx* c, int rule, double q, int32_t* tiling, int32_t N, int64_t steps,
                            int64_t seed) {
    if (!c || !tiling) return fail(RT_ERR_ARG, "rt_lozenge_walk_noalloc: null argument");
    if (steps < 0) return fail(RT_ERR_ARG, "rt_lozenge_walk_noalloc: negative steps");
    if (N < 6 || N % 3) return fail(RT_ERR_ARG, "tiling side %d must be a multiple of 3 and >= 6", N);
    for (int64_t i = 0; i < (int64_t)N * N; i++)
        if (tiling[i] < 0 || tiling[i] > 255)
            return fail(RT_ERR_ARG, "lozenge vertex state %d out of the 8-bit range", tiling[i]);
    std::string why;
    uint8_t* base = (uint8_t*)noalloc_scratch(c->device, why);
    if (!base) return fail(RT_ERR_HIP, "rt_lozenge_walk_noalloc: no scratch array: %s", why.c_str());
    rt_chain ch;
    ch.ctx = c;
    ch.model = RT_LOZENGE;
    ch.rule = rule;
    ch.N = N;
    ch.NC = 3;
    ch.W = N / 3;
    ch.esize = 1;
    ch.G = (int64_t)N * ch.W;
    ch.P = (ch.W + 15) / 16 * 16;
    ch.path = c->default_path;
    ch.cfg = c->default_cfg;
    ch.skip = false;
    ch.borrowed = true;
    if (!lozenge_rule(rule, q, ch.rp.fr)) return fail(RT_ERR_ARG, "invalid lozenge rule %d or q", rule);
    // layout: error word, exchange flags, schedule, two state copies, two generator copies
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    constexpr int64_t kSched = (int64_t)1 << 20;  // schedule bytes per run; longer walks run in pieces
    const size_t cls_bytes = (size_t)N * ch.P * 3, gen_bytes = sizeof(uint4) * (size_t)ch.G;
    size_t off = 0;
    const size_t o_misc = off;
    off += up(64);
    const size_t o_bar = off;
    off += up(kBarBytes);
    const size_t o_sched = off;
    off += up((size_t)kSched + 16);
    size_t o_cls[2], o_gen[2];
    for (int b = 0; b < 2; b++) o_cls[b] = off, off += up(cls_bytes);
    for (int b = 0; b < 2; b++) o_gen[b] = off, off += up(gen_bytes);
    if (off > kNoAllocScratchBytes)
        return fail(RT_ERR_ARG,
                    "rt_lozenge_walk_noalloc: side %d needs %zu bytes of device scratch, the array holds %zu "
                    "(rt_lozenge_walk has no limit)",
                    N, off, kNoAllocScratchBytes);
    ch.misc = (uint32_t*)(base + o_misc);
    ch.bar = (uint32_t*)(base + o_bar);
    ch.dsched = base + o_sched;
    ch.dsched_cap = kSched;
    for (int b = 0; b < 2; b++) {
        ch.cls[b] = base + o_cls[b];
        ch.gens[b] = (uint4*)(base + o_gen[b]);
    }
    HIPCHK(hipSetDevice(c->device));
    std::vector<uint8_t> h;
    split(&ch, tiling, h);
    for (int b = 0; b < 2; b++) HIPCHK(hipMemcpyAsync(ch.cls[b], h.data(), cls_bytes, hipMemcpyHostToDevice, c->stream));
    // the previous call's flags, epochs and error word are still in the array
    HIPCHK(launch_fill_u32(ch.misc, 16, 0u, c->stream));
    HIPCHK(launch_fill_u32(ch.bar, (int)(kBarBytes / sizeof(uint32_t)), 0u, c->stream));
    int rc = rt_chain_seed(&ch, RT_SEED_WALK, seed);
    for (int64_t s0 = 0; s0 < steps && rc == RT_OK; s0 += kSched) rc = rt_chain_run(&ch, std::min(kSched, steps - s0));
    if (rc == RT_OK) rc = rt_chain_read(&ch, tiling);  // synchronizes; reports the device error word
    if (rc != RT_OK) (void)hipStreamSynchronize(c->stream);  // leave nothing queued on the array
    return rc;
}

int rt_tridimer_walk(rt_ctx* ctx, int32_t* tiling, int32_t N, int64_t steps, int64_t seed) {
    return walk(ctx, RT_TRIDIMER, 0, nullptr, 0, tiling, N, &steps, &seed, 1, RT_SEED_WALK);
}

int rt_recttri_walk(rt_ctx* ctx, const float* weights16, int32_t* tiling, int32_t N, int64_t steps, int64_t seed) {
    double w[16];
    if (weights16)
        for (int k = 0; k < 16; k++) w[k] = weights16[k];
    return walk(ctx, RT_RECTTRI, 0, weights16 ? w : nullptr, weights16 ? 16 : 0, tiling, N, &steps, &seed, 1,
                RT_SEED_WALK);
}

}  // extern "C"
