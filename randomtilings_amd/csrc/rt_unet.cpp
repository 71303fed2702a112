// rt_unet.cpp — the warm-start network: a U-Net trained and evaluated on the GPU with this
// library's own kernels (bf16 matrix-core products for the convolutions, rt_mfma.hip; the
// element-wise kernels of rt_nnops.hip).  No counterpart in the reference: it predicts the
// limit-shape height of a domain from its boundary and weights, and the chain then starts
// from the prediction projected onto admissible height functions (rt_height_closure).
//
// Architecture (L levels, base width C, in_ch input channels, one output channel):
//   encoder level l (width c_l = C 2^l): e_l = relu(conv(relu(conv(x_l)))), x_0 = input,
//     x_{l+1} = avgpool2(e_l);
//   decoder level l = L-2 .. 0: s_l = relu(conv(up2(d_{l+1}))) + e_l, d_l = relu(conv(s_l)),
//     d_{L-1} = e_{L-1};
//   output y = conv(d_0) (linear).
// Every conv is 3 x 3 with zero padding: im2col to bf16 [pixels x 9 cin], then
// Y = col W (W: [9 cin x cout], bf16 copy of fp32 master weights), bias added in fp32.
// Backward: dW = col^T dY (split-K over pixels), dcol = dY W^T, dX = col2im(dcol).
// Parameters are one flat fp32 vector in layer order, each layer W then b; Adam updates
// it and refreshes the bf16 copy.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <algorithm>
#include <string>
#include <vector>

#include "rt_internal.hpp"
#include "rt_mfma.hpp"

struct rt_unet;

namespace {

struct Conv {
    int cin, cout;
    int64_t w, b;  // offsets of W (9 cin x cout) and b (cout) in the parameter vector
};

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    float* f() const { return (float*)p; }
    uint16_t* h() const { return (uint16_t*)p; }
};

}  // namespace

struct rt_unet {
    rt_ctx* ctx = nullptr;
    int device = 0;
    hipStream_t st = nullptr;
    int L = 0, C = 0, cin = 0;
    std::vector<Conv> enc_a, enc_b, dec_up, dec_a;  // enc: L each, dec: L-1 each (index = level)
    Conv out{};
    int64_t nparams = 0;
    float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t* wbf = nullptr;
    int64_t adam_t = 0;
    // activations of the current shape
    int B = 0, H = 0, W = 0;
    std::vector<Buf> pool;  // every device buffer of the workspace
    float* x0 = nullptr;
    std::vector<float*> xin, t1, e, u, v, s, d;  // per level
    float* y = nullptr;
    std::vector<float*> dt1, de, du, dv, ds, dd, dxin;
    float *dy = nullptr, *target = nullptr, *mask = nullptr, *acc = nullptr;
    uint16_t* col = nullptr;
    float* dcol = nullptr;
    uint16_t* dybf = nullptr;
    uint16_t* gws = nullptr;  // transposed-operand workspace of the matrix products
    int64_t gws_elems = 0;
};

namespace {

int width(const rt_unet* n, int l) { return n->C << l; }

hipError_t alloc(rt_unet* n, size_t bytes, void** out) {
    Buf b;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return e;
    b.bytes = bytes;
    n->pool.push_back(b);
    *out = b.p;
    return hipSuccess;
}

void free_ws(rt_unet* n) {
    for (Buf& b : n->pool) (void)hipFree(b.p);
    n->pool.clear();
    n->B = n->H = n->W = 0;
}

// Workspace for batch B of H x W inputs (H, W divisible by 2^(L-1)).
hipError_t ensure_ws(rt_unet* n, int B, int H, int W) {
    if (n->B == B && n->H == H && n->W == W) return hipSuccess;
    free_ws(n);
    const int L = n->L;
    auto px = [&](int l) { return (size_t)B * (H >> l) * (W >> l); };
    auto F = [&](size_t count, float** out) { return alloc(n, count * sizeof(float), (void**)out); };
    hipError_t e = F(px(0) * n->cin, &n->x0);
    n->xin.assign(L, nullptr);
    n->t1.assign(L, nullptr);
    n->e.assign(L, nullptr);
    n->dt1.assign(L, nullptr);
    n->de.assign(L, nullptr);
    n->dxin.assign(L, nullptr);
    n->u.assign(L, nullptr);
    n->v.assign(L, nullptr);
    n->s.assign(L, nullptr);
    n->d.assign(L, nullptr);
    n->du.assign(L, nullptr);
    n->dv.assign(L, nullptr);
    n->ds.assign(L, nullptr);
    n->dd.assign(L, nullptr);
    size_t colmax = 0, dymax = 0;
    int64_t gmax = 0;
    auto conv_ws = [&](const Conv& c, int l) {
        colmax = std::max(colmax, px(l) * 9 * c.cin);
        dymax = std::max(dymax, px(l) * c.cout);
        const int P = (int)px(l);
        gmax = std::max({gmax, rt::nn_gemm_workspace(P, c.cout, 9 * c.cin), rt::nn_gemm_workspace(9 * c.cin, c.cout, P),
                         rt::nn_gemm_workspace(P, 9 * c.cin, c.cout)});
    };
    for (int l = 0; l < L && e == hipSuccess; l++) {
        const int cl = width(n, l), cprev = l ? width(n, l - 1) : n->cin;
        if (l > 0) e = F(px(l) * cprev, &n->xin[l]);
        if (e == hipSuccess && l > 0) e = F(px(l) * cprev, &n->dxin[l]);
        if (e == hipSuccess) e = F(px(l) * cl, &n->t1[l]);
        if (e == hipSuccess) e = F(px(l) * cl, &n->e[l]);
        if (e == hipSuccess) e = F(px(l) * cl, &n->dt1[l]);
        if (e == hipSuccess) e = F(px(l) * cl, &n->de[l]);
        if (e == hipSuccess) e = F(px(l) * cl, &n->dd[l]);
        conv_ws(n->enc_a[l], l);
        conv_ws(n->enc_b[l], l);
        if (l < L - 1) {
            if (e == hipSuccess) e = F(px(l) * width(n, l + 1), &n->u[l]);
            if (e == hipSuccess) e = F(px(l) * width(n, l + 1), &n->du[l]);
            if (e == hipSuccess) e = F(px(l) * cl, &n->v[l]);
            if (e == hipSuccess) e = F(px(l) * cl, &n->dv[l]);
            if (e == hipSuccess) e = F(px(l) * cl, &n->s[l]);
            if (e == hipSuccess) e = F(px(l) * cl, &n->ds[l]);
            if (e == hipSuccess) e = F(px(l) * cl, &n->d[l]);
            conv_ws(n->dec_up[l], l);
            conv_ws(n->dec_a[l], l);
        }
    }
    n->xin[0] = n->x0;
    n->d[L - 1] = n->e[L - 1];
    conv_ws(n->out, 0);
    if (e == hipSuccess) e = F(px(0), &n->y);
    if (e == hipSuccess) e = F(px(0), &n->dy);
    if (e == hipSuccess) e = F(px(0), &n->target);
    if (e == hipSuccess) e = F(px(0), &n->mask);
    if (e == hipSuccess) e = F(64, &n->acc);
    if (e == hipSuccess) e = alloc(n, colmax * 2, (void**)&n->col);
    if (e == hipSuccess) e = F(colmax, &n->dcol);
    if (e == hipSuccess) e = alloc(n, dymax * 2, (void**)&n->dybf);
    if (e == hipSuccess) e = alloc(n, (size_t)gmax * 2, (void**)&n->gws);
    n->gws_elems = e == hipSuccess ? gmax : 0;
    if (e == hipSuccess) {
        n->B = B;
        n->H = H;
        n->W = W;
    } else {
        free_ws(n);
    }
    return e;
}

#define NNCHK(x)                                  \
    do {                                          \
        hipError_t e_ = (x);                      \
        if (e_ != hipSuccess) return e_;          \
    } while (0)

// y = conv(x) + b (relu optional), x: [B, h, w, cin], y: [B, h, w, cout]
hipError_t conv_fwd(rt_unet* n, const Conv& c, const float* x, float* y, int h, int w, bool relu) {
    const int64_t P = (int64_t)n->B * h * w;
    NNCHK(rt::nn_im2col(x, n->col, n->B, h, w, c.cin, n->st));
    rt::GemmArgs g{};
    g.A = n->col;
    g.B = n->wbf + c.w;
    g.C = y;
    g.M = (int)P;
    g.N = c.cout;
    g.K = 9 * c.cin;
    g.sam = 9 * c.cin;
    g.sak = 1;
    g.sbk = c.cout;
    g.sbn = 1;
    g.ldc = c.cout;
    g.alpha = 1.f;
    g.beta = 0.f;
    g.ws = n->gws;
    g.ws_elems = n->gws_elems;
    NNCHK(rt::nn_gemm(g, n->st));
    return rt::nn_bias_relu(y, n->w + c.b, P, c.cout, relu, n->st);
}

// Given dy (already through the activation's derivative) and the layer input x: dW, db into
// the gradient vector, dx (may be null) = the input gradient.
hipError_t conv_bwd(rt_unet* n, const Conv& c, const float* x, const float* dy, float* dx, int h, int w) {
    const int64_t P = (int64_t)n->B * h * w;
    NNCHK(rt::nn_im2col(x, n->col, n->B, h, w, c.cin, n->st));
    NNCHK(rt::nn_to_bf16(dy, n->dybf, P * c.cout, n->st));
    rt::GemmArgs g{};
    // dW [9 cin x cout] = col^T [9 cin x P] dY [P x cout]
    g.A = n->col;
    g.B = n->dybf;
    g.C = n->g + c.w;
    g.M = 9 * c.cin;
    g.N = c.cout;
    g.K = (int)P;
    g.sam = 1;
    g.sak = 9 * c.cin;
    g.sbk = c.cout;
    g.sbn = 1;
    g.ldc = c.cout;
    g.alpha = 1.f;
    g.beta = 0.f;
    g.ws = n->gws;
    g.ws_elems = n->gws_elems;
    NNCHK(rt::nn_gemm(g, n->st));
    NNCHK(rt::nn_bias_grad(dy, n->g + c.b, P, c.cout, n->st));
    if (!dx) return hipSuccess;
    // dcol [P x 9 cin] = dY [P x cout] W^T [cout x 9 cin]
    g.A = n->dybf;
    g.B = n->wbf + c.w;
    g.C = n->dcol;
    g.M = (int)P;
    g.N = 9 * c.cin;
    g.K = c.cout;
    g.sam = c.cout;
    g.sak = 1;
    g.sbk = 1;
    g.sbn = c.cout;
    g.ldc = 9 * c.cin;
    NNCHK(rt::nn_gemm(g, n->st));
    return rt::nn_col2im(n->dcol, dx, n->B, h, w, c.cin, n->st);
}

hipError_t forward(rt_unet* n) {
    const int L = n->L, H = n->H, W = n->W, B = n->B;
    for (int l = 0; l < L; l++) {
        const int h = H >> l, w = W >> l;
        if (l > 0) NNCHK(rt::nn_pool2(n->e[l - 1], n->xin[l], B, H >> (l - 1), W >> (l - 1), width(n, l - 1), n->st));
        NNCHK(conv_fwd(n, n->enc_a[l], n->xin[l], n->t1[l], h, w, true));
        NNCHK(conv_fwd(n, n->enc_b[l], n->t1[l], n->e[l], h, w, true));
    }
    for (int l = L - 2; l >= 0; l--) {
        const int h = H >> l, w = W >> l;
        NNCHK(rt::nn_up2(n->d[l + 1], n->u[l], B, h / 2, w / 2, width(n, l + 1), false, n->st));
        NNCHK(conv_fwd(n, n->dec_up[l], n->u[l], n->v[l], h, w, true));
        NNCHK(hipMemcpyAsync(n->s[l], n->v[l], sizeof(float) * (size_t)B * h * w * width(n, l), hipMemcpyDeviceToDevice,
                             n->st));
        NNCHK(rt::nn_add(n->s[l], n->e[l], (int64_t)B * h * w * width(n, l), n->st));
        NNCHK(conv_fwd(n, n->dec_a[l], n->s[l], n->d[l], h, w, true));
    }
    return conv_fwd(n, n->out, n->d[0], n->y, H, W, false);
}

// Backward from n->dy; gradients into n->g.
hipError_t backward(rt_unet* n) {
    const int L = n->L, H = n->H, W = n->W, B = n->B;
    auto count = [&](int l, int c) { return (int64_t)B * (H >> l) * (W >> l) * c; };
    for (int l = 0; l < L; l++) NNCHK(hipMemsetAsync(n->de[l], 0, sizeof(float) * count(l, width(n, l)), n->st));
    NNCHK(conv_bwd(n, n->out, n->d[0], n->dy, n->dd[0], H, W));
    for (int l = 0; l < L - 1; l++) {
        const int h = H >> l, w = W >> l, cl = width(n, l);
        // d_l = relu(conv_a(s_l))
        NNCHK(rt::nn_relu_back(n->dd[l], n->d[l], count(l, cl), n->st));
        NNCHK(conv_bwd(n, n->dec_a[l], n->s[l], n->dd[l], n->ds[l], h, w));
        // s_l = v_l + e_l
        NNCHK(rt::nn_add(n->de[l], n->ds[l], count(l, cl), n->st));
        NNCHK(hipMemcpyAsync(n->dv[l], n->ds[l], sizeof(float) * count(l, cl), hipMemcpyDeviceToDevice, n->st));
        NNCHK(rt::nn_relu_back(n->dv[l], n->v[l], count(l, cl), n->st));
        NNCHK(conv_bwd(n, n->dec_up[l], n->u[l], n->dv[l], n->du[l], h, w));
        // u_l = up2(d_{l+1}); d_{L-1} is e_{L-1}
        float* target = (l + 1 == L - 1) ? n->de[L - 1] : n->dd[l + 1];
        float* tmp = n->dt1[l + 1];  // free until the encoder pass
        NNCHK(rt::nn_up2_back(n->du[l], tmp, B, h / 2, w / 2, width(n, l + 1), n->st));
        if (l + 1 == L - 1) NNCHK(rt::nn_add(target, tmp, count(l + 1, width(n, l + 1)), n->st));
        else NNCHK(hipMemcpyAsync(target, tmp, sizeof(float) * count(l + 1, width(n, l + 1)), hipMemcpyDeviceToDevice,
                                  n->st));
    }
    for (int l = L - 1; l >= 0; l--) {
        const int h = H >> l, w = W >> l, cl = width(n, l);
        NNCHK(rt::nn_relu_back(n->de[l], n->e[l], count(l, cl), n->st));
        NNCHK(conv_bwd(n, n->enc_b[l], n->t1[l], n->de[l], n->dt1[l], h, w));
        NNCHK(rt::nn_relu_back(n->dt1[l], n->t1[l], count(l, cl), n->st));
        NNCHK(conv_bwd(n, n->enc_a[l], n->xin[l], n->dt1[l], l > 0 ? n->dxin[l] : nullptr, h, w));
        if (l > 0) {
            const int cp = width(n, l - 1);
            float* tmp = n->dt1[l - 1];  // not yet used at level l-1
            NNCHK(rt::nn_pool2_back(n->dxin[l], tmp, B, H >> (l - 1), W >> (l - 1), cp, n->st));
            NNCHK(rt::nn_add(n->de[l - 1], tmp, count(l - 1, cp), n->st));
        }
    }
    return hipSuccess;
}

int ufail(int code, const std::string& m) { return rt::set_error(code, m); }

hipError_t upload(rt_unet* n, const float* x, const float* t, const float* m) {
    const size_t px = (size_t)n->B * n->H * n->W;
    NNCHK(hipMemcpyAsync(n->x0, x, sizeof(float) * px * n->cin, hipMemcpyHostToDevice, n->st));
    if (t) NNCHK(hipMemcpyAsync(n->target, t, sizeof(float) * px, hipMemcpyHostToDevice, n->st));
    if (m) NNCHK(hipMemcpyAsync(n->mask, m, sizeof(float) * px, hipMemcpyHostToDevice, n->st));
    return hipSuccess;
}

// loss of the current forward pass against target/mask, and n->dy
hipError_t loss(rt_unet* n, double msum, float* out) {
    const int64_t px = (int64_t)n->B * n->H * n->W;
    NNCHK(rt::launch_fill_u32((uint32_t*)n->acc, 1, 0u, n->st));
    const float inv = msum > 0 ? (float)(1.0 / msum) : 0.f;
    NNCHK(rt::nn_mse(n->y, n->target, n->mask, n->dy, px, inv, n->acc, n->st));
    float s = 0.f;
    NNCHK(hipMemcpyAsync(&s, n->acc, sizeof(float), hipMemcpyDeviceToHost, n->st));
    NNCHK(hipStreamSynchronize(n->st));
    *out = (float)(s * inv);
    return hipSuccess;
}

bool shape_ok(const rt_unet* n, int B, int H, int W) {
    const int q = 1 << (n->L - 1);
    return B > 0 && H > 0 && W > 0 && H % q == 0 && W % q == 0;
}

}  // namespace

extern "C" {

int rt_unet_create(rt_ctx* ctx, int levels, int base, int in_ch, int64_t seed, rt_unet** out) {
    if (!ctx || !out || levels < 1 || levels > 8 || base < 1 || in_ch < 1)
        return ufail(RT_ERR_ARG, "rt_unet_create: bad arguments");
    rt_unet* n = new rt_unet();
    n->ctx = ctx;
    n->L = levels;
    n->C = base;
    n->cin = in_ch;
    int64_t off = 0;
    auto conv = [&](int cin, int cout) {
        Conv c{cin, cout, off, off + 9 * (int64_t)cin * cout};
        off = c.b + cout;
        return c;
    };
    for (int l = 0; l < levels; l++) {
        n->enc_a.push_back(conv(l ? base << (l - 1) : in_ch, base << l));
        n->enc_b.push_back(conv(base << l, base << l));
    }
    n->dec_up.resize(std::max(levels - 1, 0));
    n->dec_a.resize(std::max(levels - 1, 0));
    for (int l = levels - 2; l >= 0; l--) {
        n->dec_up[l] = conv(base << (l + 1), base << l);
        n->dec_a[l] = conv(base << l, base << l);
    }
    n->out = conv(base, 1);
    n->nparams = off;
    // He-normal weights, zero biases
    std::vector<float> w((size_t)off, 0.f);
    std::mt19937_64 rng((uint64_t)seed);
    auto init = [&](const Conv& c) {
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * c.cin)));
        for (int64_t k = 0; k < 9 * (int64_t)c.cin * c.cout; k++) w[(size_t)(c.w + k)] = nd(rng);
    };
    for (const Conv& c : n->enc_a) init(c);
    for (const Conv& c : n->enc_b) init(c);
    for (const Conv& c : n->dec_up) init(c);
    for (const Conv& c : n->dec_a) init(c);
    init(n->out);
    n->device = rt::ctx_device(ctx);
    n->st = rt::ctx_stream(ctx);
    hipError_t e = hipSetDevice(n->device);
    const size_t bytes = sizeof(float) * (size_t)off;
    if (e == hipSuccess) e = hipMalloc(&n->w, bytes);
    if (e == hipSuccess) e = hipMalloc(&n->g, bytes);
    if (e == hipSuccess) e = hipMalloc(&n->m1, bytes);
    if (e == hipSuccess) e = hipMalloc(&n->m2, bytes);
    if (e == hipSuccess) e = hipMalloc(&n->wbf, sizeof(uint16_t) * (size_t)off);
    if (e == hipSuccess) e = hipMemcpy(n->w, w.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(n->m1, 0, bytes);
    if (e == hipSuccess) e = hipMemset(n->m2, 0, bytes);
    if (e == hipSuccess) e = hipMemset(n->g, 0, bytes);
    if (e == hipSuccess) e = rt::nn_to_bf16(n->w, n->wbf, off, n->st);
    if (e == hipSuccess) e = hipStreamSynchronize(n->st);
    if (e != hipSuccess) {
        rt_unet_destroy(n);
        return ufail(RT_ERR_HIP, std::string("rt_unet_create: ") + hipGetErrorString(e));
    }
    *out = n;
    return RT_OK;
}

int rt_unet_destroy(rt_unet* n) {
    if (!n) return RT_OK;
    (void)hipSetDevice(n->device);
    if (n->st) (void)hipStreamSynchronize(n->st);
    free_ws(n);
    for (void* p : {(void*)n->w, (void*)n->g, (void*)n->m1, (void*)n->m2, (void*)n->wbf})
        if (p) (void)hipFree(p);
    delete n;
    return RT_OK;
}

int64_t rt_unet_param_count(rt_unet* n) { return n ? n->nparams : 0; }

int rt_unet_get_params(rt_unet* n, float* out) {
    if (!n || !out) return ufail(RT_ERR_ARG, "rt_unet_get_params: null argument");
    if (hipMemcpy(out, n->w, sizeof(float) * (size_t)n->nparams, hipMemcpyDeviceToHost) != hipSuccess)
        return ufail(RT_ERR_HIP, "rt_unet_get_params: copy failed");
    return RT_OK;
}

int rt_unet_set_params(rt_unet* n, const float* in) {
    if (!n || !in) return ufail(RT_ERR_ARG, "rt_unet_set_params: null argument");
    if (hipMemcpy(n->w, in, sizeof(float) * (size_t)n->nparams, hipMemcpyHostToDevice) != hipSuccess ||
        rt::nn_to_bf16(n->w, n->wbf, n->nparams, n->st) != hipSuccess || hipStreamSynchronize(n->st) != hipSuccess)
        return ufail(RT_ERR_HIP, "rt_unet_set_params: copy failed");
    n->adam_t = 0;
    (void)hipMemset(n->m1, 0, sizeof(float) * (size_t)n->nparams);
    (void)hipMemset(n->m2, 0, sizeof(float) * (size_t)n->nparams);
    return RT_OK;
}

// x: B x H x W x in_ch (NHWC), y: B x H x W.
int rt_unet_forward(rt_unet* n, const float* x, int B, int H, int W, float* y) {
    if (!n || !x || !y || !shape_ok(n, B, H, W)) return ufail(RT_ERR_ARG, "rt_unet_forward: bad arguments");
    (void)hipSetDevice(n->device);
    hipError_t e = ensure_ws(n, B, H, W);
    if (e == hipSuccess) e = upload(n, x, nullptr, nullptr);
    if (e == hipSuccess) e = forward(n);
    if (e == hipSuccess) e = hipMemcpyAsync(y, n->y, sizeof(float) * (size_t)B * H * W, hipMemcpyDeviceToHost, n->st);
    if (e == hipSuccess) e = hipStreamSynchronize(n->st);
    if (e != hipSuccess) return ufail(RT_ERR_HIP, std::string("rt_unet_forward: ") + hipGetErrorString(e));
    return RT_OK;
}

// Masked mean squared error against `target` over pixels with mask > 0 (mask values weight
// the residual); writes the loss and, when grad != NULL, its gradient w.r.t. every
// parameter.
int rt_unet_loss_grad(rt_unet* n, const float* x, const float* target, const float* mask, int B, int H, int W,
                      float* loss_out, float* grad) {
    if (!n || !x || !target || !mask || !loss_out || !shape_ok(n, B, H, W))
        return ufail(RT_ERR_ARG, "rt_unet_loss_grad: bad arguments");
    (void)hipSetDevice(n->device);
    double msum = 0;
    for (int64_t k = 0; k < (int64_t)B * H * W; k++) msum += mask[k] != 0.f ? 1.0 : 0.0;
    hipError_t e = ensure_ws(n, B, H, W);
    if (e == hipSuccess) e = upload(n, x, target, mask);
    if (e == hipSuccess) e = forward(n);
    if (e == hipSuccess) e = loss(n, msum, loss_out);
    if (e == hipSuccess && grad) e = backward(n);
    if (e == hipSuccess && grad)
        e = hipMemcpyAsync(grad, n->g, sizeof(float) * (size_t)n->nparams, hipMemcpyDeviceToHost, n->st);
    if (e == hipSuccess) e = hipStreamSynchronize(n->st);
    if (e != hipSuccess) return ufail(RT_ERR_HIP, std::string("rt_unet_loss_grad: ") + hipGetErrorString(e));
    return RT_OK;
}

// One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the batch; writes the loss before the
// step.
int rt_unet_train_step(rt_unet* n, const float* x, const float* target, const float* mask, int B, int H, int W, float lr,
                       float* loss_out) {
    float l = 0.f;
    int rc = rt_unet_loss_grad(n, x, target, mask, B, H, W, &l, nullptr);
    if (rc) return rc;
    hipError_t e = backward(n);
    n->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f / (1.f - std::pow(b1, (float)n->adam_t)), c2 = 1.f / (1.f - std::pow(b2, (float)n->adam_t));
    if (e == hipSuccess) e = rt::nn_adam(n->w, n->g, n->m1, n->m2, n->wbf, n->nparams, lr, b1, b2, 1e-8f, c1, c2, n->st);
    if (e == hipSuccess) e = hipStreamSynchronize(n->st);
    if (e != hipSuccess) return ufail(RT_ERR_HIP, std::string("rt_unet_train_step: ") + hipGetErrorString(e));
    if (loss_out) *loss_out = l;
    return RT_OK;
}

// C = A B (A: M x K, B: K x N, both converted to bf16), C: M x N fp32 — the matrix-core
// product on its own, for tests of the kernel.
int rt_nn_gemm_ex(rt_ctx* ctx, int M, int N, int K, int layout, const float* A, const float* B, float* C) {
    if (!ctx || !A || !B || !C || M <= 0 || N <= 0 || K <= 0 || layout < 0 || layout > 2)
        return ufail(RT_ERR_ARG, "rt_nn_gemm: bad arguments");
    hipStream_t st = rt::ctx_stream(ctx);
    (void)hipSetDevice(rt::ctx_device(ctx));
    float *a = nullptr, *b = nullptr, *c = nullptr;
    uint16_t *ah = nullptr, *bh = nullptr;
    hipError_t e = hipMalloc(&a, sizeof(float) * (size_t)M * K);
    if (e == hipSuccess) e = hipMalloc(&b, sizeof(float) * (size_t)K * N);
    if (e == hipSuccess) e = hipMalloc(&c, sizeof(float) * (size_t)M * N);
    if (e == hipSuccess) e = hipMalloc(&ah, 2 * (size_t)M * K);
    if (e == hipSuccess) e = hipMalloc(&bh, 2 * (size_t)K * N);
    uint16_t* ws = nullptr;
    const int64_t wse = rt::nn_gemm_workspace(M, N, K);
    if (e == hipSuccess) e = hipMalloc(&ws, 2 * (size_t)wse);
    if (e == hipSuccess) e = hipMemcpy(a, A, sizeof(float) * (size_t)M * K, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b, B, sizeof(float) * (size_t)K * N, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = rt::nn_to_bf16(a, ah, (int64_t)M * K, st);
    if (e == hipSuccess) e = rt::nn_to_bf16(b, bh, (int64_t)K * N, st);
    rt::GemmArgs g{ah, bh, c, M, N, K, K, 1, N, 1, N, 1.f, 0.f};
    if (layout == 1) {  // A stored [K x M]
        g.sam = 1;
        g.sak = M;
    } else if (layout == 2) {  // B stored [N x K]
        g.sbk = 1;
        g.sbn = K;
    }
    g.ws = ws;
    g.ws_elems = wse;
    if (e == hipSuccess) e = rt::nn_gemm(g, st);
    if (e == hipSuccess) e = hipMemcpyAsync(C, c, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    for (void* p : {(void*)a, (void*)b, (void*)c, (void*)ah, (void*)bh, (void*)ws})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return ufail(RT_ERR_HIP, std::string("rt_nn_gemm: ") + hipGetErrorString(e));
    return RT_OK;
}

int rt_nn_gemm(rt_ctx* ctx, int M, int N, int K, const float* A, const float* B, float* C) {
    return rt_nn_gemm_ex(ctx, M, N, K, 0, A, B, C);
}

// Device-resident timing of the matrix-core product for one operand layout (0: A and B
// row-major as in a convolution's forward product; 1: A stored transposed, the weight
// gradient; 2: B stored transposed, the input gradient), on bf16 operands filled on the
// device: milliseconds per product averaged over `reps` launches (HIP events on the context
// stream).
int rt_nn_gemm_time(rt_ctx* ctx, int M, int N, int K, int layout, int reps, double* ms) {
    if (!ctx || !ms || M <= 0 || N <= 0 || K <= 0 || reps <= 0 || layout < 0 || layout > 2)
        return ufail(RT_ERR_ARG, "rt_nn_gemm_time: bad arguments");
    hipStream_t st = rt::ctx_stream(ctx);
    (void)hipSetDevice(rt::ctx_device(ctx));
    float* f = nullptr;
    float* c = nullptr;
    uint16_t *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int64_t na = (int64_t)M * K, nb = (int64_t)K * N;
    hipError_t e = hipMalloc(&f, sizeof(float) * (size_t)std::max(na, nb));
    if (e == hipSuccess) e = hipMalloc(&a, 2 * (size_t)na);
    if (e == hipSuccess) e = hipMalloc(&b, 2 * (size_t)nb);
    if (e == hipSuccess) e = hipMalloc(&c, sizeof(float) * (size_t)M * N);
    uint16_t* ws = nullptr;
    const int64_t wse = rt::nn_gemm_workspace(M, N, K);
    if (e == hipSuccess) e = hipMalloc(&ws, 2 * (size_t)wse);
    if (e == hipSuccess) e = hipMemsetAsync(f, 0, sizeof(float) * (size_t)std::max(na, nb), st);
    if (e == hipSuccess) e = rt::nn_to_bf16(f, a, na, st);
    if (e == hipSuccess) e = rt::nn_to_bf16(f, b, nb, st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    rt::GemmArgs g{a, b, c, M, N, K, K, 1, N, 1, N, 1.f, 0.f};
    if (layout == 1) {
        g.sam = 1;
        g.sak = M;
    } else if (layout == 2) {
        g.sbk = 1;
        g.sbn = K;
    }
    g.ws = ws;
    g.ws_elems = wse;
    if (e == hipSuccess) e = rt::nn_gemm(g, st);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = rt::nn_gemm(g, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    *ms = (double)t / reps;
    for (void* p : {(void*)f, (void*)c, (void*)a, (void*)b, (void*)ws})
        if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return ufail(RT_ERR_HIP, std::string("rt_nn_gemm_time: ") + hipGetErrorString(e));
    return RT_OK;
}

}  // extern "C"
