// The learned warm start's network: 24 -> 64 -> 64 -> 3, ReLU.
//
// Inference runs on the matrix cores: each wave64 evaluates 16 contact rows with
// v_mfma_f32_16x16x32_bf16 (BF16 inputs and weights, FP32 accumulation); the hidden activations
// are re-laid out through LDS between layers (the MFMA output layout holds 4 rows x 1 column per
// lane, the input layout 1 row x 8 K values per lane). Training runs in FP32: per block a tile of
// 32 samples goes forward and backward in LDS, each block writes its partial gradients, and one
// reduction kernel sums them in block order (deterministic) and applies Adam.
#include "../../include/pxg.h"
#include "pxg_mlp.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace pxg {

namespace {


constexpr int kWavesPerBlock = 4;  // 64 rows per block

__global__ void __launch_bounds__(256) k_mlp_infer(const MlpDev* net, const float* feats, float* out,
                                                   const uint32_t* n_dev, uint32_t n_host) {
    __shared__ uint16_t s_w1[kMlpH * kMlpInPad];
    __shared__ uint16_t s_w2[kMlpH * kMlpH];
    __shared__ uint16_t s_w3[kMlpOutPad * kMlpH];
    __shared__ float s_b1[kMlpH], s_b2[kMlpH], s_b3[kMlpOutPad], s_shift[kMlpIn], s_scale[kMlpIn], s_oscale[kMlpOut];
    __shared__ uint16_t s_h[kWavesPerBlock][16][kMlpH];
    const uint32_t n = n_dev ? *n_dev : n_host;
    const int t = threadIdx.x;
    for (int i = t; i < kMlpH * kMlpInPad; i += 256) s_w1[i] = net->w1[i];
    for (int i = t; i < kMlpH * kMlpH; i += 256) s_w2[i] = net->w2[i];
    for (int i = t; i < kMlpOutPad * kMlpH; i += 256) s_w3[i] = net->w3[i];
    if (t < kMlpH) { s_b1[t] = net->b1[t]; s_b2[t] = net->b2[t]; }
    if (t < kMlpOutPad) s_b3[t] = net->b3[t];
    if (t < kMlpIn) { s_shift[t] = net->in_shift[t]; s_scale[t] = net->in_scale[t]; }
    if (t < kMlpOut) s_oscale[t] = net->out_scale[t];
    __syncthreads();
    MlpView W;
    W.w1 = s_w1; W.w2 = s_w2; W.w3 = s_w3; W.b1 = s_b1; W.b2 = s_b2; W.b3 = s_b3;
    W.shift = s_shift; W.scale = s_scale; W.oscale = s_oscale;
    const int wave = t >> 6;
    for (uint32_t row0 = (blockIdx.x * kWavesPerBlock + wave) * 16u; row0 < n; row0 += gridDim.x * kWavesPerBlock * 16u)
        mlp_rows16(W, feats, row0, n, s_h[wave], out);
}

// ---------------------------------------------------------------------------------------------
// training (FP32)
// ---------------------------------------------------------------------------------------------
constexpr int kTile = 32;  // samples per block
constexpr int kP1 = kMlpH * kMlpIn, kP2 = kMlpH * kMlpH, kP3 = kMlpOut * kMlpH;
constexpr int kNumParams = kP1 + kMlpH + kP2 + kMlpH + kP3 + kMlpOut;  // w1 b1 w2 b2 w3 b3

// params layout: w1[64][24], b1[64], w2[64][64], b2[64], w3[3][64], b3[3]
__global__ void __launch_bounds__(256) k_mlp_grad(const float* params, const float* X, const float* T,
                                                  const uint32_t* idx, uint32_t n_batch, float inv_count,
                                                  const float* shift, const float* scale, float* partial,
                                                  float* loss_partial) {
    __shared__ float s_x[kTile][kMlpIn];
    __shared__ float s_h1[kTile][kMlpH];
    __shared__ float s_h2[kTile][kMlpH];
    __shared__ float s_d2[kTile][kMlpH];
    __shared__ float s_d1[kTile][kMlpH];
    __shared__ float s_y[kTile][4];
    __shared__ float s_loss[256];
    const float* w1 = params;
    const float* b1 = w1 + kP1;
    const float* w2 = b1 + kMlpH;
    const float* b2 = w2 + kP2;
    const float* w3 = b2 + kMlpH;
    const float* b3 = w3 + kP3;
    const int t = threadIdx.x;
    const uint32_t s0 = blockIdx.x * kTile;
    for (int e = t; e < kTile * kMlpIn; e += 256) {
        int s = e / kMlpIn, k = e % kMlpIn;
        float v = 0.0f;
        if (s0 + s < n_batch) v = (X[size_t(idx[s0 + s]) * kMlpIn + k] - shift[k]) * scale[k];
        s_x[s][k] = v;
    }
    __syncthreads();
    for (int e = t; e < kTile * kMlpH; e += 256) {
        int s = e / kMlpH, j = e % kMlpH;
        float a = b1[j];
        for (int k = 0; k < kMlpIn; ++k) a += w1[j * kMlpIn + k] * s_x[s][k];
        s_h1[s][j] = fmaxf(a, 0.0f);
    }
    __syncthreads();
    for (int e = t; e < kTile * kMlpH; e += 256) {
        int s = e / kMlpH, j = e % kMlpH;
        float a = b2[j];
        for (int k = 0; k < kMlpH; ++k) a += w2[j * kMlpH + k] * s_h1[s][k];
        s_h2[s][j] = fmaxf(a, 0.0f);
    }
    __syncthreads();
    float lsum = 0.0f;
    if (t < kTile * kMlpOut) {
        int s = t / kMlpOut, o = t % kMlpOut;
        float a = b3[o];
        for (int k = 0; k < kMlpH; ++k) a += w3[o * kMlpH + k] * s_h2[s][k];
        float d = 0.0f;
        if (s0 + s < n_batch) {
            float err = a - T[size_t(idx[s0 + s]) * kMlpOut + o];
            lsum = err * err;
            d = 2.0f * err * inv_count;  // dL/dy for L = mean over samples of the squared error sum
        }
        s_y[s][o] = d;
    }
    s_loss[t] = lsum;
    __syncthreads();
    // dH2 = (dY W3) * relu'
    for (int e = t; e < kTile * kMlpH; e += 256) {
        int s = e / kMlpH, j = e % kMlpH;
        float a = 0.0f;
        for (int o = 0; o < kMlpOut; ++o) a += s_y[s][o] * w3[o * kMlpH + j];
        s_d2[s][j] = s_h2[s][j] > 0.0f ? a : 0.0f;
    }
    __syncthreads();
    // dH1 = (dH2 W2) * relu'
    for (int e = t; e < kTile * kMlpH; e += 256) {
        int s = e / kMlpH, j = e % kMlpH;
        float a = 0.0f;
        for (int k = 0; k < kMlpH; ++k) a += s_d2[s][k] * w2[k * kMlpH + j];
        s_d1[s][j] = s_h1[s][j] > 0.0f ? a : 0.0f;
    }
    __syncthreads();
    float* gp = partial + size_t(blockIdx.x) * kNumParams;
    for (int p = t; p < kNumParams; p += 256) {
        float a = 0.0f;
        if (p < kP1) {
            int j = p / kMlpIn, k = p % kMlpIn;
            for (int s = 0; s < kTile; ++s) a += s_d1[s][j] * s_x[s][k];
        } else if (p < kP1 + kMlpH) {
            int j = p - kP1;
            for (int s = 0; s < kTile; ++s) a += s_d1[s][j];
        } else if (p < kP1 + kMlpH + kP2) {
            int q = p - kP1 - kMlpH, j = q / kMlpH, k = q % kMlpH;
            for (int s = 0; s < kTile; ++s) a += s_d2[s][j] * s_h1[s][k];
        } else if (p < kP1 + 2 * kMlpH + kP2) {
            int j = p - kP1 - kMlpH - kP2;
            for (int s = 0; s < kTile; ++s) a += s_d2[s][j];
        } else if (p < kP1 + 2 * kMlpH + kP2 + kP3) {
            int q = p - kP1 - 2 * kMlpH - kP2, o = q / kMlpH, k = q % kMlpH;
            for (int s = 0; s < kTile; ++s) a += s_y[s][o] * s_h2[s][k];
        } else {
            int o = p - kP1 - 2 * kMlpH - kP2 - kP3;
            for (int s = 0; s < kTile; ++s) a += s_y[s][o];
        }
        gp[p] = a;
    }
    if (t == 0) {
        float l = 0.0f;
        for (int i = 0; i < 256; ++i) l += s_loss[i];
        loss_partial[blockIdx.x] = l;
    }
}

// sums the per-block gradients in block order; Adam step when lr > 0
__global__ void k_mlp_reduce_adam(float* params, const float* partial, uint32_t nblk, float* grad_out, float* m, float* v,
                                  float lr, float b1, float b2, float eps, float wd, float bc1, float bc2) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNumParams) return;
    float g = 0.0f;
    for (uint32_t b = 0; b < nblk; ++b) g += partial[size_t(b) * kNumParams + p];
    if (grad_out) grad_out[p] = g;
    if (lr > 0.0f) {
        float mm = b1 * m[p] + (1.0f - b1) * g;
        float vv = b2 * v[p] + (1.0f - b2) * g * g;
        m[p] = mm;
        v[p] = vv;
        float upd = (mm / bc1) / (sqrtf(vv / bc2) + eps);
        params[p] = params[p] - lr * (upd + wd * params[p]);
    }
}

uint16_t host_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

thread_local std::string g_mlp_error;
int mfail(int code, const std::string& m) {
    g_mlp_error = "[pxg mlp] " + m;
    return code;
}
#define MCHK(call)                                                                                       \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) return mfail(PXG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

void pack_params(const pxg_mlp* net, std::vector<float>& p) {
    p.resize(kNumParams);
    float* o = p.data();
    std::memcpy(o, net->w1, sizeof(float) * kP1); o += kP1;
    std::memcpy(o, net->b1, sizeof(float) * kMlpH); o += kMlpH;
    std::memcpy(o, net->w2, sizeof(float) * kP2); o += kP2;
    std::memcpy(o, net->b2, sizeof(float) * kMlpH); o += kMlpH;
    std::memcpy(o, net->w3, sizeof(float) * kP3); o += kP3;
    std::memcpy(o, net->b3, sizeof(float) * kMlpOut);
}
void unpack_params(const std::vector<float>& p, pxg_mlp* net) {
    const float* o = p.data();
    std::memcpy(net->w1, o, sizeof(float) * kP1); o += kP1;
    std::memcpy(net->b1, o, sizeof(float) * kMlpH); o += kMlpH;
    std::memcpy(net->w2, o, sizeof(float) * kP2); o += kP2;
    std::memcpy(net->b2, o, sizeof(float) * kMlpH); o += kMlpH;
    std::memcpy(net->w3, o, sizeof(float) * kP3); o += kP3;
    std::memcpy(net->b3, o, sizeof(float) * kMlpOut);
}

}  // namespace

// BF16 device image of a pxg_mlp
void mlp_to_device_image(const pxg_mlp* net, MlpDev* img) {
    std::memset(img, 0, sizeof(*img));
    for (int j = 0; j < kMlpH; ++j)
        for (int k = 0; k < kMlpIn; ++k) img->w1[j * kMlpInPad + k] = host_bf16(net->w1[j * kMlpIn + k]);
    for (int i = 0; i < kMlpH * kMlpH; ++i) img->w2[i] = host_bf16(net->w2[i]);
    for (int o = 0; o < kMlpOut; ++o)
        for (int k = 0; k < kMlpH; ++k) img->w3[o * kMlpH + k] = host_bf16(net->w3[o * kMlpH + k]);
    std::memcpy(img->b1, net->b1, sizeof(img->b1));
    std::memcpy(img->b2, net->b2, sizeof(img->b2));
    for (int o = 0; o < kMlpOut; ++o) img->b3[o] = net->b3[o];
    std::memcpy(img->in_shift, net->in_shift, sizeof(img->in_shift));
    std::memcpy(img->in_scale, net->in_scale, sizeof(img->in_scale));
    std::memcpy(img->out_scale, net->out_scale, sizeof(img->out_scale));
}

void launch_mlp_infer(const MlpDev* net, const float* feats, float* out, uint32_t n, hipStream_t st) {
    if (!n) return;
    uint32_t blocks = std::min<uint32_t>((n + 63) / 64, 2048);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, st, net, feats, out, nullptr, n);
}
void launch_mlp_infer_dev(const MlpDev* net, const float* feats, float* out, const uint32_t* n_dev, uint32_t n_bound,
                          hipStream_t st) {
    if (!n_bound) return;
    uint32_t blocks = std::min<uint32_t>((n_bound + 63) / 64, 2048);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, st, net, feats, out, n_dev, 0u);
}

}  // namespace pxg

using namespace pxg;

extern "C" const char* pxg_mlp_last_error(void) { return g_mlp_error.c_str(); }

// BF16-MFMA inference of n feature rows on the current device (allocates temporaries).
extern "C" int pxg_mlp_infer(const pxg_mlp* net, const float* features, uint32_t n, float* out3) {
    if (!net || (n && (!features || !out3))) return mfail(PXG_ERR_INVALID, "pxg_mlp_infer: null argument");
    if (!n) return PXG_OK;
    MlpDev img;
    mlp_to_device_image(net, &img);
    MlpDev* d_net = nullptr;
    float *d_x = nullptr, *d_y = nullptr;
    MCHK(hipMalloc(&d_net, sizeof(MlpDev)));
    MCHK(hipMalloc(&d_x, size_t(n) * kMlpIn * 4));
    MCHK(hipMalloc(&d_y, size_t(n) * kMlpOut * 4));
    MCHK(hipMemcpy(d_net, &img, sizeof(img), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(d_x, features, size_t(n) * kMlpIn * 4, hipMemcpyHostToDevice));
    launch_mlp_infer(d_net, d_x, d_y, n, nullptr);
    MCHK(hipGetLastError());
    MCHK(hipMemcpy(out3, d_y, size_t(n) * kMlpOut * 4, hipMemcpyDeviceToHost));
    (void)hipFree(d_net);
    (void)hipFree(d_x);
    (void)hipFree(d_y);
    return PXG_OK;
}

namespace {
struct TrainBuffers {
    float *params = nullptr, *partial = nullptr, *loss = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    float *X = nullptr, *T = nullptr, *shift = nullptr, *scale = nullptr;
    uint32_t* idx = nullptr;
    ~TrainBuffers() {
        for (void* p : {(void*)params, (void*)partial, (void*)loss, (void*)m, (void*)v, (void*)grad, (void*)X, (void*)T,
                        (void*)shift, (void*)scale, (void*)idx})
            if (p) (void)hipFree(p);
    }
};
}  // namespace

// Full-batch gradient of the loss (mean over samples of the squared error of the normalised
// outputs, targets Y / out_scale) at the given parameters; FP32 on the GPU.
extern "C" int pxg_mlp_gradients(const pxg_mlp* net, const float* X, const float* Y, uint32_t n, pxg_mlp* grad,
                                 float* loss) {
    if (!net || !X || !Y || !grad || !n) return mfail(PXG_ERR_INVALID, "pxg_mlp_gradients: bad argument");
    TrainBuffers b;
    uint32_t nblk = (n + kTile - 1) / kTile;
    std::vector<float> p, T(size_t(n) * kMlpOut);
    pack_params(net, p);
    for (uint32_t i = 0; i < n; ++i)
        for (int o = 0; o < kMlpOut; ++o) T[size_t(i) * kMlpOut + o] = Y[size_t(i) * kMlpOut + o] / net->out_scale[o];
    std::vector<uint32_t> idx(n);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    MCHK(hipMalloc(&b.params, kNumParams * 4));
    MCHK(hipMalloc(&b.partial, size_t(nblk) * kNumParams * 4));
    MCHK(hipMalloc(&b.loss, nblk * 4));
    MCHK(hipMalloc(&b.grad, kNumParams * 4));
    MCHK(hipMalloc(&b.X, size_t(n) * kMlpIn * 4));
    MCHK(hipMalloc(&b.T, size_t(n) * kMlpOut * 4));
    MCHK(hipMalloc(&b.shift, kMlpIn * 4));
    MCHK(hipMalloc(&b.scale, kMlpIn * 4));
    MCHK(hipMalloc(&b.idx, n * 4));
    MCHK(hipMemcpy(b.params, p.data(), kNumParams * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(b.X, X, size_t(n) * kMlpIn * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(b.T, T.data(), size_t(n) * kMlpOut * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(b.shift, net->in_shift, kMlpIn * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(b.scale, net->in_scale, kMlpIn * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(b.idx, idx.data(), n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(256), 0, nullptr, b.params, b.X, b.T, b.idx, n, 1.0f / float(n),
                       b.shift, b.scale, b.partial, b.loss);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((kNumParams + 255) / 256), dim3(256), 0, nullptr, b.params, b.partial,
                       nblk, b.grad, nullptr, nullptr, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f);
    MCHK(hipGetLastError());
    std::vector<float> g(kNumParams), l(nblk);
    MCHK(hipMemcpy(g.data(), b.grad, kNumParams * 4, hipMemcpyDeviceToHost));
    MCHK(hipMemcpy(l.data(), b.loss, nblk * 4, hipMemcpyDeviceToHost));
    std::memset(grad, 0, sizeof(*grad));
    unpack_params(g, grad);
    if (loss) {
        double s = 0.0;
        for (float x : l) s += x;
        *loss = float(s / n);
    }
    return PXG_OK;
}

// Adam on minibatches; `net` carries the initial weights and the normalisation on entry and the
// trained weights on return. loss_per_epoch (epochs entries) receives the mean training loss.
extern "C" int pxg_mlp_train(pxg_mlp* net, const float* X, const float* Y, uint32_t n, const pxg_mlp_train_config* cfg,
                             float* loss_per_epoch) {
    if (!net || !X || !Y || !cfg || !n || !cfg->batch) return mfail(PXG_ERR_INVALID, "pxg_mlp_train: bad argument");
    TrainBuffers b;
    uint32_t B = std::min(cfg->batch, n);
    uint32_t nblk = (B + kTile - 1) / kTile;
    std::vector<float> p, T(size_t(n) * kMlpOut), zeros(kNumParams, 0.0f);
    pack_params(net, p);
    for (uint32_t i = 0; i < n; ++i)
        for (int o = 0; o < kMlpOut; ++o) T[size_t(i) * kMlpOut + o] = Y[size_t(i) * kMlpOut + o] / net->out_scale[o];
    MCHK(hipMalloc(&b.params, kNumParams * 4));
    MCHK(hipMalloc(&b.partial, size_t(nblk) * kNumParams * 4));
    MCHK(hipMalloc(&b.loss, nblk * 4));
    MCHK(hipMalloc(&b.m, kNumParams * 4));
    MCHK(hipMalloc(&b.v, kNumParams * 4));
    MCHK(hipMalloc(&b.X, size_t(n) * kMlpIn * 4));
    MCHK(hipMalloc(&b.T, size_t(n) * kMlpOut * 4));
    MCHK(hipMalloc(&b.shift, kMlpIn * 4));
    MCHK(hipMalloc(&b.scale, kMlpIn * 4));
    MCHK(hipMalloc(&b.idx, size_t(n) * 4));
    MCHK(hipMemcpy(b.params, p.data(), kNumParams * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(b.m, zeros.data(), kNumParams * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(b.v, zeros.data(), kNumParams * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(b.X, X, size_t(n) * kMlpIn * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(b.T, T.data(), size_t(n) * kMlpOut * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(b.shift, net->in_shift, kMlpIn * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(b.scale, net->in_scale, kMlpIn * 4, hipMemcpyHostToDevice));
    std::mt19937 rng(cfg->seed);
    std::vector<uint32_t> perm(n);
    std::vector<float> l(nblk);
    uint64_t step = 0;
    float b1 = cfg->beta1 > 0.0f ? cfg->beta1 : 0.9f, b2 = cfg->beta2 > 0.0f ? cfg->beta2 : 0.999f;
    float eps = cfg->eps > 0.0f ? cfg->eps : 1e-8f;
    for (uint32_t ep = 0; ep < cfg->epochs; ++ep) {
        for (uint32_t i = 0; i < n; ++i) perm[i] = i;
        std::shuffle(perm.begin(), perm.end(), rng);
        MCHK(hipMemcpy(b.idx, perm.data(), size_t(n) * 4, hipMemcpyHostToDevice));
        double ls = 0.0;
        uint32_t sampled = 0;
        uint32_t nb_batches = n / B;
        if (nb_batches == 0) nb_batches = 1;
        float lr = cfg->lr * (cfg->lr_decay > 0.0f ? std::pow(cfg->lr_decay, float(ep)) : 1.0f);
        for (uint32_t k = 0; k < nb_batches; ++k) {
            ++step;
            float bc1 = 1.0f - std::pow(b1, float(step)), bc2 = 1.0f - std::pow(b2, float(step));
            hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(256), 0, nullptr, b.params, b.X, b.T, b.idx + size_t(k) * B,
                               B, 1.0f / float(B), b.shift, b.scale, b.partial, b.loss);
            hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((kNumParams + 255) / 256), dim3(256), 0, nullptr, b.params,
                               b.partial, nblk, nullptr, b.m, b.v, lr, b1, b2, eps, cfg->weight_decay, bc1, bc2);
            if ((k & 15) == 0 || k + 1 == nb_batches) {
                MCHK(hipMemcpy(l.data(), b.loss, nblk * 4, hipMemcpyDeviceToHost));
                double s = 0.0;
                for (float x : l) s += x;
                ls += s / B;
                ++sampled;
            }
        }
        MCHK(hipGetLastError());
        if (loss_per_epoch) loss_per_epoch[ep] = float(ls / std::max(1u, sampled));
    }
    MCHK(hipMemcpy(p.data(), b.params, kNumParams * 4, hipMemcpyDeviceToHost));
    unpack_params(p, net);
    return PXG_OK;
}
