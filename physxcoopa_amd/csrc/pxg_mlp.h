// Learned warm start: contact features, the 24-64-64-3 MLP (BF16 MFMA inference, FP32 training).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pxg_device.h"

namespace pxg {

constexpr int kMlpIn = 24, kMlpH = 64, kMlpOut = 3;
constexpr int kMlpInPad = 32, kMlpOutPad = 16;

// Device image of the network: BF16 weights row-major [out][in] (padded), FP32 biases and scales.
struct MlpDev {
    uint16_t w1[kMlpH * kMlpInPad];    // [64][32], columns 24..31 zero
    uint16_t w2[kMlpH * kMlpH];        // [64][64]
    uint16_t w3[kMlpOutPad * kMlpH];   // [16][64], rows 3..15 zero
    float b1[kMlpH], b2[kMlpH], b3[kMlpOutPad];
    float in_shift[kMlpIn], in_scale[kMlpIn];
    float out_scale[kMlpOut];
};

// The 24 input features of one contact point, computed from the pre-solve state (velocities after
// force integration) exactly as the solver's warm start sees it:
//   0-2   contact normal (world)
//   3-5   relative velocity vb - va at the point in the contact frame (n, t1, t2)
//   6     penetration depth, 7 speculative separation
//   8, 9  effective inverse masses of A and B (0 unless dynamic and awake)
//   10-12 lever arm ra in the contact frame, 13-15 rb in the contact frame
//   16-18 Ia^-1 (ra x n) in the contact frame, 19-21 Ib^-1 (rb x n) in the contact frame
//   22    friction, 23 restitution
struct FeatBody {
    pxd::v3 x, v, w;
    float im;
    pxd::m3 I;
};
__device__ __forceinline__ void contact_features(const FeatBody& A, const FeatBody& B, pxd::v3 n, pxd::v3 pos, float pen,
                                                 float sep, float mu, float e, float* f) {
    using namespace pxd;
    v3 t1, t2;
    orthonormal_basis(n, t1, t2);
    v3 ra = pos - A.x, rb = pos - B.x;
    v3 dv = (B.v + cross(B.w, rb)) - (A.v + cross(A.w, ra));
    v3 ia = mul(A.I, cross(ra, n)), ib = mul(B.I, cross(rb, n));
    f[0] = n.x; f[1] = n.y; f[2] = n.z;
    f[3] = dot(dv, n); f[4] = dot(dv, t1); f[5] = dot(dv, t2);
    f[6] = pen; f[7] = sep;
    f[8] = A.im; f[9] = B.im;
    f[10] = dot(ra, n); f[11] = dot(ra, t1); f[12] = dot(ra, t2);
    f[13] = dot(rb, n); f[14] = dot(rb, t1); f[15] = dot(rb, t2);
    f[16] = dot(ia, n); f[17] = dot(ia, t1); f[18] = dot(ia, t2);
    f[19] = dot(ib, n); f[20] = dot(ib, t1); f[21] = dot(ib, t2);
    f[22] = mu; f[23] = e;
}

// ---- inference on the matrix cores ----
typedef short mlp_i16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint16_t mlp_f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
// Where the weights are read from (LDS copies or the device image itself).
struct MlpView {
    const uint16_t *w1, *w2, *w3;
    const float *b1, *b2, *b3, *shift, *scale, *oscale;
};
__device__ __forceinline__ MlpView mlp_view(const MlpDev* n) {
    MlpView v;
    v.w1 = n->w1; v.w2 = n->w2; v.w3 = n->w3; v.b1 = n->b1; v.b2 = n->b2; v.b3 = n->b3;
    v.shift = n->in_shift; v.scale = n->in_scale; v.oscale = n->out_scale;
    return v;
}
// LDS writes of this wave visible to its own later LDS reads
__device__ __forceinline__ void mlp_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// eight consecutive BF16 weights (a row of W, K values k .. k+7) as an MFMA operand
__device__ __forceinline__ mlp_bf16x8 mlp_w8(const uint16_t* p) {
    mlp_i16x8 v;
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = short(p[q]);
    return __builtin_bit_cast(mlp_bf16x8, v);
}
// One 16x16 output tile over K = 32 with gfx950's v_mfma_f32_16x16x32_bf16: lane 16 g + r gives
// row r's K values 8g .. 8g+7 (A) and weight row (column r of the tile)'s same K values (B); the
// accumulator lane holds rows 4g .. 4g+3 of column r.
__device__ __forceinline__ mlp_f32x4 mlp_mfma(mlp_i16x8 a, mlp_bf16x8 b, mlp_f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mlp_bf16x8, a), b, acc, 0, 0, 0);
}
// One wave64 evaluates rows row0 .. row0+15 of `feats` (kMlpIn floats per row, n rows valid): each
// layer one K = 32 step per 16-wide output tile (layer 1: the 24 inputs padded to 32) or two (the
// 64-wide hidden layers), activations re-laid out through the wave's LDS buffer h between layers.
// Writes 3 outputs per valid row to out. All 64 lanes must call it.
__device__ __forceinline__ void mlp_rows16(const MlpView& W, const float* feats, uint32_t row0, uint32_t n,
                                           uint16_t (*h)[kMlpH], float* out) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    const uint32_t row = row0 + uint32_t(r);
    mlp_i16x8 a1;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int k = 8 * g + q;
        float x = 0.0f;
        if (row < n && k < kMlpIn) x = (feats[size_t(row) * kMlpIn + k] - W.shift[k]) * W.scale[k];
        a1[q] = short(mlp_f2bf(x));
    }
    for (int nt = 0; nt < 4; ++nt) {
        mlp_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mlp_mfma(a1, mlp_w8(&W.w1[(nt * 16 + r) * kMlpInPad + 8 * g]), acc);
        int col = nt * 16 + r;
        for (int i = 0; i < 4; ++i) h[4 * g + i][col] = mlp_f2bf(fmaxf(acc[i] + W.b1[col], 0.0f));
    }
    mlp_wave_sync();
    mlp_i16x8 a2[2];
    for (int ks = 0; ks < 2; ++ks)
        for (int q = 0; q < 8; ++q) a2[ks][q] = short(h[r][ks * 32 + 8 * g + q]);
    mlp_wave_sync();
    for (int nt = 0; nt < 4; ++nt) {
        mlp_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int ks = 0; ks < 2; ++ks) acc = mlp_mfma(a2[ks], mlp_w8(&W.w2[(nt * 16 + r) * kMlpH + ks * 32 + 8 * g]), acc);
        int col = nt * 16 + r;
        for (int i = 0; i < 4; ++i) h[4 * g + i][col] = mlp_f2bf(fmaxf(acc[i] + W.b2[col], 0.0f));
    }
    mlp_wave_sync();
    mlp_i16x8 a3[2];
    for (int ks = 0; ks < 2; ++ks)
        for (int q = 0; q < 8; ++q) a3[ks][q] = short(h[r][ks * 32 + 8 * g + q]);
    mlp_wave_sync();
    mlp_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int ks = 0; ks < 2; ++ks) acc = mlp_mfma(a3[ks], mlp_w8(&W.w3[r * kMlpH + ks * 32 + 8 * g]), acc);
    if (r < kMlpOut) {
        for (int i = 0; i < 4; ++i) {
            uint32_t orow = row0 + uint32_t(4 * g + i);
            if (orow < n) out[size_t(orow) * kMlpOut + r] = (acc[i] + W.b3[r]) * W.oscale[r];
        }
    }
}

// launchers (pxg_mlp.hip)
// predictions for `n` feature rows (device pointers): out[i*3 .. +3]; rows are independent
void launch_mlp_infer(const MlpDev* net, const float* feats, float* out, uint32_t n, hipStream_t st);
// n rows read from a device counter (grid sized by n_bound)
void launch_mlp_infer_dev(const MlpDev* net, const float* feats, float* out, const uint32_t* n_dev, uint32_t n_bound,
                          hipStream_t st);

}  // namespace pxg
