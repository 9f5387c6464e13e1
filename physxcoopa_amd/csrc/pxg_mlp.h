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

// launchers (pxg_mlp.hip)
// predictions for `n` feature rows (device pointers): out[i*3 .. +3]; rows are independent
void launch_mlp_infer(const MlpDev* net, const float* feats, float* out, uint32_t n, hipStream_t st);
// n rows read from a device counter (grid sized by n_bound)
void launch_mlp_infer_dev(const MlpDev* net, const float* feats, float* out, const uint32_t* n_dev, uint32_t n_bound,
                          hipStream_t st);

}  // namespace pxg
