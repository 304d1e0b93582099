// Launchers of the warm-start U-Net kernels (unet_kernels.hip).  Tensors are
// channels-last at their level's resolution: activations BF16 bit patterns
// (uint16_t), gradients FP32.  Weights: FP32 master [27][ci][co], BF16
// copies in that layout (data gradient) and as [27][co][ci] (forward).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace artss {
namespace nn {

struct Dims {
    int X, Y, Z;  // voxels along x, y, z
};

// 3x3x3 convolution with ci, co in {16,32,64} (MFMA); up: the input is at half
// resolution and read nearest-upsampled, and Wt holds weights_up_bf16's
// coarse-form weights instead of [27][co][ci]; pre (may be null):
// pre-activation; skip (may be null): added after the activation
void conv_fwd(hipStream_t s, int ci, int co, bool up, Dims D, const uint16_t *in, const uint16_t *Wt,
              const float *bias, uint16_t *out, uint16_t *pre, const uint16_t *skip, bool relu);
// d in (at the resolution of D) from d out
void conv_dgrad(hipStream_t s, int ci, int co, Dims D, const float *gout, const uint16_t *W, float *gin);
// d W += (atomically; zero gW first)
void conv_wgrad(hipStream_t s, int ci, int co, bool up, Dims D, const uint16_t *in, const float *gout, float *gW);
// first layer 2 -> 16 (FP32 input, ReLU) and last layer 16 -> 1 (no activation)
void conv_in(hipStream_t s, Dims D, const float *x, const float *W, const float *bias, uint16_t *out);
void conv_out(hipStream_t s, Dims D, const uint16_t *in, const float *W, const float *bias, float *y);
void conv_in_wgrad(hipStream_t s, Dims D, const float *x, const float *g, float *gW, float *gb);
void conv_out_grads(hipStream_t s, Dims D, const uint16_t *in, const float *W, const float *gy, float *gW,
                    float *gb, float *gin);
// sampling (Dc: coarse dims)
void down(hipStream_t s, int c, Dims Dc, const uint16_t *in, uint16_t *out);
void down_backward_add(hipStream_t s, int c, Dims Dc, const float *gc, float *gf);
void up_backward(hipStream_t s, int c, Dims Dc, const float *gf, float *gc);
// elementwise and reductions
void bias_grad(hipStream_t s, int c, long long voxels, const float *g, float *gb);
void relu_mask(hipStream_t s, long long n, float *g, const uint16_t *act);
void copy_f(hipStream_t s, long long n, float *dst, const float *src);
void fill_f(hipStream_t s, long long n, float *dst, float v);
void mse(hipStream_t s, long long n, const float *y, const float *target, float *gy, float *loss);
void adam(hipStream_t s, int n, int ci, int co, float *w, const float *g, float *m, float *v, float lr, float b1,
          float b2, float eps, float c1, float c2, uint16_t *w_bf, uint16_t *wt_bf);
void weights_bf16(hipStream_t s, int n, int ci, int co, const float *w, uint16_t *w_bf, uint16_t *wt_bf);
// the [8 parities][8 taps][co][ci] BF16 weights conv_fwd takes for up = true:
// per output parity, the fine taps that read the same coarse voxel summed
void weights_up_bf16(hipStream_t s, int ci, int co, const float *w, uint16_t *wup);

}  // namespace nn
}  // namespace artss
