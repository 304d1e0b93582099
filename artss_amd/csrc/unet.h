// Warm-start U-Net for the pressure solve: a small 3-D convolutional network
// (levels 0..3 with 16, 32, 64, 64 channels, 3x3x3 convolutions, BF16
// activations and weights, FP32 accumulation on the matrix cores) that maps
// the normalised residual of the previous pressure and the previous
// pressure increment to the next increment.  Trained on the GPU by its own
// forward and backward kernels (unet_kernels.hip) with Adam.  It has no
// counterpart in the reference; the solver iterates from its prediction to
// the same residual target as without it.
//
// Architecture (the numpy reference oracle/unet_ref.py states it):
//   h0 = relu(conv_in(x)) 2->16;  e0 = relu(conv(h0)) 16->16
//   e1 = relu(conv(down e0)) 16->32;  e2 = relu(conv(down e1)) 32->64
//   e3 = relu(conv(down e2)) 64->64
//   d2 = relu(conv(up e3)) + e2 64->64;  d1 = relu(conv(up d2)) + e1 64->32
//   d0 = relu(conv(up d1)) + e0 32->16;  y = conv_out(d0) 16->1
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "unet_kernels.h"

namespace artss {

class UNet {
  public:
    explicit UNet(hipStream_t stream, unsigned long long seed = 0);
    ~UNet();
    UNet(const UNet &) = delete;
    UNet &operator=(const UNet &) = delete;

    // parameters in layer order (in, e0, e1, e2, e3, d2, d1, d0, out), each
    // W[27][ci][co] then b[co], FP32
    static size_t param_count();
    void set_params(const float *host);
    void get_params(float *host) const;

    // y[v] = net(x)[v] for a volume of X x Y x Z voxels (each divisible by 8);
    // x: [V][2] FP32, y: [V] FP32, both device pointers
    void forward(int X, int Y, int Z, const float *x, float *y);
    // forward + mean squared error against target + backward; gradients in
    // param order into host `grads` (if not null); returns the loss
    float gradients(int X, int Y, int Z, const float *x, const float *target, float *grads);
    // one Adam step on the gradients of the last call of gradients()
    void adam_step(float lr);

  private:
    struct Layer {
        int ci, co, level;
        float *W = nullptr, *b = nullptr, *gW = nullptr, *gb = nullptr;
        float *mW = nullptr, *vW = nullptr, *mb = nullptr, *vb = nullptr;
        uint16_t *W_bf = nullptr, *Wt_bf = nullptr;  // [27][ci][co], [27][co][ci]
        uint16_t *Wup_bf = nullptr;  // layers on an upsampled input: nn::weights_up_bf16's form
    };
    void ensure(int X, int Y, int Z);
    // keep_pre: store the decoder's pre-activations, which only the backward
    // pass of gradients() reads
    void run_forward(const float *x, bool keep_pre);
    void free_buffers();
    hipStream_t m_s;
    std::vector<Layer> m_layers;
    int m_X = 0, m_Y = 0, m_Z = 0;
    std::vector<void *> m_buffers;
    // activations (BF16) and gradients (FP32) at the current size
    uint16_t *h0, *e0, *p1, *e1, *p2, *e2, *p3, *e3, *c2, *d2, *c1, *d1, *c0, *d0;
    float *y, *gy, *g0a, *g0b, *g0c, *g1a, *g1b, *g1c, *g2a, *g2b, *g2c, *g3a, *g3b;
    const float *m_x = nullptr;
    float *m_loss = nullptr, *m_loss_host = nullptr;
    int m_adam_t = 0;
};

}  // namespace artss
