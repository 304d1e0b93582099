#include "unet.h"

#include <cmath>
#include <random>
#include <stdexcept>

#include "domain.h"

namespace artss {

namespace {

struct Spec {
    int ci, co, level;
};
// layer order of the parameter vector
const Spec kLayers[] = {{2, 16, 0},  {16, 16, 0}, {16, 32, 1}, {32, 64, 2}, {64, 64, 3},
                        {64, 64, 2}, {64, 32, 1}, {32, 16, 0}, {16, 1, 0}};
enum { L_IN, L_E0, L_E1, L_E2, L_E3, L_D2, L_D1, L_D0, L_OUT, NLAYERS };
// the layers that read a nearest-upsampled input
inline bool upsampling(size_t layer) { return layer == L_D2 || layer == L_D1 || layer == L_D0; }

template <typename T>
T *dalloc(size_t n, std::vector<void *> *owner = nullptr) {
    T *p = nullptr;
    ARTSS_HIP(hipMalloc(&p, sizeof(T) * (n ? n : 1)));
    ARTSS_HIP(hipMemset(p, 0, sizeof(T) * (n ? n : 1)));
    if (owner) owner->push_back(p);
    return p;
}

}  // namespace

size_t UNet::param_count() {
    size_t n = 0;
    for (const Spec &s : kLayers) n += 27 * s.ci * s.co + s.co;
    return n;
}

UNet::UNet(hipStream_t stream, unsigned long long seed) : m_s(stream) {
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<float> host;
    for (const Spec &s : kLayers) {
        Layer L;
        L.ci = s.ci;
        L.co = s.co;
        L.level = s.level;
        const size_t nw = 27 * static_cast<size_t>(s.ci) * s.co;
        L.W = dalloc<float>(nw);
        L.b = dalloc<float>(s.co);
        L.gW = dalloc<float>(nw);
        L.gb = dalloc<float>(s.co);
        L.mW = dalloc<float>(nw);
        L.vW = dalloc<float>(nw);
        L.mb = dalloc<float>(s.co);
        L.vb = dalloc<float>(s.co);
        if (s.ci >= 16 && s.co >= 16) {
            L.W_bf = dalloc<uint16_t>(nw);
            L.Wt_bf = dalloc<uint16_t>(nw);
        }
        if (upsampling(m_layers.size())) L.Wup_bf = dalloc<uint16_t>(64 * static_cast<size_t>(s.ci) * s.co);
        // He initialisation; the output layer starts at zero, so an untrained
        // network predicts no correction
        const bool out_layer = m_layers.size() + 1 == sizeof(kLayers) / sizeof(kLayers[0]);
        const double std_ = out_layer ? 0.0 : std::sqrt(2.0 / (27.0 * s.ci));
        host.resize(nw);
        for (auto &x : host) x = static_cast<float>(std_ * nd(rng));
        ARTSS_HIP(hipMemcpy(L.W, host.data(), sizeof(float) * nw, hipMemcpyHostToDevice));
        if (L.W_bf) nn::weights_bf16(m_s, static_cast<int>(nw), s.ci, s.co, L.W, L.W_bf, L.Wt_bf);
        if (L.Wup_bf) nn::weights_up_bf16(m_s, s.ci, s.co, L.W, L.Wup_bf);
        m_layers.push_back(L);
    }
    m_loss = dalloc<float>(1);
    ARTSS_HIP(hipHostMalloc(&m_loss_host, sizeof(float)));
    ARTSS_HIP(hipStreamSynchronize(m_s));
}

UNet::~UNet() {
    (void)hipStreamSynchronize(m_s);
    free_buffers();
    for (Layer &L : m_layers)
        for (void *p : {static_cast<void *>(L.W), static_cast<void *>(L.b), static_cast<void *>(L.gW),
                        static_cast<void *>(L.gb), static_cast<void *>(L.mW), static_cast<void *>(L.vW),
                        static_cast<void *>(L.mb), static_cast<void *>(L.vb), static_cast<void *>(L.W_bf),
                        static_cast<void *>(L.Wt_bf), static_cast<void *>(L.Wup_bf)})
            if (p) (void)hipFree(p);
    (void)hipFree(m_loss);
    (void)hipHostFree(m_loss_host);
}

void UNet::set_params(const float *host) {
    size_t off = 0;
    for (Layer &L : m_layers) {
        const size_t nw = 27 * static_cast<size_t>(L.ci) * L.co;
        ARTSS_HIP(hipMemcpy(L.W, host + off, sizeof(float) * nw, hipMemcpyHostToDevice));
        off += nw;
        ARTSS_HIP(hipMemcpy(L.b, host + off, sizeof(float) * L.co, hipMemcpyHostToDevice));
        off += L.co;
        if (L.W_bf) nn::weights_bf16(m_s, static_cast<int>(nw), L.ci, L.co, L.W, L.W_bf, L.Wt_bf);
        if (L.Wup_bf) nn::weights_up_bf16(m_s, L.ci, L.co, L.W, L.Wup_bf);
        ARTSS_HIP(hipMemset(L.mW, 0, sizeof(float) * nw));
        ARTSS_HIP(hipMemset(L.vW, 0, sizeof(float) * nw));
        ARTSS_HIP(hipMemset(L.mb, 0, sizeof(float) * L.co));
        ARTSS_HIP(hipMemset(L.vb, 0, sizeof(float) * L.co));
    }
    m_adam_t = 0;
    ARTSS_HIP(hipStreamSynchronize(m_s));
}

void UNet::get_params(float *host) const {
    ARTSS_HIP(hipStreamSynchronize(m_s));
    size_t off = 0;
    for (const Layer &L : m_layers) {
        const size_t nw = 27 * static_cast<size_t>(L.ci) * L.co;
        ARTSS_HIP(hipMemcpy(host + off, L.W, sizeof(float) * nw, hipMemcpyDeviceToHost));
        off += nw;
        ARTSS_HIP(hipMemcpy(host + off, L.b, sizeof(float) * L.co, hipMemcpyDeviceToHost));
        off += L.co;
    }
}

void UNet::free_buffers() {
    for (void *p : m_buffers) (void)hipFree(p);
    m_buffers.clear();
    m_X = m_Y = m_Z = 0;
}

void UNet::ensure(int X, int Y, int Z) {
    if (X == m_X && Y == m_Y && Z == m_Z) return;
    if (X <= 0 || Y <= 0 || Z <= 0 || X % 8 || Y % 8 || Z % 8)
        throw std::invalid_argument("U-Net volume sides must be positive multiples of 8");
    ARTSS_HIP(hipStreamSynchronize(m_s));
    free_buffers();
    const size_t V0 = static_cast<size_t>(X) * Y * Z, V1 = V0 / 8, V2 = V1 / 8, V3 = V2 / 8;
    auto b16 = [&](size_t n) { return dalloc<uint16_t>(n, &m_buffers); };
    auto f32 = [&](size_t n) { return dalloc<float>(n, &m_buffers); };
    h0 = b16(V0 * 16);
    e0 = b16(V0 * 16);
    p1 = b16(V1 * 16);
    e1 = b16(V1 * 32);
    p2 = b16(V2 * 32);
    e2 = b16(V2 * 64);
    p3 = b16(V3 * 64);
    e3 = b16(V3 * 64);
    c2 = b16(V2 * 64);
    d2 = b16(V2 * 64);
    c1 = b16(V1 * 32);
    d1 = b16(V1 * 32);
    c0 = b16(V0 * 16);
    d0 = b16(V0 * 16);
    y = f32(V0);
    gy = f32(V0);
    g0a = f32(V0 * 32);
    g0b = f32(V0 * 16);
    g0c = f32(V0 * 16);
    g1a = f32(V1 * 64);
    g1b = f32(V1 * 32);
    g1c = f32(V1 * 32);
    g2a = f32(V2 * 64);
    g2b = f32(V2 * 64);
    g2c = f32(V2 * 64);
    g3a = f32(V3 * 64);
    g3b = f32(V3 * 64);
    m_X = X;
    m_Y = Y;
    m_Z = Z;
}

void UNet::run_forward(const float *x, bool keep_pre) {
    const nn::Dims D0{m_X, m_Y, m_Z}, D1{m_X / 2, m_Y / 2, m_Z / 2}, D2{m_X / 4, m_Y / 4, m_Z / 4},
        D3{m_X / 8, m_Y / 8, m_Z / 8};
    const Layer *L = m_layers.data();
    nn::conv_in(m_s, D0, x, L[L_IN].W, L[L_IN].b, h0);
    nn::conv_fwd(m_s, 16, 16, false, D0, h0, L[L_E0].Wt_bf, L[L_E0].b, e0, nullptr, nullptr, true);
    nn::down(m_s, 16, D1, e0, p1);
    nn::conv_fwd(m_s, 16, 32, false, D1, p1, L[L_E1].Wt_bf, L[L_E1].b, e1, nullptr, nullptr, true);
    nn::down(m_s, 32, D2, e1, p2);
    nn::conv_fwd(m_s, 32, 64, false, D2, p2, L[L_E2].Wt_bf, L[L_E2].b, e2, nullptr, nullptr, true);
    nn::down(m_s, 64, D3, e2, p3);
    nn::conv_fwd(m_s, 64, 64, false, D3, p3, L[L_E3].Wt_bf, L[L_E3].b, e3, nullptr, nullptr, true);
    nn::conv_fwd(m_s, 64, 64, true, D2, e3, L[L_D2].Wup_bf, L[L_D2].b, d2, keep_pre ? c2 : nullptr, e2, true);
    nn::conv_fwd(m_s, 64, 32, true, D1, d2, L[L_D1].Wup_bf, L[L_D1].b, d1, keep_pre ? c1 : nullptr, e1, true);
    nn::conv_fwd(m_s, 32, 16, true, D0, d1, L[L_D0].Wup_bf, L[L_D0].b, d0, keep_pre ? c0 : nullptr, e0, true);
    nn::conv_out(m_s, D0, d0, L[L_OUT].W, L[L_OUT].b, y);
    m_x = x;
}

void UNet::forward(int X, int Y, int Z, const float *x, float *out) {
    ensure(X, Y, Z);
    run_forward(x, false);
    ARTSS_HIP(hipMemcpyAsync(out, y, sizeof(float) * static_cast<size_t>(X) * Y * Z, hipMemcpyDeviceToDevice, m_s));
    ARTSS_HIP(hipStreamSynchronize(m_s));
}

float UNet::gradients(int X, int Y, int Z, const float *x, const float *target, float *grads) {
    ensure(X, Y, Z);
    run_forward(x, true);
    const nn::Dims D0{m_X, m_Y, m_Z}, D1{m_X / 2, m_Y / 2, m_Z / 2}, D2{m_X / 4, m_Y / 4, m_Z / 4},
        D3{m_X / 8, m_Y / 8, m_Z / 8};
    const long long V0 = static_cast<long long>(m_X) * m_Y * m_Z, V1 = V0 / 8, V2 = V1 / 8, V3 = V2 / 8;
    Layer *L = m_layers.data();
    for (Layer &l : m_layers) {
        nn::fill_f(m_s, 27LL * l.ci * l.co, l.gW, 0.f);
        nn::fill_f(m_s, l.co, l.gb, 0.f);
    }
    nn::fill_f(m_s, 1, m_loss, 0.f);
    nn::mse(m_s, V0, y, target, gy, m_loss);
    // output layer and the d0 branch
    nn::conv_out_grads(m_s, D0, d0, L[L_OUT].W, gy, L[L_OUT].gW, L[L_OUT].gb, g0b);
    nn::copy_f(m_s, V0 * 16, g0c, g0b);  // skip gradient into e0
    nn::relu_mask(m_s, V0 * 16, g0b, c0);
    nn::conv_wgrad(m_s, 32, 16, true, D0, d1, g0b, L[L_D0].gW);
    nn::bias_grad(m_s, 16, V0, g0b, L[L_D0].gb);
    nn::conv_dgrad(m_s, 32, 16, D0, g0b, L[L_D0].W_bf, g0a);
    nn::up_backward(m_s, 32, D1, g0a, g1b);
    // d1
    nn::copy_f(m_s, V1 * 32, g1c, g1b);
    nn::relu_mask(m_s, V1 * 32, g1b, c1);
    nn::conv_wgrad(m_s, 64, 32, true, D1, d2, g1b, L[L_D1].gW);
    nn::bias_grad(m_s, 32, V1, g1b, L[L_D1].gb);
    nn::conv_dgrad(m_s, 64, 32, D1, g1b, L[L_D1].W_bf, g1a);
    nn::up_backward(m_s, 64, D2, g1a, g2b);
    // d2
    nn::copy_f(m_s, V2 * 64, g2c, g2b);
    nn::relu_mask(m_s, V2 * 64, g2b, c2);
    nn::conv_wgrad(m_s, 64, 64, true, D2, e3, g2b, L[L_D2].gW);
    nn::bias_grad(m_s, 64, V2, g2b, L[L_D2].gb);
    nn::conv_dgrad(m_s, 64, 64, D2, g2b, L[L_D2].W_bf, g2a);
    nn::up_backward(m_s, 64, D3, g2a, g3a);
    // encoder, coarsest first
    nn::relu_mask(m_s, V3 * 64, g3a, e3);
    nn::conv_wgrad(m_s, 64, 64, false, D3, p3, g3a, L[L_E3].gW);
    nn::bias_grad(m_s, 64, V3, g3a, L[L_E3].gb);
    nn::conv_dgrad(m_s, 64, 64, D3, g3a, L[L_E3].W_bf, g3b);
    nn::down_backward_add(m_s, 64, D3, g3b, g2c);
    nn::relu_mask(m_s, V2 * 64, g2c, e2);
    nn::conv_wgrad(m_s, 32, 64, false, D2, p2, g2c, L[L_E2].gW);
    nn::bias_grad(m_s, 64, V2, g2c, L[L_E2].gb);
    nn::conv_dgrad(m_s, 32, 64, D2, g2c, L[L_E2].W_bf, g2a);
    nn::down_backward_add(m_s, 32, D2, g2a, g1c);
    nn::relu_mask(m_s, V1 * 32, g1c, e1);
    nn::conv_wgrad(m_s, 16, 32, false, D1, p1, g1c, L[L_E1].gW);
    nn::bias_grad(m_s, 32, V1, g1c, L[L_E1].gb);
    nn::conv_dgrad(m_s, 16, 32, D1, g1c, L[L_E1].W_bf, g1a);
    nn::down_backward_add(m_s, 16, D1, g1a, g0c);
    nn::relu_mask(m_s, V0 * 16, g0c, e0);
    nn::conv_wgrad(m_s, 16, 16, false, D0, h0, g0c, L[L_E0].gW);
    nn::bias_grad(m_s, 16, V0, g0c, L[L_E0].gb);
    nn::conv_dgrad(m_s, 16, 16, D0, g0c, L[L_E0].W_bf, g0a);
    nn::relu_mask(m_s, V0 * 16, g0a, h0);
    nn::conv_in_wgrad(m_s, D0, m_x, g0a, L[L_IN].gW, L[L_IN].gb);
    ARTSS_HIP(hipMemcpyAsync(m_loss_host, m_loss, sizeof(float), hipMemcpyDeviceToHost, m_s));
    ARTSS_HIP(hipStreamSynchronize(m_s));
    if (grads) {
        size_t off = 0;
        for (const Layer &l : m_layers) {
            const size_t nw = 27 * static_cast<size_t>(l.ci) * l.co;
            ARTSS_HIP(hipMemcpy(grads + off, l.gW, sizeof(float) * nw, hipMemcpyDeviceToHost));
            off += nw;
            ARTSS_HIP(hipMemcpy(grads + off, l.gb, sizeof(float) * l.co, hipMemcpyDeviceToHost));
            off += l.co;
        }
    }
    return *m_loss_host;
}

void UNet::adam_step(float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    m_adam_t++;
    const float c1 = 1.f - std::pow(b1, static_cast<float>(m_adam_t));
    const float c2 = 1.f - std::pow(b2, static_cast<float>(m_adam_t));
    for (Layer &l : m_layers) {
        const int nw = 27 * l.ci * l.co;
        nn::adam(m_s, nw, l.ci, l.co, l.W, l.gW, l.mW, l.vW, lr, b1, b2, eps, c1, c2, l.W_bf, l.Wt_bf);
        if (l.Wup_bf) nn::weights_up_bf16(m_s, l.ci, l.co, l.W, l.Wup_bf);
        nn::adam(m_s, l.co, 1, l.co, l.b, l.gb, l.mb, l.vb, lr, b1, b2, eps, c1, c2, nullptr, nullptr);
    }
}

}  // namespace artss
