// Kernels of the pressure warm-start U-Net (unet.{h,cpp}).
//
// Tensors are channels-last (voxel-major, channel innermost) at the
// resolution of their level, without ghost cells.  Activations and weights
// are BF16 (raw bit patterns, round-to-nearest-even), gradients FP32, and
// every product accumulates in FP32.  The 3x3x3 convolutions with 16 or more
// channels on both sides are implicit GEMMs on the gfx950 matrix cores.
// The forward pass (k_conv_fwd, the one the warm start runs every step) uses
// v_mfma_f32_16x16x32_bf16 with a register tile per wavefront (see there);
// the gradient kernels use v_mfma_f32_16x16x16_bf16: a wavefront owns 16
// consecutive voxels along x and all output channels; per tap and 16-channel
// chunk it loads one 16x16 BF16 operand of the input (8 bytes per lane,
// contiguous in the channel) and one of the weights per output tile.
// Operand layouts of the 16x16x16 form:
//   A (16x16, rows m, cols k): lane l holds A[l%16][4*(l/16) .. +3]
//   B (16x16, rows k, cols n): lane l holds B[4*(l/16) .. +3][l%16]
//   D (16x16 FP32):            lane l holds D[4*(l/16) .. +3][l%16]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>

#include "unet_kernels.h"

namespace artss {
namespace nn {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline float bf2f(uint16_t b) { return __uint_as_float(static_cast<unsigned>(b) << 16); }
__device__ inline uint16_t f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return static_cast<uint16_t>(u >> 16);
}
__device__ inline short f2bfs(float f) { return static_cast<short>(f2bf(f)); }

__device__ inline long long vox(const Dims &D, int x, int y, int z) {
    return (static_cast<long long>(z) * D.Y + y) * D.X + x;
}

__device__ inline void tap_offset(int t, int &dx, int &dy, int &dz) {
    dz = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dx = t % 3 - 1;
}

// ---------------------------------------------------------------- forward
// out[v][co] = act(b[co] + sum_t sum_ci in[v + off_t][ci] * W[t][ci][co]) (+ skip[v][co])
// Wt is [27][COUT][CIN].  UP: the input lives at half resolution and is read
// nearest-upsampled.  pre (optional): the pre-activation values.
//
// UP is evaluated on the coarse input directly: along each axis a fine
// voxel of parity p reads coarse offsets floor((p + d) / 2) for its taps
// d = -1, 0, 1, which are p - 1 and p, and a tap outside the fine volume is
// exactly a coarse offset outside the coarse one.  So the 27-tap
// convolution of the upsampled input equals, per parity (px, py, pz), a
// 2x2x2 convolution of the coarse input with the weights of the fine taps
// that meet on each coarse tap summed (weights_up_bf16: Wt is then
// [8 parities][8 taps][COUT][CIN]).  A wavefront's voxel tiles each hold 16
// voxels of one x parity (every other voxel), so the weight operand is the
// same for all columns of a tile and the tile reads 16 consecutive coarse
// voxels; y and z parities are those of the wavefront's row and plane.
//
// An implicit GEMM D[co][v] += sum_k A[co][k] B[k][v] over k = (tap, ci) on
// v_mfma_f32_16x16x32_bf16.  A wavefront owns one row of NV x 16 consecutive
// voxels along x and all COUT / 16 output-channel tiles, so per K step of 32
// it loads COUT / 16 weight operands and NV input operands (16 bytes per
// lane each, contiguous in the channel) and issues (COUT / 16) x NV matrix
// instructions on them.  A K step is 32 channels of one tap (CIN >= 32) or
// 16 channels of two consecutive taps (CIN = 16, the 28th tap reads zero).
// Operand layouts (both operands give lane l the same eight k, so their
// order within the lane does not change the sum):
//   A (16x32): lane l holds row l%16, k = 8*(l/16) .. +7
//   B (32x16): lane l holds column l%16, k = 8*(l/16) .. +7
//   D (16x16 FP32): lane l holds D[4*(l/16) .. +3][l%16]
// The epilogue stores the four channels a lane holds as one 8-byte write.
template <int CIN, int COUT, bool UP, int NV>
__global__ void __launch_bounds__(256) k_conv_fwd(Dims D, const uint16_t *__restrict__ in,
                                                  const uint16_t *__restrict__ Wt, const float *__restrict__ bias,
                                                  uint16_t *__restrict__ out, uint16_t *__restrict__ pre,
                                                  const uint16_t *__restrict__ skip, int relu) {
    static_assert(CIN % 16 == 0 && COUT % 16 == 0, "channels come in tiles of 16");
    static_assert(!UP || (CIN >= 32 && NV % 2 == 0), "the upsampling form takes 32-channel K steps, tile pairs");
    constexpr int MT = COUT / 16;
    constexpr int KSTEPS = UP ? 8 * (CIN / 32) : (CIN >= 32 ? 27 * (CIN / 32) : 14);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int y = static_cast<int>(blockIdx.y) * 4 + wave, z = blockIdx.z;
    const int x0 = static_cast<int>(blockIdx.x) * (16 * NV);
    if (y >= D.Y) return;  // no workgroup barrier in this kernel
    const int col = lane & 15, kg = lane >> 4;
    const Dims C{D.X >> 1, D.Y >> 1, D.Z >> 1};
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, uint4{0u, 0u, 0u, 0u});
    f32x4 acc[MT][NV];
#pragma unroll
    for (int ct = 0; ct < MT; ct++)
#pragma unroll
        for (int j = 0; j < NV; j++) acc[ct][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (UP) {
        const int py = y & 1, pz = z & 1;
        for (int st = 0; st < KSTEPS; st++) {
            const int t8 = st / (CIN / 32), c = (st % (CIN / 32)) * 32 + 8 * kg;
            const int ox = t8 & 1, oy = (t8 >> 1) & 1, oz = t8 >> 2;
            bf16x8 a[2][MT], b[NV];
#pragma unroll
            for (int px = 0; px < 2; px++)
#pragma unroll
                for (int ct = 0; ct < MT; ct++)
                    a[px][ct] = __builtin_bit_cast(
                        bf16x8, *reinterpret_cast<const uint4 *>(
                                    Wt + ((static_cast<long long>(((pz * 2 + py) * 2 + px) * 8 + t8)) * COUT + ct * 16 +
                                          col) * CIN + c));
            const int cy = (y >> 1) + py - 1 + oy, cz = (z >> 1) + pz - 1 + oz;
            const bool row_ok = cy >= 0 && cy < C.Y && cz >= 0 && cz < C.Z;
#pragma unroll
            for (int j = 0; j < NV; j++) {
                const int cx = x0 / 2 + 16 * (j / 2) + col + (j & 1) - 1 + ox;
                b[j] = zero8;
                if (row_ok && cx >= 0 && cx < C.X)
                    b[j] = __builtin_bit_cast(bf16x8,
                                              *reinterpret_cast<const uint4 *>(in + vox(C, cx, cy, cz) * CIN + c));
            }
#pragma unroll
            for (int ct = 0; ct < MT; ct++)
#pragma unroll
                for (int j = 0; j < NV; j++)
                    acc[ct][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[j & 1][ct], b[j], acc[ct][j], 0, 0, 0);
        }
    }
    for (int st = 0; st < (UP ? 0 : KSTEPS); st++) {
        // tap and first channel of this lane's eight k values
        int t, c;
        if (CIN >= 32) {
            t = st / (CIN / 32);
            c = (st % (CIN / 32)) * 32 + 8 * kg;
        } else {
            t = 2 * st + (kg >> 1);
            c = 8 * (kg & 1);
        }
        const bool tap_ok = t < 27;
        int dx, dy, dz;
        tap_offset(tap_ok ? t : 0, dx, dy, dz);
        bf16x8 a[MT], b[NV];
#pragma unroll
        for (int ct = 0; ct < MT; ct++)
            a[ct] = tap_ok ? __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(
                                                            Wt + (static_cast<long long>(t) * COUT + ct * 16 + col) *
                                                                     CIN + c))
                           : zero8;
        const int yy = y + dy, zz = z + dz;
        const bool row_ok = tap_ok && yy >= 0 && yy < D.Y && zz >= 0 && zz < D.Z;
#pragma unroll
        for (int j = 0; j < NV; j++) {
            const int xx = x0 + 16 * j + col + dx;
            b[j] = zero8;
            if (row_ok && xx >= 0 && xx < D.X)
                b[j] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(in + vox(D, xx, yy, zz) * CIN + c));
        }
#pragma unroll
        for (int ct = 0; ct < MT; ct++)
#pragma unroll
            for (int j = 0; j < NV; j++)
                acc[ct][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ct], b[j], acc[ct][j], 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < NV; j++) {
        const int vx = UP ? x0 + 32 * (j / 2) + 2 * col + (j & 1) : x0 + 16 * j + col;
        if (vx >= D.X) continue;
        const long long v = vox(D, vx, y, z);
#pragma unroll
        for (int ct = 0; ct < MT; ct++) {
            const int co = ct * 16 + 4 * kg;
            float c4[4];
#pragma unroll
            for (int e = 0; e < 4; e++) c4[e] = acc[ct][j][e] + bias[co + e];
            if (pre) {
                uint2 pk;
                pk.x = static_cast<unsigned>(f2bf(c4[0])) | (static_cast<unsigned>(f2bf(c4[1])) << 16);
                pk.y = static_cast<unsigned>(f2bf(c4[2])) | (static_cast<unsigned>(f2bf(c4[3])) << 16);
                *reinterpret_cast<uint2 *>(pre + v * COUT + co) = pk;
            }
            float r[4];
#pragma unroll
            for (int e = 0; e < 4; e++) r[e] = relu ? fmaxf(c4[e], 0.f) : c4[e];
            if (skip) {
                const uint2 sk = *reinterpret_cast<const uint2 *>(skip + v * COUT + co);
                r[0] += bf2f(static_cast<uint16_t>(sk.x & 0xFFFFu));
                r[1] += bf2f(static_cast<uint16_t>(sk.x >> 16));
                r[2] += bf2f(static_cast<uint16_t>(sk.y & 0xFFFFu));
                r[3] += bf2f(static_cast<uint16_t>(sk.y >> 16));
            }
            uint2 o;
            o.x = static_cast<unsigned>(f2bf(r[0])) | (static_cast<unsigned>(f2bf(r[1])) << 16);
            o.y = static_cast<unsigned>(f2bf(r[2])) | (static_cast<unsigned>(f2bf(r[3])) << 16);
            *reinterpret_cast<uint2 *>(out + v * COUT + co) = o;
        }
    }
}

// First layer: x is FP32 with CI (2) channels, 16 BF16 outputs with ReLU.
// One voxel per thread; the 27 taps are unrolled with zero operands outside
// the volume, so their loads are in flight together, and the 16 outputs
// leave as two 16-byte stores.
template <int CI, int CO>
__global__ void k_conv_in(Dims D, const float *__restrict__ x, const float *__restrict__ W,
                          const float *__restrict__ bias, uint16_t *__restrict__ out) {
    static_assert(CI == 2 && CO == 16, "two input channels, sixteen outputs");
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (i >= D.X) return;
    float2 xv[27];
#pragma unroll
    for (int t = 0; t < 27; t++) {
        int dx, dy, dz;
        tap_offset(t, dx, dy, dz);
        const int xx = i + dx, yy = y + dy, zz = z + dz;
        const bool inside = xx >= 0 && xx < D.X && yy >= 0 && yy < D.Y && zz >= 0 && zz < D.Z;
        xv[t] = inside ? *reinterpret_cast<const float2 *>(x + vox(D, xx, yy, zz) * CI) : make_float2(0.f, 0.f);
    }
    float acc[CO];
#pragma unroll
    for (int c = 0; c < CO; c++) acc[c] = bias[c];
#pragma unroll
    for (int t = 0; t < 27; t++)
#pragma unroll
        for (int c = 0; c < CO; c++) {
            acc[c] += xv[t].x * W[(t * CI + 0) * CO + c];
            acc[c] += xv[t].y * W[(t * CI + 1) * CO + c];
        }
    unsigned pk[CO / 2];
#pragma unroll
    for (int c = 0; c < CO; c += 2)
        pk[c / 2] = static_cast<unsigned>(f2bf(fmaxf(acc[c], 0.f))) |
                    (static_cast<unsigned>(f2bf(fmaxf(acc[c + 1], 0.f))) << 16);
    uint4 *o = reinterpret_cast<uint4 *>(out + vox(D, i, y, z) * CO);
    o[0] = uint4{pk[0], pk[1], pk[2], pk[3]};
    o[1] = uint4{pk[4], pk[5], pk[6], pk[7]};
}

// Last layer: CI BF16 channels -> one FP32 output, no activation.  A thread
// owns NZ voxels of one (x, y) column and marches through the input planes
// z0-1 .. z0+NZ: each tap value it loads (16 bytes per eight channels) is
// added, with the weight of its z offset, to each of the up to three
// outputs whose stencil holds that plane.  Every output still receives its
// terms in tap order (z offset, then y, x, channel), skipping taps outside
// the volume, as a per-voxel loop over the 27 taps adds them.
template <int CI, int NZ>
__global__ void k_conv_out(Dims D, const uint16_t *__restrict__ in, const float *__restrict__ W,
                           const float *__restrict__ bias, float *__restrict__ y) {
    static_assert(CI % 8 == 0, "channels in groups of eight");
    const int i = static_cast<int>(blockIdx.x) * 64 + static_cast<int>(threadIdx.x);
    const int j = static_cast<int>(blockIdx.y) * 4 + static_cast<int>(threadIdx.y);
    const int z0 = static_cast<int>(blockIdx.z) * NZ;
    if (i >= D.X || j >= D.Y) return;
    float acc[NZ];
#pragma unroll
    for (int q = 0; q < NZ; q++) acc[q] = bias[0];
#pragma unroll
    for (int r = 0; r < NZ + 2; r++) {
        const int zp = z0 - 1 + r;  // input plane; feeds outputs q = r (dz -1), r - 1 (dz 0), r - 2 (dz +1)
        if (zp < 0 || zp >= D.Z) continue;
#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
            const int yy = j + dy;
            if (yy < 0 || yy >= D.Y) continue;
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                const int xx = i + dx;
                if (xx < 0 || xx >= D.X) continue;
                const uint4 *src = reinterpret_cast<const uint4 *>(in + vox(D, xx, yy, zp) * CI);
                const int tb = (dy + 1) * 3 + (dx + 1);  // tap within its z plane
#pragma unroll
                for (int g = 0; g < CI / 8; g++) {
                    const uint4 qv = src[g];
                    const unsigned w4[4] = {qv.x, qv.y, qv.z, qv.w};
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        const float v = bf2f(static_cast<uint16_t>((w4[e >> 1] >> (16 * (e & 1))) & 0xFFFFu));
                        const int ci = 8 * g + e;
                        if (r < NZ) acc[r] += v * W[tb * CI + ci];
                        if (r >= 1 && r - 1 < NZ) acc[r - 1] += v * W[(9 + tb) * CI + ci];
                        if (r >= 2) acc[r - 2] += v * W[(18 + tb) * CI + ci];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NZ; q++)
        if (z0 + q < D.Z) y[vox(D, i, j, z0 + q)] = acc[q];
}

// 2x2x2 mean, D: the output (coarse) dims
template <int C>
__global__ void k_down(Dims D, const uint16_t *__restrict__ in, uint16_t *__restrict__ out) {
    const long long n = static_cast<long long>(D.X) * D.Y * D.Z * C;
    const Dims F{D.X * 2, D.Y * 2, D.Z * 2};
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += 256LL * gridDim.x) {
        const int c = static_cast<int>(q % C);
        const long long v = q / C;
        const int x = static_cast<int>(v % D.X), y = static_cast<int>((v / D.X) % D.Y),
                  z = static_cast<int>(v / (static_cast<long long>(D.X) * D.Y));
        float s = 0.f;
        for (int o = 0; o < 8; o++) s += bf2f(in[vox(F, 2 * x + (o & 1), 2 * y + ((o >> 1) & 1), 2 * z + (o >> 2)) * C + c]);
        out[q] = f2bf(s * 0.125f);
    }
}

// ---------------------------------------------------------------- backward
// gin[v][ci] = sum_t sum_co gout[v - off_t][co] * W[t][ci][co] (zero outside);
// W is [27][CIN][COUT]; gout FP32 (rounded to BF16 as the matrix operand).
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv_dgrad(Dims D, const float *__restrict__ gout,
                                                    const uint16_t *__restrict__ W, float *__restrict__ gin) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x0 = (blockIdx.x * 4 + wave) * 16;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x0 >= D.X) return;
    const int vx = x0 + (lane & 15), kq = lane >> 4;
    f32x4 acc[CIN / 16];
#pragma unroll
    for (int ct = 0; ct < CIN / 16; ct++) acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 27; t++) {
        int dx, dy, dz;
        tap_offset(t, dx, dy, dz);
        const int xx = vx - dx, yy = y - dy, zz = z - dz;
        const bool inside = vx < D.X && xx >= 0 && xx < D.X && yy >= 0 && yy < D.Y && zz >= 0 && zz < D.Z;
        const long long src = inside ? vox(D, xx, yy, zz) : 0;
#pragma unroll
        for (int c0 = 0; c0 < COUT; c0 += 16) {
            bf16x4 b = bf16x4{0, 0, 0, 0};
            if (inside) {
                const float *g = gout + src * COUT + c0 + 4 * kq;
                b = bf16x4{f2bfs(g[0]), f2bfs(g[1]), f2bfs(g[2]), f2bfs(g[3])};
            }
#pragma unroll
            for (int ct = 0; ct < CIN / 16; ct++) {
                const bf16x4 a = *reinterpret_cast<const bf16x4 *>(
                    W + (static_cast<long long>(t) * CIN + ct * 16 + (lane & 15)) * COUT + c0 + 4 * kq);
                acc[ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[ct], 0, 0, 0);
            }
        }
    }
    if (vx >= D.X) return;
    const long long v = vox(D, vx, y, z);
#pragma unroll
    for (int ct = 0; ct < CIN / 16; ct++)
#pragma unroll
        for (int e = 0; e < 4; e++) gin[v * CIN + ct * 16 + 4 * kq + e] = acc[ct][e];
}

// gW[t][ci][co] += sum_v in[v + off_t][ci] * gout[v][co] over this
// workgroup's x-lines (grid: line blocks x 27 taps).  Each wavefront
// accumulates its lines' 16-voxel chunks as matrix products with the voxels
// as the reduction index, then adds its tiles to gW with FP32 atomics.
// UP: the forward read `in` nearest-upsampled from half resolution.
template <int CIN, int COUT, bool UP>
__global__ void __launch_bounds__(256) k_conv_wgrad(Dims D, const uint16_t *__restrict__ in,
                                                    const float *__restrict__ gout, float *__restrict__ gW,
                                                    int lines_per_block) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = blockIdx.y;
    int dx, dy, dz;
    tap_offset(t, dx, dy, dz);
    const int kq = lane >> 4, m = lane & 15;
    f32x4 acc[COUT / 16][CIN / 16];
#pragma unroll
    for (int a = 0; a < COUT / 16; a++)
#pragma unroll
        for (int b = 0; b < CIN / 16; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int lines = D.Y * D.Z;
    const int chunks = (D.X + 15) / 16;
    const int first = blockIdx.x * lines_per_block;
    const int last = min(lines, first + lines_per_block);
    for (int line = first + wave; line < last; line += 4) {
        const int y = line % D.Y, z = line / D.Y;
        for (int ch = 0; ch < chunks; ch++) {
            // reduction index k = voxel 4*kq + e of this chunk
            bf16x4 ga[COUT / 16], ib[CIN / 16];
#pragma unroll
            for (int a = 0; a < COUT / 16; a++) ga[a] = bf16x4{0, 0, 0, 0};
#pragma unroll
            for (int b = 0; b < CIN / 16; b++) ib[b] = bf16x4{0, 0, 0, 0};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int x = ch * 16 + 4 * kq + e;
                if (x >= D.X) continue;
                const long long v = vox(D, x, y, z);
#pragma unroll
                for (int a = 0; a < COUT / 16; a++) ga[a][e] = f2bfs(gout[v * COUT + a * 16 + m]);
                const int xx = x + dx, yy = y + dy, zz = z + dz;
                if (xx < 0 || xx >= D.X || yy < 0 || yy >= D.Y || zz < 0 || zz >= D.Z) continue;
                long long s;
                if (UP) {
                    const Dims C{D.X >> 1, D.Y >> 1, D.Z >> 1};
                    s = vox(C, xx >> 1, yy >> 1, zz >> 1);
                } else {
                    s = vox(D, xx, yy, zz);
                }
#pragma unroll
                for (int b = 0; b < CIN / 16; b++) ib[b][e] = static_cast<short>(in[s * CIN + b * 16 + m]);
            }
#pragma unroll
            for (int a = 0; a < COUT / 16; a++)
#pragma unroll
                for (int b = 0; b < CIN / 16; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ga[a], ib[b], acc[a][b], 0, 0, 0);
        }
    }
    // D[m = co][n = ci]: lane holds co = a*16 + 4*kq + e, ci = b*16 + m
#pragma unroll
    for (int a = 0; a < COUT / 16; a++)
#pragma unroll
        for (int b = 0; b < CIN / 16; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int co = a * 16 + 4 * kq + e, ci = b * 16 + m;
                atomicAdd(&gW[(static_cast<long long>(t) * CIN + ci) * COUT + co], acc[a][b][e]);
            }
}

// weight gradients of the first layer (CI FP32 inputs -> CO outputs):
// one thread per weight entry (t, ci, co) sums x[v + off_t][ci] * g[v][co]
// over this workgroup's voxel range, then adds it to gW atomically.
template <int CI, int CO>
__global__ void k_conv_in_wgrad(Dims D, const float *__restrict__ x, const float *__restrict__ g,
                                float *__restrict__ gW, long long voxels_per_block) {
    const int q = threadIdx.x;
    if (q >= 27 * CI * CO) return;
    const int t = q / (CI * CO), ci = (q / CO) % CI, co = q % CO;
    int dx, dy, dz;
    tap_offset(t, dx, dy, dz);
    const long long n = static_cast<long long>(D.X) * D.Y * D.Z;
    const long long v0 = blockIdx.x * voxels_per_block, v1 = min(n, v0 + voxels_per_block);
    float acc = 0.f;
    for (long long v = v0; v < v1; v++) {
        const int i = static_cast<int>(v % D.X), y = static_cast<int>((v / D.X) % D.Y),
                  z = static_cast<int>(v / (static_cast<long long>(D.X) * D.Y));
        const int xx = i + dx, yy = y + dy, zz = z + dz;
        if (xx < 0 || xx >= D.X || yy < 0 || yy >= D.Y || zz < 0 || zz >= D.Z) continue;
        acc += x[vox(D, xx, yy, zz) * CI + ci] * g[v * CO + co];
    }
    atomicAdd(&gW[q], acc);
}

// gradients of the last layer (CI BF16 inputs -> 1 FP32 output, no activation):
// weights as above, one thread per (t, ci); d in by a separate pass
template <int CI>
__global__ void k_conv_out_wgrad(Dims D, const uint16_t *__restrict__ in, const float *__restrict__ gy,
                                 float *__restrict__ gW, long long voxels_per_block) {
    const int q = threadIdx.x;
    if (q >= 27 * CI) return;
    const int t = q / CI, ci = q % CI;
    int dx, dy, dz;
    tap_offset(t, dx, dy, dz);
    const long long n = static_cast<long long>(D.X) * D.Y * D.Z;
    const long long v0 = blockIdx.x * voxels_per_block, v1 = min(n, v0 + voxels_per_block);
    float acc = 0.f;
    for (long long v = v0; v < v1; v++) {
        const int i = static_cast<int>(v % D.X), y = static_cast<int>((v / D.X) % D.Y),
                  z = static_cast<int>(v / (static_cast<long long>(D.X) * D.Y));
        const int xx = i + dx, yy = y + dy, zz = z + dz;
        if (xx < 0 || xx >= D.X || yy < 0 || yy >= D.Y || zz < 0 || zz >= D.Z) continue;
        acc += bf2f(in[vox(D, xx, yy, zz) * CI + ci]) * gy[v];
    }
    atomicAdd(&gW[q], acc);
}
template <int CI>
__global__ void k_conv_out_dgrad(Dims D, const float *__restrict__ W, const float *__restrict__ gy,
                                 float *__restrict__ gin) {
    const long long n = static_cast<long long>(D.X) * D.Y * D.Z;
    for (long long v = blockIdx.x * 256LL + threadIdx.x; v < n; v += 256LL * gridDim.x) {
        const int i = static_cast<int>(v % D.X), y = static_cast<int>((v / D.X) % D.Y),
                  z = static_cast<int>(v / (static_cast<long long>(D.X) * D.Y));
        float acc[CI];
#pragma unroll
        for (int ci = 0; ci < CI; ci++) acc[ci] = 0.f;
        for (int t = 0; t < 27; t++) {
            int dx, dy, dz;
            tap_offset(t, dx, dy, dz);
            const int xx = i - dx, yy = y - dy, zz = z - dz;
            if (xx < 0 || xx >= D.X || yy < 0 || yy >= D.Y || zz < 0 || zz >= D.Z) continue;
            const float gs = gy[vox(D, xx, yy, zz)];
#pragma unroll
            for (int ci = 0; ci < CI; ci++) acc[ci] += gs * W[t * CI + ci];
        }
#pragma unroll
        for (int ci = 0; ci < CI; ci++) gin[v * CI + ci] = acc[ci];
    }
}

// bias gradient: gb[c] += sum over voxels of g[v][c]
template <int C>
__global__ void k_bias_grad(long long n, const float *__restrict__ g, float *__restrict__ gb) {
    __shared__ float part[C];
    for (int q = threadIdx.x; q < C; q += blockDim.x) part[q] = 0.f;
    __syncthreads();
    float acc[C];
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0.f;
    for (long long v = blockIdx.x * 256LL + threadIdx.x; v < n; v += 256LL * gridDim.x)
#pragma unroll
        for (int c = 0; c < C; c++) acc[c] += g[v * C + c];
#pragma unroll
    for (int c = 0; c < C; c++) atomicAdd(&part[c], acc[c]);
    __syncthreads();
    for (int q = threadIdx.x; q < C; q += blockDim.x) atomicAdd(&gb[q], part[q]);
}

// g *= (act > 0): the ReLU mask from stored BF16 activations
__global__ void k_relu_mask(long long n, float *__restrict__ g, const uint16_t *__restrict__ act) {
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += 256LL * gridDim.x)
        if (!(bf2f(act[q]) > 0.f)) g[q] = 0.f;
}

// dst = src (the skip gradient) ; dst += src
__global__ void k_copy_f(long long n, float *__restrict__ dst, const float *__restrict__ src) {
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += 256LL * gridDim.x) dst[q] = src[q];
}

// coarse gradient of the 2x2x2 mean: gc (coarse dims D) spread /8 onto gf (fine) and added
template <int C>
__global__ void k_down_backward_add(Dims D, const float *__restrict__ gc, float *__restrict__ gf) {
    const Dims F{D.X * 2, D.Y * 2, D.Z * 2};
    const long long n = static_cast<long long>(F.X) * F.Y * F.Z * C;
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += 256LL * gridDim.x) {
        const int c = static_cast<int>(q % C);
        const long long v = q / C;
        const int x = static_cast<int>(v % F.X), y = static_cast<int>((v / F.X) % F.Y),
                  z = static_cast<int>(v / (static_cast<long long>(F.X) * F.Y));
        gf[q] += 0.125f * gc[vox(D, x >> 1, y >> 1, z >> 1) * C + c];
    }
}

// gradient of nearest upsampling: gc[v][c] = sum of the 8 fine children (D: coarse dims)
template <int C>
__global__ void k_up_backward(Dims D, const float *__restrict__ gf, float *__restrict__ gc) {
    const Dims F{D.X * 2, D.Y * 2, D.Z * 2};
    const long long n = static_cast<long long>(D.X) * D.Y * D.Z * C;
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += 256LL * gridDim.x) {
        const int c = static_cast<int>(q % C);
        const long long v = q / C;
        const int x = static_cast<int>(v % D.X), y = static_cast<int>((v / D.X) % D.Y),
                  z = static_cast<int>(v / (static_cast<long long>(D.X) * D.Y));
        float s = 0.f;
        for (int o = 0; o < 8; o++) s += gf[vox(F, 2 * x + (o & 1), 2 * y + ((o >> 1) & 1), 2 * z + (o >> 2)) * C + c];
        gc[q] = s;
    }
}

// loss partials and dL/dy for the mean squared error over n voxels
__global__ void k_mse(long long n, const float *__restrict__ y, const float *__restrict__ target,
                      float *__restrict__ gy, float *__restrict__ loss) {
    float s = 0.f;
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += 256LL * gridDim.x) {
        const float d = y[q] - target[q];
        gy[q] = 2.f * d / static_cast<float>(n);
        s += d * d;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, s / static_cast<float>(n));
}

// Adam on FP32 master weights; refreshes the BF16 copies in both layouts
// ([t][ci][co] for the data gradient, [t][co][ci] for the forward pass).
__global__ void k_adam(int n, int ci, int co, float *__restrict__ w, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2, float eps,
                       float c1, float c2, uint16_t *__restrict__ w_bf, uint16_t *__restrict__ wt_bf) {
    for (int q = blockIdx.x * 256 + threadIdx.x; q < n; q += 256 * gridDim.x) {
        const float gq = g[q];
        const float mq = b1 * m[q] + (1.f - b1) * gq;
        const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
        m[q] = mq;
        v[q] = vq;
        const float wq = w[q] - lr * (mq / c1) / (sqrtf(vq / c2) + eps);
        w[q] = wq;
        if (w_bf) {
            const int t = q / (ci * co), r = q % (ci * co), i = r / co, o = r % co;
            w_bf[q] = f2bf(wq);
            wt_bf[(t * co + o) * ci + i] = f2bf(wq);
        }
    }
}

// BF16 copies of FP32 weights in both layouts
__global__ void k_weights_bf16(int n, int ci, int co, const float *__restrict__ w, uint16_t *__restrict__ w_bf,
                               uint16_t *__restrict__ wt_bf) {
    for (int q = blockIdx.x * 256 + threadIdx.x; q < n; q += 256 * gridDim.x) {
        const int t = q / (ci * co), r = q % (ci * co), i = r / co, o = r % co;
        w_bf[q] = f2bf(w[q]);
        wt_bf[(t * co + o) * ci + i] = f2bf(w[q]);
    }
}

// BF16 weights of the upsampling convolution's coarse form (k_conv_fwd):
// wup[par][t8][o][i] = sum of W[t][i][o] over the fine taps t that meet on
// coarse tap t8 for output parity par (bits x, y, z from the lowest), summed
// in FP32 in tap order, rounded once
__global__ void k_weights_up_bf16(int ci, int co, const float *__restrict__ w, uint16_t *__restrict__ wup) {
    const int n = 64 * ci * co;
    for (int q = blockIdx.x * 256 + threadIdx.x; q < n; q += 256 * gridDim.x) {
        const int i = q % ci, o = (q / ci) % co, t8 = (q / (ci * co)) % 8, par = q / (8 * ci * co);
        const int p[3] = {par & 1, (par >> 1) & 1, par >> 2}, oo[3] = {t8 & 1, (t8 >> 1) & 1, t8 >> 2};
        float acc = 0.f;
        for (int t = 0; t < 27; t++) {
            int d[3];
            tap_offset(t, d[0], d[1], d[2]);
            bool hit = true;
            for (int a = 0; a < 3; a++) {
                const int c = (p[a] + d[a] + 2) / 2 - 1;  // floor((p + d) / 2)
                hit = hit && c == p[a] - 1 + oo[a];
            }
            if (hit) acc += w[(t * ci + i) * co + o];
        }
        wup[q] = f2bf(acc);
    }
}

__global__ void k_fill_f(long long n, float *__restrict__ d, float v) {
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += 256LL * gridDim.x) d[q] = v;
}

inline unsigned blocks_for(long long n) {
    const long long b = (n + 255) / 256;
    return static_cast<unsigned>(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

}  // namespace

// ---------------------------------------------------------------- launchers
#define ARTSS_CONV_CASES(X) X(16, 16) X(16, 32) X(32, 64) X(64, 64) X(64, 32) X(32, 16)

void conv_fwd(hipStream_t s, int cin, int cout, bool up, Dims D, const uint16_t *in, const uint16_t *Wt,
              const float *bias, uint16_t *out, uint16_t *pre, const uint16_t *skip, bool relu) {
    // voxels per wavefront row: 16 x NV with NV = 8, 4, 2 for 16, 32, 64
    // output channels (8 matrix tiles per K step)
#define CASE(CI, CO)                                                                                             \
    if (cin == CI && cout == CO) {                                                                              \
        constexpr int NV = CO == 16 ? 8 : (CO == 32 ? 4 : 2);                                                    \
        const dim3 g((D.X + 16 * NV - 1) / (16 * NV), (D.Y + 3) / 4, D.Z);                                     \
        if (!up) {                                                                                              \
            k_conv_fwd<CI, CO, false, NV><<<g, 256, 0, s>>>(D, in, Wt, bias, out, pre, skip, relu ? 1 : 0);     \
            return;                                                                                             \
        }                                                                                                       \
        if constexpr (CI >= 32) {                                                                               \
            k_conv_fwd<CI, CO, true, NV><<<g, 256, 0, s>>>(D, in, Wt, bias, out, pre, skip, relu ? 1 : 0);      \
            return;                                                                                             \
        }                                                                                                       \
        throw std::invalid_argument("conv_fwd: an upsampled input needs at least 32 channels");               \
    }
    ARTSS_CONV_CASES(CASE)
#undef CASE
    throw std::invalid_argument("conv_fwd: unsupported channel pair");
}

void conv_dgrad(hipStream_t s, int cin, int cout, Dims D, const float *gout, const uint16_t *W, float *gin) {
    const dim3 g((D.X + 63) / 64, D.Y, D.Z);
#define CASE(CI, CO)                                                      \
    if (cin == CI && cout == CO) {                                       \
        k_conv_dgrad<CI, CO><<<g, 256, 0, s>>>(D, gout, W, gin);         \
        return;                                                          \
    }
    ARTSS_CONV_CASES(CASE)
#undef CASE
    throw std::invalid_argument("conv_dgrad: unsupported channel pair");
}

void conv_wgrad(hipStream_t s, int cin, int cout, bool up, Dims D, const uint16_t *in, const float *gout,
                float *gW) {
    const int lines = D.Y * D.Z;
    const int per = 16;
    const dim3 g((lines + per - 1) / per, 27);
#define CASE(CI, CO)                                                                            \
    if (cin == CI && cout == CO) {                                                             \
        if (up) k_conv_wgrad<CI, CO, true><<<g, 256, 0, s>>>(D, in, gout, gW, per);             \
        else k_conv_wgrad<CI, CO, false><<<g, 256, 0, s>>>(D, in, gout, gW, per);               \
        return;                                                                                \
    }
    ARTSS_CONV_CASES(CASE)
#undef CASE
    throw std::invalid_argument("conv_wgrad: unsupported channel pair");
}

void conv_in(hipStream_t s, Dims D, const float *x, const float *W, const float *bias, uint16_t *out) {
    k_conv_in<2, 16><<<dim3((D.X + 63) / 64, D.Y, D.Z), 64, 0, s>>>(D, x, W, bias, out);
}
void conv_out(hipStream_t s, Dims D, const uint16_t *in, const float *W, const float *bias, float *y) {
    k_conv_out<16, 8><<<dim3((D.X + 63) / 64, (D.Y + 3) / 4, (D.Z + 7) / 8), dim3(64, 4, 1), 0, s>>>(D, in, W, bias, y);
}
void conv_in_wgrad(hipStream_t s, Dims D, const float *x, const float *g, float *gW, float *gb) {
    const long long n = static_cast<long long>(D.X) * D.Y * D.Z;
    const long long per = 2048;
    k_conv_in_wgrad<2, 16><<<static_cast<unsigned>((n + per - 1) / per), 27 * 2 * 16, 0, s>>>(D, x, g, gW, per);
    bias_grad(s, 16, n, g, gb);
}
void conv_out_grads(hipStream_t s, Dims D, const uint16_t *in, const float *W, const float *gy, float *gW,
                    float *gb, float *gin) {
    const long long n = static_cast<long long>(D.X) * D.Y * D.Z;
    const long long per = 2048;
    k_conv_out_wgrad<16><<<static_cast<unsigned>((n + per - 1) / per), 27 * 16, 0, s>>>(D, in, gy, gW, per);
    k_conv_out_dgrad<16><<<blocks_for(n), 256, 0, s>>>(D, W, gy, gin);
    bias_grad(s, 1, n, gy, gb);
}
void down(hipStream_t s, int c, Dims Dc, const uint16_t *in, uint16_t *out) {
    const long long n = static_cast<long long>(Dc.X) * Dc.Y * Dc.Z * c;
    if (c == 16) k_down<16><<<blocks_for(n), 256, 0, s>>>(Dc, in, out);
    else if (c == 32) k_down<32><<<blocks_for(n), 256, 0, s>>>(Dc, in, out);
    else if (c == 64) k_down<64><<<blocks_for(n), 256, 0, s>>>(Dc, in, out);
    else throw std::invalid_argument("down: unsupported channels");
}
void down_backward_add(hipStream_t s, int c, Dims Dc, const float *gc, float *gf) {
    const long long n = static_cast<long long>(Dc.X) * Dc.Y * Dc.Z * 8 * c;
    if (c == 16) k_down_backward_add<16><<<blocks_for(n), 256, 0, s>>>(Dc, gc, gf);
    else if (c == 32) k_down_backward_add<32><<<blocks_for(n), 256, 0, s>>>(Dc, gc, gf);
    else if (c == 64) k_down_backward_add<64><<<blocks_for(n), 256, 0, s>>>(Dc, gc, gf);
    else throw std::invalid_argument("down_backward_add: unsupported channels");
}
void up_backward(hipStream_t s, int c, Dims Dc, const float *gf, float *gc) {
    const long long n = static_cast<long long>(Dc.X) * Dc.Y * Dc.Z * c;
    if (c == 16) k_up_backward<16><<<blocks_for(n), 256, 0, s>>>(Dc, gf, gc);
    else if (c == 32) k_up_backward<32><<<blocks_for(n), 256, 0, s>>>(Dc, gf, gc);
    else if (c == 64) k_up_backward<64><<<blocks_for(n), 256, 0, s>>>(Dc, gf, gc);
    else throw std::invalid_argument("up_backward: unsupported channels");
}
void bias_grad(hipStream_t s, int c, long long voxels, const float *g, float *gb) {
    const unsigned b = blocks_for(voxels) > 512 ? 512 : blocks_for(voxels);
    if (c == 1) k_bias_grad<1><<<b, 256, 0, s>>>(voxels, g, gb);
    else if (c == 16) k_bias_grad<16><<<b, 256, 0, s>>>(voxels, g, gb);
    else if (c == 32) k_bias_grad<32><<<b, 256, 0, s>>>(voxels, g, gb);
    else if (c == 64) k_bias_grad<64><<<b, 256, 0, s>>>(voxels, g, gb);
    else throw std::invalid_argument("bias_grad: unsupported channels");
}
void relu_mask(hipStream_t s, long long n, float *g, const uint16_t *act) {
    k_relu_mask<<<blocks_for(n), 256, 0, s>>>(n, g, act);
}
void copy_f(hipStream_t s, long long n, float *dst, const float *src) {
    k_copy_f<<<blocks_for(n), 256, 0, s>>>(n, dst, src);
}
void fill_f(hipStream_t s, long long n, float *dst, float v) { k_fill_f<<<blocks_for(n), 256, 0, s>>>(n, dst, v); }
void mse(hipStream_t s, long long n, const float *y, const float *target, float *gy, float *loss) {
    k_mse<<<blocks_for(n) > 1024 ? 1024 : blocks_for(n), 256, 0, s>>>(n, y, target, gy, loss);
}
void adam(hipStream_t s, int n, int ci, int co, float *w, const float *g, float *m, float *v, float lr, float b1,
          float b2, float eps, float c1, float c2, uint16_t *w_bf, uint16_t *wt_bf) {
    k_adam<<<blocks_for(n), 256, 0, s>>>(n, ci, co, w, g, m, v, lr, b1, b2, eps, c1, c2, w_bf, wt_bf);
}
void weights_up_bf16(hipStream_t s, int ci, int co, const float *w, uint16_t *wup) {
    k_weights_up_bf16<<<blocks_for(64LL * ci * co), 256, 0, s>>>(ci, co, w, wup);
}

void weights_bf16(hipStream_t s, int n, int ci, int co, const float *w, uint16_t *w_bf, uint16_t *wt_bf) {
    k_weights_bf16<<<blocks_for(n), 256, 0, s>>>(n, ci, co, w, w_bf, wt_bf);
}

}  // namespace nn
}  // namespace artss
