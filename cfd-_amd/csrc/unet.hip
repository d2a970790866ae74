// Convolution building blocks of the learned pressure preconditioner (U-Net)
// on gfx950: 3x3 "same" convolutions in NHWC with BF16 operands and FP32
// accumulation on the matrix cores, their backward passes,
// pooling/upsampling, the loss and the Adam update.
//
// MFMA tiling (one wavefront = one output tile of 16 pixels x 16 channels per
// accumulator).  A K step is 16 input channels of one filter tap; for it lane
// l holds
//   A (16 x 16, M x K): row l%16, k = 4*(l/16) .. +3
//   B (16 x 16, K x N): col l%16, k = 4*(l/16) .. +3
//   D (16 x 16, M x N): col l%16, rows 4*(l/16) .. +3
// (v_mfma_f32_16x16x16_bf16's layout).  Two K steps go to gfx950's double-K
// instruction v_mfma_f32_16x16x32_bf16 as one: each lane passes its 4 values
// of the first step followed by its 4 of the second, in A and in B alike, so
// the instruction's sum over its 32 K positions is the sum of both steps'
// products whatever order it pairs them in (mfma2).  Both layouts are checked
// on the device against host products (cfd_mfma_layout_check,
// cfd_mfma_layout_check_k32).
// Here M = 16 consecutive pixels of one image row, K = 16 input channels of
// one filter tap, N = 16 output channels.  Weights are stored [tap][Cout][Cin]
// so the 4 K-values a lane needs are contiguous (one 8-byte load).
#include "cfd_kernels.h"
#include <hip/hip_bf16.h>

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// acc + A0 B0 + A1 B1 for two K steps (lane fragments as for
// v_mfma_f32_16x16x16_bf16) in one v_mfma_f32_16x16x32_bf16
__device__ __forceinline__ f32x4 mfma2(bf16x4 a0, bf16x4 a1, bf16x4 b0, bf16x4 b1, f32x4 acc) {
    const bf16x8 a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
    const bf16x8 b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8v, a),
                                                   __builtin_bit_cast(bf16x8v, b), acc, 0, 0, 0);
}

__device__ __forceinline__ unsigned short f2bf(float f) {
    // round to nearest even
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16) | ((u & 0xffff) ? 0x40 : 0);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

// ---------------------------------------------------------------------------
// Layout check: D = A(16x16) * B(16x16) from row-major float inputs.
// ---------------------------------------------------------------------------
__global__ void k_mfma_layout_check(const float* A, const float* B, float* D) {
    const int l = threadIdx.x;   // one wavefront
    bf16x4 a, b;
    for (int j = 0; j < 4; ++j) {
        const int k = 4 * (l / 16) + j;
        a[j] = (short)f2bf(A[(l % 16) * 16 + k]);
        b[j] = (short)f2bf(B[k * 16 + (l % 16)]);
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    for (int j = 0; j < 4; ++j) D[(4 * (l / 16) + j) * 16 + (l % 16)] = acc[j];
}

// The same for D = A(16x32) * B(32x16) through mfma2: the two halves of K as
// the two K steps
__global__ void k_mfma_layout_check_k32(const float* A, const float* B, float* D) {
    const int l = threadIdx.x;   // one wavefront
    bf16x4 a0, a1, b0, b1;
    for (int j = 0; j < 4; ++j) {
        const int k = 4 * (l / 16) + j;
        a0[j] = (short)f2bf(A[(l % 16) * 32 + k]);
        a1[j] = (short)f2bf(A[(l % 16) * 32 + 16 + k]);
        b0[j] = (short)f2bf(B[k * 16 + (l % 16)]);
        b1[j] = (short)f2bf(B[(16 + k) * 16 + (l % 16)]);
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma2(a0, a1, b0, b1, acc);
    for (int j = 0; j < 4; ++j) D[(4 * (l / 16) + j) * 16 + (l % 16)] = acc[j];
}

// ---------------------------------------------------------------------------
// 3x3 convolution, MFMA path.  Input = concat(X1 [H][W][C1], X2 [H][W][C2]),
// C1, C2, Cout multiples of 16 (C2 may be 0).  Wt [9][Cout][C1+C2] bf16.
// mode bit 0: add bias; bit 1: ReLU; bit 2: accumulate into Y (fp32 out).
// The forward kernels return at once when *skip is set (a solver that has
// converged leaves the rest of its captured iteration idle); skip may be NULL.
// Output either bf16 (Yb) or fp32 (Yf); with fp32 output, channels >= csplit
// go to Yf2 (channel stride Cout - csplit), the others to Yf (stride csplit):
// the gradient of a concatenated input lands in its two parts.
// One work-group computes a 16-pixel tile (of one image row) times G
// 16-channel output chunks; its 4 waves split the K loop over (tap, 16-channel
// input chunk) steps in batches of 8 (operands of a batch loaded before its
// matrix-core instructions) and add their accumulators through LDS: on the
// small images of the coarse levels the chains of dependent loads, not the
// matrix cores, set the time.
// ---------------------------------------------------------------------------
// rs1 (resampling of X1): 0 X1 is H x W; 1 X1 is 2H x 2W and is read through
// 2x2 average pooling (k_avgpool2's arithmetic); 2 X1 is H/2 x W/2 and is read
// through nearest upsampling.  Hi: rows per image (H for one image; a batch
// of images stacked along H has zero padding at every image boundary).
template <int G>
__global__ void __launch_bounds__(256) k_conv3x3_mfma(
    const unsigned short* __restrict__ X1, int C1, const unsigned short* __restrict__ X2, int C2,
    const unsigned short* __restrict__ Wt, const float* __restrict__ bias, int H, int Wd, int Cout,
    unsigned short* __restrict__ Yb, float* __restrict__ Yf, float* __restrict__ Yf2, int csplit,
    int mode, const int* skip, int rs1, int Hi) {
    if (skip && *skip) return;
    __shared__ f32x4 red[3][G][64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tiles_w = (Wd + 15) / 16;
    const int ngrp = Cout / (16 * G);
    const long task = blockIdx.x;
    if (task >= (long)tiles_w * H * ngrp) return;   // uniform over the work-group
    const int grp = (int)(task % ngrp);
    const long tile = task / ngrp;
    const int h = (int)(tile / tiles_w);
    const int hl = h % Hi;   // row within its image (a batch stacks images of Hi rows)
    const int w0 = (int)(tile - (long)h * tiles_w) * 16;
    const int Cin = C1 + C2;
    const int nk = Cin / 16;
    const int nsteps = 9 * nk;
    const int m = lane % 16, kq = 4 * (lane / 16);
    f32x4 acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = (f32x4){0.f, 0.f, 0.f, 0.f};
    auto load = [&](int st, bf16x4& a, bf16x4 (&b)[G]) {
        const int t = st / nk;
        const int k0 = (st - t * nk) * 16;
        const int hh = h + t / 3 - 1;
        const int ww = w0 + m + t % 3 - 1;
        a = (bf16x4){0, 0, 0, 0};
        if (hl + t / 3 - 1 >= 0 && hl + t / 3 - 1 < Hi && ww >= 0 && ww < Wd) {
            if (k0 >= C1) {
                a = *reinterpret_cast<const bf16x4*>(X2 + ((long)hh * Wd + ww) * C2 + (k0 - C1) + kq);
            } else if (rs1 == 0) {
                a = *reinterpret_cast<const bf16x4*>(X1 + ((long)hh * Wd + ww) * C1 + k0 + kq);
            } else if (rs1 == 2) {
                a = *reinterpret_cast<const bf16x4*>(
                    X1 + ((long)(hh >> 1) * (Wd >> 1) + (ww >> 1)) * C1 + k0 + kq);
            } else {
                const long W2 = 2L * Wd;
                const unsigned short* p0 = X1 + ((long)(2 * hh) * W2 + 2 * ww) * C1 + k0 + kq;
                const bf16x4 q0 = *reinterpret_cast<const bf16x4*>(p0);
                const bf16x4 q1 = *reinterpret_cast<const bf16x4*>(p0 + C1);
                const bf16x4 q2 = *reinterpret_cast<const bf16x4*>(p0 + W2 * C1);
                const bf16x4 q3 = *reinterpret_cast<const bf16x4*>(p0 + W2 * C1 + C1);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float sm = bf2f((unsigned short)q0[e]) + bf2f((unsigned short)q1[e]) +
                                     bf2f((unsigned short)q2[e]) + bf2f((unsigned short)q3[e]);
                    a[e] = (short)f2bf(0.25f * sm);
                }
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
            b[g] = *reinterpret_cast<const bf16x4*>(
                Wt + ((long)t * Cout + (grp * G + g) * 16 + m) * Cin + k0 + kq);
    };
    constexpr int U = 8;
    static_assert(U % 2 == 0, "steps in pairs");
    for (int s0 = wave * U; s0 < nsteps; s0 += 4 * U) {
        bf16x4 a[U], b[U][G];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (s0 + u < nsteps) {
                load(s0 + u, a[u], b[u]);
            } else {
                a[u] = (bf16x4){0, 0, 0, 0};
#pragma unroll
                for (int g = 0; g < G; ++g) b[u][g] = (bf16x4){0, 0, 0, 0};
            }
        }
#pragma unroll
        for (int u = 0; u < U; u += 2)
#pragma unroll
            for (int g = 0; g < G; ++g)
                acc[g] = mfma2(a[u], a[u + 1], b[u][g], b[u + 1][g], acc[g]);
    }
    if (wave > 0)
#pragma unroll
        for (int g = 0; g < G; ++g) red[wave - 1][g][lane] = acc[g];
    __syncthreads();
    if (wave > 0) return;
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const f32x4 o = red[q][g][lane];
            acc[g][0] += o[0];
            acc[g][1] += o[1];
            acc[g][2] += o[2];
            acc[g][3] += o[3];
        }
    // D: lane holds pixels w0 + kq + j (j = 0..3) of channel (grp*G + g)*16 + m
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int co = (grp * G + g) * 16 + m;
        const float bv = (mode & 1) ? bias[co] : 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int w = w0 + kq + j;
            if (w >= Wd) continue;
            float v = acc[g][j] + bv;
            const long px = (long)h * Wd + w;
            if (Yf) {
                float* dst = co < csplit ? Yf + px * csplit + co
                                         : Yf2 + px * (Cout - csplit) + (co - csplit);
                if (mode & 4) v += *dst;
                if (mode & 2) v = fmaxf(v, 0.f);
                *dst = v;
            } else {
                if (mode & 2) v = fmaxf(v, 0.f);
                Yb[px * Cout + co] = f2bf(v);
            }
        }
    }
}

// Weight-gradient partials: dW[t][co][ci] = sum_p G[p][co] * X[p+off_t][ci]
// over the pixel chunk blockIdx.x (pix_per_block pixels).  One wave per (tap,
// 16-channel output chunk) computes all input-channel chunks (K = pixels in
// steps of 16: A = G^T (co x p) shared by the chunks, B = X shifted (p x ci));
// blockIdx.y * 4 + wave selects the (tap, output chunk).  Partials of chunk b
// at part[b][9][Cout][Cin].
__global__ void __launch_bounds__(256) k_conv3x3_wgrad(
    const unsigned short* __restrict__ G, int Cout, const unsigned short* __restrict__ X1, int C1,
    const unsigned short* __restrict__ X2, int C2, int H, int Wd, long pix_per_block,
    float* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int Cin = C1 + C2;
    const int nco = Cout / 16, nci = Cin / 16;
    const int task = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (task >= 9 * nco) return;
    const int t = task / nco, cb = task - t * nco;
    const int dh = t / 3 - 1, dw = t % 3 - 1;
    const long p0 = (long)blockIdx.x * pix_per_block;
    const long npix = (long)H * Wd;
    const long p1 = p0 + pix_per_block < npix ? p0 + pix_per_block : npix;
    const int m = lane % 16, kq = 4 * (lane / 16);
    f32x4 acc[8];
#pragma unroll
    for (int ib = 0; ib < 8; ++ib) acc[ib] = (f32x4){0.f, 0.f, 0.f, 0.f};
    // two K steps of 16 pixels per double-K instruction (pixels past p1 add 0)
    for (long p = p0; p < p1; p += 32) {
        bf16x4 a[2];
        long q[2][4];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long pp = p + 16 * s + kq + j;
                unsigned short ga = 0;
                q[s][j] = -1;
                if (pp < p1) {
                    ga = G[pp * Cout + cb * 16 + m];
                    const int hh = (int)(pp / Wd) + dh, ww = (int)(pp % Wd) + dw;
                    if (hh >= 0 && hh < H && ww >= 0 && ww < Wd) q[s][j] = (long)hh * Wd + ww;
                }
                a[s][j] = (short)ga;
            }
#pragma unroll
        for (int ib = 0; ib < 8; ++ib) {
            if (ib >= nci) break;
            const int ci = ib * 16 + m;
            bf16x4 b[2];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    unsigned short xb = 0;
                    if (q[s][j] >= 0)
                        xb = ci < C1 ? X1[q[s][j] * C1 + ci] : X2[q[s][j] * C2 + ci - C1];
                    b[s][j] = (short)xb;
                }
            acc[ib] = mfma2(a[0], a[1], b[0], b[1], acc[ib]);
        }
    }
    // D rows = co (4*(lane/16)+j within the chunk), cols = ci (lane%16)
    float* dst = part + (long)blockIdx.x * 9 * Cout * Cin;
#pragma unroll
    for (int ib = 0; ib < 8; ++ib) {
        if (ib >= nci) break;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int co = cb * 16 + kq + j, ci = ib * 16 + m;
            dst[((long)t * Cout + co) * Cin + ci] = acc[ib][j];
        }
    }
}

// sum of partials over blocks (fixed order): out[k] += sum_b part[b * stride + k]
__global__ void k_sum_partials(const float* __restrict__ part, int nb, long stride, long n,
                               float* __restrict__ out) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float s = 0.f;
    for (int b = 0; b < nb; ++b) s += part[(long)b * stride + k];
    out[k] += s;
}

// Backward-data weights: W2[t][ci][co] = W[8 - t][co][ci]
__global__ void k_flip_transpose(const unsigned short* __restrict__ W, int Cout, int Cin,
                                 unsigned short* __restrict__ W2) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = 9L * Cout * Cin;
    if (k >= n) return;
    const int t = (int)(k / ((long)Cout * Cin));
    const long r = k - (long)t * Cout * Cin;
    const int co = (int)(r / Cin), ci = (int)(r - (long)co * Cin);
    W2[((long)(8 - t) * Cin + ci) * Cout + co] = W[k];
}

// ---------------------------------------------------------------------------
// Elementwise / resampling kernels (NHWC, bf16 activations)
// ---------------------------------------------------------------------------
// G = dY * (Y > 0)   (ReLU backward), bf16 in/out
__global__ void k_relu_mask(const float* __restrict__ dY, const unsigned short* __restrict__ Y,
                            long n, unsigned short* __restrict__ G) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) G[k] = f2bf(bf2f(Y[k]) > 0.f ? dY[k] : 0.f);
}

// 2x2 average pooling, C channels: Y[h][w] = mean of X[2h..2h+1][2w..2w+1]
__global__ void k_avgpool2(const unsigned short* __restrict__ X, int H, int Wd, int C,
                           unsigned short* __restrict__ Y, const int* skip) {
    if (skip && *skip) return;
    const int H2 = H / 2, W2 = Wd / 2;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)H2 * W2 * C) return;
    const int c = (int)(k % C);
    const long p = k / C;
    const int h = (int)(p / W2), w = (int)(p % W2);
    const float s = bf2f(X[((long)(2 * h) * Wd + 2 * w) * C + c]) +
                    bf2f(X[((long)(2 * h) * Wd + 2 * w + 1) * C + c]) +
                    bf2f(X[((long)(2 * h + 1) * Wd + 2 * w) * C + c]) +
                    bf2f(X[((long)(2 * h + 1) * Wd + 2 * w + 1) * C + c]);
    Y[k] = f2bf(0.25f * s);
}

// backward of avgpool: dX[fine] += dY[coarse] / 4
__global__ void k_avgpool2_bwd(const float* __restrict__ dY, int H, int Wd, int C,
                               float* __restrict__ dX) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)H * Wd * C) return;
    const int c = (int)(k % C);
    const long p = k / C;
    const int h = (int)(p / Wd), w = (int)(p % Wd);
    dX[k] += 0.25f * dY[((long)(h / 2) * (Wd / 2) + w / 2) * C + c];
}

// nearest upsampling x2: Y[h][w] = X[h/2][w/2]
__global__ void k_upsample2(const unsigned short* __restrict__ X, int H, int Wd, int C,
                            unsigned short* __restrict__ Y, const int* skip) {
    if (skip && *skip) return;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)H * Wd * C) return;
    const int c = (int)(k % C);
    const long p = k / C;
    const int h = (int)(p / Wd), w = (int)(p % Wd);
    Y[k] = X[((long)(h / 2) * (Wd / 2) + w / 2) * C + c];
}

// backward of upsampling: dX[coarse] += sum of the four fine dY
__global__ void k_upsample2_bwd(const float* __restrict__ dY, int H, int Wd, int C,
                                float* __restrict__ dX) {
    const int H2 = H / 2, W2 = Wd / 2;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)H2 * W2 * C) return;
    const int c = (int)(k % C);
    const long p = k / C;
    const int h = (int)(p / W2), w = (int)(p % W2);
    dX[k] += dY[((long)(2 * h) * Wd + 2 * w) * C + c] + dY[((long)(2 * h) * Wd + 2 * w + 1) * C + c] +
             dY[((long)(2 * h + 1) * Wd + 2 * w) * C + c] + dY[((long)(2 * h + 1) * Wd + 2 * w + 1) * C + c];
}

// first layer (1 -> C channels), scalar: Y = relu(b + sum_t W[t][c] x[p+off_t])
__global__ void k_conv_in(const float* __restrict__ x, int H, int Wd, const float* __restrict__ W,
                          const float* __restrict__ b, int C, unsigned short* __restrict__ Y,
                          const int* skip, int Hi) {
    if (skip && *skip) return;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)H * Wd * C) return;
    const int c = (int)(k % C);
    const long p = k / C;
    const int h = (int)(p / Wd), w = (int)(p % Wd);
    const int hl = h % Hi;
    float s = b[c];
    for (int t = 0; t < 9; ++t) {
        const int hh = h + t / 3 - 1, ww = w + t % 3 - 1;
        if (hl + t / 3 - 1 >= 0 && hl + t / 3 - 1 < Hi && ww >= 0 && ww < Wd)
            s += W[t * C + c] * x[(long)hh * Wd + ww];
    }
    Y[k] = f2bf(fmaxf(s, 0.f));
}

// last layer (C -> 1), linear: y = b + sum_t sum_c W[t][c] X[p+off_t][c]
// (C a multiple of 8: 16-byte loads of 8 channels)
__global__ void k_conv_out(const unsigned short* __restrict__ X, int H, int Wd, int C,
                           const float* __restrict__ W, const float* __restrict__ b,
                           float* __restrict__ y, const int* skip, int Hi) {
    if (skip && *skip) return;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long)H * Wd) return;
    const int h = (int)(p / Wd), w = (int)(p % Wd);
    const int hl = h % Hi;
    float s = b[0];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int hh = h + t / 3 - 1, ww = w + t % 3 - 1;
        if (hl + t / 3 - 1 < 0 || hl + t / 3 - 1 >= Hi || ww < 0 || ww >= Wd) continue;
        const uint4* xp = reinterpret_cast<const uint4*>(X + ((long)hh * Wd + ww) * C);
        for (int c8 = 0; c8 < C / 8; ++c8) {
            const uint4 q = xp[c8];
            const unsigned int u[4] = {q.x, q.y, q.z, q.w};
            const float* wt = W + t * C + c8 * 8;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                s += wt[2 * e] * __uint_as_float(u[e] << 16);
                s += wt[2 * e + 1] * __uint_as_float(u[e] & 0xffff0000u);
            }
        }
    }
    y[p] = s;
}

// backward of the last layer to its input: dX[p][c] = sum_t W[t][c] dy[p - off_t]
__global__ void k_conv_out_bwd_data(const float* __restrict__ dy, int H, int Wd, int C,
                                    const float* __restrict__ W, float* __restrict__ dX) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)H * Wd * C) return;
    const int c = (int)(k % C);
    const long p = k / C;
    const int h = (int)(p / Wd), w = (int)(p % Wd);
    float s = 0.f;
    for (int t = 0; t < 9; ++t) {
        const int hh = h - (t / 3 - 1), ww = w - (t % 3 - 1);
        if (hh >= 0 && hh < H && ww >= 0 && ww < Wd) s += W[t * C + c] * dy[(long)hh * Wd + ww];
    }
    dX[k] = s;
}

// weight/bias gradient partials of the scalar layers over a pixel range:
// in-layer:  gW[t][c] = sum_p G[p][c] x[p+off_t],  gb[c] = sum_p G[p][c]
// out-layer: gW[t][c] = sum_p dy[p] X[p+off_t][c], gb = sum_p dy[p]
// Each block writes 10*C floats (9 taps + bias) of partials.
__global__ void k_scalar_layer_wgrad(int which, const float* __restrict__ xin,
                                     const unsigned short* __restrict__ Gin,
                                     const float* __restrict__ dy,
                                     const unsigned short* __restrict__ Xout, int H, int Wd, int C,
                                     long pix_per_block, float* __restrict__ part) {
    __shared__ float acc[10 * 64];
    for (int k = threadIdx.x; k < 10 * C; k += blockDim.x) acc[k] = 0.f;
    __syncthreads();
    const long npix = (long)H * Wd;
    const long p0 = (long)blockIdx.x * pix_per_block;
    const long p1 = p0 + pix_per_block < npix ? p0 + pix_per_block : npix;
    // thread handles channel c for a strided pixel subset; private sums
    const int c = threadIdx.x % C;
    const int lanes = blockDim.x / C;
    const int sub = threadIdx.x / C;
    if (sub < lanes) {
        float s[10];
        for (int q = 0; q < 10; ++q) s[q] = 0.f;
        for (long p = p0 + sub; p < p1; p += lanes) {
            const int h = (int)(p / Wd), w = (int)(p % Wd);
            const float g = which == 0 ? bf2f(Gin[p * C + c]) : dy[p];
            if (g == 0.f) continue;
            for (int t = 0; t < 9; ++t) {
                const int hh = h + t / 3 - 1, ww = w + t % 3 - 1;
                if (hh < 0 || hh >= H || ww < 0 || ww >= Wd) continue;
                const long q = (long)hh * Wd + ww;
                s[t] += g * (which == 0 ? xin[q] : bf2f(Xout[q * C + c]));
            }
            s[9] += g;
        }
        for (int q = 0; q < 10; ++q) atomicAdd(&acc[q * C + c], s[q]);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 10 * C; k += blockDim.x)
        part[(long)blockIdx.x * 10 * C + k] = acc[k];
}

// bias gradient partials of an MFMA layer: gb[c] = sum_p G[p][c]
__global__ void k_bias_grad(const unsigned short* __restrict__ G, long npix, int C,
                            long pix_per_block, float* __restrict__ part) {
    __shared__ float acc[128];
    for (int k = threadIdx.x; k < C; k += blockDim.x) acc[k] = 0.f;
    __syncthreads();
    const long p0 = (long)blockIdx.x * pix_per_block;
    const long p1 = p0 + pix_per_block < npix ? p0 + pix_per_block : npix;
    const int c = threadIdx.x % C, lanes = blockDim.x / C, sub = threadIdx.x / C;
    if (sub < lanes) {
        float s = 0.f;
        for (long p = p0 + sub; p < p1; p += lanes) s += bf2f(G[p * C + c]);
        atomicAdd(&acc[c], s);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < C; k += blockDim.x) part[(long)blockIdx.x * C + k] = acc[k];
}

// fp32 -> bf16 copy and fp32 fill
__global__ void k_to_bf16(const float* __restrict__ x, long n, unsigned short* __restrict__ y) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] = f2bf(x[k]);
}
__global__ void k_fill_f32(float* __restrict__ x, long n, float v) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = v;
}
__global__ void k_scale_f32(float* __restrict__ x, long n, float v) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] *= v;
}

// masked squared error and its gradient: loss += sum m (y - t)^2 / norm,
// dy = 2 m (y - t) / norm; per-block partial sums of the loss
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t,
                      const float* __restrict__ mask, long n, float inv_norm,
                      float* __restrict__ dy, float* __restrict__ part) {
    __shared__ float s_acc[256];
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (k < n) {
        const float m = mask ? mask[k] : 1.f;
        const float d = m * (y[k] - t[k]);
        dy[k] = 2.f * d * inv_norm;
        l = d * d * inv_norm;
    }
    s_acc[threadIdx.x] = l;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s_acc[threadIdx.x] += s_acc[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s_acc[0];
}

// energy-norm error and its gradient: e = y - t on the valid nx x ny corner
// of the H x W image, A = the 5-point Laplacian with zero-flux sides (unit
// couplings) plus eps I;  loss += e.A e / norm, dy = 2 A e / norm (0 outside)
__global__ void k_energy(const float* __restrict__ y, const float* __restrict__ t, int H, int W,
                         int nxv, int nyv, float eps, float inv_norm, float* __restrict__ dy,
                         float* __restrict__ part) {
    __shared__ float s_acc[256];
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (k < (long)H * W) {
        const int h = (int)(k / W), w = (int)(k % W);
        float g = 0.f;
        if (h < nxv && w < nyv) {
            const float e = y[k] - t[k];
            float ae = eps * e;
            if (h > 0) ae += e - (y[k - W] - t[k - W]);
            if (h < nxv - 1) ae += e - (y[k + W] - t[k + W]);
            if (w > 0) ae += e - (y[k - 1] - t[k - 1]);
            if (w < nyv - 1) ae += e - (y[k + 1] - t[k + 1]);
            l = e * ae * inv_norm;
            g = 2.f * ae * inv_norm;
        }
        dy[k] = g;
    }
    s_acc[threadIdx.x] = l;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s_acc[threadIdx.x] += s_acc[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s_acc[0];
}

// Adam on FP32 master weights; refreshes the BF16 copy used by the MFMA layers
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, unsigned short* __restrict__ wb) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float a = b1 * m1[k] + (1.f - b1) * gk;
    const float b = b2 * m2[k] + (1.f - b2) * gk * gk;
    m1[k] = a;
    m2[k] = b;
    const float v = w[k] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    w[k] = v;
    if (wb) wb[k] = f2bf(v);
}

// ---------------------------------------------------------------------------
// The whole forward pass of the network for one 32 x 32 image in one
// work-group of 16 wavefronts, activations in LDS.  Same layers, weights and
// BF16 rounding as the per-layer kernels (pooled tensors are formed once per
// layer, with the same arithmetic as the per-layer kernels' pooling read).
//
// Matrix-core operands: the B fragments (weights) come from a copy stored in
// lane order (k_swizzle_b: the 64 lanes' 4 BF16 values of one K step are 512
// contiguous bytes), loaded one batch of 9 K steps ahead of their use.  A
// tile is 16 consecutive pixels in row-major order (several rows of the small
// images), the A fragments are LDS reads at constant offsets, and the pixel
// stride of every LDS tensor is its channel count + 4 so that the 16 pixels
// of a fragment fall in different LDS banks.  Each wavefront owns one
// 16-channel output chunk and 1-4 tiles; on the 4 x 4 level, where there are
// fewer (tile, chunk) pairs than wavefronts, 4 wavefronts split the K loop of
// a pair and add their accumulators through LDS in a fixed order.
//
// LDS (BF16 elements): level 0 S0 and W0a (20480 each); during levels 1-3 the
// work tensors W1a, W2a, W2b, W3a, W3b live inside W0a, the skips S1, S2 and
// W1b in R (22784); the last level-0 tensor W0b reuses S0.  The network input
// (FP32) sits in R before level 1 starts.
// ---------------------------------------------------------------------------
// offsets of layer l's weights and bias in the flat parameter vector (per
// layer: 9 Cin Cout weights, then Cout biases), compile-time constants so
// that the one-image network keeps no per-layer offsets in registers
// (cfd_unet_create checks them against the layer table)
namespace img_layout {
constexpr int CIN[15] = {1, 16, 16, 32, 32, 64, 64, 64, 128, 64, 96, 32, 48, 16, 16};
constexpr int COUT[15] = {16, 16, 32, 32, 64, 64, 64, 64, 64, 64, 32, 32, 16, 16, 1};
constexpr long woff(int l) {
    long o = 0;
    for (int k = 0; k < l; ++k) o += 9L * CIN[k] * COUT[k] + COUT[k];
    return o;
}
constexpr long boff(int l) { return woff(l) + 9L * CIN[l] * COUT[l]; }
}  // namespace img_layout

// B operands of a matrix-core layer in lane order: fragment (ch, st) for
// output chunk ch and K step st (tap t = st / nk, input channels
// 16 (st % nk) .. +15, nk = Cin / 16) holds, for lane = m + 16 q, the weights
// W[t][16 ch + m][16 (st % nk) + 4 q .. + 3] of the [tap][Cout][Cin] layout
__global__ void k_swizzle_b(const unsigned short* __restrict__ wb, int Cin, int Cout,
                            unsigned short* __restrict__ out) {
    const long n = 9L * Cin * Cout;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int nk = Cin / 16, nsteps = 9 * nk;
    const int j = (int)(e & 3), lane = (int)((e >> 2) & 63);
    const long frag = e >> 8;
    const int st = (int)(frag % nsteps), ch = (int)(frag / nsteps);
    const int t = st / nk, k0 = (st % nk) * 16;
    const int m = lane & 15, kq = 4 * (lane >> 4);
    out[e] = wb[((long)t * Cout + ch * 16 + m) * Cin + k0 + kq + j];
}

namespace {

constexpr int IMG_WAVES = 16;   // wavefronts of the one-image work-group
constexpr int IMG_PAD = 4;      // extra BF16 elements per pixel in LDS

// pointers with their address space spelled out: the layer functions are not
// inlined, and through generic pointers every LDS access would carry a 64-bit
// address conversion and null check, every weight load a flat access
typedef __attribute__((address_space(3))) unsigned short lds_u16;
typedef __attribute__((address_space(3))) bf16x4 lds_bf16x4;
typedef __attribute__((address_space(3))) float lds_f32;
typedef __attribute__((address_space(1))) bf16x4 glb_bf16x4;
typedef __attribute__((address_space(1))) float glb_f32;

// 2x2 average pooling of a 2S x 2S x C tensor into S x S x C (BF16 in LDS,
// pixel strides C + IMG_PAD; the sum of the four values in FP32, times 0.25,
// rounded to BF16: the per-layer kernels' pooling read)
template <int S, int C>
__device__ __forceinline__ void pool_img(const lds_u16* X, lds_u16* P) {
    constexpr int S2 = 2 * S, CS = C + IMG_PAD;
    for (int e = threadIdx.x; e < S * S * C; e += blockDim.x) {
        const int c = e % C, p = e / C, h = p / S, w = p % S;
        const lds_u16* p0 = X + ((2 * h) * S2 + 2 * w) * CS + c;
        const float sm = bf2f(p0[0]) + bf2f(p0[CS]) + bf2f(p0[S2 * CS]) + bf2f(p0[S2 * CS + CS]);
        P[p * CS + c] = f2bf(0.25f * sm);
    }
    __syncthreads();
}

// 3x3 convolution of an S x S image held in LDS: input channels X1 (C1 of
// them, read directly (UP = 0) or through nearest upsampling of an
// S/2 x S/2 tensor (UP = 1)) then X2 (C2, direct; with C2 = 0 X2 is never
// read, but pass a valid LDS pointer: a null one has crashed the compiler),
// bias + ReLU, BF16 out to LDS.  Bsw: the layer's lane-ordered B fragments;
// red: LDS scratch of 4096 floats for the split K loop (S = 4 only).  Shapes
// are template constants and the K loop is unrolled; not inlined, so that the
// layers' registers are not live at once.
template <int UP, int C1, int C2, int S, int COUT>
__device__ __attribute__((noinline)) void conv_img(const lds_u16* X1, const lds_u16* X2,
                                                   const glb_bf16x4* __restrict__ Bsw,
                                                   const glb_f32* __restrict__ bias, lds_u16* Y,
                                                   lds_f32* red) {
    static_assert(COUT == 16 || COUT == 32 || COUT == 64, "output channels");
    static_assert(C1 % 16 == 0 && C2 % 16 == 0 && S >= 4, "shapes");
    constexpr int CS1 = C1 + IMG_PAD, CS2 = C2 + IMG_PAD, CSO = COUT + IMG_PAD;
    constexpr int ntile = S * S / 16, nchunk = COUT / 16, pairs = ntile * nchunk;
    constexpr int nk1 = C1 / 16, nk = (C1 + C2) / 16, nsteps = 9 * nk;
    constexpr int ksplit = pairs >= IMG_WAVES ? 1 : IMG_WAVES / pairs;
    constexpr int T = pairs >= IMG_WAVES ? pairs / IMG_WAVES : 1;   // tiles per wavefront
    constexpr int kst = nsteps / ksplit;                             // K steps per wavefront
    static_assert(T <= 4 && pairs % (IMG_WAVES / ksplit) == 0, "work split");
    static_assert(kst % 9 == 0, "K steps per wavefront");
    static_assert(ksplit == 1 || pairs * ksplit * 256 <= 4096, "reduction scratch");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kq = 4 * (lane >> 4);
    const int pw = wave % (IMG_WAVES / ksplit);       // (tile, chunk) work item
    const int kpart = wave / (IMG_WAVES / ksplit);
    const int ch = pw % nchunk, t0 = pw / nchunk;
    constexpr int tstride = IMG_WAVES / ksplit / nchunk;   // tile stride between a wave's tiles
    int ph[T], pwc[T];   // this lane's A-row pixel (h, w) per tile
#pragma unroll
    for (int q = 0; q < T; ++q) {
        const int p = (t0 + q * tstride) * 16 + m;
        ph[q] = p / S;
        pwc[q] = p % S;
    }
    f32x4 acc[T];
#pragma unroll
    for (int q = 0; q < T; ++q) acc[q] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const int sbeg = kpart * kst;
    const glb_bf16x4* B = Bsw + ((long)ch * nsteps + sbeg) * 64 + lane;
    constexpr int U = 9;   // K steps per batch of B fragments
    bf16x4 bc[U], bn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) bc[u] = B[u * 64];
#pragma unroll
    for (int s0 = 0; s0 < kst; s0 += U) {
        if (s0 + U < kst) {
#pragma unroll
            for (int u = 0; u < U; ++u) bn[u] = B[(s0 + U + u) * 64];
        }
        // the A fragment of K step u of the batch for tile q: the wave's K
        // range starts at a multiple of 9 steps; tap and channel chunk follow
        // from the absolute step
        auto afrag = [&](int u, int q) -> bf16x4 {
            const int st = sbeg + s0 + u, t = st / nk, kc = st - t * nk;
            const int hh = ph[q] + t / 3 - 1, ww = pwc[q] + t % 3 - 1;
            const bool in = hh >= 0 && hh < S && ww >= 0 && ww < S;
            const int h2 = in ? hh : 0, w2 = in ? ww : 0;
            bf16x4 a;
            if (kc >= nk1)
                a = *reinterpret_cast<const lds_bf16x4*>(X2 + (h2 * S + w2) * CS2 + (kc - nk1) * 16 + kq);
            else if (UP)
                a = *reinterpret_cast<const lds_bf16x4*>(
                    X1 + ((h2 >> 1) * (S / 2) + (w2 >> 1)) * CS1 + kc * 16 + kq);
            else
                a = *reinterpret_cast<const lds_bf16x4*>(X1 + (h2 * S + w2) * CS1 + kc * 16 + kq);
            return in ? a : (bf16x4){0, 0, 0, 0};
        };
        // steps in pairs on the double-K instruction, the odd one out paired
        // with zeros: every product into an accumulator goes through the same
        // instruction (with the single-K form after the double-K one on the
        // same accumulator, this toolchain issued the dependent instruction
        // too early: wrong sums, which the batched-forward test caught)
        const bf16x4 z4 = {0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < U; u += 2)
#pragma unroll
            for (int q = 0; q < T; ++q)
                acc[q] = u + 1 < U ? mfma2(afrag(u, q), afrag(u + 1, q), bc[u], bc[u + 1], acc[q])
                                   : mfma2(afrag(u, q), z4, bc[u], z4, acc[q]);
        if (s0 + U < kst) {
#pragma unroll
            for (int u = 0; u < U; ++u) bc[u] = bn[u];
        }
    }
    const int co = ch * 16 + m;
    const float bv = bias[co];
    if constexpr (ksplit > 1) {
        // partial accumulators of the K parts, added in part order
        lds_f32* rp = red + ((pw * ksplit + kpart) * 64 + lane) * 4;
#pragma unroll
        for (int j = 0; j < 4; ++j) rp[j] = acc[0][j];
        __syncthreads();
        if (kpart == 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float v = 0.f;
#pragma unroll
                for (int k = 0; k < ksplit; ++k) v += red[((pw * ksplit + k) * 64 + lane) * 4 + j];
                const int p = t0 * 16 + kq + j;
                Y[p * CSO + co] = f2bf(fmaxf(v + bv, 0.f));
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < T; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int p = (t0 + q * tstride) * 16 + kq + j;
                Y[p * CSO + co] = f2bf(fmaxf(acc[q][j] + bv, 0.f));
            }
    }
    __syncthreads();
}

// LDS of the one-image network (BF16 elements): S0 and W0a, and R
constexpr int IMG_LDS_S0 = 1024 * (16 + IMG_PAD);
constexpr int IMG_LDS_R = 256 * (32 + IMG_PAD) * 2 + 64 * (64 + IMG_PAD);

// The network on one 32 x 32 image: xin (FP32, 1024 values in LDS, may lie in
// R), out(p, y) called once per pixel p = 32 h + w with the network output.
// wsw: lane-ordered B fragments (at the weight offsets woff of the MFMA
// layers), w: FP32 parameters (input/output layer weights, all biases).  All
// threads of the 1024-thread work-group must call it.
template <typename Out>
__device__ __forceinline__ void unet_img32(const float* xin, unsigned short* S0,
                                           unsigned short* W0a, unsigned short* R,
                                           const unsigned short* __restrict__ wsw,
                                           const float* __restrict__ w, Out out) {
    using img_layout::woff;
    using img_layout::boff;
    constexpr int P16 = 16 + IMG_PAD, P32 = 32 + IMG_PAD, P64 = 64 + IMG_PAD;
    unsigned short* W1a = W0a;                       // 256 x 32  (inside W0a, levels 1-3)
    unsigned short* W2a = W0a + 256 * P32;           // 64 x 64
    unsigned short* W2b = W2a + 64 * P64;            // 64 x 64
    unsigned short* W3a = W2b + 64 * P64;            // 16 x 64
    unsigned short* W3b = W3a + 16 * P64;            // 16 x 64
    unsigned short* S1 = R;                          // 256 x 32
    unsigned short* W1b = R + 256 * P32;             // 256 x 32
    unsigned short* S2 = W1b + 256 * P32;            // 64 x 64
    unsigned short* W0b = S0;                        // 1024 x 16 (after S0 is dead)
    float* red = reinterpret_cast<float*>(W1a);      // split-K partials (levels 3)
    static_assert(256 * P32 + 2 * 64 * P64 + 2 * 16 * P64 <= IMG_LDS_S0, "W0a holds levels 1-3");
    // LDS and global views of the tensors and weights for the layer functions
    auto L = [](unsigned short* p) { return (lds_u16*)p; };
    // input layer 1 -> 16 (scalar), ReLU: one pixel per thread, all 16
    // channels (the weights are the same for every thread), each output
    // summed in tap order as before
    {
        const float* W = w + woff(0);
        const float* b = w + boff(0);
        const lds_f32* xl = (const lds_f32*)xin;
        for (int p = threadIdx.x; p < 1024; p += blockDim.x) {
            const int h = p / 32, ww = p % 32;
            float xv[9];
            bool in[9];
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int hh = h + t / 3 - 1, w2 = ww + t % 3 - 1;
                in[t] = hh >= 0 && hh < 32 && w2 >= 0 && w2 < 32;
                xv[t] = xl[in[t] ? hh * 32 + w2 : p];
            }
            bf16x4 o[4];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                float s = b[c];
#pragma unroll
                for (int t = 0; t < 9; ++t)
                    if (in[t]) s += W[t * 16 + c] * xv[t];
                o[c / 4][c % 4] = (short)f2bf(fmaxf(s, 0.f));
            }
            lds_bf16x4* dst = (lds_bf16x4*)(L(W0a) + p * P16);
#pragma unroll
            for (int q = 0; q < 4; ++q) dst[q] = o[q];
        }
        __syncthreads();
    }
    auto Wt = [&](int l) { return (const glb_bf16x4*)(wsw + woff(l)); };
    auto Bs = [&](int l) { return (const glb_f32*)(w + boff(l)); };
    lds_f32* rd = (lds_f32*)red;
    // pooled inputs go to regions that are free at the time (W1b, W2b, W3b)
    conv_img<0, 16, 0, 32, 16>(L(W0a), L(W0a), Wt(1), Bs(1), L(S0), rd);       // A0B
    pool_img<16, 16>(L(S0), L(W1b));                                           // P1
    conv_img<0, 16, 0, 16, 32>(L(W1b), L(W1b), Wt(2), Bs(2), L(W1a), rd);      // A1A
    conv_img<0, 32, 0, 16, 32>(L(W1a), L(W1a), Wt(3), Bs(3), L(S1), rd);       // A1B
    pool_img<8, 32>(L(S1), L(W2b));                                            // P2
    conv_img<0, 32, 0, 8, 64>(L(W2b), L(W2b), Wt(4), Bs(4), L(W2a), rd);       // A2A
    conv_img<0, 64, 0, 8, 64>(L(W2a), L(W2a), Wt(5), Bs(5), L(S2), rd);        // A2B
    pool_img<4, 64>(L(S2), L(W3b));                                            // P3
    conv_img<0, 64, 0, 4, 64>(L(W3b), L(W3b), Wt(6), Bs(6), L(W3a), rd);       // A3A
    conv_img<0, 64, 0, 4, 64>(L(W3a), L(W3a), Wt(7), Bs(7), L(W3b), rd);       // A3B
    conv_img<1, 64, 64, 8, 64>(L(W3b), L(S2), Wt(8), Bs(8), L(W2a), rd);       // B2A
    conv_img<0, 64, 0, 8, 64>(L(W2a), L(W2a), Wt(9), Bs(9), L(W2b), rd);       // B2B
    conv_img<1, 64, 32, 16, 32>(L(W2b), L(S1), Wt(10), Bs(10), L(W1a), rd);    // B1A
    conv_img<0, 32, 0, 16, 32>(L(W1a), L(W1a), Wt(11), Bs(11), L(W1b), rd);    // B1B
    conv_img<1, 32, 16, 32, 16>(L(W1b), L(S0), Wt(12), Bs(12), L(W0a), rd);    // B0A
    conv_img<0, 16, 0, 32, 16>(L(W0a), L(W0a), Wt(13), Bs(13), L(W0b), rd);    // B0B
    // output layer 16 -> 1 (linear)
    {
        const float* W = w + woff(14);
        const float b = w[boff(14)];
        for (int p = threadIdx.x; p < 1024; p += blockDim.x) {
            const int h = p / 32, ww = p % 32;
            float s = b;
            for (int t = 0; t < 9; ++t) {
                const int hh = h + t / 3 - 1, w2 = ww + t % 3 - 1;
                if (hh < 0 || hh >= 32 || w2 < 0 || w2 >= 32) continue;
                const lds_u16* xp = L(W0b) + (hh * 32 + w2) * P16;
#pragma unroll
                for (int c = 0; c < 16; ++c) s += W[t * 16 + c] * bf2f(xp[c]);
            }
            out(p, s);
        }
    }
}

}  // namespace

// batches of 32 x 32 images (image k = rows 32 k .. 32 k + 31 of x and y):
// the network on one image per work-group
__global__ void __launch_bounds__(1024) k_unet32(const float* __restrict__ x, float* __restrict__ y,
                                                 const unsigned short* __restrict__ wsw,
                                                 const float* __restrict__ w, const int* skip) {
    if (skip && *skip) return;
    __shared__ __align__(16) unsigned short S0[IMG_LDS_S0], W0a[IMG_LDS_S0];
    __shared__ __align__(16) unsigned short R[IMG_LDS_R];
    float* xin = reinterpret_cast<float*>(R);
    const float* xi = x + (long)blockIdx.x * 1024;
    for (int p = threadIdx.x; p < 1024; p += blockDim.x) xin[p] = xi[p];
    __syncthreads();
    float* yo = y + (long)blockIdx.x * 1024;
    unet_img32(xin, S0, W0a, R, wsw, w, [&](int p, float v) { yo[p] = v; });
}
