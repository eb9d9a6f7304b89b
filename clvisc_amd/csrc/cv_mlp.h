// cv_mlp.h -- the learned seed of the conserved-to-primitive root solve: a 4-64-64-2 MLP
// (tanh) trained and evaluated on the device with BF16 MFMA (FP32 accumulate).
//
// Arithmetic (what the host check restates): every matrix product takes BF16-rounded
// operands (round to nearest even) and accumulates in FP32; biases, tanh, the loss, the
// gradients' element-wise factors and Adam are FP32.  Layer 1 (K = 4) and layer 3 (N = 2)
// are VALU products; layer 2 (64 x 64) and its two backward products are
// v_mfma_f32_16x16x16_bf16.  A wave handles 16 samples, so every 64 x 64 product of a wave
// is 16 MFMAs.
//
// MFMA 16x16x16 operand layout (lane l of 64, i = 0..3):
//   A (16 x 16, rows x k):  row = l % 16, k = 4 (l / 16) + i
//   B (16 x 16, k x cols):  col = l % 16, k = 4 (l / 16) + i
//   C (16 x 16, rows x cols): col = l % 16, row = 4 (l / 16) + i
// so a C tile of one product is directly the A operand of a product that contracts over
// its rows (the weight gradient) and the B operand of one that keeps them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cv {

#define MLP_IN 4
#define MLP_H 64
#define MLP_OUT 2
// flat parameter vector: W1[64][4], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]
#define MLP_OFF_W1 0
#define MLP_OFF_B1 (MLP_OFF_W1 + MLP_H * MLP_IN)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_H)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_H * MLP_H)
#define MLP_OFF_W3 (MLP_OFF_B2 + MLP_H)
#define MLP_OFF_B3 (MLP_OFF_W3 + MLP_OUT * MLP_H)
#define MLP_NPARAM (MLP_OFF_B3 + MLP_OUT)
#define MLP_WAVES 4                      // waves per workgroup, 16 samples each

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_bits(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);           // round to nearest even (finite inputs)
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16r(float x) { return __uint_as_float((uint32_t)bf16_bits(x) << 16); }

__device__ __forceinline__ mlp_f32x4 mfma16(mlp_bf16x4 a, mlp_bf16x4 b, mlp_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ mlp_bf16x4 pack4(float a, float b, float c, float d) {
    mlp_bf16x4 r;
    r[0] = (short)bf16_bits(a); r[1] = (short)bf16_bits(b); r[2] = (short)bf16_bits(c); r[3] = (short)bf16_bits(d);
    return r;
}

// One wave: forward of 16 samples.  x[16][4] in LDS (already features), params FP32.
// Outputs in C layout per lane: h1[n][i], h2[n][i] (n = column tile 0..3, i = row 0..3 of
// the lane), y[i][o] for the lane's rows (replicated over the 16 lanes of a row group).
struct MlpFwd {
    float h1[4][4], h2[4][4];
    float y[4][MLP_OUT];
};

// sum over the 16 lanes that share l / 16 (lanes of one row group)
__device__ __forceinline__ float sum16(float v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
}
// sum over the 4 lanes that share l % 16 (same column, the four row groups)
__device__ __forceinline__ float sum_rowgroups(float v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// hbuf: per-wave LDS scratch of 16 x 64 floats (transposes C layout to A layout)
__device__ __forceinline__ void mlp_forward_wave(const float *__restrict__ P, const float xs[16][MLP_IN],
                                                 float (*hbuf)[MLP_H], MlpFwd &f) {
    const int l = threadIdx.x & 63;
    const int c0 = l & 15, rg = l >> 4;
    // layer 1 (VALU, K = 4): z1[row][col], row = 4 rg + i, col = 16 n + c0
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const int col = 16 * n + c0;
        const float w0 = bf16r(P[MLP_OFF_W1 + col * MLP_IN + 0]), w1 = bf16r(P[MLP_OFF_W1 + col * MLP_IN + 1]);
        const float w2 = bf16r(P[MLP_OFF_W1 + col * MLP_IN + 2]), w3 = bf16r(P[MLP_OFF_W1 + col * MLP_IN + 3]);
        const float b = P[MLP_OFF_B1 + col];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = 4 * rg + i;
            float z = bf16r(xs[row][0]) * w0;
            z = z + bf16r(xs[row][1]) * w1;
            z = z + bf16r(xs[row][2]) * w2;
            z = z + bf16r(xs[row][3]) * w3;
            f.h1[n][i] = tanhf(z + b);
        }
    }
    // C layout -> LDS -> A layout for layer 2
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int i = 0; i < 4; i++) hbuf[4 * rg + i][16 * n + c0] = f.h1[n][i];
    __syncthreads();
    mlp_f32x4 acc[4];
#pragma unroll
    for (int n = 0; n < 4; n++) acc[n] = mlp_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int kb = 16 * s + 4 * rg;
        mlp_bf16x4 a = pack4(hbuf[c0][kb + 0], hbuf[c0][kb + 1], hbuf[c0][kb + 2], hbuf[c0][kb + 3]);
#pragma unroll
        for (int n = 0; n < 4; n++) {
            // B[k][c] = W2[16 n + c][k]
            const float *w = P + MLP_OFF_W2 + (16 * n + c0) * MLP_H + kb;
            mlp_bf16x4 b = pack4(w[0], w[1], w[2], w[3]);
            acc[n] = mfma16(a, b, acc[n]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const float b = P[MLP_OFF_B2 + 16 * n + c0];
#pragma unroll
        for (int i = 0; i < 4; i++) f.h2[n][i] = tanhf(acc[n][i] + b);
    }
    // layer 3 (VALU, N = 2): lane partial over its 4 columns, then over the 16 lanes
#pragma unroll
    for (int o = 0; o < MLP_OUT; o++) {
        float w[4];
#pragma unroll
        for (int n = 0; n < 4; n++) w[n] = bf16r(P[MLP_OFF_W3 + o * MLP_H + 16 * n + c0]);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float s = bf16r(f.h2[0][i]) * w[0];
            s = s + bf16r(f.h2[1][i]) * w[1];
            s = s + bf16r(f.h2[2][i]) * w[2];
            s = s + bf16r(f.h2[3][i]) * w[3];
            f.y[i][o] = sum16(s) + P[MLP_OFF_B3 + o];
        }
    }
}

// Training: forward + backward of a batch; workgroup g writes its partial gradient to
// part[g][MLP_NPARAM] and its partial loss to part_loss[g].  Loss = mean over samples of
// sum_o (y_o - t_o)^2.
__global__ __launch_bounds__(64 * MLP_WAVES) void k_mlp_grad(const float *__restrict__ P, const float *__restrict__ X,
                                                           const float *__restrict__ T, int64_t nsamp, float inv_b,
                                                           float *part, float *part_loss) {
    __shared__ float s_x[MLP_WAVES][16][MLP_IN];
    __shared__ float s_h[MLP_WAVES][16][MLP_H];
    __shared__ float s_g[MLP_NPARAM];
    __shared__ float s_loss[MLP_WAVES];
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int c0 = l & 15, rg = l >> 4;
    for (int q = threadIdx.x; q < MLP_NPARAM; q += blockDim.x) s_g[q] = 0.0f;
    const int64_t base = ((int64_t)blockIdx.x * MLP_WAVES + wv) * 16;
    for (int q = l; q < 16 * MLP_IN; q += 64) {
        const int64_t r = base + q / MLP_IN;
        s_x[wv][q / MLP_IN][q % MLP_IN] = r < nsamp ? X[r * MLP_IN + q % MLP_IN] : 0.0f;
    }
    __syncthreads();
    MlpFwd f;
    mlp_forward_wave(P, s_x[wv], s_h[wv], f);
    // dy = 2 (y - t) / B for valid samples
    float dy[4][MLP_OUT];
    float loss = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int64_t r = base + 4 * rg + i;
#pragma unroll
        for (int o = 0; o < MLP_OUT; o++) {
            const float t = r < nsamp ? T[r * MLP_OUT + o] : 0.0f;
            const float d = f.y[i][o] - t;
            dy[i][o] = r < nsamp ? 2.0f * d * inv_b : 0.0f;
            if (c0 == 0 && r < nsamp) loss += d * d;
        }
    }
    loss = sum_rowgroups(loss);
    // layer 3 gradients: dW3[o][col] = sum_rows dy * bf16(h2); db3 = sum_rows dy
#pragma unroll
    for (int o = 0; o < MLP_OUT; o++) {
#pragma unroll
        for (int n = 0; n < 4; n++) {
            float s = 0.0f;
#pragma unroll
            for (int i = 0; i < 4; i++) s = s + dy[i][o] * bf16r(f.h2[n][i]);
            s = sum_rowgroups(s);
            if (rg == 0) atomicAdd(&s_g[MLP_OFF_W3 + o * MLP_H + 16 * n + c0], s);
        }
        // dy is the same on the 16 lanes of a row group: lane 0 of each group holds its rows
        float sb = dy[0][o] + dy[1][o] + dy[2][o] + dy[3][o];
        sb = sum_rowgroups(sb);
        if (l == 0) atomicAdd(&s_g[MLP_OFF_B3 + o], sb);
    }
    // dz2 = (dy W3) (1 - h2^2), C layout
    float dz2[4][4];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const float w0 = bf16r(P[MLP_OFF_W3 + 0 * MLP_H + 16 * n + c0]);
        const float w1 = bf16r(P[MLP_OFF_W3 + 1 * MLP_H + 16 * n + c0]);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float dh = dy[i][0] * w0 + dy[i][1] * w1;
            dz2[n][i] = dh * (1.0f - f.h2[n][i] * f.h2[n][i]);
        }
    }
    // db2 = sum_rows dz2; dW2 = dz2^T h1 (MFMA: A = dz2 tile m (C layout), B = h1 tile n)
#pragma unroll
    for (int n = 0; n < 4; n++) {
        float s = dz2[n][0] + dz2[n][1] + dz2[n][2] + dz2[n][3];
        s = sum_rowgroups(s);
        if (rg == 0) atomicAdd(&s_g[MLP_OFF_B2 + 16 * n + c0], s);
    }
#pragma unroll
    for (int m = 0; m < 4; m++) {
        mlp_bf16x4 a = pack4(dz2[m][0], dz2[m][1], dz2[m][2], dz2[m][3]);
#pragma unroll
        for (int n = 0; n < 4; n++) {
            mlp_bf16x4 b = pack4(f.h1[n][0], f.h1[n][1], f.h1[n][2], f.h1[n][3]);
            mlp_f32x4 acc = mfma16(a, b, mlp_f32x4{0.f, 0.f, 0.f, 0.f});
            // acc: rows = W2 row index 16 m + 4 rg + i, col = 16 n + c0
#pragma unroll
            for (int i = 0; i < 4; i++)
                atomicAdd(&s_g[MLP_OFF_W2 + (16 * m + 4 * rg + i) * MLP_H + 16 * n + c0], acc[i]);
        }
    }
    // dh1 = dz2 W2: A = dz2 in A layout (via LDS), B[k][c] = W2[k][16 n + c]
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int i = 0; i < 4; i++) s_h[wv][4 * rg + i][16 * n + c0] = dz2[n][i];
    __syncthreads();
    mlp_f32x4 dh1[4];
#pragma unroll
    for (int n = 0; n < 4; n++) dh1[n] = mlp_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int kb = 16 * s + 4 * rg;
        mlp_bf16x4 a = pack4(s_h[wv][c0][kb + 0], s_h[wv][c0][kb + 1], s_h[wv][c0][kb + 2], s_h[wv][c0][kb + 3]);
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const float *w = P + MLP_OFF_W2 + 16 * n + c0;
            mlp_bf16x4 b = pack4(w[(kb + 0) * MLP_H], w[(kb + 1) * MLP_H], w[(kb + 2) * MLP_H], w[(kb + 3) * MLP_H]);
            dh1[n] = mfma16(a, b, dh1[n]);
        }
    }
    // dz1 = dh1 (1 - h1^2); dW1[col][k] = sum_rows dz1 * bf16(x); db1 = sum_rows dz1
#pragma unroll
    for (int n = 0; n < 4; n++) {
        float dz1[4];
#pragma unroll
        for (int i = 0; i < 4; i++) dz1[i] = dh1[n][i] * (1.0f - f.h1[n][i] * f.h1[n][i]);
        float sb = dz1[0] + dz1[1] + dz1[2] + dz1[3];
        sb = sum_rowgroups(sb);
        if (rg == 0) atomicAdd(&s_g[MLP_OFF_B1 + 16 * n + c0], sb);
#pragma unroll
        for (int k = 0; k < MLP_IN; k++) {
            float s = 0.0f;
#pragma unroll
            for (int i = 0; i < 4; i++) s = s + dz1[i] * bf16r(s_x[wv][4 * rg + i][k]);
            s = sum_rowgroups(s);
            if (rg == 0) atomicAdd(&s_g[MLP_OFF_W1 + (16 * n + c0) * MLP_IN + k], s);
        }
    }
    if (l == 0) s_loss[wv] = loss;
    __syncthreads();
    for (int q = threadIdx.x; q < MLP_NPARAM; q += blockDim.x) part[(int64_t)blockIdx.x * MLP_NPARAM + q] = s_g[q];
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int w = 0; w < MLP_WAVES; w++) s += s_loss[w];
        part_loss[blockIdx.x] = s * inv_b;
    }
}

// Sum the workgroup partials (fixed order) into grad, then (if adam) one Adam step on P.
__global__ void k_mlp_reduce_adam(const float *__restrict__ part, const float *__restrict__ part_loss, int nblk,
                                  float *grad, float *loss_out, float *P, float *m1, float *m2, float lr,
                                  float beta1, float beta2, float eps, float c1, float c2, int adam) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q == 0) {
        float s = 0.0f;
        for (int b = 0; b < nblk; b++) s += part_loss[b];
        *loss_out = s;
    }
    if (q >= MLP_NPARAM) return;
    float g = 0.0f;
    for (int b = 0; b < nblk; b++) g += part[(int64_t)b * MLP_NPARAM + q];
    grad[q] = g;
    if (adam) {
        const float a = beta1 * m1[q] + (1.0f - beta1) * g;
        const float v = beta2 * m2[q] + (1.0f - beta2) * g * g;
        m1[q] = a;
        m2[q] = v;
        P[q] = P[q] - lr * (a * c1) / (sqrtf(v * c2) + eps);
    }
}

// Inference: y[n][2] for x[n][4].
__global__ __launch_bounds__(64 * MLP_WAVES) void k_mlp_forward(const float *__restrict__ P, const float *__restrict__ X,
                                                              int64_t nsamp, float *Y) {
    __shared__ float s_x[MLP_WAVES][16][MLP_IN];
    __shared__ float s_h[MLP_WAVES][16][MLP_H];
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int c0 = l & 15, rg = l >> 4;
    const int64_t base = ((int64_t)blockIdx.x * MLP_WAVES + wv) * 16;
    for (int q = l; q < 16 * MLP_IN; q += 64) {
        const int64_t r = base + q / MLP_IN;
        s_x[wv][q / MLP_IN][q % MLP_IN] = r < nsamp ? X[r * MLP_IN + q % MLP_IN] : 0.0f;
    }
    __syncthreads();
    MlpFwd f;
    mlp_forward_wave(P, s_x[wv], s_h[wv], f);
    if (c0 == 0) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t r = base + 4 * rg + i;
            if (r < nsamp)
                for (int o = 0; o < MLP_OUT; o++) Y[r * MLP_OUT + o] = f.y[i][o];
        }
    }
}

}  // namespace cv

namespace cv {

// Input features of the seed network for the conserved state (T^tt, |T^ti|, T^t eta):
// log2 T^tt / 16, |T^ti| / T^tt, T^t eta / T^tt, and the octave of T^tt (the table-row
// feature) / 24.  Outputs: e / T^tt and |v|.
__device__ __forceinline__ void seed_features(float T00, float M, float T03, float x[MLP_IN]) {
    const float t = fmaxf(T00, 1e-30f);
    int ex;
    (void)frexpf(t, &ex);
    x[0] = log2f(t) * (1.0f / 16.0f);
    x[1] = M / t;
    x[2] = T03 / t;
    x[3] = (float)(ex - 1) * (1.0f / 24.0f);
}

// Seed of the viscous root solve (split and fused stages): the clamp, the features, the
// network; the seed is the |v| output.  Throughput form of the forward pass, with the
// products transposed: Z^T = W X^T, the weights are the A operands (rows = a layer's units,
// k = its inputs) and the activations the B operands (k = inputs, columns = samples), so the
// C tiles of one layer (rows = its units, columns = 16 samples) are directly the B operands
// of the next and no transposition through LDS is needed.  Lane l = 16 rg + c0 holds sample
// c0 and units 4 rg + i (i = 0..3) of each 16-unit tile.  Layer 1 is one MFMA per 16 units
// (its 4 inputs fill k = 0..3, zero above), layer 3 (the |v| row) one MFMA per 16 units whose
// A tile holds the weights in row 0, the biases are the accumulators' initial values,
// tanh(z) = 1 - 2/(2^(2 z log2 e) + 1) is one multiply, the hardware exp2, one add, the
// hardware reciprocal and one FMA, and BF16 conversions are the hardware's round to nearest
// even.  (Folding 2 log2 e into the packed weights would save the multiply but changes their
// BF16 rounding: the seeds then move 1e-2 away from the network's, against 6e-4.)  The seed only sets where the root solve starts
// (the solve keeps the reference's bracket and stopping rule), so these approximations change
// iteration counts, not what the solve converges to.
#define MLP_SEED_TILES 8
__device__ __forceinline__ float tanh_seed(float z) {
    const float e = __builtin_amdgcn_exp2f(z * 2.88539008177792681f);   // exp(2 z)
    return fmaf(-2.0f, __builtin_amdgcn_rcpf(e + 1.0f), 1.0f);
}
__device__ __forceinline__ float bf16_hw(float x) { return (float)(__bf16)x; }
__device__ __forceinline__ mlp_bf16x4 pack4_hw(float a, float b, float c, float d) {
    typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
    typedef float f2 __attribute__((ext_vector_type(2)));
    // one v_cvt_pk_bf16_f32 per pair
    const bf2 lo = __builtin_convertvector((f2){a, b}, bf2), hi = __builtin_convertvector((f2){c, d}, bf2);
    return __builtin_bit_cast(mlp_bf16x4, make_uint2(__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi)));
}
__device__ __forceinline__ mlp_bf16x4 as_bf16x4(uint2 u) { return __builtin_bit_cast(mlp_bf16x4, u); }

// Packed operands of the transposed form (MLP_PK_WORDS 32-bit words, k_mlp_pack_t), per lane
// l = 16 rg + c0 so that one wave's load of a tile is 512 contiguous bytes:
//   W1 tile m (0..3) at uint2 64 m + l:            W1[16 m + c0][4 rg + i], zero for rg > 0
//   W2 tile (m, s) at uint2 256 + 64 (4 m + s) + l: W2[16 m + c0][16 s + 4 rg + i]
//   W3 tile m at uint2 1280 + 64 m + l:            W3[1][16 m + 4 rg + i] (the |v| row) for
//                                                  c0 = 0, zero for the other rows
#define MLP_PK_WORDS (2 * 64 * 24)

__global__ void k_mlp_pack_t(const float *__restrict__ P, uint32_t *pk) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < 24 * 64) {
        const int t = q >> 6, l = q & 63, c0 = l & 15, rg = l >> 4;
        float w[4];
        if (t < 4) {
#pragma unroll
            for (int i = 0; i < 4; i++) w[i] = rg == 0 ? P[MLP_OFF_W1 + (16 * t + c0) * MLP_IN + i] : 0.0f;
        } else if (t < 20) {
            const int m = (t - 4) >> 2, s = (t - 4) & 3;
#pragma unroll
            for (int i = 0; i < 4; i++) w[i] = P[MLP_OFF_W2 + (16 * m + c0) * MLP_H + 16 * s + 4 * rg + i];
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) w[i] = c0 == 0 ? P[MLP_OFF_W3 + MLP_H + 16 * (t - 20) + 4 * rg + i] : 0.0f;
        }
        reinterpret_cast<uint2 *>(pk)[q] = __builtin_bit_cast(uint2, pack4_hw(w[0], w[1], w[2], w[3]));
    }
}
#define MLP_PK_THREADS (24 * 64)

// |v| seeds of sample c0 of NG groups of 16 samples (y[g] on the lanes with rg = 0), the
// groups sharing each weight load.  xT[g]: the samples' BF16 features on the lanes with
// rg = 0, zero on the others.  A sample's seed depends on its own features only, whatever
// NG.  All 64 lanes of the wave must execute it.
template <int NG>
__device__ __forceinline__ void seed_groups_t(const float *__restrict__ P, const uint32_t *__restrict__ pk,
                                              const mlp_bf16x4 xT[NG], float y[NG]) {
    const int l = threadIdx.x & 63, rg = l >> 4;
    const uint2 *__restrict__ tiles = reinterpret_cast<const uint2 *>(pk);
    mlp_bf16x4 h1[NG][4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const float4 b = *reinterpret_cast<const float4 *>(P + MLP_OFF_B1 + 16 * m + 4 * rg);
        const mlp_bf16x4 w = as_bf16x4(tiles[64 * m + l]);
#pragma unroll
        for (int g = 0; g < NG; g++) {
            const mlp_f32x4 z = mfma16(w, xT[g], mlp_f32x4{b.x, b.y, b.z, b.w});
            h1[g][m] = pack4_hw(tanh_seed(z[0]), tanh_seed(z[1]), tanh_seed(z[2]), tanh_seed(z[3]));
        }
    }
    const float b3 = P[MLP_OFF_B3 + 1];
    mlp_f32x4 out[NG];
#pragma unroll
    for (int g = 0; g < NG; g++) out[g] = mlp_f32x4{b3, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const float4 b = *reinterpret_cast<const float4 *>(P + MLP_OFF_B2 + 16 * m + 4 * rg);
        mlp_f32x4 z[NG];
#pragma unroll
        for (int g = 0; g < NG; g++) z[g] = mlp_f32x4{b.x, b.y, b.z, b.w};
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const mlp_bf16x4 w = as_bf16x4(tiles[256 + 64 * (4 * m + s) + l]);
#pragma unroll
            for (int g = 0; g < NG; g++) z[g] = mfma16(w, h1[g][s], z[g]);
        }
        // layer 3: row 0 of the output tile (column = sample) accumulates the |v| row
        const mlp_bf16x4 w3 = as_bf16x4(tiles[1280 + 64 * m + l]);
#pragma unroll
        for (int g = 0; g < NG; g++)
            out[g] = mfma16(w3, pack4_hw(tanh_seed(z[g][0]), tanh_seed(z[g][1]), tanh_seed(z[g][2]),
                                         tanh_seed(z[g][3])), out[g]);
    }
#pragma unroll
    for (int g = 0; g < NG; g++) y[g] = out[g][0];
}

// Viscous stages: the network proposes the start of the root solve for cells with
// T^tt >= MLP_SEED_TMIN [GeV/fm^3]; a more dilute cell starts where the reference starts
// (v = 0.5, the plain solve converges there in 2 iterations), and a wavefront without a
// cell above the floor skips the network.  A cell's seed depends on its own T^tt and
// features only, so the fused and the split stage give the same seeds.
#define MLP_SEED_TMIN 1.0e-3f
__device__ __forceinline__ float seed_gate(float T00, float y) { return T00 >= MLP_SEED_TMIN ? y : 0.5f; }

// The features of one cell's root-solve input after the |T^ti| <= 0.999 T^tt clamp
// (kernel_visc.cl:365-372), as BF16 pairs (x0, x1) and (x2, x3).
__device__ __forceinline__ void seed_features_bf16(float T00, float M, float T03, uint32_t &lo, uint32_t &hi) {
    float x[MLP_IN];
    seed_features(T00, M, T03, x);
    const mlp_bf16x4 b = pack4_hw(x[0], x[1], x[2], x[3]);
    const uint2 u = __builtin_bit_cast(uint2, b);
    lo = u.x;
    hi = u.y;
}

// Seeds of the 64 cells of a wave (this lane's cell: features lo, hi from
// seed_features_bf16, zero for a lane without a cell): four groups of 16 samples, group g = lanes 16 g ..
// 16 g + 15.  All 64 lanes must execute it.
__device__ __forceinline__ float seed_wave_t(const float *__restrict__ P, const uint32_t *__restrict__ pk,
                                             uint32_t lo, uint32_t hi) {
    const int l = threadIdx.x & 63, c0 = l & 15, rg = l >> 4;
    float v = 0.0f;
#pragma unroll 1
    for (int g0 = 0; g0 < 4; g0 += 2) {
        mlp_bf16x4 xT[2];
#pragma unroll
        for (int g = 0; g < 2; g++) {
            uint32_t a = (uint32_t)__shfl((int)lo, 16 * (g0 + g) + c0, 64);
            uint32_t b = (uint32_t)__shfl((int)hi, 16 * (g0 + g) + c0, 64);
            if (rg != 0) { a = 0u; b = 0u; }
            xT[g] = as_bf16x4(make_uint2(a, b));
        }
        // the weights are read per pair of groups: the empty asm hides that the offset is
        // zero (and the loads loop invariant), so that the compiler does not hold all 88
        // weight registers of a lane across the host kernel's loops
        int zero = 0;
        asm volatile("" : "+s"(zero));
        float y[2];
        seed_groups_t<2>(P + zero, pk + zero, xT, y);
        // sample c0's seed is on lane c0 (rg = 0); lane 16 g + c0 holds that cell
        const float ya = __shfl(y[0], c0, 64), yb = __shfl(y[1], c0, 64);
        if (rg == g0) v = ya;
        if (rg == g0 + 1) v = yb;
    }
    return v;
}

// One block (MLP_WAVES waves, 16 MLP_WAVES MLP_SEED_TILES samples) of the seed pass over the
// conserved input of a split stage (four planes T^tt, T^tx, T^ty, T^t eta before the clamp,
// as k_visc_root reads them); seed_v = the seeds.  A sample's seed depends on its own
// features only, and equals the fused stage's seed of the same cell bit for bit.
__device__ __forceinline__ void seed_cons_block(const float *__restrict__ P, const uint32_t *__restrict__ pk,
                                                const float *__restrict__ cons, int64_t n, float *seed_v,
                                                int64_t blk) {
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int c0 = l & 15, rg = l >> 4;
    const int64_t tile0 = (blk * MLP_WAVES + wv) * MLP_SEED_TILES;
    for (int t = 0; t < MLP_SEED_TILES; t += 2) {
        // two tiles of 16 samples per pass of the weights
        mlp_bf16x4 xT[2];
        int64_t r[2];
        bool need = false;
#pragma unroll
        for (int g = 0; g < 2; g++) {
            r[g] = (tile0 + t + g) * 16 + c0;
            uint32_t lo = 0u, hi = 0u;
            if (rg == 0 && r[g] < n) {
                const int64_t q = r[g];
                float T00 = cons[q], T01 = cons[n + q], T02 = cons[2 * n + q], T03 = cons[3 * n + q];
                float M = sqrtf(T01 * T01 + T02 * T02 + T03 * T03);
                if (M > T00) {
                    T03 *= (0.999f * T00) * (1.0f / M);
                    M = 0.999f * T00;
                }
                seed_features_bf16(T00, M, T03, lo, hi);
                need = need || T00 >= MLP_SEED_TMIN;
            }
            xT[g] = as_bf16x4(make_uint2(lo, hi));
        }
        // the 32 samples of this pass all below the seed floor: no network (wave-uniform)
        float y[2] = {0.5f, 0.5f};
        if (__any(need)) seed_groups_t<2>(P, pk, xT, y);
#pragma unroll
        for (int g = 0; g < 2; g++)
            if (rg == 0 && r[g] < n) seed_v[r[g]] = seed_gate(cons[r[g]], y[g]);
    }
}

#define MLP_SEED_PER_BLOCK (16 * MLP_WAVES * MLP_SEED_TILES)

__global__ __launch_bounds__(64 * MLP_WAVES) void k_mlp_seed_cons(const float *__restrict__ P,
                                                                const uint32_t *__restrict__ pk,
                                                                const float *__restrict__ cons, int64_t n,
                                                                float *seed_v) {
    seed_cons_block(P, pk, cons, n, seed_v, blockIdx.x);
}

__global__ void k_seed_features(const float *__restrict__ T00, const float *__restrict__ M,
                                const float *__restrict__ T03, int64_t n, float *X) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    float x[MLP_IN];
    seed_features(T00[q], M[q], T03[q], x);
    for (int k = 0; k < MLP_IN; k++) X[q * MLP_IN + k] = x[k];
}

}  // namespace cv

namespace cv {
// out[q] = in[q * stride + col]
__global__ void k_take_col(const float *__restrict__ in, int64_t n, int stride, int col, float *out) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = in[q * stride + col];
}
}  // namespace cv
