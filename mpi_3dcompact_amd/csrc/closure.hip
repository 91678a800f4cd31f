// Learned SGS closure kernels: training (gradient, reduction, Adam), the
// register-resident inference path, the fused per-point closure of the solver
// step, its DSM fallback merge, and training-pair sampling from a DSM state.
//
// Matrix products use v_mfma_f32_16x16x32_bf16 (gfx950).  Operand layout of
// one 16x16x32 product D = A B on a 64-lane wave, lane l, g = l / 16:
//   A[i][k]: lane holds i = l % 16, k = 8 g + e (e = 0..7)
//   B[k][j]: lane holds j = l % 16, k = 8 g + e
//   D[i][j]: lane holds j = l % 16, i = 4 g + r (r = 0..3)
// Compiled with -ffp-contract=off: the FP32 and FP64 scalar arithmetic is
// rounded once per operation, as the checker restates it.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "closure.h"

namespace mc3d {

namespace {

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ f32x4 mfma(const s16x8 &a, const s16x8 &b, const f32x4 &c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                   c, 0, 0, 0);
}

constexpr float kG0 = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kG1 = 0.044715f;

// tanh-form GELU 0.5 z (1 + tanh u), u = k0 (z + k1 z^3), written as z s with
// s = 1 / (1 + exp(-2u)) = (1 + tanh u) / 2 and exp(-2u) = 2^(z (c1 + c2 z^2)),
// c1 = -2 k0 log2(e), c2 = c1 k1: three FP32 operations, the hardware 2^x and
// reciprocal (the network's evaluation is bound by these vector operations)
constexpr float kGc1 = (float)(-2.0 * 0.7978845608028654 * 1.4426950408889634);
constexpr float kGc2 = (float)(-2.0 * 0.7978845608028654 * 1.4426950408889634 * 0.044715);
__device__ __forceinline__ float gelu_s(float z) {
    const float z2 = z * z;
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * (kGc1 + kGc2 * z2)));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_s(z); }
// d/dz: s + z 2 s (1 - s) k0 (1 + 3 k1 z^2)
__device__ __forceinline__ float gelu_grad(float z) {
    const float s = gelu_s(z);
    return s + z * (2.0f * s * (1.0f - s)) * (kG0 * (1.0f + 3.0f * kG1 * z * z));
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// ---------------------------------------------------------------------------
// training: one workgroup = 4 waves = 64 samples; wave w owns samples
// 16w..16w+15 in the data products and output-row tile w in the weight
// products.  Weights and activations are staged in LDS as BF16.
// ---------------------------------------------------------------------------
constexpr int kT = kMlpTile;
constexpr int kLd0 = 8;  // padded input width (7 features + a zero column)

struct TrainLds {
    unsigned short w0[kMlpHid * kLd0];     // W1 padded to [64][8]
    unsigned short w1[kMlpHid * kMlpHid];  // W2
    unsigned short w2[kMlpHid * kMlpHid];  // W3
    unsigned short w3[kMlpOut * kMlpHid];  // W4
    unsigned short a0[kT * kLd0];          // standardised inputs
    unsigned short act[3][kT * kMlpHid];   // GELU outputs of the hidden layers
    unsigned short dz[kT * kMlpHid];       // BF16 delta of the current layer's pre-activation
    unsigned short dyb[kT * kMlpOut];      // BF16 dL/dy
    float dyf[kT * kMlpOut];               // FP32 dL/dy, then targets before that
    float red[4 * kMlpHid];                // cross-wave partial sums
    unsigned char valid[kT];
};

// pre-activation tile (no bias) of hidden/output layer `layer` for samples
// 16 si.. and neurons 16 nj..: in[s][k] (ld ldin, width kin) times W[j][k]
__device__ __forceinline__ f32x4 fwd_tile(const unsigned short *in, int ldin, int kin, const unsigned short *W,
                                          int ldw, int nrows, int si, int nj) {
    const int l = lane_id(), i = l & 15, g = l >> 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < kin; k0 += 32) {
        s16x8 a, b;
        const int row_w = 16 * nj + i;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + 8 * g + e;
            a[e] = k < kin ? (short)in[(16 * si + i) * ldin + k] : (short)0;
            b[e] = (k < kin && row_w < nrows) ? (short)W[row_w * ldw + k] : (short)0;
        }
        acc = mfma(a, b, acc);
    }
    return acc;
}

__device__ __forceinline__ const unsigned short *layer_w(const TrainLds &s, int l) {
    return l == 0 ? s.w0 : l == 1 ? s.w1 : l == 2 ? s.w2 : s.w3;
}
__device__ __forceinline__ const unsigned short *layer_in(const TrainLds &s, int l) {
    return l == 0 ? s.a0 : s.act[l - 1];
}

constexpr int kWOff[4] = {kMlpOffW1, kMlpOffW2, kMlpOffW3, kMlpOffW4};
constexpr int kBOff[4] = {kMlpOffB1, kMlpOffB2, kMlpOffB3, kMlpOffB4};

// column sums over the 64 samples of per-lane values v[t][r] (sample 16w+4g+r,
// column 16t + l%16) into out[col] (FP32, fixed order); all threads call it
__device__ void column_sums(TrainLds &s, const float v[4][4], int ncols, float *out) {
    const int l = lane_id(), w = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        float x = ((v[t][0] + v[t][1]) + v[t][2]) + v[t][3];
        x = x + __shfl_xor(x, 16);
        x = x + __shfl_xor(x, 32);
        if (l < 16 && 16 * t + l < ncols) s.red[w * kMlpHid + 16 * t + l] = x;
    }
    __syncthreads();
    if ((int)threadIdx.x < ncols) {
        const int c = threadIdx.x;
        out[c] = ((s.red[c] + s.red[kMlpHid + c]) + s.red[2 * kMlpHid + c]) + s.red[3 * kMlpHid + c];
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ P, const float *__restrict__ X,
                                                  const float *__restrict__ Tg, const int *__restrict__ idx,
                                                  long n, float scale, float *__restrict__ partial,
                                                  float *__restrict__ lossp) {
    __shared__ TrainLds s;
    const int tid = threadIdx.x, l = lane_id(), w = tid >> 6, g = l >> 4;
    float *gout = partial + (size_t)blockIdx.x * kMlpParams;

    // ---- stage weights and the sample tile ----
    for (int q = tid; q < kMlpHid * kLd0; q += 256) {
        const int j = q / kLd0, k = q % kLd0;
        s.w0[q] = k < kMlpIn ? f2bf(P[kMlpOffW1 + j * kMlpIn + k]) : (unsigned short)0;
    }
    for (int q = tid; q < kMlpHid * kMlpHid; q += 256) {
        s.w1[q] = f2bf(P[kMlpOffW2 + q]);
        s.w2[q] = f2bf(P[kMlpOffW3 + q]);
    }
    for (int q = tid; q < kMlpOut * kMlpHid; q += 256) s.w3[q] = f2bf(P[kMlpOffW4 + q]);
    if (tid < kT) {
        const long q = (long)blockIdx.x * kT + tid;
        const bool ok = q < n;
        const long src = ok ? (idx ? (long)idx[q] : q) : 0;
        s.valid[tid] = ok;
        for (int k = 0; k < kLd0; ++k)
            s.a0[tid * kLd0 + k] = (ok && k < kMlpIn) ? f2bf(X[src * kMlpIn + k]) : (unsigned short)0;
        for (int o = 0; o < kMlpOut; ++o) s.dyf[tid * kMlpOut + o] = ok ? Tg[src * kMlpOut + o] : 0.0f;
    }
    __syncthreads();

    // ---- forward: hidden layers (wave w: samples 16w.., all 64 neurons) ----
    for (int layer = 0; layer < 3; ++layer) {
        const int ldin = layer == 0 ? kLd0 : kMlpHid;
        const unsigned short *in = layer_in(s, layer), *W = layer_w(s, layer);
#pragma unroll
        for (int nj = 0; nj < 4; ++nj) {
            const f32x4 acc = fwd_tile(in, ldin, ldin, W, ldin, kMlpHid, w, nj);
            const float b = P[kBOff[layer] + 16 * nj + (l & 15)];
#pragma unroll
            for (int r = 0; r < 4; ++r)
                s.act[layer][(16 * w + 4 * g + r) * kMlpHid + 16 * nj + (l & 15)] = f2bf(gelu(acc[r] + b));
        }
        __syncthreads();
    }
    // ---- output layer, loss, dL/dy ----
    float sq = 0.0f;
    {
        const f32x4 acc = fwd_tile(s.act[2], kMlpHid, kMlpHid, s.w3, kMlpHid, kMlpOut, w, 0);
        const int o = l & 15;
        if (o < kMlpOut) {
            const float b = P[kMlpOffB4 + o];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int smp = 16 * w + 4 * g + r;
                const float y = acc[r] + b;
                const float d = s.valid[smp] ? y - s.dyf[smp * kMlpOut + o] : 0.0f;
                sq = sq + d * d;
                const float dy = scale * d;
                s.dyb[smp * kMlpOut + o] = f2bf(dy);
                s.red[w * kMlpHid + (4 * g + r) * kMlpOut + o] = dy;  // staged, copied below
            }
        }
    }
    // wave sums of squares -> lossp
    for (int off = 32; off > 0; off >>= 1) sq = sq + __shfl_xor(sq, off);
    __syncthreads();
    if (tid < kT * kMlpOut) s.dyf[tid] = s.red[(tid / 32) * kMlpHid + tid % 32];
    __shared__ float wsq[4];
    if (l == 0) wsq[w] = sq;
    __syncthreads();
    if (tid == 0) lossp[blockIdx.x] = ((wsq[0] + wsq[1]) + wsq[2]) + wsq[3];
    if (tid < kMlpOut) {  // db4
        float acc = 0.0f;
        for (int smp = 0; smp < kT; ++smp) acc = acc + s.dyf[smp * kMlpOut + tid];
        gout[kMlpOffB4 + tid] = acc;
    }

    // ---- backward through the output layer ----
    {
        // dW4[o][k] = sum_s dy[s][o] a3[s][k]: wave w -> k tile w
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < kT; k0 += 32) {
            s16x8 a, b;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int smp = k0 + 8 * g + e;
                a[e] = (l & 15) < kMlpOut ? (short)s.dyb[smp * kMlpOut + (l & 15)] : (short)0;
                b[e] = (short)s.act[2][smp * kMlpHid + 16 * w + (l & 15)];
            }
            acc = mfma(a, b, acc);
        }
        if (g == 0) {
#pragma unroll
            for (int r = 0; r < kMlpOut; ++r) gout[kMlpOffW4 + r * kMlpHid + 16 * w + (l & 15)] = acc[r];
        }
    }
    float dzv[4][4];
    {
        // dh3[s][k] = sum_o dy[s][o] W4[o][k], then dz = dh3 * gelu'(z3)
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            s16x8 a, b;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int o = 8 * g + e;
                a[e] = o < kMlpOut ? (short)s.dyb[(16 * w + (l & 15)) * kMlpOut + o] : (short)0;
                b[e] = o < kMlpOut ? (short)s.w3[o * kMlpHid + 16 * kt + (l & 15)] : (short)0;
            }
            const f32x4 dh = mfma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
            const f32x4 z = fwd_tile(s.act[1], kMlpHid, kMlpHid, s.w2, kMlpHid, kMlpHid, w, kt);
            const float bias = P[kMlpOffB3 + 16 * kt + (l & 15)];
#pragma unroll
            for (int r = 0; r < 4; ++r) dzv[kt][r] = dh[r] * gelu_grad(z[r] + bias);
        }
    }
    // dz of layer 2 (W3): store, bias gradient
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int r = 0; r < 4; ++r) s.dz[(16 * w + 4 * g + r) * kMlpHid + 16 * kt + (l & 15)] = f2bf(dzv[kt][r]);
    column_sums(s, dzv, kMlpHid, gout + kBOff[2]);

    // ---- hidden layers, top down ----
    for (int layer = 2; layer >= 0; --layer) {
        const unsigned short *in = layer_in(s, layer);
        const int kin = layer == 0 ? kMlpIn : kMlpHid;
        const int ldin = layer == 0 ? kLd0 : kMlpHid;
        // dW[j][k] = sum_s dz[s][j] in[s][k]: wave w -> row tile w
        for (int kt = 0; kt < (layer == 0 ? 1 : 4); ++kt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < kT; k0 += 32) {
                s16x8 a, b;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int smp = k0 + 8 * g + e;
                    a[e] = (short)s.dz[smp * kMlpHid + 16 * w + (l & 15)];
                    const int k = 16 * kt + (l & 15);
                    b[e] = k < ldin ? (short)in[smp * ldin + k] : (short)0;
                }
                acc = mfma(a, b, acc);
            }
            const int k = 16 * kt + (l & 15);
            if (k < kin) {
#pragma unroll
                for (int r = 0; r < 4; ++r) gout[kWOff[layer] + (16 * w + 4 * g + r) * kin + k] = acc[r];
            }
        }
        if (layer == 0) break;
        // dh[s][k] = sum_j dz[s][j] W[j][k]; dz_prev = dh * gelu'(z_prev)
        const unsigned short *W = layer_w(s, layer);
        const unsigned short *inp = layer_in(s, layer - 1), *Wp = layer_w(s, layer - 1);
        const int ldp = layer - 1 == 0 ? kLd0 : kMlpHid;
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            f32x4 dh = {0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < kMlpHid; k0 += 32) {
                s16x8 a, b;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int j = k0 + 8 * g + e;
                    a[e] = (short)s.dz[(16 * w + (l & 15)) * kMlpHid + j];
                    b[e] = (short)W[j * kMlpHid + 16 * kt + (l & 15)];
                }
                dh = mfma(a, b, dh);
            }
            const f32x4 z = fwd_tile(inp, ldp, ldp, Wp, ldp, kMlpHid, w, kt);
            const float bias = P[kBOff[layer - 1] + 16 * kt + (l & 15)];
#pragma unroll
            for (int r = 0; r < 4; ++r) dzv[kt][r] = dh[r] * gelu_grad(z[r] + bias);
        }
        __syncthreads();  // every read of the old dz is done
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                s.dz[(16 * w + 4 * g + r) * kMlpHid + 16 * kt + (l & 15)] = f2bf(dzv[kt][r]);
        column_sums(s, dzv, kMlpHid, gout + kBOff[layer - 1]);
    }
}

// Sum of the nb partial gradients of every parameter and of the partial
// losses, in a fixed order (the result does not depend on scheduling): a
// workgroup takes 64 parameters, its four waves each sum a quarter of the
// partials in order (rows of 64 consecutive floats), and the four sums are
// added in wave order.  Workgroup 0 also sums the partial losses: strided
// per thread, then a fixed tree.  (One thread per parameter summing all nb
// partials kept 35 workgroups busy and took 35 us per batch.)
constexpr int kReduceCols = 64;
__global__ __launch_bounds__(256) void k_mlp_reduce(const float *__restrict__ partial,
                                                    const float *__restrict__ lossp, int nb, float inv_norm,
                                                    float *__restrict__ grad, float *__restrict__ loss) {
    __shared__ float part[4][kReduceCols];
    __shared__ float lsum[256];
    const int t = threadIdx.x % kReduceCols, g = threadIdx.x / kReduceCols;
    const int p = blockIdx.x * kReduceCols + t;
    const int per = (nb + 3) / 4, b0 = g * per, b1 = b0 + per < nb ? b0 + per : nb;
    float acc = 0.0f;
    if (p < kMlpParams) {
#pragma unroll 8
        for (int b = b0; b < b1; ++b) acc = acc + partial[(size_t)b * kMlpParams + p];
    }
    part[g][t] = acc;
    if (blockIdx.x == 0) {
        float l = 0.0f;
        for (int b = threadIdx.x; b < nb; b += 256) l = l + lossp[b];
        lsum[threadIdx.x] = l;
    }
    __syncthreads();
    if (g == 0 && p < kMlpParams) grad[p] = ((part[0][t] + part[1][t]) + part[2][t]) + part[3][t];
    if (blockIdx.x == 0) {
        for (int h = 128; h > 0; h >>= 1) {
            if (threadIdx.x < h) lsum[threadIdx.x] = lsum[threadIdx.x] + lsum[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) loss[0] = lsum[0] * inv_norm;
    }
}

__global__ void k_mlp_adam(float *__restrict__ P, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ grad, AdamArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float gr = grad[p];
    const float mn = a.b1 * m[p] + (1.0f - a.b1) * gr;
    const float vn = a.b2 * v[p] + (1.0f - a.b2) * (gr * gr);
    m[p] = mn;
    v[p] = vn;
    const float mh = mn / a.c1, vh = vn / a.c2;
    P[p] = P[p] - a.lr * mh / (sqrtf(vh) + a.eps);
}

// ---------------------------------------------------------------------------
// inference: the network lives in registers as MFMA A operands.  Products are
// computed transposed, D = W h^T (rows = neurons, columns = 16 samples), so a
// layer's output D tiles are already the next layer's B operand once the
// contraction index is permuted: for k-block b, lane group g, element e the
// neuron is kappa(b, g, e) = 16 (2b + e/4) + 4g + e%4, which is exactly what
// the lane holds in its output tiles 2b and 2b+1.  No LDS is needed.
// ---------------------------------------------------------------------------
struct InferW {
    s16x8 w1[4];
    s16x8 w2[4][2], w3[4][2];
    s16x8 w4[2];
    float b1[16], b2[16], b3[16];  // [t * 4 + r]: neuron 16t + 4g + r
    float b4[4];                   // [r]: output 4g + r
};

__device__ __forceinline__ int kappa(int b, int g, int e) { return 16 * (2 * b + (e >> 2)) + 4 * g + (e & 3); }

__device__ void load_infer(const float *__restrict__ P, InferW &W) {
    const int l = lane_id(), i = l & 15, g = l >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = 8 * g + e;
            W.w1[t][e] = k < kMlpIn ? (short)f2bf(P[kMlpOffW1 + (16 * t + i) * kMlpIn + k]) : (short)0;
        }
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                W.w2[t][b][e] = (short)f2bf(P[kMlpOffW2 + (16 * t + i) * kMlpHid + kappa(b, g, e)]);
                W.w3[t][b][e] = (short)f2bf(P[kMlpOffW3 + (16 * t + i) * kMlpHid + kappa(b, g, e)]);
            }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            W.b1[t * 4 + r] = P[kMlpOffB1 + 16 * t + 4 * g + r];
            W.b2[t * 4 + r] = P[kMlpOffB2 + 16 * t + 4 * g + r];
            W.b3[t * 4 + r] = P[kMlpOffB3 + 16 * t + 4 * g + r];
        }
    }
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 8; ++e)
            W.w4[b][e] = i < kMlpOut ? (short)f2bf(P[kMlpOffW4 + i * kMlpHid + kappa(b, g, e)]) : (short)0;
#pragma unroll
    for (int r = 0; r < 4; ++r) W.b4[r] = (g == 0 && r < kMlpOut) ? P[kMlpOffB4 + r] : 0.0f;
}

__device__ __forceinline__ void activate(const f32x4 acc[4], const float *bias, s16x8 h[2]) {
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int t = 2 * b + (e >> 2), r = e & 3;
            h[b][e] = (short)f2bf(gelu(acc[t][r] + bias[t * 4 + r]));
        }
}

// 16 samples: x = layer-1 B operand (lane l, g = 0: features of sample l % 16;
// other lanes zero).  Returns D of the output layer: lanes with g = 0 hold
// outputs 0 and 1 of sample l % 16 in elements 0 and 1.
__device__ __forceinline__ f32x4 infer16(const InferW &W, const s16x8 &x) {
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 acc[4];
    s16x8 h[2];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = mfma(W.w1[t], x, zero);
    activate(acc, W.b1, h);
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = mfma(W.w2[t][1], h[1], mfma(W.w2[t][0], h[0], zero));
    activate(acc, W.b2, h);
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = mfma(W.w3[t][1], h[1], mfma(W.w3[t][0], h[0], zero));
    activate(acc, W.b3, h);
    f32x4 y = mfma(W.w4[1], h[1], mfma(W.w4[0], h[0], zero));
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] = y[r] + W.b4[r];
    return y;
}

// the network on the 64 samples of a wave (lane L: standardised inputs x[7]);
// returns this lane's two outputs
__device__ __forceinline__ void infer64(const InferW &W, const float x[kMlpIn], float &y0, float &y1) {
    const int l = lane_id(), g = l >> 4;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        s16x8 op;
        const int src = 16 * b + (l & 15);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float v = e < kMlpIn ? __shfl(x[e < kMlpIn ? e : 0], src) : 0.0f;
            op[e] = (g == 0 && e < kMlpIn) ? (short)f2bf(v) : (short)0;
        }
        const f32x4 y = infer16(W, op);
        const float o0 = __shfl(y[0], l & 15), o1 = __shfl(y[1], l & 15);
        if (g == b) { y0 = o0; y1 = o1; }
    }
}

__global__ __launch_bounds__(256) void k_mlp_infer(const float *__restrict__ P, const float *__restrict__ X,
                                                   float *__restrict__ Y, long n) {
    InferW W;
    load_infer(P, W);
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long ngroups = (n + 63) / 64;
    for (long grp = wave; grp < ngroups; grp += nwaves) {
        const long q = grp * 64 + lane_id();
        float x[kMlpIn];
#pragma unroll
        for (int k = 0; k < kMlpIn; ++k) x[k] = q < n ? X[q * kMlpIn + k] : 0.0f;
        float y0 = 0.0f, y1 = 0.0f;
        infer64(W, x, y0, y1);
        if (q < n) {
            Y[q * kMlpOut] = y0;
            Y[q * kMlpOut + 1] = y1;
        }
    }
}

// ---------------------------------------------------------------------------
// per-point features (FP64 -> FP32), shared by the solver closure and sampling
// ---------------------------------------------------------------------------
struct GradIn {
    const double *dU[3], *dV[3], *dW[3];
    Metrics m;
};

// d u_i / d x_j from d u_i / d xi_k, src/CurvilinearCSolver_Core.cpp:548-561;
// without metric terms (m.J11 null) the inputs are the Cartesian gradient itself
__device__ __forceinline__ void cart_grad(const GradIn &a, long p, double A[3][3]) {
    if (!a.m.J11) {
        const double *const *d[3] = {a.dU, a.dV, a.dW};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) A[i][j] = d[i][j][p];
        return;
    }
    const double J11 = a.m.J11[p], J12 = a.m.J12[p], J13 = a.m.J13[p];
    const double J21 = a.m.J21[p], J22 = a.m.J22[p], J23 = a.m.J23[p];
    const double J31 = a.m.J31[p], J32 = a.m.J32[p], J33 = a.m.J33[p];
    const double *const *d[3] = {a.dU, a.dV, a.dW};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double d0 = d[i][0][p], d1 = d[i][1][p], d2 = d[i][2][p];
        A[i][0] = J11 * d0 + J21 * d1 + J31 * d2;
        A[i][1] = J12 * d0 + J22 * d1 + J32 * d2;
        A[i][2] = J13 * d0 + J23 * d1 + J33 * d2;
    }
}

// raw features and |S| = sqrt(2 S_ij S_ij):
// theta/a, (s2 - w2)/a^2, s3/a^3, sw2/a^3, s2w2/a^4, log a, log(vol/J)/3
__device__ __forceinline__ void features(const double A[3][3], double vol_J, float f[kMlpIn], double &Smag) {
    double S[3][3], Wm[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            S[i][j] = 0.5 * (A[i][j] + A[j][i]);
            Wm[i][j] = 0.5 * (A[i][j] - A[j][i]);
        }
    double s2 = 0.0, w2 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            s2 = s2 + S[i][j] * S[i][j];
            w2 = w2 + Wm[i][j] * Wm[i][j];
        }
    double SS[3][3], WW[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            SS[i][j] = (S[i][0] * S[0][j] + S[i][1] * S[1][j]) + S[i][2] * S[2][j];
            WW[i][j] = (Wm[i][0] * Wm[0][j] + Wm[i][1] * Wm[1][j]) + Wm[i][2] * Wm[2][j];
        }
    double s3 = 0.0, sw2 = 0.0, s2w2 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            s3 = s3 + SS[i][j] * S[j][i];
            sw2 = sw2 + S[i][j] * WW[j][i];
            s2w2 = s2w2 + SS[i][j] * WW[j][i];
        }
    const double theta = (A[0][0] + A[1][1]) + A[2][2];
    const double a2 = s2 + w2;
    Smag = sqrt(2.0 * s2);
    if (a2 > 0.0) {
        const double a = sqrt(a2), inv = 1.0 / a;
        const double inv2 = inv * inv, inv3 = inv2 * inv;
        f[0] = (float)(theta * inv);
        f[1] = (float)((s2 - w2) * inv2);
        f[2] = (float)(s3 * inv3);
        f[3] = (float)(sw2 * inv3);
        f[4] = (float)(s2w2 * (inv2 * inv2));
        f[5] = (float)log(a);
    } else {
        f[0] = f[1] = f[2] = f[3] = f[4] = 0.0f;
        f[5] = (float)log(1e-300);
    }
    f[6] = (float)(log(vol_J) / 3.0);
}

// The solver's closure keeps the network in LDS and loads a wave's weight
// operands into registers only around the products of each 64-point group:
// the FP64 feature phase and the network phase then do not hold their
// registers at the same time (4 waves per SIMD instead of 1).  Per-lane
// operands are stored [operand][lane] as 16-byte entries (w1 t = 0..3, w2 t,b,
// w3 t,b, w4 b), biases by neuron.
constexpr int kWOps = 4 + 8 + 8 + 2;
struct InferLds {
    s16x8 w[kWOps][64];
    float b[3][kMlpHid];
    float b4[4];
};

// the per-lane operands of load_infer, written entry by entry by the whole
// workgroup (no register copy of the network)
__device__ void stage_infer(const float *__restrict__ P, InferLds &s) {
    for (int idx = threadIdx.x; idx < kWOps * 64 * 8; idx += blockDim.x) {
        const int e = idx & 7, l = (idx >> 3) & 63, op = idx >> 9;
        const int i = l & 15, g = l >> 4;
        float v;
        if (op < 4) {
            const int k = 8 * g + e;
            v = k < kMlpIn ? P[kMlpOffW1 + (16 * op + i) * kMlpIn + k] : 0.0f;
        } else if (op < 20) {
            const int q = op < 12 ? op - 4 : op - 12, t = q >> 1, b = q & 1;
            v = P[(op < 12 ? kMlpOffW2 : kMlpOffW3) + (16 * t + i) * kMlpHid + kappa(b, g, e)];
        } else {
            v = i < kMlpOut ? P[kMlpOffW4 + i * kMlpHid + kappa(op - 20, g, e)] : 0.0f;
        }
        reinterpret_cast<short *>(&s.w[op][l])[e] = (short)f2bf(v);
    }
    for (int k = threadIdx.x; k < 3 * kMlpHid; k += blockDim.x) {
        const int layer = k / kMlpHid, j = k % kMlpHid;
        s.b[layer][j] = P[(layer == 0 ? kMlpOffB1 : layer == 1 ? kMlpOffB2 : kMlpOffB3) + j];
    }
    if (threadIdx.x < 4) s.b4[threadIdx.x] = threadIdx.x < kMlpOut ? P[kMlpOffB4 + threadIdx.x] : 0.0f;
    __syncthreads();
}

// FP32 -> BF16 by the hardware conversion (round to nearest even, as f2bf for
// finite values)
__device__ __forceinline__ short bf16_bits(float v) { return __builtin_bit_cast(short, (__bf16)v); }

__device__ __forceinline__ void activate_lds(const f32x4 acc[4], const float *bias, int g, s16x8 h[2]) {
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int t = 2 * b + (e >> 2), r = e & 3;
            h[b][e] = bf16_bits(gelu(acc[t][r] + bias[16 * t + 4 * g + r]));
        }
}

// The network on the 64 points of a wave (lane L: standardised inputs x[7]),
// layer by layer over the wave's four 16-point blocks: only one layer's weight
// operands are in registers at a time, loaded from LDS right before its
// products.  Returns this lane's two outputs.
__device__ __forceinline__ void infer64_lds(const InferLds &s, const float x[kMlpIn], float &y0, float &y1) {
    const int l = lane_id(), g = l >> 4;
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    // compiler-only barrier: the operand loads below stay here instead of being
    // hoisted out of the caller's loop (which would keep them live through the
    // feature phase)
    __asm__ volatile("" ::: "memory");
    s16x8 h[4][2];
    {  // layer 1: block b's inputs are the features of points 16 b .. 16 b + 15
        s16x8 w[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) w[t] = s.w[t][l];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            s16x8 op;
            const int src = 16 * b + (l & 15);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float v = e < kMlpIn ? __shfl(x[e < kMlpIn ? e : 0], src) : 0.0f;
                op[e] = (g == 0 && e < kMlpIn) ? bf16_bits(v) : (short)0;
            }
            f32x4 acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = mfma(w[t], op, zero);
            activate_lds(acc, s.b[0], g, h[b]);
        }
    }
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {  // layers 2 and 3
        s16x8 w[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) w[k] = s.w[4 + 8 * layer + k][l];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            f32x4 acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = mfma(w[2 * t + 1], h[b][1], mfma(w[2 * t], h[b][0], zero));
            activate_lds(acc, s.b[1 + layer], g, h[b]);
        }
    }
    const s16x8 w40 = s.w[20][l], w41 = s.w[21][l];
    const float b4 = g == 0 ? s.b4[0] : 0.0f, b4b = g == 0 ? s.b4[1] : 0.0f;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const f32x4 y = mfma(w41, h[b][1], mfma(w40, h[b][0], zero));
        const float o0 = __shfl(y[0] + b4, l & 15), o1 = __shfl(y[1] + b4b, l & 15);
        if (g == b) { y0 = o0; y1 = o1; }
    }
}

__global__ __launch_bounds__(256) void k_closure_features(FeatureArgs a, long n) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const GradIn gi{{a.dU[0], a.dU[1], a.dU[2]}, {a.dV[0], a.dV[1], a.dV[2]}, {a.dW[0], a.dW[1], a.dW[2]}, a.m};
    double A[3][3];
    cart_grad(gi, p, A);
    float f[kMlpIn];
    double Smag;
    const double vol_J = a.vol / a.m.J[p];
    features(A, vol_J, f, Smag);
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) __builtin_nontemporal_store((double)f[k], a.feat[k] + p);
    if (a.smag) {
        const double d = cbrt(vol_J);
        __builtin_nontemporal_store(Smag, a.smag + p);
        __builtin_nontemporal_store(d * d, a.delta2 + p);
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_learned_sgs(ClosureArgs a, long n) {
    __shared__ InferLds net;
    // the input range and standardisation in LDS: read per group instead of
    // held in scalar registers, which spilled to vector-register lanes
    __shared__ MlpScaling sc;
    if (threadIdx.x < sizeof(MlpScaling) / sizeof(float))
        reinterpret_cast<float *>(&sc)[threadIdx.x] = reinterpret_cast<const float *>(&a.sc)[threadIdx.x];
    stage_infer(a.params, net);
    __shared__ int wcount[4];
    const int l = lane_id(), w = threadIdx.x >> 6;
    const long nwaves = (long)gridDim.x * 4;
    const long ngroups = (n + 63) / 64;
    const GradIn gi{{a.dU[0], a.dU[1], a.dU[2]}, {a.dV[0], a.dV[1], a.dV[2]}, {a.dW[0], a.dW[1], a.dW[2]}, a.m};
    int count = 0;
    for (long grp = (long)blockIdx.x * 4 + w; grp < ngroups; grp += nwaves) {
        const long p = grp * 64 + l;
        const bool live = p < n;
        float x[kMlpIn];
        double base = 0.0;
        bool out = false;
        if (live) {
            float f[kMlpIn];
            double Smag, d2;
            if (a.feat[0]) {  // averaged inputs; |S| and Delta^2 of the point from the feature pass
#pragma unroll
                for (int k = 0; k < kMlpIn; ++k) f[k] = (float)a.feat[k][p];
                Smag = a.smag[p];
                d2 = a.delta2[p];
            } else {
                double A[3][3];
                cart_grad(gi, p, A);
                const double vol_J = a.vol / a.m.J[p];
                features(A, vol_J, f, Smag);
                const double d = cbrt(vol_J);
                d2 = d * d;
            }
#pragma unroll
            for (int k = 0; k < kMlpIn; ++k) {
                const bool in = f[k] >= sc.lo[k] && f[k] <= sc.hi[k];
                out = out || !in;
                // outside the trusted range the input is clamped to it
                const float fc = in ? f[k] : (f[k] < sc.lo[k] ? sc.lo[k] : sc.hi[k]);
                x[k] = (fc - sc.in_mean[k]) / sc.in_std[k];
            }
            base = (a.rho[p] * d2) * Smag;
        } else {
#pragma unroll
            for (int k = 0; k < kMlpIn; ++k) x[k] = 0.0f;
        }
        float y0 = 0.0f, y1 = 0.0f;
        infer64_lds(net, x, y0, y1);
        if (live) {
            const float C = y0 * sc.out_std[0] + sc.out_mean[0];
            const float D = y1 * sc.out_std[1] + sc.out_mean[1];
            a.mu_sgs[p] = base * (double)C;
            a.k_sgs[p] = (a.cp * base) * (double)D;
            a.flag[p] = out ? 1 : 0;
        }
        count += __popcll(__ballot(live && out));
    }
    if (l == 0) wcount[w] = count;
    __syncthreads();
    if (threadIdx.x == 0) a.blockcount[blockIdx.x] = ((wcount[0] + wcount[1]) + wcount[2]) + wcount[3];
}

// the network on lines (LineClosureArgs); the clamp, standardisation and
// de-standardisation of k_learned_sgs
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_line_closure(LineClosureArgs a) {
    __shared__ InferLds net;
    stage_infer(a.params, net);
    __shared__ int wcount[4];
    const int l = lane_id(), w = threadIdx.x >> 6;
    const long nwaves = (long)gridDim.x * 4;
    const long ngroups = (a.nlines + 63) / 64;
    int count = 0;
    for (long grp = (long)blockIdx.x * 4 + w; grp < ngroups; grp += nwaves) {
        const long q = grp * 64 + l;
        const bool live = q < a.nlines;
        float x[kMlpIn];
        bool out = false;
#pragma unroll
        for (int k = 0; k < kMlpIn; ++k) {
            const float f = live ? (float)a.feat[k][q] : 0.0f;
            const bool in = f >= a.sc.lo[k] && f <= a.sc.hi[k];
            out = out || !in;
            const float fc = in ? f : (f < a.sc.lo[k] ? a.sc.lo[k] : a.sc.hi[k]);
            x[k] = live ? (fc - a.sc.in_mean[k]) / a.sc.in_std[k] : 0.0f;
        }
        float y0 = 0.0f, y1 = 0.0f;
        infer64_lds(net, x, y0, y1);
        if (live) {
            a.C[q] = y0 * a.sc.out_std[0] + a.sc.out_mean[0];
            a.D[q] = y1 * a.sc.out_std[1] + a.sc.out_mean[1];
            a.lflag[q] = out ? 1 : 0;
        }
        count += __popcll(__ballot(live && out));
    }
    if (l == 0) wcount[w] = count;
    __syncthreads();
    if (threadIdx.x == 0) a.blockcount[blockIdx.x] = (((wcount[0] + wcount[1]) + wcount[2]) + wcount[3]) * a.line_len;
}

__global__ __launch_bounds__(256) void k_closure_apply(ClosureApplyArgs a, long n) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const long q = xi_line_of(p, a.dir, a.Nx, a.Ny);
    const double base = (ldnt(a.rho, p) * ldnt(a.delta2, p)) * ldnt(a.smag, p);
    stnt(a.mu_sgs, p, base * (double)a.C[q]);
    stnt(a.k_sgs, p, (a.cp * base) * (double)a.D[q]);
    a.flag[p] = a.lflag[q];
}

__global__ void k_count_sum(const int *__restrict__ bc, int nb, int *__restrict__ count) {
    __shared__ int part[256];
    int acc = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += bc[b];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) count[0] = part[0];
}

__global__ void k_closure_fallback(FallbackArgs a, long n) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !a.flag[p]) return;
    a.mu_sgs[p] = a.mu_dsm[p];
    a.k_sgs[p] = a.k_dsm[p];
}

__global__ void k_closure_samples(SampleArgs a) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    const long p = a.offset + i * a.stride;
    const GradIn gi{{a.dU[0], a.dU[1], a.dU[2]}, {a.dV[0], a.dV[1], a.dV[2]}, {a.dW[0], a.dW[1], a.dW[2]}, a.m};
    float f[kMlpIn];
    if (a.feat[0]) {
#pragma unroll
        for (int k = 0; k < kMlpIn; ++k) f[k] = (float)a.feat[k][p];
    } else {
        double A[3][3];
        cart_grad(gi, p, A);
        double Smag;
        features(A, a.vol / a.m.J[p], f, Smag);
    }
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) a.X[i * kMlpIn + k] = f[k];
    const double C = a.C[p];
    a.T[i * kMlpOut] = (float)C;
    a.T[i * kMlpOut + 1] = (float)(C / a.Prt[p]);
}

}  // namespace

hipError_t launch_closure_features(const FeatureArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_closure_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n);
    return hipGetLastError();
}

hipError_t launch_mlp_grad_partials(const float *params, const float *X, const float *T, const int *idx, long n,
                                    long n_total, float *partial, float *lossp, hipStream_t st) {
    if (n <= 0) return hipErrorInvalidValue;
    const unsigned nb = (unsigned)((n + kT - 1) / kT);
    const float scale = (float)(2.0 / ((double)n_total * kMlpOut));
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(256), 0, st, params, X, T, idx, n, scale, partial, lossp);
    return hipGetLastError();
}

hipError_t launch_mlp_reduce(const float *partial, const float *lossp, int nblocks, long n_total, float *grad,
                             float *loss, hipStream_t st) {
    const float inv_norm = (float)(1.0 / ((double)n_total * kMlpOut));
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + kReduceCols - 1) / kReduceCols), dim3(256), 0, st, partial,
                       lossp, nblocks, inv_norm, grad, loss);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float *params, float *m, float *v, const float *grad, AdamArgs a, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, params, m, v, grad, a);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const float *params, const float *X, float *Y, long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const long groups = (n + 63) / 64;
    const unsigned nb = (unsigned)std::min<long>((groups + 3) / 4, 2048);
    hipLaunchKernelGGL(k_mlp_infer, dim3(nb), dim3(256), 0, st, params, X, Y, n);
    return hipGetLastError();
}

int closure_blocks(long n) {
    const long groups = (n + 63) / 64;
    return (int)std::max<long>(1, std::min<long>((groups + 3) / 4, 2048));
}

int line_closure_blocks(long nlines) { return closure_blocks(nlines); }

hipError_t launch_line_closure(const LineClosureArgs &a, hipStream_t st) {
    if (a.nlines < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_line_closure, dim3((unsigned)line_closure_blocks(a.nlines)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_closure_apply(const ClosureApplyArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_closure_apply, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n);
    return hipGetLastError();
}

hipError_t launch_learned_sgs(const ClosureArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_learned_sgs, dim3((unsigned)closure_blocks(n)), dim3(256), 0, st, a, n);
    return hipGetLastError();
}

hipError_t launch_count_sum(const int *blockcount, int nb, int *count, hipStream_t st) {
    hipLaunchKernelGGL(k_count_sum, dim3(1), dim3(256), 0, st, blockcount, nb, count);
    return hipGetLastError();
}

hipError_t launch_closure_fallback(const FallbackArgs &a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_closure_fallback, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n);
    return hipGetLastError();
}

hipError_t launch_closure_samples(const SampleArgs &a, hipStream_t st) {
    if (a.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_closure_samples, dim3((unsigned)((a.count + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace mc3d
