// Learned diagonal preconditioner for Davidson: a 64-128-128-1 tanh MLP with
// bf16 weights and FP32 accumulation on MFMA (the forward pass on gfx950's
// v_mfma_f32_16x16x32_bf16, the backward products on
// v_mfma_f32_16x16x16_bf16), trained and evaluated on the GPU.  No reference
// counterpart.
//
// The network maps per-state features to the log of a diagonal
// preconditioner, t_i = r_i e^{m_i}, replacing Jacobi's 1 / (theta - H_ii).
// Training maximises the one-step Ritz gain of t (loss kernels below) on
// snapshots of Jacobi Davidson runs on small lattices.
//
// MFMA 16x16x16 operand layout (per 64-lane wave): A[i][k] with i = lane % 16,
// k = 4 (lane / 16) + j; B[k][n] with n = lane % 16, k = 4 (lane / 16) + j;
// D[i][n] with n = lane % 16, i = 4 (lane / 16) + j; j = 0..3 the element
// within the lane's vector.  16x16x32: A and B hold eight elements per lane,
// element j of lane group g = lane / 16 multiplying element j of the same
// group in the other operand (the contraction index of slot (g, j) is the
// caller's choice, the same in both operands); D as for 16x16x16.
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "tcp_internal.hpp"

namespace {

constexpr uint32_t kF = 64;   // input features
constexpr uint32_t kH = 128;  // hidden width
constexpr uint32_t kP = kF * kH + kH + kH * kH + kH + kH + 1;
constexpr uint32_t kOffW1 = 0, kOffB1 = kF * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
                   kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH;
constexpr uint32_t kThreads = 256;
// The forward pass: eight waves share one LDS copy of the weights (48 KiB);
// at its register count two workgroups fit a CU, four waves per SIMD.
constexpr uint32_t kFwdThreads = 512;
// The Davidson callback re-evaluates the network when theta has moved by more
// than kRefresh times the coupling scale since its last evaluation (default;
// tcp_learned_precond_set_refresh changes it).  A preconditioner for a
// somewhat different theta is as good a search direction: on the seven
// sectors the network is trained on (4x4, 3x6 and 4x5 tori) the learned
// Davidson iteration counts at subspaces 12/4 and 24/4 are the same for
// every threshold from 0 (evaluate at every iteration) to 4, the largest
// such value of those tried (at 8 a few sectors take one iteration more;
// evaluating once, at the start vector's theta, does not converge), while the
// evaluations fall several-fold.
constexpr double kRefresh = 4.0;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double dvec2 __attribute__((ext_vector_type(2)));

__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

// tanh x = 1 - 2 / (e^{2x} + 1) from the hardware base-2 exponential and
// reciprocal (absolute error ~1e-7; the result is rounded to bf16 or feeds an
// FP32 dot product).  Saturates correctly: e^{2x} -> inf gives 1, -> 0 gives -1.
// The forward kernel applies it as 1 - 2 rcp(2^{a k + b k} + 1) with k =
// kTwoLog2e and the bias b pre-scaled by k, one FMA in place of add and mul.
constexpr float kTwoLog2e = 2.8853900817779268f;  // 2 / ln 2
__device__ __forceinline__ float tanh_denominator_rcp(float scaled) {  // 1 / (2^scaled + 1)
    return __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(scaled) + 1.0f);
}

typedef unsigned u2 __attribute__((ext_vector_type(2)));
// bf16(lo) | bf16(hi) << 16, round to nearest even: one v_cvt_pk_bf16_f32
// (written out because the compiler splits a converted float pair into two
// half-used conversions and a permute).
__device__ __forceinline__ unsigned cvt_pk_bf16(float lo, float hi) {
    unsigned r;
    asm("v_cvt_pk_bf16_f32 %0, %1, %2" : "=v"(r) : "v"(lo), "v"(hi));
    return r;
}
__device__ __forceinline__ s4 pack_bf16(f4 v) {
    return __builtin_bit_cast(s4, (u2{cvt_pk_bf16(v.x, v.y), cvt_pk_bf16(v.z, v.w)}));
}

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f4 mfma32(s8 a, s8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0,
                                                   0);
}
// Eight bf16 values: lo[0..3], hi[0..3].
__device__ __forceinline__ s8 pack_bf16x8(f4 lo, f4 hi) {
    return __builtin_bit_cast(
        s8, (u4{cvt_pk_bf16(lo.x, lo.y), cvt_pk_bf16(lo.z, lo.w), cvt_pk_bf16(hi.x, hi.y), cvt_pk_bf16(hi.z, hi.w)}));
}

// Forward-pass fragments (16x16x32).  Layer 1, block n of 16 hidden units and
// block ks of 32 features: slot (lane, j) holds W1[32 ks + 8 g + j][16 n + lane % 16]
// (features 8g..8g+7 of a row are one 16-byte load of X).  Layer 2, block m of
// 16 outputs and block q of 32 hidden inputs: slot (lane, j) holds
// W2[k2(q, g, j)][16 m + lane % 16] with k2 = 32 q + 4 g + j for j < 4 and
// 32 q + 16 + 4 g + (j - 4) for j >= 4 -- the units a lane holds in the layer-1
// results of blocks 2q and 2q + 1, so those results are the layer-2 B operand
// as they stand.
__host__ __device__ constexpr uint32_t w2_input(uint32_t q, uint32_t g, uint32_t j) {
    return 32 * q + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

struct MlpDev {
    const s8 *w1f;   // [8 n][2 ks][64 lane] forward fragments of W1 (bf16 x8)
    const s8 *w2f;   // [8 m][4 q][64 lane] forward fragments of W2
    const uint16_t *w2b;  // W2 bf16, row-major [in][out]
    const float *params;  // FP32 master parameters
};

// Fragment-major bf16 copies of W1, W2 and a plain bf16 copy of W2.
__global__ void swizzle_kernel(const float *__restrict__ params, uint16_t *__restrict__ w1f,
                               uint16_t *__restrict__ w2f, uint16_t *__restrict__ w2b) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    // W1: 8 * 2 * 64 * 8 = 8192 entries
    if (tid < 8 * 2 * 64 * 8) {
        const uint32_t j = tid & 7, lane = (tid >> 3) & 63, ks = (tid >> 9) & 1, n = tid >> 10;
        const uint32_t k = ks * 32 + 8 * (lane >> 4) + j, col = n * 16 + (lane & 15);
        w1f[tid] = f2bf(params[kOffW1 + k * kH + col]);
    }
    // W2: 8 * 4 * 64 * 8 = 16384 entries
    if (tid < 8 * 4 * 64 * 8) {
        const uint32_t j = tid & 7, lane = (tid >> 3) & 63, q = (tid >> 9) & 3, m = tid >> 11;
        const uint32_t k = w2_input(q, lane >> 4, j), col = m * 16 + (lane & 15);
        w2f[tid] = f2bf(params[kOffW2 + k * kH + col]);
        w2b[tid] = f2bf(params[kOffW2 + tid]);
    }
}

// Feature 0 of every state: (diag[r] - theta) * inv_scale, in place.
__global__ void fill_feature0_kernel(uint16_t *__restrict__ X, const double *__restrict__ diag, double theta,
                                     double inv_scale, uint64_t B) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < B;
         r += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        X[r * kF] = f2bf(static_cast<float>((diag[r] - theta) * inv_scale));
}

// y[r] = MLP(X[r]); X is bf16 [B x 64].  With h1/h2, the post-activation
// hidden layers are stored (FP32 [B x 128]) for the backward pass.  With diag,
// feature 0 is taken as (diag[r] - theta) * inv_scale instead of X[r][0].
//
// Both layers run transposed, H^T = W^T X^T: the weight fragments are the A
// operand and 16 rows of X (or of H1) the B operand.  The MFMA result then
// holds, in each lane, hidden units 4 (lane / 16) + j of one 16-block for row
// lane % 16; two such blocks are the next layer's B fragment (the W2
// fragments follow that unit order), so the first hidden layer goes from
// accumulator to bf16 fragment in registers (no LDS staging, no barrier).
// Each wave takes R tiles of 16 rows per weight-fragment read from LDS.  The
// output layer is a dot product over the lanes that share a row.
// Without STORE (inference) the output needs only sum_u w3_u tanh z_u =
// sum_u w3_u - 2 sum_u w3_u / (e^{2 z_u} + 1): each lane starts its partial
// sum at its share of sum_u w3_u and adds one FMA per hidden unit.
template <bool STORE, int R>
__global__ __launch_bounds__(kFwdThreads) void mlp_forward_kernel(MlpDev p, const uint16_t *__restrict__ X,
                                                               uint64_t B, float *__restrict__ y,
                                                               float *__restrict__ h1, float *__restrict__ h2,
                                                               const double *__restrict__ diag, double theta,
                                                               double inv_scale) {
    __shared__ s8 w1s[8 * 2 * 64];
    __shared__ s8 w2s[8 * 4 * 64];
    __shared__ float bs1[kH], bs2[kH], ws3[kH];  // biases times kTwoLog2e; w3 (times -2 without STORE)
    for (uint32_t i = threadIdx.x; i < 8 * 2 * 64; i += kFwdThreads) w1s[i] = p.w1f[i];
    for (uint32_t i = threadIdx.x; i < 8 * 4 * 64; i += kFwdThreads) w2s[i] = p.w2f[i];
    for (uint32_t i = threadIdx.x; i < kH; i += kFwdThreads) {
        bs1[i] = p.params[kOffB1 + i] * kTwoLog2e;
        bs2[i] = p.params[kOffB2 + i] * kTwoLog2e;
        ws3[i] = STORE ? p.params[kOffW3 + i] : -2.0f * p.params[kOffW3 + i];
    }
    const float b3 = p.params[kOffB3];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t li = lane & 15, lg = lane >> 4;
    float base = 0.f;
    if (!STORE)
        for (int n = 0; n < 8; ++n)
            for (int j = 0; j < 4; ++j) base += p.params[kOffW3 + n * 16 + lg * 4 + j];
    constexpr uint64_t kRows = (kFwdThreads / 64) * 16 * R;  // rows per workgroup step: R tiles of 16 per wave
    const uint64_t ntile = (B + kRows - 1) / kRows;
    // The next step's rows are loaded while this step computes (with three
    // waves per SIMD the HBM latency of X would otherwise stall each step).
    s8 xn[R][2];
    double dn[R];
    auto load_rows = [&](uint64_t tile) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint64_t rr = tile * kRows + (wave * R + r) * 16 + li;
            const bool ok = tile < ntile && rr < B;
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                xn[r][ks] = ok ? *reinterpret_cast<const s8 *>(X + rr * kF + ks * 32 + lg * 8)
                               : s8{0, 0, 0, 0, 0, 0, 0, 0};
            dn[r] = diag && ok && lg == 0 ? diag[rr] : 0.0;
        }
    };
    load_rows(blockIdx.x);
    for (uint64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        // keeps the weight fragments in LDS: without it the compiler hoists the
        // loop-invariant fragment loads into registers (occupancy 1)
        asm volatile("" ::: "memory");
        uint64_t row[R];
        bool valid[R];
        s8 x[R][2];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            row[r] = tile * kRows + (wave * R + r) * 16 + li;
            valid[r] = row[r] < B;
            x[r][0] = xn[r][0];
            x[r][1] = xn[r][1];
            if (diag && valid[r] && lg == 0)
                x[r][0][0] = static_cast<short>(f2bf(static_cast<float>((dn[r] - theta) * inv_scale)));
        }
        load_rows(tile + gridDim.x);
        f4 acc[R][8];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int n = 0; n < 8; ++n) acc[r][n] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                const s8 a = w1s[(n * 2 + ks) * 64 + lane];
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r][n] = mfma32(a, x[r][ks], acc[r][n]);
            }
        s8 hb[R][4];  // layer-1 activations of blocks 2q, 2q + 1 as the layer-2 B fragment q
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                f4 v[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int n = 2 * q + h;
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        v[h][j] = fmaf(-2.0f,
                                       tanh_denominator_rcp(fmaf(acc[r][n][j], kTwoLog2e, bs1[n * 16 + lg * 4 + j])),
                                       1.0f);
                    if (STORE && valid[r]) *reinterpret_cast<f4 *>(h1 + row[r] * kH + n * 16 + lg * 4) = v[h];
                }
                hb[r][q] = pack_bf16x8(v[0], v[1]);
            }
        f4 acc2[R][8];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int m = 0; m < 8; ++m) acc2[r][m] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const s8 a = w2s[(m * 4 + q) * 64 + lane];
#pragma unroll
                for (int r = 0; r < R; ++r) acc2[r][m] = mfma32(a, hb[r][q], acc2[r][m]);
            }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float part = base;
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                f4 v;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t u = m * 16 + lg * 4 + j;
                    const float rc = tanh_denominator_rcp(fmaf(acc2[r][m][j], kTwoLog2e, bs2[u]));
                    if (STORE) {
                        v[j] = fmaf(-2.0f, rc, 1.0f);
                        part = fmaf(v[j], ws3[u], part);
                    } else {
                        part = fmaf(ws3[u], rc, part);
                    }
                }
                if (STORE && valid[r]) *reinterpret_cast<f4 *>(h2 + row[r] * kH + m * 16 + lg * 4) = v;
            }
            part += __shfl_xor(part, 16);
            part += __shfl_xor(part, 32);
            if (lg == 0 && valid[r]) y[row[r]] = part + b3;
        }
    }
}

// dA2 = dy * w3 * (1 - h2^2) (bf16 out); per-block partial sums of
// dw3 = sum dy h2, db2 = sum dA2, db3 = sum dy.
__global__ __launch_bounds__(kThreads) void act2_backward_kernel(const float *__restrict__ params,
                                                                 const float *__restrict__ dy,
                                                                 const float *__restrict__ h2, uint64_t B,
                                                                 uint16_t *__restrict__ da2,
                                                                 float *__restrict__ partial) {
    // 256 threads: column u = tid % 128, row phase = tid / 128 (2 rows in flight)
    const uint32_t u = threadIdx.x & (kH - 1), phase = threadIdx.x >> 7;
    const float w3 = params[kOffW3 + u];
    float gw3 = 0.f, gb2 = 0.f, gb3 = 0.f;
    for (uint64_t r = static_cast<uint64_t>(blockIdx.x) * 2 + phase; r < B; r += static_cast<uint64_t>(gridDim.x) * 2) {
        const float d = dy[r];
        const float v = h2[r * kH + u];
        const float g = d * w3 * (1.f - v * v);
        da2[r * kH + u] = f2bf(g);
        gw3 += d * v;
        gb2 += g;
        if (u == 0) gb3 += d;
    }
    __shared__ float red[2][3 * kH + 1];
    red[phase][u] = gw3;
    red[phase][kH + u] = gb2;
    if (u == 0) red[phase][3 * kH] = gb3;
    __syncthreads();
    if (phase == 0) {
        float* out = partial + static_cast<uint64_t>(blockIdx.x) * (2 * kH + 1);
        out[u] = red[0][u] + red[1][u];
        out[kH + u] = red[0][kH + u] + red[1][kH + u];
        if (u == 0) out[2 * kH] = red[0][3 * kH] + red[1][3 * kH];
    }
}

// dA1 = (dA2 W2^T) * (1 - h1^2), bf16 out; per-block partial db1 = sum dA1.
__global__ __launch_bounds__(kThreads) void h1_backward_kernel(MlpDev p, const uint16_t *__restrict__ da2,
                                                               const float *__restrict__ h1, uint64_t B,
                                                               uint16_t *__restrict__ da1,
                                                               float *__restrict__ partial) {
    __shared__ float colsum[4][kH];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t li = lane & 15, lg = lane >> 4;
    float gb1[8][4] = {};
    const uint64_t ntile = (B + 63) / 64;
    for (uint64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const uint64_t row0 = tile * 64 + wave * 16;
        const uint64_t row = row0 + li;
        f4 acc[8];
#pragma unroll
        for (int n = 0; n < 8; ++n) acc[n] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            const s4 a = row < B ? *reinterpret_cast<const s4 *>(da2 + row * kH + ks * 16 + lg * 4) : s4{0, 0, 0, 0};
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                // B[k = u][n = i] = W2[i][u]: contiguous along u
                const s4 b = *reinterpret_cast<const s4 *>(p.w2b + (n * 16 + li) * kH + ks * 16 + lg * 4);
                acc[n] = mfma(a, b, acc[n]);
            }
        }
#pragma unroll
        for (int n = 0; n < 8; ++n)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint64_t r = row0 + lg * 4 + j;
                if (r < B) {
                    const float v = h1[r * kH + n * 16 + li];
                    const float g = acc[n][j] * (1.f - v * v);
                    da1[r * kH + n * 16 + li] = f2bf(g);
                    gb1[n][j] += g;
                }
            }
    }
    // column sums: reduce over j and over the 4 lane groups (same li), then waves
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        float s = gb1[n][0] + gb1[n][1] + gb1[n][2] + gb1[n][3];
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (lg == 0) colsum[wave][n * 16 + li] = s;
    }
    __syncthreads();
    if (threadIdx.x < kH)
        partial[static_cast<uint64_t>(blockIdx.x) * kH + threadIdx.x] =
            colsum[0][threadIdx.x] + colsum[1][threadIdx.x] + colsum[2][threadIdx.x] + colsum[3][threadIdx.x];
}

// Weight gradient G[M x 128] = sum_r A[r][m] * D[r][n] over this block's rows
// (A^T D with the batch as the contraction), A given as bf16 (Abf) or FP32
// (Af) rows of length M, D bf16 rows of length 128.  Each wave owns
// M/64 row tiles x 8 column tiles; partial[block][M x 128].
template <uint32_t M>
__global__ __launch_bounds__(kThreads) void wgrad_kernel(const uint16_t *__restrict__ Abf, const float *__restrict__ Af,
                                                         const uint16_t *__restrict__ Dm, uint64_t B, uint64_t rows_per_block,
                                                         float *__restrict__ partial) {
    constexpr uint32_t TM = M / 16 / 4;  // row tiles per wave
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t li = lane & 15, lg = lane >> 4;
    f4 acc[TM][8];
#pragma unroll
    for (uint32_t a = 0; a < TM; ++a)
#pragma unroll
        for (int n = 0; n < 8; ++n) acc[a][n] = f4{0.f, 0.f, 0.f, 0.f};
    const uint64_t r_begin = static_cast<uint64_t>(blockIdx.x) * rows_per_block;
    const uint64_t r_end = r_begin + rows_per_block < B ? r_begin + rows_per_block : B;
    for (uint64_t r0 = r_begin; r0 < r_end; r0 += 16) {
        s4 bfrag[8];
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            s4 b;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint64_t r = r0 + lg * 4 + j;
                b[j] = r < r_end ? static_cast<short>(Dm[r * kH + n * 16 + li]) : 0;
            }
            bfrag[n] = b;
        }
#pragma unroll
        for (uint32_t a = 0; a < TM; ++a) {
            const uint32_t m = (wave * TM + a) * 16 + li;
            s4 afrag;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint64_t r = r0 + lg * 4 + j;
                short v = 0;
                if (r < r_end) v = Abf ? static_cast<short>(Abf[r * M + m]) : static_cast<short>(f2bf(Af[r * M + m]));
                afrag[j] = v;
            }
#pragma unroll
            for (int n = 0; n < 8; ++n) acc[a][n] = mfma(afrag, bfrag[n], acc[a][n]);
        }
    }
    float *out = partial + static_cast<uint64_t>(blockIdx.x) * M * kH;
#pragma unroll
    for (uint32_t a = 0; a < TM; ++a)
#pragma unroll
        for (int n = 0; n < 8; ++n)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t m = (wave * TM + a) * 16 + lg * 4 + j;
                out[m * kH + n * 16 + li] = acc[a][n][j];
            }
}

// out[i] = scale * sum_b partial[b * width + i] (fixed order)
__global__ void reduce_columns_kernel(const float *__restrict__ partial, uint32_t nblk, uint32_t width,
                                      float *__restrict__ out, uint32_t out_offset) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= width) return;
    double s = 0.0;
    for (uint32_t b = 0; b < nblk; ++b) s += partial[static_cast<uint64_t>(b) * width + i];
    out[out_offset + i] = static_cast<float>(s);
}

__global__ void adam_kernel(float *__restrict__ params, const float *__restrict__ grad, float *__restrict__ m,
                            float *__restrict__ v, uint32_t n, float lr, float beta1, float beta2, float eps,
                            float corr1, float corr2, float weight_decay) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i] + weight_decay * params[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * g;
    const float vi = beta2 * v[i] + (1.f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr * (mi / corr1) / (sqrtf(vi / corr2) + eps);
}

// ---- preconditioner loss: one-step Ritz gain ------------------------------
//
// For a Davidson state (normalised Ritz vector u, Ritz value theta, residual
// r = H u - theta u) and a diagonal preconditioner t_i = r_i e^{m_i}, the new
// direction q = t - u <u, t> extends the subspace span{u} to span{u, q}; the
// lowest Ritz value of that 2 x 2 problem is theta - Delta with
//   gamma = <q, H q> / |q|^2,  |rho|^2 = |<q, r>|^2 / |q|^2,  a = (gamma - theta) / 2,
//   Delta = sqrt(a^2 + |rho|^2) - a.
// The loss is L = -log Delta (scale-free in t).  Its gradient with respect to
// q (real embedding, g = dL/dRe q + i dL/dIm q) is
//   g_q = A1 (Hq - gamma q) + A2 conj(c) r + A3 q,   c = <q, r>, n2 = |q|^2,
//   A1 = (1 - a/R) / (n2 Delta),  A2 = -1 / (n2 R Delta),  A3 = |c|^2 / (n2^2 R Delta),
// R = sqrt(a^2 + |rho|^2); then g_t = g_q - u <u, g_q> and dL/dm_i = Re(conj(g_t,i) t_i).

__device__ __forceinline__ dvec2 cmulc(dvec2 a, dvec2 b) {  // conj(a) * b
    return dvec2{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}

// Block sum of two complex values, written to partial[2 * blockIdx + {0, 1}].
__device__ __forceinline__ void block_sum2(dvec2 v0, dvec2 v1, dvec2 *__restrict__ partial) {
    __shared__ dvec2 red[2][kThreads / 64];
    for (int off = 32; off > 0; off >>= 1) {
        v0.x += __shfl_xor(v0.x, off);
        v0.y += __shfl_xor(v0.y, off);
        v1.x += __shfl_xor(v1.x, off);
        v1.y += __shfl_xor(v1.y, off);
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wave] = v0;
        red[1][wave] = v1;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        dvec2 s{0.0, 0.0};
        for (uint32_t w = 0; w < kThreads / 64; ++w) s += red[threadIdx.x][w];
        partial[2 * blockIdx.x + threadIdx.x] = s;
    }
}

// t = P r (learned: e^{m_i}; Jacobi: 1 / (theta - d_i), floored); partial <u, t>
__global__ __launch_bounds__(kThreads) void ritz_t_kernel(const float *__restrict__ logm,
                                                          const double *__restrict__ d, double theta,
                                                          const dvec2 *__restrict__ r, const dvec2 *__restrict__ u,
                                                          uint64_t D, dvec2 *__restrict__ t,
                                                          dvec2 *__restrict__ partial) {
    dvec2 ut{0.0, 0.0};
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        double scale;
        if (logm) {
            scale = exp(static_cast<double>(logm[i]));
        } else {
            double den = theta - d[i];
            if (fabs(den) < 1e-8) den = den < 0.0 ? -1e-8 : 1e-8;
            scale = 1.0 / den;
        }
        const dvec2 ti = r[i] * scale;
        t[i] = ti;
        ut += cmulc(u[i], ti);
    }
    block_sum2(ut, dvec2{0.0, 0.0}, partial);
}

// q = t - alpha u; partials |q|^2 and <q, r>
__global__ __launch_bounds__(kThreads) void ritz_q_kernel(const dvec2 *__restrict__ t, const dvec2 *__restrict__ u,
                                                          dvec2 alpha, const dvec2 *__restrict__ r, uint64_t D,
                                                          dvec2 *__restrict__ q, dvec2 *__restrict__ partial) {
    dvec2 qq{0.0, 0.0}, qr{0.0, 0.0};
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const dvec2 ui = u[i];
        const dvec2 qi = t[i] - dvec2{alpha.x * ui.x - alpha.y * ui.y, alpha.x * ui.y + alpha.y * ui.x};
        q[i] = qi;
        qq.x += qi.x * qi.x + qi.y * qi.y;
        qr += cmulc(qi, r[i]);
    }
    block_sum2(qq, qr, partial);
}

// partial <q, Hq>
__global__ __launch_bounds__(kThreads) void ritz_gamma_kernel(const dvec2 *__restrict__ q,
                                                              const dvec2 *__restrict__ hq, uint64_t D,
                                                              dvec2 *__restrict__ partial) {
    dvec2 s{0.0, 0.0};
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        s += cmulc(q[i], hq[i]);
    block_sum2(s, dvec2{0.0, 0.0}, partial);
}

struct RitzCoef {
    double A1, gamma, A3;
    dvec2 A2c;  // A2 * conj(c)
};

__device__ __forceinline__ dvec2 ritz_gq(const RitzCoef &k, dvec2 q, dvec2 hq, dvec2 r) {
    const dvec2 a2r{k.A2c.x * r.x - k.A2c.y * r.y, k.A2c.x * r.y + k.A2c.y * r.x};
    return (hq - q * k.gamma) * k.A1 + a2r + q * k.A3;
}

// partial <u, g_q>
__global__ __launch_bounds__(kThreads) void ritz_beta_kernel(RitzCoef k, const dvec2 *__restrict__ q,
                                                             const dvec2 *__restrict__ hq,
                                                             const dvec2 *__restrict__ r, const dvec2 *__restrict__ u,
                                                             uint64_t D, dvec2 *__restrict__ partial) {
    dvec2 s{0.0, 0.0};
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        s += cmulc(u[i], ritz_gq(k, q[i], hq[i], r[i]));
    block_sum2(s, dvec2{0.0, 0.0}, partial);
}

// dL/dm_i = Re(conj(g_q,i - beta u_i) t_i)
__global__ __launch_bounds__(kThreads) void ritz_grad_kernel(RitzCoef k, dvec2 beta, const dvec2 *__restrict__ q,
                                                             const dvec2 *__restrict__ hq,
                                                             const dvec2 *__restrict__ r, const dvec2 *__restrict__ u,
                                                             const dvec2 *__restrict__ t, uint64_t D,
                                                             float *__restrict__ dm) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const dvec2 ui = u[i];
        const dvec2 g = ritz_gq(k, q[i], hq[i], r[i]) -
                        dvec2{beta.x * ui.x - beta.y * ui.y, beta.x * ui.y + beta.y * ui.x};
        const dvec2 ti = t[i];
        dm[i] = static_cast<float>(g.x * ti.x + g.y * ti.y);
    }
}

using tcp::fail;

}  // namespace

struct tcp_mlp {
    int device = 0;
    tcp::DeviceBuffer params, grad, adam_m, adam_v, w1f, w2f, w2b;
    uint32_t step = 0;     // Adam steps taken
    uint32_t version = 0;  // bumped whenever the weights change
    // training scratch (grown on demand)
    tcp::DeviceBuffer h1, h2, y, da2, da1, partial, small;
    MlpDev dev() const {
        return MlpDev{static_cast<const s8 *>(w1f.ptr), static_cast<const s8 *>(w2f.ptr),
                      static_cast<const uint16_t *>(w2b.ptr), static_cast<const float *>(params.ptr)};
    }
    int refresh(hipStream_t s) {
        hipLaunchKernelGGL(swizzle_kernel, dim3((8 * 8 * 64 * 4 + 255) / 256), dim3(256), 0, s,
                           static_cast<const float *>(params.ptr), static_cast<uint16_t *>(w1f.ptr),
                           static_cast<uint16_t *>(w2f.ptr), static_cast<uint16_t *>(w2b.ptr));
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? TCP_OK : tcp::hip_fail(e, "mlp swizzle");
    }
};

struct tcp_learned_precond {
    const tcp_mlp *mlp = nullptr;
    const tcp_sector_plan *plan = nullptr;  // borrowed: must outlive this object
    int device = 0;
    uint64_t dim = 0;
    double scale = 1.0;
    tcp::DeviceBuffer features, diag, logm;
    tcp::DeviceBuffer t, q, hq, partial;  // Ritz-gain loss workspace (allocated on first use)
    // the callback's last evaluation: reused while theta stays within
    // refresh * scale of the theta it was made for and the weights are
    // unchanged
    double refresh = kRefresh;
    bool cached = false;
    double cached_theta = 0.0;
    uint32_t cached_step = 0;
};

namespace {

// Deterministic sum of per-block partials (fixed order on the host).
int sum_partials(const tcp::DeviceBuffer &partial, uint32_t nb, hipStream_t s, dvec2 &v0, dvec2 &v1) {
    std::vector<dvec2> h(2 * nb);
    TCP_HIP_CHECK(hipMemcpyAsync(h.data(), partial.ptr, sizeof(dvec2) * 2 * nb, hipMemcpyDeviceToHost, s));
    TCP_HIP_CHECK(hipStreamSynchronize(s));
    v0 = dvec2{0.0, 0.0};
    v1 = dvec2{0.0, 0.0};
    for (uint32_t b = 0; b < nb; ++b) {
        v0 += h[2 * b];
        v1 += h[2 * b + 1];
    }
    return TCP_OK;
}

int ensure(tcp::DeviceBuffer &b, size_t bytes) {
    if (b.bytes >= bytes) return TCP_OK;
    return b.allocate(bytes);
}

// Row tiles of 16 per wave and weight read (inference; training stores the
// hidden layers and keeps one).
constexpr int kInferTiles = 1, kTrainTiles = 1;

// Exactly the workgroups the device holds at once (a persistent grid: one
// more partial round of blocks would leave most of the device idle), fewer
// for small batches.
template <typename K>
uint32_t forward_blocks(K kernel, uint64_t B, uint64_t rows_per_block) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kFwdThreads, 0) != hipSuccess || per_cu < 1)
        cus = 256, per_cu = 2;
    const uint64_t tiles = (B + rows_per_block - 1) / rows_per_block;
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(tiles, static_cast<uint64_t>(cus) * per_cu)));
}

int forward(const tcp_mlp *mlp, const uint16_t *X, uint64_t B, float *y, float *h1, float *h2, hipStream_t s,
            const double *diag = nullptr, double theta = 0.0, double inv_scale = 0.0) {
    if (B == 0) return TCP_OK;
    if (h1 || h2) {
        if (!h1 || !h2) return fail(TCP_ERR_INVALID_ARGUMENT, "mlp forward stores both hidden layers or neither");
        const uint32_t blocks = forward_blocks(mlp_forward_kernel<true, kTrainTiles>, B, 16 * (kFwdThreads / 64) * kTrainTiles);
        hipLaunchKernelGGL((mlp_forward_kernel<true, kTrainTiles>), dim3(blocks), dim3(kFwdThreads), 0, s, mlp->dev(), X,
                           B, y, h1, h2, diag, theta, inv_scale);
    } else {
        const uint32_t blocks = forward_blocks(mlp_forward_kernel<false, kInferTiles>, B, 16 * (kFwdThreads / 64) * kInferTiles);
        hipLaunchKernelGGL((mlp_forward_kernel<false, kInferTiles>), dim3(blocks), dim3(kFwdThreads), 0, s, mlp->dev(),
                           X, B, y, h1, h2, diag, theta, inv_scale);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : tcp::hip_fail(e, "mlp forward");
}

// grad (device, kP floats) = dL/dparams for dL/dy = dy over a batch of B rows.
int backward(tcp_mlp *mlp, const uint16_t *X, uint64_t B, const float *dy, hipStream_t s) {
    int st;
    if ((st = ensure(mlp->h1, 4 * B * kH)) || (st = ensure(mlp->h2, 4 * B * kH)) || (st = ensure(mlp->y, 4 * B)) ||
        (st = ensure(mlp->da2, 2 * B * kH)) || (st = ensure(mlp->da1, 2 * B * kH)))
        return st;
    auto *h1 = static_cast<float *>(mlp->h1.ptr);
    auto *h2 = static_cast<float *>(mlp->h2.ptr);
    if ((st = forward(mlp, X, B, static_cast<float *>(mlp->y.ptr), h1, h2, s))) return st;
    const uint32_t nb_act = static_cast<uint32_t>(std::min<uint64_t>((B + 1) / 2, 1024));
    const uint32_t nb_h1 = static_cast<uint32_t>(std::min<uint64_t>((B + 63) / 64, 1024));
    const uint64_t rows_per_block = 2048;
    const uint32_t nb_w = static_cast<uint32_t>((B + rows_per_block - 1) / rows_per_block);
    const size_t partial_floats = std::max<size_t>({static_cast<size_t>(nb_act) * (2 * kH + 1),
                                                    static_cast<size_t>(nb_h1) * kH,
                                                    static_cast<size_t>(nb_w) * kH * kH});
    if ((st = ensure(mlp->partial, 4 * partial_floats)) || (st = ensure(mlp->small, 4 * (2 * kH + 1)))) return st;
    auto *partial = static_cast<float *>(mlp->partial.ptr);
    auto *grad = static_cast<float *>(mlp->grad.ptr);
    auto *da2 = static_cast<uint16_t *>(mlp->da2.ptr);
    auto *da1 = static_cast<uint16_t *>(mlp->da1.ptr);
    hipLaunchKernelGGL(act2_backward_kernel, dim3(nb_act), dim3(kThreads), 0, s,
                       static_cast<const float *>(mlp->params.ptr), dy, h2, B, da2, partial);
    // gw3 -> [kOffW3, +128), gb2 -> [kOffB2, +128), gb3 -> kOffB3: columns of width 2*128+1
    hipLaunchKernelGGL(reduce_columns_kernel, dim3((2 * kH + 1 + 127) / 128), dim3(128), 0, s, partial, nb_act,
                       2 * kH + 1,
                       static_cast<float *>(mlp->small.ptr), 0u);
    // move the three segments into place (small device copies)
    TCP_HIP_CHECK(hipMemcpyAsync(grad + kOffW3, static_cast<float *>(mlp->small.ptr), 4 * kH, hipMemcpyDeviceToDevice, s));
    TCP_HIP_CHECK(hipMemcpyAsync(grad + kOffB2, static_cast<float *>(mlp->small.ptr) + kH, 4 * kH, hipMemcpyDeviceToDevice, s));
    TCP_HIP_CHECK(hipMemcpyAsync(grad + kOffB3, static_cast<float *>(mlp->small.ptr) + 2 * kH, 4, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(h1_backward_kernel, dim3(nb_h1), dim3(kThreads), 0, s, mlp->dev(), da2, h1, B, da1, partial);
    hipLaunchKernelGGL(reduce_columns_kernel, dim3(1), dim3(128), 0, s, partial, nb_h1, kH, grad, kOffB1);
    hipLaunchKernelGGL((wgrad_kernel<kH>), dim3(nb_w), dim3(kThreads), 0, s, nullptr, h1, da2, B, rows_per_block,
                       partial);
    hipLaunchKernelGGL(reduce_columns_kernel, dim3((kH * kH + 255) / 256), dim3(256), 0, s, partial, nb_w, kH * kH,
                       grad, kOffW2);
    hipLaunchKernelGGL((wgrad_kernel<kF>), dim3(nb_w), dim3(kThreads), 0, s, X, nullptr, da1, B, rows_per_block,
                       partial);
    hipLaunchKernelGGL(reduce_columns_kernel, dim3((kF * kH + 255) / 256), dim3(256), 0, s, partial, nb_w, kF * kH,
                       grad, kOffW1);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : tcp::hip_fail(e, "mlp backward");
}

}  // namespace

extern "C" {

uint32_t tcp_mlp_parameter_count(void) { return kP; }

int tcp_mlp_create(uint64_t seed, tcp_mlp **out) {
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    auto *m = new tcp_mlp();
    std::unique_ptr<tcp_mlp> guard(m);
    std::vector<float> p(kP, 0.0f);
    std::mt19937_64 rng(seed);
    auto glorot = [&](uint32_t off, uint32_t fan_in, uint32_t fan_out) {
        const float lim = std::sqrt(6.0f / static_cast<float>(fan_in + fan_out));
        std::uniform_real_distribution<float> dist(-lim, lim);
        for (uint32_t i = 0; i < fan_in * fan_out; ++i) p[off + i] = dist(rng);
    };
    glorot(kOffW1, kF, kH);
    glorot(kOffW2, kH, kH);
    glorot(kOffW3, kH, 1);
    for (uint32_t i = 0; i < kH; ++i) p[kOffW3 + i] *= 0.1f;  // start near the plain Jacobi correction
    TCP_HIP_CHECK(hipGetDevice(&m->device));
    int st;
    std::vector<float> zeros(kP, 0.0f);
    if ((st = m->params.upload(p.data(), 4 * kP)) || (st = m->grad.upload(zeros.data(), 4 * kP)) ||
        (st = m->adam_m.upload(zeros.data(), 4 * kP)) || (st = m->adam_v.upload(zeros.data(), 4 * kP)) ||
        (st = m->w1f.allocate(2 * kF * kH)) || (st = m->w2f.allocate(2 * kH * kH)) ||
        (st = m->w2b.allocate(2 * kH * kH)) || (st = m->refresh(nullptr)))
        return st;
    TCP_HIP_CHECK(hipDeviceSynchronize());
    *out = guard.release();
    return tcp::ok();
}

int tcp_mlp_get_parameters(const tcp_mlp *mlp, float *params) {
    if (!mlp || !params) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    tcp::DeviceGuard dg;
    int st = dg.enter(mlp->device);
    if (st) return st;
    TCP_HIP_CHECK(hipMemcpy(params, mlp->params.ptr, 4 * kP, hipMemcpyDeviceToHost));
    return tcp::ok();
}

int tcp_mlp_set_parameters(tcp_mlp *mlp, const float *params) {
    if (!mlp || !params) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    tcp::DeviceGuard dg;
    int st = dg.enter(mlp->device);
    if (st) return st;
    TCP_HIP_CHECK(hipMemcpy(mlp->params.ptr, params, 4 * kP, hipMemcpyHostToDevice));
    ++mlp->version;
    if ((st = mlp->refresh(nullptr))) return st;
    TCP_HIP_CHECK(hipDeviceSynchronize());
    return tcp::ok();
}

int tcp_mlp_forward(const tcp_mlp *mlp, const void *features, uint64_t batch, float *out, void *stream) {
    if (!mlp || (batch && (!features || !out))) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    tcp::DeviceGuard dg;
    int st = dg.enter(mlp->device);
    if (st) return st;
    st = forward(mlp, static_cast<const uint16_t *>(features), batch, out, nullptr, nullptr,
                 static_cast<hipStream_t>(stream));
    return st ? st : tcp::ok();
}

int tcp_mlp_gradient(tcp_mlp *mlp, const void *features, uint64_t batch, const float *dloss_dout, float *grad_host,
                     void *stream) {
    if (!mlp || !batch || !features || !dloss_dout) return fail(TCP_ERR_INVALID_ARGUMENT, "bad argument");
    tcp::DeviceGuard dg;
    int st = dg.enter(mlp->device);
    if (st) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((st = backward(mlp, static_cast<const uint16_t *>(features), batch, dloss_dout, s))) return st;
    if (grad_host) {
        TCP_HIP_CHECK(hipMemcpyAsync(grad_host, mlp->grad.ptr, 4 * kP, hipMemcpyDeviceToHost, s));
        TCP_HIP_CHECK(hipStreamSynchronize(s));
    }
    return tcp::ok();
}

int tcp_mlp_adam_step(tcp_mlp *mlp, float lr, float weight_decay, void *stream) {
    if (!mlp) return fail(TCP_ERR_INVALID_ARGUMENT, "null mlp");
    tcp::DeviceGuard dg;
    int st = dg.enter(mlp->device);
    if (st) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    ++mlp->step;
    ++mlp->version;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, static_cast<float>(mlp->step));
    const float c2 = 1.f - std::pow(b2, static_cast<float>(mlp->step));
    hipLaunchKernelGGL(adam_kernel, dim3((kP + 255) / 256), dim3(256), 0, s, static_cast<float *>(mlp->params.ptr),
                       static_cast<const float *>(mlp->grad.ptr), static_cast<float *>(mlp->adam_m.ptr),
                       static_cast<float *>(mlp->adam_v.ptr), kP, lr, b1, b2, 1e-8f, c1, c2, weight_decay);
    if ((st = mlp->refresh(s))) return st;
    return tcp::ok();
}

void tcp_mlp_destroy(tcp_mlp *mlp) { delete mlp; }

int tcp_learned_precond_ritz_loss(tcp_learned_precond *p, double theta, const void *u, const void *r,
                                  const float *logm, float *dlogm, double *loss, double *gain, void *stream) {
    if (!p || !u || !r || !loss) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    if (logm && !dlogm) return fail(TCP_ERR_INVALID_ARGUMENT, "learned multipliers need a gradient buffer");
    tcp::DeviceGuard dg;
    int st = dg.enter(p->device);
    if (st) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint64_t D = p->dim;
    const size_t vb = 16 * D;
    const uint32_t nb = static_cast<uint32_t>(std::min<uint64_t>((D + kThreads - 1) / kThreads, 1024));
    if ((st = ensure(p->t, vb)) || (st = ensure(p->q, vb)) || (st = ensure(p->hq, vb)) ||
        (st = ensure(p->partial, sizeof(dvec2) * 2 * nb)))
        return st;
    auto *T = static_cast<dvec2 *>(p->t.ptr);
    auto *Q = static_cast<dvec2 *>(p->q.ptr);
    auto *HQ = static_cast<dvec2 *>(p->hq.ptr);
    auto *U = static_cast<const dvec2 *>(u);
    auto *Rv = static_cast<const dvec2 *>(r);
    auto *part = static_cast<dvec2 *>(p->partial.ptr);
    dvec2 alpha, unused, qq, qr, qhq;
    hipLaunchKernelGGL(ritz_t_kernel, dim3(nb), dim3(kThreads), 0, s, logm,
                       static_cast<const double *>(p->diag.ptr), theta, Rv, U, D, T, part);
    if ((st = sum_partials(p->partial, nb, s, alpha, unused))) return st;
    hipLaunchKernelGGL(ritz_q_kernel, dim3(nb), dim3(kThreads), 0, s, T, U, alpha, Rv, D, Q, part);
    if ((st = sum_partials(p->partial, nb, s, qq, qr))) return st;
    const double n2 = qq.x;
    if (!(n2 > 0.0)) return fail(TCP_ERR_INVALID_ARGUMENT, "degenerate snapshot: the correction lies along u");
    if ((st = tcp::plan_apply(p->plan, Q, HQ, TCP_C128, s))) return st;
    hipLaunchKernelGGL(ritz_gamma_kernel, dim3(nb), dim3(kThreads), 0, s, Q, HQ, D, part);
    if ((st = sum_partials(p->partial, nb, s, qhq, unused))) return st;
    const double c2 = qr.x * qr.x + qr.y * qr.y;
    const double gamma = qhq.x / n2, rho2 = c2 / n2, a = 0.5 * (gamma - theta);
    const double R = std::sqrt(a * a + rho2);
    const double delta = a > 0.0 ? rho2 / (R + a) : R - a;
    if (!(delta > 0.0)) return fail(TCP_ERR_INVALID_ARGUMENT, "degenerate snapshot: zero Ritz gain");
    *loss = -std::log(delta);
    if (gain) *gain = delta;
    if (logm) {
        RitzCoef k;
        k.A1 = (1.0 - a / R) / (n2 * delta);
        k.gamma = gamma;
        k.A3 = c2 / (n2 * n2 * R * delta);
        const double A2 = -1.0 / (n2 * R * delta);
        k.A2c = dvec2{A2 * qr.x, -A2 * qr.y};
        dvec2 beta;
        hipLaunchKernelGGL(ritz_beta_kernel, dim3(nb), dim3(kThreads), 0, s, k, Q, HQ, Rv, U, D, part);
        if ((st = sum_partials(p->partial, nb, s, beta, unused))) return st;
        hipLaunchKernelGGL(ritz_grad_kernel, dim3(nb), dim3(kThreads), 0, s, k, beta, Q, HQ, Rv, U, T, D, dlogm);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return tcp::hip_fail(err, "Ritz-gain loss");
    TCP_HIP_CHECK(hipStreamSynchronize(s));
    return tcp::ok();
}

int tcp_learned_precond_create(const tcp_mlp *mlp, const tcp_sector_plan *plan, tcp_learned_precond **out) {
    if (!mlp || !plan || !out) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    auto *p = new tcp_learned_precond();
    std::unique_ptr<tcp_learned_precond> guard(p);
    p->mlp = mlp;
    p->plan = plan;
    p->dim = tcp::plan_dimension(plan);
    p->scale = tcp::plan_energy_scale(plan);
    p->device = tcp::plan_device(plan);
    tcp::DeviceGuard dg;
    int st = dg.enter(p->device);
    if (st) return st;
    if ((st = p->features.allocate(2 * kF * p->dim)) || (st = p->diag.allocate(8 * p->dim)) ||
        (st = p->logm.allocate(4 * p->dim)))
        return st;
    if ((st = tcp::plan_features(plan, p->scale, static_cast<uint16_t *>(p->features.ptr),
                                 static_cast<double *>(p->diag.ptr), nullptr)))
        return st;
    TCP_HIP_CHECK(hipDeviceSynchronize());
    *out = guard.release();
    return tcp::ok();
}

int tcp_learned_precond_prepare(tcp_learned_precond *p, double theta, void *stream) {
    if (!p) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((p->dim + 255) / 256, 4096));
    hipLaunchKernelGGL(fill_feature0_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<uint16_t *>(p->features.ptr), static_cast<const double *>(p->diag.ptr), theta,
                       1.0 / p->scale, p->dim);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? tcp::ok() : tcp::hip_fail(e, "feature fill");
}

// tcp_log_multiplier_fn: log multipliers for the current theta (feature 0 =
// (H_ii - theta) / scale, then MLP inference over every state).
int tcp_learned_precond_set_refresh(tcp_learned_precond *precond, double threshold) {
    if (!precond) return fail(TCP_ERR_INVALID_ARGUMENT, "null preconditioner");
    if (!(threshold >= 0.0) || !std::isfinite(threshold))
        return fail(TCP_ERR_INVALID_ARGUMENT, "refresh threshold must be finite and >= 0");
    precond->refresh = threshold;
    precond->cached = false;
    return TCP_OK;
}

const void *tcp_learned_precond_log_multiplier(void *ctx, double theta, void *stream) {
    auto *p = static_cast<tcp_learned_precond *>(ctx);
    if (!p) return nullptr;
    if (p->cached && p->cached_step == p->mlp->version && std::fabs(theta - p->cached_theta) < p->refresh * p->scale)
        return p->logm.ptr;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int st = forward(p->mlp, static_cast<const uint16_t *>(p->features.ptr), p->dim,
                     static_cast<float *>(p->logm.ptr), nullptr, nullptr, s,
                     static_cast<const double *>(p->diag.ptr), theta, 1.0 / p->scale);
    if (st) return nullptr;
    p->cached = true;
    p->cached_theta = theta;
    p->cached_step = p->mlp->version;
    return p->logm.ptr;
}

int tcp_learned_precond_buffers(const tcp_learned_precond *p, const void **features, const double **diag,
                                double *scale) {
    if (!p) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    if (features) *features = p->features.ptr;
    if (diag) *diag = static_cast<const double *>(p->diag.ptr);
    if (scale) *scale = p->scale;
    return tcp::ok();
}

void tcp_learned_precond_destroy(tcp_learned_precond *p) { delete p; }

}  // extern "C"
