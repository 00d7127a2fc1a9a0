// Shared pieces of the GPU eigensolvers (Davidson and thick-restart
// Lanczos): complex helpers in FP64/FP32 storage, deterministic block
// reductions, basis combinations, start vectors, and the host-side dense
// Hermitian eigensolver of the projected problem.  Everything sits in an
// anonymous namespace: each solver's translation unit has its own copy.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

namespace {

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

constexpr uint32_t kThreads = 256;
constexpr uint32_t kWaves = kThreads / 64;
constexpr uint32_t kStreamBlocks = 1024;
constexpr uint32_t kMaxBasis = 32;  // basis columns

__device__ __forceinline__ dvec2 widen(dvec2 x) { return x; }
__device__ __forceinline__ dvec2 widen(fvec2 x) { return dvec2{x.x, x.y}; }
template <typename ST>
__device__ __forceinline__ ST narrow(dvec2 x);
template <>
__device__ __forceinline__ dvec2 narrow<dvec2>(dvec2 x) {
    return x;
}
template <>
__device__ __forceinline__ fvec2 narrow<fvec2>(dvec2 x) {
    return fvec2{static_cast<float>(x.x), static_cast<float>(x.y)};
}

__device__ __forceinline__ dvec2 cmul(dvec2 a, dvec2 b) {
    return dvec2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ dvec2 cmulc(dvec2 a, dvec2 b) {  // conj(a) * b
    return dvec2{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}

template <typename ST>
struct Cols {
    ST *p[kMaxBasis + 1];
};

// Per-block sums of acc[0..nval) into partial[block * nval + j].
template <int NV>
__device__ __forceinline__ void block_partials(dvec2 (&acc)[NV], uint32_t nval, dvec2 *__restrict__ partial) {
    __shared__ dvec2 red[kWaves][NV];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        if (j < static_cast<int>(nval)) {
            dvec2 v = acc[j];
            for (int off = 32; off > 0; off >>= 1) {
                v.x += __shfl_xor(v.x, off);
                v.y += __shfl_xor(v.y, off);
            }
            if (lane == 0) red[wave][j] = v;
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < nval; j += kThreads) {
        dvec2 s = red[0][j];
        for (uint32_t w = 1; w < kWaves; ++w) s += red[w][j];
        partial[static_cast<uint64_t>(blockIdx.x) * nval + j] = s;
    }
}

// out_k = sum_j Y[k * MAXM + j] in_j for k < K (the kept Ritz vectors of a
// restart, or one Ritz vector formed in another precision); Y is zero past m.
template <typename SI, typename SO, int MAXM>
__global__ __launch_bounds__(kThreads) void combine_kernel(Cols<SI> in, uint32_t m, const dvec2 *__restrict__ Y,
                                                           uint32_t K, Cols<SO> out, uint64_t D) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kThreads) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * kThreads) {
        SI x[MAXM];
#pragma unroll
        for (int j = 0; j < MAXM; ++j) x[j] = in.p[j][i];
        for (uint32_t k = 0; k < K; ++k) {
            dvec2 acc{0.0, 0.0};
#pragma unroll
            for (int j = 0; j < MAXM; ++j) acc += cmul(Y[k * MAXM + j], widen(x[j]));
            out.p[k][i] = narrow<SO>(acc);
        }
    }
}

template <typename ST>
__global__ void scale_kernel(ST *__restrict__ x, double alpha, uint64_t D) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        x[i] = narrow<ST>(widen(x[i]) * alpha);
}

// Deterministic start vector: splitmix64 hash of (seed, i) -> U(-1, 1), real.
template <typename ST>
__global__ void init_kernel(ST *__restrict__ x, uint64_t seed, uint64_t D) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        uint64_t z = seed + 0x9e3779b97f4a7c15ull * (i + 1);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        x[i] = narrow<ST>(dvec2{static_cast<double>(z >> 11) * (2.0 / 9007199254740992.0) - 1.0, 0.0});
    }
}

// ---------------------------------------------------------------- host side

inline uint32_t stream_blocks(uint64_t D) {
    return static_cast<uint32_t>(std::min<uint64_t>((D + kThreads - 1) / kThreads, kStreamBlocks));
}

// Cyclic Jacobi eigensolver for a real symmetric n x n matrix (row-major).
// Eigenvalues ascending; eigenvector k is column k of `vecs`.
void jacobi_eigh(int n, std::vector<double> a, std::vector<double> &vals, std::vector<double> &vecs) {
    vecs.assign(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) vecs[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, scale = 0.0;
        for (int i = 0; i < n; ++i) {
            scale += a[i * n + i] * a[i * n + i];
            for (int j = i + 1; j < n; ++j) off += a[i * n + j] * a[i * n + j];
        }
        if (off <= 1e-32 * (scale + 1e-300)) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double tau = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(tau * tau + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    const double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    const double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = vecs[k * n + p], vkq = vecs[k * n + q];
                    vecs[k * n + p] = c * vkp - s * vkq;
                    vecs[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return a[x * n + x] < a[y * n + y]; });
    vals.resize(n);
    std::vector<double> sorted(static_cast<size_t>(n) * n);
    for (int k = 0; k < n; ++k) {
        vals[k] = a[order[k] * n + order[k]];
        for (int i = 0; i < n; ++i) sorted[i * n + k] = vecs[i * n + order[k]];
    }
    vecs.swap(sorted);
}

// Lowest `count` eigenpairs of a Hermitian m x m matrix G (interleaved
// complex, row-major).  Real matrices are solved directly; complex ones via
// the real symmetric embedding [[Re, -Im], [Im, Re]] whose spectrum is G's,
// each eigenvalue twice (one copy per pair is kept).
void hermitian_lowest(int m, const std::vector<double> &g, int count, std::vector<double> &vals,
                      std::vector<double> &vecs /* m x count complex, column-major by pair */) {
    bool real = true;
    for (int i = 0; i < m * m; ++i)
        if (g[2 * i + 1] != 0.0) real = false;
    std::vector<double> ev, evec;
    vals.assign(count, 0.0);
    vecs.assign(static_cast<size_t>(2) * m * count, 0.0);
    if (real) {
        std::vector<double> a(static_cast<size_t>(m) * m);
        for (int i = 0; i < m * m; ++i) a[i] = g[2 * i];
        jacobi_eigh(m, a, ev, evec);
        for (int k = 0; k < count; ++k) {
            vals[k] = ev[k];
            for (int i = 0; i < m; ++i) vecs[2 * (k * m + i)] = evec[i * m + k];
        }
        return;
    }
    const int n = 2 * m;
    std::vector<double> a(static_cast<size_t>(n) * n);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            const double re = g[2 * (i * m + j)], im = g[2 * (i * m + j) + 1];
            a[i * n + j] = re;
            a[(i + m) * n + (j + m)] = re;
            a[i * n + (j + m)] = -im;
            a[(i + m) * n + j] = im;
        }
    jacobi_eigh(n, a, ev, evec);
    int found = 0;
    std::vector<std::vector<double>> kept;
    for (int k = 0; k < n && found < count; ++k) {
        std::vector<double> x(2 * m);
        double norm = 0.0;
        for (int i = 0; i < m; ++i) {
            x[2 * i] = evec[i * n + k];
            x[2 * i + 1] = evec[(i + m) * n + k];
        }
        // Gram-Schmidt against kept vectors of the same pairs (drop the duplicate copy)
        for (const auto &y : kept) {
            double pr = 0.0, pi = 0.0;
            for (int i = 0; i < m; ++i) {
                pr += y[2 * i] * x[2 * i] + y[2 * i + 1] * x[2 * i + 1];
                pi += y[2 * i] * x[2 * i + 1] - y[2 * i + 1] * x[2 * i];
            }
            for (int i = 0; i < m; ++i) {
                x[2 * i] -= pr * y[2 * i] - pi * y[2 * i + 1];
                x[2 * i + 1] -= pr * y[2 * i + 1] + pi * y[2 * i];
            }
        }
        for (double v : x) norm += v * v;
        if (norm < 1e-12) continue;
        norm = std::sqrt(norm);
        for (double &v : x) v /= norm;
        kept.push_back(x);
        vals[found] = ev[k];
        for (int i = 0; i < 2 * m; ++i) vecs[static_cast<size_t>(found) * 2 * m + i] = x[i];
        ++found;
    }
}

// The smallest register-array bound that holds m columns.
template <class F>
int with_maxm(uint32_t m, F &&f) {
    if (m <= 4) return f(std::integral_constant<int, 4>());
    if (m <= 8) return f(std::integral_constant<int, 8>());
    if (m <= 12) return f(std::integral_constant<int, 12>());
    if (m <= 16) return f(std::integral_constant<int, 16>());
    if (m <= 20) return f(std::integral_constant<int, 20>());
    if (m <= 24) return f(std::integral_constant<int, 24>());
    if (m <= 28) return f(std::integral_constant<int, 28>());
    return f(std::integral_constant<int, 32>());
}

}  // namespace
