// Shared pieces of the GPU eigensolvers (Davidson and thick-restart
// Lanczos): complex helpers in FP64/FP32 storage, deterministic block
// reductions, basis combinations, start vectors, and the host-side dense
// Hermitian eigensolver of the projected problem.  Everything sits in an
// anonymous namespace: each solver's translation unit has its own copy.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <limits>
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

inline double conj_of(double x) { return x; }
inline std::complex<double> conj_of(const std::complex<double> &x) { return std::conj(x); }
inline double abs2_of(double x) { return x * x; }
inline double abs2_of(const std::complex<double> &x) { return std::norm(x); }
inline double real_of(double x) { return x; }
inline double real_of(const std::complex<double> &x) { return x.real(); }
inline double phase_of(double x) { return x < 0.0 ? -1.0 : 1.0; }
inline std::complex<double> phase_of(const std::complex<double> &x) {
    const double r = std::abs(x);
    return r > 0.0 ? x / r : std::complex<double>(1.0, 0.0);
}

// Eigen-decomposition of a Hermitian n x n matrix `a` (row-major; S = double
// for real symmetric, std::complex<double> for complex Hermitian input):
// Householder reduction to tridiagonal form, A = Q T Q^H, a diagonal phase
// that makes T real (T = D T' D^H), then the implicit QL iteration with
// Wilkinson shifts on T' (Bowdler, Martin, Reinsch and Wilkinson, Numer.
// Math. 11 (1968) 293-306), its plane rotations applied to the columns of
// Q D.  Eigenvalues ascending in `vals`; eigenvector k is column k of the
// row-major n x n `vecs`.  O(n^3) with a small constant: for n = 24 it takes
// a few percent of the time of a cyclic Jacobi sweep sequence.
template <typename S>
void tridiagonal_eigh(int n, std::vector<S> a, std::vector<double> &vals, std::vector<S> &vecs) {
    auto A = [&](int i, int j) -> S & { return a[static_cast<size_t>(i) * n + j]; };
    std::vector<S> q(static_cast<size_t>(n) * n, S(0.0));
    auto Q = [&](int i, int j) -> S & { return q[static_cast<size_t>(i) * n + j]; };
    for (int i = 0; i < n; ++i) Q(i, i) = S(1.0);
    std::vector<S> v(n), p(n);
    // Householder steps: H_k = I - beta v v^H zeroes column k below the subdiagonal
    for (int k = 0; k + 2 < n; ++k) {
        double sigma = 0.0;
        for (int i = k + 1; i < n; ++i) sigma += abs2_of(A(i, k));
        const double tail = sigma - abs2_of(A(k + 1, k));
        if (!(tail > 0.0)) continue;  // already tridiagonal in this column
        const double norm = std::sqrt(sigma);
        const S alpha = -phase_of(A(k + 1, k)) * norm;
        for (int i = k + 1; i < n; ++i) v[i] = A(i, k);
        v[k + 1] -= alpha;
        double h = 0.0;
        for (int i = k + 1; i < n; ++i) h += abs2_of(v[i]);
        const double beta = 2.0 / h;
        // p = beta B v, K = beta (v^H p) / 2, q = p - K v, B <- B - v q^H - q v^H
        double vp = 0.0;
        for (int i = k + 1; i < n; ++i) {
            S acc(0.0);
            for (int j = k + 1; j < n; ++j) acc += A(i, j) * v[j];
            p[i] = beta * acc;
            vp += real_of(conj_of(v[i]) * p[i]);
        }
        const double K = 0.5 * beta * vp;
        for (int i = k + 1; i < n; ++i) p[i] -= K * v[i];
        for (int i = k + 1; i < n; ++i)
            for (int j = k + 1; j < n; ++j) A(i, j) -= v[i] * conj_of(p[j]) + p[i] * conj_of(v[j]);
        A(k + 1, k) = alpha;
        A(k, k + 1) = conj_of(alpha);
        for (int i = k + 2; i < n; ++i) A(i, k) = A(k, i) = S(0.0);
        // Q <- Q H_k (columns k+1..n-1)
        for (int r = 0; r < n; ++r) {
            S acc(0.0);
            for (int j = k + 1; j < n; ++j) acc += Q(r, j) * v[j];
            acc *= beta;
            for (int j = k + 1; j < n; ++j) Q(r, j) -= acc * conj_of(v[j]);
        }
    }
    // real tridiagonal T' = D^H T D; the columns of Q pick up the phases
    std::vector<double> d(n), e(n, 0.0);
    S ph(1.0);
    for (int i = 0; i < n; ++i) {
        d[i] = real_of(A(i, i));
        if (i > 0)
            for (int r = 0; r < n; ++r) Q(r, i) *= ph;
        if (i + 1 < n) {
            const S sub = A(i + 1, i);
            e[i] = std::sqrt(abs2_of(sub));
            ph *= phase_of(sub);
        }
    }
    // implicit QL with Wilkinson shifts; e[i] couples d[i] and d[i + 1]
    const double eps = std::numeric_limits<double>::epsilon();
    for (int l = 0; l < n; ++l) {
        for (int iter = 0; iter < 64; ++iter) {
            int m = l;
            while (m + 1 < n && std::fabs(e[m]) > eps * (std::fabs(d[m]) + std::fabs(d[m + 1]))) ++m;
            if (m == l) break;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, pp = 0.0;
            bool split = false;
            for (int i = m - 1; i >= l; --i) {
                const double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {  // the rotation vanished: T split at i + 1
                    d[i + 1] -= pp;
                    e[m] = 0.0;
                    split = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - pp;
                r = (d[i] - g) * s + 2.0 * c * b;
                pp = s * r;
                d[i + 1] = g + pp;
                g = c * r - b;
                for (int row = 0; row < n; ++row) {
                    const S zi = Q(row, i), zj = Q(row, i + 1);
                    Q(row, i + 1) = s * zi + c * zj;
                    Q(row, i) = c * zi - s * zj;
                }
            }
            if (split) continue;
            d[l] -= pp;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
    vals.resize(n);
    vecs.assign(static_cast<size_t>(n) * n, S(0.0));
    for (int k = 0; k < n; ++k) {
        vals[k] = d[order[k]];
        for (int i = 0; i < n; ++i) vecs[static_cast<size_t>(i) * n + k] = Q(i, order[k]);
    }
}

// Lowest `count` eigenpairs of a Hermitian m x m matrix G (interleaved
// complex, row-major); a matrix whose imaginary parts are all zero is solved
// in real arithmetic.
void hermitian_lowest(int m, const std::vector<double> &g, int count, std::vector<double> &vals,
                      std::vector<double> &vecs /* m x count complex, column-major by pair */) {
    bool real = true;
    for (int i = 0; i < m * m; ++i)
        if (g[2 * i + 1] != 0.0) real = false;
    std::vector<double> ev;
    vals.assign(count, 0.0);
    vecs.assign(static_cast<size_t>(2) * m * count, 0.0);
    if (real) {
        std::vector<double> a(static_cast<size_t>(m) * m), evec;
        for (int i = 0; i < m * m; ++i) a[i] = g[2 * i];
        tridiagonal_eigh<double>(m, std::move(a), ev, evec);
        for (int k = 0; k < count; ++k) {
            vals[k] = ev[k];
            for (int i = 0; i < m; ++i) vecs[2 * (k * m + i)] = evec[i * m + k];
        }
        return;
    }
    std::vector<std::complex<double>> a(static_cast<size_t>(m) * m), evec;
    for (int i = 0; i < m * m; ++i) a[i] = std::complex<double>(g[2 * i], g[2 * i + 1]);
    tridiagonal_eigh<std::complex<double>>(m, std::move(a), ev, evec);
    for (int k = 0; k < count; ++k) {
        vals[k] = ev[k];
        for (int i = 0; i < m; ++i) {
            vecs[2 * (k * m + i)] = evec[i * m + k].real();
            vecs[2 * (k * m + i) + 1] = evec[i * m + k].imag();
        }
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
