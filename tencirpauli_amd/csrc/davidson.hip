// Davidson eigensolver for the lowest eigenpair of an MVP plan, on the GPU.
//
// No reference counterpart: the reference's eigen entry point is SciPy
// eigsh over a LinearOperator (py/integrations/scipy.py:13-47).  This solver
// keeps the search basis V, the products W = H V and every vector operation on
// the device in FP64 (orthogonalisation, Rayleigh-quotient matrix entries,
// residual norms); only the m x m projected eigenproblem is solved on the
// host.  The MVP may run with complex64 storage (FP32 coefficients, FP64
// accumulation) in a first phase; once the residual reaches the switch
// tolerance the solver restarts from the Ritz vector with FP64 MVPs, so the
// final eigenvalue belongs to the FP64 operator.
//
// Correction: t = P (r), P_i = e^{m_i} / (theta - H_ii) -- the Jacobi
// (diagonal) preconditioner, optionally scaled by a per-state log multiplier
// m_i produced by the learned preconditioner for the current theta.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "tcp_internal.hpp"

namespace {

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

constexpr uint32_t kThreads = 256;
constexpr uint32_t kChunk = 4096;      // elements per block of the block-dot kernel
constexpr uint32_t kStreamBlocks = 1024;
constexpr uint32_t kMaxSubspace = 64;

__device__ __forceinline__ dvec2 cmul(dvec2 a, dvec2 b) {
    return dvec2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ dvec2 cmulc(dvec2 a, dvec2 b) {  // conj(a) * b
    return dvec2{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}

// Sum over the block; the result is valid in thread 0.
__device__ __forceinline__ dvec2 block_sum(dvec2 v, dvec2 *scratch) {
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off);
        v.y += __shfl_xor(v.y, off);
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    dvec2 total{0.0, 0.0};
    if (threadIdx.x == 0)
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) total += scratch[w];
    __syncthreads();
    return total;
}

// partial[b * m + j] = sum over chunk b of conj(V_j[i]) w[i]
__global__ __launch_bounds__(kThreads) void dot_partials_kernel(const dvec2 *__restrict__ V, uint32_t m,
                                                                uint64_t ld, const dvec2 *__restrict__ w,
                                                                uint64_t D, dvec2 *__restrict__ partial) {
    __shared__ dvec2 ws[kChunk];
    __shared__ dvec2 red[kThreads / 64];
    const uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * kChunk;
    const uint32_t len = static_cast<uint32_t>(D - i0 < kChunk ? D - i0 : kChunk);
    for (uint32_t e = threadIdx.x; e < len; e += kThreads) ws[e] = w[i0 + e];
    __syncthreads();
    for (uint32_t j = 0; j < m; ++j) {
        const dvec2 *vj = V + j * ld + i0;
        dvec2 acc{0.0, 0.0};
        for (uint32_t e = threadIdx.x; e < len; e += kThreads) acc += cmulc(vj[e], ws[e]);
        acc = block_sum(acc, red);
        if (threadIdx.x == 0) partial[static_cast<uint64_t>(blockIdx.x) * m + j] = acc;
    }
}

// out[j] = sum over blocks (fixed order: deterministic)
__global__ void reduce_partials_kernel(const dvec2 *__restrict__ partial, uint32_t nblk, uint32_t m,
                                       dvec2 *__restrict__ out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    dvec2 s{0.0, 0.0};
    for (uint32_t b = 0; b < nblk; ++b) s += partial[static_cast<uint64_t>(b) * m + j];
    out[j] = s;
}

// out = beta * out + sum_j c_j V_j
__global__ __launch_bounds__(kThreads) void combine_kernel(const dvec2 *__restrict__ V, uint32_t m, uint64_t ld,
                                                           const dvec2 *__restrict__ c, double beta,
                                                           dvec2 *__restrict__ out, uint64_t D) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        dvec2 acc = beta != 0.0 ? out[i] * beta : dvec2{0.0, 0.0};
        for (uint32_t j = 0; j < m; ++j) acc += cmul(c[j], V[j * ld + i]);
        out[i] = acc;
    }
}

// u = V s, hu = W s, r = hu - theta u; partial[b] = block sum of |r|^2
__global__ __launch_bounds__(kThreads) void residual_kernel(const dvec2 *__restrict__ V,
                                                            const dvec2 *__restrict__ W, uint32_t m,
                                                            uint64_t ld, const dvec2 *__restrict__ s,
                                                            double theta, dvec2 *__restrict__ u,
                                                            dvec2 *__restrict__ hu, dvec2 *__restrict__ r,
                                                            uint64_t D, double *__restrict__ partial) {
    __shared__ dvec2 red[kThreads / 64];
    double local = 0.0;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        dvec2 uu{0.0, 0.0}, hh{0.0, 0.0};
        for (uint32_t j = 0; j < m; ++j) {
            uu += cmul(s[j], V[j * ld + i]);
            hh += cmul(s[j], W[j * ld + i]);
        }
        const dvec2 rr = hh - uu * theta;
        u[i] = uu;
        hu[i] = hh;
        r[i] = rr;
        local += rr.x * rr.x + rr.y * rr.y;
    }
    const dvec2 total = block_sum(dvec2{local, 0.0}, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = total.x;
}

__global__ __launch_bounds__(kThreads) void norm_kernel(const dvec2 *__restrict__ x, uint64_t D,
                                                        double *__restrict__ partial) {
    __shared__ dvec2 red[kThreads / 64];
    double local = 0.0;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        local += x[i].x * x[i].x + x[i].y * x[i].y;
    const dvec2 total = block_sum(dvec2{local, 0.0}, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = total.x;
}

__global__ void scale_kernel(dvec2 *__restrict__ x, double alpha, uint64_t D) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        x[i] *= alpha;
}

// t_i = r_i e^{m_i} / (theta - d_i), the denominator kept away from zero.
__global__ void precondition_kernel(const dvec2 *__restrict__ r, const double *__restrict__ diag,
                                    const float *__restrict__ logm, double theta, double floor_,
                                    dvec2 *__restrict__ t, uint64_t D) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        double den = theta - diag[i];
        if (fabs(den) < floor_) den = den < 0.0 ? -floor_ : (den > 0.0 ? floor_ : -floor_);
        double scale = 1.0 / den;
        if (logm) scale *= exp(static_cast<double>(logm[i]));
        t[i] = r[i] * scale;
    }
}

__global__ void to_c64_kernel(const dvec2 *__restrict__ x, fvec2 *__restrict__ y, uint64_t D) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        y[i] = fvec2{static_cast<float>(x[i].x), static_cast<float>(x[i].y)};
}

__global__ void to_c128_kernel(const fvec2 *__restrict__ x, dvec2 *__restrict__ y, uint64_t D) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        y[i] = dvec2{static_cast<double>(x[i].x), static_cast<double>(x[i].y)};
}

// Deterministic start vector: splitmix64 hash of (seed, i) -> U(-1, 1), real.
__global__ void init_kernel(dvec2 *__restrict__ x, uint64_t seed, uint64_t D) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        uint64_t z = seed + 0x9e3779b97f4a7c15ull * (i + 1);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        x[i] = dvec2{static_cast<double>(z >> 11) * (2.0 / 9007199254740992.0) - 1.0, 0.0};
    }
}

// ---------------------------------------------------------------- host side
using tcp::fail;

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

struct Work {
    uint64_t D = 0;
    uint32_t nblk_dot = 0;
    tcp::DeviceBuffer V, W, Vn, Wn, u, hu, r, x32, y32, diag, partial, small, out;
    std::vector<double> partial_host;
};

template <class Plan>
struct Solver {
    const Plan *plan;
    const tcp_davidson_options &opt;
    hipStream_t stream;
    Work w;
    dvec2 *Vp(uint32_t j) { return static_cast<dvec2 *>(w.V.ptr) + j * w.D; }
    dvec2 *Wp(uint32_t j) { return static_cast<dvec2 *>(w.W.ptr) + j * w.D; }

    Solver(const Plan *p, const tcp_davidson_options &o, hipStream_t s) : plan(p), opt(o), stream(s) {}

    int dots(const dvec2 *V, uint32_t m, const dvec2 *x, std::vector<double> &out) {
        hipLaunchKernelGGL(dot_partials_kernel, dim3(w.nblk_dot), dim3(kThreads), 0, stream, V, m, w.D, x, w.D,
                           static_cast<dvec2 *>(w.partial.ptr));
        hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kMaxSubspace), 0, stream,
                           static_cast<const dvec2 *>(w.partial.ptr), w.nblk_dot, m,
                           static_cast<dvec2 *>(w.out.ptr));
        out.resize(2 * m);
        TCP_HIP_CHECK(hipMemcpyAsync(out.data(), w.out.ptr, 16 * m, hipMemcpyDeviceToHost, stream));
        TCP_HIP_CHECK(hipStreamSynchronize(stream));
        return TCP_OK;
    }
    int norm(const dvec2 *x, double &value) {
        const uint32_t nb = stream_blocks(w.D);
        hipLaunchKernelGGL(norm_kernel, dim3(nb), dim3(kThreads), 0, stream, x, w.D,
                           static_cast<double *>(w.partial.ptr));
        w.partial_host.resize(nb);
        TCP_HIP_CHECK(hipMemcpyAsync(w.partial_host.data(), w.partial.ptr, 8 * nb, hipMemcpyDeviceToHost, stream));
        TCP_HIP_CHECK(hipStreamSynchronize(stream));
        double s = 0.0;
        for (double v : w.partial_host) s += v;
        value = std::sqrt(s);
        return TCP_OK;
    }
    int upload_small(const std::vector<double> &c) {
        TCP_HIP_CHECK(hipMemcpyAsync(w.small.ptr, c.data(), 8 * c.size(), hipMemcpyHostToDevice, stream));
        return TCP_OK;
    }
    int combine(const dvec2 *V, uint32_t m, const std::vector<double> &c, double beta, dvec2 *out) {
        int st = upload_small(c);
        if (st) return st;
        hipLaunchKernelGGL(combine_kernel, dim3(stream_blocks(w.D)), dim3(kThreads), 0, stream, V, m, w.D,
                           static_cast<const dvec2 *>(w.small.ptr), beta, out, w.D);
        return TCP_OK;
    }
    // y = H x in the requested precision (complex128 in and out).
    int apply(const dvec2 *x, dvec2 *y, int dtype, tcp_davidson_result &res) {
        if (dtype == TCP_C64) {
            hipLaunchKernelGGL(to_c64_kernel, dim3(stream_blocks(w.D)), dim3(kThreads), 0, stream, x,
                               static_cast<fvec2 *>(w.x32.ptr), w.D);
            int st = tcp::plan_apply(plan, w.x32.ptr, w.y32.ptr, TCP_C64, stream);
            if (st) return st;
            hipLaunchKernelGGL(to_c128_kernel, dim3(stream_blocks(w.D)), dim3(kThreads), 0, stream,
                               static_cast<const fvec2 *>(w.y32.ptr), y, w.D);
            ++res.mvps_c64;
            return TCP_OK;
        }
        ++res.mvps_c128;
        return tcp::plan_apply(plan, x, y, TCP_C128, stream);
    }

    int run(void *evec_out, tcp_davidson_result &res) {
        using clock = std::chrono::steady_clock;
        const auto t0 = clock::now();
        const uint64_t D = tcp::plan_dimension(plan);
        w.D = D;
        const uint32_t mmax = std::min<uint32_t>(std::max<uint32_t>(opt.max_subspace ? opt.max_subspace : 24, 4),
                                                 kMaxSubspace);
        const uint32_t keep = std::min<uint32_t>(std::max<uint32_t>(opt.restart_keep ? opt.restart_keep : 4, 1),
                                                 mmax - 2);
        const uint32_t max_it = opt.max_iterations ? opt.max_iterations : 1000;
        const double tol = opt.residual_tol > 0 ? opt.residual_tol : 1e-8;
        int phase = opt.mvp_dtype == TCP_C64 ? TCP_C64 : TCP_C128;
        const double switch_tol = opt.switch_tol > 0 ? opt.switch_tol : 1e-4;
        w.nblk_dot = static_cast<uint32_t>((D + kChunk - 1) / kChunk);
        int st;
        const size_t vb = static_cast<size_t>(D) * 16;
        if ((st = w.V.allocate(vb * mmax)) || (st = w.W.allocate(vb * mmax)) || (st = w.Vn.allocate(vb * keep)) ||
            (st = w.Wn.allocate(vb * keep)) || (st = w.u.allocate(vb)) || (st = w.hu.allocate(vb)) ||
            (st = w.r.allocate(vb)) || (st = w.diag.allocate(8 * D)) ||
            (st = w.partial.allocate(16 * std::max<uint64_t>(static_cast<uint64_t>(w.nblk_dot) * mmax,
                                                               kStreamBlocks))) ||
            (st = w.small.allocate(16 * kMaxSubspace)) || (st = w.out.allocate(16 * kMaxSubspace)))
            return st;
        if (phase == TCP_C64 && ((st = w.x32.allocate(vb / 2)) || (st = w.y32.allocate(vb / 2)))) return st;
        if ((st = tcp::plan_diagonal(plan, static_cast<double *>(w.diag.ptr), stream))) return st;

        const uint32_t nb = stream_blocks(D);
        hipLaunchKernelGGL(init_kernel, dim3(nb), dim3(kThreads), 0, stream, Vp(0), opt.seed, D);
        double nv;
        if ((st = norm(Vp(0), nv))) return st;
        hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(kThreads), 0, stream, Vp(0), 1.0 / nv, D);
        if ((st = apply(Vp(0), Wp(0), phase, res))) return st;
        std::vector<double> G(2 * kMaxSubspace * kMaxSubspace, 0.0), col, vals, vecs;
        auto Gat = [&](uint32_t i, uint32_t j) -> double * { return &G[2 * (i * kMaxSubspace + j)]; };
        if ((st = dots(Vp(0), 1, Wp(0), col))) return st;
        Gat(0, 0)[0] = col[0];
        Gat(0, 0)[1] = 0.0;
        uint32_t m = 1;
        double theta = col[0], rnorm = 0.0;
        res.trace_len = 0;
        for (uint32_t it = 0; it < max_it; ++it) {
            std::vector<double> g(2 * m * m);
            for (uint32_t i = 0; i < m; ++i)
                for (uint32_t j = 0; j < m; ++j) {
                    g[2 * (i * m + j)] = Gat(i, j)[0];
                    g[2 * (i * m + j) + 1] = Gat(i, j)[1];
                }
            const int want = static_cast<int>(std::min<uint32_t>(m, keep));
            hermitian_lowest(static_cast<int>(m), g, want, vals, vecs);
            theta = vals[0];
            std::vector<double> s(vecs.begin(), vecs.begin() + 2 * m);
            if ((st = upload_small(s))) return st;
            hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(kThreads), 0, stream, Vp(0), Wp(0), m, D,
                               static_cast<const dvec2 *>(w.small.ptr), theta, static_cast<dvec2 *>(w.u.ptr),
                               static_cast<dvec2 *>(w.hu.ptr), static_cast<dvec2 *>(w.r.ptr), D,
                               static_cast<double *>(w.partial.ptr));
            w.partial_host.resize(nb);
            TCP_HIP_CHECK(hipMemcpyAsync(w.partial_host.data(), w.partial.ptr, 8 * nb, hipMemcpyDeviceToHost, stream));
            TCP_HIP_CHECK(hipStreamSynchronize(stream));
            double r2 = 0.0;
            for (double v : w.partial_host) r2 += v;
            rnorm = std::sqrt(r2);
            res.iterations = it + 1;
            const double elapsed = std::chrono::duration<double>(clock::now() - t0).count();
            if (opt.trace && res.trace_len < opt.trace_capacity) {
                double *row = opt.trace + 4 * res.trace_len;
                row[0] = elapsed;
                row[1] = theta;
                row[2] = rnorm;
                row[3] = phase == TCP_C64 ? 32.0 : 64.0;
                ++res.trace_len;
            }
            if (phase == TCP_C128 && rnorm <= tol) {
                res.converged = 1;
                break;
            }
            if (phase == TCP_C64 && rnorm <= switch_tol) {
                // restart from the Ritz vector with FP64 MVPs
                phase = TCP_C128;
                TCP_HIP_CHECK(hipMemcpyAsync(Vp(0), w.u.ptr, 16 * D, hipMemcpyDeviceToDevice, stream));
                if ((st = norm(Vp(0), nv))) return st;
                hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(kThreads), 0, stream, Vp(0), 1.0 / nv, D);
                if ((st = apply(Vp(0), Wp(0), phase, res))) return st;
                if ((st = dots(Vp(0), 1, Wp(0), col))) return st;
                Gat(0, 0)[0] = col[0];
                Gat(0, 0)[1] = 0.0;
                m = 1;
                ++res.restarts;
                continue;
            }
            if (m + 1 > mmax) {
                // thick restart: keep the `want` lowest Ritz vectors
                for (int k = 0; k < want; ++k) {
                    std::vector<double> c(vecs.begin() + 2 * m * k, vecs.begin() + 2 * m * (k + 1));
                    if ((st = combine(Vp(0), m, c, 0.0, static_cast<dvec2 *>(w.Vn.ptr) + k * D))) return st;
                    if ((st = combine(Wp(0), m, c, 0.0, static_cast<dvec2 *>(w.Wn.ptr) + k * D))) return st;
                }
                TCP_HIP_CHECK(hipMemcpyAsync(Vp(0), w.Vn.ptr, 16 * D * want, hipMemcpyDeviceToDevice, stream));
                TCP_HIP_CHECK(hipMemcpyAsync(Wp(0), w.Wn.ptr, 16 * D * want, hipMemcpyDeviceToDevice, stream));
                m = static_cast<uint32_t>(want);
                for (uint32_t j = 0; j < m; ++j) {
                    if ((st = dots(Vp(0), m, Wp(j), col))) return st;
                    for (uint32_t i = 0; i < m; ++i) {
                        Gat(i, j)[0] = col[2 * i];
                        Gat(i, j)[1] = col[2 * i + 1];
                    }
                }
                for (uint32_t i = 0; i < m; ++i)
                    for (uint32_t j = i + 1; j < m; ++j) {  // Hermitian average
                        const double re = 0.5 * (Gat(i, j)[0] + Gat(j, i)[0]);
                        const double im = 0.5 * (Gat(i, j)[1] - Gat(j, i)[1]);
                        Gat(i, j)[0] = re;
                        Gat(i, j)[1] = im;
                        Gat(j, i)[0] = re;
                        Gat(j, i)[1] = -im;
                    }
                for (uint32_t i = 0; i < m; ++i) Gat(i, i)[1] = 0.0;
                ++res.restarts;
            }
            // correction vector into the next basis slot
            dvec2 *t = Vp(m);
            const float *logm = nullptr;
            if (opt.log_multiplier_fn) {
                logm = static_cast<const float *>(opt.log_multiplier_fn(opt.log_multiplier_ctx, theta));
                if (!logm) return fail(TCP_ERR_INVALID_ARGUMENT, "learned preconditioner callback failed");
            }
            hipLaunchKernelGGL(precondition_kernel, dim3(nb), dim3(kThreads), 0, stream,
                               static_cast<const dvec2 *>(w.r.ptr), static_cast<const double *>(w.diag.ptr), logm,
                               theta, 1e-8, t, D);
            for (int pass = 0; pass < 2; ++pass) {  // classical Gram-Schmidt, twice
                if ((st = dots(Vp(0), m, t, col))) return st;
                for (double &v : col) v = -v;
                if ((st = combine(Vp(0), m, col, 1.0, t))) return st;
            }
            double nt;
            if ((st = norm(t, nt))) return st;
            if (!(nt > 1e-14)) break;  // correction lies in the basis: stagnated
            hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(kThreads), 0, stream, t, 1.0 / nt, D);
            if ((st = apply(t, Wp(m), phase, res))) return st;
            if ((st = dots(Vp(0), m + 1, Wp(m), col))) return st;
            for (uint32_t i = 0; i <= m; ++i) {
                Gat(i, m)[0] = col[2 * i];
                Gat(i, m)[1] = col[2 * i + 1];
                Gat(m, i)[0] = col[2 * i];
                Gat(m, i)[1] = -col[2 * i + 1];
            }
            Gat(m, m)[1] = 0.0;
            ++m;
        }
        res.eigenvalue = theta;
        res.residual_norm = rnorm;
        res.final_fp64 = phase == TCP_C128 ? 1 : 0;
        if (evec_out) {
            double nu;
            if ((st = norm(static_cast<dvec2 *>(w.u.ptr), nu))) return st;
            hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(kThreads), 0, stream, static_cast<dvec2 *>(w.u.ptr),
                               1.0 / nu, D);
            TCP_HIP_CHECK(hipMemcpyAsync(evec_out, w.u.ptr, 16 * D, hipMemcpyDeviceToDevice, stream));
        }
        TCP_HIP_CHECK(hipStreamSynchronize(stream));
        res.seconds = std::chrono::duration<double>(clock::now() - t0).count();
        return TCP_OK;
    }
};

template <class Plan>
int davidson(const Plan *plan, const tcp_davidson_options *opt, void *evec, tcp_davidson_result *res) {
    if (!plan || !opt || !res) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    if (opt->mvp_dtype != TCP_C128 && opt->mvp_dtype != TCP_C64)
        return fail(TCP_ERR_INVALID_ARGUMENT, "mvp_dtype must be TCP_C128 or TCP_C64");
    *res = tcp_davidson_result{};
    if (tcp::plan_dimension(plan) < 2) return fail(TCP_ERR_INVALID_ARGUMENT, "dimension must be at least 2");
    tcp::DeviceGuard dg;
    int st = dg.enter(tcp::plan_device(plan));
    if (st) return st;
    Solver<Plan> solver(plan, *opt, static_cast<hipStream_t>(opt->stream));
    st = solver.run(evec, *res);
    return st ? st : tcp::ok();
}

}  // namespace

extern "C" {

void tcp_davidson_default_options(tcp_davidson_options *opt) {
    if (!opt) return;
    *opt = tcp_davidson_options{};
    opt->max_subspace = 24;
    opt->restart_keep = 4;
    opt->max_iterations = 1000;
    opt->residual_tol = 1e-8;
    opt->mvp_dtype = TCP_C128;
    opt->switch_tol = 1e-4;
    opt->seed = 20260801;
}

int tcp_davidson_sector(const tcp_sector_plan *plan, const tcp_davidson_options *opt, void *eigenvector,
                        tcp_davidson_result *result) {
    return davidson(plan, opt, eigenvector, result);
}

int tcp_davidson_full(const tcp_mvp_plan *plan, const tcp_davidson_options *opt, void *eigenvector,
                      tcp_davidson_result *result) {
    return davidson(plan, opt, eigenvector, result);
}

int tcp_sector_plan_diagonal(const tcp_sector_plan *plan, double *diag, void *stream) {
    if (!plan || !diag) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    tcp::DeviceGuard dg;
    int st = dg.enter(tcp::plan_device(plan));
    if (st) return st;
    st = tcp::plan_diagonal(plan, diag, static_cast<hipStream_t>(stream));
    return st ? st : tcp::ok();
}

int tcp_mvp_plan_diagonal(const tcp_mvp_plan *plan, double *diag, void *stream) {
    if (!plan || !diag) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    tcp::DeviceGuard dg;
    int st = dg.enter(tcp::plan_device(plan));
    if (st) return st;
    st = tcp::plan_diagonal(plan, diag, static_cast<hipStream_t>(stream));
    return st ? st : tcp::ok();
}

}  // extern "C"
