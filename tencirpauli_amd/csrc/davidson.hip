// Davidson eigensolver for the lowest eigenpair of an MVP plan, on the GPU.
//
// No reference counterpart: the reference's eigen entry point is SciPy
// eigsh over a LinearOperator (py/integrations/scipy.py:13-47).
//
// Storage and precision.  The search basis V and the products W = H V live on
// the device in the storage precision of the current phase: complex128, or
// complex64 in a first phase whose MVPs use FP32 coefficients with FP64
// accumulation.  Every reduction (residual norms, Gram-Schmidt coefficients,
// Rayleigh-quotient entries) accumulates in FP64 in a fixed order, and only
// the m x m projected problem is solved on the host.  When the complex64 phase
// reaches `switch_tol`, its Ritz vector is formed in FP64 and one FP64 MVP
// gives the FP64 Rayleigh quotient and residual; the solve ends there if that
// residual meets `residual_tol`, otherwise FP64 iterations continue from it.
// Reported eigenvalues of a finished solve therefore belong to the FP64
// operator.
//
// One iteration is three streaming passes plus the MVP, with one host
// synchronisation:
//   K1  u = V s, Hu = W s, r = Hu - theta u, t = P r, and the Gram-Schmidt
//       coefficients <V_j, t> in the same sweep (V_j[i] stays in registers);
//   K2  t <- t - sum_j <V_j, t> V_j, with |t|^2 and <V_j, t> of the result; a
//       second projection (K2 again) runs only if the norm fell below
//       1/sqrt(2) of its value (the test is made on the device);
//   MVP W_m = H t;
//   K4  the new column <V_j, W_m> of the projected matrix.
// Columns are kept unnormalised; their norms live on the host and enter every
// combination as coefficients.  A thick restart writes the kept Ritz vectors
// into free column slots (no copies).
//
// Correction: t_i = r_i / (theta - H_ii) (Jacobi), or t_i = r_i e^{m_i} with
// per-state log values m_i from a learned preconditioner.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "krylov.hpp"
#include "tcp_internal.hpp"

namespace {

constexpr uint32_t kMaxKeep = 8;    // Ritz vectors kept at a restart
constexpr double kEta2 = 0.5;  // second projection when |t'|^2 < kEta2 |t|^2
constexpr uint32_t kStallIterations = 20;  // complex64 phase: iterations without a new best residual
constexpr double kStallFloor = 1e-5;       // ... counted as a stall only below kStallFloor |theta|

// Device scalar slots (dvec2 each): reductions land here, kernels read their
// coefficients from here, and one copy per iteration brings them to the host.
constexpr uint32_t kScR2 = 0;                    // |r|^2
constexpr uint32_t kScT2 = 1;                    // |t|^2 after preconditioning
constexpr uint32_t kScC = 2;                     // <V_j, t>, j < m
constexpr uint32_t kScP1 = kScC + kMaxBasis;     // |t'|^2, then <V_j, t'>
constexpr uint32_t kScP2 = kScP1 + 1 + kMaxBasis;  // |t''|^2, then <V_j, t''>
constexpr uint32_t kScE = kScP2 + 1 + kMaxBasis;   // <V_j, W_m>, j <= m
constexpr uint32_t kScX = kScE + kMaxBasis + 1;    // three spare slots (pair dots)
constexpr uint32_t kScCount = kScX + 3;

struct Coef {
    dvec2 c[kMaxBasis];
};
struct Sig2 {
    double s[kMaxBasis];
};

typedef float fvec4 __attribute__((ext_vector_type(4)));

// Elements per thread of the streaming passes K1, K2 and K4: complex64
// columns go in pairs (one 16-byte access, the size of a complex128 element),
// so a complex64 pass keeps as many bytes in flight per wave as a complex128
// one instead of half as many.  Above 24 columns the pairs would spill.
template <typename ST, int MAXM>
struct Run {
    static constexpr int E = std::is_same<ST, fvec2>::value && MAXM <= 24 ? 2 : 1;
};

// x[0..E) = p[i..i+E) for i a multiple of E.  The pair that starts at the
// last index of an odd-length column reads one element past its end: every
// column these passes touch is a solver slot of 16 D bytes whose complex64
// data fills the first half, so the read stays inside the slot, and the extra
// element is never used.
template <typename ST>
__device__ __forceinline__ void load_run(const ST *p, uint64_t i, ST (&x)[1]) {
    x[0] = p[i];
}
__device__ __forceinline__ void load_run(const fvec2 *p, uint64_t i, fvec2 (&x)[2]) {
    const fvec4 v = *reinterpret_cast<const fvec4 *>(p + i);
    x[0] = fvec2{v.x, v.y};
    x[1] = fvec2{v.z, v.w};
}
template <typename ST>
__device__ __forceinline__ void store_run(ST *p, uint64_t i, const ST (&x)[1], bool) {
    p[i] = x[0];
}
__device__ __forceinline__ void store_run(fvec2 *p, uint64_t i, const fvec2 (&x)[2], bool pair) {
    if (pair)
        *reinterpret_cast<fvec4 *>(p + i) = fvec4{x[0].x, x[0].y, x[1].x, x[1].y};
    else
        p[i] = x[0];
}

// Second projection gate: true when |t'|^2 >= kEta2 |t|^2 (no second pass).
__device__ __forceinline__ bool gate_skips(const dvec2 *__restrict__ sc) {
    return sc[kScP1].x >= kEta2 * sc[kScT2].x;
}

// out[j] = sum over blocks of partial[b * nval + j], fixed order (deterministic).
__global__ __launch_bounds__(kThreads) void reduce_kernel(const dvec2 *__restrict__ partial, uint32_t nblk,
                                                          uint32_t nval, dvec2 *__restrict__ out,
                                                          const dvec2 *__restrict__ gate) {
    if (gate && gate_skips(gate)) return;
    __shared__ dvec2 red[kThreads];
    const uint32_t j = blockIdx.x;
    dvec2 s{0.0, 0.0};
    for (uint32_t b = threadIdx.x; b < nblk; b += kThreads) s += partial[static_cast<uint64_t>(b) * nval + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

// K1: residual, preconditioner and the first Gram-Schmidt coefficients.
// Partials: [|r|^2, |t|^2, <V_0, t>, ..., <V_{m-1}, t>].
template <typename ST, int MAXM, bool LEARNED>
__global__ __launch_bounds__(kThreads) void residual_kernel(Cols<ST> V, Cols<ST> W, uint32_t m, Coef s,
                                                            double theta, const double *__restrict__ diag,
                                                            const float *__restrict__ logm, ST *__restrict__ t,
                                                            uint64_t D, dvec2 *__restrict__ partial) {
    constexpr int E = Run<ST, MAXM>::E;
    dvec2 acc[MAXM + 2];
#pragma unroll
    for (int j = 0; j < MAXM + 2; ++j) acc[j] = dvec2{0.0, 0.0};
    for (uint64_t i0 = E * (blockIdx.x * static_cast<uint64_t>(kThreads) + threadIdx.x); i0 < D;
         i0 += E * static_cast<uint64_t>(gridDim.x) * kThreads) {
        // Every load is issued unconditionally (columns past m alias column 0
        // with zero coefficients), so all of them are in flight together.
        ST vv[MAXM][E], ww[MAXM][E];
#pragma unroll
        for (int j = 0; j < MAXM; ++j) {
            load_run(V.p[j], i0, vv[j]);
            load_run(W.p[j], i0, ww[j]);
        }
        ST ts[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint64_t i = i0 + e;
            const bool in = e == 0 || i < D;
            dvec2 u{0.0, 0.0}, hu{0.0, 0.0};
#pragma unroll
            for (int j = 0; j < MAXM; ++j) {
                u += cmul(s.c[j], widen(vv[j][e]));
                hu += cmul(s.c[j], widen(ww[j][e]));
            }
            const dvec2 r = hu - u * theta;
            double scale;
            if (LEARNED) {
                scale = in ? exp(static_cast<double>(logm[i])) : 0.0;
            } else {
                double den = theta - (in ? diag[i] : 0.0);
                if (fabs(den) < 1e-8) den = den < 0.0 ? -1e-8 : 1e-8;
                scale = 1.0 / den;
            }
            ts[e] = narrow<ST>(r * scale);
            if (in) {
                const dvec2 tw = widen(ts[e]);
                acc[0].x += r.x * r.x + r.y * r.y;
                acc[1].x += tw.x * tw.x + tw.y * tw.y;
#pragma unroll
                for (int j = 0; j < MAXM; ++j) acc[2 + j] += cmulc(widen(vv[j][e]), tw);
            }
        }
        store_run(t, i0, ts, E == 1 || i0 + 1 < D);
    }
    block_partials<MAXM + 2>(acc, m + 2, partial);
}

// K2: t <- t - sum_j sig2_j c_j V_j with c_j = sc[coef + j]; partials
// [|t|^2, <V_0, t>, ..., <V_{m-1}, t>].  With `gate`, does nothing when the
// second projection is not needed.
template <typename ST, int MAXM>
__global__ __launch_bounds__(kThreads) void project_kernel(Cols<ST> V, uint32_t m, Sig2 sig2,
                                                           const dvec2 *__restrict__ sc, uint32_t coef,
                                                           ST *__restrict__ t, uint64_t D,
                                                           dvec2 *__restrict__ partial, bool gate) {
    if (gate && gate_skips(sc)) return;
    constexpr int E = Run<ST, MAXM>::E;
    __shared__ dvec2 c[MAXM];  // uniform coefficients: LDS broadcasts, not per-lane registers
    if (threadIdx.x < MAXM)
        c[threadIdx.x] = threadIdx.x < m ? sc[coef + threadIdx.x] * sig2.s[threadIdx.x] : dvec2{0.0, 0.0};
    __syncthreads();
    dvec2 acc[MAXM + 1];
#pragma unroll
    for (int j = 0; j < MAXM + 1; ++j) acc[j] = dvec2{0.0, 0.0};
    for (uint64_t i0 = E * (blockIdx.x * static_cast<uint64_t>(kThreads) + threadIdx.x); i0 < D;
         i0 += E * static_cast<uint64_t>(gridDim.x) * kThreads) {
        ST vv[MAXM][E], tt[E];
#pragma unroll
        for (int j = 0; j < MAXM; ++j) load_run(V.p[j], i0, vv[j]);
        load_run(t, i0, tt);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const bool in = e == 0 || i0 + e < D;
            dvec2 x = widen(tt[e]);
#pragma unroll
            for (int j = 0; j < MAXM; ++j) x -= cmul(c[j], widen(vv[j][e]));
            tt[e] = narrow<ST>(x);
            if (in) {
                const dvec2 xw = widen(tt[e]);
                acc[0].x += xw.x * xw.x + xw.y * xw.y;
#pragma unroll
                for (int j = 0; j < MAXM; ++j) acc[1 + j] += cmulc(widen(vv[j][e]), xw);
            }
        }
        store_run(t, i0, tt, E == 1 || i0 + 1 < D);
    }
    block_partials<MAXM + 1>(acc, m + 1, partial);
}

// K4: partials <V_j, w>, j = 0..m (V has m + 1 columns here).
template <typename ST, int MAXM>
__global__ __launch_bounds__(kThreads) void column_kernel(Cols<ST> V, uint32_t ncols, const ST *__restrict__ w,
                                                          uint64_t D, dvec2 *__restrict__ partial) {
    constexpr int E = Run<ST, MAXM>::E;
    dvec2 acc[MAXM + 1];
#pragma unroll
    for (int j = 0; j < MAXM + 1; ++j) acc[j] = dvec2{0.0, 0.0};
    for (uint64_t i0 = E * (blockIdx.x * static_cast<uint64_t>(kThreads) + threadIdx.x); i0 < D;
         i0 += E * static_cast<uint64_t>(gridDim.x) * kThreads) {
        ST ww[E], vv[MAXM + 1][E];
        load_run(w, i0, ww);
#pragma unroll
        for (int j = 0; j < MAXM + 1; ++j) load_run(V.p[j], i0, vv[j]);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (e > 0 && i0 + e >= D) break;
            const dvec2 wi = widen(ww[e]);
#pragma unroll
            for (int j = 0; j < MAXM + 1; ++j) acc[j] += cmulc(widen(vv[j][e]), wi);
        }
    }
    block_partials<MAXM + 1>(acc, ncols, partial);
}

// Partials [<a, a>, <a, b>, <b, b>] of two vectors.
template <typename ST>
__global__ __launch_bounds__(kThreads) void pair_kernel(const ST *__restrict__ a, const ST *__restrict__ b, uint64_t D,
                                                        dvec2 *__restrict__ partial) {
    dvec2 acc[3] = {dvec2{0.0, 0.0}, dvec2{0.0, 0.0}, dvec2{0.0, 0.0}};
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kThreads) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * kThreads) {
        const dvec2 x = widen(a[i]), y = widen(b[i]);
        acc[0].x += x.x * x.x + x.y * x.y;
        acc[1] += cmulc(x, y);
        acc[2].x += y.x * y.x + y.y * y.y;
    }
    block_partials<3>(acc, 3, partial);
}

// Partials [|b - theta a|^2]: the explicit residual norm of a Rayleigh pair.
__global__ __launch_bounds__(kThreads) void pair_residual_kernel(const dvec2 *__restrict__ a,
                                                                 const dvec2 *__restrict__ b, double theta,
                                                                 uint64_t D, dvec2 *__restrict__ partial) {
    dvec2 acc[1] = {dvec2{0.0, 0.0}};
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kThreads) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * kThreads) {
        const dvec2 r = b[i] - a[i] * theta;
        acc[0].x += r.x * r.x + r.y * r.y;
    }
    block_partials<1>(acc, 1, partial);
}

// ---------------------------------------------------------------- host side
using tcp::fail;

template <class Plan>
struct Solver {
    const Plan *plan;
    const tcp_davidson_options &opt;
    hipStream_t stream;
    uint64_t D = 0;
    uint32_t nb = 0;
    // column slot pools (each slot holds a complex128 column; complex64
    // columns use its first half)
    tcp::DeviceBuffer vpool, wpool, diag, partial, sc, ybuf, eig;
    std::vector<uint32_t> vslot, wslot;  // column j -> slot
    std::vector<uint32_t> vfree, wfree;
    std::vector<double> sig;  // column j holds sig[j]^-1 times its normalised vector
    std::vector<dvec2> hsc;
    int phase = TCP_C128;
    tcp_davidson_result *res = nullptr;

    Solver(const Plan *p, const tcp_davidson_options &o, hipStream_t s) : plan(p), opt(o), stream(s) {}

    void *vptr(uint32_t slot) { return static_cast<char *>(vpool.ptr) + static_cast<size_t>(slot) * 16 * D; }
    void *wptr(uint32_t slot) { return static_cast<char *>(wpool.ptr) + static_cast<size_t>(slot) * 16 * D; }
    dvec2 *scp(uint32_t k) { return static_cast<dvec2 *>(sc.ptr) + k; }
    dvec2 *part() { return static_cast<dvec2 *>(partial.ptr); }

    // Column pointers; entries past n alias column 0 (the kernels load every
    // entry and weight the extra ones by zero).
    template <typename ST>
    Cols<ST> vcols(uint32_t n) {
        Cols<ST> c{};
        for (uint32_t j = 0; j <= kMaxBasis; ++j) c.p[j] = static_cast<ST *>(vptr(vslot[j < n ? j : 0]));
        return c;
    }
    template <typename ST>
    Cols<ST> wcols(uint32_t n) {
        Cols<ST> c{};
        for (uint32_t j = 0; j <= kMaxBasis; ++j) c.p[j] = static_cast<ST *>(wptr(wslot[j < n ? j : 0]));
        return c;
    }
    uint32_t take(std::vector<uint32_t> &pool) {
        const uint32_t s = pool.back();
        pool.pop_back();
        return s;
    }

    int reduce(uint32_t nval, uint32_t dst, bool gated) {
        hipLaunchKernelGGL(reduce_kernel, dim3(nval), dim3(kThreads), 0, stream, part(), nb, nval, scp(dst),
                           gated ? scp(0) : nullptr);
        return TCP_OK;
    }
    int fetch() {
        hsc.resize(kScCount);
        TCP_HIP_CHECK(hipMemcpyAsync(hsc.data(), sc.ptr, sizeof(dvec2) * kScCount, hipMemcpyDeviceToHost, stream));
        TCP_HIP_CHECK(hipStreamSynchronize(stream));
        return TCP_OK;
    }
    int apply(const void *x, void *y, tcp_davidson_result &r) {
        if (phase == TCP_C64)
            ++r.mvps_c64;
        else
            ++r.mvps_c128;
        return tcp::plan_apply(plan, x, y, phase, stream);
    }
    // <a, a>, <a, b>, <b, b> into sc[kScX..kScX+3)
    template <typename ST>
    int pair_dots(const void *a, const void *b) {
        hipLaunchKernelGGL((pair_kernel<ST>), dim3(nb), dim3(kThreads), 0, stream, static_cast<const ST *>(a),
                           static_cast<const ST *>(b), D, part());
        return reduce(3, kScX, false);
    }
    int pair_dots_phase(const void *a, const void *b) {
        return phase == TCP_C64 ? pair_dots<fvec2>(a, b) : pair_dots<dvec2>(a, b);
    }
    // out_k = sum_j Y[k][j] * (column j of the basis, normalised), k < K
    // out_k = sum_j Y[k * m + j] * (column j of V, or of W with `products`), k < K
    template <typename SI, typename SO>
    int combine(uint32_t m, const std::vector<dvec2> &Y, uint32_t K, const Cols<SO> &out, bool products = false) {
        const Cols<SI> in = products ? wcols<SI>(m) : vcols<SI>(m);
        return with_maxm(m, [&](auto MM) -> int {
            constexpr int M = decltype(MM)::value;
            std::vector<dvec2> Yp(static_cast<size_t>(K) * M, dvec2{0.0, 0.0});
            for (uint32_t k = 0; k < K; ++k)
                for (uint32_t j = 0; j < m; ++j) Yp[k * M + j] = Y[k * m + j];
            TCP_HIP_CHECK(
                hipMemcpyAsync(ybuf.ptr, Yp.data(), sizeof(dvec2) * Yp.size(), hipMemcpyHostToDevice, stream));
            hipLaunchKernelGGL((combine_kernel<SI, SO, M>), dim3(nb), dim3(kThreads), 0, stream, in, m,
                               static_cast<const dvec2 *>(ybuf.ptr), K, out, D);
            return TCP_OK;
        });
    }

    // One iteration's device work after the projected solve: K1, K2 (x2,
    // gated), MVP, K4.  Column m is the new direction.
    template <typename ST>
    int iterate(uint32_t m, const Coef &s, double theta, const float *logm) {
        // K1/K2 see the m basis columns only (the new slot m may hold stale
        // data, which a zero weight would not cancel if it were not finite)
        const Cols<ST> V = vcols<ST>(m), W = wcols<ST>(m), V1 = vcols<ST>(m + 1);
        ST *t = static_cast<ST *>(vptr(vslot[m]));
        ST *wm = static_cast<ST *>(wptr(wslot[m]));
        Sig2 sig2{};
        for (uint32_t j = 0; j < m; ++j) sig2.s[j] = sig[j] * sig[j];
        const double *dg = static_cast<const double *>(diag.ptr);
        int st = with_maxm(m, [&](auto MM) {
            constexpr int M = decltype(MM)::value;
            if (logm)
                hipLaunchKernelGGL((residual_kernel<ST, M, true>), dim3(nb), dim3(kThreads), 0, stream, V, W, m, s,
                                   theta, dg, logm, t, D, part());
            else
                hipLaunchKernelGGL((residual_kernel<ST, M, false>), dim3(nb), dim3(kThreads), 0, stream, V, W, m, s,
                                   theta, dg, logm, t, D, part());
            reduce(m + 2, kScR2, false);
            hipLaunchKernelGGL((project_kernel<ST, M>), dim3(nb), dim3(kThreads), 0, stream, V, m, sig2, scp(0), kScC,
                               t, D, part(), false);
            reduce(m + 1, kScP1, false);
            hipLaunchKernelGGL((project_kernel<ST, M>), dim3(nb), dim3(kThreads), 0, stream, V, m, sig2, scp(0),
                               kScP1 + 1, t, D, part(), true);
            reduce(m + 1, kScP2, true);
            return TCP_OK;
        });
        if (st) return st;
        if ((st = apply(t, wm, *res))) return st;
        return with_maxm(m + 1, [&](auto MM) {
            constexpr int M = decltype(MM)::value;
            hipLaunchKernelGGL((column_kernel<ST, M>), dim3(nb), dim3(kThreads), 0, stream, V1, m + 1, wm, D,
                               part());
            return reduce(m + 1, kScE, false);
        });
    }

    int run(void *evec_out, tcp_davidson_result &result) {
        using clock = std::chrono::steady_clock;
        const auto t0 = clock::now();
        res = &result;
        D = tcp::plan_dimension(plan);
        nb = stream_blocks(D);
        const uint32_t mmax = std::min<uint32_t>(std::max<uint32_t>(opt.max_subspace ? opt.max_subspace : 24, 4),
                                                 kMaxBasis);
        const uint32_t keep = std::min<uint32_t>(
            std::min<uint32_t>(std::max<uint32_t>(opt.restart_keep ? opt.restart_keep : 4, 1), mmax - 2), kMaxKeep);
        const uint32_t max_it = opt.max_iterations ? opt.max_iterations : 1000;
        const double tol = opt.residual_tol > 0 ? opt.residual_tol : 1e-8;
        const double switch_tol = opt.switch_tol > 0 ? opt.switch_tol : 1e-4;
        phase = opt.mvp_dtype == TCP_C64 ? TCP_C64 : TCP_C128;
        const uint32_t slots = mmax + keep;
        int st;
        if ((st = vpool.allocate(static_cast<size_t>(16) * D * slots)) ||
            (st = wpool.allocate(static_cast<size_t>(16) * D * slots)) || (st = diag.allocate(8 * D)) ||
            (st = partial.allocate(sizeof(dvec2) * static_cast<size_t>(nb) * (kMaxBasis + 2))) ||
            (st = sc.allocate(sizeof(dvec2) * kScCount)) ||
            (st = ybuf.allocate(sizeof(dvec2) * kMaxBasis * kMaxKeep)))
            return st;
        TCP_HIP_CHECK(hipMemsetAsync(sc.ptr, 0, sizeof(dvec2) * kScCount, stream));
        // the Jacobi correction's diagonal (a learned preconditioner brings its own multipliers)
        if (!opt.log_multiplier_fn && (st = tcp::plan_diagonal(plan, static_cast<double *>(diag.ptr), stream)))
            return st;
        for (uint32_t k = slots; k-- > 0;) {
            vfree.push_back(k);
            wfree.push_back(k);
        }

        // start vector and its product
        vslot.assign(1, take(vfree));
        wslot.assign(1, take(wfree));
        if (phase == TCP_C64)
            hipLaunchKernelGGL((init_kernel<fvec2>), dim3(nb), dim3(kThreads), 0, stream,
                               static_cast<fvec2 *>(vptr(vslot[0])), opt.seed, D);
        else
            hipLaunchKernelGGL((init_kernel<dvec2>), dim3(nb), dim3(kThreads), 0, stream,
                               static_cast<dvec2 *>(vptr(vslot[0])), opt.seed, D);
        if ((st = apply(vptr(vslot[0]), wptr(wslot[0]), result))) return st;
        if ((st = pair_dots_phase(vptr(vslot[0]), wptr(wslot[0]))) || (st = fetch())) return st;
        sig.assign(1, 1.0 / std::sqrt(hsc[kScX].x));
        std::vector<double> G(2 * kMaxBasis * kMaxBasis, 0.0), vals, vecs;
        auto Gat = [&](uint32_t i, uint32_t j) -> double * { return &G[2 * (i * kMaxBasis + j)]; };
        Gat(0, 0)[0] = hsc[kScX + 1].x * sig[0] * sig[0];
        uint32_t m = 1;
        double theta = Gat(0, 0)[0], rnorm = 0.0;
        std::vector<double> s_host;  // current Ritz coefficients (interleaved complex)
        result.trace_len = 0;
        auto trace = [&](double th, double rn, int bits) {
            if (opt.trace && result.trace_len < opt.trace_capacity) {
                double *row = opt.trace + 4 * result.trace_len;
                row[0] = std::chrono::duration<double>(clock::now() - t0).count();
                row[1] = th;
                row[2] = rn;
                row[3] = bits;
                ++result.trace_len;
            }
        };
        double best32 = HUGE_VAL;  // best complex64-phase residual and its iteration
        uint32_t best32_it = 0;
        bool have_u64 = false;  // the refined FP64 Ritz vector sits in vector slot u64_slot
        uint32_t u64_slot = 0;
        for (uint32_t it = 0; it < max_it; ++it) {
            // projected problem
            std::vector<double> g(2 * m * m);
            for (uint32_t i = 0; i < m; ++i)
                for (uint32_t j = 0; j < m; ++j) {
                    g[2 * (i * m + j)] = Gat(i, j)[0];
                    g[2 * (i * m + j) + 1] = Gat(i, j)[1];
                }
            const int want = static_cast<int>(std::min<uint32_t>(m, keep));
            hermitian_lowest(static_cast<int>(m), g, want, vals, vecs);
            theta = vals[0];
            if (m == mmax) {
                // thick restart: the `want` lowest Ritz vectors into free slots
                std::vector<dvec2> Y(static_cast<size_t>(want) * m);
                for (int k = 0; k < want; ++k)
                    for (uint32_t j = 0; j < m; ++j)
                        Y[k * m + j] = dvec2{vecs[2 * (k * m + j)] * sig[j], vecs[2 * (k * m + j) + 1] * sig[j]};
                std::vector<uint32_t> nv(want), nw(want);
                for (int k = 0; k < want; ++k) {
                    nv[k] = take(vfree);
                    nw[k] = take(wfree);
                }
                if (phase == TCP_C64) {
                    Cols<fvec2> ov{}, ow{};
                    for (int k = 0; k < want; ++k) {
                        ov.p[k] = static_cast<fvec2 *>(vptr(nv[k]));
                        ow.p[k] = static_cast<fvec2 *>(wptr(nw[k]));
                    }
                    if ((st = combine<fvec2, fvec2>(m, Y, want, ov)) || (st = combine<fvec2, fvec2>(m, Y, want, ow, true)))
                        return st;
                } else {
                    Cols<dvec2> ov{}, ow{};
                    for (int k = 0; k < want; ++k) {
                        ov.p[k] = static_cast<dvec2 *>(vptr(nv[k]));
                        ow.p[k] = static_cast<dvec2 *>(wptr(nw[k]));
                    }
                    if ((st = combine<dvec2, dvec2>(m, Y, want, ov)) || (st = combine<dvec2, dvec2>(m, Y, want, ow, true)))
                        return st;
                }
                for (uint32_t j = 0; j < m; ++j) {
                    vfree.push_back(vslot[j]);
                    wfree.push_back(wslot[j]);
                }
                vslot = nv;
                wslot = nw;
                m = static_cast<uint32_t>(want);
                sig.assign(m, 1.0);
                std::fill(G.begin(), G.end(), 0.0);
                for (uint32_t k = 0; k < m; ++k) Gat(k, k)[0] = vals[k];
                vecs.assign(static_cast<size_t>(2) * m, 0.0);
                vecs[0] = 1.0;
                ++result.restarts;
            }
            Coef s{};
            for (uint32_t j = 0; j < m; ++j) s.c[j] = dvec2{vecs[2 * j] * sig[j], vecs[2 * j + 1] * sig[j]};
            const float *logm = nullptr;
            if (opt.log_multiplier_fn) {
                logm = static_cast<const float *>(opt.log_multiplier_fn(opt.log_multiplier_ctx, theta, stream));
                if (!logm) return fail(TCP_ERR_INVALID_ARGUMENT, "learned preconditioner callback failed");
            }
            vslot.push_back(take(vfree));
            wslot.push_back(take(wfree));
            st = phase == TCP_C64 ? iterate<fvec2>(m, s, theta, logm) : iterate<dvec2>(m, s, theta, logm);
            if (st || (st = fetch())) return st;
            rnorm = std::sqrt(hsc[kScR2].x);
            result.iterations = it + 1;
            trace(theta, rnorm, phase == TCP_C64 ? 32 : 64);
            if (phase == TCP_C64 && rnorm < best32) {
                best32 = rnorm;
                best32_it = it;
            }
            const bool done64 = phase == TCP_C128 && rnorm <= tol;
            // the complex64 phase also ends when its residual has stopped
            // improving near the storage rounding floor (~1e-7 |theta|, which
            // may lie above switch_tol); a stall far above it is ordinary
            // Davidson behaviour and the phase goes on
            const bool stalled = it - best32_it >= kStallIterations && best32 <= kStallFloor * std::fabs(theta);
            const bool switch32 = phase == TCP_C64 && (rnorm <= switch_tol || stalled);
            if (done64 || switch32) {
                // drop the direction just built; form the Ritz vector in FP64
                vfree.push_back(vslot.back());
                wfree.push_back(wslot.back());
                vslot.pop_back();
                wslot.pop_back();
                std::vector<dvec2> Y(m);
                for (uint32_t j = 0; j < m; ++j) Y[j] = s.c[j];
                u64_slot = take(vfree);
                Cols<dvec2> out{};
                out.p[0] = static_cast<dvec2 *>(vptr(u64_slot));
                st = phase == TCP_C64 ? combine<fvec2, dvec2>(m, Y, 1, out) : combine<dvec2, dvec2>(m, Y, 1, out);
                if (st) return st;
                have_u64 = true;
                if (done64) {
                    result.converged = 1;
                    break;
                }
                // FP64 Rayleigh quotient and residual of the FP32-phase Ritz vector
                phase = TCP_C128;
                const uint32_t hu_slot = take(wfree);
                if ((st = apply(vptr(u64_slot), wptr(hu_slot), result))) return st;
                if ((st = pair_dots<dvec2>(vptr(u64_slot), wptr(hu_slot))) || (st = fetch())) return st;
                const double uu = hsc[kScX].x;
                theta = hsc[kScX + 1].x / uu;
                hipLaunchKernelGGL(pair_residual_kernel, dim3(nb), dim3(kThreads), 0, stream,
                                   static_cast<const dvec2 *>(vptr(u64_slot)),
                                   static_cast<const dvec2 *>(wptr(hu_slot)), theta, D, part());
                if ((st = reduce(1, kScX, false)) || (st = fetch())) return st;
                rnorm = std::sqrt(hsc[kScX].x / uu);
                trace(theta, rnorm, 64);
                ++result.restarts;
                // continue in FP64 from the refined vector (also when converged:
                // the basis below is what a caller resuming would need)
                for (uint32_t j = 0; j < m; ++j) {
                    vfree.push_back(vslot[j]);
                    wfree.push_back(wslot[j]);
                }
                vslot.assign(1, u64_slot);
                wslot.assign(1, hu_slot);
                sig.assign(1, 1.0 / std::sqrt(uu));
                m = 1;
                std::fill(G.begin(), G.end(), 0.0);
                Gat(0, 0)[0] = theta;
                if (rnorm <= tol) {
                    result.converged = 1;
                    break;
                }
                have_u64 = false;
                continue;
            }
            // accept the new direction
            const double t2 = hsc[kScT2].x, p1 = hsc[kScP1].x;
            const double n2 = p1 >= kEta2 * t2 ? p1 : hsc[kScP2].x;
            if (!(n2 > 1e-28 * t2) || !(n2 > 0.0)) {
                // the correction lies in the basis: stagnated
                vfree.push_back(vslot.back());
                wfree.push_back(wslot.back());
                vslot.pop_back();
                wslot.pop_back();
                break;
            }
            sig.push_back(1.0 / std::sqrt(n2));
            for (uint32_t j = 0; j <= m; ++j) {
                const dvec2 e = hsc[kScE + j];
                const double f = sig[j] * sig[m];
                Gat(j, m)[0] = e.x * f;
                Gat(j, m)[1] = e.y * f;
                Gat(m, j)[0] = e.x * f;
                Gat(m, j)[1] = -e.y * f;
            }
            Gat(m, m)[1] = 0.0;
            ++m;
        }
        result.eigenvalue = theta;
        result.residual_norm = rnorm;
        result.final_fp64 = phase == TCP_C128 ? 1 : 0;
        if (evec_out) {
            if (!have_u64) {
                // Ritz vector of the current basis
                std::vector<double> gm(2 * m * m);
                for (uint32_t i = 0; i < m; ++i)
                    for (uint32_t j = 0; j < m; ++j) {
                        gm[2 * (i * m + j)] = Gat(i, j)[0];
                        gm[2 * (i * m + j) + 1] = Gat(i, j)[1];
                    }
                hermitian_lowest(static_cast<int>(m), gm, 1, vals, vecs);
                std::vector<dvec2> Y(m);
                for (uint32_t j = 0; j < m; ++j) Y[j] = dvec2{vecs[2 * j] * sig[j], vecs[2 * j + 1] * sig[j]};
                u64_slot = take(vfree);
                Cols<dvec2> out{};
                out.p[0] = static_cast<dvec2 *>(vptr(u64_slot));
                st = phase == TCP_C64 ? combine<fvec2, dvec2>(m, Y, 1, out) : combine<dvec2, dvec2>(m, Y, 1, out);
                if (st) return st;
            }
            if ((st = pair_dots<dvec2>(vptr(u64_slot), vptr(u64_slot))) || (st = fetch())) return st;
            hipLaunchKernelGGL((scale_kernel<dvec2>), dim3(nb), dim3(kThreads), 0, stream,
                               static_cast<dvec2 *>(vptr(u64_slot)), 1.0 / std::sqrt(hsc[kScX].x), D);
            TCP_HIP_CHECK(hipMemcpyAsync(evec_out, vptr(u64_slot), 16 * D, hipMemcpyDeviceToDevice, stream));
        }
        TCP_HIP_CHECK(hipGetLastError());
        TCP_HIP_CHECK(hipStreamSynchronize(stream));
        result.seconds = std::chrono::duration<double>(clock::now() - t0).count();
        return TCP_OK;
    }
};

template <class Plan>
int davidson(const Plan *plan, const tcp_davidson_options *opt, void *evec, tcp_davidson_result *res) {
    if (!plan || !opt || !res) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    if (opt->mvp_dtype != TCP_C128 && opt->mvp_dtype != TCP_C64)
        return fail(TCP_ERR_INVALID_ARGUMENT, "mvp_dtype must be TCP_C128 or TCP_C64");
    if (opt->max_subspace > kMaxBasis)
        return fail(TCP_ERR_INVALID_ARGUMENT, "max_subspace must be at most " + std::to_string(kMaxBasis));
    *res = tcp_davidson_result{};
    if (tcp::plan_dimension(plan) < 2) return fail(TCP_ERR_INVALID_ARGUMENT, "dimension must be at least 2");
    if (opt->log_multiplier_fn && opt->log_multiplier_len != tcp::plan_dimension(plan))
        return fail(TCP_ERR_INVALID_STRUCTURE_LENGTH,
                    "expected structure length " + std::to_string(tcp::plan_dimension(plan)) + ", got " +
                        std::to_string(opt->log_multiplier_len) + " (learned preconditioner multipliers)");
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

int tcp_davidson_csr(const tcp_csr_plan *plan, const tcp_davidson_options *opt, void *eigenvector,
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

int tcp_csr_plan_diagonal(const tcp_csr_plan *plan, double *diag, void *stream) {
    if (!plan || !diag) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    tcp::DeviceGuard dg;
    int st = dg.enter(tcp::plan_device(plan));
    if (st) return st;
    st = tcp::plan_diagonal(plan, diag, static_cast<hipStream_t>(stream));
    return st ? st : tcp::ok();
}

}  // extern "C"
