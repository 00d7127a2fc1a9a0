// Thick-restart Lanczos for a few extreme eigenpairs of an MVP plan, on the
// GPU, in FP64 (TRLan: K. Wu and H. Simon, SIAM J. Matrix Anal. Appl. 22
// (2000) 602-616; the Krylov-Schur form of G. W. Stewart, SIAM J. Matrix Anal.
// Appl. 23 (2001) 601-614).
//
// No reference counterpart: the reference's eigen entry point is SciPy eigsh
// (ARPACK implicitly restarted Lanczos on the host) over a LinearOperator of
// its native plan (py/integrations/scipy.py:13-47, called as in
// examples/research/fermi_hubbard/run_tencirpauli.py:194-211).  Here the
// basis, the products and every reduction stay on the device; the host sees
// one column of the projected matrix per step and solves the ncv x ncv
// problem at each restart.
//
// Step j (columns v_0..v_j orthonormal, v_j the newest):
//   MVP  w = A v_j, into the slot of column j + 1;
//   D    h_i = <v_i, w> (i <= j) and |w|^2 in one sweep;
//   O    w <- w - sum_i h_i v_i, with |w'|^2 and <v_i, w'>;
//   O'   only when |w'|^2 < |w|^2 / 2 (tested on the device): the same
//        projection again with the new coefficients (DGKS reorthogonalisation);
//   F    on the device: column j of T = V^H A V is h (plus the second-pass
//        correction) and beta_j = |w|, written to step j's row of a history;
//        v_{j+1} = w / beta_j (scaled by zero on a breakdown).
// The steps of one expansion run back to back with no host synchronisation;
// the host reads the history once, when the basis is full, and assembles T
// (a breakdown found there truncates the expansion at its step and continues
// from a fresh vector as before).
// When the basis holds ncv columns, the Ritz pairs (theta, y) of T give the
// residual estimates |beta_{ncv-1}| |y_{ncv-1}|; converged when every wanted
// pair meets tol * max(|theta|, eps^(2/3)) (ARPACK's test).  Otherwise the
// `keep` best Ritz vectors V y replace the basis in place, T becomes
// diag(theta), and v_ncv continues as the next column: its first step's dots
// give the arrow couplings A V y = theta V y + beta y_{ncv-1} v_ncv.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <vector>

#include "krylov.hpp"
#include "tcp_internal.hpp"

namespace {

using tcp::fail;

constexpr double kEta2 = 0.5;  // second projection when |w'|^2 < kEta2 |w|^2
// Scalar slots: [0] |w|^2, [1, 1 + n) first dots; kP1: |w'|^2 then second
// dots; kP2: |w''|^2 then third dots (unused values of a skipped pass).
constexpr uint32_t kP1 = 1 + kMaxBasis;
constexpr uint32_t kP2 = kP1 + 1 + kMaxBasis;
constexpr uint32_t kScalars = kP2 + 1 + kMaxBasis;

__device__ __forceinline__ bool second_pass_needed(const dvec2 *__restrict__ sc) {
    return sc[kP1].x < kEta2 * sc[0].x;
}

// Partials [|w|^2, <V_0, w>, ..., <V_{n-1}, w>]; columns past n alias column 0.
template <int MAXM>
__global__ __launch_bounds__(kThreads) void dots_kernel(Cols<dvec2> V, uint32_t n, const dvec2 *__restrict__ w,
                                                        uint64_t D, dvec2 *__restrict__ partial) {
    dvec2 acc[MAXM + 1];
#pragma unroll
    for (int j = 0; j < MAXM + 1; ++j) acc[j] = dvec2{0.0, 0.0};
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kThreads) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * kThreads) {
        dvec2 vv[MAXM];
#pragma unroll
        for (int j = 0; j < MAXM; ++j) vv[j] = V.p[j][i];
        const dvec2 wi = w[i];
        acc[0].x += wi.x * wi.x + wi.y * wi.y;
#pragma unroll
        for (int j = 0; j < MAXM; ++j) acc[1 + j] += cmulc(vv[j], wi);
    }
    block_partials<MAXM + 1>(acc, n + 1, partial);
}

// w <- w - sum_j c_j V_j with c_j = sc[coef + j]; partials [|w|^2, <V_j, w>].
// With `gated`, returns at once unless the second projection is needed.
template <int MAXM>
__global__ __launch_bounds__(kThreads) void orth_kernel(Cols<dvec2> V, uint32_t n, const dvec2 *__restrict__ sc,
                                                        uint32_t coef, dvec2 *__restrict__ w, uint64_t D,
                                                        dvec2 *__restrict__ partial, bool gated) {
    if (gated && !second_pass_needed(sc)) return;
    __shared__ dvec2 c[MAXM];
    if (threadIdx.x < MAXM) c[threadIdx.x] = threadIdx.x < n ? sc[coef + threadIdx.x] : dvec2{0.0, 0.0};
    __syncthreads();
    dvec2 acc[MAXM + 1];
#pragma unroll
    for (int j = 0; j < MAXM + 1; ++j) acc[j] = dvec2{0.0, 0.0};
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kThreads) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * kThreads) {
        dvec2 vv[MAXM];
#pragma unroll
        for (int j = 0; j < MAXM; ++j) vv[j] = V.p[j][i];
        dvec2 x = w[i];
#pragma unroll
        for (int j = 0; j < MAXM; ++j) x -= cmul(c[j], vv[j]);
        w[i] = x;
        acc[0].x += x.x * x.x + x.y * x.y;
#pragma unroll
        for (int j = 0; j < MAXM; ++j) acc[1 + j] += cmulc(vv[j], x);
    }
    block_partials<MAXM + 1>(acc, n + 1, partial);
}

// out[j] = sum over blocks of partial[b * nval + j] in a fixed order.
__global__ __launch_bounds__(kThreads) void sum_kernel(const dvec2 *__restrict__ partial, uint32_t nblk, uint32_t nval,
                                                       dvec2 *__restrict__ out, const dvec2 *__restrict__ gate) {
    if (gate && !second_pass_needed(gate)) return;
    __shared__ dvec2 red[kThreads];
    const uint32_t j = blockIdx.x;
    dvec2 s{0.0, 0.0};
    for (uint32_t b = threadIdx.x; b < nblk; b += kThreads) s += partial[static_cast<uint64_t>(b) * nval + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t h = kThreads / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

// Row layout of the step history: h_0..h_{n-1}, then {|w|^2 before, after}
// and {1 / beta (0 on a breakdown or a non-finite norm), 0}.
constexpr uint32_t kHistRow = kMaxBasis + 2;
constexpr double kBreakdown = 1e-24;  // |w'|^2 <= kBreakdown |w|^2: an invariant subspace

// Step j's row of the history from the reduced scalars of its projections
// (the device's own DGKS test decides which pass's values count).
__global__ void finalize_kernel(const dvec2 *__restrict__ sc, uint32_t n, dvec2 *__restrict__ row) {
    const bool second = second_pass_needed(sc);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        row[i] = second ? sc[1 + i] + sc[kP1 + 1 + i] : sc[1 + i];
    if (threadIdx.x == 0) {
        const double before2 = sc[0].x, after2 = second ? sc[kP2].x : sc[kP1].x;
        row[kMaxBasis] = dvec2{before2, after2};
        const bool ok = after2 > kBreakdown * before2 && after2 > 0.0 && after2 < HUGE_VAL && before2 < HUGE_VAL;
        row[kMaxBasis + 1] = dvec2{ok ? 1.0 / sqrt(after2) : 0.0, 0.0};
    }
}

// x *= *alpha (a factor the device computed).
__global__ void scale_by_kernel(dvec2 *__restrict__ x, const dvec2 *__restrict__ alpha, uint64_t D) {
    const double a = alpha->x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        x[i] *= a;
}

// |A x - theta x|^2 partials for the explicit residual of a returned pair.
__global__ __launch_bounds__(kThreads) void residual2_kernel(const dvec2 *__restrict__ x, const dvec2 *__restrict__ ax,
                                                             double theta, uint64_t D, dvec2 *__restrict__ partial) {
    dvec2 acc[1] = {dvec2{0.0, 0.0}};
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kThreads) + threadIdx.x; i < D;
         i += static_cast<uint64_t>(gridDim.x) * kThreads) {
        const dvec2 r = ax[i] - x[i] * theta;
        acc[0].x += r.x * r.x + r.y * r.y;
    }
    block_partials<1>(acc, 1, partial);
}

// Order of the Ritz values (ascending) by how much they are wanted.
std::vector<int> wanted_order(const std::vector<double> &vals, int which) {
    std::vector<int> order(vals.size());
    std::iota(order.begin(), order.end(), 0);
    if (which == TCP_WHICH_LA)
        std::reverse(order.begin(), order.end());
    else if (which == TCP_WHICH_LM)
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return std::fabs(vals[a]) > std::fabs(vals[b]); });
    return order;
}

template <class Plan>
struct Lanczos {
    const Plan *plan;
    const tcp_lanczos_options &opt;
    hipStream_t stream;
    uint64_t D = 0;
    uint32_t nb = 0;
    tcp::DeviceBuffer pool, partial, sc, ybuf, hist;
    std::vector<uint32_t> slot;  // column j -> vector slot
    std::vector<dvec2> hsc;
    tcp_lanczos_result *res = nullptr;

    Lanczos(const Plan *p, const tcp_lanczos_options &o, hipStream_t s) : plan(p), opt(o), stream(s) {}

    dvec2 *vec(uint32_t s) { return static_cast<dvec2 *>(pool.ptr) + static_cast<size_t>(s) * D; }
    dvec2 *scp(uint32_t k) { return static_cast<dvec2 *>(sc.ptr) + k; }
    dvec2 *part() { return static_cast<dvec2 *>(partial.ptr); }
    Cols<dvec2> cols(uint32_t n) {
        Cols<dvec2> c{};
        for (uint32_t j = 0; j <= kMaxBasis; ++j) c.p[j] = vec(slot[j < n ? j : 0]);
        return c;
    }
    void sum(uint32_t nval, uint32_t dst, bool gated) {
        hipLaunchKernelGGL(sum_kernel, dim3(nval), dim3(kThreads), 0, stream, part(), nb, nval, scp(dst),
                           gated ? scp(0) : nullptr);
    }
    int fetch() {
        hsc.resize(kScalars);
        TCP_HIP_CHECK(hipMemcpyAsync(hsc.data(), sc.ptr, sizeof(dvec2) * kScalars, hipMemcpyDeviceToHost, stream));
        TCP_HIP_CHECK(hipStreamSynchronize(stream));
        return TCP_OK;
    }

    // Orthogonalise w against columns 0..n-1 (classical Gram-Schmidt with a
    // DGKS second pass); h receives <v_i, w> including the correction, and
    // the squared norms before and after.
    int orthogonalize(uint32_t n, dvec2 *w, std::vector<dvec2> &h, double &before2, double &after2) {
        if (n == 0) {
            hipLaunchKernelGGL(dots_kernel<4>, dim3(nb), dim3(kThreads), 0, stream, cols(1), 0u, w, D, part());
            sum(1, 0, false);
            int st = fetch();
            if (st) return st;
            h.clear();
            before2 = after2 = hsc[0].x;
            return TCP_OK;
        }
        const Cols<dvec2> V = cols(n);
        int st = with_maxm(n, [&](auto MM) -> int {
            constexpr int M = decltype(MM)::value;
            hipLaunchKernelGGL(dots_kernel<M>, dim3(nb), dim3(kThreads), 0, stream, V, n, w, D, part());
            sum(n + 1, 0, false);
            hipLaunchKernelGGL(orth_kernel<M>, dim3(nb), dim3(kThreads), 0, stream, V, n, scp(0), 1u, w, D, part(),
                               false);
            sum(n + 1, kP1, false);
            hipLaunchKernelGGL(orth_kernel<M>, dim3(nb), dim3(kThreads), 0, stream, V, n, scp(0), kP1 + 1, w, D,
                               part(), true);
            sum(n + 1, kP2, true);
            return TCP_OK;
        });
        if (st || (st = fetch())) return st;
        const bool second = hsc[kP1].x < kEta2 * hsc[0].x;  // the device's test on the same values
        h.resize(n);
        for (uint32_t i = 0; i < n; ++i) h[i] = second ? hsc[1 + i] + hsc[kP1 + 1 + i] : hsc[1 + i];
        before2 = hsc[0].x;
        after2 = second ? hsc[kP2].x : hsc[kP1].x;
        return TCP_OK;
    }

    // The projections of orthogonalize (n >= 1) and the history row of step
    // n - 1, with no host synchronisation.
    int orthogonalize_async(uint32_t n, dvec2 *w) {
        const Cols<dvec2> V = cols(n);
        int st = with_maxm(n, [&](auto MM) -> int {
            constexpr int M = decltype(MM)::value;
            hipLaunchKernelGGL(dots_kernel<M>, dim3(nb), dim3(kThreads), 0, stream, V, n, w, D, part());
            sum(n + 1, 0, false);
            hipLaunchKernelGGL(orth_kernel<M>, dim3(nb), dim3(kThreads), 0, stream, V, n, scp(0), 1u, w, D, part(),
                               false);
            sum(n + 1, kP1, false);
            hipLaunchKernelGGL(orth_kernel<M>, dim3(nb), dim3(kThreads), 0, stream, V, n, scp(0), kP1 + 1, w, D,
                               part(), true);
            sum(n + 1, kP2, true);
            return TCP_OK;
        });
        if (st) return st;
        dvec2 *row = static_cast<dvec2 *>(hist.ptr) + static_cast<size_t>(n - 1) * kHistRow;
        hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, stream, static_cast<const dvec2 *>(sc.ptr), n,
                           row);
        hipLaunchKernelGGL(scale_by_kernel, dim3(nb), dim3(kThreads), 0, stream, w, row + kMaxBasis + 1, D);
        return TCP_OK;
    }

    int scale(dvec2 *x, double alpha) {
        hipLaunchKernelGGL(scale_kernel<dvec2>, dim3(nb), dim3(kThreads), 0, stream, x, alpha, D);
        return TCP_OK;
    }

    int run(double *evals, void *evecs, double *residuals, tcp_lanczos_result &result) {
        using clock = std::chrono::steady_clock;
        const auto t0 = clock::now();
        res = &result;
        D = tcp::plan_dimension(plan);
        nb = stream_blocks(D);
        const uint32_t nev = opt.nev;
        const uint32_t ncv = opt.ncv;
        const double tol = opt.tol > 0 ? opt.tol : 2.220446049250313e-16;
        const double eps23 = std::pow(2.220446049250313e-16, 2.0 / 3.0);
        const uint32_t keep = opt.keep ? opt.keep : std::min(ncv - 1, nev + (ncv - nev) / 2);
        int st;
        if ((st = pool.allocate(static_cast<size_t>(16) * D * (ncv + 3))) ||
            (st = partial.allocate(sizeof(dvec2) * static_cast<size_t>(nb) * (kMaxBasis + 1))) ||
            (st = sc.allocate(sizeof(dvec2) * kScalars)) || (st = ybuf.allocate(sizeof(dvec2) * kMaxBasis * kMaxBasis)) ||
            (st = hist.allocate(sizeof(dvec2) * kMaxBasis * kHistRow)))
            return st;
        TCP_HIP_CHECK(hipMemsetAsync(sc.ptr, 0, sizeof(dvec2) * kScalars, stream));
        std::vector<uint32_t> free_slots;
        for (uint32_t s = ncv + 3; s-- > 0;) free_slots.push_back(s);
        auto take = [&]() {
            const uint32_t s = free_slots.back();
            free_slots.pop_back();
            return s;
        };

        // start vector
        slot.assign(1, take());
        if (opt.v0)
            TCP_HIP_CHECK(hipMemcpyAsync(vec(slot[0]), opt.v0, 16 * D, hipMemcpyDeviceToDevice, stream));
        else
            hipLaunchKernelGGL(init_kernel<dvec2>, dim3(nb), dim3(kThreads), 0, stream, vec(slot[0]), opt.seed, D);
        std::vector<dvec2> h;
        double before2 = 0.0, after2 = 0.0;
        if ((st = orthogonalize(0, vec(slot[0]), h, before2, after2))) return st;
        if (!(after2 > 0.0) || !std::isfinite(after2)) return fail(TCP_ERR_INVALID_ARGUMENT, "start vector is zero");
        scale(vec(slot[0]), 1.0 / std::sqrt(after2));

        const uint32_t max_restarts = opt.max_restarts;
        std::vector<double> T(2 * kMaxBasis * kMaxBasis, 0.0);
        auto Tat = [&](uint32_t i, uint32_t j) -> double * { return &T[2 * (i * kMaxBasis + j)]; };
        uint32_t m = 0;       // finished columns of T
        double beta = 0.0;    // |w| of the last step: coupling of v_m to the basis
        bool exhausted = false;
        uint64_t fresh = 0;   // replacement vectors drawn after a breakdown
        std::vector<double> vals, vecs;
        std::vector<int> order;
        for (uint32_t restart = 0;; ++restart) {
            // expand to ncv columns: the steps back to back on the device,
            // then one read of their history
            while (m < ncv && !exhausted) {
                const uint32_t j0 = m;
                for (uint32_t j = j0; j < ncv; ++j) {
                    slot.push_back(take());
                    ++result.mvps;
                    if ((st = tcp::plan_apply(plan, vec(slot[j]), vec(slot[j + 1]), TCP_C128, stream))) return st;
                    if ((st = orthogonalize_async(j + 1, vec(slot[j + 1])))) return st;
                }
                std::vector<dvec2> rows(static_cast<size_t>(ncv - j0) * kHistRow);
                TCP_HIP_CHECK(hipMemcpyAsync(rows.data(), static_cast<dvec2 *>(hist.ptr) + static_cast<size_t>(j0) * kHistRow,
                                             sizeof(dvec2) * rows.size(), hipMemcpyDeviceToHost, stream));
                TCP_HIP_CHECK(hipStreamSynchronize(stream));
                for (uint32_t j = j0; j < ncv; ++j) {
                    const dvec2 *row = &rows[static_cast<size_t>(j - j0) * kHistRow];
                    before2 = row[kMaxBasis].x;
                    after2 = row[kMaxBasis].y;
                    if (!std::isfinite(after2) || !std::isfinite(before2))
                        return fail(TCP_ERR_INVALID_ARGUMENT, "non-finite value in the Lanczos recurrence");
                    for (uint32_t i = 0; i <= j; ++i) {
                        Tat(i, j)[0] = row[i].x;
                        Tat(i, j)[1] = row[i].y;
                        Tat(j, i)[0] = row[i].x;
                        Tat(j, i)[1] = -row[i].y;
                    }
                    Tat(j, j)[1] = 0.0;
                    m = j + 1;
                    if (after2 > kBreakdown * before2 && after2 > 0.0) {
                        beta = std::sqrt(after2);  // the device scaled v_{j+1} by 1 / beta
                        continue;
                    }
                    // invariant subspace at step j: the later steps ran on a zero
                    // vector and are dropped; continue from a fresh vector
                    // orthogonal to the basis (coupling zero), or stop when the
                    // space is spanned
                    for (uint32_t k = j + 2; k < slot.size(); ++k) free_slots.push_back(slot[k]);
                    slot.resize(j + 2);
                    beta = 0.0;
                    if (m >= D) {
                        exhausted = true;
                        break;
                    }
                    dvec2 *w = vec(slot[j + 1]);
                    hipLaunchKernelGGL(init_kernel<dvec2>, dim3(nb), dim3(kThreads), 0, stream, w,
                                       opt.seed + 0x9e3779b97f4a7c15ull * (++fresh), D);
                    std::vector<dvec2> discard;
                    if ((st = orthogonalize(m, w, discard, before2, after2))) return st;
                    if (!(after2 > 1e-20 * before2) || !(after2 > 0.0)) {
                        exhausted = true;
                        break;
                    }
                    scale(w, 1.0 / std::sqrt(after2));
                    break;
                }
            }

            // Ritz pairs of T (ascending), residual estimates of the wanted ones
            std::vector<double> g(2 * m * m);
            for (uint32_t i = 0; i < m; ++i)
                for (uint32_t j = 0; j < m; ++j) {
                    g[2 * (i * m + j)] = Tat(i, j)[0];
                    g[2 * (i * m + j) + 1] = Tat(i, j)[1];
                }
            hermitian_lowest(static_cast<int>(m), g, static_cast<int>(m), vals, vecs);
            order = wanted_order(vals, opt.which);
            uint32_t nconv = 0;
            for (uint32_t k = 0; k < nev; ++k) {
                const int c = order[k];
                const double *y = &vecs[static_cast<size_t>(2) * c * m];
                const double est = beta * std::hypot(y[2 * (m - 1)], y[2 * (m - 1) + 1]);
                if (est <= tol * std::max(std::fabs(vals[c]), eps23)) ++nconv;
            }
            result.converged = nconv;
            result.restarts = restart;
            if (nconv == nev || exhausted || restart >= max_restarts) break;

            // thick restart in place: columns 0..keep-1 <- V y of the best Ritz
            // vectors, T <- diag(theta), v_m continues as column `keep`
            const uint32_t K = keep;
            st = with_maxm(m, [&](auto MM) -> int {
                constexpr int M = decltype(MM)::value;
                std::vector<dvec2> Y(static_cast<size_t>(K) * M, dvec2{0.0, 0.0});
                for (uint32_t k = 0; k < K; ++k) {
                    const double *y = &vecs[static_cast<size_t>(2) * order[k] * m];
                    for (uint32_t j = 0; j < m; ++j) Y[k * M + j] = dvec2{y[2 * j], y[2 * j + 1]};
                }
                TCP_HIP_CHECK(
                    hipMemcpyAsync(ybuf.ptr, Y.data(), sizeof(dvec2) * Y.size(), hipMemcpyHostToDevice, stream));
                // each element's m inputs are read before its K <= m outputs
                // are written, by the same thread: in place is safe
                const Cols<dvec2> V = cols(m);
                hipLaunchKernelGGL((combine_kernel<dvec2, dvec2, M>), dim3(nb), dim3(kThreads), 0, stream, V, m,
                                   static_cast<const dvec2 *>(ybuf.ptr), K, V, D);
                return TCP_OK;
            });
            if (st) return st;
            const uint32_t next = slot[m];
            for (uint32_t j = K; j < m; ++j) free_slots.push_back(slot[j]);
            slot.resize(K);
            slot.push_back(next);
            std::fill(T.begin(), T.end(), 0.0);
            for (uint32_t k = 0; k < K; ++k) Tat(k, k)[0] = vals[order[k]];
            m = K;
        }

        // wanted pairs in ascending eigenvalue order
        std::vector<int> chosen(order.begin(), order.begin() + nev);
        std::sort(chosen.begin(), chosen.end(), [&](int a, int b) { return vals[a] < vals[b]; });
        for (uint32_t k = 0; k < nev; ++k) evals[k] = vals[chosen[k]];
        if (evecs || residuals) {
            // one Ritz vector at a time through the two spare slots (the basis
            // holds at most ncv + 1 of the ncv + 3)
            const uint32_t out_slot = take();
            const uint32_t ax_slot = take();
            for (uint32_t k = 0; k < nev; ++k) {
                st = with_maxm(m, [&](auto MM) -> int {
                    constexpr int M = decltype(MM)::value;
                    std::vector<dvec2> Y(M, dvec2{0.0, 0.0});
                    const double *y = &vecs[static_cast<size_t>(2) * chosen[k] * m];
                    for (uint32_t j = 0; j < m; ++j) Y[j] = dvec2{y[2 * j], y[2 * j + 1]};
                    TCP_HIP_CHECK(
                        hipMemcpyAsync(ybuf.ptr, Y.data(), sizeof(dvec2) * M, hipMemcpyHostToDevice, stream));
                    Cols<dvec2> out{};
                    out.p[0] = vec(out_slot);
                    hipLaunchKernelGGL((combine_kernel<dvec2, dvec2, M>), dim3(nb), dim3(kThreads), 0, stream,
                                       cols(m), m, static_cast<const dvec2 *>(ybuf.ptr), 1u, out, D);
                    return TCP_OK;
                });
                if (st) return st;
                if (evecs)
                    TCP_HIP_CHECK(hipMemcpyAsync(static_cast<dvec2 *>(evecs) + static_cast<size_t>(k) * D,
                                                 vec(out_slot), 16 * D, hipMemcpyDeviceToDevice, stream));
                if (residuals) {
                    ++result.mvps;
                    if ((st = tcp::plan_apply(plan, vec(out_slot), vec(ax_slot), TCP_C128, stream))) return st;
                    hipLaunchKernelGGL(residual2_kernel, dim3(nb), dim3(kThreads), 0, stream, vec(out_slot),
                                       vec(ax_slot), evals[k], D, part());
                    sum(1, 0, false);
                    if ((st = fetch())) return st;
                    residuals[k] = std::sqrt(hsc[0].x);
                }
            }
        }
        TCP_HIP_CHECK(hipGetLastError());
        TCP_HIP_CHECK(hipStreamSynchronize(stream));
        result.seconds = std::chrono::duration<double>(clock::now() - t0).count();
        return TCP_OK;
    }
};

template <class Plan>
int lanczos(const Plan *plan, const tcp_lanczos_options *opt, double *evals, void *evecs, double *residuals,
            tcp_lanczos_result *res) {
    if (!plan || !opt || !res || !evals) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    *res = tcp_lanczos_result{};
    const uint64_t D = tcp::plan_dimension(plan);
    if (opt->which != TCP_WHICH_SA && opt->which != TCP_WHICH_LA && opt->which != TCP_WHICH_LM)
        return fail(TCP_ERR_INVALID_ARGUMENT, "which must be TCP_WHICH_SA, TCP_WHICH_LA or TCP_WHICH_LM");
    if (opt->nev < 1 || static_cast<uint64_t>(opt->nev) >= D)
        return fail(TCP_ERR_INVALID_ARGUMENT, "nev must satisfy 1 <= nev < dimension");
    if (opt->ncv <= opt->nev || opt->ncv > kMaxBasis || opt->ncv > D)
        return fail(TCP_ERR_INVALID_ARGUMENT,
                    "ncv must satisfy nev < ncv <= min(dimension, " + std::to_string(kMaxBasis) + ")");
    if (!(opt->tol >= 0.0)) return fail(TCP_ERR_INVALID_ARGUMENT, "tol must be non-negative");
    if (opt->keep && (opt->keep < opt->nev || opt->keep >= opt->ncv))
        return fail(TCP_ERR_INVALID_ARGUMENT, "keep must satisfy nev <= keep < ncv");
    tcp::DeviceGuard dg;
    int st = dg.enter(tcp::plan_device(plan));
    if (st) return st;
    Lanczos<Plan> solver(plan, *opt, static_cast<hipStream_t>(opt->stream));
    st = solver.run(evals, evecs, residuals, *res);
    return st ? st : tcp::ok();
}

}  // namespace

extern "C" {

void tcp_lanczos_default_options(tcp_lanczos_options *opt) {
    if (!opt) return;
    *opt = tcp_lanczos_options{};
    opt->nev = 1;
    opt->ncv = 20;
    opt->which = TCP_WHICH_SA;
    opt->max_restarts = 1000;
    opt->tol = 0.0;
    opt->seed = 20260801;
}

int tcp_lanczos_full(const tcp_mvp_plan *plan, const tcp_lanczos_options *opt, double *eigenvalues,
                     void *eigenvectors, double *residual_norms, tcp_lanczos_result *result) {
    return lanczos(plan, opt, eigenvalues, eigenvectors, residual_norms, result);
}

int tcp_lanczos_sector(const tcp_sector_plan *plan, const tcp_lanczos_options *opt, double *eigenvalues,
                       void *eigenvectors, double *residual_norms, tcp_lanczos_result *result) {
    return lanczos(plan, opt, eigenvalues, eigenvectors, residual_norms, result);
}

int tcp_lanczos_csr(const tcp_csr_plan *plan, const tcp_lanczos_options *opt, double *eigenvalues,
                    void *eigenvectors, double *residual_norms, tcp_lanczos_result *result) {
    return lanczos(plan, opt, eigenvalues, eigenvectors, residual_norms, result);
}

}  // extern "C"
