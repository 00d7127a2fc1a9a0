// Full-space matrix-free Pauli MVP: plan builder and gfx950 kernels.
//
// Replaces MvpPlan::from_operator / apply_into (core/hamiltonian.rs:63-76,
// 204-270).  The reference loops every row over every canonical term:
//     y[r] = sum_t w_t (-1)^popc(z_t & (r ^ x_t)) v[r ^ x_t].
// Here the sign is moved to the destination index,
//     popc(z & (r ^ x)) = popc(z & r) + popc(z & x)  (mod 2),
// so w'_t = w_t (-1)^popc(z_t & x_t) and every term of one x-mask group
// shares one source amplitude v[r ^ x].
//
// Layout.  The index space is cut into tiles of 2^L amplitudes (L = 12 for
// n >= 12): a 512-thread workgroup owns one output tile, each thread holding
// A = 2^(L-9) amplitudes l = t + a*512 as FP64 accumulators in registers.
// Group g reads source tile (tile ^ (x_g >> L)) permuted by x_g & (2^L-1).
// Groups are sorted by that high part ("class"): the source tile of a class
// is streamed from HBM into LDS once and every group of the class gathers
// from LDS, so HBM reads per apply are 2^n * (classes) amplitudes instead of
// 2^n * (groups).  The next class's tile is prefetched into registers while
// the current one is consumed.  Term masks and coefficients are wave-uniform
// and are read through scalar loads.
//
// Per term only the thread-part parity popc(z & t) is per-lane; the tile-part
// parity popc((z >> L) & tile) is scalar, and the slot part (bits 9..L-1,
// amplitude a of the thread) is folded by sorting each group's terms by those
// z bits and applying a length-A Walsh-Hadamard transform once per group.
#include <algorithm>
#include <memory>
#include <mutex>

#include "tcp_internal.hpp"

namespace {

constexpr uint32_t kThreads = 512;  // threads per workgroup (fixed stride of slots)
constexpr uint32_t kThreadBits = 9;
constexpr uint32_t kMaxSlotBits = 3;  // up to 8 amplitudes per thread

struct MvpArgs {
    const uint64_t *cls_hi;    // [H]   x >> L of each class
    const uint32_t *cls_goff;  // [H+1] group range of each class
    const uint32_t *g_lo;      // [G]   x & (2^L - 1)
    const uint32_t *g_off;     // [G*A + 1] term start per (group, slot class)
    const uint32_t *t_zlo;     // [T]   z & (2^tb - 1)
    const uint64_t *t_zhi;     // [T]   z >> L
    const void *t_w;           // [T]   (re, im) of w * (-1)^popc(z & x)
    uint32_t H;
    uint32_t L;
    uint32_t active;  // threads owning an amplitude: min(512, 2^n)
    uint64_t dim;     // 2^n, the stride between right-hand sides
};

// Plan tables are read through the constant address space: every index into
// them is wave-uniform, and constant-space loads become scalar loads (s_load)
// whose completion is tracked apart from the in-flight tile prefetch.
template <class T>
using cptr = const __attribute__((address_space(4))) T *;

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

template <typename VT>
struct Amp;
template <>
struct Amp<dvec2> {
    __device__ static double re(const dvec2 &v) { return v.x; }
    __device__ static double im(const dvec2 &v) { return v.y; }
    __device__ static dvec2 make(double r, double i) { return dvec2{r, i}; }
};
template <>
struct Amp<fvec2> {
    __device__ static double re(const fvec2 &v) { return static_cast<double>(v.x); }
    __device__ static double im(const fvec2 &v) { return static_cast<double>(v.y); }
    __device__ static fvec2 make(double r, double i) {
        return fvec2{static_cast<float>(r), static_cast<float>(i)};
    }
};

__device__ __forceinline__ double flip_sign(double x, uint32_t parity) {
    return __longlong_as_double(__double_as_longlong(x) ^ (static_cast<long long>(parity) << 63));
}

// PARTIAL: fewer than 512 amplitudes (n < 9); threads t >= 2^n idle.
template <int A, typename VT, bool PARTIAL>
__global__ __launch_bounds__(kThreads) void mvp_tile_kernel(const VT *__restrict__ v,
                                                            VT *__restrict__ y, MvpArgs p) {
    __shared__ VT tile[kThreads * A];
    cptr<uint64_t> cls_hi = (cptr<uint64_t>)p.cls_hi;
    cptr<uint32_t> cls_goff = (cptr<uint32_t>)p.cls_goff;
    cptr<uint32_t> g_lo = (cptr<uint32_t>)p.g_lo;
    cptr<uint32_t> g_off = (cptr<uint32_t>)p.g_off;
    cptr<uint32_t> t_zlo = (cptr<uint32_t>)p.t_zlo;
    cptr<uint64_t> t_zhi = (cptr<uint64_t>)p.t_zhi;
    cptr<dvec2> t_w = (cptr<dvec2>)p.t_w;

    const uint32_t t = threadIdx.x;
    const uint64_t tile_id = blockIdx.x;
    const uint64_t base = tile_id << p.L;
    const uint64_t rhs = static_cast<uint64_t>(blockIdx.y) * p.dim;
    const VT *__restrict__ vin = v + rhs + t;
    const bool on = !PARTIAL || t < p.active;

    double accr[A], acci[A];
    VT pre[A];
#pragma unroll
    for (int a = 0; a < A; ++a) {
        accr[a] = 0.0;
        acci[a] = 0.0;
    }
    {
        const uint64_t src = base ^ (cls_hi[0] << p.L);
        if (on) {
#pragma unroll
            for (int a = 0; a < A; ++a) pre[a] = vin[src + a * kThreads];
        }
    }
    for (uint32_t h = 0; h < p.H; ++h) {
        __syncthreads();  // every wave is done reading the previous class's tile
        if (on) {
#pragma unroll
            for (int a = 0; a < A; ++a) tile[t + a * kThreads] = pre[a];
        }
        __syncthreads();
        const uint32_t hn = h + 1 < p.H ? h + 1 : h;
        const uint64_t src = base ^ (cls_hi[hn] << p.L);
        if (h + 1 < p.H && on) {  // prefetch the next class's source tile
#pragma unroll
            for (int a = 0; a < A; ++a) pre[a] = vin[src + a * kThreads];
        }
        const uint32_t g0 = cls_goff[h], g1 = cls_goff[h + 1];
        for (uint32_t g = g0; g < g1; ++g) {
            const uint32_t lo = g_lo[g];
            double sr[A], si[A];
#pragma unroll
            for (int c = 0; c < A; ++c) {
                sr[c] = 0.0;
                si[c] = 0.0;
                const uint32_t k0 = g_off[g * A + c], k1 = g_off[g * A + c + 1];
                for (uint32_t k = k0; k < k1; ++k) {
                    const uint32_t tile_par = __popcll(t_zhi[k] & tile_id);
                    const uint32_t par = (__popc(t_zlo[k] & t) + tile_par) & 1u;
                    const dvec2 w = t_w[k];
                    sr[c] += flip_sign(w.x, par);
                    si[c] += flip_sign(w.y, par);
                }
            }
            // d[a] = sum_c (-1)^popc(c & a) S[c]
#pragma unroll
            for (int m = 1; m < A; m <<= 1) {
#pragma unroll
                for (int c = 0; c < A; ++c) {
                    if (!(c & m)) {
                        const double ur = sr[c], ui = si[c], vr = sr[c | m], vi = si[c | m];
                        sr[c] = ur + vr;
                        si[c] = ui + vi;
                        sr[c | m] = ur - vr;
                        si[c | m] = ui - vi;
                    }
                }
            }
#pragma unroll
            for (int a = 0; a < A; ++a) {
                const VT s = tile[(t + a * kThreads) ^ lo];
                const double s_r = Amp<VT>::re(s), s_i = Amp<VT>::im(s);
                accr[a] = fma(sr[a], s_r, accr[a]);
                accr[a] = fma(-si[a], s_i, accr[a]);
                acci[a] = fma(sr[a], s_i, acci[a]);
                acci[a] = fma(si[a], s_r, acci[a]);
            }
        }
    }
    if (on) {
        VT *__restrict__ out = y + rhs + base + t;
#pragma unroll
        for (int a = 0; a < A; ++a) out[a * kThreads] = Amp<VT>::make(accr[a], acci[a]);
    }
}

// Fill kernel used instead of hipMemsetAsync (safe under graph capture).
template <typename VT>
__global__ void zero_kernel(VT *__restrict__ y, uint64_t count) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        y[i] = Amp<VT>::make(0.0, 0.0);
}

// Eager transition count: number of (row, group) pairs whose diagonal sum is
// nonzero (MvpPlan::transition_count, core/hamiltonian.rs:165-180; the
// destination-form diagonal is a permutation of the source-form one).
__global__ void mvp_count_kernel(MvpArgs p, uint32_t A, uint32_t G, uint64_t dim,
                                 unsigned long long *count) {
    unsigned long long local = 0;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < dim;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t tile_id = i >> p.L;
        const uint32_t l = static_cast<uint32_t>(i & ((1ull << p.L) - 1));
        const uint32_t t = l & (kThreads - 1);
        const uint32_t slot = l >> kThreadBits;
        for (uint32_t g = 0; g < G; ++g) {
            double dr = 0.0, di = 0.0;
            for (uint32_t c = 0; c < A; ++c) {
                const uint32_t cpar = __popc(c & slot);
                for (uint32_t k = p.g_off[g * A + c]; k < p.g_off[g * A + c + 1]; ++k) {
                    const uint32_t par =
                        (__popc(p.t_zlo[k] & t) + __popcll(p.t_zhi[k] & tile_id) + cpar) & 1u;
                    const double2 w = reinterpret_cast<const double2 *>(p.t_w)[k];
                    dr += flip_sign(w.x, par);
                    di += flip_sign(w.y, par);
                }
            }
            local += (dr != 0.0 || di != 0.0);
        }
    }
    atomicAdd(count, local);
}

// Real part of the diagonal H_ii: the x = 0 group (group 0 when present).
__global__ void mvp_diagonal_kernel(MvpArgs p, uint32_t A, uint32_t has_diag, uint64_t dim,
                                    double *__restrict__ diag) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < dim;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t tile_id = i >> p.L;
        const uint32_t l = static_cast<uint32_t>(i & ((1ull << p.L) - 1));
        const uint32_t t = l & (kThreads - 1);
        const uint32_t slot = l >> kThreadBits;
        double d = 0.0;
        if (has_diag) {
            for (uint32_t c = 0; c < A; ++c) {
                const uint32_t cpar = __popc(c & slot);
                for (uint32_t k = p.g_off[c]; k < p.g_off[c + 1]; ++k) {
                    const uint32_t par =
                        (__popc(p.t_zlo[k] & t) + __popcll(p.t_zhi[k] & tile_id) + cpar) & 1u;
                    d += flip_sign(reinterpret_cast<const double2 *>(p.t_w)[k].x, par);
                }
            }
        }
        diag[i] = d;
    }
}

}  // namespace

struct tcp_mvp_plan {
    uint32_t n = 0;
    uint64_t T = 0, G = 0, H = 0;
    int storage = TCP_STORAGE_LAZY;
    uint64_t transitions = UINT64_MAX;
    uint32_t tb = 0, ab = 0, L = 0;
    int device = 0;
    tcp::DeviceBuffer cls_hi, cls_goff, g_lo, g_off, t_zlo, t_zhi, t_w64, t_w32;
    uint64_t retained_bytes = 0;
    bool has_diagonal = false;  // group 0 has x = 0
    std::mutex staging_mutex;
    tcp::DeviceBuffer stage_in, stage_out;

    uint64_t dim() const { return 1ull << n; }
    MvpArgs args(int dtype) const {
        MvpArgs a;
        a.cls_hi = static_cast<const uint64_t *>(cls_hi.ptr);
        a.cls_goff = static_cast<const uint32_t *>(cls_goff.ptr);
        a.g_lo = static_cast<const uint32_t *>(g_lo.ptr);
        a.g_off = static_cast<const uint32_t *>(g_off.ptr);
        a.t_zlo = static_cast<const uint32_t *>(t_zlo.ptr);
        a.t_zhi = static_cast<const uint64_t *>(t_zhi.ptr);
        a.t_w = static_cast<const double2 *>(dtype == TCP_C64 ? t_w32.ptr : t_w64.ptr);
        a.H = static_cast<uint32_t>(H);
        a.L = L;
        a.active = n >= kThreadBits ? kThreads : (1u << n);
        a.dim = dim();
        return a;
    }
};

namespace {

using tcp::fail;

int build_plan(const tcp::Operator &op, int storage, uint64_t max_bytes, tcp_mvp_plan **out) {
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (storage != TCP_STORAGE_LAZY && storage != TCP_STORAGE_EAGER)
        return fail(TCP_ERR_INVALID_ARGUMENT, "storage must be either 'eager' or 'lazy'");
    const uint32_t n = op.nqubits;
    if (n >= 64) return fail(TCP_ERR_OVERFLOW, "integer overflow while computing matrix dimension");
    const uint64_t T = op.term_count();
    if (T >= (1ull << 31)) return fail(TCP_ERR_UNSUPPORTED, "more than 2^31 terms");

    auto plan = new tcp_mvp_plan();
    std::unique_ptr<tcp_mvp_plan> guard(plan);
    plan->n = n;
    plan->T = T;
    plan->storage = storage;
    plan->tb = std::min(n, kThreadBits);
    plan->ab = n > kThreadBits ? std::min(n - kThreadBits, kMaxSlotBits) : 0;
    plan->L = plan->tb + plan->ab;
    const uint32_t L = plan->L, A = 1u << plan->ab;
    const uint64_t lmask = (1ull << L) - 1;

    std::vector<tcp::IndexTerm> terms = tcp::index_terms(op);
    struct Key {
        uint64_t hi, lo;
        uint32_t cls;
        uint64_t index;
    };
    std::vector<Key> keys(T);
    for (uint64_t k = 0; k < T; ++k)
        keys[k] = Key{terms[k].x >> L, terms[k].x & lmask,
                      static_cast<uint32_t>((terms[k].z >> plan->tb) & (A - 1)), k};
    std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
        if (a.hi != b.hi) return a.hi < b.hi;
        if (a.lo != b.lo) return a.lo < b.lo;
        if (a.cls != b.cls) return a.cls < b.cls;
        return a.index < b.index;
    });

    std::vector<uint64_t> cls_hi;
    std::vector<uint32_t> cls_goff, g_lo, g_off, t_zlo;
    std::vector<uint64_t> t_zhi;
    std::vector<double2> t_w64, t_w32;
    t_zlo.reserve(T);
    t_zhi.reserve(T);
    t_w64.reserve(T);
    t_w32.reserve(T);
    uint64_t k = 0;
    while (k < T) {
        if (cls_hi.empty() || cls_hi.back() != keys[k].hi) {
            cls_hi.push_back(keys[k].hi);
            cls_goff.push_back(static_cast<uint32_t>(g_lo.size()));
        }
        uint64_t e = k;
        while (e < T && keys[e].hi == keys[k].hi && keys[e].lo == keys[k].lo) ++e;
        g_lo.push_back(static_cast<uint32_t>(keys[k].lo));
        uint64_t j = k;
        for (uint32_t c = 0; c < A; ++c) {
            g_off.push_back(static_cast<uint32_t>(j));
            while (j < e && keys[j].cls == c) ++j;
        }
        for (uint64_t q = k; q < e; ++q) {
            const tcp::IndexTerm &term = terms[keys[q].index];
            const bool neg = __builtin_popcountll(term.z & term.x) & 1;
            const double wr = neg ? -term.w.re : term.w.re;
            const double wi = neg ? -term.w.im : term.w.im;
            t_zlo.push_back(static_cast<uint32_t>(term.z & ((1ull << plan->tb) - 1)));
            t_zhi.push_back(term.z >> L);
            t_w64.push_back(make_double2(wr, wi));
            t_w32.push_back(make_double2(static_cast<double>(static_cast<float>(wr)),
                                         static_cast<double>(static_cast<float>(wi))));
        }
        k = e;
    }
    g_off.push_back(static_cast<uint32_t>(T));
    cls_goff.push_back(static_cast<uint32_t>(g_lo.size()));
    plan->G = g_lo.size();
    plan->H = cls_hi.size();
    plan->has_diagonal = !cls_hi.empty() && cls_hi[0] == 0 && g_lo[0] == 0;
    if (cls_hi.empty()) cls_hi.push_back(0);

    uint64_t device_bytes = cls_hi.size() * 8 + cls_goff.size() * 4 + g_lo.size() * 4 +
                            g_off.size() * 4 + T * (4 + 8 + 16 + 16);
    plan->retained_bytes = device_bytes + sizeof(tcp_mvp_plan);
    if (tcp::exceeds_budget(plan->retained_bytes, max_bytes))
        return tcp::memory_limit(plan->retained_bytes, max_bytes);

    TCP_HIP_CHECK(hipGetDevice(&plan->device));
    int st;
    if ((st = plan->cls_hi.upload(cls_hi.data(), cls_hi.size() * 8))) return st;
    if ((st = plan->cls_goff.upload(cls_goff.data(), cls_goff.size() * 4))) return st;
    if ((st = plan->g_lo.upload(g_lo.data(), g_lo.size() * 4))) return st;
    if ((st = plan->g_off.upload(g_off.data(), g_off.size() * 4))) return st;
    if ((st = plan->t_zlo.upload(t_zlo.data(), t_zlo.size() * 4))) return st;
    if ((st = plan->t_zhi.upload(t_zhi.data(), t_zhi.size() * 8))) return st;
    if ((st = plan->t_w64.upload(t_w64.data(), t_w64.size() * 16))) return st;
    if ((st = plan->t_w32.upload(t_w32.data(), t_w32.size() * 16))) return st;

    if (storage == TCP_STORAGE_EAGER) {
        unsigned long long *dcount = nullptr;
        TCP_HIP_CHECK(hipMalloc(&dcount, sizeof(unsigned long long)));
        TCP_HIP_CHECK(hipMemset(dcount, 0, sizeof(unsigned long long)));
        const uint64_t dim = plan->dim();
        const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((dim + 255) / 256, 4096));
        hipLaunchKernelGGL(mvp_count_kernel, dim3(blocks), dim3(256), 0, nullptr, plan->args(TCP_C128),
                           A, static_cast<uint32_t>(plan->G), dim, dcount);
        hipError_t e = hipGetLastError();
        unsigned long long hcount = 0;
        if (e == hipSuccess) e = hipMemcpy(&hcount, dcount, sizeof hcount, hipMemcpyDeviceToHost);
        (void)hipFree(dcount);
        if (e != hipSuccess) return tcp::hip_fail(e, "eager transition count");
        plan->transitions = hcount;
    }
    *out = guard.release();
    return tcp::ok();
}

template <int A, typename VT, bool PARTIAL = false>
hipError_t launch_tile(const tcp_mvp_plan *plan, const void *in, void *out, uint32_t nrhs,
                       int dtype, hipStream_t stream) {
    MvpArgs args = plan->args(dtype);
    const uint64_t tiles = plan->dim() >> plan->L;
    hipLaunchKernelGGL((mvp_tile_kernel<A, VT, PARTIAL>), dim3(static_cast<uint32_t>(tiles), nrhs),
                       dim3(kThreads), 0, stream, static_cast<const VT *>(in),
                       static_cast<VT *>(out), args);
    return hipGetLastError();
}

template <typename VT>
hipError_t launch_typed(const tcp_mvp_plan *plan, const void *in, void *out, uint32_t nrhs,
                        int dtype, hipStream_t stream) {
    if (plan->H == 0) {  // empty operator: H v = 0
        const uint64_t count = plan->dim() * nrhs;
        const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((count + 255) / 256, 65536));
        hipLaunchKernelGGL(zero_kernel<VT>, dim3(blocks), dim3(256), 0, stream, static_cast<VT *>(out),
                           count);
        return hipGetLastError();
    }
    if (plan->n < kThreadBits) return launch_tile<1, VT, true>(plan, in, out, nrhs, dtype, stream);
    switch (plan->ab) {
        case 0: return launch_tile<1, VT>(plan, in, out, nrhs, dtype, stream);
        case 1: return launch_tile<2, VT>(plan, in, out, nrhs, dtype, stream);
        case 2: return launch_tile<4, VT>(plan, in, out, nrhs, dtype, stream);
        default: return launch_tile<8, VT>(plan, in, out, nrhs, dtype, stream);
    }
}

int apply_device(const tcp_mvp_plan *plan, const void *in, void *out, uint64_t nrhs, int dtype,
                 void *stream) {
    if (dtype != TCP_C128 && dtype != TCP_C64)
        return fail(TCP_ERR_INVALID_ARGUMENT, "dtype must be TCP_C128 or TCP_C64");
    if (!in || !out) return fail(TCP_ERR_INVALID_ARGUMENT, "null state buffer");
    if (nrhs == 0) return tcp::ok();
    if (nrhs > 65535) return fail(TCP_ERR_UNSUPPORTED, "at most 65535 right-hand sides per call");
    const size_t amp = dtype == TCP_C128 ? 16 : 8;
    const size_t bytes = static_cast<size_t>(plan->dim()) * amp * nrhs;
    if (tcp::ranges_overlap(in, bytes, out, bytes))
        return fail(TCP_ERR_INVALID_ARGUMENT, "apply_into input and output must not overlap");
    if ((plan->dim() >> plan->L) > 0x7fffffffull)
        return fail(TCP_ERR_UNSUPPORTED, "too many tiles for one launch");
    tcp::DeviceGuard dg;
    int st = dg.enter(plan->device);
    if (st) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = dtype == TCP_C128
                       ? launch_typed<dvec2>(plan, in, out, static_cast<uint32_t>(nrhs), dtype, s)
                       : launch_typed<fvec2>(plan, in, out, static_cast<uint32_t>(nrhs), dtype, s);
    if (e != hipSuccess) return tcp::hip_fail(e, "mvp kernel launch");
    return tcp::ok();
}

}  // namespace

namespace tcp {
uint64_t plan_dimension(const tcp_mvp_plan *plan) { return plan->dim(); }
int plan_device(const tcp_mvp_plan *plan) { return plan->device; }
int plan_apply(const tcp_mvp_plan *plan, const void *in, void *out, int dtype, hipStream_t stream) {
    return apply_device(plan, in, out, 1, dtype, stream);
}
int plan_diagonal(const tcp_mvp_plan *plan, double *diag, hipStream_t stream) {
    const uint64_t dim = plan->dim();
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((dim + 255) / 256, 65536));
    hipLaunchKernelGGL(mvp_diagonal_kernel, dim3(blocks), dim3(256), 0, stream, plan->args(TCP_C128),
                       1u << plan->ab, plan->has_diagonal ? 1u : 0u, dim, diag);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "diagonal kernel");
}
}  // namespace tcp

namespace {

int check_len(uint64_t expected, uint64_t actual) {
    if (expected != actual)
        return fail(TCP_ERR_INVALID_STRUCTURE_LENGTH, "expected structure length " +
                                                          std::to_string(expected) + ", got " +
                                                          std::to_string(actual));
    return TCP_OK;
}

}  // namespace

extern "C" {

int tcp_mvp_plan_create(const tcp_operator *op, int storage, uint64_t max_bytes,
                        tcp_mvp_plan **out) {
    if (!op) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    return build_plan(tcp::operator_of(op), storage, max_bytes, out);
}

int tcp_mvp_plan_create_from_words(uint32_t nqubits, uint64_t nterms, const uint64_t *x_words,
                                   const uint64_t *z_words, const double *coef, int storage,
                                   uint64_t max_bytes, tcp_mvp_plan **out) {
    tcp_operator *op = nullptr;
    int st = tcp_operator_from_words(nqubits, nterms, x_words, z_words, coef, &op);
    if (st) return st;
    st = build_plan(tcp::operator_of(op), storage, max_bytes, out);
    tcp_operator_destroy(op);
    return st;
}

uint32_t tcp_mvp_plan_nqubits(const tcp_mvp_plan *plan) { return plan ? plan->n : 0; }
uint64_t tcp_mvp_plan_term_count(const tcp_mvp_plan *plan) { return plan ? plan->T : 0; }
uint64_t tcp_mvp_plan_group_count(const tcp_mvp_plan *plan) { return plan ? plan->G : 0; }
uint64_t tcp_mvp_plan_class_count(const tcp_mvp_plan *plan) { return plan ? plan->H : 0; }
uint64_t tcp_mvp_plan_dimension(const tcp_mvp_plan *plan) { return plan ? plan->dim() : 0; }
int tcp_mvp_plan_storage(const tcp_mvp_plan *plan) { return plan ? plan->storage : -1; }
uint64_t tcp_mvp_plan_transition_count(const tcp_mvp_plan *plan) {
    return plan ? plan->transitions : UINT64_MAX;
}
uint64_t tcp_mvp_plan_estimated_bytes(const tcp_mvp_plan *plan) { return plan ? plan->retained_bytes : 0; }

uint64_t tcp_mvp_plan_bytes_alg(const tcp_mvp_plan *plan, int dtype) {
    if (!plan) return 0;
    const uint64_t s = dtype == TCP_C64 ? 8 : 16;
    const uint64_t W = tcp::word_count(plan->n);
    return plan->dim() * (plan->G * s + s) + plan->T * (16 * W + s);
}

int tcp_mvp_plan_apply_into(const tcp_mvp_plan *plan, const void *input, void *output,
                            uint64_t input_len, uint64_t output_len, int dtype, void *stream) {
    if (!plan) return fail(TCP_ERR_INVALID_ARGUMENT, "null plan");
    int st;
    if ((st = check_len(plan->dim(), input_len))) return st;
    if ((st = check_len(plan->dim(), output_len))) return st;
    return apply_device(plan, input, output, 1, dtype, stream);
}

int tcp_mvp_plan_apply_batch(const tcp_mvp_plan *plan, const void *input, void *output,
                             uint64_t nrhs, int dtype, void *stream) {
    if (!plan) return fail(TCP_ERR_INVALID_ARGUMENT, "null plan");
    return apply_device(plan, input, output, nrhs, dtype, stream);
}

int tcp_mvp_plan_apply_host(const tcp_mvp_plan *plan, const double *input, double *output,
                            uint64_t input_len, uint64_t output_len) {
    if (!plan) return fail(TCP_ERR_INVALID_ARGUMENT, "null plan");
    int st;
    if ((st = check_len(plan->dim(), input_len))) return st;
    if ((st = check_len(plan->dim(), output_len))) return st;
    if (!input || !output) return fail(TCP_ERR_INVALID_ARGUMENT, "null state buffer");
    const size_t bytes = static_cast<size_t>(plan->dim()) * 16;
    if (tcp::ranges_overlap(input, bytes, output, bytes))
        return fail(TCP_ERR_INVALID_ARGUMENT, "apply_into input and output must not overlap");
    tcp::DeviceGuard dg;
    if ((st = dg.enter(plan->device))) return st;
    auto *mp = const_cast<tcp_mvp_plan *>(plan);
    std::lock_guard<std::mutex> lock(mp->staging_mutex);
    if (mp->stage_in.bytes < bytes && (st = mp->stage_in.allocate(bytes))) return st;
    if (mp->stage_out.bytes < bytes && (st = mp->stage_out.allocate(bytes))) return st;
    TCP_HIP_CHECK(hipMemcpy(mp->stage_in.ptr, input, bytes, hipMemcpyHostToDevice));
    if ((st = apply_device(plan, mp->stage_in.ptr, mp->stage_out.ptr, 1, TCP_C128, nullptr))) return st;
    TCP_HIP_CHECK(hipMemcpy(output, mp->stage_out.ptr, bytes, hipMemcpyDeviceToHost));
    return tcp::ok();
}

void tcp_mvp_plan_destroy(tcp_mvp_plan *plan) { delete plan; }

int tcp_mvp(const tcp_operator *op, const double *input, double *output, uint64_t input_len,
            uint64_t output_len) {
    tcp_mvp_plan *plan = nullptr;
    int st = tcp_mvp_plan_create(op, TCP_STORAGE_LAZY, 0, &plan);
    if (st) return st;
    st = tcp_mvp_plan_apply_host(plan, input, output, input_len, output_len);
    tcp_mvp_plan_destroy(plan);
    return st;
}

}  // extern "C"
