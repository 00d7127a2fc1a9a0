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

#include "tile_tables.hpp"

namespace {

constexpr uint32_t kThreads = 512;  // threads per workgroup (fixed stride of slots)
constexpr uint32_t kThreadBits = 9;
constexpr uint32_t kMaxSlotBits = 3;  // up to 8 amplitudes per thread

using tcp::Amp;
using tcp::cptr;
using tcp::dvec2;
using tcp::flip_sign;
using tcp::fvec2;
using MvpArgs = tcp::TileArgs;

// One workgroup per output tile (blockIdx.x) and right-hand side (blockIdx.y).
// PARTIAL: fewer than 512 amplitudes (n < 9); threads t >= 2^n idle.
template <int A, typename VT, bool PARTIAL>
__global__ __launch_bounds__(kThreads) void mvp_tile_kernel(const VT *__restrict__ v, VT *__restrict__ y,
                                                            MvpArgs p, uint32_t active, uint64_t dim) {
    __shared__ VT tile[kThreads * A];
    const uint64_t rhs = static_cast<uint64_t>(blockIdx.y) * dim;
    tcp::tile_apply<A, VT, PARTIAL, kThreads>(v + rhs, y + rhs, p, blockIdx.x, active, tile);
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
        a.t_w = dtype == TCP_C64 ? t_w32.ptr : t_w64.ptr;
        a.H = static_cast<uint32_t>(H);
        a.L = L;
        return a;
    }
    uint32_t active() const { return n >= kThreadBits ? kThreads : (1u << n); }
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
    const uint32_t A = 1u << plan->ab;
    tcp::TileTables tab;
    tab.append(op, plan->tb, plan->ab);
    plan->G = tab.G;
    plan->H = tab.H;
    plan->has_diagonal = tab.has_diagonal;
    std::vector<uint64_t> &cls_hi = tab.cls_hi;
    std::vector<uint32_t> &cls_goff = tab.cls_goff, &g_lo = tab.g_lo, &g_off = tab.g_off, &t_zlo = tab.t_zlo;
    std::vector<uint64_t> &t_zhi = tab.t_zhi;
    std::vector<double> &t_w64 = tab.t_w64, &t_w32 = tab.t_w32;
    if (cls_hi.empty()) cls_hi.push_back(0);
    if (g_lo.empty()) g_lo.push_back(0);
    if (t_zlo.empty()) {
        t_zlo.push_back(0);
        t_zhi.push_back(0);
        t_w64.assign(2, 0.0);
        t_w32.assign(2, 0.0);
    }

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
    if ((st = plan->t_w64.upload(t_w64.data(), t_w64.size() * 8))) return st;
    if ((st = plan->t_w32.upload(t_w32.data(), t_w32.size() * 8))) return st;

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
                       static_cast<VT *>(out), args, plan->active(), plan->dim());
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
