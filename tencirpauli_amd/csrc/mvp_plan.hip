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
// Layout.  The index space is cut into tiles of 2^L amplitudes: a workgroup
// of TPB threads owns one output tile, each thread holding A = 2^L / TPB
// amplitudes l = t + a*TPB as FP64 accumulators in registers (geometries
// 512x8 with L = 12, 1024x8 and 512x16 with L = 13; smaller spaces use 512
// threads and 1-4 amplitudes), or two output tiles (twin geometries,
// tile_tables.hpp twin_apply).  Group g reads source tile (tile ^ (x_g >> L))
// permuted by x_g & (2^L-1).  Groups are sorted by that high part ("class"):
// the source tile of a class is streamed from HBM into LDS once (widened to
// FP64) and every group of the class gathers from LDS, so HBM reads per apply
// are 2^n * (classes) amplitudes instead of 2^n * (groups).  The next class's
// tile is prefetched into registers while the current one is consumed.
// A class's passes (its groups split by slot pattern) and their terms travel
// as a class block staged in LDS beside the tile; tile_tables.hpp describes
// the blocks, the passes and their kinds.
#include <algorithm>
#include <memory>
#include <mutex>

#include "tile_tables.hpp"

namespace {

using tcp::Amp;
using tcp::dvec2;
using tcp::flip_sign;
using tcp::fvec2;
using MvpArgs = tcp::TileArgs;

// Tile geometries: threads per workgroup x amplitudes per thread, and
// whether a workgroup owns twin tiles (two output tiles sharing each staged
// source half, tile_tables.hpp twin_apply).
struct Geometry {
    uint32_t threads, amps, twin;
};
constexpr Geometry kGeometries[] = {{0, 0, 0},   {512, 8, 0},  {1024, 8, 0}, {512, 16, 0},
                                    {512, 4, 0}, {1024, 4, 0}, {1024, 8, 1}, {512, 16, 1}};
constexpr uint32_t kGeometryCount = sizeof(kGeometries) / sizeof(kGeometries[0]);
constexpr uint32_t kSmallThreads = 512;
// Host-pointer applies stage through two device buffers; up to this size they
// stay cached on the plan between calls (and count in its retained bytes),
// larger ones are freed after each call.
constexpr size_t kStagingKeepBytes = size_t(64) << 20;  // below 2^12 amplitudes: 512 threads, 1-4 amplitudes each
constexpr uint32_t kSmallBits = 9;

uint32_t log2u(uint32_t x) { return 31u - static_cast<uint32_t>(__builtin_clz(x)); }

constexpr uint32_t kNoXcdOrder = ~0u;

// Source tiles in flight per workgroup encoded in a tiling word (field 0 = one).
constexpr uint32_t tile_prefetch_depth(uint32_t tiling) {
    return ((tiling & TCP_TILE_PREFETCH_MASK) >> TCP_TILE_PREFETCH_SHIFT) + 1u;
}

// One workgroup per output tile and right-hand side (blockIdx.y).
// Workgroups are dealt to the eight XCDs round-robin (XCD = b mod 8).  With
// an XCD order (xcd_shift != kNoXcdOrder) tile bits xcd_shift .. +2 name the
// XCD and j = b / 8 fills the others: the workgroups one XCD runs together
// differ in the low xcd_shift tile bits (the bit map's window, so their
// source tiles coincide class after class, shared in that XCD's L2), and at
// one moment the eight XCDs work on tiles that differ in the bit map's XCD
// bits, so they share source tiles through the last-level cache.  With
// xcd_shift = log2(tiles) - 3 each XCD walks one contiguous eighth.
// PARTIAL: the whole space is smaller than one tile; l >= dim idle.
// Waves per SIMD the register budget is sized for: as many workgroups per CU
// as their LDS (the FP64 tile, 16 B per amplitude, and the class block, 8 B
// per thread) fits in the CU's 160 KiB, spread over the 4 SIMDs, at most 8
// (512 x 8: 4 waves, <= 128 VGPRs; 512 x 4 and 1024 x 4: 8 waves, <= 64
// VGPRs, 6 for complex128 storage, whose four FP64 accumulators and
// prefetched amplitudes do not fit in 64; 512 x 16: 2 waves).
template <int A, uint32_t TPB, typename VT>
constexpr int waves_per_simd() {
    constexpr int wgs = (160 * 1024) / ((2 * A + 1) * TPB * 8 + 32);
    constexpr int w = wgs * static_cast<int>(TPB / 64) / 4;
    constexpr int cap = (sizeof(VT) == 16 && A == 4) ? 6 : 8;
    return w < 1 ? 1 : (w > cap ? cap : w);
}

// PF: source tiles in flight (register ring depth); NT: non-temporal tile loads.
template <int A, typename VT, uint32_t TPB, bool PARTIAL, int PF = 1, bool NT = false>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(waves_per_simd<A, TPB, VT>())))
void mvp_tile_kernel(const VT *__restrict__ v, VT *__restrict__ y, MvpArgs p,
                                                       uint64_t dim, uint32_t xcd_shift) {
    __shared__ dvec2 lds[TPB * A];
    __shared__ tcp::u32x4 lblk[tcp::block_words(TPB) + 2];
    const uint64_t rhs = static_cast<uint64_t>(blockIdx.y) * dim;
    uint64_t tile = blockIdx.x;
    if (xcd_shift != kNoXcdOrder) {
        const uint64_t j = tile >> 3;
        tile = (j & ((1ull << xcd_shift) - 1)) | ((tile & 7u) << xcd_shift) | ((j >> xcd_shift) << (xcd_shift + 3));
    }
    const uint32_t active = PARTIAL ? static_cast<uint32_t>(dim) : TPB * A;
    tcp::tile_apply<A, VT, TPB, PARTIAL, PF, NT, !PARTIAL>(v + rhs, y + rhs, p, tile, active, lds, lblk);
}

// Twin-tile kernel: workgroup b owns output tiles 2*pair and 2*pair + 1, with
// pair = b remapped by the XCD order as in mvp_tile_kernel.
template <int A, typename VT, uint32_t TPB, int CR = 0>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(waves_per_simd<A, TPB, VT>())))
void mvp_twin_kernel(const VT *__restrict__ v, VT *__restrict__ y, MvpArgs p, uint64_t dim, uint32_t xcd_shift) {
    __shared__ dvec2 lds[TPB * A];
    __shared__ tcp::u32x4 lblk[tcp::block_words(TPB) + 2];
    const uint64_t rhs = static_cast<uint64_t>(blockIdx.y) * dim;
    uint64_t pair = blockIdx.x;
    if (xcd_shift != kNoXcdOrder) {
        const uint64_t j = pair >> 3;
        pair = (j & ((1ull << xcd_shift) - 1)) | ((pair & 7u) << xcd_shift) | ((j >> xcd_shift) << (xcd_shift + 3));
    }
    tcp::twin_apply<A, VT, TPB, CR>(v + rhs, y + rhs, p, pair, lds, lblk);
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
__global__ void mvp_count_kernel(MvpArgs p, uint32_t tb, uint32_t passes, uint64_t dim,
                                 unsigned long long *count) {
    unsigned long long local = 0;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < dim;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t k = tcp::kernel_index(p, i);
        const uint64_t tile_id = k >> p.L;
        const uint32_t l = static_cast<uint32_t>(k & ((1ull << p.L) - 1));
        const uint32_t t = l & ((1u << tb) - 1);
        const uint32_t slot = l >> tb;
        // in x-mask order the passes of one x-mask are consecutive (`same`
        // marks the continuation)
        dvec2 d = dvec2{0.0, 0.0};
        for (uint32_t i = 0; i < passes; ++i) {
            const uint32_t g = p.xorder[i];
            if (i > 0 && !p.grp[g].same) local += (d.x != 0.0 || d.y != 0.0);
            if (i == 0 || !p.grp[g].same) d = dvec2{0.0, 0.0};
            d += tcp::group_sum(p, g, t, slot, tile_id);
        }
        if (passes > 0) local += (d.x != 0.0 || d.y != 0.0);
    }
    atomicAdd(count, local);
}

// Real part of the diagonal H_ii: the passes of the x = 0 group (the first
// `diag_passes` passes in x-mask order).
__global__ void mvp_diagonal_kernel(MvpArgs p, uint32_t tb, uint32_t diag_passes, uint64_t dim,
                                    double *__restrict__ diag) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < dim;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t k = tcp::kernel_index(p, i);
        const uint64_t tile_id = k >> p.L;
        const uint32_t l = static_cast<uint32_t>(k & ((1ull << p.L) - 1));
        const uint32_t t = l & ((1u << tb) - 1);
        double d = 0.0;
        for (uint32_t i = 0; i < diag_passes; ++i) d += tcp::group_sum(p, p.xorder[i], t, l >> tb, tile_id).x;
        diag[i] = d;
    }
}

}  // namespace

// Device tables of one tile geometry.
struct PlanTiles;
uint32_t tile_xcd_shift(const PlanTiles &pt, uint64_t dim);

struct PlanTiles {
    uint32_t tiling = 0;  // geometry index (0: small space) | TCP_TILE_XCD_ORDER
    uint32_t tb = 0, ab = 0, L = 0;
    uint32_t twin = 0;    // 1: twin tiles (workgroups own 2^(L+1) amplitudes)
    uint64_t H = 0, passes = 0;
    uint32_t diag_passes = 0;  // passes with x = 0
    uint32_t n = 0;
    bool permuted = false;     // kernel coordinates differ from the index bits
    uint8_t pos[64] = {};
    uint32_t xcd_bits_at = kNoXcdOrder;  // tile bit of the bit map's XCD bits (kNoXcdOrder: none chosen)
    tcp::DeviceBuffer cls, blk64, blk32, xorder, grp64, grp32, term64, term32;
    uint64_t bytes = 0;

    MvpArgs args(int dtype) const {
        MvpArgs a;
        a.n = n;
        a.permuted = permuted ? 1u : 0u;
        for (int c = 0; c < 64; ++c) a.pos[c] = pos[c];
        a.cls = static_cast<const tcp::TileClass *>(cls.ptr);
        a.blk = static_cast<const tcp::u32x4 *>(dtype == TCP_C64 ? blk32.ptr : blk64.ptr);
        a.xorder = static_cast<const uint32_t *>(xorder.ptr);
        a.grp = static_cast<const tcp::TileGroup *>(dtype == TCP_C64 ? grp32.ptr : grp64.ptr);
        a.term = static_cast<const tcp::TileTerm *>(dtype == TCP_C64 ? term32.ptr : term64.ptr);
        a.H = static_cast<uint32_t>(H);
        a.L = L;
        return a;
    }
};

// XCD-order shift of a launch (see mvp_tile_kernel): the bit map's XCD bits
// when it chose them and the space has tiles above them, else the top three
// tile bits; kNoXcdOrder without an XCD order or with fewer than 8 tiles.
uint32_t tile_xcd_shift(const PlanTiles &pt, uint64_t dim) {
    const uint64_t wgs = dim >> (pt.L + pt.twin);
    if (!(pt.tiling & TCP_TILE_XCD_ORDER) || wgs < 8) return kNoXcdOrder;
    const uint32_t top = 63u - static_cast<uint32_t>(__builtin_clzll(wgs)) - 3u;
    return pt.xcd_bits_at <= top ? pt.xcd_bits_at : top;
}

struct tcp_mvp_plan {
    uint32_t n = 0;
    uint64_t T = 0, G = 0;
    int storage = TCP_STORAGE_LAZY;
    uint64_t transitions = UINT64_MAX;
    int device = 0;
    // complex128 tables, and complex64 tables when that storage uses another geometry
    PlanTiles t128, t64;
    bool own64 = false;
    uint64_t retained_bytes = 0;
    std::mutex staging_mutex;
    tcp::DeviceBuffer stage_in, stage_out;

    uint64_t dim() const { return 1ull << n; }
    const PlanTiles &tiles(int dtype) const { return dtype == TCP_C64 && own64 ? t64 : t128; }
    MvpArgs args(int dtype) const { return tiles(dtype).args(dtype); }
    uint32_t xcd_shift(const PlanTiles &pt) const { return tile_xcd_shift(pt, dim()); }
};

namespace {

using tcp::fail;

// Geometry index for `tiling` on an n-qubit space; 0 means the small-space
// geometry (512 threads, 1-4 amplitudes each).  Automatic, measured on the
// 28-qubit lattice workloads and random 22-27-qubit operators: complex128
// from 2^22 amplitudes uses 1024 x 8 twin tiles (its FP64 source tiles
// stream into LDS at a fixed rate per CU, so the fewer classes of 2^14
// amplitudes win: 6-16 % faster than single 1024 x 8 tiles), complex64
// storage from 2^24 uses 512 x 8 (bound by the pass loop, which the eight
// amplitudes per thread amortise); smaller spaces use 512 x 8 and 512 x 4,
// which give more workgroups.
uint32_t resolve_geometry(uint32_t n, uint32_t tiling, int dtype) {
    uint32_t g = tiling & 0xffu;
    if (g == TCP_TILE_AUTO) {
        if (dtype == TCP_C64) g = n >= 24 ? TCP_TILE_512x8 : TCP_TILE_512x4;
        else g = n >= 22 ? TCP_TILE_1024x8x2 : TCP_TILE_512x8;
    }
    const uint32_t L = log2u(kGeometries[g].threads) + log2u(kGeometries[g].amps) + kGeometries[g].twin;
    if (n < L) g = n >= 13 ? TCP_TILE_1024x8 : (n >= 12 ? TCP_TILE_512x8 : (n >= 11 ? TCP_TILE_512x4 : 0));
    return g;
}

int build_tiles(const tcp::Operator &op, uint32_t geo, bool xcd, uint32_t pipeline, PlanTiles &pt,
                tcp::TileTables &tab) {
    const uint32_t n = op.nqubits;
    pt.tiling = geo | (xcd ? TCP_TILE_XCD_ORDER : 0u) | pipeline;
    if (geo == 0) {
        pt.tb = std::min(n, kSmallBits);
        pt.ab = n > kSmallBits ? n - kSmallBits : 0;
    } else {
        pt.tb = log2u(kGeometries[geo].threads);
        pt.ab = log2u(kGeometries[geo].amps);
        pt.twin = kGeometries[geo].twin;
    }
    pt.L = pt.tb + pt.ab;
    pt.n = n;
    for (uint32_t c = 0; c < 64; ++c) pt.pos[c] = static_cast<uint8_t>(c);
    if (geo != 0) {
        // tile bits that vary across the workgroups resident on one XCD (32 CUs)
        const uint32_t threads = kGeometries[geo].threads, amps = kGeometries[geo].amps;
        const uint32_t per_cu = std::max<uint32_t>(
            1, std::min<uint32_t>((160u * 1024u) / ((2 * amps + 1) * threads * 8u + 32u), 2048u / threads));
        // window: half the workgroups one XCD runs at once (measured 2-3% faster
        // on the 28-qubit lattice than all of them, whose source tiles fill the
        // 4 MiB L2 exactly); three XCD bits above it
        const uint32_t window = log2u(32u * per_cu) - 1;
        uint8_t pos[64];
        if (tcp::choose_bit_map(op, pt.tb, pt.ab, window, 3, pos, pt.twin)) {
            for (uint32_t c = 0; c < n; ++c) {
                pt.pos[c] = pos[c];
                pt.permuted = pt.permuted || pos[c] != c;
            }
            pt.xcd_bits_at = window;
        }
    }
    tab.append(op, pt.tb, pt.ab, pt.permuted ? pt.pos : nullptr, geo == 0 ? kSmallThreads : kGeometries[geo].threads,
               pt.twin != 0, pt.permuted ? pt.xcd_bits_at : 0u);
    pt.H = tab.H;
    pt.passes = tab.passes;
    pt.diag_passes = static_cast<uint32_t>(tab.diag_passes);
    tab.pad_empty();
    pt.bytes = tab.device_bytes();
    return TCP_OK;
}

int upload_tiles(PlanTiles &pt, const tcp::TileTables &tab) {
    int st;
    if ((st = pt.cls.upload(tab.cls.data(), tab.cls.size() * sizeof(tcp::TileClass)))) return st;
    if ((st = pt.blk64.upload(tab.blk64.data(), tab.blk64.size() * sizeof(tcp::u32x4)))) return st;
    if ((st = pt.blk32.upload(tab.blk32.data(), tab.blk32.size() * sizeof(tcp::u32x4)))) return st;
    if ((st = pt.xorder.upload(tab.xorder.data(), tab.xorder.size() * sizeof(uint32_t)))) return st;
    if ((st = pt.grp64.upload(tab.grp64.data(), tab.grp64.size() * sizeof(tcp::TileGroup)))) return st;
    if ((st = pt.grp32.upload(tab.grp32.data(), tab.grp32.size() * sizeof(tcp::TileGroup)))) return st;
    if ((st = pt.term64.upload(tab.term64.data(), tab.term64.size() * sizeof(tcp::TileTerm)))) return st;
    return pt.term32.upload(tab.term32.data(), tab.term32.size() * sizeof(tcp::TileTerm));
}

int build_plan(const tcp::Operator &op, int storage, uint64_t max_bytes, uint32_t tiling, tcp_mvp_plan **out) {
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (storage != TCP_STORAGE_LAZY && storage != TCP_STORAGE_EAGER)
        return fail(TCP_ERR_INVALID_ARGUMENT, "storage must be either 'eager' or 'lazy'");
    if ((tiling & 0xffu) >= kGeometryCount ||
        (tiling & ~(0xffu | TCP_TILE_XCD_ORDER | TCP_TILE_LINEAR_ORDER | TCP_TILE_PREFETCH_MASK | TCP_TILE_NT_LOADS)) ||
        tile_prefetch_depth(tiling) > 3 ||
        ((tiling & TCP_TILE_XCD_ORDER) && (tiling & TCP_TILE_LINEAR_ORDER)))
        return fail(TCP_ERR_INVALID_ARGUMENT, "unknown tile geometry");
    const uint32_t n = op.nqubits;
    if (n >= 64) return fail(TCP_ERR_OVERFLOW, "integer overflow while computing matrix dimension");
    const uint64_t T = op.term_count();
    if (T >= (1ull << 31)) return fail(TCP_ERR_UNSUPPORTED, "more than 2^31 terms");

    auto plan = new tcp_mvp_plan();
    std::unique_ptr<tcp_mvp_plan> guard(plan);
    plan->n = n;
    plan->T = T;
    plan->storage = storage;
    // automatic tiling walks the tiles XCD by XCD (measured faster) unless told otherwise
    const bool xcd = (tiling & TCP_TILE_XCD_ORDER) ||
                     ((tiling & 0xffu) == TCP_TILE_AUTO && !(tiling & TCP_TILE_LINEAR_ORDER));
    const uint32_t g128 = resolve_geometry(n, tiling, TCP_C128), g64 = resolve_geometry(n, tiling, TCP_C64);
    tcp::TileTables tab128, tab64;
    // prefetch depth and load policy apply to the geometries that have them
    const uint32_t pipeline = tiling & (TCP_TILE_PREFETCH_MASK | TCP_TILE_NT_LOADS);
    auto pipelined = [](uint32_t g) { return g == TCP_TILE_512x8 || g == TCP_TILE_512x4 || g == TCP_TILE_1024x4; };
    build_tiles(op, g128, xcd, pipelined(g128) ? pipeline : 0u, plan->t128, tab128);
    plan->G = tab128.G;
    plan->own64 = g64 != g128;
    if (plan->own64) build_tiles(op, g64, xcd, pipelined(g64) ? pipeline : 0u, plan->t64, tab64);

    plan->retained_bytes = plan->t128.bytes + (plan->own64 ? plan->t64.bytes : 0) + sizeof(tcp_mvp_plan);
    if (tcp::exceeds_budget(plan->retained_bytes, max_bytes))
        return tcp::memory_limit(plan->retained_bytes, max_bytes);

    TCP_HIP_CHECK(hipGetDevice(&plan->device));
    int st;
    if ((st = upload_tiles(plan->t128, tab128))) return st;
    if (plan->own64 && (st = upload_tiles(plan->t64, tab64))) return st;

    if (storage == TCP_STORAGE_EAGER) {
        unsigned long long *dcount = nullptr;
        TCP_HIP_CHECK(hipMalloc(&dcount, sizeof(unsigned long long)));
        TCP_HIP_CHECK(hipMemset(dcount, 0, sizeof(unsigned long long)));
        const uint64_t dim = plan->dim();
        const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((dim + 255) / 256, 4096));
        hipLaunchKernelGGL(mvp_count_kernel, dim3(blocks), dim3(256), 0, nullptr, plan->args(TCP_C128),
                           plan->t128.tb, static_cast<uint32_t>(plan->t128.passes), dim, dcount);
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

// Everything one apply launch needs: the space, the geometry word and the
// device tables (a plan's own, or tables staged in the static workspace).
struct LaunchSpec {
    uint64_t dim;
    uint32_t n, tiling, L, ab;
    uint64_t H;
    uint32_t xcd_shift;  // kNoXcdOrder: launch order
    MvpArgs args;
};

LaunchSpec launch_spec(const tcp_mvp_plan *plan, int dtype) {
    const PlanTiles &pt = plan->tiles(dtype);
    return LaunchSpec{plan->dim(), plan->n, pt.tiling, pt.L, pt.ab, pt.H, plan->xcd_shift(pt), pt.args(dtype)};
}

template <int A, typename VT, uint32_t TPB, bool PARTIAL = false, int PF = 1, bool NT = false>
hipError_t launch_tile(const LaunchSpec &ls, const void *in, void *out, uint32_t nrhs, hipStream_t stream) {
    const uint64_t tiles = PARTIAL ? 1 : ls.dim >> ls.L;
    hipLaunchKernelGGL((mvp_tile_kernel<A, VT, TPB, PARTIAL, PF, NT>), dim3(static_cast<uint32_t>(tiles), nrhs),
                       dim3(TPB), 0, stream, static_cast<const VT *>(in), static_cast<VT *>(out), ls.args, ls.dim,
                       ls.xcd_shift);
    return hipGetLastError();
}

template <int A, typename VT, uint32_t TPB, int CR = 0>
hipError_t launch_twin(const LaunchSpec &ls, const void *in, void *out, uint32_t nrhs, hipStream_t stream) {
    const uint64_t pairs = ls.dim >> (ls.L + 1);
    hipLaunchKernelGGL((mvp_twin_kernel<A, VT, TPB, CR>), dim3(static_cast<uint32_t>(pairs), nrhs), dim3(TPB), 0, stream,
                       static_cast<const VT *>(in), static_cast<VT *>(out), ls.args, ls.dim, ls.xcd_shift);
    return hipGetLastError();
}

// Geometries with a selectable prefetch depth (1-3 tiles in flight) and load policy.
template <int A, typename VT, uint32_t TPB>
hipError_t launch_pipelined(const LaunchSpec &ls, const void *in, void *out, uint32_t nrhs, hipStream_t stream) {
    const uint32_t depth = tile_prefetch_depth(ls.tiling);
    const bool nt = (ls.tiling & TCP_TILE_NT_LOADS) != 0;
    if (depth == 3) return nt ? launch_tile<A, VT, TPB, false, 3, true>(ls, in, out, nrhs, stream)
                              : launch_tile<A, VT, TPB, false, 3, false>(ls, in, out, nrhs, stream);
    if (depth == 2) return nt ? launch_tile<A, VT, TPB, false, 2, true>(ls, in, out, nrhs, stream)
                              : launch_tile<A, VT, TPB, false, 2, false>(ls, in, out, nrhs, stream);
    return nt ? launch_tile<A, VT, TPB, false, 1, true>(ls, in, out, nrhs, stream)
              : launch_tile<A, VT, TPB, false, 1, false>(ls, in, out, nrhs, stream);
}

template <typename VT>
hipError_t launch_typed(const LaunchSpec &ls, const void *in, void *out, uint32_t nrhs, hipStream_t stream) {
    if (ls.H == 0) {  // empty operator: H v = 0
        const uint64_t count = ls.dim * nrhs;
        const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((count + 255) / 256, 65536));
        hipLaunchKernelGGL(zero_kernel<VT>, dim3(blocks), dim3(256), 0, stream, static_cast<VT *>(out), count);
        return hipGetLastError();
    }
    switch (ls.tiling & 0xffu) {
        case TCP_TILE_512x8: return launch_pipelined<8, VT, 512>(ls, in, out, nrhs, stream);
        case TCP_TILE_1024x8: return launch_tile<8, VT, 1024>(ls, in, out, nrhs, stream);
        case TCP_TILE_512x16: return launch_tile<16, VT, 512>(ls, in, out, nrhs, stream);
        case TCP_TILE_512x4: return launch_pipelined<4, VT, 512>(ls, in, out, nrhs, stream);
        case TCP_TILE_1024x4: return launch_pipelined<4, VT, 1024>(ls, in, out, nrhs, stream);
        case TCP_TILE_1024x8x2: return launch_twin<8, VT, 1024, 1>(ls, in, out, nrhs, stream);
        case TCP_TILE_512x16x2: return launch_twin<16, VT, 512>(ls, in, out, nrhs, stream);
        default: break;
    }
    if (ls.n < kSmallBits) return launch_tile<1, VT, kSmallThreads, true>(ls, in, out, nrhs, stream);
    switch (ls.ab) {
        case 0: return launch_tile<1, VT, kSmallThreads>(ls, in, out, nrhs, stream);
        case 1: return launch_tile<2, VT, kSmallThreads>(ls, in, out, nrhs, stream);
        default: return launch_tile<4, VT, kSmallThreads>(ls, in, out, nrhs, stream);
    }
}

hipError_t launch_any(const LaunchSpec &ls, const void *in, void *out, uint32_t nrhs, int dtype, hipStream_t stream) {
    return dtype == TCP_C128 ? launch_typed<dvec2>(ls, in, out, nrhs, stream)
                             : launch_typed<fvec2>(ls, in, out, nrhs, stream);
}

int apply_device(const tcp_mvp_plan *plan, const void *in, void *out, uint64_t nrhs, int dtype, void *stream) {
    if (dtype != TCP_C128 && dtype != TCP_C64)
        return fail(TCP_ERR_INVALID_ARGUMENT, "dtype must be TCP_C128 or TCP_C64");
    if (!in || !out) return fail(TCP_ERR_INVALID_ARGUMENT, "null state buffer");
    if (nrhs == 0) return tcp::ok();
    if (nrhs > 65535) return fail(TCP_ERR_UNSUPPORTED, "at most 65535 right-hand sides per call");
    const size_t amp = dtype == TCP_C128 ? 16 : 8;
    const size_t bytes = static_cast<size_t>(plan->dim()) * amp * nrhs;
    if (tcp::ranges_overlap(in, bytes, out, bytes))
        return fail(TCP_ERR_INVALID_ARGUMENT, "apply_into input and output must not overlap");
    if ((plan->dim() >> plan->tiles(dtype).L) > 0x7fffffffull)
        return fail(TCP_ERR_UNSUPPORTED, "too many tiles for one launch");
    tcp::DeviceGuard dg;
    int st = dg.enter(plan->device);
    if (st) return st;
    hipError_t e = launch_any(launch_spec(plan, dtype), in, out, static_cast<uint32_t>(nrhs), dtype,
                              static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return tcp::hip_fail(e, "mvp kernel launch");
    return tcp::ok();
}

}  // namespace

// Workspace of tcp_mvp_apply_static: the class table and class blocks of the
// operator being applied are staged here instead of in allocated memory.
// Each device holds its own copy until the process ends; host code reaches
// it only through the address hipGetSymbolAddress returns on that device.
constexpr size_t kStaticTableBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticTableBytes];

namespace {

std::mutex g_scratch_mutex;
std::vector<void *> g_scratch_ptr;  // per device ordinal

// Address of the current device's workspace (looked up once per device).
int scratch_address(void **out) {
    int device = 0;
    TCP_HIP_CHECK(hipGetDevice(&device));
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr.size() <= static_cast<size_t>(device)) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        TCP_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return TCP_OK;
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
                       plan->t128.tb, plan->t128.diag_passes, dim, diag);
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
    return build_plan(tcp::operator_of(op), storage, max_bytes, TCP_TILE_AUTO, out);
}

int tcp_mvp_plan_create_tiled(const tcp_operator *op, int storage, uint64_t max_bytes, uint32_t tiling,
                              tcp_mvp_plan **out) {
    if (!op) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    return build_plan(tcp::operator_of(op), storage, max_bytes, tiling, out);
}

uint32_t tcp_mvp_plan_tiling(const tcp_mvp_plan *plan, int dtype) { return plan ? plan->tiles(dtype).tiling : 0; }
uint32_t tcp_mvp_plan_tile_bits(const tcp_mvp_plan *plan, int dtype) { return plan ? plan->tiles(dtype).L : 0; }

int tcp_mvp_plan_create_from_words(uint32_t nqubits, uint64_t nterms, const uint64_t *x_words,
                                   const uint64_t *z_words, const double *coef, int storage,
                                   uint64_t max_bytes, tcp_mvp_plan **out) {
    tcp_operator *op = nullptr;
    int st = tcp_operator_from_words(nqubits, nterms, x_words, z_words, coef, &op);
    if (st) return st;
    st = build_plan(tcp::operator_of(op), storage, max_bytes, TCP_TILE_AUTO, out);
    tcp_operator_destroy(op);
    return st;
}

uint32_t tcp_mvp_plan_nqubits(const tcp_mvp_plan *plan) { return plan ? plan->n : 0; }
uint64_t tcp_mvp_plan_term_count(const tcp_mvp_plan *plan) { return plan ? plan->T : 0; }
uint64_t tcp_mvp_plan_group_count(const tcp_mvp_plan *plan) { return plan ? plan->G : 0; }
// Twin-tile tables list two stages (halves) per class.
uint64_t tcp_mvp_plan_class_count(const tcp_mvp_plan *plan) { return plan ? plan->t128.H >> plan->t128.twin : 0; }
uint64_t tcp_mvp_plan_pass_count(const tcp_mvp_plan *plan) { return plan ? plan->t128.passes : 0; }
uint64_t tcp_mvp_plan_class_count_for(const tcp_mvp_plan *plan, int dtype) {
    return plan ? plan->tiles(dtype).H >> plan->tiles(dtype).twin : 0;
}
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

int tcp_mvp_apply_static(const tcp_operator *op, const void *input, void *output, uint64_t input_len,
                         uint64_t output_len, int dtype, void *stream) {
    if (!op) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    if (dtype != TCP_C128 && dtype != TCP_C64)
        return fail(TCP_ERR_INVALID_ARGUMENT, "dtype must be TCP_C128 or TCP_C64");
    const tcp::Operator &o = tcp::operator_of(op);
    const uint32_t n = o.nqubits;
    if (n >= 64) return fail(TCP_ERR_OVERFLOW, "integer overflow while computing matrix dimension");
    const uint64_t dim = 1ull << n;
    int st;
    if ((st = check_len(dim, input_len)) || (st = check_len(dim, output_len))) return st;
    if (!input || !output) return fail(TCP_ERR_INVALID_ARGUMENT, "null state buffer");
    const size_t vec_bytes = static_cast<size_t>(dim) * (dtype == TCP_C128 ? 16 : 8);
    if (tcp::ranges_overlap(input, vec_bytes, output, vec_bytes))
        return fail(TCP_ERR_INVALID_ARGUMENT, "apply_into input and output must not overlap");
    // the plan builder's tables, on the host
    PlanTiles pt;
    tcp::TileTables tab;
    build_tiles(o, resolve_geometry(n, TCP_TILE_AUTO, dtype), true, 0u, pt, tab);
    if ((dim >> pt.L) > 0x7fffffffull) return fail(TCP_ERR_UNSUPPORTED, "too many tiles for one launch");
    const std::vector<tcp::u32x4> &blk = dtype == TCP_C64 ? tab.blk32 : tab.blk64;
    const size_t cls_bytes = tab.cls.size() * sizeof(tcp::TileClass);
    const size_t blk_off = (cls_bytes + 255) & ~size_t(255);
    const size_t blk_bytes = blk.size() * sizeof(tcp::u32x4);
    if (blk_off + blk_bytes > kStaticTableBytes)
        return fail(TCP_ERR_UNSUPPORTED, "tcp_mvp_apply_static: the operator's tables need " +
                                             std::to_string(blk_off + blk_bytes) + " bytes, more than the " +
                                             std::to_string(kStaticTableBytes) + "-byte static workspace");
    void *ws = nullptr;
    if ((st = scratch_address(&ws))) return st;
    char *base = static_cast<char *>(ws);
    hipStream_t s = static_cast<hipStream_t>(stream);
    TCP_HIP_CHECK(hipMemcpyAsync(base, tab.cls.data(), cls_bytes, hipMemcpyHostToDevice, s));
    TCP_HIP_CHECK(hipMemcpyAsync(base + blk_off, blk.data(), blk_bytes, hipMemcpyHostToDevice, s));
    TCP_HIP_CHECK(hipStreamSynchronize(s));  // the host tables are released on return
    MvpArgs a = pt.args(dtype);
    a.cls = reinterpret_cast<const tcp::TileClass *>(base);
    a.blk = reinterpret_cast<const tcp::u32x4 *>(base + blk_off);
    a.grp = nullptr;  // pass and term tables serve the plan builders only
    a.term = nullptr;
    a.xorder = nullptr;
    const LaunchSpec ls{dim, n, pt.tiling, pt.L, pt.ab, pt.H, tile_xcd_shift(pt, dim), a};
    hipError_t e = launch_any(ls, input, output, 1, dtype, s);
    if (e != hipSuccess) return tcp::hip_fail(e, "mvp kernel launch");
    return tcp::ok();
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
    st = apply_device(plan, mp->stage_in.ptr, mp->stage_out.ptr, 1, TCP_C128, nullptr);
    hipError_t e = st ? hipSuccess : hipMemcpy(output, mp->stage_out.ptr, bytes, hipMemcpyDeviceToHost);
    if (bytes > kStagingKeepBytes) {  // large staging is released, never held for the plan's lifetime
        mp->stage_in.release();
        mp->stage_out.release();
    }
    if (st) return st;
    if (e != hipSuccess) return tcp::hip_fail(e, "hipMemcpy device to host");
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
