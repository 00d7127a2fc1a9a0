// Batched MVP: many Pauli operators of uneven size, each applied to a block of
// right-hand sides, in one persistent launch (no reference counterpart; the
// reference applies one plan per call, core/hamiltonian.rs:204-270).
//
// Every (operator, right-hand side, output tile) is one work item.  Items of
// the costliest operators come first (longest-processing-time order) and a
// persistent grid sized to the device's resident workgroups pulls items from
// an atomic counter, so uneven operators share the machine without a launch
// per operator.  Each item runs the same tile routine as the single-operator
// kernel (classes of x-groups, source tiles staged in LDS).  The counter is
// cleared by a one-thread kernel, so the whole apply can be captured into a
// hipGraph and replayed.  The autotuner times each tile geometry (threads per
// workgroup x amplitudes per thread) on the caller's buffers and keeps the
// fastest, then the number of classes per work item on that geometry: an
// operator with more classes than that is split into several items per
// output tile (split-K over its classes), whose partial tiles are added to
// the output with atomic adds -- more items to balance few heavy operators
// over the device, at the price of a cleared output, atomic traffic and a
// summation order that varies from run to run (within rounding).
#include <algorithm>
#include <memory>
#include <numeric>
#include <vector>

#include "tile_tables.hpp"

namespace {

using tcp::dvec2;
using tcp::fvec2;
using tcp::TileArgs;
using tcp::fail;

struct BatchProblem {
    uint64_t in_off;      // amplitude offset of this operator's right-hand-side block
    uint64_t dim;         // 2^n
    uint64_t item_begin;  // first work item
    uint32_t tiles;       // output tiles per right-hand side
    uint32_t H;           // classes
    uint32_t L;           // tile bits
    uint32_t cls_base;    // first entry in the class table
    uint32_t active;      // amplitudes per tile (2^L, or 2^n when the space is smaller than a tile)
    uint32_t chunk;       // classes per work item (H: one item per tile)
    uint32_t chunks;      // work items per output tile, ceil(H / chunk)
    uint32_t pad;
};

struct Config {
    uint32_t threads, amps;
};
constexpr Config kConfigs[] = {{256, 2}, {256, 4}, {512, 2}, {128, 8}, {512, 4}};
constexpr uint32_t kConfigCount = sizeof(kConfigs) / sizeof(kConfigs[0]);
// classes per work item the autotuner tries (0: all of an operator's classes)
constexpr uint32_t kSplits[] = {0, 256, 64, 16};
constexpr uint32_t kSplitCount = sizeof(kSplits) / sizeof(kSplits[0]);

__global__ void reset_counter_kernel(unsigned int *counter) { *counter = 0u; }

// Clears the output before split items add into it (a kernel, not a memset,
// so the apply can be captured).
template <typename VT>
__global__ void clear_kernel(VT *__restrict__ y, uint64_t count) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        y[i] = VT{0, 0};
}

template <int A, typename VT, uint32_t TPB>
__global__ __launch_bounds__(TPB) void batch_kernel(const VT *__restrict__ v, VT *__restrict__ y,
                                                    const BatchProblem *__restrict__ probs, uint32_t nprob,
                                                    TileArgs shared, uint64_t total_items,
                                                    unsigned int *__restrict__ counter) {
    __shared__ dvec2 lds[TPB * A];
    __shared__ tcp::u32x4 lblk[tcp::block_words(TPB) + 2];
    __shared__ unsigned int item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(counter, 1u);
        __syncthreads();
        const uint64_t item = item_s;
        __syncthreads();  // item read by every thread before it is overwritten
        if (item >= total_items) break;
        uint32_t lo = 0, hi = nprob;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (probs[mid].item_begin <= item) lo = mid;
            else hi = mid;
        }
        const BatchProblem P = probs[lo];
        const uint64_t local = item - P.item_begin;
        const uint64_t per_rhs = static_cast<uint64_t>(P.tiles) * P.chunks;
        const uint64_t rhs = local / per_rhs;
        const uint64_t rest = local - rhs * per_rhs;
        const uint32_t chunk = static_cast<uint32_t>(rest / P.tiles);  // chunk-major: fewer concurrent adds per tile
        const uint64_t tile_id = rest - static_cast<uint64_t>(chunk) * P.tiles;
        TileArgs a = shared;
        a.cls += P.cls_base + chunk * P.chunk;
        a.H = P.chunks > 1 ? min(P.chunk, P.H - chunk * P.chunk) : P.H;
        a.L = P.L;
        const uint64_t off = P.in_off + rhs * P.dim;
        tcp::tile_apply<A, VT, TPB, true>(v + off, y + off, a, tile_id, P.active, lds, lblk, P.chunks > 1);
    }
}

struct Tables {
    bool built = false;
    tcp::DeviceBuffer cls, blk64, blk32, grp64, grp32, term64, term32;
    tcp::DeviceBuffer probs[kSplitCount];  // work items for each classes-per-item setting
    uint64_t total_items[kSplitCount] = {};
    bool any_split[kSplitCount] = {};
};

}  // namespace

struct tcp_batch_plan {
    std::vector<tcp::Operator> ops;
    std::vector<uint64_t> offsets;  // per operator, in input order
    uint32_t nrhs = 1;
    uint64_t total = 0;             // amplitudes of the whole batch buffer
    double work = 0;                // sum of terms x 2^n x nrhs
    int device = 0;
    uint32_t config = 0;
    uint32_t split = 0;  // index into kSplits
    float split_times[kSplitCount] = {-1.0f, -1.0f, -1.0f, -1.0f};
    Tables tables[kConfigCount];
    tcp::DeviceBuffer counter;
    uint32_t grid[kConfigCount] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    ~tcp_batch_plan() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

template <int A, typename VT, uint32_t TPB>
const void *kernel_ptr() {
    return reinterpret_cast<const void *>(&batch_kernel<A, VT, TPB>);
}

template <typename VT>
const void *kernel_for(uint32_t cfg) {
    switch (cfg) {
        case 0: return kernel_ptr<2, VT, 256>();
        case 1: return kernel_ptr<4, VT, 256>();
        case 2: return kernel_ptr<2, VT, 512>();
        case 3: return kernel_ptr<8, VT, 128>();
        default: return kernel_ptr<4, VT, 512>();
    }
}

uint32_t log2u(uint32_t x) { return 31u - static_cast<uint32_t>(__builtin_clz(x)); }

int build_tables(tcp_batch_plan *plan, uint32_t cfg) {
    Tables &tb = plan->tables[cfg];
    if (tb.built) return TCP_OK;
    const uint32_t tbits = log2u(kConfigs[cfg].threads), abits = log2u(kConfigs[cfg].amps);
    const uint32_t L = tbits + abits;
    tcp::TileTables tab;
    const size_t P = plan->ops.size();
    std::vector<BatchProblem> probs(P);
    std::vector<double> cost(P);
    for (size_t p = 0; p < P; ++p) {
        tab.append(plan->ops[p], tbits, abits);
        BatchProblem &bp = probs[p];
        bp.in_off = plan->offsets[p];
        bp.dim = 1ull << plan->ops[p].nqubits;
        bp.tiles = static_cast<uint32_t>(std::max<uint64_t>(1, bp.dim >> L));
        bp.H = static_cast<uint32_t>(tab.H);
        bp.L = L;
        bp.cls_base = static_cast<uint32_t>(tab.cls_base);
        bp.active = static_cast<uint32_t>(std::min<uint64_t>(bp.dim, 1ull << L));
        cost[p] = static_cast<double>(plan->ops[p].term_count() + 1) * static_cast<double>(bp.dim);
    }
    // longest processing time first
    std::vector<size_t> order(P);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    int st;
    for (uint32_t k = 0; k < kSplitCount; ++k) {
        std::vector<BatchProblem> sorted(P);
        uint64_t item = 0;
        bool split = false;
        for (size_t i = 0; i < P; ++i) {
            BatchProblem &bp = sorted[i];
            bp = probs[order[i]];
            bp.chunk = kSplits[k] && bp.H > kSplits[k] ? kSplits[k] : std::max<uint32_t>(bp.H, 1u);
            bp.chunks = bp.H > bp.chunk ? (bp.H + bp.chunk - 1) / bp.chunk : 1u;
            split = split || bp.chunks > 1;
            bp.item_begin = item;
            item += static_cast<uint64_t>(bp.tiles) * bp.chunks * plan->nrhs;
        }
        if (item >= (1ull << 32)) return fail(TCP_ERR_UNSUPPORTED, "more than 2^32 work items");
        tb.total_items[k] = item;
        tb.any_split[k] = split;
        if ((st = tb.probs[k].upload(sorted.data(), sizeof(BatchProblem) * std::max<size_t>(sorted.size(), 1))))
            return st;
    }
    tab.pad_empty();
    if ((st = tb.cls.upload(tab.cls.data(), sizeof(tcp::TileClass) * tab.cls.size())) ||
        (st = tb.blk64.upload(tab.blk64.data(), sizeof(tcp::u32x4) * tab.blk64.size())) ||
        (st = tb.blk32.upload(tab.blk32.data(), sizeof(tcp::u32x4) * tab.blk32.size())) ||
        (st = tb.grp64.upload(tab.grp64.data(), sizeof(tcp::TileGroup) * tab.grp64.size())) ||
        (st = tb.grp32.upload(tab.grp32.data(), sizeof(tcp::TileGroup) * tab.grp32.size())) ||
        (st = tb.term64.upload(tab.term64.data(), sizeof(tcp::TileTerm) * tab.term64.size())) ||
        (st = tb.term32.upload(tab.term32.data(), sizeof(tcp::TileTerm) * tab.term32.size())))
        return st;
    // persistent grid: every resident workgroup of the device
    hipDeviceProp_t prop;
    TCP_HIP_CHECK(hipGetDeviceProperties(&prop, plan->device));
    int per_cu_d = 0, per_cu_f = 0;
    TCP_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_d, kernel_for<dvec2>(cfg),
                                                               static_cast<int>(kConfigs[cfg].threads), 0));
    TCP_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_f, kernel_for<fvec2>(cfg),
                                                               static_cast<int>(kConfigs[cfg].threads), 0));
    const int per_cu = std::max(1, std::min(per_cu_d, per_cu_f));
    plan->grid[cfg] = static_cast<uint32_t>(prop.multiProcessorCount * per_cu);
    tb.built = true;
    return TCP_OK;
}

template <int A, typename VT, uint32_t TPB>
void launch_cfg(const tcp_batch_plan *plan, uint32_t cfg, uint32_t split, const void *in, void *out, int dtype,
                hipStream_t s) {
    const Tables &tb = plan->tables[cfg];
    TileArgs a;
    a.cls = static_cast<const tcp::TileClass *>(tb.cls.ptr);
    a.blk = static_cast<const tcp::u32x4 *>(dtype == TCP_C64 ? tb.blk32.ptr : tb.blk64.ptr);
    a.xorder = nullptr;
    a.grp = static_cast<const tcp::TileGroup *>(dtype == TCP_C64 ? tb.grp32.ptr : tb.grp64.ptr);
    a.term = static_cast<const tcp::TileTerm *>(dtype == TCP_C64 ? tb.term32.ptr : tb.term64.ptr);
    a.H = 0;
    a.L = 0;
    a.n = 0;
    a.permuted = 0;
    for (int c = 0; c < 64; ++c) a.pos[c] = static_cast<uint8_t>(c);
    if (tb.any_split[split]) {
        const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((plan->total + 255) / 256, 256u * 64u));
        hipLaunchKernelGGL(clear_kernel<VT>, dim3(blocks), dim3(256), 0, s, static_cast<VT *>(out), plan->total);
    }
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(plan->grid[cfg], tb.total_items[split]));
    hipLaunchKernelGGL((batch_kernel<A, VT, TPB>), dim3(grid), dim3(TPB), 0, s, static_cast<const VT *>(in),
                       static_cast<VT *>(out), static_cast<const BatchProblem *>(tb.probs[split].ptr),
                       static_cast<uint32_t>(plan->ops.size()), a, tb.total_items[split],
                       static_cast<unsigned int *>(plan->counter.ptr));
}

template <typename VT>
void launch_typed(const tcp_batch_plan *plan, uint32_t cfg, uint32_t split, const void *in, void *out, int dtype,
                  hipStream_t s) {
    switch (cfg) {
        case 0: launch_cfg<2, VT, 256>(plan, cfg, split, in, out, dtype, s); break;
        case 1: launch_cfg<4, VT, 256>(plan, cfg, split, in, out, dtype, s); break;
        case 2: launch_cfg<2, VT, 512>(plan, cfg, split, in, out, dtype, s); break;
        case 3: launch_cfg<8, VT, 128>(plan, cfg, split, in, out, dtype, s); break;
        default: launch_cfg<4, VT, 512>(plan, cfg, split, in, out, dtype, s); break;
    }
}

int enqueue(const tcp_batch_plan *plan, uint32_t cfg, uint32_t split, const void *in, void *out, int dtype,
            hipStream_t s) {
    if (plan->tables[cfg].total_items[split] == 0) return TCP_OK;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, static_cast<unsigned int *>(plan->counter.ptr));
    if (dtype == TCP_C128) launch_typed<dvec2>(plan, cfg, split, in, out, dtype, s);
    else launch_typed<fvec2>(plan, cfg, split, in, out, dtype, s);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : tcp::hip_fail(e, "batched kernel launch");
}

int check_buffers(const tcp_batch_plan *plan, const void *in, void *out, int dtype) {
    if (dtype != TCP_C128 && dtype != TCP_C64)
        return fail(TCP_ERR_INVALID_ARGUMENT, "dtype must be TCP_C128 or TCP_C64");
    if (!in || !out) return fail(TCP_ERR_INVALID_ARGUMENT, "null state buffer");
    const size_t bytes = static_cast<size_t>(plan->total) * (dtype == TCP_C128 ? 16 : 8);
    if (tcp::ranges_overlap(in, bytes, out, bytes))
        return fail(TCP_ERR_INVALID_ARGUMENT, "apply_into input and output must not overlap");
    return TCP_OK;
}

}  // namespace

extern "C" {

int tcp_batch_plan_create(const tcp_operator *const *ops, uint32_t count, uint32_t nrhs, tcp_batch_plan **out) {
    if (!out || (count && !ops) || nrhs == 0) return fail(TCP_ERR_INVALID_ARGUMENT, "bad argument");
    *out = nullptr;
    auto *plan = new tcp_batch_plan();
    std::unique_ptr<tcp_batch_plan> guard(plan);
    plan->nrhs = nrhs;
    for (uint32_t p = 0; p < count; ++p) {
        if (!ops[p]) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator in batch");
        const tcp::Operator &op = tcp::operator_of(ops[p]);
        if (op.nqubits >= 40) return fail(TCP_ERR_UNSUPPORTED, "batched operators must have fewer than 40 qubits");
        plan->ops.push_back(op);
        plan->offsets.push_back(plan->total);
        plan->total += (1ull << op.nqubits) * nrhs;
        plan->work += static_cast<double>(op.term_count()) * static_cast<double>(1ull << op.nqubits) * nrhs;
    }
    TCP_HIP_CHECK(hipGetDevice(&plan->device));
    int st = plan->counter.allocate(sizeof(unsigned int));
    if (st) return st;
    // default: 512 x 4 tiles (operators smaller than a tile use a partial one)
    plan->config = kConfigCount - 1;
    if (count && (st = build_tables(plan, plan->config))) return st;
    *out = guard.release();
    return tcp::ok();
}

uint32_t tcp_batch_plan_count(const tcp_batch_plan *plan) { return plan ? static_cast<uint32_t>(plan->ops.size()) : 0; }
uint64_t tcp_batch_plan_offset(const tcp_batch_plan *plan, uint32_t p) {
    return plan && p < plan->offsets.size() ? plan->offsets[p] : UINT64_MAX;
}
uint64_t tcp_batch_plan_total_amplitudes(const tcp_batch_plan *plan) { return plan ? plan->total : 0; }
double tcp_batch_plan_work(const tcp_batch_plan *plan) { return plan ? plan->work : 0.0; }
uint32_t tcp_batch_config_count(void) { return kConfigCount; }

int tcp_batch_config(uint32_t index, uint32_t *threads, uint32_t *amplitudes) {
    if (index >= kConfigCount) return fail(TCP_ERR_INVALID_ARGUMENT, "config index out of range");
    if (threads) *threads = kConfigs[index].threads;
    if (amplitudes) *amplitudes = kConfigs[index].amps;
    return tcp::ok();
}

uint32_t tcp_batch_plan_selected(const tcp_batch_plan *plan) { return plan ? plan->config : 0; }

int tcp_batch_plan_select(tcp_batch_plan *plan, uint32_t config) {
    if (!plan || config >= kConfigCount) return fail(TCP_ERR_INVALID_ARGUMENT, "bad config");
    tcp::DeviceGuard dg;
    int st = dg.enter(plan->device);
    if (st) return st;
    if ((st = build_tables(plan, config))) return st;
    plan->config = config;
    return tcp::ok();
}

uint32_t tcp_batch_split_count(void) { return kSplitCount; }

int tcp_batch_split(uint32_t index, uint32_t *classes_per_item) {
    if (index >= kSplitCount) return fail(TCP_ERR_INVALID_ARGUMENT, "split index out of range");
    if (classes_per_item) *classes_per_item = kSplits[index];
    return tcp::ok();
}

uint32_t tcp_batch_plan_selected_split(const tcp_batch_plan *plan) { return plan ? plan->split : 0; }

int tcp_batch_plan_select_split(tcp_batch_plan *plan, uint32_t index) {
    if (!plan || index >= kSplitCount) return fail(TCP_ERR_INVALID_ARGUMENT, "bad split");
    plan->split = index;
    return tcp::ok();
}

int tcp_batch_plan_split_times(const tcp_batch_plan *plan, float *times_ms) {
    if (!plan || !times_ms) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    for (uint32_t k = 0; k < kSplitCount; ++k) times_ms[k] = plan->split_times[k];
    return tcp::ok();
}

int tcp_batch_plan_apply(const tcp_batch_plan *plan, const void *input, void *output, int dtype, void *stream) {
    if (!plan) return fail(TCP_ERR_INVALID_ARGUMENT, "null plan");
    int st = check_buffers(plan, input, output, dtype);
    if (st) return st;
    if (plan->ops.empty()) return tcp::ok();
    tcp::DeviceGuard dg;
    if ((st = dg.enter(plan->device))) return st;
    st = enqueue(plan, plan->config, plan->split, input, output, dtype, static_cast<hipStream_t>(stream));
    return st ? st : tcp::ok();
}

int tcp_batch_plan_autotune(tcp_batch_plan *plan, const void *input, void *output, int dtype, uint32_t repeats,
                            float *times_ms, uint32_t *chosen) {
    if (!plan) return fail(TCP_ERR_INVALID_ARGUMENT, "null plan");
    int st = check_buffers(plan, input, output, dtype);
    if (st) return st;
    tcp::DeviceGuard dg;
    if ((st = dg.enter(plan->device))) return st;
    if (repeats == 0) repeats = 3;
    hipEvent_t e0, e1;
    TCP_HIP_CHECK(hipEventCreate(&e0));
    TCP_HIP_CHECK(hipEventCreate(&e1));
    // per-apply time of (geometry c, split k), after one warm-up
    auto time_one = [&](uint32_t c, uint32_t k, float &ms) -> int {
        int s2;
        if ((s2 = build_tables(plan, c))) return s2;
        if ((s2 = enqueue(plan, c, k, input, output, dtype, nullptr))) return s2;
        TCP_HIP_CHECK(hipEventRecord(e0, nullptr));
        for (uint32_t r = 0; r < repeats; ++r)
            if ((s2 = enqueue(plan, c, k, input, output, dtype, nullptr))) return s2;
        TCP_HIP_CHECK(hipEventRecord(e1, nullptr));
        TCP_HIP_CHECK(hipEventSynchronize(e1));
        TCP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        ms /= static_cast<float>(repeats);
        return TCP_OK;
    };
    // the geometry with whole operators per tile, then the classes per item on it
    float best = 3.4e38f;
    uint32_t best_cfg = plan->config, best_split = 0;
    for (uint32_t c = 0; c < kConfigCount && !st; ++c) {
        if (times_ms) times_ms[c] = -1.0f;
        float ms = 0.0f;
        if ((st = time_one(c, 0, ms))) break;
        if (times_ms) times_ms[c] = ms;
        if (ms < best) {
            best = ms;
            best_cfg = c;
        }
    }
    for (uint32_t k = 0; k < kSplitCount; ++k) plan->split_times[k] = -1.0f;
    if (!st) plan->split_times[0] = best;
    for (uint32_t k = 1; k < kSplitCount && !st; ++k) {
        if (!plan->tables[best_cfg].any_split[k]) continue;  // splits nothing: same items as whole operators
        float ms = 0.0f;
        if ((st = time_one(best_cfg, k, ms))) break;
        plan->split_times[k] = ms;
        if (ms < best) {
            best = ms;
            best_split = k;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    plan->config = best_cfg;
    plan->split = best_split;
    if (chosen) *chosen = best_cfg;
    return tcp::ok();
}

int tcp_batch_plan_capture(tcp_batch_plan *plan, const void *input, void *output, int dtype, void *stream) {
    if (!plan) return fail(TCP_ERR_INVALID_ARGUMENT, "null plan");
    int st = check_buffers(plan, input, output, dtype);
    if (st) return st;
    tcp::DeviceGuard dg;
    if ((st = dg.enter(plan->device))) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    bool own_stream = false;
    if (!s) {  // capture needs a non-default stream
        TCP_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        own_stream = true;
    }
    if (plan->graph_exec) {
        (void)hipGraphExecDestroy(plan->graph_exec);
        plan->graph_exec = nullptr;
    }
    if (plan->graph) {
        (void)hipGraphDestroy(plan->graph);
        plan->graph = nullptr;
    }
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        st = plan->ops.empty() ? TCP_OK : enqueue(plan, plan->config, plan->split, input, output, dtype, s);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(s, &g);
        if (!st && e2 == hipSuccess) {
            plan->graph = g;
            e = hipGraphInstantiate(&plan->graph_exec, g, nullptr, nullptr, 0);
        } else if (e2 != hipSuccess) {
            e = e2;
        }
    }
    if (own_stream) (void)hipStreamDestroy(s);
    if (st) return st;
    if (e != hipSuccess) return tcp::hip_fail(e, "graph capture of the batched apply");
    return tcp::ok();
}

int tcp_batch_plan_replay(const tcp_batch_plan *plan, void *stream) {
    if (!plan || !plan->graph_exec) return fail(TCP_ERR_INVALID_ARGUMENT, "no captured graph");
    tcp::DeviceGuard dg;
    int st = dg.enter(plan->device);
    if (st) return st;
    TCP_HIP_CHECK(hipGraphLaunch(plan->graph_exec, static_cast<hipStream_t>(stream)));
    return tcp::ok();
}

void tcp_batch_plan_destroy(tcp_batch_plan *plan) { delete plan; }

}  // extern "C"
