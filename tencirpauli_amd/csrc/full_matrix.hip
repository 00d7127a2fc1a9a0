// Full-space CSR of a Pauli operator on the GPU (core/hamiltonian.rs:360-460,
// fill_sparse_row :725-747).  One thread per row.  The count pass evaluates
// every X group at column r ^ x and counts the nonzero values; the fill pass
// walks a binary trie over the ascending X masks, visiting at each branching
// bit b first the child whose masks share bit b of r (those columns have bit b
// clear), so entries leave the walk already in ascending column order.
#include <algorithm>
#include <numeric>

#include "full_matrix.hpp"

namespace {

using tcp::fail;

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

constexpr uint32_t kFullThreads = 256;
constexpr uint32_t kLeaf = 0x80000000u;
// Pending subtrees of one walk: at most one per branching bit above the
// current node plus the node itself; branching bits strictly decrease along a
// path, so n <= 32 qubits needs at most 33.
constexpr int kStack = 34;

struct FullArgs {
    uint64_t dim;
    uint32_t G;
    uint32_t root;
    const uint64_t *gx;
    const uint32_t *goff;
    const uint64_t *tz;
    const dvec2 *tw;
    const uint4 *nodes;
};

__device__ __forceinline__ double flip_sign(double x, uint32_t parity) {
    return __longlong_as_double(__double_as_longlong(x) ^ (static_cast<long long>(parity) << 63));
}

// Value of group g at `column`: the terms added to zero in canonical order,
// each negated when z & column has odd parity (core/hamiltonian.rs:728-736).
__device__ __forceinline__ dvec2 group_value(const FullArgs &a, uint32_t g, uint64_t column) {
    dvec2 v = {0.0, 0.0};
    const uint32_t end = a.goff[g + 1];
    for (uint32_t t = a.goff[g]; t < end; ++t) {
        const dvec2 w = a.tw[t];
        const uint32_t parity = static_cast<uint32_t>(__popcll(a.tz[t] & column)) & 1u;
        v.x += flip_sign(w.x, parity);
        v.y += flip_sign(w.y, parity);
    }
    return v;
}

__global__ __launch_bounds__(kFullThreads) void full_count_kernel(FullArgs a, uint32_t *__restrict__ counts) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kFullThreads) + threadIdx.x; r < a.dim;
         r += static_cast<uint64_t>(gridDim.x) * kFullThreads) {
        uint32_t c = 0;
        for (uint32_t g = 0; g < a.G; ++g) {
            const dvec2 v = group_value(a, g, r ^ a.gx[g]);
            c += (v.x != 0.0 || v.y != 0.0) ? 1u : 0u;
        }
        counts[r] = c;
    }
}

__global__ __launch_bounds__(kFullThreads) void full_fill_kernel(FullArgs a, const uint64_t *__restrict__ indptr,
                                                                 uint32_t *__restrict__ cols, dvec2 *__restrict__ vals,
                                                                 fvec2 *__restrict__ vals32) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kFullThreads) + threadIdx.x; r < a.dim;
         r += static_cast<uint64_t>(gridDim.x) * kFullThreads) {
        uint64_t pos = indptr[r];
        const uint64_t end = indptr[r + 1];
        uint32_t stack[kStack];
        int sp = 0;
        stack[sp++] = a.root;
        while (sp > 0) {
            const uint32_t code = stack[--sp];
            if (code & kLeaf) {
                const uint32_t g = code & ~kLeaf;
                const uint64_t column = r ^ a.gx[g];
                const dvec2 v = group_value(a, g, column);
                if ((v.x != 0.0 || v.y != 0.0) && pos < end) {
                    cols[pos] = static_cast<uint32_t>(column);
                    vals[pos] = v;
                    vals32[pos] = fvec2{static_cast<float>(v.x), static_cast<float>(v.y)};
                    ++pos;
                }
            } else {
                const uint4 node = a.nodes[code];  // {bit, left (bit clear), right (bit set), 0}
                const bool set = (r >> node.x) & 1u;
                // LIFO: push the later subtree first
                stack[sp++] = set ? node.y : node.z;
                stack[sp++] = set ? node.z : node.y;
            }
        }
    }
}

FullArgs args_of(const tcp::FullGroups &f) {
    FullArgs a{};
    a.dim = f.dim;
    a.G = f.G;
    a.root = f.root;
    a.gx = static_cast<const uint64_t *>(f.gx.ptr);
    a.goff = static_cast<const uint32_t *>(f.goff.ptr);
    a.tz = static_cast<const uint64_t *>(f.tz.ptr);
    a.tw = static_cast<const dvec2 *>(f.tw.ptr);
    a.nodes = static_cast<const uint4 *>(f.nodes.ptr);
    return a;
}

uint32_t grid_for(uint64_t dim) {
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((dim + kFullThreads - 1) / kFullThreads,
                                                                          256 * 32)));
}

// Trie over gx[lo, hi) (ascending, distinct): an internal node branches on the
// highest bit where its masks differ; returns the node index or kLeaf | group.
uint32_t build_trie(const std::vector<uint64_t> &gx, uint32_t lo, uint32_t hi, uint32_t depth,
                    std::vector<uint4> &nodes, uint32_t &max_depth) {
    max_depth = std::max(max_depth, depth);
    if (hi - lo == 1) return kLeaf | lo;
    const uint32_t bit = 63u - static_cast<uint32_t>(__builtin_clzll(gx[lo] ^ gx[hi - 1]));
    uint32_t split = lo;
    while (!((gx[split] >> bit) & 1u)) ++split;  // the masks share every bit above `bit`
    const uint32_t id = static_cast<uint32_t>(nodes.size());
    nodes.push_back(uint4{bit, 0u, 0u, 0u});
    const uint32_t left = build_trie(gx, lo, split, depth + 1, nodes, max_depth);
    const uint32_t right = build_trie(gx, split, hi, depth + 1, nodes, max_depth);
    nodes[id].y = left;
    nodes[id].z = right;
    return id;
}

}  // namespace

namespace tcp {

int full_groups_build(const Operator &op, FullGroups &f) {
    if (op.nqubits > kFullCsrMaxQubits)
        return fail(TCP_ERR_UNSUPPORTED, "full-space CSR supports at most 32 qubits (32-bit column indices)");
    const std::vector<IndexTerm> terms = index_terms(op);
    const uint64_t T = terms.size();
    if (T >= kLeaf) return fail(TCP_ERR_UNSUPPORTED, "too many terms for the full-space CSR");
    std::vector<uint32_t> order(T);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return terms[a].x < terms[b].x; });
    std::vector<uint64_t> gx, tz;
    std::vector<uint32_t> goff;
    std::vector<Complex> tw;
    for (uint64_t i = 0; i < T; ++i) {
        const IndexTerm &t = terms[order[i]];
        if (gx.empty() || gx.back() != t.x) {
            gx.push_back(t.x);
            goff.push_back(static_cast<uint32_t>(i));
        }
        tz.push_back(t.z);
        tw.push_back(t.w);
    }
    goff.push_back(static_cast<uint32_t>(T));
    f.n = op.nqubits;
    f.dim = 1ull << op.nqubits;
    f.G = static_cast<uint32_t>(gx.size());
    if (f.G == 0) return TCP_OK;
    std::vector<uint4> nodes;
    uint32_t max_depth = 0;
    f.root = build_trie(gx, 0, f.G, 0, nodes, max_depth);
    if (max_depth + 1 > static_cast<uint32_t>(kStack)) return fail(TCP_ERR_UNSUPPORTED, "full-space CSR trie too deep");
    int st;
    if ((st = f.gx.upload(gx.data(), 8 * gx.size()))) return st;
    if ((st = f.goff.upload(goff.data(), 4 * goff.size()))) return st;
    if ((st = f.tz.upload(tz.data(), 8 * tz.size()))) return st;
    if ((st = f.tw.upload(tw.data(), 16 * tw.size()))) return st;
    if (!nodes.empty() && (st = f.nodes.upload(nodes.data(), sizeof(uint4) * nodes.size()))) return st;
    return TCP_OK;
}

int full_csr_counts(const FullGroups &f, uint32_t *counts, hipStream_t stream) {
    if (f.dim == 0) return TCP_OK;
    if (f.G == 0) {
        TCP_HIP_CHECK(hipMemsetAsync(counts, 0, 4 * f.dim, stream));
        return TCP_OK;
    }
    hipLaunchKernelGGL(full_count_kernel, dim3(grid_for(f.dim)), dim3(kFullThreads), 0, stream, args_of(f), counts);
    TCP_HIP_CHECK(hipGetLastError());
    return TCP_OK;
}

int full_csr_fill(const FullGroups &f, const uint64_t *indptr, uint32_t *cols, void *vals64, void *vals32,
                  hipStream_t stream) {
    if (f.dim == 0 || f.G == 0) return TCP_OK;
    hipLaunchKernelGGL(full_fill_kernel, dim3(grid_for(f.dim)), dim3(kFullThreads), 0, stream, args_of(f), indptr,
                       cols, static_cast<dvec2 *>(vals64), static_cast<fvec2 *>(vals32));
    TCP_HIP_CHECK(hipGetLastError());
    return TCP_OK;
}

}  // namespace tcp
