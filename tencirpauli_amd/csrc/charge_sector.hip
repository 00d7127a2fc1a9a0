// General additive-charge sectors: host suffix-count tables and the lazy /
// eager restricted-MVP kernels on them (charge_sector.hpp describes the
// order and the tables).  Same semantics as the block-sector plan: restricted
// y = P H P^T v gathered per destination, x-groups in first-appearance order,
// exact-zero aggregated values skipped, and the reference's leakage error
// raised when its serial loop would raise it (core/charge/generic.rs:949-1045).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <unordered_map>

#include "charge_sector.hpp"

namespace {

using tcp::fail;

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

constexpr uint32_t kChargeThreads = 256;
constexpr uint32_t kChargeLdsLimit = 64 * 1024;  // count tables up to this size are copied to LDS

struct ChargeArgs {
    uint32_t n;
    uint32_t m;
    uint32_t G;
    uint64_t dim;
    uint64_t box;
    int64_t enc_target, enc_zero;
    const uint32_t *count;   // [(n + 1) * box]
    const int64_t *enc;      // [n]
    const int64_t *wt;       // [n * m] weight of axis a in constraint j
    int64_t target[tcp::kChargeMaxConstraints];
    int64_t lo[tcp::kChargeMaxConstraints], hi[tcp::kChargeMaxConstraints];
    const uint64_t *gx;      // [G]
    const uint32_t *goff;    // [G + 1]
    const uint64_t *tz;      // [T]
    const void *tw;          // [T] (re, im)
    const uint64_t *tc1;     // [T] fermion words: modes that must be occupied in the source
    const uint64_t *tc0;     // [T] fermion words: modes that must be empty in the source
    const int32_t *enc32;    // [n] the encoded weights as int32 (narrow sectors), or null
};

__device__ __forceinline__ double flip_sign(double x, uint32_t parity) {
    return __longlong_as_double(__double_as_longlong(x) ^ (static_cast<long long>(parity) << 63));
}

// Rank of bitstring s (axis a = bit n-1-a), or UINT64_MAX when s is outside
// the sector: one table lookup per set bit, axis order.
__host__ __device__ __forceinline__ uint64_t gen_rank(uint32_t n, uint64_t box, int64_t enc_target, int64_t enc_zero,
                                                      const uint32_t *count, const int64_t *enc, uint64_t s) {
    int64_t idx = enc_target;
    uint64_t rank = 0;
    while (s) {
        const uint32_t bit = 63u - static_cast<uint32_t>(__builtin_clzll(s));
        const uint32_t a = n - 1 - bit;
        rank += count[static_cast<uint64_t>(a + 1) * box + static_cast<uint64_t>(idx)];
        idx -= enc[a];
        s &= ~(1ull << bit);
    }
    return idx == enc_zero ? rank : UINT64_MAX;
}

// Rank of an arbitrary bitstring s: the walk of gen_rank, returning
// UINT64_MAX as soon as a prefix remainder leaves the box (no completion).
__host__ __device__ __forceinline__ uint64_t gen_rank_checked(uint32_t n, uint32_t m, uint64_t box, int64_t enc_target,
                                                              int64_t enc_zero, const uint32_t *count,
                                                              const int64_t *enc, const int64_t *wt,
                                                              const int64_t *target, const int64_t *lo,
                                                              const int64_t *hi, uint64_t s) {
    int64_t R[tcp::kChargeMaxConstraints];
    for (uint32_t j = 0; j < m; ++j) {
        R[j] = target[j];
        if (R[j] < lo[j] || R[j] > hi[j]) return UINT64_MAX;  // empty sector
    }
    int64_t idx = enc_target;
    uint64_t rank = 0;
    while (s) {
        const uint32_t bit = 63u - static_cast<uint32_t>(__builtin_clzll(s));
        const uint32_t a = n - 1 - bit;
        rank += count[static_cast<uint64_t>(a + 1) * box + static_cast<uint64_t>(idx)];
        idx -= enc[a];
        for (uint32_t j = 0; j < m; ++j) {
            R[j] -= wt[static_cast<uint64_t>(a) * m + j];
            if (R[j] < lo[j] || R[j] > hi[j]) return UINT64_MAX;
        }
        s &= ~(1ull << bit);
    }
    return idx == enc_zero ? rank : UINT64_MAX;
}

// State of rank r < dim.
__host__ __device__ __forceinline__ uint64_t gen_unrank(uint32_t n, uint64_t box, int64_t enc_target,
                                                        const uint32_t *count, const int64_t *enc, uint64_t r) {
    int64_t idx = enc_target;
    uint64_t bits = 0;
    for (uint32_t a = 0; a < n; ++a) {
        const uint64_t c0 = count[static_cast<uint64_t>(a + 1) * box + static_cast<uint64_t>(idx)];
        if (r >= c0) {
            r -= c0;
            bits |= 1ull << (n - 1 - a);
            idx -= enc[a];
        }
    }
    return bits;
}

// gen_rank / gen_unrank in 32-bit arithmetic (every encoded value fits in
// int32; ranks are below 2^32): the walk of the restricted gather, with the
// count table and the weights in LDS.
__device__ __forceinline__ uint32_t gen_rank32(uint32_t n, uint32_t box, int32_t idx, int32_t enc_zero,
                                               const uint32_t *count, const int32_t *enc, uint64_t s) {
    uint32_t rank = 0;
    while (s) {
        const uint32_t bit = 63u - static_cast<uint32_t>(__builtin_clzll(s));
        const uint32_t a = n - 1 - bit;
        rank += count[(a + 1) * box + static_cast<uint32_t>(idx)];
        idx -= enc[a];
        s &= ~(1ull << bit);
    }
    return idx == enc_zero ? rank : UINT32_MAX;
}

__device__ __forceinline__ uint64_t gen_unrank32(uint32_t n, uint32_t box, int32_t idx, const uint32_t *count,
                                                 const int32_t *enc, uint32_t r) {
    uint64_t bits = 0;
    for (uint32_t a = 0; a < n; ++a) {
        const uint32_t c0 = count[(a + 1) * box + static_cast<uint32_t>(idx)];
        if (r >= c0) {
            r -= c0;
            bits |= 1ull << (n - 1 - a);
            idx -= enc[a];
        }
    }
    return bits;
}

// Whether b ^ x keeps every charge of the in-sector state b: per constraint,
// the weights of the bits x flips, + where b has the bit clear and - where
// set, sum to zero.  (x is the same in every lane: a scalar loop over its
// bits.)  Then every prefix of b ^ x stays in the box and gen_rank needs no
// checks.
__device__ __forceinline__ bool keeps_charges(const ChargeArgs &p, uint64_t b, uint64_t x) {
    int64_t d[tcp::kChargeMaxConstraints] = {0, 0, 0, 0};
    for (uint64_t r = x; r; r &= r - 1) {
        const uint32_t bit = static_cast<uint32_t>(__builtin_ctzll(r));
        const uint32_t a = p.n - 1 - bit;
        const bool set = (b >> bit) & 1ull;
#pragma unroll
        for (uint32_t j = 0; j < tcp::kChargeMaxConstraints; ++j)
            if (j < p.m) {
                const int64_t w = p.wt[static_cast<uint64_t>(a) * p.m + j];
                d[j] += set ? -w : w;
            }
    }
    bool ok = true;
#pragma unroll
    for (uint32_t j = 0; j < tcp::kChargeMaxConstraints; ++j) ok &= j >= p.m || d[j] == 0;
    return ok;
}

template <bool COND>
__device__ __forceinline__ bool acts(const ChargeArgs &p, uint32_t k, uint64_t s) {
    if (!COND) return true;
    const uint64_t need = p.tc1[k];
    return (s & need) == need && (s & p.tc0[k]) == 0;
}

// Sum of group g's terms at source s: sum_k w_k (-1)^popc(z_k & s) over the
// terms that act on s (core/sector.rs:1051-1099, core/charge/schedule.rs:645-726).
template <bool COND>
__device__ __forceinline__ dvec2 group_value(const ChargeArgs &p, uint32_t g, uint64_t s) {
    const dvec2 *tw = static_cast<const dvec2 *>(p.tw);
    double dr = 0.0, di = 0.0;
    for (uint32_t k = p.goff[g]; k < p.goff[g + 1]; ++k) {
        if (!acts<COND>(p, k, s)) continue;
        const uint32_t par = __popcll(p.tz[k] & s) & 1u;
        dr += flip_sign(tw[k].x, par);
        di += flip_sign(tw[k].y, par);
    }
    return dvec2{dr, di};
}

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
    __device__ static fvec2 make(double r, double i) { return fvec2{static_cast<float>(r), static_cast<float>(i)}; }
};

// The count table from global memory, or its LDS copy when LDS is set; with
// narrow weights, their int32 copy (in LDS after the table when LDS is set).
template <bool LDS>
__device__ __forceinline__ const uint32_t *stage_counts(const ChargeArgs &p, uint32_t *lds, const int32_t *&enc32) {
    enc32 = p.enc32;
    if (!LDS) return p.count;
    const uint64_t words = (p.n + 1) * p.box;
    for (uint64_t i = threadIdx.x; i < words; i += blockDim.x) lds[i] = p.count[i];
    if (p.enc32) {
        int32_t *le = reinterpret_cast<int32_t *>(lds + words);
        for (uint32_t i = threadIdx.x; i < p.n; i += blockDim.x) le[i] = p.enc32[i];
        enc32 = le;
    }
    __syncthreads();
    return lds;
}

// One destination rank per thread: unrank, then for every group the source
// s = b ^ x, its rank (UINT64_MAX outside the sector), the group's value at s
// and the gather.  FP64 accumulation; the output is written once.
template <typename VT, bool COND, bool LDS, bool N32>
__global__ __launch_bounds__(kChargeThreads) void charge_gather_kernel(const VT *__restrict__ v, VT *__restrict__ y,
                                                                       ChargeArgs p) {
    extern __shared__ uint32_t lds_count[];
    const int32_t *enc32 = nullptr;
    const uint32_t *count = stage_counts<LDS>(p, lds_count, enc32);
    const uint32_t box32 = static_cast<uint32_t>(p.box);
    const int32_t et32 = static_cast<int32_t>(p.enc_target), ez32 = static_cast<int32_t>(p.enc_zero);
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < p.dim;
         r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t b = N32 ? gen_unrank32(p.n, box32, et32, count, enc32, static_cast<uint32_t>(r))
                               : gen_unrank(p.n, p.box, p.enc_target, count, p.enc, r);
        double ar = 0.0, ai = 0.0;
        for (uint32_t g = 0; g < p.G; ++g) {
            const uint64_t x = p.gx[g];
            if (x != 0 && !keeps_charges(p, b, x)) continue;
            const uint64_t s = b ^ x;
            const dvec2 w = group_value<COND>(p, g, s);
            if (w.x == 0.0 && w.y == 0.0) continue;
            uint64_t rs = r;
            if (x != 0) {
                if (N32) {
                    const uint32_t r32 = gen_rank32(p.n, box32, et32, ez32, count, enc32, s);
                    rs = r32 == UINT32_MAX ? UINT64_MAX : r32;
                } else {
                    rs = gen_rank(p.n, p.box, p.enc_target, p.enc_zero, count, p.enc, s);
                }
            }
            if (rs >= p.dim) continue;  // unreachable for an in-sector s; keeps the gather in bounds
            const VT sv = v[rs];
            const double sr = Amp<VT>::re(sv), si = Amp<VT>::im(sv);
            ar = fma(w.x, sr, ar);
            ar = fma(-w.y, si, ar);
            ai = fma(w.x, si, ai);
            ai = fma(w.y, sr, ai);
        }
        y[r] = Amp<VT>::make(ar, ai);
    }
}

// Leakage scan in the reference's scatter order: the smallest
// rank * G + group whose destination leaves the sector with a value that is
// not an exact zero (the first error the serial loop raises).
template <bool COND>
__global__ __launch_bounds__(kChargeThreads) void charge_leak_kernel(ChargeArgs p, unsigned long long *first) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < p.dim;
         r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t b = gen_unrank(p.n, p.box, p.enc_target, p.count, p.enc, r);
        for (uint32_t g = 0; g < p.G; ++g) {
            if (keeps_charges(p, b, p.gx[g])) continue;  // the destination b ^ x is in the sector
            const dvec2 w = group_value<COND>(p, g, b);
            if (w.x != 0.0 || w.y != 0.0) {
                atomicMin(first, static_cast<unsigned long long>(r) * p.G + g);
                break;
            }
        }
    }
}

template <bool COND>
__global__ void charge_diag_kernel(ChargeArgs p, uint32_t diag_group, double *__restrict__ diag) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < p.dim;
         r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t b = gen_unrank(p.n, p.box, p.enc_target, p.count, p.enc, r);
        diag[r] = diag_group < p.G ? group_value<COND>(p, diag_group, b).x : 0.0;
    }
}

__global__ void charge_rank_kernel(ChargeArgs p, const uint64_t *__restrict__ bits, uint64_t *__restrict__ ranks,
                                   uint64_t count) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t s = bits[i];
        const bool inside = p.n == 64 || (s >> p.n) == 0;
        ranks[i] = inside ? gen_rank_checked(p.n, p.m, p.box, p.enc_target, p.enc_zero, p.count, p.enc, p.wt,
                                             p.target, p.lo, p.hi, s)
                          : UINT64_MAX;
    }
}

__global__ void charge_unrank_kernel(ChargeArgs p, const uint64_t *__restrict__ ranks, uint64_t *__restrict__ bits,
                                     uint64_t count, uint64_t first, int walk) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t r = walk ? first + i : ranks[i];
        bits[i] = r < p.dim ? gen_unrank(p.n, p.box, p.enc_target, p.count, p.enc, r) : UINT64_MAX;
    }
}

// Eager rows: row r holds (rank(s), value) for every group whose source
// s = b ^ x is in the sector with a nonzero value, columns ascending.
template <bool COND>
__global__ __launch_bounds__(kChargeThreads) void charge_count_kernel(ChargeArgs p, uint32_t *__restrict__ counts) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < p.dim;
         r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t b = gen_unrank(p.n, p.box, p.enc_target, p.count, p.enc, r);
        uint32_t c = 0;
        for (uint32_t g = 0; g < p.G; ++g) {
            const uint64_t x = p.gx[g];
            if (x != 0 && !keeps_charges(p, b, x)) continue;
            const dvec2 w = group_value<COND>(p, g, b ^ x);
            c += (w.x != 0.0 || w.y != 0.0);
        }
        counts[r] = c;
    }
}

template <bool COND>
__global__ __launch_bounds__(kChargeThreads) void charge_fill_kernel(ChargeArgs p, const uint64_t *__restrict__ indptr,
                                                                     uint32_t *__restrict__ cols,
                                                                     dvec2 *__restrict__ vals,
                                                                     fvec2 *__restrict__ vals32) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < p.dim;
         r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t b = gen_unrank(p.n, p.box, p.enc_target, p.count, p.enc, r);
        const uint64_t pos = indptr[r];
        const uint64_t len = indptr[r + 1] - pos;
        uint32_t k = 0;
        for (uint32_t g = 0; g < p.G && k < len; ++g) {
            const uint64_t x = p.gx[g];
            if (x != 0 && !keeps_charges(p, b, x)) continue;
            const uint64_t s = b ^ x;
            const dvec2 w = group_value<COND>(p, g, s);
            if (w.x == 0.0 && w.y == 0.0) continue;
            const uint64_t rs = x == 0 ? r : gen_rank(p.n, p.box, p.enc_target, p.enc_zero, p.count, p.enc, s);
            // insertion into the row's sorted prefix (rows hold at most G entries)
            uint32_t i = k;
            while (i > 0 && cols[pos + i - 1] > rs) {
                cols[pos + i] = cols[pos + i - 1];
                vals[pos + i] = vals[pos + i - 1];
                --i;
            }
            cols[pos + i] = static_cast<uint32_t>(rs);
            vals[pos + i] = w;
            ++k;
        }
        for (uint32_t q = 0; q < k; ++q) {
            const dvec2 w = vals[pos + q];
            vals32[pos + q] = fvec2{static_cast<float>(w.x), static_cast<float>(w.y)};
        }
    }
}

ChargeArgs args_of(const tcp::ChargeSector &c, int dtype) {
    ChargeArgs a{};
    a.n = c.geo.n;
    a.G = c.G;
    a.dim = c.geo.dim;
    a.box = c.geo.box;
    a.enc_target = c.geo.enc_target;
    a.enc_zero = c.geo.enc_zero;
    a.count = static_cast<const uint32_t *>(c.count.ptr);
    a.enc = static_cast<const int64_t *>(c.enc.ptr);
    a.m = c.geo.m;
    a.wt = static_cast<const int64_t *>(c.wt.ptr);
    for (uint32_t j = 0; j < tcp::kChargeMaxConstraints; ++j) {
        a.target[j] = c.geo.target[j];
        a.lo[j] = c.geo.lo[j];
        a.hi[j] = c.geo.hi[j];
    }
    a.gx = static_cast<const uint64_t *>(c.gx.ptr);
    a.goff = static_cast<const uint32_t *>(c.goff.ptr);
    a.tz = static_cast<const uint64_t *>(c.tz.ptr);
    a.tw = dtype == TCP_C64 ? c.tw32.ptr : c.tw64.ptr;
    a.tc1 = c.cond ? static_cast<const uint64_t *>(c.tc1.ptr) : nullptr;
    a.tc0 = c.cond ? static_cast<const uint64_t *>(c.tc0.ptr) : nullptr;
    a.enc32 = c.narrow ? static_cast<const int32_t *>(c.enc32.ptr) : nullptr;
    return a;
}

uint32_t grid_for(uint64_t dim) {
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((dim + kChargeThreads - 1) / kChargeThreads,
                                                                           256 * 16)));
}

}  // namespace

namespace tcp {

int charge_geometry(uint32_t n, uint32_t m, const int64_t *levels, const int64_t *targets, uint64_t max_bytes,
                    ChargeGeometry &g) {
    if (m == 0) return fail(TCP_ERR_INVALID_ARGUMENT, "a charge sector needs at least one constraint");
    if (m > kChargeMaxConstraints)
        return fail(TCP_ERR_UNSUPPORTED, "GPU charge sectors support at most 4 simultaneous constraints");
    if (n > kChargeMaxAxes) return fail(TCP_ERR_UNSUPPORTED, "GPU charge sectors support at most 63 binary axes");
    if (!levels || !targets) return fail(TCP_ERR_INVALID_ARGUMENT, "null charge levels or targets");
    const int64_t kLimit = int64_t(1) << 40;  // keeps every sum below 2^46
    g = ChargeGeometry{};
    g.n = n;
    g.m = m;
    g.weight.assign(static_cast<size_t>(n) * m, 0);
    for (uint32_t j = 0; j < m; ++j) {
        int64_t t = targets[j];
        if (t > kLimit || t < -kLimit) return fail(TCP_ERR_OVERFLOW, "charge target is too large");
        for (uint32_t a = 0; a < n; ++a) {
            const int64_t l0 = levels[(static_cast<size_t>(j) * n + a) * 2];
            const int64_t l1 = levels[(static_cast<size_t>(j) * n + a) * 2 + 1];
            if (l0 > kLimit || l0 < -kLimit || l1 > kLimit || l1 < -kLimit)
                return fail(TCP_ERR_OVERFLOW, "charge level is too large");
            t -= l0;
            g.weight[static_cast<size_t>(a) * m + j] = l1 - l0;
        }
        g.target[j] = t;
    }
    // remainder box per constraint and its mixed-radix encoding
    unsigned __int128 box = 1;
    for (uint32_t j = 0; j < m; ++j) {
        int64_t pos = 0, neg = 0;
        for (uint32_t a = 0; a < n; ++a) {
            const int64_t w = g.weight[static_cast<size_t>(a) * m + j];
            (w > 0 ? pos : neg) += w;
        }
        g.lo[j] = std::max(g.target[j] - pos, neg);
        g.hi[j] = std::min(g.target[j] - neg, pos);
        if (g.hi[j] < g.lo[j]) g.hi[j] = g.lo[j];  // an empty sector (its target is out of reach)
        g.stride[j] = static_cast<uint64_t>(box);
        box *= static_cast<unsigned __int128>(g.hi[j] - g.lo[j] + 1);
        if (box * (n + 1) > kChargeMaxTableEntries)
            return fail(TCP_ERR_UNSUPPORTED, "charge sector rank tables would exceed 2^24 entries (charge levels "
                                             "too large for the GPU suffix-count tables)");
    }
    g.box = static_cast<uint64_t>(box);
    const uint64_t table_bytes = 4ull * (n + 1) * g.box;
    if (exceeds_budget(table_bytes, max_bytes)) return memory_limit(table_bytes, max_bytes);
    auto encode = [&](const int64_t *r) {
        int64_t e = 0;
        for (uint32_t j = 0; j < m; ++j) e += (r[j] - g.lo[j]) * static_cast<int64_t>(g.stride[j]);
        return e;
    };
    g.enc.assign(n, 0);
    for (uint32_t a = 0; a < n; ++a) {
        int64_t e = 0;
        for (uint32_t j = 0; j < m; ++j) e += g.weight[static_cast<size_t>(a) * m + j] * static_cast<int64_t>(g.stride[j]);
        g.enc[a] = e;
    }
    g.enc_target = encode(g.target);
    const int64_t zero[kChargeMaxConstraints] = {0, 0, 0, 0};
    g.enc_zero = encode(zero);
    // count[p][R] over the box, p = n .. 0 (u64 while building, checked < 2^32)
    std::vector<uint64_t> cur(g.box, 0), next(g.box, 0);
    g.count.assign(static_cast<size_t>(n + 1) * g.box, 0);
    // a remainder of zero lies in the box (it is the target minus every weight's contribution, or reachable)
    auto in_box = [&](const int64_t *r) {
        for (uint32_t j = 0; j < m; ++j) {
            const uint64_t extent = (j + 1 < m ? g.stride[j + 1] : g.box) / g.stride[j];
            if (r[j] < g.lo[j] || r[j] >= g.lo[j] + static_cast<int64_t>(extent)) return false;
        }
        return true;
    };
    if (in_box(zero)) next[static_cast<uint64_t>(g.enc_zero)] = 1;
    for (uint64_t i = 0; i < g.box; ++i) g.count[static_cast<size_t>(n) * g.box + i] = static_cast<uint32_t>(next[i]);
    for (int a = static_cast<int>(n) - 1; a >= 0; --a) {
        // count[a][R] = count[a+1][R] + count[a+1][R - w_a] (out of the box: none)
        std::vector<int64_t> coord(m, 0);
        for (uint64_t i = 0; i < g.box; ++i) {
            uint64_t rest = i;
            bool ok = true;
            int64_t shifted = 0;
            for (int j = static_cast<int>(m) - 1; j >= 0; --j) {
                const uint64_t c = rest / g.stride[j];
                rest -= c * g.stride[j];
                const int64_t r = g.lo[j] + static_cast<int64_t>(c);
                const int64_t rm = r - g.weight[static_cast<size_t>(a) * m + j];
                const uint64_t extent = (static_cast<uint32_t>(j) + 1 < m ? g.stride[j + 1] : g.box) / g.stride[j];
                if (rm < g.lo[j] || rm >= g.lo[j] + static_cast<int64_t>(extent)) ok = false;
                shifted += (rm - g.lo[j]) * static_cast<int64_t>(g.stride[j]);
            }
            uint64_t v = next[i];
            if (ok) v += next[static_cast<uint64_t>(shifted)];
            cur[i] = v;
        }
        for (uint64_t i = 0; i < g.box; ++i) {
            if (cur[i] >= (1ull << 32))
                return fail(TCP_ERR_UNSUPPORTED, "GPU sector kernels support dimensions below 2^32");
            g.count[static_cast<size_t>(a) * g.box + i] = static_cast<uint32_t>(cur[i]);
        }
        std::swap(cur, next);
    }
    bool reachable = true;
    for (uint32_t j = 0; j < m; ++j) reachable &= g.target[j] >= g.lo[j] && g.target[j] <= g.hi[j];
    g.dim = reachable ? g.count[static_cast<uint64_t>(g.enc_target)] : 0;
    return TCP_OK;
}

int charge_rank(const ChargeGeometry &g, uint64_t bits, uint64_t &rank) {
    if (g.n < 64 && (bits >> g.n))
        return fail(TCP_ERR_INVALID_INDEX, "occupation is outside the finite sector layout");
    rank = g.dim == 0 ? UINT64_MAX
                      : gen_rank_checked(g.n, g.m, g.box, g.enc_target, g.enc_zero, g.count.data(), g.enc.data(),
                                         g.weight.data(), g.target, g.lo, g.hi, bits);
    if (rank == UINT64_MAX)
        return fail(TCP_ERR_INVALID_SECTOR, "invalid sector: occupation does not satisfy every charge constraint");
    return TCP_OK;
}

int charge_unrank(const ChargeGeometry &g, uint64_t rank, uint64_t &bits) {
    if (rank >= g.dim) return fail(TCP_ERR_INVALID_INDEX, "invalid index: sector index is out of range");
    bits = gen_unrank(g.n, g.box, g.enc_target, g.count.data(), g.enc.data(), rank);
    return TCP_OK;
}

int charge_build(const Operator *op, const std::vector<CondTerm> *fermion_terms, const ChargeGeometry &geo,
                 uint64_t max_bytes, bool require_conserved, ChargeSector &out) {
    const uint32_t n = geo.n, m = geo.m;
    std::vector<CondTerm> terms;
    if (fermion_terms) {
        terms = *fermion_terms;
    } else {
        if (op->nqubits != n)
            return fail(TCP_ERR_INCOMPATIBLE_QUBIT_COUNTS,
                        "incompatible qubit counts: " + std::to_string(op->nqubits) + " and " + std::to_string(n));
        for (const IndexTerm &t : index_terms(*op)) terms.push_back(CondTerm{t.x, t.z, 0, 0, t.w});
    }
    if (require_conserved) {
        for (uint32_t j = 0; j < m; ++j) {
            bool conserved = true;
            if (!fermion_terms) {
                std::vector<double> levels(2 * n, 0.0);
                for (uint32_t a = 0; a < n; ++a) levels[2 * a + 1] = static_cast<double>(geo.weight[a * m + j]);
                uint64_t nonzero = 0;
                int st = analyze_charge(*op, levels.data(), conserved, nonzero);
                if (st) return st;
            } else {
                // a ladder word fills the modes of c0 & x and empties those of c1 & x
                for (const CondTerm &t : terms) {
                    int64_t change = 0;
                    for (uint32_t a = 0; a < n; ++a) {
                        const uint64_t bit = 1ull << (n - 1 - a);
                        if (t.c0 & t.x & bit) change += geo.weight[a * m + j];
                        if (t.c1 & t.x & bit) change -= geo.weight[a * m + j];
                    }
                    conserved &= change == 0;
                }
            }
            if (!conserved)
                return fail(TCP_ERR_SECTOR_LEAKAGE, "selected charge sector requires an exactly conserved operator");
        }
    }
    out.geo = geo;
    out.cond = fermion_terms != nullptr;
    // x-mask groups in first-appearance order; a group's terms keep canonical order
    std::vector<uint64_t> gx;
    std::vector<std::vector<size_t>> members;
    std::unordered_map<uint64_t, size_t> group_of;
    for (size_t k = 0; k < terms.size(); ++k) {
        auto it = group_of.emplace(terms[k].x, gx.size()).first;
        if (it->second == gx.size()) {
            gx.push_back(terms[k].x);
            members.emplace_back();
        }
        members[it->second].push_back(k);
    }
    std::vector<uint32_t> goff;
    std::vector<uint64_t> tz, tc1, tc0;
    std::vector<dvec2> tw64, tw32;
    double scale = 0.0;
    for (size_t g = 0; g < gx.size(); ++g) {
        goff.push_back(static_cast<uint32_t>(tz.size()));
        double sum = 0.0;
        for (size_t k : members[g]) {
            tz.push_back(terms[k].z);
            tc1.push_back(terms[k].c1);
            tc0.push_back(terms[k].c0);
            tw64.push_back(dvec2{terms[k].w.re, terms[k].w.im});
            tw32.push_back(dvec2{static_cast<double>(static_cast<float>(terms[k].w.re)),
                                 static_cast<double>(static_cast<float>(terms[k].w.im))});
            sum += std::hypot(terms[k].w.re, terms[k].w.im);
        }
        if (gx[g] == 0) out.diag_group = static_cast<uint32_t>(g);
        else scale = std::max(scale, sum);
    }
    goff.push_back(static_cast<uint32_t>(tz.size()));
    out.G = static_cast<uint32_t>(gx.size());
    out.T = static_cast<uint32_t>(tz.size());
    out.energy_scale = scale > 0.0 ? scale : 1.0;
    if (gx.empty()) gx.push_back(0);
    if (tz.empty()) {
        tz.push_back(0);
        tc1.push_back(0);
        tc0.push_back(0);
        tw64.push_back(dvec2{0, 0});
        tw32.push_back(dvec2{0, 0});
    }
    out.retained_bytes = gx.size() * 8 + goff.size() * 4 + tz.size() * (out.cond ? 8 + 32 + 16 : 8 + 32) +
                         geo.count.size() * 4 + geo.enc.size() * 8 + geo.weight.size() * 8;
    if (exceeds_budget(out.retained_bytes, max_bytes)) return memory_limit(out.retained_bytes, max_bytes);
    int st;
    if ((st = out.gx.upload(gx.data(), gx.size() * 8)) || (st = out.goff.upload(goff.data(), goff.size() * 4)) ||
        (st = out.tz.upload(tz.data(), tz.size() * 8)) || (st = out.tw64.upload(tw64.data(), tw64.size() * 16)) ||
        (st = out.tw32.upload(tw32.data(), tw32.size() * 16)) ||
        (st = out.count.upload(geo.count.data(), geo.count.size() * 4)) ||
        (st = out.enc.upload(geo.enc.data(), std::max<size_t>(1, geo.enc.size()) * 8)) ||
        (st = out.wt.upload(geo.weight.data(), std::max<size_t>(1, geo.weight.size()) * 8)))
        return st;
    if (out.cond && ((st = out.tc1.upload(tc1.data(), tc1.size() * 8)) ||
                     (st = out.tc0.upload(tc0.data(), tc0.size() * 8))))
        return st;
    // 32-bit walk: every encoded weight, the encoded target and zero, and the
    // table's row offsets fit in 32 bits
    const int64_t lim = int64_t(1) << 30;
    bool narrow = std::llabs(geo.enc_target) < lim && std::llabs(geo.enc_zero) < lim && geo.box < (1ull << 31) &&
                  (static_cast<uint64_t>(n) + 1) * geo.box < (1ull << 32);
    std::vector<int32_t> e32(std::max<size_t>(1, geo.enc.size()), 0);
    for (size_t a = 0; a < geo.enc.size(); ++a) {
        narrow = narrow && std::llabs(geo.enc[a]) < lim;
        e32[a] = static_cast<int32_t>(geo.enc[a]);
    }
    if (narrow && (st = out.enc32.upload(e32.data(), 4 * e32.size()))) return st;
    out.narrow = narrow;
    // leakage scan (state-independent, once per plan)
    if (out.G == 0 || geo.dim == 0) return TCP_OK;
    unsigned long long *dfirst = nullptr;
    TCP_HIP_CHECK(hipMalloc(&dfirst, sizeof(unsigned long long)));
    const unsigned long long none = ~0ull;
    hipError_t e = hipMemcpy(dfirst, &none, sizeof none, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const ChargeArgs a = args_of(out, TCP_C128);
        if (out.cond)
            hipLaunchKernelGGL(charge_leak_kernel<true>, dim3(grid_for(geo.dim)), dim3(kChargeThreads), 0, nullptr, a,
                               dfirst);
        else
            hipLaunchKernelGGL(charge_leak_kernel<false>, dim3(grid_for(geo.dim)), dim3(kChargeThreads), 0, nullptr,
                               a, dfirst);
        e = hipGetLastError();
    }
    unsigned long long first = none;
    if (e == hipSuccess) e = hipMemcpy(&first, dfirst, sizeof first, hipMemcpyDeviceToHost);
    (void)hipFree(dfirst);
    if (e != hipSuccess) return hip_fail(e, "charge sector leakage scan");
    if (first != none) {
        out.leaks = true;
        out.leak_source = first / out.G;
    }
    return TCP_OK;
}

int charge_apply(const ChargeSector &c, const void *in, void *out, int dtype, hipStream_t stream) {
    if (c.geo.dim == 0) return TCP_OK;
    const ChargeArgs a = args_of(c, dtype);
    const uint64_t table = 4ull * (c.geo.n + 1) * c.geo.box + (c.narrow ? 4ull * c.geo.n : 0);
    const bool lds = table <= kChargeLdsLimit;
    const uint32_t lds_bytes = static_cast<uint32_t>(table);
    const uint32_t grid = grid_for(c.geo.dim);
    auto launch = [&](auto vt, auto cond) {
        using VT = typename decltype(vt)::type;
        constexpr bool COND = decltype(cond)::value;
        if (lds && c.narrow)
            hipLaunchKernelGGL((charge_gather_kernel<VT, COND, true, true>), dim3(grid), dim3(kChargeThreads),
                               lds_bytes, stream, static_cast<const VT *>(in), static_cast<VT *>(out), a);
        else if (lds)
            hipLaunchKernelGGL((charge_gather_kernel<VT, COND, true, false>), dim3(grid), dim3(kChargeThreads),
                               lds_bytes, stream, static_cast<const VT *>(in), static_cast<VT *>(out), a);
        else if (c.narrow)
            hipLaunchKernelGGL((charge_gather_kernel<VT, COND, false, true>), dim3(grid), dim3(kChargeThreads), 0,
                               stream, static_cast<const VT *>(in), static_cast<VT *>(out), a);
        else
            hipLaunchKernelGGL((charge_gather_kernel<VT, COND, false, false>), dim3(grid), dim3(kChargeThreads), 0,
                               stream, static_cast<const VT *>(in), static_cast<VT *>(out), a);
    };
    struct D {
        using type = dvec2;
    };
    struct F {
        using type = fvec2;
    };
    if (dtype == TCP_C128) c.cond ? launch(D{}, std::true_type()) : launch(D{}, std::false_type());
    else c.cond ? launch(F{}, std::true_type()) : launch(F{}, std::false_type());
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "charge sector kernel launch");
}

int charge_diagonal(const ChargeSector &c, double *diag, hipStream_t stream) {
    if (c.geo.dim == 0) return TCP_OK;
    const ChargeArgs a = args_of(c, TCP_C128);
    if (c.cond)
        hipLaunchKernelGGL(charge_diag_kernel<true>, dim3(grid_for(c.geo.dim)), dim3(kChargeThreads), 0, stream, a,
                           c.diag_group, diag);
    else
        hipLaunchKernelGGL(charge_diag_kernel<false>, dim3(grid_for(c.geo.dim)), dim3(kChargeThreads), 0, stream, a,
                           c.diag_group, diag);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "charge sector diagonal kernel");
}

int charge_rank_device(const ChargeSector &c, const uint64_t *bits, uint64_t *ranks, uint64_t count,
                       hipStream_t stream) {
    hipLaunchKernelGGL(charge_rank_kernel, dim3(grid_for(count)), dim3(kChargeThreads), 0, stream,
                       args_of(c, TCP_C128), bits, ranks, count);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "charge sector rank kernel");
}

int charge_unrank_device(const ChargeSector &c, const uint64_t *ranks, uint64_t *bits, uint64_t count,
                         hipStream_t stream) {
    hipLaunchKernelGGL(charge_unrank_kernel, dim3(grid_for(count)), dim3(kChargeThreads), 0, stream,
                       args_of(c, TCP_C128), ranks, bits, count, uint64_t(0), 0);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "charge sector unrank kernel");
}

int charge_walk_device(const ChargeSector &c, uint64_t first, uint64_t *bits, uint64_t count, hipStream_t stream) {
    hipLaunchKernelGGL(charge_unrank_kernel, dim3(grid_for(count)), dim3(kChargeThreads), 0, stream,
                       args_of(c, TCP_C128), nullptr, bits, count, first, 1);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "charge sector walk kernel");
}

int charge_csr_counts(const ChargeSector &c, uint32_t *counts, hipStream_t stream) {
    if (c.geo.dim == 0) return TCP_OK;
    const ChargeArgs a = args_of(c, TCP_C128);
    if (c.cond)
        hipLaunchKernelGGL(charge_count_kernel<true>, dim3(grid_for(c.geo.dim)), dim3(kChargeThreads), 0, stream, a,
                           counts);
    else
        hipLaunchKernelGGL(charge_count_kernel<false>, dim3(grid_for(c.geo.dim)), dim3(kChargeThreads), 0, stream, a,
                           counts);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "charge sector csr count kernel");
}

int charge_csr_fill(const ChargeSector &c, const uint64_t *indptr, uint32_t *cols, void *vals64, void *vals32,
                    hipStream_t stream) {
    if (c.geo.dim == 0) return TCP_OK;
    const ChargeArgs a = args_of(c, TCP_C128);
    if (c.cond)
        hipLaunchKernelGGL(charge_fill_kernel<true>, dim3(grid_for(c.geo.dim)), dim3(kChargeThreads), 0, stream, a,
                           indptr, cols, static_cast<dvec2 *>(vals64), static_cast<fvec2 *>(vals32));
    else
        hipLaunchKernelGGL(charge_fill_kernel<false>, dim3(grid_for(c.geo.dim)), dim3(kChargeThreads), 0, stream, a,
                           indptr, cols, static_cast<dvec2 *>(vals64), static_cast<fvec2 *>(vals32));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "charge sector csr fill kernel");
}

}  // namespace tcp
