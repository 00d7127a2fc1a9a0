// Wide U(1) sectors (64..1024 qubits): host rank/unrank of packed words and
// the lazy/eager restricted-MVP kernels on them.  Same semantics as the
// narrow sector plan (restricted y = P H P^T v, x-groups in first-appearance
// order, exact-zero aggregated values skipped, leakage raised at build), with
// bitstrings of W 64-bit words and ranks from a choose table instead of byte
// lookup tables.
#include <algorithm>
#include <cmath>

#include "wide_sector.hpp"

namespace {

using tcp::fail;

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

constexpr uint32_t kWideThreads = 256;

struct WideArgs {
    uint32_t n, k, kc, comp;
    uint32_t words;  // real words of a bitstring (<= the kernel's template width)
    uint32_t G;
    uint32_t diag_group;
    uint64_t dim;
    const uint64_t *gx;      // [G][stride] x-masks
    const uint32_t *goff;    // [G + 1]
    const uint64_t *tz;      // [T][stride] z-masks
    const void *tw;          // [T] (re, im) of w = c i^(#Y mod 4)
    const uint64_t *choose;  // [(n + 1) (kc + 1)]
};

__device__ __forceinline__ double flip_sign(double x, uint32_t parity) {
    return __longlong_as_double(__double_as_longlong(x) ^ (static_cast<long long>(parity) << 63));
}

// Mask of the valid bits of word w.
__host__ __device__ __forceinline__ uint64_t word_mask(uint32_t n, uint32_t w) {
    const uint32_t lo = 64 * w;
    if (n >= lo + 64) return ~0ull;
    if (n <= lo) return 0;
    return (1ull << (n - lo)) - 1;
}

// Rank of a packed bitstring of weight k: positions p = n - 1 - q of the
// ranked pattern (the complement when comp) visited in ascending order, the
// i-th adding C(p, i + 1).  Returns UINT64_MAX when the weight is not k.
template <int WT>
__host__ __device__ uint64_t rank_words(const WideArgs &a, const uint64_t *s) {
    uint64_t r = 0;
    uint32_t i = 0;
    const uint32_t stride = a.kc + 1;
    for (int w = static_cast<int>(a.words) - 1; w >= 0; --w) {
        uint64_t v = s[w];
        if (a.comp) v = ~v & word_mask(a.n, static_cast<uint32_t>(w));
        while (v) {
            const uint32_t bit = 63u - static_cast<uint32_t>(__builtin_clzll(v));
            v &= ~(1ull << bit);
            if (i >= a.kc) return UINT64_MAX;
            const uint32_t p = a.n - 1 - (64u * static_cast<uint32_t>(w) + bit);
            r += a.choose[p * stride + i + 1];
            ++i;
        }
    }
    if (i != a.kc) return UINT64_MAX;
    return a.comp ? a.dim - 1 - r : r;
}

template <int WT>
__host__ __device__ void unrank_words(const WideArgs &a, uint64_t rank, uint64_t *s) {
    for (uint32_t w = 0; w < WT; ++w) s[w] = 0;
    uint64_t r = a.comp ? a.dim - 1 - rank : rank;
    const uint32_t stride = a.kc + 1;
    int pmax = static_cast<int>(a.n) - 1;
    for (int i = static_cast<int>(a.kc) - 1; i >= 0; --i) {
        // largest p in [i, pmax] with C(p, i + 1) <= r (C(i, i + 1) = 0 always qualifies)
        int lo = i, hi = pmax;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (a.choose[static_cast<uint32_t>(mid) * stride + static_cast<uint32_t>(i) + 1] <= r)
                lo = mid;
            else
                hi = mid - 1;
        }
        r -= a.choose[static_cast<uint32_t>(lo) * stride + static_cast<uint32_t>(i) + 1];
        const uint32_t q = a.n - 1 - static_cast<uint32_t>(lo);
        s[q >> 6] |= 1ull << (q & 63);
        pmax = lo - 1;
    }
    if (a.comp)
        for (uint32_t w = 0; w < a.words; ++w) s[w] = ~s[w] & word_mask(a.n, w);
}

template <int WT>
__device__ __forceinline__ uint32_t weight(const uint64_t *s, uint32_t words) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < WT; ++w)
        if (w < static_cast<int>(words)) c += __popcll(s[w]);
    return c;
}

// Aggregated value of group g on source s: sum of w (-1)^{popcount(z & s)}
// (a.tw holds the FP64 or the FP32-rounded coefficients).
template <int WT>
__device__ __forceinline__ dvec2 group_value(const WideArgs &a, uint32_t g, const uint64_t *s) {
    double dr = 0.0, di = 0.0;
    const uint32_t k1 = a.goff[g + 1];
    const dvec2 *tw = static_cast<const dvec2 *>(a.tw);
    for (uint32_t t = a.goff[g]; t < k1; ++t) {
        uint32_t par = 0;
#pragma unroll
        for (int w = 0; w < WT; ++w) par += __popcll(a.tz[static_cast<uint64_t>(t) * WT + w] & s[w]);
        par &= 1u;
        const dvec2 wt = tw[t];
        dr += flip_sign(wt.x, par);
        di += flip_sign(wt.y, par);
    }
    return dvec2{dr, di};
}

template <typename VT>
struct WAmp;
template <>
struct WAmp<dvec2> {
    __device__ static dvec2 load(const dvec2 *v, uint64_t i) { return v[i]; }
    __device__ static dvec2 make(double r, double i) { return dvec2{r, i}; }
};
template <>
struct WAmp<fvec2> {
    __device__ static dvec2 load(const fvec2 *v, uint64_t i) { return dvec2{v[i].x, v[i].y}; }
    __device__ static fvec2 make(double r, double i) { return fvec2{static_cast<float>(r), static_cast<float>(i)}; }
};

// Destination-major gather: y[r] = sum over groups of value(g, s) v[rank(s)],
// s = b ^ x_g in the sector.
template <typename VT, int WT>
__global__ __launch_bounds__(kWideThreads) void wide_gather_kernel(const VT *__restrict__ v, VT *__restrict__ y,
                                                                   WideArgs a) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kWideThreads) + threadIdx.x; r < a.dim;
         r += static_cast<uint64_t>(gridDim.x) * kWideThreads) {
        uint64_t b[WT], s[WT];
        unrank_words<WT>(a, r, b);
        double ar = 0.0, ai = 0.0;
        for (uint32_t g = 0; g < a.G; ++g) {
#pragma unroll
            for (int w = 0; w < WT; ++w) s[w] = b[w] ^ a.gx[static_cast<uint64_t>(g) * WT + w];
            if (weight<WT>(s, a.words) != a.k) continue;
            const dvec2 val = group_value<WT>(a, g, s);
            if (val.x == 0.0 && val.y == 0.0) continue;  // exact-zero skip (core/sector.rs:1099)
            const uint64_t src = g == a.diag_group ? r : rank_words<WT>(a, s);
            if (src >= a.dim) continue;  // unreachable for in-sector s; keeps the gather in bounds
            const dvec2 x = WAmp<VT>::load(v, src);
            ar = fma(val.x, x.x, ar);
            ar = fma(-val.y, x.y, ar);
            ai = fma(val.x, x.y, ai);
            ai = fma(val.y, x.x, ai);
        }
        y[r] = WAmp<VT>::make(ar, ai);
    }
}

// Leakage in the reference's scatter order: the smallest (source rank, group)
// whose destination leaves the sector with a nonzero aggregated value.
template <int WT>
__global__ __launch_bounds__(kWideThreads) void wide_leak_kernel(WideArgs a, unsigned long long *first) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kWideThreads) + threadIdx.x; r < a.dim;
         r += static_cast<uint64_t>(gridDim.x) * kWideThreads) {
        uint64_t b[WT], s[WT];
        unrank_words<WT>(a, r, b);
        for (uint32_t g = 0; g < a.G; ++g) {
#pragma unroll
            for (int w = 0; w < WT; ++w) s[w] = b[w] ^ a.gx[static_cast<uint64_t>(g) * WT + w];
            if (weight<WT>(s, a.words) == a.k) continue;
            const dvec2 val = group_value<WT>(a, g, b);
            if (val.x != 0.0 || val.y != 0.0) {
                atomicMin(first, static_cast<unsigned long long>(r) * a.G + g);
                break;
            }
        }
    }
}

template <int WT>
__global__ void wide_diag_kernel(WideArgs a, double *__restrict__ diag) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < a.dim;
         r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        double d = 0.0;
        if (a.diag_group < a.G) {
            uint64_t b[WT];
            unrank_words<WT>(a, r, b);
            d = group_value<WT>(a, a.diag_group, b).x;
        }
        diag[r] = d;
    }
}

template <int WT>
__global__ __launch_bounds__(kWideThreads) void wide_count_kernel(WideArgs a, uint32_t *__restrict__ counts) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kWideThreads) + threadIdx.x; r < a.dim;
         r += static_cast<uint64_t>(gridDim.x) * kWideThreads) {
        uint64_t b[WT], s[WT];
        unrank_words<WT>(a, r, b);
        uint32_t c = 0;
        for (uint32_t g = 0; g < a.G; ++g) {
#pragma unroll
            for (int w = 0; w < WT; ++w) s[w] = b[w] ^ a.gx[static_cast<uint64_t>(g) * WT + w];
            if (weight<WT>(s, a.words) != a.k) continue;
            const dvec2 val = group_value<WT>(a, g, s);
            if (val.x != 0.0 || val.y != 0.0) ++c;
        }
        counts[r] = c;
    }
}

// Row r's (column, value) pairs, columns ascending (the order of the
// reference's CSR rows, core/sector.rs:345-432).
template <int WT>
__global__ __launch_bounds__(kWideThreads) void wide_fill_kernel(WideArgs a, const uint64_t *__restrict__ indptr,
                                                                 uint32_t *__restrict__ cols, dvec2 *__restrict__ vals,
                                                                 fvec2 *__restrict__ vals32) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kWideThreads) + threadIdx.x; r < a.dim;
         r += static_cast<uint64_t>(gridDim.x) * kWideThreads) {
        uint64_t b[WT], s[WT];
        unrank_words<WT>(a, r, b);
        const uint64_t pos = indptr[r];
        const uint64_t len = indptr[r + 1] - pos;
        uint64_t m = 0;
        for (uint32_t g = 0; g < a.G && m < len; ++g) {
#pragma unroll
            for (int w = 0; w < WT; ++w) s[w] = b[w] ^ a.gx[static_cast<uint64_t>(g) * WT + w];
            if (weight<WT>(s, a.words) != a.k) continue;
            const dvec2 val = group_value<WT>(a, g, s);
            if (val.x == 0.0 && val.y == 0.0) continue;
            const uint64_t src = g == a.diag_group ? r : rank_words<WT>(a, s);
            // insertion into the sorted prefix (rows are short: a few hops per particle)
            uint64_t j = m;
            while (j > 0 && cols[pos + j - 1] > static_cast<uint32_t>(src)) {
                cols[pos + j] = cols[pos + j - 1];
                vals[pos + j] = vals[pos + j - 1];
                --j;
            }
            cols[pos + j] = static_cast<uint32_t>(src);
            vals[pos + j] = val;
            ++m;
        }
        for (uint64_t q = 0; q < len; ++q) {
            const dvec2 v = vals[pos + q];
            vals32[pos + q] = fvec2{static_cast<float>(v.x), static_cast<float>(v.y)};
        }
    }
}

WideArgs args_of(const tcp::WideSector &w, int dtype) {
    WideArgs a{};
    a.n = w.geo.n;
    a.k = w.geo.k;
    a.kc = w.geo.kc;
    a.comp = w.geo.comp ? 1u : 0u;
    a.words = w.geo.words;
    a.G = w.G;
    a.diag_group = w.diag_group;
    a.dim = w.geo.dim;
    a.gx = static_cast<const uint64_t *>(w.gx.ptr);
    a.goff = static_cast<const uint32_t *>(w.goff.ptr);
    a.tz = static_cast<const uint64_t *>(w.tz.ptr);
    a.tw = dtype == TCP_C64 ? w.tw32.ptr : w.tw64.ptr;
    a.choose = static_cast<const uint64_t *>(w.choose.ptr);
    return a;
}

// The kernels' template width: the smallest of 2, 4, 8, 16 words holding n qubits.
template <class F>
int with_width(uint32_t stride, F &&f) {
    switch (stride) {
        case 2: return f(std::integral_constant<int, 2>());
        case 4: return f(std::integral_constant<int, 4>());
        case 8: return f(std::integral_constant<int, 8>());
        default: return f(std::integral_constant<int, 16>());
    }
}

uint32_t width_for(uint32_t words) {
    uint32_t s = 2;
    while (s < words) s *= 2;
    return s;
}

uint32_t grid_for(uint64_t dim) {
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((dim + kWideThreads - 1) / kWideThreads,
                                                                           256 * 16)));
}

bool choose_fits(uint32_t n, uint32_t k, uint64_t &out) {
    if (k > n) {
        out = 0;
        return true;
    }
    if (k > n - k) k = n - k;
    unsigned __int128 value = 1;
    for (uint32_t i = 1; i <= k; ++i) {
        value = value * (n - k + i) / i;
        if (value > static_cast<unsigned __int128>(UINT64_MAX)) return false;
    }
    out = static_cast<uint64_t>(value);
    return true;
}

}  // namespace

namespace tcp {

int wide_geometry(uint32_t n, uint32_t k, WideGeometry &g) {
    if (k > n) return fail(TCP_ERR_INVALID_SECTOR, "invalid sector: particle_number must be between 0 and nqubits");
    if (n > kWideMaxQubits)
        return fail(TCP_ERR_UNSUPPORTED, "U1 sectors are supported up to " + std::to_string(kWideMaxQubits) +
                                             " qubits");
    g = WideGeometry{};
    g.n = n;
    g.k = k;
    g.comp = 2 * k > n;
    g.kc = g.comp ? n - k : k;
    g.words = word_count(n);
    if (!choose_fits(n, k, g.dim))
        return fail(TCP_ERR_OVERFLOW, "integer overflow while computing U1 sector dimension");
    const uint32_t stride = g.kc + 1;
    g.choose.assign(static_cast<size_t>(n + 1) * stride, 0);
    for (uint32_t m = 0; m <= n; ++m) {
        g.choose[m * stride] = 1;
        for (uint32_t j = 1; j <= std::min(m, g.kc); ++j) {
            // C(m, j) <= C(n, kc) = dim for j <= kc <= n / 2: no overflow
            const uint64_t above = j <= m - 1 ? g.choose[(m - 1) * stride + j] : 0;
            g.choose[m * stride + j] = g.choose[(m - 1) * stride + j - 1] + above;
        }
    }
    return TCP_OK;
}

int wide_rank(const WideGeometry &g, const uint64_t *words, uint64_t &rank) {
    for (uint32_t w = 0; w < g.words; ++w)
        if (words[w] & ~word_mask(g.n, w))
            return fail(TCP_ERR_INVALID_INDEX, "invalid index: bitstring has bits beyond nqubits");
    WideArgs a{};
    a.n = g.n;
    a.k = g.k;
    a.kc = g.kc;
    a.comp = g.comp;
    a.words = g.words;
    a.dim = g.dim;
    a.choose = g.choose.data();
    rank = rank_words<16>(a, words);
    if (rank == UINT64_MAX)
        return fail(TCP_ERR_INVALID_INDEX, "invalid index: bitstring has the wrong Hamming weight");
    return TCP_OK;
}

int wide_unrank(const WideGeometry &g, uint64_t rank, uint64_t *words) {
    if (rank >= g.dim) return fail(TCP_ERR_INVALID_INDEX, "invalid index: restricted basis index is out of range");
    WideArgs a{};
    a.n = g.n;
    a.k = g.k;
    a.kc = g.kc;
    a.comp = g.comp;
    a.words = g.words;
    a.dim = g.dim;
    a.choose = g.choose.data();
    std::vector<uint64_t> tmp(kWideMaxQubits / 64, 0);
    unrank_words<kWideMaxQubits / 64>(a, rank, tmp.data());
    std::copy(tmp.begin(), tmp.begin() + g.words, words);
    return TCP_OK;
}

int wide_build(const Operator &op, uint32_t k, uint64_t max_bytes, bool require_conserved, WideSector &out) {
    const uint32_t n = op.nqubits;
    int st = wide_geometry(n, k, out.geo);
    if (st) return st;
    if (out.geo.dim >= (1ull << 32))
        return fail(TCP_ERR_UNSUPPORTED, "GPU sector kernels support dimensions below 2^32");
    if (require_conserved) {
        std::vector<double> levels(2 * n, 0.0);
        for (uint32_t q = 0; q < n; ++q) levels[2 * q + 1] = 1.0;
        bool conserved = false;
        uint64_t nonzero = 0;
        if ((st = analyze_charge(op, levels.data(), conserved, nonzero))) return st;
        if (!conserved)
            return fail(TCP_ERR_SECTOR_LEAKAGE,
                        "selected U1 sector requires an exactly conserved operator; nonzero U(1) sector leakage "
                        "was detected");
    }
    const uint32_t words = out.geo.words;
    const uint32_t stride = width_for(words);
    out.stride = stride;
    // packed masks (qubit q in word q / 64, bit q % 64) and w = c i^(#Y mod 4)
    const uint64_t T = op.term_count();
    std::vector<uint64_t> x(T * stride, 0), z(T * stride, 0);
    std::vector<Complex> w(T);
    for (uint64_t t = 0; t < T; ++t) {
        unsigned y = 0;
        const uint8_t *c = op.codes.data() + t * n;
        for (uint32_t q = 0; q < n; ++q) {
            const uint64_t bit = 1ull << (q & 63);
            if (c[q] == 1 || c[q] == 2) x[t * stride + (q >> 6)] |= bit;
            if (c[q] == 2 || c[q] == 3) z[t * stride + (q >> 6)] |= bit;
            if (c[q] == 2) ++y;
        }
        const Complex a = op.coef[t];
        switch (y % 4) {
            case 0: w[t] = a; break;
            case 1: w[t] = Complex{-a.im, a.re}; break;
            case 2: w[t] = Complex{-a.re, -a.im}; break;
            default: w[t] = Complex{a.im, -a.re}; break;
        }
    }
    // x-groups in first-appearance order; terms keep canonical order
    std::vector<uint64_t> gx;
    std::vector<std::vector<uint64_t>> members;
    for (uint64_t t = 0; t < T; ++t) {
        size_t g = 0;
        for (; g < members.size(); ++g)
            if (std::equal(x.begin() + t * stride, x.begin() + (t + 1) * stride, gx.begin() + g * stride)) break;
        if (g == members.size()) {
            gx.insert(gx.end(), x.begin() + t * stride, x.begin() + (t + 1) * stride);
            members.emplace_back();
        }
        members[g].push_back(t);
    }
    std::vector<uint32_t> goff;
    std::vector<uint64_t> tz;
    std::vector<dvec2> tw64, tw32;
    double scale = 0.0;
    for (size_t g = 0; g < members.size(); ++g) {
        goff.push_back(static_cast<uint32_t>(tw64.size()));
        bool diag = std::all_of(gx.begin() + g * stride, gx.begin() + (g + 1) * stride, [](uint64_t v) { return v == 0; });
        if (diag) out.diag_group = static_cast<uint32_t>(g);
        double sum = 0.0;
        for (uint64_t t : members[g]) {
            tz.insert(tz.end(), z.begin() + t * stride, z.begin() + (t + 1) * stride);
            tw64.push_back(dvec2{w[t].re, w[t].im});
            tw32.push_back(dvec2{static_cast<double>(static_cast<float>(w[t].re)),
                                 static_cast<double>(static_cast<float>(w[t].im))});
            sum += std::hypot(w[t].re, w[t].im);
        }
        if (!diag) scale = std::max(scale, sum);
    }
    goff.push_back(static_cast<uint32_t>(tw64.size()));
    out.G = static_cast<uint32_t>(members.size());
    out.T = static_cast<uint32_t>(tw64.size());
    out.energy_scale = scale > 0.0 ? scale : 1.0;
    if (gx.empty()) gx.assign(stride, 0);
    if (tz.empty()) {
        tz.assign(stride, 0);
        tw64.push_back(dvec2{0, 0});
        tw32.push_back(dvec2{0, 0});
    }
    out.retained_bytes = sizeof(WideSector) + gx.size() * 8 + goff.size() * 4 + tz.size() * 8 + tw64.size() * 32 +
                         out.geo.choose.size() * 8;
    if (exceeds_budget(out.retained_bytes, max_bytes)) return memory_limit(out.retained_bytes, max_bytes);
    if ((st = out.gx.upload(gx.data(), gx.size() * 8)) || (st = out.goff.upload(goff.data(), goff.size() * 4)) ||
        (st = out.tz.upload(tz.data(), tz.size() * 8)) || (st = out.tw64.upload(tw64.data(), tw64.size() * 16)) ||
        (st = out.tw32.upload(tw32.data(), tw32.size() * 16)) ||
        (st = out.choose.upload(out.geo.choose.data(), out.geo.choose.size() * 8)))
        return st;
    // leakage scan
    if (out.G == 0 || out.geo.dim == 0) return TCP_OK;
    unsigned long long *dfirst = nullptr;
    TCP_HIP_CHECK(hipMalloc(&dfirst, sizeof(unsigned long long)));
    const unsigned long long none = ~0ull;
    hipError_t e = hipMemcpy(dfirst, &none, sizeof none, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const WideArgs a = args_of(out, TCP_C128);
        with_width(stride, [&](auto WT) {
            hipLaunchKernelGGL((wide_leak_kernel<decltype(WT)::value>), dim3(grid_for(a.dim)), dim3(kWideThreads), 0,
                               nullptr, a, dfirst);
            return 0;
        });
        e = hipGetLastError();
    }
    unsigned long long first = none;
    if (e == hipSuccess) e = hipMemcpy(&first, dfirst, sizeof first, hipMemcpyDeviceToHost);
    (void)hipFree(dfirst);
    if (e != hipSuccess) return hip_fail(e, "wide sector leakage scan");
    if (first != none) {
        out.leaks = true;
        out.leak_source = first / out.G;
        const uint64_t g = first % out.G;
        std::vector<uint64_t> b(words);
        if ((st = wide_unrank(out.geo, out.leak_source, b.data()))) return st;
        uint32_t weight = 0;
        for (uint32_t i = 0; i < words; ++i) weight += static_cast<uint32_t>(__builtin_popcountll(b[i] ^ gx[g * stride + i]));
        out.leak_weight = weight;
    }
    return TCP_OK;
}

int wide_apply(const WideSector &w, const void *in, void *out, int dtype, hipStream_t stream) {
    const WideArgs a = args_of(w, dtype);
    with_width(w.stride, [&](auto WT) {
        constexpr int K = decltype(WT)::value;
        if (dtype == TCP_C64)
            hipLaunchKernelGGL((wide_gather_kernel<fvec2, K>), dim3(grid_for(a.dim)), dim3(kWideThreads), 0, stream,
                               static_cast<const fvec2 *>(in), static_cast<fvec2 *>(out), a);
        else
            hipLaunchKernelGGL((wide_gather_kernel<dvec2, K>), dim3(grid_for(a.dim)), dim3(kWideThreads), 0, stream,
                               static_cast<const dvec2 *>(in), static_cast<dvec2 *>(out), a);
        return 0;
    });
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "wide sector kernel launch");
}

int wide_diagonal(const WideSector &w, double *diag, hipStream_t stream) {
    if (w.geo.dim == 0) return TCP_OK;
    const WideArgs a = args_of(w, TCP_C128);
    with_width(w.stride, [&](auto WT) {
        hipLaunchKernelGGL((wide_diag_kernel<decltype(WT)::value>), dim3(grid_for(a.dim)), dim3(256), 0, stream, a,
                           diag);
        return 0;
    });
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "wide sector diagonal kernel");
}

int wide_csr_counts(const WideSector &w, uint32_t *counts, hipStream_t stream) {
    const WideArgs a = args_of(w, TCP_C128);
    with_width(w.stride, [&](auto WT) {
        hipLaunchKernelGGL((wide_count_kernel<decltype(WT)::value>), dim3(grid_for(a.dim)), dim3(kWideThreads), 0,
                           stream, a, counts);
        return 0;
    });
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "wide csr count kernel");
}

int wide_csr_fill(const WideSector &w, const uint64_t *indptr, uint32_t *cols, void *vals64, void *vals32,
                  hipStream_t stream) {
    const WideArgs a = args_of(w, TCP_C128);
    with_width(w.stride, [&](auto WT) {
        hipLaunchKernelGGL((wide_fill_kernel<decltype(WT)::value>), dim3(grid_for(a.dim)), dim3(kWideThreads), 0,
                           stream, a, indptr, cols, static_cast<dvec2 *>(vals64), static_cast<fvec2 *>(vals32));
        return 0;
    });
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "wide csr fill kernel");
}

}  // namespace tcp

extern "C" {

int tcp_u1_rank_words(uint32_t nqubits, uint32_t particle_number, const uint64_t *words, uint64_t *rank) {
    if (!words || !rank) return tcp::fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    tcp::WideGeometry g;
    int st = tcp::wide_geometry(nqubits, particle_number, g);
    if (!st) st = tcp::wide_rank(g, words, *rank);
    return st ? st : tcp::ok();
}

int tcp_u1_unrank_words(uint32_t nqubits, uint32_t particle_number, uint64_t index, uint64_t *words) {
    if (!words) return tcp::fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    tcp::WideGeometry g;
    int st = tcp::wide_geometry(nqubits, particle_number, g);
    if (!st) st = tcp::wide_unrank(g, index, words);
    return st ? st : tcp::ok();
}

}  // extern "C"
