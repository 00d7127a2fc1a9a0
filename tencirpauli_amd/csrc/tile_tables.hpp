// Shared pieces of the full-space MVP kernels: the x-group/class tables built
// on the host and the device routine that applies them to one output tile.
//
// Tile geometry: a workgroup of TPB = 2^tb threads owns 2^L amplitudes,
// L = tb + ab, thread t holding A = 2^ab amplitudes l = t + a * TPB ("slot"
// a).  Terms are grouped by x-mask; groups are ordered by their high part
// x >> L (the "class": one source tile per class).  Coefficients carry the
// destination-index sign, w' = w (-1)^popc(z & x), so a group contributes
//     y[l] += D(l) v[l ^ lo],   D(l) = sum_k w'_k (-1)^popc(z_k & l)
// to every amplitude of the tile.  The sign of a term splits into a tile part
// (popc((z >> L) & tile), wave-uniform), a thread part (popc(z & (TPB-1) & t))
// and a slot part (popc(((z >> tb) & (A-1)) & a)).
//
// Each x-group is split by slot pattern zs = (z >> tb) & (A-1) into
// "passes" whose terms share zs (a 2D lattice with random terms has ~1.15
// passes per x-group): a pass contributes D(l) = S(t) (-1)^popc(zs & a) with
// one per-thread sum S.  Inside a class the passes are ordered by zs and the
// kernel runs one loop per slot pattern, so the slot signs are compile-time
// constants of straight-line code.
#pragma once

#include <algorithm>
#include <vector>

#include "tcp_internal.hpp"

namespace tcp {

template <class T>
using cptr = const __attribute__((address_space(4))) T *;

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

template <typename VT>
struct Amp;
template <>
struct Amp<dvec2> {
    __device__ static dvec2 widen(const dvec2 &v) { return v; }
    __device__ static dvec2 make(double r, double i) { return dvec2{r, i}; }
};
template <>
struct Amp<fvec2> {
    __device__ static dvec2 widen(const fvec2 &v) {
        return dvec2{static_cast<double>(v.x), static_cast<double>(v.y)};
    }
    __device__ static fvec2 make(double r, double i) {
        return fvec2{static_cast<float>(r), static_cast<float>(i)};
    }
};

__device__ __forceinline__ double flip_sign(double x, uint32_t parity) {
    return __longlong_as_double(__double_as_longlong(x) ^ (static_cast<long long>(parity) << 63));
}

// One source tile per class; its passes are [gbegin, gend), and those with
// slot pattern z are [zoff[c*S + z], zoff[c*S + z + 1]) with the record
// stride S = zoff_stride(A) (one 64- or 128-byte scalar load per class).
struct alignas(16) TileClass {
    uint64_t hi;  // x >> L
    uint32_t gbegin, gend;
};

// 64-byte pass descriptor (one scalar load).  The first term is stored
// inline, the others are [tbegin, tend) of the term array.
struct alignas(64) TileGroup {
    uint32_t lo;    // x & (2^L - 1)
    uint32_t zs;    // the shared slot pattern
    uint32_t same;  // 1 if the previous pass in x-mask order (xorder) has the same x-mask
    uint32_t zlo;   // first term: z & (TPB - 1)
    uint64_t zhi;   // first term: z >> L
    double wr, wi;  // first term: w'
    uint32_t tbegin, tend;
    uint32_t pad[4];
};

struct alignas(32) TileTerm {
    uint32_t zlo;  // z & (TPB - 1)
    uint32_t zs;   // (z >> tb) & (A - 1)
    uint64_t zhi;  // z >> L
    double wr, wi;
};

typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));

// A pass descriptor in one 64-byte scalar load.
__device__ __forceinline__ TileGroup load_group(cptr<TileGroup> p) {
    const u32x16 w = *reinterpret_cast<cptr<u32x16>>(p);
    TileGroup g;
    g.lo = w[0];
    g.zs = w[1];
    g.same = w[2];
    g.zlo = w[3];
    g.zhi = w[4] | (static_cast<uint64_t>(w[5]) << 32);
    g.wr = __longlong_as_double(static_cast<long long>(w[6] | (static_cast<uint64_t>(w[7]) << 32)));
    g.wi = __longlong_as_double(static_cast<long long>(w[8] | (static_cast<uint64_t>(w[9]) << 32)));
    g.tbegin = w[10];
    g.tend = w[11];
    return g;
}
__device__ __forceinline__ TileTerm load_term(cptr<TileTerm> p) {
    TileTerm t;
    t.zlo = p->zlo;
    t.zs = p->zs;
    t.zhi = p->zhi;
    t.wr = p->wr;
    t.wi = p->wi;
    return t;
}

// Device view of one operator's tables.  Group and term indices are absolute
// (several operators may share the arrays).
struct TileArgs {
    const TileClass *cls;   // [H]
    const uint32_t *zoff;   // [H * zoff_stride(A)] pass ranges per slot pattern
    const TileGroup *grp;   // passes
    const TileTerm *term;   // extra terms
    const uint32_t *xorder; // passes in x-mask order (plan builders only)
    uint32_t H;
    uint32_t L;
};

// Words per class in the pass-range table: A + 1 entries padded to 16 or 32.
__host__ __device__ constexpr uint32_t zoff_stride(uint32_t A) { return A + 1 <= 16 ? 16 : 32; }

template <int A>
struct SlotSign {  // (-1)^popc(zs & a) as a compile-time bit
    static constexpr bool neg(int zs, int a) { return __builtin_popcount(zs & a) & 1; }
};

// Byte offset in the LDS tile of source amplitude (t + a*TPB) ^ lo, given
// tbase = (t ^ (lo & (TPB-1))) * 16 and the slot part ls = lo >> tb:
// (a ^ ls) selects the slot, a wave-uniform offset.
template <uint32_t TPB>
__device__ __forceinline__ uint32_t src_offset(uint32_t tbase, uint32_t ls, int a) {
    return tbase + ((static_cast<uint32_t>(a) ^ ls) * TPB * static_cast<uint32_t>(sizeof(dvec2)));
}

// acc[a] += (-1)^popc(ZS & a) * S * src[a] for every pass of slot pattern ZS
// in [g0, g1) (straight-line code: the slot signs are compile-time).  The A
// LDS reads of a pass are issued together, then the per-thread sum S of the
// pass's terms, then the FMAs.
template <int ZS, int A, uint32_t TPB>
__device__ __forceinline__ void passes_uniform(cptr<TileGroup> grp, cptr<TileTerm> term, uint32_t g0, uint32_t g1,
                                               uint32_t t, uint32_t t16, uint64_t tile_id, const char *tile,
                                               double *accr, double *acci) {
    constexpr uint32_t TB = __builtin_ctz(TPB);
    for (uint32_t g = g0; g < g1; ++g) {
        const TileGroup G = load_group(grp + g);
        const uint32_t tbase = t16 ^ ((G.lo & (TPB - 1)) * static_cast<uint32_t>(sizeof(dvec2)));
        const uint32_t ls = G.lo >> TB;
        dvec2 s[A];
#pragma unroll
        for (int a = 0; a < A; ++a) s[a] = *reinterpret_cast<const dvec2 *>(tile + src_offset<TPB>(tbase, ls, a));
        uint32_t par = __popc(G.zlo & t) + __popcll(G.zhi & tile_id);
        double sr = flip_sign(G.wr, par & 1u), si = flip_sign(G.wi, par & 1u);
        for (uint32_t k = G.tbegin; k < G.tend; ++k) {
            const TileTerm T = load_term(term + k);
            par = __popc(T.zlo & t) + __popcll(T.zhi & tile_id);
            sr += flip_sign(T.wr, par & 1u);
            si += flip_sign(T.wi, par & 1u);
        }
#pragma unroll
        for (int a = 0; a < A; ++a) {
            if (SlotSign<A>::neg(ZS, a)) {
                accr[a] = fma(-sr, s[a].x, accr[a]);
                accr[a] = fma(si, s[a].y, accr[a]);
                acci[a] = fma(-sr, s[a].y, acci[a]);
                acci[a] = fma(-si, s[a].x, acci[a]);
            } else {
                accr[a] = fma(sr, s[a].x, accr[a]);
                accr[a] = fma(-si, s[a].y, accr[a]);
                acci[a] = fma(sr, s[a].y, acci[a]);
                acci[a] = fma(si, s[a].x, acci[a]);
            }
        }
    }
}

// All passes of one class, slot pattern by slot pattern (zr: the class's
// pass-range record).
template <int A, uint32_t TPB, int ZS = 0>
__device__ __forceinline__ void passes_all(cptr<TileGroup> grp, cptr<TileTerm> term, cptr<uint32_t> zr, uint32_t t,
                                           uint32_t t16, uint64_t tile_id, const char *tile, double *accr,
                                           double *acci) {
    passes_uniform<ZS, A, TPB>(grp, term, zr[ZS], zr[ZS + 1], t, t16, tile_id, tile, accr, acci);
    if constexpr (ZS + 1 < A) passes_all<A, TPB, ZS + 1>(grp, term, zr, t, t16, tile_id, tile, accr, acci);
}

// y[tile] = sum over classes/passes of the gathered source tiles.  `lds` is
// the workgroup's buffer of TPB * A FP64 amplitudes (FP32 storage is widened
// once per class).  `active` amplitudes of the tile exist (2^L except when
// the whole space is smaller than one tile; then only l < active are read
// and written).  The next class's source tile is loaded into registers while
// the current class is applied.  Every thread of the workgroup must call
// this (it synchronises).
template <int A, typename VT, uint32_t TPB, bool PARTIAL>
__device__ __forceinline__ void tile_apply(const VT *__restrict__ vin, VT *__restrict__ yout, const TileArgs &p,
                                           uint64_t tile_id, uint32_t active, dvec2 *lds) {
    cptr<TileClass> cls = (cptr<TileClass>)p.cls;
    cptr<TileGroup> grp = (cptr<TileGroup>)p.grp;
    cptr<TileTerm> term = (cptr<TileTerm>)p.term;
    cptr<uint32_t> zoff = (cptr<uint32_t>)p.zoff;

    const uint32_t t = threadIdx.x;
    const uint64_t base = tile_id << p.L;
    const VT *__restrict__ src_base = vin + t;
    auto live = [&](int a) { return !PARTIAL || t + a * TPB < active; };
    if (p.H == 0) {  // empty operator: H v = 0
#pragma unroll
        for (int a = 0; a < A; ++a)
            if (live(a)) yout[base + t + a * TPB] = Amp<VT>::make(0.0, 0.0);
        return;
    }

    const uint32_t t16 = t * static_cast<uint32_t>(sizeof(dvec2));
    const char *tile = reinterpret_cast<const char *>(lds);
    double accr[A], acci[A];
    VT pre[A];
#pragma unroll
    for (int a = 0; a < A; ++a) {
        accr[a] = 0.0;
        acci[a] = 0.0;
        pre[a] = Amp<VT>::make(0.0, 0.0);
    }
    {
        const uint64_t src = base ^ (static_cast<uint64_t>(cls[0].hi) << p.L);
#pragma unroll
        for (int a = 0; a < A; ++a)
            if (live(a)) pre[a] = src_base[src + a * TPB];
    }
    constexpr uint32_t ZSTRIDE = zoff_stride(A);
    for (uint32_t h = 0; h < p.H; ++h) {
        __syncthreads();  // every wave is done reading the previous class's tile
#pragma unroll
        for (int a = 0; a < A; ++a) lds[t + a * TPB] = Amp<VT>::widen(pre[a]);
        __syncthreads();
        if (h + 1 < p.H) {  // prefetch the next class's source tile
            const uint64_t src = base ^ (static_cast<uint64_t>(cls[h + 1].hi) << p.L);
#pragma unroll
            for (int a = 0; a < A; ++a)
                if (live(a)) pre[a] = src_base[src + a * TPB];
        }
        passes_all<A, TPB>(grp, term, zoff + h * ZSTRIDE, t, t16, tile_id, tile, accr, acci);
    }
    VT *__restrict__ out = yout + base + t;
#pragma unroll
    for (int a = 0; a < A; ++a)
        if (live(a)) out[a * TPB] = Amp<VT>::make(accr[a], acci[a]);
}

// Diagonal sum of pass g at tile-local position (t, slot) of tile `tile_id`,
// in FP64 (plan builders: eager transition counts, diagonals).
__device__ inline dvec2 group_sum(const TileArgs &p, uint32_t g, uint32_t t, uint32_t slot, uint64_t tile_id) {
    const TileGroup &G = p.grp[g];
    const uint32_t spar = __popc(G.zs & slot);
    uint32_t par = __popc(G.zlo & t) + __popcll(G.zhi & tile_id) + spar;
    double dr = flip_sign(G.wr, par & 1u), di = flip_sign(G.wi, par & 1u);
    for (uint32_t k = G.tbegin; k < G.tend; ++k) {
        const TileTerm &T = p.term[k];
        par = __popc(T.zlo & t) + __popcll(T.zhi & tile_id) + spar;
        dr += flip_sign(T.wr, par & 1u);
        di += flip_sign(T.wi, par & 1u);
    }
    return dvec2{dr, di};
}

// Host-side tables of one or more operators for tile geometry (tb, ab).
struct TileTables {
    std::vector<TileClass> cls;
    std::vector<uint32_t> zoff;            // (A+1) per class
    std::vector<uint32_t> xorder;          // passes in x-mask order
    std::vector<TileGroup> grp64, grp32;   // grp32: coefficients rounded to FP32
    std::vector<TileTerm> term64, term32;
    uint64_t G = 0, H = 0;      // of the last appended operator: distinct x-masks, classes
    uint64_t passes = 0;        // its passes (x-mask, slot pattern)
    uint64_t diag_passes = 0;   // its leading passes with x = 0
    size_t cls_base = 0;        // its first entry in cls
    size_t grp_base = 0;        // its first pass

    // Appends the tables of `op`.
    void append(const Operator &op, uint32_t tb, uint32_t ab) {
        const uint32_t L = tb + ab, A = 1u << ab;
        const uint64_t lmask = (1ull << L) - 1, tmask = (1ull << tb) - 1;
        const std::vector<IndexTerm> terms = index_terms(op);
        const uint64_t T = terms.size();
        struct Key {
            uint64_t hi, lo;
            uint32_t zs;
            uint64_t index;
        };
        std::vector<Key> keys(T);
        for (uint64_t k = 0; k < T; ++k)
            keys[k] = Key{terms[k].x >> L, terms[k].x & lmask, static_cast<uint32_t>((terms[k].z >> tb) & (A - 1)), k};
        // kernel order: class, slot pattern, x-mask, canonical term order
        std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
            if (a.hi != b.hi) return a.hi < b.hi;
            if (a.zs != b.zs) return a.zs < b.zs;
            if (a.lo != b.lo) return a.lo < b.lo;
            return a.index < b.index;
        });
        cls_base = cls.size();
        grp_base = grp64.size();
        auto make_term = [&](const IndexTerm &term, bool fp32) {
            const bool neg = __builtin_popcountll(term.z & term.x) & 1;
            double wr = neg ? -term.w.re : term.w.re, wi = neg ? -term.w.im : term.w.im;
            if (fp32) {
                wr = static_cast<double>(static_cast<float>(wr));
                wi = static_cast<double>(static_cast<float>(wi));
            }
            return TileTerm{static_cast<uint32_t>(term.z & tmask), static_cast<uint32_t>((term.z >> tb) & (A - 1)),
                            term.z >> L, wr, wi};
        };
        struct PassKey {
            uint64_t hi, lo;
            uint32_t zs, index;
        };
        std::vector<PassKey> pkeys;
        uint64_t k = 0;
        while (k < T) {
            if (cls.size() == cls_base || cls.back().hi != keys[k].hi) {
                if (cls.size() > cls_base) cls.back().gend = static_cast<uint32_t>(grp64.size());
                cls.push_back(TileClass{keys[k].hi, static_cast<uint32_t>(grp64.size()), 0});
            }
            uint64_t e = k;
            while (e < T && keys[e].hi == keys[k].hi && keys[e].zs == keys[k].zs && keys[e].lo == keys[k].lo) ++e;
            pkeys.push_back(PassKey{keys[k].hi, keys[k].lo, keys[k].zs, static_cast<uint32_t>(grp64.size())});
            for (int fp32 = 0; fp32 < 2; ++fp32) {
                std::vector<TileGroup> &grp = fp32 ? grp32 : grp64;
                std::vector<TileTerm> &tv = fp32 ? term32 : term64;
                TileGroup P{};
                const TileTerm head = make_term(terms[keys[k].index], fp32);
                P.lo = static_cast<uint32_t>(keys[k].lo);
                P.zs = keys[k].zs;
                P.zlo = head.zlo;
                P.zhi = head.zhi;
                P.wr = head.wr;
                P.wi = head.wi;
                P.tbegin = static_cast<uint32_t>(tv.size());
                for (uint64_t q = k + 1; q < e; ++q) tv.push_back(make_term(terms[keys[q].index], fp32));
                P.tend = static_cast<uint32_t>(tv.size());
                grp.push_back(P);
            }
            k = e;
        }
        if (cls.size() > cls_base) cls.back().gend = static_cast<uint32_t>(grp64.size());
        H = cls.size() - cls_base;
        passes = grp64.size() - grp_base;
        // pass ranges per slot pattern
        const uint32_t stride = zoff_stride(A);
        for (size_t c = cls_base; c < cls.size(); ++c) {
            uint32_t g = cls[c].gbegin;
            for (uint32_t z = 0; z < stride; ++z) {
                while (z <= A && g < cls[c].gend && grp64[g].zs < z) ++g;
                zoff.push_back(z <= A ? g : cls[c].gend);
            }
        }
        // x-mask order for the plan builders
        std::sort(pkeys.begin(), pkeys.end(), [](const PassKey &a, const PassKey &b) {
            if (a.hi != b.hi) return a.hi < b.hi;
            if (a.lo != b.lo) return a.lo < b.lo;
            return a.zs < b.zs;
        });
        G = 0;
        diag_passes = 0;
        for (size_t i = 0; i < pkeys.size(); ++i) {
            const bool same = i > 0 && pkeys[i].hi == pkeys[i - 1].hi && pkeys[i].lo == pkeys[i - 1].lo;
            grp64[pkeys[i].index].same = grp32[pkeys[i].index].same = same ? 1u : 0u;
            if (!same) ++G;
            if (pkeys[i].hi == 0 && pkeys[i].lo == 0) ++diag_passes;
            xorder.push_back(pkeys[i].index);
        }
    }

    // Keeps every device array non-empty (a zero-size upload has no pointer).
    void pad_empty() {
        if (cls.empty()) cls.push_back(TileClass{0, 0, 0});
        if (zoff.empty()) zoff.push_back(0);
        if (xorder.empty()) xorder.push_back(0);
        if (grp64.empty()) {
            grp64.push_back(TileGroup{});
            grp32.push_back(TileGroup{});
        }
        if (term64.empty()) {
            term64.push_back(TileTerm{});
            term32.push_back(TileTerm{});
        }
    }
    uint64_t device_bytes() const {
        return cls.size() * sizeof(TileClass) + (zoff.size() + xorder.size()) * sizeof(uint32_t) +
               (grp64.size() + grp32.size()) * sizeof(TileGroup) +
               (term64.size() + term32.size()) * sizeof(TileTerm);
    }
};

}  // namespace tcp
