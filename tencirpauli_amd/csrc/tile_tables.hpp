// Shared pieces of the full-space MVP kernels: the x-group/class tables built
// on the host and the device routine that applies them to one output tile.
//
// Tile geometry: a workgroup of TPB = 2^tb threads owns 2^L amplitudes,
// L = tb + ab, thread t holding A = 2^ab amplitudes l = t + a * TPB.  Terms
// are grouped by x-mask; groups are ordered by their high part x >> L (the
// "class": one source tile per class) and, inside a group, terms by the slot
// bits (z >> tb) & (A - 1) so one length-A Walsh-Hadamard transform per group
// folds the slot-dependent signs.  Coefficients carry the destination-index
// sign: w' = w (-1)^popc(z & x).
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

// Device view of one operator's tables.  Group and term indices stored in
// cls_goff / g_off are absolute (several operators may share the arrays).
struct TileArgs {
    const uint64_t *cls_hi;    // [H]    x >> L of each class
    const uint32_t *cls_goff;  // [H+1]  group range of each class
    const uint32_t *g_lo;      // [G]    x & (2^L - 1)
    const uint32_t *g_off;     // [G*A+1] term start per (group, slot class)
    const uint32_t *t_zlo;     // [T]    z & (2^tb - 1)
    const uint64_t *t_zhi;     // [T]    z >> L
    const void *t_w;           // [T]    (re, im) of w * (-1)^popc(z & x), as doubles
    uint32_t H;
    uint32_t L;
};

// y[tile] = sum over classes/groups of the gathered source tiles.  `tile` is
// the workgroup's LDS buffer of TPB * A amplitudes; `active` threads own an
// amplitude (fewer than TPB only when the whole space is smaller than TPB).
// Every thread of the workgroup must call this (it synchronises).
template <int A, typename VT, bool PARTIAL, uint32_t TPB>
__device__ __forceinline__ void tile_apply(const VT *__restrict__ vin, VT *__restrict__ yout, const TileArgs &p,
                                           uint64_t tile_id, uint32_t active, VT *tile) {
    cptr<uint64_t> cls_hi = (cptr<uint64_t>)p.cls_hi;
    cptr<uint32_t> cls_goff = (cptr<uint32_t>)p.cls_goff;
    cptr<uint32_t> g_lo = (cptr<uint32_t>)p.g_lo;
    cptr<uint32_t> g_off = (cptr<uint32_t>)p.g_off;
    cptr<uint32_t> t_zlo = (cptr<uint32_t>)p.t_zlo;
    cptr<uint64_t> t_zhi = (cptr<uint64_t>)p.t_zhi;
    cptr<dvec2> t_w = (cptr<dvec2>)p.t_w;

    const uint32_t t = threadIdx.x;
    const uint64_t base = tile_id << p.L;
    const VT *__restrict__ src_base = vin + t;
    const bool on = !PARTIAL || t < active;
    if (p.H == 0) {  // empty operator: H v = 0
        if (on) {
#pragma unroll
            for (int a = 0; a < A; ++a) yout[base + t + a * TPB] = Amp<VT>::make(0.0, 0.0);
        }
        return;
    }

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
            for (int a = 0; a < A; ++a) pre[a] = src_base[src + a * TPB];
        }
    }
    for (uint32_t h = 0; h < p.H; ++h) {
        __syncthreads();  // every wave is done reading the previous class's tile
        if (on) {
#pragma unroll
            for (int a = 0; a < A; ++a) tile[t + a * TPB] = pre[a];
        }
        __syncthreads();
        const uint32_t hn = h + 1 < p.H ? h + 1 : h;
        const uint64_t src = base ^ (cls_hi[hn] << p.L);
        if (h + 1 < p.H && on) {  // prefetch the next class's source tile
#pragma unroll
            for (int a = 0; a < A; ++a) pre[a] = src_base[src + a * TPB];
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
                const VT s = tile[(t + a * TPB) ^ lo];
                const double s_r = Amp<VT>::re(s), s_i = Amp<VT>::im(s);
                accr[a] = fma(sr[a], s_r, accr[a]);
                accr[a] = fma(-si[a], s_i, accr[a]);
                acci[a] = fma(sr[a], s_i, acci[a]);
                acci[a] = fma(si[a], s_r, acci[a]);
            }
        }
    }
    if (on) {
        VT *__restrict__ out = yout + base + t;
#pragma unroll
        for (int a = 0; a < A; ++a) out[a * TPB] = Amp<VT>::make(accr[a], acci[a]);
    }
}

// Host-side tables of one operator for tile geometry (tb, ab).
struct TileTables {
    std::vector<uint64_t> cls_hi;
    std::vector<uint32_t> cls_goff, g_lo, g_off, t_zlo;
    std::vector<uint64_t> t_zhi;
    std::vector<double> t_w64, t_w32;  // interleaved (re, im); t_w32 rounded to FP32
    uint64_t G = 0, H = 0;      // of the last appended operator
    bool has_diagonal = false;  // its first group has x = 0
    size_t cls_base = 0;        // its first entry in cls_hi
    size_t goff_base = 0;       // its first entry in cls_goff

    // Appends the tables of `op`; group and term indices are offset by what
    // the arrays already hold.
    void append(const Operator &op, uint32_t tb, uint32_t ab) {
        const uint32_t L = tb + ab, A = 1u << ab;
        const uint64_t lmask = (1ull << L) - 1;
        const std::vector<IndexTerm> terms = index_terms(op);
        const uint64_t T = terms.size();
        struct Key {
            uint64_t hi, lo;
            uint32_t cls;
            uint64_t index;
        };
        std::vector<Key> keys(T);
        for (uint64_t k = 0; k < T; ++k)
            keys[k] = Key{terms[k].x >> L, terms[k].x & lmask, static_cast<uint32_t>((terms[k].z >> tb) & (A - 1)), k};
        std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
            if (a.hi != b.hi) return a.hi < b.hi;
            if (a.lo != b.lo) return a.lo < b.lo;
            if (a.cls != b.cls) return a.cls < b.cls;
            return a.index < b.index;
        });
        const size_t cls_start = cls_hi.size();
        cls_base = cls_start;
        goff_base = cls_goff.size();
        const uint32_t term_base = static_cast<uint32_t>(t_zlo.size());
        if (!g_off.empty()) g_off.pop_back();  // the previous operator's end marker
        uint64_t k = 0;
        while (k < T) {
            if (cls_hi.size() == cls_start || cls_hi.back() != keys[k].hi) {
                cls_hi.push_back(keys[k].hi);
                cls_goff.push_back(static_cast<uint32_t>(g_lo.size()));
            }
            uint64_t e = k;
            while (e < T && keys[e].hi == keys[k].hi && keys[e].lo == keys[k].lo) ++e;
            g_lo.push_back(static_cast<uint32_t>(keys[k].lo));
            uint64_t j = k;
            for (uint32_t c = 0; c < A; ++c) {
                g_off.push_back(static_cast<uint32_t>(term_base + j));
                while (j < e && keys[j].cls == c) ++j;
            }
            for (uint64_t q = k; q < e; ++q) {
                const IndexTerm &term = terms[keys[q].index];
                const bool neg = __builtin_popcountll(term.z & term.x) & 1;
                const double wr = neg ? -term.w.re : term.w.re;
                const double wi = neg ? -term.w.im : term.w.im;
                t_zlo.push_back(static_cast<uint32_t>(term.z & ((1ull << tb) - 1)));
                t_zhi.push_back(term.z >> L);
                t_w64.push_back(wr);
                t_w64.push_back(wi);
                t_w32.push_back(static_cast<double>(static_cast<float>(wr)));
                t_w32.push_back(static_cast<double>(static_cast<float>(wi)));
            }
            k = e;
        }
        g_off.push_back(static_cast<uint32_t>(t_zlo.size()));
        cls_goff.push_back(static_cast<uint32_t>(g_lo.size()));
        H = cls_hi.size() - cls_start;
        G = cls_goff.back() - cls_goff[goff_base];
        has_diagonal = H > 0 && cls_hi[cls_start] == 0 && g_lo[cls_goff[goff_base]] == 0;
    }
};

}  // namespace tcp
