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
//
// Class blocks.  The terms of a class's passes are stored as 32-byte records
// in one block (a header of pass ranges per slot pattern, then the records),
// at most TPB 16-byte words.  Each thread loads one word of the next class's
// block together with its part of the next source tile, and the block is
// staged in LDS beside the tile, so the pass loop reads its terms from LDS
// (the same address in every lane) and issues no scalar memory loads: on
// gfx9 a scalar load and the LDS reads share one wait counter, and a
// descriptor fetched per pass through the scalar cache serialised every
// pass behind a scalar-cache round trip.
//
// Kernel coordinates.  All of the above is stated on the kernel's own index
// k, whose bits are, from the lowest, the thread bits, the slot bits and the
// tile bits.  Coordinate c of k is bit pos[c] of the amplitude's index in
// memory (a permutation of the index bits; identity unless the plan builder
// chose another).  Masks in the tables are in kernel coordinates; a class
// also carries its source tile's offset in memory.  The builder keeps the
// lowest four coordinates on the lowest four index bits, so that every group
// of 16 lanes loads 16 consecutive amplitudes (whole 128-byte lines), and
// picks the other tile-local bits and the six tile bits that vary fastest
// across the workgroups of one XCD so that few distinct source tiles are
// loaded per output tile and few distinct tiles are live in one XCD's L2.
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
    __device__ static void atomic_add(dvec2 *y, double r, double i) {
        double *d = reinterpret_cast<double *>(y);
        unsafeAtomicAdd(d, r);
        unsafeAtomicAdd(d + 1, i);
    }
};
template <>
struct Amp<fvec2> {
    __device__ static dvec2 widen(const fvec2 &v) {
        return dvec2{static_cast<double>(v.x), static_cast<double>(v.y)};
    }
    __device__ static fvec2 make(double r, double i) {
        return fvec2{static_cast<float>(r), static_cast<float>(i)};
    }
    __device__ static void atomic_add(fvec2 *y, double r, double i) {
        float *f = reinterpret_cast<float *>(y);
        unsafeAtomicAdd(f, static_cast<float>(r));
        unsafeAtomicAdd(f + 1, static_cast<float>(i));
    }
};

__device__ __forceinline__ double flip_sign(double x, uint32_t parity) {
    return __longlong_as_double(__double_as_longlong(x) ^ (static_cast<long long>(parity) << 63));
}

// One source tile per class, and the class block holding its passes.
struct alignas(16) TileClass {
    uint64_t off;          // source tile of this class relative to the output tile, as an index offset in memory
    uint32_t boff, bwords; // its block: words [boff, boff + bwords) of the block array
};

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// A term of a pass in a class block (two 16-byte words).
struct alignas(16) TermRecord {
    uint32_t lo_last;  // bits 0..15: the pass's x & (2^L - 1); bits 29..30: its kind; bit 31: its last term
    uint32_t zlo;      // z & (TPB - 1)
    uint64_t zhi;      // z >> L
    double wr, wi;     // w' (rounded to FP32 in the complex64 blocks)
};

// Record slots (32 bytes) taken by a block's header: A + 1 u16 record indices
// (slot-pattern z's passes are records [zr[z], zr[z+1])).
__host__ __device__ constexpr uint32_t header_slots(uint32_t A) { return ((A + 1) * 2 + 31) / 32; }
// 16-byte words of one class block in a TPB-thread workgroup (threads
// t < block_words(TPB) each load and stage one), and the most term records
// it holds.  The LDS copy has two more words, so that reading one record past
// a block's last stays inside it.
__host__ __device__ constexpr uint32_t block_words(uint32_t TPB) { return TPB / 2; }
__host__ __device__ constexpr uint32_t block_records(uint32_t TPB, uint32_t A) {
    return block_words(TPB) / 2 - header_slots(A);
}

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
    const u32x4 *blk;       // class blocks
    const TileGroup *grp;   // passes (plan builders)
    const TileTerm *term;   // extra terms (plan builders)
    const uint32_t *xorder; // passes in x-mask order (plan builders only)
    uint32_t H;
    uint32_t L;
    uint32_t n;             // index bits (kernel coordinates 0..n-1)
    uint32_t permuted;      // 0: pos is the identity
    uint8_t pos[64];        // kernel coordinate c -> index bit pos[c]
};

// Kernel coordinates of the in-memory index `phys` (plan builders).
__device__ __forceinline__ uint64_t kernel_index(const TileArgs &p, uint64_t phys) {
    if (!p.permuted) return phys;
    uint64_t k = 0;
    for (uint32_t c = 0; c < p.n; ++c) k |= ((phys >> p.pos[c]) & 1ull) << c;
    return k;
}

// In-memory index bits of kernel coordinates [first, first + count) holding `bits`.
__host__ __device__ __forceinline__ uint64_t deposit(const uint8_t *pos, uint32_t first, uint32_t count,
                                                     uint64_t bits) {
    uint64_t out = 0;
    for (uint32_t j = 0; j < count; ++j) out |= ((bits >> j) & 1ull) << pos[first + j];
    return out;
}

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

struct Rec {
    uint32_t lo_last, zlo;
    uint64_t zhi;
    double wr, wi;
};

// Record r of a block staged in LDS (every lane reads the same address).
__device__ __forceinline__ Rec read_rec(const u32x4 *blk, uint32_t r) {
    const u32x4 a = blk[2 * r], b = blk[2 * r + 1];
    Rec R;
    R.lo_last = a[0];
    R.zlo = a[1];
    R.zhi = a[2] | (static_cast<uint64_t>(a[3]) << 32);
    R.wr = __longlong_as_double(static_cast<long long>(b[0] | (static_cast<uint64_t>(b[1]) << 32)));
    R.wi = __longlong_as_double(static_cast<long long>(b[2] | (static_cast<uint64_t>(b[3]) << 32)));
    return R;
}

// Pass kinds, bits 29-30 of a pass's records' lo_last: the sum S of a pass
// whose terms all have real w' (imaginary w') is real (imaginary), and its
// complex multiply-add takes two FMAs per amplitude instead of four.
constexpr uint32_t kPassComplex = 0, kPassReal = 1, kPassImag = 2;
constexpr uint32_t kKindShift = 29;

// S += w' (-1)^popc(z & l) for one record, the parts KIND can have.  With
// SCALAR (FIRST: the pass's first term) S is set to the first term instead of
// 0 + term (stored weights are never zero, so the two are equal) and the tile
// part of the parity, popc(zhi & tile_id), the same in every lane, is formed
// in scalar registers: fewer vector instructions, measured 3 % faster in the
// twin kernel and 0.3-0.8 % slower in the single-tile kernels.
template <uint32_t KIND, bool SCALAR = false, bool FIRST = false>
__device__ __forceinline__ void add_term_k(const Rec &R, uint32_t t, uint64_t tile_id, double &sr, double &si) {
    uint32_t par;
    if constexpr (SCALAR) {
        const uint64_t zhi =
            static_cast<uint64_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(R.zhi))) |
            (static_cast<uint64_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(R.zhi >> 32))) << 32);
        par = (__popc(R.zlo & t) + static_cast<uint32_t>(__popcll(zhi & tile_id))) & 1u;
    } else {
        par = (__popc(R.zlo & t) + __popcll(R.zhi & tile_id)) & 1u;
    }
    if constexpr (SCALAR && FIRST) {
        if constexpr (KIND != kPassImag) sr = flip_sign(R.wr, par);
        if constexpr (KIND != kPassReal) si = flip_sign(R.wi, par);
    } else {
        if constexpr (KIND != kPassImag) sr += flip_sign(R.wr, par);
        if constexpr (KIND != kPassReal) si += flip_sign(R.wi, par);
    }
}

// One pass of kind KIND starting at record r of a pattern's range [r, r1)
// (head = the pass's first lo_last, uniform): acc[a] += (-1)^popc(ZS & a) *
// S * src[a].  The LDS reads of the first C slots are issued, then the
// terms' per-thread sum S is formed, then the FMAs run C slots at a time.
// With AHEAD the pass's first record is read here, together with the
// source reads (head, read ahead, gives the source offset), and the next
// pass's head is read before the FMAs, which cover its latency; head then
// becomes the next pass's (if any).  Without AHEAD, R is the pass's first
// record, read by the caller.  Returns the record after the pass.
template <uint32_t KIND, int ZS, int A, int C, uint32_t TPB, bool AHEAD, bool SCALAR>
__device__ __forceinline__ uint32_t pass_body(const u32x4 *blk, uint32_t r, uint32_t r1, Rec R, uint32_t &head,
                                              uint32_t t, uint32_t t16, uint64_t tile_id, const char *tile,
                                              double *accr, double *acci) {
    constexpr uint32_t TB = __builtin_ctz(TPB);
    const uint32_t lo = head & 0xFFFFu;
    const uint32_t tbase = t16 ^ ((lo & (TPB - 1)) * static_cast<uint32_t>(sizeof(dvec2)));
    const uint32_t ls = lo >> TB;
    dvec2 s[C];
#pragma unroll
    for (int a = 0; a < C; ++a) s[a] = *reinterpret_cast<const dvec2 *>(tile + src_offset<TPB>(tbase, ls, a));
    if constexpr (AHEAD) R = read_rec(blk, r);
    double sr = 0.0, si = 0.0;
    add_term_k<KIND, SCALAR, true>(R, t, tile_id, sr, si);
    ++r;
    for (uint32_t last = head >> 31; !last; ++r) {
        R = read_rec(blk, r);
        last = __builtin_amdgcn_readfirstlane(R.lo_last) >> 31;
        add_term_k<KIND, SCALAR>(R, t, tile_id, sr, si);
    }
    uint32_t next = 0;
    if constexpr (AHEAD) {
        if (r < r1) next = reinterpret_cast<const uint32_t *>(blk)[8 * r];  // lo_last of record r
    }
#pragma unroll
    for (int a0 = 0; a0 < A; a0 += C) {
        if (a0 > 0) {
#pragma unroll
            for (int c = 0; c < C; ++c)
                s[c] = *reinterpret_cast<const dvec2 *>(tile + src_offset<TPB>(tbase, ls, a0 + c));
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int a = a0 + c;
            const double pr = SlotSign<A>::neg(ZS, a) ? -sr : sr;
            const double pi = SlotSign<A>::neg(ZS, a) ? -si : si;
            if constexpr (KIND == kPassReal) {
                accr[a] = fma(pr, s[c].x, accr[a]);
                acci[a] = fma(pr, s[c].y, acci[a]);
            } else if constexpr (KIND == kPassImag) {
                accr[a] = fma(-pi, s[c].y, accr[a]);
                acci[a] = fma(pi, s[c].x, acci[a]);
            } else {
                accr[a] = fma(pr, s[c].x, accr[a]);
                accr[a] = fma(-pi, s[c].y, accr[a]);
                acci[a] = fma(pr, s[c].y, acci[a]);
                acci[a] = fma(pi, s[c].x, acci[a]);
            }
        }
    }
    if constexpr (AHEAD) head = __builtin_amdgcn_readfirstlane(next);
    return r;
}

// Every pass of slot pattern ZS, records [r0, r1) of the staged block
// (straight-line code: the slot signs are compile-time).  The plan builder
// orders a pattern's passes real, imaginary, complex; with KINDS three loops
// run them one after the other (no live registers shared between their
// bodies), the last one on the general code, which is right for a pass of
// any kind.  KINDS and AHEAD (pass_body) serve the single-tile kernels
// (measured 2-4 % faster with each); the twin kernel, at its register limit
// with one LDS read in flight, runs without both (each measured 3-5 %
// slower there).
template <int ZS, int A, uint32_t TPB, int CR = 0, bool KINDS = true, bool AHEAD = true, bool SCALAR = false>
__device__ __forceinline__ void passes_uniform(const u32x4 *blk, uint32_t r0, uint32_t r1, uint32_t t, uint32_t t16,
                                               uint64_t tile_id, const char *tile, double *accr, double *acci) {
    // slots read from LDS at a time (bounds live registers; CR > 0 overrides)
    constexpr int C = CR > 0 ? CR : (A >= 8 ? 4 : (A >= 4 ? 2 : A));
    uint32_t r = r0;
    Rec R{};
    uint32_t head = 0;
    if constexpr (AHEAD) {
        if (r0 >= r1) return;
        head = __builtin_amdgcn_readfirstlane(reinterpret_cast<const uint32_t *>(blk)[8 * r0]);
    }
    // the kind of the pass at r (AHEAD: from its head; else its first record is read here)
    auto kind_at = [&]() {
        if constexpr (!AHEAD) {
            R = read_rec(blk, r);
            head = __builtin_amdgcn_readfirstlane(R.lo_last);
        }
        return (head >> kKindShift) & 3u;
    };
    if constexpr (KINDS) {
        while (r < r1 && kind_at() == kPassReal)
            r = pass_body<kPassReal, ZS, A, C, TPB, AHEAD, SCALAR>(blk, r, r1, R, head, t, t16, tile_id, tile, accr,
                                                                   acci);
        while (r < r1 && kind_at() == kPassImag)
            r = pass_body<kPassImag, ZS, A, C, TPB, AHEAD, SCALAR>(blk, r, r1, R, head, t, t16, tile_id, tile, accr,
                                                                   acci);
    }
    while (r < r1) {
        if constexpr (!AHEAD) kind_at();
        r = pass_body<kPassComplex, ZS, A, C, TPB, AHEAD, SCALAR>(blk, r, r1, R, head, t, t16, tile_id, tile, accr,
                                                                  acci);
    }
}

// All passes of one class, slot pattern by slot pattern (zr: the block
// header, record index where slot pattern z's passes begin, zr[A] the end).
template <int A, uint32_t TPB, int ZS = 0, int CR = 0, bool KINDS = true, bool AHEAD = true, bool SCALAR = false>
__device__ __forceinline__ void passes_all(const u32x4 *blk, const uint32_t *zr, uint32_t t, uint32_t t16,
                                           uint64_t tile_id, const char *tile, double *accr, double *acci) {
    passes_uniform<ZS, A, TPB, CR, KINDS, AHEAD, SCALAR>(blk, zr[ZS], zr[ZS + 1], t, t16, tile_id, tile, accr, acci);
    if constexpr (ZS + 1 < A)
        passes_all<A, TPB, ZS + 1, CR, KINDS, AHEAD, SCALAR>(blk, zr, t, t16, tile_id, tile, accr, acci);
}

// Loads of one source tile into registers; NT marks them non-temporal.
template <bool NT, typename VT>
__device__ __forceinline__ VT load_amp(const VT *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// y[tile] = sum over classes/passes of the gathered source tiles.  `lds` is
// the workgroup's buffer of TPB * A FP64 amplitudes (FP32 storage is widened
// once per class).  `active` amplitudes of the tile exist (2^L except when
// the whole space is smaller than one tile; then only l < active are read
// and written).  Source tiles of the next PF classes are in flight in
// registers (a ring of PF buffers) while the current class is applied.
// Every thread of the workgroup must call this (it synchronises).  With
// `accumulate` the tile is added to y with atomic adds (the classes of one
// output tile split over several calls; y cleared beforehand).
template <int A, typename VT, uint32_t TPB, bool PARTIAL, int PF = 1, bool NT = false, bool PERM = false>
__device__ __forceinline__ void tile_apply(const VT *__restrict__ vin, VT *__restrict__ yout, const TileArgs &p,
                                           uint64_t tile_id, uint32_t active, dvec2 *lds, u32x4 *lblk,
                                           bool accumulate = false) {
    cptr<TileClass> cls = (cptr<TileClass>)p.cls;
    constexpr uint32_t TB = __builtin_ctz(TPB), AB = __builtin_ctz(static_cast<uint32_t>(A));

    const uint32_t t = threadIdx.x;
    // in-memory index = base | toff | slot_off(a) (disjoint bits)
    uint64_t base, toff;
    if constexpr (PERM) {
        base = deposit(p.pos, p.L, p.n - p.L, tile_id);
        toff = deposit(p.pos, 0, TB, t);
    } else {
        base = tile_id << p.L;
        toff = t;
    }
    auto slot_off = [&](int a) -> uint64_t {
        if constexpr (PERM) return deposit(p.pos, TB, AB, static_cast<uint64_t>(a));
        else return static_cast<uint64_t>(a) * TPB;
    };
    const VT *__restrict__ src_base = vin + toff;
    auto live = [&](int a) { return !PARTIAL || t + a * TPB < active; };
    if (p.H == 0) {  // empty operator: H v = 0
#pragma unroll
        for (int a = 0; a < A; ++a)
            if (live(a)) yout[base + toff + slot_off(a)] = Amp<VT>::make(0.0, 0.0);
        return;
    }

    const uint32_t t16 = t * static_cast<uint32_t>(sizeof(dvec2));
    const char *tile = reinterpret_cast<const char *>(lds);
    double accr[A], acci[A];
    VT pre[PF][A];
    u32x4 bpre[PF];  // this thread's word of the in-flight classes' blocks
#pragma unroll
    for (int a = 0; a < A; ++a) {
        accr[a] = 0.0;
        acci[a] = 0.0;
    }
    auto fetch = [&](VT *buf, u32x4 &bw, uint32_t h) {
        const uint64_t src = base ^ cls[h].off;
        const uint32_t boff = cls[h].boff, bwords = cls[h].bwords;
#pragma unroll
        for (int a = 0; a < A; ++a)
            buf[a] = live(a) ? load_amp<NT>(src_base + (src | slot_off(a))) : Amp<VT>::make(0.0, 0.0);
        bw = t < bwords ? p.blk[boff + t] : u32x4{0u, 0u, 0u, 0u};  // bwords <= block_words(TPB)
    };
#pragma unroll
    for (int k = 0; k < PF; ++k)
        if (static_cast<uint32_t>(k) < p.H) fetch(pre[k], bpre[k], k);
    for (uint32_t h0 = 0; h0 < p.H; h0 += PF) {
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const uint32_t h = h0 + k;
            if (h >= p.H) break;
            __syncthreads();  // every wave is done reading the previous class's tile and block
#pragma unroll
            for (int a = 0; a < A; ++a) lds[t + a * TPB] = Amp<VT>::widen(pre[k][a]);
            if (t < block_words(TPB)) lblk[t] = bpre[k];
            __syncthreads();
            if (h + PF < p.H) fetch(pre[k], bpre[k], h + PF);  // the class PF ahead into the freed buffers
            uint32_t zr[A + 1];  // block header: u16 entries, read once into scalar registers
#pragma unroll
            for (int z = 0; z <= A; ++z) {
                const uint32_t word = reinterpret_cast<const uint32_t *>(lblk)[z >> 1];
                zr[z] = __builtin_amdgcn_readfirstlane((word >> ((z & 1) * 16)) & 0xFFFFu);
            }
            passes_all<A, TPB>(lblk, zr, t, t16, tile_id, tile, accr, acci);
        }
    }
    VT *__restrict__ out = yout + base + toff;
    if (accumulate) {
#pragma unroll
        for (int a = 0; a < A; ++a)
            if (live(a)) Amp<VT>::atomic_add(out + slot_off(a), accr[a], acci[a]);
        return;
    }
#pragma unroll
    for (int a = 0; a < A; ++a)
        if (live(a)) out[slot_off(a)] = Amp<VT>::make(accr[a], acci[a]);
}

// Twin tiles: the workgroup owns the two output tiles 2*pair and 2*pair + 1
// (kernel tile bit 0, coordinate L, tells them apart), 2A FP64 accumulators
// per thread.  A class is an x-mask part above coordinate L; its source is
// the 2^(L+1) tile pair ^ (class), staged through LDS as two halves.  A pass
// whose x has bit L = px, applied while half h is staged, reads that half for
// output tile h ^ px, so every staged half serves both output tiles: source
// loads per amplitude follow the class count of 2^(L+1)-amplitude tiles while
// LDS holds 2^L amplitudes.  Class blocks order passes by (px, slot pattern),
// header entries [0, A] for px = 0 and [A, 2A] for px = 1.  The class table
// lists stages (TileTables::append): p.H is their count.  Every thread of
// the workgroup must call this (it synchronises).
template <int A, typename VT, uint32_t TPB, int CR = 0>
__device__ __forceinline__ void twin_apply(const VT *__restrict__ vin, VT *__restrict__ yout, const TileArgs &p,
                                           uint64_t pair, dvec2 *lds, u32x4 *lblk) {
    cptr<TileClass> cls = (cptr<TileClass>)p.cls;
    constexpr uint32_t TB = __builtin_ctz(TPB), AB = __builtin_ctz(static_cast<uint32_t>(A));
    const uint32_t t = threadIdx.x;
    const uint64_t base = deposit(p.pos, p.L + 1, p.n - p.L - 1, pair);  // output tile 2*pair
    const uint64_t half_off = 1ull << p.pos[p.L];                       // its twin, and a class's second half
    const uint64_t toff = deposit(p.pos, 0, TB, t);
    auto slot_off = [&](int a) -> uint64_t { return deposit(p.pos, TB, AB, static_cast<uint64_t>(a)); };
    const VT *__restrict__ src_base = vin + toff;
    const uint64_t tile0 = pair << 1, tile1 = tile0 | 1u;
    const uint32_t t16 = t * static_cast<uint32_t>(sizeof(dvec2));
    const char *tile = reinterpret_cast<const char *>(lds);
    double accr0[A], acci0[A], accr1[A], acci1[A];
#pragma unroll
    for (int a = 0; a < A; ++a) {
        accr0[a] = 0.0;
        acci0[a] = 0.0;
        accr1[a] = 0.0;
        acci1[a] = 0.0;
    }
    VT pre[A];
    u32x4 bpre;
    uint32_t pre_half = 0;
    // stage s (a class's half, folded into its offset) into the registers
    auto fetch = [&](uint32_t s) {
        const uint64_t src = base ^ cls[s].off;
        const uint32_t boff = cls[s].boff, bwords = cls[s].bwords & 0x7FFFFFFFu;
        pre_half = cls[s].bwords >> 31;
#pragma unroll
        for (int a = 0; a < A; ++a) pre[a] = load_amp<false>(src_base + (src | slot_off(a)));
        bpre = t < bwords ? p.blk[boff + t] : u32x4{0u, 0u, 0u, 0u};  // bwords <= block_words(TPB)
    };
    auto stage = [&]() {
        __syncthreads();  // every wave is done reading the previous half and block
#pragma unroll
        for (int a = 0; a < A; ++a) lds[t + a * TPB] = Amp<VT>::widen(pre[a]);
        if (t < block_words(TPB)) lblk[t] = bpre;
        __syncthreads();
    };
    auto header = [&](uint32_t *zr) {  // u16 entries, read once into scalar registers
#pragma unroll
        for (int z = 0; z <= 2 * A; ++z) {
            const uint32_t word = reinterpret_cast<const uint32_t *>(lblk)[z >> 1];
            zr[z] = __builtin_amdgcn_readfirstlane((word >> ((z & 1) * 16)) & 0xFFFFu);
        }
    };
    fetch(0);
    for (uint32_t s = 0; s < p.H; ++s) {
        uint32_t zr[2 * A + 1];
        const uint32_t half = pre_half;
        stage();
        if (s + 1 < p.H) fetch(s + 1);
        header(zr);
        // passes for output tile 0 have px = half, those for tile 1 px = 1 - half
        uint32_t z0[A + 1], z1[A + 1];
#pragma unroll
        for (int z = 0; z <= A; ++z) {
            z0[z] = half ? zr[A + z] : zr[z];
            z1[z] = half ? zr[z] : zr[A + z];
        }
        passes_all<A, TPB, 0, CR, false, false, true>(lblk, z0, t, t16, tile0, tile, accr0, acci0);
        passes_all<A, TPB, 0, CR, false, false, true>(lblk, z1, t, t16, tile1, tile, accr1, acci1);
    }
    VT *__restrict__ out0 = yout + base + toff;
    VT *__restrict__ out1 = yout + (base ^ half_off) + toff;
#pragma unroll
    for (int a = 0; a < A; ++a) {
        out0[slot_off(a)] = Amp<VT>::make(accr0[a], acci0[a]);
        out1[slot_off(a)] = Amp<VT>::make(accr1[a], acci1[a]);
    }
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

// Kernel coordinates for the tile kernel of geometry (tb, ab) on `op`, with
// `window` tile bits varying fastest across the workgroups that run together
// on one XCD and `xcd` tile bits (placed right above the window) that tell
// the XCDs apart.  Coordinates 0..3 stay on index bits 0..3; the other
// tile-local bits are picked greedily to leave the fewest distinct x-mask
// parts outside the tile (source tiles per output tile), then the window bits
// to leave the fewest outside tile and window (distinct source tiles an XCD
// streams while its workgroups run one class each), then the XCD bits the
// same way (source tiles the eight XCDs share through the last-level cache),
// then, among the tile-local bits, the slot bits that split the fewest
// x-groups into passes.  With `twin` one more tile-local bit is chosen and
// placed at coordinate tb + ab, the lowest tile bit: it tells apart the two
// output tiles a twin-tile workgroup owns.  Ties go to the lowest bit.  Returns false
// (identity) for spaces too small to tile, and for operators with more than
// 16,384 distinct x-masks (where the search would cost more than it saves).
inline bool choose_bit_map(const Operator &op, uint32_t tb, uint32_t ab, uint32_t window, uint32_t xcd,
                           uint8_t *pos, uint32_t twin = 0) {
    const uint32_t n = op.nqubits, L = tb + ab + twin;
    if (n > 64 || tb < 4 || n < L + window + xcd) return false;
    const std::vector<IndexTerm> terms = index_terms(op);
    std::vector<uint64_t> xs(terms.size());
    for (size_t i = 0; i < terms.size(); ++i) xs[i] = terms[i].x;
    std::sort(xs.begin(), xs.end());
    xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
    if (xs.size() > 16384) return false;
    std::vector<uint64_t> scratch;
    auto distinct_outside = [&](uint64_t inside) {
        scratch.assign(xs.begin(), xs.end());
        for (uint64_t &v : scratch) v &= ~inside;
        std::sort(scratch.begin(), scratch.end());
        return static_cast<size_t>(std::unique(scratch.begin(), scratch.end()) - scratch.begin());
    };
    auto grow = [&](uint64_t &set, uint32_t count, uint64_t fixed) {
        for (uint32_t step = 0; step < count; ++step) {
            size_t best = SIZE_MAX;
            uint32_t best_b = 0;
            for (uint32_t b = 0; b < n; ++b) {
                if (((set | fixed) >> b) & 1ull) continue;
                const size_t c = distinct_outside(fixed | set | (1ull << b));
                if (c < best) {
                    best = c;
                    best_b = b;
                }
            }
            set |= 1ull << best_b;
        }
    };
    uint64_t local = 0xFull;
    grow(local, L - 4, 0);
    uint64_t win = 0;
    grow(win, window, local);
    uint64_t xb = 0;
    grow(xb, xcd, local | win);
    // slot bits: fewest distinct (x, z restricted to the slot bits) pairs
    std::vector<std::pair<uint64_t, uint64_t>> pairs(terms.size());
    uint64_t slot = 0;
    for (uint32_t step = 0; step < ab; ++step) {
        size_t best = SIZE_MAX;
        uint32_t best_b = 0;
        for (uint32_t b = 4; b < n; ++b) {
            if (!((local >> b) & 1ull) || ((slot >> b) & 1ull)) continue;
            const uint64_t m = slot | (1ull << b);
            for (size_t i = 0; i < terms.size(); ++i) pairs[i] = {terms[i].x, terms[i].z & m};
            std::sort(pairs.begin(), pairs.end());
            const size_t c = static_cast<size_t>(std::unique(pairs.begin(), pairs.end()) - pairs.begin());
            if (c < best) {
                best = c;
                best_b = b;
            }
        }
        slot |= 1ull << best_b;
    }
    // twin tiles: the highest tile-local bit outside the slot and lowest four
    // bits tells the two output tiles apart (any choice loads the same tiles)
    uint64_t pair = 0;
    if (twin) {
        const uint64_t cand = local & ~slot & ~0xFull;
        pair = 1ull << (63 - __builtin_clzll(cand));
    }
    uint32_t c = 0;
    auto place = [&](uint64_t set) {
        for (uint32_t b = 0; b < n; ++b)
            if ((set >> b) & 1ull) pos[c++] = static_cast<uint8_t>(b);
    };
    place(local & ~slot & ~pair);  // thread coordinates (bits 0..3 first)
    place(slot);
    place(pair);
    place(win);
    place(xb);
    const uint64_t all = n == 64 ? ~0ull : (1ull << n) - 1;
    place(all & ~(local | win | xb));
    return true;
}

// Host-side tables of one or more operators for tile geometry (tb, ab).
struct TileTables {
    std::vector<TileClass> cls;
    std::vector<u32x4> blk64, blk32;       // class blocks (blk32: coefficients rounded to FP32)
    std::vector<uint32_t> xorder;          // passes in x-mask order
    std::vector<TileGroup> grp64, grp32;   // grp32: coefficients rounded to FP32
    std::vector<TileTerm> term64, term32;
    uint64_t G = 0, H = 0;      // of the last appended operator: distinct x-masks, classes
    uint64_t passes = 0;        // its passes (x-mask, slot pattern)
    uint64_t diag_passes = 0;   // its leading passes with x = 0
    size_t cls_base = 0;        // its first entry in cls
    size_t grp_base = 0;        // its first pass

    // Appends the tables of `op`, with kernel coordinate c at index bit
    // pos[c] (pos = nullptr: the identity), for a kernel of `wg_threads`
    // threads per workgroup (0: 2^tb), which sets the class block size.
    // With `twin` the classes are the x-mask parts above coordinate L (one
    // 2^(L+1) source tile, staged as two halves) and a class block orders its
    // passes by (bit L of x, slot pattern): its header has 2A + 1 entries.
    // The class table then lists stages, one per (class, half), with the half
    // folded into the offset and flagged in bit 31 of bwords, ordered run by
    // run (a run: the classes whose x-mask parts above L differ only in the
    // `window` lowest bits, the bits that vary across the workgroups one XCD
    // runs together), the first halves of a run before its second halves: a
    // source tile that one workgroup stages is staged by another at most a
    // run's length of stages later, while it is still in L2.
    void append(const Operator &op, uint32_t tb, uint32_t ab, const uint8_t *pos = nullptr,
                uint32_t wg_threads = 0, bool twin = false, uint32_t window = 0) {
        const uint32_t L = tb + ab, A = 1u << ab, CL = L + (twin ? 1u : 0u);
        const uint64_t lmask = (1ull << L) - 1, tmask = (1ull << tb) - 1;
        std::vector<IndexTerm> terms = index_terms(op);
        const uint64_t T = terms.size();
        const uint32_t n = op.nqubits;
        if (pos) {
            auto to_kernel = [&](uint64_t phys) {
                uint64_t k = 0;
                for (uint32_t c = 0; c < n; ++c) k |= ((phys >> pos[c]) & 1ull) << c;
                return k;
            };
            for (IndexTerm &term : terms) {  // popc(z & x) is the same in both coordinates
                term.x = to_kernel(term.x);
                term.z = to_kernel(term.z);
            }
        }
        auto class_offset = [&](uint64_t hi) {
            return pos ? deposit(pos, CL, n > CL ? n - CL : 0, hi) : hi << CL;
        };
        struct Key {
            uint64_t hi, lo;
            uint32_t pat;  // block pattern: slot pattern, plus A * (bit L of x) for twin tiles
            uint32_t zs;
            uint64_t index;
            uint32_t rank = 0;  // pass kind order: real, imaginary, complex
        };
        std::vector<Key> keys(T);
        for (uint64_t k = 0; k < T; ++k) {
            const uint32_t zs = static_cast<uint32_t>((terms[k].z >> tb) & (A - 1));
            const uint32_t px = twin ? static_cast<uint32_t>((terms[k].x >> L) & 1u) : 0u;
            keys[k] = Key{terms[k].x >> CL, terms[k].x & lmask, zs + A * px, zs, k, 0u};
        }
        // kernel order: class, block pattern, x-mask, canonical term order
        std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
            if (a.hi != b.hi) return a.hi < b.hi;
            if (a.pat != b.pat) return a.pat < b.pat;
            if (a.lo != b.lo) return a.lo < b.lo;
            return a.index < b.index;
        });
        // then, within a class and pattern, passes whose terms' w' are all
        // real first, all imaginary next, the others last (the kernel's pass
        // loops run in that order)
        for (uint64_t k0 = 0; k0 < T;) {
            uint64_t k1 = k0;
            bool real = true, imag = true;
            while (k1 < T && keys[k1].hi == keys[k0].hi && keys[k1].pat == keys[k0].pat && keys[k1].lo == keys[k0].lo) {
                const IndexTerm &term = terms[keys[k1].index];
                real = real && term.w.im == 0.0;  // w' = w (-1)^popc(z & x): the same parts are zero
                imag = imag && term.w.re == 0.0;
                ++k1;
            }
            const uint32_t rank = real ? 0u : (imag ? 1u : 2u);
            for (uint64_t k = k0; k < k1; ++k) keys[k].rank = rank;
            k0 = k1;
        }
        std::stable_sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
            if (a.hi != b.hi) return a.hi < b.hi;
            if (a.pat != b.pat) return a.pat < b.pat;
            return a.rank < b.rank;
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
        std::vector<uint32_t> pats;  // block pattern of each pass
        struct ClassPasses {
            uint64_t hi, off;
            size_t gbegin, gend;
        };
        std::vector<ClassPasses> classes;
        uint64_t k = 0;
        while (k < T) {
            if (classes.empty() || keys[k].hi != keys[k - 1].hi) {
                if (!classes.empty()) classes.back().gend = grp64.size();
                classes.push_back(ClassPasses{keys[k].hi, class_offset(keys[k].hi), grp64.size(), 0});
            }
            uint64_t e = k;
            while (e < T && keys[e].hi == keys[k].hi && keys[e].pat == keys[k].pat && keys[e].lo == keys[k].lo) ++e;
            // x-mask order keys use the part above L (twin classes hold two of them)
            pkeys.push_back(PassKey{terms[keys[k].index].x >> L, keys[k].lo, keys[k].zs,
                                    static_cast<uint32_t>(grp64.size())});
            pats.push_back(keys[k].pat);
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
        if (!classes.empty()) classes.back().gend = grp64.size();
        std::vector<uint64_t> run_of;  // run of each class-table entry (twin stage order)
        for (size_t i = 0; i < classes.size(); ++i) {
            const ClassPasses &c = classes[i];
            append_blocks(c.off, c.gbegin, c.gend, pats.data() + (c.gbegin - grp_base), twin ? 2 * A : A,
                          wg_threads ? wg_threads : 1u << tb);
            run_of.resize(cls.size() - cls_base, c.hi >> window);
        }
        if (twin) {
            const uint64_t half_off = pos ? 1ull << pos[L] : 1ull << L;
            std::vector<TileClass> entries(cls.begin() + cls_base, cls.end()), stages;
            for (size_t r0 = 0; r0 < entries.size();) {
                size_t r1 = r0;
                while (r1 < entries.size() && run_of[r1] == run_of[r0]) ++r1;
                for (uint32_t half = 0; half < 2; ++half)
                    for (size_t e = r0; e < r1; ++e)
                        stages.push_back(TileClass{entries[e].off ^ (half ? half_off : 0ull), entries[e].boff,
                                                   entries[e].bwords | (half << 31)});
                r0 = r1;
            }
            cls.resize(cls_base);
            cls.insert(cls.end(), stages.begin(), stages.end());
        }
        H = cls.size() - cls_base;
        passes = grp64.size() - grp_base;
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

    // The blocks of one class: passes [gb, ge) (ordered by block pattern,
    // pat[g - gb] in [0, P)) as term records, packed into blocks of at most
    // block_records(wg_threads, P) records; a class that does not fit in one
    // block becomes several classes with the same source tile, and a pass
    // with more terms than a block holds is split into passes of consecutive
    // terms (the same sum, accumulated in two steps).
    void append_blocks(uint64_t off, size_t gb, size_t ge, const uint32_t *pat, uint32_t P, uint32_t wg_threads) {
        const uint32_t A = P, HS = header_slots(A);
        // 2 * (HS + records) words per block <= block_words(wg_threads); the kernels run
        // 64 or more threads per workgroup, so a block holds 14 or more records
        const uint32_t RMAX = block_records(std::max<uint32_t>(wg_threads, 64), A);
        struct Chunk {
            uint32_t zs;
            std::vector<TermRecord> r64, r32;
        };
        std::vector<Chunk> chunks;
        for (size_t g = gb; g < ge; ++g) {
            auto record = [](const TileGroup &P, const TileTerm &T) {
                return TermRecord{P.lo, T.zlo, T.zhi, T.wr, T.wi};
            };
            std::vector<TermRecord> all64, all32;
            const TileGroup &P64 = grp64[g], &P32 = grp32[g];
            all64.push_back(record(P64, TileTerm{P64.zlo, P64.zs, P64.zhi, P64.wr, P64.wi}));
            all32.push_back(record(P32, TileTerm{P32.zlo, P32.zs, P32.zhi, P32.wr, P32.wi}));
            for (uint32_t q = P64.tbegin; q < P64.tend; ++q) {
                all64.push_back(record(P64, term64[q]));
                all32.push_back(record(P32, term32[q]));
            }
            // a chunk is one pass of the kernel: its kind from its records' coefficients
            auto mark_kind = [](std::vector<TermRecord> &rs) {
                bool real = true, imag = true;
                for (const TermRecord &q : rs) {
                    real = real && q.wi == 0.0;
                    imag = imag && q.wr == 0.0;
                }
                const uint32_t kind = real ? kPassReal : (imag ? kPassImag : kPassComplex);
                for (TermRecord &q : rs) q.lo_last |= kind << kKindShift;
                rs.back().lo_last |= 1u << 31;
            };
            for (size_t i = 0; i < all64.size(); i += RMAX) {
                Chunk c{pat[g - gb], {}, {}};
                const size_t e = std::min(all64.size(), i + RMAX);
                c.r64.assign(all64.begin() + i, all64.begin() + e);
                c.r32.assign(all32.begin() + i, all32.begin() + e);
                mark_kind(c.r64);
                mark_kind(c.r32);
                chunks.push_back(std::move(c));
            }
        }
        auto push_words = [](std::vector<u32x4> &blk, const void *data, size_t words) {
            const u32x4 *w = static_cast<const u32x4 *>(data);
            blk.insert(blk.end(), w, w + words);
        };
        size_t i = 0;
        while (i < chunks.size()) {
            size_t j = i, nrec = 0;
            while (j < chunks.size() && nrec + chunks[j].r64.size() <= RMAX) nrec += chunks[j++].r64.size();
            std::vector<uint16_t> header(HS * 16, 0);
            uint32_t r = HS;
            size_t q = i;
            for (uint32_t z = 0; z <= A; ++z) {
                while (q < j && chunks[q].zs < z) r += static_cast<uint32_t>(chunks[q++].r64.size());
                header[z] = static_cast<uint16_t>(z == A ? HS + nrec : r);
            }
            cls.push_back(TileClass{off, static_cast<uint32_t>(blk64.size()), static_cast<uint32_t>(2 * (HS + nrec))});
            push_words(blk64, header.data(), 2 * HS);
            push_words(blk32, header.data(), 2 * HS);
            for (size_t c = i; c < j; ++c) {
                push_words(blk64, chunks[c].r64.data(), 2 * chunks[c].r64.size());
                push_words(blk32, chunks[c].r32.data(), 2 * chunks[c].r32.size());
            }
            i = j;
        }
    }

    // Keeps every device array non-empty (a zero-size upload has no pointer).
    void pad_empty() {
        if (cls.empty()) cls.push_back(TileClass{0, 0, 0});
        if (blk64.empty()) {
            blk64.push_back(u32x4{0u, 0u, 0u, 0u});
            blk32.push_back(u32x4{0u, 0u, 0u, 0u});
        }
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
        return cls.size() * sizeof(TileClass) + (blk64.size() + blk32.size()) * sizeof(u32x4) +
               xorder.size() * sizeof(uint32_t) +
               (grp64.size() + grp32.size()) * sizeof(TileGroup) +
               (term64.size() + term32.size()) * sizeof(TileTerm);
    }
};

}  // namespace tcp
