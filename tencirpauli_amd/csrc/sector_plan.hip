// Restricted (fixed particle number) MVP: U(1) sectors and contiguous
// fixed-population block sectors, plan builder and gfx950 kernels.
//
// Replaces U1LazyMvpPlan::apply_into (core/sector.rs:579-642) and the
// factorized charge path of apply_charge_mvp_from_prepared_plan_into
// (core/charge/generic.rs:854-1051) for contiguous fixed-population blocks
// (core/charge_sector.rs:150-217, 419-456).
//
// The reference scatters serially: for each source state, for each x-mask
// group, value = sum_k w_k (-1)^popc(z_k & src); exact zeros skipped;
// output[rank(src ^ x)] += value * state[src].  Here each thread owns R
// consecutive destination ranks and gathers: for destination bits b and
// group x the source is s = b ^ x; s contributes iff it lies in the sector
// (the pair (s, b) is then exactly one of the reference's transitions), the
// same exact-zero skip applies, and rank(s) selects the input amplitude.
// No atomics; the output is written once.  Exact conservation is checked when
// the plan is built (py/symmetry.py:675-681, py/charge.py:1128-1132), so no
// transition can leave the sector.
//
// Ranks.  Block j occupies a contiguous field of the index (qubit 0 = MSB,
// block 0 highest) and holds pop_j ones; the sector rank is mixed radix over
// blocks with block 0 most significant, and a block's rank is the
// combinatorial number system sum_i C(p_i, i) over its set bits p_1 < p_2 < ..
// (ascending integer order: the order of U1Sector::rank, core/sector.rs:127-167,
// and FixedPopulationIndex::rank_contiguous_binary_word,
// core/charge_sector.rs:973-1007).  The kernel evaluates it one byte at a
// time from an LDS table lut[chunk][ones below][byte]; unranking walks the
// bits from the top against C(p, needed) and later ranks of a thread follow
// by the next-larger-combination step (ascending order).
#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <type_traits>

#include "full_matrix.hpp"
#include "tcp_internal.hpp"
#include "charge_sector.hpp"
#include "wide_sector.hpp"

namespace {

constexpr uint32_t kMaxBlocks = 4;
constexpr uint32_t kMaxChunksPerBlock = 8;  // byte chunks of a block of up to 64 qubits
constexpr uint32_t kSectorThreads = 256;
// The gather kernel's workgroups hold 8 waves, so the LDS copy of the rank
// tables (one per workgroup) no longer caps the waves per CU below 8 per SIMD.
constexpr uint32_t kGatherThreads = 512;
// Consecutive destination ranks per thread (one unrank, then successor
// steps): 2 against 1 and 4 is 3-6% faster on the 5x5 XXZ and 4x4 Hubbard
// sectors.
constexpr uint32_t kRanksPerThread = 2;
constexpr uint32_t kLdsLutLimit = 64 * 1024;  // bytes

template <class T>
using cptr = const __attribute__((address_space(4))) T *;
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

struct SectorArgs {
    uint32_t nblocks;
    uint32_t shift[kMaxBlocks];    // field position of block j in the index
    uint32_t size[kMaxBlocks];
    uint32_t pop[kMaxBlocks];
    uint32_t nchunk[kMaxBlocks];   // bytes per block field
    uint32_t lut_off[kMaxBlocks];  // offset of block j's table in lut
    uint32_t stride[kMaxBlocks];   // mixed-radix stride of block j
    uint64_t bmask[kMaxBlocks];    // (1 << size) - 1
    uint64_t dim;
    uint32_t G;
    uint32_t lut_words;
    const uint64_t *gx;      // [G] index-space x-masks
    const uint32_t *goff;    // [G+1]
    const uint64_t *tz;      // [T] index-space z-masks
    const void *tw;          // [T] (re, im) of w = c * i^(#Y mod 4)
    const uint64_t *tc1;     // [T] fermion terms: modes that must be occupied in the source (null: Pauli)
    const uint64_t *tc0;     // [T] fermion terms: modes that must be empty in the source
    const uint32_t *lut;     // rank tables
    const uint32_t *cbase;   // [kMaxBlocks][8] table base of (block, chunk), offset by its lowest row
    const uint64_t *choose;  // [64 * 64] C(p, j), p, j < 64
};

__device__ __forceinline__ double flip_sign(double x, uint32_t parity) {
    return __longlong_as_double(__double_as_longlong(x) ^ (static_cast<long long>(parity) << 63));
}

// A fermion ladder word acts on source s only when its annihilated modes are
// occupied and its other created modes empty; then its value is
// w (-1)^popc(z & s) like a Pauli term's.  Pauli terms (COND false) always act.
template <bool COND>
__device__ __forceinline__ bool term_acts(const SectorArgs &p, uint32_t k, uint64_t s) {
    if (!COND) return true;
    const uint64_t need = ((cptr<uint64_t>)p.tc1)[k];
    return (s & need) == need && (s & ((cptr<uint64_t>)p.tc0)[k]) == 0;
}

// The rank/unrank/successor helpers are __host__ __device__ so the same code
// can be exercised on the CPU against the oracle.
__host__ __device__ __forceinline__ bool in_sector(const SectorArgs &p, uint64_t s) {
    bool ok = true;
    for (uint32_t j = 0; j < p.nblocks; ++j)
        ok &= static_cast<uint32_t>(__builtin_popcountll((s >> p.shift[j]) & p.bmask[j])) == p.pop[j];
    return ok;
}

// Rank of block j's bits of s among the block's fixed-population patterns,
// times the block's mixed-radix stride (the rank-table entries carry it).
__host__ __device__ __forceinline__ uint32_t block_rank(const SectorArgs &p, const uint32_t *lut, uint32_t j,
                                                         uint64_t s) {
    const uint64_t w = (s >> p.shift[j]) & p.bmask[j];
    const uint32_t *base = p.cbase + j * kMaxChunksPerBlock;
    uint32_t local = 0, below = 0;
    for (uint32_t c = 0; c < p.nchunk[j]; ++c) {
        const uint32_t byte = static_cast<uint32_t>(w >> (8 * c)) & 0xffu;
        local += lut[base[c] + below * 256 + byte];
        below += __builtin_popcount(byte);
    }
    return local;
}

__host__ __device__ __forceinline__ uint32_t sector_rank(const SectorArgs &p, const uint32_t *lut, uint64_t s) {
    uint32_t rank = 0;
    for (uint32_t j = 0; j < p.nblocks; ++j) rank += block_rank(p, lut, j, s);  // entries carry the stride
    return rank;
}

// NB > 0 fixes the block count at compile time (0: p.nblocks).
template <uint32_t NB = 0>
__host__ __device__ uint64_t sector_unrank(const SectorArgs &p, uint64_t rank) {
    const uint64_t *choose = p.choose;
    uint64_t bits = 0;
    const uint32_t nb = NB ? NB : p.nblocks;
    for (uint32_t j = 0; j < nb; ++j) {
        uint64_t local = rank / p.stride[j];
        rank -= local * p.stride[j];
        uint32_t needed = p.pop[j];
        uint64_t w = 0;
        for (int q = static_cast<int>(p.size[j]) - 1; q >= 0 && needed > 0; --q) {
            const uint64_t c = choose[q * 64 + needed];  // C(q, needed)
            if (local >= c) {
                local -= c;
                w |= 1ull << q;
                --needed;
            }
        }
        bits |= w << p.shift[j];
    }
    return bits;
}

// Next sector state in ascending rank order: advance the lowest block to its
// next combination (Gosper's step); on overflow reset it and carry upwards.
template <uint32_t NB = 0>
__host__ __device__ __forceinline__ uint64_t sector_next(const SectorArgs &p, uint64_t bits) {
    for (int j = static_cast<int>(NB ? NB : p.nblocks) - 1; j >= 0; --j) {
        const uint64_t w = (bits >> p.shift[j]) & p.bmask[j];
        uint64_t nw = ~0ull;
        if (p.pop[j] > 0) {
            const uint64_t c = w & (~w + 1);
            const uint64_t r = w + c;
            nw = r | (((r ^ w) >> 2) >> (__builtin_ffsll(static_cast<long long>(c)) - 1));
        }
        bits &= ~(p.bmask[j] << p.shift[j]);
        if (p.pop[j] > 0 && nw <= p.bmask[j]) return bits | (nw << p.shift[j]);
        const uint64_t first = p.pop[j] ? ((1ull << p.pop[j]) - 1) : 0;
        bits |= first << p.shift[j];
    }
    return bits;
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
    __device__ static fvec2 make(double r, double i) {
        return fvec2{static_cast<float>(r), static_cast<float>(i)};
    }
};

// Groups evaluated before their gathers are issued together.  Against one
// group at a time, 4 takes 9% off the 5x5 XXZ sector MVP and adds 4% to the
// 4x4 Hubbard one (whose groups mostly leave the sector); 8 adds 19% there.
constexpr uint32_t kGatherChunk = 4;

// The gather kernel's own tables (built once per plan by gather_tables):
//  * byte chunks flattened over the blocks -- chunk t covers index bits
//    cshift[t] .. cshift[t] + 7 of one block -- so a destination keeps each
//    chunk's rank-table entry in a register and a group re-ranks, in each
//    block, only the chunks from the lowest to the highest one its x-mask
//    flips: the chunks below are unchanged, and above them the ones-below
//    count is unchanged because the flip keeps the block's population;
//  * one 64-byte record per group (a few scalar loads): Pauli terms whose z
//    covers all of x or none of it are summed per z & ~x, their signs on an
//    in-sector source being fixed up to popc(z & ~x & s) -- the XX + YY of a
//    bond becomes one constant, a Jordan-Wigner hop one term -- and one
//    term sits inline; further terms and fermion words are in a term list;
//  * groups that cannot keep every block's population (an odd number of
//    flipped bits in a block, or more ones or zeros to move than the block
//    holds) are left out: they never reach the sector;
//  * the diagonal stored per state (when the memory budget allows), which
//    takes the x = 0 group out of the loop.
constexpr uint32_t kMaxFlatChunks = 12;  // sum of ceil(size / 8) over <= 4 blocks of <= 63 qubits
// sum of ceil(size / 8) over NB blocks of 32 bits in all: ceil(s / 8) <= (s + 7) / 8
__host__ __device__ constexpr uint32_t narrow_chunk_count(uint32_t nb) { return 3 + nb; }
constexpr uint32_t kInlineTerm = 1u << 31;

struct GatherGroup {
    uint64_t x;
    uint32_t chunks;  // bits 0..11: flat chunks to re-rank; 12..23: those that start a run; 24..27: blocks x touches
    uint32_t halves;  // block q: bits 8q .. 8q+7 half the number of bits x flips in block q
    double cre, cim;  // the folded z = 0 terms
    uint64_t z1;      // the inline term (when nterms has kInlineTerm)
    double w1re, w1im;
    uint32_t toff;    // further terms: term list [toff, toff + (nterms & ~kInlineTerm))
    uint32_t nterms;
};
static_assert(sizeof(GatherGroup) == 64, "one scalar load per group");

struct GatherArgs {
    uint32_t G;       // groups the kernel visits
    uint32_t nflat;   // flat chunks
    uint32_t cfirst;  // bit t: chunk t is its block's lowest
    // chunk t is byte (index >> cshift[t]) & 0xff: a block's last chunk may
    // carry low bits of the next block up, which its rank-table entries
    // ignore (they count positions inside the block only), and no chunk of
    // the same block follows it to need its ones count
    uint32_t cshift[kMaxFlatChunks];
    uint32_t cbase[kMaxFlatChunks];    // rank-table row base of chunk t (entries carry the block's stride)
    uint64_t blockmask[kMaxBlocks];    // index bits of block q
    const GatherGroup *groups;         // [G]
    const uint64_t *tz;
    const void *tw;
    const uint64_t *tc1, *tc0;  // fermion words: occupancy conditions (null: Pauli)
    const double *diag;         // [dim] stored diagonal, or null
};

// A group record read field by field (one scalar load after merging).
__device__ __forceinline__ GatherGroup load_group(cptr<GatherGroup> groups, uint32_t g) {
    GatherGroup r;
    r.x = groups[g].x;
    r.chunks = groups[g].chunks;
    r.halves = groups[g].halves;
    r.cre = groups[g].cre;
    r.cim = groups[g].cim;
    r.z1 = groups[g].z1;
    r.w1re = groups[g].w1re;
    r.w1im = groups[g].w1im;
    r.toff = groups[g].toff;
    r.nterms = groups[g].nterms;
    return r;
}

template <bool COND, typename IT>
__device__ __forceinline__ bool gather_term_acts(const GatherArgs &q, uint32_t k, IT s) {
    if (!COND) return true;
    const IT need = static_cast<IT>(((cptr<uint64_t>)q.tc1)[k]);
    return (s & need) == need && (s & static_cast<IT>(((cptr<uint64_t>)q.tc0)[k])) == 0;
}

__device__ __forceinline__ uint32_t popcount_of(uint32_t x) { return static_cast<uint32_t>(__popc(x)); }
__device__ __forceinline__ uint32_t popcount_of(uint64_t x) { return static_cast<uint32_t>(__popcll(x)); }

// NB = number of blocks, fixed at compile time so the per-group loops keep
// their block and chunk fields in registers.  RPT = consecutive destination
// ranks per thread (one unrank, then successor steps): 1 for one-block (U(1))
// sectors, 2 for block sectors.  IT holds a state's index bits: uint32_t for
// spaces of up to 32 qubits (half the integer instructions of the per-group
// bit work), else uint64_t.
template <typename VT, bool LDS, bool COND, uint32_t NB, uint32_t RPT, typename IT>
__global__ __launch_bounds__(kGatherThreads) void sector_gather_kernel(const VT *__restrict__ v,
                                                                       VT *__restrict__ y, SectorArgs p,
                                                                       GatherArgs q) {
    // >= sum of ceil(size / 8) over the blocks: at most 3 + NB chunks fit 32 index bits
    constexpr uint32_t F = sizeof(IT) == 4 ? narrow_chunk_count(NB) : (7 + NB < kMaxFlatChunks ? 7 + NB : kMaxFlatChunks);
    // groups in flight together: with several blocks, four 64-byte records
    // at once would not fit the scalar registers
    constexpr uint32_t GC = NB == 1 ? kGatherChunk : 2;
    extern __shared__ uint32_t lds_lut[];
    const uint32_t *lut = p.lut;
    if (LDS) {
        for (uint32_t i = threadIdx.x; i < p.lut_words; i += blockDim.x) lds_lut[i] = p.lut[i];
        __syncthreads();
        lut = lds_lut;
    }
    cptr<GatherGroup> groups = (cptr<GatherGroup>)q.groups;
    cptr<uint64_t> tz = (cptr<uint64_t>)q.tz;
    cptr<dvec2> tw = (cptr<dvec2>)q.tw;

    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x * RPT;
    for (uint64_t r0 = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * RPT;
         r0 < p.dim; r0 += stride) {
        uint64_t b64 = sector_unrank<NB>(p, r0);
        for (uint32_t j = 0; j < RPT; ++j) {
            const IT b = static_cast<IT>(b64);
            const uint64_t r = r0 + j;
            if (r >= p.dim) break;
            // rank-table entry and ones-below count of every chunk of b
            uint32_t contrib[F], below_at[F];
            uint32_t below = 0;
#pragma unroll
            for (uint32_t t = 0; t < F; ++t) {
                contrib[t] = 0;
                below_at[t] = 0;
                if (t < q.nflat) {
                    if ((q.cfirst >> t) & 1u) below = 0;
                    const uint32_t byte = static_cast<uint32_t>(b >> q.cshift[t]) & 0xffu;
                    contrib[t] = lut[q.cbase[t] + below * 256u + byte];
                    below_at[t] = below;
                    below += __popc(byte);
                }
            }
            double ar = 0.0, ai = 0.0;
            if (q.diag) {
                const double d = q.diag[r];
                const VT sv = v[r];
                ar = d * Amp<VT>::re(sv);
                ai = d * Amp<VT>::im(sv);
            }
            // GC groups at a time: their source ranks and weights
            // first, then all their gathers in flight together (a group that
            // leaves the sector or sums to an exact zero -- skipped as
            // core/sector.rs:1099 does -- gathers v[r] with weight zero)
            for (uint32_t g0 = 0; g0 < q.G; g0 += GC) {
                uint64_t rs[GC];
                double wr[GC], wi[GC];
                uint32_t used = 0;  // bit u: group g0 + u contributes
#pragma unroll
                for (uint32_t u = 0; u < GC; ++u) {
                    const uint32_t g = g0 + u;
                    const bool live = g < q.G;
                    const GatherGroup gr = load_group(groups, live ? g : 0);
                    const IT gx = static_cast<IT>(gr.x);
                    const IT s = b ^ gx;
                    // s keeps every block's population when, inside each
                    // block, half of the bits x flips are ones in b
                    bool ok = live;
#pragma unroll
                    for (uint32_t k = 0; k < NB; ++k) {
                        if (!((gr.chunks >> (24 + k)) & 1u)) continue;
                        const IT xk = NB == 1 ? gx : gx & static_cast<IT>(q.blockmask[k]);
                        ok &= popcount_of(b & xk) == ((gr.halves >> (8 * k)) & 0xffu);
                    }
                    double dr = gr.cre, di = gr.cim;
                    if (ok && gr.nterms != 0) {  // else: a nonzero constant
                        if (gr.nterms & kInlineTerm) {
                            const uint32_t par = popcount_of(static_cast<IT>(gr.z1) & s) & 1u;
                            dr += flip_sign(gr.w1re, par);
                            di += flip_sign(gr.w1im, par);
                        }
                        const uint32_t k1 = gr.toff + (gr.nterms & ~kInlineTerm);
                        for (uint32_t k = gr.toff; k < k1; ++k) {
                            if (!gather_term_acts<COND>(q, k, s)) continue;
                            const uint32_t par = popcount_of(static_cast<IT>(tz[k]) & s) & 1u;
                            const dvec2 w = tw[k];
                            dr += flip_sign(w.x, par);
                            di += flip_sign(w.y, par);
                        }
                        ok = dr != 0.0 || di != 0.0;
                    }
                    uint64_t src = r;
                    if (ok && gx != 0) {
                        // the chunks of each block from the lowest to the highest
                        // one x flips; a run starts from b's ones-below count
                        uint32_t delta = 0, nb = 0;
#pragma unroll
                        for (uint32_t t = 0; t < F; ++t) {
                            if ((gr.chunks >> t) & 1u) {
                                if ((gr.chunks >> (12 + t)) & 1u) nb = below_at[t];
                                const uint32_t byte = static_cast<uint32_t>(s >> q.cshift[t]) & 0xffu;
                                delta += lut[q.cbase[t] + nb * 256u + byte] - contrib[t];
                                nb += __popc(byte);
                            }
                        }
                        src = static_cast<uint32_t>(static_cast<uint32_t>(r) + delta);
                        ok = src < p.dim;  // unreachable for in-sector s; keeps the gather in bounds
                        if (!ok) src = r;
                    }
                    // a group that does not contribute keeps its (finite) weight: its
                    // gathered value is replaced by zero below
                    rs[u] = src;
                    wr[u] = dr;
                    wi[u] = di;
                    used |= static_cast<uint32_t>(ok) << u;
                }
                VT sv[GC];
#pragma unroll
                for (uint32_t u = 0; u < GC; ++u) sv[u] = v[rs[u]];
#pragma unroll
                for (uint32_t u = 0; u < GC; ++u) {
                    // unused slots add an exact zero whatever v[r] holds
                    const bool on = (used >> u) & 1u;
                    const double sr = on ? Amp<VT>::re(sv[u]) : 0.0, si = on ? Amp<VT>::im(sv[u]) : 0.0;
                    ar = fma(wr[u], sr, ar);
                    ar = fma(-wi[u], si, ar);
                    ai = fma(wr[u], si, ai);
                    ai = fma(wi[u], sr, ai);
                }
            }
            y[r] = Amp<VT>::make(ar, ai);
            b64 = sector_next<NB>(p, b64);
        }
    }
}

__global__ void sector_rank_kernel(SectorArgs p, const uint64_t *__restrict__ bits,
                                   uint64_t *__restrict__ ranks, uint64_t count, uint32_t nqubits) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t s = bits[i];
        const bool inside = (nqubits == 64 || (s >> nqubits) == 0) && in_sector(p, s);
        ranks[i] = inside ? sector_rank(p, p.lut, s) : ~0ull;
    }
}

__global__ void sector_unrank_kernel(SectorArgs p, const uint64_t *__restrict__ ranks,
                                     uint64_t *__restrict__ bits, uint64_t count) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t r = ranks[i];
        bits[i] = r < p.dim ? sector_unrank(p, r) : ~0ull;
    }
}

// Consecutive ranks through the successor step, for bit-exact sweeps of the
// iterator (core/sector.rs:751-834): bits[i] = state of rank first + i.
__global__ void sector_walk_kernel(SectorArgs p, uint64_t first, uint64_t *__restrict__ bits,
                                   uint64_t count, uint32_t run) {
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    const uint64_t start = t * run;
    if (start >= count) return;
    uint64_t b = sector_unrank(p, first + start);
    for (uint32_t j = 0; j < run && start + j < count; ++j) {
        bits[start + j] = b;
        b = sector_next(p, b);
    }
}

// Leakage scan in the reference's scatter order (core/sector.rs:1051-1127,
// core/charge/generic.rs:1013-1041): for every source state and x-group whose
// destination leaves the sector, the group's summed value must be an exact
// zero.  The smallest (source rank, group) pair that is not is recorded as
// rank * G + group, which is the first error the reference's serial loop
// raises.  State-independent, so it runs once per plan.
template <bool COND>
__global__ __launch_bounds__(kSectorThreads) void sector_leak_kernel(SectorArgs p,
                                                                     unsigned long long *first) {
    cptr<uint64_t> gx = (cptr<uint64_t>)p.gx;
    cptr<uint32_t> goff = (cptr<uint32_t>)p.goff;
    cptr<uint64_t> tz = (cptr<uint64_t>)p.tz;
    cptr<dvec2> tw = (cptr<dvec2>)p.tw;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x * kRanksPerThread;
    for (uint64_t r0 = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * kRanksPerThread;
         r0 < p.dim; r0 += stride) {
        uint64_t b = sector_unrank(p, r0);
        for (uint32_t j = 0; j < kRanksPerThread; ++j) {
            const uint64_t r = r0 + j;
            if (r >= p.dim) break;
            for (uint32_t g = 0; g < p.G; ++g) {
                const uint64_t s = b ^ gx[g];
                if (in_sector(p, s)) continue;
                double dr = 0.0, di = 0.0;
                const uint32_t k1 = goff[g + 1];
                for (uint32_t k = goff[g]; k < k1; ++k) {
                    if (!term_acts<COND>(p, k, b)) continue;
                    const uint32_t par = __popcll(tz[k] & b) & 1u;
                    const dvec2 w = tw[k];
                    dr += flip_sign(w.x, par);
                    di += flip_sign(w.y, par);
                }
                if (dr != 0.0 || di != 0.0) {
                    atomicMin(first, static_cast<unsigned long long>(r) * p.G + g);
                    break;
                }
            }
            b = sector_next(p, b);
        }
    }
}

// Real part of the diagonal H_rr: the x = 0 group evaluated on each state.
template <bool COND>
__global__ void sector_diagonal_kernel(SectorArgs p, uint32_t diag_group, double *__restrict__ diag) {
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    const uint64_t start = t * kRanksPerThread;
    if (start >= p.dim) return;
    uint64_t b = sector_unrank(p, start);
    for (uint32_t j = 0; j < kRanksPerThread && start + j < p.dim; ++j) {
        double d = 0.0;
        if (diag_group < p.G) {
            const dvec2 *tw = static_cast<const dvec2 *>(p.tw);
            for (uint32_t k = p.goff[diag_group]; k < p.goff[diag_group + 1]; ++k)
                if (term_acts<COND>(p, k, b)) d += flip_sign(tw[k].x, __popcll(p.tz[k] & b) & 1u);
        }
        diag[start + j] = d;
        b = sector_next(p, b);
    }
}

// Per-state features of the learned preconditioner (64 bf16 per state, state-
// local and independent of the lattice size).  For state b with diagonal D_b:
// every x-group g whose source s = b ^ x_g stays in the sector contributes a
// coupling c = |sum_k w_k (-1)^popc(z_k & s)| and an excitation energy
// dE = D_s - D_b.  Slots: 0 = (D_b - theta) / scale, filled at inference;
// 1 = #couplings / 16; 2 = sum c^2 / scale^2 / 16; 3..26 = c^2-weighted
// histogram of dE / scale over 24 bins on [-6, 6); 27..50 = unweighted
// histogram / 8; 51..53 = mean, min, max of dE / scale; 54 = sum c^2 / dE
// over dE > 0 (second-order shift) / scale; the rest zero.
constexpr uint32_t kFeatures = 64;

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

template <bool COND>
__device__ double diag_at(const SectorArgs &p, uint32_t diag_group, uint64_t s) {
    double d = 0.0;
    if (diag_group < p.G) {
        const dvec2 *tw = static_cast<const dvec2 *>(p.tw);
        for (uint32_t k = p.goff[diag_group]; k < p.goff[diag_group + 1]; ++k)
            if (term_acts<COND>(p, k, s)) d += flip_sign(tw[k].x, __popcll(p.tz[k] & s) & 1u);
    }
    return d;
}

template <bool COND>
__global__ void sector_features_kernel(SectorArgs p, uint32_t diag_group, double inv_scale,
                                       uint16_t *__restrict__ features, double *__restrict__ diag) {
    const uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (r >= p.dim) return;
    const uint64_t b = sector_unrank(p, r);
    const double db = diag_at<COND>(p, diag_group, b);
    float f[kFeatures];
    for (uint32_t i = 0; i < kFeatures; ++i) f[i] = 0.0f;
    double ncoup = 0.0, c2sum = 0.0, desum = 0.0, demin = 1e300, demax = -1e300, shift = 0.0;
    const dvec2 *tw = static_cast<const dvec2 *>(p.tw);
    for (uint32_t g = 0; g < p.G; ++g) {
        if (g == diag_group) continue;
        const uint64_t s = b ^ p.gx[g];
        if (!in_sector(p, s)) continue;
        double vr = 0.0, vi = 0.0;
        for (uint32_t k = p.goff[g]; k < p.goff[g + 1]; ++k) {
            if (!term_acts<COND>(p, k, s)) continue;
            const uint32_t par = __popcll(p.tz[k] & s) & 1u;
            vr += flip_sign(tw[k].x, par);
            vi += flip_sign(tw[k].y, par);
        }
        const double c2 = (vr * vr + vi * vi) * inv_scale * inv_scale;
        if (c2 == 0.0) continue;
        const double de = (diag_at<COND>(p, diag_group, s) - db) * inv_scale;
        ncoup += 1.0;
        c2sum += c2;
        desum += de;
        demin = fmin(demin, de);
        demax = fmax(demax, de);
        if (de > 1e-3) shift += c2 / de;
        int bin = static_cast<int>(floor((de + 6.0) * 2.0));
        bin = bin < 0 ? 0 : (bin > 23 ? 23 : bin);
        f[3 + bin] += static_cast<float>(c2);
        f[27 + bin] += 1.0f / 8.0f;
    }
    f[1] = static_cast<float>(ncoup / 16.0);
    f[2] = static_cast<float>(c2sum / 16.0);
    if (ncoup > 0) {
        f[51] = static_cast<float>(desum / ncoup);
        f[52] = static_cast<float>(demin);
        f[53] = static_cast<float>(demax);
    }
    f[54] = static_cast<float>(shift);
    uint16_t *out = features + r * kFeatures;
    for (uint32_t i = 0; i < kFeatures; ++i) out[i] = to_bf16(f[i]);
    diag[r] = db;
}

// ---- eager restricted operator: CSR with rows = destination ranks ----------
//
// Replaces U1RestrictedOperator::new / CsrGraph (core/sector.rs:345-432,
// core/csr.rs:196-240) and the charge eager build (core/charge/eager.rs).
// Row r holds, for every in-sector source s = b ^ x_g with a nonzero
// aggregated value, (rank(s), value); columns ascend inside a row, the order
// the reference's source-major two-pass build produces.

constexpr uint32_t kCsrSortCap = 64;  // rows up to this length are sorted in registers

template <bool LDS, bool COND>
__global__ __launch_bounds__(kSectorThreads) void csr_count_kernel(SectorArgs p, uint32_t *__restrict__ counts) {
    extern __shared__ uint32_t lds_lut[];
    if (LDS) {
        for (uint32_t i = threadIdx.x; i < p.lut_words; i += blockDim.x) lds_lut[i] = p.lut[i];
        __syncthreads();
    }
    cptr<uint64_t> gx = (cptr<uint64_t>)p.gx;
    cptr<uint32_t> goff = (cptr<uint32_t>)p.goff;
    cptr<uint64_t> tz = (cptr<uint64_t>)p.tz;
    cptr<dvec2> tw = (cptr<dvec2>)p.tw;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x * kRanksPerThread;
    for (uint64_t r0 = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * kRanksPerThread;
         r0 < p.dim; r0 += stride) {
        uint64_t b = sector_unrank(p, r0);
        for (uint32_t j = 0; j < kRanksPerThread; ++j) {
            const uint64_t r = r0 + j;
            if (r >= p.dim) break;
            uint32_t count = 0;
            for (uint32_t g = 0; g < p.G; ++g) {
                const uint64_t s = b ^ gx[g];
                if (!in_sector(p, s)) continue;
                double dr = 0.0, di = 0.0;
                for (uint32_t k = goff[g]; k < goff[g + 1]; ++k) {
                    if (!term_acts<COND>(p, k, s)) continue;
                    const uint32_t par = __popcll(tz[k] & s) & 1u;
                    dr += flip_sign(tw[k].x, par);
                    di += flip_sign(tw[k].y, par);
                }
                count += (dr != 0.0 || di != 0.0);
            }
            counts[r] = count;
            b = sector_next(p, b);
        }
    }
}

template <bool LDS, bool COND>
__global__ __launch_bounds__(kSectorThreads) void csr_fill_kernel(SectorArgs p, const uint64_t *__restrict__ indptr,
                                                                  uint32_t *__restrict__ cols, dvec2 *__restrict__ vals,
                                                                  fvec2 *__restrict__ vals32) {
    extern __shared__ uint32_t lds_lut[];
    const uint32_t *lut = p.lut;
    if (LDS) {
        for (uint32_t i = threadIdx.x; i < p.lut_words; i += blockDim.x) lds_lut[i] = p.lut[i];
        __syncthreads();
        lut = lds_lut;
    }
    cptr<uint64_t> gx = (cptr<uint64_t>)p.gx;
    cptr<uint32_t> goff = (cptr<uint32_t>)p.goff;
    cptr<uint64_t> tz = (cptr<uint64_t>)p.tz;
    cptr<dvec2> tw = (cptr<dvec2>)p.tw;
    // Sum of group g's terms at source s (zero when none acts).
    auto group_value = [&](uint32_t g, uint64_t s) -> dvec2 {
        double dr = 0.0, di = 0.0;
        for (uint32_t k = goff[g]; k < goff[g + 1]; ++k) {
            if (!term_acts<COND>(p, k, s)) continue;
            const uint32_t par = __popcll(tz[k] & s) & 1u;
            dr += flip_sign(tw[k].x, par);
            di += flip_sign(tw[k].y, par);
        }
        return dvec2{dr, di};
    };
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x * kRanksPerThread;
    for (uint64_t r0 = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * kRanksPerThread;
         r0 < p.dim; r0 += stride) {
        uint64_t b = sector_unrank(p, r0);
        for (uint32_t j = 0; j < kRanksPerThread; ++j) {
            const uint64_t r = r0 + j;
            if (r >= p.dim) break;
            const uint64_t pos = indptr[r];
            const uint64_t len = indptr[r + 1] - pos;
            if (len <= kCsrSortCap) {
                // Rows up to kCsrSortCap entries: keys column << 32 | group
                // kept sorted in registers (every index below is static, so
                // the array stays out of private memory), then emitted in
                // order with each group's value recomputed.
                uint64_t keys[kCsrSortCap];
#pragma unroll
                for (uint32_t q = 0; q < kCsrSortCap; ++q) keys[q] = ~0ull;
                uint32_t n = 0;
                for (uint32_t g = 0; g < p.G && n < len; ++g) {
                    const uint64_t x = gx[g];
                    const uint64_t s = b ^ x;
                    if (!in_sector(p, s)) continue;
                    const dvec2 v = group_value(g, s);
                    if (v.x == 0.0 && v.y == 0.0) continue;
                    const uint32_t c = x == 0 ? static_cast<uint32_t>(r) : sector_rank(p, lut, s);
                    uint64_t k = (static_cast<uint64_t>(c) << 32) | g;
#pragma unroll
                    for (uint32_t q = 0; q < kCsrSortCap; ++q) {  // insertion by compare-exchange
                        const uint64_t lo = keys[q] < k ? keys[q] : k;
                        k = keys[q] < k ? k : keys[q];
                        keys[q] = lo;
                    }
                    ++n;
                }
                for (uint32_t q = 0; q < n; ++q) {
                    const uint32_t c = static_cast<uint32_t>(keys[0] >> 32);
                    const uint32_t g = static_cast<uint32_t>(keys[0]);
#pragma unroll
                    for (uint32_t i = 0; i + 1 < kCsrSortCap; ++i) keys[i] = keys[i + 1];
                    const dvec2 v = group_value(g, b ^ gx[g]);
                    cols[pos + q] = c;
                    vals[pos + q] = v;
                    vals32[pos + q] = fvec2{static_cast<float>(v.x), static_cast<float>(v.y)};
                }
            } else {  // long row: written in group order, then insertion-sorted in place
                uint32_t n = 0;
                for (uint32_t g = 0; g < p.G && n < len; ++g) {
                    const uint64_t x = gx[g];
                    const uint64_t s = b ^ x;
                    if (!in_sector(p, s)) continue;
                    const dvec2 v = group_value(g, s);
                    if (v.x == 0.0 && v.y == 0.0) continue;
                    cols[pos + n] = x == 0 ? static_cast<uint32_t>(r) : sector_rank(p, lut, s);
                    vals[pos + n] = v;
                    ++n;
                }
                for (uint32_t q = 1; q < n; ++q) {
                    const uint32_t c = cols[pos + q];
                    const dvec2 v = vals[pos + q];
                    uint32_t i = q;
                    while (i > 0 && cols[pos + i - 1] > c) {
                        cols[pos + i] = cols[pos + i - 1];
                        vals[pos + i] = vals[pos + i - 1];
                        --i;
                    }
                    cols[pos + i] = c;
                    vals[pos + i] = v;
                }
                for (uint32_t q = 0; q < n; ++q) {
                    const dvec2 v = vals[pos + q];
                    vals32[pos + q] = fvec2{static_cast<float>(v.x), static_cast<float>(v.y)};
                }
            }
            b = sector_next(p, b);
        }
    }
}

// y[row] = sum_k vals[k] v[cols[k]] over the row, LANES lanes per row (an
// aligned group inside one wave), FP64 accumulation, lanes reduced by
// cross-lane exchange; every output row is written once.
template <typename VT, typename WT, uint32_t LANES, uint32_t kUnroll = 4>
__global__ __launch_bounds__(256) void csr_spmv_kernel(const uint64_t *__restrict__ indptr,
                                                       const uint32_t *__restrict__ cols,
                                                       const WT *__restrict__ vals, const VT *__restrict__ v,
                                                       VT *__restrict__ y, uint64_t dim) {
    const uint32_t lane = threadIdx.x % LANES;
    const uint64_t rows_per_step = static_cast<uint64_t>(gridDim.x) * (blockDim.x / LANES);
    // the LANES lanes of a row share `row`, so they enter and leave the loop
    // together and every exchange partner is active
    for (uint64_t row = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / LANES) + threadIdx.x / LANES; row < dim;
         row += rows_per_step) {
        double ar = 0.0, ai = 0.0;
        const uint64_t k0 = indptr[row], k1 = indptr[row + 1];
        // kUnroll entries per lane per step: their index and value loads, then
        // their gathers, are all in flight together (slots past the row end
        // read column 0 with a zero weight)
        for (uint64_t kb = k0 + lane; kb < k1; kb += kUnroll * LANES) {
            uint32_t c[kUnroll];
            WT w[kUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
                const uint64_t k = kb + u * LANES;
                const bool in = k < k1;
                c[u] = in ? cols[k] : 0u;
                w[u] = in ? vals[k] : WT{0, 0};
            }
            VT x[kUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) x[u] = v[c[u]];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
                const bool in = kb + u * LANES < k1;  // padded slots add an exact zero
                const double cr = static_cast<double>(w[u].x), ci = static_cast<double>(w[u].y);
                const double sr = in ? static_cast<double>(x[u].x) : 0.0, si = in ? static_cast<double>(x[u].y) : 0.0;
                ar = fma(cr, sr, ar);
                ar = fma(-ci, si, ar);
                ai = fma(cr, si, ai);
                ai = fma(ci, sr, ai);
            }
        }
#pragma unroll
        for (uint32_t off = LANES / 2; off > 0; off >>= 1) {
            ar += __shfl_xor(ar, off, LANES);
            ai += __shfl_xor(ai, off, LANES);
        }
        if (lane == 0) y[row] = Amp<VT>::make(ar, ai);
    }
}

// ---------------------------------------------------------------- compact CSR
// Lattice operators have few distinct off-diagonal values (the 5x5 XXZ sector
// has one), so the off-diagonal values become an 8-bit index into a table of
// at most kDictMax values, packed with a 24-bit column into one word per
// entry (4 bytes instead of 4 + 16); index kDiagIndex marks the row's diagonal
// entry, whose value sits in a dense per-row array.  The table is found by a
// hash set of kDictSlots slots (hash of the value's bits; 0 marks an empty
// slot), filled on the device.
constexpr uint32_t kDictSlots = 1024;
constexpr uint32_t kDictMax = 255;
constexpr uint32_t kDiagIndex = 255;
constexpr uint32_t kColBits = 24;
constexpr uint32_t kColMask = (1u << kColBits) - 1;

// Bit patterns of a value.  Vector elements are passed by value: clang's
// __builtin_bit_cast applied directly to an ext_vector element such as v.y
// reads element 0 instead (which made v and conj(v) hash and compare equal).
__device__ __forceinline__ unsigned long long bits_of(double d) { return __builtin_bit_cast(unsigned long long, d); }
__device__ __forceinline__ uint32_t bits_of(float f) { return __builtin_bit_cast(uint32_t, f); }

__device__ __forceinline__ unsigned long long value_hash(dvec2 v) {
    unsigned long long z = bits_of(v.x) ^ (bits_of(v.y) * 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return (z ^ (z >> 31)) | 1ull;  // never 0
}
__device__ __forceinline__ uint32_t slot_of(unsigned long long h) {
    return static_cast<uint32_t>(h >> 40) & (kDictSlots - 1);
}

// Inserts hash h (value v) into the set; state[0] counts distinct values,
// state[1] is raised when they exceed kDictMax or the set is full.
__device__ void dict_insert(unsigned long long h, dvec2 v, unsigned long long *slots, dvec2 *slot_vals,
                            uint32_t *state) {
    uint32_t s = slot_of(h);
    for (uint32_t probe = 0; probe < kDictSlots; ++probe, s = (s + 1) & (kDictSlots - 1)) {
        const unsigned long long cur = __hip_atomic_load(&slots[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == h) return;
        if (cur == 0) {
            const unsigned long long prev = atomicCAS(&slots[s], 0ull, h);
            if (prev == 0) {
                slot_vals[s] = v;
                if (atomicAdd(&state[0], 1u) + 1 > kDictMax) atomicExch(&state[1], 1u);
                return;
            }
            if (prev == h) return;
        }
    }
    atomicExch(&state[1], 1u);
}

// One thread per row over its off-diagonal entries; lanes of a wave holding
// the same hash are served by one insertion (the wave's first such lane).
__global__ __launch_bounds__(256) void csr_dict_insert_kernel(const uint64_t *__restrict__ indptr,
                                                              const uint32_t *__restrict__ cols,
                                                              const dvec2 *__restrict__ vals, uint64_t dim,
                                                              unsigned long long *slots, dvec2 *slot_vals,
                                                              uint32_t *state) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x; base < dim;
         base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t row = base + threadIdx.x;
        uint64_t k = row < dim ? indptr[row] : 0;
        const uint64_t k1 = row < dim ? indptr[row + 1] : 0;
        for (;;) {  // wave-uniform: every lane stays until no lane has entries left
            if (k < k1 && cols[k] == row) ++k;  // columns ascend: at most one diagonal entry
            const bool has = k < k1;
            if (!__any(has)) break;
            if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(&state[1], __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_AGENT)))
                return;  // too many distinct values: the caller keeps plain CSR
            const dvec2 v = has ? vals[k] : dvec2{0.0, 0.0};
            const unsigned long long h = has ? value_hash(v) : 0ull;
            bool todo = has;
            for (;;) {
                const unsigned long long m = __ballot(todo);
                if (!m) break;
                const int first = __ffsll(static_cast<long long>(m)) - 1;
                const unsigned long long lead = __shfl(h, first);
                if (static_cast<int>(lane) == first) dict_insert(h, v, slots, slot_vals, state);
                if (todo && h == lead) todo = false;
            }
            if (has) ++k;
        }
    }
}

// Packs every entry as column | index << kColBits and writes the diagonal
// values; raises state[1] if an entry's value is not bit-identical to its
// table value (a hash collision) in FP64 or FP32.
__global__ __launch_bounds__(256) void csr_pack_kernel(const uint64_t *__restrict__ indptr,
                                                       const uint32_t *__restrict__ cols,
                                                       const dvec2 *__restrict__ vals,
                                                       const fvec2 *__restrict__ vals32, uint64_t dim,
                                                       const unsigned long long *__restrict__ slots,
                                                       const uint32_t *__restrict__ slot_index,
                                                       const dvec2 *__restrict__ table64,
                                                       const fvec2 *__restrict__ table32, uint32_t *__restrict__ packed,
                                                       dvec2 *__restrict__ diag64, fvec2 *__restrict__ diag32,
                                                       uint32_t *state) {
    for (uint64_t row = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; row < dim;
         row += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        for (uint64_t k = indptr[row]; k < indptr[row + 1]; ++k) {
            const uint32_t c = cols[k];
            const dvec2 v = vals[k];
            if (c == row) {
                packed[k] = c | (kDiagIndex << kColBits);
                diag64[row] = v;
                diag32[row] = vals32[k];
                continue;
            }
            const unsigned long long h = value_hash(v);
            uint32_t s = slot_of(h), probe = 0;
            while (slots[s] != h && ++probe < kDictSlots) s = (s + 1) & (kDictSlots - 1);
            const uint32_t idx = probe < kDictSlots ? slot_index[s] : kDiagIndex;
            const dvec2 t64 = idx < kDictMax ? table64[idx] : dvec2{0.0, 0.0};
            const fvec2 t32 = idx < kDictMax ? table32[idx] : fvec2{0.f, 0.f};
            const fvec2 v32 = vals32[k];
            const bool same = idx < kDictMax && bits_of(t64.x) == bits_of(v.x) && bits_of(t64.y) == bits_of(v.y) &&
                              bits_of(t32.x) == bits_of(v32.x) && bits_of(t32.y) == bits_of(v32.y);
            if (!same) {
                atomicExch(&state[1], 1u);
                return;
            }
            packed[k] = c | (idx << kColBits);
        }
    }
}

// The SpMV of csr_spmv_kernel over the packed entries: the value table sits
// in LDS, the diagonal entry reads the per-row array.
template <typename VT, uint32_t LANES, uint32_t kUnroll = 4>
__global__ __launch_bounds__(256) void csr_compact_spmv_kernel(const uint64_t *__restrict__ indptr,
                                                               const uint32_t *__restrict__ packed,
                                                               const VT *__restrict__ table, uint32_t ntable,
                                                               const VT *__restrict__ diag,
                                                               const VT *__restrict__ v, VT *__restrict__ y,
                                                               uint64_t dim) {
    __shared__ VT tab[kDictMax];
    for (uint32_t i = threadIdx.x; i < kDictMax; i += blockDim.x) tab[i] = i < ntable ? table[i] : VT{0, 0};
    __syncthreads();
    const uint32_t lane = threadIdx.x % LANES;
    const uint64_t rows_per_step = static_cast<uint64_t>(gridDim.x) * (blockDim.x / LANES);
    for (uint64_t row = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / LANES) + threadIdx.x / LANES; row < dim;
         row += rows_per_step) {
        double ar = 0.0, ai = 0.0;
        const uint64_t k0 = indptr[row], k1 = indptr[row + 1];
        for (uint64_t kb = k0 + lane; kb < k1; kb += kUnroll * LANES) {
            uint32_t c[kUnroll];
            VT w[kUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
                const uint64_t k = kb + u * LANES;
                const uint32_t p = k < k1 ? packed[k] : 0u;
                c[u] = p & kColMask;
                const uint32_t idx = p >> kColBits;
                w[u] = idx == kDiagIndex ? diag[c[u]] : tab[idx];
            }
            VT x[kUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) x[u] = v[c[u]];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
                const bool in = kb + u * LANES < k1;  // padded slots add an exact zero
                const double cr = in ? static_cast<double>(w[u].x) : 0.0, ci = in ? static_cast<double>(w[u].y) : 0.0;
                const double sr = in ? static_cast<double>(x[u].x) : 0.0, si = in ? static_cast<double>(x[u].y) : 0.0;
                ar = fma(cr, sr, ar);
                ar = fma(-ci, si, ar);
                ai = fma(cr, si, ai);
                ai = fma(ci, sr, ai);
            }
        }
#pragma unroll
        for (uint32_t off = LANES / 2; off > 0; off >>= 1) {
            ar += __shfl_xor(ar, off, LANES);
            ai += __shfl_xor(ai, off, LANES);
        }
        if (lane == 0) y[row] = Amp<VT>::make(ar, ai);
    }
}

// ---------------------------------------------------------------- slab form
// The compact entries regrouped by x-group: slot q's words for consecutive
// rows are consecutive in memory (within a chunk of kSlabRows rows), and one
// thread per row visits the slots in turn, so a wave reads 64 packed words of
// one group in one access and its
// gathers, the sources of consecutive rows under one x-mask, fall in few
// cache lines (the CSR's lanes gather unrelated columns).  Rows without a
// transition in a group hold kSlabEmpty.
constexpr uint32_t kSlabEmpty = 0xFFFFFFFFu;
constexpr uint32_t kSlabUnroll = 8;  // slots whose loads are in flight together
// Rows per chunk (= the SpMV's workgroup): a chunk's words lie slot after
// slot in one contiguous run, so a workgroup streams one block of memory
// (slot-major over all rows put each slot 4 D bytes apart: 5-7 % slower).
constexpr uint32_t kSlabRows = 256;
__host__ __device__ __forceinline__ uint64_t slab_word(uint64_t row, uint32_t q, uint32_t nslab) {
    return (row / kSlabRows) * nslab * kSlabRows + static_cast<uint64_t>(q) * kSlabRows + row % kSlabRows;
}

__global__ void fill_words_kernel(uint32_t *__restrict__ p, uint64_t count, uint32_t value) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        p[i] = value;
}

// Index of x in the ascending xs[0..nx), or nx.
__device__ __forceinline__ uint32_t find_mask(const uint64_t *__restrict__ xs, uint32_t nx, uint64_t x) {
    uint32_t lo = 0, hi = nx;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (xs[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo < nx && xs[lo] == x ? lo : nx;
}

// Pass 1 (slot == nullptr): marks used[g] for the x-mask g of every
// off-diagonal entry (x = state(row) ^ state(column)), counts those entries
// into *offdiag and raises *bad when an x-mask is not in xs.  Pass 2: writes
// each entry into its row of slot slot[g].
__global__ __launch_bounds__(256) void slab_scatter_kernel(const uint64_t *__restrict__ indptr,
                                                           const uint32_t *__restrict__ packed,
                                                           const uint64_t *__restrict__ states, uint64_t dim,
                                                           const uint64_t *__restrict__ xs, uint32_t nx,
                                                           uint32_t *__restrict__ used,
                                                           const uint32_t *__restrict__ slot, uint32_t nslab,
                                                           uint32_t *__restrict__ slab,
                                                           unsigned long long *offdiag, uint32_t *bad) {
    unsigned long long local = 0;
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < dim;
         r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        for (uint64_t k = indptr[r]; k < indptr[r + 1]; ++k) {
            const uint32_t w = packed[k];
            if ((w >> kColBits) == kDiagIndex) continue;
            const uint32_t c = w & kColMask;
            const uint32_t g = c < dim ? find_mask(xs, nx, states[r] ^ states[c]) : nx;
            if (g == nx) {
                atomicExch(bad, 1u);
                continue;
            }
            if (slot) slab[slab_word(r, slot[g], nslab)] = w;
            else used[g] = 1u;
            ++local;
        }
    }
    if (!slot && local) atomicAdd(offdiag, local);
}

// y = H v over the slab form: one row per thread, kSlabUnroll slots' words
// and gathers in flight together, FP64 accumulation; the value table is in LDS.
template <typename VT>
__global__ __launch_bounds__(256) void slab_spmv_kernel(const uint32_t *__restrict__ slab, uint32_t nslab,
                                                        const VT *__restrict__ table, uint32_t ntable,
                                                        const VT *__restrict__ diag, const VT *__restrict__ v,
                                                        VT *__restrict__ y, uint64_t dim) {
    __shared__ VT tab[kDictMax];
    for (uint32_t i = threadIdx.x; i < kDictMax; i += blockDim.x) tab[i] = i < ntable ? table[i] : VT{0, 0};
    __syncthreads();
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < dim;
         r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const VT d = diag[r], vr = v[r];
        double ar = static_cast<double>(d.x) * vr.x - static_cast<double>(d.y) * vr.y;
        double ai = static_cast<double>(d.x) * vr.y + static_cast<double>(d.y) * vr.x;
        for (uint32_t q0 = 0; q0 < nslab; q0 += kSlabUnroll) {
            uint32_t w[kSlabUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kSlabUnroll; ++u)
                w[u] = q0 + u < nslab ? slab[slab_word(r, q0 + u, nslab)] : kSlabEmpty;
            VT x[kSlabUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kSlabUnroll; ++u) x[u] = v[w[u] != kSlabEmpty ? (w[u] & kColMask) : r];
#pragma unroll
            for (uint32_t u = 0; u < kSlabUnroll; ++u) {
                const bool on = w[u] != kSlabEmpty;  // an empty slot adds an exact zero
                const VT c = tab[on ? (w[u] >> kColBits) : 0u];
                const double cr = on ? static_cast<double>(c.x) : 0.0, ci = on ? static_cast<double>(c.y) : 0.0;
                const double sr = static_cast<double>(x[u].x), si = static_cast<double>(x[u].y);
                ar = fma(cr, sr, ar);
                ar = fma(-ci, si, ar);
                ai = fma(cr, si, ai);
                ai = fma(ci, sr, ai);
            }
        }
        y[r] = Amp<VT>::make(ar, ai);
    }
}

}  // namespace

// Host rank authority of a general charge sector.
struct tcp_charge_sector {
    tcp::ChargeGeometry geo;
};

struct tcp_sector_plan {
    uint32_t n = 0;
    uint32_t nblocks = 0;
    uint32_t sizes[kMaxBlocks] = {}, pops[kMaxBlocks] = {};
    uint64_t dim = 0, T = 0, G = 0;
    int device = 0;
    bool lut_in_lds = true;
    bool u1 = true;
    uint32_t diag_group = UINT32_MAX;  // index of the x = 0 group
    double energy_scale = 1.0;         // largest sum |w| of an off-diagonal group
    bool leaks = false;        // the reference's apply would raise
    uint64_t leak_source = 0;  // first leaking source rank
    uint32_t leak_weight = 0;  // its destination weight (U(1) message)
    SectorArgs host_args{};
    bool cond = false;  // terms are fermion ladder words with occupancy conditions
    tcp::DeviceBuffer gx, goff, tz, tw64, tw32, lut, choose, tc1, tc0, cbase;
    // the gather kernel's tables (see GatherArgs)
    GatherArgs host_gather{};
    tcp::DeviceBuffer grec64, grec32, gtz, gtw64, gtw32, gtc1, gtc0, diag64;
    std::unique_ptr<tcp::WideSector> wide;        // U(1) sectors of 64 or more qubits
    std::unique_ptr<tcp::ChargeSector> general;   // charge sectors that are not contiguous blocks
    uint64_t retained_bytes = 0;
    std::mutex staging_mutex;
    tcp::DeviceBuffer stage_in, stage_out;

    SectorArgs args(int dtype) const {
        SectorArgs a = host_args;
        a.gx = static_cast<const uint64_t *>(gx.ptr);
        a.goff = static_cast<const uint32_t *>(goff.ptr);
        a.tz = static_cast<const uint64_t *>(tz.ptr);
        a.tw = dtype == TCP_C64 ? tw32.ptr : tw64.ptr;
        a.tc1 = cond ? static_cast<const uint64_t *>(tc1.ptr) : nullptr;
        a.tc0 = cond ? static_cast<const uint64_t *>(tc0.ptr) : nullptr;
        a.lut = static_cast<const uint32_t *>(lut.ptr);
        a.choose = static_cast<const uint64_t *>(choose.ptr);
        a.cbase = static_cast<const uint32_t *>(cbase.ptr);
        return a;
    }
    GatherArgs gather_args(int dtype) const {
        GatherArgs q = host_gather;
        q.groups = static_cast<const GatherGroup *>(dtype == TCP_C64 ? grec32.ptr : grec64.ptr);
        q.tz = static_cast<const uint64_t *>(gtz.ptr);
        q.tw = dtype == TCP_C64 ? gtw32.ptr : gtw64.ptr;
        q.tc1 = cond ? static_cast<const uint64_t *>(gtc1.ptr) : nullptr;
        q.tc0 = cond ? static_cast<const uint64_t *>(gtc0.ptr) : nullptr;
        q.diag = static_cast<const double *>(diag64.ptr);
        return q;
    }
};

namespace {

using tcp::fail;

// C(n, k) with overflow detection (core/sector.rs:1342-1362).
bool choose_checked(uint32_t n, uint32_t k, uint64_t &out) {
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

struct BlockSpec {
    uint32_t nblocks;
    uint32_t sizes[kMaxBlocks], pops[kMaxBlocks];
};

// Host-side sector geometry shared by rank/unrank entry points and plans.
int sector_geometry(const BlockSpec &spec, SectorArgs &a, uint64_t &dim, std::vector<uint32_t> &lut,
                    std::vector<uint64_t> &choose_table, std::vector<uint32_t> &cbase) {
    if (spec.nblocks == 0) return fail(TCP_ERR_INVALID_SECTOR, "invalid sector: no blocks");
    if (spec.nblocks > kMaxBlocks)
        return fail(TCP_ERR_UNSUPPORTED, "at most 4 contiguous blocks are supported on the GPU");
    uint32_t n = 0;
    for (uint32_t j = 0; j < spec.nblocks; ++j) {
        if (spec.pops[j] > spec.sizes[j])
            return fail(TCP_ERR_INVALID_SECTOR, "invalid sector: particle_number must be between 0 and nqubits");
        n += spec.sizes[j];
    }
    if (n > 64) return fail(TCP_ERR_UNSUPPORTED, "sector bitstrings are limited to 64 qubits");
    a = SectorArgs{};
    a.nblocks = spec.nblocks;
    uint32_t shift = n;
    unsigned __int128 total = 1;
    uint64_t bdim[kMaxBlocks];
    for (uint32_t j = 0; j < spec.nblocks; ++j) {
        shift -= spec.sizes[j];
        a.shift[j] = shift;
        a.size[j] = spec.sizes[j];
        a.pop[j] = spec.pops[j];
        a.bmask[j] = spec.sizes[j] == 0 ? 0 : (spec.sizes[j] >= 64 ? ~0ull : (1ull << spec.sizes[j]) - 1);
        a.nchunk[j] = (spec.sizes[j] + 7) / 8;
        if (!choose_checked(spec.sizes[j], spec.pops[j], bdim[j]))
            return fail(TCP_ERR_OVERFLOW, "integer overflow while computing U1 sector dimension");
        total *= bdim[j];
        if (total > static_cast<unsigned __int128>(UINT64_MAX))
            return fail(TCP_ERR_OVERFLOW, "integer overflow while computing U1 sector dimension");
    }
    dim = static_cast<uint64_t>(total);
    a.dim = dim;
    uint64_t stride = 1;
    for (int j = static_cast<int>(spec.nblocks) - 1; j >= 0; --j) {
        a.stride[j] = static_cast<uint32_t>(stride);
        stride *= bdim[j];
    }
    choose_table.assign(64 * 64, 0);
    for (uint32_t q = 0; q < 64; ++q)
        for (uint32_t k = 0; k < 64; ++k) {
            uint64_t c = 0;
            if (!choose_checked(q, k, c)) c = UINT64_MAX;
            choose_table[q * 64 + k] = c;
        }
    // Rows only for the ones-below counts an in-sector pattern can have in
    // front of chunk c: max(0, k - (size - 8c)) .. min(k, 8c).  cbase[j][c]
    // is the row base minus lo * 256 (unsigned wrap-around), so the entry of
    // (below, byte) sits at cbase + below * 256 + byte.
    lut.clear();
    cbase.assign(kMaxBlocks * kMaxChunksPerBlock, 0);
    for (uint32_t j = 0; j < spec.nblocks; ++j) {
        a.lut_off[j] = static_cast<uint32_t>(lut.size());
        const uint32_t k = spec.pops[j];
        for (uint32_t c = 0; c < a.nchunk[j]; ++c) {
            const uint32_t rest = spec.sizes[j] - 8 * c;  // positions from chunk c upwards
            const uint32_t lo = k > rest ? k - rest : 0;
            const uint32_t hi = std::min(k, 8 * c);
            cbase[j * kMaxChunksPerBlock + c] = static_cast<uint32_t>(lut.size()) - lo * 256u;
            for (uint32_t below = lo; below <= hi; ++below)
                for (uint32_t byte = 0; byte < 256; ++byte) {
                    uint64_t value = 0;
                    uint32_t seen = 0;
                    for (uint32_t q = 0; q < 8; ++q) {
                        if (!((byte >> q) & 1)) continue;
                        const uint32_t pos = 8 * c + q;
                        const uint32_t index = below + seen + 1;
                        ++seen;
                        if (pos < spec.sizes[j] && index <= k) value += choose_table[pos * 64 + index];
                    }
                    lut.push_back(static_cast<uint32_t>(value * a.stride[j]));  // < dim < 2^32 in sector
                }
        }
    }
    a.lut_words = static_cast<uint32_t>(lut.size());
    return TCP_OK;
}

int scan_leakage(tcp_sector_plan *plan, const std::vector<uint64_t> &gx);
int store_diagonal(tcp_sector_plan *plan);

// The gather kernel's tables of a block-sector plan (see GatherArgs): flat
// byte chunks, one record per group it visits (FP64 and FP32 weights), the
// further terms and, when the budget allows dim * 8 more bytes, the stored
// diagonal in place of the x = 0 group.
int gather_tables(tcp_sector_plan *plan, const std::vector<tcp::CondTerm> &terms,
                  const std::vector<std::vector<uint64_t>> &members, const std::vector<uint64_t> &gx,
                  const std::vector<uint32_t> &cbase, uint64_t max_bytes) {
    const SectorArgs &a = plan->host_args;
    GatherArgs q{};
    uint32_t first_chunk[kMaxBlocks] = {};
    for (uint32_t j = 0; j < a.nblocks; ++j) {
        first_chunk[j] = q.nflat;
        q.blockmask[j] = a.bmask[j] << a.shift[j];
        for (uint32_t c = 0; c < a.nchunk[j]; ++c) {
            const uint32_t t = q.nflat++;
            if (t >= kMaxFlatChunks) return fail(TCP_ERR_UNSUPPORTED, "too many rank-table chunks");
            q.cshift[t] = a.shift[j] + 8 * c;
            q.cbase[t] = cbase[j * kMaxChunksPerBlock + c];
            if (c == 0) q.cfirst |= 1u << t;
        }
    }
    const bool cond = plan->cond;
    const bool store = plan->diag_group != UINT32_MAX && plan->dim > 0 &&
                       !tcp::exceeds_budget(plan->retained_bytes + plan->dim * 8, max_bytes);
    std::vector<GatherGroup> rec64, rec32;
    std::vector<uint64_t> gtz, gtc1, gtc0;
    std::vector<dvec2> gtw64, gtw32;
    for (size_t g = 0; g < gx.size() && g < members.size(); ++g) {
        if (store && g == plan->diag_group) continue;
        GatherGroup r{};
        r.x = gx[g];
        bool reachable = true;
        for (uint32_t j = 0; j < kMaxBlocks; ++j) {
            const uint64_t xj = j < a.nblocks ? (gx[g] >> a.shift[j]) & a.bmask[j] : 0;
            if (!xj) {
                    continue;
            }
            const uint32_t flips = static_cast<uint32_t>(__builtin_popcountll(xj));
            if ((flips & 1u) || flips / 2 > a.pop[j] || flips / 2 > a.size[j] - a.pop[j]) reachable = false;
            const uint32_t lo = first_chunk[j] + static_cast<uint32_t>(__builtin_ctzll(xj)) / 8;
            const uint32_t hi = first_chunk[j] + (63u - static_cast<uint32_t>(__builtin_clzll(xj))) / 8;
            for (uint32_t t = lo; t <= hi; ++t) r.chunks |= 1u << t;
            r.chunks |= (1u << (12 + lo)) | (1u << (24 + j));
            r.halves |= (flips / 2) << (8 * j);
        }
        if (!reachable) continue;  // never maps a sector state into the sector
        GatherGroup r32 = r;
        r.toff = r32.toff = static_cast<uint32_t>(gtz.size());
        // On an in-sector source s = b ^ x, x & s = x & ~b holds exactly
        // H = popc(x) / 2 ones, so a Pauli term whose z covers all of x or
        // none of it has the sign (-1)^(H [z & x = x]) (-1)^popc(z' & s),
        // z' = z & ~x: such terms are summed per z' (z' = 0: the group's
        // constant, the XX and YY of a bond), each sum kept as one term.
        // Other terms and fermion words (whose conditions read s) stay as
        // they are.
        const uint32_t half_x = static_cast<uint32_t>(__builtin_popcountll(gx[g])) / 2;
        std::vector<uint64_t> keys;
        std::vector<dvec2> sum64, sum32;
        std::vector<const tcp::CondTerm *> rest;
        for (uint64_t k : members[g]) {
            const tcp::CondTerm &t = terms[k];
            const uint64_t zx = t.z & gx[g];
            if (cond || (zx != 0 && zx != gx[g])) {
                rest.push_back(&t);
                continue;
            }
            const double sign = (zx != 0 && (half_x & 1u)) ? -1.0 : 1.0;
            const uint64_t key = t.z & ~gx[g];
            const size_t i = std::find(keys.begin(), keys.end(), key) - keys.begin();
            if (i == keys.size()) {
                keys.push_back(key);
                sum64.push_back(dvec2{0.0, 0.0});
                sum32.push_back(dvec2{0.0, 0.0});
            }
            sum64[i] += sign * dvec2{t.w.re, t.w.im};
            sum32[i] += sign * dvec2{static_cast<double>(static_cast<float>(t.w.re)),
                                     static_cast<double>(static_cast<float>(t.w.im))};
        }
        auto add_term = [&](uint64_t z, uint64_t c1, uint64_t c0, dvec2 w64, dvec2 w32) {
            if (!(r.nterms & kInlineTerm) && !cond) {
                r.z1 = r32.z1 = z;
                r.w1re = w64.x;
                r.w1im = w64.y;
                r32.w1re = w32.x;
                r32.w1im = w32.y;
                r.nterms |= kInlineTerm;
                r32.nterms |= kInlineTerm;
                return;
            }
            gtz.push_back(z);
            gtc1.push_back(c1);
            gtc0.push_back(c0);
            gtw64.push_back(w64);
            gtw32.push_back(w32);
            ++r.nterms;
            ++r32.nterms;
        };
        for (size_t i = 0; i < keys.size(); ++i) {
            if (keys[i] == 0) {
                r.cre = sum64[i].x;
                r.cim = sum64[i].y;
                r32.cre = sum32[i].x;
                r32.cim = sum32[i].y;
            } else if (sum64[i].x != 0.0 || sum64[i].y != 0.0 || sum32[i].x != 0.0 || sum32[i].y != 0.0) {
                add_term(keys[i], 0, 0, sum64[i], sum32[i]);
            }
        }
        for (const tcp::CondTerm *t : rest)
            add_term(t->z, t->c1, t->c0, dvec2{t->w.re, t->w.im},
                     dvec2{static_cast<double>(static_cast<float>(t->w.re)),
                           static_cast<double>(static_cast<float>(t->w.im))});
        const bool constant = r.nterms == 0;
        if (constant && r.cre == 0.0 && r.cim == 0.0 && r32.cre == 0.0 && r32.cim == 0.0)
            continue;  // sums to an exact zero on every in-sector source
        rec64.push_back(r);
        rec32.push_back(r32);
    }
    q.G = static_cast<uint32_t>(rec64.size());
    if (rec64.empty()) {
        rec64.push_back(GatherGroup{});
        rec32.push_back(GatherGroup{});
    }
    if (gtz.empty()) {
        gtz.push_back(0);
        gtc1.push_back(0);
        gtc0.push_back(0);
        gtw64.push_back(dvec2{0, 0});
        gtw32.push_back(dvec2{0, 0});
    }
    plan->host_gather = q;
    int st;
    if ((st = plan->grec64.upload(rec64.data(), rec64.size() * sizeof(GatherGroup))) ||
        (st = plan->grec32.upload(rec32.data(), rec32.size() * sizeof(GatherGroup))) ||
        (st = plan->gtz.upload(gtz.data(), gtz.size() * 8)) ||
        (st = plan->gtw64.upload(gtw64.data(), gtw64.size() * 16)) ||
        (st = plan->gtw32.upload(gtw32.data(), gtw32.size() * 16)))
        return st;
    if (cond && ((st = plan->gtc1.upload(gtc1.data(), gtc1.size() * 8)) ||
                 (st = plan->gtc0.upload(gtc0.data(), gtc0.size() * 8))))
        return st;
    plan->retained_bytes += rec64.size() * 2 * sizeof(GatherGroup) + gtz.size() * (cond ? 56 : 40);
    if (store) {
        if ((st = store_diagonal(plan))) return st;
        plan->retained_bytes += plan->dim * 8;
    }
    return TCP_OK;
}

// require_conserved: lazy restriction's precondition (exact conservation of
// every block's particle number).  The eager builders skip it, as the
// reference's do (core/sector.rs:345-432 checks each aggregated transition
// instead), and rely on the leakage scan.
// A sector plan over the terms of a Pauli operator (`op`) or of a fermion
// operator (`fermion`, ladder words with occupancy conditions); exactly one is
// given.
int build_sector_plan_impl(const tcp::Operator *op, const tcp_fermion *fermion, const BlockSpec &spec, bool u1,
                           uint64_t max_bytes, tcp_sector_plan **out, bool require_conserved) {
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    uint32_t n = 0;
    for (uint32_t j = 0; j < spec.nblocks && j < kMaxBlocks; ++j) n += spec.sizes[j];
    const uint32_t width = op ? op->nqubits : fermion->n_modes;
    if (width != n)
        return fail(TCP_ERR_INCOMPATIBLE_QUBIT_COUNTS, "incompatible qubit counts: " + std::to_string(width) +
                                                            " and " + std::to_string(n));
    if (fermion && n > 63) return fail(TCP_ERR_UNSUPPORTED, "GPU fermion sector kernels support at most 63 modes");
    if (u1 && spec.nblocks == 1 && n >= 64) {
        const tcp::Operator &opr = *op;
        // wide U(1) sector: packed-word bitstrings (see wide_sector.hpp)
        auto plan = new tcp_sector_plan();
        std::unique_ptr<tcp_sector_plan> guard(plan);
        plan->wide.reset(new tcp::WideSector());
        int st = tcp::wide_build(opr, spec.pops[0], max_bytes, require_conserved, *plan->wide);
        if (st) return st;
        const tcp::WideSector &w = *plan->wide;
        plan->n = n;
        plan->nblocks = 1;
        plan->sizes[0] = n;
        plan->pops[0] = spec.pops[0];
        plan->dim = w.geo.dim;
        plan->T = w.T;
        plan->G = w.G;
        plan->diag_group = w.diag_group;
        plan->energy_scale = w.energy_scale;
        plan->leaks = w.leaks;
        plan->leak_source = w.leak_source;
        plan->leak_weight = w.leak_weight;
        plan->u1 = true;
        plan->retained_bytes = sizeof(tcp_sector_plan) + w.retained_bytes;
        TCP_HIP_CHECK(hipGetDevice(&plan->device));
        *out = guard.release();
        return tcp::ok();
    }
    SectorArgs a;
    uint64_t dim = 0;
    std::vector<uint32_t> lut;
    std::vector<uint64_t> choose_table;
    std::vector<uint32_t> cbase;
    int st = sector_geometry(spec, a, dim, lut, choose_table, cbase);
    if (st) return st;
    if (n > 63) return fail(TCP_ERR_UNSUPPORTED, "GPU sector kernels support at most 63 qubits");
    if (dim >= (1ull << 32))
        return fail(TCP_ERR_UNSUPPORTED, "GPU sector kernels support dimensions below 2^32");

    // Exact conservation of each block's particle number: the precondition
    // of lazy restriction (py/symmetry.py:675-681, py/charge.py:1128-1132).
    // Fermion words conserve a block's number when they create as many
    // particles in it as they annihilate.
    std::vector<tcp::CondTerm> terms;
    if (fermion) terms = tcp::fermion_index_terms(*fermion);
    for (uint32_t j = 0; require_conserved && j < spec.nblocks; ++j) {
        bool conserved = true;
        if (op) {
            std::vector<double> levels(2 * n, 0.0);
            const uint32_t first = n - a.shift[j] - spec.sizes[j];
            for (uint32_t q = first; q < first + spec.sizes[j]; ++q) levels[2 * q + 1] = 1.0;
            uint64_t nonzero = 0;
            if ((st = tcp::analyze_charge(*op, levels.data(), conserved, nonzero))) return st;
        } else {
            const uint64_t block = a.bmask[j] << a.shift[j];
            for (const tcp::CondTerm &t : terms)
                conserved &= __builtin_popcountll(t.c1 & block) == __builtin_popcountll((t.c0 | (t.c1 & ~t.x)) & block);
        }
        if (!conserved)
            return fail(TCP_ERR_SECTOR_LEAKAGE,
                        u1 ? "selected U1 sector requires an exactly conserved operator; nonzero "
                             "U(1) sector leakage was detected"
                           : "selected charge sector requires an exactly conserved operator");
    }

    auto plan = new tcp_sector_plan();
    std::unique_ptr<tcp_sector_plan> guard(plan);
    plan->n = n;
    plan->nblocks = spec.nblocks;
    for (uint32_t j = 0; j < spec.nblocks; ++j) {
        plan->sizes[j] = spec.sizes[j];
        plan->pops[j] = spec.pops[j];
    }
    plan->dim = dim;

    // x-mask groups in first-appearance order over the canonical terms; a
    // group's terms keep canonical order (compile_terms, core/sector.rs:926-1021).
    if (op)
        for (const tcp::IndexTerm &t : tcp::index_terms(*op)) terms.push_back(tcp::CondTerm{t.x, t.z, 0, 0, t.w});
    std::vector<uint64_t> gx;
    std::vector<std::vector<uint64_t>> members;
    for (uint64_t k = 0; k < terms.size(); ++k) {
        size_t g = std::find(gx.begin(), gx.end(), terms[k].x) - gx.begin();
        if (g == gx.size()) {
            gx.push_back(terms[k].x);
            members.emplace_back();
        }
        members[g].push_back(k);
    }
    std::vector<uint32_t> goff;
    std::vector<uint64_t> tz, tc1, tc0;
    std::vector<dvec2> tw64, tw32;
    for (size_t g = 0; g < gx.size(); ++g) {
        goff.push_back(static_cast<uint32_t>(tz.size()));
        for (uint64_t k : members[g]) {
            tz.push_back(terms[k].z);
            tc1.push_back(terms[k].c1);
            tc0.push_back(terms[k].c0);
            tw64.push_back(dvec2{terms[k].w.re, terms[k].w.im});
            tw32.push_back(dvec2{static_cast<double>(static_cast<float>(terms[k].w.re)),
                                 static_cast<double>(static_cast<float>(terms[k].w.im))});
        }
    }
    goff.push_back(static_cast<uint32_t>(tz.size()));
    plan->T = tz.size();
    plan->G = gx.size();
    double scale = 0.0;
    for (size_t g = 0; g < gx.size(); ++g) {
        if (gx[g] == 0) {
            plan->diag_group = static_cast<uint32_t>(g);
            continue;
        }
        double sum = 0.0;
        for (uint64_t k : members[g]) sum += std::hypot(terms[k].w.re, terms[k].w.im);
        scale = std::max(scale, sum);
    }
    plan->energy_scale = scale > 0.0 ? scale : 1.0;
    a.G = static_cast<uint32_t>(gx.size());
    if (gx.empty()) gx.push_back(0);
    if (tz.empty()) {
        tz.push_back(0);
        tc1.push_back(0);
        tc0.push_back(0);
        tw64.push_back(dvec2{0, 0});
        tw32.push_back(dvec2{0, 0});
    }
    plan->lut_in_lds = lut.size() * 4 <= kLdsLutLimit;
    plan->host_args = a;
    plan->retained_bytes = sizeof(tcp_sector_plan) + gx.size() * 8 + goff.size() * 4 +
                           tz.size() * (fermion ? 8 + 32 + 16 : 8 + 32) + lut.size() * 4 + choose_table.size() * 8;
    if (tcp::exceeds_budget(plan->retained_bytes, max_bytes))
        return tcp::memory_limit(plan->retained_bytes, max_bytes);

    TCP_HIP_CHECK(hipGetDevice(&plan->device));
    if ((st = plan->gx.upload(gx.data(), gx.size() * 8))) return st;
    if ((st = plan->goff.upload(goff.data(), goff.size() * 4))) return st;
    if ((st = plan->tz.upload(tz.data(), tz.size() * 8))) return st;
    if ((st = plan->tw64.upload(tw64.data(), tw64.size() * 16))) return st;
    if ((st = plan->tw32.upload(tw32.data(), tw32.size() * 16))) return st;
    if ((st = plan->lut.upload(lut.data(), lut.size() * 4))) return st;
    if ((st = plan->choose.upload(choose_table.data(), choose_table.size() * 8))) return st;
    if ((st = plan->cbase.upload(cbase.data(), cbase.size() * 4))) return st;
    if (fermion) {
        if ((st = plan->tc1.upload(tc1.data(), tc1.size() * 8))) return st;
        if ((st = plan->tc0.upload(tc0.data(), tc0.size() * 8))) return st;
        plan->cond = true;
    }
    plan->u1 = u1;
    if ((st = scan_leakage(plan, gx))) return st;
    // plans built for the eager CSR (no conservation precondition) never run the gather
    if (require_conserved && (st = gather_tables(plan, terms, members, gx, cbase, max_bytes))) return st;
    *out = guard.release();
    return tcp::ok();
}

// A sector plan on a general charge sector (see charge_sector.hpp) over the
// terms of a Pauli operator or of a fermion operator.
int build_charge_plan(const tcp::Operator *op, const tcp_fermion *fermion, const tcp::ChargeGeometry &geo,
                      uint64_t max_bytes, bool require_conserved, tcp_sector_plan **out) {
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    const uint32_t width = op ? op->nqubits : fermion->n_modes;
    if (width != geo.n)
        return fail(TCP_ERR_INCOMPATIBLE_QUBIT_COUNTS, "incompatible qubit counts: " + std::to_string(width) +
                                                            " and " + std::to_string(geo.n));
    auto plan = new tcp_sector_plan();
    std::unique_ptr<tcp_sector_plan> guard(plan);
    plan->general.reset(new tcp::ChargeSector());
    std::vector<tcp::CondTerm> fterms;
    if (fermion) fterms = tcp::fermion_index_terms(*fermion);
    int st = tcp::charge_build(op, fermion ? &fterms : nullptr, geo, max_bytes, require_conserved, *plan->general);
    if (st) return st;
    const tcp::ChargeSector &c = *plan->general;
    plan->n = geo.n;
    plan->nblocks = 0;
    plan->dim = geo.dim;
    plan->T = c.T;
    plan->G = c.G;
    plan->diag_group = c.diag_group;
    plan->energy_scale = c.energy_scale;
    plan->leaks = c.leaks;
    plan->leak_source = c.leak_source;
    plan->u1 = false;
    plan->cond = fermion != nullptr;
    plan->retained_bytes = sizeof(tcp_sector_plan) + c.retained_bytes;
    TCP_HIP_CHECK(hipGetDevice(&plan->device));
    *out = guard.release();
    return tcp::ok();
}

int build_sector_plan(const tcp::Operator &op, const BlockSpec &spec, bool u1, uint64_t max_bytes,
                      tcp_sector_plan **out, bool require_conserved = true) {
    return build_sector_plan_impl(&op, nullptr, spec, u1, max_bytes, out, require_conserved);
}

int host_unrank(const BlockSpec &spec, uint64_t index, uint64_t &bits);

int scan_leakage(tcp_sector_plan *plan, const std::vector<uint64_t> &gx) {
    if (plan->G == 0 || plan->dim == 0) return TCP_OK;
    unsigned long long *dfirst = nullptr;
    TCP_HIP_CHECK(hipMalloc(&dfirst, sizeof(unsigned long long)));
    const unsigned long long none = ~0ull;
    hipError_t e = hipMemcpy(dfirst, &none, sizeof none, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const uint64_t per_block = static_cast<uint64_t>(kSectorThreads) * kRanksPerThread;
        const uint64_t blocks = std::min<uint64_t>((plan->dim + per_block - 1) / per_block, 256 * 16);
        if (plan->cond)
            hipLaunchKernelGGL(sector_leak_kernel<true>, dim3(static_cast<uint32_t>(blocks)), dim3(kSectorThreads),
                               0, nullptr, plan->args(TCP_C128), dfirst);
        else
            hipLaunchKernelGGL(sector_leak_kernel<false>, dim3(static_cast<uint32_t>(blocks)), dim3(kSectorThreads),
                               0, nullptr, plan->args(TCP_C128), dfirst);
        e = hipGetLastError();
    }
    unsigned long long first = none;
    if (e == hipSuccess) e = hipMemcpy(&first, dfirst, sizeof first, hipMemcpyDeviceToHost);
    (void)hipFree(dfirst);
    if (e != hipSuccess) return tcp::hip_fail(e, "sector leakage scan");
    if (first != none) {
        plan->leaks = true;
        plan->leak_source = first / plan->G;
        const uint64_t g = first % plan->G;
        uint64_t bits = 0;
        std::vector<uint32_t> lut;
        std::vector<uint64_t> ch;
        BlockSpec spec{};
        spec.nblocks = plan->nblocks;
        for (uint32_t j = 0; j < plan->nblocks; ++j) {
            spec.sizes[j] = plan->sizes[j];
            spec.pops[j] = plan->pops[j];
        }
        if (host_unrank(spec, plan->leak_source, bits)) return TCP_ERR_INVALID_INDEX;
        plan->leak_weight = static_cast<uint32_t>(__builtin_popcountll(bits ^ gx[g]));
    }
    return TCP_OK;
}

int leak_error(const tcp_sector_plan *plan) {
    if (plan->u1)
        return fail(TCP_ERR_SECTOR_LEAKAGE,
                    "U(1) sector leakage from restricted source index " + std::to_string(plan->leak_source) +
                        ": expected particle number " + std::to_string(plan->pops[0]) +
                        ", got destination weight " + std::to_string(plan->leak_weight));
    return fail(TCP_ERR_INVALID_SECTOR, "invalid sector: operator leaks outside the selected charge sector");
}

int apply_device(const tcp_sector_plan *plan, const void *in, void *out, int dtype, void *stream);

}  // namespace

namespace tcp {
uint64_t plan_dimension(const tcp_sector_plan *plan) { return plan->dim; }
int plan_device(const tcp_sector_plan *plan) { return plan->device; }
int plan_apply(const tcp_sector_plan *plan, const void *in, void *out, int dtype, hipStream_t stream) {
    return apply_device(plan, in, out, dtype, stream);
}
double plan_energy_scale(const tcp_sector_plan *plan) { return plan->energy_scale; }
int plan_features(const tcp_sector_plan *plan, double scale, uint16_t *features, double *diag, hipStream_t stream) {
    if (plan->wide)
        return fail(TCP_ERR_UNSUPPORTED, "learned-preconditioner features support sectors of at most 63 qubits");
    if (plan->general)
        return fail(TCP_ERR_UNSUPPORTED, "learned-preconditioner features support U(1) and block sectors only");
    if (plan->dim == 0) return TCP_OK;
    const uint32_t blocks = static_cast<uint32_t>((plan->dim + 255) / 256);
    if (plan->cond)
        hipLaunchKernelGGL(sector_features_kernel<true>, dim3(blocks), dim3(256), 0, stream, plan->args(TCP_C128),
                           plan->diag_group, 1.0 / scale, features, diag);
    else
        hipLaunchKernelGGL(sector_features_kernel<false>, dim3(blocks), dim3(256), 0, stream, plan->args(TCP_C128),
                           plan->diag_group, 1.0 / scale, features, diag);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "sector features kernel");
}
int plan_diagonal(const tcp_sector_plan *plan, double *diag, hipStream_t stream) {
    if (plan->wide) return wide_diagonal(*plan->wide, diag, stream);
    if (plan->general) return charge_diagonal(*plan->general, diag, stream);
    if (plan->dim == 0) return TCP_OK;
    const uint64_t threads = (plan->dim + kRanksPerThread - 1) / kRanksPerThread;
    const uint32_t blocks = static_cast<uint32_t>((threads + 255) / 256);
    if (plan->cond)
        hipLaunchKernelGGL(sector_diagonal_kernel<true>, dim3(blocks), dim3(256), 0, stream, plan->args(TCP_C128),
                           plan->diag_group, diag);
    else
        hipLaunchKernelGGL(sector_diagonal_kernel<false>, dim3(blocks), dim3(256), 0, stream, plan->args(TCP_C128),
                           plan->diag_group, diag);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "sector diagonal kernel");
}
}  // namespace tcp

namespace {

// The diagonal H_rr of every state, kept for the gather kernel.
int store_diagonal(tcp_sector_plan *plan) {
    int st = plan->diag64.allocate(plan->dim * 8);
    if (st) return st;
    if ((st = tcp::plan_diagonal(plan, static_cast<double *>(plan->diag64.ptr), nullptr))) return st;
    TCP_HIP_CHECK(hipDeviceSynchronize());
    return TCP_OK;
}

int check_len(uint64_t expected, uint64_t actual) {
    if (expected != actual)
        return fail(TCP_ERR_INVALID_STRUCTURE_LENGTH, "expected structure length " +
                                                          std::to_string(expected) + ", got " +
                                                          std::to_string(actual));
    return TCP_OK;
}

template <typename VT, bool COND>
hipError_t launch_gather(const tcp_sector_plan *plan, const void *in, void *out, int dtype,
                         hipStream_t stream) {
    SectorArgs a = plan->args(dtype);
    const GatherArgs q = plan->gather_args(dtype);
    // index bits fit 32-bit words (and so do the byte chunks: checked, not assumed)
    const bool narrow = plan->n <= 32 && q.nflat <= narrow_chunk_count(a.nblocks);
    auto launch = [&](auto nb) {
        constexpr uint32_t NB = decltype(nb)::value;
        constexpr uint32_t RPT = NB == 1 ? 1 : 2;
        const uint64_t per_block = static_cast<uint64_t>(kGatherThreads) * RPT;
        const uint64_t blocks = std::min<uint64_t>((plan->dim + per_block - 1) / per_block, 256 * 16);
        const dim3 grid(static_cast<uint32_t>(blocks)), block(kGatherThreads);
        const size_t lds = plan->lut_in_lds ? a.lut_words * 4 : 0;
        const VT *vin = static_cast<const VT *>(in);
        VT *vout = static_cast<VT *>(out);
        if (plan->lut_in_lds) {
            if (narrow)
                hipLaunchKernelGGL((sector_gather_kernel<VT, true, COND, NB, RPT, uint32_t>), grid, block, lds, stream,
                                   vin, vout, a, q);
            else
                hipLaunchKernelGGL((sector_gather_kernel<VT, true, COND, NB, RPT, uint64_t>), grid, block, lds, stream,
                                   vin, vout, a, q);
        } else {
            if (narrow)
                hipLaunchKernelGGL((sector_gather_kernel<VT, false, COND, NB, RPT, uint32_t>), grid, block, 0, stream,
                                   vin, vout, a, q);
            else
                hipLaunchKernelGGL((sector_gather_kernel<VT, false, COND, NB, RPT, uint64_t>), grid, block, 0, stream,
                                   vin, vout, a, q);
        }
    };
    switch (a.nblocks) {
        case 1: launch(std::integral_constant<uint32_t, 1>()); break;
        case 2: launch(std::integral_constant<uint32_t, 2>()); break;
        case 3: launch(std::integral_constant<uint32_t, 3>()); break;
        default: launch(std::integral_constant<uint32_t, 4>()); break;
    }
    return hipGetLastError();
}

int apply_device(const tcp_sector_plan *plan, const void *in, void *out, int dtype, void *stream) {
    if (dtype != TCP_C128 && dtype != TCP_C64)
        return fail(TCP_ERR_INVALID_ARGUMENT, "dtype must be TCP_C128 or TCP_C64");
    if (plan->dim == 0) return TCP_OK;  // an empty sector: nothing to read or write
    if (!in || !out) return fail(TCP_ERR_INVALID_ARGUMENT, "null state buffer");
    if (plan->leaks) return leak_error(plan);
    const size_t bytes = static_cast<size_t>(plan->dim) * (dtype == TCP_C128 ? 16 : 8);
    if (tcp::ranges_overlap(in, bytes, out, bytes))
        return fail(TCP_ERR_INVALID_ARGUMENT, "apply_into input and output must not overlap");
    tcp::DeviceGuard dg;
    int st = dg.enter(plan->device);
    if (st) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (plan->wide) return (st = tcp::wide_apply(*plan->wide, in, out, dtype, s)) ? st : tcp::ok();
    if (plan->general) return (st = tcp::charge_apply(*plan->general, in, out, dtype, s)) ? st : tcp::ok();
    hipError_t e;
    if (plan->cond)
        e = dtype == TCP_C128 ? launch_gather<dvec2, true>(plan, in, out, dtype, s)
                              : launch_gather<fvec2, true>(plan, in, out, dtype, s);
    else
        e = dtype == TCP_C128 ? launch_gather<dvec2, false>(plan, in, out, dtype, s)
                              : launch_gather<fvec2, false>(plan, in, out, dtype, s);
    if (e != hipSuccess) return tcp::hip_fail(e, "sector kernel launch");
    return tcp::ok();
}

BlockSpec u1_spec(uint32_t n, uint32_t k) {
    BlockSpec s{};
    s.nblocks = 1;
    s.sizes[0] = n;
    s.pops[0] = k;
    return s;
}

int blocks_spec(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops, BlockSpec &s) {
    if (nblocks == 0 || !sizes || !pops) return fail(TCP_ERR_INVALID_SECTOR, "invalid sector: no blocks");
    if (nblocks > kMaxBlocks)
        return fail(TCP_ERR_UNSUPPORTED, "at most 4 contiguous blocks are supported on the GPU");
    s = BlockSpec{};
    s.nblocks = nblocks;
    for (uint32_t j = 0; j < nblocks; ++j) {
        s.sizes[j] = sizes[j];
        s.pops[j] = pops[j];
    }
    return TCP_OK;
}

// Host rank/unrank with the same tables as the kernels (used by the public
// scalar entry points).
int host_rank(const BlockSpec &spec, uint64_t bits, uint64_t &rank) {
    SectorArgs a;
    uint64_t dim;
    std::vector<uint32_t> lut;
    std::vector<uint64_t> ch;
    std::vector<uint32_t> cb;
    int st = sector_geometry(spec, a, dim, lut, ch, cb);
    if (st) return st;
    uint32_t n = 0;
    for (uint32_t j = 0; j < spec.nblocks; ++j) n += spec.sizes[j];
    if (n < 64 && (bits >> n))
        return fail(TCP_ERR_INVALID_INDEX, "bitstring is outside the computational basis");
    unsigned __int128 r = 0;
    for (uint32_t j = 0; j < spec.nblocks; ++j) {
        const uint64_t w = (bits >> a.shift[j]) & a.bmask[j];
        if (static_cast<uint32_t>(__builtin_popcountll(w)) != spec.pops[j])
            return fail(TCP_ERR_INVALID_INDEX, "invalid index: packed basis word has the wrong Hamming weight");
        uint64_t local = 0;
        uint32_t seen = 0;
        for (uint32_t q = 0; q < spec.sizes[j]; ++q)
            if ((w >> q) & 1) local += ch[q * 64 + (++seen)];
        uint64_t stride = 1;
        for (uint32_t i = j + 1; i < spec.nblocks; ++i) {
            uint64_t c;
            choose_checked(spec.sizes[i], spec.pops[i], c);
            stride *= c;
        }
        r += static_cast<unsigned __int128>(local) * stride;
    }
    rank = static_cast<uint64_t>(r);
    return TCP_OK;
}

int host_unrank(const BlockSpec &spec, uint64_t index, uint64_t &bits) {
    SectorArgs a;
    uint64_t dim;
    std::vector<uint32_t> lut;
    std::vector<uint64_t> ch;
    std::vector<uint32_t> cb;
    int st = sector_geometry(spec, a, dim, lut, ch, cb);
    if (st) return st;
    if (index >= dim) return fail(TCP_ERR_INVALID_INDEX, "invalid index: restricted basis index is out of range");
    bits = 0;
    for (uint32_t j = 0; j < spec.nblocks; ++j) {
        uint64_t stride = 1;
        for (uint32_t i = j + 1; i < spec.nblocks; ++i) {
            uint64_t c;
            choose_checked(spec.sizes[i], spec.pops[i], c);
            stride *= c;
        }
        uint64_t local = index / stride;
        index -= local * stride;
        uint32_t needed = spec.pops[j];
        uint64_t w = 0;
        for (int q = static_cast<int>(spec.sizes[j]) - 1; q >= 0 && needed > 0; --q) {
            const uint64_t c = ch[q * 64 + needed];
            if (local >= c) {
                local -= c;
                w |= 1ull << q;
                --needed;
            }
        }
        bits |= w << a.shift[j];
    }
    return TCP_OK;
}

}  // namespace

extern "C" {

int tcp_u1_dimension(uint32_t nqubits, uint32_t particle_number, uint64_t *dim) {
    if (!dim) return fail(TCP_ERR_INVALID_ARGUMENT, "null output");
    if (particle_number > nqubits)
        return fail(TCP_ERR_INVALID_SECTOR, "invalid sector: particle_number must be between 0 and nqubits");
    if (!choose_checked(nqubits, particle_number, *dim))
        return fail(TCP_ERR_OVERFLOW, "integer overflow while computing U1 sector dimension");
    return tcp::ok();
}

int tcp_u1_rank(uint32_t nqubits, uint32_t particle_number, uint64_t bitstring, uint64_t *rank) {
    if (!rank) return fail(TCP_ERR_INVALID_ARGUMENT, "null output");
    int st = host_rank(u1_spec(nqubits, particle_number), bitstring, *rank);
    return st ? st : tcp::ok();
}

int tcp_u1_unrank(uint32_t nqubits, uint32_t particle_number, uint64_t index, uint64_t *bitstring) {
    if (!bitstring) return fail(TCP_ERR_INVALID_ARGUMENT, "null output");
    int st = host_unrank(u1_spec(nqubits, particle_number), index, *bitstring);
    return st ? st : tcp::ok();
}

int tcp_u1_basis_words(uint32_t nqubits, uint32_t particle_number, uint64_t *out, uint64_t rows) {
    tcp::WideGeometry g;
    int st = tcp::wide_geometry(nqubits, particle_number, g);
    if (st) return st;
    if (!out && g.dim) return fail(TCP_ERR_INVALID_ARGUMENT, "null output");
    if (rows != g.dim) return check_len(g.dim, rows);
    if (nqubits <= 64) {
        const BlockSpec spec = u1_spec(nqubits, particle_number);
        for (uint64_t i = 0; i < g.dim; ++i)
            if ((st = host_unrank(spec, i, out[i]))) return st;
    } else {
        for (uint64_t i = 0; i < g.dim; ++i)
            if ((st = tcp::wide_unrank(g, i, out + i * g.words))) return st;
    }
    return tcp::ok();
}

int tcp_blocks_dimension(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops, uint64_t *dim) {
    if (!dim) return fail(TCP_ERR_INVALID_ARGUMENT, "null output");
    BlockSpec s;
    int st = blocks_spec(nblocks, sizes, pops, s);
    if (st) return st;
    SectorArgs a;
    std::vector<uint32_t> lut;
    std::vector<uint64_t> ch;
    std::vector<uint32_t> cb;
    st = sector_geometry(s, a, *dim, lut, ch, cb);
    return st ? st : tcp::ok();
}

int tcp_blocks_rank(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops, uint64_t bitstring,
                    uint64_t *rank) {
    if (!rank) return fail(TCP_ERR_INVALID_ARGUMENT, "null output");
    BlockSpec s;
    int st = blocks_spec(nblocks, sizes, pops, s);
    if (!st) st = host_rank(s, bitstring, *rank);
    return st ? st : tcp::ok();
}

int tcp_blocks_unrank(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops, uint64_t index,
                      uint64_t *bitstring) {
    if (!bitstring) return fail(TCP_ERR_INVALID_ARGUMENT, "null output");
    BlockSpec s;
    int st = blocks_spec(nblocks, sizes, pops, s);
    if (!st) st = host_unrank(s, index, *bitstring);
    return st ? st : tcp::ok();
}

int tcp_sector_plan_create_u1(const tcp_operator *op, uint32_t particle_number, uint64_t max_bytes,
                              tcp_sector_plan **out) {
    if (!op) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    const tcp::Operator &o = tcp::operator_of(op);
    if (particle_number > o.nqubits)
        return fail(TCP_ERR_INVALID_SECTOR, "invalid sector: particle_number must be between 0 and nqubits");
    return build_sector_plan(o, u1_spec(o.nqubits, particle_number), true, max_bytes, out);
}

int tcp_sector_plan_create_blocks(const tcp_operator *op, uint32_t nblocks, const uint32_t *sizes,
                                  const uint32_t *pops, uint64_t max_bytes, tcp_sector_plan **out) {
    if (!op) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    BlockSpec s;
    int st = blocks_spec(nblocks, sizes, pops, s);
    if (st) return st;
    return build_sector_plan(tcp::operator_of(op), s, false, max_bytes, out);
}

int tcp_sector_plan_create_fermion(const tcp_fermion *f, uint32_t nblocks, const uint32_t *sizes,
                                   const uint32_t *pops, uint64_t max_bytes, tcp_sector_plan **out) {
    if (!f) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    BlockSpec s;
    int st = blocks_spec(nblocks, sizes, pops, s);
    if (!st) st = build_sector_plan_impl(nullptr, f, s, false, max_bytes, out, true);
    return st ? st : tcp::ok();
}

int tcp_charge_sector_create(uint32_t naxes, uint32_t nconstraints, const int64_t *levels, const int64_t *targets,
                             uint64_t max_bytes, tcp_charge_sector **out) {
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    auto sector = new tcp_charge_sector();
    std::unique_ptr<tcp_charge_sector> guard(sector);
    int st = tcp::charge_geometry(naxes, nconstraints, levels, targets, max_bytes, sector->geo);
    if (st) return st;
    *out = guard.release();
    return tcp::ok();
}

uint64_t tcp_charge_sector_dimension(const tcp_charge_sector *sector) { return sector ? sector->geo.dim : 0; }

int tcp_charge_sector_rank(const tcp_charge_sector *sector, uint64_t bitstring, uint64_t *rank) {
    if (!sector || !rank) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    const int st = tcp::charge_rank(sector->geo, bitstring, *rank);
    return st ? st : tcp::ok();
}

int tcp_charge_sector_unrank(const tcp_charge_sector *sector, uint64_t index, uint64_t *bitstring) {
    if (!sector || !bitstring) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    const int st = tcp::charge_unrank(sector->geo, index, *bitstring);
    return st ? st : tcp::ok();
}

int tcp_charge_sector_basis_words(const tcp_charge_sector *sector, uint64_t first, uint64_t count, uint64_t *out) {
    if (!sector || (!out && count)) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    if (first > sector->geo.dim || count > sector->geo.dim - first)
        return fail(TCP_ERR_INVALID_INDEX, "invalid index: sector index is out of range");
    for (uint64_t i = 0; i < count; ++i) {
        const int st = tcp::charge_unrank(sector->geo, first + i, out[i]);
        if (st) return st;
    }
    return tcp::ok();
}

void tcp_charge_sector_destroy(tcp_charge_sector *sector) { delete sector; }

int tcp_sector_plan_create_charge(const tcp_operator *op, const tcp_charge_sector *sector, uint64_t max_bytes,
                                  tcp_sector_plan **out) {
    if (!op || !sector) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator or sector");
    const int st = build_charge_plan(&tcp::operator_of(op), nullptr, sector->geo, max_bytes, true, out);
    return st ? st : tcp::ok();
}

int tcp_sector_plan_create_fermion_charge(const tcp_fermion *f, const tcp_charge_sector *sector,
                                          uint64_t max_bytes, tcp_sector_plan **out) {
    if (!f || !sector) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator or sector");
    const int st = build_charge_plan(nullptr, f, sector->geo, max_bytes, true, out);
    return st ? st : tcp::ok();
}

uint64_t tcp_sector_plan_dimension(const tcp_sector_plan *plan) { return plan ? plan->dim : 0; }
uint32_t tcp_sector_plan_nqubits(const tcp_sector_plan *plan) { return plan ? plan->n : 0; }
uint64_t tcp_sector_plan_term_count(const tcp_sector_plan *plan) { return plan ? plan->T : 0; }
uint64_t tcp_sector_plan_group_count(const tcp_sector_plan *plan) { return plan ? plan->G : 0; }
uint64_t tcp_sector_plan_estimated_bytes(const tcp_sector_plan *plan) {
    return plan ? plan->retained_bytes : 0;
}

int tcp_sector_plan_apply_into(const tcp_sector_plan *plan, const void *input, void *output,
                               uint64_t input_len, uint64_t output_len, int dtype, void *stream) {
    if (!plan) return fail(TCP_ERR_INVALID_ARGUMENT, "null plan");
    int st;
    if ((st = check_len(plan->dim, input_len))) return st;
    if ((st = check_len(plan->dim, output_len))) return st;
    return apply_device(plan, input, output, dtype, stream);
}

int tcp_sector_plan_apply_host(const tcp_sector_plan *plan, const double *input, double *output,
                               uint64_t input_len, uint64_t output_len) {
    if (!plan) return fail(TCP_ERR_INVALID_ARGUMENT, "null plan");
    int st;
    if ((st = check_len(plan->dim, input_len))) return st;
    if ((st = check_len(plan->dim, output_len))) return st;
    if (plan->dim == 0) return tcp::ok();
    if (!input || !output) return fail(TCP_ERR_INVALID_ARGUMENT, "null state buffer");
    const size_t bytes = static_cast<size_t>(plan->dim) * 16;
    if (tcp::ranges_overlap(input, bytes, output, bytes))
        return fail(TCP_ERR_INVALID_ARGUMENT, "apply_into input and output must not overlap");
    tcp::DeviceGuard dg;
    if ((st = dg.enter(plan->device))) return st;
    auto *mp = const_cast<tcp_sector_plan *>(plan);
    std::lock_guard<std::mutex> lock(mp->staging_mutex);
    if (mp->stage_in.bytes < bytes && (st = mp->stage_in.allocate(bytes))) return st;
    if (mp->stage_out.bytes < bytes && (st = mp->stage_out.allocate(bytes))) return st;
    TCP_HIP_CHECK(hipMemcpy(mp->stage_in.ptr, input, bytes, hipMemcpyHostToDevice));
    if ((st = apply_device(plan, mp->stage_in.ptr, mp->stage_out.ptr, TCP_C128, nullptr))) return st;
    TCP_HIP_CHECK(hipMemcpy(output, mp->stage_out.ptr, bytes, hipMemcpyDeviceToHost));
    return tcp::ok();
}

int tcp_sector_plan_rank_device(const tcp_sector_plan *plan, const uint64_t *bitstrings, uint64_t *ranks,
                                uint64_t count, void *stream) {
    if (!plan || (count && (!bitstrings || !ranks))) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    if (plan->wide)
        return fail(TCP_ERR_UNSUPPORTED, "device rank/unrank of sectors of 64 or more qubits is not supported; "
                                         "use tcp_u1_rank_words / tcp_u1_unrank_words");
    if (count == 0) return tcp::ok();
    tcp::DeviceGuard dg;
    int st = dg.enter(plan->device);
    if (st) return st;
    if (plan->general) return (st = tcp::charge_rank_device(*plan->general, bitstrings, ranks, count, static_cast<hipStream_t>(stream))) ? st : tcp::ok();
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((count + 255) / 256, 8192));
    hipLaunchKernelGGL(sector_rank_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       plan->args(TCP_C128), bitstrings, ranks, count, plan->n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? tcp::ok() : tcp::hip_fail(e, "sector rank kernel");
}

int tcp_sector_plan_unrank_device(const tcp_sector_plan *plan, const uint64_t *ranks, uint64_t *bitstrings,
                                  uint64_t count, void *stream) {
    if (!plan || (count && (!bitstrings || !ranks))) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    if (plan->wide)
        return fail(TCP_ERR_UNSUPPORTED, "device rank/unrank of sectors of 64 or more qubits is not supported; "
                                         "use tcp_u1_rank_words / tcp_u1_unrank_words");
    if (count == 0) return tcp::ok();
    tcp::DeviceGuard dg;
    int st = dg.enter(plan->device);
    if (st) return st;
    if (plan->general) return (st = tcp::charge_unrank_device(*plan->general, ranks, bitstrings, count, static_cast<hipStream_t>(stream))) ? st : tcp::ok();
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((count + 255) / 256, 8192));
    hipLaunchKernelGGL(sector_unrank_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       plan->args(TCP_C128), ranks, bitstrings, count);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? tcp::ok() : tcp::hip_fail(e, "sector unrank kernel");
}

int tcp_sector_plan_walk_device(const tcp_sector_plan *plan, uint64_t first, uint64_t *bitstrings,
                                uint64_t count, uint32_t run, void *stream) {
    if (!plan || (count && !bitstrings) || run == 0) return fail(TCP_ERR_INVALID_ARGUMENT, "bad argument");
    if (plan->wide)
        return fail(TCP_ERR_UNSUPPORTED, "device rank/unrank of sectors of 64 or more qubits is not supported; "
                                         "use tcp_u1_rank_words / tcp_u1_unrank_words");
    if (count == 0) return tcp::ok();
    if (first > plan->dim || count > plan->dim - first)
        return fail(TCP_ERR_INVALID_INDEX, "invalid index: restricted basis index is out of range");
    tcp::DeviceGuard dg;
    int st = dg.enter(plan->device);
    if (st) return st;
    if (plan->general)
        return (st = tcp::charge_walk_device(*plan->general, first, bitstrings, count, static_cast<hipStream_t>(stream)))
                   ? st
                   : tcp::ok();
    const uint64_t threads = (count + run - 1) / run;
    const uint32_t blocks = static_cast<uint32_t>((threads + 255) / 256);
    hipLaunchKernelGGL(sector_walk_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       plan->args(TCP_C128), first, bitstrings, count, run);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? tcp::ok() : tcp::hip_fail(e, "sector walk kernel");
}

void tcp_sector_plan_destroy(tcp_sector_plan *plan) { delete plan; }

}  // extern "C"

// ---- eager restricted operator (CSR) ------------------------------------

struct tcp_csr_plan {
    uint32_t n = 0;
    uint64_t dim = 0, nnz = 0;
    int device = 0;
    tcp::DeviceBuffer indptr, cols, vals64, vals32;
    // compact form (used by the SpMV when `compact`): packed column | value
    // index per entry, the value table and the dense diagonal, FP64 and FP32
    bool compact = false;
    uint32_t ntable = 0;
    tcp::DeviceBuffer packed, table64, table32, diag64, diag32;
    // slab form (used by the SpMV when `slab`; built on the compact form,
    // whose packed entries it replaces): rows in chunks of kSlabRows, word
    // slab_word(row, q, nslab) holds row's entry of off-diagonal x-group slot
    // q, packed as in the compact form, or kSlabEmpty when the row has no
    // transition in that group
    bool slab = false;
    uint32_t nslab = 0;
    tcp::DeviceBuffer slab_words;
    uint64_t retained_bytes = 0;
    std::mutex staging_mutex;
    tcp::DeviceBuffer stage_in, stage_out;
};

namespace {

constexpr uint32_t kCsrLanes = 8;  // lanes per row in the SpMV (rows average ~27 entries on lattice models)

// The compact form is optional: when its buffers cannot be allocated the plan
// keeps plain CSR, as when it does not apply.
void drop_compact(tcp_csr_plan *plan) {
    plan->packed.release();
    plan->diag64.release();
    plan->diag32.release();
    plan->table64.release();
    plan->table32.release();
    (void)hipGetLastError();  // the failed allocation is not this call's error
}

// Builds the compact form of a filled plan when it applies: fewer than 2^24
// rows, at most kDictMax distinct off-diagonal values, room in the budget
// while both forms exist, and device memory for it.  Otherwise the plan keeps
// plain CSR (not an error).  On success the FP32 values are dropped (the
// compact form carries them).
int compact_csr(tcp_csr_plan *plan, uint64_t max_bytes) {
    const uint64_t dim = plan->dim, nnz = plan->nnz;
    if (dim == 0 || nnz == 0 || dim > kColMask + 1ull) return TCP_OK;
    const uint64_t extra = 4 * nnz + 24 * dim + 64 * kDictSlots;
    if (tcp::exceeds_budget(plan->retained_bytes + extra, max_bytes)) return TCP_OK;
    tcp::DeviceBuffer slots, slot_vals, state, slot_index;
    if (slots.allocate(8 * kDictSlots) || slot_vals.allocate(16 * kDictSlots) || state.allocate(8)) {
        drop_compact(plan);
        return TCP_OK;
    }
    TCP_HIP_CHECK(hipMemset(slots.ptr, 0, 8 * kDictSlots));
    TCP_HIP_CHECK(hipMemset(state.ptr, 0, 8));
    const auto *ip = static_cast<const uint64_t *>(plan->indptr.ptr);
    const auto *cp = static_cast<const uint32_t *>(plan->cols.ptr);
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((dim + 255) / 256, 256 * 16));
    hipLaunchKernelGGL(csr_dict_insert_kernel, dim3(blocks), dim3(256), 0, nullptr, ip, cp,
                       static_cast<const dvec2 *>(plan->vals64.ptr), dim,
                       static_cast<unsigned long long *>(slots.ptr), static_cast<dvec2 *>(slot_vals.ptr),
                       static_cast<uint32_t *>(state.ptr));
    TCP_HIP_CHECK(hipGetLastError());
    uint32_t hstate[2];
    TCP_HIP_CHECK(hipMemcpy(hstate, state.ptr, 8, hipMemcpyDeviceToHost));
    if (hstate[1]) return TCP_OK;
    std::vector<unsigned long long> hslots(kDictSlots);
    std::vector<dvec2> hvals(kDictSlots);
    TCP_HIP_CHECK(hipMemcpy(hslots.data(), slots.ptr, 8 * kDictSlots, hipMemcpyDeviceToHost));
    TCP_HIP_CHECK(hipMemcpy(hvals.data(), slot_vals.ptr, 16 * kDictSlots, hipMemcpyDeviceToHost));
    std::vector<uint32_t> index(kDictSlots, kDiagIndex);
    std::vector<dvec2> t64;
    std::vector<fvec2> t32;
    for (uint32_t k = 0; k < kDictSlots; ++k)
        if (hslots[k]) {
            index[k] = static_cast<uint32_t>(t64.size());
            t64.push_back(hvals[k]);
            t32.push_back(fvec2{static_cast<float>(hvals[k].x), static_cast<float>(hvals[k].y)});
        }
    if (t64.size() > kDictMax) return TCP_OK;
    const uint32_t ntable = static_cast<uint32_t>(t64.size());
    const size_t tb = std::max<size_t>(ntable, 1);
    t64.resize(tb, dvec2{0.0, 0.0});
    t32.resize(tb, fvec2{0.f, 0.f});
    if (slot_index.upload(index.data(), 4 * kDictSlots) || plan->table64.upload(t64.data(), 16 * tb) ||
        plan->table32.upload(t32.data(), 8 * tb) || plan->packed.allocate(4 * nnz) ||
        plan->diag64.allocate(16 * dim) || plan->diag32.allocate(8 * dim)) {
        drop_compact(plan);
        return TCP_OK;
    }
    TCP_HIP_CHECK(hipMemset(plan->diag64.ptr, 0, 16 * dim));
    TCP_HIP_CHECK(hipMemset(plan->diag32.ptr, 0, 8 * dim));
    hipLaunchKernelGGL(csr_pack_kernel, dim3(blocks), dim3(256), 0, nullptr, ip, cp,
                       static_cast<const dvec2 *>(plan->vals64.ptr), static_cast<const fvec2 *>(plan->vals32.ptr), dim,
                       static_cast<const unsigned long long *>(slots.ptr), static_cast<const uint32_t *>(slot_index.ptr),
                       static_cast<const dvec2 *>(plan->table64.ptr), static_cast<const fvec2 *>(plan->table32.ptr),
                       static_cast<uint32_t *>(plan->packed.ptr), static_cast<dvec2 *>(plan->diag64.ptr),
                       static_cast<fvec2 *>(plan->diag32.ptr), static_cast<uint32_t *>(state.ptr));
    TCP_HIP_CHECK(hipGetLastError());
    TCP_HIP_CHECK(hipMemcpy(hstate, state.ptr, 8, hipMemcpyDeviceToHost));
    if (hstate[1]) {  // a hash collision: keep plain CSR
        drop_compact(plan);
        return TCP_OK;
    }
    plan->compact = true;
    plan->ntable = ntable;
    plan->vals32.release();
    plan->retained_bytes = plan->retained_bytes - 8 * nnz + 4 * nnz + 24 * dim + 24 * tb;
    return TCP_OK;
}

// Builds the slab form of a compact plan when it pays: the slots (distinct
// off-diagonal x-masks that hold entries) times the rows may exceed the
// off-diagonal entries by at most kSlabMaxPad (the empty words are read too),
// and the budget and device memory must allow it.  `states(bits)` writes the
// bitstring of every row; `xs` lists the operator's x-masks.  Otherwise, or
// when an entry's x-mask is not among them, the plan stays compact (not an
// error).  On success the packed entries are dropped (the slab carries them).
constexpr double kSlabMaxPad = 8.0 / 3.0;  // slab words per off-diagonal entry (fill >= 3/8)

template <class States>
int slab_csr(tcp_csr_plan *plan, States states, std::vector<uint64_t> xs, uint64_t max_bytes) {
    if (!plan->compact || plan->dim == 0) return TCP_OK;
    std::sort(xs.begin(), xs.end());
    xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
    xs.erase(std::remove(xs.begin(), xs.end(), 0ull), xs.end());
    if (xs.empty() || xs.size() > (1u << 20)) return TCP_OK;
    const uint64_t dim = plan->dim;
    const uint32_t nx = static_cast<uint32_t>(xs.size());
    tcp::DeviceBuffer bits, dxs, used, slot, flags;
    auto quit = [&]() {  // the slab is optional: a failed allocation is not this call's error
        (void)hipGetLastError();
        return TCP_OK;
    };
    if (tcp::exceeds_budget(plan->retained_bytes + 8 * dim + 12 * static_cast<uint64_t>(nx), max_bytes)) return TCP_OK;
    if (bits.allocate(8 * dim) || dxs.upload(xs.data(), 8 * xs.size()) || used.allocate(4 * nx) ||
        slot.allocate(4 * nx) || flags.allocate(16))
        return quit();
    int st;
    if ((st = states(static_cast<uint64_t *>(bits.ptr)))) return st;
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((dim + 255) / 256, 256 * 16));
    hipLaunchKernelGGL(fill_words_kernel, dim3(std::max<uint32_t>(1, std::min<uint32_t>((nx + 255) / 256, 1024))),
                       dim3(256), 0, nullptr, static_cast<uint32_t *>(used.ptr), uint64_t(nx), 0u);
    hipLaunchKernelGGL(fill_words_kernel, dim3(1), dim3(64), 0, nullptr, static_cast<uint32_t *>(flags.ptr),
                       uint64_t(4), 0u);
    auto *offdiag = static_cast<unsigned long long *>(flags.ptr);
    auto *bad = static_cast<uint32_t *>(flags.ptr) + 2;
    const auto *ip = static_cast<const uint64_t *>(plan->indptr.ptr);
    const auto *pk = static_cast<const uint32_t *>(plan->packed.ptr);
    hipLaunchKernelGGL(slab_scatter_kernel, dim3(blocks), dim3(256), 0, nullptr, ip, pk,
                       static_cast<const uint64_t *>(bits.ptr), dim, static_cast<const uint64_t *>(dxs.ptr), nx,
                       static_cast<uint32_t *>(used.ptr), nullptr, 0u, nullptr, offdiag, bad);
    TCP_HIP_CHECK(hipGetLastError());
    uint32_t hflags[4];
    std::vector<uint32_t> hused(nx);
    TCP_HIP_CHECK(hipMemcpy(hflags, flags.ptr, 16, hipMemcpyDeviceToHost));
    TCP_HIP_CHECK(hipMemcpy(hused.data(), used.ptr, 4 * nx, hipMemcpyDeviceToHost));
    if (hflags[2]) return TCP_OK;  // an entry outside the listed x-masks: stay compact
    const uint64_t noff = static_cast<uint64_t>(hflags[0]) | (static_cast<uint64_t>(hflags[1]) << 32);
    std::vector<uint32_t> hslot(nx, 0);
    uint32_t nslab = 0;
    for (uint32_t g = 0; g < nx; ++g)
        if (hused[g]) hslot[g] = nslab++;
    if (nslab == 0 || static_cast<double>(nslab) * static_cast<double>(dim) > kSlabMaxPad * static_cast<double>(noff))
        return TCP_OK;
    const uint64_t words = static_cast<uint64_t>(nslab) * ((dim + kSlabRows - 1) / kSlabRows) * kSlabRows;
    const uint64_t retained = plan->retained_bytes - 4 * plan->nnz + 4 * words;
    if (tcp::exceeds_budget(plan->retained_bytes + 4 * words, max_bytes)) return TCP_OK;
    if (slot.upload(hslot.data(), 4 * nx) || plan->slab_words.allocate(4 * words)) {
        plan->slab_words.release();
        return quit();
    }
    hipLaunchKernelGGL(fill_words_kernel, dim3(static_cast<uint32_t>(std::min<uint64_t>((words + 255) / 256, 256 * 64))),
                       dim3(256), 0, nullptr, static_cast<uint32_t *>(plan->slab_words.ptr), words, kSlabEmpty);
    hipLaunchKernelGGL(slab_scatter_kernel, dim3(blocks), dim3(256), 0, nullptr, ip, pk,
                       static_cast<const uint64_t *>(bits.ptr), dim, static_cast<const uint64_t *>(dxs.ptr), nx,
                       static_cast<uint32_t *>(used.ptr), static_cast<const uint32_t *>(slot.ptr), nslab,
                       static_cast<uint32_t *>(plan->slab_words.ptr), offdiag, bad);
    TCP_HIP_CHECK(hipGetLastError());
    TCP_HIP_CHECK(hipDeviceSynchronize());
    plan->slab = true;
    plan->nslab = nslab;
    plan->packed.release();
    plan->retained_bytes = retained;
    return TCP_OK;
}

// Two-pass CSR assembly: `count` writes the entries of every row, the host
// scans them into indptr, `fill` writes each row's columns (ascending) and
// values in FP64 and FP32.
struct NoSlab {
    int operator()(tcp_csr_plan *) const { return TCP_OK; }
};

template <class Count, class Fill, class Slab = NoSlab>
int assemble_csr(uint32_t n, uint64_t dim, int device, uint64_t max_bytes, Count count, Fill fill,
                 tcp_csr_plan **out, Slab slab = Slab()) {
    auto plan = new tcp_csr_plan();
    std::unique_ptr<tcp_csr_plan> guard(plan);
    plan->n = n;
    plan->dim = dim;
    plan->device = device;

    // pass 1: entries per row; exclusive scan on the host
    tcp::DeviceBuffer counts;
    int st;
    if ((st = counts.allocate(4 * std::max<uint64_t>(dim, 1)))) return st;
    if ((st = count(static_cast<uint32_t *>(counts.ptr)))) return st;
    std::vector<uint32_t> hcounts(dim);
    TCP_HIP_CHECK(hipMemcpy(hcounts.data(), counts.ptr, 4 * dim, hipMemcpyDeviceToHost));
    counts.release();
    std::vector<uint64_t> hptr(dim + 1, 0);
    for (uint64_t r = 0; r < dim; ++r) hptr[r + 1] = hptr[r] + hcounts[r];
    plan->nnz = hptr[dim];
    plan->retained_bytes = sizeof(tcp_csr_plan) + (dim + 1) * 8 + plan->nnz * (4 + 16 + 8);
    if (tcp::exceeds_budget(plan->retained_bytes, max_bytes))
        return tcp::memory_limit(plan->retained_bytes, max_bytes);

    // pass 2: fill rows, columns ascending
    if ((st = plan->indptr.upload(hptr.data(), 8 * (dim + 1)))) return st;
    if ((st = plan->cols.allocate(4 * std::max<uint64_t>(plan->nnz, 1)))) return st;
    if ((st = plan->vals64.allocate(16 * std::max<uint64_t>(plan->nnz, 1)))) return st;
    if ((st = plan->vals32.allocate(8 * std::max<uint64_t>(plan->nnz, 1)))) return st;
    if ((st = fill(static_cast<const uint64_t *>(plan->indptr.ptr), static_cast<uint32_t *>(plan->cols.ptr),
                   plan->vals64.ptr, plan->vals32.ptr)))
        return st;
    TCP_HIP_CHECK(hipDeviceSynchronize());
    if ((st = compact_csr(plan, max_bytes))) return st;
    if ((st = slab(plan))) return st;
    *out = guard.release();
    return TCP_OK;
}

int build_csr(const tcp_sector_plan *sp, uint64_t max_bytes, tcp_csr_plan **out) {
    if (sp->leaks) return leak_error(sp);
    const uint64_t dim = sp->dim;
    const SectorArgs a = sp->args(TCP_C128);
    const uint64_t per_block = static_cast<uint64_t>(kSectorThreads) * kRanksPerThread;
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((dim + per_block - 1) / per_block, 256 * 16));
    const size_t lds = sp->lut_in_lds ? a.lut_words * 4 : 0;
    auto count_with = [&](auto cond, uint32_t *counts) -> int {
        constexpr bool C = decltype(cond)::value;
        if (sp->lut_in_lds)
            hipLaunchKernelGGL((csr_count_kernel<true, C>), dim3(blocks), dim3(kSectorThreads), lds, nullptr, a,
                               counts);
        else
            hipLaunchKernelGGL((csr_count_kernel<false, C>), dim3(blocks), dim3(kSectorThreads), 0, nullptr, a,
                               counts);
        TCP_HIP_CHECK(hipGetLastError());
        return TCP_OK;
    };
    auto fill_with = [&](auto cond, const uint64_t *indptr, uint32_t *cols, void *v64, void *v32) -> int {
        constexpr bool C = decltype(cond)::value;
        if (sp->lut_in_lds)
            hipLaunchKernelGGL((csr_fill_kernel<true, C>), dim3(blocks), dim3(kSectorThreads), lds, nullptr, a,
                               indptr, cols, static_cast<dvec2 *>(v64), static_cast<fvec2 *>(v32));
        else
            hipLaunchKernelGGL((csr_fill_kernel<false, C>), dim3(blocks), dim3(kSectorThreads), 0, nullptr, a,
                               indptr, cols, static_cast<dvec2 *>(v64), static_cast<fvec2 *>(v32));
        TCP_HIP_CHECK(hipGetLastError());
        return TCP_OK;
    };
    auto count = [&](uint32_t *counts) -> int {
        if (sp->wide) return tcp::wide_csr_counts(*sp->wide, counts, nullptr);
        if (sp->general) return tcp::charge_csr_counts(*sp->general, counts, nullptr);
        return sp->cond ? count_with(std::true_type(), counts) : count_with(std::false_type(), counts);
    };
    auto fill = [&](const uint64_t *indptr, uint32_t *cols, void *v64, void *v32) -> int {
        if (sp->wide) return tcp::wide_csr_fill(*sp->wide, indptr, cols, v64, v32, nullptr);
        if (sp->general) return tcp::charge_csr_fill(*sp->general, indptr, cols, v64, v32, nullptr);
        return sp->cond ? fill_with(std::true_type(), indptr, cols, v64, v32)
                        : fill_with(std::false_type(), indptr, cols, v64, v32);
    };
    // the slab form's rows and x-masks: the plan's walk and its x-groups
    auto slab = [&](tcp_csr_plan *plan) -> int {
        if (sp->wide) return TCP_OK;
        const tcp::DeviceBuffer &gxb = sp->general ? sp->general->gx : sp->gx;
        const uint64_t ng = sp->general ? sp->general->G : sp->G;
        std::vector<uint64_t> xs(ng);
        if (ng) TCP_HIP_CHECK(hipMemcpy(xs.data(), gxb.ptr, 8 * ng, hipMemcpyDeviceToHost));
        auto states = [&](uint64_t *bits) -> int { return tcp_sector_plan_walk_device(sp, 0, bits, dim, 64, nullptr); };
        return slab_csr(plan, states, std::move(xs), max_bytes);
    };
    return assemble_csr(sp->n, dim, sp->device, max_bytes, count, fill, out, slab);
}

int csr_apply_device(const tcp_csr_plan *plan, const void *in, void *out, int dtype, void *stream) {
    if (dtype != TCP_C128 && dtype != TCP_C64)
        return fail(TCP_ERR_INVALID_ARGUMENT, "dtype must be TCP_C128 or TCP_C64");
    if (plan->dim == 0) return TCP_OK;
    if (!in || !out) return fail(TCP_ERR_INVALID_ARGUMENT, "null state buffer");
    const size_t bytes = static_cast<size_t>(plan->dim) * (dtype == TCP_C128 ? 16 : 8);
    if (tcp::ranges_overlap(in, bytes, out, bytes))
        return fail(TCP_ERR_INVALID_ARGUMENT, "apply_into input and output must not overlap");
    tcp::DeviceGuard dg;
    int st = dg.enter(plan->device);
    if (st) return st;
    const uint64_t rows_per_block = 256 / kCsrLanes;
    const uint32_t blocks = static_cast<uint32_t>(
        std::max<uint64_t>(1, std::min<uint64_t>((plan->dim + rows_per_block - 1) / rows_per_block, 256 * 64)));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint64_t *ip = static_cast<const uint64_t *>(plan->indptr.ptr);
    const uint32_t *cp = static_cast<const uint32_t *>(plan->cols.ptr);
    if (plan->slab) {
        const uint32_t sblocks = static_cast<uint32_t>((plan->dim + kSlabRows - 1) / kSlabRows);
        const auto *sw = static_cast<const uint32_t *>(plan->slab_words.ptr);
        if (dtype == TCP_C128)
            hipLaunchKernelGGL(slab_spmv_kernel<dvec2>, dim3(sblocks), dim3(256), 0, s, sw, plan->nslab,
                               static_cast<const dvec2 *>(plan->table64.ptr), plan->ntable,
                               static_cast<const dvec2 *>(plan->diag64.ptr), static_cast<const dvec2 *>(in),
                               static_cast<dvec2 *>(out), plan->dim);
        else
            hipLaunchKernelGGL(slab_spmv_kernel<fvec2>, dim3(sblocks), dim3(256), 0, s, sw, plan->nslab,
                               static_cast<const fvec2 *>(plan->table32.ptr), plan->ntable,
                               static_cast<const fvec2 *>(plan->diag32.ptr), static_cast<const fvec2 *>(in),
                               static_cast<fvec2 *>(out), plan->dim);
    } else if (plan->compact) {
        const auto *pk = static_cast<const uint32_t *>(plan->packed.ptr);
        if (dtype == TCP_C128)
            hipLaunchKernelGGL((csr_compact_spmv_kernel<dvec2, kCsrLanes>), dim3(blocks), dim3(256), 0, s, ip, pk,
                               static_cast<const dvec2 *>(plan->table64.ptr), plan->ntable,
                               static_cast<const dvec2 *>(plan->diag64.ptr), static_cast<const dvec2 *>(in),
                               static_cast<dvec2 *>(out), plan->dim);
        else
            hipLaunchKernelGGL((csr_compact_spmv_kernel<fvec2, kCsrLanes>), dim3(blocks), dim3(256), 0, s, ip, pk,
                               static_cast<const fvec2 *>(plan->table32.ptr), plan->ntable,
                               static_cast<const fvec2 *>(plan->diag32.ptr), static_cast<const fvec2 *>(in),
                               static_cast<fvec2 *>(out), plan->dim);
    } else if (dtype == TCP_C128)
        hipLaunchKernelGGL((csr_spmv_kernel<dvec2, dvec2, kCsrLanes>), dim3(blocks), dim3(256), 0, s, ip, cp,
                           static_cast<const dvec2 *>(plan->vals64.ptr), static_cast<const dvec2 *>(in),
                           static_cast<dvec2 *>(out), plan->dim);
    else
        hipLaunchKernelGGL((csr_spmv_kernel<fvec2, fvec2, kCsrLanes>), dim3(blocks), dim3(256), 0, s, ip, cp,
                           static_cast<const fvec2 *>(plan->vals32.ptr), static_cast<const fvec2 *>(in),
                           static_cast<fvec2 *>(out), plan->dim);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? tcp::ok() : tcp::hip_fail(e, "csr kernel launch");
}

// out[r] = Re x[r] (the dense diagonal of the compact and slab forms).
__global__ void real_part_kernel(const dvec2 *__restrict__ x, uint64_t dim, double *__restrict__ out) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < dim;
         r += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        out[r] = x[r].x;
}

// Real part of the stored diagonal entry of every row (0 where the row has none).
__global__ void csr_diagonal_kernel(const uint64_t *__restrict__ indptr, const uint32_t *__restrict__ cols,
                                    const dvec2 *__restrict__ vals, uint64_t dim, double *__restrict__ diag) {
    for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < dim;
         r += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        double d = 0.0;
        for (uint64_t k = indptr[r]; k < indptr[r + 1]; ++k)
            if (cols[k] == r) {
                d = vals[k].x;
                break;
            }
        diag[r] = d;
    }
}

int csr_from_spec(const tcp::Operator &op, const BlockSpec &spec, bool u1, uint64_t max_bytes, tcp_csr_plan **out) {
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    tcp_sector_plan *sp = nullptr;
    int st = build_sector_plan(op, spec, u1, max_bytes, &sp, false);
    if (st) return st;
    std::unique_ptr<tcp_sector_plan> guard(sp);
    tcp::DeviceGuard dg;
    if ((st = dg.enter(sp->device))) return st;
    return build_csr(sp, max_bytes, out);
}

}  // namespace

namespace tcp {
uint64_t plan_dimension(const tcp_csr_plan *plan) { return plan->dim; }
int plan_device(const tcp_csr_plan *plan) { return plan->device; }
int plan_apply(const tcp_csr_plan *plan, const void *in, void *out, int dtype, hipStream_t stream) {
    return csr_apply_device(plan, in, out, dtype, stream);
}
int plan_diagonal(const tcp_csr_plan *plan, double *diag, hipStream_t stream) {
    if (plan->dim == 0) return TCP_OK;
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((plan->dim + 255) / 256, 4096));
    if (plan->compact)  // the compact and slab forms hold the diagonal densely
        hipLaunchKernelGGL(real_part_kernel, dim3(blocks), dim3(256), 0, stream,
                           static_cast<const dvec2 *>(plan->diag64.ptr), plan->dim, diag);
    else
        hipLaunchKernelGGL(csr_diagonal_kernel, dim3(blocks), dim3(256), 0, stream,
                           static_cast<const uint64_t *>(plan->indptr.ptr),
                           static_cast<const uint32_t *>(plan->cols.ptr), static_cast<const dvec2 *>(plan->vals64.ptr),
                           plan->dim, diag);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TCP_OK : hip_fail(e, "csr diagonal kernel");
}
}  // namespace tcp

extern "C" {

int tcp_csr_plan_create_u1(const tcp_operator *op, uint32_t particle_number, uint64_t max_bytes, tcp_csr_plan **out) {
    if (!op) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    const tcp::Operator &o = tcp::operator_of(op);
    if (particle_number > o.nqubits)
        return fail(TCP_ERR_INVALID_SECTOR, "invalid sector: particle_number must be between 0 and nqubits");
    const int st = csr_from_spec(o, u1_spec(o.nqubits, particle_number), true, max_bytes, out);
    return st ? st : tcp::ok();
}

int tcp_csr_plan_create_blocks(const tcp_operator *op, uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops,
                               uint64_t max_bytes, tcp_csr_plan **out) {
    if (!op) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    BlockSpec s;
    int st = blocks_spec(nblocks, sizes, pops, s);
    if (!st) st = csr_from_spec(tcp::operator_of(op), s, false, max_bytes, out);
    return st ? st : tcp::ok();
}

int tcp_csr_plan_create_full(const tcp_operator *op, uint64_t max_bytes, tcp_csr_plan **out) {
    if (!op) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    const tcp::Operator &o = tcp::operator_of(op);
    tcp::FullGroups f;
    int st = tcp::full_groups_build(o, f);
    if (st) return st;
    int device = 0;
    TCP_HIP_CHECK(hipGetDevice(&device));
    auto count = [&](uint32_t *counts) { return tcp::full_csr_counts(f, counts, nullptr); };
    auto fill = [&](const uint64_t *indptr, uint32_t *cols, void *v64, void *v32) {
        return tcp::full_csr_fill(f, indptr, cols, v64, v32, nullptr);
    };
    st = assemble_csr(o.nqubits, f.dim, device, max_bytes, count, fill, out);
    return st ? st : tcp::ok();
}

int tcp_csr_plan_create_fermion(const tcp_fermion *f, uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops,
                                uint64_t max_bytes, tcp_csr_plan **out) {
    if (!f) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    BlockSpec s;
    int st = blocks_spec(nblocks, sizes, pops, s);
    if (st) return st;
    tcp_sector_plan *sp = nullptr;
    if ((st = build_sector_plan_impl(nullptr, f, s, false, max_bytes, &sp, false))) return st;
    std::unique_ptr<tcp_sector_plan> guard(sp);
    tcp::DeviceGuard dg;
    if ((st = dg.enter(sp->device))) return st;
    st = build_csr(sp, max_bytes, out);
    return st ? st : tcp::ok();
}

}  // extern "C"

namespace {
int csr_from_charge(const tcp::Operator *op, const tcp_fermion *f, const tcp_charge_sector *sector,
                    uint64_t max_bytes, tcp_csr_plan **out) {
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (!sector) return fail(TCP_ERR_INVALID_ARGUMENT, "null sector");
    tcp_sector_plan *sp = nullptr;
    int st = build_charge_plan(op, f, sector->geo, max_bytes, false, &sp);
    if (st) return st;
    std::unique_ptr<tcp_sector_plan> guard(sp);
    tcp::DeviceGuard dg;
    if ((st = dg.enter(sp->device))) return st;
    return build_csr(sp, max_bytes, out);
}
}  // namespace

extern "C" {

int tcp_csr_plan_create_charge(const tcp_operator *op, const tcp_charge_sector *sector, uint64_t max_bytes,
                               tcp_csr_plan **out) {
    if (!op) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    const int st = csr_from_charge(&tcp::operator_of(op), nullptr, sector, max_bytes, out);
    return st ? st : tcp::ok();
}

int tcp_csr_plan_create_fermion_charge(const tcp_fermion *f, const tcp_charge_sector *sector, uint64_t max_bytes,
                                       tcp_csr_plan **out) {
    if (!f) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    const int st = csr_from_charge(nullptr, f, sector, max_bytes, out);
    return st ? st : tcp::ok();
}

uint64_t tcp_csr_plan_dimension(const tcp_csr_plan *plan) { return plan ? plan->dim : 0; }
uint64_t tcp_csr_plan_transition_count(const tcp_csr_plan *plan) { return plan ? plan->nnz : 0; }
uint64_t tcp_csr_plan_estimated_bytes(const tcp_csr_plan *plan) { return plan ? plan->retained_bytes : 0; }
int tcp_csr_plan_is_compact(const tcp_csr_plan *plan) { return plan && plan->compact ? 1 : 0; }
int tcp_csr_plan_storage_form(const tcp_csr_plan *plan) {
    return !plan ? -1 : plan->slab ? TCP_CSR_SLAB : plan->compact ? TCP_CSR_COMPACT : TCP_CSR_PLAIN;
}
uint32_t tcp_csr_plan_slab_groups(const tcp_csr_plan *plan) { return plan && plan->slab ? plan->nslab : 0; }

int tcp_csr_plan_apply_into(const tcp_csr_plan *plan, const void *input, void *output, uint64_t input_len,
                            uint64_t output_len, int dtype, void *stream) {
    if (!plan) return fail(TCP_ERR_INVALID_ARGUMENT, "null plan");
    int st;
    if ((st = check_len(plan->dim, input_len))) return st;
    if ((st = check_len(plan->dim, output_len))) return st;
    return csr_apply_device(plan, input, output, dtype, stream);
}

int tcp_csr_plan_apply_host(const tcp_csr_plan *plan, const double *input, double *output, uint64_t input_len,
                            uint64_t output_len) {
    if (!plan) return fail(TCP_ERR_INVALID_ARGUMENT, "null plan");
    int st;
    if ((st = check_len(plan->dim, input_len))) return st;
    if ((st = check_len(plan->dim, output_len))) return st;
    if (plan->dim == 0) return tcp::ok();
    if (!input || !output) return fail(TCP_ERR_INVALID_ARGUMENT, "null state buffer");
    const size_t bytes = static_cast<size_t>(plan->dim) * 16;
    if (tcp::ranges_overlap(input, bytes, output, bytes))
        return fail(TCP_ERR_INVALID_ARGUMENT, "apply_into input and output must not overlap");
    tcp::DeviceGuard dg;
    if ((st = dg.enter(plan->device))) return st;
    auto *mp = const_cast<tcp_csr_plan *>(plan);
    std::lock_guard<std::mutex> lock(mp->staging_mutex);
    if (mp->stage_in.bytes < bytes && (st = mp->stage_in.allocate(bytes))) return st;
    if (mp->stage_out.bytes < bytes && (st = mp->stage_out.allocate(bytes))) return st;
    TCP_HIP_CHECK(hipMemcpy(mp->stage_in.ptr, input, bytes, hipMemcpyHostToDevice));
    if ((st = csr_apply_device(plan, mp->stage_in.ptr, mp->stage_out.ptr, TCP_C128, nullptr))) return st;
    TCP_HIP_CHECK(hipMemcpy(output, mp->stage_out.ptr, bytes, hipMemcpyDeviceToHost));
    return tcp::ok();
}

int tcp_csr_plan_export(const tcp_csr_plan *plan, uint64_t *indptr, uint64_t *indices, double *values) {
    if (!plan) return fail(TCP_ERR_INVALID_ARGUMENT, "null plan");
    if (!indptr || (plan->nnz && (!indices || !values))) return fail(TCP_ERR_INVALID_ARGUMENT, "null output array");
    tcp::DeviceGuard dg;
    int st = dg.enter(plan->device);
    if (st) return st;
    TCP_HIP_CHECK(hipMemcpy(indptr, plan->indptr.ptr, 8 * (plan->dim + 1), hipMemcpyDeviceToHost));
    if (plan->nnz) {
        std::vector<uint32_t> c(plan->nnz);
        TCP_HIP_CHECK(hipMemcpy(c.data(), plan->cols.ptr, 4 * plan->nnz, hipMemcpyDeviceToHost));
        for (uint64_t k = 0; k < plan->nnz; ++k) indices[k] = c[k];
        TCP_HIP_CHECK(hipMemcpy(values, plan->vals64.ptr, 16 * plan->nnz, hipMemcpyDeviceToHost));
    }
    return tcp::ok();
}

void tcp_csr_plan_destroy(tcp_csr_plan *plan) { delete plan; }

}  // extern "C"
