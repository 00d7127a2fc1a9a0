// General additive-charge sectors on binary axes (qubits or fermion modes):
// sectors that are not a product of contiguous fixed-population blocks, such
// as total number and 2 S_z on interleaved spin orbitals, or charges with
// several levels per axis.
//
// The public order is the reference's: lexicographic over axes with axis 0
// most significant (operator_space_axis0_msb_mixed_radix), axis a being index
// bit n-1-a; for binary axes this is ascending integer order.  Rank and
// unrank follow the suffix-count plan of ChargeSectorPlan
// (core/charge_sector.rs:8-16, 1426-1580, build_charge_sector_plan :1620):
// count[p][R] is the number of assignments of axes p..n-1 whose contributions
// sum to the remainder vector R, and
//     rank(b) = sum over set bits a (axis order) of count[a+1][R_a],
// R_a being the target minus the contributions of axes 0..a-1.  Levels are
// shifted so every axis contributes 0 when its bit is clear (the constant
// part moves into the target); a set bit of axis a then subtracts w_a.
// Remainders are stored densely over the box, per constraint, of the values
// a prefix remainder of an in-sector state can take: within [target - sum of
// positive weights, target - sum of negative weights] (what the prefix can
// have used) and within [sum of negative weights, sum of positive weights]
// (what a suffix can still supply), encoded mixed radix, so the update
// R -= w_a is one subtraction of the encoded weight.  A state whose prefix
// leaves the box is outside the sector: ranks of arbitrary bitstrings check
// each step; the kernels first test that a flip keeps every charge, after
// which the walk stays in the box.
#pragma once

#include <cstdint>
#include <vector>

#include "tcp_internal.hpp"

namespace tcp {

constexpr uint32_t kChargeMaxConstraints = 4;
constexpr uint32_t kChargeMaxAxes = 63;
constexpr uint64_t kChargeMaxTableEntries = 1ull << 24;  // (n + 1) x box counts, 64 MiB of u32

// Host geometry and rank tables of one general sector.
struct ChargeGeometry {
    uint32_t n = 0, m = 0;
    std::vector<int64_t> weight;  // [n * m]: level_1 - level_0 of axis a, constraint j
    int64_t target[kChargeMaxConstraints] = {};  // after removing every axis's level_0
    int64_t lo[kChargeMaxConstraints] = {};      // remainder box [lo, hi]
    int64_t hi[kChargeMaxConstraints] = {};
    uint64_t stride[kChargeMaxConstraints] = {};
    uint64_t box = 0;                            // encoded remainders per axis position
    std::vector<uint32_t> count;                 // [(n + 1) * box] completions
    std::vector<int64_t> enc;                    // [n] encoded weight of each axis
    int64_t enc_target = 0, enc_zero = 0;
    uint64_t dim = 0;
};

// levels: [m][n][2] integer levels (offset already folded into targets).
int charge_geometry(uint32_t n, uint32_t m, const int64_t *levels, const int64_t *targets, uint64_t max_bytes,
                    ChargeGeometry &g);
int charge_rank(const ChargeGeometry &g, uint64_t bits, uint64_t &rank);
int charge_unrank(const ChargeGeometry &g, uint64_t rank, uint64_t &bits);

// Device side of a restricted plan on a general sector: x-mask groups in
// first-appearance order over the canonical terms (Pauli terms, or fermion
// ladder words with occupancy conditions), the rank tables, and the leakage
// found at build time.
struct ChargeSector {
    ChargeGeometry geo;
    uint32_t G = 0, T = 0;
    uint32_t diag_group = UINT32_MAX;
    double energy_scale = 1.0;
    bool cond = false;
    bool leaks = false;
    uint64_t leak_source = 0;
    DeviceBuffer gx, goff, tz, tw64, tw32, tc1, tc0, count, enc, wt;
    // the encoded weights as int32 when every encoded value fits (the gather
    // then walks in 32-bit arithmetic with the weights in LDS)
    bool narrow = false;
    DeviceBuffer enc32;
    uint64_t retained_bytes = 0;
};

// require_conserved: lazy restriction's precondition (every charge exactly
// conserved by the operator, py/charge.py:1128-1132); eager restriction skips
// it and relies on the leakage scan, as the reference does.
int charge_build(const Operator *op, const std::vector<CondTerm> *fermion_terms, const ChargeGeometry &geo,
                 uint64_t max_bytes, bool require_conserved, ChargeSector &out);
int charge_apply(const ChargeSector &c, const void *in, void *out, int dtype, hipStream_t stream);
int charge_diagonal(const ChargeSector &c, double *diag, hipStream_t stream);
int charge_rank_device(const ChargeSector &c, const uint64_t *bits, uint64_t *ranks, uint64_t count,
                       hipStream_t stream);
int charge_unrank_device(const ChargeSector &c, const uint64_t *ranks, uint64_t *bits, uint64_t count,
                         hipStream_t stream);
int charge_walk_device(const ChargeSector &c, uint64_t first, uint64_t *bits, uint64_t count, hipStream_t stream);
// Eager restriction: per-destination counts, then rows with ascending columns.
int charge_csr_counts(const ChargeSector &c, uint32_t *counts, hipStream_t stream);
int charge_csr_fill(const ChargeSector &c, const uint64_t *indptr, uint32_t *cols, void *vals64, void *vals32,
                    hipStream_t stream);

}  // namespace tcp
