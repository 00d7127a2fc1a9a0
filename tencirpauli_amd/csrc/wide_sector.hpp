// Wide U(1) sectors: more than 63 qubits (up to 1024), bitstrings held as
// packed 64-bit words with the operator packing of core/word.rs:53-110 (qubit
// q in word q / 64, bit q % 64).  Restricted order is the reference's public
// order (ascending integer value with qubit 0 as the most significant bit,
// U1Sector::rank / rank_words, core/sector.rs:127-238): the combinatorial
// number system over positions p = n - 1 - q, through the complement when
// k > n / 2.
#pragma once

#include <cstdint>
#include <vector>

#include "tcp_internal.hpp"

namespace tcp {

constexpr uint32_t kWideMaxQubits = 1024;

// Host geometry of one wide sector.
struct WideGeometry {
    uint32_t n = 0, k = 0;
    uint32_t kc = 0;    // min(k, n - k): particles of the ranked pattern
    bool comp = false;  // rank the complement (k > n / 2)
    uint32_t words = 0;
    uint64_t dim = 0;
    std::vector<uint64_t> choose;  // C(m, j), m <= n, j <= kc: choose[m * (kc + 1) + j]
};

int wide_geometry(uint32_t n, uint32_t k, WideGeometry &g);
// rank of packed words (TCP_ERR_INVALID_INDEX when the weight is not k or a
// padding bit is set) and its inverse.
int wide_rank(const WideGeometry &g, const uint64_t *words, uint64_t &rank);
int wide_unrank(const WideGeometry &g, uint64_t rank, uint64_t *words);

// Device side of a wide lazy plan.
struct WideSector {
    WideGeometry geo;
    uint32_t stride = 0;  // words per stored mask (the kernels' template width)
    uint32_t G = 0, T = 0;
    uint32_t diag_group = UINT32_MAX;
    double energy_scale = 1.0;
    bool leaks = false;
    uint64_t leak_source = 0;
    uint32_t leak_weight = 0;
    DeviceBuffer gx, goff, tz, tw64, tw32, choose;
    uint64_t retained_bytes = 0;
};

// Builds the tables and runs the leakage scan (the reference raises on the
// first leaking aggregated transition, core/sector.rs:1051-1127).
// require_conserved: lazy restriction's exact-conservation precondition.
int wide_build(const Operator &op, uint32_t k, uint64_t max_bytes, bool require_conserved, WideSector &out);
int wide_apply(const WideSector &w, const void *in, void *out, int dtype, hipStream_t stream);
int wide_diagonal(const WideSector &w, double *diag, hipStream_t stream);
// Eager restriction: per-destination counts, then rows with ascending columns.
int wide_csr_counts(const WideSector &w, uint32_t *counts, hipStream_t stream);
int wide_csr_fill(const WideSector &w, const uint64_t *indptr, uint32_t *cols, void *vals64, void *vals32,
                  hipStream_t stream);

}  // namespace tcp
