// The full-space matrix of a Pauli operator as device CSR: the eager
// materialisation behind PauliOperator::coo_matrix / csr_matrix
// (core/hamiltonian.rs:360-460, 573-582, 725-747).  Row r holds one entry per
// distinct X mask x, at column r ^ x, with value
// sum_t w_t (-1)^popcount(z_t & column) over the group's terms in canonical
// order; exact zeros are dropped and columns ascend.
#pragma once

#include <cstdint>

#include "tcp_internal.hpp"

namespace tcp {

// Columns are stored as 32-bit indices, as in the restricted CSR plans.
constexpr uint32_t kFullCsrMaxQubits = 32;

// Groups by X mask (ascending, the BTreeMap order of core/hamiltonian.rs:573-582)
// and the binary trie over the sorted masks that yields each row's columns in
// ascending order without a sort.
struct FullGroups {
    uint32_t n = 0;
    uint64_t dim = 0;
    uint32_t G = 0;     // distinct X masks
    uint32_t root = 0;  // trie root (a leaf code when G == 1)
    DeviceBuffer gx;    // [G] x masks, ascending
    DeviceBuffer goff;  // [G + 1] term offsets
    DeviceBuffer tz;    // [T] z masks, canonical order within a group
    DeviceBuffer tw;    // [T] (re, im) of w = c i^(#Y mod 4)
    DeviceBuffer nodes; // [G - 1] internal trie nodes {bit, left, right, 0}
};

int full_groups_build(const Operator &op, FullGroups &out);
int full_csr_counts(const FullGroups &f, uint32_t *counts, hipStream_t stream);
int full_csr_fill(const FullGroups &f, const uint64_t *indptr, uint32_t *cols, void *vals64, void *vals32,
                  hipStream_t stream);

}  // namespace tcp
