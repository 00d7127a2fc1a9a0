// Internal declarations shared by the host plan builders and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tencirpauli_amd.h"

namespace tcp {

// Records `message` as the calling thread's last error and returns `status`.
int fail(int status, const std::string &message);
// Records a HIP runtime failure (status TCP_ERR_HIP).
int hip_fail(hipError_t error, const char *what);
// Clears the calling thread's last error and returns TCP_OK.
int ok();

#define TCP_HIP_CHECK(call)                                    \
    do {                                                       \
        hipError_t tcp_hip_status_ = (call);                   \
        if (tcp_hip_status_ != hipSuccess)                     \
            return ::tcp::hip_fail(tcp_hip_status_, #call);    \
    } while (0)

constexpr uint64_t kMinEnforcedMaxBytes = 1ull << 30;  // py/_native_mvp.py:28

// max_bytes is advisory below 1 GiB (py/_native_mvp.py:19-58).
inline bool exceeds_budget(uint64_t requested, uint64_t max_bytes) {
    return max_bytes >= kMinEnforcedMaxBytes && requested > max_bytes;
}
int memory_limit(uint64_t requested, uint64_t limit);

inline uint32_t word_count(uint32_t n) { return n / 64 + (n % 64 != 0); }

struct Complex {
    double re, im;
};

// A canonical Pauli operator (core/operator.rs:117-155): codes are
// term_count x nqubits, sorted by code tuple, duplicates aggregated.
struct Operator {
    uint32_t nqubits = 0;
    std::vector<uint8_t> codes;
    std::vector<Complex> coef;
    uint64_t term_count() const { return coef.size(); }
};

// The canonical operator behind a public handle.
const Operator &operator_of(const tcp_operator *handle);

// Restores the caller's current device when it goes out of scope.
struct DeviceGuard {
    int previous = -1;
    bool switched = false;
    int enter(int device) {
        if (hipGetDevice(&previous) != hipSuccess) previous = -1;
        if (previous != device) {
            hipError_t e = hipSetDevice(device);
            if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
            switched = true;
        }
        return TCP_OK;
    }
    ~DeviceGuard() {
        if (switched && previous >= 0) (void)hipSetDevice(previous);
    }
};

// Canonicalises raw (codes, coef) into `out`; validates codes and finiteness.
int canonicalize(uint32_t nqubits, uint64_t nterms, const uint8_t *codes, const double *coef,
                 Operator &out);

// Index-space masks of one canonical term (qubit q -> bit n-1-q) and the
// weighted phase w = c * i^(#Y mod 4) (core/pauli_action.rs:24-88).
struct IndexTerm {
    uint64_t x, z;
    Complex w;
};
std::vector<IndexTerm> index_terms(const Operator &op);

// Index-space form of a canonical fermion ladder word (mode m -> bit n-1-m):
// on a source s with every bit of c1 set and every bit of c0 clear, the word
// maps s to s ^ x with value w (-1)^popcount(z & s); elsewhere it gives zero.
// z is the Jordan-Wigner parity mask of the word and w carries its constant
// sign.  Pauli terms are the special case c1 = c0 = 0.
struct CondTerm {
    uint64_t x, z, c1, c0;
    Complex w;
};

// Exact conservation analysis of a diagonal qubit charge (core/operator.rs:396-445).
int analyze_charge(const Operator &op, const double *levels, bool &conserved, uint64_t &nonzero);

// Operator access for solvers built on the plans (device buffers, async on `stream`).
uint64_t plan_dimension(const tcp_mvp_plan *plan);
uint64_t plan_dimension(const tcp_sector_plan *plan);
int plan_device(const tcp_mvp_plan *plan);
int plan_device(const tcp_sector_plan *plan);
int plan_apply(const tcp_mvp_plan *plan, const void *in, void *out, int dtype, hipStream_t stream);
int plan_apply(const tcp_sector_plan *plan, const void *in, void *out, int dtype, hipStream_t stream);
uint64_t plan_dimension(const tcp_csr_plan *plan);
int plan_device(const tcp_csr_plan *plan);
int plan_apply(const tcp_csr_plan *plan, const void *in, void *out, int dtype, hipStream_t stream);
// Real part of the diagonal matrix elements H_ii (FP64, `dimension` entries).
int plan_diagonal(const tcp_mvp_plan *plan, double *diag, hipStream_t stream);
int plan_diagonal(const tcp_sector_plan *plan, double *diag, hipStream_t stream);
int plan_diagonal(const tcp_csr_plan *plan, double *diag, hipStream_t stream);

// Learned-preconditioner inputs of a sector plan: 64 bf16 features per state
// (slot 0 left for (H_ii - theta) / scale) and the diagonal.
double plan_energy_scale(const tcp_sector_plan *plan);
int plan_features(const tcp_sector_plan *plan, double scale, uint16_t *features, double *diag, hipStream_t stream);

// RAII device buffer.
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    int upload(const void *host, size_t n);
    int allocate(size_t n);
};

inline bool ranges_overlap(const void *a, size_t abytes, const void *b, size_t bbytes) {
    auto pa = reinterpret_cast<uintptr_t>(a);
    auto pb = reinterpret_cast<uintptr_t>(b);
    return pa < pb + bbytes && pb < pa + abytes;
}

}  // namespace tcp

// A canonical fermion operator: word t is creation[t] (increasing modes)
// followed by annihilation[t] (decreasing modes), with coefficient coef[t].
struct tcp_fermion {
    uint32_t n_modes = 0;
    std::vector<std::vector<uint32_t>> creation, annihilation;
    std::vector<tcp::Complex> coef;
};

namespace tcp {
// The words of a fermion operator on at most 64 modes as CondTerms, in
// canonical order.
std::vector<CondTerm> fermion_index_terms(const tcp_fermion &f);
}  // namespace tcp
