// Host side of the operator: canonicalisation, packing and charge analysis,
// plus the library-wide error state.  No device code.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>

#include "tcp_internal.hpp"

struct tcp_operator {
    tcp::Operator op;
};

namespace tcp {

namespace {
thread_local std::string g_last_error;
}

int fail(int status, const std::string &message) {
    g_last_error = message;
    return status;
}

int hip_fail(hipError_t error, const char *what) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "HIP error %d (%s) in %s", static_cast<int>(error),
                  hipGetErrorString(error), what);
    g_last_error = buf;
    return TCP_ERR_HIP;
}

int ok() {
    g_last_error.clear();
    return TCP_OK;
}

int memory_limit(uint64_t requested, uint64_t limit) {
    // PauliError::MemoryLimit display (core/error.rs:76-197)
    return fail(TCP_ERR_MEMORY_LIMIT, "requested " + std::to_string(requested) +
                                          " bytes exceeds memory limit " + std::to_string(limit));
}

int DeviceBuffer::allocate(size_t n) {
    release();
    if (n == 0) n = 1;
    TCP_HIP_CHECK(hipMalloc(&ptr, n));
    bytes = n;
    return TCP_OK;
}

int DeviceBuffer::upload(const void *host, size_t n) {
    int st = allocate(n);
    if (st) return st;
    if (n) TCP_HIP_CHECK(hipMemcpy(ptr, host, n, hipMemcpyHostToDevice));
    return TCP_OK;
}

static uint64_t bits_of(double x) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return u;
}

// PauliOperator::from_terms (core/operator.rs:117-155).  Words are compared
// by their code tuple from qubit 0 (core/word.rs:240-254); since the codes
// 0..3 already encode I<X<Y<Z, the byte rows compare lexicographically.
// Duplicate contributions are summed after sorting by (re bits, im bits) so
// the result does not depend on input order; only exact zeros are dropped.
int canonicalize(uint32_t nqubits, uint64_t nterms, const uint8_t *codes, const double *coef,
                 Operator &out) {
    if (nterms && (!coef || (nqubits && !codes)))
        return fail(TCP_ERR_INVALID_ARGUMENT, "null term arrays");
    for (uint64_t t = 0; t < nterms; ++t) {
        if (!std::isfinite(coef[2 * t]) || !std::isfinite(coef[2 * t + 1]))
            return fail(TCP_ERR_NON_FINITE_COEFFICIENT,
                        "coefficient at index " + std::to_string(t) + " is not finite");
        for (uint32_t q = 0; q < nqubits; ++q) {
            uint8_t c = codes[t * nqubits + q];
            if (c > 3)
                return fail(TCP_ERR_INVALID_CODE, "invalid Pauli code " + std::to_string(c) +
                                                      " at index " + std::to_string(q) +
                                                      "; expected 0..=3");
        }
    }
    std::vector<uint64_t> order(nterms);
    std::iota(order.begin(), order.end(), 0);
    auto row = [&](uint64_t t) { return codes + t * nqubits; };
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
        return nqubits && std::memcmp(row(a), row(b), nqubits) < 0;
    });
    out.nqubits = nqubits;
    out.codes.clear();
    out.coef.clear();
    std::vector<Complex> values;
    uint64_t start = 0;
    while (start < nterms) {
        uint64_t end = start + 1;
        while (end < nterms && (nqubits == 0 || std::memcmp(row(order[end]), row(order[start]), nqubits) == 0))
            ++end;
        values.clear();
        for (uint64_t j = start; j < end; ++j)
            values.push_back(Complex{coef[2 * order[j]], coef[2 * order[j] + 1]});
        std::sort(values.begin(), values.end(), [](const Complex &a, const Complex &b) {
            uint64_t ar = bits_of(a.re), br = bits_of(b.re);
            if (ar != br) return ar < br;
            return bits_of(a.im) < bits_of(b.im);
        });
        Complex sum{0.0, 0.0};
        for (const Complex &v : values) {
            sum.re = sum.re + v.re;
            sum.im = sum.im + v.im;
        }
        if (!(sum.re == 0.0 && sum.im == 0.0)) {
            out.codes.insert(out.codes.end(), row(order[start]), row(order[start]) + nqubits);
            out.coef.push_back(sum);
        }
        start = end;
    }
    return ok();
}

std::vector<IndexTerm> index_terms(const Operator &op) {
    const uint32_t n = op.nqubits;
    std::vector<IndexTerm> terms(op.term_count());
    for (uint64_t t = 0; t < op.term_count(); ++t) {
        uint64_t x = 0, z = 0;
        unsigned y = 0;
        const uint8_t *c = op.codes.data() + t * n;
        for (uint32_t q = 0; q < n; ++q) {
            uint64_t bit = 1ull << (n - 1 - q);
            if (c[q] == 1 || c[q] == 2) x |= bit;
            if (c[q] == 2 || c[q] == 3) z |= bit;
            if (c[q] == 2) ++y;
        }
        // i^(#Y mod 4) times the coefficient, as the textbook complex product
        const Complex a = op.coef[t];
        Complex w;
        switch (y % 4) {
            case 0: w = a; break;
            case 1: w = Complex{-a.im, a.re}; break;
            case 2: w = Complex{-a.re, -a.im}; break;
            default: w = Complex{a.im, -a.re}; break;
        }
        terms[t] = IndexTerm{x, z, w};
    }
    return terms;
}

// analyze_charge (core/operator.rs:396-445): the commutator with a diagonal
// qubit charge only has contributions from X/Y letters on qubits whose two
// levels differ; each swaps that letter (X<->Y) with coefficient
// c * i * (-d for X, +d for Y), d = level0 - level1.  The charge is conserved
// exactly when every aggregated commutator word cancels to an exact zero.
int analyze_charge(const Operator &op, const double *levels, bool &conserved, uint64_t &nonzero) {
    const uint32_t n = op.nqubits;
    std::vector<uint8_t> wcodes;
    std::vector<double> wcoef;
    for (uint64_t t = 0; t < op.term_count(); ++t) {
        const uint8_t *c = op.codes.data() + t * n;
        for (uint32_t q = 0; q < n; ++q) {
            if (c[q] != 1 && c[q] != 2) continue;
            double difference = levels[2 * q] - levels[2 * q + 1];
            if (difference == 0.0) continue;
            wcodes.insert(wcodes.end(), c, c + n);
            wcodes[wcodes.size() - n + q] = c[q] == 1 ? 2 : 1;
            double scale = c[q] == 1 ? -difference : difference;
            const Complex a = op.coef[t];
            // (a.re + i a.im) * (0 + i scale)
            wcoef.push_back(a.re * 0.0 - a.im * scale);
            wcoef.push_back(a.re * scale + a.im * 0.0);
        }
    }
    Operator commutator;
    int st = canonicalize(n, wcoef.size() / 2, wcodes.data(), wcoef.data(), commutator);
    if (st) return st;
    nonzero = commutator.term_count();
    conserved = nonzero == 0;
    return ok();
}

}  // namespace tcp

using tcp::fail;

extern "C" {

const char *tcp_version(void) { return "tencirpauli_amd 0.1.0 (gfx950)"; }

size_t tcp_last_error(char *buf, size_t len) {
    const std::string &m = tcp::g_last_error;
    if (buf && len) {
        size_t k = std::min(len - 1, m.size());
        std::memcpy(buf, m.data(), k);
        buf[k] = '\0';
    }
    return m.size();
}

int tcp_device_count(int *count) {
    if (!count) return fail(TCP_ERR_INVALID_ARGUMENT, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *count = e == hipSuccess ? c : 0;
    return tcp::ok();
}

int tcp_operator_from_codes(uint32_t nqubits, uint64_t nterms, const uint8_t *codes,
                            const double *coef, tcp_operator **out) {
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    auto *h = new tcp_operator();
    int st = tcp::canonicalize(nqubits, nterms, codes, coef, h->op);
    if (st) {
        delete h;
        return st;
    }
    *out = h;
    return tcp::ok();
}

int tcp_operator_from_words(uint32_t nqubits, uint64_t nterms, const uint64_t *x_words,
                            const uint64_t *z_words, const double *coef, tcp_operator **out) {
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    const uint32_t W = tcp::word_count(nqubits);
    if (nterms && W && (!x_words || !z_words)) return fail(TCP_ERR_INVALID_ARGUMENT, "null word arrays");
    std::vector<uint8_t> codes(nterms * nqubits);
    for (uint64_t t = 0; t < nterms; ++t) {
        for (uint32_t q = 0; q < nqubits; ++q) {
            // padding bits beyond nqubits are ignored (core/word.rs:63-91)
            uint64_t x = (x_words[t * W + q / 64] >> (q % 64)) & 1;
            uint64_t z = (z_words[t * W + q / 64] >> (q % 64)) & 1;
            codes[t * nqubits + q] = x ? (z ? 2 : 1) : (z ? 3 : 0);
        }
    }
    return tcp_operator_from_codes(nqubits, nterms, codes.data(), coef, out);
}

uint32_t tcp_operator_nqubits(const tcp_operator *op) { return op ? op->op.nqubits : 0; }

uint64_t tcp_operator_term_count(const tcp_operator *op) { return op ? op->op.term_count() : 0; }

uint64_t tcp_operator_distinct_x_mask_count(const tcp_operator *op) {
    if (!op) return 0;
    const uint32_t n = op->op.nqubits;
    std::vector<std::vector<uint8_t>> xs;
    xs.reserve(op->op.term_count());
    for (uint64_t t = 0; t < op->op.term_count(); ++t) {
        std::vector<uint8_t> x(n);
        for (uint32_t q = 0; q < n; ++q) {
            uint8_t c = op->op.codes[t * n + q];
            x[q] = c == 1 || c == 2;
        }
        xs.push_back(std::move(x));
    }
    std::sort(xs.begin(), xs.end());
    return static_cast<uint64_t>(std::unique(xs.begin(), xs.end()) - xs.begin());
}

int tcp_operator_export_codes(const tcp_operator *op, uint8_t *codes, double *coef) {
    if (!op) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    const auto &o = op->op;
    if (codes && !o.codes.empty()) std::memcpy(codes, o.codes.data(), o.codes.size());
    if (coef)
        for (uint64_t t = 0; t < o.term_count(); ++t) {
            coef[2 * t] = o.coef[t].re;
            coef[2 * t + 1] = o.coef[t].im;
        }
    return tcp::ok();
}

int tcp_operator_export_words(const tcp_operator *op, uint64_t *x_words, uint64_t *z_words,
                              double *coef) {
    if (!op) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    const auto &o = op->op;
    const uint32_t n = o.nqubits, W = tcp::word_count(n);
    for (uint64_t t = 0; t < o.term_count(); ++t) {
        for (uint32_t w = 0; w < W; ++w) {
            if (x_words) x_words[t * W + w] = 0;
            if (z_words) z_words[t * W + w] = 0;
        }
        for (uint32_t q = 0; q < n; ++q) {
            uint8_t c = o.codes[t * n + q];
            if (x_words && (c == 1 || c == 2)) x_words[t * W + q / 64] |= 1ull << (q % 64);
            if (z_words && (c == 2 || c == 3)) z_words[t * W + q / 64] |= 1ull << (q % 64);
        }
        if (coef) {
            coef[2 * t] = o.coef[t].re;
            coef[2 * t + 1] = o.coef[t].im;
        }
    }
    return tcp::ok();
}

int tcp_operator_analyze_charge(const tcp_operator *op, const double *qubit_levels,
                                int *is_conserved, uint64_t *nonzero) {
    if (!op || (!qubit_levels && op->op.nqubits))
        return fail(TCP_ERR_INVALID_ARGUMENT, "null operator or levels");
    bool conserved = false;
    uint64_t nz = 0;
    int st = tcp::analyze_charge(op->op, qubit_levels, conserved, nz);
    if (st) return st;
    if (is_conserved) *is_conserved = conserved ? 1 : 0;
    if (nonzero) *nonzero = nz;
    return tcp::ok();
}

// ---- operator algebra (core/operator.rs:217-393) --------------------------

namespace {

int compatible(const tcp_operator *a, const tcp_operator *b) {
    if (!a || !b) return tcp::fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    if (a->op.nqubits != b->op.nqubits)
        return tcp::fail(TCP_ERR_INCOMPATIBLE_QUBIT_COUNTS, "incompatible qubit counts: " +
                                                                std::to_string(a->op.nqubits) + " and " +
                                                                std::to_string(b->op.nqubits));
    return TCP_OK;
}

tcp::Complex cmul(tcp::Complex a, tcp::Complex b) {
    return tcp::Complex{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

int wrap(const std::vector<uint8_t> &codes, const std::vector<double> &coef, uint32_t n, tcp_operator **out) {
    auto *h = new tcp_operator();
    const int st = tcp::canonicalize(n, coef.size() / 2, codes.data(), coef.data(), h->op);
    if (st) {
        delete h;
        return st;
    }
    *out = h;
    return TCP_OK;
}

}  // namespace

int tcp_operator_add(const tcp_operator *a, const tcp_operator *b, tcp_operator **out) {
    if (!out) return tcp::fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    int st = compatible(a, b);
    if (st) return st;
    // equal words add their two coefficients (an IEEE sum is commutative, so
    // the canonical bit-pattern order gives the reference's left + right)
    std::vector<uint8_t> codes(a->op.codes);
    codes.insert(codes.end(), b->op.codes.begin(), b->op.codes.end());
    std::vector<double> coef;
    for (const auto *o : {&a->op, &b->op})
        for (const tcp::Complex &c : o->coef) {
            coef.push_back(c.re);
            coef.push_back(c.im);
        }
    st = wrap(codes, coef, a->op.nqubits, out);
    return st ? st : tcp::ok();
}

int tcp_operator_scale(const tcp_operator *a, double re, double im, tcp_operator **out) {
    if (!a || !out) return tcp::fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (!std::isfinite(re) || !std::isfinite(im))
        return tcp::fail(TCP_ERR_NON_FINITE_COEFFICIENT, "coefficient at index 0 is not finite");
    auto *h = new tcp_operator();
    h->op.nqubits = a->op.nqubits;
    if (re != 0.0 || im != 0.0) {
        const uint32_t n = a->op.nqubits;
        for (uint64_t t = 0; t < a->op.term_count(); ++t) {
            const tcp::Complex c = cmul(a->op.coef[t], tcp::Complex{re, im});
            if (c.re == 0.0 && c.im == 0.0) continue;
            h->op.codes.insert(h->op.codes.end(), a->op.codes.begin() + t * n, a->op.codes.begin() + (t + 1) * n);
            h->op.coef.push_back(c);
        }
    }
    *out = h;
    return tcp::ok();
}

int tcp_operator_multiply(const tcp_operator *a, const tcp_operator *b, uint64_t max_bytes, tcp_operator **out) {
    if (!out) return tcp::fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    int st = compatible(a, b);
    if (st) return st;
    const uint32_t n = a->op.nqubits;
    const uint64_t Ta = a->op.term_count(), Tb = b->op.term_count();
    const unsigned __int128 pairs = static_cast<unsigned __int128>(Ta) * Tb;
    const unsigned __int128 bytes = pairs * (static_cast<unsigned __int128>(n) + 16);
    if (bytes > static_cast<unsigned __int128>(UINT64_MAX))
        return tcp::fail(TCP_ERR_OVERFLOW, "integer overflow while estimating Pauli operator product");
    if (tcp::exceeds_budget(static_cast<uint64_t>(bytes), max_bytes))
        return tcp::memory_limit(static_cast<uint64_t>(bytes), max_bytes);
    // product code and phase exponent (powers of i) of two single-qubit Paulis:
    // XY = iZ, YZ = iX, ZX = iY, the reversed orders -i
    static const uint8_t code[4][4] = {{0, 1, 2, 3}, {1, 0, 3, 2}, {2, 3, 0, 1}, {3, 2, 1, 0}};
    static const uint8_t power[4][4] = {{0, 0, 0, 0}, {0, 0, 1, 3}, {0, 3, 0, 1}, {0, 1, 3, 0}};
    static const tcp::Complex phase[4] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}, {0.0, -1.0}};
    std::vector<uint8_t> codes(static_cast<size_t>(pairs) * n);
    std::vector<double> coef(2 * static_cast<size_t>(pairs));
    size_t k = 0;
    for (uint64_t i = 0; i < Ta; ++i)
        for (uint64_t j = 0; j < Tb; ++j, ++k) {
            const uint8_t *x = a->op.codes.data() + i * n, *y = b->op.codes.data() + j * n;
            uint32_t p = 0;
            for (uint32_t q = 0; q < n; ++q) {
                codes[k * n + q] = code[x[q]][y[q]];
                p += power[x[q]][y[q]];
            }
            const tcp::Complex c = cmul(cmul(a->op.coef[i], b->op.coef[j]), phase[p & 3]);
            coef[2 * k] = c.re;
            coef[2 * k + 1] = c.im;
        }
    st = wrap(codes, coef, n, out);
    return st ? st : tcp::ok();
}

int tcp_operator_adjoint(const tcp_operator *a, tcp_operator **out) {
    if (!a || !out) return tcp::fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    auto *h = new tcp_operator();
    h->op = a->op;
    for (tcp::Complex &c : h->op.coef) c.im = -c.im;
    *out = h;
    return tcp::ok();
}

void tcp_operator_destroy(tcp_operator *op) { delete op; }

}  // extern "C"

// Internal accessor for the plan builders.
namespace tcp {
const Operator &operator_of(const tcp_operator *h) { return h->op; }
}  // namespace tcp
