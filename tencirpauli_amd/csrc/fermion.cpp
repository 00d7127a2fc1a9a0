// Fermion operators on the occupation basis: CAR canonicalisation of raw
// ladder products and the Jordan-Wigner map onto the Pauli path.
//
// canonicalize_fermion_terms (core/structured/integrals.rs:2-54): each raw
// product is rewritten with the anticommutation relations into canonical
// words (creations with increasing modes, then annihilations with decreasing
// modes; fermion_rewrite, core/structured/boson.rs:177-239) with integer
// weights; words are aggregated over all terms, sorted lexicographically by
// (creation modes, annihilation modes), every word's contributions summed in
// IEEE-bit-pattern order (deterministic_sum, core/structured/common.rs:225-230)
// and exact zeros dropped.
//
// jordan_wigner_terms (core/structured/fermion.rs:141-183, 261-283): mode m is
// qubit m; a_m = Z_0 ... Z_{m-1} (X_m + i Y_m) / 2 and
// a+_m = Z_0 ... Z_{m-1} (X_m - i Y_m) / 2, multiplied out factor by factor in
// the canonical word's order, then aggregated over terms like the words above.
// The Pauli image serves the full-space targets; restricted plans use the
// words directly (fermion_index_terms below, the sector gather's occupancy-
// conditioned terms).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "tcp_internal.hpp"

namespace {

using tcp::Complex;
using tcp::fail;

using Word = std::pair<std::vector<uint32_t>, std::vector<uint32_t>>;  // (creations, annihilations)
using Factor = std::pair<uint32_t, bool>;                              // (mode, is_creation)
using IntegerMap = std::map<Word, __int128>;

constexpr uint64_t kBytesPerItem = 192;  // check_structured_bytes_u128 (core/structured/common.rs:142-156)

Complex cmul(Complex a, Complex b) { return Complex{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
Complex cadd(Complex a, Complex b) { return Complex{a.re + b.re, a.im + b.im}; }
bool exact_zero(Complex a) { return a.re == 0.0 && a.im == 0.0; }

int check_items(uint64_t count, uint64_t max_bytes, const char *context) {
    if (count > UINT64_MAX / kBytesPerItem) return fail(TCP_ERR_OVERFLOW, std::string("integer overflow while ") + context);
    const uint64_t requested = count * kBytesPerItem;
    if (requested > max_bytes) return tcp::memory_limit(requested, max_bytes);
    return TCP_OK;
}

void add_scaled(IntegerMap &target, const IntegerMap &source, __int128 factor) {
    for (const auto &kv : source) target[kv.first] += kv.second * factor;
}

// A product in which no annihilation precedes a creation of the same mode
// never contracts: it is zero when a mode repeats among its creations or its
// annihilations, and otherwise the canonical word with the sign of the
// permutation that sorts it (the one-key result of
// canonical_without_contractions, core/structured/boson.rs:241-300).
bool rewrite_without_contractions(const std::vector<Factor> &seq, IntegerMap &out) {
    for (size_t i = 0; i < seq.size(); ++i)
        if (!seq[i].second)
            for (size_t j = i + 1; j < seq.size(); ++j)
                if (seq[j].second && seq[j].first == seq[i].first) return false;
    Word w;
    for (const auto &[mode, creation] : seq) (creation ? w.first : w.second).push_back(mode);
    // canonical position key: creations by increasing mode, then annihilations by decreasing mode
    auto key = [](const Factor &f) { return f.second ? std::make_pair(0, static_cast<int64_t>(f.first))
                                                     : std::make_pair(1, -static_cast<int64_t>(f.first)); };
    uint64_t inversions = 0;
    for (size_t i = 0; i < seq.size(); ++i)
        for (size_t j = i + 1; j < seq.size(); ++j) {
            if (key(seq[i]) == key(seq[j])) {  // a repeated mode of one kind: nilpotent
                out.clear();
                return true;
            }
            inversions += key(seq[j]) < key(seq[i]);
        }
    std::sort(w.first.begin(), w.first.end());
    std::sort(w.second.begin(), w.second.end(), std::greater<uint32_t>());
    out.clear();
    out.emplace(std::move(w), (inversions & 1) ? -1 : 1);
    return true;
}

// Normal-orders one product: the first adjacent pair that is nilpotent makes
// the word vanish; the first adjacent inversion is swapped (weight -1) and,
// for a_m a+_m, also contracted (weight +1); a word with neither is canonical.
int rewrite(const std::vector<Factor> &seq, uint64_t max_bytes, IntegerMap &out) {
    if (rewrite_without_contractions(seq, out)) return TCP_OK;
    out.clear();
    for (size_t i = 0; i + 1 < seq.size(); ++i) {
        const auto [lm, lc] = seq[i];
        const auto [rm, rc] = seq[i + 1];
        if (lc == rc && lm == rm) return TCP_OK;  // a a = a+ a+ = 0
        const bool inversion = (lc && rc) ? lm > rm : (!lc && !rc) ? lm < rm : (!lc && rc);
        if (!inversion) continue;
        std::vector<Factor> swapped = seq;
        std::swap(swapped[i], swapped[i + 1]);
        IntegerMap sub;
        int st = rewrite(swapped, max_bytes, sub);
        if (st) return st;
        add_scaled(out, sub, -1);
        if (!lc && rc && lm == rm) {
            std::vector<Factor> contracted;
            contracted.reserve(seq.size() - 2);
            for (size_t j = 0; j < seq.size(); ++j)
                if (j != i && j != i + 1) contracted.push_back(seq[j]);
            if ((st = rewrite(contracted, max_bytes, sub))) return st;
            add_scaled(out, sub, 1);
        }
        for (auto it = out.begin(); it != out.end();) it = it->second == 0 ? out.erase(it) : std::next(it);
        return check_items(out.size(), max_bytes, "estimating fermion expansion");
    }
    Word w;
    for (const auto &[mode, creation] : seq) (creation ? w.first : w.second).push_back(mode);
    out.emplace(std::move(w), 1);
    return TCP_OK;
}

// Sum in ascending order of the (real, imaginary) IEEE bit patterns.
Complex deterministic_sum(std::vector<Complex> values) {
    auto bits = [](double x) {
        uint64_t b;
        std::memcpy(&b, &x, 8);
        return b;
    };
    std::sort(values.begin(), values.end(), [&](const Complex &a, const Complex &b) {
        const uint64_t ar = bits(a.re), br = bits(b.re);
        return ar != br ? ar < br : bits(a.im) < bits(b.im);
    });
    Complex s{0.0, 0.0};
    for (const Complex &v : values) s = cadd(s, v);
    return s;
}

// Product of two Pauli codes (0 I, 1 X, 2 Y, 3 Z) with its phase.
std::pair<uint8_t, Complex> pauli_product(uint8_t a, uint8_t b) {
    if (a == 0) return {b, Complex{1.0, 0.0}};
    if (b == 0 || a == b) return {static_cast<uint8_t>(a == b ? 0 : a), Complex{1.0, 0.0}};
    // XY = iZ, YZ = iX, ZX = iY and the reversed products with -i
    const uint8_t c = static_cast<uint8_t>(6 - a - b);
    const bool cyclic = (a == 1 && b == 2) || (a == 2 && b == 3) || (a == 3 && b == 1);
    return {c, Complex{0.0, cyclic ? 1.0 : -1.0}};
}

// Jordan-Wigner words of one canonical product (coefficient 1).
std::vector<std::pair<std::vector<uint8_t>, Complex>> jw_expand(uint32_t n, const std::vector<Factor> &seq) {
    std::vector<std::pair<std::vector<uint8_t>, Complex>> current{{std::vector<uint8_t>(n, 0), Complex{1.0, 0.0}}};
    for (const auto &[mode, creation] : seq) {
        const Complex ycoef{0.0, creation ? -0.5 : 0.5};
        std::map<std::vector<uint8_t>, Complex> next;
        std::vector<uint8_t> right(n, 0);
        for (uint32_t q = 0; q < mode; ++q) right[q] = 3;
        for (const auto &[left, lcoef] : current) {
            for (int k = 0; k < 2; ++k) {
                right[mode] = k == 0 ? 1 : 2;
                const Complex rcoef = k == 0 ? Complex{0.5, 0.0} : ycoef;
                std::vector<uint8_t> word(n);
                Complex phase{1.0, 0.0};
                for (uint32_t q = 0; q < n; ++q) {
                    const auto [c, ph] = pauli_product(left[q], right[q]);
                    word[q] = c;
                    phase = cmul(phase, ph);
                }
                auto it = next.try_emplace(std::move(word), Complex{0.0, 0.0}).first;
                it->second = cadd(it->second, cmul(cmul(lcoef, rcoef), phase));
            }
        }
        current.clear();
        for (auto &kv : next)
            if (!exact_zero(kv.second)) current.emplace_back(kv.first, kv.second);
    }
    return current;
}

std::vector<Factor> sequence_of(const tcp_fermion &f, uint64_t t) {
    std::vector<Factor> seq;
    for (uint32_t m : f.creation[t]) seq.emplace_back(m, true);
    for (uint32_t m : f.annihilation[t]) seq.emplace_back(m, false);
    return seq;
}

}  // namespace

namespace tcp {

// Applying the word to |s> runs its factors right to left; the factor on mode
// m contributes (-1)^popcount(state & below(m)) with state = s ^ F, F the bits
// flipped so far, i.e. popcount(s & below(m)) + popcount(F & below(m)).  The
// first parts add up to the parity mask z, the second to the constant sign.
std::vector<CondTerm> fermion_index_terms(const tcp_fermion &f) {
    const uint32_t n = f.n_modes;
    auto bit = [n](uint32_t m) { return 1ull << (n - 1 - m); };
    auto below = [n](uint32_t m) {  // modes l < m: index bits n-m .. n-1
        const uint64_t all = n >= 64 ? ~0ull : (1ull << n) - 1;
        const uint64_t low = (n - m) >= 64 ? ~0ull : (1ull << (n - m)) - 1;
        return all & ~low;
    };
    std::vector<CondTerm> out(f.coef.size());
    for (size_t t = 0; t < f.coef.size(); ++t) {
        uint64_t flipped = 0, parity = 0, amask = 0, cmask = 0;
        uint32_t sign = 0;
        std::vector<uint32_t> order;  // application order: annihilations last-to-first, then creations
        for (size_t i = f.annihilation[t].size(); i-- > 0;) order.push_back(f.annihilation[t][i]);
        const size_t nann = order.size();
        for (size_t i = f.creation[t].size(); i-- > 0;) order.push_back(f.creation[t][i]);
        for (size_t i = 0; i < order.size(); ++i) {
            const uint32_t m = order[i];
            parity ^= below(m);
            sign ^= static_cast<uint32_t>(__builtin_popcountll(flipped & below(m))) & 1u;
            flipped ^= bit(m);
            (i < nann ? amask : cmask) |= bit(m);
        }
        const Complex c = f.coef[t];
        out[t] = CondTerm{flipped, parity, amask, cmask & ~amask, sign ? Complex{-c.re, -c.im} : c};
    }
    return out;
}

}  // namespace tcp

extern "C" {

int tcp_fermion_from_ladder(uint32_t n_modes, uint64_t nterms, const uint64_t *offsets, const uint32_t *modes,
                            const uint8_t *actions, const double *coef, uint64_t max_bytes, tcp_fermion **out) {
    if (!out) return fail(TCP_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (nterms && (!offsets || !coef)) return fail(TCP_ERR_INVALID_ARGUMENT, "null term arrays");
    if (nterms && offsets[nterms] && (!modes || !actions)) return fail(TCP_ERR_INVALID_ARGUMENT, "null factor arrays");
    std::map<Word, std::vector<Complex>> aggregate;
    uint64_t total = 0;
    IntegerMap words;
    for (uint64_t t = 0; t < nterms; ++t) {
        const Complex c{coef[2 * t], coef[2 * t + 1]};
        if (!std::isfinite(c.re) || !std::isfinite(c.im))
            return fail(TCP_ERR_NON_FINITE_COEFFICIENT, "coefficient at index " + std::to_string(t) + " is not finite");
        if (offsets[t + 1] < offsets[t]) return fail(TCP_ERR_INVALID_ARGUMENT, "factor offsets must not decrease");
        int st = check_items(offsets[t + 1] - offsets[t], max_bytes, "estimating fermion raw expansion");
        if (st) return st;
        std::vector<Factor> seq;
        for (uint64_t k = offsets[t]; k < offsets[t + 1]; ++k) {
            if (modes[k] >= n_modes) return fail(TCP_ERR_INVALID_INDEX, "invalid sector index: fermion mode");
            if (actions[k] > 1)
                return fail(TCP_ERR_INVALID_CODE, "invalid ladder action " + std::to_string(actions[k]) +
                                                      " in term " + std::to_string(t) + "; expected 0 or 1");
            seq.emplace_back(modes[k], actions[k] == 0);
        }
        if ((st = rewrite(seq, max_bytes, words))) return st;
        for (const auto &[word, weight] : words) {
            const double w = static_cast<double>(weight);
            aggregate[word].push_back(Complex{c.re * w, c.im * w});
            if ((st = check_items(++total, max_bytes, "estimating structured aggregation"))) return st;
        }
    }
    auto f = new tcp_fermion();
    f->n_modes = n_modes;
    for (auto &[word, values] : aggregate) {
        const Complex v = deterministic_sum(values);
        if (exact_zero(v)) continue;
        f->creation.push_back(word.first);
        f->annihilation.push_back(word.second);
        f->coef.push_back(v);
    }
    *out = f;
    return tcp::ok();
}

uint32_t tcp_fermion_n_modes(const tcp_fermion *f) { return f ? f->n_modes : 0; }
uint64_t tcp_fermion_term_count(const tcp_fermion *f) { return f ? f->coef.size() : 0; }

uint64_t tcp_fermion_factor_count(const tcp_fermion *f) {
    if (!f) return 0;
    uint64_t n = 0;
    for (size_t t = 0; t < f->coef.size(); ++t) n += f->creation[t].size() + f->annihilation[t].size();
    return n;
}

int tcp_fermion_export(const tcp_fermion *f, uint64_t *offsets, uint32_t *modes, uint8_t *actions, double *coef) {
    if (!f) return fail(TCP_ERR_INVALID_ARGUMENT, "null operator");
    uint64_t k = 0;
    if (offsets) offsets[0] = 0;
    for (size_t t = 0; t < f->coef.size(); ++t) {
        for (const Factor &fac : sequence_of(*f, t)) {
            if (modes) modes[k] = fac.first;
            if (actions) actions[k] = fac.second ? 0 : 1;
            ++k;
        }
        if (offsets) offsets[t + 1] = k;
        if (coef) {
            coef[2 * t] = f->coef[t].re;
            coef[2 * t + 1] = f->coef[t].im;
        }
    }
    return tcp::ok();
}

int tcp_fermion_jordan_wigner(const tcp_fermion *f, uint64_t max_bytes, tcp_operator **out) {
    if (!f || !out) return fail(TCP_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const uint32_t n = f->n_modes;
    std::map<std::vector<uint8_t>, std::vector<Complex>> aggregate;
    uint64_t total = 0;
    int st;
    for (size_t t = 0; t < f->coef.size(); ++t) {
        const std::vector<Factor> seq = sequence_of(*f, t);
        // each factor at most doubles the word list (core/structured/fermion.rs:270-273)
        if (seq.size() < 63 && (st = check_items(1ull << seq.size(), max_bytes, "estimating Jordan-Wigner expansion")))
            return st;
        for (const auto &[word, value] : jw_expand(n, seq)) {
            aggregate[word].push_back(cmul(f->coef[t], value));
            if ((st = check_items(++total, max_bytes, "estimating structured aggregation"))) return st;
        }
    }
    std::vector<uint8_t> codes;
    std::vector<double> coef;
    uint64_t T = 0;
    for (auto &[word, values] : aggregate) {
        const Complex v = deterministic_sum(values);
        if (exact_zero(v)) continue;
        codes.insert(codes.end(), word.begin(), word.end());
        coef.push_back(v.re);
        coef.push_back(v.im);
        ++T;
    }
    st = tcp_operator_from_codes(n, T, codes.data(), coef.data(), out);
    return st ? st : tcp::ok();
}

void tcp_fermion_destroy(tcp_fermion *f) { delete f; }

}  // extern "C"
