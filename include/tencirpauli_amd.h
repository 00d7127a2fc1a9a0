/*
 * tencirpauli_amd.h -- C ABI of the MI355X (gfx950) matrix-free Pauli-sum MVP.
 *
 * Drop-in boundary for TenCirPauli's native MVP path.  Each entry point names
 * the reference interface it replaces (paths relative to the reference
 * repository; core/ = crates/tencir-pauli-core/src/, native/ =
 * crates/tencirpauli-native/src/, py/ = python/tencirpauli/).
 *
 * Conventions kept from the reference:
 *   - Pauli codes 0=I 1=X 2=Y 3=Z; packed words put qubit q in word q/64,
 *     bit q%64, and Y = (x=1, z=1)                       (core/word.rs:53-110)
 *   - qubit 0 is the most significant bit of a computational index
 *                                                      (core/pauli_action.rs:65)
 *   - operators are canonical: duplicates summed in IEEE-bit-pattern order,
 *     sorted by code tuple from qubit 0, exact zeros dropped
 *                                                  (core/operator.rs:117-155)
 *   - apply overwrites its output, never accumulates     (core/tests.rs:7-23)
 *   - max_bytes is advisory below 1 GiB                (py/_native_mvp.py:19-58)
 *   - complex numbers are interleaved (re, im)
 *
 * All functions return a tcp_status; on failure tcp_last_error() returns the
 * message of the calling thread's last error.  Plans are immutable after
 * creation: concurrent applies on distinct output buffers are allowed.
 * Device pointers and hipStream_t handles are passed as void*.
 */
#ifndef TENCIRPAULI_AMD_H
#define TENCIRPAULI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One code per PauliError variant used on this path (core/error.rs:5-74);
 * native/convert.rs:35-42 maps MEMORY_LIMIT to MemoryError, OVERFLOW to
 * OverflowError and the rest to ValueError. */
typedef enum tcp_status {
    TCP_OK = 0,
    TCP_ERR_INVALID_STRUCTURE_LENGTH = 1, /* InvalidStructureLength{expected, actual} */
    TCP_ERR_INVALID_CODE = 2,             /* InvalidCode{code, index} */
    TCP_ERR_NON_FINITE_COEFFICIENT = 3,   /* NonFiniteCoefficient{index} */
    TCP_ERR_MEMORY_LIMIT = 4,             /* MemoryLimit{requested, limit} */
    TCP_ERR_OVERFLOW = 5,                 /* Overflow{context} */
    TCP_ERR_INVALID_SECTOR = 6,           /* InvalidSector{context} */
    TCP_ERR_SECTOR_LEAKAGE = 7,           /* SectorLeakage{source_index, expected, actual} */
    TCP_ERR_INVALID_INDEX = 8,            /* InvalidIndex{context} */
    TCP_ERR_INVALID_ARGUMENT = 9,         /* null pointer, bad dtype/storage, overlap */
    TCP_ERR_INCOMPATIBLE_QUBIT_COUNTS = 10, /* IncompatibleQubitCounts{left, right} */
    TCP_ERR_INVALID_WORD_LENGTH = 11,     /* InvalidWordLength{expected, actual} */
    TCP_ERR_HIP = 12,                     /* HIP runtime failure (no reference counterpart) */
    TCP_ERR_UNSUPPORTED = 13              /* valid input outside the GPU kernels' limits */
} tcp_status;

typedef enum tcp_dtype {
    TCP_C128 = 0, /* complex128 amplitudes, FP64 coefficients, FP64 accumulation */
    TCP_C64 = 1   /* complex64 amplitudes, coefficients rounded to FP32, FP64 accumulation */
} tcp_dtype;

typedef enum tcp_storage {
    TCP_STORAGE_LAZY = 0, /* "lazy" (py/pauli.py:985-1036) */
    TCP_STORAGE_EAGER = 1 /* "eager": same GPU kernels, transition count precomputed */
} tcp_storage;

typedef struct tcp_operator tcp_operator;
typedef struct tcp_mvp_plan tcp_mvp_plan;
typedef struct tcp_sector_plan tcp_sector_plan;

/* ---- library ----------------------------------------------------------- */

/* Library version string (static storage). */
const char *tcp_version(void);

/* Copy the calling thread's last error message into buf (NUL-terminated,
 * truncated to len); returns the full message length. */
size_t tcp_last_error(char *buf, size_t len);

/* Number of visible HIP devices (0 when none). */
int tcp_device_count(int *count);

/* ---- operator: PauliOperator::from_terms (core/operator.rs:117-155),
 *      reached from native/operator.rs:206-243 pauli_operator_native_array -- */

/* codes: nterms x nqubits row-major u8; coef: 2*nterms doubles. */
int tcp_operator_from_codes(uint32_t nqubits, uint64_t nterms, const uint8_t *codes,
                            const double *coef, tcp_operator **out);

/* Packed words (nterms x ceil(n/64) each, core/word.rs:63-91 from_words:
 * padding bits are cleared), then canonicalised like from_terms. */
int tcp_operator_from_words(uint32_t nqubits, uint64_t nterms, const uint64_t *x_words,
                            const uint64_t *z_words, const double *coef, tcp_operator **out);

uint32_t tcp_operator_nqubits(const tcp_operator *op);
uint64_t tcp_operator_term_count(const tcp_operator *op);
/* distinct_x_mask_count (native/operator.rs:50-59) */
uint64_t tcp_operator_distinct_x_mask_count(const tcp_operator *op);

/* Canonical terms as codes (term_count x nqubits) and coefficients. */
int tcp_operator_export_codes(const tcp_operator *op, uint8_t *codes, double *coef);

/* Canonical terms as BackendMvpPlan arrays (core/hamiltonian.rs:501-524):
 * x_words/z_words are term_count x ceil(n/64). */
int tcp_operator_export_words(const tcp_operator *op, uint64_t *x_words, uint64_t *z_words,
                              double *coef);

/* PauliOperator::analyze_charge (core/operator.rs:396-445) for a diagonal
 * qubit charge: qubit_levels holds (level of |0>, level of |1>) per qubit. */
int tcp_operator_analyze_charge(const tcp_operator *op, const double *qubit_levels,
                                int *is_conserved, uint64_t *nonzero);

/* Operator algebra of PauliOperator (core/operator.rs:217-393; py/pauli.py:
 * 597-668): add merges equal words (left + right, exact zeros dropped); scale
 * multiplies every coefficient (zero scalar -> empty operator); multiply forms
 * every product word with its exact Pauli phase and sums equal words in
 * IEEE-bit-pattern order (max_bytes bounds term pairs x (nqubits + 16) bytes);
 * adjoint conjugates the coefficients.  Operands must have equal qubit
 * counts (TCP_ERR_INCOMPATIBLE_QUBIT_COUNTS). */
int tcp_operator_add(const tcp_operator *a, const tcp_operator *b, tcp_operator **out);
int tcp_operator_scale(const tcp_operator *a, double re, double im, tcp_operator **out);
int tcp_operator_multiply(const tcp_operator *a, const tcp_operator *b, uint64_t max_bytes, tcp_operator **out);
int tcp_operator_adjoint(const tcp_operator *a, tcp_operator **out);

void tcp_operator_destroy(tcp_operator *op);

/* ---- fermion operators: FermionOperator.from_terms / map_fermions
 *      (py/_structured/fermion.py:51-244) over structured_fermion_canonicalize
 *      and jordan_wigner (core/structured/integrals.rs:2-54,
 *      core/structured/fermion.rs:141-183, 261-283) ------------------------- */

typedef struct tcp_fermion tcp_fermion;

/* CAR-canonicalise raw ladder products: term t is the product of factors
 * offsets[t] .. offsets[t+1]-1, factor k acting on modes[k] with actions[k]
 * 0 = create, 1 = annihilate; coef interleaved (re, im).  Canonical words put
 * creations (increasing modes) before annihilations (decreasing modes); equal
 * words are summed in IEEE-bit-pattern order, sorted by (creation modes,
 * annihilation modes), exact zeros dropped.  Errors: mode >= n_modes
 * (TCP_ERR_INVALID_INDEX), action > 1 (TCP_ERR_INVALID_CODE), non-finite
 * coefficient, and 192 bytes per expansion item against max_bytes
 * (TCP_ERR_MEMORY_LIMIT). */
int tcp_fermion_from_ladder(uint32_t n_modes, uint64_t nterms, const uint64_t *offsets, const uint32_t *modes,
                            const uint8_t *actions, const double *coef, uint64_t max_bytes, tcp_fermion **out);
uint32_t tcp_fermion_n_modes(const tcp_fermion *f);
uint64_t tcp_fermion_term_count(const tcp_fermion *f);
uint64_t tcp_fermion_factor_count(const tcp_fermion *f); /* factors over all canonical words */
/* Canonical words as factor lists (offsets[term_count + 1], modes/actions
 * [factor_count], coef[2 term_count]); any output may be NULL. */
int tcp_fermion_export(const tcp_fermion *f, uint64_t *offsets, uint32_t *modes, uint8_t *actions, double *coef);
/* Jordan-Wigner image (mode m = qubit m, a_m = Z_<m (X_m + iY_m)/2) as a
 * canonical Pauli operator: the input of every MVP plan and sector plan. */
int tcp_fermion_jordan_wigner(const tcp_fermion *f, uint64_t max_bytes, tcp_operator **out);
void tcp_fermion_destroy(tcp_fermion *f);

/* ---- full-space plan: MvpPlan (core/hamiltonian.rs:52-270) via
 *      pauli_mvp_plan_handle (native/hamiltonian.rs:270-288) ------------- */

/* Build a reusable plan on the current HIP device: index-space masks with the
 * Y phase folded in, terms regrouped by x-mask, tables uploaded to HBM.
 * n >= 64 is TCP_ERR_OVERFLOW (core/hamiltonian.rs:544-555).  max_bytes is
 * checked only when >= 1 GiB, against the retained plan bytes. */
int tcp_mvp_plan_create(const tcp_operator *op, int storage, uint64_t max_bytes,
                        tcp_mvp_plan **out);

/* Same, from canonical BackendMvpPlan arrays (no re-canonicalisation). */
int tcp_mvp_plan_create_from_words(uint32_t nqubits, uint64_t nterms, const uint64_t *x_words,
                                   const uint64_t *z_words, const double *coef, int storage,
                                   uint64_t max_bytes, tcp_mvp_plan **out);

/* Tile geometry of a full-space plan (no reference counterpart: the reference
 * has one CPU loop).  Threads per workgroup x amplitudes per thread; OR-ing
 * TCP_TILE_XCD_ORDER walks the tiles XCD by XCD (the default of
 * TCP_TILE_AUTO; TCP_TILE_LINEAR_ORDER turns it off).  Spaces smaller than the
 * geometry's tile fall back to a smaller one.  tcp_mvp_plan_create uses
 * TCP_TILE_AUTO: complex128 from 22 qubits 1024x8x2 (below, 512x8), complex64
 * storage from 24 qubits 512x8 (below, 512x4), all with the XCD order.  Every tiled
 * geometry maps its kernel index onto the amplitude index through a bit
 * permutation the plan builder chooses per operator (fewer source tiles per
 * output tile, more L2 reuse); results differ only in summation order. */
enum {
    TCP_TILE_AUTO = 0,
    TCP_TILE_512x8 = 1,  /* 2^12-amplitude tiles, two workgroups per CU */
    TCP_TILE_1024x8 = 2, /* 2^13-amplitude tiles */
    TCP_TILE_512x16 = 3, /* 2^13-amplitude tiles */
    TCP_TILE_512x4 = 4,  /* 2^11-amplitude tiles, four workgroups per CU */
    TCP_TILE_1024x4 = 5,  /* 2^12-amplitude tiles, 16 waves per workgroup, two per CU */
    TCP_TILE_1024x8x2 = 6, /* twin tiles: a workgroup owns two 2^13-amplitude output tiles and
                              stages each 2^14-amplitude source tile as two halves that serve both */
    TCP_TILE_512x16x2 = 7  /* twin tiles of 2^13 amplitudes, 512 threads */
};
#define TCP_TILE_XCD_ORDER 0x100u
#define TCP_TILE_LINEAR_ORDER 0x200u /* plain tile order, also for TCP_TILE_AUTO */
/* Source tiles in flight per workgroup for the 512x8, 512x4 and 1024x4
 * geometries: (depth - 1) << TCP_TILE_PREFETCH_SHIFT, depth 1..3 (default 1). */
#define TCP_TILE_PREFETCH_SHIFT 10
#define TCP_TILE_PREFETCH_MASK 0xC00u
#define TCP_TILE_NT_LOADS 0x1000u /* non-temporal source-tile loads (same geometries) */
int tcp_mvp_plan_create_tiled(const tcp_operator *op, int storage, uint64_t max_bytes, uint32_t tiling,
                              tcp_mvp_plan **out);
/* Geometry the given dtype's applies use (after the small-space fallback;
 * 0 = small-space kernel) and log2 of its tile (twin geometries: of each of
 * the two output tiles, the size of a staged source half). */
uint32_t tcp_mvp_plan_tiling(const tcp_mvp_plan *plan, int dtype);
uint32_t tcp_mvp_plan_tile_bits(const tcp_mvp_plan *plan, int dtype);

uint32_t tcp_mvp_plan_nqubits(const tcp_mvp_plan *plan);
uint64_t tcp_mvp_plan_term_count(const tcp_mvp_plan *plan);
uint64_t tcp_mvp_plan_group_count(const tcp_mvp_plan *plan);
/* number of distinct high-part classes of the x-masks above the tile bits */
uint64_t tcp_mvp_plan_class_count(const tcp_mvp_plan *plan);
/* Passes of the tile kernel: (x-mask, slot pattern) pairs (>= group_count). */
uint64_t tcp_mvp_plan_pass_count(const tcp_mvp_plan *plan);
/* Classes of the tables the given dtype's applies use (complex64 storage may use another tile). */
uint64_t tcp_mvp_plan_class_count_for(const tcp_mvp_plan *plan, int dtype);
uint64_t tcp_mvp_plan_dimension(const tcp_mvp_plan *plan);
int tcp_mvp_plan_storage(const tcp_mvp_plan *plan);
/* Nonzero transitions of the eager graph (MvpPlan::transition_count,
 * core/hamiltonian.rs:165-180); UINT64_MAX for lazy plans. */
uint64_t tcp_mvp_plan_transition_count(const tcp_mvp_plan *plan);
/* Retained plan bytes (host + device tables). */
uint64_t tcp_mvp_plan_estimated_bytes(const tcp_mvp_plan *plan);
/* Grouped-gather algorithmic bytes of one apply:
 * 2^n * (G * s_in + s_out) + T * (16 * ceil(n/64) + s_coef). */
uint64_t tcp_mvp_plan_bytes_alg(const tcp_mvp_plan *plan, int dtype);

/* MvpPlan::apply_into (core/hamiltonian.rs:204-270) on device buffers:
 * output <- H * input, enqueued on `stream` (NULL = default stream),
 * asynchronous.  Lengths are in amplitudes and must equal 2^n
 * (TCP_ERR_INVALID_STRUCTURE_LENGTH otherwise); input and output must not
 * overlap. */
int tcp_mvp_plan_apply_into(const tcp_mvp_plan *plan, const void *input, void *output,
                            uint64_t input_len, uint64_t output_len, int dtype, void *stream);

/* Batched right-hand sides: nrhs vectors stored contiguously (vector r at
 * offset r * 2^n amplitudes) in input and output. */
int tcp_mvp_plan_apply_batch(const tcp_mvp_plan *plan, const void *input, void *output,
                             uint64_t nrhs, int dtype, void *stream);

/* NativeMvpPlan.apply_into (native/hamiltonian.rs:347-365) on host complex128
 * buffers: stages through device memory and returns when done. */
int tcp_mvp_plan_apply_host(const tcp_mvp_plan *plan, const double *input, double *output,
                            uint64_t input_len, uint64_t output_len);

void tcp_mvp_plan_destroy(tcp_mvp_plan *plan);

/* PauliOperator::mvp (core/hamiltonian.rs:481-483, native/hamiltonian.rs:254-268):
 * one-shot plan + apply on host complex128 buffers. */
int tcp_mvp(const tcp_operator *op, const double *input, double *output, uint64_t input_len,
            uint64_t output_len);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): y = op v on device buffers of dtype TCP_C128 or TCP_C64
 * with no plan and no device allocation.  The operator's tile tables (the
 * ones tcp_mvp_plan_create builds) are staged in a static 8 MiB workspace
 * that each device holds until the process ends; operators whose tables need
 * more (roughly 250,000 terms) are TCP_ERR_UNSUPPORTED.  The call returns
 * once the tables are copied (it synchronises `stream` once); the MVP then
 * runs on `stream`.  Every call on one device shares the workspace: calls
 * that overlap on one device race, so issue them on one stream or serialise
 * them.  Same results as tcp_mvp_plan_apply_into on a plan of `op`. */
int tcp_mvp_apply_static(const tcp_operator *op, const void *input, void *output, uint64_t input_len,
                         uint64_t output_len, int dtype, void *stream);

/* ---- sectors ------------------------------------------------------------ */

/* U1Sector (core/sector.rs:79-238): dimension C(n, k); rank/unrank of an
 * integer bitstring (qubit 0 = MSB) in ascending integer order. n <= 64. */
int tcp_u1_dimension(uint32_t nqubits, uint32_t particle_number, uint64_t *dim);
int tcp_u1_rank(uint32_t nqubits, uint32_t particle_number, uint64_t bitstring, uint64_t *rank);
int tcp_u1_unrank(uint32_t nqubits, uint32_t particle_number, uint64_t index,
                  uint64_t *bitstring);
/* U1Sector::rank_words / unrank_into (core/sector.rs:127-238, wide sectors):
 * the same order on packed words, ceil(n / 64) of them with qubit q in word
 * q / 64, bit q % 64 (core/word.rs:53-110); n <= 1024.  A wrong weight or a
 * set padding bit is TCP_ERR_INVALID_INDEX.  Sector plans of 64 or more
 * qubits (U(1) only) run the same lazy, eager and Davidson paths on these
 * words; their device rank/unrank/walk entry points are TCP_ERR_UNSUPPORTED. */
int tcp_u1_rank_words(uint32_t nqubits, uint32_t particle_number, const uint64_t *words,
                      uint64_t *rank);
int tcp_u1_unrank_words(uint32_t nqubits, uint32_t particle_number, uint64_t index,
                        uint64_t *words);
/* U1Sector.basis_words_packed (py/symmetry.py:230-240): the whole basis in
 * restricted order, `rows` = dimension rows of one integer bitstring (n <= 64)
 * or of ceil(n / 64) packed words. */
int tcp_u1_basis_words(uint32_t nqubits, uint32_t particle_number, uint64_t *out, uint64_t rows);

/* Contiguous fixed-population blocks, the factorized ChargeSector layout
 * (core/charge_sector.rs:193-217, 419-456, 973-1038): block j covers the next
 * sizes[j] qubits from qubit 0 and holds pops[j] ones; rank is mixed radix
 * with block 0 most significant.  A single block is the U(1) sector. */
int tcp_blocks_dimension(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops,
                         uint64_t *dim);
int tcp_blocks_rank(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops,
                    uint64_t bitstring, uint64_t *rank);
int tcp_blocks_unrank(uint32_t nblocks, const uint32_t *sizes, const uint32_t *pops,
                      uint64_t index, uint64_t *bitstring);

/* Restricted plan for U1Sector(n, k): U1LazyMvpPlan::new (core/sector.rs:515-534)
 * behind pauli_restrict_u1_lazy_handle (native/symmetry.rs:389-402), with the
 * exact-conservation precondition of py/symmetry.py:675-681 checked on the
 * host (TCP_ERR_SECTOR_LEAKAGE when violated). */
int tcp_sector_plan_create_u1(const tcp_operator *op, uint32_t particle_number,
                              uint64_t max_bytes, tcp_sector_plan **out);

/* Restricted plan for contiguous fixed-population blocks: the factorized
 * charge path prepare_charge_mvp_plan (core/charge/generic.rs:1087-1110)
 * behind compile_mvp_pauli_handle (native/charge_sector.rs:342-365); each
 * block's number charge must be exactly conserved (TCP_ERR_INVALID_SECTOR). */
int tcp_sector_plan_create_blocks(const tcp_operator *op, uint32_t nblocks, const uint32_t *sizes,
                                  const uint32_t *pops, uint64_t max_bytes,
                                  tcp_sector_plan **out);
/* The same block sectors for a fermion operator, gathered from its ladder
 * words directly (the raw-fermion restricted MVP of the reference,
 * FermionMaskDescriptor gathers, core/fermion_mask.rs and
 * core/charge/generic.rs:642-809, behind FermionOperator.restrict_charge,
 * py/_structured/fermion.py:230-244): a word acts on a source whose
 * annihilated modes are occupied and other created modes empty, with its
 * Jordan-Wigner sign; mode m is bit n-1-m of the occupation index.  Every word
 * must conserve every block's number (TCP_ERR_INVALID_SECTOR); n <= 63. */
int tcp_sector_plan_create_fermion(const tcp_fermion *f, uint32_t nblocks, const uint32_t *sizes,
                                   const uint32_t *pops, uint64_t max_bytes, tcp_sector_plan **out);

/* General additive-charge sectors on binary axes (qubits or fermion modes):
 * ChargeSectorPlan's suffix-count rank authority (core/charge_sector.rs:8-16,
 * rank/unrank :1426-1580, build_charge_sector_plan :1620) for sectors that are
 * not contiguous fixed-population blocks -- e.g. total number and 2 S_z on
 * interleaved spin orbitals.  Axis a is index bit n-1-a; constraint j reads
 * levels[(j * naxes + a) * 2 + bit], an integer, and targets[j] (charge
 * offsets already subtracted).  Order: lexicographic over axes, axis 0 most
 * significant (ascending integer order).  At most 63 axes and 4 constraints;
 * the rank tables hold (naxes + 1) x prod_j (sum_a |level_1 - level_0| + 1)
 * counts (TCP_ERR_UNSUPPORTED above 2^24); dimensions below 2^32.  Host only. */
typedef struct tcp_charge_sector tcp_charge_sector;
int tcp_charge_sector_create(uint32_t naxes, uint32_t nconstraints, const int64_t *levels, const int64_t *targets,
                             uint64_t max_bytes, tcp_charge_sector **out);
uint64_t tcp_charge_sector_dimension(const tcp_charge_sector *sector);
/* rank of an occupation (TCP_ERR_INVALID_SECTOR when a constraint fails) */
int tcp_charge_sector_rank(const tcp_charge_sector *sector, uint64_t bitstring, uint64_t *rank);
int tcp_charge_sector_unrank(const tcp_charge_sector *sector, uint64_t index, uint64_t *bitstring);
/* states of ranks first .. first + count - 1 (ChargeSector.basis_states, py/charge.py:770-791) */
int tcp_charge_sector_basis_words(const tcp_charge_sector *sector, uint64_t first, uint64_t count, uint64_t *out);
void tcp_charge_sector_destroy(tcp_charge_sector *sector);
/* Restricted plan of a Pauli operator on a general sector: the generic charge
 * loop apply_charge_mvp_from_prepared_plan_into (core/charge/generic.rs:854-1051)
 * as a destination gather.  Every charge must be exactly conserved
 * (TCP_ERR_SECTOR_LEAKAGE otherwise); an aggregated transition that still
 * leaves the sector makes applies fail with TCP_ERR_INVALID_SECTOR, as the
 * reference's loop does. */
int tcp_sector_plan_create_charge(const tcp_operator *op, const tcp_charge_sector *sector, uint64_t max_bytes,
                                  tcp_sector_plan **out);
/* The same for a fermion operator's ladder words (as tcp_sector_plan_create_fermion). */
int tcp_sector_plan_create_fermion_charge(const tcp_fermion *f, const tcp_charge_sector *sector,
                                          uint64_t max_bytes, tcp_sector_plan **out);

uint64_t tcp_sector_plan_dimension(const tcp_sector_plan *plan);
uint32_t tcp_sector_plan_nqubits(const tcp_sector_plan *plan);
uint64_t tcp_sector_plan_term_count(const tcp_sector_plan *plan);
uint64_t tcp_sector_plan_group_count(const tcp_sector_plan *plan);
uint64_t tcp_sector_plan_estimated_bytes(const tcp_sector_plan *plan);

/* U1LazyMvpPlan::apply_into (core/sector.rs:579-642) /
 * PreparedChargeMvpPlan::apply_into (core/charge/generic.rs:1231-1248) on
 * device buffers of `dimension` amplitudes; output overwritten. */
int tcp_sector_plan_apply_into(const tcp_sector_plan *plan, const void *input, void *output,
                               uint64_t input_len, uint64_t output_len, int dtype, void *stream);

/* Host complex128 convenience variant (stages through device memory). */
int tcp_sector_plan_apply_host(const tcp_sector_plan *plan, const double *input, double *output,
                               uint64_t input_len, uint64_t output_len);

/* Device-side rank/unrank of `count` values (u64 device arrays), bit-exact
 * with U1Sector::rank / unrank and ChargeSectorPlan::rank / unrank.
 * Bitstrings outside the sector rank to UINT64_MAX. */
int tcp_sector_plan_rank_device(const tcp_sector_plan *plan, const uint64_t *bitstrings,
                                uint64_t *ranks, uint64_t count, void *stream);
int tcp_sector_plan_unrank_device(const tcp_sector_plan *plan, const uint64_t *ranks,
                                  uint64_t *bitstrings, uint64_t count, void *stream);

/* States of ranks first .. first+count-1 produced by the kernels' successor
 * step (each thread unranks once, then advances `run` times), for bit-exact
 * sweeps of the in-kernel basis iteration against U1BasisIterator
 * (core/sector.rs:751-834). */
int tcp_sector_plan_walk_device(const tcp_sector_plan *plan, uint64_t first, uint64_t *bitstrings,
                                uint64_t count, uint32_t run, void *stream);

void tcp_sector_plan_destroy(tcp_sector_plan *plan);

/* ---- eager restricted operator: U1RestrictedOperator::new + CsrGraph
 *      (core/sector.rs:345-432, core/csr.rs:196-240) and the charge eager
 *      build (core/charge/eager.rs, native/charge_sector.rs:569-845) ------ */

typedef struct tcp_csr_plan tcp_csr_plan;

/* Materialise the sector-restricted operator as CSR on the device: rows are
 * destination ranks, columns ascend inside a row, exact zeros are skipped.
 * No conservation precondition (as the reference's eager builders): every
 * aggregated transition is checked and leakage fails the build with the
 * reference's error (TCP_ERR_SECTOR_LEAKAGE for U(1), TCP_ERR_INVALID_SECTOR
 * for charge blocks).  max_bytes (>= 1 GiB) bounds the retained CSR bytes. */
int tcp_csr_plan_create_u1(const tcp_operator *op, uint32_t particle_number, uint64_t max_bytes,
                           tcp_csr_plan **out);
int tcp_csr_plan_create_blocks(const tcp_operator *op, uint32_t nblocks, const uint32_t *sizes,
                               const uint32_t *pops, uint64_t max_bytes, tcp_csr_plan **out);
/* The full-space matrix as the same device CSR: replaces
 * PauliOperator::coo_matrix / csr_matrix (core/hamiltonian.rs:360-460) and the
 * pauli_coo_handle / pauli_csr_handle bindings behind py/pauli.py:884-928.
 * Row r holds, for each distinct X mask x (ascending), column r ^ x with value
 * sum_t w_t (-1)^popcount(z_t & column) over the group's terms in canonical
 * order; exact zeros are dropped and columns ascend.  nqubits <= 32
 * (TCP_ERR_UNSUPPORTED beyond: columns are 32-bit). */
int tcp_csr_plan_create_full(const tcp_operator *op, uint64_t max_bytes, tcp_csr_plan **out);
/* Eager CSR of a fermion operator on block sectors (as
 * tcp_sector_plan_create_fermion; no conservation precondition, a word that
 * leaves the sector fails the build with TCP_ERR_INVALID_SECTOR). */
int tcp_csr_plan_create_fermion(const tcp_fermion *f, uint32_t nblocks, const uint32_t *sizes,
                                const uint32_t *pops, uint64_t max_bytes, tcp_csr_plan **out);
/* Eager CSR on a general charge sector (compile_charge_csr_from_prepared_plan,
 * core/charge/eager.rs:133-204): no conservation precondition, leakage fails
 * the build with TCP_ERR_INVALID_SECTOR. */
int tcp_csr_plan_create_charge(const tcp_operator *op, const tcp_charge_sector *sector, uint64_t max_bytes,
                               tcp_csr_plan **out);
int tcp_csr_plan_create_fermion_charge(const tcp_fermion *f, const tcp_charge_sector *sector, uint64_t max_bytes,
                                       tcp_csr_plan **out);
uint64_t tcp_csr_plan_dimension(const tcp_csr_plan *plan);
uint64_t tcp_csr_plan_transition_count(const tcp_csr_plan *plan); /* stored entries */
uint64_t tcp_csr_plan_estimated_bytes(const tcp_csr_plan *plan);
/* 1 when the plan stores its entries in the compact form: a 24-bit column and
 * an 8-bit index into a table of at most 255 distinct off-diagonal values per
 * entry, the diagonal in a dense array (chosen when the matrix has fewer than
 * 2^24 rows and at most 255 distinct off-diagonal values); 0 for plain CSR.
 * No reference counterpart (the reference's CsrGraph, core/csr.rs, stores a
 * value per entry); results are the same either way. */
int tcp_csr_plan_is_compact(const tcp_csr_plan *plan);
/* The form the SpMV reads: TCP_CSR_PLAIN, TCP_CSR_COMPACT, or TCP_CSR_SLAB --
 * the compact entries regrouped by x-mask, one 32-bit word per (x-group,
 * row) with 0xFFFFFFFF where the row has no transition in the group, read one
 * row per lane so that a wave's gathers under one x-mask coalesce (chosen
 * for restricted operators of 64 qubits or fewer when the x-groups times the
 * rows are at most 8/3 of the off-diagonal entries and the budget allows);
 * -1 for a null plan.  No reference counterpart; results agree with plain
 * CSR to FP64 rounding (the summation order differs). */
#define TCP_CSR_PLAIN 0
#define TCP_CSR_COMPACT 1
#define TCP_CSR_SLAB 2
int tcp_csr_plan_storage_form(const tcp_csr_plan *plan);
/* x-group slots of the slab form (its words: slots x dimension); 0 for the
 * other forms.  No reference counterpart. */
uint32_t tcp_csr_plan_slab_groups(const tcp_csr_plan *plan);
/* y = A v: row-parallel CSR gather (CsrGraph::apply_into), FP64 accumulation;
 * TCP_C64 uses complex64 vectors and FP32-rounded values. */
int tcp_csr_plan_apply_into(const tcp_csr_plan *plan, const void *input, void *output, uint64_t input_len,
                            uint64_t output_len, int dtype, void *stream);
int tcp_csr_plan_apply_host(const tcp_csr_plan *plan, const double *input, double *output, uint64_t input_len,
                            uint64_t output_len);
/* CsrGraph::to_u64_arrays: host arrays of dimension+1, nnz and 2*nnz entries. */
int tcp_csr_plan_export(const tcp_csr_plan *plan, uint64_t *indptr, uint64_t *indices, double *values);
void tcp_csr_plan_destroy(tcp_csr_plan *plan);

/* ---- batched MVP (no reference counterpart): many operators of uneven size,
 *      each applied to a block of right-hand sides, in one persistent launch -- */

typedef struct tcp_batch_plan tcp_batch_plan;

/* The batch buffer holds, for operator p in order, `nrhs` vectors of 2^n_p
 * amplitudes back to back starting at amplitude tcp_batch_plan_offset(p);
 * input and output use the same layout.  Operators need >= 9 qubits. */
int tcp_batch_plan_create(const tcp_operator *const *ops, uint32_t count, uint32_t nrhs,
                          tcp_batch_plan **out);
uint32_t tcp_batch_plan_count(const tcp_batch_plan *plan);
uint64_t tcp_batch_plan_offset(const tcp_batch_plan *plan, uint32_t index);
uint64_t tcp_batch_plan_total_amplitudes(const tcp_batch_plan *plan);
/* sum over operators of terms x 2^n x nrhs (term*amp*rhs per apply) */
double tcp_batch_plan_work(const tcp_batch_plan *plan);
/* Tile geometries the autotuner chooses from: threads per workgroup and
 * amplitudes per thread. */
uint32_t tcp_batch_config_count(void);
int tcp_batch_config(uint32_t index, uint32_t *threads, uint32_t *amplitudes);
uint32_t tcp_batch_plan_selected(const tcp_batch_plan *plan);
int tcp_batch_plan_select(tcp_batch_plan *plan, uint32_t config);
/* Classes per work item (split-K over an operator's classes; 0 = all):
 * an operator with more classes is applied as several items per output
 * tile whose partial tiles are added with atomic adds to a cleared output,
 * so the summation order, and the last bits of the result, vary between
 * runs.  The plan starts at index 0 (no split). */
uint32_t tcp_batch_split_count(void);
int tcp_batch_split(uint32_t index, uint32_t *classes_per_item);
uint32_t tcp_batch_plan_selected_split(const tcp_batch_plan *plan);
int tcp_batch_plan_select_split(tcp_batch_plan *plan, uint32_t index);
/* Times every geometry on the given buffers (after one warm-up each) with
 * whole operators per item, keeps the fastest, then times every split on it
 * and keeps the fastest; times_ms (optional, tcp_batch_config_count()
 * entries) gets the geometry times, -1 for geometries that do not fit;
 * tcp_batch_plan_split_times the split times (tcp_batch_split_count()
 * entries, -1 for splits not timed because they split no operator). */
int tcp_batch_plan_autotune(tcp_batch_plan *plan, const void *input, void *output, int dtype,
                            uint32_t repeats, float *times_ms, uint32_t *chosen);
int tcp_batch_plan_split_times(const tcp_batch_plan *plan, float *times_ms);
/* output <- H_p input for every operator and right-hand side, on `stream`. */
int tcp_batch_plan_apply(const tcp_batch_plan *plan, const void *input, void *output, int dtype,
                         void *stream);
/* Capture one apply into a hipGraph kept by the plan; replay launches it. */
int tcp_batch_plan_capture(tcp_batch_plan *plan, const void *input, void *output, int dtype,
                           void *stream);
int tcp_batch_plan_replay(const tcp_batch_plan *plan, void *stream);
void tcp_batch_plan_destroy(tcp_batch_plan *plan);

/* ---- Davidson eigensolver (no reference counterpart: the reference's
 *      eigen entry point is SciPy eigsh over a LinearOperator,
 *      py/integrations/scipy.py:13-47) ------------------------------------- */

/* Learned diagonal preconditioner: per-state log values m_i of the correction
 * t_i = r_i e^{m_i} (replacing Jacobi's r_i / (theta - H_ii)).  Returns a device
 * float[dimension] array written in order on `stream` (the solver's), or NULL
 * on failure. */
typedef const void *(*tcp_log_multiplier_fn)(void *ctx, double theta, void *stream);

typedef struct tcp_davidson_options {
    uint32_t max_subspace;    /* basis size before a thick restart (4..32; default 24) */
    uint32_t restart_keep;    /* Ritz vectors kept at a restart (1..8, default 4) */
    uint32_t max_iterations;  /* default 1000 */
    double residual_tol;      /* converged when ||H u - theta u|| <= residual_tol (FP64 phase) */
    int mvp_dtype;            /* TCP_C128, or TCP_C64: a first phase with complex64 basis
                                 storage and complex64 MVPs (FP64 accumulation throughout) */
    double switch_tol;        /* the TCP_C64 phase ends at ||r|| <= switch_tol: its Ritz vector
                                 is formed in FP64 and checked with one FP64 MVP (converged if
                                 that residual meets residual_tol, else FP64 iterations follow) */
    uint64_t seed;            /* deterministic start vector */
    tcp_log_multiplier_fn log_multiplier_fn; /* NULL: plain Jacobi preconditioner */
    void *log_multiplier_ctx;
    uint64_t log_multiplier_len; /* floats the callback's buffer holds; must equal the plan's
                                    dimension when log_multiplier_fn is set (else
                                    TCP_ERR_INVALID_STRUCTURE_LENGTH before any launch) */
    double *trace;            /* host [trace_capacity x 4]: seconds, theta, ||r||, MVP bits (32/64) */
    uint32_t trace_capacity;
    void *stream;             /* hipStream_t (NULL = default stream) */
} tcp_davidson_options;

typedef struct tcp_davidson_result {
    double eigenvalue;
    double residual_norm;
    double seconds;           /* wall time of the solve, host clock */
    uint32_t iterations;
    uint32_t mvps_c64;
    uint32_t mvps_c128;
    uint32_t restarts;
    uint32_t trace_len;
    uint32_t converged;
    uint32_t final_fp64;
} tcp_davidson_result;

void tcp_davidson_default_options(tcp_davidson_options *opt);

/* Lowest eigenpair of a restricted (lazy or eager CSR) / full-space plan;
 * `eigenvector` is an optional device complex128 buffer of `dimension`
 * amplitudes (normalised Ritz vector on return).  Device memory: two pools of
 * (max_subspace + restart_keep) complex128 columns. */
int tcp_davidson_sector(const tcp_sector_plan *plan, const tcp_davidson_options *opt,
                        void *eigenvector, tcp_davidson_result *result);
int tcp_davidson_full(const tcp_mvp_plan *plan, const tcp_davidson_options *opt, void *eigenvector,
                      tcp_davidson_result *result);
int tcp_davidson_csr(const tcp_csr_plan *plan, const tcp_davidson_options *opt, void *eigenvector,
                     tcp_davidson_result *result);

/* ---- thick-restart Lanczos (no reference counterpart: replaces the
 *      scipy.sparse.linalg.eigsh call the reference makes over its
 *      LinearOperator, examples/research/fermi_hubbard/run_tencirpauli.py:
 *      194-211, py/integrations/scipy.py:13-47) ------------------------------ */

enum {
    TCP_WHICH_SA = 0, /* smallest algebraic */
    TCP_WHICH_LA = 1, /* largest algebraic */
    TCP_WHICH_LM = 2  /* largest magnitude */
};

typedef struct tcp_lanczos_options {
    uint32_t nev;          /* wanted eigenpairs: 1 <= nev < dimension */
    uint32_t ncv;          /* basis size: nev < ncv <= min(dimension, 32) (default 20) */
    uint32_t which;        /* TCP_WHICH_* */
    uint32_t max_restarts; /* default 1000 */
    double tol;            /* residual estimate <= tol * max(|theta|, eps^(2/3)); 0 = machine epsilon */
    const void *v0;        /* optional device complex128 start vector (NULL: deterministic hash) */
    uint64_t seed;
    void *stream;          /* hipStream_t (NULL = default stream) */
    uint32_t keep;         /* Ritz vectors kept at a restart, nev <= keep < ncv (0: nev + (ncv - nev) / 2) */
} tcp_lanczos_options;

typedef struct tcp_lanczos_result {
    uint32_t converged;    /* wanted pairs that met tol (== nev on success) */
    uint32_t restarts;
    uint64_t mvps;
    double seconds;        /* wall time, host clock */
} tcp_lanczos_result;

void tcp_lanczos_default_options(tcp_lanczos_options *opt);

/* The `nev` wanted eigenvalues in ascending order into host double[nev];
 * optional outputs: `eigenvectors` a device complex128 [nev][dimension] array
 * (row k the normalised Ritz vector of eigenvalue k) and `residual_norms` a
 * host double[nev] of explicit ||A x_k - lambda_k x_k|| (one MVP each).  A
 * solve that stops at max_restarts returns TCP_OK with converged < nev.
 * Device memory: ncv + 3 complex128 vectors. */
int tcp_lanczos_full(const tcp_mvp_plan *plan, const tcp_lanczos_options *opt, double *eigenvalues,
                     void *eigenvectors, double *residual_norms, tcp_lanczos_result *result);
int tcp_lanczos_sector(const tcp_sector_plan *plan, const tcp_lanczos_options *opt, double *eigenvalues,
                       void *eigenvectors, double *residual_norms, tcp_lanczos_result *result);
int tcp_lanczos_csr(const tcp_csr_plan *plan, const tcp_lanczos_options *opt, double *eigenvalues,
                    void *eigenvectors, double *residual_norms, tcp_lanczos_result *result);

/* Real diagonal H_ii of a plan into a device double[dimension] array. */
int tcp_sector_plan_diagonal(const tcp_sector_plan *plan, double *diag, void *stream);
int tcp_mvp_plan_diagonal(const tcp_mvp_plan *plan, double *diag, void *stream);
int tcp_csr_plan_diagonal(const tcp_csr_plan *plan, double *diag, void *stream);

/* ---- learned preconditioner (no reference counterpart): a 64-128-128-1 tanh
 *      MLP, bf16 weights, FP32 accumulation on MFMA, trained on the GPU ---- */

typedef struct tcp_mlp tcp_mlp;
typedef struct tcp_learned_precond tcp_learned_precond;

/* Flat FP32 parameter vector: W1[64][128], b1[128], W2[128][128], b2[128],
 * w3[128], b3 (in-major weight matrices). */
uint32_t tcp_mlp_parameter_count(void);
int tcp_mlp_create(uint64_t seed, tcp_mlp **out);
int tcp_mlp_get_parameters(const tcp_mlp *mlp, float *params);
int tcp_mlp_set_parameters(tcp_mlp *mlp, const float *params);
/* out[b] = MLP(features[b]) for bf16 features [batch x 64] (device buffers). */
int tcp_mlp_forward(const tcp_mlp *mlp, const void *features, uint64_t batch, float *out, void *stream);
/* Parameter gradient of sum_b dloss_dout[b] * MLP(features[b]) into the
 * MLP's gradient buffer (and grad_host when not NULL). */
int tcp_mlp_gradient(tcp_mlp *mlp, const void *features, uint64_t batch, const float *dloss_dout,
                     float *grad_host, void *stream);
/* One Adam step (beta 0.9/0.999) with the last gradient. */
int tcp_mlp_adam_step(tcp_mlp *mlp, float lr, float weight_decay, void *stream);
void tcp_mlp_destroy(tcp_mlp *mlp);

/* Per-state features of a sector plan (bf16 [dim x 64]) for `mlp`; the
 * callback below is a tcp_log_multiplier_fn for tcp_davidson_options. */
int tcp_learned_precond_create(const tcp_mlp *mlp, const tcp_sector_plan *plan,
                               tcp_learned_precond **out);
const void *tcp_learned_precond_log_multiplier(void *ctx, double theta, void *stream);
/* The callback re-evaluates the network only when theta has moved by more
 * than threshold * scale (scale: the operator's coupling scale) since its last
 * evaluation; 0 re-evaluates on every call.  Default 4. */
int tcp_learned_precond_set_refresh(tcp_learned_precond *precond, double threshold);
/* Writes feature 0, (H_ii - theta) / scale, of every state (the callback does
 * this itself; training calls it before tcp_mlp_forward / tcp_mlp_gradient on
 * the features buffer). */
int tcp_learned_precond_prepare(tcp_learned_precond *precond, double theta, void *stream);
int tcp_learned_precond_buffers(const tcp_learned_precond *precond, const void **features,
                                const double **diag, double *scale);
/* One-step Ritz-gain loss of a Davidson state (normalised Ritz vector u,
 * Ritz value theta, residual r = H u - theta u; device complex128 [dim]) for the
 * correction t_i = r_i e^{logm_i}: with q = t - u <u, t>, the lowest Ritz value
 * of span{u, q} is theta - gain, and loss = -log(gain).  dlogm (device float
 * [dim]) receives dloss/dlogm.  With logm NULL the Jacobi correction
 * r_i / (theta - H_ii) is scored instead and dlogm is not written.  The
 * sector plan the object was created from must still exist. */
int tcp_learned_precond_ritz_loss(tcp_learned_precond *precond, double theta, const void *u,
                                  const void *r, const float *logm, float *dlogm, double *loss,
                                  double *gain, void *stream);
void tcp_learned_precond_destroy(tcp_learned_precond *precond);

#ifdef __cplusplus
}
#endif

#endif /* TENCIRPAULI_AMD_H */
