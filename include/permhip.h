/*
 * permhip — exact modular permanents on AMD Instinct MI355X (gfx950).
 *
 * C ABI of libpermhip.so.  Plain pointers and sizes only; every entry point
 * returns a negative value on failure (the reference CUDA ABI's -1 sentinel,
 * perm_cuda.py:125-129, perm_binary_cuda.py:115-123) and leaves a message
 * for phip_last_error().  Host buffers belong to the caller; the library
 * owns all device memory.  Calls may come from several threads: each device
 * context serialises its own calls, and no call poisons later ones.
 *
 * Reference interfaces replaced (file:line in igorrivin/rivin-perm):
 *   cuda_runtime_available       perm_modp_cuda.cu:731-735, perm_binary_cuda.cu:465-469
 *   permanent_modp_cuda          perm_modp_cuda.cu:740-742 (and perm_c_ext.permanent_mod_p,
 *                                perm_c_omp.c:1104-1107, reached through perm.py:908-930)
 *   schur_permanent_modp_cuda    perm_modp_cuda.cu:1266 (and perm_c_ext.schur_permanent_mod_p,
 *                                perm_c_omp.c:942-1060)
 *   binary_permanent_modp_cuda   perm_binary_cuda.cu:471-472 (and
 *                                perm_binary_ext.binary_permanent_mod_p, perm_binary_omp.c:386-519)
 *   permanent_mod3_matrix        perm_mod3_cuda.cu:300 (and perm_mod3_omp.permanent_mod3,
 *                                perm_mod3_omp.c:220-301)
 *   permanent_mod3_cuda          perm_mod3_cuda.cu:164-166
 */
#ifndef PERMHIP_H
#define PERMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 1 when a HIP device is usable, 0 otherwise (never negative).
 * Replaces cuda_runtime_available (perm_modp_cuda.cu:731-735). */
int phip_available(void);

/* Copies the last error message of the calling thread into buf (always
 * NUL-terminated when len > 0).  Returns the message length. */
int phip_last_error(char *buf, size_t len);

/* Writes "<device name> <gcnArchName> CUs=<n>" into buf.  Returns 0 or -1. */
int phip_device_info(char *buf, size_t len);

/* Permanent of one n x n integer matrix modulo each of k moduli, all moduli
 * in one device pass.  A is row-major int64; prime_stride is the element
 * offset between per-modulus copies of the matrix (0: the same matrix for
 * every modulus; n*n: a k x n x n array of matrices already reduced by the
 * caller, the reference wrappers' normalise-before-narrowing rule,
 * perm_cuda.py:104-122).  Any 0 < m < 2^63 is accepted, 0 <= n < 64.
 * Odd parts below 2^31 at orders 5..46 run on the persistent 31-bit
 * engine; other odd parts on the wide Glynn kernels (32-bit Montgomery
 * below 2^32, 61-bit pseudo-Mersenne for 2^61 - c with odd c < 2^28,
 * 64-bit Montgomery otherwise); even parts add a Ryser pass modulo 2^64.
 * out[i] receives perm(A) mod moduli[i] in [0, m).  Returns 0 or -1.
 * Multi-modulus form of permanent_modp_cuda (perm_modp_cuda.cu:740-742). */
int phip_permanent_residues(const int64_t *A, int64_t prime_stride, int n, const int64_t *moduli,
                            int k, int64_t *out);

/* phip_permanent_residues without any device allocation: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results, restricted to the persistent engine's inputs:
 * orders 5..46, 1..64 odd moduli in [3, 2^31).  The plan's device image
 * lives in a fixed 8 MiB device array per GPU (held until the process
 * ends); a plan that does not fit is refused with an error.  Calls on one
 * GPU share that array and are serialised by the library; callers outside
 * the library that run it concurrently on one GPU would race.  0 or -1. */
int phip_permanent_residues_noalloc(const int64_t *A, int64_t prime_stride, int n, const int64_t *moduli,
                                    int k, int64_t *out);

/* Coefficients of per(C + x L) modulo each of k odd moduli 3 <= m < 2^63
 * (not necessarily prime), 0 <= n <= 63 for moduli below 2^30, n <= 40
 * below 2^32 and n <= 32 for larger ones.  C and L are row-major int64;
 * prime_stride as in phip_permanent_residues (0: one pair for every modulus;
 * n*n: k pairs already reduced by the caller).  out[i*(n+1) + d] receives
 * the coefficient of x^d modulo moduli[i] in [0, m).  Returns 0 or -1.
 * Device form of perm_polynomial_ext.affine_permanent_mod
 * (perm_polynomial_omp.c:1207-1260; perm_polynomial.py:402-443), all
 * moduli in one pass.  Orders 41..63 take 2^40 and more Gray states of
 * degree-n products per modulus (hours at order 44). */
int phip_affine_permanent_residues(const int64_t *C, const int64_t *L, int64_t prime_stride, int n,
                                   const int64_t *moduli, int k, int64_t *out);

/* Host preprocessing (no device work): a perfect matching of the 0/1 matrix
 * given as n <= 63 row masks (bit j = column j) and the irreducible blocks
 * of its matching digraph (the strongly connected components), which
 * replaces _matching_blocks (perm_binary.py:709-752).  Blocks are numbered
 * by their smallest row; a block's rows and columns keep ascending order,
 * which fixes the local column labels.  Returns 1 with *block_count blocks
 * (block_sizes[b] rows each; block_masks holds their rows' local masks,
 * block after block, n words in all), 0 when no perfect matching exists,
 * -1 on invalid input. */
int phip_binary_blocks(const uint64_t *masks, int n, int *block_count, int *block_sizes,
                       uint64_t *block_masks);

/* The same decomposition for a 0/1 matrix of any order in CSR form (row r's
 * columns are col_idx[row_ptr[r] .. row_ptr[r+1]), no repeats): a perfect
 * matching (row_match[r] = its column; row_match may be null), and for
 * every row and column the block it belongs to (blocks numbered by their
 * smallest row).  Returns 1, 0 when no perfect matching exists, -1 on
 * invalid input.  Host code only.  Replaces _matching_blocks and
 * _perfect_matching (perm_binary.py:609-655, 709-752) for orders beyond
 * the 64-bit masks of phip_binary_blocks. */
int phip_binary_blocks_csr(const int32_t *row_ptr, const int32_t *col_idx, int n, int32_t *row_match,
                           int32_t *row_block, int32_t *col_block, int *block_count);

/* Live-column frontier DP of a 0/1 matrix (perm_binary.py:389-441) on the
 * device, counts modulo k <= 64 odd moduli below 2^31.  The reference's DP
 * runs in Python integers with no FFI; this entry replaces its state loop
 * (perm_binary.py:412-441).  Rows come in processing order as masks over
 * `slots` <= 64 live-column slots: row_masks[i] has the slots of row i's
 * columns, expiring[i] (a subset of row_masks[i]) the slots of the columns
 * whose last row it is; a slot may be reused by a column whose first row
 * comes after the last row of the column that held it.  Returns 0 with
 * out[q] = count mod primes[q]; 1 when a layer would hold more than
 * max_states states (info[0] = rows processed including that one,
 * info[1] = max_states + 1, as FrontierStateLimitExceeded reports); -1 on
 * invalid input or a layer with 2^31 or more successor pairs. */
int phip_binary_frontier_residues(const uint64_t *row_masks, const uint64_t *expiring, int n, int slots,
                                  const int64_t *primes, int k, int64_t max_states, int64_t *out,
                                  int64_t *info);

/* Processing order of a 0/1 matrix's rows for the frontier DP, in place of
 * _frontier_greedy_order (perm_binary.py:282-386): Cuthill-McKee level
 * structures refined by sifting, the order with the least (maximum width,
 * frontier area) among those tried.  The matrix is in CSR form: row r's
 * columns are col_idx[row_ptr[r] .. row_ptr[r+1]) (distinct, in [0, n),
 * any order).  order_out receives n row indices, *max_width the largest
 * live-column frontier and *area the sum of the widths.  Host code only.
 * Returns 0 or -1. */
int phip_frontier_order(const int32_t *row_ptr, const int32_t *col_idx, int n, int32_t *order_out, int *max_width,
                        int64_t *area);

/* perm(A) mod p for one modulus; -1 on failure.
 * Replaces permanent_modp_cuda (perm_modp_cuda.cu:740-742). */
int64_t phip_permanent_modp(const int64_t *A, int n, int64_t p);

/* perm(S_n) mod p with S_n[j][k] = omega^(jk); n == 0 or 1 gives 1 mod p.
 * Odd n >= 3, odd p and omega^n == 1 use the cofactor identity
 * per(S_n) = n * per(minor_(n-1)) (perm_c_omp.c:971-975); otherwise the
 * full matrix is used.  For a prime order n in [necklace min order, 47] and
 * an odd prime p < 2^32 the minor is evaluated by necklace orbits of its
 * circulant form (perm_c_omp.c:79-507), else by Glynn.  Replaces
 * schur_permanent_modp_cuda (perm_modp_cuda.cu:1266) and
 * perm_c_ext.schur_permanent_mod_p. */
int64_t phip_schur_permanent_modp(int n, int64_t omega, int64_t p);

/* The necklace walk of the Schur path for an odd prime order n, with no
 * terms formed: *representatives = orbits of sign words of length n - 1
 * under rotation and complement that the walk's representatives make up
 * (the walk visits a weight-(n-1)/2 necklace and its complement's necklace
 * separately; such a pair counts once), *words = sign words they stand for
 * (2^(n-1) when the walk covers every word once).  Replaces perm_c_ext.schur_necklace_orbit_stats
 * (perm_c_omp.c:1062-1080, 314-318).  The device walks orders up to 47;
 * larger prime orders (2^52 words and more) are refused.  0 or -1. */
int phip_schur_necklace_orbit_stats(int n, uint64_t *representatives, uint64_t *words);

/* Permanent of a 0/1 matrix (row-major bytes) modulo k odd primes p with
 * n < p < 2^63.  A may be NULL only for n == 0.  Entries other than 0/1,
 * composite or too small moduli and n > 63 are rejected with -1 before any
 * device work.  Returns 0 or -1.  Multi-prime form of
 * binary_permanent_modp_cuda (perm_binary_cuda.cu:471-472). */
int phip_binary_permanent_residues(const uint8_t *A, int n, const int64_t *primes, int k,
                                   int64_t *out);

/* One-prime form; -1 on failure.  Replaces binary_permanent_modp_cuda. */
int64_t phip_binary_permanent_modp(const uint8_t *A, int n, int64_t p);

/* perm(A) mod 3 in {0, 1, 2} for an int matrix (row-major, any values;
 * reduced mod 3 inside), 0 <= n < 64; -1 on failure.
 * Replaces permanent_mod3_matrix (perm_mod3_cuda.cu:300). */
int phip_permanent_mod3_matrix(const int *A, int n);

/* perm mod 3 from the bipedal column encoding (bit i of mags[j] set when
 * a_ij != 0 mod 3, bit i of sgns[j] set when a_ij == 2 mod 3).
 * Replaces permanent_mod3_cuda (perm_mod3_cuda.cu:164-166). */
int phip_permanent_mod3(const uint64_t *mags, const uint64_t *sgns, int n);

/* ---------------------------------------------------------------------- */
/* Persistent batch (no reference counterpart: the reference issues one    */
/* synchronous call per (matrix, prime), perm_cuda.py:196-198).            */
/* ---------------------------------------------------------------------- */

/* Build a device-resident plan for `count` matrices: entries holds the
 * row-major int64 matrices back to back (sum of orders[i]^2 values),
 * moduli holds moduli_counts[i] odd moduli in [3, 2^31) per matrix.  Orders
 * must lie in [5, 46].  *out receives an opaque handle.  Returns 0 or -1. */
int phip_batch_create(const int64_t *entries, const int32_t *orders, int count, const int64_t *moduli,
                      const int32_t *moduli_counts, void **out);

/* Run the whole plan (one persistent launch per row-sum class, replayed
 * from a captured hipGraph).  residues_out, when not NULL, receives
 * perm(A_i) mod each of A_i's moduli in input order.  Returns 0 or -1. */
int phip_batch_run(void *plan, int64_t *residues_out);

/* Asynchronous form: phip_batch_launch enqueues a run and returns at once;
 * phip_batch_query returns 1 once it has finished (copying the residues to
 * residues_out when not NULL), 0 while it runs (block == 0), or -1.  Runs of
 * different plans proceed concurrently on their own streams. */
int phip_batch_launch(void *plan);
int phip_batch_query(void *plan, int64_t *residues_out, int block);

#define PHIP_BATCH_BLOCKS_PER_CU 1 /* persistent workgroups per CU (1..64) */
#define PHIP_BATCH_USE_GRAPH 2     /* 0: plain stream launches, else hipGraph replay */
#define PHIP_BATCH_ITEMS_LOG2 3    /* re-size the chunks for about 2^value work items (0..30) */
int phip_batch_set_option(void *plan, int option, int64_t value);

/* out[0] last run ms (persistent kernels only without a graph; whole graph
 * otherwise), [1] Gray states and [2] algorithmic mulmods (2^(n-1) and
 * n 2^(n-1) per residue), [3] jobs (one table with a set of primes that
 * share it), [4] work items, [5] persistent launches, [6] persistent
 * wavefronts, [7] row-sum group size and [9] prime-set size of the first
 * launch, [8] residues, [10] and [11] smallest and largest chunk (log2 of
 * the Gray states per lane), [12] work-item target (log2) the chunks were
 * sized for, [13] workgroups per CU, [14] kernel variant of the first
 * launch: 0 plain, 1 pairs consecutive states (prime sets below
 * 1.25 * 2^30), 2 skips zero terms (0/1 tables whose states are mostly
 * zero), 3 exact accumulation, 4 compact exact accumulation, 5 the
 * zero-skipping block walk with exact accumulation, 6 exact accumulation in
 * signed limbs, [15] most workgroups
 * per CU that any of the plan's kernels keeps
 * resident (register and LDS use; each class's grid is capped at its own
 * count; -1 if unknown), [16] table words the zero test of the first job
 * reads (zero-skipping jobs: all of them, or the leading half when every
 * row of even degree lies there; 0 for other jobs), [17] rows the first
 * job's 32-state block test reads (zero-skipping jobs walked in blocks;
 * -1 otherwise).  Returns the number of values written or -1. */
int phip_batch_stats(void *plan, double *out, int len);

/* The kernel class of every residue of a batch, in input order, as
 * phip_batch_create would plan it: (group * 32 + prime-set size) * 8 +
 * variant (0 plain, 1 paired states, 2 zero-skipping, 3 exact accumulation,
 * 4 compact exact accumulation, 5 zero-skipping block walk with exact
 * accumulation, 6 exact accumulation in signed limbs; the set size of an exact class is its limb column count).  Same description
 * arguments as phip_batch_create; host code only (no device).  Writes
 * min(len, residues) entries and returns the residue count, or -1.  No
 * counterpart in the reference (plan introspection for tests and
 * profiles). */
int phip_batch_plan_classes(const int64_t *entries, const int32_t *orders, int count, const int64_t *moduli,
                            const int32_t *moduli_counts, int32_t *out, int len);

/* The tables the block zero-skipping walk reads for one matrix of order
 * 8..46 (row-major int64 entries), as a plan builds them for the matrix as
 * given (a plan walks a 0/1 matrix's transpose instead when its estimated
 * share of nonzero terms is smaller): with mod3 == 0 a 0/1 matrix's, with
 * mod3 != 0 the modulo-3 walk's for entries in {-1, 0, 1}.  meta[0] rows the block test reads, [1] rows of even degree
 * (first in the table's row order), [2] row-sum group, [3] n.  Writes the
 * table words to out when cap holds them and returns their count, or -1.
 * Host code only.  No counterpart in the reference (table introspection
 * for tests). */
int64_t phip_block_tables(const int64_t *a, int n, int mod3, int32_t *out, int64_t cap, int32_t *meta);

int phip_batch_destroy(void *plan);

/* ---------------------------------------------------------------------- */
/* Tuning and measurement (no reference counterpart).                      */
/* ---------------------------------------------------------------------- */

#define PHIP_TUNE_LANES_LOG2 1       /* single-matrix launch: target lanes (log2) */
#define PHIP_TUNE_MIN_CBITS 2        /* single-matrix launch: minimum chunk bits */
#define PHIP_TUNE_BATCH_ITEMS_LOG2 3 /* batch: target work items (log2), read at create */
#define PHIP_TUNE_BATCH_BLOCKS 4     /* batch: default persistent workgroups per CU (default 4; each class's
                                        grid is capped at what its kernel keeps resident) */
#define PHIP_TUNE_NECKLACE_MIN_ORDER 5 /* Schur: smallest prime order using necklace orbits (5..64) */
#define PHIP_TUNE_NECKLACE_DEPTH 6   /* Schur necklaces: largest FKM prefix depth of the tasks (2..30) */
#define PHIP_TUNE_WIDE_MIN_TIER 7    /* wide-modulus Glynn: smallest row-register tier (8..64, a multiple of 4) */
#define PHIP_TUNE_AFFINE_MIN_TIER 8  /* affine Glynn: smallest row-register tier (8..40 a multiple of 4, or 48,
                                        56, 64: taken by moduli below 2^30, others use their largest tier) */
#define PHIP_TUNE_SKIP_ZEROS_PERMILLE 9 /* 0/1 tables: largest estimated share of nonzero Gray terms, in
                                           permille, that runs on the zero-skipping kernels (0..1000, 0 off;
                                           read when a plan is built) */
#define PHIP_TUNE_EXACT_MIN_PRIMES 10   /* fewest primes of one table (of at most 8 row groups) that run on the
                                           exact-accumulation kernels, one walk for all of them (0 off, 1..1024;
                                           read when a plan is built) */
#define PHIP_TUNE_EXACT_WIDE_MIN_PRIMES 11 /* the same for tables of 9..16 row groups whose row-bound product is
                                              below 2^372 (0 off, 1..1024; read when a plan is built; the
                                              exact variant is off for every table while
                                              PHIP_TUNE_EXACT_MIN_PRIMES is 0) */
#define PHIP_TUNE_EXACT_COMPACT 12 /* 1 (default): tables of 9..16 groups of 2 rows whose partial bound
                                      products fit take the compact exact classes (product tree sized for
                                      22-bit group values); 0: the 31-bit exact classes.  Read when a plan is
                                      built */
#define PHIP_TUNE_SKIP_BLOCKS 13   /* 1 (default): tables on the zero-skipping kernels of order 8 and above
                                      walk in blocks of 32 Gray states, one mask lookup per row and block;
                                      0: one zero test per state.  Read when a plan is built */
#define PHIP_TUNE_MOD3_BLOCKS 14   /* 1 (default): permanents modulo 3 of orders 8..46 run on the modulo-3
                                      block walk (32 states per row lookup, a term is zero when a row sum
                                      is a multiple of 3); 0: the bit-sliced kernel for every order */
#define PHIP_TUNE_SKIP_EXACT_MIN_PRIMES 15  /* fewest primes (default 5; 0 off) of a 0/1 block table of at most
                                      eight row groups that take the block walk with exact accumulation (one
                                      limb product per taken state for all primes instead of a Montgomery
                                      tree per prime of a padded set of 4 or 8).  Read when a plan is
                                      built */
#define PHIP_TUNE_EXACT_SIGNED 16  /* 1 (default): exact-accumulation tables of at most eight groups of 3 or 4
                                      rows whose four-group bound products are below 2^122 multiply their
                                      signed group products in signed limbs (no magnitudes or sign per
                                      state); 0: the magnitude classes.  Read when a plan is built */
int phip_set_tuning(int key, int64_t value);
int64_t phip_get_tuning(int key);

/* Event timing of the single-matrix Glynn kernel launches: enable (and
 * reset) with phip_profile_enable(1); phip_profile_read returns the summed
 * kernel milliseconds and the launch count. */
int phip_profile_enable(int on);
int phip_profile_read(double *kernel_ms, uint64_t *launches);

/* Whole-device throughput (lane operations per second) of v_add_u32,
 * v_mul_i32_i24, v_mul_lo_u32, v_mul_hi_i32, v_mad_i64_i32, v_mad_u64_u32,
 * one signed Montgomery product, v_pk_add_u16, v_mul_i32_i24 with SDWA
 * word selects, v_lshl_add_u64 and v_ashrrev_i32, in that order.  Returns
 * the count. */
int phip_valu_microbench(double *lane_ops_per_s, int len);

/* Issue rate of an instruction mix: each iteration of every lane issues
 * rounds[t] rounds of 8 independent instructions of type t = v_mad_i64_i32,
 * v_mul_lo_u32, v_pk_add_u16, SDWA v_mul_i32_i24, v_add_u32, interleaved
 * evenly.  Only the mixes of compiled Glynn classes are available (per 8
 * Gray states of n = 32, groups of 4: {92, 56, 18, 24, 4} for a set of 8
 * primes, {15, 7, 18, 24, 4} for one prime); others return -1.
 * *lane_iters_per_s receives lane-iterations per second over the whole
 * device.  Returns 0 or -1. */
int phip_mix_microbench(const int32_t *rounds, int iters, double *lane_iters_per_s);

/* `reps` launches of a plain copy kernel over `bytes` bytes. */
int phip_copy_calibration(uint64_t bytes, int reps);

#ifdef __cplusplus
}
#endif

#endif /* PERMHIP_H */
