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
 * out[i] receives perm(A) mod moduli[i] in [0, m).  Returns 0 or -1.
 * Multi-modulus form of permanent_modp_cuda (perm_modp_cuda.cu:740-742). */
int phip_permanent_residues(const int64_t *A, int64_t prime_stride, int n, const int64_t *moduli,
                            int k, int64_t *out);

/* perm(A) mod p for one modulus; -1 on failure.
 * Replaces permanent_modp_cuda (perm_modp_cuda.cu:740-742). */
int64_t phip_permanent_modp(const int64_t *A, int n, int64_t p);

/* perm(S_n) mod p with S_n[j][k] = omega^(jk); n == 0 or 1 gives 1 mod p.
 * Odd n >= 3, odd p and omega^n == 1 use the cofactor identity
 * per(S_n) = n * per(minor_(n-1)) (perm_c_omp.c:971-975); otherwise the
 * full matrix is used.  Replaces schur_permanent_modp_cuda
 * (perm_modp_cuda.cu:1266) and perm_c_ext.schur_permanent_mod_p. */
int64_t phip_schur_permanent_modp(int n, int64_t omega, int64_t p);

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

#ifdef __cplusplus
}
#endif

#endif /* PERMHIP_H */
