/*
 * ORACLE — test infrastructure only.  Nothing in the product links or calls
 * this file; it is used by tests/, by __graft_entry__.smoke() as the checker,
 * and by bench.py's cpu_baseline leg.
 *
 * A CPU restatement (C + OpenMP) of the reference's residue kernels for the
 * exact modular permanent path, written from the reference's algorithms:
 *
 *   oracle_glynn_mod_p    fixed-sign Glynn with contiguous Gray blocks per
 *                         thread and 64-bit '%' row products
 *                         (perm_c_omp.c:510-639, 36-62)
 *   oracle_ryser_mod_p    Ryser over 2^n subsets, any 0 < p < 2^63, with the
 *                         (-1)^n correction (perm_c_omp.c:666-835)
 *   oracle_permanent_mod_p  the reference's algorithm choice: Glynn iff
 *                         n >= 4, p odd and p <= 3037000499 (perm_c_omp.c:846-854)
 *   oracle_binary_mod_p   0/1 fixed-sign Glynn with the densest column
 *                         fixed, ascending-degree Gray bits, orientation by
 *                         update cost, ratio-product and zero counter,
 *                         32-bit Montgomery for p < 2^31, 128-bit products
 *                         otherwise (perm_binary_omp.c:95-365)
 *   oracle_binary_update_cost  Sum_b deg_b 2^(n-2-b) over ascending column
 *                         degrees (perm_binary_omp.c:203-223)
 *   oracle_mod3           bit-sliced (bipedal) Ryser over F_3
 *                         (perm_mod3_omp.c:21-215)
 *
 * Each function is pinned against the reference's own golden values and
 * against closed forms / brute force in tests/test_oracle.py.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_MAX_ORDER 63

static int thread_count(int requested) {
#ifdef _OPENMP
    return requested > 0 ? requested : omp_get_max_threads();
#else
    (void)requested;
    return 1;
#endif
}

static inline int64_t addm(int64_t a, int64_t b, int64_t p) { return a >= p - b ? a - (p - b) : a + b; }
static inline int64_t subm(int64_t a, int64_t b, int64_t p) { return a >= b ? a - b : p - (b - a); }

static inline int64_t mulm(int64_t a, int64_t b, int64_t p) {
    return (int64_t)(((unsigned __int128)(uint64_t)a * (uint64_t)b) % (uint64_t)p);
}

static int64_t powm(int64_t b, uint64_t e, int64_t p) {
    int64_t r = 1 % p;
    b %= p;
    while (e) {
        if (e & 1u) r = mulm(r, b, p);
        b = mulm(b, b, p);
        e >>= 1;
    }
    return r;
}

/* Contiguous split of [0, total) for thread tid of team. */
static void split(uint64_t total, int tid, int team, uint64_t *start, uint64_t *end) {
    const uint64_t per = total / (uint64_t)team, extra = total % (uint64_t)team;
    *start = (uint64_t)tid * per + ((uint64_t)tid < extra ? (uint64_t)tid : extra);
    *end = *start + per + ((uint64_t)tid < extra ? 1u : 0u);
}

/* Row product: plain 64-bit '%' when p*p fits (the reference's direct tier,
 * perm_c_omp.c:40-47), 128-bit products otherwise. */
static inline int64_t row_product(const int64_t *rs, int n, int64_t p) {
    int64_t prod = 1 % p;
    if (p <= INT64_C(3037000499)) {
        for (int i = 0; i < n; ++i) prod = (prod * rs[i]) % p;
    } else {
        for (int i = 0; i < n; ++i) prod = mulm(prod, rs[i], p);
    }
    return prod;
}

/* A: row-major n x n, already reduced into [0, p); p odd. */
int64_t oracle_glynn_mod_p(const int64_t *A, int n, int64_t p, int threads) {
    if (n <= 0) return 1 % p;
    if (n == 1) return A[0] % p;
    const uint64_t states = UINT64_C(1) << (n - 1);
    int nt = thread_count(threads);
    if ((uint64_t)nt > states) nt = (int)states;
    int64_t total = 0;
#pragma omp parallel num_threads(nt)
    {
        int tid = 0, team = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        team = omp_get_num_threads();
#endif
        uint64_t start, end;
        split(states, tid, team, &start, &end);
        int64_t rs[ORACLE_MAX_ORDER + 1];
        int64_t local = 0;
        if (start < end) {
            for (int i = 0; i < n; ++i) {
                int64_t s = 0;
                for (int j = 0; j < n; ++j) s = addm(s, A[i * n + j], p);
                rs[i] = s;
            }
            uint64_t g = start ^ (start >> 1);
            for (int j = 0; j < n - 1; ++j) {
                if ((g >> j) & 1u) {
                    for (int i = 0; i < n; ++i) rs[i] = subm(rs[i], addm(A[i * n + j], A[i * n + j], p), p);
                }
            }
            int negative = __builtin_popcountll(g) & 1;
            for (uint64_t s = start; s < end; ++s) {
                if (s != start) {
                    const int j = __builtin_ctzll(s);
                    g = s ^ (s >> 1);
                    const int sub = (int)((g >> j) & 1u);
                    for (int i = 0; i < n; ++i) {
                        const int64_t d = addm(A[i * n + j], A[i * n + j], p);
                        rs[i] = sub ? subm(rs[i], d, p) : addm(rs[i], d, p);
                    }
                    negative = !negative;
                }
                const int64_t prod = row_product(rs, n, p);
                local = negative ? subm(local, prod, p) : addm(local, prod, p);
            }
        }
#pragma omp critical(oracle_glynn)
        total = addm(total, local, p);
    }
    return mulm(total, powm((p >> 1) + 1, (uint64_t)(n - 1), p), p);
}

/* A: row-major n x n, already reduced into [0, p); any 0 < p < 2^63. */
int64_t oracle_ryser_mod_p(const int64_t *A, int n, int64_t p, int threads) {
    if (n <= 0) return 1 % p;
    if (n == 1) return A[0] % p;
    const uint64_t subsets = UINT64_C(1) << n;
    int nt = thread_count(threads);
    if ((uint64_t)nt > subsets) nt = (int)subsets;
    int64_t total = 0;
#pragma omp parallel num_threads(nt)
    {
        int tid = 0, team = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        team = omp_get_num_threads();
#endif
        uint64_t start, end;
        split(subsets, tid, team, &start, &end);
        if (start == 0) start = 1;
        int64_t rs[ORACLE_MAX_ORDER + 1];
        int64_t local = 0;
        if (start < end) {
            uint64_t g = start ^ (start >> 1);
            for (int i = 0; i < n; ++i) rs[i] = 0;
            for (int j = 0; j < n; ++j) {
                if ((g >> j) & 1u) {
                    for (int i = 0; i < n; ++i) rs[i] = addm(rs[i], A[i * n + j], p);
                }
            }
            int negative = __builtin_popcountll(g) & 1;
            for (uint64_t s = start; s < end; ++s) {
                if (s != start) {
                    const int j = __builtin_ctzll(s);
                    g = s ^ (s >> 1);
                    const int add = (int)((g >> j) & 1u);
                    for (int i = 0; i < n; ++i) rs[i] = add ? addm(rs[i], A[i * n + j], p) : subm(rs[i], A[i * n + j], p);
                    negative = !negative;
                }
                const int64_t prod = row_product(rs, n, p);
                local = negative ? subm(local, prod, p) : addm(local, prod, p);
            }
        }
#pragma omp critical(oracle_ryser)
        total = addm(total, local, p);
    }
    if (n & 1) total = total == 0 ? 0 : p - total;
    return total;
}

int64_t oracle_permanent_mod_p(const int64_t *A, int n, int64_t p, int threads) {
    if (n >= 4 && p > 1 && (p & 1) && p <= INT64_C(3037000499)) return oracle_glynn_mod_p(A, n, p, threads);
    return oracle_ryser_mod_p(A, n, p, threads);
}

/* ------------------------------------------------------------------ */
/* Binary (0/1) ratio-product Glynn.                                   */
/* ------------------------------------------------------------------ */

typedef struct {
    uint32_t p, nprime, r1; /* modulus, -p^-1 mod 2^32, 2^32 mod p */
} Mont32;

static Mont32 mont32_make(uint32_t p) {
    uint32_t x = p;
    for (int i = 0; i < 5; ++i) x *= 2u - p * x;
    Mont32 m = {p, 0u - x, (uint32_t)((UINT64_C(1) << 32) % p)};
    return m;
}

static inline uint32_t mont32_mul(uint32_t a, uint32_t b, Mont32 m) {
    const uint64_t t = (uint64_t)a * b;
    const uint32_t q = (uint32_t)t * m.nprime;
    uint64_t r = (t + (uint64_t)q * m.p) >> 32;
    return (uint32_t)(r >= m.p ? r - m.p : r);
}

static inline uint32_t mont32_in(uint64_t v, Mont32 m) { return (uint32_t)((v % m.p) * (uint64_t)m.r1 % m.p); }

static int degree_of(uint64_t mask) { return __builtin_popcountll(mask); }

/* Fixed column = densest (first on ties); the others in ascending degree,
 * stable. */
static void order_columns(const uint64_t *cols, int n, uint64_t *ordered) {
    int fixed = 0;
    for (int j = 1; j < n; ++j)
        if (degree_of(cols[j]) > degree_of(cols[fixed])) fixed = j;
    int c = 0;
    for (int j = 0; j < n; ++j)
        if (j != fixed) ordered[c++] = cols[j];
    for (int i = 1; i < c; ++i) {
        const uint64_t v = ordered[i];
        int k = i;
        while (k > 0 && degree_of(ordered[k - 1]) > degree_of(v)) {
            ordered[k] = ordered[k - 1];
            --k;
        }
        ordered[k] = v;
    }
}

/* Sum over Gray bits b of deg_b * 2^(n-2-b), degrees ascending; returned as
 * two 64-bit halves (hi, lo). */
void oracle_binary_update_cost(const uint64_t *cols, int n, uint64_t *hi, uint64_t *lo) {
    int deg[ORACLE_MAX_ORDER];
    for (int j = 0; j < n; ++j) deg[j] = degree_of(cols[j]);
    for (int i = 1; i < n; ++i) {
        const int v = deg[i];
        int k = i;
        while (k > 0 && deg[k - 1] > v) {
            deg[k] = deg[k - 1];
            --k;
        }
        deg[k] = v;
    }
    unsigned __int128 cost = 0;
    for (int b = 0; b < n - 1; ++b) cost += (unsigned __int128)deg[b] << (n - 2 - b);
    *hi = (uint64_t)(cost >> 64);
    *lo = (uint64_t)cost;
}

static unsigned __int128 update_cost(const uint64_t *cols, int n) {
    uint64_t hi, lo;
    oracle_binary_update_cost(cols, n, &hi, &lo);
    return ((unsigned __int128)hi << 64) | lo;
}

/* Generic modular product helpers for the ratio kernel: Montgomery form
 * when p < 2^31, plain 128-bit residues otherwise. */
typedef struct {
    int mont;
    Mont32 m32;
    uint64_t p;
} Arith;

static inline uint64_t ar_mul(uint64_t a, uint64_t b, const Arith *ar) {
    return ar->mont ? mont32_mul((uint32_t)a, (uint32_t)b, ar->m32)
                    : (uint64_t)(((unsigned __int128)a * b) % ar->p);
}
static inline uint64_t ar_in(uint64_t v, const Arith *ar) { return ar->mont ? mont32_in(v, ar->m32) : v % ar->p; }
static inline uint64_t ar_one(const Arith *ar) { return ar->mont ? ar->m32.r1 % ar->p : 1 % ar->p; }
static inline uint64_t ar_out(uint64_t v, const Arith *ar) { return ar->mont ? mont32_mul((uint32_t)v, 1u, ar->m32) : v; }

static uint64_t smod(int v, uint64_t p) { return v < 0 ? p - (uint64_t)(-v) : (uint64_t)v; }

/* A: row-major n x n bytes with entries 0/1; p odd prime, n < p < 2^63. */
int64_t oracle_binary_mod_p(const uint8_t *A, int n, int64_t p_signed, int threads) {
    const uint64_t p = (uint64_t)p_signed;
    if (n <= 0) return 1;
    uint64_t colm[ORACLE_MAX_ORDER] = {0}, rowm[ORACLE_MAX_ORDER] = {0};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (A[i * n + j]) {
                colm[j] |= UINT64_C(1) << i;
                rowm[i] |= UINT64_C(1) << j;
            }
    if (n == 1) return (int64_t)(colm[0] & 1u);
    for (int i = 0; i < n; ++i)
        if (!colm[i] || !rowm[i]) return 0;
    const uint64_t *input = update_cost(rowm, n) < update_cost(colm, n) ? rowm : colm;
    uint64_t cols[ORACLE_MAX_ORDER];
    order_columns(input, n, cols);

    Arith ar;
    ar.mont = p < (UINT64_C(1) << 31);
    ar.p = p;
    if (ar.mont) ar.m32 = mont32_make((uint32_t)p);
    /* ratio[dir][old + n] = (old +- 2) / old, or (old +- 2) when old == 0. */
    uint64_t ratio[2][2 * ORACLE_MAX_ORDER + 1];
    for (int old = -n; old <= n; ++old) {
        for (int dir = 0; dir < 2; ++dir) {
            const int next = old + (dir ? 2 : -2);
            uint64_t r = 0;
            if (next >= -n && next <= n) {
                if (old == 0) {
                    r = smod(next, p);
                } else {
                    const uint64_t inv = (uint64_t)powm((int64_t)smod(old, p), p - 2, (int64_t)p);
                    r = next == 0 ? inv : (uint64_t)mulm((int64_t)smod(next, p), (int64_t)inv, (int64_t)p);
                }
            }
            ratio[dir][old + n] = ar_in(r, &ar);
        }
    }
    const uint64_t states = UINT64_C(1) << (n - 1);
    int nt = thread_count(threads);
    if ((uint64_t)nt > states) nt = (int)states;
    uint64_t total = 0;
#pragma omp parallel num_threads(nt)
    {
        int tid = 0, team = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        team = omp_get_num_threads();
#endif
        uint64_t start, end;
        split(states, tid, team, &start, &end);
        int rs[ORACLE_MAX_ORDER] = {0};
        uint64_t local = 0;
        if (start < end) {
            for (int j = 0; j < n; ++j)
                for (uint64_t r = input[j]; r; r &= r - 1) rs[__builtin_ctzll(r)]++;
            uint64_t g = start ^ (start >> 1);
            for (uint64_t sel = g; sel; sel &= sel - 1)
                for (uint64_t r = cols[__builtin_ctzll(sel)]; r; r &= r - 1) rs[__builtin_ctzll(r)] -= 2;
            int zeros = 0;
            uint64_t prod = ar_one(&ar);
            for (int i = 0; i < n; ++i) {
                if (rs[i] == 0)
                    ++zeros;
                else
                    prod = ar_mul(prod, ar_in(smod(rs[i], p), &ar), &ar);
            }
            int negative = __builtin_popcountll(g) & 1;
            for (uint64_t s = start; s < end; ++s) {
                if (s != start) {
                    const int j = __builtin_ctzll(s);
                    g = s ^ (s >> 1);
                    const int adding = ((g >> j) & 1u) ? 0 : 1;
                    for (uint64_t r = cols[j]; r; r &= r - 1) {
                        const int row = __builtin_ctzll(r);
                        const int old = rs[row];
                        const int next = old + (adding ? 2 : -2);
                        prod = ar_mul(prod, ratio[adding][old + n], &ar);
                        if (old == 0) --zeros;
                        if (next == 0) ++zeros;
                        rs[row] = next;
                    }
                    negative = !negative;
                }
                if (zeros == 0) {
                    local = negative ? (uint64_t)subm((int64_t)local, (int64_t)prod, (int64_t)p)
                                     : (uint64_t)addm((int64_t)local, (int64_t)prod, (int64_t)p);
                }
            }
        }
#pragma omp critical(oracle_binary)
        total = (uint64_t)addm((int64_t)total, (int64_t)local, (int64_t)p);
    }
    const uint64_t scale = (uint64_t)powm((int64_t)((p >> 1) + 1), (uint64_t)(n - 1), (int64_t)p);
    const uint64_t t = ar_out(total, &ar);
    return (int64_t)(((unsigned __int128)t * scale) % p);
}

/* ------------------------------------------------------------------ */
/* Bipedal F_3 Ryser.                                                  */
/* ------------------------------------------------------------------ */

static inline void f3_add(uint64_t *m1, uint64_t *s1, uint64_t m2, uint64_t s2) {
    const uint64_t t = *m1 ^ *s1 ^ s2;
    const uint64_t mag = (m2 & t) | (*m1 ^ m2);
    *s1 = (m2 & t) ^ *s1;
    *m1 = mag;
}

static inline void f3_sub(uint64_t *m1, uint64_t *s1, uint64_t m2, uint64_t s2) {
    const uint64_t t = *s1 ^ s2;
    const uint64_t mag = (*m1 & t) | (*m1 ^ m2);
    *s1 = (*m1 & t) ^ (m2 ^ s2);
    *m1 = mag;
}

/* A: row-major n x n int64 (any values, reduced mod 3 here); n < 64. */
int oracle_mod3(const int64_t *A, int n, int threads) {
    if (n <= 0) return 1;
    uint64_t mags[ORACLE_MAX_ORDER] = {0}, sgns[ORACLE_MAX_ORDER] = {0};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const int v = (int)(((A[i * n + j] % 3) + 3) % 3);
            if (v) {
                mags[j] |= UINT64_C(1) << i;
                if (v == 2) sgns[j] |= UINT64_C(1) << i;
            }
        }
    if (n == 1) return (int)(((A[0] % 3) + 3) % 3);
    const uint64_t mask = (UINT64_C(1) << n) - 1;
    const uint64_t subsets = UINT64_C(1) << n;
    int nt = thread_count(threads);
    if ((uint64_t)nt > subsets) nt = (int)subsets;
    int64_t total = 0;
#pragma omp parallel num_threads(nt) reduction(+ : total)
    {
        int tid = 0, team = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        team = omp_get_num_threads();
#endif
        uint64_t start, end;
        split(subsets, tid, team, &start, &end);
        if (start == 0) start = 1;
        int64_t local = 0;
        if (start < end) {
            uint64_t g = start ^ (start >> 1), rm = 0, rsg = 0;
            for (int j = 0; j < n; ++j)
                if ((g >> j) & 1u) f3_add(&rm, &rsg, mags[j], sgns[j]);
            int negative = __builtin_popcountll(g) & 1;
            for (uint64_t s = start; s < end; ++s) {
                if (s != start) {
                    const int j = __builtin_ctzll(s);
                    g = s ^ (s >> 1);
                    if ((g >> j) & 1u)
                        f3_add(&rm, &rsg, mags[j], sgns[j]);
                    else
                        f3_sub(&rm, &rsg, mags[j], sgns[j]);
                    negative = !negative;
                }
                if ((rm & mask) == mask) {
                    const int odd = (__builtin_popcountll(rsg & mask) & 1) ^ negative;
                    local += odd ? -1 : 1;
                }
            }
        }
        total += local;
    }
    if (n & 1) total = -total;
    int r = (int)(total % 3);
    return r < 0 ? r + 3 : r;
}

/* Number of OpenMP threads a default parallel region would use. */
/* Coefficients of per(C + x L) mod p (odd p < 2^63), C and L row-major
 * n x n already reduced into [0, p): the coefficient-valued fixed-sign Glynn
 * sum of perm_polynomial.py:268-330 (row sums of both matrices, one affine
 * factor per row multiplied into a coefficient vector, signed accumulation,
 * division by 2^(n-1)), split over OpenMP threads by Gray state.  out
 * receives n + 1 coefficients in ascending degree. */
void oracle_affine_mod_p(const int64_t *C, const int64_t *L, int n, int64_t p, int threads, int64_t *out) {
    if (n <= 0) {
        out[0] = 1 % p;
        return;
    }
    if (n == 1) {
        out[0] = C[0] % p;
        out[1] = L[0] % p;
        return;
    }
    const uint64_t states = UINT64_C(1) << (n - 1);
    int nt = thread_count(threads);
    if ((uint64_t)nt > states) nt = (int)states;
    for (int d = 0; d <= n; ++d) out[d] = 0;
#pragma omp parallel num_threads(nt)
    {
        int tid = 0, team = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        team = omp_get_num_threads();
#endif
        uint64_t start, end;
        split(states, tid, team, &start, &end);
        int64_t cs[ORACLE_MAX_ORDER + 1], ls[ORACLE_MAX_ORDER + 1];
        int64_t prod[ORACLE_MAX_ORDER + 2], local[ORACLE_MAX_ORDER + 2];
        for (int d = 0; d <= n; ++d) local[d] = 0;
        if (start < end) {
            for (int i = 0; i < n; ++i) {
                int64_t a = 0, b = 0;
                for (int j = 0; j < n; ++j) {
                    a = addm(a, C[i * n + j], p);
                    b = addm(b, L[i * n + j], p);
                }
                cs[i] = a;
                ls[i] = b;
            }
            uint64_t g = start ^ (start >> 1);
            for (int j = 0; j < n - 1; ++j) {
                if ((g >> j) & 1u) {
                    for (int i = 0; i < n; ++i) {
                        cs[i] = subm(cs[i], addm(C[i * n + j], C[i * n + j], p), p);
                        ls[i] = subm(ls[i], addm(L[i * n + j], L[i * n + j], p), p);
                    }
                }
            }
            int negative = __builtin_popcountll(g) & 1;
            for (uint64_t s = start; s < end; ++s) {
                if (s != start) {
                    const int j = __builtin_ctzll(s);
                    g = s ^ (s >> 1);
                    const int sub = (int)((g >> j) & 1u);
                    for (int i = 0; i < n; ++i) {
                        const int64_t dc = addm(C[i * n + j], C[i * n + j], p);
                        const int64_t dl = addm(L[i * n + j], L[i * n + j], p);
                        cs[i] = sub ? subm(cs[i], dc, p) : addm(cs[i], dc, p);
                        ls[i] = sub ? subm(ls[i], dl, p) : addm(ls[i], dl, p);
                    }
                    negative = !negative;
                }
                /* prod = product over rows of (cs[i] + x ls[i]). */
                prod[0] = 1 % p;
                for (int i = 0; i < n; ++i) {
                    prod[i + 1] = mulm(prod[i], ls[i], p);
                    for (int d = i; d >= 1; --d) prod[d] = addm(mulm(prod[d], cs[i], p), mulm(prod[d - 1], ls[i], p), p);
                    prod[0] = mulm(prod[0], cs[i], p);
                }
                for (int d = 0; d <= n; ++d) local[d] = negative ? subm(local[d], prod[d], p) : addm(local[d], prod[d], p);
            }
        }
#pragma omp critical(oracle_affine)
        for (int d = 0; d <= n; ++d) out[d] = addm(out[d], local[d], p);
    }
    const int64_t inv = powm((p >> 1) + 1, (uint64_t)(n - 1), p);
    for (int d = 0; d <= n; ++d) out[d] = mulm(out[d], inv, p);
}

int oracle_max_threads(void) { return thread_count(0); }
