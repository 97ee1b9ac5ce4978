/*
 * dogqc_amd synthetic data generation on the device (libdqcgen.so).
 *
 * No counterpart in the reference, whose tables come from TPC-H dbgen .tbl
 * files imported into mmdb column files by dbio (dogqc/dbio.py:117-253).
 * These kernels fill HBM-resident columns of the benchmark workloads at
 * sizes (tens of GB) that would take minutes to generate on the host and
 * copy over.  Every value is a pure function of (seed, column id, row), so
 * a column is reproducible, any row range can be regenerated or copied back
 * for checking, and the fill runs at HBM bandwidth.
 *
 * The random word of row i of column c is splitmix64's output function
 * applied to  seed ^ (c * 0x9E3779B97F4A7C15) + (i + 1) * 0xBF58476D1CE4E5B9.
 * All functions enqueue on `stream` (NULL = the null stream) and return 0 or
 * a negative DQC_E* code (include/dogqc_amd.h).
 */
#ifndef DOGQC_AMD_DATAGEN_H
#define DOGQC_AMD_DATAGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { DQC_GEN_I8 = 0, DQC_GEN_I32 = 1, DQC_GEN_U32 = 2, DQC_GEN_I64 = 3 };

/* out[i] = lo + floor(u * (hi - lo + 1)), u uniform in [0,1) from 32 random
 * bits; hi - lo < 2^32; element type DQC_GEN_*. */
int dqc_gen_uniform(void* out, int32_t dtype, int64_t n, uint64_t seed, uint32_t column, int64_t lo, int64_t hi,
                    void* stream);

/* out[i] = (double)uniform(lo, hi) / divisor  (money in cents, hundredths) */
int dqc_gen_decimal(double* out, int64_t n, uint64_t seed, uint32_t column, int64_t lo, int64_t hi,
                    double divisor, void* stream);

/* out[i] = table[uniform(0, ntable - 1)]  (dates from a day table) */
int dqc_gen_pick_u32(uint32_t* out, int64_t n, uint64_t seed, uint32_t column, const uint32_t* table,
                     int64_t ntable, void* stream);

/* out[i] = (a * i + b) mod n: a permutation of 0..n-1 when gcd(a, n) = 1
 * (a, b < n < 2^31). */
int dqc_gen_affine_perm(int32_t* out, int64_t n, uint64_t a, uint64_t b, void* stream);

/* Foreign keys whose every prefix stays inside the referenced table's
 * prefix: out[i] = (a * r + b) mod nkeys, the key of row r of a table whose
 * keys are dqc_gen_affine_perm(nkeys, a, b), with r uniform over that
 * table's first ceil(nkeys * (i + 1) / n) rows.  A query that reads the
 * first f * n rows of this column and the first f * nkeys rows of the
 * referenced table then finds every key, so join selectivities are the same
 * at every prefix size (low rows are referenced more often: row r by about
 * (n / nkeys) * ln(nkeys / r) rows).  nkeys < 2^31, a, b < nkeys. */
int dqc_gen_prefix_key(int32_t* out, int64_t n, uint64_t seed, uint32_t column, int64_t nkeys, uint64_t a,
                       uint64_t b, void* stream);

/* out[i] = values[r], r the first index with cdf[r] > u (u uniform in
 * [0,1) from 53 random bits): inverse-CDF sampling of a discrete
 * distribution (Zipf) given its cumulative table. */
int dqc_gen_inverse_cdf(int32_t* out, int64_t n, uint64_t seed, uint32_t column, const double* cdf,
                        const int32_t* values, int32_t nvalues, void* stream);

/* STRING column of fixed-width rows: offs[i] = i * width (i = 0..n) and row
 * i = `words` vocabulary words chosen uniformly (word j of row i from the
 * random word of column `column + j`), joined by single spaces and padded
 * with spaces.  The vocabulary is given as a char heap with offsets
 * voffs[0..nvocab]; every row must fit `width`. */
int dqc_gen_text(uint64_t* offs, char* chars, int64_t n, int32_t width, uint64_t seed, uint32_t column,
                 const char* vchars, const int32_t* voffs, int32_t nvocab, int32_t words, void* stream);

#ifdef __cplusplus
}
#endif
#endif
