/*
 * ck_oracle.h -- CPU restatement of the reference cuckoo Monte Carlo trial.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * cuckoo_bench executable, the Python host mirror) may link or call this
 * code.  It is the checker that tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py compare the GPU path against.
 *
 * Every function names the reference file:line it restates.  Pinning (what
 * fixes each function independently of this restatement) is listed in the
 * header of ck_oracle.c.
 */
#ifndef CK_ORACLE_H
#define CK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* AES-128 round keys as 44 big-endian words (reference AesKey,
 * cuckoo_bench.cu:79). */
typedef struct { uint32_t w[44]; } ock_key;

/* Shape of one experiment.  mode: 1 = bgh25, 2 = drrt18 (journal encoding,
 * cuckoo_bench.cu:607).  k = candidate buckets per item (the reference is
 * fixed at 3), stash = stash slots (the reference has none). */
typedef struct {
    uint32_t mode, k, pnum, buckets, kick, stash, ell_i, input_mode;
    uint32_t half_bits, ell_d;
    uint64_t bucket_size, domain, mask, range_mask;
} ock_params;

/* Per-trial derived state (cuckoo_bench.cu:222-238). */
typedef struct {
    ock_key key;           /* PRP key schedule */
    uint64_t walk;         /* SplitMix64 state after the first advance */
    uint64_t randomness;   /* first walk output */
    uint64_t shift, step;  /* affine input map */
} ock_trial;

/* Per-trial counters.  probes = reads of a table slot by the insertion
 * chain; evals = bucket-function evaluations in the reference's order;
 * stashed = items that went to the stash. */
typedef struct {
    uint64_t probes, evals, stashed, placed;
} ock_stats;

#define OCK_MODE_BGH25 1u
#define OCK_MODE_DRRT18 2u
#define OCK_DERIVED 0xffffffffu

/* Math helpers of parameters() (cuckoo_bench.cu:483-498). */
unsigned ock_bit_width(uint64_t v);
float ock_original_log2(uint64_t n);

/* parameters() (cuckoo_bench.cu:500-526) plus the documented extension:
 * k choices, explicit bucket count, explicit kick bound, stash.  Pass
 * OCK_DERIVED for buckets/kick to get the reference's mode-derived values.
 * Returns 0, or a negative code for an invalid shape. */
int ock_parameters(uint32_t mode, uint32_t pnum, uint32_t ell_i, uint32_t k,
                   uint32_t buckets, uint32_t kick, uint32_t stash,
                   uint32_t input_mode, ock_params* out);

/* AES-128 and the key derivations. */
void ock_sbox(uint8_t out[256]);
void ock_expand_key(uint64_t lo, uint64_t hi, ock_key* out);          /* cuckoo_bench.cu:86-99 */
void ock_expand_words(const uint32_t words[4], ock_key* out);          /* cuckoo_bench.cu:136-149 */
void ock_aes_bytewise(const ock_key* k, const uint32_t in[4], uint32_t out[4]); /* :360-392 */
void ock_aes(const ock_key* k, const uint32_t in[4], uint32_t out[4]); /* AES-NI when present */
int ock_has_aesni(void);
void ock_batch_key(uint64_t seed, uint64_t batch, ock_key* out);      /* cuckoo_bench.cu:393-404 */
uint64_t ock_splitmix(uint64_t* state);                                /* cuckoo_core.hpp:110-114 */

/* Per-trial derivation and the Feistel PRP bucket function. */
void ock_trial_setup(const ock_key* batch, uint32_t offset, const ock_params* p, ock_trial* t);
uint64_t ock_element(const ock_trial* t, const ock_params* p, uint32_t ordinal);
uint32_t ock_bucket_direct(const ock_trial* t, const ock_params* p, uint32_t ordinal, uint32_t which);
void ock_build_rounds(const ock_trial* t, const ock_params* p, uint16_t* rounds);
uint32_t ock_bucket_cached(const ock_trial* t, const ock_params* p, const uint16_t* rounds,
                           uint32_t ordinal, uint32_t which);
/* map[o*k+w] = bucket(o, w) for every ordinal (cached when half_bits<=16). */
void ock_bucket_map(const ock_key* batch, uint32_t offset, const ock_params* p, uint32_t* map);

/* One trial.  bins has p->buckets entries and receives the final table,
 * entry = ordinal | choice<<30, empty = 0xffffffff.  rounds may be NULL
 * (direct AES path) or hold 8<<half_bits entries (cached path).
 * Returns 1 when the trial fails. */
int ock_trial_run(const ock_key* batch, uint32_t offset, const ock_params* p,
                  uint32_t* bins, uint16_t* rounds, ock_stats* st);

/* Trials [first, first+count) of one batch on `threads` host threads.
 * failed receives the failing offsets in ascending order (capacity count);
 * probes (optional) receives per-trial probe counts indexed by offset-first.
 * Returns the number of failures. */
uint32_t ock_run_range(const ock_key* batch, uint32_t first, uint32_t count,
                       const ock_params* p, int threads, int cached,
                       uint32_t* failed, uint32_t* probes, ock_stats* totals);

#ifdef __cplusplus
}
#endif
#endif
