/*
 * pz_oracle.h -- CPU restatement of phantom-zone's bootstrapping (PBS) path.
 *
 * ORACLE / TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg do, and only as the checker (or as the timed CPU baseline).
 *
 * Every function follows one reference routine loop for loop (same Shoup /
 * Barrett formulas, same lazy ranges, same stage order, same bucket order,
 * same f64 moddown); the citation is on each definition in pz_oracle.c.
 *
 * Parity status per function (how each is pinned) is listed in the header
 * comment of pz_oracle.c.
 */
#ifndef PZ_ORACLE_H
#define PZ_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Parameter set (src/bool/parameters.rs:96-167). */
typedef struct pzo_params {
    uint64_t rlwe_q;        /* Q (prime, non-native)                          */
    uint32_t rlwe_n;        /* N (ring size)                                  */
    uint32_t lwe_n;         /* n (LWE dimension)                              */
    uint64_t lwe_q;         /* q_ks (power of two)                            */
    uint32_t br_q;          /* blind-rotation modulus                         */
    uint32_t lwe_logb, lwe_d;               /* LWE key switch decomposer      */
    uint32_t rlrg_logb, rlrg_da, rlrg_db;   /* RLWE x RGSW decomposer (a, b)  */
    uint32_t auto_logb, auto_d;             /* automorphism decomposer        */
    uint32_t g, w;                          /* generator, window              */
    uint32_t lwe_secret_gaussian;           /* 1: error dist, 0: ternary      */
    uint32_t uitos_logb, uitos_d;           /* u_j -> s RLWE key switch (NI)  */
    uint32_t rgrg_logb, rgrg_da, rgrg_db;   /* RGSW x RGSW (multi-party)      */
} pzo_params;

/* Named parameter sets: 0 NI_2P, 1 NI_4P_HB_FR, 2 NI_8P, 3 SP_TEST_BOOL_PARAMS,
 * 4 I_2P_LB_SR, 5 I_4P, 6 I_8P, 7 NI_4P_LB_SR. Returns 0 on success. */
int pzo_params_named(int which, pzo_params *out);

/* ---------------- ChaCha (rand_chacha 0.3.1 state layout) ---------------- */
void pzo_chacha_block(const uint8_t key[32], uint64_t counter, uint64_t stream,
                      int rounds, uint32_t out[16]);
typedef struct pzo_rng pzo_rng;
pzo_rng *pzo_rng_new(const uint8_t seed[32]);
void pzo_rng_free(pzo_rng *r);
uint32_t pzo_rng_next_u32(pzo_rng *r);
uint64_t pzo_rng_next_u64(pzo_rng *r);
void pzo_rng_fill_bytes(pzo_rng *r, uint8_t *out, size_t n);
uint64_t pzo_rng_uniform(pzo_rng *r, uint64_t range);   /* [0, range) */
int64_t pzo_rng_gauss_round(pzo_rng *r);
int64_t pzo_rng_gauss_trunc(pzo_rng *r);
void pzo_ternary_hw(pzo_rng *r, int32_t *out, uint32_t n, uint32_t hw);

/* ---------------- scalar arithmetic ---------------- */
uint64_t pzo_shoup_repr(uint64_t v, uint64_t q);
uint64_t pzo_shoup_mul(uint64_t a, uint64_t b, uint64_t b_shoup, uint64_t q);
uint64_t pzo_mul_mod_fast(uint64_t q, uint64_t a, uint64_t b);
uint64_t pzo_mul_mod_fast_lazy(uint64_t q, uint64_t a, uint64_t b);
uint64_t pzo_mod_exponent(uint64_t a, uint64_t b, uint64_t q);
void pzo_shoup_matrix_fma(uint64_t q, uint64_t *out, const uint64_t *a,
                          const uint64_t *a_shoup, const uint64_t *b,
                          uint32_t d, uint32_t n);

/* ---------------- NTT (src/ntt.rs) ---------------- */
typedef struct pzo_ntt pzo_ntt;
pzo_ntt *pzo_ntt_new(uint64_t q, uint32_t n);
void pzo_ntt_free(pzo_ntt *t);
uint64_t pzo_ntt_psi(const pzo_ntt *t);
/* tables: [psi_bo, psi_bo_shoup, psi_inv_bo, psi_inv_bo_shoup] each n, then
 * n_inv, n_inv_shoup  (4n + 2 words) */
void pzo_ntt_tables(const pzo_ntt *t, uint64_t *out);
void pzo_ntt_forward_lazy(const pzo_ntt *t, uint64_t *v);
void pzo_ntt_forward(const pzo_ntt *t, uint64_t *v);
void pzo_ntt_backward_lazy(const pzo_ntt *t, uint64_t *v);
void pzo_ntt_backward(const pzo_ntt *t, uint64_t *v);
void pzo_ntt_rows(const pzo_ntt *t, int which, uint64_t *v, uint32_t rows);
void pzo_negacyclic_mul_schoolbook(uint64_t q, const uint64_t *a,
                                   const uint64_t *b, uint64_t *out, uint32_t n);

/* ---------------- decomposer (src/decomposer.rs) ---------------- */
/* digits out: count x d, row-major per value */
void pzo_decompose_iter(uint64_t q, uint32_t logb, uint32_t d,
                        const uint64_t *v, uint64_t *digits, size_t count);
void pzo_decompose_to_vec(uint64_t q, uint32_t logb, uint32_t d,
                          const uint64_t *v, uint64_t *digits, size_t count);
uint64_t pzo_round_value(uint64_t v, uint32_t ignore_bits);
uint64_t pzo_recompose(uint64_t q, uint32_t logb, uint32_t d, const uint64_t *limbs);
uint32_t pzo_decomposer_ignore_bits(uint64_t q, uint32_t logb, uint32_t d);

/* ---------------- evaluator (BoolEvaluator + BoolPbsInfo) ---------------- */
typedef struct pzo_eval pzo_eval;
pzo_eval *pzo_eval_new(const pzo_params *p);
void pzo_eval_free(pzo_eval *e);
const pzo_ntt *pzo_eval_ntt(const pzo_eval *e);
void pzo_eval_dlog_map(const pzo_eval *e, uint32_t *out);            /* br_q */
int64_t pzo_eval_auto_element(const pzo_eval *e, uint32_t k);
void pzo_eval_auto_map(const pzo_eval *e, uint32_t k, uint32_t *idx, uint8_t *sign);
/* gate ids: 0 NAND, 1 AND, 2 OR, 3 NOR, 4 XOR, 5 XNOR */
void pzo_eval_test_vec(const pzo_eval *e, uint32_t gate, uint64_t *out);  /* br_q/2 */
void pzo_generate_auto_map(uint32_t ring_size, int64_t k, uint32_t *idx, uint8_t *sign);

/* key sizes in u64 words */
typedef struct pzo_key_sizes {
    size_t rgsw_rows;       /* 2*da + 2*db */
    size_t rgsw_words;      /* lwe_n * rgsw_rows * N */
    size_t auto_rows;       /* 2*d_auto */
    size_t auto_words;      /* (w+1) * auto_rows * N */
    size_t ksk_rows;        /* N * lwe_d */
    size_t ksk_words;       /* ksk_rows * (lwe_n + 1) */
} pzo_key_sizes;
void pzo_key_sizes_get(const pzo_eval *e, pzo_key_sizes *out);

/* In-memory server key, reference layout: rgsw[i][row][N] and auto[k][row][N]
 * in evaluation domain with Shoup twins; lwe_ksk[row][n+1]. Caller owns. */
typedef struct pzo_key {
    uint64_t *rgsw, *rgsw_shoup;
    uint64_t *auto_key, *auto_shoup;
    uint64_t *lwe_ksk;
} pzo_key;

/* seeded key generation (the oracle's own procedure, see pz_oracle.c) */
void pzo_keygen_secrets(const pzo_eval *e, const uint8_t seed[32], uint32_t parties,
                        int32_t *s_rlwe, int32_t *z_lwe);
void pzo_keygen_server(const pzo_eval *e, const uint8_t seed[32],
                       const int32_t *s_rlwe, const int32_t *z_lwe, pzo_key *key);
void pzo_key_shoup(const pzo_eval *e, pzo_key *key);

/* runtime ops (src/rgsw/runtime.rs, src/lwe.rs, src/pbs.rs) */
void pzo_rlwe_by_rgsw_shoup(const pzo_eval *e, uint64_t *rlwe, const uint64_t *rgsw,
                            const uint64_t *rgsw_shoup, int is_trivial);
void pzo_rlwe_by_rgsw(const pzo_eval *e, uint64_t *rlwe, const uint64_t *rgsw,
                      int is_trivial);
void pzo_rlwe_auto_shoup(const pzo_eval *e, uint64_t *rlwe, const uint64_t *ksk,
                         const uint64_t *ksk_shoup, uint32_t k, int is_trivial);
void pzo_rlwe_auto(const pzo_eval *e, uint64_t *rlwe, const uint64_t *ksk,
                   uint32_t k, int is_trivial);
void pzo_lwe_key_switch(const pzo_eval *e, uint64_t *lwe_out, const uint64_t *lwe_in,
                        const uint64_t *ksk);
uint64_t pzo_moddown(const pzo_eval *e, uint64_t v);
uint32_t pzo_mod_switch_odd(double v, double from_q, double to_q);
void pzo_sample_extract(const pzo_eval *e, uint64_t *lwe_out, const uint64_t *rlwe,
                        uint32_t index);
void pzo_monomial_mul(uint64_t q, const uint64_t *p_in, uint64_t *p_out,
                      uint32_t mon_exp, int mon_sign, uint32_t ring_size);

/* PBS statistics: counts of each operation in the blind rotation */
typedef struct pzo_pbs_stats {
    uint32_t ext_products;
    uint32_t autos;
    uint32_t trivial_autos;
} pzo_pbs_stats;
void pzo_pbs(const pzo_eval *e, const pzo_key *key, uint32_t gate, uint64_t *lwe,
             pzo_pbs_stats *stats);
/* Blind-rotation schedule of one PBS input (after key switch): a list of
 * ops, each (kind, arg, trivial) packed as kind | arg<<8 | trivial<<31:
 * kind 0 = RLWE x RGSW with LWE index arg, kind 1 = auto with key index arg.
 * Returns the number of ops written (cap = max ops). */
uint32_t pzo_br_schedule(const pzo_eval *e, const uint64_t *lwe_ks, uint32_t *ops,
                         uint32_t cap);
/* The gate pre-processing + PBS (src/bool/evaluator.rs:2118-2250). */
void pzo_gate(const pzo_eval *e, const pzo_key *key, uint32_t gate, uint64_t *c0,
              const uint64_t *c1);
void pzo_not(const pzo_eval *e, uint64_t *c0);
/* batched PBS timing helper for the CPU baseline: returns seconds */
double pzo_pbs_batch_threads(const pzo_eval *e, const pzo_key *key, uint32_t gate,
                             uint64_t *lwes, uint32_t count, uint32_t threads);
/* gates[i] of (c0[i], c1[i]) into c0[i] for count items on `threads` threads
 * (the same words as pzo_gate on each); returns seconds */
double pzo_gates_batch_threads(const pzo_eval *e, const pzo_key *key, const uint32_t *gates,
                               uint64_t *c0, const uint64_t *c1, uint32_t count,
                               uint32_t threads);

/* encryption helpers (tests) */
/* ---- input path of the non-interactive protocol ---- */
/* party RLWE secret u_j (ternary, Hamming weight N/2), the oracle's own
 * seeded draw */
void pzo_keygen_user_secret(const pzo_eval *e, const uint8_t seed[32], int32_t *u);
/* Ksk(u -> s), [2d][N] evaluation domain: rows A_j uniform, then
 * B_j = A_j*s + e_j + 2^(logb*j + ignore_bits) * u (the oracle's own stream:
 * per j, N uniform then N noise draws) */
void pzo_keygen_rlwe_ksk(const pzo_eval *e, const uint8_t seed[32], const int32_t *u,
                         const int32_t *s, uint64_t *ksk);
/* rlwe_key_switch, src/rgsw/runtime.rs:1001-1063: rlwe_in [2][N] under u to
 * out [2][N] under s with ksk [2d][N] (evaluation domain) and its Shoup rows */
void pzo_rlwe_key_switch(const pzo_eval *e, const uint64_t *rlwe_in, const uint64_t *ksk,
                         const uint64_t *ksk_shoup, uint64_t *out);
/* bool batch under u as ceil(count/N) RLWEs [.][2][N]
 * (src/bool/ni_mp_api.rs:294-352, seeded_secret_key_encrypt_rlwe
 * src/rgsw/keygen.rs:439-481); a from the stream of seed_a, noise from the
 * stream of seed_e, both derived from seed (first and second 32 bytes) */
void pzo_encrypt_rlwe_bools(const pzo_eval *e, const uint8_t seed[32], const int32_t *u,
                            const uint8_t *bits, size_t count, uint64_t *out);

/* rgsw_by_rgsw_inplace, src/rgsw/runtime.rs:883-985: rgsw0 [2da+2db][N]
 * (RLWE x RGSW decomposer, coefficient domain, overwritten) times rgsw1
 * [2da'+2db'][N] (RGSW x RGSW decomposer, evaluation domain) */
void pzo_rgsw_by_rgsw(const pzo_eval *e, uint64_t *rgsw0, const uint64_t *rgsw1);

/* ---- non-interactive server key: shares and aggregation (the seed tree of
 * NonInteractiveMultiPartyCrs, src/bool/evaluator.rs:97-160; the oracle's own
 * noise streams) ---- */
void pzo_puncture(const uint8_t seed[32], uint32_t times, uint8_t out[32]);
uint32_t pzo_ni_leader(uint32_t l, uint32_t total, uint32_t lwe_n);
/* party self's u->s parts: ksk_b [d][N] (a_{i,self}*s_j + e + beta_i*u_j) and
 * zero encryptions a_{i,o}*s_j + e for every other party o, ascending,
 * [(total-1)][d][N]; coefficient domain (src/multi_party.rs:186-291) */
void pzo_ni_share_uitos(const pzo_eval *e, const uint8_t crs[32], const uint8_t noise_seed[32],
                        uint32_t self, uint32_t total, const int32_t *s, const int32_t *u,
                        uint64_t *ksk_b, uint64_t *zero_others);
/* party self's RGSW parts for LWE indices [l0, l0 + count)
 * (src/bool/evaluator.rs:1697-1937): zero [count][da][N] (a_{i,l}*s_j + e for
 * i in [dmax - da, dmax)) and cts [count][dmax][N] (u_j*a_{i,l} + e +
 * beta*X^{emb*z_j[l]}; leader indices fill only the last max(da', db') rows,
 * the rest are zero); RGSW x RGSW counts (da, db), dmax = max(da, db) */
void pzo_ni_share_rgsw(const pzo_eval *e, const uint8_t crs[32], const uint8_t noise_seed[32],
                       uint32_t self, uint32_t total, const int32_t *s, const int32_t *u,
                       const int32_t *z, uint32_t l0, uint32_t count, uint64_t *zero,
                       uint64_t *cts);
/* Ksk(u_user -> s) [2d][N] evaluation domain from the user's ksk_b and the
 * zero encryptions for the user from every other party (ascending)
 * (src/bool/evaluator.rs:1266-1337) */
void pzo_ni_uitos_ksk(const pzo_eval *e, const uint8_t crs[32], uint32_t total, uint32_t user,
                      const uint64_t *ksk_b, const uint64_t *zero_for_user, uint64_t *ksk);
/* produce_rgsw_ciphertext_from_ni_rgsw (src/bool/evaluator.rs:480-608) for
 * index l: leader = 1 gives [2da'+2db'][N] coefficient domain, leader = 0
 * gives [2da+2db][N] evaluation domain */
void pzo_ni_produce(const pzo_eval *e, const uint8_t crs[32], uint32_t total, uint32_t l,
                    const uint64_t *ksk, const uint64_t *zero_l, const uint64_t *ct, int leader,
                    uint64_t *out);
/* RGSW part of the aggregation: out [count][2da'+2db'][N] coefficient domain;
 * ksks [total][2d][N], zero [total][count][da][N], cts [total][count][dmax][N] */
void pzo_ni_aggregate_rgsw(const pzo_eval *e, const uint8_t crs[32], uint32_t total,
                           const uint64_t *ksks, uint32_t l0, uint32_t count,
                           const uint64_t *zero, const uint64_t *cts, uint64_t *out);

void pzo_encrypt_lwe(const pzo_eval *e, const uint8_t seed[32], const int32_t *s,
                     uint32_t dim, uint64_t m, uint64_t *out);
uint64_t pzo_decrypt_lwe(const pzo_eval *e, const uint64_t *lwe, const int32_t *s,
                         uint32_t dim);
int pzo_decode_bool(const pzo_eval *e, uint64_t m);
uint64_t pzo_true_el(const pzo_eval *e);
uint64_t pzo_false_el(const pzo_eval *e);

/* non-interactive automorphism-key and LWE-KSK shares (b parts, coefficient
 * domain / mod q_ks), their sums, and the expansion of a seeded server key
 * (coefficient-domain RGSWs + summed b parts) into an evaluation-domain key */
void pzo_ni_share_auto_ksk(const pzo_eval *e, const uint8_t crs[32], const uint8_t noise_seed[32],
                           const int32_t *s, const int32_t *z, uint64_t *auto_b, uint64_t *ksk_b);
void pzo_ni_aggregate_auto_ksk(const pzo_eval *e, uint32_t total, const uint64_t *auto_shares,
                               const uint64_t *ksk_shares, uint64_t *auto_b, uint64_t *ksk_b);
void pzo_ni_expand_server_key(const pzo_eval *e, const uint8_t crs[32], const uint64_t *rgsw,
                              const uint64_t *auto_b, const uint64_t *ksk_b, pzo_key *key);

#ifdef __cplusplus
}
#endif
#endif
