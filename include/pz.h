/*
 * pz.h -- C ABI of the MI355X (gfx950) bootstrapping engine for
 * phantom-zone's boolean FHE.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns
 * an int status (PZ_OK = 0, negative on error; pz_last_error() gives the
 * message) and never aborts or throws across the boundary.  Buffers named
 * d_* are device pointers (HBM) and all such calls are ordered on `stream`
 * (a hipStream_t, NULL = the context's stream).  Buffers named h_* are host
 * pointers.  A context is single-stream and not thread-safe, like the
 * reference's thread-local evaluator; a key handle is immutable and may be
 * shared by contexts on the same device, like the reference's OnceLock key.
 *
 * Each declaration cites the reference interface it replaces
 * (gausslabs/phantom-zone, Rust).  Layouts are the reference's: an RLWE
 * ciphertext is [a | b] (2 x N u64), an LWE ciphertext is [b, a_1..a_n],
 * matrices are row-major rows of N words.
 */
#ifndef PZ_H
#define PZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PZ_OK 0
#define PZ_ERR_ARG (-1)
#define PZ_ERR_HIP (-2)
#define PZ_ERR_UNSUPPORTED (-3)
#define PZ_ERR_STATE (-4)

/* BoolParameters (src/bool/parameters.rs:96-167). */
typedef struct pz_params {
    uint64_t rlwe_q;      /* Q, prime, < 2^54, = 1 mod 2N             */
    uint32_t rlwe_n;      /* N = 512 .. 4096                          */
    uint32_t lwe_n;       /* n                                        */
    uint64_t lwe_q;       /* q_ks, power of two <= 2^17              */
    uint32_t br_q;        /* blind-rotation modulus                   */
    uint32_t lwe_logb, lwe_d;
    uint32_t rlrg_logb, rlrg_da, rlrg_db;
    uint32_t auto_logb, auto_d;
    uint32_t g, w;
    uint32_t lwe_secret_gaussian;
    uint32_t uitos_logb, uitos_d; /* u_j -> s RLWE key switch (NI sets; 0 = none) */
    uint32_t rgrg_logb, rgrg_da, rgrg_db; /* RGSW x RGSW (multi-party sets; 0 = none) */
} pz_params;

typedef struct pz_ctx pz_ctx;
typedef struct pz_key pz_key;

/* Gate ids of the BooleanGates seam (src/bool/mod.rs:22-71). */
enum { PZ_NAND = 0, PZ_AND = 1, PZ_OR = 2, PZ_NOR = 3, PZ_XOR = 4, PZ_XNOR = 5 };

/* Named constant parameter sets, src/bool/parameters.rs:526-726:
 * "NI_2P", "NI_4P_HB_FR", "NI_8P", "SP_TEST", "I_2P_LB_SR", "I_4P", "I_8P",
 * "NI_4P_LB_SR". */
int pz_params_named(const char *name, pz_params *out);

const char *pz_last_error(void);

/* ---- context: BoolEvaluator::new (src/bool/evaluator.rs:651-784) with
 * NttInit::new (src/ntt.rs:388-394), ModInit::new
 * (src/backend/modulus_u64.rs:18-43) and the decomposers
 * (src/decomposer.rs:158-178). */
int pz_ctx_create(const pz_params *p, int device, pz_ctx **out);
int pz_ctx_destroy(pz_ctx *ctx);
/* psi (the 2N-th root NttBackendU64 derives from ChaCha8 seed 0,
 * src/ntt.rs:323-328), and the number of compute units / resident PBS work
 * groups the context uses. */
int pz_ctx_info(pz_ctx *ctx, uint64_t *psi, uint32_t *num_cu, uint32_t *pbs_groups_per_cu);
/* Host copy of the NTT tables: psi_bo, psi_bo_shoup, psi_inv_bo,
 * psi_inv_bo_shoup (N each), n_inv, n_inv_shoup (src/ntt.rs:310-386). */
int pz_ctx_ntt_tables(pz_ctx *ctx, uint64_t *h_out);
int pz_ctx_stream(pz_ctx *ctx, void **stream);

/* ---- device memory helpers ---- */
int pz_malloc(pz_ctx *ctx, size_t bytes, void **d_ptr);
int pz_free(pz_ctx *ctx, void *d_ptr);
int pz_memcpy_htod(pz_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int pz_memcpy_dtoh(pz_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
int pz_sync(pz_ctx *ctx);

/* ---- NTT: Ntt::{forward_lazy, forward, backward_lazy, backward}
 * (src/ntt.rs:16-22, impl :396-442), in place on nrows rows of N words.
 * Output ranges as the reference: forward [0,q), forward_lazy [0,2q),
 * backward [0,q), backward_lazy [0,2q). */
int pz_ntt_forward(pz_ctx *ctx, uint64_t *d_rows, size_t nrows, void *stream);
int pz_ntt_forward_lazy(pz_ctx *ctx, uint64_t *d_rows, size_t nrows, void *stream);
int pz_ntt_backward(pz_ctx *ctx, uint64_t *d_rows, size_t nrows, void *stream);
int pz_ntt_backward_lazy(pz_ctx *ctx, uint64_t *d_rows, size_t nrows, void *stream);

/* ---- ShoupMatrixFMA::shoup_matrix_fma (src/backend/mod.rs:137-141, impl
 * src/backend/modulus_u64.rs:231-264): d_out[j] += sum_r a[r][j]*b[r][j],
 * a/a_shoup/b are d rows of n words; d_out stays in [0,2q). */
int pz_shoup_matrix_fma(pz_ctx *ctx, uint64_t *d_out, const uint64_t *d_a,
                        const uint64_t *d_a_shoup, const uint64_t *d_b, uint32_t d, size_t n,
                        void *stream);

/* ---- Decomposer::decompose_iter over `count` values, digits scattered as
 * decompose_r does (src/rgsw/runtime.rs:382-399): d_digits[r*count + i] is
 * digit r of value i.  which: 0 RLWE x RGSW part a, 1 part b, 2 auto,
 * 3 LWE key switch (src/bool/parameters.rs decomposer getters). */
int pz_decompose(pz_ctx *ctx, int which, const uint64_t *d_in, uint64_t *d_digits,
                 size_t count, void *stream);

/* ---- server key: NonInteractiveServerKeyEvaluationDomain +
 * ShoupNonInteractiveServerKeyEvaluationDomain (src/bool/keys.rs:851-1005,
 * :1068-1173) implementing PbsKey (src/pbs.rs:16-27).  Host inputs in the
 * reference layout: h_rgsw[n][2da+2db][N] and h_auto[w+1][2*d_auto][N] in the
 * evaluation domain, h_lwe_ksk[N*d_ks][n+1].  Shoup twins are computed on the
 * device. */
int pz_key_upload(pz_ctx *ctx, const uint64_t *h_rgsw, const uint64_t *h_auto,
                  const uint64_t *h_lwe_ksk, pz_key **out);
int pz_key_free(pz_key *key);
/* Device bytes held by the key (serving admission by HBM footprint). */
int pz_key_bytes(pz_key *key, size_t *bytes);

/* Copy a key back to the host in the pz_key_upload layout (h_lwe_ksk as u64
 * rows [N*d_ks][n+1]).  Any pointer may be NULL to skip that part. */
int pz_key_download(pz_key *key, uint64_t *h_rgsw, uint64_t *h_auto, uint64_t *h_lwe_ksk);

/* ---- seeded client-side key generation (the library's own seeded
 * procedure; structure of ClientKey::sk_rlwe / sk_lwe, src/bool/keys.rs:
 * 75-154, secret_key_encrypt_rgsw, seeded_auto_key_gen, seeded_lwe_ksk_keygen,
 * src/rgsw/keygen.rs:68-430, src/lwe.rs:46-101).  The ideal secret is the sum
 * of `parties` parties' secrets: h_s[N] (RLWE, ternary per party) and h_z[n]
 * (LWE).  pz_keygen_server builds the evaluation-domain server key on the
 * device under that ideal secret. */
int pz_keygen_secrets(pz_ctx *ctx, const uint8_t *seed32, uint32_t parties, int32_t *h_s,
                      int32_t *h_z);
int pz_keygen_server(pz_ctx *ctx, const uint8_t *seed32, const int32_t *h_s, const int32_t *h_z,
                     pz_key **out);

/* ---- client-side boolean LWE encryption / decryption under h_s (dim
 * coefficients): encrypt_lwe / decrypt_lwe (src/lwe.rs:104-162) with the
 * bool encoding +-Q/8 and decode (src/bool/evaluator.rs:209-258).
 * Ciphertext i uses the seed (seed32 + i) as a 256-bit little-endian integer.
 * pz_decrypt_bools returns PZ_ERR_STATE if any value decodes outside {0,1}. */
int pz_encrypt_bools(pz_ctx *ctx, const int32_t *h_s, uint32_t dim, const uint8_t *seed32,
                     const uint8_t *h_bits, size_t count, uint64_t *h_out);
int pz_decrypt_bools(pz_ctx *ctx, const int32_t *h_s, uint32_t dim, const uint64_t *h_cts,
                     size_t count, uint8_t *h_bits);

/* ---- rlwe_by_rgsw_shoup (src/rgsw/runtime.rs:754-855) on B RLWE
 * ciphertexts d_rlwe[B][2][N]; RLWE b is multiplied by RGSW(X^{z_i}) with
 * i = d_rgsw_idx[b]; d_trivial[b] != 0 skips part a (trivial RLWE).  An index
 * outside [0, lwe_n) leaves its RLWE unchanged (the reference panics). */
int pz_rlwe_by_rgsw(pz_ctx *ctx, pz_key *key, uint64_t *d_rlwe, const uint32_t *d_rgsw_idx,
                    const uint8_t *d_trivial, size_t B, void *stream);

/* ---- rlwe_auto_shoup (src/rgsw/runtime.rs:530-643): automorphism with key
 * index d_auto_idx[b] in [0, w] (0 = -g, k = g^k; src/bool/evaluator.rs:
 * 730-743); an index outside [0, w] leaves its RLWE unchanged. */
int pz_rlwe_auto(pz_ctx *ctx, pz_key *key, uint64_t *d_rlwe, const uint32_t *d_auto_idx,
                 const uint8_t *d_trivial, size_t B, void *stream);

/* ---- lwe_key_switch (src/lwe.rs:14-44): d_in[B][N+1] (mod q_ks) ->
 * d_out[B][n+1] (mod q_ks). */
int pz_lwe_key_switch(pz_ctx *ctx, pz_key *key, const uint64_t *d_in, uint64_t *d_out,
                      size_t B, void *stream);

/* ---- pbs (src/pbs.rs:104-233): B independent bootstraps of
 * d_lwe[B][N+1] (mod Q, overwritten), test vector d_gate[b] (gate id).
 * Runs moddown + key switch + persistent blind rotation + sample extract. */
int pz_pbs(pz_ctx *ctx, pz_key *key, const uint32_t *d_gate, uint64_t *d_lwe, size_t B,
           void *stream);

/* ---- BooleanGates batch (src/bool/evaluator.rs:2136-2250): d_out[b] =
 * gate d_ops[b] of (d_c0[b], d_c1[b]).  The first call for a buffer set and
 * batch size captures the three launches in a hipGraph; later calls with the
 * same buffers replay it. */
int pz_gates(pz_ctx *ctx, pz_key *key, const uint32_t *d_ops, const uint64_t *d_c0,
             const uint64_t *d_c1, uint64_t *d_out, size_t B, void *stream);
/* pz_gates that allocates nothing.  No counterpart in the reference (its
 * evaluator allocates per gate): for callers that cannot call hipMalloc --
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation.  The pipeline scratch is a fixed 16 MiB device array per GPU,
 * held until the process ends, so B is at most what it holds (576 at NI_2P,
 * 512 at NI_8P; larger B: PZ_ERR_ARG with the limit in the message).  The
 * launches are eager (no captured graph) and give the same words as
 * pz_gates.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (from any context or thread) race. */
int pz_gates_noalloc(pz_ctx *ctx, pz_key *key, const uint32_t *d_ops, const uint64_t *d_c0,
                     const uint64_t *d_c1, uint64_t *d_out, size_t B, void *stream);
/* not_inplace (src/bool/evaluator.rs:2236-2239) on B ciphertexts. */
int pz_not(pz_ctx *ctx, uint64_t *d_c, size_t B, void *stream);

/* ---- FheUint8 circuits (src/shortint/ops.rs:5-356, operators of
 * src/shortint/mod.rs:30-293) built as gate DAGs over the BooleanGates seam
 * and evaluated level by level: all ready gates of a level, across every
 * operation in the circuit, are bootstrapped in one batched launch.  Wires
 * are ciphertext ids; bits are little endian (8 wires per FheUint8). */
typedef struct pz_circuit pz_circuit;
enum {
    PZ_GATE_NOT = 6, /* pz_circuit_gate op for NOT (no bootstrap) */
    PZ_U8_ADD = 0, PZ_U8_SUB, PZ_U8_MUL, PZ_U8_DIV, PZ_U8_REM, PZ_U8_DIV_REM,
    PZ_U8_EQ, PZ_U8_NEQ, PZ_U8_LT, PZ_U8_GT, PZ_U8_LE, PZ_U8_GE, PZ_U8_MUX, PZ_U8_MAX,
    PZ_U8_MIN, PZ_U8_OVERFLOWING_ADD, PZ_U8_OVERFLOWING_SUB
};
/* ctx may be NULL for a build-only circuit (no pz_circuit_run). */
int pz_circuit_create(pz_ctx *ctx, pz_circuit **out);
int pz_circuit_destroy(pz_circuit *c);
int pz_circuit_inputs(pz_circuit *c, uint32_t count, uint32_t *wires_out);
/* one BooleanGates gate (op PZ_NAND..PZ_XNOR, or PZ_GATE_NOT with b unused) */
int pz_circuit_gate(pz_circuit *c, uint32_t op, uint32_t a, uint32_t b, uint32_t *out);
/* one FheUint8 operation on a8/b8 (and the selector wire for PZ_U8_MUX);
 * writes *nout wires to out: 8 for arithmetic and mux/max/min, 16 for
 * DIV_REM (quotient then remainder), 9 for OVERFLOWING_* (value, flag), 1 for
 * comparisons.  DIV and DIV_REM also update the circuit's div-by-zero flag
 * (src/shortint/mod.rs:42-56). */
int pz_circuit_u8(pz_circuit *c, uint32_t op, const uint32_t *a8, const uint32_t *b8,
                  const uint32_t *sel, uint32_t *out, uint32_t *nout);
int pz_circuit_div_zero_flag(pz_circuit *c, uint32_t *wire);
int pz_circuit_reset_flags(pz_circuit *c);
int pz_circuit_stats(pz_circuit *c, uint32_t *n_wires, uint32_t *n_pbs, uint32_t *depth);
int pz_circuit_level_sizes(pz_circuit *c, uint32_t *h_sizes, uint32_t cap);
int pz_circuit_nodes(pz_circuit *c, uint32_t *h_out, size_t cap);
/* Level of every node in the batch schedule (input wires 0; a NOT shares the
 * level of the gate producing its input).  Levels: each gate's latest level at
 * the circuit's depth, filled least slack first up to whole rounds of the
 * resident blind-rotation grid (512 bootstraps without a device). */
int pz_circuit_node_levels(pz_circuit *c, uint32_t *h_levels, size_t cap);
/* d_inputs[i] is the ciphertext of the i-th input wire ([N+1] u64 each).
 * Every level is queued on the stream (NULL: the context's) without host
 * synchronisation between levels; the schedule goes to the device once per
 * circuit shape.  Runs of one circuit must not overlap (they share its wire
 * pool), and a run is complete once the stream has drained. */
int pz_circuit_run(pz_circuit *c, pz_key *key, const uint64_t *d_inputs, void *stream);
/* d_out[i] = the ciphertext of wire wires[i] from the last pz_circuit_run;
 * PZ_ERR_STATE before any run or for a wire added after the last run. */
int pz_circuit_read(pz_circuit *c, const uint32_t *wires, size_t count, uint64_t *d_out,
                    void *stream);

/* ---- serving loop: FheUint8 programs from many sessions.  A session's key
 * (host rows in the pz_key_upload layout, kept alive by the caller) is
 * uploaded by pz_server_add_session while the resident keys fit
 * hbm_budget_bytes, otherwise when one of its requests is admitted (idle
 * sessions are then evicted LRU to make room).  Each step runs the next
 * ready gate level of every active request, across sessions and up to
 * max_batch_pbs bootstraps, as one batched launch; requests arriving while
 * others run join at the next step.  Steps are queued without waiting for
 * the previous one (up to two in flight per context).  Arrival times are
 * seconds from the start of pz_server_run.  pz_server_add_session registers a session of the
 * creating context's parameter set; pz_server_add_session_ctx one of another
 * context's (same device, same ring size, else PZ_ERR_UNSUPPORTED), whose
 * gates a step batches apart and launches on that context's stream.  Every
 * context a server uses must outlive it. */
typedef struct pz_server pz_server;
typedef struct pz_server_config {
    size_t hbm_budget_bytes; /* resident key bytes allowed             */
    uint32_t pool_rows;      /* ciphertext rows of the device wire pool */
    uint32_t max_batch_pbs;  /* bootstraps per launch (soft cap)        */
    uint32_t max_active;     /* requests in flight                      */
} pz_server_config;
int pz_server_create(pz_ctx *ctx, const pz_server_config *cfg, pz_server **out);
int pz_server_destroy(pz_server *s);
int pz_server_add_session(pz_server *s, const uint64_t *h_rgsw, const uint64_t *h_auto,
                          const uint64_t *h_lwe_ksk, uint32_t *session_id);
int pz_server_add_session_ctx(pz_server *s, pz_ctx *ctx, const uint64_t *h_rgsw,
                              const uint64_t *h_auto, const uint64_t *h_lwe_ksk,
                              uint32_t *session_id);
/* prog: a circuit (may be build-only); h_inputs: one [N+1] ciphertext per
 * input wire in order; out_wires: results to return; n_ops: FheUint8
 * operations in the program (for per-op latency).  Thread-safe with respect
 * to a running pz_server_run (the only call that may overlap it; sessions are
 * added before).  arrival < 0: the request arrives when submitted. */
int pz_server_submit(pz_server *s, uint32_t session, pz_circuit *prog, const uint64_t *h_inputs,
                     const uint32_t *out_wires, uint32_t n_out, uint32_t n_ops, double arrival,
                     uint32_t *request_id);
/* Intake open (open = 1): pz_server_run keeps serving, waiting for
 * submissions when idle, until the intake is closed (open = 0, from any
 * thread) and every request is complete.  Closed (the default): it returns
 * once every request submitted so far is complete.  No counterpart in the
 * reference (it has no server). */
int pz_server_intake(pz_server *s, int open);
int pz_server_run(pz_server *s);
/* Result ciphertexts of a completed request ([n_out][N+1]); PZ_ERR_STATE
 * while it is not complete (may be called while the loop runs). */
int pz_server_result(pz_server *s, uint32_t request_id, uint64_t *h_out);
/* Times of a request (seconds from the start of the run; -1 until admitted /
 * finished); may be called while the loop runs. */
int pz_server_request_times(pz_server *s, uint32_t request_id, double *arrival, double *admitted,
                            double *finished, uint32_t *n_pbs, uint32_t *n_ops);
/* Totals of the last run (read after pz_server_run has returned). */
int pz_server_stats(pz_server *s, uint64_t *total_pbs, uint64_t *batches, uint64_t *key_uploads,
                    uint64_t *evictions, double *wall_s);

/* ---- input path of the non-interactive protocol: a party encrypts bits
 * under its own RLWE secret u_j as batched RLWE ciphertexts
 * (src/bool/ni_mp_api.rs:294-352); the server switches them to the ideal
 * secret s (KeySwitchWithId, src/bool/ni_mp_api.rs:395-440, rlwe_key_switch
 * src/rgsw/runtime.rs:1001-1063) and extracts one LWE per bit
 * (sample_extract, src/pbs.rs:423-449; SampleExtractor,
 * src/shortint/enc_dec.rs:100-135).  Parameter sets with uitos_d = 0 have
 * no such path (PZ_ERR_UNSUPPORTED). */
typedef struct pz_rlwe_ksk pz_rlwe_ksk;
/* party secret u_j[N] (ternary, Hamming weight N/2; the library's own seeded
 * draw) */
int pz_keygen_user_secret(pz_ctx *ctx, const uint8_t *seed32, int32_t *h_u);
/* Ksk(u -> s) (ui_to_s_ksk, src/bool/keys.rs:965-995): the library's own
 * seeded generation of rows A_j uniform, B_j = A_j*s + e + 2^(logb*j + ib)*u,
 * sent to the evaluation domain */
int pz_keygen_rlwe_ksk(pz_ctx *ctx, const uint8_t *seed32, const int32_t *h_u, const int32_t *h_s,
                       pz_rlwe_ksk **out);
/* upload a Ksk given as host rows [2*uitos_d][N] in the evaluation domain
 * (A rows then B rows); Shoup twins are computed on the device */
int pz_rlwe_ksk_upload(pz_ctx *ctx, const uint64_t *h_rows, pz_rlwe_ksk **out);
int pz_rlwe_ksk_download(pz_rlwe_ksk *ksk, uint64_t *h_rows);
int pz_rlwe_ksk_free(pz_rlwe_ksk *ksk);
/* seeded RLWE encryption of a bool batch under h_u: ceil(count/N)
 * ciphertexts [.][2][N], bit t in coefficient t % N of ciphertext t / N
 * (+-Q/8, unused coefficients 0); a from one ChaCha8 stream and the noise from
 * another, both derived from seed32 */
int pz_encrypt_rlwe_bools(pz_ctx *ctx, const int32_t *h_u, const uint8_t *seed32,
                          const uint8_t *h_bits, size_t count, uint64_t *h_out);
/* d_out[b] = rlwe_key_switch(d_in[b]) for B ciphertexts [B][2][N]
 * (d_out may equal d_in) */
int pz_rlwe_key_switch(pz_ctx *ctx, pz_rlwe_ksk *ksk, const uint64_t *d_in, uint64_t *d_out,
                       size_t B, void *stream);
/* d_lwe[t] = sample_extract(d_rlwe[d_src[2t]], index d_src[2t+1]) ([N+1]);
 * the RLWE indices must lie inside d_rlwe, an index d_src[2t+1] >= N gives an
 * all-zero row */
int pz_sample_extract(pz_ctx *ctx, const uint64_t *d_rlwe, const uint32_t *d_src, size_t count,
                      uint64_t *d_lwe, void *stream);

/* ---- RGSW x RGSW product of the non-interactive server-key aggregation:
 * rgsw_by_rgsw_inplace (src/rgsw/runtime.rs:883-985) on B pairs.
 * d_rgsw0[B][2 rlrg_da + 2 rlrg_db][N] (coefficient domain) is overwritten
 * by its product with d_rgsw1[B][2 rgrg_da + 2 rgrg_db][N] (evaluation
 * domain), as the aggregation does per LWE index and party
 * (src/bool/evaluator.rs:1505-1535).  PZ_ERR_UNSUPPORTED where the
 * parameter set has no RGSW x RGSW decomposer. */
int pz_rgsw_by_rgsw(pz_ctx *ctx, uint64_t *d_rgsw0, const uint64_t *d_rgsw1, size_t B,
                    void *stream);

/* ---- Non-interactive multi-party server key: shares and their aggregation.
 * Seeds follow the tree of NonInteractiveMultiPartyCrs
 * (src/bool/evaluator.rs:97-160) from the 32-byte common reference seed crs;
 * noise comes from the party's own noise_seed (the library's own streams).
 * LWE index l is led by party l / ceil(lwe_n / total)
 * (src/bool/evaluator.rs:613-626).  dmax = max(rgrg_da, rgrg_db).
 * PZ_ERR_UNSUPPORTED for sets without uitos_d or rgrg counts; at most
 * PZ_NI_MAX_PARTIES parties. */
#define PZ_NI_MAX_PARTIES 8
/* Party `self`'s u -> s parts (non_interactive_ksk_gen and the zero
 * encryptions for the others, src/multi_party.rs:186-291,
 * src/bool/evaluator.rs:1650-1690): h_ksk_b[uitos_d][N] and
 * h_zero_others[total-1][uitos_d][N] (ascending party order), coefficient
 * domain. */
int pz_ni_share_uitos(pz_ctx *ctx, const uint8_t *crs, const uint8_t *noise_seed, uint32_t self,
                      uint32_t total, const int32_t *h_s, const int32_t *h_u, uint64_t *h_ksk_b,
                      uint64_t *h_zero_others);
/* Party `self`'s RGSW parts for LWE indices [l0, l0 + count)
 * (src/bool/evaluator.rs:1697-1937): h_zero[count][rgrg_da][N] and
 * h_cts[count][dmax][N], row i of an index holding the row built on
 * a_{i,l}; where `self` leads, only the last max(rlrg_da, rlrg_db) rows are
 * filled and the others are zero. */
int pz_ni_share_rgsw(pz_ctx *ctx, const uint8_t *crs, const uint8_t *noise_seed, uint32_t self,
                     uint32_t total, const int32_t *h_s, const int32_t *h_u, const int32_t *h_z,
                     uint32_t l0, uint32_t count, uint64_t *h_zero, uint64_t *h_cts);
/* Server: Ksk(u_user -> s) from the user's ksk_b and the zero encryptions
 * for the user from the other parties, d_zero_for_user[total-1][uitos_d][N]
 * in ascending party order (src/bool/evaluator.rs:1266-1337). */
int pz_ni_uitos_ksk(pz_ctx *ctx, const uint8_t *crs, uint32_t total, uint32_t user,
                    const uint64_t *d_ksk_b, const uint64_t *d_zero_for_user, pz_rlwe_ksk **out);
/* Server: the RGSW part of aggregate_non_interactive_multi_party_server_key_
 * shares (src/bool/evaluator.rs:1338-1565) for LWE indices [l0, l0+count):
 * every party's produced RGSW (produce_rgsw_ciphertext_from_ni_rgsw,
 * src/bool/evaluator.rs:480-608) and the leader's times the others' in
 * ascending party order (rgsw_by_rgsw_inplace).  ksks[total] from
 * pz_ni_uitos_ksk; d_zero[total][count][rgrg_da][N],
 * d_cts[total][count][dmax][N] (each party's share arrays);
 * d_out[count][2 rlrg_da + 2 rlrg_db][N], coefficient domain. */
int pz_ni_aggregate_rgsw(pz_ctx *ctx, const uint8_t *crs, uint32_t total,
                         pz_rlwe_ksk *const *ksks, uint32_t l0, uint32_t count,
                         const uint64_t *d_zero, const uint64_t *d_cts, uint64_t *d_out,
                         void *stream);

/* Party side: the automorphism-key and LWE-KSK shares
 * (_common_rountine_multi_party_auto_keys_share_gen and
 * _common_rountine_multi_party_lwe_ksk_share_gen, src/bool/evaluator.rs:
 * 1952-1965, 1981-2046, over seeded_auto_key_gen src/rgsw/keygen.rs:380-430
 * and seeded_lwe_ksk_keygen src/lwe.rs:46-101) of a party with secrets h_s
 * (RLWE) and h_z (LWE): b parts only, h_auto_b[w+1][auto_d][N] (coefficient
 * domain, keys for -g, g, .., g^w) and h_ksk_b[N * lwe_d] (mod q_ks); the a
 * parts are the CRS streams that pz_ni_expand_server_key regenerates. */
int pz_ni_share_auto_ksk(pz_ctx *ctx, const uint8_t *crs, const uint8_t *noise_seed,
                         const int32_t *h_s, const int32_t *h_z, uint64_t *h_auto_b,
                         uint64_t *h_ksk_b);
/* Server: the automorphism-key and LWE-KSK parts of the aggregation
 * (src/bool/evaluator.rs:1574-1612): d_auto_b = sum of
 * d_auto_shares[total][w+1][auto_d][N] mod Q, d_ksk_b = sum of
 * d_ksk_shares[total][N * lwe_d] mod q_ks. */
int pz_ni_aggregate_auto_ksk(pz_ctx *ctx, uint32_t total, const uint64_t *d_auto_shares,
                             const uint64_t *d_ksk_shares, uint64_t *d_auto_b,
                             uint64_t *d_ksk_b, void *stream);
/* Server: the evaluation-domain server key of a seeded non-interactive
 * server key (NonInteractiveServerKeyEvaluationDomain::from, src/bool/keys.rs:
 * 851-945, with the Shoup twins of :1088-1155): d_rgsw[n][2 rlrg_da +
 * 2 rlrg_db][N] (coefficient domain, from pz_ni_aggregate_rgsw), d_auto_b and
 * d_ksk_b (the summed b parts); the a parts are expanded on the device from
 * the CRS seeds.  The inputs are left unchanged. */
int pz_ni_expand_server_key(pz_ctx *ctx, const uint8_t *crs, const uint64_t *d_rgsw,
                            const uint64_t *d_auto_b, const uint64_t *d_ksk_b, pz_key **out);

/* ---- seeded wire formats of the non-interactive server key.  A blob is a
 * pz_ni_blob_header followed by its sections: arrays of little-endian u64
 * words at 8-byte aligned offsets (offset[i] bytes from the blob start,
 * words[i] long; unused entries 0).  Only b parts travel; every a part is
 * regenerated from the common reference seed.  Blob buffers must be 8-byte
 * aligned.
 *   A party's share ("PZNISHR", CommonReferenceSeededNonInteractive
 *   MultiPartyServerKeyShare, src/bool/keys.rs:1240-1394), sections:
 *     0 self_leader_ni_rgsw_cts  [lead][max(rlrg_da, rlrg_db)][N], the LWE
 *       indices [lb, le) the party leads (src/bool/evaluator.rs:613-626)
 *     1 not_self_leader_ni_rgsw_cts [lwe_n - lead][max(rgrg_da, rgrg_db)][N],
 *       the other indices in ascending order
 *     2 ni_rgsw_zero_encs        [lwe_n][rgrg_da][N]
 *     3 ui_to_s_ksk (b parts)    [uitos_d][N]
 *     4 ksk_zero_encs_for_others [total - 1][uitos_d][N], ascending party
 *       order without the party itself
 *     5 auto_keys_share          [w + 1][auto_d][N]
 *     6 lwe_ksk_share            [N * lwe_d]
 *   The seeded server key ("PZNIKEY", SeededNonInteractiveMultiPartyServerKey,
 *   src/bool/keys.rs:1029-1063), sections:
 *     0 ui_to_s_ksks (b parts)   [total][uitos_d][N], coefficient domain
 *     1 rgsw_cts                 [lwe_n][2 rlrg_da + 2 rlrg_db][N],
 *       coefficient domain
 *     2 auto_keys (b parts)      [w + 1][auto_d][N]
 *     3 lwe_ksk (b parts)        [N * lwe_d]
 */
#define PZ_NI_SECTIONS 8
#define PZ_NI_BLOB_VERSION 1
typedef struct pz_ni_blob_header {
    char magic[8];       /* "PZNISHR" or "PZNIKEY", NUL-padded             */
    uint32_t version;    /* PZ_NI_BLOB_VERSION                              */
    uint32_t user;       /* the party's id (share); 0 (seeded server key)    */
    uint32_t total;      /* parties                                         */
    uint32_t nsections;
    pz_params params;    /* the parameter set (88 bytes, no padding)        */
    uint8_t crs[32];     /* common reference seed                           */
    uint64_t offset[PZ_NI_SECTIONS];
    uint64_t words[PZ_NI_SECTIONS];
} pz_ni_blob_header; /* 272 bytes */
/* Blob sizes in bytes (no device needed). */
int pz_ni_share_bytes(const pz_params *p, uint32_t self, uint32_t total, size_t *bytes);
int pz_ni_seeded_key_bytes(const pz_params *p, uint32_t total, size_t *bytes);
/* Party `self` of `total`: its whole share into h_share
 * (gen_non_interactive_multi_party_key_share, src/bool/evaluator.rs:
 * 1624-1979; the parts as pz_ni_share_uitos, pz_ni_share_rgsw over every
 * index and pz_ni_share_auto_ksk make them, from the same noise_seed). */
int pz_ni_key_share(pz_ctx *ctx, const uint8_t *crs, const uint8_t *noise_seed, uint32_t self,
                    uint32_t total, const int32_t *h_s, const int32_t *h_u, const int32_t *h_z,
                    uint8_t *h_share);
/* Server: the seeded server key of `total` shares, in any order
 * (aggregate_non_interactive_multi_party_server_key_shares,
 * src/bool/evaluator.rs:1222-1622), computed on the device.  PZ_ERR_ARG if
 * the shares disagree on the parameter set, party count or seed, or two come
 * from one party. */
int pz_ni_aggregate_shares(pz_ctx *ctx, const uint8_t *const *h_shares, uint32_t total,
                           uint8_t *h_seeded_key);
/* Server: the evaluation-domain server key (and, if ksks is not NULL, the
 * parties' Ksk(u_j -> s), ksks[total]) of a seeded server key
 * (src/bool/keys.rs:851-1005, :1088-1155), expanded on the device. */
int pz_ni_server_key_from_seeded(pz_ctx *ctx, const uint8_t *h_seeded_key, pz_key **key,
                                 pz_rlwe_ksk **ksks);

/* Autotuner: times every blind-rotation launch shape that fits the device
 * (kernel variant x resident work groups per CU) on B synthetic gates with
 * this key, and uses the fastest for batches in B's power-of-two bucket
 * (automatic mode only).  h_choice (may be null) receives the variant, the
 * work groups per CU and the measured microseconds per batch.  No
 * counterpart in the reference (its evaluator runs one gate at a time). */
int pz_autotune(pz_ctx *ctx, pz_key *key, size_t B, uint32_t *h_choice);

/* Blind-rotation kernel choice.  No counterpart in the reference (its
 * evaluator is single-threaded): 0 = automatic, 1 = one work group per
 * bootstrap, 2 = two half work groups per bootstrap, 8 coefficients per
 * thread, 3 = the same with 4 coefficients per thread (twice the waves;
 * 2 and 3 only where d_a = d_b = d_auto = 1; PZ_ERR_UNSUPPORTED otherwise).
 * Modes 2 and 3 run their latency variant while every bootstrap of a batch
 * has a CU of its own and their throughput variant (two bootstraps per CU)
 * beyond that.  Automatic (where 2 and 3 are available, else 1) takes the
 * latency variant of 3 and the throughput variant of 2, unless pz_autotune
 * chose another shape for the batch size.  Outputs are the same words in
 * every mode. */
int pz_set_blind_rotate_mode(pz_ctx *ctx, int mode);

/* Automorphism count of each bootstrap of the last pz_pbs / pz_gates batch
 * (the blind rotation always runs n RLWE x RGSW products; the number of
 * automorphisms depends on the input).  Used for the algorithmic work count. */
int pz_last_auto_counts(pz_ctx *ctx, uint32_t *h_counts, size_t B);

/* Per-kernel timing of the last pz_gates / pz_pbs call made with timing
 * enabled (milliseconds from HIP events on the launch stream: pz_gates records
 * them as nodes of its captured graph, around each of its kernels; pz_pbs
 * around eager launches).  With timing enabled the call waits for its
 * kernels. */
int pz_set_timing(pz_ctx *ctx, int enabled);
int pz_last_timing(pz_ctx *ctx, float *ms_prep, float *ms_ks, float *ms_blind_rotate);

/* VALU issue-rate probe.  No counterpart in the reference: it measures the
 * integer VALU rates the PBS kernels' roofline rests on.  kind: 0 v_add_u32,
 * 1 v_mul_lo_u32, 2 v_mul_hi_u32, 3 v_mad_u64_u32 (lane operations per
 * second), 4 one 54-bit Shoup multiply (multiplies per second), 5
 * v_lshl_add_u64 (lane operations per second), each over 8 independent
 * chains per thread on every CU. */
int pz_valu_probe(pz_ctx *ctx, int kind, uint32_t iters, double *ops_per_s);

#ifdef __cplusplus
}
#endif
#endif
