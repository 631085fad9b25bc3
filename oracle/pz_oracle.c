/*
 * pz_oracle.c -- CPU restatement of phantom-zone's PBS path (C11).
 *
 * ORACLE / TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library, and only as the
 * checker (or as the timed CPU baseline).  The product never routes through it.
 *
 * Reference: gausslabs/phantom-zone (Rust).  Citations are file:line in that
 * repository.  Each routine restates the reference's arithmetic exactly:
 * Shoup/Barrett formulas, lazy ranges, NTT stage order and twiddle indexing,
 * decomposition carries, dlog bucketing and the blind-rotation op order, and
 * the f64 moddown.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   ChaCha block         RFC 7539 / OpenSSL ChaCha20 keystream KATs at 20
 *                        rounds; the 8-round variant only changes the round
 *                        count (rand_chacha's ChaCha8Rng).
 *   Shoup / Barrett /    brute force against exact integer arithmetic
 *   shoup_matrix_fma     (Python ints), range invariants [0,2q).
 *   NTT fwd / inv        INTT(NTT(a).NTT(b)) == schoolbook negacyclic product;
 *                        round trips; output ranges; psi^N == -1 (mod Q);
 *                        an independent Python radix-2 restatement over
 *                        Python ints for the lazy values.
 *   decomposer           recomposition within +-1 of round(v), digit bounds,
 *                        decompose_iter == decompose_to_vec, the lwe_q wrap.
 *   ext product / auto   Shoup path == Barrett path bit for bit (reference
 *                        tests src/rgsw/mod.rs:650, :790); decryption equals
 *                        the schoolbook product / the permuted message.
 *   LWE key switch       brute force over Python ints; decryption under z.
 *   moddown / odd switch Python floats (IEEE doubles) at rounding boundaries.
 *   PBS / gates          gate truth tables after decryption for all 4 inputs
 *                        of all 6 gates (NAND / XOR on the interactive sets);
 *                        the recorded op order equals an independent Python
 *                        restatement of the bucket loop for every named set;
 *                        dlog maps against powers of g; multi-threaded gate
 *                        batches equal one-at-a-time gates.
 *   u -> s key switch,   exact big-int closed form of every output word;
 *   sample extract       the extracted LWE decrypts to the RLWE phase.
 *   RGSW x RGSW          exact big-int closed form of every row pair.
 *   NI shares, Ksk from  CRS seed-tree bytes against the Python ChaCha8
 *   shares, produced     stream; leader segments by brute force; Ksk rows,
 *   RGSWs, aggregation   produced RGSW rows and the aggregate decrypt under the
 *                        ideal secret with noise below the smallest gadget
 *                        power.
 *   NI auto-key / LWE-   the a parts equal the CRS streams drawn by the Python
 *   KSK shares, sums,    ChaCha8 / UniformInt restatement; share rows decrypt
 *   key expansion        to RLWE'(s(X^k)) and beta * s_i; sums exact; a key
 *                        from shares alone bootstraps correctly.
 *   RNG streams (uniform / Gaussian / ternary sampling) and therefore the key
 *   material:            PARITY UNPINNED -- this is the project's own seeded
 *                        procedure (rand / rand_distr sources are not
 *                        available), not the Rust crate's stream.  The CRS
 *                        seed tree and UniformInt draws follow rand 0.8.5's
 *                        published algorithm (pinned as above); the Gaussian
 *                        noise words of keys and shares stay unpinned.
 */
#define _GNU_SOURCE
#include "pz_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------- */
/* Parameter sets: src/bool/parameters.rs:526-726                            */
/* ------------------------------------------------------------------------- */
/* rgrg_decomposer_params, src/bool/parameters.rs:526-726 (SP_TEST: none) */
static const uint32_t PZO_RGRG[8][3] = {{4, 10, 9}, {3, 13, 12}, {2, 20, 18}, {0, 0, 0},
                                        {7, 6, 5},  {6, 7, 6},   {5, 9, 8},   {4, 10, 9}};

static int params_base(int which, pzo_params *o);

int pzo_params_named(int which, pzo_params *o)
{
    int rc = params_base(which, o);
    if (rc == 0) {
        o->rgrg_logb = PZO_RGRG[which][0];
        o->rgrg_da = PZO_RGRG[which][1];
        o->rgrg_db = PZO_RGRG[which][2];
    }
    return rc;
}

static int params_base(int which, pzo_params *o)
{
    memset(o, 0, sizeof(*o));
    o->rlwe_q = 18014398509404161ULL;
    o->rlwe_n = 2048;
    o->rlrg_logb = 17; o->rlrg_da = 1; o->rlrg_db = 1;
    o->auto_logb = 24; o->auto_d = 1;
    o->lwe_logb = 1;
    o->g = 5; o->w = 10;
    switch (which) {
    case 0: /* NI_2P :598-623 */
        o->lwe_q = 1ULL << 16; o->br_q = 1u << 12; o->lwe_n = 520; o->lwe_d = 13;
        o->lwe_secret_gaussian = 1; o->uitos_logb = 1; o->uitos_d = 50; return 0;
    case 1: /* NI_4P_HB_FR :625-650 */
        o->lwe_q = 1ULL << 16; o->br_q = 1u << 11; o->lwe_n = 620; o->lwe_d = 13;
        o->uitos_logb = 1; o->uitos_d = 50; return 0;
    case 2: /* NI_8P :679-704 */
        o->lwe_q = 1ULL << 17; o->br_q = 1u << 12; o->lwe_n = 660; o->lwe_d = 14;
        o->uitos_logb = 1; o->uitos_d = 50; return 0;
    case 3: /* SP_TEST_BOOL_PARAMS :706-726 */
        o->rlwe_q = 268369921ULL; o->rlwe_n = 512; o->lwe_q = 1ULL << 16;
        o->br_q = 1u << 9; o->lwe_n = 100; o->lwe_logb = 4; o->lwe_d = 4;
        o->rlrg_logb = 7; o->rlrg_da = 4; o->rlrg_db = 4;
        o->auto_logb = 7; o->auto_d = 4; o->g = 5; o->w = 5;
        o->lwe_secret_gaussian = 1; return 0;
    case 4: /* I_2P_LB_SR :526-549 */
        o->lwe_q = 1ULL << 15; o->br_q = 1u << 11; o->lwe_n = 580; o->lwe_d = 12;
        return 0;
    case 5: /* I_4P :551-574 */
        o->lwe_q = 1ULL << 16; o->br_q = 1u << 11; o->lwe_n = 620; o->lwe_d = 13;
        return 0;
    case 6: /* I_8P :576-596 */
        o->lwe_q = 1ULL << 17; o->br_q = 1u << 12; o->lwe_n = 660; o->lwe_d = 14;
        return 0;
    case 7: /* NI_4P_LB_SR :652-677 */
        o->lwe_q = 1ULL << 16; o->br_q = 1u << 12; o->lwe_n = 620; o->lwe_d = 13;
        o->uitos_logb = 1; o->uitos_d = 50; return 0;
    default:
        return -1;
    }
}

/* ------------------------------------------------------------------------- */
/* ChaCha block function with rand_chacha's layout: words 12..13 hold a      */
/* 64-bit block counter, 14..15 a 64-bit stream id (RFC 7539 with a 64-bit    */
/* counter).  ChaCha8Rng = 8 rounds (src/random.rs:66-91, src/ntt.rs:325).    */
/* ------------------------------------------------------------------------- */
#define ROTL32(v, c) (((v) << (c)) | ((v) >> (32 - (c))))
#define QR(a, b, c, d)                                   \
    do {                                                 \
        a += b; d ^= a; d = ROTL32(d, 16);               \
        c += d; b ^= c; b = ROTL32(b, 12);               \
        a += b; d ^= a; d = ROTL32(d, 8);                \
        c += d; b ^= c; b = ROTL32(b, 7);                \
    } while (0)

void pzo_chacha_block(const uint8_t key[32], uint64_t counter, uint64_t stream,
                      int rounds, uint32_t out[16])
{
    uint32_t s[16], x[16];
    s[0] = 0x61707865u; s[1] = 0x3320646eu; s[2] = 0x79622d32u; s[3] = 0x6b206574u;
    for (int i = 0; i < 8; i++)
        s[4 + i] = (uint32_t)key[4 * i] | ((uint32_t)key[4 * i + 1] << 8) |
                   ((uint32_t)key[4 * i + 2] << 16) | ((uint32_t)key[4 * i + 3] << 24);
    s[12] = (uint32_t)counter; s[13] = (uint32_t)(counter >> 32);
    s[14] = (uint32_t)stream;  s[15] = (uint32_t)(stream >> 32);
    memcpy(x, s, sizeof(s));
    for (int r = 0; r < rounds; r += 2) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

struct pzo_rng {
    uint8_t key[32];
    uint64_t counter;
    uint32_t buf[16];
    int idx;
};

pzo_rng *pzo_rng_new(const uint8_t seed[32])
{
    pzo_rng *r = (pzo_rng *)calloc(1, sizeof(*r));
    memcpy(r->key, seed, 32);
    r->idx = 16;
    return r;
}

void pzo_rng_free(pzo_rng *r) { free(r); }

/* rand_core BlockRng: words are consumed in order; next_u64 = lo | hi<<32
 * of the next two words (rand_core 0.6.4 BlockRng::next_u64). */
uint32_t pzo_rng_next_u32(pzo_rng *r)
{
    if (r->idx >= 16) {
        pzo_chacha_block(r->key, r->counter, 0, 8, r->buf);
        r->counter++;
        r->idx = 0;
    }
    return r->buf[r->idx++];
}

uint64_t pzo_rng_next_u64(pzo_rng *r)
{
    uint64_t lo = pzo_rng_next_u32(r);
    uint64_t hi = pzo_rng_next_u32(r);
    return lo | (hi << 32);
}

/* byte stream: whole words, little endian (rand_core fill_via_u32_chunks) */
void pzo_rng_fill_bytes(pzo_rng *r, uint8_t *out, size_t n)
{
    size_t i = 0;
    while (i < n) {
        uint32_t w = pzo_rng_next_u32(r);
        for (int b = 0; b < 4 && i < n; b++, i++) out[i] = (uint8_t)(w >> (8 * b));
    }
}

/* rand 0.8.5 UniformInt<u64>::sample: widening multiply with rejection zone
 * (published algorithm, restated; stream parity unpinned). */
uint64_t pzo_rng_uniform(pzo_rng *r, uint64_t range)
{
    uint64_t ints_to_reject = (UINT64_MAX - range + 1) % range;
    uint64_t zone = UINT64_MAX - ints_to_reject;
    for (;;) {
        uint64_t v = pzo_rng_next_u64(r);
        u128 m = (u128)v * range;
        uint64_t lo = (uint64_t)m, hi = (uint64_t)(m >> 64);
        if (lo <= zone) return hi;
    }
}

static const uint64_t PZO_CDT_ROUND[42] = {
    0x1fe283c57e62b000ULL, 0x5c9f0a1d68211000ULL, 0x9118b59b30188000ULL,
    0xba394ad0e9bb6000ULL, 0xd776ad2ff2044000ULL, 0xea5237ec9550e000ULL,
    0xf55a79de6976e000ULL, 0xfb3549a53711a000ULL, 0xfe06d0641332f000ULL,
    0xff41e43843be2000ULL, 0xffbeb2df59999000ULL, 0xffeb8b424b33c000ULL,
    0xfffa291785952000ULL, 0xfffe7b5bd6539000ULL, 0xffffa4085a2f0000ULL,
    0xffffec331b5fb000ULL, 0xfffffc1f7ad22000ULL, 0xffffff4f4fc0a000ULL,
    0xffffffe369496000ULL, 0xfffffffbcbf48000ULL, 0xffffffff7045c000ULL,
    0xffffffffee900000ULL, 0xfffffffffe142000ULL, 0xffffffffffcec000ULL,
    0xfffffffffffb8000ULL, 0xffffffffffffa000ULL, 0xffffffffffffffffULL,
    0xffffffffffffffffULL, 0xffffffffffffffffULL, 0xffffffffffffffffULL,
    0xffffffffffffffffULL, 0xffffffffffffffffULL, 0xffffffffffffffffULL,
    0xffffffffffffffffULL, 0xffffffffffffffffULL, 0xffffffffffffffffULL,
    0xffffffffffffffffULL, 0xffffffffffffffffULL, 0xffffffffffffffffULL,
    0xffffffffffffffffULL, 0xffffffffffffffffULL, 0xffffffffffffffffULL,
};
static const uint64_t PZO_CDT_TRUNC[42] = {
    0x3eff57680c66a400ULL, 0x7824f735e7db0000ULL, 0xa72b8006276f2000ULL,
    0xca45da4d592f4000ULL, 0xe20ade4707b23000ULL, 0xf0a4971d1c5a8000ULL,
    0xf8c73f4b85aae000ULL, 0xfce3eaa49015f000ULL, 0xfec68fa732ff1000ULL,
    0xff8f4ba19fa56000ULL, 0xffdb06ac14e44000ULL, 0xfff4f143d774e000ULL,
    0xfffcfcc974786000ULL, 0xffff40b3b3cca000ULL, 0xffffd4d199668000ULL,
    0xfffff72233732000ULL, 0xfffffe5835846000ULL, 0xffffffb8129ae000ULL,
    0xfffffff4e7dfa000ULL, 0xfffffffe71fbb000ULL, 0xffffffffcd54e000ULL,
    0xfffffffffa245000ULL, 0xffffffffff628000ULL, 0xfffffffffff10000ULL,
    0xfffffffffffeb000ULL, 0xffffffffffffe000ULL, 0xffffffffffffffffULL,
    0xffffffffffffffffULL, 0xffffffffffffffffULL, 0xffffffffffffffffULL,
    0xffffffffffffffffULL, 0xffffffffffffffffULL, 0xffffffffffffffffULL,
    0xffffffffffffffffULL, 0xffffffffffffffffULL, 0xffffffffffffffffULL,
    0xffffffffffffffffULL, 0xffffffffffffffffULL, 0xffffffffffffffffULL,
    0xffffffffffffffffULL, 0xffffffffffffffffULL, 0xffffffffffffffffULL,
};

static int64_t cdt_draw(pzo_rng *r, const uint64_t *t)
{
    uint64_t u = pzo_rng_next_u64(r);
    int64_t k = 41;
    for (int i = 0; i < 42; i++) {
        if (u < t[i]) { k = i; break; }
    }
    if (k > 0 && (pzo_rng_next_u32(r) & 1u)) k = -k;
    return k;
}

int64_t pzo_rng_gauss_round(pzo_rng *r) { return cdt_draw(r, PZO_CDT_ROUND); }
int64_t pzo_rng_gauss_trunc(pzo_rng *r) { return cdt_draw(r, PZO_CDT_TRUNC); }

/* fill_random_ternary_secret_with_hamming_weight, src/utils.rs:75-110.
 * Bytes come from Uniform<u8> (one u32 word per byte). */
void pzo_ternary_hw(pzo_rng *r, int32_t *out, uint32_t n, uint32_t hw)
{
    uint32_t nbytes = (hw + 7) / 8;
    uint8_t *bytes = (uint8_t *)malloc(nbytes ? nbytes : 1);
    for (uint32_t i = 0; i < nbytes; i++) bytes[i] = (uint8_t)pzo_rng_next_u32(r);
    uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; i++) { idx[i] = i; out[i] = 0; }
    uint32_t len = n, bit = 0, byte = 0;
    for (uint32_t i = 0; i < hw; i++) {
        uint32_t s = (uint32_t)pzo_rng_uniform(r, len);
        uint32_t cur = (bytes[byte] >> bit) & 1u;
        out[idx[s]] = cur ? 1 : -1;
        idx[s] = idx[len - 1];
        len--;
        if (bit == 7) { bit = 0; byte++; } else bit++;
    }
    free(idx);
    free(bytes);
}

/* ------------------------------------------------------------------------- */
/* Scalar arithmetic                                                         */
/* ------------------------------------------------------------------------- */

/* ShoupMul::representation, src/utils.rs:37-39 */
uint64_t pzo_shoup_repr(uint64_t v, uint64_t q)
{
    return (uint64_t)(((u128)v << 64) / q);
}

/* ShoupMul::mul, src/utils.rs:43-46: b*a - q*floor(b_shoup*a / 2^64), wrapping */
uint64_t pzo_shoup_mul(uint64_t a, uint64_t b, uint64_t b_shoup, uint64_t q)
{
    uint64_t quot = (uint64_t)(((u128)b_shoup * a) >> 64);
    return b * a - q * quot;
}

/* ModularOpsU64 Barrett constants, src/backend/modulus_u64.rs:23-43 */
typedef struct {
    uint64_t q, q_twice;
    uint32_t logq, alpha;
    u128 mu;
} modq;

static void modq_init(modq *m, uint64_t q)
{
    m->q = q;
    m->q_twice = q << 1;
    m->logq = 64 - (uint32_t)__builtin_clzll(q + 1);
    m->mu = ((u128)1 << (m->logq * 2 + 3)) / q;
    m->alpha = m->logq + 3;
}

/* mul_mod_fast, src/backend/modulus_u64.rs:106-130 */
static inline uint64_t mm_mul(const modq *m, uint64_t a, uint64_t b)
{
    u128 ab = (u128)a * b;
    u128 tmp = ab >> (m->logq - 2);
    u128 k = (tmp * m->mu) >> (m->alpha + 2);
    uint64_t out = (uint64_t)(ab - k * m->q);
    if (out >= m->q) out -= m->q;
    return out;
}

/* mul_mod_fast_lazy, :83-100 */
static inline uint64_t mm_mul_lazy(const modq *m, uint64_t a, uint64_t b)
{
    u128 ab = (u128)a * b;
    u128 tmp = ab >> (m->logq - 2);
    u128 k = (tmp * m->mu) >> (m->alpha + 2);
    return (uint64_t)(ab - k * m->q);
}

/* add_mod_fast :46-55, sub_mod_fast :68-77, neg :147-149 */
static inline uint64_t mm_add(const modq *m, uint64_t a, uint64_t b)
{
    uint64_t o = a + b;
    if (o >= m->q) o -= m->q;
    return o;
}
static inline uint64_t mm_sub(const modq *m, uint64_t a, uint64_t b)
{
    return a >= b ? a - b : (m->q + a) - b;
}
static inline uint64_t mm_neg(const modq *m, uint64_t a) { return m->q - a; }

uint64_t pzo_mul_mod_fast(uint64_t q, uint64_t a, uint64_t b)
{
    modq m; modq_init(&m, q);
    return mm_mul(&m, a, b);
}

uint64_t pzo_mul_mod_fast_lazy(uint64_t q, uint64_t a, uint64_t b)
{
    modq m; modq_init(&m, q);
    return mm_mul_lazy(&m, a, b);
}

/* mod_exponent, src/utils.rs:145-165 */
uint64_t pzo_mod_exponent(uint64_t a, uint64_t b, uint64_t q)
{
    uint64_t acc = a, out = 1;
    while (b != 0) {
        if (b & 1) out = (uint64_t)(((u128)acc * out) % q);
        acc = (uint64_t)(((u128)acc * acc) % q);
        b >>= 1;
    }
    return out;
}

/* ShoupMatrixFMA::shoup_matrix_fma, src/backend/modulus_u64.rs:231-264.
 * a, a_shoup, b: d rows x n, row-major. out in [0,2q). */
void pzo_shoup_matrix_fma(uint64_t q, uint64_t *out, const uint64_t *a,
                          const uint64_t *a_shoup, const uint64_t *b,
                          uint32_t d, uint32_t n)
{
    uint64_t q_twice = q << 1;
    for (uint32_t r = 0; r < d; r++) {
        const uint64_t *ar = a + (size_t)r * n, *asr = a_shoup + (size_t)r * n,
                       *br = b + (size_t)r * n;
        for (uint32_t j = 0; j < n; j++) {
            uint64_t quot = (uint64_t)(((u128)asr[j] * br[j]) >> 64);
            uint64_t v = ar[j] * br[j] + out[j];
            v = v - q * quot;
            if (v >= q_twice) v -= q_twice;
            out[j] = v;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* NTT, src/ntt.rs                                                           */
/* ------------------------------------------------------------------------- */
struct pzo_ntt {
    uint64_t q, q_twice;
    uint32_t n;
    uint64_t n_inv, n_inv_shoup, psi;
    uint64_t *psi_bo, *psi_bo_shoup, *psi_inv_bo, *psi_inv_bo_shoup;
};

/* forward_butterly_0_to_4q, src/ntt.rs:31-49 */
static inline void fb_4q(uint64_t *px, uint64_t *py, uint64_t w, uint64_t ws,
                         uint64_t q, uint64_t q2)
{
    uint64_t x = *px;
    if (x >= q2) x -= q2;
    uint64_t t = pzo_shoup_mul(*py, w, ws, q);
    *px = x + t;
    *py = x + q2 - t;
}

/* forward_butterly_0_to_2q, src/ntt.rs:51-75 */
static inline void fb_2q(uint64_t *px, uint64_t *py, uint64_t w, uint64_t ws,
                         uint64_t q, uint64_t q2)
{
    uint64_t x = *px;
    if (x >= q2) x -= q2;
    uint64_t t = pzo_shoup_mul(*py, w, ws, q);
    uint64_t ox = x + t, oy = x - t;
    uint64_t ox2 = ox - q2, oy2 = oy + q2;
    *px = ox < ox2 ? ox : ox2;
    *py = oy < oy2 ? oy : oy2;
}

/* inverse_butterfly_0_to_2q, src/ntt.rs:84-104 */
static inline void ib_2q(uint64_t *px, uint64_t *py, uint64_t w, uint64_t ws,
                         uint64_t q, uint64_t q2)
{
    uint64_t x = *px, y = *py;
    uint64_t xd = x + y;
    if (xd >= q2) xd -= q2;
    uint64_t t = x + q2 - y;
    *px = xd;
    *py = pzo_shoup_mul(t, w, ws, q);
}

static inline uint64_t min_sub(uint64_t v, uint64_t q)
{
    uint64_t v2 = v - q;
    return v < v2 ? v : v2;
}

/* ntt_lazy :110-143 (strict = 0) and ntt :146-180 (strict = 1) */
static void ntt_fwd(const pzo_ntt *t, uint64_t *a, int strict)
{
    uint32_t n = t->n, tt = n, m = 1;
    uint64_t q = t->q, q2 = t->q_twice;
    while (m < n) {
        tt >>= 1;
        const uint64_t *w = t->psi_bo + m, *ws = t->psi_bo_shoup + m;
        if (tt == 1) {
            for (uint32_t k = 0; k < n / 2; k++) {
                fb_2q(&a[2 * k], &a[2 * k + 1], w[k], ws[k], q, q2);
                if (strict) {
                    a[2 * k] = min_sub(a[2 * k], q);
                    a[2 * k + 1] = min_sub(a[2 * k + 1], q);
                }
            }
        } else {
            for (uint32_t i = 0; i < m; i++) {
                uint64_t *blk = a + 2 * (size_t)i * tt;
                for (uint32_t j = 0; j < tt; j++)
                    fb_4q(&blk[j], &blk[j + tt], w[i], ws[i], q, q2);
            }
        }
        m <<= 1;
    }
}

/* ntt_inv_lazy :187-230 (strict = 0) and ntt_inv :233-278 (strict = 1) */
static void ntt_inv(const pzo_ntt *t, uint64_t *a, int strict)
{
    uint32_t n = t->n, m = n >> 1, tt = 1;
    uint64_t q = t->q, q2 = t->q_twice;
    while (m > 0) {
        if (m == 1) {
            for (uint32_t j = 0; j < tt; j++) {
                uint64_t *x = &a[j], *y = &a[j + tt];
                ib_2q(x, y, t->psi_inv_bo[1], t->psi_inv_bo_shoup[1], q, q2);
                uint64_t ox = pzo_shoup_mul(*x, t->n_inv, t->n_inv_shoup, q);
                uint64_t oy = pzo_shoup_mul(*y, t->n_inv, t->n_inv_shoup, q);
                if (strict) { ox = min_sub(ox, q); oy = min_sub(oy, q); }
                *x = ox; *y = oy;
            }
        } else {
            const uint64_t *w = t->psi_inv_bo + m, *ws = t->psi_inv_bo_shoup + m;
            for (uint32_t i = 0; i < m; i++) {
                uint64_t *blk = a + 2 * (size_t)i * tt;
                for (uint32_t j = 0; j < tt; j++)
                    ib_2q(&blk[j], &blk[j + tt], w[i], ws[i], q, q2);
            }
        }
        tt *= 2;
        m >>= 1;
    }
}

/* find_primitive_root :283-307 with ChaCha8Rng::from_seed([0;32]) :325-327 */
static uint64_t find_primitive_root(uint64_t q, uint64_t n)
{
    uint8_t seed[32] = {0};
    pzo_rng *r = pzo_rng_new(seed);
    uint64_t t = (q - 1) / n, res = 0;
    for (int it = 0; it < 100; it++) {
        uint64_t omega = pzo_rng_next_u64(r) % q;
        omega = pzo_mod_exponent(omega, t, q);
        if (pzo_mod_exponent(omega, n >> 1, q) == 1) continue;
        res = omega;
        break;
    }
    pzo_rng_free(r);
    return res;
}

static uint32_t bitrev(uint32_t i, uint32_t logn)
{
    uint32_t r = 0;
    for (uint32_t b = 0; b < logn; b++) r |= ((i >> b) & 1u) << (logn - 1 - b);
    return r;
}

/* NttBackendU64::_new, src/ntt.rs:323-386 */
pzo_ntt *pzo_ntt_new(uint64_t q, uint32_t n)
{
    pzo_ntt *t = (pzo_ntt *)calloc(1, sizeof(*t));
    t->q = q; t->q_twice = 2 * q; t->n = n;
    uint64_t psi = find_primitive_root(q, 2 * (uint64_t)n);
    uint64_t psi_inv = pzo_mod_exponent(psi, q - 2, q);
    t->psi = psi;
    modq m; modq_init(&m, q);
    t->psi_bo = (uint64_t *)malloc(8 * n);
    t->psi_inv_bo = (uint64_t *)malloc(8 * n);
    t->psi_bo_shoup = (uint64_t *)malloc(8 * n);
    t->psi_inv_bo_shoup = (uint64_t *)malloc(8 * n);
    uint32_t logn = 0;
    while ((1u << logn) < n) logn++;
    uint64_t rp = 1, rpi = 1;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t bo = bitrev(i, logn);
        t->psi_bo[bo] = rp;
        t->psi_inv_bo[bo] = rpi;
        rp = mm_mul(&m, rp, psi);
        rpi = mm_mul(&m, rpi, psi_inv);
    }
    for (uint32_t i = 0; i < n; i++) {
        t->psi_bo_shoup[i] = pzo_shoup_repr(t->psi_bo[i], q);
        t->psi_inv_bo_shoup[i] = pzo_shoup_repr(t->psi_inv_bo[i], q);
    }
    t->n_inv = pzo_mod_exponent(n, q - 2, q);
    t->n_inv_shoup = pzo_shoup_repr(t->n_inv, q);
    return t;
}

void pzo_ntt_free(pzo_ntt *t)
{
    if (!t) return;
    free(t->psi_bo); free(t->psi_inv_bo); free(t->psi_bo_shoup); free(t->psi_inv_bo_shoup);
    free(t);
}

uint64_t pzo_ntt_psi(const pzo_ntt *t) { return t->psi; }

void pzo_ntt_tables(const pzo_ntt *t, uint64_t *out)
{
    uint32_t n = t->n;
    memcpy(out, t->psi_bo, 8 * n);
    memcpy(out + n, t->psi_bo_shoup, 8 * n);
    memcpy(out + 2 * n, t->psi_inv_bo, 8 * n);
    memcpy(out + 3 * n, t->psi_inv_bo_shoup, 8 * n);
    out[4 * n] = t->n_inv;
    out[4 * n + 1] = t->n_inv_shoup;
}

void pzo_ntt_forward_lazy(const pzo_ntt *t, uint64_t *v) { ntt_fwd(t, v, 0); }
void pzo_ntt_forward(const pzo_ntt *t, uint64_t *v) { ntt_fwd(t, v, 1); }
void pzo_ntt_backward_lazy(const pzo_ntt *t, uint64_t *v) { ntt_inv(t, v, 0); }
void pzo_ntt_backward(const pzo_ntt *t, uint64_t *v) { ntt_inv(t, v, 1); }

void pzo_ntt_rows(const pzo_ntt *t, int which, uint64_t *v, uint32_t rows)
{
    for (uint32_t r = 0; r < rows; r++) {
        uint64_t *row = v + (size_t)r * t->n;
        switch (which) {
        case 0: ntt_fwd(t, row, 0); break;
        case 1: ntt_fwd(t, row, 1); break;
        case 2: ntt_inv(t, row, 0); break;
        default: ntt_inv(t, row, 1); break;
        }
    }
}

/* negacyclic_mul, src/utils.rs:171-192 (mul = u128 %) */
void pzo_negacyclic_mul_schoolbook(uint64_t q, const uint64_t *a, const uint64_t *b,
                                   uint64_t *r, uint32_t n)
{
    for (uint32_t i = 0; i < n; i++) {
        uint64_t acc = 0;
        for (uint32_t j = 0; j <= i; j++)
            acc = (uint64_t)(((u128)acc + (uint64_t)(((u128)a[j] * b[i - j]) % q)) % q);
        for (uint32_t j = i + 1; j < n; j++)
            acc = (uint64_t)(((u128)acc + q - (uint64_t)(((u128)a[j] * b[n - j + i]) % q)) % q);
        r[i] = acc;
    }
}

/* ------------------------------------------------------------------------- */
/* Decomposer, src/decomposer.rs                                             */
/* ------------------------------------------------------------------------- */
typedef struct {
    uint64_t q, b, b_mask, bby2;
    uint32_t logq, logb, d, ignore_bits;
} decomp;

/* utils::log2, src/utils.rs:230-237 */
static uint32_t ref_log2(uint64_t v)
{
    if ((v & (v - 1)) == 0) return 64 - (uint32_t)__builtin_clzll(v) - 1;
    return 64 - (uint32_t)__builtin_clzll(v);
}

/* DefaultDecomposer::new :158-178 */
static void decomp_init(decomp *dc, uint64_t q, uint32_t logb, uint32_t d)
{
    dc->q = q; dc->logq = ref_log2(q); dc->logb = logb; dc->d = d;
    dc->b = 1ULL << logb; dc->b_mask = dc->b - 1; dc->bby2 = 1ULL << (logb - 1);
    dc->ignore_bits = dc->logq - logb * d;
}

uint32_t pzo_decomposer_ignore_bits(uint64_t q, uint32_t logb, uint32_t d)
{
    decomp dc; decomp_init(&dc, q, logb, d);
    return dc.ignore_bits;
}

/* round_value :302-309 */
uint64_t pzo_round_value(uint64_t v, uint32_t ib)
{
    if (ib == 0) return v;
    uint64_t msb = (v & ((1ULL << ib) - 1)) >> (ib - 1);
    return (v >> ib) + msb;
}

/* decompose_iter :217-234 + DecomposerIter::next :264-299 (release-build
 * wrapping: for v == q the !(q - v) + 1 wraps to 0). */
static inline void decomp_iter(const decomp *dc, uint64_t value, uint64_t *out,
                               size_t stride)
{
    if (value >= (dc->q >> 1)) value = ~(dc->q - value) + 1;
    value = pzo_round_value(value, dc->ignore_bits);
    for (uint32_t i = 0; i < dc->d; i++) {
        uint64_t k = value & dc->b_mask;
        value = (value - k) >> dc->logb;
        uint64_t carry = (k > dc->bby2) || (k == dc->bby2 && (value & 1) == 1);
        uint64_t neg_carry = 0 - carry;
        value = value + carry;
        out[i * stride] = (neg_carry & dc->q) + k - (carry << dc->logb);
    }
}

void pzo_decompose_iter(uint64_t q, uint32_t logb, uint32_t d, const uint64_t *v,
                        uint64_t *digits, size_t count)
{
    decomp dc; decomp_init(&dc, q, logb, d);
    for (size_t i = 0; i < count; i++) decomp_iter(&dc, v[i], digits + i * d, 1);
}

/* decompose_to_vec :180-207 */
void pzo_decompose_to_vec(uint64_t q, uint32_t logb, uint32_t d, const uint64_t *v,
                          uint64_t *digits, size_t count)
{
    decomp dc; decomp_init(&dc, q, logb, d);
    for (size_t c = 0; c < count; c++) {
        uint64_t value = v[c];
        if (value >= (q >> 1)) value = ~(q - value) + 1;
        value = pzo_round_value(value, dc.ignore_bits);
        for (uint32_t i = 0; i < d; i++) {
            uint64_t k = value & dc.b_mask;
            value = (value - k) >> logb;
            if (k > dc.bby2 || (k == dc.bby2 && (value & 1) == 1)) {
                digits[c * d + i] = q - (dc.b - k);
                value = value + 1;
            } else {
                digits[c * d + i] = k;
            }
        }
    }
}

/* recompose :129-140 with gadget_vector :13-21 and ModularOpsU64 add/mul */
uint64_t pzo_recompose(uint64_t q, uint32_t logb, uint32_t d, const uint64_t *limbs)
{
    decomp dc; decomp_init(&dc, q, logb, d);
    modq m; modq_init(&m, q);
    uint64_t value = 0;
    for (uint32_t i = 0; i < d; i++) {
        uint64_t beta = 1ULL << (logb * i + dc.ignore_bits);
        value = mm_add(&m, value, mm_mul(&m, limbs[i], beta));
    }
    return value;
}

/* ------------------------------------------------------------------------- */
/* Evaluator: BoolEvaluator::new, src/bool/evaluator.rs:651-784              */
/* ------------------------------------------------------------------------- */
struct pzo_eval {
    pzo_params p;
    pzo_ntt *ntt;
    modq mq;                 /* RLWE modulus Q */
    uint64_t lwe_mask;
    decomp dec_lwe, dec_a, dec_b, dec_auto, dec_uitos;
    uint32_t *dlog;          /* [br_q] */
    uint32_t embedding;
    uint64_t qby4, true_el, false_el;
    int64_t *auto_el;        /* [w+1] */
    uint32_t *auto_idx;      /* [w+1][N] */
    uint8_t *auto_sign;      /* [w+1][N] */
    uint64_t *test_vec;      /* [6][br_q/2] */
    uint32_t rgsw_rows, auto_rows, scratch_rows;
};

/* generate_auto_map, src/rgsw/keygen.rs:16-41 */
void pzo_generate_auto_map(uint32_t ring_size, int64_t k, uint32_t *idx, uint8_t *sign)
{
    uint64_t two_n = 2ULL * ring_size;
    uint64_t kk = k < 0 ? two_n - ((uint64_t)(-k) % two_n) : (uint64_t)k;
    for (uint32_t i = 0; i < ring_size; i++) {
        uint64_t to = ((uint64_t)i * kk) % two_n;
        uint8_t s = 1;
        if (to >= ring_size) { to -= ring_size; s = 0; }
        idx[i] = (uint32_t)to;
        sign[i] = s;
    }
}

static uint64_t ipow(uint64_t g, uint32_t e)
{
    uint64_t r = 1;
    while (e--) r *= g;
    return r;
}

pzo_eval *pzo_eval_new(const pzo_params *p)
{
    pzo_eval *e = (pzo_eval *)calloc(1, sizeof(*e));
    e->p = *p;
    uint32_t N = p->rlwe_n, brq = p->br_q;
    e->ntt = pzo_ntt_new(p->rlwe_q, N);
    modq_init(&e->mq, p->rlwe_q);
    e->lwe_mask = p->lwe_q - 1;
    decomp_init(&e->dec_lwe, p->lwe_q, p->lwe_logb, p->lwe_d);
    decomp_init(&e->dec_a, p->rlwe_q, p->rlrg_logb, p->rlrg_da);
    decomp_init(&e->dec_b, p->rlwe_q, p->rlrg_logb, p->rlrg_db);
    decomp_init(&e->dec_auto, p->rlwe_q, p->auto_logb, p->auto_d);
    if (p->uitos_d) decomp_init(&e->dec_uitos, p->rlwe_q, p->uitos_logb, p->uitos_d);

    /* dlog map :664-673 */
    e->dlog = (uint32_t *)calloc(brq, sizeof(uint32_t));
    for (uint32_t i = 0; i < brq / 4; i++) {
        uint32_t v = (uint32_t)pzo_mod_exponent(p->g, i, brq);
        e->dlog[v] = i;
        e->dlog[brq - v] = i + brq / 4;
    }
    e->embedding = (2 * N) / brq;

    /* BoolEncoding for CiphertextModulus, :209-270 */
    e->qby4 = p->rlwe_q >> 2;
    e->true_el = p->rlwe_q >> 3;
    e->false_el = (p->rlwe_q - 1) - e->true_el + 1;

    /* test vectors :680-726 */
    uint32_t qby2 = brq >> 1, qby8 = brq >> 3;
    uint32_t *tv_idx = (uint32_t *)malloc(4 * qby2);
    uint8_t *tv_sign = (uint8_t *)malloc(qby2);
    pzo_generate_auto_map(qby2, -(int64_t)p->g, tv_idx, tv_sign);
    e->test_vec = (uint64_t *)calloc(6 * (size_t)qby2, 8);
    struct { uint32_t part; uint64_t before, after; } tv[6] = {
        {3 * qby8, e->true_el, e->false_el},  /* nand */
        {3 * qby8, e->false_el, e->true_el},  /* and */
        {qby8, e->false_el, e->true_el},      /* or */
        {qby8, e->true_el, e->false_el},      /* nor */
        {qby8, e->false_el, e->true_el},      /* xor */
        {qby8, e->true_el, e->false_el},      /* xnor */
    };
    for (int g = 0; g < 6; g++) {
        uint64_t *out = e->test_vec + (size_t)g * qby2;
        for (uint32_t i = 0; i < qby2; i++) {
            uint64_t v = i < tv[g].part ? tv[g].before : tv[g].after;
            out[tv_idx[i]] = tv_sign[i] ? v : mm_neg(&e->mq, v);
        }
    }
    free(tv_idx); free(tv_sign);

    /* auto maps [-g, g^1, ..., g^w] :730-743 */
    e->auto_el = (int64_t *)malloc(sizeof(int64_t) * (p->w + 1));
    e->auto_idx = (uint32_t *)malloc(4 * (size_t)(p->w + 1) * N);
    e->auto_sign = (uint8_t *)malloc((size_t)(p->w + 1) * N);
    for (uint32_t k = 0; k <= p->w; k++) {
        e->auto_el[k] = k == 0 ? -(int64_t)p->g : (int64_t)(ipow(p->g, k) % brq);
        pzo_generate_auto_map(N, e->auto_el[k], e->auto_idx + (size_t)k * N,
                              e->auto_sign + (size_t)k * N);
    }
    e->rgsw_rows = 2 * p->rlrg_da + 2 * p->rlrg_db;
    e->auto_rows = 2 * p->auto_d;
    uint32_t dmax = p->rlrg_da > p->rlrg_db ? p->rlrg_da : p->rlrg_db;
    if (p->auto_d > dmax) dmax = p->auto_d;
    e->scratch_rows = dmax + 2;
    return e;
}

void pzo_eval_free(pzo_eval *e)
{
    if (!e) return;
    pzo_ntt_free(e->ntt);
    free(e->dlog); free(e->test_vec); free(e->auto_el); free(e->auto_idx);
    free(e->auto_sign);
    free(e);
}

const pzo_ntt *pzo_eval_ntt(const pzo_eval *e) { return e->ntt; }
void pzo_eval_dlog_map(const pzo_eval *e, uint32_t *out)
{
    memcpy(out, e->dlog, 4 * (size_t)e->p.br_q);
}
int64_t pzo_eval_auto_element(const pzo_eval *e, uint32_t k) { return e->auto_el[k]; }
void pzo_eval_auto_map(const pzo_eval *e, uint32_t k, uint32_t *idx, uint8_t *sign)
{
    memcpy(idx, e->auto_idx + (size_t)k * e->p.rlwe_n, 4 * (size_t)e->p.rlwe_n);
    memcpy(sign, e->auto_sign + (size_t)k * e->p.rlwe_n, e->p.rlwe_n);
}
void pzo_eval_test_vec(const pzo_eval *e, uint32_t gate, uint64_t *out)
{
    memcpy(out, e->test_vec + (size_t)gate * (e->p.br_q / 2), 8 * (size_t)(e->p.br_q / 2));
}
uint64_t pzo_true_el(const pzo_eval *e) { return e->true_el; }
uint64_t pzo_false_el(const pzo_eval *e) { return e->false_el; }

void pzo_key_sizes_get(const pzo_eval *e, pzo_key_sizes *o)
{
    uint32_t N = e->p.rlwe_n;
    o->rgsw_rows = e->rgsw_rows;
    o->rgsw_words = (size_t)e->p.lwe_n * e->rgsw_rows * N;
    o->auto_rows = e->auto_rows;
    o->auto_words = (size_t)(e->p.w + 1) * e->auto_rows * N;
    o->ksk_rows = (size_t)N * e->p.lwe_d;
    o->ksk_words = o->ksk_rows * (e->p.lwe_n + 1);
}

/* ------------------------------------------------------------------------- */
/* Runtime ops                                                               */
/* ------------------------------------------------------------------------- */

/* decompose_r, src/rgsw/runtime.rs:382-399 */
static void decompose_r(const decomp *dc, const uint64_t *r, uint64_t *out, uint32_t n)
{
    for (uint32_t i = 0; i < n; i++) decomp_iter(dc, r[i], out + i, n);
}

/* rlwe_by_rgsw_shoup, src/rgsw/runtime.rs:754-855.  rlwe = [a | b] (2 x N),
 * rgsw rows [A(-sm) x da, B(-sm) x da, A(m) x db, B(m) x db] (eval domain). */
static void ext_shoup(const pzo_eval *e, uint64_t *rlwe, const uint64_t *rgsw,
                      const uint64_t *rgsw_s, int is_trivial, uint64_t *scr)
{
    uint32_t N = e->p.rlwe_n, da = e->p.rlrg_da, db = e->p.rlrg_db;
    uint64_t q = e->p.rlwe_q;
    uint32_t dmax = da > db ? da : db;
    uint64_t *dec = scr, *ta = scr + (size_t)dmax * N, *tb = ta + N;
    memset(ta, 0, 16 * (size_t)N);
    if (!is_trivial) {
        decompose_r(&e->dec_a, rlwe, dec, N);
        for (uint32_t r = 0; r < da; r++) ntt_fwd(e->ntt, dec + (size_t)r * N, 0);
        pzo_shoup_matrix_fma(q, ta, rgsw, rgsw_s, dec, da, N);
        pzo_shoup_matrix_fma(q, tb, rgsw + (size_t)da * N, rgsw_s + (size_t)da * N, dec, da, N);
    }
    decompose_r(&e->dec_b, rlwe + N, dec, N);
    for (uint32_t r = 0; r < db; r++) ntt_fwd(e->ntt, dec + (size_t)r * N, 0);
    size_t off = (size_t)2 * da * N;
    pzo_shoup_matrix_fma(q, ta, rgsw + off, rgsw_s + off, dec, db, N);
    off += (size_t)db * N;
    pzo_shoup_matrix_fma(q, tb, rgsw + off, rgsw_s + off, dec, db, N);
    ntt_inv(e->ntt, ta, 1);
    ntt_inv(e->ntt, tb, 1);
    memcpy(rlwe, ta, 16 * (size_t)N);
}

/* poly_fma_routine, src/rgsw/runtime.rs:363-372 with elwise_fma_mut
 * (src/backend/modulus_u64.rs:209-213) */
static void poly_fma(const modq *m, uint64_t *out, const uint64_t *ma, const uint64_t *mb,
                     uint32_t rows, uint32_t n)
{
    for (uint32_t r = 0; r < rows; r++)
        for (uint32_t j = 0; j < n; j++)
            out[j] = mm_add(m, out[j], mm_mul(m, ma[(size_t)r * n + j], mb[(size_t)r * n + j]));
}

/* rlwe_by_rgsw (Barrett twin), src/rgsw/runtime.rs:651-747 */
static void ext_plain(const pzo_eval *e, uint64_t *rlwe, const uint64_t *rgsw,
                      int is_trivial, uint64_t *scr)
{
    uint32_t N = e->p.rlwe_n, da = e->p.rlrg_da, db = e->p.rlrg_db;
    uint32_t dmax = da > db ? da : db;
    uint64_t *dec = scr, *ta = scr + (size_t)dmax * N, *tb = ta + N;
    memset(ta, 0, 16 * (size_t)N);
    if (!is_trivial) {
        decompose_r(&e->dec_a, rlwe, dec, N);
        for (uint32_t r = 0; r < da; r++) ntt_fwd(e->ntt, dec + (size_t)r * N, 1);
        poly_fma(&e->mq, ta, dec, rgsw, da, N);
        poly_fma(&e->mq, tb, dec, rgsw + (size_t)da * N, da, N);
    }
    decompose_r(&e->dec_b, rlwe + N, dec, N);
    for (uint32_t r = 0; r < db; r++) ntt_fwd(e->ntt, dec + (size_t)r * N, 1);
    size_t off = (size_t)2 * da * N;
    poly_fma(&e->mq, ta, dec, rgsw + off, db, N);
    poly_fma(&e->mq, tb, dec, rgsw + off + (size_t)db * N, db, N);
    ntt_inv(e->ntt, ta, 1);
    ntt_inv(e->ntt, tb, 1);
    memcpy(rlwe, ta, 16 * (size_t)N);
}

/* rlwe_auto_shoup, src/rgsw/runtime.rs:530-643 (shoup = 1) and rlwe_auto
 * :412-522 (shoup = 0).  k indexes the auto maps [-g, g, ..., g^w]. */
static void auto_op(const pzo_eval *e, uint64_t *rlwe, const uint64_t *ksk,
                    const uint64_t *ksk_s, uint32_t k, int is_trivial, int shoup,
                    uint64_t *scr)
{
    uint32_t N = e->p.rlwe_n, d = e->p.auto_d;
    uint64_t q = e->p.rlwe_q;
    const uint32_t *to = e->auto_idx + (size_t)k * N;
    const uint8_t *sg = e->auto_sign + (size_t)k * N;
    const modq *m = &e->mq;
    uint64_t *a = rlwe, *b = rlwe + N;
    if (!is_trivial) {
        uint64_t *dec = scr, *ta = scr + (size_t)d * N, *tb = ta + N;
        memset(ta, 0, 16 * (size_t)N);
        for (uint32_t i = 0; i < N; i++) {
            uint64_t el = sg[i] ? a[i] : mm_neg(m, a[i]);
            decomp_iter(&e->dec_auto, el, dec + to[i], N);
        }
        for (uint32_t r = 0; r < d; r++) ntt_fwd(e->ntt, dec + (size_t)r * N, shoup ? 0 : 1);
        if (shoup) {
            pzo_shoup_matrix_fma(q, ta, ksk, ksk_s, dec, d, N);
            pzo_shoup_matrix_fma(q, tb, ksk + (size_t)d * N, ksk_s + (size_t)d * N, dec, d, N);
        } else {
            poly_fma(m, ta, dec, ksk, d, N);
            poly_fma(m, tb, dec, ksk + (size_t)d * N, d, N);
        }
        ntt_inv(e->ntt, ta, 1);
        ntt_inv(e->ntt, tb, 1);
        for (uint32_t i = 0; i < N; i++) {
            if (!sg[i]) tb[to[i]] = mm_sub(m, tb[to[i]], b[i]);
            else tb[to[i]] = mm_add(m, tb[to[i]], b[i]);
        }
        memcpy(rlwe, ta, 16 * (size_t)N);
    } else {
        uint64_t *row = scr;
        for (uint32_t i = 0; i < N; i++) row[to[i]] = sg[i] ? b[i] : mm_neg(m, b[i]);
        memcpy(b, row, 8 * (size_t)N);
    }
}

static uint64_t *scratch_new(const pzo_eval *e)
{
    return (uint64_t *)malloc(8 * (size_t)e->scratch_rows * e->p.rlwe_n);
}

void pzo_rlwe_by_rgsw_shoup(const pzo_eval *e, uint64_t *rlwe, const uint64_t *rgsw,
                            const uint64_t *rgsw_shoup, int is_trivial)
{
    uint64_t *s = scratch_new(e);
    ext_shoup(e, rlwe, rgsw, rgsw_shoup, is_trivial, s);
    free(s);
}

void pzo_rlwe_by_rgsw(const pzo_eval *e, uint64_t *rlwe, const uint64_t *rgsw, int is_trivial)
{
    uint64_t *s = scratch_new(e);
    ext_plain(e, rlwe, rgsw, is_trivial, s);
    free(s);
}

void pzo_rlwe_auto_shoup(const pzo_eval *e, uint64_t *rlwe, const uint64_t *ksk,
                         const uint64_t *ksk_shoup, uint32_t k, int is_trivial)
{
    uint64_t *s = scratch_new(e);
    auto_op(e, rlwe, ksk, ksk_shoup, k, is_trivial, 1, s);
    free(s);
}

void pzo_rlwe_auto(const pzo_eval *e, uint64_t *rlwe, const uint64_t *ksk, uint32_t k,
                   int is_trivial)
{
    uint64_t *s = scratch_new(e);
    auto_op(e, rlwe, ksk, NULL, k, is_trivial, 0, s);
    free(s);
}

/* lwe_key_switch, src/lwe.rs:14-44 with ModulusPowerOf2 ops
 * (src/backend/power_of_2.rs:16-18, :64-72).  lwe_out is zeroed first, as
 * pbs() does at src/pbs.rs:136. */
void pzo_lwe_key_switch(const pzo_eval *e, uint64_t *out, const uint64_t *in,
                        const uint64_t *ksk)
{
    uint32_t N = e->p.rlwe_n, n1 = e->p.lwe_n + 1, d = e->p.lwe_d;
    uint64_t mask = e->lwe_mask;
    uint64_t digits[64];
    memset(out, 0, 8 * (size_t)n1);
    for (uint32_t i = 0; i < N; i++) {
        decomp_iter(&e->dec_lwe, in[1 + i], digits, 1);
        for (uint32_t j = 0; j < d; j++) {
            const uint64_t *row = ksk + ((size_t)i * d + j) * n1;
            uint64_t c = digits[j];
            for (uint32_t k = 0; k < n1; k++) out[k] = (out[k] + row[k] * c) & mask;
        }
    }
    out[0] = (out[0] + in[0]) & mask;
}

/* moddown Q -> q_ks in f64, src/pbs.rs:128-132 */
uint64_t pzo_moddown(const pzo_eval *e, uint64_t v)
{
    double lwe_qf = (double)e->p.lwe_q, rlwe_qf = (double)e->p.rlwe_q;
    double x = ((double)v * lwe_qf) / rlwe_qf;
    return (uint64_t)round(x);
}

/* mod_switch_odd, src/pbs.rs:416-420 */
uint32_t pzo_mod_switch_odd(double v, double from_q, double to_q)
{
    uint64_t odd = (uint64_t)floor((v * to_q) / from_q);
    return (uint32_t)(odd + ((odd & 1) ^ 1));
}

/* sample_extract, src/pbs.rs:423-449 */
void pzo_sample_extract(const pzo_eval *e, uint64_t *out, const uint64_t *rlwe, uint32_t index)
{
    uint32_t N = e->p.rlwe_n;
    uint64_t *to = out + 1;
    const uint64_t *from = rlwe;
    for (uint32_t i = 0; i < index + 1; i++) to[i] = from[index - i];
    for (uint32_t i = index + 1; i < N; i++) to[i] = mm_neg(&e->mq, from[N + index - i]);
    out[0] = rlwe[N + index];
}

/* monomial_mul, src/pbs.rs:454-482 */
void pzo_monomial_mul(uint64_t q, const uint64_t *p_in, uint64_t *p_out, uint32_t mon_exp,
                      int mon_sign, uint32_t ring_size)
{
    for (uint32_t i = 0; i < ring_size; i++) {
        uint32_t to = i + mon_exp;
        int s = mon_sign;
        if (to >= ring_size) { to -= ring_size; s = !s; }
        p_out[to] = s ? p_in[i] : q - p_in[i];
    }
}

/* Dlog buckets of the key-switched LWE, src/pbs.rs:146-161.  Returns
 * bucket_start[br_q/2 + 1] and idx[n] (indices ascending within bucket). */
static void make_buckets(const pzo_eval *e, const uint64_t *lwe_ks, uint32_t *start,
                         uint32_t *idx)
{
    uint32_t n = e->p.lwe_n, nb = e->p.br_q >> 1;
    double lq = (double)e->p.lwe_q, bq = (double)e->p.br_q;
    uint32_t *k = (uint32_t *)malloc(4 * (size_t)n);
    memset(start, 0, 4 * (size_t)(nb + 1));
    for (uint32_t i = 0; i < n; i++) {
        uint32_t odd = pzo_mod_switch_odd((double)lwe_ks[1 + i], lq, bq);
        k[i] = e->dlog[odd];
        start[k[i] + 1]++;
    }
    for (uint32_t b = 0; b < nb; b++) start[b + 1] += start[b];
    uint32_t *fill = (uint32_t *)malloc(4 * (size_t)nb);
    memcpy(fill, start, 4 * (size_t)nb);
    for (uint32_t i = 0; i < n; i++) idx[fill[k[i]]++] = i;
    free(fill);
    free(k);
}

/* blind_rotation, src/pbs.rs:240-414.  When ops != NULL the op sequence is
 * recorded instead of executed. */
typedef struct {
    const pzo_eval *e;
    const pzo_key *key;
    uint64_t *rlwe, *scr;
    uint32_t *ops, cap, nops;
    pzo_pbs_stats st;
} br_ctx;

static void br_ext(br_ctx *c, uint32_t s_index, int *is_trivial)
{
    const pzo_eval *e = c->e;
    if (c->ops) {
        if (c->nops < c->cap)
            c->ops[c->nops] = 0u | (s_index << 8) | ((uint32_t)(*is_trivial != 0) << 31);
        c->nops++;
    } else {
        size_t stride = (size_t)e->rgsw_rows * e->p.rlwe_n;
        ext_shoup(e, c->rlwe, c->key->rgsw + s_index * stride,
                  c->key->rgsw_shoup + s_index * stride, *is_trivial, c->scr);
    }
    c->st.ext_products++;
    *is_trivial = 0;
}

static void br_auto(br_ctx *c, uint32_t k, int is_trivial)
{
    const pzo_eval *e = c->e;
    if (c->ops) {
        if (c->nops < c->cap)
            c->ops[c->nops] = 1u | (k << 8) | ((uint32_t)(is_trivial != 0) << 31);
        c->nops++;
    } else {
        size_t stride = (size_t)e->auto_rows * e->p.rlwe_n;
        auto_op(e, c->rlwe, c->key->auto_key + k * stride, c->key->auto_shoup + k * stride,
                k, is_trivial, 1, c->scr);
    }
    c->st.autos++;
    if (is_trivial) c->st.trivial_autos++;
}

static void blind_rotation(br_ctx *c, const uint32_t *start, const uint32_t *idx)
{
    uint32_t q = c->e->p.br_q, w = c->e->p.w;
    int is_trivial = 1;
    uint32_t q_by_4 = q >> 2;
#define BUCKET_LEN(b) (start[(b) + 1] - start[(b)])
    /* -(g^k) */
    uint32_t v = 0;
    for (uint32_t i = q_by_4 - 1; i >= 1; i--) {
        for (uint32_t t = start[q_by_4 + i]; t < start[q_by_4 + i + 1]; t++)
            br_ext(c, idx[t], &is_trivial);
        v += 1;
        if (BUCKET_LEN(q_by_4 + i - 1) != 0 || v == w || i == 1) {
            br_auto(c, v, is_trivial);
            v = 0;
        }
    }
    /* -(g^0) */
    for (uint32_t t = start[q_by_4]; t < start[q_by_4 + 1]; t++) br_ext(c, idx[t], &is_trivial);
    br_auto(c, 0, is_trivial);
    /* +(g^k) */
    v = 0;
    for (uint32_t i = q_by_4 - 1; i >= 1; i--) {
        for (uint32_t t = start[i]; t < start[i + 1]; t++) br_ext(c, idx[t], &is_trivial);
        v += 1;
        if (BUCKET_LEN(i - 1) != 0 || v == w || i == 1) {
            br_auto(c, v, is_trivial);
            v = 0;
        }
    }
    /* +(g^0) */
    for (uint32_t t = start[0]; t < start[1]; t++) br_ext(c, idx[t], &is_trivial);
#undef BUCKET_LEN
}

uint32_t pzo_br_schedule(const pzo_eval *e, const uint64_t *lwe_ks, uint32_t *ops, uint32_t cap)
{
    uint32_t nb = e->p.br_q >> 1;
    uint32_t *start = (uint32_t *)malloc(4 * (size_t)(nb + 1));
    uint32_t *idx = (uint32_t *)malloc(4 * (size_t)e->p.lwe_n);
    make_buckets(e, lwe_ks, start, idx);
    br_ctx c;
    memset(&c, 0, sizeof(c));
    c.e = e; c.ops = ops; c.cap = cap;
    blind_rotation(&c, start, idx);
    free(start); free(idx);
    return c.nops;
}

/* Working memory of one PBS stream, allocated once per stream (the
 * reference's evaluator likewise owns its scratch, src/bool/evaluator.rs:
 * 186-206). */
typedef struct {
    uint64_t *scr;     /* scratch_rows x N */
    uint64_t *ks;      /* n + 1            */
    uint32_t *start;   /* br_q/2 + 1       */
    uint32_t *idx;     /* n                */
    uint64_t *rlwe;    /* 2 x N            */
} pbs_scratch;

static void pbs_scratch_init(const pzo_eval *e, pbs_scratch *w)
{
    w->scr = scratch_new(e);
    w->ks = (uint64_t *)malloc(8 * (size_t)(e->p.lwe_n + 1));
    w->start = (uint32_t *)malloc(4 * (size_t)((e->p.br_q >> 1) + 1));
    w->idx = (uint32_t *)malloc(4 * (size_t)e->p.lwe_n);
    w->rlwe = (uint64_t *)malloc(16 * (size_t)e->p.rlwe_n);
}

static void pbs_scratch_free(pbs_scratch *w)
{
    free(w->scr); free(w->ks); free(w->start); free(w->idx); free(w->rlwe);
}

/* pbs, src/pbs.rs:104-233.  lwe: [N+1] mod Q, overwritten with the output. */
static void pbs_impl(const pzo_eval *e, const pzo_key *key, uint32_t gate, uint64_t *lwe,
                     pzo_pbs_stats *stats, pbs_scratch *w)
{
    uint32_t N = e->p.rlwe_n, brq = e->p.br_q, nb = brq >> 1;
    uint64_t *scr = w->scr, *ks = w->ks;
    uint32_t *start = w->start, *idx = w->idx;
    /* (1) moddown Q -> q_ks */
    for (uint32_t i = 0; i < N + 1; i++) lwe[i] = pzo_moddown(e, lwe[i]);
    /* (2) key switch */
    pzo_lwe_key_switch(e, ks, lwe, key->lwe_ksk);
    /* (3) odd moddown + buckets */
    make_buckets(e, ks, start, idx);
    /* (4) b: rotate the test vector by X^{g*b} */
    double lq = (double)e->p.lwe_q, bq = (double)brq;
    uint32_t g_times_b = (uint32_t)(((uint64_t)e->p.g * pzo_mod_switch_odd((double)ks[0], lq, bq)) % brq);
    uint32_t exp = g_times_b;
    int sign = 1;
    if (exp > nb) { exp -= nb; sign = 0; }
    uint64_t *rlwe = w->rlwe;
    memset(rlwe, 0, 16 * (size_t)N);
    const uint64_t *tvec = e->test_vec + (size_t)gate * nb;
    if (e->embedding == 1) {
        pzo_monomial_mul(e->p.rlwe_q, tvec, rlwe + N, exp, sign, nb);
    } else {
        pzo_monomial_mul(e->p.rlwe_q, tvec, lwe, exp, sign, nb);
        for (uint32_t i = 0; i < nb; i++) rlwe[N + e->embedding * i] = lwe[i];
    }
    /* (5) blind rotation */
    br_ctx c;
    memset(&c, 0, sizeof(c));
    c.e = e; c.key = key; c.rlwe = rlwe; c.scr = scr;
    blind_rotation(&c, start, idx);
    if (stats) *stats = c.st;
    /* (6) sample extract at index 0 */
    pzo_sample_extract(e, lwe, rlwe, 0);
}

void pzo_pbs(const pzo_eval *e, const pzo_key *key, uint32_t gate, uint64_t *lwe,
             pzo_pbs_stats *stats)
{
    pbs_scratch w;
    pbs_scratch_init(e, &w);
    pbs_impl(e, key, gate, lwe, stats, &w);
    pbs_scratch_free(&w);
}

/* _add_and_shift_lwe_cts :2118-2123, _subtract_double_lwe_cts :2126-2133,
 * gates :2163-2250 */
static void gate_linear(const pzo_eval *e, uint32_t gate, uint64_t *c0, const uint64_t *c1)
{
    uint32_t N1 = e->p.rlwe_n + 1;
    const modq *m = &e->mq;
    if (gate <= 3) {
        for (uint32_t i = 0; i < N1; i++) c0[i] = mm_add(m, c0[i], c1[i]);
        c0[0] = mm_add(m, c0[0], e->qby4);
    } else {
        for (uint32_t i = 0; i < N1; i++) c0[i] = mm_sub(m, c0[i], c1[i]);
        for (uint32_t i = 0; i < N1; i++) c0[i] = mm_add(m, c0[i], c0[i]);
    }
}

void pzo_gate(const pzo_eval *e, const pzo_key *key, uint32_t gate, uint64_t *c0,
              const uint64_t *c1)
{
    gate_linear(e, gate, c0, c1);
    pzo_pbs(e, key, gate, c0, NULL);
}

/* not_inplace :2236-2239 */
void pzo_not(const pzo_eval *e, uint64_t *c0)
{
    for (uint32_t i = 0; i < e->p.rlwe_n + 1; i++) c0[i] = mm_neg(&e->mq, c0[i]);
}

/* Independent PBS streams, one per thread (the reference evaluator is
 * single-threaded; this is the multi-core CPU baseline and a fast checker).
 * gates: one gate id per item, or NULL for `gate` throughout; c1: the second
 * operand of each gate's linear step, or NULL for a bare PBS of lwes[i]. */
typedef struct {
    const pzo_eval *e;
    const pzo_key *key;
    uint32_t gate;
    const uint32_t *gates;
    uint64_t *lwes;
    const uint64_t *c1;
    uint32_t begin, end;
} batch_job;

static void *batch_worker(void *arg)
{
    batch_job *j = (batch_job *)arg;
    const size_t N1 = (size_t)j->e->p.rlwe_n + 1;
    pbs_scratch w;
    pbs_scratch_init(j->e, &w);
    for (uint32_t i = j->begin; i < j->end; i++) {
        const uint32_t g = j->gates ? j->gates[i] : j->gate;
        if (j->c1) gate_linear(j->e, g, j->lwes + i * N1, j->c1 + i * N1);
        pbs_impl(j->e, j->key, g, j->lwes + i * N1, NULL, &w);
    }
    pbs_scratch_free(&w);
    return NULL;
}

static double run_batch(const pzo_eval *e, const pzo_key *key, uint32_t gate,
                        const uint32_t *gates, uint64_t *lwes, const uint64_t *c1,
                        uint32_t count, uint32_t threads)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    batch_job jobs[256];
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (uint32_t t = 0; t < threads; t++) {
        jobs[t].e = e; jobs[t].key = key; jobs[t].gate = gate; jobs[t].gates = gates;
        jobs[t].lwes = lwes; jobs[t].c1 = c1;
        jobs[t].begin = (uint32_t)(((uint64_t)count * t) / threads);
        jobs[t].end = (uint32_t)(((uint64_t)count * (t + 1)) / threads);
        pthread_create(&th[t], NULL, batch_worker, &jobs[t]);
    }
    for (uint32_t t = 0; t < threads; t++) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

double pzo_pbs_batch_threads(const pzo_eval *e, const pzo_key *key, uint32_t gate,
                             uint64_t *lwes, uint32_t count, uint32_t threads)
{
    return run_batch(e, key, gate, NULL, lwes, NULL, count, threads);
}

double pzo_gates_batch_threads(const pzo_eval *e, const pzo_key *key, const uint32_t *gates,
                               uint64_t *c0, const uint64_t *c1, uint32_t count,
                               uint32_t threads)
{
    return run_batch(e, key, 0, gates, c0, c1, count, threads);
}

/* ------------------------------------------------------------------------- */
/* Seeded key generation -- the project's own procedure (RNG parity with the */
/* Rust crates is unpinned).  Structure follows the reference:               */
/*   secrets  ClientKey::sk_rlwe/sk_lwe (src/bool/keys.rs:75-154), summed     */
/*            over parties into the ideal secret;                            */
/*   RGSW     secret_key_encrypt_rgsw (src/rgsw/keygen.rs:68-160) with both   */
/*            RLWE'(m) halves explicit, then forward NTT of every row         */
/*            (src/bool/keys.rs:881-885);                                    */
/*   auto     seeded_auto_key_gen / seeded_rlwe_ksk_gen                      */
/*            (src/rgsw/keygen.rs:306-430), rows [a x d, b x d], forward NTT; */
/*   LWE KSK  seeded_lwe_ksk_keygen (src/lwe.rs:46-101), rows [b, a_1..a_n]. */
/* Sampling order: each RLWE row pair draws a (N uniform) then e (N noise).   */
/* ------------------------------------------------------------------------- */
static uint64_t map_i64(uint64_t q, int64_t v)
{
    return v < 0 ? q - (uint64_t)(-v) : (uint64_t)v;
}

void pzo_keygen_secrets(const pzo_eval *e, const uint8_t seed[32], uint32_t parties,
                        int32_t *s_rlwe, int32_t *z_lwe)
{
    uint32_t N = e->p.rlwe_n, n = e->p.lwe_n;
    pzo_rng *master = pzo_rng_new(seed);
    int32_t *sp = (int32_t *)malloc(4 * (size_t)N), *zp = (int32_t *)malloc(4 * (size_t)n);
    memset(s_rlwe, 0, 4 * (size_t)N);
    memset(z_lwe, 0, 4 * (size_t)n);
    for (uint32_t p = 0; p < parties; p++) {
        uint8_t ps[32];
        pzo_rng_fill_bytes(master, ps, 32);
        pzo_rng *r = pzo_rng_new(ps);
        pzo_ternary_hw(r, sp, N, N >> 1);
        if (e->p.lwe_secret_gaussian) {
            for (uint32_t i = 0; i < n; i++) zp[i] = (int32_t)pzo_rng_gauss_trunc(r);
        } else {
            pzo_ternary_hw(r, zp, n, n >> 1);
        }
        for (uint32_t i = 0; i < N; i++) s_rlwe[i] += sp[i];
        for (uint32_t i = 0; i < n; i++) z_lwe[i] += zp[i];
        pzo_rng_free(r);
    }
    free(sp); free(zp);
    pzo_rng_free(master);
}

/* out = a*s (negacyclic) via forward / Barrett pointwise / backward, as the
 * reference's keygen computes it (src/rgsw/keygen.rs:113-116). */
static void mul_by_s(const pzo_eval *e, const uint64_t *a, const uint64_t *s_eval, uint64_t *out)
{
    uint32_t N = e->p.rlwe_n;
    memcpy(out, a, 8 * (size_t)N);
    ntt_fwd(e->ntt, out, 1);
    for (uint32_t i = 0; i < N; i++) out[i] = mm_mul(&e->mq, out[i], s_eval[i]);
    ntt_inv(e->ntt, out, 1);
}

/* one RLWE row pair under s: a uniform, b = a*s + e (+ msg_b); a (+ msg_a) */
static void rlwe_pair(const pzo_eval *e, pzo_rng *r, const uint64_t *s_eval,
                      const uint64_t *msg_a, const uint64_t *msg_b, uint64_t beta,
                      uint64_t *row_a, uint64_t *row_b, uint64_t *tmp)
{
    uint32_t N = e->p.rlwe_n;
    uint64_t q = e->p.rlwe_q;
    const modq *m = &e->mq;
    for (uint32_t i = 0; i < N; i++) row_a[i] = pzo_rng_uniform(r, q);
    for (uint32_t i = 0; i < N; i++) row_b[i] = map_i64(q, pzo_rng_gauss_round(r));
    mul_by_s(e, row_a, s_eval, tmp);
    for (uint32_t i = 0; i < N; i++) row_b[i] = mm_add(m, row_b[i], tmp[i]);
    if (msg_b)
        for (uint32_t i = 0; i < N; i++) row_b[i] = mm_add(m, row_b[i], mm_mul(m, msg_b[i], beta));
    if (msg_a)
        for (uint32_t i = 0; i < N; i++) row_a[i] = mm_add(m, row_a[i], mm_mul(m, msg_a[i], beta));
}

void pzo_keygen_server(const pzo_eval *e, const uint8_t seed[32], const int32_t *s_rlwe,
                       const int32_t *z_lwe, pzo_key *key)
{
    const pzo_params *p = &e->p;
    uint32_t N = p->rlwe_n, n = p->lwe_n, da = p->rlrg_da, db = p->rlrg_db, dau = p->auto_d;
    uint64_t q = p->rlwe_q;
    const modq *m = &e->mq;
    pzo_rng *r = pzo_rng_new(seed);
    uint64_t *s_eval = (uint64_t *)malloc(8 * (size_t)N), *tmp = (uint64_t *)malloc(8 * (size_t)N);
    uint64_t *msg = (uint64_t *)malloc(8 * (size_t)N);
    for (uint32_t i = 0; i < N; i++) s_eval[i] = map_i64(q, s_rlwe[i]);
    ntt_fwd(e->ntt, s_eval, 1);

    /* RGSW(X^{emb * z_i}) */
    size_t rs = (size_t)e->rgsw_rows * N;
    for (uint32_t i = 0; i < n; i++) {
        uint64_t *ct = key->rgsw + i * rs;
        memset(msg, 0, 8 * (size_t)N);
        int64_t si = (int64_t)e->embedding * z_lwe[i];
        if (si < 0) msg[N - (uint32_t)(-si)] = q - 1;
        else msg[si] = 1;
        for (uint32_t j = 0; j < da; j++) {
            uint64_t beta = 1ULL << (p->rlrg_logb * j + e->dec_a.ignore_bits);
            rlwe_pair(e, r, s_eval, msg, NULL, beta, ct + (size_t)j * N,
                      ct + (size_t)(da + j) * N, tmp);
        }
        for (uint32_t j = 0; j < db; j++) {
            uint64_t beta = 1ULL << (p->rlrg_logb * j + e->dec_b.ignore_bits);
            rlwe_pair(e, r, s_eval, NULL, msg, beta, ct + (size_t)(2 * da + j) * N,
                      ct + (size_t)(2 * da + db + j) * N, tmp);
        }
        for (uint32_t row = 0; row < e->rgsw_rows; row++) ntt_fwd(e->ntt, ct + (size_t)row * N, 1);
    }

    /* auto keys for [-g, g^1..g^w]: ksk from -s(X^k) to s */
    size_t as = (size_t)e->auto_rows * N;
    for (uint32_t k = 0; k <= p->w; k++) {
        uint64_t *ct = key->auto_key + k * as;
        const uint32_t *to = e->auto_idx + (size_t)k * N;
        const uint8_t *sg = e->auto_sign + (size_t)k * N;
        for (uint32_t i = 0; i < N; i++) {
            uint64_t el = map_i64(q, s_rlwe[i]);
            msg[to[i]] = sg[i] ? mm_neg(m, el) : el;
        }
        for (uint32_t j = 0; j < dau; j++) {
            uint64_t beta = 1ULL << (p->auto_logb * j + e->dec_auto.ignore_bits);
            rlwe_pair(e, r, s_eval, NULL, msg, beta, ct + (size_t)j * N,
                      ct + (size_t)(dau + j) * N, tmp);
        }
        for (uint32_t row = 0; row < e->auto_rows; row++) ntt_fwd(e->ntt, ct + (size_t)row * N, 1);
    }

    /* LWE KSK from s (dim N) to z (dim n), mod lwe_q */
    uint64_t lq = p->lwe_q, mask = e->lwe_mask;
    size_t n1 = (size_t)n + 1;
    for (uint32_t i = 0; i < N; i++) {
        uint64_t neg_si = (0 - map_i64(lq, s_rlwe[i])) & mask;
        for (uint32_t j = 0; j < p->lwe_d; j++) {
            uint64_t beta = 1ULL << (p->lwe_logb * j + e->dec_lwe.ignore_bits);
            uint64_t *row = key->lwe_ksk + ((size_t)i * p->lwe_d + j) * n1;
            uint64_t az = 0;
            for (uint32_t k = 0; k < n; k++) {
                row[1 + k] = pzo_rng_uniform(r, lq);
                az = (az + row[1 + k] * map_i64(lq, z_lwe[k])) & mask;
            }
            uint64_t b = (az + beta * neg_si) & mask;
            b = (b + map_i64(lq, pzo_rng_gauss_round(r))) & mask;
            row[0] = b;
        }
    }
    free(s_eval); free(tmp); free(msg);
    pzo_rng_free(r);
    pzo_key_shoup(e, key);
}

/* ShoupNonInteractiveServerKeyEvaluationDomain::from, src/bool/keys.rs:1088-1155 */
void pzo_key_shoup(const pzo_eval *e, pzo_key *key)
{
    pzo_key_sizes ks;
    pzo_key_sizes_get(e, &ks);
    uint64_t q = e->p.rlwe_q;
    for (size_t i = 0; i < ks.rgsw_words; i++) key->rgsw_shoup[i] = pzo_shoup_repr(key->rgsw[i], q);
    for (size_t i = 0; i < ks.auto_words; i++) key->auto_shoup[i] = pzo_shoup_repr(key->auto_key[i], q);
}

/* ------------------------------------------------------------------------- */
/* Input path of the non-interactive protocol.                               */
/* ------------------------------------------------------------------------- */

void pzo_keygen_user_secret(const pzo_eval *e, const uint8_t seed[32], int32_t *u)
{
    pzo_rng *r = pzo_rng_new(seed);
    pzo_ternary_hw(r, u, e->p.rlwe_n, e->p.rlwe_n >> 1);
    pzo_rng_free(r);
}

void pzo_keygen_rlwe_ksk(const pzo_eval *e, const uint8_t seed[32], const int32_t *u,
                         const int32_t *s, uint64_t *ksk)
{
    uint32_t N = e->p.rlwe_n, d = e->p.uitos_d;
    uint64_t q = e->p.rlwe_q;
    pzo_rng *r = pzo_rng_new(seed);
    uint64_t *s_eval = (uint64_t *)malloc(8 * (size_t)N), *tmp = (uint64_t *)malloc(8 * (size_t)N);
    uint64_t *msg = (uint64_t *)malloc(8 * (size_t)N);
    for (uint32_t i = 0; i < N; i++) s_eval[i] = map_i64(q, s[i]);
    ntt_fwd(e->ntt, s_eval, 1);
    for (uint32_t i = 0; i < N; i++) msg[i] = map_i64(q, u[i]);
    for (uint32_t j = 0; j < d; j++) {
        uint64_t beta = 1ULL << (e->p.uitos_logb * j + e->dec_uitos.ignore_bits);
        rlwe_pair(e, r, s_eval, NULL, msg, beta, ksk + (size_t)j * N, ksk + (size_t)(d + j) * N,
                  tmp);
    }
    for (uint32_t row = 0; row < 2 * d; row++) ntt_fwd(e->ntt, ksk + (size_t)row * N, 1);
    free(s_eval); free(tmp); free(msg);
    pzo_rng_free(r);
}

void pzo_rlwe_key_switch(const pzo_eval *e, const uint64_t *rlwe_in, const uint64_t *ksk,
                         const uint64_t *ksk_shoup, uint64_t *out)
{
    uint32_t N = e->p.rlwe_n, d = e->p.uitos_d;
    uint64_t q = e->p.rlwe_q;
    const modq *m = &e->mq;
    uint64_t *tmp_row = (uint64_t *)calloc(N, 8);
    uint64_t *tmp = (uint64_t *)calloc((size_t)d * N, 8);
    /* -A, decomposed (decompose_r) and forward_lazy per row */
    for (uint32_t i = 0; i < N; i++) tmp_row[i] = mm_neg(m, rlwe_in[i]);
    decompose_r(&e->dec_uitos, tmp_row, tmp, N);
    for (uint32_t row = 0; row < d; row++) ntt_fwd(e->ntt, tmp + (size_t)row * N, 0);
    /* A' = decomp(-A) . Ksk_A, B' = decomp(-A) . Ksk_B (shoup_matrix_fma) */
    memset(out, 0, 16 * (size_t)N);
    pzo_shoup_matrix_fma(q, out, ksk, ksk_shoup, tmp, d, N);
    pzo_shoup_matrix_fma(q, out + N, ksk + (size_t)d * N, ksk_shoup + (size_t)d * N, tmp, d, N);
    ntt_inv(e->ntt, out, 1);
    ntt_inv(e->ntt, out + N, 1);
    /* B' + B */
    for (uint32_t i = 0; i < N; i++) out[N + i] = mm_add(m, out[N + i], rlwe_in[N + i]);
    free(tmp_row); free(tmp);
}

void pzo_encrypt_rlwe_bools(const pzo_eval *e, const uint8_t seed[32], const int32_t *u,
                            const uint8_t *bits, size_t count, uint64_t *out)
{
    uint32_t N = e->p.rlwe_n;
    uint64_t q = e->p.rlwe_q;
    const modq *m = &e->mq;
    uint8_t sa[32], se[32];
    pzo_rng *master = pzo_rng_new(seed);
    pzo_rng_fill_bytes(master, sa, 32);
    pzo_rng_fill_bytes(master, se, 32);
    pzo_rng_free(master);
    pzo_rng *ra = pzo_rng_new(sa), *re = pzo_rng_new(se);
    uint64_t *u_eval = (uint64_t *)malloc(8 * (size_t)N), *tmp = (uint64_t *)malloc(8 * (size_t)N);
    for (uint32_t i = 0; i < N; i++) u_eval[i] = map_i64(q, u[i]);
    ntt_fwd(e->ntt, u_eval, 1);
    size_t nr = (count + N - 1) / N;
    for (size_t k = 0; k < nr; k++) {
        uint64_t *a = out + k * 2 * N, *b = a + N;
        for (uint32_t i = 0; i < N; i++) a[i] = pzo_rng_uniform(ra, q);
        /* a * u: forward, Barrett pointwise (elwise_mul_mut), backward */
        memcpy(tmp, a, 8 * (size_t)N);
        ntt_fwd(e->ntt, tmp, 1);
        for (uint32_t i = 0; i < N; i++) tmp[i] = mm_mul(m, tmp[i], u_eval[i]);
        ntt_inv(e->ntt, tmp, 1);
        /* b = e + m + a*u, in that order of additions */
        for (uint32_t i = 0; i < N; i++) {
            size_t t = k * N + i;
            uint64_t msg = t < count ? (bits[t] ? e->true_el : e->false_el) : 0;
            uint64_t v = map_i64(q, pzo_rng_gauss_round(re));
            v = mm_add(m, v, msg);
            b[i] = mm_add(m, v, tmp[i]);
        }
    }
    free(u_eval); free(tmp);
    pzo_rng_free(ra); pzo_rng_free(re);
}

/* poly_fma_routine, src/rgsw/runtime.rs:363-372: out += sum_r a[r] . b[r]
 * (elwise_fma_mut: add_mod_fast(out, mul_mod_fast(a, b))) */
static void poly_fma_rows(const modq *m, uint64_t *out, const uint64_t *a, const uint64_t *b,
                          uint32_t rows, uint32_t N)
{
    for (uint32_t r = 0; r < rows; r++)
        for (uint32_t i = 0; i < N; i++)
            out[i] = mm_add(m, out[i], mm_mul(m, a[(size_t)r * N + i], b[(size_t)r * N + i]));
}

void pzo_rgsw_by_rgsw(const pzo_eval *e, uint64_t *rgsw0, const uint64_t *rgsw1)
{
    const pzo_params *p = &e->p;
    uint32_t N = p->rlwe_n, da0 = p->rlrg_da, db0 = p->rlrg_db;
    uint32_t da1 = p->rgrg_da, db1 = p->rgrg_db;
    decomp dca, dcb;
    decomp_init(&dca, p->rlwe_q, p->rgrg_logb, da1);
    decomp_init(&dcb, p->rlwe_q, p->rgrg_logb, db1);
    const uint32_t r0 = 2 * da0 + 2 * db0;
    uint32_t dmax = da1 > db1 ? da1 : db1;
    uint64_t *space = (uint64_t *)calloc((size_t)r0 * N, 8);
    uint64_t *dec = (uint64_t *)calloc((size_t)dmax * N, 8);
    /* rgsw1 parts: nsm A, nsm B (da1 rows each), m A, m B (db1 rows each) */
    const uint64_t *nsm_a = rgsw1, *nsm_b = rgsw1 + (size_t)da1 * N;
    const uint64_t *m_a = rgsw1 + (size_t)2 * da1 * N, *m_b = m_a + (size_t)db1 * N;
    /* RLWE pairs of rgsw0: nsm rows (i, da0 + i), then m rows (2da0 + i, 2da0 + db0 + i) */
    for (uint32_t k = 0; k < da0 + db0; k++) {
        uint32_t ra = k < da0 ? k : 2 * da0 + (k - da0);
        uint32_t rb = k < da0 ? da0 + k : 2 * da0 + db0 + (k - da0);
        const uint64_t *rlwe_a = rgsw0 + (size_t)ra * N, *rlwe_b = rgsw0 + (size_t)rb * N;
        uint64_t *out_a = space + (size_t)ra * N, *out_b = space + (size_t)rb * N;
        decompose_r(&dca, rlwe_a, dec, N);
        for (uint32_t r = 0; r < da1; r++) ntt_fwd(e->ntt, dec + (size_t)r * N, 1);
        poly_fma_rows(&e->mq, out_a, dec, nsm_a, da1, N);
        poly_fma_rows(&e->mq, out_b, dec, nsm_b, da1, N);
        decompose_r(&dcb, rlwe_b, dec, N);
        for (uint32_t r = 0; r < db1; r++) ntt_fwd(e->ntt, dec + (size_t)r * N, 1);
        poly_fma_rows(&e->mq, out_a, dec, m_a, db1, N);
        poly_fma_rows(&e->mq, out_b, dec, m_b, db1, N);
    }
    memcpy(rgsw0, space, 8 * (size_t)r0 * N);
    for (uint32_t r = 0; r < r0; r++) ntt_inv(e->ntt, rgsw0 + (size_t)r * N, 1);
    free(space); free(dec);
}

/* ---- non-interactive multi-party server key: shares and aggregation ----
 * The seed tree follows NonInteractiveMultiPartyCrs
 * (src/bool/evaluator.rs:97-160): puncture(seed, t) draws t 32-byte seeds
 * from ChaCha8(seed), one stream word per byte, and keeps the last
 * (puncture_p_rng, src/utils.rs:219-228);
 *   index l:  puncture(puncture(puncture(crs, 1), 1), l + 1)
 *   user j:   puncture(puncture(crs, 2), j + 1)
 * Noise streams are the oracle's own: ChaCha8(puncture(noise_seed, 1)) for
 * the u->s parts, ChaCha8(puncture(puncture(noise_seed, 2), l + 1)) for LWE
 * index l. */
void pzo_puncture(const uint8_t seed[32], uint32_t times, uint8_t out[32])
{
    /* each of the 32 bytes is Uniform::new_inclusive(0u8, 255u8) (RandomFill<[u8; 32]>,
     * src/random.rs:162-175): over the full range rand 0.8.5 samples Standard, the low
     * byte of next_u32 */
    pzo_rng *r = pzo_rng_new(seed);
    memset(out, 0, 32);
    for (uint32_t t = 0; t < times; t++)
        for (int i = 0; i < 32; i++) out[i] = (uint8_t)pzo_rng_next_u32(r);
    pzo_rng_free(r);
}

static void ni_seed_index(const uint8_t crs[32], uint32_t l, uint8_t out[32])
{
    uint8_t k[32], m[32];
    pzo_puncture(crs, 1, k);
    pzo_puncture(k, 1, m);
    pzo_puncture(m, l + 1, out);
}

static void ni_seed_user(const uint8_t crs[32], uint32_t j, uint8_t out[32])
{
    uint8_t k[32];
    pzo_puncture(crs, 2, k);
    pzo_puncture(k, j + 1, out);
}

/* multi_party_user_id_lwe_segment, src/bool/evaluator.rs:613-626: the
 * leader of LWE index l (ceil(lwe_n / total) indices per party) */
uint32_t pzo_ni_leader(uint32_t l, uint32_t total, uint32_t lwe_n)
{
    uint32_t per = (lwe_n + total - 1) / total;
    return l / per;
}

/* out = a*x + e + beta*msg (msg may be NULL): the row recipe of
 * non_interactive_ksk_gen / zero encryptions (src/multi_party.rs:186-291)
 * and of the NI RGSW cts (src/bool/evaluator.rs:1744-1937); e drawn N times
 * from `noise` */
static void ni_row(const pzo_eval *e, const uint64_t *a, const uint64_t *x_eval, pzo_rng *noise,
                   const uint64_t *msg, uint64_t beta, uint64_t *out)
{
    uint32_t N = e->p.rlwe_n;
    const modq *m = &e->mq;
    mul_by_s(e, a, x_eval, out);
    for (uint32_t i = 0; i < N; i++)
        out[i] = mm_add(m, out[i], map_i64(e->p.rlwe_q, pzo_rng_gauss_round(noise)));
    if (msg)
        for (uint32_t i = 0; i < N; i++) out[i] = mm_add(m, out[i], mm_mul(m, msg[i], beta));
}

static uint64_t *secret_eval(const pzo_eval *e, const int32_t *x)
{
    uint32_t N = e->p.rlwe_n;
    uint64_t *v = (uint64_t *)malloc(8 * (size_t)N);
    for (uint32_t i = 0; i < N; i++) v[i] = map_i64(e->p.rlwe_q, x[i]);
    ntt_fwd(e->ntt, v, 1);
    return v;
}

void pzo_ni_share_uitos(const pzo_eval *e, const uint8_t crs[32], const uint8_t noise_seed[32],
                        uint32_t self, uint32_t total, const int32_t *s, const int32_t *u,
                        uint64_t *ksk_b, uint64_t *zero_others)
{
    uint32_t N = e->p.rlwe_n, d = e->p.uitos_d;
    uint64_t q = e->p.rlwe_q;
    uint8_t ns[32], sd[32];
    pzo_puncture(noise_seed, 1, ns);
    pzo_rng *noise = pzo_rng_new(ns);
    uint64_t *s_eval = secret_eval(e, s), *msg = (uint64_t *)malloc(8 * (size_t)N);
    uint64_t *a = (uint64_t *)malloc(8 * (size_t)N);
    for (uint32_t i = 0; i < N; i++) msg[i] = map_i64(q, u[i]);
    ni_seed_user(crs, self, sd);
    pzo_rng *pr = pzo_rng_new(sd);
    for (uint32_t j = 0; j < d; j++) {
        for (uint32_t i = 0; i < N; i++) a[i] = pzo_rng_uniform(pr, q);
        uint64_t beta = 1ULL << (e->p.uitos_logb * j + e->dec_uitos.ignore_bits);
        ni_row(e, a, s_eval, noise, msg, beta, ksk_b + (size_t)j * N);
    }
    pzo_rng_free(pr);
    uint32_t slot = 0;
    for (uint32_t o = 0; o < total; o++) {
        if (o == self) continue;
        ni_seed_user(crs, o, sd);
        pr = pzo_rng_new(sd);
        for (uint32_t j = 0; j < d; j++) {
            for (uint32_t i = 0; i < N; i++) a[i] = pzo_rng_uniform(pr, q);
            ni_row(e, a, s_eval, noise, NULL, 0, zero_others + ((size_t)slot * d + j) * N);
        }
        pzo_rng_free(pr);
        slot++;
    }
    pzo_rng_free(noise);
    free(s_eval); free(msg); free(a);
}

/* a_{i,l}, i in [0, d_max): the rows of the index-l stream (the RGSW x RGSW
 * decomposer's larger count) */
static void ni_index_rows(const pzo_eval *e, const uint8_t crs[32], uint32_t l, uint32_t dmax,
                          uint64_t *rows)
{
    uint32_t N = e->p.rlwe_n;
    uint8_t sd[32];
    ni_seed_index(crs, l, sd);
    pzo_rng *pr = pzo_rng_new(sd);
    for (size_t i = 0; i < (size_t)dmax * N; i++) rows[i] = pzo_rng_uniform(pr, e->p.rlwe_q);
    pzo_rng_free(pr);
}

static uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

void pzo_ni_share_rgsw(const pzo_eval *e, const uint8_t crs[32], const uint8_t noise_seed[32],
                       uint32_t self, uint32_t total, const int32_t *s, const int32_t *u,
                       const int32_t *z, uint32_t l0, uint32_t count, uint64_t *zero,
                       uint64_t *cts)
{
    const pzo_params *p = &e->p;
    uint32_t N = p->rlwe_n, da = p->rgrg_da, dmax = umax(p->rgrg_da, p->rgrg_db);
    uint32_t dmax1 = umax(p->rlrg_da, p->rlrg_db);
    uint32_t lq = ref_log2(p->rlwe_q);
    uint64_t q = p->rlwe_q;
    uint8_t nr[32], nl[32];
    pzo_puncture(noise_seed, 2, nr);
    uint64_t *s_eval = secret_eval(e, s), *u_eval = secret_eval(e, u);
    uint64_t *a = (uint64_t *)malloc(8 * (size_t)dmax * N), *msg = (uint64_t *)malloc(8 * (size_t)N);
    for (uint32_t t = 0; t < count; t++) {
        uint32_t l = l0 + t;
        ni_index_rows(e, crs, l, dmax, a);
        pzo_puncture(nr, l + 1, nl);
        pzo_rng *noise = pzo_rng_new(nl);
        /* zero encryptions a_i*s_j + e, i in [dmax - da, dmax) */
        for (uint32_t k = 0; k < da; k++)
            ni_row(e, a + (size_t)(dmax - da + k) * N, s_eval, noise, NULL, 0,
                   zero + ((size_t)t * da + k) * N);
        /* u_j*a_i + e + beta*X^{emb*z_j[l]} (encode_x_pow_si_with_emebedding_factor,
         * src/utils.rs:195-217) */
        memset(msg, 0, 8 * (size_t)N);
        int64_t si = (int64_t)e->embedding * z[l];
        if (si < 0) msg[N - (uint32_t)(-si)] = q - 1;
        else msg[si] = 1;
        uint64_t *ct = cts + (size_t)t * dmax * N;
        memset(ct, 0, 8 * (size_t)dmax * N);
        if (pzo_ni_leader(l, total, p->lwe_n) == self) {
            uint32_t ib = lq - p->rlrg_logb * dmax1;
            for (uint32_t k = 0; k < dmax1; k++)
                ni_row(e, a + (size_t)(dmax - dmax1 + k) * N, u_eval, noise, msg,
                       1ULL << (ib + p->rlrg_logb * k), ct + (size_t)(dmax - dmax1 + k) * N);
        } else {
            uint32_t ib = lq - p->rgrg_logb * dmax;
            for (uint32_t k = 0; k < dmax; k++)
                ni_row(e, a + (size_t)k * N, u_eval, noise, msg, 1ULL << (ib + p->rgrg_logb * k),
                       ct + (size_t)k * N);
        }
        pzo_rng_free(noise);
    }
    free(s_eval); free(u_eval); free(a); free(msg);
}

void pzo_ni_uitos_ksk(const pzo_eval *e, const uint8_t crs[32], uint32_t total, uint32_t user,
                      const uint64_t *ksk_b, const uint64_t *zero_for_user, uint64_t *ksk)
{
    uint32_t N = e->p.rlwe_n, d = e->p.uitos_d;
    uint8_t sd[32];
    ni_seed_user(crs, user, sd);
    pzo_rng *pr = pzo_rng_new(sd);
    for (size_t i = 0; i < (size_t)d * N; i++) ksk[i] = pzo_rng_uniform(pr, e->p.rlwe_q);
    pzo_rng_free(pr);
    uint64_t *b = ksk + (size_t)d * N;
    memcpy(b, ksk_b, 8 * (size_t)d * N);
    for (uint32_t o = 0; o + 1 < total; o++)
        for (size_t i = 0; i < (size_t)d * N; i++)
            b[i] = mm_add(&e->mq, b[i], zero_for_user[(size_t)o * d * N + i]);
    for (uint32_t row = 0; row < 2 * d; row++) ntt_fwd(e->ntt, ksk + (size_t)row * N, 1);
}

/* produce_rgsw_ciphertext_from_ni_rgsw, src/bool/evaluator.rs:480-608, for
 * LWE index l and one party: counts (da_o, db_o) of the output decomposer
 * (leader: RLWE x RGSW, coefficient domain; others: RGSW x RGSW, evaluation
 * domain).  zero_l [total][da][N] are every party's zero encryptions of
 * index l; ct [dmax][N] the party's NI RGSW rows, indexed by the a_{i,l} row
 * i; ksk [2d][N] the party's Ksk(u_j -> s) in the evaluation domain. */
void pzo_ni_produce(const pzo_eval *e, const uint8_t crs[32], uint32_t total, uint32_t l,
                    const uint64_t *ksk, const uint64_t *zero_l, const uint64_t *ct, int leader,
                    uint64_t *out)
{
    const pzo_params *p = &e->p;
    const modq *m = &e->mq;
    uint32_t N = p->rlwe_n, d = p->uitos_d, da = p->rgrg_da, db = p->rgrg_db;
    uint32_t dmax = umax(da, db);
    uint32_t dao = leader ? p->rlrg_da : da, dbo = leader ? p->rlrg_db : db;
    uint64_t *a = (uint64_t *)malloc(8 * (size_t)dmax * N);
    uint64_t *c = (uint64_t *)malloc(8 * (size_t)N), *dig = (uint64_t *)malloc(8 * (size_t)d * N);
    uint64_t *tmp = (uint64_t *)malloc(8 * (size_t)N);
    ni_index_rows(e, crs, l, dmax, a);
    for (uint32_t k = 0; k < dao + dbo; k++) {
        int nsm = k < dao;
        uint32_t i = nsm ? dmax - dao + k : dmax - dbo + (k - dao);
        uint64_t *oa = out + (size_t)(nsm ? k : 2 * dao + (k - dao)) * N;
        uint64_t *ob = out + (size_t)(nsm ? dao + k : 2 * dao + dbo + (k - dao)) * N;
        if (nsm) {
            /* sum_k s_k a_{i,l} + e: aggregated zero encryptions */
            memset(c, 0, 8 * (size_t)N);
            for (uint32_t o = 0; o < total; o++)
                for (uint32_t x = 0; x < N; x++)
                    c[x] = mm_add(m, c[x], zero_l[((size_t)o * da + (i - (dmax - da))) * N + x]);
        } else {
            for (uint32_t x = 0; x < N; x++) c[x] = mm_neg(m, a[(size_t)i * N + x]) % p->rlwe_q;
        }
        /* key switch with Ksk(u_j -> s): decomposed, forward, elwise_fma_mut */
        decompose_r(&e->dec_uitos, c, dig, N);
        for (uint32_t r = 0; r < d; r++) ntt_fwd(e->ntt, dig + (size_t)r * N, 1);
        memset(oa, 0, 8 * (size_t)N);
        memset(ob, 0, 8 * (size_t)N);
        poly_fma_rows(m, oa, dig, ksk, d, N);
        poly_fma_rows(m, ob, dig, ksk + (size_t)d * N, d, N);
        uint64_t *dst = nsm ? oa : ob;
        memcpy(tmp, ct + (size_t)i * N, 8 * (size_t)N);
        if (leader) {
            ntt_inv(e->ntt, oa, 1);
            ntt_inv(e->ntt, ob, 1);
        } else {
            ntt_fwd(e->ntt, tmp, 1);
        }
        for (uint32_t x = 0; x < N; x++) dst[x] = mm_add(m, dst[x], tmp[x]);
    }
    free(a); free(c); free(dig); free(tmp);
}

/* The RGSW part of aggregate_non_interactive_multi_party_server_key_shares
 * (src/bool/evaluator.rs:1222-1565) for LWE indices [l0, l0 + count):
 * the leader's RGSW times, in ascending party order, every other party's. */
void pzo_ni_aggregate_rgsw(const pzo_eval *e, const uint8_t crs[32], uint32_t total,
                           const uint64_t *ksks, uint32_t l0, uint32_t count,
                           const uint64_t *zero, const uint64_t *cts, uint64_t *out)
{
    const pzo_params *p = &e->p;
    uint32_t N = p->rlwe_n, d = p->uitos_d, da = p->rgrg_da, dmax = umax(p->rgrg_da, p->rgrg_db);
    size_t r0 = e->rgsw_rows, r1 = 2 * (size_t)(p->rgrg_da + p->rgrg_db);
    uint64_t *zl = (uint64_t *)malloc(8 * (size_t)total * da * N);
    uint64_t *other = (uint64_t *)malloc(8 * r1 * N);
    for (uint32_t t = 0; t < count; t++) {
        uint32_t l = l0 + t, lead = pzo_ni_leader(l, total, p->lwe_n);
        for (uint32_t o = 0; o < total; o++)
            memcpy(zl + (size_t)o * da * N, zero + ((size_t)o * count + t) * da * N,
                   8 * (size_t)da * N);
        uint64_t *rg = out + (size_t)t * r0 * N;
        pzo_ni_produce(e, crs, total, l, ksks + (size_t)lead * 2 * d * N, zl,
                       cts + ((size_t)lead * count + t) * dmax * N, 1, rg);
        for (uint32_t o = 0; o < total; o++) {
            if (o == lead) continue;
            pzo_ni_produce(e, crs, total, l, ksks + (size_t)o * 2 * d * N, zl,
                           cts + ((size_t)o * count + t) * dmax * N, 0, other);
            pzo_rgsw_by_rgsw(e, rg, other);
        }
    }
    free(zl); free(other);
}

/* encrypt_lwe, src/lwe.rs:104-137 (operator = RLWE modop, secret of dim) */
void pzo_encrypt_lwe(const pzo_eval *e, const uint8_t seed[32], const int32_t *s, uint32_t dim,
                     uint64_t m, uint64_t *out)
{
    uint64_t q = e->p.rlwe_q;
    const modq *mq = &e->mq;
    pzo_rng *r = pzo_rng_new(seed);
    uint64_t sa = 0;
    for (uint32_t i = 0; i < dim; i++) out[1 + i] = pzo_rng_uniform(r, q);
    for (uint32_t i = 0; i < dim; i++) sa = mm_add(mq, mm_mul(mq, out[1 + i], map_i64(q, s[i])), sa);
    uint64_t err = map_i64(q, pzo_rng_gauss_round(r));
    out[0] = mm_add(mq, mm_add(mq, sa, err), m);
    pzo_rng_free(r);
}

/* decrypt_lwe, src/lwe.rs:139-162 */
uint64_t pzo_decrypt_lwe(const pzo_eval *e, const uint64_t *lwe, const int32_t *s, uint32_t dim)
{
    uint64_t q = e->p.rlwe_q;
    const modq *mq = &e->mq;
    uint64_t sa = 0;
    for (uint32_t i = 0; i < dim; i++) sa = mm_add(mq, mm_mul(mq, lwe[1 + i], map_i64(q, s[i])), sa);
    return mm_sub(mq, lwe[0], sa);
}

/* BoolEncoding::decode, src/bool/evaluator.rs:245-258; -1 on a bad value */
int pzo_decode_bool(const pzo_eval *e, uint64_t m)
{
    uint64_t qby8 = e->true_el;
    double x = ((double)(m + qby8) * 4.0) / (double)e->p.rlwe_q;
    uint64_t v = ((uint64_t)round(x)) % 4;
    if (v == 0) return 0;
    if (v == 1) return 1;
    return -1;
}

/* ---- non-interactive automorphism-key and LWE-KSK shares, their sums and
 * the seeded server key's expansion ----
 * Seeds (NonInteractiveMultiPartyCrs, src/bool/evaluator.rs:148-156):
 *   auto keys a parts:  puncture(puncture(crs, 1), 2)
 *   LWE KSK a parts:    puncture(puncture(crs, 1), 3)
 * Noise streams are the oracle's own: ChaCha8(puncture(noise_seed, 3)) for
 * the automorphism keys, ChaCha8(puncture(noise_seed, 4)) for the LWE KSK. */
static void ni_seed_key_part(const uint8_t crs[32], uint32_t part, uint8_t out[32])
{
    uint8_t k[32];
    pzo_puncture(crs, 1, k);
    pzo_puncture(k, part, out);
}

/* _common_rountine_multi_party_auto_keys_share_gen and
 * _common_rountine_multi_party_lwe_ksk_share_gen (src/bool/evaluator.rs:
 * 1952-1965, 1981-2046) over seeded_auto_key_gen / seeded_rlwe_ksk_gen
 * (src/rgsw/keygen.rs:306-430) and seeded_lwe_ksk_keygen (src/lwe.rs:46-101):
 *   auto_b[k][j] = a_{k,j}*s + e + beta_j*(-s(X^{el_k})), k over the dlogs
 *                  [0, 1, .., w] (el_0 = -g, el_k = g^k), j < d_auto, the
 *                  a rows drawn in that order from the auto-key seed;
 *   ksk_b[i*d + j] = <a_{i,j}, z> + beta_j*(-s_i) + e mod q_ks, the n
 *                  entries of each a row drawn in row order from the KSK
 *                  seed. */
void pzo_ni_share_auto_ksk(const pzo_eval *e, const uint8_t crs[32], const uint8_t noise_seed[32],
                           const int32_t *s, const int32_t *z, uint64_t *auto_b, uint64_t *ksk_b)
{
    const pzo_params *p = &e->p;
    uint32_t N = p->rlwe_n, n = p->lwe_n, dau = p->auto_d;
    uint64_t q = p->rlwe_q;
    const modq *m = &e->mq;
    uint8_t sd[32], ns[32];
    uint64_t *s_eval = secret_eval(e, s);
    uint64_t *a = (uint64_t *)malloc(8 * (size_t)N), *msg = (uint64_t *)malloc(8 * (size_t)N);
    /* automorphism keys */
    ni_seed_key_part(crs, 2, sd);
    pzo_rng *pr = pzo_rng_new(sd);
    pzo_puncture(noise_seed, 3, ns);
    pzo_rng *noise = pzo_rng_new(ns);
    for (uint32_t k = 0; k <= p->w; k++) {
        const uint32_t *to = e->auto_idx + (size_t)k * N;
        const uint8_t *sg = e->auto_sign + (size_t)k * N;
        for (uint32_t i = 0; i < N; i++) {  /* -s(X^k): the sign opposite to the map's */
            uint64_t el = map_i64(q, s[i]);
            msg[to[i]] = sg[i] ? mm_neg(m, el) : el;
        }
        for (uint32_t j = 0; j < dau; j++) {
            for (uint32_t i = 0; i < N; i++) a[i] = pzo_rng_uniform(pr, q);
            uint64_t beta = 1ULL << (p->auto_logb * j + e->dec_auto.ignore_bits);
            ni_row(e, a, s_eval, noise, msg, beta, auto_b + ((size_t)k * dau + j) * N);
        }
    }
    pzo_rng_free(pr);
    pzo_rng_free(noise);
    /* LWE key-switching key from s (dimension N) to z (dimension n) */
    ni_seed_key_part(crs, 3, sd);
    pr = pzo_rng_new(sd);
    pzo_puncture(noise_seed, 4, ns);
    noise = pzo_rng_new(ns);
    uint64_t lq = p->lwe_q, mask = e->lwe_mask;
    for (uint32_t i = 0; i < N; i++) {
        uint64_t neg_si = (0 - map_i64(lq, s[i])) & mask;
        for (uint32_t j = 0; j < p->lwe_d; j++) {
            uint64_t beta = 1ULL << (p->lwe_logb * j + e->dec_lwe.ignore_bits);
            uint64_t az = 0;
            for (uint32_t k = 0; k < n; k++)
                az = (az + pzo_rng_uniform(pr, lq) * map_i64(lq, z[k])) & mask;
            uint64_t b = (az + beta * neg_si) & mask;
            ksk_b[(size_t)i * p->lwe_d + j] = (b + map_i64(lq, pzo_rng_gauss_round(noise))) & mask;
        }
    }
    pzo_rng_free(pr);
    pzo_rng_free(noise);
    free(s_eval); free(a); free(msg);
}

/* The automorphism-key and LWE-KSK parts of
 * aggregate_non_interactive_multi_party_server_key_shares
 * (src/bool/evaluator.rs:1574-1612): sums of the parties' b parts, mod Q and
 * mod q_ks.  shares: [total][(w+1) d_auto N] and [total][N d_ks]. */
void pzo_ni_aggregate_auto_ksk(const pzo_eval *e, uint32_t total, const uint64_t *auto_shares,
                               const uint64_t *ksk_shares, uint64_t *auto_b, uint64_t *ksk_b)
{
    const pzo_params *p = &e->p;
    size_t aw = (size_t)(p->w + 1) * p->auto_d * p->rlwe_n, kw = (size_t)p->rlwe_n * p->lwe_d;
    memset(auto_b, 0, 8 * aw);
    memset(ksk_b, 0, 8 * kw);
    for (uint32_t t = 0; t < total; t++) {
        for (size_t i = 0; i < aw; i++) auto_b[i] = mm_add(&e->mq, auto_b[i], auto_shares[t * aw + i]);
        for (size_t i = 0; i < kw; i++) ksk_b[i] = (ksk_b[i] + ksk_shares[t * kw + i]) & e->lwe_mask;
    }
}

/* NonInteractiveServerKeyEvaluationDomain::from(&SeededNonInteractive
 * MultiPartyServerKey) (src/bool/keys.rs:851-945) and its Shoup twins
 * (:1088-1155): the RGSW rows (coefficient domain) sent to the evaluation
 * domain; each automorphism key's a rows drawn from the auto-key seed, its b
 * rows copied, every row sent to the evaluation domain; each LWE-KSK row's n
 * a entries drawn from the KSK seed after its b. */
void pzo_ni_expand_server_key(const pzo_eval *e, const uint8_t crs[32], const uint64_t *rgsw,
                              const uint64_t *auto_b, const uint64_t *ksk_b, pzo_key *key)
{
    const pzo_params *p = &e->p;
    uint32_t N = p->rlwe_n, n = p->lwe_n, dau = p->auto_d;
    size_t rw = (size_t)n * e->rgsw_rows * N;
    memcpy(key->rgsw, rgsw, 8 * rw);
    for (size_t r = 0; r < (size_t)n * e->rgsw_rows; r++) ntt_fwd(e->ntt, key->rgsw + r * N, 1);
    uint8_t sd[32];
    ni_seed_key_part(crs, 2, sd);
    pzo_rng *pr = pzo_rng_new(sd);
    size_t as = (size_t)e->auto_rows * N;
    for (uint32_t k = 0; k <= p->w; k++) {
        uint64_t *ct = key->auto_key + k * as;
        for (size_t i = 0; i < (size_t)dau * N; i++) ct[i] = pzo_rng_uniform(pr, p->rlwe_q);
        memcpy(ct + (size_t)dau * N, auto_b + (size_t)k * dau * N, 8 * (size_t)dau * N);
        for (uint32_t r = 0; r < e->auto_rows; r++) ntt_fwd(e->ntt, ct + (size_t)r * N, 1);
    }
    pzo_rng_free(pr);
    ni_seed_key_part(crs, 3, sd);
    pr = pzo_rng_new(sd);
    size_t n1 = (size_t)n + 1;
    for (size_t r = 0; r < (size_t)N * p->lwe_d; r++) {
        uint64_t *row = key->lwe_ksk + r * n1;
        for (uint32_t k = 0; k < n; k++) row[1 + k] = pzo_rng_uniform(pr, p->lwe_q);
        row[0] = ksk_b[r];
    }
    pzo_rng_free(pr);
    pzo_key_shoup(e, key);
}
