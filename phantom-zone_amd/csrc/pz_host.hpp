// pz_host.hpp -- host-side set-up of an evaluator: parameter sets, the
// ChaCha8 generator that fixes the NTT root, NTT tables, Barrett constants,
// dlog map, automorphism exponents and the gate test vectors.
//
// These mirror BoolEvaluator::new (src/bool/evaluator.rs:651-784) and
// NttBackendU64::_new (src/ntt.rs:323-386); they run once per context.
#pragma once
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pz.h"

namespace pz {
namespace host {

typedef unsigned __int128 u128;

// Named parameter sets, src/bool/parameters.rs:526-726.
inline int params_named(const char *name, pz_params *o)
{
    memset(o, 0, sizeof(*o));
    o->rlwe_q = 18014398509404161ULL;
    o->rlwe_n = 2048;
    o->rlrg_logb = 17; o->rlrg_da = 1; o->rlrg_db = 1;
    o->auto_logb = 24; o->auto_d = 1;
    o->lwe_logb = 1;
    o->g = 5; o->w = 10;
    std::string s(name);
    if (s == "NI_2P") {
        o->lwe_q = 1ULL << 16; o->br_q = 1u << 12; o->lwe_n = 520; o->lwe_d = 13;
        o->lwe_secret_gaussian = 1; o->uitos_logb = 1; o->uitos_d = 50;
    } else if (s == "NI_4P_HB_FR") {
        o->lwe_q = 1ULL << 16; o->br_q = 1u << 11; o->lwe_n = 620; o->lwe_d = 13;
        o->uitos_logb = 1; o->uitos_d = 50;
    } else if (s == "NI_8P") {
        o->lwe_q = 1ULL << 17; o->br_q = 1u << 12; o->lwe_n = 660; o->lwe_d = 14;
        o->uitos_logb = 1; o->uitos_d = 50;
    } else if (s == "SP_TEST") {
        o->rlwe_q = 268369921ULL; o->rlwe_n = 512; o->lwe_q = 1ULL << 16;
        o->br_q = 1u << 9; o->lwe_n = 100; o->lwe_logb = 4; o->lwe_d = 4;
        o->rlrg_logb = 7; o->rlrg_da = 4; o->rlrg_db = 4;
        o->auto_logb = 7; o->auto_d = 4; o->w = 5; o->lwe_secret_gaussian = 1;
    } else if (s == "I_2P_LB_SR") {
        o->lwe_q = 1ULL << 15; o->br_q = 1u << 11; o->lwe_n = 580; o->lwe_d = 12;
    } else if (s == "I_4P") {
        o->lwe_q = 1ULL << 16; o->br_q = 1u << 11; o->lwe_n = 620; o->lwe_d = 13;
    } else if (s == "I_8P") {
        o->lwe_q = 1ULL << 17; o->br_q = 1u << 12; o->lwe_n = 660; o->lwe_d = 14;
    } else if (s == "NI_4P_LB_SR") {
        o->lwe_q = 1ULL << 16; o->br_q = 1u << 12; o->lwe_n = 620; o->lwe_d = 13;
        o->uitos_logb = 1; o->uitos_d = 50;
    } else {
        return -1;
    }
    // RGSW x RGSW decomposers (rgrg_decomposer_params, src/bool/parameters.rs:526-726)
    struct { const char *name; uint32_t logb, da, db; } rg[] = {
        {"NI_2P", 4, 10, 9}, {"NI_4P_HB_FR", 3, 13, 12}, {"NI_8P", 2, 20, 18},
        {"I_2P_LB_SR", 7, 6, 5}, {"I_4P", 6, 7, 6}, {"I_8P", 5, 9, 8}, {"NI_4P_LB_SR", 4, 10, 9}};
    for (auto &e : rg)
        if (s == e.name) {
            o->rgrg_logb = e.logb;
            o->rgrg_da = e.da;
            o->rgrg_db = e.db;
        }
    return 0;
}

// ChaCha with rand_chacha's state layout (64-bit counter in words 12-13,
// 64-bit stream in 14-15); 8 rounds = ChaCha8Rng.  Words are consumed in
// order; next_u64 = lo | hi << 32 (rand_core BlockRng).
struct ChaCha8 {
    uint32_t key[8];
    uint64_t counter = 0;
    uint32_t buf[16];
    int idx = 16;

    explicit ChaCha8(const uint8_t seed[32])
    {
        for (int i = 0; i < 8; i++)
            key[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) |
                     ((uint32_t)seed[4 * i + 2] << 16) | ((uint32_t)seed[4 * i + 3] << 24);
    }
    static uint32_t rotl(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }
    static void qr(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d)
    {
        a += b; d ^= a; d = rotl(d, 16);
        c += d; b ^= c; b = rotl(b, 12);
        a += b; d ^= a; d = rotl(d, 8);
        c += d; b ^= c; b = rotl(b, 7);
    }
    void refill()
    {
        uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int i = 0; i < 8; i++) s[4 + i] = key[i];
        s[12] = (uint32_t)counter; s[13] = (uint32_t)(counter >> 32);
        s[14] = 0; s[15] = 0;
        uint32_t x[16];
        memcpy(x, s, sizeof(s));
        for (int r = 0; r < 8; r += 2) {
            qr(x[0], x[4], x[8], x[12]); qr(x[1], x[5], x[9], x[13]);
            qr(x[2], x[6], x[10], x[14]); qr(x[3], x[7], x[11], x[15]);
            qr(x[0], x[5], x[10], x[15]); qr(x[1], x[6], x[11], x[12]);
            qr(x[2], x[7], x[8], x[13]); qr(x[3], x[4], x[9], x[14]);
        }
        for (int i = 0; i < 16; i++) buf[i] = x[i] + s[i];
        counter++;
        idx = 0;
    }
    uint32_t next_u32()
    {
        if (idx >= 16) refill();
        return buf[idx++];
    }
    uint64_t next_u64()
    {
        uint64_t lo = next_u32();
        uint64_t hi = next_u32();
        return lo | (hi << 32);
    }
    void fill_bytes(uint8_t *out, size_t n)
    {
        size_t i = 0;
        while (i < n) {
            uint32_t w = next_u32();
            for (int b = 0; b < 4 && i < n; b++, i++) out[i] = (uint8_t)(w >> (8 * b));
        }
    }
    // UniformInt<u64> sampling (widening multiply + rejection zone)
    uint64_t uniform(uint64_t range)
    {
        uint64_t reject = (UINT64_MAX - range + 1) % range;
        uint64_t zone = UINT64_MAX - reject;
        for (;;) {
            uint64_t v = next_u64();
            u128 m = (u128)v * range;
            if ((uint64_t)m <= zone) return (uint64_t)(m >> 64);
        }
    }
};

inline uint64_t mod_exp(uint64_t a, uint64_t b, uint64_t q)
{
    uint64_t acc = a, out = 1;
    while (b) {
        if (b & 1) out = (uint64_t)(((u128)acc * out) % q);
        acc = (uint64_t)(((u128)acc * acc) % q);
        b >>= 1;
    }
    return out;
}

inline uint64_t shoup_repr(uint64_t v, uint64_t q) { return (uint64_t)(((u128)v << 64) / q); }

// find_primitive_root (src/ntt.rs:283-307) with ChaCha8Rng::from_seed([0;32])
inline uint64_t find_psi(uint64_t q, uint64_t two_n)
{
    uint8_t seed[32] = {0};
    ChaCha8 r(seed);
    uint64_t t = (q - 1) / two_n;
    for (int it = 0; it < 100; it++) {
        uint64_t omega = r.next_u64() % q;
        omega = mod_exp(omega, t, q);
        if (mod_exp(omega, two_n >> 1, q) == 1) continue;
        return omega;
    }
    return 0;
}

struct HostTables {
    uint64_t psi = 0;
    std::vector<uint64_t> psi_bo, psi_bo_s, ipsi_bo, ipsi_bo_s;
    uint64_t n_inv = 0, n_inv_s = 0;
};

// NttBackendU64::_new (src/ntt.rs:323-386).  Powers are accumulated with
// exact modular products (the reference's Barrett mul is exact).
inline HostTables make_tables(uint64_t q, uint32_t n)
{
    HostTables t;
    t.psi = find_psi(q, 2ull * n);
    uint64_t psi_inv = mod_exp(t.psi, q - 2, q);
    uint32_t logn = 0;
    while ((1u << logn) < n) logn++;
    t.psi_bo.resize(n); t.ipsi_bo.resize(n); t.psi_bo_s.resize(n); t.ipsi_bo_s.resize(n);
    uint64_t rp = 1, rpi = 1;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t bo = 0;
        for (uint32_t b = 0; b < logn; b++) bo |= ((i >> b) & 1u) << (logn - 1 - b);
        t.psi_bo[bo] = rp;
        t.ipsi_bo[bo] = rpi;
        rp = (uint64_t)(((u128)rp * t.psi) % q);
        rpi = (uint64_t)(((u128)rpi * psi_inv) % q);
    }
    for (uint32_t i = 0; i < n; i++) {
        t.psi_bo_s[i] = shoup_repr(t.psi_bo[i], q);
        t.ipsi_bo_s[i] = shoup_repr(t.ipsi_bo[i], q);
    }
    t.n_inv = mod_exp(n, q - 2, q);
    t.n_inv_s = shoup_repr(t.n_inv, q);
    return t;
}

// utils::log2 (src/utils.rs:230-237)
inline uint32_t ref_log2(uint64_t v)
{
    if ((v & (v - 1)) == 0) return 64 - (uint32_t)__builtin_clzll(v) - 1;
    return 64 - (uint32_t)__builtin_clzll(v);
}

// generate_auto_map element k reduced into [0, 2*ring) (src/rgsw/keygen.rs:16-24)
inline uint64_t auto_exp(int64_t k, uint64_t ring)
{
    uint64_t two_n = 2 * ring;
    return k < 0 ? two_n - ((uint64_t)(-k) % two_n) : (uint64_t)k % two_n;
}

// Test vectors of BoolEvaluator::new (src/bool/evaluator.rs:680-726), each
// mapped X -> X^{-g} over br_q/2 coefficients. Order: nand, and, or, nor,
// xor, xnor.
inline std::vector<uint64_t> test_vectors(const pz_params &p)
{
    uint64_t q = p.rlwe_q;
    uint32_t qby2 = p.br_q >> 1, qby8 = p.br_q >> 3;
    uint64_t tru = q >> 3, fal = (q - 1) - tru + 1;
    struct TV { uint32_t part; uint64_t before, after; } tv[6] = {
        {3 * qby8, tru, fal}, {3 * qby8, fal, tru}, {qby8, fal, tru},
        {qby8, tru, fal},     {qby8, fal, tru},     {qby8, tru, fal},
    };
    std::vector<uint64_t> out(6 * (size_t)qby2);
    uint64_t k = auto_exp(-(int64_t)p.g, qby2);
    for (int g = 0; g < 6; g++) {
        for (uint32_t i = 0; i < qby2; i++) {
            uint64_t v = i < tv[g].part ? tv[g].before : tv[g].after;
            uint64_t to = ((uint64_t)i * k) % (2ull * qby2);
            bool pos = true;
            if (to >= qby2) { to -= qby2; pos = false; }
            out[(size_t)g * qby2 + to] = pos ? v : q - v;
        }
    }
    return out;
}

// dlog map (src/bool/evaluator.rs:664-673)
inline std::vector<uint32_t> dlog_map(const pz_params &p)
{
    uint32_t q = p.br_q;
    std::vector<uint32_t> d(q, 0);
    for (uint32_t i = 0; i < q / 4; i++) {
        uint32_t v = (uint32_t)mod_exp(p.g, i, q);
        d[v] = i;
        d[q - v] = i + q / 4;
    }
    return d;
}

// RandomFill<[u8; 32]> (src/random.rs:162-175): 32 draws of
// Uniform::new_inclusive(0u8, 255u8).  Over the full u8 range rand 0.8.5's
// UniformInt stores range 256 as 0 and samples Standard instead, which for u8
// is the low byte of next_u32: one stream word per byte.
inline void fill_seed(ChaCha8 &r, uint8_t out[32])
{
    for (int i = 0; i < 32; i++) out[i] = (uint8_t)r.next_u32();
}

// puncture_p_rng (src/utils.rs:219-228): the t-th seed drawn from ChaCha8(seed)
inline void puncture(const uint8_t seed[32], uint32_t times, uint8_t out[32])
{
    ChaCha8 r(seed);
    memset(out, 0, 32);
    for (uint32_t t = 0; t < times; t++) fill_seed(r, out);
}

// Seed tree of NonInteractiveMultiPartyCrs (src/bool/evaluator.rs:97-160):
// LWE index l's RGSW seed and user j's u -> s seed.
inline void ni_seed_index(const uint8_t crs[32], uint32_t l, uint8_t out[32])
{
    uint8_t k[32], m[32];
    puncture(crs, 1, k);
    puncture(k, 1, m);
    puncture(m, l + 1, out);
}

// the seeds of indices [l0, l0 + count) at once: puncture(m, l + 1) is the
// (l + 1)-th seed of the one stream ChaCha8(m)
inline void ni_index_seeds(const uint8_t crs[32], uint32_t l0, uint32_t count, uint8_t *out)
{
    uint8_t k[32], m[32], skip[32];
    puncture(crs, 1, k);
    puncture(k, 1, m);
    ChaCha8 r(m);
    for (uint32_t l = 0; l < l0; l++) fill_seed(r, skip);
    for (uint32_t t = 0; t < count; t++) fill_seed(r, out + 32 * (size_t)t);
}

inline void ni_seed_user(const uint8_t crs[32], uint32_t j, uint8_t out[32])
{
    uint8_t k[32];
    puncture(crs, 2, k);
    puncture(k, j + 1, out);
}

// seeds of the server key's seeded parts (src/bool/evaluator.rs:148-156):
// part 2 the automorphism keys' a rows, part 3 the LWE KSK's a entries
inline void ni_seed_key_part(const uint8_t crs[32], uint32_t part, uint8_t out[32])
{
    uint8_t k[32];
    puncture(crs, 1, k);
    puncture(k, part, out);
}

// multi_party_user_id_lwe_segment (src/bool/evaluator.rs:613-626): indices
// per party; the leader of index l is l / ni_per_party
inline uint32_t ni_per_party(uint32_t lwe_n, uint32_t total) { return (lwe_n + total - 1) / total; }

// Runs f(i) for i in [0, n) on up to 16 host threads.
template <class F>
inline void parallel_for(size_t n, F f)
{
    unsigned hw = std::thread::hardware_concurrency();
    size_t nt = hw == 0 ? 1 : (hw > 16 ? 16 : hw);
    if (nt > n) nt = n;
    if (nt <= 1) {
        for (size_t i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<size_t> next(0);
    std::vector<std::thread> pool;
    for (size_t t = 0; t < nt; t++)
        pool.emplace_back([&]() {
            for (size_t i = next++; i < n; i = next++) f(i);
        });
    for (auto &th : pool) th.join();
}

}  // namespace host
}  // namespace pz

namespace pz {
namespace host {

// Noise tables: T[k] = floor(2^64 * P(|X| <= k)), X ~ N(0, 3.19^2), |X|
// rounded half away from zero (ROUND, the error distribution of
// src/random.rs:112-128 after map_element_from_f64) or truncated toward zero
// (TRUNC, the Gaussian LWE secret of src/random.rs:145-160).
static const uint64_t PZ_CDT_ROUND[42] = {
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
static const uint64_t PZ_CDT_TRUNC[42] = {
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

// One noise draw: u64 against the table, then a sign bit for k > 0.
inline int64_t cdt_draw(ChaCha8 &r, const uint64_t *t)
{
    uint64_t u = r.next_u64();
    int64_t k = 41;
    for (int i = 0; i < 42; i++) {
        if (u < t[i]) {
            k = i;
            break;
        }
    }
    if (k > 0 && (r.next_u32() & 1u)) k = -k;
    return k;
}
inline int64_t gauss_round(ChaCha8 &r) { return cdt_draw(r, PZ_CDT_ROUND); }
inline int64_t gauss_trunc(ChaCha8 &r) { return cdt_draw(r, PZ_CDT_TRUNC); }

// fill_random_ternary_secret_with_hamming_weight (src/utils.rs:75-110); the
// sign bytes are Uniform<u8> draws (one word each).
inline void ternary_hw(ChaCha8 &r, int32_t *out, uint32_t n, uint32_t hw)
{
    std::vector<uint8_t> bytes((hw + 7) / 8 + 1);
    for (uint32_t i = 0; i < (hw + 7) / 8; i++) bytes[i] = (uint8_t)r.next_u32();
    std::vector<uint32_t> idx(n);
    for (uint32_t i = 0; i < n; i++) {
        idx[i] = i;
        out[i] = 0;
    }
    uint32_t len = n, bit = 0, byte = 0;
    for (uint32_t i = 0; i < hw; i++) {
        uint32_t s = (uint32_t)r.uniform(len);
        out[idx[s]] = ((bytes[byte] >> bit) & 1u) ? 1 : -1;
        idx[s] = idx[len - 1];
        len--;
        if (bit == 7) {
            bit = 0;
            byte++;
        } else {
            bit++;
        }
    }
}

inline uint64_t map_i64(uint64_t q, int64_t v) { return v < 0 ? q - (uint64_t)(-v) : (uint64_t)v; }

// Secrets of `parties` parties summed into the ideal secret (ClientKey::
// sk_rlwe / sk_lwe, src/bool/keys.rs:75-154; one seed per party drawn from
// the master stream).
inline void keygen_secrets(const pz_params &p, const uint8_t seed[32], uint32_t parties,
                           int32_t *s_out, int32_t *z_out)
{
    uint32_t N = p.rlwe_n, n = p.lwe_n;
    ChaCha8 master(seed);
    std::vector<int32_t> sp(N), zp(n);
    for (uint32_t i = 0; i < N; i++) s_out[i] = 0;
    for (uint32_t i = 0; i < n; i++) z_out[i] = 0;
    for (uint32_t q = 0; q < parties; q++) {
        uint8_t ps[32];
        master.fill_bytes(ps, 32);
        ChaCha8 r(ps);
        ternary_hw(r, sp.data(), N, N >> 1);
        if (p.lwe_secret_gaussian) {
            for (uint32_t i = 0; i < n; i++) zp[i] = (int32_t)gauss_trunc(r);
        } else {
            ternary_hw(r, zp.data(), n, n >> 1);
        }
        for (uint32_t i = 0; i < N; i++) s_out[i] += sp[i];
        for (uint32_t i = 0; i < n; i++) z_out[i] += zp[i];
    }
}

}  // namespace host
}  // namespace pz
