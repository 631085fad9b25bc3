// pz_internal.hpp -- declarations shared by the kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pz.h"
#include "pz_device.hpp"

typedef uint8_t u8;
typedef uint16_t u16;

#ifndef PZ_KS_CT
#define PZ_KS_CT 32
#endif

namespace pz {

enum { PREP_GATE = 0, PREP_PBS = 1, PREP_KSONLY = 2 };

// LWE key-switch tiling: threads per group (one output column each),
// ciphertexts per group, KSK rows staged per LDS round.
constexpr int KS_CT = PZ_KS_CT;  // ciphertexts per work group
constexpr int KS_ROWS = 256;  // KSK rows per LDS round
constexpr int KS_SPLIT = 2;   // minimum row splits (summed with atomics)
constexpr int KS_GROUPS = 2048;  // work groups a key-switch launch aims for

void launch_ntt(int logn, int mode, u64 *rows, size_t nrows, const Tables &tb, hipStream_t s);
void launch_fma(u64 *out, const u64 *a, const u64 *as, const u64 *b, u32 d, size_t n, u64 q,
                hipStream_t s);
void launch_decompose(const u64 *in, u64 *digits, size_t count, const Decomp &dc,
                      hipStream_t s);
void launch_ext(int logn, bool is_auto, u64 *rlwe, const u32 *idx, const u8 *triv, size_t B,
                const DevCtx &c, const DevKey &k, hipStream_t s, const u32 *order = nullptr,
                int grid = 0);
// resident work groups per CU of the split step (d = 1 batched operators)
int step_split_per_cu(int logn);
// order[] = the batch's positions sorted by key index (idx >= nkeys last),
// one work group; false (nothing launched) for more than 8191 keys
bool launch_key_order(const u32 *idx, size_t B, u32 nkeys, u32 *order, hipStream_t s);
// rowmajor: digits as [ciphertext][row] (the MFMA key switch's layout)
// instead of tile-major [B / KS_CT][row][KS_CT]
void launch_prep(int mode, const u64 *c0, const u64 *c1, const u32 *idx0, const u32 *idx1,
                 const u32 *ops, u32 N1, size_t B, int8_t *digits, u32 *ks_b, u32 *ks_out,
                 const DevCtx &c, hipStream_t s, bool rowmajor = false);
// With planes (the key's ksk_planes) the MFMA key switch, which takes
// row-major digits (ks_rowmajor); otherwise the VALU kernel on tile-major ones.
void launch_ks(const int8_t *digits, const u32 *ks_b, const void *ksk, u32 ksk_bytes, u32 rows,
               u32 n1, u32 ldk, size_t seg_begin, size_t seg_end, u32 *out, hipStream_t s,
               const int8_t *planes = nullptr);
#ifndef PZ_KS_MFMA
#define PZ_KS_MFMA 1
#endif
// the digit layout a single-key batch with this key uses
inline bool ks_rowmajor(const DevKey &k) { return PZ_KS_MFMA && k.ksk_planes != nullptr; }
size_t br_lds_bytes(int logn, bool split, u32 br_q, u32 n_lwe);
// blind-rotation kernel variants: one work group per bootstrap, or two half
// groups (d = 1) with the register budget of 2 (latency) or 4 waves per SIMD
// (BR_SPLIT4*: 4 coefficients per thread)
enum { BR_ONE = 0, BR_SPLIT_LAT = 1, BR_SPLIT = 2, BR_SPLIT4_LAT = 3, BR_SPLIT4 = 4, BR_NVAR = 5 };
const void *blind_rotate_kernel(int logn, int var);
bool br_split(const DevCtx &c);
int br_threads(int logn, int var);
void launch_ks_multi(const int8_t *digits, const u32 *ks_b, const DevKey *keys, const u32 *work,
                     size_t n_work, u32 ksk_bytes, u32 rows, u32 n1, u32 ldk, u32 *out,
                     hipStream_t s);
void launch_blind_rotate(int logn, int var, int grid, const u32 *ks, const u32 *gate, size_t B,
                         u64 *out, u32 *stats, const DevCtx &c, const DevKey &k,
                         const DevKey *keys, const u32 *key_idx, const u32 *out_idx,
                         hipStream_t s);
constexpr int PROBE_CHAINS = 8;  // independent chains per thread in k_valu_probe
void launch_valu_probe(int kind, int grid, u64 *out, u32 iters, u32 c, u64 w, u64 ws, u64 q,
                       hipStream_t s);
DevCtx rgrg_devctx(const pz_params &p, const DevCtx &dc);
// S ChaCha8 streams (keys [S][8] words) of `draws` UniformInt<u64> draws
// mod q each; out[s][k - keep_from] for k >= keep_from; reject[s] set when a
// draw of stream s falls outside the acceptance zone (the positions after it
// are then not the sequential stream's)
void launch_expand_uniform(const u32 *keys, u32 S, u64 draws, u64 keep_from, u64 q, u64 *out,
                           u32 *reject, hipStream_t s);
void launch_rgsw_rows(int logn, u64 *rgsw0, u32 r0, u32 da0, u32 db0, size_t B, const DevCtx &c,
                      const u64 *K1, const u64 *K1s, u32 r1, hipStream_t s);
void launch_scale(u64 *x, size_t n, u64 c, u64 c_s, u64 q, hipStream_t s);
void launch_shoup(const u64 *in, u64 *out, size_t n, u64 q, hipStream_t s);
void launch_not(u64 *c0, size_t n, u64 q, hipStream_t s);
void launch_widen(const u32 *in, u64 *out, size_t n, u32 mask, hipStream_t s);
void launch_pack_ksk(const u64 *in, void *out, size_t rows, u32 n1, u32 ldk, u32 bytes,
                     hipStream_t s);
// KSK rows [rows][n1] u64 -> KS_PLANES 7-bit planes [plane][ldk][rows] int8
void launch_pack_ksk_planes(const u64 *in, int8_t *out, size_t rows, u32 n1, u32 ldk,
                            hipStream_t s);

}  // namespace pz

// Sets the message pz_last_error() returns ("<fn>: <what>"; for PZ_ERR_HIP
// the pending HIP error's text is appended) and returns `code`.
int pz_internal_fail(int code, const char *fn, const char *what);
#define PZ_FAIL(code, what) pz_internal_fail((code), __func__, (what))
// PZ_OK, or PZ_ERR_HIP with the error's text when the last launch failed
int pz_internal_launch_status(const char *fn);
#define PZ_LAUNCH_STATUS() pz_internal_launch_status(__func__)

// Host-blocking copy ordered on stream st.  The contexts' streams are
// non-blocking: a plain hipMemcpy (null stream) neither waits for the work
// queued on them nor is waited for by it, so every copy next to such work --
// reading what a kernel wrote, or refilling a buffer a kernel may still read
// -- is queued on the same stream and waited for.
inline hipError_t pz_copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind,
                               hipStream_t st)
{
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
    return e != hipSuccess ? e : hipStreamSynchronize(st);
}

// Uniform rows of S seeded ChaCha8 streams (UniformInt<u64> mod q, rand
// 0.8.5) into d_out[S][draws - keep_from], expanded on the device; d_kf is
// device workspace of 12*S words.  A stream with a rejected draw is
// regenerated sequentially on the host, so the rows always equal the
// sequential streams.  Synchronises the stream.
bool pz_internal_expand_uniform(hipStream_t st, const uint8_t *seeds, u32 S, u64 draws,
                                u64 keep_from, u64 q, u64 *d_out, u32 *d_kf);

// Key-switching key for RLWE ciphertexts (Ksk(u -> s)): [2d][N] rows in the
// evaluation domain, A_0..A_{d-1} then B_0..B_{d-1}, with Shoup twins.
struct pz_rlwe_ksk {
    pz_ctx *owner = nullptr;
    u64 *rows = nullptr, *rows_s = nullptr;
    u32 d = 0;
};
// takes ownership of device rows d_rows (evaluation domain)
int pz_internal_rlwe_ksk_adopt(pz_ctx *c, u64 *d_rows, u32 d, pz_rlwe_ksk **out);
// same device and parameter set: keys made under one serve the other
bool pz_internal_ctx_compatible(const pz_ctx *a, const pz_ctx *b);
bool pz_internal_key_fits(const pz_ctx *c, const pz_key *k);
size_t pz_internal_key_footprint(const pz_ctx *c);
int pz_internal_ctx_device(const pz_ctx *c);
// releases the batch executor's staging buffers of a context being destroyed
void pz_internal_exec_forget(const pz_ctx *c);

