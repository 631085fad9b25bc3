// pz_kernels.hip -- element-wise and row kernels of the bootstrapping path
// (gfx950). The RLWE operators and the blind rotation are in pz_rlwe.hip.
//
//   k_ntt_rows        batched Ntt::{forward,forward_lazy,backward,backward_lazy}
//   k_fma             ShoupMatrixFMA::shoup_matrix_fma
//   k_decompose       decompose_r (all digits)
//   k_prep            gate linear step + f64 moddown + LWE-KS digits
//   k_ks              LWE key switch (digits x KSK, mod lwe_q)
//   k_shoup           Shoup twins of key rows (ToShoup)
//   k_not / k_*       small element-wise helpers (NOT gate, keygen)
#include "pz_internal.hpp"

namespace pz {

// ---------------------------------------------------------------------------
// NTT rows: one work group (N/8 threads) per row.
// ---------------------------------------------------------------------------
template <int LOGN, int MODE>
__global__ __launch_bounds__(1 << (LOGN - 3)) void k_ntt_rows(u64 *rows, size_t nrows, Tables tb)
{
    constexpr int N = 1 << LOGN;
    extern __shared__ u64 lds[];
    const u32 j = threadIdx.x;
    for (size_t row = blockIdx.x; row < nrows; row += gridDim.x) {
        u64 *p = rows + row * N;
        u64 v[1][8];
        if (MODE <= 1) {
#pragma unroll
            for (int r = 0; r < 8; r++) v[0][r] = p[coef_x<LOGN>(j, r)];
            if (MODE == 0) ntt_fwd<LOGN, 1, FWD_LAZY>(v, j, lds, tb);
            else ntt_fwd<LOGN, 1, FWD_STRICT>(v, j, lds, tb);
#pragma unroll
            for (int r = 0; r < 8; r++) p[eval_x<LOGN>(j, r)] = v[0][r];
        } else {
#pragma unroll
            for (int r = 0; r < 8; r++) v[0][r] = p[eval_x<LOGN>(j, r)];
            if (MODE == 2) ntt_inv<LOGN, 1, false>(v, j, lds, tb);
            else ntt_inv<LOGN, 1, true>(v, j, lds, tb);
#pragma unroll
            for (int r = 0; r < 8; r++) p[coef_x<LOGN>(j, r)] = v[0][r];
        }
    }
}

// ---------------------------------------------------------------------------
// shoup_matrix_fma over d rows of length n (element-wise, grid-stride).
// ---------------------------------------------------------------------------
__global__ void k_fma(u64 *out, const u64 *a, const u64 *as, const u64 *b, u32 d, size_t n,
                      u64 q)
{
    const u64 q2 = q << 1;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        u64 o = out[i];
        for (u32 r = 0; r < d; r++) {
            size_t k = (size_t)r * n + i;
            o = shoup_fma(o, a[k], as[k], b[k], q, q2);
        }
        out[i] = o;
    }
}

// decompose_r: digits[r][i] = digit r of in[i]
__global__ void k_decompose(const u64 *in, u64 *digits, size_t count, Decomp dc)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x) {
        u64 value = decomp_prepare(dc, in[i]);
        for (u32 r = 0; r < dc.d; r++) digits[(size_t)r * count + i] = decomp_next(dc, value);
    }
}

// ---------------------------------------------------------------------------
// k_prep: per ciphertext coefficient -- gate linear step (src/bool/
// evaluator.rs:2118-2133), f64 moddown Q -> q_ks (src/pbs.rs:128-132) and
// the LWE-KS decomposition (src/lwe.rs:31-35) into signed int8 digits.
// mode: PREP_GATE (c0, c1, ops), PREP_PBS (in), PREP_KSONLY (in is mod q_ks).
// ---------------------------------------------------------------------------
__global__ void k_prep(int mode, const u64 *c0, const u64 *c1, const u32 *ops, u32 N1,
                       size_t B, int8_t *digits, u32 *ks_b, DevCtx c)
{
    const size_t bb = blockIdx.x;
    const u32 col = blockIdx.y * blockDim.x + threadIdx.x;
    if (bb >= B || col >= N1) return;
    const u64 q = c.tb.q;
    u64 v = c0[bb * N1 + col];
    if (mode == PREP_GATE) {
        u64 v1 = c1[bb * N1 + col];
        u32 op = ops[bb];
        if (op <= 3) {
            v = add_mod(v, v1, q);
            if (col == 0) v = add_mod(v, c.qby4, q);
        } else {
            v = sub_mod(v, v1, q);
            v = add_mod(v, v, q);
        }
    }
    u64 md = v;
    if (mode != PREP_KSONLY) {
        double x = ((double)v * c.lwe_qf) / c.rlwe_qf;
        md = (u64)round(x);
    }
    if (col == 0) {
        ks_b[bb] = (u32)md;
        return;
    }
    const u32 d = c.lwe_d;
    const u64 lq = c.lwe_mask + 1;
    int8_t *out = digits + bb * (size_t)(N1 - 1) * d + (size_t)(col - 1) * d;
    u64 value = decomp_prepare(c.dec_lwe, md);
    for (u32 r = 0; r < d; r++) {
        u64 dg = decomp_next(c.dec_lwe, value);
        long long sv = dg >= (lq >> 1) ? (long long)dg - (long long)lq : (long long)dg;
        out[r] = (int8_t)sv;
    }
}

// ---------------------------------------------------------------------------
// k_ks: out[b][col] = (sum_rows digit[b][row] * ksk[row][col] + (col==0)*b_in) mod q_ks
// grid: (ceil((n+1)/KS_TPB), ceil(B/KS_CT)); block KS_TPB; each thread one
// output column for KS_CT ciphertexts; digits staged through LDS.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(KS_TPB) void k_ks(const int8_t *digits, const u32 *ks_b,
                                               const T *ksk, u32 rows, u32 n1, size_t B,
                                               u32 *out, u32 mask)
{
    __shared__ int8_t dig[KS_ROWS][KS_CT];
    const u32 col = blockIdx.x * KS_TPB + threadIdx.x;
    const size_t b0 = (size_t)blockIdx.y * KS_CT;
    u32 acc[KS_CT];
#pragma unroll
    for (int t = 0; t < KS_CT; t++) acc[t] = 0;
    for (u32 r0 = 0; r0 < rows; r0 += KS_ROWS) {
        __syncthreads();
        for (u32 e = threadIdx.x; e < KS_ROWS * KS_CT; e += KS_TPB) {
            u32 rr = e % KS_ROWS, t = e / KS_ROWS;
            size_t bb = b0 + t;
            int8_t dv = 0;
            if (bb < B && r0 + rr < rows) dv = digits[bb * rows + r0 + rr];
            dig[rr][t] = dv;
        }
        __syncthreads();
        if (col < n1) {
            const u32 rmax = (rows - r0) < KS_ROWS ? (rows - r0) : KS_ROWS;
            for (u32 rr = 0; rr < rmax; rr++) {
                int kv = (int)ksk[(size_t)(r0 + rr) * n1 + col];
#pragma unroll
                for (int t = 0; t < KS_CT; t++) acc[t] += (u32)((int)dig[rr][t] * kv);
            }
        }
    }
    if (col < n1) {
#pragma unroll
        for (int t = 0; t < KS_CT; t++) {
            size_t bb = b0 + t;
            if (bb < B) {
                u32 v = acc[t] + (col == 0 ? ks_b[bb] : 0u);
                out[bb * n1 + col] = v & mask;
            }
        }
    }
}

// Shoup twins: out[i] = floor(in[i] * 2^64 / q) (ToShoup, src/utils.rs:49-73)
// by restoring long division of (v << 64) by q, v < q.
__global__ void k_shoup(const u64 *in, u64 *out, size_t n, u64 q)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        u64 rem = in[i], quo = 0;
        for (int bit = 0; bit < 64; bit++) {
            // rem < q < 2^63: shifting left cannot overflow
            rem <<= 1;
            quo <<= 1;
            if (rem >= q) {
                rem -= q;
                quo |= 1;
            }
        }
        out[i] = quo;
    }
}

// NOT gate: every element negated (src/bool/evaluator.rs:2236-2239)
__global__ void k_not(u64 *c0, size_t n, u64 q)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        c0[i] = neg_mod(c0[i], q);
}

// u32 LWE words -> u64
__global__ void k_widen(const u32 *in, u64 *out, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

// KSK rows (u64, reference layout) -> packed device words
template <typename T>
__global__ void k_pack_ksk(const u64 *in, T *out, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = (T)in[i];
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static int grid_for(size_t n, int tpb, int cap = 65535 * 8)
{
    size_t g = (n + tpb - 1) / tpb;
    if (g < 1) g = 1;
    if (g > (size_t)cap) g = cap;
    return (int)g;
}

template <int LOGN>
static void launch_ntt_t(int mode, u64 *rows, size_t nrows, const Tables &tb, hipStream_t s)
{
    constexpr int TPB = 1 << (LOGN - 3);
    size_t lds = 8 * (size_t)Geom<LOGN>::LDS_WORDS;
    int grid = nrows < 65535 ? (int)nrows : 65535;
    if (grid < 1) return;
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_ntt_rows<LOGN, 0>), grid, TPB, lds, s, rows, nrows, tb); break;
    case 1: hipLaunchKernelGGL((k_ntt_rows<LOGN, 1>), grid, TPB, lds, s, rows, nrows, tb); break;
    case 2: hipLaunchKernelGGL((k_ntt_rows<LOGN, 2>), grid, TPB, lds, s, rows, nrows, tb); break;
    default: hipLaunchKernelGGL((k_ntt_rows<LOGN, 3>), grid, TPB, lds, s, rows, nrows, tb); break;
    }
}

void launch_ntt(int logn, int mode, u64 *rows, size_t nrows, const Tables &tb, hipStream_t s)
{
    switch (logn) {
    case 9: launch_ntt_t<9>(mode, rows, nrows, tb, s); break;
    case 11: launch_ntt_t<11>(mode, rows, nrows, tb, s); break;
    }
}

void launch_fma(u64 *out, const u64 *a, const u64 *as, const u64 *b, u32 d, size_t n, u64 q,
                hipStream_t s)
{
    hipLaunchKernelGGL(k_fma, grid_for(n, 256), 256, 0, s, out, a, as, b, d, n, q);
}

void launch_decompose(const u64 *in, u64 *digits, size_t count, const Decomp &dc, hipStream_t s)
{
    hipLaunchKernelGGL(k_decompose, grid_for(count, 256), 256, 0, s, in, digits, count, dc);
}

void launch_prep(int mode, const u64 *c0, const u64 *c1, const u32 *ops, u32 N1, size_t B,
                 int8_t *digits, u32 *ks_b, const DevCtx &c, hipStream_t s)
{
    if (B == 0) return;
    dim3 grid((unsigned)B, (N1 + 255) / 256);
    hipLaunchKernelGGL(k_prep, grid, 256, 0, s, mode, c0, c1, ops, N1, B, digits, ks_b, c);
}

void launch_ks(const int8_t *digits, const u32 *ks_b, const void *ksk, u32 ksk_bytes, u32 rows,
               u32 n1, size_t B, u32 *out, u32 mask, hipStream_t s)
{
    if (B == 0) return;
    dim3 grid((n1 + KS_TPB - 1) / KS_TPB, (unsigned)((B + KS_CT - 1) / KS_CT));
    if (ksk_bytes == 2)
        hipLaunchKernelGGL(k_ks<uint16_t>, grid, KS_TPB, 0, s, digits, ks_b,
                           (const uint16_t *)ksk, rows, n1, B, out, mask);
    else
        hipLaunchKernelGGL(k_ks<uint32_t>, grid, KS_TPB, 0, s, digits, ks_b,
                           (const uint32_t *)ksk, rows, n1, B, out, mask);
}

void launch_shoup(const u64 *in, u64 *out, size_t n, u64 q, hipStream_t s)
{
    hipLaunchKernelGGL(k_shoup, grid_for(n, 256), 256, 0, s, in, out, n, q);
}

void launch_not(u64 *c0, size_t n, u64 q, hipStream_t s)
{
    hipLaunchKernelGGL(k_not, grid_for(n, 256), 256, 0, s, c0, n, q);
}

void launch_widen(const u32 *in, u64 *out, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_widen, grid_for(n, 256), 256, 0, s, in, out, n);
}

void launch_pack_ksk(const u64 *in, void *out, size_t n, u32 bytes, hipStream_t s)
{
    if (bytes == 2)
        hipLaunchKernelGGL(k_pack_ksk<uint16_t>, grid_for(n, 256), 256, 0, s, in,
                           (uint16_t *)out, n);
    else
        hipLaunchKernelGGL(k_pack_ksk<uint32_t>, grid_for(n, 256), 256, 0, s, in,
                           (uint32_t *)out, n);
}

}  // namespace pz
