// Large-n mixed-precision SYRK: H += B diag(w) B^T as a signed symmetric
// rank-K update of X = sqrt|w| B on split-BF16 MFMA.
//
// Replaces dsyr2k (ScaledMatrixElements.cpp:942-949,970) for basis sizes
// where the upper triangle has thousands of 256 x 256 tiles (n >= ~4000).
//
// With s_p = sign(w_p) and x_p = sqrt|w_p| b_p,  B diag(w) B^T = sum_p s_p
// x_p x_p^T, so one operand array serves both sides of the product and the
// sign is applied to the right operand as it leaves LDS (an XOR of the BF16
// sign bits; the right operand is the smaller per-wave fragment, and the
// left one then feeds the MFMAs without a VALU step in between).  x is FP32 and is split exactly into hi = bf16(x),
// lo = bf16(x - hi); the product is hi.hi + hi.lo + lo.hi (relative error
// ~2^-16), accumulated in FP32 by v_mfma_f32_16x16x32_bf16.
//
// Layout (written by k_basis_split): Xh, Xl are [K/32][rows_pad][32] BF16,
// i.e. each 32-point chunk of every row is 64 contiguous bytes and the
// chunks of consecutive rows are contiguous, so a workgroup's 256 rows x 32
// points are one 16 KiB contiguous block per plane.  sgn[K/32] holds the
// sign bits of each 32-point chunk.
//
// Tiling: a workgroup (8 waves, 2 x 4) owns a 256 x 256 tile (ti <= tj);
// each wave owns 128 x 64 = 8 x 4 MFMA blocks (128 FP32 accumulators per
// lane), two waves per SIMD.  Per 32-point step the workgroup stages A-hi, A-lo, B-hi, B-lo
// (64 KiB) in LDS, double buffered (128 KiB).  LDS rows are 64 bytes (four
// 16-byte chunks); chunk c of row r is stored at position c ^ ((r >> 1) & 3)
// so that an MFMA operand read (16 lanes on rows r..r+15, one chunk each) is
// free of bank conflicts (measured with SQ_LDS_BANK_CONFLICT: the swizzle
// (r >> 2) & 3 costs 4 conflict cycles per ds_read_b128 for this pattern,
// (r >> 1) & 3 none).  The epilogue folds each tile's FP32 sum
// into the FP64 running H with Kahan compensation (Hc), or writes split-K
// partials.
#include "scp3g_internal.h"

#include <mutex>
#include <tuple>

namespace scp3g {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kFT = 256;  // workgroup tile

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// An MFMA operand fragment read from LDS with the weight signs of its points
// applied (sign-bit XOR), as one vector load: a uint4 read is split into
// four i32 loads of which only the first keeps the LDS variable's alias
// scope, so the re-merged ds_read_b128 then looks as if it could alias the
// operand DMA in flight and the compiler waits for that DMA first.
__device__ __forceinline__ bf16x8 lds_frag_xor(const uint4 &src, u32x4 f)
{
    const u32x4 v = *reinterpret_cast<const u32x4 *>(&src);
    return __builtin_bit_cast(bf16x8, v ^ f);
}

__device__ __forceinline__ unsigned short bf16_rne_u(float f)
{
    const unsigned u = __float_as_uint(f);
    const unsigned bias = 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)((u + bias) >> 16);
}

// Sign words: bit j of sgn[c] is set when the weight of point 32 c + j is
// negative (points beyond npts count as positive, their x is zero).
__global__ void __launch_bounds__(256)
    k_sign_words(const double *__restrict__ w, int64_t npts, int64_t K, unsigned *__restrict__ sgn)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool neg = i < npts && w[i] < 0.0;
    const unsigned long long m = __ballot(neg);
    const int lane = threadIdx.x & 63;
    if (i < K) {
        if (lane == 0) sgn[i >> 5] = (unsigned)m;
        if (lane == 32) sgn[i >> 5] = (unsigned)(m >> 32);
    }
}

// x = sqrt|w| * bra in FP32, split into hi/lo BF16 planes in the k-blocked
// layout (Xl null: the hi plane only, for a one-product SYRK).  A block holds 64 points (two 32-point chunk blocks) and a range of
// rows.  Phase 1 puts the Hermite values h1..h3 of the selected modes and
// sqrt|w| of its points in LDS.  Phase 2: thread t takes chunk block t / 128,
// points 8 (t % 4) .. +7 of it, and row (t % 128) / 4 of every 32-row pass;
// it forms the 8 products of its row and stores them as
// one 16-byte run per plane, so a wave writes 16 consecutive rows x 64 B =
// 1 KiB of contiguous memory per plane.  Product order (f0 f1) f2 times
// sqrt|w|, as in get_bra (ScaledMatrixElements.cpp:870-930).
__global__ void __launch_bounds__(256)
    k_basis_split(const double *__restrict__ y, int64_t ldy, const int *__restrict__ modes, int nsel,
                  const RowDesc *__restrict__ desc, int64_t n, int64_t rows_pad, int rows_per_block,
                  int64_t npts, int64_t K, const double *__restrict__ w,
                  uint4 *__restrict__ Xh, uint4 *__restrict__ Xl)
{
    constexpr int PB = 64;
    extern __shared__ float hsf[];  // [3][nsel][PB], then sqrt|w| [PB]
    float *sws = hsf + (size_t)3 * nsel * PB;
    const int tid = threadIdx.x;
    {
        const int lane = tid & (PB - 1), grp = tid >> 6;
        const int64_t i = (int64_t)blockIdx.x * PB + lane;
        const bool live = i < npts;
        for (int m = grp; m < nsel; m += 4) {
            const float q = live ? (float)y[(size_t)modes[m] * ldy + i] : 0.f;
            hsf[(0 * nsel + m) * PB + lane] = 1.41421356f * q;
            hsf[(1 * nsel + m) * PB + lane] = (2.f * q * q - 1.f) * 0.70710678f;
            hsf[(2 * nsel + m) * PB + lane] = (2.f * q * q - 3.f) * q * 0.57735027f;
        }
        if (grp == 0) sws[lane] = live ? (float)sqrt(fabs(w[i])) : 0.f;
    }
    __syncthreads();
    const int half = tid >> 7, idx = tid & 127, rl = idx >> 2, pg = idx & 3;
    const int p0 = half * 32 + pg * 8;  // first of this thread's 8 points in the block
    const int64_t kb = (int64_t)blockIdx.x * 2 + half;
    if (kb * 32 >= K) return;
    float sw[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) sw[j] = sws[p0 + j];
    const int64_t a0 = (int64_t)blockIdx.y * rows_per_block;
    int64_t a1 = a0 + rows_per_block;
    if (a1 > rows_pad) a1 = rows_pad;
    uint4 *ph = Xh + (size_t)kb * rows_pad * 4 + pg;
    uint4 *pl = Xl ? Xl + (size_t)kb * rows_pad * 4 + pg : nullptr;
    for (int64_t a = a0 + rl; a < a1; a += 32) {
        float v[8];
        if (a < n) {
            const RowDesc d = desc[a];
            float f[3][8];
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const int mode = d.f[t] & 0xffff, deg = d.f[t] >> 16;
                if (deg) {
                    const float4 *src = reinterpret_cast<const float4 *>(hsf + ((deg - 1) * nsel + mode) * PB + p0);
                    const float4 u0 = src[0], u1 = src[1];
                    f[t][0] = u0.x; f[t][1] = u0.y; f[t][2] = u0.z; f[t][3] = u0.w;
                    f[t][4] = u1.x; f[t][5] = u1.y; f[t][6] = u1.z; f[t][7] = u1.w;
                } else {
#pragma unroll
                    for (int j = 0; j < 8; ++j) f[t][j] = 1.f;
                }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = sw[j] * ((f[0][j] * f[1][j]) * f[2][j]);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = 0.f;
        }
        unsigned hh[8], ll[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            hh[j] = bf16_rne_u(v[j]);
            ll[j] = bf16_rne_u(v[j] - __uint_as_float(hh[j] << 16));
        }
        ph[a * 4] = make_uint4(hh[0] | (hh[1] << 16), hh[2] | (hh[3] << 16), hh[4] | (hh[5] << 16),
                               hh[6] | (hh[7] << 16));
        if (Xl)
            pl[a * 4] = make_uint4(ll[0] | (ll[1] << 16), ll[2] | (ll[3] << 16), ll[4] | (ll[5] << 16),
                                   ll[6] | (ll[7] << 16));
    }
}

// The gradient estimator's symmetric rank-2 term Psi phi^T + phi Psi^T as
// one signed SYRK over 2 Kh contraction slots: x = sqrt(wp) (phi + s Psi)
// in k-blocks [0, Kh/32) with sign +, x = sqrt(wp) (phi - s Psi) in
// [Kh/32, Kh/16) with sign -, wp = 1/(2s) (so wp [(phi+sPsi)(.)^T -
// (phi-sPsi)(.)^T] = Psi phi^T + phi Psi^T).  Layout, threads and split as
// k_basis_split; phase 1 puts q and s V_m of every selected mode and s V of
// the block's 64 points in LDS (2 x 4 bytes per mode and point, so 174
// modes fit), and each row's factors h_deg(q) and s-factors (see grad.hip)
// are evaluated from them in FP32.
__global__ void __launch_bounds__(256)
    k_grad_split(const double *__restrict__ y, int64_t ldy, const int *__restrict__ modes, int nsel,
                 const RowDesc *__restrict__ desc, int64_t n, int64_t rows_pad, int rows_per_block,
                 int64_t npts, int64_t Kh, const double *__restrict__ w, const double *__restrict__ Vq,
                 int64_t ldv, const double *__restrict__ sc, const double *__restrict__ wp,
                 uint4 *__restrict__ Xh, uint4 *__restrict__ Xl)
{
    constexpr int PB = 64;
    extern __shared__ float qsg[];  // q [nsel][PB], s V_m [nsel][PB], s V [PB], sqrt(wp) [PB]
    float *tvs = qsg + (size_t)nsel * PB, *tV = tvs + (size_t)nsel * PB, *sws = tV + PB;
    const int tid = threadIdx.x;
    {
        const int lane = tid & (PB - 1), grp = tid >> 6;
        const int64_t i = (int64_t)blockIdx.x * PB + lane;
        const bool live = i < npts;
        const double t = live ? sc[i] : 0.0;
        for (int m = grp; m < nsel; m += 4) {
            qsg[m * PB + lane] = live ? (float)y[(size_t)modes[m] * ldy + i] : 0.f;
            tvs[m * PB + lane] = live ? (float)(t * Vq[(size_t)m * ldv + i]) : 0.f;
        }
        if (grp == 0) {
            tV[lane] = live ? (float)(t * w[i]) : 0.f;
            sws[lane] = live ? (float)sqrt(wp[i]) : 0.f;
        }
    }
    __syncthreads();
    const int half = tid >> 7, idx = tid & 127, rl = idx >> 2, pg = idx & 3;
    const int p0 = half * 32 + pg * 8;
    const int64_t kb = (int64_t)blockIdx.x * 2 + half;
    if (kb * 32 >= Kh) return;
    float sw[8], tv0[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        sw[j] = sws[p0 + j];
        tv0[j] = tV[p0 + j];
    }
    const int64_t a0 = (int64_t)blockIdx.y * rows_per_block;
    int64_t a1 = a0 + rows_per_block;
    if (a1 > rows_pad) a1 = rows_pad;
    const int64_t minus = Kh / 32 * rows_pad * 4;  // uint4 offset of the second half
    uint4 *ph = Xh + (size_t)kb * rows_pad * 4 + pg;
    uint4 *pl = Xl + (size_t)kb * rows_pad * 4 + pg;
    for (int64_t a = a0 + rl; a < a1; a += 32) {
        float vp[8], vm[8];
        if (a < n) {
            const RowDesc d = desc[a];
            float f[3][8], g[3][8];
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const int mode = d.f[t] & 0xffff, deg = d.f[t] >> 16;
                const float *qm = qsg + mode * PB + p0, *vm_ = tvs + mode * PB + p0;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const float q = deg ? qm[j] : 0.f, v = deg ? vm_[j] : 0.f;
                    if (deg == 1) {
                        f[t][j] = 1.41421356f * q;
                        g[t][j] = 0.70710678f * v;
                    } else if (deg == 2) {
                        f[t][j] = (2.f * q * q - 1.f) * 0.70710678f;
                        g[t][j] = 0.70710678f * v * q;
                    } else if (deg == 3) {
                        f[t][j] = (2.f * q * q - 3.f) * q * 0.57735027f;
                        g[t][j] = 1.15470054f * tv0[j] * q + 0.28867513f * v * (2.f * q * q - 3.f);
                    } else {
                        f[t][j] = 1.f;
                        g[t][j] = 0.f;
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float phi = (f[0][j] * f[1][j]) * f[2][j];
                const float psi = g[0][j] * f[1][j] * f[2][j] + f[0][j] * g[1][j] * f[2][j] +
                                  f[0][j] * f[1][j] * g[2][j];
                vp[j] = sw[j] * (phi + psi);
                vm[j] = sw[j] * (phi - psi);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) vp[j] = vm[j] = 0.f;
        }
        unsigned hh[8], ll[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            hh[j] = bf16_rne_u(vp[j]);
            ll[j] = bf16_rne_u(vp[j] - __uint_as_float(hh[j] << 16));
        }
        ph[a * 4] = make_uint4(hh[0] | (hh[1] << 16), hh[2] | (hh[3] << 16), hh[4] | (hh[5] << 16),
                               hh[6] | (hh[7] << 16));
        pl[a * 4] = make_uint4(ll[0] | (ll[1] << 16), ll[2] | (ll[3] << 16), ll[4] | (ll[5] << 16),
                               ll[6] | (ll[7] << 16));
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            hh[j] = bf16_rne_u(vm[j]);
            ll[j] = bf16_rne_u(vm[j] - __uint_as_float(hh[j] << 16));
        }
        ph[minus + a * 4] = make_uint4(hh[0] | (hh[1] << 16), hh[2] | (hh[3] << 16), hh[4] | (hh[5] << 16),
                                       hh[6] | (hh[7] << 16));
        pl[minus + a * 4] = make_uint4(ll[0] | (ll[1] << 16), ll[2] | (ll[3] << 16), ll[4] | (ll[5] << 16),
                                       ll[6] | (ll[7] << 16));
    }
}

// Sign words of the pair: 0 for the first Kh/32 words, all ones after.
__global__ void k_grad_sign_words(int64_t Kh, unsigned *__restrict__ sgn)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Kh / 16) sgn[i] = i < Kh / 32 ? 0u : 0xffffffffu;
}

// The gradient estimator's block correction of mode d (see grad.hip) as a
// signed SYRK over the rows A_d (the basis functions exciting d) and 2 Kh
// slots: x = sqrt|w_d| u^d in k-blocks [0, Kh/32) with the sign of w_d =
// V (4 q_d^2 + 2) - 2 q_d V_d, x = sqrt|V| phi in [Kh/32, Kh/16) with the
// sign of -V.  rows[off + r] is the basis row of A_d's r-th member with the
// factor slot of mode d in bits 30-31.  Layout and threads as k_basis_split.
__global__ void __launch_bounds__(256)
    k_corr_split(const double *__restrict__ y, int64_t ldy, const int *__restrict__ modes, int nsel,
                 const RowDesc *__restrict__ desc, const int *__restrict__ rows, int off, int na,
                 int64_t rows_pad, int rows_per_block, int64_t npts, int64_t Kh, int d,
                 const double *__restrict__ w, const double *__restrict__ Vq, int64_t ldv,
                 uint4 *__restrict__ Xh, uint4 *__restrict__ Xl)
{
    constexpr int PB = 64;
    extern __shared__ float hsc[];  // [3][nsel][PB], e_1..3(q_d) [3][PB], sqrt|w_d|, sqrt|V| [2][PB]
    float *es = hsc + (size_t)3 * nsel * PB, *su = es + 3 * PB, *sv = su + PB;
    const int tid = threadIdx.x;
    {
        const int lane = tid & (PB - 1), grp = tid >> 6;
        const int64_t i = (int64_t)blockIdx.x * PB + lane;
        const bool live = i < npts;
        for (int m = grp; m < nsel; m += 4) {
            const double q = live ? y[(size_t)modes[m] * ldy + i] : 0.0;
            hsc[(0 * nsel + m) * PB + lane] = (float)(1.4142135623730951 * q);
            hsc[(1 * nsel + m) * PB + lane] = (float)((2.0 * q * q - 1.0) * 0.70710678118654752);
            hsc[(2 * nsel + m) * PB + lane] = (float)((2.0 * q * q - 3.0) * q * 0.57735026918962576);
        }
        if (grp == 0) {
            double qd = 0.0, V = 0.0, wd = 0.0;
            if (live) {
                qd = y[(size_t)modes[d] * ldy + i];
                V = w[i];
                wd = V * (4.0 * qd * qd + 2.0) - 2.0 * qd * Vq[(size_t)d * ldv + i];
            }
            es[lane] = 0.70710678118654752f;
            es[PB + lane] = (float)(qd * 0.70710678118654752);
            es[2 * PB + lane] = (float)((2.0 * qd * qd - 3.0) * 0.28867513459481288);
            su[lane] = (float)sqrt(fabs(wd));
            sv[lane] = (float)sqrt(fabs(V));
        }
    }
    __syncthreads();
    const int half = tid >> 7, idx = tid & 127, rl = idx >> 2, pg = idx & 3;
    const int p0 = half * 32 + pg * 8;
    const int64_t kb = (int64_t)blockIdx.x * 2 + half;
    if (kb * 32 >= Kh) return;
    const int64_t a0 = (int64_t)blockIdx.y * rows_per_block;
    int64_t a1 = a0 + rows_per_block;
    if (a1 > rows_pad) a1 = rows_pad;
    const int64_t second = Kh / 32 * rows_pad * 4;
    uint4 *ph = Xh + (size_t)kb * rows_pad * 4 + pg;
    uint4 *pl = Xl + (size_t)kb * rows_pad * 4 + pg;
    for (int64_t a = a0 + rl; a < a1; a += 32) {
        float vu[8], vf[8];
        if (a < na) {
            const int packed = rows[off + a];
            const RowDesc ds = desc[packed & 0x3fffffff];
            const int slot = (unsigned)packed >> 30;
            float u[8], f[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) u[j] = f[j] = 1.f;
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const int mode = ds.f[t] & 0xffff, deg = ds.f[t] >> 16;
                if (!deg) continue;
                const float *hv = hsc + ((deg - 1) * nsel + mode) * PB + p0;
                const float *ev = es + (deg - 1) * PB + p0;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    f[j] *= hv[j];
                    u[j] *= (t == slot) ? ev[j] : hv[j];
                }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                vu[j] = su[p0 + j] * u[j];
                vf[j] = sv[p0 + j] * f[j];
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) vu[j] = vf[j] = 0.f;
        }
        unsigned hh[8], ll[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            hh[j] = bf16_rne_u(vu[j]);
            ll[j] = bf16_rne_u(vu[j] - __uint_as_float(hh[j] << 16));
        }
        ph[a * 4] = make_uint4(hh[0] | (hh[1] << 16), hh[2] | (hh[3] << 16), hh[4] | (hh[5] << 16),
                               hh[6] | (hh[7] << 16));
        pl[a * 4] = make_uint4(ll[0] | (ll[1] << 16), ll[2] | (ll[3] << 16), ll[4] | (ll[5] << 16),
                               ll[6] | (ll[7] << 16));
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            hh[j] = bf16_rne_u(vf[j]);
            ll[j] = bf16_rne_u(vf[j] - __uint_as_float(hh[j] << 16));
        }
        ph[second + a * 4] = make_uint4(hh[0] | (hh[1] << 16), hh[2] | (hh[3] << 16), hh[4] | (hh[5] << 16),
                                        hh[6] | (hh[7] << 16));
        pl[second + a * 4] = make_uint4(ll[0] | (ll[1] << 16), ll[2] | (ll[3] << 16), ll[4] | (ll[5] << 16),
                                        ll[6] | (ll[7] << 16));
    }
}

// Sign words of mode d's correction: w_d < 0 in the first Kh slots, V > 0 in
// the next Kh (points beyond npts count as positive, their x is zero).
__global__ void __launch_bounds__(256)
    k_corr_sign_words(const double *__restrict__ y, int64_t ldy, const int *__restrict__ modes, int d,
                      const double *__restrict__ w, const double *__restrict__ Vq, int64_t ldv,
                      int64_t npts, int64_t Kh, unsigned *__restrict__ sgn)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool first = i < Kh;
    const int64_t p = first ? i : i - Kh;
    bool neg = false;
    if (i < 2 * Kh && p < npts) {
        const double V = w[p];
        if (first) {
            const double qd = y[(size_t)modes[d] * ldy + p];
            neg = V * (4.0 * qd * qd + 2.0) - 2.0 * qd * Vq[(size_t)d * ldv + p] < 0.0;
        } else {
            neg = V > 0.0;
        }
    }
    const unsigned long long m = __ballot(neg);
    const int lane = threadIdx.x & 63;
    if (i < 2 * Kh) {
        if (lane == 0) sgn[i >> 5] = (unsigned)m;
        if (lane == 32) sgn[i >> 5] = (unsigned)(m >> 32);
    }
}

__device__ __forceinline__ void kahan_add2(double &h, double &c, double v)
{
    const double yv = v - c;
    const double t = h + yv;
    c = (t - h) - yv;
    h = t;
}

// Work item of k_syrk_x3: x = row block | column block << 16; y = the
// k-range's index (k-blocks y kbs .. min((y + 1) kbs, nkb)) when w >= 0; w =
// the slot of the FP32 partial that the item writes, or -1: the item covers
// the whole K range and Kahan-adds its tile into H.
//
// TERMS = 3: hi.hi + hi.lo + lo.hi; TERMS = 1: hi.hi only (relative error
// ~2^-8 per product; used for the surrogate term of the control variate and
// for w_nn in its correction term, where the two roundings cancel)
template <int TERMS>
__global__ void __launch_bounds__(512, 1)
    k_syrk_x3(const uint4 *__restrict__ Xh, const uint4 *__restrict__ Xl,
              const unsigned *__restrict__ sgn, int64_t rows_pad, const int4 *__restrict__ work,
              int kbs, int nkb, int64_t n, double *__restrict__ H, double *__restrict__ Hc,
              float *__restrict__ part)
{
    // two stages [plane Ah, Al, Bh, Bl][row * 4 + swizzled chunk], as two
    // arrays: the steps are unrolled by two so that every LDS access names
    // its stage and the compiler does not wait for the other stage's DMA
    __shared__ uint4 lds0[4][kFT * 4];
    __shared__ uint4 lds1[4][kFT * 4];
    const int4 wk = work[blockIdx.x];
    const int2 t = make_int2(wk.x & 0xffff, wk.x >> 16);
    const int out = wk.w;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 2, wc = wave & 3;
    const int li = lane & 15, lg = lane >> 4;
    int kb0 = 0, kb1 = nkb;
    if (out >= 0) {
        kb0 = wk.y * kbs;
        kb1 = kb0 + kbs < nkb ? kb0 + kbs : nkb;
    }
    const size_t rowA = (size_t)t.x * kFT, rowB = (size_t)t.y * kFT;
    const size_t planeStride = (size_t)rows_pad * 4;  // uint4 per 32-point chunk plane

    f4 acc[8][4];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};

    // Operands move global -> LDS by DMA, one stage ahead, in 1 KiB blocks
    // of 16 rows of one plane slot, 64 blocks (8 per wave) per stage.
    // TERMS = 3: a stage is one 32-point k-block, slots A-hi, A-lo, B-hi,
    // B-lo.  TERMS = 1 needs only the hi planes, so a stage holds two
    // k-blocks, slots A-hi[kb], A-hi[kb + 1], B-hi[kb], B-hi[kb + 1] (half the
    // barriers per MFMA).  Lane l of a block writes LDS position 16 l, i.e.
    // row l >> 2, slot l & 3, which holds chunk (l & 3) ^ ((row >> 1) & 3) of
    // that row; so each lane fetches that chunk, and a block is still one
    // contiguous KiB of global memory.
    //
    // The DMA is buffer_load ... lds (MUBUF), not global_load_lds: the
    // compiler models global_load_lds as a FLAT access that may touch LDS,
    // and while one is in flight it turns every LDS wait into a full drain
    // (lgkmcnt(0)), which would serialise the fragment reads below.  A
    // wave's 8 blocks lie in one plane slot (blocks 8 wv .. 8 wv + 7 of the
    // 64, 16 per slot): the buffer resource's base is that plane at the
    // stage's k-block, the lane's byte offset is one VGPR, block i adds
    // 1 KiB through soffset, and the bound is one k-block of the plane.
    constexpr int KS = TERMS == 3 ? 1 : 2;  // k-blocks per stage
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int wslot = wv >> 1;
    // row r = rb 16 + (lane >> 2) of a block has (r >> 1) & 3 = (lane >> 3) & 3
    const int lane_off = (lane >> 2) * 4 + ((lane & 3) ^ ((lane >> 3) & 3));
    const unsigned short *wplane =
        reinterpret_cast<const unsigned short *>(TERMS == 3 ? ((wslot & 1) ? Xl : Xh) : Xh);
    const int voff = (int)(((wslot < 2 ? rowA : rowB) + (size_t)(wv & 1) * 128) * 64) + lane_off * 16;
    const int nrec = (int)(rows_pad * 64);
    int loff[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) loff[i] = wslot * (kFT * 4) + ((wv & 1) * 8 + i) * 64;
    // DMA of the stage that starts at k-block kbv (the second k-block of a
    // TERMS = 1 stage is clamped to the last one of the range; it is then
    // loaded but not used)
#define X3_DMA(arr, kbv)                                                                          \
    do {                                                                                          \
        const int k0_ = (kbv);                                                                    \
        const int k1_ = k0_ + 1 < kb1 ? k0_ + 1 : k0_;                                            \
        const int kk_ = (TERMS == 1 && (wslot & 1)) ? k1_ : k0_;                                  \
        const __amdgpu_buffer_rsrc_t rs_ = __builtin_amdgcn_make_buffer_rsrc(                     \
            (void *)(wplane + (size_t)kk_ * planeStride * 8), (short)0, nrec, 0x00020000);        \
        _Pragma("unroll") for (int i = 0; i < 8; ++i) {                                           \
            __builtin_amdgcn_raw_ptr_buffer_load_lds(                                             \
                rs_, (__attribute__((address_space(3))) void *)(&(arr)[0][0] + loff[i]), 16, voff, \
                i * 1024, 0, 0);                                                                  \
        }                                                                                         \
    } while (0)
    // The weight signs of this lane's 8 points of a k-block (8 lg .. 8 lg + 7)
    // are byte lg of the k-block's sign word; the byte is a per-lane load
    // issued with the DMA one stage ahead, so it has landed by the barrier
    // (a scalar load at the top of a stage would be waited for with
    // lgkmcnt(0), which also drains every LDS read).  The flips go onto the
    // B fragments as a sign-bit XOR.
    const unsigned char *sgb = reinterpret_cast<const unsigned char *>(sgn) + lg;
#define X3_SIGNS(s0_, s1_, kbv)                                                                   \
    do {                                                                                          \
        (s0_) = sgb[4 * (size_t)(kbv)];                                                           \
        if (TERMS == 1) (s1_) = sgb[4 * ((size_t)(kbv) + 1)];                                     \
    } while (0)
#define X3_FLIP(fl_, bits_)                                                                       \
    do {                                                                                          \
        const unsigned b_ = (bits_);                                                              \
        (fl_).x = ((b_ & 1u) << 15) | (((b_ >> 1) & 1u) << 31);                                   \
        (fl_).y = (((b_ >> 2) & 1u) << 15) | (((b_ >> 3) & 1u) << 31);                            \
        (fl_).z = (((b_ >> 4) & 1u) << 15) | (((b_ >> 5) & 1u) << 31);                            \
        (fl_).w = (((b_ >> 6) & 1u) << 15) | (((b_ >> 7) & 1u) << 31);                            \
    } while (0)
    // Within a stage the 24 fragment reads are issued in the order of first
    // use and interleaved with the MFMAs: 12 reads ahead, then one read per
    // G MFMAs, so the first MFMA waits for the first few reads rather than
    // for all of them and every later read is issued well before its use.
#define X3_SCHED(G, nmfma)                                                                        \
    do {                                                                                          \
        __builtin_amdgcn_sched_group_barrier(0x100, 12, 0);                                       \
        _Pragma("unroll") for (int g_ = 0; g_ < 12; ++g_) {                                       \
            __builtin_amdgcn_sched_group_barrier(0x008, (G), 0);                                  \
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                                    \
        }                                                                                         \
        __builtin_amdgcn_sched_group_barrier(0x008, (nmfma) - 12 * (G), 0);                       \
    } while (0)
    // LDS positions of this lane's fragment rows: row r = wr 128 + 16 a + li
    // (A) or wc 64 + 16 b + li (B) holds chunk lg at r 4 + (lg ^ ((r >> 1) & 3)),
    // and (r >> 1) & 3 = (li >> 1) & 3 for every a and b, so fragment a sits at
    // pA + 64 a.  The empty asm makes pA and pB opaque: the optimiser would
    // split each address into a chain of getelementptrs deeper than the LDS
    // lowering pass follows when it tags the reads with their array's alias
    // scope, and an untagged read waits for the operand DMA in flight.
    int pA = (wr * 128 + li) * 4 + (lg ^ ((li >> 1) & 3));
    int pB = (wc * 64 + li) * 4 + (lg ^ ((li >> 1) & 3));
    asm("" : "+v"(pA));
    asm("" : "+v"(pB));
    // TERMS = 3: one k-block, slots A-hi 0, A-lo 1, B-hi 2, B-lo 3.  The 8 B
    // fragments stay in registers and the A fragments stream through: row
    // block a gets hi.hi, hi.lo, lo.hi on its 4 accumulators (4 independent
    // MFMAs between two updates of one accumulator), and the reads of row
    // block a >= 2 are issued 12 and 18 MFMAs ahead of their first use.
#define X3_STEP3(arr, sbits)                                                                      \
    do {                                                                                          \
        u32x4 fl_;                                                                                \
        X3_FLIP(fl_, sbits);                                                                      \
        bf16x8 ah_[8], al_[8], bh_[4], bl_[4];                                                    \
        _Pragma("unroll") for (int b = 0; b < 4; ++b) bh_[b] = lds_frag_xor((arr)[2][pB + 64 * b], fl_); \
        ah_[0] = __builtin_bit_cast(bf16x8, (arr)[0][pA]);                                        \
        _Pragma("unroll") for (int b = 0; b < 4; ++b) bl_[b] = lds_frag_xor((arr)[3][pB + 64 * b], fl_); \
        al_[0] = __builtin_bit_cast(bf16x8, (arr)[1][pA]);                                        \
        _Pragma("unroll") for (int a = 1; a < 8; ++a) {                                           \
            ah_[a] = __builtin_bit_cast(bf16x8, (arr)[0][pA + 64 * a]);                           \
            al_[a] = __builtin_bit_cast(bf16x8, (arr)[1][pA + 64 * a]);                           \
        }                                                                                         \
        _Pragma("unroll") for (int a = 0; a < 8; ++a)                                             \
            _Pragma("unroll") for (int term = 0; term < 3; ++term)                                \
                _Pragma("unroll") for (int b = 0; b < 4; ++b) {                                   \
                    const bf16x8 x_ = term == 2 ? al_[a] : ah_[a];                                \
                    const bf16x8 y_ = term == 1 ? bl_[b] : bh_[b];                                \
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x_, y_, acc[a][b], 0, 0, 0); \
                }                                                                                 \
        X3_SCHED(6, 96);                                                                          \
    } while (0)
    // TERMS = 1: two k-blocks, slots A[kb] 0, A[kb + 1] 1, B[kb] 2,
    // B[kb + 1] 3; k-block kb first for every accumulator
#define X3_STEP1(arr, sbits0, sbits1)                                                             \
    do {                                                                                          \
        u32x4 f0_, f1_;                                                                           \
        X3_FLIP(f0_, sbits0);                                                                     \
        X3_FLIP(f1_, sbits1);                                                                     \
        bf16x8 a0_[8], a1_[8], b0_[4], b1_[4];                                                    \
        _Pragma("unroll") for (int b = 0; b < 4; ++b) b0_[b] = lds_frag_xor((arr)[2][pB + 64 * b], f0_); \
        _Pragma("unroll") for (int a = 0; a < 8; ++a)                                             \
            a0_[a] = __builtin_bit_cast(bf16x8, (arr)[0][pA + 64 * a]);                           \
        _Pragma("unroll") for (int b = 0; b < 4; ++b) b1_[b] = lds_frag_xor((arr)[3][pB + 64 * b], f1_); \
        _Pragma("unroll") for (int a = 0; a < 8; ++a)                                             \
            a1_[a] = __builtin_bit_cast(bf16x8, (arr)[1][pA + 64 * a]);                           \
        _Pragma("unroll") for (int k_ = 0; k_ < 2; ++k_)                                          \
            _Pragma("unroll") for (int a = 0; a < 8; ++a)                                         \
                _Pragma("unroll") for (int b = 0; b < 4; ++b)                                     \
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(                          \
                        k_ ? a1_[a] : a0_[a], k_ ? b1_[b] : b0_[b], acc[a][b], 0, 0, 0);         \
        X3_SCHED(3, 64);                                                                          \
    } while (0)
#define X3_STEP(arr, s0_, s1_)                                                                    \
    do {                                                                                          \
        if constexpr (TERMS == 3) X3_STEP3(arr, s0_);                                             \
        else X3_STEP1(arr, s0_, s1_);                                                             \
    } while (0)
    // Each barrier is preceded by a wait for all of this wave's outstanding
    // vector memory operations, so that once every wave has passed it, every
    // wave's DMA into the stage about to be read has landed; the barrier
    // also orders every wave's reads of the other stage before the next DMA
    // overwrites it.  The wait is explicit: the compiler's own waits cover
    // only this wave's reads of its own DMA blocks, and it may leave the last
    // DMAs of a stage outstanding at the barrier and wait for them after it,
    // while other waves already read the rows those DMAs write.  The
    // sched_barrier after each DMA keeps its issue at the top of the stage:
    // left alone, the scheduler sinks the DMA below the stage's MFMAs, next
    // to the barrier that waits for it, and its latency is exposed.
#define X3_BARRIER()                                                                              \
    do {                                                                                          \
        __builtin_amdgcn_s_waitcnt(0x0f70); /* vmcnt(0); expcnt, lgkmcnt not waited */           \
        __syncthreads();                                                                          \
    } while (0)
    unsigned sa0 = 0, sb0 = 0, sa1 = 0, sb1 = 0;  // sign bytes of the stages in lds0, lds1
    if (kb0 < kb1) {
        X3_DMA(lds0, kb0);
        X3_SIGNS(sa0, sb0, kb0);
    }
    int kb = kb0;
    for (; kb + 2 * KS <= kb1; kb += 2 * KS) {
        X3_BARRIER();
        X3_DMA(lds1, kb + KS);
        X3_SIGNS(sa1, sb1, kb + KS);
        __builtin_amdgcn_sched_barrier(0);
        X3_STEP(lds0, sa0, sb0);
        X3_BARRIER();
        const int nx = kb + 2 * KS < kb1 ? kb + 2 * KS : kb + KS;
        X3_DMA(lds0, nx);
        X3_SIGNS(sa0, sb0, nx);
        __builtin_amdgcn_sched_barrier(0);
        X3_STEP(lds1, sa1, sb1);
    }
    // the launcher makes every k-range a multiple of KS, so at most one
    // stage is left (already in lds0)
    if (kb < kb1) {
        X3_BARRIER();
        X3_STEP(lds0, sa0, sb0);
    }
#undef X3_DMA
#undef X3_SIGNS
#undef X3_FLIP
#undef X3_SCHED
#undef X3_STEP3
#undef X3_STEP1
#undef X3_STEP
#undef X3_BARRIER

    // D[i][j] of a 16x16 block: lane j + 16*(i/4), element i%4
    const int oj = lane & 15, oi = 4 * (lane >> 4);
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int tr0 = wr * 128 + 16 * a + oi;
            const int tc = wc * 64 + 16 * b + oj;
            const int64_t col = (int64_t)rowB + tc;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int tr = tr0 + e;
                const int64_t row = (int64_t)rowA + tr;
                const float v = acc[a][b][e];
                if (out >= 0) {
                    part[(size_t)out * (kFT * kFT) + (size_t)tr * kFT + tc] = v;
                } else if (row <= col && col < n) {
                    const size_t idx = (size_t)col * n + row;
                    double h = H[idx], c = Hc[idx];
                    kahan_add2(h, c, (double)v);
                    H[idx] = h;
                    Hc[idx] = c;
                }
            }
        }
}

// One thread per tile element: the split-K partials are summed in FP64 in a
// fixed order (four interleaved partial sums, then (s0 + s1) + (s2 + s3)) so
// that the loads of several splits are in flight at once, and the sum is
// Kahan-added into H.
__global__ void __launch_bounds__(256)
    k_syrk_x3_fold(const float *__restrict__ part, const int2 *__restrict__ tiles, int ntiles,
                   int nsplit, int64_t n, double *__restrict__ H, double *__restrict__ Hc)
{
    const int tile = blockIdx.x;
    const int2 t = tiles[tile];
    const int e = blockIdx.y * 256 + threadIdx.x;
    const int tr = e / kFT, tc = e % kFT;
    const int64_t row = (int64_t)t.x * kFT + tr, col = (int64_t)t.y * kFT + tc;
    if (row > col || col >= n) return;
    const size_t stride = (size_t)ntiles * (kFT * kFT);
    const float *p = part + (size_t)tile * (kFT * kFT) + e;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int sp = 0;
    for (; sp + 4 <= nsplit; sp += 4) {
        s0 += (double)p[(size_t)sp * stride];
        s1 += (double)p[(size_t)(sp + 1) * stride];
        s2 += (double)p[(size_t)(sp + 2) * stride];
        s3 += (double)p[(size_t)(sp + 3) * stride];
    }
    for (; sp < nsplit; ++sp) s0 += (double)p[(size_t)sp * stride];
    const double s = (s0 + s1) + (s2 + s3);
    const size_t idx = (size_t)col * n + row;
    double h = H[idx], c = Hc[idx];
    kahan_add2(h, c, s);
    H[idx] = h;
    Hc[idx] = c;
}

// --------------------------------------------------------------------------

struct WorkOrder {
    int64_t n = -1;
    int nsplit = 0;
    int4 *d = nullptr;
    int count = 0;
    // unsplit order: the last round's tiles, each split over K into
    // tail_split items whose FP32 partials k_syrk_x3_fold sums into H
    int ntail = 0, tail_split = 0;
    int2 *d_tail = nullptr;
};
// one entry per (context, n, nsplit): the gradient estimator alternates
// between the basis and its per-mode blocks
static std::map<std::tuple<scp3g_ctx *, int64_t, int>, WorkOrder> g_work;

struct FastTiles {
    int64_t n = -1;
    int2 *d = nullptr;
    int ntiles = 0;
};
static std::map<std::pair<scp3g_ctx *, int64_t>, FastTiles> g_ftiles;
// shared by all contexts (one per host thread): lookups and releases lock
static std::mutex g_fast_lock;

static hipError_t fast_tiles(scp3g_ctx *ctx, int64_t n, const int2 **tiles, int *ntiles)
{
    std::lock_guard<std::mutex> guard_(g_fast_lock);
    FastTiles &ft = g_ftiles[std::make_pair(ctx, n)];
    if (ft.n != n) {
        if (ft.d) hipFree(ft.d);
        ft.d = nullptr;
        const int T = (int)((n + kFT - 1) / kFT);
        std::vector<int2> v;
        for (int j = 0; j < T; ++j)
            for (int i = 0; i <= j; ++i) v.push_back(make_int2(i, j));
        hipError_t e = hipMalloc(&ft.d, v.size() * sizeof(int2));
        if (e != hipSuccess) return e;
        e = hipMemcpy(ft.d, v.data(), v.size() * sizeof(int2), hipMemcpyHostToDevice);
        if (e != hipSuccess) return e;
        ft.n = n;
        ft.ntiles = (int)v.size();
    }
    *tiles = ft.d;
    *ntiles = ft.ntiles;
    return hipSuccess;
}

void release_fast_tiles(scp3g_ctx *ctx)
{
    std::lock_guard<std::mutex> guard_(g_fast_lock);
    for (auto iw = g_work.begin(); iw != g_work.end();) {
        if (std::get<0>(iw->first) == ctx) {
            if (iw->second.d) hipFree(iw->second.d);
            if (iw->second.d_tail) hipFree(iw->second.d_tail);
            iw = g_work.erase(iw);
        } else {
            ++iw;
        }
    }
    for (auto it = g_ftiles.begin(); it != g_ftiles.end();) {
        if (it->first.first == ctx) {
            if (it->second.d) hipFree(it->second.d);
            it = g_ftiles.erase(it);
        } else {
            ++it;
        }
    }
}

int64_t fast_rows_pad(int64_t n) { return (n + kFT - 1) / kFT * kFT; }

hipError_t launch_basis_split(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts, int64_t K,
                              const double *w, unsigned short *Xh, unsigned short *Xl, unsigned *sgn)
{
    const int64_t n = ctx->nbasis, rows_pad = fast_rows_pad(n);
    const int nsel = (int)ctx->modes.size();
    // 64 points per block: Hermite values of every selected mode and the
    // point weights (up to 134 KiB for 174 modes; the CU has 160 KiB)
    const size_t smem = ((size_t)3 * nsel * 64 + 64) * sizeof(float);
    if (smem > 160 * 1024 || K % 64) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_basis_split),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sign_words, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, ctx->stream, w, npts, K,
                       sgn);
    const int64_t pblocks = K / 64;
    int64_t rpb = (rows_pad + 31) / 32 * 32;
    while (rpb > 256 && pblocks * ((rows_pad + rpb - 1) / rpb) < 2048) rpb = (rpb / 2 + 31) / 32 * 32;
    dim3 grid((unsigned)pblocks, (unsigned)((rows_pad + rpb - 1) / rpb));
    hipLaunchKernelGGL(k_basis_split, grid, dim3(256), smem, ctx->stream, y, ldy, ctx->d_modes, nsel,
                       ctx->d_desc, n, rows_pad, (int)rpb, npts, K, w, reinterpret_cast<uint4 *>(Xh),
                       reinterpret_cast<uint4 *>(Xl));
    return hipGetLastError();
}

hipError_t launch_grad_split(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts, int64_t Kh,
                             const double *w, const double *Vq, int64_t ldv, const double *sc,
                             const double *wp, unsigned short *Xh, unsigned short *Xl, unsigned *sgn)
{
    const int64_t n = ctx->nbasis, rows_pad = fast_rows_pad(n);
    const int nsel = (int)ctx->modes.size();
    const size_t smem = ((size_t)2 * nsel * 64 + 2 * 64) * sizeof(float);
    if (smem > 160 * 1024 || Kh % 64) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_grad_split),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_grad_sign_words, dim3((unsigned)((Kh / 16 + 255) / 256)), dim3(256), 0, ctx->stream,
                       Kh, sgn);
    const int64_t pblocks = Kh / 64;
    int64_t rpb = (rows_pad + 31) / 32 * 32;
    while (rpb > 256 && pblocks * ((rows_pad + rpb - 1) / rpb) < 2048) rpb = (rpb / 2 + 31) / 32 * 32;
    dim3 grid((unsigned)pblocks, (unsigned)((rows_pad + rpb - 1) / rpb));
    hipLaunchKernelGGL(k_grad_split, grid, dim3(256), smem, ctx->stream, y, ldy, ctx->d_modes, nsel,
                       ctx->d_desc, n, rows_pad, (int)rpb, npts, Kh, w, Vq, ldv, sc, wp,
                       reinterpret_cast<uint4 *>(Xh), reinterpret_cast<uint4 *>(Xl));
    return hipGetLastError();
}

hipError_t launch_corr_split(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts, int64_t Kh, int d,
                             const int *rows, int off, int na, const double *w, const double *Vq, int64_t ldv,
                             unsigned short *Xh, unsigned short *Xl, unsigned *sgn)
{
    const int64_t rows_pad = fast_rows_pad(na);
    const int nsel = (int)ctx->modes.size();
    const size_t smem = ((size_t)3 * nsel * 64 + 5 * 64) * sizeof(float);
    if (smem > 160 * 1024 || Kh % 64) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_corr_split),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_corr_sign_words, dim3((unsigned)((2 * Kh + 255) / 256)), dim3(256), 0, ctx->stream, y,
                       ldy, ctx->d_modes, d, w, Vq, ldv, npts, Kh, sgn);
    const int64_t pblocks = Kh / 64;
    int64_t rpb = (rows_pad + 31) / 32 * 32;
    while (rpb > 256 && pblocks * ((rows_pad + rpb - 1) / rpb) < 2048) rpb = (rpb / 2 + 31) / 32 * 32;
    dim3 grid((unsigned)pblocks, (unsigned)((rows_pad + rpb - 1) / rpb));
    hipLaunchKernelGGL(k_corr_split, grid, dim3(256), smem, ctx->stream, y, ldy, ctx->d_modes, nsel, ctx->d_desc,
                       rows, off, na, rows_pad, (int)rpb, npts, Kh, d, w, Vq, ldv, reinterpret_cast<uint4 *>(Xh),
                       reinterpret_cast<uint4 *>(Xl));
    return hipGetLastError();
}

// Dispatch order of the work items.  Consecutive workgroup ids go to the 8
// XCDs in turn and each XCD has its own L2.
//  * split K (few tiles): the items are listed split-major (the tiles of one
//    split read the same few row blocks of that K range), the list is cut
//    into 8 contiguous chunks, and chunk x is dealt to the workgroup ids
//    congruent to x mod 8;
//  * unsplit (many tiles, each over the whole K range): the triangle is cut
//    into 8 x 8 super-tiles of 256 x 256 tiles, and the tiles of a super-tile
//    are interleaved by sub-block (2 row groups x 4 column groups of tiles,
//    one per XCD), so that the XCDs run the same super-tile together: each
//    XCD's L2 holds its sub-block's 6 row blocks, and the Infinity Cache the
//    16 row blocks that the XCDs share.  (Measured at 20825 states: 3.5%
//    faster than 4 x 8 super-tiles per XCD chunk.)  One workgroup fills a
//    CU, so the tiles run in rounds of ncu; when the last round would be
//    partly idle (ntiles % ncu = r tiles), those r tiles are split over K
//    into s = ncu / r items each, which fit one round of 1/s the length,
//    and their FP32 partials are folded into H afterwards.  (At 20825
//    states, 3403 tiles on 256 CUs: 13 rounds and a 75-tile one that took
//    as long as a full round; now 13 rounds and a third.)

static hipError_t work_order(scp3g_ctx *ctx, int64_t n, int nsplit, int ncu, WorkOrder **out)
{
    WorkOrder &wo = g_work[std::make_tuple(ctx, n, nsplit)];
    if (wo.n != n || wo.nsplit != nsplit) {
        if (wo.d) hipFree(wo.d);
        if (wo.d_tail) hipFree(wo.d_tail);
        wo.d = nullptr;
        wo.d_tail = nullptr;
        wo.n = -1;
        wo.ntail = wo.tail_split = 0;
        const int T = (int)((n + kFT - 1) / kFT);
        if (T > 0xffff) return hipErrorInvalidValue;
        const int ntiles = T * (T + 1) / 2;
        auto item = [](int i, int j, int sp, int slot) { return make_int4(i | (j << 16), sp, 0, slot); };
        std::vector<int4> v;
        std::vector<int2> tail;
        v.reserve((size_t)ntiles * nsplit);
        if (nsplit > 1) {
            std::vector<int4> grouped;
            grouped.reserve((size_t)ntiles * nsplit);
            for (int sp = 0; sp < nsplit; ++sp)
                for (int j = 0; j < T; ++j)
                    for (int i = 0; i <= j; ++i) grouped.push_back(item(i, j, sp, sp * ntiles + j * (j + 1) / 2 + i));
            const int total = (int)grouped.size();
            const int q = total / 8, r = total % 8;
            v.resize((size_t)total);
            std::vector<char> seen((size_t)total, 0);
            for (int x = 0; x < 8; ++x) {
                const int start = x * q + std::min(x, r), len = q + (x < r ? 1 : 0);
                for (int k = 0; k < len; ++k) {
                    const int L = 8 * k + x;
                    if (L >= total || seen[L]) return hipErrorInvalidValue;
                    seen[L] = 1;
                    v[L] = grouped[start + k];
                }
            }
        } else {
            constexpr int S = 8, RI = S / 4, CJ = S / 2;
            for (int J0 = 0; J0 < T; J0 += S)
                for (int I0 = 0; I0 <= J0; I0 += S) {
                    std::vector<int4> sub[8];
                    for (int j = J0; j < std::min(J0 + S, T); ++j)
                        for (int i = I0; i < std::min(I0 + S, j + 1); ++i)
                            sub[((i - I0) / RI) * 2 + (j - J0) / CJ].push_back(item(i, j, 0, -1));
                    size_t mx = 0;
                    for (auto &q : sub) mx = std::max(mx, q.size());
                    for (size_t k = 0; k < mx; ++k)
                        for (int x = 0; x < 8; ++x)
                            if (k < sub[x].size()) v.push_back(sub[x][k]);
                }
            std::vector<char> seen((size_t)ntiles, 0);
            for (const int4 &w : v) {
                const int i = w.x & 0xffff, j = w.x >> 16;
                const int c = j * (j + 1) / 2 + i;
                if (i > j || j >= T || seen[c]) return hipErrorInvalidValue;
                seen[c] = 1;
            }
            if ((int64_t)v.size() != ntiles) return hipErrorInvalidValue;
            const int r = ntiles % ncu, s = r ? ncu / r : 0;
            if (ntiles > ncu && r > 0 && s >= 2) {
                std::vector<int4> last(v.end() - r, v.end());
                v.resize(v.size() - r);
                for (int q = 0; q < r; ++q) {
                    const int i = last[q].x & 0xffff, j = last[q].x >> 16;
                    tail.push_back(make_int2(i, j));
                    for (int sp = 0; sp < s; ++sp) v.push_back(item(i, j, sp, sp * r + q));
                }
                wo.ntail = r;
                wo.tail_split = s;
            }
        }
        hipError_t e = hipMalloc(&wo.d, v.size() * sizeof(int4));
        if (e != hipSuccess) return e;
        e = hipMemcpy(wo.d, v.data(), v.size() * sizeof(int4), hipMemcpyHostToDevice);
        if (e != hipSuccess) return e;
        if (!tail.empty()) {
            if ((e = hipMalloc(&wo.d_tail, tail.size() * sizeof(int2))) != hipSuccess) return e;
            e = hipMemcpy(wo.d_tail, tail.data(), tail.size() * sizeof(int2), hipMemcpyHostToDevice);
            if (e != hipSuccess) return e;
        }
        wo.n = n;
        wo.nsplit = nsplit;
        wo.count = (int)v.size();
    }
    *out = &wo;
    return hipSuccess;
}

static int device_cus(int device)
{
    static std::mutex m;
    static std::map<int, int> cus;
    std::lock_guard<std::mutex> guard_(m);
    auto it = cus.find(device);
    if (it != cus.end()) return it->second;
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c <= 0) c = 256;
    cus[device] = c;
    return c;
}

static hipError_t ensure_part(scp3g_ctx *ctx, size_t need)
{
    if (ctx->part_bytes >= need) return hipSuccess;
    if (ctx->d_part) hipFree(ctx->d_part);
    ctx->d_part = nullptr;
    ctx->part_bytes = 0;
    hipError_t e = hipMalloc(&ctx->d_part, need);
    if (e != hipSuccess) return e;
    ctx->part_bytes = need;
    return hipSuccess;
}

hipError_t launch_syrk_x3(scp3g_ctx *ctx, int64_t n, const unsigned short *Xh,
                          const unsigned short *Xl, const unsigned *sgn, int64_t K, double *H,
                          double *Hc, int terms)
{
    const int2 *tiles;
    int ntiles;
    hipError_t e = fast_tiles(ctx, n, &tiles, &ntiles);
    if (e != hipSuccess) return e;
    const int nkb = (int)(K / 32);
    if (terms == 1 && (nkb & 1)) return hipErrorInvalidValue;  // K is a multiple of 64
    int nsplit = 1;
    while (ntiles * nsplit < 512 && nkb / (nsplit * 2) >= 8) nsplit *= 2;
    int kbs = (nkb + nsplit - 1) / nsplit;
    if (terms == 1) kbs += kbs & 1;  // k-ranges of whole two-block stages
    nsplit = (nkb + kbs - 1) / kbs;
    WorkOrder *wo;
    {
        std::lock_guard<std::mutex> guard_(g_fast_lock);
        if ((e = work_order(ctx, n, nsplit, device_cus(ctx->device), &wo)) != hipSuccess) return e;
    }
    // the split-K partials, or those of the last round's split tiles
    const int fold_tiles = nsplit > 1 ? ntiles : wo->ntail;
    const int fold_split = nsplit > 1 ? nsplit : wo->tail_split;
    const int2 *fold_list = nsplit > 1 ? tiles : wo->d_tail;
    if (nsplit == 1 && fold_tiles) {
        kbs = (nkb + fold_split - 1) / fold_split;
        kbs += kbs & 1;
    }
    float *part = nullptr;
    if (fold_tiles) {
        if ((e = ensure_part(ctx, (size_t)fold_split * fold_tiles * kFT * kFT * sizeof(float))) != hipSuccess)
            return e;
        part = reinterpret_cast<float *>(ctx->d_part);
    }
    const int64_t rows_pad = fast_rows_pad(n);
    const uint4 *xh = reinterpret_cast<const uint4 *>(Xh), *xl = reinterpret_cast<const uint4 *>(Xl);
    hipEvent_t ev;
    timer_begin(ctx, "syrk_mma", &ev);
    if (terms == 1)
        hipLaunchKernelGGL(k_syrk_x3<1>, dim3((unsigned)wo->count), dim3(512), 0, ctx->stream, xh, xl, sgn,
                           rows_pad, wo->d, kbs, nkb, n, H, Hc, part);
    else
        hipLaunchKernelGGL(k_syrk_x3<3>, dim3((unsigned)wo->count), dim3(512), 0, ctx->stream, xh, xl, sgn,
                           rows_pad, wo->d, kbs, nkb, n, H, Hc, part);
    timer_end(ctx, "syrk_mma", ev);
    e = hipGetLastError();
    if (e != hipSuccess || !fold_tiles) return e;
    hipLaunchKernelGGL(k_syrk_x3_fold, dim3((unsigned)fold_tiles, kFT * kFT / 256), dim3(256), 0, ctx->stream,
                       part, fold_list, fold_tiles, fold_split, n, H, Hc);
    return hipGetLastError();
}

}  // namespace scp3g
