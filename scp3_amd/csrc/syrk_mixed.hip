// H += B diag(w) B^T, mixed form: 3-term split-BF16 MFMA, FP32 tile sums,
// compensated fold into the FP64 running H.
//
// Reference arithmetic being replaced: dsyr2k in FP64
// (ScaledMatrixElements.cpp:942-949,970).  The basis values B and the
// weighted values W = w * B are FP32; each is split exactly as
// x = hi + lo + O(2^-16 x) with hi = bf16(x), lo = bf16(x - hi), and
// sum_p B_a W_b is formed as hi.hi + hi.lo + lo.hi on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation (the dropped lo.lo term
// is O(2^-16) relative).  The tile's FP32 sum is added to the FP64 running
// H with Kahan compensation (Hc holds the running compensation).
//
// Tiling as in the FP64 kernel: 64 x 64 tile per workgroup, 2 x 2 waves of
// 32 x 32, split-K with partials when the triangle has few tiles.  Lane
// (i + 16 g) feeds row i at points p0 + 8 g .. 8 g + 7 of a 32-point chunk.
#include "scp3g_internal.h"

namespace scp3g {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kMTile = 64;

__device__ __forceinline__ unsigned short bf16_rne(float f)
{
    const unsigned u = __float_as_uint(f);
    const unsigned bias = 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)((u + bias) >> 16);
}

__device__ __forceinline__ void split8(const float *x, bf16x8 &hi, bf16x8 &lo)
{
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const unsigned short h = bf16_rne(x[e]);
        const float hf = __uint_as_float((unsigned)h << 16);
        const unsigned short l = bf16_rne(x[e] - hf);
        hi[e] = __builtin_bit_cast(__bf16, h);
        lo[e] = __builtin_bit_cast(__bf16, l);
    }
}

__device__ __forceinline__ void load8(const float *__restrict__ p, float *x)
{
    const float4 u = *reinterpret_cast<const float4 *>(p);
    const float4 v = *reinterpret_cast<const float4 *>(p + 4);
    x[0] = u.x; x[1] = u.y; x[2] = u.z; x[3] = u.w;
    x[4] = v.x; x[5] = v.y; x[6] = v.z; x[7] = v.w;
}

__device__ __forceinline__ void kahan_add(double &h, double &c, double v)
{
    const double yv = v - c;
    const double t = h + yv;
    c = (t - h) - yv;
    h = t;
}

__global__ void __launch_bounds__(256)
    k_syrk_mixed(const float *__restrict__ B, int64_t ldb, const float *__restrict__ wf,
                 const int2 *__restrict__ tiles, int ntiles, int64_t kchunk, int64_t kend,
                 int64_t n, double *__restrict__ H, double *__restrict__ Hc,
                 float *__restrict__ part)
{
    const int2 t = tiles[blockIdx.x];
    const int split = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int64_t r0 = (int64_t)t.x * kMTile + (wave >> 1) * 32;
    const int64_t c0 = (int64_t)t.y * kMTile + (wave & 1) * 32;
    const int64_t kb = (int64_t)split * kchunk;
    int64_t ke = kb + kchunk;
    if (ke > kend) ke = kend;

    const float *__restrict__ A0 = B + (size_t)(r0 + li) * ldb + 8 * lg;
    const float *__restrict__ A1 = A0 + (size_t)16 * ldb;
    const float *__restrict__ B0 = B + (size_t)(c0 + li) * ldb + 8 * lg;
    const float *__restrict__ B1 = B0 + (size_t)16 * ldb;
    const float *__restrict__ W = wf + 8 * lg;

    f4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};

    for (int64_t p = kb; p < ke; p += 32) {
        float xa0[8], xa1[8], xb0[8], xb1[8], ww[8];
        load8(A0 + p, xa0);
        load8(A1 + p, xa1);
        load8(B0 + p, xb0);
        load8(B1 + p, xb1);
        load8(W + p, ww);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            xb0[e] *= ww[e];
            xb1[e] *= ww[e];
        }
        bf16x8 ah[2], al[2], bh[2], bl[2];
        split8(xa0, ah[0], al[0]);
        split8(xa1, ah[1], al[1]);
        split8(xb0, bh[0], bl[0]);
        split8(xb1, bh[1], bl[1]);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
            }
    }

    const int oj = lane & 15;
    const int oi = 4 * (lane >> 4);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int64_t col = c0 + 16 * b + oj;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t row = r0 + 16 * a + oi + e;
                const float v = acc[a][b][e];
                if (part) {
                    const int tr = (int)(row - (int64_t)t.x * kMTile);
                    const int tc = (int)(col - (int64_t)t.y * kMTile);
                    part[((size_t)split * ntiles + blockIdx.x) * (kMTile * kMTile) + tr * kMTile + tc] = v;
                } else if (row <= col && col < n) {
                    const size_t idx = (size_t)col * n + row;
                    double h = H[idx], c = Hc[idx];
                    kahan_add(h, c, (double)v);
                    H[idx] = h;
                    Hc[idx] = c;
                }
            }
        }
}

__global__ void __launch_bounds__(256)
    k_syrk_mixed_fold(const float *__restrict__ part, const int2 *__restrict__ tiles,
                      int ntiles, int nsplit, int64_t n, double *__restrict__ H,
                      double *__restrict__ Hc)
{
    const int tile = blockIdx.x;
    const int2 t = tiles[tile];
    for (int e = threadIdx.x; e < kMTile * kMTile; e += blockDim.x) {
        const int tr = e >> 6, tc = e & 63;
        const int64_t row = (int64_t)t.x * kMTile + tr;
        const int64_t col = (int64_t)t.y * kMTile + tc;
        if (row > col || col >= n) continue;
        double s = 0.0;
        for (int sp = 0; sp < nsplit; ++sp)
            s += (double)part[((size_t)sp * ntiles + tile) * (kMTile * kMTile) + e];
        const size_t idx = (size_t)col * n + row;
        double h = H[idx], c = Hc[idx];
        kahan_add(h, c, s);
        H[idx] = h;
        Hc[idx] = c;
    }
}

__global__ void k_weights_f32(const double *__restrict__ w, int64_t npts, float *__restrict__ wf)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npts) wf[i] = (float)w[i];
}

hipError_t get_tiles_public(scp3g_ctx *ctx, int64_t n, const int2 **tiles, int *ntiles);

hipError_t launch_syrk_mixed(scp3g_ctx *ctx, int64_t n, const float *B, int64_t ldb,
                             const double *w, int64_t npts, double *H, double *Hc)
{
    const int2 *tiles;
    int ntiles;
    hipError_t e = get_tiles_public(ctx, n, &tiles, &ntiles);
    if (e != hipSuccess) return e;
    const int64_t kend = (npts + 31) / 32 * 32;
    int64_t nsplit = 1;
    while ((int64_t)ntiles * nsplit < 1024 && kend / (nsplit * 2) >= 256) nsplit *= 2;
    const int64_t kchunk = ((kend + nsplit - 1) / nsplit + 31) / 32 * 32;
    nsplit = (kend + kchunk - 1) / kchunk;
    // FP32 weights (points beyond npts are zero in w by construction),
    // then the split-K partials, in one scratch allocation
    const size_t off = ((size_t)kend * sizeof(float) + 255) / 256 * 256;
    size_t need = off;
    if (nsplit > 1) need += (size_t)nsplit * ntiles * kMTile * kMTile * sizeof(float);
    if (ctx->part_bytes < need) {
        if (ctx->d_part) hipFree(ctx->d_part);
        ctx->d_part = nullptr;
        ctx->part_bytes = 0;
        e = hipMalloc(&ctx->d_part, need);
        if (e != hipSuccess) return e;
        ctx->part_bytes = need;
    }
    float *wf = reinterpret_cast<float *>(ctx->d_part);
    float *part = nsplit > 1 ? reinterpret_cast<float *>((char *)ctx->d_part + off) : nullptr;
    hipLaunchKernelGGL(k_weights_f32, dim3((unsigned)((kend + 255) / 256)), dim3(256), 0,
                       ctx->stream, w, kend, wf);
    hipLaunchKernelGGL(k_syrk_mixed, dim3((unsigned)ntiles, (unsigned)nsplit), dim3(256), 0,
                       ctx->stream, B, ldb, wf, tiles, ntiles, kchunk, kend, n, H, Hc, part);
    e = hipGetLastError();
    if (e != hipSuccess || !part) return e;
    hipLaunchKernelGGL(k_syrk_mixed_fold, dim3((unsigned)ntiles), dim3(256), 0, ctx->stream, part,
                       tiles, ntiles, (int)nsplit, n, H, Hc);
    return hipGetLastError();
}

}  // namespace scp3g
