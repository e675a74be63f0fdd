// Hermite product basis table B[a][i] = bra_a(q_i).
//
// Reference: ScaledMatrixElements::get_bra (ScaledMatrixElements.cpp:870-930)
// with ho_basis[] (HarmonicOscillatorBasis.cpp:16-46).  Row a of the basis is
// a product of at most three Hermite factors; the host expands get_bra's
// loop nest once into a row descriptor (mode, degree) x 3 in get_bra's order
// and multiplication order ((f0 * f1) * f2), so the device values are
// bit-identical to get_bra (factors of degree 0 are exactly 1.0).
//
// Layout: B is [row][point] with points contiguous (the K dimension of the
// SYRK), rows padded to a multiple of 64 with zeros.  A block holds PB
// points (64, or fewer when many modes are selected); its threads first
// evaluate h1, h2, h3 of every selected mode of their point into LDS as
// [degree][mode][point], then 256/PB row groups sweep the block's rows, each
// thread writing one point of a row (coalesced stores).
#include "scp3g_internal.h"

namespace scp3g {

__device__ __forceinline__ double ho_val(int deg, double q)
{
#pragma clang fp contract(off)
    // HarmonicOscillatorBasis.cpp:16-42 in the same operation order
    if (deg == 1) return 1.4142135623730951 * q;
    if (deg == 2) return (2.0 * q * q - 1.0) / 1.4142135623730951;
    return (2.0 * q * q - 3.0) * q / 1.7320508075688772;
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_basis(const double *__restrict__ y, int64_t ldy, const int *__restrict__ modes,
            int nsel, const RowDesc *__restrict__ desc, int64_t n, int64_t nrows_pad,
            int rows_per_block, int64_t npts, T *__restrict__ B, int64_t ldb, int pb_log2)
{
#pragma clang fp contract(off)
    extern __shared__ double hs[];  // [3][nsel][PB], degree 1..3
    const int PB = 1 << pb_log2;    // points per block
    const int ng = 256 >> pb_log2;  // row groups
    const int lane = threadIdx.x & (PB - 1);
    const int grp = threadIdx.x >> pb_log2;
    const int64_t i = (int64_t)blockIdx.x * PB + lane;
    const bool live = i < npts;
    for (int m = grp; m < nsel; m += ng) {
        const double q = live ? y[(size_t)modes[m] * ldy + i] : 0.0;
        hs[(0 * nsel + m) * PB + lane] = ho_val(1, q);
        hs[(1 * nsel + m) * PB + lane] = ho_val(2, q);
        hs[(2 * nsel + m) * PB + lane] = ho_val(3, q);
    }
    __syncthreads();
    const int64_t a0 = (int64_t)blockIdx.y * rows_per_block;
    int64_t a1 = a0 + rows_per_block;
    if (a1 > nrows_pad) a1 = nrows_pad;
    if (i >= ldb) return;
    for (int64_t a = a0 + grp; a < a1; a += ng) {
        double v = 0.0;
        if (a < n && live) {
            const RowDesc d = desc[a];
            double f[3];
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const int mode = d.f[t] & 0xffff;
                const int deg = d.f[t] >> 16;
                f[t] = deg ? hs[((deg - 1) * nsel + mode) * PB + lane] : 1.0;
            }
            v = (f[0] * f[1]) * f[2];
        }
        B[(size_t)a * ldb + i] = (T)v;
    }
}

hipError_t launch_basis(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts, void *B,
                        int64_t ldb, int fp32)
{
    if (npts <= 0) return hipSuccess;
    const int64_t n = ctx->nbasis;
    const int64_t nrows_pad = (n + 63) / 64 * 64;
    const int nsel = (int)ctx->modes.size();
    // points per block: the largest power of two <= 64 whose Hermite table
    // fits the 64 KiB LDS a workgroup may always allocate
    int pb_log2 = 6;
    while (pb_log2 > 3 && (size_t)3 * nsel * (1u << pb_log2) * sizeof(double) > 64 * 1024) --pb_log2;
    const size_t smem = (size_t)3 * nsel * (1u << pb_log2) * sizeof(double);
    if (smem > 64 * 1024) return hipErrorInvalidValue;
    const int64_t pblocks = (ldb + (1 << pb_log2) - 1) >> pb_log2;
    int rows_per_block = 256;
    while (rows_per_block > 16 && pblocks * ((nrows_pad + rows_per_block - 1) / rows_per_block) < 2048)
        rows_per_block /= 2;
    dim3 grid((unsigned)pblocks, (unsigned)((nrows_pad + rows_per_block - 1) / rows_per_block));
    if (fp32)
        hipLaunchKernelGGL(k_basis<float>, grid, dim3(256), smem, ctx->stream, y, ldy,
                           ctx->d_modes, nsel, ctx->d_desc, n, nrows_pad, rows_per_block, npts,
                           (float *)B, ldb, pb_log2);
    else
        hipLaunchKernelGGL(k_basis<double>, grid, dim3(256), smem, ctx->stream, y, ldy,
                           ctx->d_modes, nsel, ctx->d_desc, n, nrows_pad, rows_per_block, npts,
                           (double *)B, ldb, pb_log2);
    return hipGetLastError();
}

}  // namespace scp3g
