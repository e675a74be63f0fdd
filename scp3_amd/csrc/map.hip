// Normal-mode -> Cartesian map and the anharmonic weight.
//
// r = bohr * (x0 + MUa * y)            SCP3.cpp:208 (per point GEMV)
// bV = V/autokcalpmol - 0.5 y.(w o y)  SCP3.cpp:219-223
//
// The map is a (9nw x M) * (M x tile) product with a tiny left operand:
// each lane owns one point and keeps a chunk of NC Cartesian accumulators
// in registers; the chunk's columns of MUa are staged in LDS and read as
// broadcasts, y with coalesced vector loads, each y element once per chunk.
#include "scp3g_internal.h"

namespace scp3g {

template <int NC, typename T>
__global__ void __launch_bounds__(256)
    k_map(const double *__restrict__ y, int64_t ldy, int nmodes0,
          const double *__restrict__ MUa, const double *__restrict__ x0, int n3,
          int64_t npts, T *__restrict__ crd, int64_t ldc)
{
    // this chunk's MUa columns, staged in LDS (read as broadcasts)
    extern __shared__ double mu[];  // [nmodes0][NC]
    const int c0 = blockIdx.y * NC;
    for (int e = threadIdx.x; e < nmodes0 * NC; e += blockDim.x) {
        const int k = e / NC, c = e - k * NC;
        mu[e] = c0 + c < n3 ? MUa[(size_t)k * n3 + c0 + c] : 0.0;
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;
#pragma unroll 4
    for (int k = 0; k < nmodes0; ++k) {
        const double yk = y[(size_t)k * ldy + i];
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] = fma(mu[k * NC + c], yk, acc[c]);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
        if (c0 + c < n3) crd[(size_t)(c0 + c) * ldc + i] = (T)(kBohr * (x0[c0 + c] + acc[c]));
}

hipError_t launch_map(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts, void *crd,
                      int64_t ldc, int fp32)
{
    if (npts <= 0) return hipSuccess;
    const int n3 = 9 * ctx->nw;
    const int nc = n3 <= 18 ? 18 : (n3 <= 54 ? 54 : 30);
    if ((size_t)ctx->nmodes0 * nc * sizeof(double) > 64 * 1024) return hipErrorInvalidValue;  // > 31 waters
    // chunk of Cartesians per lane: all of them up to the hexamer (y read
    // once), 30 at a time beyond (the 20-water cluster: 6 chunks), so that
    // the chunk's MUa columns fit in 64 KiB of LDS
#define MAP_LAUNCH(NC_)                                                                            \
    do {                                                                                           \
        dim3 grid((unsigned)((npts + 255) / 256), (unsigned)((n3 + (NC_) - 1) / (NC_)));           \
        const size_t smem = sizeof(double) * ctx->nmodes0 * (NC_);                                 \
        if (fp32)                                                                                  \
            hipLaunchKernelGGL((k_map<NC_, float>), grid, dim3(256), smem, ctx->stream, y, ldy,    \
                               ctx->nmodes0, ctx->d_MUa, ctx->d_x0, n3, npts, (float *)crd, ldc);  \
        else                                                                                       \
            hipLaunchKernelGGL((k_map<NC_, double>), grid, dim3(256), smem, ctx->stream, y, ldy,   \
                               ctx->nmodes0, ctx->d_MUa, ctx->d_x0, n3, npts, (double *)crd, ldc); \
    } while (0)
    if (nc == 18)
        MAP_LAUNCH(18);
    else if (nc == 54)
        MAP_LAUNCH(54);
    else
        MAP_LAUNCH(30);
#undef MAP_LAUNCH
    return hipGetLastError();
}

// bV = E/autokcalpmol - 0.5 * sum_k y_k (omega_k y_k); SCP3.cpp:219-223;
// times `scale` (1 leaves bV bit for bit).
__global__ void __launch_bounds__(256)
    k_weight(const double *__restrict__ E, const double *__restrict__ y, int64_t ldy,
             const double *__restrict__ omega, int nmodes0, int64_t npts, int64_t npad, double scale,
             double *__restrict__ w)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    if (i >= npts) {
        w[i] = 0.0;  // padded points carry no weight
        return;
    }
    double h = 0.0;
    for (int k = 0; k < nmodes0; ++k) {
        const double yk = y[(size_t)k * ldy + i];
        h += yk * (omega[k] * yk);
    }
    const double bV = E[i] / kAutokcalpmol - 0.5 * h;
    w[i] = scale * bV;
}

hipError_t launch_weight(scp3g_ctx *ctx, const double *E, const double *y, int64_t ldy,
                         int64_t npts, double *w, double scale)
{
    const int64_t npad = (npts + 255) / 256 * 256;
    if (npad <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_weight, dim3((unsigned)(npad / 256)), dim3(256), 0, ctx->stream, E, y,
                       ldy, ctx->d_omega, ctx->nmodes0, npts, npad, scale, w);
    return hipGetLastError();
}

}  // namespace scp3g
