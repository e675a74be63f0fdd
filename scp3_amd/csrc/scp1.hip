// SCP1's Gaussian averages (SCP1.cpp:33-86) on one GPU.
//
// For Sobol points i = skip .. skip + npts - 1 in 3N = 9nw dimensions:
//   z     = std_normal(3N, i)                      SCP1.cpp:57
//   dq    = isqrtM_U (z o sqrt(d))                 :60-62
//   r     = bohr (q + dq)           (Angstrom)     :63
//   V, g  = PES(r)                  (kcal/mol, /A) :64
//   <U>   = sum V / autokcalpmol / npts            :64, :81
//   <UX>  = sum g bohr/autokcalpmol / npts         :65-67, :82
//   <UX q^T> = sum (g bohr/autokcalpmol) dq^T / npts  :68, :83
// The reference splits the index range over MPI ranks and sums with three
// Allreduce calls; here one device takes the whole range in tiles of up to
// 2^17 points:
//   k_sobol     normals, coordinate-major [3N x T]
//   k_scp1_map  dq and r, one lane per point (isqrtM_U columns in LDS)
//   PES gradient kernel (one configuration per wave)
//   k_scp1_moments  the (3N+1) x (3N+1) moment matrix
//                   [g; V] [dq; 1]^T over 1024-point chunks, FP64
//   k_scp1_fold     chunk partials summed in a fixed order into the running
//                   FP64 sum (the result does not depend on scheduling)
#include "scp3g_internal.h"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace scp3g {

namespace {

constexpr int kTileMax = 1 << 17;   // points per tile
constexpr int kChunk = 1024;        // points per moment partial
constexpr int kMt = 32;             // moment tile (rows and columns)
constexpr int kSub = 64;            // points per LDS stage

// dq = U (z o sd), r = bohr (q + dq); NC Cartesian outputs per lane
template <int NC>
__global__ void __launch_bounds__(256)
    k_scp1_map(const double *__restrict__ z, int64_t ld, int n3, const double *__restrict__ sd,
               const double *__restrict__ U, const double *__restrict__ q, int64_t npts,
               double *__restrict__ dq, double *__restrict__ r)
{
    extern __shared__ double u[];  // [n3 columns][NC rows of this chunk]
    const int c0 = blockIdx.y * NC;
    for (int e = threadIdx.x; e < n3 * NC; e += blockDim.x) {
        const int k = e / NC, c = e - k * NC;
        u[e] = c0 + c < n3 ? U[(size_t)k * n3 + c0 + c] : 0.0;
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;
#pragma unroll 2
    for (int k = 0; k < n3; ++k) {
        const double yk = z[(size_t)k * ld + i] * sd[k];  // y %= sqrt(d), SCP1.cpp:60
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] = fma(u[k * NC + c], yk, acc[c]);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
        if (c0 + c < n3) {
            dq[(size_t)(c0 + c) * ld + i] = acc[c];
            r[(size_t)(c0 + c) * ld + i] = kBohr * (q[c0 + c] + acc[c]);
        }
}

// part[chunk][j][i] = sum over the chunk's points of a_i b_j, where
// a = [g_0 .. g_{n3-1}, V] and b = [dq_0 .. dq_{n3-1}, 1]; rows beyond n3
// and points beyond npts read as zero.  grid: x = i tiles, y = j tiles,
// z = chunks; 256 lanes, each 2 x 2 outputs of a 32 x 32 tile.
__global__ void __launch_bounds__(256)
    k_scp1_moments(const double *__restrict__ G, const double *__restrict__ E,
                   const double *__restrict__ dq, int64_t ld, int n3, int64_t npts,
                   double *__restrict__ part)
{
    __shared__ double As[kMt][kSub + 1];
    __shared__ double Bs[kMt][kSub + 1];
    const int m = n3 + 1;
    const int i0 = blockIdx.x * kMt, j0 = blockIdx.y * kMt;
    const int64_t p0 = (int64_t)blockIdx.z * kChunk;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int s = 0; s < kChunk; s += kSub) {
        // stage 32 rows x 64 points of a and b; lanes run along points
        for (int e = threadIdx.x; e < kMt * kSub; e += blockDim.x) {
            const int row = e / kSub, pp = e - row * kSub;
            const int64_t p = p0 + s + pp;
            const int ia = i0 + row, jb = j0 + row;
            double a = 0.0, b = 0.0;
            if (p < npts) {
                a = ia < n3 ? G[(size_t)ia * ld + p] : (ia == n3 ? E[p] : 0.0);
                b = jb < n3 ? dq[(size_t)jb * ld + p] : (jb == n3 ? 1.0 : 0.0);
            }
            As[row][pp] = a;
            Bs[row][pp] = b;
        }
        __syncthreads();
#pragma unroll 8
        for (int pp = 0; pp < kSub; ++pp) {
            const double a0 = As[tx][pp], a1 = As[tx + 16][pp];
            const double b0 = Bs[ty][pp], b1 = Bs[ty + 16][pp];
            acc[0][0] = fma(a0, b0, acc[0][0]);
            acc[0][1] = fma(a0, b1, acc[0][1]);
            acc[1][0] = fma(a1, b0, acc[1][0]);
            acc[1][1] = fma(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
    double *__restrict__ out = part + (size_t)blockIdx.z * m * m;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int i = i0 + tx + 16 * a, j = j0 + ty + 16 * b;
            if (i < m && j < m) out[(size_t)j * m + i] = acc[a][b];
        }
}

// acc[e] += sum_c part[c][e], chunks in order
__global__ void __launch_bounds__(256)
    k_scp1_fold(const double *__restrict__ part, int nchunks, int64_t mm, double *__restrict__ acc)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mm) return;
    double s = 0.0;
    for (int c = 0; c < nchunks; ++c) s += part[(size_t)c * mm + e];
    acc[e] += s;
}

struct Scp1Work {
    int dev = -1;
    int n3 = 0;
    int64_t T = 0;
    double *buf = nullptr;     // z/G, dq, r: 3 x n3 x T; E: T; partials; acc
    uint64_t *dirs = nullptr;  // n3 x 62 direction integers
    double *aux = nullptr;     // sqrt(d) (n3), q (n3), isqrtM_U (n3 x n3)
};

std::mutex g_scp1_lock;
std::map<const scp3g_ctx *, Scp1Work> g_scp1;

int partial_chunks(int64_t T) { return (int)((T + kChunk - 1) / kChunk); }

}  // namespace

void scp1_release(scp3g_ctx *ctx)
{
    std::lock_guard<std::mutex> g(g_scp1_lock);
    auto it = g_scp1.find(ctx);
    if (it == g_scp1.end()) return;
    hipFree(it->second.buf);
    hipFree(it->second.dirs);
    hipFree(it->second.aux);
    g_scp1.erase(it);
}

}  // namespace scp3g

using namespace scp3g;

int scp3g_scp1_average(scp3g_ctx *ctx, int pes, int nw, const double *q, const double *d,
                       const double *isqrtM_U, int64_t skip, int64_t npts, double *Upot,
                       double *UX, double *UX_q)
{
    if (!ctx) return -1;
    if (pes != SCP3G_PES_QTIP4PF && pes != SCP3G_PES_TTM3F) return fail(ctx, -4, "unknown potential id");
    if (nw < 1 || nw > 64) return fail(ctx, -4, "nw must be in 1..64");
    if (npts < 1 || skip < 0) return fail(ctx, -4, "bad point range");
    if (skip + npts + 64 * 16 >= ((int64_t)1 << 60)) return fail(ctx, -4, "SEED too large");
    if (!q || !d || !isqrtM_U || !Upot || !UX || !UX_q) return fail(ctx, -4, "null argument");
    const int n3 = 9 * nw;
    for (int k = 0; k < n3; ++k)
        if (!(d[k] >= 0.0) || !std::isfinite(d[k])) return fail(ctx, -4, "d must be finite and >= 0");
    const int nc = n3 <= 18 ? 18 : (n3 <= 54 ? 54 : 30);
    // the map's LDS holds n3 x nc doubles of isqrtM_U
    if ((size_t)n3 * nc * sizeof(double) > 64 * 1024) return fail(ctx, -4, "at most 30 waters");
    hipSetDevice(ctx->device);

    const int64_t T = std::min<int64_t>(kTileMax, (npts + 255) / 256 * 256);
    const int nch = partial_chunks(T);
    const int m = n3 + 1;
    const size_t mm = (size_t)m * m;
    const size_t nbuf = 3 * (size_t)n3 * T + T + (size_t)nch * mm + mm;

    // the lock guards the map; a context's own entry is used by that
    // context's calls only (a context is not shared between threads), and
    // map nodes stay in place while other contexts add theirs
    Scp1Work *wp;
    {
        std::lock_guard<std::mutex> guard(g_scp1_lock);
        wp = &g_scp1[ctx];
    }
    Scp1Work &w = *wp;
    int rc = 0;
    if (w.dev != ctx->device || w.n3 != n3 || w.T < T) {
        hipFree(w.buf);
        hipFree(w.dirs);
        hipFree(w.aux);
        w = Scp1Work();
        std::vector<uint64_t> dirs((size_t)n3 * kSobolCols);
        sobol_directions(n3, dirs.data());
        if ((rc = check_hip(ctx, hipMalloc(&w.buf, nbuf * sizeof(double)), "alloc scp1 work")) ||
            (rc = check_hip(ctx, hipMalloc(&w.dirs, dirs.size() * 8), "alloc scp1 dirs")) ||
            (rc = check_hip(ctx, hipMalloc(&w.aux, (2 * (size_t)n3 + (size_t)n3 * n3) * 8), "alloc scp1 aux"))) {
            hipFree(w.buf);
            hipFree(w.dirs);
            hipFree(w.aux);
            w = Scp1Work();
            return rc;
        }
        if ((rc = check_hip(ctx, hipMemcpy(w.dirs, dirs.data(), dirs.size() * 8, hipMemcpyHostToDevice),
                            "upload scp1 dirs")))
            return rc;
        w.dev = ctx->device;
        w.n3 = n3;
        w.T = T;
    }
    const int64_t ld = w.T;
    double *zG = w.buf;
    double *dq = zG + (size_t)n3 * ld;
    double *r = dq + (size_t)n3 * ld;
    double *E = r + (size_t)n3 * ld;
    double *part = E + ld;
    double *acc = part + (size_t)partial_chunks(w.T) * mm;

    std::vector<double> aux(2 * (size_t)n3 + (size_t)n3 * n3);
    // declared after aux: an error return waits for its upload before aux goes
    SyncOnError sync_guard(ctx, rc);
    for (int k = 0; k < n3; ++k) {
        aux[k] = std::sqrt(d[k]);
        aux[n3 + k] = q[k];
    }
    std::copy(isqrtM_U, isqrtM_U + (size_t)n3 * n3, aux.begin() + 2 * n3);
    double *sd = w.aux, *dq0 = w.aux + n3, *Ud = w.aux + 2 * n3;
    if ((rc = check_hip(ctx, hipMemcpyAsync(w.aux, aux.data(), aux.size() * 8, hipMemcpyHostToDevice, ctx->stream),
                        "upload scp1 inputs")) ||
        (rc = check_hip(ctx, hipMemsetAsync(acc, 0, mm * 8, ctx->stream), "clear scp1 sums")))
        return rc;

    hipEvent_t ev, ek;
    timer_begin(ctx, "scp1", &ev);
    for (int64_t done = 0; done < npts && !rc; done += T) {
        const int64_t np = std::min<int64_t>(T, npts - done);
        timer_begin(ctx, "scp1_sobol", &ek);
        rc = check_hip(ctx, launch_sobol(ctx, w.dirs, n3, skip + done, np, 0, zG, ld, nullptr), "scp1 sobol");
        timer_end(ctx, "scp1_sobol", ek);
        if (rc) break;
        timer_begin(ctx, "scp1_map", &ek);
        dim3 mgrid((unsigned)((np + 255) / 256), (unsigned)((n3 + nc - 1) / nc));
        const size_t smem = sizeof(double) * n3 * nc;
        if (nc == 18)
            hipLaunchKernelGGL(k_scp1_map<18>, mgrid, dim3(256), smem, ctx->stream, zG, ld, n3, sd, Ud, dq0, np, dq, r);
        else if (nc == 54)
            hipLaunchKernelGGL(k_scp1_map<54>, mgrid, dim3(256), smem, ctx->stream, zG, ld, n3, sd, Ud, dq0, np, dq, r);
        else
            hipLaunchKernelGGL(k_scp1_map<30>, mgrid, dim3(256), smem, ctx->stream, zG, ld, n3, sd, Ud, dq0, np, dq, r);
        timer_end(ctx, "scp1_map", ek);
        if ((rc = check_hip(ctx, hipGetLastError(), "scp1 map"))) break;
        // the normals are dead once mapped: their buffer takes the gradients
        timer_begin(ctx, "scp1_pes", &ek);
        rc = pes == SCP3G_PES_TTM3F
                 ? check_hip(ctx, launch_ttm3f_grad(ctx, nw, np, r, ld, E, zG, ld), "scp1 ttm3f")
                 : check_hip(ctx, launch_qtip4pf_grad(ctx, nw, np, r, ld, E, zG, ld), "scp1 qtip4pf");
        timer_end(ctx, "scp1_pes", ek);
        if (rc) break;
        const int ch = partial_chunks(np);
        const unsigned tiles = (unsigned)((m + kMt - 1) / kMt);
        timer_begin(ctx, "scp1_moments", &ek);
        hipLaunchKernelGGL(k_scp1_moments, dim3(tiles, tiles, (unsigned)ch), dim3(256), 0, ctx->stream, zG, E, dq,
                           ld, n3, np, part);
        timer_end(ctx, "scp1_moments", ek);
        if ((rc = check_hip(ctx, hipGetLastError(), "scp1 moments"))) break;
        hipLaunchKernelGGL(k_scp1_fold, dim3((unsigned)((mm + 255) / 256)), dim3(256), 0, ctx->stream, part, ch,
                           (int64_t)mm, acc);
        rc = check_hip(ctx, hipGetLastError(), "scp1 fold");
    }
    timer_end(ctx, "scp1", ev);
    if (rc) return rc;
    std::vector<double> S(mm);
    if ((rc = check_hip(ctx, hipMemcpyAsync(S.data(), acc, mm * 8, hipMemcpyDeviceToHost, ctx->stream),
                        "download scp1 sums")) ||
        (rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "scp1 sync")))
        return rc;
    // S(i, j) = sum a_i b_j, column-major with leading dimension m
    const double c = kBohr / kAutokcalpmol, inv = 1.0 / (double)npts;
    *Upot = S[(size_t)n3 * m + n3] / kAutokcalpmol * inv;
    for (int i = 0; i < n3; ++i) UX[i] = S[(size_t)n3 * m + i] * c * inv;
    for (int j = 0; j < n3; ++j)
        for (int i = 0; i < n3; ++i) UX_q[(size_t)j * n3 + i] = S[(size_t)j * m + i] * c * inv;
    return 0;
}

int scp3g_hessian(scp3g_ctx *ctx, int pes, int nw, const double *q, double s, double *H)
{
    if (!ctx) return -1;
    if (pes != SCP3G_PES_QTIP4PF && pes != SCP3G_PES_TTM3F) return fail(ctx, -4, "unknown potential id");
    if (nw < 1 || nw > 64) return fail(ctx, -4, "nw must be in 1..64");
    if (!q || !H || !(s > 0.0)) return fail(ctx, -4, "bad arguments");
    hipSetDevice(ctx->device);
    const int n3 = 9 * nw;
    const int64_t np = 2 * (int64_t)n3;  // column i: +s at 2i, -s at 2i + 1
    std::vector<double> crd((size_t)n3 * np);
    for (int64_t p = 0; p < np; ++p) {
        const int i = (int)(p >> 1);
        for (int c = 0; c < n3; ++c) {
            double v = q[c] * kBohr;
            if (c == i) v = q[c] * kBohr + ((p & 1) ? -s * kBohr : s * kBohr);  // Hessian.cpp:24-28
            crd[(size_t)c * np + p] = v;
        }
    }
    double *buf = nullptr;
    const size_t nb = (size_t)n3 * np * 2 + np;
    int rc = check_hip(ctx, hipMalloc(&buf, nb * sizeof(double)), "alloc hessian");
    if (rc) return rc;
    double *dcrd = buf, *G = buf + (size_t)n3 * np, *E = G + (size_t)n3 * np;
    std::vector<double> g((size_t)n3 * np);
    rc = check_hip(ctx, hipMemcpyAsync(dcrd, crd.data(), crd.size() * 8, hipMemcpyHostToDevice, ctx->stream),
                   "upload hessian points");
    if (!rc)
        rc = pes == SCP3G_PES_TTM3F
                 ? check_hip(ctx, launch_ttm3f_grad(ctx, nw, np, dcrd, np, E, G, np), "hessian ttm3f")
                 : check_hip(ctx, launch_qtip4pf_grad(ctx, nw, np, dcrd, np, E, G, np), "hessian qtip4pf");
    if (!rc)
        rc = check_hip(ctx, hipMemcpyAsync(g.data(), G, g.size() * 8, hipMemcpyDeviceToHost, ctx->stream),
                       "download hessian gradients");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hessian sync");
    hipFree(buf);
    if (rc) return rc;
    // H(:, i) = (g(+) - g(-)) / (2 s), the gradient taken to Hartree/bohr
    // (Hessian.cpp:30 divides kcal/mol/A by a step in bohr)
    for (int i = 0; i < n3; ++i)
        for (int c = 0; c < n3; ++c)
            H[(size_t)i * n3 + c] = (g[(size_t)c * np + 2 * i] - g[(size_t)c * np + 2 * i + 1]) * kBohr /
                                    (2.0 * s * kAutokcalpmol);
    // the upper triangle symmetrised, the lower one left as computed (:38-41)
    for (int i = 1; i < n3; ++i)
        for (int j = 0; j < i; ++j)
            H[(size_t)i * n3 + j] = 0.5 * (H[(size_t)i * n3 + j] + H[(size_t)j * n3 + i]);
    return 0;
}
