// H += B diag(w) B^T on the upper triangle, FP64 MFMA form.
//
// Reference: ScaledMatrixElements::addEpot(q, V, M) (ScaledMatrixElements.cpp:
// 951-971): ket = bra * V, dsyr2k('U','N',0.5,bra,ket,1,M), i.e. the upper
// triangle of M gains bra diag(V) bra^T.  Here the weight is folded into the
// right operand when it is loaded, so one weighted SYRK does the work of the
// reference's syr2k at half the flops.
//
// Tiling: a workgroup (4 waves, 2 x 2) owns a 64 x 64 tile (ti <= tj) of H and
// a range of points (split-K when the triangle has too few tiles to fill the
// chip).  Each wave owns 32 x 32 = 2 x 2 blocks of v_mfma_f64_16x16x4f64.
// Lane (i + 16 k) feeds row i of both operands at points p0 + 4k + s for the
// four MFMA steps s of a 16-point chunk; any permutation of the contraction
// index that is the same on both operands leaves the sum unchanged, so each
// lane reads 4 consecutive points (32 bytes) of a row per chunk.
#include "scp3g_internal.h"

#include <mutex>

namespace scp3g {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;

__global__ void __launch_bounds__(256)
    k_syrk_f64(const double *__restrict__ B, int64_t ldb, const double *__restrict__ w,
               const int2 *__restrict__ tiles, int ntiles, int64_t kchunk, int64_t kend,
               int64_t n, double *__restrict__ H, double *__restrict__ part)
{
    const int2 t = tiles[blockIdx.x];
    const int split = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t r0 = (int64_t)t.x * kTile + (wave >> 1) * 32;
    const int64_t c0 = (int64_t)t.y * kTile + (wave & 1) * 32;
    const int64_t kb = (int64_t)split * kchunk;
    int64_t ke = kb + kchunk;
    if (ke > kend) ke = kend;

    const double *__restrict__ A0 = B + (size_t)(r0 + li) * ldb + 4 * lk;
    const double *__restrict__ A1 = A0 + (size_t)16 * ldb;
    const double *__restrict__ B0 = B + (size_t)(c0 + li) * ldb + 4 * lk;
    const double *__restrict__ B1 = B0 + (size_t)16 * ldb;
    const double *__restrict__ W = w + 4 * lk;

    d4 acc00 = {0, 0, 0, 0}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
    for (int64_t p = kb; p < ke; p += 16) {
        const d4 a0 = *reinterpret_cast<const d4 *>(A0 + p);
        const d4 a1 = *reinterpret_cast<const d4 *>(A1 + p);
        const d4 ww = *reinterpret_cast<const d4 *>(W + p);
        const d4 b0 = *reinterpret_cast<const d4 *>(B0 + p) * ww;
        const d4 b1 = *reinterpret_cast<const d4 *>(B1 + p) * ww;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[s], b0[s], acc00, 0, 0, 0);
            acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[s], b1[s], acc01, 0, 0, 0);
            acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[s], b0[s], acc10, 0, 0, 0);
            acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[s], b1[s], acc11, 0, 0, 0);
        }
    }

    // v_mfma_f64_16x16x4f64 result: D[i][j] sits in lane j + 16*(i%4),
    // element i/4 (rows interleave across the four lane groups)
    const int oj = lane & 15;
    const int oi = lane >> 4;
    const d4 *accs[4] = {&acc00, &acc01, &acc10, &acc11};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int bi = b >> 1, bj = b & 1;
        const d4 v = *accs[b];
        const int64_t col = c0 + 16 * bj + oj;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t row = r0 + 16 * bi + 4 * e + oi;
            if (part) {
                const int tr = (int)(row - (int64_t)t.x * kTile);
                const int tc = (int)(col - (int64_t)t.y * kTile);
                part[((size_t)split * ntiles + blockIdx.x) * (kTile * kTile) + tr * kTile + tc] = v[e];
            } else if (row <= col && col < n) {
                H[(size_t)col * n + row] += v[e];
            }
        }
    }
}

// Sum split-K partials in split order and add into H (deterministic).
__global__ void __launch_bounds__(256)
    k_syrk_reduce(const double *__restrict__ part, const int2 *__restrict__ tiles, int ntiles,
                  int nsplit, int64_t n, double *__restrict__ H)
{
    const int tile = blockIdx.x;
    const int2 t = tiles[tile];
    for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
        const int tr = e >> 6, tc = e & 63;
        const int64_t row = (int64_t)t.x * kTile + tr;
        const int64_t col = (int64_t)t.y * kTile + tc;
        if (row > col || col >= n) continue;
        double s = 0.0;
        for (int sp = 0; sp < nsplit; ++sp)
            s += part[((size_t)sp * ntiles + tile) * (kTile * kTile) + e];
        H[(size_t)col * n + row] += s;
    }
}

// Mout = H / P (+ diag) on the upper triangle, zero below (SCP3.cpp:237-239).
__global__ void __launch_bounds__(256)
    k_finalize(const double *__restrict__ H, int64_t n, double P,
               const double *__restrict__ diag_add, const double *__restrict__ offset,
               double *__restrict__ out)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    const int64_t c = idx / n, r = idx - c * n;
    double v = 0.0;
    if (r <= c) {
        v = H[idx] / P;
        if (offset) v += offset[idx];
        if (r == c && diag_add) v += diag_add[r];
    }
    out[idx] = v;
}

static std::vector<int2> upper_tiles(int64_t n)
{
    const int T = (int)((n + kTile - 1) / kTile);
    std::vector<int2> v;
    v.reserve((size_t)T * (T + 1) / 2);
    for (int j = 0; j < T; ++j)
        for (int i = 0; i <= j; ++i) v.push_back(make_int2(i, j));
    return v;
}

struct TileCache {
    int64_t n = -1;
    int2 *d_tiles = nullptr;
    int ntiles = 0;
};
static std::map<scp3g_ctx *, TileCache> g_tiles;
// shared by all contexts (one per host thread): lookups and releases lock
static std::mutex g_tiles_lock;

static hipError_t get_tiles(scp3g_ctx *ctx, int64_t n, const int2 **tiles, int *ntiles)
{
    std::lock_guard<std::mutex> guard_(g_tiles_lock);
    TileCache &tc = g_tiles[ctx];
    if (tc.n != n) {
        if (tc.d_tiles) hipFree(tc.d_tiles);
        tc.d_tiles = nullptr;
        std::vector<int2> v = upper_tiles(n);
        hipError_t e = hipMalloc(&tc.d_tiles, v.size() * sizeof(int2));
        if (e != hipSuccess) return e;
        e = hipMemcpy(tc.d_tiles, v.data(), v.size() * sizeof(int2), hipMemcpyHostToDevice);
        if (e != hipSuccess) return e;
        tc.n = n;
        tc.ntiles = (int)v.size();
    }
    *tiles = tc.d_tiles;
    *ntiles = tc.ntiles;
    return hipSuccess;
}

hipError_t get_tiles_public(scp3g_ctx *ctx, int64_t n, const int2 **tiles, int *ntiles)
{
    return get_tiles(ctx, n, tiles, ntiles);
}

void release_tiles(scp3g_ctx *ctx)
{
    std::lock_guard<std::mutex> guard_(g_tiles_lock);
    auto it = g_tiles.find(ctx);
    if (it != g_tiles.end()) {
        if (it->second.d_tiles) hipFree(it->second.d_tiles);
        g_tiles.erase(it);
    }
}

hipError_t launch_syrk_mixed(scp3g_ctx *ctx, int64_t n, const float *B, int64_t ldb,
                             const double *w, int64_t npts, double *H, double *Hc);

hipError_t launch_syrk(scp3g_ctx *ctx, int mode, int64_t n, const void *B, int64_t ldb,
                       const double *w, int64_t npts, double *H, double *Hc)
{
    if (npts <= 0) return hipSuccess;
    if (mode == SCP3G_MIXED)
        return launch_syrk_mixed(ctx, n, (const float *)B, ldb, w, npts, H, Hc);
    if (ctx->mixed_fast) return launch_syrk_f64_lds(ctx, n, (const double *)B, ldb, w, npts, H);
    const int2 *tiles;
    int ntiles;
    hipError_t e = get_tiles(ctx, n, &tiles, &ntiles);
    if (e != hipSuccess) return e;
    const int64_t kend = (npts + 15) / 16 * 16;
    // split K so that the grid has >= ~1024 workgroups, chunks >= 256 points
    int64_t nsplit = 1;
    while ((int64_t)ntiles * nsplit < 1024 && kend / (nsplit * 2) >= 256) nsplit *= 2;
    const int64_t kchunk = ((kend + nsplit - 1) / nsplit + 15) / 16 * 16;
    nsplit = (kend + kchunk - 1) / kchunk;
    double *part = nullptr;
    if (nsplit > 1) {
        const size_t need = (size_t)nsplit * ntiles * kTile * kTile * sizeof(double);
        if (ctx->part_bytes < need) {
            if (ctx->d_part) hipFree(ctx->d_part);
            ctx->d_part = nullptr;
            ctx->part_bytes = 0;
            e = hipMalloc(&ctx->d_part, need);
            if (e != hipSuccess) return e;
            ctx->part_bytes = need;
        }
        part = ctx->d_part;
    }
    hipLaunchKernelGGL(k_syrk_f64, dim3((unsigned)ntiles, (unsigned)nsplit), dim3(256), 0,
                       ctx->stream, (const double *)B, ldb, w, tiles, ntiles, kchunk, kend, n, H,
                       part);
    e = hipGetLastError();
    if (e != hipSuccess || !part) return e;
    hipLaunchKernelGGL(k_syrk_reduce, dim3((unsigned)ntiles), dim3(256), 0, ctx->stream, part,
                       tiles, ntiles, (int)nsplit, n, H);
    return hipGetLastError();
}

hipError_t launch_finalize(scp3g_ctx *ctx, const double *H, int64_t n, double scale,
                           const double *diag_add, double *out, const double *offset)
{
    const int64_t tot = n * n;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       ctx->stream, H, n, scale, diag_add, offset, out);
    return hipGetLastError();
}

}  // namespace scp3g
