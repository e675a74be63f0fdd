// FP64 SYRK H += B diag(w) B^T on 128 x 128 tiles staged through LDS.
//
// Replaces dsyr2k (ScaledMatrixElements.cpp:942-949,970) in the plain FP64
// form.  A workgroup (4 waves, 2 x 2) owns a 128 x 128 tile (ti <= tj) of
// the upper triangle; each wave owns 64 x 64 = 4 x 4 blocks of
// v_mfma_f64_16x16x4f64 (64 FP64 accumulators per lane).  Per 16-point step
// the workgroup copies its 128 rows of B (left operand) and 128 rows of
// w o B (right operand; the weight is applied on the copy) into LDS, double
// buffered (2 x 32 KiB).  Lane (i + 16 g) feeds row i at points 4g .. 4g+3
// of the step over the four MFMA sub-steps; any permutation of the
// contraction index shared by both operands leaves the sum unchanged.
// B is the point-contiguous [row][point] table of basis.hip, so one row's 16
// points are one 128-byte line.
#include "scp3g_internal.h"

namespace scp3g {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int kT64 = 128;

__global__ void __launch_bounds__(256, 2)
    k_syrk_f64_lds(const double *__restrict__ B, int64_t ldb, const double *__restrict__ w,
                   const int2 *__restrict__ tiles, int ntiles, int64_t kchunk, int64_t kend,
                   int64_t n, double *__restrict__ H, double *__restrict__ part)
{
    // [stage][operand A, W][row][16 points]
    __shared__ double lds[2][2][kT64][16];
    const int2 t = tiles[blockIdx.x];
    const int split = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int li = lane & 15, lg = lane >> 4;
    const int64_t kb = (int64_t)split * kchunk;
    int64_t ke = kb + kchunk;
    if (ke > kend) ke = kend;
    const int64_t rA = (int64_t)t.x * kT64, rB = (int64_t)t.y * kT64;

    d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = d4{0, 0, 0, 0};

    // loader: 128 rows x 16 points per operand = 1024 double2; thread tid
    // moves double2 q = tid + 256 j (row q / 8, points 2 (q % 8) .. +1)
    double2 sa[4], sb[4];
    auto gload = [&](int64_t p0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int q = tid + 256 * j, r = q >> 3, c = (q & 7) * 2;
            sa[j] = *reinterpret_cast<const double2 *>(B + (size_t)(rA + r) * ldb + p0 + c);
            double2 v = *reinterpret_cast<const double2 *>(B + (size_t)(rB + r) * ldb + p0 + c);
            const double2 ww = *reinterpret_cast<const double2 *>(w + p0 + c);
            v.x *= ww.x;
            v.y *= ww.y;
            sb[j] = v;
        }
    };
    auto lstore = [&](int s) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int q = tid + 256 * j, r = q >> 3, c = (q & 7) * 2;
            *reinterpret_cast<double2 *>(&lds[s][0][r][c]) = sa[j];
            *reinterpret_cast<double2 *>(&lds[s][1][r][c]) = sb[j];
        }
    };
    if (kb < ke) {
        gload(kb);
        lstore(0);
    }
    __syncthreads();
    int s = 0;
    for (int64_t p = kb; p < ke; p += 16) {
        const bool more = p + 16 < ke;
        if (more) gload(p + 16);
        d4 av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const double *row = &lds[s][0][wr * 64 + 16 * a + li][4 * lg];
            av[a] = d4{row[0], row[1], row[2], row[3]};
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const double *row = &lds[s][1][wc * 64 + 16 * b + li][4 * lg];
            bv[b] = d4{row[0], row[1], row[2], row[3]};
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a][k], bv[b][k], acc[a][b], 0, 0, 0);
        if (more) lstore(s ^ 1);
        __syncthreads();
        s ^= 1;
    }

    // v_mfma_f64_16x16x4f64 result: D[i][j] sits in lane j + 16*(i%4),
    // element i/4 (rows interleave across the four lane groups)
    const int oj = lane & 15, oi = lane >> 4;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int tc = wc * 64 + 16 * b + oj;
            const int64_t col = rB + tc;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int tr = wr * 64 + 16 * a + 4 * e + oi;
                const int64_t row = rA + tr;
                if (part)
                    part[((size_t)split * ntiles + blockIdx.x) * (kT64 * kT64) + (size_t)tr * kT64 + tc] =
                        acc[a][b][e];
                else if (row <= col && col < n)
                    H[(size_t)col * n + row] += acc[a][b][e];
            }
        }
}

__global__ void __launch_bounds__(256)
    k_syrk_f64_lds_reduce(const double *__restrict__ part, const int2 *__restrict__ tiles, int ntiles,
                          int nsplit, int64_t n, double *__restrict__ H)
{
    const int tile = blockIdx.x;
    const int2 t = tiles[tile];
    for (int e = blockIdx.y * 256 + threadIdx.x; e < kT64 * kT64; e += 256 * gridDim.y) {
        const int tr = e / kT64, tc = e % kT64;
        const int64_t row = (int64_t)t.x * kT64 + tr, col = (int64_t)t.y * kT64 + tc;
        if (row > col || col >= n) continue;
        double s = 0.0;
        for (int sp = 0; sp < nsplit; ++sp) s += part[((size_t)sp * ntiles + tile) * (kT64 * kT64) + e];
        H[(size_t)col * n + row] += s;
    }
}

struct Tiles128 {
    int64_t n = -1;
    int2 *d = nullptr;
    int ntiles = 0;
};
static std::map<scp3g_ctx *, Tiles128> g_t128;

void release_tiles128(scp3g_ctx *ctx)
{
    auto it = g_t128.find(ctx);
    if (it != g_t128.end()) {
        if (it->second.d) hipFree(it->second.d);
        g_t128.erase(it);
    }
}

int64_t rows_pad128(int64_t n) { return (n + kT64 - 1) / kT64 * kT64; }

// B: rows padded to a multiple of 128 (zero rows), ldb a multiple of 16
// covering round_up(npts, 16); w zero beyond npts.
hipError_t launch_syrk_f64_lds(scp3g_ctx *ctx, int64_t n, const double *B, int64_t ldb,
                               const double *w, int64_t npts, double *H)
{
    Tiles128 &tc = g_t128[ctx];
    hipError_t e;
    if (tc.n != n) {
        if (tc.d) hipFree(tc.d);
        tc.d = nullptr;
        const int T = (int)((n + kT64 - 1) / kT64);
        std::vector<int2> v;
        for (int j = 0; j < T; ++j)
            for (int i = 0; i <= j; ++i) v.push_back(make_int2(i, j));
        if ((e = hipMalloc(&tc.d, v.size() * sizeof(int2))) != hipSuccess) return e;
        if ((e = hipMemcpy(tc.d, v.data(), v.size() * sizeof(int2), hipMemcpyHostToDevice)) != hipSuccess)
            return e;
        tc.n = n;
        tc.ntiles = (int)v.size();
    }
    const int ntiles = tc.ntiles;
    const int64_t kend = (npts + 15) / 16 * 16;
    int64_t nsplit = 1;
    while ((int64_t)ntiles * nsplit < 1024 && kend / (nsplit * 2) >= 256) nsplit *= 2;
    const int64_t kchunk = ((kend + nsplit - 1) / nsplit + 15) / 16 * 16;
    nsplit = (kend + kchunk - 1) / kchunk;
    double *part = nullptr;
    if (nsplit > 1) {
        const size_t need = (size_t)nsplit * ntiles * kT64 * kT64 * sizeof(double);
        if (ctx->part_bytes < need) {
            if (ctx->d_part) hipFree(ctx->d_part);
            ctx->d_part = nullptr;
            ctx->part_bytes = 0;
            if ((e = hipMalloc(&ctx->d_part, need)) != hipSuccess) return e;
            ctx->part_bytes = need;
        }
        part = ctx->d_part;
    }
    hipEvent_t ev;
    timer_begin(ctx, "syrk_mma", &ev);
    hipLaunchKernelGGL(k_syrk_f64_lds, dim3((unsigned)ntiles, (unsigned)nsplit), dim3(256), 0, ctx->stream, B,
                       ldb, w, tc.d, ntiles, kchunk, kend, n, H, part);
    timer_end(ctx, "syrk_mma", ev);
    e = hipGetLastError();
    if (e != hipSuccess || !part) return e;
    hipLaunchKernelGGL(k_syrk_f64_lds_reduce, dim3((unsigned)ntiles, 8), dim3(256), 0, ctx->stream, part, tc.d,
                       ntiles, (int)nsplit, n, H);
    return hipGetLastError();
}

}  // namespace scp3g
