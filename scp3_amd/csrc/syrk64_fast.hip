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

#include <mutex>

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
    // moves double2 q = tid + 256 j (row q / 8, points 2 (q % 8) .. +1),
    // j = 0..3, into named registers (an array here would be promoted to LDS
    // scratch).  Step p+16 is fetched before step p's MFMAs and stored after
    // them; the load is unconditional (the last step re-reads its own
    // points into the idle buffer) and sched_barriers keep the order.
    const int lr = tid >> 3, lc = (tid & 7) * 2;  // row and point of j = 0
    double2 sa0, sa1, sa2, sa3, sb0, sb1, sb2, sb3;
    double2 ww;  // the step's weights for this thread's two points
#define F64_GLOAD1(j, sa, sb, p0)                                                                     \
    do {                                                                                              \
        sa = *reinterpret_cast<const double2 *>(B + (size_t)(rA + lr + 32 * (j)) * ldb + (p0) + lc); \
        sb = *reinterpret_cast<const double2 *>(B + (size_t)(rB + lr + 32 * (j)) * ldb + (p0) + lc); \
    } while (0)
#define F64_GLOAD(p0)                                                  \
    do {                                                               \
        F64_GLOAD1(0, sa0, sb0, p0);                                   \
        F64_GLOAD1(1, sa1, sb1, p0);                                   \
        F64_GLOAD1(2, sa2, sb2, p0);                                   \
        F64_GLOAD1(3, sa3, sb3, p0);                                   \
        ww = *reinterpret_cast<const double2 *>(w + (p0) + lc);        \
    } while (0)
    // the weight is applied to the right operand on its way into LDS
#define F64_STB(sg, rr, sb) \
    *reinterpret_cast<double2 *>(&lds[sg][1][rr][lc]) = double2{(sb).x * ww.x, (sb).y * ww.y}
#define F64_LSTORE(sg)                                                                  \
    do {                                                                                \
        *reinterpret_cast<double2 *>(&lds[sg][0][lr][lc]) = sa0;                        \
        *reinterpret_cast<double2 *>(&lds[sg][0][lr + 32][lc]) = sa1;                   \
        *reinterpret_cast<double2 *>(&lds[sg][0][lr + 64][lc]) = sa2;                   \
        *reinterpret_cast<double2 *>(&lds[sg][0][lr + 96][lc]) = sa3;                   \
        F64_STB(sg, lr, sb0);                                                           \
        F64_STB(sg, lr + 32, sb1);                                                      \
        F64_STB(sg, lr + 64, sb2);                                                      \
        F64_STB(sg, lr + 96, sb3);                                                      \
    } while (0)
    if (kb < ke) {
        F64_GLOAD(kb);
        F64_LSTORE(0);
    }
    __syncthreads();
    int s = 0;
    for (int64_t p = kb; p < ke; p += 16) {
        const int64_t pn = p + 16 < ke ? p + 16 : p;
        F64_GLOAD(pn);
        __builtin_amdgcn_sched_barrier(0);
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
        __builtin_amdgcn_sched_barrier(0);
        F64_LSTORE(s ^ 1);
        __syncthreads();
        s ^= 1;
    }
#undef F64_GLOAD1
#undef F64_GLOAD
#undef F64_STB
#undef F64_LSTORE

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
// shared by all contexts (one per host thread): lookups and releases lock
static std::mutex g_t128_lock;

void release_tiles128(scp3g_ctx *ctx)
{
    std::lock_guard<std::mutex> guard_(g_t128_lock);
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
    std::lock_guard<std::mutex> guard_(g_t128_lock);
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
