// Lowest levels of Mout on the GPU: block Davidson in FP64.
//
// Replaces the eigensolver's dsyevr (RANGE='I', IL=1, IU=K;
// eigensolver.cpp:20-80) and SCP3's eig_sym calls for the lowest level
// (SCP3.cpp:242,246).  Mout = H/P + HO diagonal is strongly diagonally
// dominant (the HO energies set the diagonal, the sampled potential couples
// states weakly), so a Davidson iteration preconditioned by the diagonal
// (Davidson 1975; block form of Liu 1978) converges in a few tens of
// iterations, each of which reads Mout once from HBM:
//
//   V orthonormal (n x m), W = A V, G = V^T W, G s = theta s (host, m small),
//   X = V S, R = W S - X Theta, T = R_i / (theta_i - diag), T orthogonalised
//   against V (classical Gram-Schmidt twice) and within itself (CholQR2),
//   V <- [V T]; restart V <- X when m would exceed the subspace limit.
//
// A is the full symmetric matrix (both triangles) stored column-major, so
// A X is computed column by column of A (A(:, r) . X = (A X)(r, :)).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "scp3g_internal.h"

namespace scp3g {

constexpr int kMaxB = 32;     // Davidson block size limit
constexpr int kLdA = 32;      // A's leading dimension is a multiple of this
constexpr int kMaxSplit = 16; // split-K limit of A X

int64_t lda_of(int64_t n) { return (n + kLdA - 1) / kLdA * kLdA; }

// A(r, c) = Mout(min, max) / P (+ diag on r == c) from the upper-triangle sum
// H, through 32 x 32 LDS tiles so that both the read of the mirrored
// triangle and the write stay coalesced.  A has leading dimension lda; its
// padding rows n..lda-1 are left as they are (zero from allocation).
__global__ void __launch_bounds__(256)
    k_sym_full(const double *__restrict__ H, int64_t ldh, int64_t n, double P,
               const double *__restrict__ diag, double *__restrict__ A, int64_t lda)
{
    __shared__ double t[32][33];
    const int64_t bc = (int64_t)blockIdx.x * 32, br = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    if (br <= bc) {
        // tile lies in the upper triangle (rows br.., cols bc..): direct copy
        for (int k = ty; k < 32; k += 8) {
            const int64_t r = br + tx, c = bc + k;
            if (r < n && c < n) {
                double v = (r <= c) ? H[(size_t)c * ldh + r] / P : 0.0;
                if (r > c) v = H[(size_t)r * ldh + c] / P;
                if (r == c) v += diag[r];
                A[(size_t)c * lda + r] = v;
            }
        }
    } else {
        // lower tile: A(r, c) = H(c, r) for r > c; read the mirror tile
        // (rows bc.., cols br..) coalesced, transpose through LDS
        for (int k = ty; k < 32; k += 8) {
            const int64_t r = bc + tx, c = br + k;  // element H(r, c), r < c
            t[k][tx] = (r < n && c < n) ? H[(size_t)c * ldh + r] / P : 0.0;
        }
        __syncthreads();
        for (int k = ty; k < 32; k += 8) {
            const int64_t r = br + tx, c = bc + k;
            if (r < n && c < n) A[(size_t)c * lda + r] = t[tx][k];
        }
    }
}

// Y(n x b) = A X for symmetric A on v_mfma_f64_16x16x4f64, computed as
// Y^T = X^T A so that a 16-lane row of the result is 16 consecutive rows of
// Y (128-byte stores).  A workgroup (4 waves) owns 128 rows r of Y and one
// range of the contraction index k (split-K, partial sums to `part`); wave w
// owns rows 32w..32w+31 (two 16-row blocks) and all NB x 16 columns.  Per
// 32-k chunk the workgroup stages X (NB*16 columns x 32 k) in LDS and each
// lane loads A(r, k) for its row directly: row r of A is column r (A is
// symmetric), contiguous in k, and lda is a multiple of 32, so lane
// (i + 16g) reads the double2 at k + 8u + 2g (u = 0..3), 64 contiguous bytes
// per row and instruction.  MFMA step (u, v) contracts over
// k + 8u + 2g + v, g = 0..3; both operands use that assignment.
// The result element e of lane (i + 16g) is Y^T(4e + g, i) of the block.
template <int NB>
__global__ void __launch_bounds__(256)
    k_symm_mfma(const double *__restrict__ A, int64_t lda, int64_t n, const double *__restrict__ X,
                int64_t ldx, int b, int64_t kchunk, double *__restrict__ Y, int64_t ldy,
                double *__restrict__ part)
{
    typedef double d4 __attribute__((ext_vector_type(4)));
    __shared__ double xs[NB * 16][34];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int64_t r0 = (int64_t)blockIdx.x * 128 + wave * 32;
    const int split = blockIdx.y;
    const int64_t kpad = (n + 31) / 32 * 32;
    const int64_t kb = (int64_t)split * kchunk;
    const int64_t ke = kb + kchunk < kpad ? kb + kchunk : kpad;
    d4 acc[NB][2];
#pragma unroll
    for (int c = 0; c < NB; ++c)
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) acc[c][rb] = d4{0, 0, 0, 0};
    const double *arow[2];
    bool rok[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        const int64_t r = r0 + 16 * rb + li;
        rok[rb] = r < n;
        arow[rb] = A + (size_t)(rok[rb] ? r : 0) * lda;
    }
    for (int64_t k = kb; k < ke; k += 32) {
        // X chunk -> LDS: element q = tid + 256 j is column q / 32, k + q % 32
#pragma unroll
        for (int j = 0; j < NB * 2; ++j) {
            const int q = tid + 256 * j, col = q >> 5, kk = q & 31;
            xs[col][kk] = (col < b && k + kk < n) ? X[(size_t)col * ldx + k + kk] : 0.0;
        }
        double2 av[2][4];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                av[rb][u] = rok[rb] ? *reinterpret_cast<const double2 *>(arow[rb] + k + 8 * u + 2 * g)
                                    : double2{0.0, 0.0};
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            double2 xv[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) xv[c] = *reinterpret_cast<const double2 *>(&xs[16 * c + li][8 * u + 2 * g]);
#pragma unroll
            for (int c = 0; c < NB; ++c)
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) {
                    acc[c][rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[c].x, av[rb][u].x, acc[c][rb], 0, 0, 0);
                    acc[c][rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[c].y, av[rb][u].y, acc[c][rb], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    double *dst = part ? part + (size_t)split * kMaxB * n : Y;
    const int64_t ld = part ? n : ldy;
#pragma unroll
    for (int c = 0; c < NB; ++c)
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
            const int64_t r = r0 + 16 * rb + li;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int j = 16 * c + 4 * e + g;
                if (j < b && r < n) dst[(size_t)j * ld + r] = acc[c][rb][e];
            }
        }
}

// Y(r, j) = sum over splits of the partial sums, in split order.
__global__ void __launch_bounds__(256)
    k_symm_reduce(const double *__restrict__ part, int nsplit, int64_t n, double *__restrict__ Y,
                  int64_t ldy)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (r >= n) return;
    double s = 0.0;
    for (int sp = 0; sp < nsplit; ++sp) s += part[((size_t)sp * kMaxB + j) * n + r];
    Y[(size_t)j * ldy + r] = s;
}

// C(p x q) = P^T Q for tall P (n x p, ld ldp) and Q (n x q, ld ldq).  One
// wave owns a 4 x 4 block of C: lanes stride over n, then a butterfly of
// shuffles reduces the 16 sums (fixed order, no LDS, no barrier; the
// Davidson sizes make this a latency problem, not a bandwidth one).
__global__ void __launch_bounds__(64)
    k_gemm_tn(const double *__restrict__ Pm, int64_t ldp, int p, const double *__restrict__ Q,
              int64_t ldq, int q, int64_t n, double *__restrict__ C, int ldc)
{
    const int i0 = blockIdx.x * 4, j0 = blockIdx.y * 4;
    const int lane = threadIdx.x;
    const double *pc[4], *qc[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        pc[a] = Pm + (size_t)(i0 + a < p ? i0 + a : i0) * ldp;
        qc[a] = Q + (size_t)(j0 + a < q ? j0 + a : j0) * ldq;
    }
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;
#pragma unroll 2
    for (int64_t k = lane; k < n; k += 64) {
        double pv[4], qv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            pv[a] = pc[a][k];
            qv[a] = qc[a][k];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[a][c] = fma(pv[a], qv[c], acc[a][c]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double v = acc[a][c];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            acc[a][c] = v;
        }
    if (lane < 16) {
        const int a = lane >> 2, c = lane & 3;
        double v = 0.0;
#pragma unroll
        for (int aa = 0; aa < 4; ++aa)
#pragma unroll
            for (int cc = 0; cc < 4; ++cc)
                if (aa == a && cc == c) v = acc[aa][cc];
        if (i0 + a < p && j0 + c < q) C[(size_t)(j0 + c) * ldc + i0 + a] = v;
    }
}

// Y(n x q) = beta*Y + alpha * P(n x p) S(p x q), q <= kMaxB; S column-major
// with ld lds.  One wave owns 64 rows and 8 columns: the 8 columns of S are
// staged in LDS once, each lane streams its row of P (one load per k).
__global__ void __launch_bounds__(64)
    k_gemm_nn(const double *__restrict__ Pm, int64_t ldp, int p, const double *__restrict__ S,
              int lds, int q, double alpha, double beta, double *__restrict__ Y, int64_t ldy,
              int64_t n)
{
    constexpr int kMaxP = 256;
    __shared__ double ss[kMaxP][8];
    const int64_t r = (int64_t)blockIdx.x * 64 + threadIdx.x;
    const int j0 = blockIdx.y * 8;
    for (int e = threadIdx.x; e < p * 8; e += 64) {
        const int k = e >> 3, j = e & 7;
        ss[k][j] = (j0 + j < q) ? S[(size_t)(j0 + j) * lds + k] : 0.0;
    }
    __syncthreads();
    if (r >= n) return;
    double acc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = 0.0;
#pragma unroll 4
    for (int k = 0; k < p; ++k) {
        const double pv = Pm[(size_t)k * ldp + r];
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] = fma(pv, ss[k][j], acc[j]);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
        if (j0 + j < q) {
            double *yp = Y + (size_t)(j0 + j) * ldy + r;
            *yp = (beta == 0.0 ? 0.0 : beta * *yp) + alpha * acc[j];
        }
}

__global__ void k_diag_of(const double *__restrict__ A, int64_t lda, int64_t n, double *__restrict__ d)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r < n) d[r] = A[(size_t)r * lda + r];
}

// R = AX - X theta; T = R / (theta - diag) (Davidson correction).
__global__ void __launch_bounds__(256)
    k_residual(const double *__restrict__ AX, const double *__restrict__ X,
               const double *__restrict__ theta, const double *__restrict__ diag, int64_t n,
               int64_t ld, double *__restrict__ R, double *__restrict__ T)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (r >= n) return;
    const size_t idx = (size_t)j * ld + r;
    const double res = AX[idx] - X[idx] * theta[j];
    R[idx] = res;
    double den = theta[j] - diag[r];
    if (fabs(den) < 1e-8) den = den < 0 ? -1e-8 : 1e-8;
    T[idx] = res / den;
}

// host-side small dense helpers ---------------------------------------------

static bool cholesky_upper(int b, std::vector<double> &G)  // G = U^T U, U upper in place
{
    for (int j = 0; j < b; ++j) {
        for (int i = 0; i <= j; ++i) {
            double s = G[(size_t)j * b + i];
            for (int k = 0; k < i; ++k) s -= G[(size_t)i * b + k] * G[(size_t)j * b + k];
            if (i == j) {
                if (s <= 0) return false;
                G[(size_t)j * b + j] = std::sqrt(s);
            } else {
                G[(size_t)j * b + i] = s / G[(size_t)i * b + i];
            }
        }
        for (int i = j + 1; i < b; ++i) G[(size_t)j * b + i] = 0.0;
    }
    return true;
}

static std::vector<double> upper_inverse(int b, const std::vector<double> &U)
{
    std::vector<double> Ui((size_t)b * b, 0.0);
    for (int j = 0; j < b; ++j) {
        Ui[(size_t)j * b + j] = 1.0 / U[(size_t)j * b + j];
        for (int i = j - 1; i >= 0; --i) {
            double s = 0.0;
            for (int k = i + 1; k <= j; ++k) s += U[(size_t)k * b + i] * Ui[(size_t)j * b + k];
            Ui[(size_t)j * b + i] = -s / U[(size_t)i * b + i];
        }
    }
    return Ui;
}

}  // namespace scp3g

using namespace scp3g;

namespace {

struct Dav {
    scp3g_ctx *ctx;
    int64_t n, lda;
    const double *A;
    double *V, *W, *X, *AX, *R, *T, *AT, *S, *G, *theta, *diag, *part;
    int mmax, b;

    hipError_t symm(const double *Xp, int nb, double *Yp)
    {
        hipEvent_t ev;
        timer_begin(ctx, "symv", &ev);
        const int64_t kpad = (n + 31) / 32 * 32;
        const unsigned rblocks = (unsigned)((n + 127) / 128);
        // split k so the grid has ~1024 workgroups, >= 256 k per split
        int nsplit = 1;
        while ((int64_t)rblocks * nsplit < 1024 && kpad / (nsplit * 2) >= 256 && nsplit < kMaxSplit)
            nsplit *= 2;
        const int64_t kchunk = ((kpad + nsplit - 1) / nsplit + 31) / 32 * 32;
        nsplit = (int)((kpad + kchunk - 1) / kchunk);
        double *pp = nsplit > 1 ? part : nullptr;
        if (nb <= 16)
            hipLaunchKernelGGL(k_symm_mfma<1>, dim3(rblocks, (unsigned)nsplit), dim3(256), 0, ctx->stream, A,
                               lda, n, Xp, n, nb, kchunk, Yp, n, pp);
        else
            hipLaunchKernelGGL(k_symm_mfma<2>, dim3(rblocks, (unsigned)nsplit), dim3(256), 0, ctx->stream, A,
                               lda, n, Xp, n, nb, kchunk, Yp, n, pp);
        if (pp)
            hipLaunchKernelGGL(k_symm_reduce, dim3((unsigned)((n + 255) / 256), (unsigned)nb), dim3(256), 0,
                               ctx->stream, pp, nsplit, n, Yp, n);
        timer_end(ctx, "symv", ev);
        return hipGetLastError();
    }
    hipError_t tn(const double *Pm, int p, const double *Q, int q, double *C, int ldc)
    {
        hipLaunchKernelGGL(k_gemm_tn, dim3((unsigned)((p + 3) / 4), (unsigned)((q + 3) / 4)), dim3(64), 0,
                           ctx->stream, Pm, n, p, Q, n, q, n, C, ldc);
        return hipGetLastError();
    }
    hipError_t nn(const double *Pm, int p, const double *Sm, int lds, int q, double alpha,
                  double beta, double *Y)
    {
        if (p > 256) return hipErrorInvalidValue;  // k_gemm_nn stages at most 256 rows of S
        hipLaunchKernelGGL(k_gemm_nn, dim3((unsigned)((n + 63) / 64), (unsigned)((q + 7) / 8)), dim3(64), 0,
                           ctx->stream, Pm, n, p,
                           Sm, lds, q, alpha, beta, Y, n, n);
        return hipGetLastError();
    }
};

}  // namespace

// Lowest k eigenpairs of A = sym(Hsrc[0:n, 0:n] / P) + diag, where Hsrc is a
// device upper triangle with leading dimension ldh.
static int davidson_run(scp3g_ctx *ctx, const double *Hsrc, int64_t ldh, int64_t n, double P,
                        const std::vector<double> &dg, int k, double tol, double *evals,
                        double *evecs)
{
    if (k < 1 || k > 28 || k > n) return fail(ctx, -4, "k must be in 1..min(28, n)");
    if (!evals) return fail(ctx, -4, "null output");
    if (tol <= 0) tol = 1e-9;
    hipSetDevice(ctx->device);
    int rc;
    const int b = std::min<int64_t>(k + 4, std::min<int64_t>(kMaxB, n));
    const int mmax = (int)std::min<int64_t>(n, std::max(4 * b, 64));
    const size_t nn_ = (size_t)n;
    const int64_t lda = lda_of(n);
    double *d_A = nullptr, *d_buf = nullptr;
    const size_t nbuf = nn_ * (2 * mmax + 5 * kMaxB) + (size_t)mmax * mmax * 2 + 2 * nn_ +
                        (size_t)kMaxSplit * kMaxB * nn_;
    if ((rc = check_hip(ctx, hipMalloc(&d_A, (size_t)lda * nn_ * 8), "alloc A"))) return rc;
    if ((rc = check_hip(ctx, hipMalloc(&d_buf, nbuf * 8), "alloc davidson"))) {
        hipFree(d_A);
        return rc;
    }
    Dav dv;
    dv.ctx = ctx;
    dv.n = n;
    dv.lda = lda;
    dv.A = d_A;
    dv.mmax = mmax;
    dv.b = b;
    double *p = d_buf;
    dv.V = p; p += nn_ * mmax;
    dv.W = p; p += nn_ * mmax;
    dv.X = p; p += nn_ * kMaxB;
    dv.AX = p; p += nn_ * kMaxB;
    dv.R = p; p += nn_ * kMaxB;
    dv.T = p; p += nn_ * kMaxB;
    dv.AT = p; p += nn_ * kMaxB;
    dv.S = p; p += (size_t)mmax * mmax;
    dv.G = p; p += (size_t)mmax * mmax;
    dv.theta = p; p += nn_;
    dv.diag = p; p += nn_;
    dv.part = p; p += (size_t)kMaxSplit * kMaxB * nn_;

    auto cleanup = [&]() {
        hipStreamSynchronize(ctx->stream);
        hipFree(d_A);
        hipFree(d_buf);
    };
#define DV_CHECK(expr, what)                                    \
    do {                                                        \
        if ((rc = check_hip(ctx, (expr), what))) {              \
            cleanup();                                          \
            return rc;                                          \
        }                                                       \
    } while (0)

    DV_CHECK(hipMemcpyAsync(dv.diag, dg.data(), nn_ * 8, hipMemcpyHostToDevice, ctx->stream), "diag");
    // A's padding rows n..lda-1 take part in the k-chunks of A X: zero them
    DV_CHECK(hipMemsetAsync(d_A, 0, (size_t)lda * nn_ * 8, ctx->stream), "clear A");
    {
        hipEvent_t ev;
        timer_begin(ctx, "symfill", &ev);
        dim3 grid((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32));
        hipLaunchKernelGGL(k_sym_full, grid, dim3(256), 0, ctx->stream, Hsrc, ldh, n, P, dv.diag,
                           d_A, lda);
        timer_end(ctx, "symfill", ev);
        DV_CHECK(hipGetLastError(), "symfill");
        // the preconditioner and the start vectors use the diagonal of A itself
        hipLaunchKernelGGL(k_diag_of, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                           d_A, lda, n, dv.diag);
        DV_CHECK(hipGetLastError(), "diag");
    }
    if (n < 4 * b) {
        // too small for a subspace iteration (the restarted subspace plus a
        // block must fit in n): solve the dense problem on the host
        std::vector<double> Ah((size_t)lda * nn_), Gd(nn_ * nn_), w, Z;
        DV_CHECK(hipMemcpyAsync(Ah.data(), d_A, Ah.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "A");
        DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
        for (int64_t c = 0; c < n; ++c)
            std::copy(Ah.begin() + (size_t)c * lda, Ah.begin() + (size_t)c * lda + n, Gd.begin() + (size_t)c * n);
        small_eigh((int)n, Gd, k, w, Z);
        for (int j = 0; j < k; ++j) evals[j] = w[j];
        if (evecs) std::copy(Z.begin(), Z.begin() + nn_ * k, evecs);
        cleanup();
        return 0;
    }
    std::vector<double> dA(nn_);
    DV_CHECK(hipMemcpyAsync(dA.data(), dv.diag, nn_ * 8, hipMemcpyDeviceToHost, ctx->stream), "diag");
    DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
    // start: unit vectors on the b lowest diagonal entries
    std::vector<int64_t> order(nn_);
    std::iota(order.begin(), order.end(), 0);
    std::partial_sort(order.begin(), order.begin() + b, order.end(),
                      [&](int64_t x, int64_t y) { return dA[x] < dA[y]; });
    {
        std::vector<double> V0(nn_ * b, 0.0);
        for (int j = 0; j < b; ++j) V0[(size_t)j * n + order[j]] = 1.0;
        DV_CHECK(hipMemcpyAsync(dv.V, V0.data(), V0.size() * 8, hipMemcpyHostToDevice, ctx->stream),
                 "V0");
        DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
    }
    // G = V^T W is kept on the host (ld mmax) and extended by the columns
    // of each new block; after a restart it is recomputed (keep x keep).
    int m = b, m0 = 0;
    DV_CHECK(dv.symm(dv.V, b, dv.W), "AV");
    std::vector<double> Gh((size_t)mmax * mmax, 0.0), Cn, theta, Sb, Sp, Gm;
    int sp_rows = 0;  // rows of Sp (previous Ritz coefficients), 0 = none
    std::vector<double> resn(b);
    int iter = 0;
    bool converged = false;
    const int maxiter = 400;
    for (; iter < maxiter; ++iter) {
        // new columns m0..m-1 of G
        const int nc = m - m0;
        DV_CHECK(dv.tn(dv.V, m, dv.W + (size_t)m0 * n, nc, dv.G, m), "VtW");
        Cn.resize((size_t)m * nc);
        DV_CHECK(hipMemcpyAsync(Cn.data(), dv.G, Cn.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "G");
        DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
        for (int c = 0; c < nc; ++c)
            for (int r = 0; r < m; ++r) {
                const double v = Cn[(size_t)c * m + r];
                Gh[(size_t)(m0 + c) * mmax + r] = v;
                if (r < m0) Gh[(size_t)r * mmax + m0 + c] = v;
            }
        for (int c = m0; c < m; ++c)
            for (int r = m0; r < c; ++r) {
                const double v = 0.5 * (Gh[(size_t)c * mmax + r] + Gh[(size_t)r * mmax + c]);
                Gh[(size_t)c * mmax + r] = Gh[(size_t)r * mmax + c] = v;
            }
        Gm.resize((size_t)m * m);
        for (int c = 0; c < m; ++c)
            std::copy(Gh.begin() + (size_t)c * mmax, Gh.begin() + (size_t)c * mmax + m, Gm.begin() + (size_t)c * m);
        small_eigh(m, Gm, b, theta, Sb);
        // X = V S_b, AX = W S_b
        DV_CHECK(hipMemcpyAsync(dv.S, Sb.data(), (size_t)m * b * 8, hipMemcpyHostToDevice, ctx->stream), "S");
        DV_CHECK(hipMemcpyAsync(dv.theta, theta.data(), (size_t)b * 8, hipMemcpyHostToDevice, ctx->stream),
                 "theta");
        DV_CHECK(dv.nn(dv.V, m, dv.S, m, b, 1.0, 0.0, dv.X), "X");
        DV_CHECK(dv.nn(dv.W, m, dv.S, m, b, 1.0, 0.0, dv.AX), "AX");
        hipLaunchKernelGGL(k_residual, dim3((unsigned)((n + 255) / 256), (unsigned)b), dim3(256), 0,
                           ctx->stream, dv.AX, dv.X, dv.theta, dv.diag, n, n, dv.R, dv.T);
        DV_CHECK(hipGetLastError(), "residual");
        // residual norms
        DV_CHECK(dv.tn(dv.R, b, dv.R, b, dv.G, b), "RtR");
        std::vector<double> RR((size_t)b * b);
        DV_CHECK(hipMemcpyAsync(RR.data(), dv.G, RR.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "RR");
        DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
        double worst = 0.0;
        for (int j = 0; j < b; ++j) {
            resn[j] = std::sqrt(std::max(0.0, RR[(size_t)j * b + j]));
            if (j < k) worst = std::max(worst, resn[j]);
        }
        if (worst < tol || m >= n) {
            converged = true;
            break;
        }
        if (m + b > mmax) {
            // thick restart (GD+k; Stathopoulos & Saad 1998): keep the current
            // Ritz vectors and the previous iteration's, orthonormalised in
            // the m-dimensional coefficient space, so V <- V C, W <- W C
            std::vector<double> C((size_t)m * kMaxB, 0.0);
            std::copy(Sb.begin(), Sb.begin() + (size_t)m * b, C.begin());
            int keep = b;
            const int kprev = sp_rows > 0 ? std::min(b, kMaxB - b) : 0;
            for (int q = 0; q < kprev; ++q) {
                std::vector<double> x((size_t)m, 0.0);
                for (int r = 0; r < sp_rows; ++r) x[r] = Sp[(size_t)q * sp_rows + r];
                for (int pass = 0; pass < 2; ++pass)
                    for (int c = 0; c < keep; ++c) {
                        const double *cc = &C[(size_t)c * m];
                        double s = 0.0;
                        for (int r = 0; r < m; ++r) s += cc[r] * x[r];
                        for (int r = 0; r < m; ++r) x[r] -= s * cc[r];
                    }
                double nr = 0.0;
                for (int r = 0; r < m; ++r) nr += x[r] * x[r];
                nr = std::sqrt(nr);
                if (nr < 1e-8) continue;  // no new direction
                for (int r = 0; r < m; ++r) C[(size_t)keep * m + r] = x[r] / nr;
                ++keep;
            }
            DV_CHECK(hipMemcpyAsync(dv.S, C.data(), (size_t)m * keep * 8, hipMemcpyHostToDevice, ctx->stream),
                     "C");
            DV_CHECK(dv.nn(dv.V, m, dv.S, m, keep, 1.0, 0.0, dv.AT), "VC");
            DV_CHECK(hipMemcpyAsync(dv.V, dv.AT, nn_ * keep * 8, hipMemcpyDeviceToDevice, ctx->stream), "V");
            DV_CHECK(dv.nn(dv.W, m, dv.S, m, keep, 1.0, 0.0, dv.AT), "WC");
            DV_CHECK(hipMemcpyAsync(dv.W, dv.AT, nn_ * keep * 8, hipMemcpyDeviceToDevice, ctx->stream), "W");
            m = keep;
            m0 = 0;
            sp_rows = 0;
        } else {
            Sp = Sb;
            sp_rows = m;
        }
        const bool restarted = sp_rows == 0;
        // orthogonalise T against V twice: T -= V (V^T T)
        for (int pass = 0; pass < 2; ++pass) {
            DV_CHECK(dv.tn(dv.V, m, dv.T, b, dv.G, m), "VtT");
            DV_CHECK(dv.nn(dv.V, m, dv.G, m, b, -1.0, 1.0, dv.T), "T-=VC");
        }
        // CholQR2 within T (drop to fewer columns if rank deficient); the
        // second pass writes the new block straight into V's next columns
        int nb = b;
        {
            const double *src = dv.T;
            for (int pass = 0; pass < 2; ++pass) {
                double *dst = pass == 0 ? dv.AT : dv.V + (size_t)m * n;
                DV_CHECK(dv.tn(src, nb, src, nb, dv.G, nb), "TtT");
                std::vector<double> U((size_t)nb * nb);
                DV_CHECK(hipMemcpyAsync(U.data(), dv.G, U.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "TtT");
                DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
                if (!cholesky_upper(nb, U)) {
                    nb = 0;
                    break;
                }
                std::vector<double> Ui = upper_inverse(nb, U);
                DV_CHECK(hipMemcpyAsync(dv.S, Ui.data(), Ui.size() * 8, hipMemcpyHostToDevice, ctx->stream), "Ui");
                DV_CHECK(dv.nn(src, nb, dv.S, nb, nb, 1.0, 0.0, dst), "T U^-1");
                src = dst;
            }
        }
        if (nb == 0) {
            // correction collapsed: fall back to the (orthogonalised) residuals
            DV_CHECK(hipMemcpyAsync(dv.T, dv.R, nn_ * b * 8, hipMemcpyDeviceToDevice, ctx->stream), "T=R");
            for (int pass = 0; pass < 2; ++pass) {
                DV_CHECK(dv.tn(dv.V, m, dv.T, b, dv.G, m), "VtR");
                DV_CHECK(dv.nn(dv.V, m, dv.G, m, b, -1.0, 1.0, dv.T), "R-=VC");
            }
            nb = b;
            DV_CHECK(dv.tn(dv.T, nb, dv.T, nb, dv.G, nb), "RtR2");
            std::vector<double> U((size_t)nb * nb);
            DV_CHECK(hipMemcpyAsync(U.data(), dv.G, U.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "RtR2");
            DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
            if (!cholesky_upper(nb, U)) {
                rc = fail(ctx, -8, "Davidson: search space collapsed");
                cleanup();
                return rc;
            }
            std::vector<double> Ui = upper_inverse(nb, U);
            DV_CHECK(hipMemcpyAsync(dv.S, Ui.data(), Ui.size() * 8, hipMemcpyHostToDevice, ctx->stream), "Ui");
            DV_CHECK(dv.nn(dv.T, nb, dv.S, nb, nb, 1.0, 0.0, dv.V + (size_t)m * n), "R U^-1");
        }
        DV_CHECK(dv.symm(dv.V + (size_t)m * n, nb, dv.W + (size_t)m * n), "AT");
        m0 = restarted ? 0 : m;
        m += nb;
    }
    if (!converged) {
        rc = fail(ctx, -9, "Davidson did not converge");
        cleanup();
        return rc;
    }
    for (int j = 0; j < k; ++j) evals[j] = theta[j];
    if (evecs) {
        std::vector<double> Xh(nn_ * b);
        DV_CHECK(hipMemcpyAsync(Xh.data(), dv.X, Xh.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "X");
        DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
        std::copy(Xh.begin(), Xh.begin() + nn_ * k, evecs);
    }
    cleanup();
#undef DV_CHECK
    return 0;
}

extern "C" int scp3g_lowest_levels(scp3g_ctx *ctx, int k, double tol, double *evals,
                                   double *evecs)
{
    if (!ctx) return -1;
    if (ctx->nbasis <= 0 || !ctx->d_H) return fail(ctx, -2, "no Hamiltonian accumulated");
    if (ctx->points_done <= 0) return fail(ctx, -6, "no points accumulated");
    return davidson_run(ctx, ctx->d_H, ctx->nbasis, ctx->nbasis, ctx->norm_sum,
                        ho_diagonal(ctx), k, tol, evals, evecs);
}

extern "C" int scp3g_lowest_levels_sub(scp3g_ctx *ctx, int64_t nsub, int k, double tol,
                                       double *evals)
{
    if (!ctx) return -1;
    if (ctx->nbasis <= 0 || !ctx->d_H) return fail(ctx, -2, "no Hamiltonian accumulated");
    if (ctx->points_done <= 0) return fail(ctx, -6, "no points accumulated");
    if (nsub < 1 || nsub > ctx->nbasis) return fail(ctx, -4, "nsub out of range");
    std::vector<double> dg = ho_diagonal(ctx);
    dg.resize((size_t)nsub);
    return davidson_run(ctx, ctx->d_H, ctx->nbasis, nsub, ctx->norm_sum, dg, k, tol,
                        evals, nullptr);
}

extern "C" int scp3g_matrix_lowest_levels(scp3g_ctx *ctx, int64_t n, const double *M_host, int k,
                                          double tol, double *evals, double *evecs)
{
    if (!ctx) return -1;
    if (n < 1 || !M_host) return fail(ctx, -4, "bad matrix");
    hipSetDevice(ctx->device);
    // eigensolver.cpp:27-30: the upper triangle unless M(0, n-1) == 0, then
    // the lower one; the device copy always holds the chosen triangle as upper
    const bool lower = n > 1 && M_host[(size_t)(n - 1) * n] == 0.0;
    std::vector<double> U((size_t)n * n, 0.0);
    for (int64_t c = 0; c < n; ++c)
        for (int64_t r = 0; r <= c; ++r)
            U[(size_t)c * n + r] = lower ? M_host[(size_t)r * n + c] : M_host[(size_t)c * n + r];
    double *d_M = nullptr;
    int rc = check_hip(ctx, hipMalloc(&d_M, U.size() * 8), "alloc M");
    if (rc) return rc;
    rc = check_hip(ctx, hipMemcpy(d_M, U.data(), U.size() * 8, hipMemcpyHostToDevice), "upload M");
    if (!rc) rc = davidson_run(ctx, d_M, n, n, 1.0, std::vector<double>((size_t)n, 0.0), k, tol, evals, evecs);
    hipFree(d_M);
    return rc;
}
