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

// A(r, c) = Mout(min, max) / P (+ diag on r == c) from the upper-triangle sum
// H, through 32 x 32 LDS tiles so that both the read of the mirrored
// triangle and the write stay coalesced.
__global__ void __launch_bounds__(256)
    k_sym_full(const double *__restrict__ H, int64_t ldh, int64_t n, double P,
               const double *__restrict__ diag, double *__restrict__ A)
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
                A[(size_t)c * n + r] = v;
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
            if (r < n && c < n) A[(size_t)c * n + r] = t[tx][k];
        }
    }
}

// Y(n x b) = A X for symmetric A: Y(r, j) = sum_c A(c, r) X(c, j).
// A block owns 64 output rows r; it streams c in chunks of 64 through LDS.
constexpr int kMvRows = 64;
constexpr int kMvChunk = 64;
constexpr int kMaxB = 32;

__global__ void __launch_bounds__(256)
    k_symm_block(const double *__restrict__ A, int64_t n, const double *__restrict__ X,
                 int64_t ldx, int b, double *__restrict__ Y, int64_t ldy)
{
    __shared__ double As[kMvChunk][kMvRows + 1];
    __shared__ double Xs[kMvChunk][kMaxB];
    const int64_t r0 = (int64_t)blockIdx.x * kMvRows;
    const int ri = threadIdx.x & 63;  // output row within the block
    const int jg = threadIdx.x >> 6;  // column group: j = jg, jg+4, ...
    double acc[kMaxB / 4];
#pragma unroll
    for (int q = 0; q < kMaxB / 4; ++q) acc[q] = 0.0;
    for (int64_t c0 = 0; c0 < n; c0 += kMvChunk) {
        // A(c0 + cc, r0 + rr): column r0+rr is contiguous in c
        for (int e = threadIdx.x; e < kMvChunk * kMvRows; e += 256) {
            const int cc = e & 63, rr = e >> 6;
            const int64_t c = c0 + cc, r = r0 + rr;
            As[cc][rr] = (c < n && r < n) ? A[(size_t)r * n + c] : 0.0;
        }
        for (int e = threadIdx.x; e < kMvChunk * kMaxB; e += 256) {
            const int cc = e % kMvChunk, j = e / kMvChunk;
            const int64_t c = c0 + cc;
            Xs[cc][j] = (c < n && j < b) ? X[(size_t)j * ldx + c] : 0.0;
        }
        __syncthreads();
        for (int cc = 0; cc < kMvChunk; ++cc) {
            const double a = As[cc][ri];
#pragma unroll
            for (int q = 0; q < kMaxB / 4; ++q) acc[q] = fma(a, Xs[cc][jg + 4 * q], acc[q]);
        }
        __syncthreads();
    }
    const int64_t r = r0 + ri;
    if (r < n)
#pragma unroll
        for (int q = 0; q < kMaxB / 4; ++q) {
            const int j = jg + 4 * q;
            if (j < b) Y[(size_t)j * ldy + r] = acc[q];
        }
}

// C(p x q) = P^T Q for tall P (n x p, ld ldp) and Q (n x q, ld ldq);
// block (i, j-group of 8) reduces over n.
__global__ void __launch_bounds__(256)
    k_gemm_tn(const double *__restrict__ Pm, int64_t ldp, const double *__restrict__ Q,
              int64_t ldq, int64_t n, int q, double *__restrict__ C, int ldc)
{
    __shared__ double red[8][256];
    const int i = blockIdx.x;
    const int j0 = blockIdx.y * 8;
    double acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = 0.0;
    const double *__restrict__ pc = Pm + (size_t)i * ldp;
    for (int64_t k = threadIdx.x; k < n; k += 256) {
        const double pv = pc[k];
#pragma unroll
        for (int t = 0; t < 8; ++t)
            if (j0 + t < q) acc[t] = fma(pv, Q[(size_t)(j0 + t) * ldq + k], acc[t]);
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) red[t][threadIdx.x] = acc[t];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int t = 0; t < 8; ++t) red[t][threadIdx.x] += red[t][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 8 && j0 + (int)threadIdx.x < q)
        C[(size_t)(j0 + threadIdx.x) * ldc + i] = red[threadIdx.x][0];
}

// Y(n x q) = beta*Y + alpha * P(n x p) S(p x q); S column-major with ld lds.
__global__ void __launch_bounds__(256)
    k_gemm_nn(const double *__restrict__ Pm, int64_t ldp, int p, const double *__restrict__ S,
              int lds, int q, double alpha, double beta, double *__restrict__ Y, int64_t ldy,
              int64_t n)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (r >= n) return;
    double acc = 0.0;
    for (int k = 0; k < p; ++k) acc = fma(Pm[(size_t)k * ldp + r], S[(size_t)j * lds + k], acc);
    double *yp = Y + (size_t)j * ldy + r;
    *yp = (beta == 0.0 ? 0.0 : beta * *yp) + alpha * acc;
}

__global__ void k_diag_of(const double *__restrict__ A, int64_t n, double *__restrict__ d)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r < n) d[r] = A[(size_t)r * n + r];
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
    int64_t n;
    const double *A;
    double *V, *W, *X, *AX, *R, *T, *AT, *S, *G, *theta, *diag;
    int mmax, b;

    hipError_t symm(const double *Xp, int nb, double *Yp)
    {
        hipEvent_t ev;
        timer_begin(ctx, "symv", &ev);
        hipLaunchKernelGGL(k_symm_block, dim3((unsigned)((n + kMvRows - 1) / kMvRows)), dim3(256), 0,
                           ctx->stream, A, n, Xp, n, nb, Yp, n);
        timer_end(ctx, "symv", ev);
        return hipGetLastError();
    }
    hipError_t tn(const double *Pm, int p, const double *Q, int q, double *C, int ldc)
    {
        hipLaunchKernelGGL(k_gemm_tn, dim3((unsigned)p, (unsigned)((q + 7) / 8)), dim3(256), 0,
                           ctx->stream, Pm, n, Q, n, n, q, C, ldc);
        return hipGetLastError();
    }
    hipError_t nn(const double *Pm, int p, const double *Sm, int lds, int q, double alpha,
                  double beta, double *Y)
    {
        hipLaunchKernelGGL(k_gemm_nn, dim3((unsigned)((n + 255) / 256), (unsigned)q), dim3(256), 0,
                           ctx->stream, Pm, n, p, Sm, lds, q, alpha, beta, Y, n, n);
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
    double *d_A = nullptr, *d_buf = nullptr;
    const size_t nbuf = nn_ * (2 * mmax + 5 * kMaxB) + (size_t)mmax * mmax * 2 + 2 * nn_;
    if ((rc = check_hip(ctx, hipMalloc(&d_A, nn_ * nn_ * 8), "alloc A"))) return rc;
    if ((rc = check_hip(ctx, hipMalloc(&d_buf, nbuf * 8), "alloc davidson"))) {
        hipFree(d_A);
        return rc;
    }
    Dav dv;
    dv.ctx = ctx;
    dv.n = n;
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
    {
        hipEvent_t ev;
        timer_begin(ctx, "symfill", &ev);
        dim3 grid((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32));
        hipLaunchKernelGGL(k_sym_full, grid, dim3(256), 0, ctx->stream, Hsrc, ldh, n, P, dv.diag,
                           d_A);
        timer_end(ctx, "symfill", ev);
        DV_CHECK(hipGetLastError(), "symfill");
        // the preconditioner and the start vectors use the diagonal of A itself
        hipLaunchKernelGGL(k_diag_of, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                           d_A, n, dv.diag);
        DV_CHECK(hipGetLastError(), "diag");
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
    int m = b;
    DV_CHECK(dv.symm(dv.V, b, dv.W), "AV");
    std::vector<double> Gh((size_t)mmax * mmax), theta, Sfull, Gm;
    std::vector<double> resn(b);
    int iter = 0;
    bool converged = false;
    const int maxiter = 400;
    for (; iter < maxiter; ++iter) {
        // G = V^T W (m x m)
        DV_CHECK(dv.tn(dv.V, m, dv.W, m, dv.G, m), "VtW");
        DV_CHECK(hipMemcpyAsync(Gh.data(), dv.G, (size_t)m * m * 8, hipMemcpyDeviceToHost, ctx->stream),
                 "G");
        DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
        Gm.assign(Gh.begin(), Gh.begin() + (size_t)m * m);
        for (int c = 0; c < m; ++c)
            for (int r = 0; r < c; ++r) {
                const double v = 0.5 * (Gm[(size_t)c * m + r] + Gm[(size_t)r * m + c]);
                Gm[(size_t)c * m + r] = Gm[(size_t)r * m + c] = v;
            }
        jacobi_eigh(m, Gm, theta, Sfull);
        // X = V S_b, AX = W S_b
        DV_CHECK(hipMemcpyAsync(dv.S, Sfull.data(), (size_t)m * b * 8, hipMemcpyHostToDevice, ctx->stream),
                 "S");
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
        // restart if the new block does not fit
        if (m + b > mmax) {
            DV_CHECK(hipMemcpyAsync(dv.V, dv.X, nn_ * b * 8, hipMemcpyDeviceToDevice, ctx->stream), "restart V");
            DV_CHECK(hipMemcpyAsync(dv.W, dv.AX, nn_ * b * 8, hipMemcpyDeviceToDevice, ctx->stream), "restart W");
            m = b;
        }
        // orthogonalise T against V twice: T -= V (V^T T)
        for (int pass = 0; pass < 2; ++pass) {
            DV_CHECK(dv.tn(dv.V, m, dv.T, b, dv.G, m), "VtT");
            DV_CHECK(dv.nn(dv.V, m, dv.G, m, b, -1.0, 1.0, dv.T), "T-=VC");
        }
        // CholQR2 within T (drop to fewer columns if rank deficient)
        int nb = b;
        for (int pass = 0; pass < 2; ++pass) {
            DV_CHECK(dv.tn(dv.T, nb, dv.T, nb, dv.G, nb), "TtT");
            std::vector<double> U((size_t)nb * nb);
            DV_CHECK(hipMemcpyAsync(U.data(), dv.G, U.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "TtT");
            DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
            if (!cholesky_upper(nb, U)) {
                nb = 0;
                break;
            }
            std::vector<double> Ui = upper_inverse(nb, U);
            DV_CHECK(hipMemcpyAsync(dv.S, Ui.data(), Ui.size() * 8, hipMemcpyHostToDevice, ctx->stream), "Ui");
            DV_CHECK(dv.nn(dv.T, nb, dv.S, nb, nb, 1.0, 0.0, dv.AT), "T U^-1");
            DV_CHECK(hipMemcpyAsync(dv.T, dv.AT, nn_ * nb * 8, hipMemcpyDeviceToDevice, ctx->stream), "T");
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
            DV_CHECK(dv.nn(dv.T, nb, dv.S, nb, nb, 1.0, 0.0, dv.AT), "R U^-1");
            DV_CHECK(hipMemcpyAsync(dv.T, dv.AT, nn_ * nb * 8, hipMemcpyDeviceToDevice, ctx->stream), "T");
        }
        // append
        DV_CHECK(hipMemcpyAsync(dv.V + (size_t)m * n, dv.T, nn_ * nb * 8, hipMemcpyDeviceToDevice, ctx->stream),
                 "append V");
        DV_CHECK(dv.symm(dv.V + (size_t)m * n, nb, dv.W + (size_t)m * n), "AT");
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
