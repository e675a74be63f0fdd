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

// A(r, c) = Mout(min, max) / P (+ O(min, max), + diag on r == c) from the
// upper-triangle sum H and the optional fixed term O (same layout), through
// 32 x 32 LDS tiles so that both the read of the mirrored triangle and the
// write stay coalesced.  A has leading dimension lda (a multiple of 32); its
// padding rows n..lda-1 are set to zero.
__global__ void __launch_bounds__(256)
    k_sym_full(const double *__restrict__ H, const double *__restrict__ O, int64_t ldh, int64_t n, double P,
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
                const size_t e = r <= c ? (size_t)c * ldh + r : (size_t)r * ldh + c;
                double v = H[e] / P;
                if (O) v += O[e];
                if (r == c) v += diag[r];
                A[(size_t)c * lda + r] = v;
            } else if (r < lda && c < n) {
                A[(size_t)c * lda + r] = 0.0;
            }
        }
    } else {
        // lower tile: A(r, c) = H(c, r) for r > c; read the mirror tile
        // (rows bc.., cols br..) coalesced, transpose through LDS
        for (int k = ty; k < 32; k += 8) {
            const int64_t r = bc + tx, c = br + k;  // element H(r, c), r < c
            t[k][tx] = (r < n && c < n) ? H[(size_t)c * ldh + r] / P + (O ? O[(size_t)c * ldh + r] : 0.0) : 0.0;
        }
        __syncthreads();
        for (int k = ty; k < 32; k += 8) {
            const int64_t r = br + tx, c = bc + k;
            if (r < lda && c < n) A[(size_t)c * lda + r] = r < n ? t[tx][k] : 0.0;
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

__global__ void k_diag_of(const double *__restrict__ A, int64_t lda, int64_t n, double *__restrict__ d)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r < n) d[r] = A[(size_t)r * lda + r];
}

// C(p x q) = P^T Q for tall P = [P0 | P1] (n x (p0 + p1), both with ld ldp)
// and Q (n x q, ld ldq).  One 256-thread workgroup owns a 4 x 4 block of C:
// the threads stride over n, each wave reduces its 16 sums with shuffles,
// and the four wave sums are added in wave order (a fixed order, so results
// do not vary between runs).
__global__ void __launch_bounds__(256)
    k_gemm_tn(const double *__restrict__ P0, int p0, const double *__restrict__ P1, int p1, int64_t ldp,
              const double *__restrict__ Q, int64_t ldq, int q, int64_t n, double *__restrict__ C, int ldc)
{
    __shared__ double red[4][16];
    const int p = p0 + p1;
    const int i0 = blockIdx.x * 4, j0 = blockIdx.y * 4;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double *pc[4], *qc[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int i = i0 + a < p ? i0 + a : i0;
        pc[a] = i < p0 ? P0 + (size_t)i * ldp : P1 + (size_t)(i - p0) * ldp;
        qc[a] = Q + (size_t)(j0 + a < q ? j0 + a : j0) * ldq;
    }
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;
#pragma unroll 2
    for (int64_t k = tid; k < n; k += 256) {
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
    if (lane == 0) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) red[wave][4 * a + c] = acc[a][c];
    }
    __syncthreads();
    if (tid < 16) {
        const int a = tid >> 2, c = tid & 3;
        const double v = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
        if (i0 + a < p && j0 + c < q) C[(size_t)(j0 + c) * ldc + i0 + a] = v;
    }
}

// Partial products of one 256-thread workgroup: 64 rows r and 8 columns
// j0.. of P S for P = [P0 | P1] (p0 + p1 columns, ld ldp) and S (p x q, ld
// lds, staged in LDS); the four waves take consecutive quarters of the
// contraction index and their sums meet in LDS, added in wave order.  The
// result is valid in wave 0 (acc[j] for column j0 + j of row r).
struct NNBlock {
    double acc[8];
};
__device__ __forceinline__ void nn_block(const double *__restrict__ P0, int p0, const double *__restrict__ P1,
                                         int p1, int64_t ldp, const double *__restrict__ S, int lds, int q,
                                         int64_t n, int64_t r, int j0, double *ss, double *red, NNBlock &out)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int p = p0 + p1;
    for (int e = tid; e < p * 8; e += 256) {
        const int k = e >> 3, j = e & 7;
        ss[k * 8 + j] = (j0 + j < q) ? S[(size_t)(j0 + j) * lds + k] : 0.0;
    }
    __syncthreads();
    const int kq = (p + 3) / 4, k0 = wave * kq, k1 = k0 + kq < p ? k0 + kq : p;
#pragma unroll
    for (int j = 0; j < 8; ++j) out.acc[j] = 0.0;
    if (r < n) {
#pragma unroll 8
        for (int k = k0; k < k1; ++k) {
            const double pv = k < p0 ? P0[(size_t)k * ldp + r] : P1[(size_t)(k - p0) * ldp + r];
#pragma unroll
            for (int j = 0; j < 8; ++j) out.acc[j] = fma(pv, ss[k * 8 + j], out.acc[j]);
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int j = 0; j < 8; ++j) red[((wave - 1) * 8 + j) * 64 + lane] = out.acc[j];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
            out.acc[j] = (out.acc[j] + red[j * 64 + lane]) + (red[(8 + j) * 64 + lane] + red[(16 + j) * 64 + lane]);
    }
}

constexpr int kMaxP = kMaxB + 128;  // contraction length limit of k_gemm_nn / k_ritz

// Y(n x q) = P S for P = [P0 | P1] (p0 + p1 <= kMaxP columns), q <= kMaxB;
// S column-major with ld lds.  Y may alias P1 (each workgroup reads all of
// its rows before it writes them).
__global__ void __launch_bounds__(256)
    k_gemm_nn(const double *P0, int p0, const double *P1, int p1, int64_t ldp, const double *__restrict__ S,
              int lds, int q, double *Y, int64_t ldy, int64_t n)
{
    __shared__ double ss[kMaxP * 8];
    __shared__ double red[3 * 8 * 64];
    const int64_t r = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int j0 = blockIdx.y * 8;
    NNBlock o;
    nn_block(P0, p0, P1, p1, ldp, S, lds, q, n, r, j0, ss, red, o);
    if (threadIdx.x < 64 && r < n) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (j0 + j < q) Y[(size_t)(j0 + j) * ldy + r] = o.acc[j];
    }
}

// Ritz vectors and Davidson corrections in one pass: X = V S, AX = W S
// (S: p x b, ld p), R = AX - X theta, T = R / (theta - diag) (theta follows
// S in memory: theta[j] = S[p * b + j]).  X, R and T are written; the sum
// of the squares of R over the workgroup's 64 rows of column j goes to
// part[j * gridDim.x + blockIdx.x].
__global__ void __launch_bounds__(256)
    k_ritz(const double *__restrict__ V, const double *__restrict__ W, int p, const double *__restrict__ S,
           int b, const double *__restrict__ diag, int64_t n, double *__restrict__ X, double *__restrict__ R,
           double *__restrict__ T, double *__restrict__ part)
{
    __shared__ double ss[kMaxP * 8];
    __shared__ double red[3 * 8 * 64];
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 64 + lane;
    const int j0 = blockIdx.y * 8;
    NNBlock x, ax;
    nn_block(V, p, nullptr, 0, n, S, p, b, n, r, j0, ss, red, x);
    __syncthreads();
    nn_block(W, p, nullptr, 0, n, S, p, b, n, r, j0, ss, red, ax);
    if (threadIdx.x >= 64) return;
    const double dr = r < n ? diag[r] : 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        double res = 0.0;
        if (j0 + j < b && r < n) {
            const double th = S[(size_t)p * b + j0 + j];
            res = ax.acc[j] - x.acc[j] * th;
            double den = th - dr;
            if (fabs(den) < 1e-8) den = den < 0 ? -1e-8 : 1e-8;
            const size_t idx = (size_t)(j0 + j) * n + r;
            X[idx] = x.acc[j];
            R[idx] = res;
            T[idx] = res / den;
        }
        double s2 = res * res;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s2 += __shfl_xor(s2, off);
        if (lane == 0 && j0 + j < b) part[(size_t)(j0 + j) * gridDim.x + blockIdx.x] = s2;
    }
}

// One pass of block Gram-Schmidt plus CholQR for a block T (n x nb) against
// an orthonormal V (n x p0), from the Gram matrix of [V T] with T:
// Gf = [V T]^T T ((p0 + nb) x nb, ld p0 + nb), C = V^T T its first p0 rows.
// Since V^T V = I, (T - V C)^T (T - V C) = T^T T - C^T C = U^T U (Cholesky),
// and Q = (T - V C) U^-1 = [V T] [-C U^-1; U^-1] is orthonormal; the
// coefficients Sf = [-C U^-1; U^-1] ((p0 + nb) x nb, ld p0 + nb) are the
// output.  A column whose Cholesky pivot is at most 1e-12 of its squared
// norm in T lies in the span of V and its predecessors (to 1e-6 relative):
// it is dropped, i.e. its coefficient column is zero and Q has a zero
// column there.  NB >= nb; 256 threads, the factorisation in wave 0 (lane
// c owns column c; the broadcasts are v_readlane of a fixed lane).
template <int NB>
__global__ void __launch_bounds__(256)
    k_chol_block(const double *__restrict__ Gf, int p0, int nb, double *__restrict__ Sf)
{
    __shared__ double Cs[128][NB + 1];
    __shared__ double G22[NB][NB + 1];
    __shared__ double U[NB][NB + 1];
    __shared__ double Ui[NB][NB + 1];
    const int tid = threadIdx.x;
    const int ld = p0 + nb;
    for (int e = tid; e < p0 * nb; e += 256) {
        const int k = e % p0, c = e / p0;
        Cs[k][c] = Gf[(size_t)c * ld + k];
    }
    __syncthreads();
    // G22 = T^T T - C^T C
    for (int e = tid; e < nb * nb; e += 256) {
        const int rr = e % nb, c = e / nb;
        double s = 0.0;
        for (int k = 0; k < p0; ++k) s = fma(Cs[k][rr], Cs[k][c], s);
        G22[rr][c] = Gf[(size_t)c * ld + p0 + rr] - s;
    }
    __syncthreads();
    if (tid < 64) {
        const int c = tid;
        double col[NB], gref = 0.0;
#pragma unroll
        for (int r = 0; r < NB; ++r) col[r] = (c < nb && r < nb) ? G22[r][c] : 0.0;
        if (c < nb) gref = Gf[(size_t)c * ld + p0 + c];
        auto bcast = [](double v, int l) {
            const long long u = __double_as_longlong(v);
            const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffff), l);
            const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
            return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
        };
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < nb) {
                const double pivot = bcast(col[j], j);
                const double gjj = bcast(gref, j);
                const bool keep = pivot > 1e-12 * gjj && pivot > 0.0;
                const double ujj = keep ? sqrt(pivot) : 0.0;
                const double ujc = !keep ? 0.0 : (c > j ? col[j] / ujj : (c == j ? ujj : 0.0));
                if (c < NB) U[j][c] = ujc;
#pragma unroll
                for (int r = j + 1; r < NB; ++r) col[r] -= bcast(ujc, r) * ujc;
            }
        }
    }
    __syncthreads();
    // column c of U^-1 by back substitution (zero for dropped rows/columns)
    if (tid < nb) {
        const int c = tid;
        double x[NB];
#pragma unroll
        for (int i = NB - 1; i >= 0; --i) {
            double v = 0.0;
            if (i <= c) {
                const double uii = U[i][i];
                if (uii != 0.0) {
                    double s = i == c ? 1.0 : 0.0;
#pragma unroll
                    for (int k = i + 1; k < NB; ++k)
                        if (k <= c) s -= U[i][k] * x[k];
                    v = s / uii;
                }
            }
            x[i] = v;
        }
#pragma unroll
        for (int i = 0; i < NB; ++i)
            if (i < nb) Ui[i][c] = x[i];
    }
    __syncthreads();
    // Sf = [-C U^-1; U^-1]
    for (int e = tid; e < ld * nb; e += 256) {
        const int rr = e % ld, c = e / ld;
        double v;
        if (rr < p0) {
            double s = 0.0;
            for (int k = 0; k < nb; ++k) s = fma(Cs[rr][k], Ui[k][c], s);
            v = -s;
        } else {
            v = Ui[rr - p0][c];
        }
        Sf[(size_t)c * ld + rr] = v;
    }
}

}  // namespace scp3g

using namespace scp3g;

void scp3g::release_davidson(scp3g_ctx *ctx)
{
    if (ctx->d_dav) hipFree(ctx->d_dav);
    if (ctx->h_dav) hipHostFree(ctx->h_dav);
    ctx->d_dav = nullptr;
    ctx->h_dav = nullptr;
    ctx->dav_bytes = ctx->hdav_bytes = 0;
}

namespace {

struct Dav {
    scp3g_ctx *ctx;
    int64_t n, lda;
    const double *A;
    double *V, *W, *X, *R, *T, *AT, *S, *SC, *G, *diag, *rr, *part;
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
    // C = [P0 | P1]^T Q, ld ldc
    hipError_t tn(const double *P0, int p0, const double *P1, int p1, const double *Q, int q, double *C, int ldc)
    {
        const int p = p0 + p1;
        hipLaunchKernelGGL(k_gemm_tn, dim3((unsigned)((p + 3) / 4), (unsigned)((q + 3) / 4)), dim3(256), 0,
                           ctx->stream, P0, p0, P1, p1, n, Q, n, q, n, C, ldc);
        return hipGetLastError();
    }
    // Y = [P0 | P1] S (S: (p0 + p1) x q, ld lds)
    hipError_t nn(const double *P0, int p0, const double *P1, int p1, const double *Sm, int lds, int q, double *Y)
    {
        if (p0 + p1 > kMaxP || q > kMaxB) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_gemm_nn, dim3((unsigned)((n + 63) / 64), (unsigned)((q + 7) / 8)), dim3(256), 0,
                           ctx->stream, P0, p0, P1, p1, n, Sm, lds, q, Y, n, n);
        return hipGetLastError();
    }
    // X = V S, R, T and the residual partial sums (S: m x b then theta)
    hipError_t ritz(int m)
    {
        if (m > kMaxP) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_ritz, dim3((unsigned)((n + 63) / 64), (unsigned)((b + 7) / 8)), dim3(256), 0,
                           ctx->stream, V, W, m, S, b, diag, n, X, R, T, rr);
        return hipGetLastError();
    }
    hipError_t chol_block(int p0, int nb)
    {
        if (p0 > 128 || nb > kMaxB) return hipErrorInvalidValue;
        if (nb <= 8)
            hipLaunchKernelGGL(k_chol_block<8>, dim3(1), dim3(256), 0, ctx->stream, G, p0, nb, S);
        else if (nb <= 16)
            hipLaunchKernelGGL(k_chol_block<16>, dim3(1), dim3(256), 0, ctx->stream, G, p0, nb, S);
        else if (nb <= 24)
            hipLaunchKernelGGL(k_chol_block<24>, dim3(1), dim3(256), 0, ctx->stream, G, p0, nb, S);
        else
            hipLaunchKernelGGL(k_chol_block<kMaxB>, dim3(1), dim3(256), 0, ctx->stream, G, p0, nb, S);
        return hipGetLastError();
    }
    // The block T (n x nb) orthonormalised against V[:, :m] and within
    // itself into dst: two passes of k_chol_block's Gram-matrix form of
    // block Gram-Schmidt + CholQR (dropped columns come out zero)
    hipError_t orth_cholqr(int m, int nb, double *dst)
    {
        hipError_t e;
        const double *src = T;
        for (int pass = 0; pass < 2; ++pass) {
            double *out = pass == 0 ? AT : dst;
            if ((e = tn(V, m, src, nb, src, nb, G, m + nb)) != hipSuccess) return e;
            if ((e = chol_block(m, nb)) != hipSuccess) return e;
            if ((e = nn(V, m, src, nb, S, m + nb, nb, out)) != hipSuccess) return e;
            src = out;
        }
        return hipSuccess;
    }
};

}  // namespace

// Lowest k eigenpairs of A = sym(Hsrc[0:n, 0:n] / P) + diag, where Hsrc is a
// device upper triangle with leading dimension ldh.
//
// One host synchronisation per iteration: the new columns of G = V^T W and
// the residual norms of the previous iteration arrive together, so the
// convergence test lags one iteration behind the GPU (the iterate it accepts,
// X and theta of the previous iteration, is still in place when it is
// tested).  The correction block is orthonormalised on the GPU (CholQR2 with
// column dropping); a dropped column stays in V as a zero column, which the
// host leaves out of the projected problem.
static int davidson_run(scp3g_ctx *ctx, const double *Hsrc, const double *Osrc, int64_t ldh, int64_t n,
                        double P, const std::vector<double> &dg, int k, double tol, double *evals,
                        double *evecs)
{
    if (k < 1 || k > 28 || k > n) return fail(ctx, -4, "k must be in 1..min(28, n)");
    if (!evals) return fail(ctx, -4, "null output");
    if (tol <= 0) tol = 1e-9;
    hipSetDevice(ctx->device);
    int rc;
    const int b = std::min<int64_t>(k + 4, std::min<int64_t>(kMaxB, n));
    // subspace limit: a larger subspace takes fewer iterations, the host's
    // projected eigenproblem costs O(m^3) per iteration (for 10 levels of the
    // 1225-state hexamer: 27 iterations at 64, 23 at 84, 22 at 128)
    const int mmax = (int)std::min<int64_t>(n, std::min(128, std::max(4 * b, 64)));
    const size_t nn_ = (size_t)n;
    const int64_t lda = lda_of(n);
    // device work space (kept in the context between solves)
    const size_t nA = (size_t)lda * nn_;
    // residual partial sums: 64-row chunks per column
    const int nchunk = (int)((n + 63) / 64);
    const size_t nrr = (size_t)kMaxB * nchunk, nG = (size_t)mmax * mmax, nS = (size_t)mmax * kMaxB + kMaxB,
                 nC = (size_t)mmax * 2 * kMaxB;
    const size_t nbuf = nn_ * (2 * mmax + 4 * kMaxB) + std::max(nS, nG) + nrr + nG + nC + nn_ +
                        (size_t)kMaxSplit * kMaxB * nn_;
    const size_t dbytes = (nA + nbuf) * 8;
    if (ctx->dav_bytes < dbytes) {
        if (ctx->d_dav) hipFree(ctx->d_dav);
        ctx->d_dav = nullptr;
        ctx->dav_bytes = 0;
        if ((rc = check_hip(ctx, hipMalloc(&ctx->d_dav, dbytes), "alloc davidson"))) return rc;
        ctx->dav_bytes = dbytes;
    }
    // pinned host staging: [residual partial sums | G columns] (one
    // download), [S | theta] (one upload), C
    const size_t hbytes = (nrr + nG + nS + nC) * 8;
    if (ctx->hdav_bytes < hbytes) {
        if (ctx->h_dav) hipHostFree(ctx->h_dav);
        ctx->h_dav = nullptr;
        ctx->hdav_bytes = 0;
        if ((rc = check_hip(ctx, hipHostMalloc(reinterpret_cast<void **>(&ctx->h_dav), hbytes), "pinned davidson")))
            return rc;
        ctx->hdav_bytes = hbytes;
    }
    double *const pRR = ctx->h_dav, *const pG = pRR + nrr, *const pS = pG + nG, *const pC = pS + nS;
    Dav dv;
    dv.ctx = ctx;
    dv.n = n;
    dv.lda = lda;
    dv.mmax = mmax;
    dv.b = b;
    double *p = ctx->d_dav;
    double *d_A = p;
    dv.A = d_A; p += nA;
    dv.V = p; p += nn_ * mmax;
    dv.W = p; p += nn_ * mmax;
    dv.X = p; p += nn_ * kMaxB;
    dv.R = p; p += nn_ * kMaxB;
    dv.T = p; p += nn_ * kMaxB;
    dv.AT = p; p += nn_ * kMaxB;
    dv.S = p; p += std::max(nS, nG);
    dv.rr = p; p += nrr;
    dv.G = p; p += nG;  // directly behind rr: one download takes both
    dv.SC = p; p += nC;
    dv.diag = p; p += nn_;
    dv.part = p; p += (size_t)kMaxSplit * kMaxB * nn_;

#define DV_CHECK(expr, what)                                    \
    do {                                                        \
        if ((rc = check_hip(ctx, (expr), what))) {              \
            hipStreamSynchronize(ctx->stream);                  \
            return rc;                                          \
        }                                                       \
    } while (0)

    DV_CHECK(hipMemcpyAsync(dv.diag, dg.data(), nn_ * 8, hipMemcpyHostToDevice, ctx->stream), "diag");
    {
        hipEvent_t ev;
        timer_begin(ctx, "symfill", &ev);
        dim3 grid((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32));
        hipLaunchKernelGGL(k_sym_full, grid, dim3(256), 0, ctx->stream, Hsrc, Osrc, ldh, n, P, dv.diag,
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
        std::vector<double> Ah(nA), Gd(nn_ * nn_), w, Z;
        DV_CHECK(hipMemcpyAsync(Ah.data(), d_A, Ah.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "A");
        DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
        for (int64_t c = 0; c < n; ++c)
            std::copy(Ah.begin() + (size_t)c * lda, Ah.begin() + (size_t)c * lda + n, Gd.begin() + (size_t)c * n);
        small_eigh((int)n, Gd, k, w, Z);
        for (int j = 0; j < k; ++j) evals[j] = w[j];
        if (evecs) std::copy(Z.begin(), Z.begin() + nn_ * k, evecs);
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
    // of each new block; after a restart it is recomputed.  live[c] marks
    // the columns of V that are not zero (dropped by CholQR).
    int m = b, m0 = 0;
    DV_CHECK(dv.symm(dv.V, b, dv.W), "AV");
    DV_CHECK(dv.tn(dv.V, m, nullptr, 0, dv.W, m, dv.G, m), "VtW");
    DV_CHECK(hipMemcpyAsync(pG, dv.G, (size_t)m * m * 8, hipMemcpyDeviceToHost, ctx->stream), "G");
    std::vector<double> Gh((size_t)mmax * mmax, 0.0), theta, Sb, Sp, Gm, thetaPrev;
    std::vector<char> live((size_t)mmax, 1);
    std::vector<int> idx;
    int sp_rows = 0;  // rows of Sp (previous Ritz coefficients), 0 = none
    bool have_prev = false, converged = false, fallback_done = false;
    const int maxiter = 400;
    for (int iter = 0; iter < maxiter; ++iter) {
        DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
        // residual norms of the previous iteration's Ritz pairs
        if (have_prev) {
            double worst = 0.0;
            for (int j = 0; j < k; ++j) {
                double s2 = 0.0;
                for (int q = 0; q < nchunk; ++q) s2 += pRR[(size_t)j * nchunk + q];
                worst = std::max(worst, std::sqrt(s2));
            }
            if (worst < tol) {
                converged = true;
                break;
            }
        }
        // new columns m0..m-1 of G
        const int nc = m - m0;
        for (int c = 0; c < nc; ++c)
            for (int r = 0; r < m; ++r) {
                const double v = pG[(size_t)c * m + r];
                Gh[(size_t)(m0 + c) * mmax + r] = v;
                if (r < m0) Gh[(size_t)r * mmax + m0 + c] = v;
            }
        for (int c = m0; c < m; ++c)
            for (int r = m0; r < c; ++r) {
                const double v = 0.5 * (Gh[(size_t)c * mmax + r] + Gh[(size_t)r * mmax + c]);
                Gh[(size_t)c * mmax + r] = Gh[(size_t)r * mmax + c] = v;
            }
        // a dropped (zero) column of V has a zero column in G
        int nlive_new = 0;
        for (int c = m0; c < m; ++c) {
            live[c] = Gh[(size_t)c * mmax + c] != 0.0;
            nlive_new += live[c];
        }
        if (nlive_new == 0 && m0 > 0) {
            // the whole correction block was dropped: fall back once to the
            // previous residuals, orthonormalised against V
            if (fallback_done) {
                rc = fail(ctx, -8, "Davidson: search space collapsed");
                hipStreamSynchronize(ctx->stream);
                return rc;
            }
            fallback_done = true;
            m = m0;
            DV_CHECK(hipMemcpyAsync(dv.T, dv.R, nn_ * b * 8, hipMemcpyDeviceToDevice, ctx->stream), "T=R");
            DV_CHECK(dv.orth_cholqr(m, b, dv.V + (size_t)m * n), "R orth");
            DV_CHECK(dv.symm(dv.V + (size_t)m * n, b, dv.W + (size_t)m * n), "AR");
            DV_CHECK(dv.tn(dv.V, m + b, nullptr, 0, dv.W + (size_t)m * n, b, dv.G, m + b), "VtW");
            DV_CHECK(hipMemcpyAsync(pG, dv.G, (size_t)(m + b) * b * 8, hipMemcpyDeviceToHost, ctx->stream), "G");
            m0 = m;
            m += b;
            --iter;
            continue;
        }
        fallback_done = false;
        idx.clear();
        for (int c = 0; c < m; ++c)
            if (live[c]) idx.push_back(c);
        const int ml = (int)idx.size();
        if (ml < b) {
            rc = fail(ctx, -8, "Davidson: search space collapsed");
            hipStreamSynchronize(ctx->stream);
            return rc;
        }
        Gm.resize((size_t)ml * ml);
        for (int c = 0; c < ml; ++c)
            for (int r = 0; r < ml; ++r) Gm[(size_t)c * ml + r] = Gh[(size_t)idx[c] * mmax + idx[r]];
        std::vector<double> Sl;
        small_eigh(ml, Gm, b, theta, Sl);
        // coefficients on all m columns (zero rows for the dropped ones)
        Sb.assign((size_t)m * b, 0.0);
        for (int j = 0; j < b; ++j)
            for (int r = 0; r < ml; ++r) Sb[(size_t)j * m + idx[r]] = Sl[(size_t)j * ml + r];
        // X = V S_b, residuals R = W S_b - X theta, corrections T and the
        // partial sums of |R|^2 (downloaded with the next G columns)
        std::copy(Sb.begin(), Sb.end(), pS);
        std::copy(theta.begin(), theta.begin() + b, pS + (size_t)m * b);
        DV_CHECK(hipMemcpyAsync(dv.S, pS, ((size_t)m * b + b) * 8, hipMemcpyHostToDevice, ctx->stream), "S");
        DV_CHECK(dv.ritz(m), "ritz");
        thetaPrev = theta;
        have_prev = true;
        if (ml >= n) {
            // the subspace is the whole space: the Ritz pairs are exact
            DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
            converged = true;
            break;
        }
        if (m + b > mmax) {
            // thick restart (GD+k; Stathopoulos & Saad 1998): keep the current
            // Ritz vectors and the previous iteration's, orthonormalised in
            // the m-dimensional coefficient space, so V <- V C, W <- W C
            std::vector<double> C((size_t)m * 2 * kMaxB, 0.0);
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
            std::copy(C.begin(), C.begin() + (size_t)m * keep, pC);
            DV_CHECK(hipMemcpyAsync(dv.SC, pC, (size_t)m * keep * 8, hipMemcpyHostToDevice, ctx->stream), "C");
            DV_CHECK(dv.nn(dv.V, m, nullptr, 0, dv.SC, m, keep, dv.AT), "VC");
            DV_CHECK(hipMemcpyAsync(dv.V, dv.AT, nn_ * keep * 8, hipMemcpyDeviceToDevice, ctx->stream), "V");
            DV_CHECK(dv.nn(dv.W, m, nullptr, 0, dv.SC, m, keep, dv.AT), "WC");
            DV_CHECK(hipMemcpyAsync(dv.W, dv.AT, nn_ * keep * 8, hipMemcpyDeviceToDevice, ctx->stream), "W");
            m = keep;
            m0 = 0;
            sp_rows = 0;
            std::fill(live.begin(), live.end(), 1);
        } else {
            Sp = Sb;
            sp_rows = m;
            m0 = m;
        }
        // correction block: orthonormalised against V and within itself,
        // written into V's next b columns, then its products with A
        DV_CHECK(dv.orth_cholqr(m, b, dv.V + (size_t)m * n), "T orth");
        DV_CHECK(dv.symm(dv.V + (size_t)m * n, b, dv.W + (size_t)m * n), "AT");
        m += b;
        DV_CHECK(dv.tn(dv.V, m, nullptr, 0, dv.W + (size_t)m0 * n, m - m0, dv.G, m), "VtW");
        DV_CHECK(hipMemcpyAsync(pRR, dv.rr, (nrr + (size_t)m * (m - m0)) * 8, hipMemcpyDeviceToHost, ctx->stream),
                 "RR, G");
    }
    if (!converged) {
        rc = fail(ctx, -9, "Davidson did not converge");
        hipStreamSynchronize(ctx->stream);
        return rc;
    }
    for (int j = 0; j < k; ++j) evals[j] = thetaPrev[j];
    if (evecs) {
        std::vector<double> Xh(nn_ * k);
        DV_CHECK(hipMemcpyAsync(Xh.data(), dv.X, Xh.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "X");
        DV_CHECK(hipStreamSynchronize(ctx->stream), "sync");
        std::copy(Xh.begin(), Xh.end(), evecs);
    }
#undef DV_CHECK
    return 0;
}

extern "C" int scp3g_lowest_levels(scp3g_ctx *ctx, int k, double tol, double *evals,
                                   double *evecs)
{
    if (!ctx) return -1;
    if (ctx->nbasis <= 0 || !ctx->d_H) return fail(ctx, -2, "no Hamiltonian accumulated");
    if (ctx->points_done <= 0) return fail(ctx, -6, "no points accumulated");
    const double *Hs = nullptr;
    int rc = read_sums(ctx, &Hs);
    if (rc) return rc;
    return davidson_run(ctx, Hs, ctx->d_offset, ctx->nbasis, ctx->nbasis, ctx->norm_sum,
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
    const double *Hs = nullptr;
    int rc = read_sums(ctx, &Hs);
    if (rc) return rc;
    return davidson_run(ctx, Hs, ctx->d_offset, ctx->nbasis, nsub, ctx->norm_sum, dg, k, tol,
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
    if (!rc) rc = davidson_run(ctx, d_M, nullptr, n, n, 1.0, std::vector<double>((size_t)n, 0.0), k, tol, evals, evecs);
    hipFree(d_M);
    return rc;
}
