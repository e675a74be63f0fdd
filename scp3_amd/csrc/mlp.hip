// MLP surrogate V_nn(q) of the potential in normal coordinates, trained and
// evaluated by this library's own kernels, used as a QMC control variate.
//
// Network: x = sqrt(2) y (unit-variance inputs, padded to a multiple of 64
// with zeros) -> 3 hidden layers of width h with SiLU -> scalar.  The output
// models the anharmonic part of the potential,
//     V(y) = V0 + 1/2 sum_k omega_k y_k^2 + t_mean + t_std * f(x),
// so that the control-variate weight is w_nn = V0 + t_mean + t_std f(x)
// (SCP3.cpp:222-223 with V replaced by V_nn).
//
// Arithmetic: BF16 weights and activations on v_mfma_f32_16x16x32_bf16 with
// FP32 accumulation; FP32 master weights, gradients and Adam moments.
// GEMMs are "NT": C[m][n] = sum_k A[m][k] B[n][k] with both operands
// k-contiguous; weights are kept as W[out][in] (forward) and W^T[in][out]
// (backward data), activations and deltas as [batch][feature] and
// [feature][batch] (weight gradients).  Epilogues fuse bias + SiLU
// (forward), SiLU' (backward) and the layout copies.
//
// No reference counterpart: the reference has no surrogate (the learned part
// and the control variate are new in this build).
#include <cmath>
#include <cstring>
#include <vector>

#include "scp3g_internal.h"

namespace scp3g {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline unsigned short f2bf(float f)
{
    unsigned u;
    memcpy(&u, &f, 4);
    const unsigned bias = 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)((u + bias) >> 16);
}

__device__ inline float bf2f(unsigned short h)
{
    return __uint_as_float((unsigned)h << 16);
}

__device__ inline float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

// Sum of the 256 threads' values s, valid in thread 0: the tree
// red[t] += red[t + k], k = 128, 64, ..., 1, with the steps below 64 as
// shuffles inside the first wave (the same additions in the same order).
__device__ inline float block_sum_256(float s, float *red)
{
    const int t = threadIdx.x;
    red[t] = s;
    __syncthreads();
    if (t < 128) red[t] += red[t + 128];
    __syncthreads();
    float v = 0.f;
    if (t < 64) {
        v = red[t] + red[t + 64];
#pragma unroll
        for (int k = 32; k > 0; k >>= 1) v += __shfl_down(v, k, 64);
    }
    return v;
}

enum { EPI_FWD = 0, EPI_BWD = 1, EPI_WGRAD = 2 };

// C = A B^T on a 64 x 64 tile per workgroup (2 x 2 waves of 32 x 32, 2 x 2
// blocks of 16 x 16 x 32 each).  M, N multiples of 64, K multiple of 32.
// The weight-gradient form splits K over blockIdx.z: split z contracts
// k in [z K, (z + 1) K) and writes its partial sums at G + z gsplit.
//
// The operands pass through LDS in K-panels of up to kPanel: every thread
// loads its share of the panel's 64 rows of A and of B into registers (all
// loads in flight at once), writes them to LDS, and the waves then run the
// panel's MFMAs from LDS while the next panel's loads are in flight.  These
// GEMMs are small (a few MFLOP per workgroup), so one memory latency per
// panel instead of one per 32 k bounds them.  LDS rows are 32 chunks of
// 16 bytes (8 k); chunk c of row r sits at c ^ (r & 15), which spreads the
// 16 rows of an MFMA operand read over all banks.  Each wave accumulates k
// in ascending order as before.
constexpr int kPanel = 256;
constexpr int kPanelChunks = kPanel / 8;
template <int EPI>
__global__ void __launch_bounds__(256)
    k_gemm_nt(const unsigned short *__restrict__ A, int64_t lda, const unsigned short *__restrict__ B,
              int64_t ldb, int K, const float *__restrict__ bias, float *__restrict__ Z,
              unsigned short *__restrict__ Out, unsigned short *__restrict__ OutT, int64_t ldo,
              int64_t ldoT, const float *__restrict__ Zprev, float *__restrict__ G, int64_t ldg,
              int64_t gsplit)
{
    __shared__ uint4 sA[64 * kPanelChunks];
    __shared__ uint4 sB[64 * kPanelChunks];
    if (EPI == EPI_WGRAD) {
        A += (size_t)blockIdx.z * K;
        B += (size_t)blockIdx.z * K;
        G += (size_t)blockIdx.z * gsplit;
    }
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int64_t mb = (int64_t)blockIdx.y * 64, nb = (int64_t)blockIdx.x * 64;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int64_t m0 = mb + wr, n0 = nb + wc;
    f4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};

    // panel staging: position q = tid + 256 j of an operand's 64 x cpr chunks
    // (row q / cpr, chunk q % cpr), j < cpr / 4 <= 8, in named registers (an
    // array of them is kept in scratch memory by the compiler)
    uint4 ra0, ra1, ra2, ra3, ra4, ra5, ra6, ra7, rb0, rb1, rb2, rb3, rb4, rb5, rb6, rb7;
    // a slot beyond the panel re-reads the thread's first chunk (same lines),
    // so that every slot is assigned
#define GEMM_LD1(j, k0_, cpr_)                                                               \
    do {                                                                                     \
        const int q = tid + 256 * ((j) < (cpr_) / 4 ? (j) : 0), r = q / (cpr_), c = q - r * (cpr_); \
        ra##j = *reinterpret_cast<const uint4 *>(A + (size_t)(mb + r) * lda + (k0_) + 8 * c); \
        rb##j = *reinterpret_cast<const uint4 *>(B + (size_t)(nb + r) * ldb + (k0_) + 8 * c); \
    } while (0)
#define GEMM_LOAD(k0_)                                                                       \
    do {                                                                                     \
        const int cpr_ = min(kPanel, K - (k0_)) / 8;                                         \
        GEMM_LD1(0, k0_, cpr_); GEMM_LD1(1, k0_, cpr_); GEMM_LD1(2, k0_, cpr_);              \
        GEMM_LD1(3, k0_, cpr_); GEMM_LD1(4, k0_, cpr_); GEMM_LD1(5, k0_, cpr_);              \
        GEMM_LD1(6, k0_, cpr_); GEMM_LD1(7, k0_, cpr_);                                      \
    } while (0)
#define GEMM_ST1(j)                                                                          \
    do {                                                                                     \
        if ((j) < cpr / 4) {                                                                 \
            const int q = tid + 256 * (j), r = q / cpr, c = q - r * cpr;                     \
            const int pos = r * kPanelChunks + (c ^ (r & 15));                               \
            sA[pos] = ra##j;                                                                 \
            sB[pos] = rb##j;                                                                 \
        }                                                                                    \
    } while (0)
    GEMM_LOAD(0);
    for (int k0 = 0; k0 < K; k0 += kPanel) {
        const int cpr = min(kPanel, K - k0) / 8;
        if (k0 > 0) __syncthreads();  // every wave is done with the previous panel
        GEMM_ST1(0); GEMM_ST1(1); GEMM_ST1(2); GEMM_ST1(3);
        GEMM_ST1(4); GEMM_ST1(5); GEMM_ST1(6); GEMM_ST1(7);
        __syncthreads();
        if (k0 + kPanel < K) GEMM_LOAD(k0 + kPanel);
        // k-steps in pairs: both steps' fragment reads are issued before the
        // first step's MFMAs (each accumulator still takes k in order)
        const int nks = cpr / 4;
#define GEMM_FRAGS(ks_, av_, bv_)                                                            \
    do {                                                                                     \
        _Pragma("unroll") for (int a = 0; a < 2; ++a) {                                      \
            const int r = wr + 16 * a + li;                                                  \
            av_[a] = __builtin_bit_cast(bf16x8, sA[r * kPanelChunks + ((4 * (ks_) + lg) ^ (r & 15))]); \
        }                                                                                    \
        _Pragma("unroll") for (int b = 0; b < 2; ++b) {                                      \
            const int r = wc + 16 * b + li;                                                  \
            bv_[b] = __builtin_bit_cast(bf16x8, sB[r * kPanelChunks + ((4 * (ks_) + lg) ^ (r & 15))]); \
        }                                                                                    \
    } while (0)
#define GEMM_MFMAS(av_, bv_)                                                                 \
    do {                                                                                     \
        _Pragma("unroll") for (int a = 0; a < 2; ++a)                                        \
            _Pragma("unroll") for (int b = 0; b < 2; ++b)                                    \
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av_[a], bv_[b], acc[a][b], 0, 0, 0); \
    } while (0)
        int ks = 0;
        for (; ks + 1 < nks; ks += 2) {
            bf16x8 av0[2], bv0[2], av1[2], bv1[2];
            GEMM_FRAGS(ks, av0, bv0);
            GEMM_FRAGS(ks + 1, av1, bv1);
            GEMM_MFMAS(av0, bv0);
            GEMM_MFMAS(av1, bv1);
        }
        if (ks < nks) {
            bf16x8 av0[2], bv0[2];
            GEMM_FRAGS(ks, av0, bv0);
            GEMM_MFMAS(av0, bv0);
        }
#undef GEMM_FRAGS
#undef GEMM_MFMAS
    }
#undef GEMM_LD1
#undef GEMM_LOAD
#undef GEMM_ST1
    // D[i][j]: lane j + 16*(i/4), element i%4; the lane's 4 elements are 4
    // consecutive m of one n, so each transposed copy gets one 8-byte store
    const int oj = lane & 15, oi = 4 * (lane >> 4);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int64_t mr = m0 + 16 * a + oi, n = n0 + 16 * b + oj;
            unsigned hv[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t m = mr + e;
                float v = acc[a][b][e];
                if (EPI == EPI_FWD) {
                    v += bias[n];
                    Z[(size_t)m * ldo + n] = v;
                    const unsigned short h = f2bf(v * sigm(v));
                    Out[(size_t)m * ldo + n] = h;
                    hv[e] = h;
                } else if (EPI == EPI_BWD) {
                    const float z = Zprev[(size_t)m * ldo + n];
                    const float s = sigm(z);
                    const unsigned short h = f2bf(v * s * (1.0f + z * (1.0f - s)));
                    Out[(size_t)m * ldo + n] = h;
                    hv[e] = h;
                } else {
                    G[(size_t)m * ldg + n] = v;
                }
            }
            if (EPI != EPI_WGRAD && OutT)
                *reinterpret_cast<uint2 *>(OutT + (size_t)n * ldoT + mr) =
                    make_uint2(hv[0] | (hv[1] << 16), hv[2] | (hv[3] << 16));
        }
}

// x[p][k] = bf16(sqrt(2) y[k][p]) for k < M, zero up to dpad; optional x^T.
__global__ void k_mlp_input(const double *__restrict__ y, int64_t ldy, int M, int dpad,
                            int64_t npts, unsigned short *__restrict__ x,
                            unsigned short *__restrict__ xT, int64_t ldxT)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    for (int k = 0; k < dpad; ++k) {
        const float v = k < M ? (float)(1.4142135623730951 * y[(size_t)k * ldy + p]) : 0.0f;
        const unsigned short h = f2bf(v);
        x[(size_t)p * dpad + k] = h;
        if (xT) xT[(size_t)k * ldxT + p] = h;
    }
}

// out[p] = a3[p] . w3 + b3.  A workgroup takes kHeadPts consecutive points
// (npts is a multiple of kHeadPts), each wave a quarter of them together:
// lane-strided partial sums over j (per point in ascending j), then a
// butterfly.  Training mode also forms the loss terms and the head's
// backward, dout = 2 (out - t) / B and delta3 = dout w3 silu'(z3), written
// row-major [p][j] and transposed [j][p]: thread j packs its points into
// 16-byte stores.
constexpr int kHeadPts = 8;
constexpr int kHeadWavePts = kHeadPts / 4;
__global__ void __launch_bounds__(256)
    k_mlp_head(const unsigned short *__restrict__ a3, const float *__restrict__ z3, int h,
               const float *__restrict__ w3, const float *__restrict__ b3, int64_t npts,
               const float *__restrict__ target, float inv_batch, float *__restrict__ out,
               float *__restrict__ sqerr, float *__restrict__ dout,
               unsigned short *__restrict__ d3, unsigned short *__restrict__ d3T, int64_t ldT)
{
    __shared__ float gsh[kHeadPts];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t p0 = (int64_t)blockIdx.x * kHeadPts;
    const int64_t pw = p0 + wave * kHeadWavePts;
    float sv[kHeadWavePts];
#pragma unroll
    for (int i = 0; i < kHeadWavePts; ++i) sv[i] = 0.f;
    for (int j = lane; j < h; j += 64) {
        const float wj = w3[j];
#pragma unroll
        for (int i = 0; i < kHeadWavePts; ++i) sv[i] += bf2f(a3[(size_t)(pw + i) * h + j]) * wj;
    }
#pragma unroll
    for (int i = 0; i < kHeadWavePts; ++i) {
        float s = sv[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        const float o = s + b3[0];
        const int64_t p = pw + i;
        if (lane == 0) {
            out[p] = o;
            if (target) {
                const float r = o - target[p];
                const float g = 2.0f * r * inv_batch;
                sqerr[p] = r * r;
                dout[p] = g;
                gsh[wave * kHeadWavePts + i] = g;
            }
        }
    }
    if (!target) return;
    __syncthreads();
    for (int j = tid; j < h; j += 256) {
        const float wj = w3[j];
        unsigned pk[kHeadPts / 2];
#pragma unroll
        for (int pl = 0; pl < kHeadPts; ++pl) {
            const int64_t p = p0 + pl;
            const float z = z3[(size_t)p * h + j];
            const float sg = sigm(z);
            const unsigned short v = f2bf(gsh[pl] * wj * sg * (1.0f + z * (1.0f - sg)));
            d3[(size_t)p * h + j] = v;
            if (pl & 1) pk[pl >> 1] |= (unsigned)v << 16;
            else pk[pl >> 1] = v;
        }
        uint4 *dst = reinterpret_cast<uint4 *>(d3T + (size_t)j * ldT + p0);
#pragma unroll
        for (int c = 0; c < kHeadPts / 8; ++c)
            dst[c] = make_uint4(pk[4 * c], pk[4 * c + 1], pk[4 * c + 2], pk[4 * c + 3]);
    }
}

// The gradient sums of one step in one launch, each as 256 strided partial
// sums (unrolled so that the loads are in flight together; the adds keep
// their order) and a tree: blocks [0, 3h) the hidden layers' bias gradients
// (row r of layer l's delta^T, block l h + r), blocks [3h, 4h) the head's
// weight gradient sum_p dout[p] a3[p][j], block 4h its bias gradient
// sum_p dout[p], block 4h + 1 the batch's summed squared error (loss).
struct GradSums {
    const unsigned short *dT[3];  // delta^T of the hidden layers, [h][ld]
    float *gb[3];                 // their bias gradients
    const unsigned short *a3T;    // last activations^T, [h][ld]
    const float *dout, *sqerr;
    float *gw3, *gb3, *loss;
};
__global__ void __launch_bounds__(256)
    k_grad_sums(GradSums gs, int h, int64_t ld, int64_t batch)
{
    __shared__ float red[256];
    const int b = blockIdx.x;
    float s = 0.f;
    if (b < 3 * h) {
        const int l = b / h, r = b - l * h;
        const unsigned short *X = gs.dT[l] + (size_t)r * ld;
#pragma unroll 16
        for (int64_t c = threadIdx.x; c < batch; c += 256) s += bf2f(X[c]);
    } else if (b < 4 * h) {
        const unsigned short *X = gs.a3T + (size_t)(b - 3 * h) * ld;
#pragma unroll 16
        for (int64_t p = threadIdx.x; p < batch; p += 256) s += gs.dout[p] * bf2f(X[p]);
    } else if (b == 4 * h) {
#pragma unroll 16
        for (int64_t p = threadIdx.x; p < batch; p += 256) s += gs.dout[p] * 1.0f;
    } else {
#pragma unroll 16
        for (int64_t p = threadIdx.x; p < batch; p += 256) s += gs.sqerr[p];
    }
    s = block_sum_256(s, red);
    if (threadIdx.x == 0) {
        if (b < 3 * h) gs.gb[b / h][b % h] = s;
        else if (b < 4 * h) gs.gw3[b - 3 * h] = s;
        else if (b == 4 * h) gs.gb3[0] = s;
        else gs.loss[0] = s;
    }
}

// One training update in one launch: the split-K partial gradients summed in
// split order (as k_grad_reduce), the Adam step (Kingma & Ba 2015) on the
// FP32 master parameters, and the BF16 copies W[o][i], W^T[i][o] of an
// updated hidden-layer weight (as k_to_bf16).  One parameter per thread: a
// 2-D tiling that writes W^T in longer runs leaves too few workgroups (77
// for the hexamer network) and was 5x slower.
struct WeightCopies {
    int64_t w_off[3];
    int in[3];
    int out;
    unsigned short *Wb[3], *WbT[3];
};
__global__ void __launch_bounds__(256)
    k_adam_update(float *__restrict__ P, const float *__restrict__ part, int nsplit, float *__restrict__ G,
                  float *__restrict__ m, float *__restrict__ v, int64_t n, float lr, float b1, float b2,
                  float eps, float c1, float c2, WeightCopies wc)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = part[i];
#pragma unroll 16
    for (int sp = 1; sp < nsplit; ++sp) g += part[(size_t)sp * n + i];
    G[i] = g;
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float p = P[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    P[i] = p;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        const int64_t e = i - wc.w_off[l];
        if (e >= 0 && e < (int64_t)wc.out * wc.in[l]) {
            const int r = (int)(e / wc.in[l]), c = (int)(e - (int64_t)r * wc.in[l]);
            const unsigned short h = f2bf(p);
            wc.Wb[l][e] = h;
            wc.WbT[l][(size_t)c * wc.out + r] = h;
        }
    }
}

// G[i] = sum over splits s of part[s][i], in split order.
__global__ void k_grad_reduce(const float *__restrict__ part, int nsplit, int64_t n, float *__restrict__ G)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = part[i];
    for (int sp = 1; sp < nsplit; ++sp) g += part[(size_t)sp * n + i];
    G[i] = g;
}

// bf16 copies W[o][i] and W^T[i][o] of an FP32 [o][i] matrix.
__global__ void k_to_bf16(const float *__restrict__ W, int rows, int cols,
                          unsigned short *__restrict__ Wb, unsigned short *__restrict__ WbT)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)rows * cols) return;
    const int r = (int)(e / cols), c = (int)(e - (int64_t)r * cols);
    const unsigned short h = f2bf(W[e]);
    Wb[e] = h;
    WbT[(size_t)c * rows + r] = h;
}

// mode 0: the surrogate weight w_nn = V0 + t_mean + t_std * f; mode 1: the
// correction weight w - w_nn = V - V_nn from the energies E (kcal/mol,
// SCP3.cpp:222-223); mode 2: the surrogate potential V_nn = w_nn + 1/2
// sum_k omega_k y_k^2 (hartree).  Every mode multiplies by `scale`.
__global__ void k_mlp_weights(const float *__restrict__ f, const double *__restrict__ y,
                              int64_t ldy, const double *__restrict__ omega, int M, double base,
                              double tstd, int64_t npts, int64_t npad, const double *__restrict__ E,
                              int mode, double scale, double *__restrict__ w)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    if (i >= npts) {
        w[i] = 0.0;
        return;
    }
    const double wnn = base + tstd * (double)f[i];
    if (mode == 0) {
        w[i] = scale * wnn;
        return;
    }
    double hsum = 0.0;
    for (int k = 0; k < M; ++k) {
        const double yk = y[(size_t)k * ldy + i];
        hsum += yk * (omega[k] * yk);
    }
    if (mode == 2) {
        w[i] = scale * (wnn + 0.5 * hsum);
        return;
    }
    const double bV = E[i] / kAutokcalpmol - 0.5 * hsum;
    w[i] = scale * (bV - wnn);
}

// training targets t = (V - V0 - 1/2 sum omega y^2 - mean) / std, FP32
__global__ void k_mlp_targets(const double *__restrict__ E, const double *__restrict__ y,
                              int64_t ldy, const double *__restrict__ omega, int M, double V0,
                              int64_t npts, double *__restrict__ raw)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    double hsum = 0.0;
    for (int k = 0; k < M; ++k) {
        const double yk = y[(size_t)k * ldy + i];
        hsum += yk * (omega[k] * yk);
    }
    raw[i] = E[i] / kAutokcalpmol - 0.5 * hsum - V0;
}

// ---------------------------------------------------------------------------
// host side

struct Layer {
    int in, out;            // out x in weights (in padded to a multiple of 64)
    int64_t w_off, b_off;   // offsets in the FP32 parameter vector
    unsigned short *Wb = nullptr, *WbT = nullptr;
};

}  // namespace scp3g

struct MlpState {
    int M = 0, dpad = 0, h = 0;
    scp3g::Layer L[3];
    int64_t w3_off = 0, b3_off = 0, nparam = 0;
    float *P = nullptr, *G = nullptr, *m = nullptr, *v = nullptr;  // device
    // split-K partial gradients [kMaxGradSplit][nparam]; entries that only
    // split 0 writes (biases, head) stay zero in the other splits
    float *Gpart = nullptr;
    int gsplit = 1;           // splits summed by k_grad_reduce
    int lsplit[3] = {1, 1, 1};  // splits of each layer's last weight gradient
    int64_t step = 0;
    double V0 = 0.0, t_mean = 0.0, t_std = 1.0;
    // work buffers for a batch of `cap` points
    int64_t cap = 0;
    unsigned short *x = nullptr, *xT = nullptr, *act[3] = {}, *actT[3] = {}, *dl[3] = {},
                   *dlT[3] = {};
    float *z[3] = {}, *out = nullptr, *sq = nullptr, *dout = nullptr, *lossd = nullptr;
    // training set
    int64_t ntrain = 0;
    unsigned short *Xtr = nullptr, *XtrT = nullptr;
    float *Ttr = nullptr;
};

namespace scp3g {

static void mfree(void *&p)
{
    if (p) hipFree(p);
    p = nullptr;
}

template <typename T>
static void mfree(T *&p)
{
    void *q = p;
    mfree(q);
    p = nullptr;
}

constexpr int kMaxGradSplit = 16;

static void free_batch(MlpState *s)
{
    mfree(s->x);
    mfree(s->xT);
    for (int l = 0; l < 3; ++l) {
        mfree(s->act[l]);
        mfree(s->actT[l]);
        mfree(s->dl[l]);
        mfree(s->dlT[l]);
        mfree(s->z[l]);
    }
    mfree(s->out);
    mfree(s->sq);
    mfree(s->dout);
    mfree(s->lossd);
    s->cap = 0;
}

void mlp_release(scp3g_ctx *ctx)
{
    MlpState *s = ctx->mlp;
    if (!s) return;
    free_batch(s);
    for (int l = 0; l < 3; ++l) {
        mfree(s->L[l].Wb);
        mfree(s->L[l].WbT);
    }
    mfree(s->P);
    mfree(s->G);
    mfree(s->Gpart);
    mfree(s->m);
    mfree(s->v);
    mfree(s->Xtr);
    mfree(s->XtrT);
    mfree(s->Ttr);
    delete s;
    ctx->mlp = nullptr;
}

static hipError_t ensure_batch(MlpState *s, int64_t cap)
{
    if (s->cap >= cap) return hipSuccess;
    free_batch(s);
    const size_t B = (size_t)cap, h = (size_t)s->h;
    hipError_t e = hipSuccess;
    auto A = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    A((void **)&s->x, B * s->dpad * 2);
    A((void **)&s->xT, B * s->dpad * 2);
    for (int l = 0; l < 3; ++l) {
        A((void **)&s->act[l], B * h * 2);
        A((void **)&s->actT[l], B * h * 2);
        A((void **)&s->dl[l], B * h * 2);
        A((void **)&s->dlT[l], B * h * 2);
        A((void **)&s->z[l], B * h * 4);
    }
    A((void **)&s->out, B * 4);
    A((void **)&s->sq, B * 4);
    A((void **)&s->dout, B * 4);
    A((void **)&s->lossd, 64 * 1024 * 4);
    if (e == hipSuccess) s->cap = cap;
    return e;
}

static hipError_t refresh_bf16(scp3g_ctx *ctx, MlpState *s)
{
    for (int l = 0; l < 3; ++l) {
        const int64_t ne = (int64_t)s->L[l].out * s->L[l].in;
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, ctx->stream,
                           s->P + s->L[l].w_off, s->L[l].out, s->L[l].in, s->L[l].Wb, s->L[l].WbT);
    }
    return hipGetLastError();
}

// forward over `np` points (np a multiple of 64) whose inputs are in x (or
// the given pointers); leaves z, act, out.
static hipError_t forward(scp3g_ctx *ctx, MlpState *s, const unsigned short *x, int64_t np,
                          bool keepT, const unsigned short *xT_unused)
{
    (void)xT_unused;
    const unsigned short *in = x;
    int64_t ldin = s->dpad;
    for (int l = 0; l < 3; ++l) {
        const Layer &Ly = s->L[l];
        dim3 grid((unsigned)(Ly.out / 64), (unsigned)(np / 64));
        hipLaunchKernelGGL(k_gemm_nt<EPI_FWD>, grid, dim3(256), 0, ctx->stream, in, ldin, Ly.Wb,
                           (int64_t)Ly.in, Ly.in, s->P + Ly.b_off, s->z[l], s->act[l],
                           keepT ? s->actT[l] : (unsigned short *)nullptr, (int64_t)s->h,
                           (int64_t)s->cap, (const float *)nullptr, (float *)nullptr, (int64_t)0, (int64_t)0);
        in = s->act[l];
        ldin = s->h;
    }
    return hipGetLastError();
}

static int mlp_ready(scp3g_ctx *ctx)
{
    if (!ctx) return -1;
    if (!ctx->mlp) return fail(ctx, -2, "no surrogate (scp3g_mlp_init)");
    return 0;
}

// inputs of np points from mode-major y
static hipError_t inputs(scp3g_ctx *ctx, MlpState *s, const double *y, int64_t ldy, int64_t np,
                         unsigned short *x, unsigned short *xT, int64_t ldxT)
{
    hipLaunchKernelGGL(k_mlp_input, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream, y,
                       ldy, s->M, s->dpad, np, x, xT, ldxT);
    return hipGetLastError();
}

// f(x) for np points of y (normalised network output) into s->out
static hipError_t eval_f(scp3g_ctx *ctx, MlpState *s, const double *y, int64_t ldy, int64_t np)
{
    const int64_t npad = (np + 63) / 64 * 64;
    hipError_t e = ensure_batch(s, npad);
    if (e != hipSuccess) return e;
    if (npad > np)  // padded inputs are finite (zero rows)
        hipMemsetAsync(s->x + (size_t)np * s->dpad, 0, (size_t)(npad - np) * s->dpad * 2, ctx->stream);
    if ((e = inputs(ctx, s, y, ldy, np, s->x, nullptr, 0)) != hipSuccess) return e;
    if ((e = forward(ctx, s, s->x, npad, false, nullptr)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_head, dim3((unsigned)(npad / kHeadPts)), dim3(256), 0, ctx->stream,
                       s->act[2], s->z[2], s->h, s->P + s->w3_off, s->P + s->b3_off, npad,
                       (const float *)nullptr, 0.f, s->out, (float *)nullptr, (float *)nullptr,
                       (unsigned short *)nullptr, (unsigned short *)nullptr, (int64_t)0);
    return hipGetLastError();
}

hipError_t launch_mlp_weights(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts,
                              const double *E, int mode, double scale, double *w)
{
    MlpState *s = ctx->mlp;
    if (!s) return hipErrorInvalidValue;
    hipError_t e = eval_f(ctx, s, y, ldy, npts);
    if (e != hipSuccess) return e;
    const int64_t npad = (npts + 255) / 256 * 256;
    hipLaunchKernelGGL(k_mlp_weights, dim3((unsigned)(npad / 256)), dim3(256), 0, ctx->stream,
                       s->out, y, ldy, ctx->d_omega, ctx->nmodes0, s->V0 + s->t_mean, s->t_std,
                       npts, npad, E, mode, scale, w);
    return hipGetLastError();
}

// one forward + backward on batch [b0, b0 + B) of the training set; leaves
// the split-K partial gradients in s->Gpart (and, with reduce, their sum in
// s->G) and the batch's summed squared error in *lossd.
static hipError_t grad_step(scp3g_ctx *ctx, MlpState *s, int64_t b0, int64_t B, float *lossd, bool reduce)
{
    hipError_t e;
    const unsigned short *x = s->Xtr + (size_t)b0 * s->dpad;
    const unsigned short *xT = s->XtrT + b0;  // ld = ntrain
    if ((e = forward(ctx, s, x, B, true, nullptr)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_head, dim3((unsigned)(B / kHeadPts)), dim3(256), 0, ctx->stream,
                       s->act[2], s->z[2], s->h, s->P + s->w3_off, s->P + s->b3_off, B,
                       s->Ttr + b0, 1.0f / (float)B, s->out, s->sq, s->dout, s->dl[2], s->dlT[2],
                       (int64_t)s->cap);
    for (int l = 2; l >= 0; --l) {
        const Layer &Ly = s->L[l];
        // weight gradient G[o][i] = sum_p delta[p][o] a_in[p][i], the batch
        // split so that the grid has >= 256 workgroups
        const unsigned short *inT = l > 0 ? s->actT[l - 1] : xT;
        const int64_t ldinT = l > 0 ? s->cap : s->ntrain;
        const int ntile = (Ly.in / 64) * (Ly.out / 64);
        int nsplit = 1;
        while (nsplit < kMaxGradSplit && ntile * nsplit < 256 && (B / (2 * nsplit)) % 32 == 0) nsplit *= 2;
        hipLaunchKernelGGL(k_gemm_nt<EPI_WGRAD>,
                           dim3((unsigned)(Ly.in / 64), (unsigned)(Ly.out / 64), (unsigned)nsplit),
                           dim3(256), 0, ctx->stream, s->dlT[l], (int64_t)s->cap, inT, ldinT, (int)(B / nsplit),
                           (const float *)nullptr, (float *)nullptr, (unsigned short *)nullptr,
                           (unsigned short *)nullptr, (int64_t)0, (int64_t)0, (const float *)nullptr,
                           s->Gpart + Ly.w_off, (int64_t)Ly.in, s->nparam);
        if (nsplit < s->lsplit[l]) {
            // a smaller batch than before: clear this layer's stale partials
            for (int sp = nsplit; sp < s->lsplit[l]; ++sp)
                hipMemsetAsync(s->Gpart + (size_t)sp * s->nparam + Ly.w_off, 0, (size_t)Ly.out * Ly.in * 4,
                               ctx->stream);
        }
        s->lsplit[l] = nsplit;
        if (nsplit > s->gsplit) s->gsplit = nsplit;
        if (l > 0) {
            // delta_{l-1}[p][i] = sum_o delta_l[p][o] W[o][i] * silu'(z_{l-1}[p][i])
            hipLaunchKernelGGL(k_gemm_nt<EPI_BWD>, dim3((unsigned)(Ly.in / 64), (unsigned)(B / 64)),
                               dim3(256), 0, ctx->stream, s->dl[l], (int64_t)s->h, Ly.WbT,
                               (int64_t)Ly.out, Ly.out, (const float *)nullptr, (float *)nullptr,
                               s->dl[l - 1], s->dlT[l - 1], (int64_t)s->h, (int64_t)s->cap,
                               s->z[l - 1], (float *)nullptr, (int64_t)0, (int64_t)0);
        }
    }
    // bias gradients (row sums of each layer's delta^T), the head's gradient
    // and the loss
    GradSums gs;
    for (int l = 0; l < 3; ++l) {
        gs.dT[l] = s->dlT[l];
        gs.gb[l] = s->Gpart + s->L[l].b_off;
    }
    gs.a3T = s->actT[2];
    gs.dout = s->dout;
    gs.sqerr = s->sq;
    gs.gw3 = s->Gpart + s->w3_off;
    gs.gb3 = s->Gpart + s->b3_off;
    gs.loss = lossd;
    hipLaunchKernelGGL(k_grad_sums, dim3((unsigned)(4 * s->h + 2)), dim3(256), 0, ctx->stream, gs, s->h,
                       (int64_t)s->cap, B);
    if (reduce)
        hipLaunchKernelGGL(k_grad_reduce, dim3((unsigned)((s->nparam + 255) / 256)), dim3(256), 0,
                           ctx->stream, s->Gpart, s->gsplit, s->nparam, s->G);
    return hipGetLastError();
}

static uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace scp3g

using namespace scp3g;

// Training set from device inputs y (M x npts, mode-major, ld npts) and host
// targets t: BF16 inputs in both layouts, targets on the device.
static int load_training_set(scp3g_ctx *ctx, MlpState *s, const double *d_y, const float *t_host, int64_t npts)
{
    mfree(s->Xtr);
    mfree(s->XtrT);
    mfree(s->Ttr);
    s->ntrain = 0;
    hipError_t e = hipMalloc((void **)&s->Xtr, (size_t)npts * s->dpad * 2);
    if (e == hipSuccess) e = hipMalloc((void **)&s->XtrT, (size_t)npts * s->dpad * 2);
    if (e == hipSuccess) e = hipMalloc((void **)&s->Ttr, (size_t)npts * 4);
    int rc = check_hip(ctx, e, "alloc training set");
    if (rc) return rc;
    e = hipMemcpyAsync(s->Ttr, t_host, (size_t)npts * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = inputs(ctx, s, d_y, npts, npts, s->Xtr, s->XtrT, npts);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if ((rc = check_hip(ctx, e, "inputs"))) return rc;
    s->ntrain = npts;
    return 0;
}

extern "C" {

int scp3g_mlp_init(scp3g_ctx *ctx, int hidden, uint64_t seed)
{
    if (!ctx) return -1;
    if (ctx->nmodes0 <= 0) return fail(ctx, -2, "system not set (scp3g_set_system)");
    if (hidden < 64 || hidden % 64) return fail(ctx, -4, "hidden width must be a positive multiple of 64");
    hipSetDevice(ctx->device);
    mlp_release(ctx);
    MlpState *s = new MlpState();
    ctx->mlp = s;
    s->M = ctx->nmodes0;
    s->dpad = (s->M + 63) / 64 * 64;
    s->h = hidden;
    int64_t off = 0;
    const int ins[3] = {s->dpad, hidden, hidden};
    for (int l = 0; l < 3; ++l) {
        s->L[l].in = ins[l];
        s->L[l].out = hidden;
        s->L[l].w_off = off;
        off += (int64_t)hidden * ins[l];
        s->L[l].b_off = off;
        off += hidden;
    }
    s->w3_off = off;
    off += hidden;
    s->b3_off = off;
    off += 1;
    s->nparam = off;
    // Glorot-uniform weights from a splitmix64 stream; zero biases; inputs
    // beyond the M normal coordinates get zero weights
    std::vector<float> P((size_t)off, 0.0f);
    uint64_t st = seed;
    auto uni = [&]() { return (double)(splitmix64(st) >> 11) * 0x1p-53; };
    for (int l = 0; l < 3; ++l) {
        const int fan_in = l == 0 ? s->M : hidden;
        const double a = std::sqrt(6.0 / (fan_in + hidden));
        for (int o = 0; o < hidden; ++o)
            for (int i = 0; i < ins[l]; ++i)
                P[(size_t)s->L[l].w_off + (size_t)o * ins[l] + i] =
                    (l == 0 && i >= s->M) ? 0.0f : (float)((2.0 * uni() - 1.0) * a);
    }
    const double a3 = std::sqrt(6.0 / (hidden + 1));
    for (int j = 0; j < hidden; ++j) P[(size_t)s->w3_off + j] = (float)((2.0 * uni() - 1.0) * a3);
    int rc;
    hipError_t e = hipSuccess;
    for (float **p : {&s->P, &s->G, &s->m, &s->v})
        if (e == hipSuccess) e = hipMalloc((void **)p, (size_t)off * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&s->Gpart, (size_t)kMaxGradSplit * off * 4);
    for (int l = 0; l < 3; ++l) {
        if (e == hipSuccess) e = hipMalloc((void **)&s->L[l].Wb, (size_t)hidden * ins[l] * 2);
        if (e == hipSuccess) e = hipMalloc((void **)&s->L[l].WbT, (size_t)hidden * ins[l] * 2);
    }
    if ((rc = check_hip(ctx, e, "alloc mlp"))) return rc;
    hipMemcpy(s->P, P.data(), P.size() * 4, hipMemcpyHostToDevice);
    hipMemset(s->G, 0, P.size() * 4);
    hipMemset(s->Gpart, 0, (size_t)kMaxGradSplit * P.size() * 4);
    hipMemset(s->m, 0, P.size() * 4);
    hipMemset(s->v, 0, P.size() * 4);
    // V0: the potential at the reference geometry, hartree (FP64 CPU)
    std::vector<double> xa(ctx->x0.size());
    for (size_t i = 0; i < xa.size(); ++i) xa[i] = kBohr * ctx->x0[i];
    s->V0 = (ctx->pes == SCP3G_PES_TTM3F ? ttm3f_energy_host(ctx->nw, xa.data(), nullptr)
                                         : qtip4pf_energy_host(ctx->nw, xa.data())) /
            kAutokcalpmol;
    return check_hip(ctx, refresh_bf16(ctx, s), "mlp bf16");
}

int scp3g_mlp_param_count(const scp3g_ctx *ctx, int64_t *n)
{
    if (!ctx || !n) return -1;
    *n = ctx->mlp ? ctx->mlp->nparam : 0;
    return 0;
}

int scp3g_mlp_get_params(scp3g_ctx *ctx, float *P_host, double *norm_host)
{
    int rc = mlp_ready(ctx);
    if (rc) return rc;
    MlpState *s = ctx->mlp;
    hipStreamSynchronize(ctx->stream);
    if (P_host) hipMemcpy(P_host, s->P, (size_t)s->nparam * 4, hipMemcpyDeviceToHost);
    if (norm_host) {
        norm_host[0] = s->V0;
        norm_host[1] = s->t_mean;
        norm_host[2] = s->t_std;
    }
    return 0;
}

int scp3g_mlp_set_params(scp3g_ctx *ctx, const float *P_host, const double *norm_host)
{
    int rc = mlp_ready(ctx);
    if (rc) return rc;
    MlpState *s = ctx->mlp;
    if (P_host) hipMemcpy(s->P, P_host, (size_t)s->nparam * 4, hipMemcpyHostToDevice);
    if (norm_host) {
        s->V0 = norm_host[0];
        s->t_mean = norm_host[1];
        s->t_std = norm_host[2];
    }
    return check_hip(ctx, refresh_bf16(ctx, s), "mlp bf16");
}

// Load a training set from host arrays: y_host[npts x M] point-major normal
// coordinates (already divided by sqrt(2)), t_host[npts] normalised targets.
int scp3g_mlp_load_data(scp3g_ctx *ctx, const double *y_host, const float *t_host, int64_t npts)
{
    int rc = mlp_ready(ctx);
    if (rc) return rc;
    MlpState *s = ctx->mlp;
    if (npts <= 0 || npts % 64) return fail(ctx, -4, "training points must be a positive multiple of 64");
    double *d_y = nullptr;
    if ((rc = check_hip(ctx, hipMalloc(&d_y, (size_t)npts * s->M * 8), "alloc y"))) return rc;
    std::vector<double> ym((size_t)npts * s->M);
    for (int64_t p = 0; p < npts; ++p)
        for (int k = 0; k < s->M; ++k) ym[(size_t)k * npts + p] = y_host[(size_t)p * s->M + k];
    rc = check_hip(ctx, hipMemcpy(d_y, ym.data(), ym.size() * 8, hipMemcpyHostToDevice), "upload y");
    if (!rc) rc = load_training_set(ctx, s, d_y, t_host, npts);
    hipFree(d_y);
    return rc;
}

// Sample a training set with the sampler itself: Sobol seeds skip..skip+npts-1,
// the context's potential in FP64, targets normalised to zero mean and unit
// variance (their mean and scale are kept for evaluation).
int scp3g_mlp_sample_data(scp3g_ctx *ctx, int64_t skip, int64_t npts)
{
    int rc = mlp_ready(ctx);
    if (rc) return rc;
    MlpState *s = ctx->mlp;
    if (npts <= 0 || npts % 64) return fail(ctx, -4, "training points must be a positive multiple of 64");
    hipSetDevice(ctx->device);
    // the whole set stays on the device: y (mode-major, ld npts) in tiles of
    // T points through Sobol -> map -> potential (FP64) -> targets; only the
    // targets visit the host, for their mean and scale in FP64
    const int M = s->M;
    const int64_t T = std::min<int64_t>(npts, 65536);
    double *d_y = nullptr, *d_crd = nullptr, *d_E = nullptr, *d_raw = nullptr;
    hipError_t e = hipMalloc(&d_y, (size_t)M * npts * 8);
    if (e == hipSuccess) e = hipMalloc(&d_crd, (size_t)9 * ctx->nw * T * 8);
    if (e == hipSuccess) e = hipMalloc(&d_E, (size_t)T * 8);
    if (e == hipSuccess) e = hipMalloc(&d_raw, (size_t)npts * 8);
    auto release = [&]() {
        hipStreamSynchronize(ctx->stream);
        hipFree(d_y);
        hipFree(d_crd);
        hipFree(d_E);
        hipFree(d_raw);
    };
    if ((rc = check_hip(ctx, e, "alloc sampling"))) {
        release();
        return rc;
    }
    for (int64_t done = 0; done < npts && e == hipSuccess; done += T) {
        const int64_t np = std::min<int64_t>(T, npts - done);
        e = launch_sobol(ctx, ctx->d_sobol, M, skip + done, np, 1, d_y + done, npts, nullptr);
        if (e == hipSuccess) e = launch_map(ctx, d_y + done, npts, np, d_crd, T, 0);
        if (e == hipSuccess)
            e = ctx->pes == SCP3G_PES_TTM3F ? launch_ttm3f(ctx, ctx->nw, np, d_crd, T, 0, d_E)
                                            : launch_qtip4pf(ctx, ctx->nw, np, d_crd, T, 0, d_E);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_mlp_targets, dim3((unsigned)((np + 255) / 256)), dim3(256), 0,
                               ctx->stream, d_E, d_y + done, npts, ctx->d_omega, M, s->V0, np, d_raw + done);
            e = hipGetLastError();
        }
    }
    std::vector<double> raw(npts);
    if (e == hipSuccess) e = hipMemcpyAsync(raw.data(), d_raw, (size_t)npts * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if ((rc = check_hip(ctx, e, "sample training set"))) {
        release();
        return rc;
    }
    double mean = 0.0, var = 0.0;
    for (double r : raw) mean += r;
    mean /= (double)npts;
    for (double r : raw) var += (r - mean) * (r - mean);
    var /= (double)npts;
    s->t_mean = mean;
    s->t_std = var > 0 ? std::sqrt(var) : 1.0;
    std::vector<float> t(npts);
    for (int64_t p = 0; p < npts; ++p) t[p] = (float)((raw[p] - mean) / s->t_std);
    rc = load_training_set(ctx, s, d_y, t.data(), npts);
    release();
    return rc;
}

// Adam on batches of `batch` consecutive training points (cycling through
// the set); loss_host (nullable, `steps` entries) gets each step's mean
// squared error in normalised units.
int scp3g_mlp_train(scp3g_ctx *ctx, int steps, int batch, double lr, float *loss_host)
{
    int rc = mlp_ready(ctx);
    if (rc) return rc;
    MlpState *s = ctx->mlp;
    if (!s->Xtr) return fail(ctx, -2, "no training set");
    if (batch <= 0 || batch % 64 || batch > s->ntrain) return fail(ctx, -4, "bad batch size");
    if ((rc = check_hip(ctx, ensure_batch(s, batch), "alloc batch"))) return rc;
    const int64_t nb = s->ntrain / batch;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    WeightCopies wc;
    wc.out = s->h;
    for (int l = 0; l < 3; ++l) {
        wc.w_off[l] = s->L[l].w_off;
        wc.in[l] = s->L[l].in;
        wc.Wb[l] = s->L[l].Wb;
        wc.WbT[l] = s->L[l].WbT;
    }
    hipEvent_t ev;
    timer_begin(ctx, "mlp_train", &ev);
    std::vector<float> losses;
    for (int it = 0; it < steps; ++it) {
        const int64_t b0 = (it % nb) * batch;
        float *ld = s->lossd + (it % (64 * 1024));
        hipError_t e = grad_step(ctx, s, b0, batch, ld, false);
        ++s->step;
        const float c1 = 1.f - std::pow(b1, (float)s->step), c2 = 1.f - std::pow(b2, (float)s->step);
        hipLaunchKernelGGL(k_adam_update, dim3((unsigned)((s->nparam + 255) / 256)), dim3(256), 0,
                           ctx->stream, s->P, s->Gpart, s->gsplit, s->G, s->m, s->v, s->nparam, (float)lr, b1,
                           b2, eps, c1, c2, wc);
        if (e == hipSuccess) e = hipGetLastError();
        if ((rc = check_hip(ctx, e, "train step"))) return rc;
        if (loss_host && ((it + 1) % (64 * 1024) == 0 || it + 1 == steps)) {
            const int64_t first = (it / (64 * 1024)) * (64 * 1024);
            const int64_t cnt = it + 1 - first;
            std::vector<float> tmp(cnt);
            hipMemcpyAsync(tmp.data(), s->lossd, cnt * 4, hipMemcpyDeviceToHost, ctx->stream);
            hipStreamSynchronize(ctx->stream);
            for (int64_t j = 0; j < cnt; ++j) loss_host[first + j] = tmp[j] / (float)batch;
        }
    }
    timer_end(ctx, "mlp_train", ev);
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), "train");
}

// Gradient of the batch loss at the current parameters (no update) over the
// first `batch` training points; grad_host[nparam], *loss = mean squared error.
int scp3g_mlp_grad(scp3g_ctx *ctx, int batch, float *grad_host, float *loss)
{
    int rc = mlp_ready(ctx);
    if (rc) return rc;
    MlpState *s = ctx->mlp;
    if (!s->Xtr || batch <= 0 || batch % 64 || batch > s->ntrain) return fail(ctx, -4, "bad batch");
    if ((rc = check_hip(ctx, ensure_batch(s, batch), "alloc batch"))) return rc;
    if ((rc = check_hip(ctx, grad_step(ctx, s, 0, batch, s->lossd, true), "grad"))) return rc;
    hipStreamSynchronize(ctx->stream);
    if (grad_host) hipMemcpy(grad_host, s->G, (size_t)s->nparam * 4, hipMemcpyDeviceToHost);
    if (loss) {
        float l;
        hipMemcpy(&l, s->lossd, 4, hipMemcpyDeviceToHost);
        *loss = l / (float)batch;
    }
    return 0;
}

// V_nn (hartree) of npts points y_dev (mode-major, ld ldy, halved normals).
int scp3g_mlp_eval(scp3g_ctx *ctx, const double *y_dev, int64_t ldy, int64_t npts, double *V_dev)
{
    int rc = mlp_ready(ctx);
    if (rc) return rc;
    MlpState *s = ctx->mlp;
    if (npts <= 0 || ldy < npts || !y_dev || !V_dev) return fail(ctx, -4, "bad arguments");
    const int64_t T = 1 << 16;
    for (int64_t done = 0; done < npts && !rc; done += T) {
        const int64_t np = std::min<int64_t>(T, npts - done);
        hipError_t e = eval_f(ctx, s, y_dev + done, ldy, np);
        if (e == hipSuccess) {
            // V_nn = w_nn + 1/2 sum omega y^2 (mode 2)
            hipLaunchKernelGGL(k_mlp_weights, dim3((unsigned)((np + 255) / 256)), dim3(256), 0,
                               ctx->stream, s->out, y_dev + done, ldy, ctx->d_omega, s->M,
                               s->V0 + s->t_mean, s->t_std, np, np, (const double *)nullptr, 2, 1.0,
                               V_dev + done);
            e = hipGetLastError();
        }
        rc = check_hip(ctx, e, "mlp eval");
    }
    return rc;
}

// Control-variate estimator of H: the surrogate weights over Sobol seeds
// skip_nn .. skip_nn+P_nn-1, each scaled 1/P_nn, plus the correction
// bV - w_nn over seeds skip_t .. skip_t+P_t-1, scaled 1/P_t.  The running
// sum then normalises to 1 (scp3g_h_download divides by it).
int scp3g_accumulate_surrogate(scp3g_ctx *ctx, int64_t skip, int64_t npts)
{
    int rc = mlp_ready(ctx);
    if (rc) return rc;
    if (npts < 0 || skip < 0) return fail(ctx, -4, "negative skip or point count");
    return scp3g_accumulate_mode(ctx, skip, npts, 1, 1.0);
}

int scp3g_accumulate_correction(scp3g_ctx *ctx, int64_t skip, int64_t npts)
{
    int rc = mlp_ready(ctx);
    if (rc) return rc;
    if (npts < 0 || skip < 0) return fail(ctx, -4, "negative skip or point count");
    if (ctx->estimator != SCP3G_EST_ENERGY)
        return fail(ctx, -4, "the control variate needs the energy (-G) estimator");
    // weights bV - w_nn, normaliser += npts
    return scp3g_accumulate_mode(ctx, skip, npts, 5, 1.0);
}

int scp3g_accumulate_cv(scp3g_ctx *ctx, int64_t skip_nn, int64_t P_nn, int64_t skip_t, int64_t P_t)
{
    int rc = mlp_ready(ctx);
    if (rc) return rc;
    if (P_nn <= 0 || P_t <= 0 || skip_nn < 0 || skip_t < 0) return fail(ctx, -4, "bad point counts");
    for (int pass = 0; pass < 2; ++pass) {
        const int64_t skip = pass ? skip_t : skip_nn, P = pass ? P_t : P_nn;
        if ((rc = scp3g_accumulate_mode(ctx, skip, P, pass ? 2 : 1, 1.0 / (double)P))) return rc;
    }
    return 0;
}

}  // extern "C"
