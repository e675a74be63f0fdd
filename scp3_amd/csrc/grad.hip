// Gradient (Stein-identity) estimator of the vibrational Hamiltonian.
//
// Reference: SCP3_a with use_gradient (SCP3.cpp:203-217) and
// ScaledMatrixElements::addEpot(q, V, Vq, M) (ScaledMatrixElements.cpp:
// 100-725), whose per-element formulas (TripleExcitations.h) are all one
// rule: for <a|V|b> with P = phi_a phi_b and D = the modes excited in a or b,
//
//   est(a,b) = 1/|D| sum_{d in D} [V (P|_{q_d=0} + d_d g_d) + V_d g_d],
//   g_d = (P - P|_{q_d=0}) / (2 q_d),
//
// V the weight bV and V_d = d bV / d q_d.  Splitting each Hermite factor as
// h_i(x) = z_i + 2x e_i(x) (z_i = h_i(0)) turns the sum over d into
//
//   |D| est(a,b) = Psi_a phi_b + phi_a Psi_b
//                + sum_{d in S_a n S_b} [w_d u^d_a u^d_b - V phi_a phi_b],
//
// with S_a the modes of a, Psi_a = sum_{d in S_a} s_{a_d}(q_d) prod_{f != d} h_f
// (s_1 = V_d/sqrt2, s_2 = V_d x/sqrt2, s_3 = 2Vx/sqrt3 + V_d (2x^2-3)/(2sqrt3)),
// u^d_a = phi_a with its mode-d factor replaced by e_{a_d}(q_d)
// (e_1 = 1/sqrt2, e_2 = x/sqrt2, e_3 = (2x^2-3)/(2sqrt3)) and
// w_d = V (4 q_d^2 + 2) - 2 q_d V_d.  On the GPU:
//
//   * Psi phi^T + phi Psi^T as a SYRK pair: (phi + s Psi) with weight 1/(2s)
//     and (phi - s Psi) with weight -1/(2s), s per point = 1/max(|V|, |V_d|)
//     so both terms are O(phi); in FP64 two launches of the LDS-tiled FP64
//     SYRK on the tables of k_grad_table, in the mixed form one signed
//     split-BF16 SYRK over both halves (k_grad_split, syrk_fast.hip);
//   * the block corrections: for each mode d, the |A_d| x |A_d| upper
//     triangle over the basis functions A_d that excite d (k_grad_blocks,
//     FP64 VALU on 64 x 64 or 128 x 128 tiles; in the mixed form with
//     blocks over 64 functions, one signed split-BF16 SYRK per mode,
//     k_corr_split, added by k_corr_scatter);
//   * M(0,0) = sum V (ScaledMatrixElements.cpp:717) and the reference's own
//     element placements (below) as a short list of per-point sums
//     (k_grad_fix);
//   * at read time the running sum is divided by |D(a,b)| and the fixed
//     elements are placed (k_grad_effective, k_grad_compat).
//
// The reference's placement (form SCP3G_EST_GRADIENT) differs from the rule
// in five element families: column h1(k)h1(l) row h1(q_k) holds the rule's
// row-h2(q_k) value and row h2(q_k) the disjoint-mode formula m2_f_k1_l1 at
// m = k (ScaledMatrixElements.cpp:192-199); column h3(k) row h1(q_k) holds
// the rule's row-h2(q_k) value and row h2(q_k) m2_f_k3 at m = k (:261-265);
// the diagonal h3(k) uses m3_f_m3 (TripleExcitations.h:770), the rule plus
// V (4.5 q^2 - 3 q^4).  SCP3G_EST_GRADIENT_STEIN uses the rule throughout.
#include "scp3g_internal.h"

#include <algorithm>
#include <vector>

namespace scp3g {

constexpr double kR2 = 1.4142135623730951;  // sqrt(2)
constexpr double kR3 = 1.7320508075688772;  // sqrt(3)
constexpr double kR6 = 2.4494897427831781;  // sqrt(6)
constexpr int kGP = 16;                     // points per staged batch
constexpr int kFixPerThread = 24;           // per-point sums per thread of k_grad_fix

__device__ __forceinline__ double hval(int deg, double q)
{
    if (deg == 1) return kR2 * q;
    if (deg == 2) return (2.0 * q * q - 1.0) / kR2;
    if (deg == 3) return (2.0 * q * q - 3.0) * q / kR3;
    return 1.0;
}

// e_i(x) = (h_i(x) - h_i(0)) / (2x)
__device__ __forceinline__ double eval_e(int deg, double q)
{
    if (deg == 1) return 1.0 / kR2;
    if (deg == 2) return q / kR2;
    return (2.0 * q * q - 3.0) / (2.0 * kR3);
}

// ---- per point: Vq, the SYRK scale and weights ------------------------------
//
// Vq_m = MUa(:, k)^T (grad V * bohr/autokcal) - omega_k y_k over the selected
// modes k = modes[m] (SCP3.cpp:213-217); the point's gradient is staged in
// LDS as [coordinate][lane] and MUa is read as uniform (scalar) loads.
__global__ void __launch_bounds__(64)
    k_vq(const double *__restrict__ G, int64_t ldg, int n3, const double *__restrict__ MUa,
         const double *__restrict__ omega, const int *__restrict__ modes, int nsel,
         const double *__restrict__ y, int64_t ldy, const double *__restrict__ w, int64_t npts,
         int64_t npad, double *__restrict__ Vq, int64_t ldv, double *__restrict__ wp,
         double *__restrict__ wm, double *__restrict__ sc)
{
    extern __shared__ double gs[];  // [n3][L]
    const int L = blockDim.x, lane = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * L + lane;
    const int64_t ic = i < npts ? i : npts - 1;
    const double f = kBohr / kAutokcalpmol;
    for (int c = 0; c < n3; ++c) gs[c * L + lane] = G[(size_t)c * ldg + ic] * f;
    if (i >= npad) return;
    if (i >= npts) {  // padded points carry no weight
        for (int m = 0; m < nsel; ++m) Vq[(size_t)m * ldv + i] = 0.0;
        wp[i] = 0.0;
        wm[i] = 0.0;
        sc[i] = 0.0;
        return;
    }
    double vmax = fabs(w[i]);
    for (int m = 0; m < nsel; ++m) {
        const int k = modes[m];
        const double *mu = MUa + (size_t)k * n3;
        double s = 0.0;
        for (int c = 0; c < n3; ++c) s = fma(mu[c], gs[c * L + lane], s);
        const double v = s - omega[k] * y[(size_t)k * ldy + i];
        Vq[(size_t)m * ldv + i] = v;
        vmax = fmax(vmax, fabs(v));
    }
    const double s = vmax > 0.0 ? 1.0 / vmax : 1.0;
    sc[i] = s;
    wp[i] = 0.5 / s;
    wm[i] = -0.5 / s;
}

// ---- B = phi +- s Psi ---------------------------------------------------------
//
// Layout as the -G basis table: [row][point], rows padded with zeros.  A
// block holds PB points; its threads first evaluate h_1..3 and s_1..3 (times
// +-s) of every selected mode of their point into LDS, then 256/PB row
// groups sweep the block's rows.
__global__ void __launch_bounds__(256)
    k_grad_table(const double *__restrict__ y, int64_t ldy, const int *__restrict__ modes, int nsel,
                 const RowDesc *__restrict__ desc, int64_t n, int64_t nrows_pad, int rows_per_block,
                 int64_t npts, const double *__restrict__ w, const double *__restrict__ Vq,
                 int64_t ldv, const double *__restrict__ sc, double sgn, double *__restrict__ B,
                 int64_t ldb, int pb_log2)
{
    extern __shared__ double hs[];  // [6][nsel][PB]: h1 h2 h3, then +-s * (s1 s2 s3)
    const int PB = 1 << pb_log2;
    const int ng = 256 >> pb_log2;
    const int lane = threadIdx.x & (PB - 1);
    const int grp = threadIdx.x >> pb_log2;
    const int64_t i = (int64_t)blockIdx.x * PB + lane;
    const bool live = i < npts;
    for (int m = grp; m < nsel; m += ng) {
        double q = 0.0, V = 0.0, vd = 0.0, t = 0.0;
        if (live) {
            q = y[(size_t)modes[m] * ldy + i];
            V = w[i];
            vd = Vq[(size_t)m * ldv + i];
            t = sgn * sc[i];
        }
        hs[(0 * nsel + m) * PB + lane] = hval(1, q);
        hs[(1 * nsel + m) * PB + lane] = hval(2, q);
        hs[(2 * nsel + m) * PB + lane] = hval(3, q);
        hs[(3 * nsel + m) * PB + lane] = t * (vd / kR2);
        hs[(4 * nsel + m) * PB + lane] = t * (vd * q / kR2);
        hs[(5 * nsel + m) * PB + lane] =
            t * (2.0 * V * q / kR3 + vd * (2.0 * q * q - 3.0) / (2.0 * kR3));
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
            double f[3], s[3];
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const int mode = d.f[t] & 0xffff;
                const int deg = d.f[t] >> 16;
                f[t] = deg ? hs[((deg - 1) * nsel + mode) * PB + lane] : 1.0;
                s[t] = deg ? hs[((deg + 2) * nsel + mode) * PB + lane] : 0.0;
            }
            const double phi = (f[0] * f[1]) * f[2];
            const double psi = s[0] * f[1] * f[2] + f[0] * s[1] * f[2] + f[0] * f[1] * s[2];
            v = phi + psi;
        }
        B[(size_t)a * ldb + i] = v;
    }
}

// ---- block corrections --------------------------------------------------------
//
// Work item (d, ti, tj): rows ti*TG.. and columns tj*TG.. (ti <= tj) of the
// upper triangle of A_d; blockIdx.y takes a chunk of kchunk points.  Thread
// (tr, tc) = (tid / 16, tid % 16) owns the RT x RT elements (tr + 16i,
// tc + 16j), RT = TG / 16.  Per batch of 16 points the block stages q (all
// selected modes) and w_d, V; each thread then evaluates u^d and phi of the
// rows tr + 16j (I rows for j < RT, J rows after) at point tc, storing
// w_d u and V phi for the I side, and the batch is contracted from LDS.
// TG = 64 suits the small A_d of a doubles-only basis (|A_d| = nmodes2 + 1),
// TG = 128 the large ones of a basis with triples (fewer LDS reads per
// FMA).  LDS (dynamic): q [nsel][16], row descriptors [2 TG], then
// [16][TG + 1] x 4.  Partial sums go to part[split][work][TG x TG].
template <int TG>
__global__ void __launch_bounds__(256)
    k_grad_blocks(const double *__restrict__ y, int64_t ldy, const int *__restrict__ modes, int nsel,
                  const RowDesc *__restrict__ desc, const int *__restrict__ rows,
                  const int4 *__restrict__ work, int nwork, const double *__restrict__ w,
                  const double *__restrict__ Vq, int64_t ldv, int64_t kchunk, int64_t npts,
                  double *__restrict__ part)
{
    constexpr int RT = TG / 16, LD = TG + 1;
    extern __shared__ double gsm[];
    double *qs = gsm;                                        // [nsel][kGP]
    uint4 *rd = reinterpret_cast<uint4 *>(qs + nsel * kGP);  // [2 TG]: f0 f1 f2 slot
    double *aI = reinterpret_cast<double *>(rd + 2 * TG);    // [kGP][LD] x 4
    double *bI = aI + kGP * LD, *uJ = bI + kGP * LD, *fJ = uJ + kGP * LD;
    __shared__ double wd[kGP], vv[kGP];
    const int4 wk = work[blockIdx.x];  // {offset of A_d, |A_d|, d, ti | tj << 16}
    const int off = wk.x, na = wk.y, d = wk.z, ti = wk.w & 0xffff, tj = wk.w >> 16;
    const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
    for (int r = tid; r < 2 * TG; r += 256) {
        const int e = ((r < TG) ? ti : tj) * TG + (r & (TG - 1));
        uint4 v = make_uint4(0, 0, 0, 0xffffffffu);
        if (e < na) {
            const int packed = rows[off + e];
            const RowDesc ds = desc[packed & 0x3fffffff];
            v = make_uint4(ds.f[0], ds.f[1], ds.f[2], (unsigned)packed >> 30);
        }
        rd[r] = v;
    }
    const int64_t p0 = (int64_t)blockIdx.y * kchunk;
    int64_t p1 = p0 + kchunk;
    if (p1 > npts) p1 = npts;
    double acc[RT][RT];
#pragma unroll
    for (int i = 0; i < RT; ++i)
#pragma unroll
        for (int j = 0; j < RT; ++j) acc[i][j] = 0.0;
    for (int64_t pb = p0; pb < p1; pb += kGP) {
        __syncthreads();  // the previous batch is consumed (and rd is written)
        for (int e = tid; e < nsel * kGP; e += 256) {
            const int m = e / kGP, p = e - m * kGP;
            qs[e] = pb + p < p1 ? y[(size_t)modes[m] * ldy + pb + p] : 0.0;
        }
        if (tid < kGP) {
            const int64_t p = pb + tid;
            double wv = 0.0, V = 0.0;
            if (p < p1) {
                V = w[p];
                const double qd = y[(size_t)modes[d] * ldy + p];
                wv = V * (4.0 * qd * qd + 2.0) - 2.0 * qd * Vq[(size_t)d * ldv + p];
            }
            wd[tid] = wv;
            vv[tid] = V;
        }
        __syncthreads();
        const int p = tc;
        for (int j = 0; j < 2 * RT; ++j) {
            const int r = tr + 16 * j;  // 0 .. 2 TG - 1
            const uint4 ds = rd[r];
            double u = 0.0, ph = 0.0;
            if (ds.w != 0xffffffffu) {
                const uint32_t fs[3] = {ds.x, ds.y, ds.z};
                ph = 1.0;
                u = 1.0;
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const int deg = fs[t] >> 16;
                    const double q = deg ? qs[(fs[t] & 0xffff) * kGP + p] : 0.0;
                    const double h = hval(deg, q);
                    ph *= h;
                    u *= (t == (int)ds.w) ? eval_e(deg, q) : h;
                }
            }
            if (r < TG) {
                aI[p * LD + r] = wd[p] * u;
                bI[p * LD + r] = vv[p] * ph;
            } else {
                uJ[p * LD + r - TG] = u;
                fJ[p * LD + r - TG] = ph;
            }
        }
        __syncthreads();
#pragma unroll 2
        for (int q = 0; q < kGP; ++q) {
            double a[RT], b[RT], u[RT], f[RT];
#pragma unroll
            for (int i = 0; i < RT; ++i) {
                a[i] = aI[q * LD + tr + 16 * i];
                b[i] = bI[q * LD + tr + 16 * i];
                u[i] = uJ[q * LD + tc + 16 * i];
                f[i] = fJ[q * LD + tc + 16 * i];
            }
#pragma unroll
            for (int i = 0; i < RT; ++i)
#pragma unroll
                for (int j = 0; j < RT; ++j) acc[i][j] += a[i] * u[j] - b[i] * f[j];
        }
    }
    double *dst = part + ((size_t)blockIdx.y * nwork + blockIdx.x) * (TG * TG);
#pragma unroll
    for (int i = 0; i < RT; ++i)
#pragma unroll
        for (int j = 0; j < RT; ++j) dst[(tr + 16 * i) * TG + tc + 16 * j] = acc[i][j];
}

// Sum the splits in order and add each element to H; two modes' blocks can
// share an element (a, b exciting both), so the add is atomic.
__global__ void __launch_bounds__(256)
    k_grad_blocks_reduce(const double *__restrict__ part, const int *__restrict__ rows,
                         const int4 *__restrict__ work, int nwork, int nsplit, int tg, int64_t n,
                         double *__restrict__ H)
{
    const int4 wk = work[blockIdx.x];
    const int off = wk.x, na = wk.y, ti = wk.w & 0xffff, tj = wk.w >> 16;
    const int tt = tg * tg;
    for (int e = threadIdx.x; e < tt; e += 256) {
        const int r = e / tg, c = e - r * tg;
        const int er = ti * tg + r, ec = tj * tg + c;
        if (er >= na || ec >= na || (ti == tj && r > c)) continue;
        double s = 0.0;
        for (int k = 0; k < nsplit; ++k) s += part[((size_t)k * nwork + blockIdx.x) * tt + e];
        const int64_t a = rows[off + er] & 0x3fffffff, b = rows[off + ec] & 0x3fffffff;
        atomicAdd(&H[(size_t)b * n + a], s);
    }
}

// ---- per-point sums of the fixed elements ------------------------------------
//
// fix = {type, k, l, col}: 0 = sum V (element (0,0)); 2 = m2_f_k1_l1 at
// m = k, row h2(q_k) of column h1(k)h1(l); 4 = m2_f_k3 at m = k, row h2(q_k)
// of column h3(k); 5 = V (4.5 q_k^2 - 3 q_k^4), the diagonal h3(k).  Block
// (x, y) holds fixes from y * 256 kFixPerThread on (kFixPerThread per
// thread) for chunk x of the points.
__device__ __forceinline__ double fix_value(int type, double qk, double ql, double vk, double vl,
                                            double V)
{
    const double x2 = qk * qk;
    if (type == 0) return V;
    if (type == 2)  // (2 q_m q_k q_l V_m + (q_l V_k + q_k V_l)(2 q_m^2 - 1)) / (3 sqrt2), m = k
        return (2.0 * qk * qk * ql * vk + (ql * vk + qk * vl) * (2.0 * x2 - 1.0)) / (3.0 * kR2);
    if (type == 4)  // (4 V q_k (2 q_m^2 - 1) + (2 q_k^2 - 3)(2 q_k q_m V_m + V_k (2 q_m^2 - 1))) / (4 sqrt6)
        return (4.0 * V * qk * (2.0 * x2 - 1.0) + (2.0 * x2 - 3.0) * (2.0 * x2 * vk + vk * (2.0 * x2 - 1.0))) /
               (4.0 * kR6);
    return V * (4.5 * x2 - 3.0 * x2 * x2);
}

__global__ void __launch_bounds__(256)
    k_grad_fix(const int4 *__restrict__ fix, int nfix, const double *__restrict__ y, int64_t ldy,
               const int *__restrict__ modes, int nsel, const double *__restrict__ w,
               const double *__restrict__ Vq, int64_t ldv, int64_t kchunk, int64_t npts,
               double *__restrict__ part)
{
    extern __shared__ double st[];  // q [nsel][kGP], Vq [nsel][kGP], V [kGP]
    double *qs = st, *vs = st + nsel * kGP, *Vs = vs + nsel * kGP;
    const int tid = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * kchunk;
    int64_t p1 = p0 + kchunk;
    if (p1 > npts) p1 = npts;
    double acc[kFixPerThread];
#pragma unroll
    for (int j = 0; j < kFixPerThread; ++j) acc[j] = 0.0;
    for (int64_t pb = p0; pb < p1; pb += kGP) {
        __syncthreads();
        for (int e = tid; e < nsel * kGP; e += 256) {
            const int m = e / kGP, p = e - m * kGP;
            const bool ok = pb + p < p1;
            qs[e] = ok ? y[(size_t)modes[m] * ldy + pb + p] : 0.0;
            vs[e] = ok ? Vq[(size_t)m * ldv + pb + p] : 0.0;
        }
        if (tid < kGP) Vs[tid] = pb + tid < p1 ? w[pb + tid] : 0.0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kFixPerThread; ++j) {
            const int f = blockIdx.y * 256 * kFixPerThread + tid + 256 * j;
            if (f >= nfix) break;
            const int4 fx = fix[f];
            for (int p = 0; p < kGP; ++p)
                acc[j] += fix_value(fx.x, qs[fx.y * kGP + p], qs[fx.z * kGP + p], vs[fx.y * kGP + p],
                                    vs[fx.z * kGP + p], Vs[p]);
        }
    }
#pragma unroll
    for (int j = 0; j < kFixPerThread; ++j) {
        const int f = blockIdx.y * 256 * kFixPerThread + tid + 256 * j;
        if (f < nfix) part[(size_t)blockIdx.x * nfix + f] = acc[j];
    }
}

__global__ void k_grad_fix_reduce(const double *__restrict__ part, int nfix, int nsplit,
                                  double *__restrict__ F)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfix) return;
    double s = 0.0;
    for (int k = 0; k < nsplit; ++k) s += part[(size_t)k * nfix + f];
    F[f] += s;
}

// ---- the effective running sum ------------------------------------------------
//
// dst (upper) = S(a,b) / |D(a,b)|, element (0,0) = F[0]; lower triangle zero.
__global__ void __launch_bounds__(256)
    k_grad_effective(const double *__restrict__ S, int64_t n, const RowDesc *__restrict__ desc,
                     const double *__restrict__ F, double *__restrict__ dst)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    const int64_t c = idx / n, r = idx - c * n;
    double v = 0.0;
    if (r <= c) {
        if (c == 0) {
            v = F[0];
        } else {
            const RowDesc a = desc[r], b = desc[c];
            int md[6], nd = 0;
#pragma unroll
            for (int t = 0; t < 6; ++t) {
                const uint32_t fct = t < 3 ? a.f[t] : b.f[t - 3];
                if (!(fct >> 16)) continue;
                const int m = fct & 0xffff;
                bool seen = false;
                for (int s = 0; s < nd; ++s) seen |= md[s] == m;
                if (!seen) md[nd++] = m;
            }
            v = S[idx] / (double)nd;
        }
    }
    dst[idx] = v;
}

// Place the reference's elements: fix {type, k, l, col} with rows h1(q_k) =
// 1 + k and h2(q_k) = 1 + nsel + k; types 2 and 4 first copy the rule's row
// h2 value into row h1, then write their own sum into row h2.
__global__ void k_grad_compat(const int4 *__restrict__ fix, int nfix, int nsel, int64_t n,
                              const double *__restrict__ F, double *__restrict__ dst)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfix || f == 0) return;
    const int4 fx = fix[f];
    const int64_t col = fx.w;
    if (fx.x == 5) {
        dst[(size_t)col * n + col] += F[f];
        return;
    }
    const int64_t r1 = 1 + fx.y, r2 = 1 + nsel + fx.y;
    dst[(size_t)col * n + r1] = dst[(size_t)col * n + r2];
    dst[(size_t)col * n + r2] = F[f];
}

// Mode d's block sum Hd (na x na, column-major upper, from the split-BF16
// SYRK) added into H at the rows of A_d; the modes run one after another on
// the stream, so the adds need no atomics.
__global__ void __launch_bounds__(256)
    k_corr_scatter(const double *__restrict__ Hd, int na, const int *__restrict__ rows, int off, int64_t n,
                   double *__restrict__ H)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)na * na) return;
    const int c = (int)(idx / na), r = (int)(idx - (int64_t)c * na);
    if (r > c) return;
    const int64_t a = rows[off + r] & 0x3fffffff, b = rows[off + c] & 0x3fffffff;
    H[(size_t)b * n + a] += Hd[idx];
}

// ---- host -----------------------------------------------------------------------

void grad_release(scp3g_ctx *ctx)
{
    if (ctx->d_grows) hipFree(ctx->d_grows);
    if (ctx->d_gwork) hipFree(ctx->d_gwork);
    if (ctx->d_gfix) hipFree(ctx->d_gfix);
    if (ctx->d_gpart) hipFree(ctx->d_gpart);
    if (ctx->d_gcx) hipFree(ctx->d_gcx);
    ctx->d_gcx = nullptr;
    ctx->gcx_bytes = 0;
    ctx->gblock_off.clear();
    ctx->gblock_na.clear();
    ctx->d_grows = nullptr;
    ctx->d_gwork = nullptr;
    ctx->d_gfix = nullptr;
    ctx->d_gpart = nullptr;
    ctx->gpart_bytes = 0;
    ctx->ngwork = 0;
    ctx->ngfix = 0;
    ctx->grad_n = 0;
}

// A_d lists, the correction work items and the fix list for the current basis.
int grad_setup(scp3g_ctx *ctx)
{
    const int64_t n = ctx->nbasis;
    const int nsel = (int)ctx->modes.size();
    if (ctx->grad_n == n && ctx->grad_form == ctx->estimator && ctx->d_gwork) return 0;
    grad_release(ctx);
    std::vector<std::vector<int>> A((size_t)nsel);
    for (int64_t a = 0; a < n; ++a)
        for (int t = 0; t < 3; ++t) {
            const uint32_t f = ctx->desc[a].f[t];
            if (f >> 16) A[f & 0xffff].push_back((int)a | (t << 30));
        }
    std::vector<int> rows;
    std::vector<int4> work;
    size_t amax = 0;
    for (const auto &v : A) amax = std::max(amax, v.size());
    const int tg = amax > 64 ? 128 : 64;
    for (int d = 0; d < nsel; ++d) {
        const int off = (int)rows.size(), na = (int)A[d].size();
        ctx->gblock_off.push_back(off);
        ctx->gblock_na.push_back(na);
        rows.insert(rows.end(), A[d].begin(), A[d].end());
        const int nt = (na + tg - 1) / tg;
        for (int tj = 0; tj < nt; ++tj)
            for (int ti = 0; ti <= tj; ++ti) work.push_back(make_int4(off, na, d, ti | (tj << 16)));
    }
    std::vector<int4> fix;
    fix.push_back(make_int4(0, 0, 0, 0));
    if (ctx->estimator == SCP3G_EST_GRADIENT) {
        const int n2 = ctx->nmodes2, n3 = ctx->nmodes3;
        const int64_t kl0 = 1 + nsel + n2, h30 = kl0 + (int64_t)n2 * (n2 - 1) / 2;
        int64_t c = kl0;
        for (int k = 0; k < n2; ++k)
            for (int l = k + 1; l < n2; ++l, ++c) {
                fix.push_back(make_int4(2, k, l, (int)c));
                fix.push_back(make_int4(2, l, k, (int)c));
            }
        for (int k = 0; k < n3; ++k) {
            fix.push_back(make_int4(4, k, k, (int)(h30 + k)));
            fix.push_back(make_int4(5, k, k, (int)(h30 + k)));
        }
    }
    int rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_grows, rows.size() * sizeof(int)), "alloc grad rows")))
        return rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_gwork, work.size() * sizeof(int4)), "alloc grad work")))
        return rc;
    if ((rc = check_hip(ctx, hipMalloc(&ctx->d_gfix, fix.size() * sizeof(int4)), "alloc grad fix")))
        return rc;
    hipMemcpy(ctx->d_grows, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice);
    hipMemcpy(ctx->d_gwork, work.data(), work.size() * sizeof(int4), hipMemcpyHostToDevice);
    hipMemcpy(ctx->d_gfix, fix.data(), fix.size() * sizeof(int4), hipMemcpyHostToDevice);
    ctx->ngwork = (int)work.size();
    ctx->grad_tile = tg;
    ctx->ngfix = (int)fix.size();
    ctx->grad_n = n;
    ctx->grad_form = ctx->estimator;
    return 0;
}

static hipError_t ensure_gpart(scp3g_ctx *ctx, size_t need)
{
    if (ctx->gpart_bytes >= need) return hipSuccess;
    if (ctx->d_gpart) hipFree(ctx->d_gpart);
    ctx->d_gpart = nullptr;
    ctx->gpart_bytes = 0;
    hipError_t e = hipMalloc(&ctx->d_gpart, need);
    if (e == hipSuccess) ctx->gpart_bytes = need;
    return e;
}

hipError_t launch_vq(scp3g_ctx *ctx, const double *G, int64_t ldg, const double *y, int64_t ldy,
                     const double *w, int64_t npts, double *Vq, int64_t ldv, double *wp, double *wm,
                     double *sc)
{
    const int64_t npad = (npts + 255) / 256 * 256;
    if (npad <= 0) return hipSuccess;
    const int n3 = 9 * ctx->nw;
    int L = 64;
    while (L > 16 && (size_t)n3 * L * sizeof(double) > 64 * 1024) L /= 2;
    const size_t smem = (size_t)n3 * L * sizeof(double);
    if (smem > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_vq, dim3((unsigned)((npad + L - 1) / L)), dim3(L), smem, ctx->stream, G, ldg,
                       n3, ctx->d_MUa, ctx->d_omega, ctx->d_modes, (int)ctx->modes.size(), y, ldy, w,
                       npts, npad, Vq, ldv, wp, wm, sc);
    return hipGetLastError();
}

hipError_t launch_grad_table(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts,
                             const double *w, const double *Vq, int64_t ldv, const double *sc,
                             double sgn, double *B, int64_t ldb)
{
    if (npts <= 0) return hipSuccess;
    const int64_t n = ctx->nbasis;
    const int64_t nrows_pad = (n + 63) / 64 * 64;
    const int nsel = (int)ctx->modes.size();
    // points per block: the largest power of two <= 64 whose six value
    // tables fit the LDS of a CU (160 KiB; 16 points for 174 modes)
    int pb_log2 = 6;
    while (pb_log2 > 3 && (size_t)6 * nsel * (1u << pb_log2) * sizeof(double) > 160 * 1024) --pb_log2;
    const size_t smem = (size_t)6 * nsel * (1u << pb_log2) * sizeof(double);
    if (smem > 160 * 1024) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_grad_table),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    const int64_t pblocks = (ldb + (1 << pb_log2) - 1) >> pb_log2;
    int rows_per_block = 256;
    while (rows_per_block > 16 && pblocks * ((nrows_pad + rows_per_block - 1) / rows_per_block) < 2048)
        rows_per_block /= 2;
    dim3 grid((unsigned)pblocks, (unsigned)((nrows_pad + rows_per_block - 1) / rows_per_block));
    hipLaunchKernelGGL(k_grad_table, grid, dim3(256), smem, ctx->stream, y, ldy, ctx->d_modes, nsel,
                       ctx->d_desc, n, nrows_pad, rows_per_block, npts, w, Vq, ldv, sc, sgn, B, ldb,
                       pb_log2);
    return hipGetLastError();
}

hipError_t launch_grad_corrections(scp3g_ctx *ctx, const double *y, int64_t ldy, int64_t npts,
                                   const double *w, const double *Vq, int64_t ldv, double *H,
                                   double *F)
{
    if (npts <= 0) return hipSuccess;
    const int nsel = (int)ctx->modes.size();
    const int64_t n = ctx->nbasis;
    if ((size_t)2 * nsel * kGP * sizeof(double) + kGP * 8 > 64 * 1024) return hipErrorInvalidValue;
    hipError_t e;
    if (ctx->precision == SCP3G_MIXED && ctx->grad_tile == 128) {
        // large A_d (a basis with triples): each mode's block as a signed
        // split-BF16 SYRK over 2 Kh slots into a compact FP64 sum, then added
        // into H.  The inner SYRK launches run untimed (the caller times the
        // whole correction).
        const int64_t Kh = (npts + 63) / 64 * 64, T = ctx->work_tile;
        int namax = 0;
        for (int v : ctx->gblock_na) namax = std::max(namax, v);
        const int64_t rp = fast_rows_pad(namax);
        const size_t plane = (size_t)rp * 2 * T * sizeof(unsigned short);
        const size_t hd = (size_t)namax * namax * sizeof(double);
        const size_t need = 2 * plane + (size_t)T / 4 + 256 + 2 * hd;
        if (ctx->gcx_bytes < need) {
            if (ctx->d_gcx) hipFree(ctx->d_gcx);
            ctx->d_gcx = nullptr;
            ctx->gcx_bytes = 0;
            if ((e = hipMalloc(&ctx->d_gcx, need)) != hipSuccess) return e;
            ctx->gcx_bytes = need;
        }
        char *base = reinterpret_cast<char *>(ctx->d_gcx);
        unsigned short *Xh = reinterpret_cast<unsigned short *>(base);
        unsigned short *Xl = reinterpret_cast<unsigned short *>(base + plane);
        double *Hd = reinterpret_cast<double *>(base + 2 * plane);
        double *Hcd = Hd + (size_t)namax * namax;
        unsigned *sgn = reinterpret_cast<unsigned *>(Hcd + (size_t)namax * namax);
        const bool timing = ctx->timing;
        ctx->timing = false;
        e = hipSuccess;
        for (int d = 0; d < nsel && e == hipSuccess; ++d) {
            const int na = ctx->gblock_na[d], off = ctx->gblock_off[d];
            if (na == 0) continue;
            const size_t bytes = (size_t)na * na * sizeof(double);
            if ((e = hipMemsetAsync(Hd, 0, bytes, ctx->stream)) != hipSuccess) break;
            if ((e = hipMemsetAsync(Hcd, 0, bytes, ctx->stream)) != hipSuccess) break;
            if ((e = launch_corr_split(ctx, y, ldy, npts, Kh, d, ctx->d_grows, off, na, w, Vq, ldv, Xh, Xl,
                                       sgn)) != hipSuccess)
                break;
            if ((e = launch_syrk_x3(ctx, na, Xh, Xl, sgn, 2 * Kh, Hd, Hcd, 3)) != hipSuccess) break;
            hipLaunchKernelGGL(k_corr_scatter, dim3((unsigned)(((int64_t)na * na + 255) / 256)), dim3(256), 0,
                               ctx->stream, Hd, na, ctx->d_grows, off, n, H);
            e = hipGetLastError();
        }
        ctx->timing = timing;
        if (e != hipSuccess) return e;
        goto fixes;
    }
    {
    // blocks: split the points so the grid has >= ~2048 workgroups, chunks >= 1024 points
    int64_t nsplit = 1;
    while ((int64_t)ctx->ngwork * nsplit < 2048 && npts / (nsplit * 2) >= 1024) nsplit *= 2;
    int64_t kchunk = ((npts + nsplit - 1) / nsplit + kGP - 1) / kGP * kGP;
    nsplit = (npts + kchunk - 1) / kchunk;
    const int tg = ctx->grad_tile;
    const size_t need = (size_t)nsplit * ctx->ngwork * tg * tg * sizeof(double);
    e = ensure_gpart(ctx, need);
    if (e != hipSuccess) return e;
    const size_t smem = (size_t)nsel * kGP * sizeof(double) + 2 * tg * sizeof(uint4) +
                        (size_t)4 * kGP * (tg + 1) * sizeof(double);
    if (smem > 150 * 1024) return hipErrorInvalidValue;
    const void *kfn = tg == 128 ? reinterpret_cast<const void *>(k_grad_blocks<128>)
                                : reinterpret_cast<const void *>(k_grad_blocks<64>);
    if ((e = hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem)) != hipSuccess)
        return e;
    if (tg == 128)
        hipLaunchKernelGGL(k_grad_blocks<128>, dim3((unsigned)ctx->ngwork, (unsigned)nsplit), dim3(256), smem,
                           ctx->stream, y, ldy, ctx->d_modes, nsel, ctx->d_desc, ctx->d_grows, ctx->d_gwork,
                           ctx->ngwork, w, Vq, ldv, kchunk, npts, ctx->d_gpart);
    else
        hipLaunchKernelGGL(k_grad_blocks<64>, dim3((unsigned)ctx->ngwork, (unsigned)nsplit), dim3(256), smem,
                           ctx->stream, y, ldy, ctx->d_modes, nsel, ctx->d_desc, ctx->d_grows, ctx->d_gwork,
                           ctx->ngwork, w, Vq, ldv, kchunk, npts, ctx->d_gpart);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_grad_blocks_reduce, dim3((unsigned)ctx->ngwork), dim3(256), 0, ctx->stream,
                       ctx->d_gpart, ctx->d_grows, ctx->d_gwork, ctx->ngwork, (int)nsplit, tg, n, H);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    }
fixes:
    // fixes: one block per chunk of >= 2048 points, at most 512 blocks
    int64_t fsplit = std::min<int64_t>(512, std::max<int64_t>(1, npts / 2048));
    int64_t fchunk = ((npts + fsplit - 1) / fsplit + kGP - 1) / kGP * kGP;
    fsplit = (npts + fchunk - 1) / fchunk;
    if ((e = ensure_gpart(ctx, (size_t)fsplit * ctx->ngfix * sizeof(double))) != hipSuccess) return e;
    const unsigned fy = (unsigned)((ctx->ngfix + 256 * kFixPerThread - 1) / (256 * kFixPerThread));
    hipLaunchKernelGGL(k_grad_fix, dim3((unsigned)fsplit, fy), dim3(256),
                       ((size_t)2 * nsel * kGP + kGP) * sizeof(double), ctx->stream, ctx->d_gfix,
                       ctx->ngfix, y, ldy, ctx->d_modes, nsel, w, Vq, ldv, fchunk, npts, ctx->d_gpart);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_grad_fix_reduce, dim3((unsigned)((ctx->ngfix + 255) / 256)), dim3(256), 0,
                       ctx->stream, ctx->d_gpart, ctx->ngfix, (int)fsplit, F);
    return hipGetLastError();
}

hipError_t launch_grad_effective(scp3g_ctx *ctx, const double *S, const double *F, double *dst)
{
    const int64_t n = ctx->nbasis, tot = n * n;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grad_effective, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream,
                       S, n, ctx->d_desc, F, dst);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ctx->ngfix <= 1) return e;
    hipLaunchKernelGGL(k_grad_compat, dim3((unsigned)((ctx->ngfix + 255) / 256)), dim3(256), 0,
                       ctx->stream, ctx->d_gfix, ctx->ngfix, (int)ctx->modes.size(), n, F, dst);
    return hipGetLastError();
}

}  // namespace scp3g
