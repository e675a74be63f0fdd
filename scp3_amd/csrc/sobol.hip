// Sobol sampling + Beasley-Springer-Moro inverse normal on the GPU.
//
// Reference: sobol::std_normal (sobol.cpp:28706-28714) = i8_sobol
// (sobol.cpp:14198-27867) followed by beasley_springer_moro
// (beasley_springer_moro.cpp:14-46), then SCP3.cpp:201 divides by sqrt(2).
//
// The reference advances one point per call with the Antonov-Saleev update
// lastq ^= v[lo0(seed)-1].  Here each lane owns one point of a 64-point
// group and a wave steps the whole group by 64 seeds at a time: adding 2^6
// to s flips, in the Gray code s ^ (s >> 1), exactly bit 5 and bit
// 6 + ctz(~(s >> 6)), so the state advances by two XORs.  The starting
// state of each lane is the XOR of v[j] over the set bits of its Gray code,
// which equals the reference's catch-up sum.  The integers are therefore
// bit-identical to i8_sobol's lastq; the normals match the reference up to
// the device log() used in BSM's tail branch.
#include "scp3g_internal.h"

#include <mutex>

#include "sobol_dirnums.inc"

namespace scp3g {

// Direction integers scaled by 2^(61-j) (sobol.cpp:27700-27780).
void sobol_directions(int dims, uint64_t *v)
{
    for (int d = 0; d < dims; d++) {
        uint64_t *row = v + (size_t)d * kSobolCols;
        if (d == 0) {
            for (int j = 0; j < kSobolCols; j++) row[j] = 1;
        } else {
            const unsigned p = scp3_sobol_poly[d];
            const int m = 31 - __builtin_clz(p);  // degree of the polynomial
            for (int j = 0; j < m; j++) row[j] = scp3_sobol_init[d][j];
            for (int j = m; j < kSobolCols; j++) {
                // Bratley-Fox recurrence: coefficient a_k of x^(m-1-k) is
                // bit (m-1-k) of p; term scaled by 2^(k+1)
                uint64_t nv = row[j - m];
                for (int k = 0; k < m; k++)
                    if ((p >> (m - 1 - k)) & 1u) nv ^= row[j - k - 1] << (k + 1);
                row[j] = nv;
            }
        }
        for (int j = 0; j < kSobolCols; j++) row[j] <<= (kSobolCols - 1 - j);
    }
}

// Affine matrix scrambling of the Sobol digits (Matousek 1998; the class of
// MATLAB's scramble(p, 'MatousekAffineOwen') behind the reference's -r
// files): per dimension a random lower-triangular binary matrix L with unit
// diagonal in digit order (output digit i depends on input digits 0..i,
// digit 0 the most significant, bit 61) and a random digital shift e, so a
// point x becomes L x + e over GF(2).  L is linear, so it is applied once to
// the direction integers (L of the XOR is the XOR of L) and e starts each
// lane's state.  Random bits: splitmix64 seeded with seed ^ (0x9e3779b97f4a7c15
// * (d + 1)), one draw per digit (the bits of L's row above the diagonal),
// then one for e.
static uint64_t splitmix64(uint64_t &st)
{
    uint64_t z = (st += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

void sobol_scramble(int dims, uint64_t seed, uint64_t *v, uint64_t *shift)
{
    const uint64_t all = (1ull << kSobolCols) - 1;
    for (int d = 0; d < dims; d++) {
        uint64_t st = seed ^ (0x9e3779b97f4a7c15ull * (uint64_t)(d + 1));
        uint64_t mask[kSobolCols];
        for (int i = 0; i < kSobolCols; i++) {
            const int b = kSobolCols - 1 - i;
            const uint64_t above = all & ~((2ull << b) - 1);
            mask[i] = (1ull << b) | (splitmix64(st) & above);
        }
        shift[d] = splitmix64(st) & all;
        uint64_t *row = v + (size_t)d * kSobolCols;
        for (int j = 0; j < kSobolCols; j++) {
            uint64_t out = 0;
            for (int i = 0; i < kSobolCols; i++)
                out |= (uint64_t)(__builtin_popcountll(row[j] & mask[i]) & 1) << (kSobolCols - 1 - i);
            row[j] = out;
        }
    }
}

// beasley_springer_moro.cpp:18-23
__constant__ double kBsmA[4] = {2.50662823884, -18.61500062529, 41.39119773534,
                                -25.44106049637};
__constant__ double kBsmB[4] = {-8.47351093090, 23.08336743743, -21.06224101826,
                                3.13082909833};
__constant__ double kBsmC[9] = {0.3374754822726147, 0.9761690190917186, 0.1607979714918209,
                                0.0276438810333863, 0.0038405729373609, 0.0003951896511919,
                                0.0000321767881768, 0.0000002888167364, 0.0000003960315187};

// Same operation order as beasley_springer_moro.cpp:25-44, no contraction:
// the central branch (|u - 1/2| < 0.42) and the tail branch as two
// functions, so that the kernel can run the tail values of a wave together.
__device__ __forceinline__ double bsm_central(double yj)
{
#pragma clang fp contract(off)
    const double r = yj * yj;
    return yj * (((kBsmA[3] * r + kBsmA[2]) * r + kBsmA[1]) * r + kBsmA[0]) /
           ((((kBsmB[3] * r + kBsmB[2]) * r + kBsmB[1]) * r + kBsmB[0]) * r + 1);
}

__device__ __forceinline__ double bsm_tail(double u)
{
#pragma clang fp contract(off)
    const double yj = u - 0.5;
    // one pair of logs per lane: the operand is chosen first
    const double q = (yj > 0) ? 1 - u : u;
    const double r = log(-log(q));
    double x = kBsmC[0] +
               r * (kBsmC[1] +
                    r * (kBsmC[2] +
                         r * (kBsmC[3] +
                              r * (kBsmC[4] +
                                   r * (kBsmC[5] + r * (kBsmC[6] + r * (kBsmC[7] + r * kBsmC[8])))))));
    if (yj < 0) x = -x;
    return x;
}

// grid: x = groups of (waves per block * 64 * PPT) points, y = dimension;
// PPT points per lane (strided by 64): each lane pays ~log2(seed) XORs for
// its starting state, so large tiles take 16 points per lane and small ones
// 4, to keep enough waves in flight (launch_sobol).
//
// The tail branch of the inverse normal (two FP64 logarithms and a 9-term
// polynomial) costs several times the central one, and about 16% of the
// values take it; a wave that branched per value would run it for nearly
// every value (some lane almost always needs it).  So each lane writes its
// central values at once and appends its tail values to a queue of its
// wave in LDS (ballot + prefix count), and the wave then runs the tail
// branch over the queue, 64 values at a time.  The queue is flushed before
// it could overflow (at most 64 entries join per point step).
template <int PPT>
__global__ void __launch_bounds__(256)
    k_sobol(const uint64_t *__restrict__ dirs, const uint64_t *__restrict__ shift, int64_t skip,
            int64_t npts, int halve, double *__restrict__ y, int64_t ldy, uint64_t *__restrict__ ints)
{
#pragma clang fp contract(off)
    constexpr int QCAP = 256;  // queue entries per wave
    __shared__ double qu[4][QCAP];
    __shared__ int qi[4][QCAP];
    const int d = blockIdx.y;
    const uint64_t *__restrict__ v = dirs + (size_t)d * kSobolCols;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + wave) * (64 * PPT);
    if (base >= npts) return;  // wave-uniform
    const int64_t i0 = base + lane;
    uint64_t s = (uint64_t)(skip + i0);
    uint64_t g = s ^ (s >> 1);
    uint64_t st = shift ? shift[d] : 0;
    while (g) {
        st ^= v[__builtin_ctzll(g)];
        g &= g - 1;
    }
    const uint64_t v5 = v[5];
    double *__restrict__ yd = y + (size_t)d * ldy;
    uint64_t *__restrict__ id = ints ? ints + (size_t)d * ldy : nullptr;
    double *__restrict__ wq = qu[wave];
    int *__restrict__ wi = qi[wave];
    int qn = 0;  // queue length (wave-uniform)
    auto flush = [&]() {
        for (int j = lane; j < qn; j += 64) {
            double x = bsm_tail(wq[j]);
            if (halve) x = x / 1.4142135623730951;  // by /= sqrt(2.0), SCP3.cpp:201
            yd[base + wi[j]] = x;
        }
        qn = 0;
    };
    for (int k = 0; k < PPT; ++k) {
        const int64_t i = i0 + 64 * k;
        const bool live = i < npts;
        const double u = (double)st * 0x1p-62;  // recipd, sobol.cpp:27780,27846
        const double yj = u - 0.5;
        const bool tail = live && !(fabs(yj) < 0.42);
        if (live && !tail) {
            double x = bsm_central(yj);
            if (halve) x = x / 1.4142135623730951;
            yd[i] = x;
        }
        if (live && id) id[i] = st;
        const uint64_t m = __builtin_amdgcn_ballot_w64(tail);
        if (tail) {
            const int pos = qn + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            wq[pos] = u;
            wi[pos] = 64 * k + lane;
        }
        qn += __builtin_popcountll(m);
        if (qn > QCAP - 64) flush();
        const int c = __builtin_ctzll(~(s >> 6));
        st ^= v5 ^ v[6 + c];
        s += 64;
    }
    flush();
}

hipError_t launch_sobol(scp3g_ctx *ctx, const uint64_t *dirs, int dims, int64_t skip,
                        int64_t npts, int halve, double *y, int64_t ldy, uint64_t *ints)
{
    if (npts <= 0) return hipSuccess;
    const int waves = 4;
    // the context's own table carries the digital shifts of a scrambled
    // sequence behind its direction integers (scp3g_set_scramble)
    const uint64_t *shift = (dirs == ctx->d_sobol && ctx->scramble_on) ? dirs + (size_t)dims * kSobolCols : nullptr;
    // 16 points per lane when that still gives >= 8 waves per SIMD (8192)
    const int64_t groups16 = (npts + waves * 64 * 16 - 1) / (waves * 64 * 16);
    if (groups16 * waves * dims >= 8192) {
        hipLaunchKernelGGL(k_sobol<16>, dim3((unsigned)groups16, (unsigned)dims), dim3(64 * waves), 0, ctx->stream,
                           dirs, shift, skip, npts, halve, y, ldy, ints);
    } else {
        const int64_t groups4 = (npts + waves * 64 * 4 - 1) / (waves * 64 * 4);
        hipLaunchKernelGGL(k_sobol<4>, dim3((unsigned)groups4, (unsigned)dims), dim3(64 * waves), 0, ctx->stream,
                           dirs, shift, skip, npts, halve, y, ldy, ints);
    }
    return hipGetLastError();
}

// ---- the allocation-free sampler (scp3g_sobol_normals_noalloc) -----------
//
// scp3g_sobol_normals stages the direction integers of its call in a buffer
// it allocates; callers that cannot allocate (graph capture, serving loops)
// get the same kernels with that staging in a fixed device array instead.
// `scratch` holds the direction integers of the call's dimensions (at most
// the table's 1111 x 62 words, 551 KiB per GPU, kept until the process
// ends); a kernel rebuilds them from the Bratley-Fox base table on every
// call, so nothing in it needs clearing.  Calls that overlap on one GPU
// race on it.
constexpr int kBaseWords = 1 + SCP3_SOBOL_INIT_MAX;  // polynomial, initial numbers
constexpr int kMaxDevices = 64;
static __device__ unsigned short sobol_base[SCP3_SOBOL_DIM_MAX * kBaseWords];
static __device__ __align__(256) uint64_t scratch[SCP3_SOBOL_DIM_MAX * kSobolCols];

struct StaticTables {
    bool ready = false;
    uint64_t *scratch = nullptr;      // device address of `scratch`
    unsigned short *base = nullptr;   // device address of `sobol_base`
};
static StaticTables g_static[kMaxDevices];
static std::mutex g_static_lock;

// Once per device, with the device current: the device addresses of the two
// arrays (the host names are stand-ins, not device addresses) and the base
// table's contents.
hipError_t sobol_static_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(g_static_lock);
    StaticTables &t = g_static[device];
    if (t.ready) return hipSuccess;
    void *ps = nullptr, *pb = nullptr;
    hipError_t e = hipGetSymbolAddress(&ps, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if ((e = hipGetSymbolAddress(&pb, HIP_SYMBOL(sobol_base))) != hipSuccess) return e;
    std::vector<unsigned short> base((size_t)SCP3_SOBOL_DIM_MAX * kBaseWords, 0);
    for (int d = 0; d < SCP3_SOBOL_DIM_MAX; ++d) {
        base[(size_t)d * kBaseWords] = scp3_sobol_poly[d];
        for (int j = 0; j < SCP3_SOBOL_INIT_MAX; ++j)
            base[(size_t)d * kBaseWords + 1 + j] = scp3_sobol_init[d][j];
    }
    if ((e = hipMemcpy(pb, base.data(), base.size() * sizeof(unsigned short), hipMemcpyHostToDevice)) != hipSuccess)
        return e;
    t.scratch = static_cast<uint64_t *>(ps);
    t.base = static_cast<unsigned short *>(pb);
    t.ready = true;
    return hipSuccess;
}

// sobol_directions on the device: one lane per dimension runs the
// Bratley-Fox recurrence (sobol.cpp:27700-27780) in its own row of v and
// then scales column j by 2^(61 - j).
__global__ void __launch_bounds__(64)
    k_sobol_dirs(const unsigned short *__restrict__ base, int dims, uint64_t *__restrict__ v)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= dims) return;
    uint64_t *row = v + (size_t)d * kSobolCols;
    if (d == 0) {
        for (int j = 0; j < kSobolCols; j++) row[j] = 1;
    } else {
        const unsigned p = base[(size_t)d * kBaseWords];
        const int m = 31 - __builtin_clz(p);
        for (int j = 0; j < m; j++) row[j] = base[(size_t)d * kBaseWords + 1 + j];
        for (int j = m; j < kSobolCols; j++) {
            uint64_t nv = row[j - m];
            for (int k = 0; k < m; k++)
                if ((p >> (m - 1 - k)) & 1u) nv ^= row[j - k - 1] << (k + 1);
            row[j] = nv;
        }
    }
    for (int j = 0; j < kSobolCols; j++) row[j] <<= (kSobolCols - 1 - j);
}

}  // namespace scp3g

extern "C" int scp3g_sobol_normals_noalloc(scp3g_ctx *ctx, int dims, int64_t skip, int64_t npts, int halve,
                                           double *y_dev, int64_t ldy, uint64_t *ints_dev)
{
    using namespace scp3g;
    if (!ctx) return -1;
    if (dims < 1 || dims > SCP3_SOBOL_DIM_MAX)
        return fail(ctx, -4, "dims must be in 1..1111 (the static direction table holds 1111 dimensions)");
    if (npts < 0 || skip < 0 || ldy < npts) return fail(ctx, -4, "bad sizes");
    if (skip + npts + 64 * 16 >= ((int64_t)1 << 60)) return fail(ctx, -4, "SEED too large");
    if (!y_dev) return fail(ctx, -4, "null output");
    const StaticTables &t = g_static[ctx->device];
    if (!t.ready) return fail(ctx, -2, "static Sobol tables not set up on this device");
    hipSetDevice(ctx->device);
    hipLaunchKernelGGL(k_sobol_dirs, dim3((unsigned)((dims + 63) / 64)), dim3(64), 0, ctx->stream, t.base, dims,
                       t.scratch);
    int rc = check_hip(ctx, hipGetLastError(), "sobol directions");
    if (rc) return rc;
    return check_hip(ctx, launch_sobol(ctx, t.scratch, dims, skip, npts, halve, y_dev, ldy, ints_dev), "sobol");
}
