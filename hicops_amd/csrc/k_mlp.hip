// k_mlp.hip -- tail-parameter surrogate: a 128-128-128-16 MLP on BF16 MFMA
// (v_mfma_f32_16x16x16_bf16, FP32 accumulate) with hand-written forward,
// backward and Adam kernels.
//
// Input per spectrum (128 features, from its score histogram): the log10
// survival function S(b) = 1 - C(b)/cpsms (C: the counts through bin b, the
// survival function of expeRT.cpp:600-640), clipped at 1e-7, at the 124
// bins b_f = f * hyp / 124 below the top score's bin hyp = (int)(hs*10 +
// 0.5), then log10(cpsms)/5, hyp/1000 and the first and last occupied bin
// below hyp over hyp.  Outputs (3 of 16 used), standardised: the plain fit's
// window ends over hyp -- the first bin at or below 22 % of the log-survival
// drop and the last at or above 87 % (expeRT.cpp:657-681) -- and the window's
// width over hyp.  The e-value is not predicted: k_sur_fit locates the
// predicted ends exactly among their 64 neighbours and fits the line through
// the window, giving the plain fit's ints, or runs the plain fit where the
// prediction misses.
//
// One wavefront handles 16 spectra (the MFMA's M); activations stay in LDS.
// MFMA operand layout (16x16x16, wave64): lane l supplies A[m = l%16][k =
// 4*(l/16) .. +3] and B[k = 4*(l/16) .. +3][n = l%16]; it receives
// D[m = 4*(l/16) + i][n = l%16], i = 0..3 (checked by hcx_selftest_mfma).
#include <cmath>
#include <cstring>
#include <vector>

#include "hcx_internal.h"

namespace hcx {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kF = 128;          // features
constexpr int kH = 128;          // hidden width
constexpr int kO = 16;           // outputs (kOut used)
constexpr int kOut = 3;
constexpr int kRows = 16;        // spectra per wavefront
constexpr int kPts = 124;        // survival-function points below hyp

// parameter block (FP32 master copy): W1[kH][kF] b1[kH] W2[kH][kH] b2[kH] W3[kO][kH] b3[kO]
constexpr int oW1 = 0, ob1 = oW1 + kH * kF, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, oW3 = ob2 + kH,
              ob3 = oW3 + kO * kH, kParams = ob3 + kO;
// BF16 copies: W1b[kH][kF] W2b[kH][kH] W3b[kO][kH] and transposes W2t[kH][kH] W3t[kH][kO]
constexpr int qW1 = 0, qW2 = qW1 + kH * kF, qW3 = qW2 + kH * kH, qW2t = qW3 + kO * kH, qW3t = qW2t + kH * kH,
              kBf = qW3t + kH * kO;

__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ bf16x4 ld4(const uint16_t *p) {
    const uint2 v = *(const uint2 *)p;   // 8-byte aligned by construction
    bf16x4 r;
    r[0] = (short)(v.x & 0xFFFFu);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xFFFFu);
    r[3] = (short)(v.y >> 16);
    return r;
}

// OUT[16][N] = IN[16][K] * W^T, IN row-major in LDS (bf16), W row-major
// [N][K] in global memory (bf16).  acc[nt] holds the 16x16 tile nt.
template <int K, int N>
__device__ __forceinline__ void wave_gemm(const uint16_t *in, const uint16_t *W, f32x4 (&acc)[N / 16]) {
    const int l = threadIdx.x & 63;
    const int m = l & 15, kq = (l >> 4) * 4;
#pragma unroll
    for (int nt = 0; nt < N / 16; nt++) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < K / 16; ks++) {
        const bf16x4 a = ld4(in + m * K + ks * 16 + kq);
#pragma unroll
        for (int nt = 0; nt < N / 16; nt++) {
            const bf16x4 b = ld4(W + (size_t)(nt * 16 + m) * K + ks * 16 + kq);
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
        }
    }
}

// G[j][k] += sum over the 16 rows r of A[r][j] * B[r][k]: A^T B with the rows
// as the MFMA's K.  A: [16][NA] f32 in LDS, B: [16][NB] bf16 in LDS.
template <int NA, int NB>
__device__ __forceinline__ void wave_outer_add(const float *A, const uint16_t *B, float *G, int jmax) {
    const int l = threadIdx.x & 63;
    const int m = l & 15, kq = (l >> 4) * 4;
#pragma unroll
    for (int jt = 0; jt < NA / 16; jt++) {
        bf16x4 a;
#pragma unroll
        for (int i = 0; i < 4; i++) a[i] = (short)f2bf(A[(kq + i) * NA + jt * 16 + m]);
#pragma unroll
        for (int kt = 0; kt < NB / 16; kt++) {
            bf16x4 b;
#pragma unroll
            for (int i = 0; i < 4; i++) b[i] = (short)B[(kq + i) * NB + kt * 16 + m];
            f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int j = jt * 16 + kq + i, k = kt * 16 + m;
                if (j < jmax) atomicAdd(&G[j * NB + k], d[i]);
            }
        }
    }
}

struct MlpW {
    const uint16_t *bf;   // kBf bf16 values
    const float *p;       // kParams f32 (biases read from here)
};

// Features and (with win, the plain fit's window ends) training targets,
// one wave per spectrum.  X: [npad][kF] bf16; T: [npad][4] f32 (t0, t1, t2,
// weight).
__global__ __launch_bounds__(256) void k_sur_features(const hcx_result *res, const uint32_t *hist, uint32_t hbins,
                                                      int32_t n, int32_t npad, uint32_t min_cpsm, const int32_t *win,
                                                      uint16_t *X, float *T) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int32_t s = blockIdx.x * 4 + w;
    if (s >= npad) return;
    uint16_t *x = X + (size_t)s * kF;
    float *t = T ? T + (size_t)s * 4 : nullptr;
    const hcx_result r = s < n ? res[s] : hcx_result{};
    const bool valid = s < n && r.cpsms >= min_cpsm && r.cpsms > 0;
    const int32_t hyp = valid ? (int32_t)((double)(r.hyperscore * 10.0f) + 0.5) : 0;
    if (!valid || hyp < 1 || hyp >= (int32_t)hbins) {
        x[lane] = 0;
        x[lane + 64] = 0;
        if (t && lane < 4) t[lane] = 0.0f;
        return;
    }
    // running counts over bins [0, hyp) into this wave's LDS row, then
    // feature f reads C(b_f) (lane f % 64)
    extern __shared__ uint32_t cs_raw[];
    uint32_t *cs = cs_raw + (size_t)w * hbins;
    const uint32_t *h = hist + (size_t)s * hbins;
    uint32_t carry = 0;
    int first = -1, last = -1;
    for (int B = 0; B < hyp; B += 64) {
        const int b = B + lane;
        const uint32_t v0 = b < hyp ? h[b] : 0u;
        const unsigned long long occ = __ballot(v0 != 0u);
        if (occ) {
            if (first < 0) first = B + (__ffsll((long long)occ) - 1);
            last = B + 63 - __clzll((long long)occ);
        }
        uint32_t v = v0;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(v, off, 64);
            if (lane >= off) v += y;
        }
        if (b < hyp) cs[b] = carry + v;
        carry += __shfl(v, 63, 64);
    }
    __syncwarp();
    const uint32_t mine0 = cs[(lane * hyp) / kPts];
    const uint32_t mine1 = lane + 64 < kPts ? cs[((lane + 64) * hyp) / kPts] : 0u;
    const double inv = 1.0 / (double)r.cpsms;
    auto feat = [&](uint32_t c) {
        const double S = 1.0 - (double)c * inv;
        return f2bf((float)log10(S > 1e-7 ? S : 1e-7));
    };
    x[lane] = feat(mine0);
    if (lane + 64 < kPts) x[lane + 64] = feat(mine1);
    if (lane < kF - kPts) {
        float v = 0.0f;
        if (lane == 0) v = (float)(log10((double)r.cpsms) / 5.0);
        if (lane == 1) v = (float)hyp / 1000.0f;
        if (lane == 2) v = (float)(first < 0 ? 0 : first) / (float)hyp;
        if (lane == 3) v = (float)(last < 0 ? 0 : last) / (float)hyp;
        x[kPts + lane] = f2bf(v);
    }
    if (t && lane == 0) {
        const int32_t lo = win[2 * s], hi = win[2 * s + 1];
        const bool ok = lo >= 0 && hi >= 0;
        t[0] = ok ? (float)lo / (float)hyp : 0.0f;
        t[1] = ok ? (float)hi / (float)hyp : 0.0f;
        t[2] = ok ? (float)(hi - lo) / (float)hyp : 0.0f;
        t[3] = ok ? 1.0f : 0.0f;
    }
}

// Forward pass of 16 rows per wave.  TRAIN keeps H1, H2 (bf16) for backward.
template <bool TRAIN>
__global__ __launch_bounds__(64) void k_mlp_fwd(const uint16_t *X, MlpW w, int32_t npad, float *Y, uint16_t *H1,
                                                uint16_t *H2) {
    __shared__ __align__(16) uint16_t xs[kRows * kF];
    __shared__ __align__(16) uint16_t h1[kRows * kH];
    __shared__ __align__(16) uint16_t h2[kRows * kH];
    const int l = threadIdx.x;
    const int32_t r0 = blockIdx.x * kRows;
    if (r0 >= npad) return;
    for (int i = l; i < kRows * kF / 4; i += 64)
        ((uint2 *)xs)[i] = ((const uint2 *)(X + (size_t)r0 * kF))[i];
    __syncthreads();
    const int n = l & 15, mq = (l >> 4) * 4;
    {
        f32x4 acc[kH / 16];
        wave_gemm<kF, kH>(xs, w.bf + qW1, acc);
#pragma unroll
        for (int nt = 0; nt < kH / 16; nt++) {
            const int c = nt * 16 + n;
            const float b = w.p[ob1 + c];
#pragma unroll
            for (int i = 0; i < 4; i++) h1[(mq + i) * kH + c] = f2bf(fmaxf(acc[nt][i] + b, 0.0f));
        }
    }
    __syncthreads();
    {
        f32x4 acc[kH / 16];
        wave_gemm<kH, kH>(h1, w.bf + qW2, acc);
#pragma unroll
        for (int nt = 0; nt < kH / 16; nt++) {
            const int c = nt * 16 + n;
            const float b = w.p[ob2 + c];
#pragma unroll
            for (int i = 0; i < 4; i++) h2[(mq + i) * kH + c] = f2bf(fmaxf(acc[nt][i] + b, 0.0f));
        }
    }
    __syncthreads();
    {
        f32x4 acc[1];
        wave_gemm<kH, kO>(h2, w.bf + qW3, acc);
        const float b = w.p[ob3 + n];
#pragma unroll
        for (int i = 0; i < 4; i++) Y[(size_t)(r0 + mq + i) * kO + n] = acc[0][i] + b;
    }
    if (TRAIN) {
        for (int i = l; i < kRows * kH / 4; i += 64) {
            ((uint2 *)(H1 + (size_t)r0 * kH))[i] = ((const uint2 *)h1)[i];
            ((uint2 *)(H2 + (size_t)r0 * kH))[i] = ((const uint2 *)h2)[i];
        }
    }
}

// Backward pass of 16 rows per wave for the loss
//   L = sum_r w_r sum_{j<3} (Y[r][j] - T[r][j])^2 * inv_wsum
// gradients accumulate (atomics) into G, laid out like the parameter block.
__global__ __launch_bounds__(64) void k_mlp_bwd(const uint16_t *X, const uint16_t *H1g, const uint16_t *H2g,
                                                const float *Y, const float *T, MlpW w, int32_t npad,
                                                float inv_wsum, float *G, float *loss) {
    __shared__ __align__(16) uint16_t xs[kRows * kF];
    __shared__ __align__(16) uint16_t h1[kRows * kH];
    __shared__ __align__(16) uint16_t h2[kRows * kH];
    __shared__ __align__(16) float dy[kRows * kO];
    __shared__ __align__(16) uint16_t dyb[kRows * kO];
    __shared__ __align__(16) float dh[kRows * kH];
    __shared__ __align__(16) uint16_t dhb[kRows * kH];
    const int l = threadIdx.x;
    const int32_t r0 = blockIdx.x * kRows;
    if (r0 >= npad) return;
    for (int i = l; i < kRows * kF / 4; i += 64) {
        ((uint2 *)xs)[i] = ((const uint2 *)(X + (size_t)r0 * kF))[i];
        ((uint2 *)h1)[i] = ((const uint2 *)(H1g + (size_t)r0 * kH))[i];
        ((uint2 *)h2)[i] = ((const uint2 *)(H2g + (size_t)r0 * kH))[i];
    }
    // dL/dY
    float lsum = 0.0f;
    for (int i = l; i < kRows * kO; i += 64) {
        const int r = i / kO, j = i % kO;
        float g = 0.0f;
        if (j < kOut) {
            const float wr = T[(size_t)(r0 + r) * 4 + 3];
            const float d = Y[(size_t)(r0 + r) * kO + j] - T[(size_t)(r0 + r) * 4 + j];
            g = 2.0f * wr * d * inv_wsum;
            lsum += wr * d * d * inv_wsum;
        }
        dy[i] = g;
        dyb[i] = f2bf(g);
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
    if (l == 0) atomicAdd(loss, lsum);
    __syncthreads();
    // layer 3: dW3 = dY^T H2, db3 = sum dY
    wave_outer_add<kO, kH>(dy, h2, G + oW3, kOut);
    if (l < kOut) {
        float s = 0.0f;
        for (int r = 0; r < kRows; r++) s += dy[r * kO + l];
        atomicAdd(&G[ob3 + l], s);
    }
    const int n = l & 15, mq = (l >> 4) * 4;
    // dH2 = dY W3, masked by relu
    {
        f32x4 acc[kH / 16];
        wave_gemm<kO, kH>(dyb, w.bf + qW3t, acc);
#pragma unroll
        for (int nt = 0; nt < kH / 16; nt++) {
            const int c = nt * 16 + n;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = mq + i;
                const float g = bf2f(h2[r * kH + c]) > 0.0f ? acc[nt][i] : 0.0f;
                dh[r * kH + c] = g;
                dhb[r * kH + c] = f2bf(g);
            }
        }
    }
    __syncthreads();
    // layer 2: dW2 = dH2^T H1, db2
    wave_outer_add<kH, kH>(dh, h1, G + oW2, kH);
    for (int c = l; c < kH; c += 64) {
        float s = 0.0f;
        for (int r = 0; r < kRows; r++) s += dh[r * kH + c];
        atomicAdd(&G[ob2 + c], s);
    }
    // dH1 = dH2 W2, masked by relu
    f32x4 acc[kH / 16];
    wave_gemm<kH, kH>(dhb, w.bf + qW2t, acc);
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < kH / 16; nt++) {
        const int c = nt * 16 + n;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = mq + i;
            dh[r * kH + c] = bf2f(h1[r * kH + c]) > 0.0f ? acc[nt][i] : 0.0f;
        }
    }
    __syncthreads();
    // layer 1: dW1 = dH1^T X, db1
    wave_outer_add<kH, kF>(dh, xs, G + oW1, kH);
    for (int c = l; c < kH; c += 64) {
        float s = 0.0f;
        for (int r = 0; r < kRows; r++) s += dh[r * kH + c];
        atomicAdd(&G[ob1 + c], s);
    }
}

// Adam step on the FP32 master copy (update == 0: cast only), then the BF16
// copies and transposes; G is cleared for the next step.
__global__ void k_mlp_adam(float *P, float *G, float *M, float *V, uint16_t *bf, float lr, float b1, float b2,
                           float eps, float c1, float c2, int update) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    float p = P[i];
    if (update) {
        const float g = G[i];
        const float m = b1 * M[i] + (1.0f - b1) * g;
        const float v = b2 * V[i] + (1.0f - b2) * g * g;
        M[i] = m;
        V[i] = v;
        p -= lr * (m / c1) / (sqrtf(v / c2) + eps);
        P[i] = p;
    }
    G[i] = 0.0f;
    const uint16_t q = f2bf(p);
    if (i >= oW1 && i < ob1) {
        bf[qW1 + (i - oW1)] = q;
    } else if (i >= oW2 && i < ob2) {
        const int o = (i - oW2) / kH, k = (i - oW2) % kH;
        bf[qW2 + o * kH + k] = q;
        bf[qW2t + k * kH + o] = q;
    } else if (i >= oW3 && i < ob3) {
        const int o = (i - oW3) / kH, k = (i - oW3) % kH;
        bf[qW3 + o * kH + k] = q;
        bf[qW3t + k * kO + o] = q;
    }
}

// MFMA lane-layout self-check: D = A B for 16x16 operands (row-major f32,
// rounded to bf16), through wave_gemm (B supplied transposed, as W).
__global__ void k_selftest_mfma(const float *A, const float *B, float *D) {
    __shared__ __align__(16) uint16_t a[16 * 16];
    __shared__ __align__(16) uint16_t bt[16 * 16];
    const int l = threadIdx.x;
    for (int i = l; i < 256; i += 64) {
        a[i] = f2bf(A[i]);
        const int r = i / 16, c = i % 16;
        bt[c * 16 + r] = f2bf(B[i]);   // W[n][k] = B[k][n]
    }
    __syncthreads();
    f32x4 acc[1];
    wave_gemm<16, 16>(a, bt, acc);
    const int n = l & 15, mq = (l >> 4) * 4;
    for (int i = 0; i < 4; i++) D[(mq + i) * 16 + n] = acc[0][i];
}

}  // namespace

// ---- host side -------------------------------------------------------------

uint32_t mlp_param_count() { return kParams; }

// device layout of ctx->sur: P f32 | G f32 | M f32 | V f32 | norm f32[8] | bf u16
static size_t sur_bytes() { return 4ull * (4 * (size_t)kParams + 8) + 2ull * kBf + 64; }
static float *surP(hcx_ctx *c) { return (float *)c->sur.p; }
static float *surG(hcx_ctx *c) { return surP(c) + kParams; }
static float *surM(hcx_ctx *c) { return surP(c) + 2 * kParams; }
static float *surV(hcx_ctx *c) { return surP(c) + 3 * kParams; }
static float *surN(hcx_ctx *c) { return surP(c) + 4 * kParams; }
static uint16_t *surB(hcx_ctx *c) { return (uint16_t *)(surN(c) + 8); }

int mlp_alloc(hcx_ctx *c) {
    if (c->sur.p) return HCX_SUCCESS;
    if (c->sur.ensure(sur_bytes())) return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "surrogate: allocation failed");
    HCX_HIP(hipMemsetAsync(c->sur.p, 0, sur_bytes(), c->stream));
    return HCX_SUCCESS;
}

int mlp_cast(hcx_ctx *c) {
    k_mlp_adam<<<(kParams + 255) / 256, 256, 0, c->stream>>>(surP(c), surG(c), surM(c), surV(c), surB(c), 0.f, 0.f,
                                                            0.f, 1.f, 1.f, 1.f, 0);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

int mlp_set(hcx_ctx *c, const float *params, const float *norm6) {
    int rc = mlp_alloc(c);
    if (rc) return rc;
    std::vector<float> nrm(8, 0.0f);
    for (int i = 0; i < 6; i++) nrm[i] = norm6[i];
    HCX_HIP(hipMemcpyAsync(surP(c), params, 4ull * kParams, hipMemcpyHostToDevice, c->stream));
    HCX_HIP(hipMemcpyAsync(surN(c), nrm.data(), 32, hipMemcpyHostToDevice, c->stream));
    HCX_HIP(hipMemsetAsync(surM(c), 0, 8ull * kParams, c->stream));   // Adam moments
    rc = mlp_cast(c);
    if (rc) return rc;
    HCX_HIP(hipStreamSynchronize(c->stream));
    c->sur_ready = true;
    c->sur_step = 0;
    return HCX_SUCCESS;
}

int mlp_get(hcx_ctx *c, float *params, float *norm6) {
    if (!c->sur_ready) return set_error_code(HCX_ERR_INVLD_PARAM, "surrogate: not initialised");
    float nrm[8];
    HCX_HIP(hipMemcpyAsync(params, surP(c), 4ull * kParams, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipMemcpyAsync(nrm, surN(c), 32, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < 6; i++) norm6[i] = nrm[i];
    return HCX_SUCCESS;
}

// features (+ targets) of a searched batch into the context's scratch:
// X [npad][kF] bf16 | T [npad][4] f32 | Y [npad][kO] f32 | [npad][kH] f32 | win [npad][2] i32
static int features(hcx_ctx *c, hcx_qbatch *qb, bool targets, int32_t *npad_out) {
    const int32_t n = qb->nspec, npad = (n + kRows - 1) / kRows * kRows;
    const size_t need = 2ull * npad * kF + 4ull * npad * 4 + 4ull * npad * kO + 4ull * npad * kH + 8ull * npad + 256;
    if (c->sur_tmp.ensure(need)) return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "surrogate: scratch allocation failed");
    uint16_t *X = (uint16_t *)c->sur_tmp.p;
    float *T = (float *)(X + (size_t)npad * kF);
    int32_t *win = (int32_t *)(T + (size_t)npad * 4 + (size_t)npad * kO + (size_t)npad * kH);
    if (targets) {
        int rc = launch_fit_windows(c, (const hcx_result *)qb->res.p, (const uint32_t *)qb->hist.p, c->hbins, n,
                                    c->params.min_cpsm, win);
        if (rc) return rc;
    }
    const size_t sm = 4ull * 4 * c->hbins;   // four waves' running counts
    HCX_HIP(hipFuncSetAttribute((const void *)k_sur_features, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    k_sur_features<<<(npad + 3) / 4, 256, sm, c->stream>>>((const hcx_result *)qb->res.p, (const uint32_t *)qb->hist.p,
                                                         c->hbins, n, npad, c->params.min_cpsm, targets ? win : nullptr,
                                                         X, targets ? T : nullptr);
    HCX_HIP(hipGetLastError());
    *npad_out = npad;
    return HCX_SUCCESS;
}

int mlp_infer(hcx_ctx *c, hcx_qbatch *qb, uint32_t *fix, uint32_t *cnt) {
    if (!c->sur_ready) return set_error_code(HCX_ERR_INVLD_PARAM, "surrogate: not trained");
    int32_t npad = 0;
    int rc = features(c, qb, false, &npad);
    if (rc) return rc;
    uint16_t *X = (uint16_t *)c->sur_tmp.p;
    float *Y = (float *)(X + (size_t)npad * kF) + (size_t)npad * 4;
    const MlpW w = {surB(c), surP(c)};
    k_mlp_fwd<false><<<npad / kRows, 64, 0, c->stream>>>(X, w, npad, Y, nullptr, nullptr);
    HCX_HIP(hipGetLastError());
    return launch_sur_fit(c, Y, kO, surN(c), (hcx_result *)qb->res.p, (const uint32_t *)qb->hist.p, c->hbins,
                          qb->nspec, c->params.min_cpsm, c->params.expect_max, fix, c->force_host_fit, cnt);
}

// one forward + backward over rows [0, npad) of X/T in scratch; G += grads
static int fwd_bwd(hcx_ctx *c, const uint16_t *X, const float *T, int32_t rows, float inv_wsum, float *loss_dev) {
    float *Y = c->mlp_Y;
    uint16_t *H1 = c->mlp_H1, *H2 = c->mlp_H2;
    const MlpW w = {surB(c), surP(c)};
    k_mlp_fwd<true><<<rows / kRows, 64, 0, c->stream>>>(X, w, rows, Y, H1, H2);
    HCX_HIP(hipGetLastError());
    k_mlp_bwd<<<rows / kRows, 64, 0, c->stream>>>(X, H1, H2, Y, T, w, rows, inv_wsum, surG(c), loss_dev);
    HCX_HIP(hipGetLastError());
    return HCX_SUCCESS;
}

// training buffers sized for `rows` rows (Y, H1, H2 and the loss scalar)
static int train_bufs(hcx_ctx *c, int32_t rows) {
    const size_t need = 4ull * rows * kO + 2ull * 2 * rows * kH + 64;
    if (c->mlp_buf.ensure(need)) return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "surrogate: training allocation failed");
    c->mlp_loss = (float *)c->mlp_buf.p;
    c->mlp_Y = c->mlp_loss + 16;
    c->mlp_H1 = (uint16_t *)(c->mlp_Y + (size_t)rows * kO);
    c->mlp_H2 = c->mlp_H1 + (size_t)rows * kH;
    return HCX_SUCCESS;
}

int mlp_gradient(hcx_ctx *c, const float *Xh, const float *Th, uint32_t n, float *grad, float *loss) {
    if (!c->sur_ready) return set_error_code(HCX_ERR_INVLD_PARAM, "surrogate: not initialised");
    if (n == 0 || n % kRows) return set_error_code(HCX_ERR_INVLD_SIZE, "surrogate: rows must be a multiple of 16");
    int rc = train_bufs(c, (int32_t)n);
    if (rc) return rc;
    if (c->sur_tmp.ensure(2ull * n * kF + 16ull * n + 256))
        return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "surrogate: scratch allocation failed");
    std::vector<uint16_t> xb((size_t)n * kF);
    for (size_t i = 0; i < xb.size(); i++) {
        uint32_t u;
        std::memcpy(&u, &Xh[i], 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        xb[i] = (uint16_t)(u >> 16);
    }
    uint16_t *X = (uint16_t *)c->sur_tmp.p;
    float *T = (float *)(X + (size_t)n * kF);
    HCX_HIP(hipMemcpyAsync(X, xb.data(), 2ull * n * kF, hipMemcpyHostToDevice, c->stream));
    HCX_HIP(hipMemcpyAsync(T, Th, 16ull * n, hipMemcpyHostToDevice, c->stream));
    HCX_HIP(hipMemsetAsync(c->mlp_loss, 0, 4, c->stream));
    rc = mlp_cast(c);   // clears G
    if (rc) return rc;
    double ws = 0.0;
    for (uint32_t r = 0; r < n; r++) ws += Th[4 * r + 3];
    rc = fwd_bwd(c, X, T, (int32_t)n, ws > 0 ? (float)(1.0 / ws) : 0.0f, c->mlp_loss);
    if (rc) return rc;
    HCX_HIP(hipMemcpyAsync(grad, surG(c), 4ull * kParams, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipMemcpyAsync(loss, c->mlp_loss, 4, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipStreamSynchronize(c->stream));
    return HCX_SUCCESS;
}

// Train on a batch searched with the plain fit: targets are its window ends
// and width over hyp.  Mini-batches of `batch` rows (multiple of 16), `steps`
// Adam steps at learning rate lr.  *loss_out: mean loss of the last 10% of steps.
int mlp_train(hcx_ctx *c, hcx_qbatch *qb, uint32_t steps, uint32_t batch, float lr, uint64_t seed, float *loss_out) {
    if (!c->sur_ready) return set_error_code(HCX_ERR_INVLD_PARAM, "surrogate: not initialised");
    if (!qb->res.p || !qb->hist.p) return set_error_code(HCX_ERR_INVLD_PARAM, "surrogate: batch has no search results");
    int32_t npad = 0;
    int rc = features(c, qb, true, &npad);
    if (rc) return rc;
    uint16_t *X = (uint16_t *)c->sur_tmp.p;
    float *T = (float *)(X + (size_t)npad * kF);
    // standardise the targets over the rows with weight 1 (host side, small)
    std::vector<float> th((size_t)npad * 4);
    HCX_HIP(hipMemcpyAsync(th.data(), T, 16ull * npad, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipStreamSynchronize(c->stream));
    double sum[3] = {0, 0, 0}, sq[3] = {0, 0, 0}, wn = 0;
    for (int32_t r = 0; r < npad; r++) {
        if (th[4 * r + 3] <= 0.0f) continue;
        wn += 1;
        for (int j = 0; j < 3; j++) sum[j] += th[4 * r + j], sq[j] += (double)th[4 * r + j] * th[4 * r + j];
    }
    if (wn < 16) return set_error_code(HCX_ERR_NOT_ENOUGH_DATA, "surrogate: fewer than 16 fitted spectra");
    float nrm[8] = {0};
    for (int j = 0; j < 3; j++) {
        const double m = sum[j] / wn, v = sq[j] / wn - m * m;
        nrm[j] = (float)m;
        nrm[3 + j] = (float)(v > 1e-12 ? std::sqrt(v) : 1.0);
    }
    for (int32_t r = 0; r < npad; r++)
        for (int j = 0; j < 3; j++) th[4 * r + j] = (th[4 * r + j] - nrm[j]) / nrm[3 + j];
    HCX_HIP(hipMemcpyAsync(T, th.data(), 16ull * npad, hipMemcpyHostToDevice, c->stream));
    HCX_HIP(hipMemcpyAsync(surN(c), nrm, 32, hipMemcpyHostToDevice, c->stream));
    const int32_t B = (int32_t)std::min<uint32_t>(std::max<uint32_t>(16, batch / kRows * kRows), (uint32_t)npad);
    rc = train_bufs(c, B);
    if (rc) return rc;
    // per mini-batch weight sums: contiguous windows starting at multiples of 16
    std::vector<double> pre((size_t)npad + 1, 0.0);
    for (int32_t r = 0; r < npad; r++) pre[r + 1] = pre[r] + th[4 * r + 3];
    uint64_t st = seed ? seed : 0x5EED;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    HCX_HIP(hipMemsetAsync(c->mlp_loss, 0, 4, c->stream));
    rc = mlp_cast(c);   // clears G
    if (rc) return rc;
    const uint32_t tail0 = steps - std::max<uint32_t>(1, steps / 10);
    const float b1 = 0.9f, b2 = 0.999f;
    for (uint32_t it = 0; it < steps; it++) {
        const int32_t nwin = (npad - B) / kRows + 1;
        const int32_t r0 = (int32_t)(next() % (uint64_t)nwin) * kRows;
        const double ws = pre[r0 + B] - pre[r0];
        if (it == tail0) HCX_HIP(hipMemsetAsync(c->mlp_loss, 0, 4, c->stream));
        rc = fwd_bwd(c, X + (size_t)r0 * kF, T + (size_t)r0 * 4, B, ws > 0 ? (float)(1.0 / ws) : 0.0f, c->mlp_loss);
        if (rc) return rc;
        c->sur_step++;
        const float c1 = 1.0f - std::pow(b1, (float)c->sur_step), c2 = 1.0f - std::pow(b2, (float)c->sur_step);
        k_mlp_adam<<<(kParams + 255) / 256, 256, 0, c->stream>>>(surP(c), surG(c), surM(c), surV(c), surB(c), lr, b1,
                                                                b2, 1e-8f, c1, c2, 1);
        HCX_HIP(hipGetLastError());
    }
    float lsum = 0.0f;
    HCX_HIP(hipMemcpyAsync(&lsum, c->mlp_loss, 4, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipStreamSynchronize(c->stream));
    if (loss_out) *loss_out = lsum / (float)(steps - tail0);
    return HCX_SUCCESS;
}

int selftest_mfma(hcx_ctx *c, const float *A, const float *B, float *D) {
    DeviceBuf buf;
    if (buf.ensure(3 * 256 * 4)) return set_error_code(HCX_ERR_BAD_MEM_ALLOC, "selftest: allocation failed");
    float *a = (float *)buf.p, *b = a + 256, *d = b + 256;
    HCX_HIP(hipMemcpyAsync(a, A, 1024, hipMemcpyHostToDevice, c->stream));
    HCX_HIP(hipMemcpyAsync(b, B, 1024, hipMemcpyHostToDevice, c->stream));
    k_selftest_mfma<<<1, 64, 0, c->stream>>>(a, b, d);
    HCX_HIP(hipGetLastError());
    HCX_HIP(hipMemcpyAsync(D, d, 1024, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipStreamSynchronize(c->stream));
    return HCX_SUCCESS;
}

}  // namespace hcx
