// kifi_learn.hip -- learned ICP initialisation: a small BF16 CNN that predicts the
// rigid twist between two depth frames, trained on the GPU with hand-written
// forward and backward kernels on CDNA4 matrix cores.
//
// No reference counterpart (the reference's ICP starts from the previous pose,
// Pipeline.cpp:37-39).  Network: input 2 x 106 x 128 (previous and current
// frame, 4x4-block mean depth / 4 m), five 3x3 stride-2 convolutions with
// padding 1 (16-32-64-64-128 channels, ReLU), fully connected 2048 -> 256 (ReLU)
// -> 6.  Weights and activations are bfloat16, accumulation is binary32
// (v_mfma_f32_16x16x32_bf16), master weights and Adam moments binary32.
//
// Every layer is one implicit-GEMM kernel (forward, backward-data and
// backward-weight), 64 x 64 output tile per workgroup, 32 x 32 per wave as 2 x 2
// MFMA 16x16x32 tiles fed straight from the gathered operands; weight gradients
// use a fixed split of the reduction dimension and a fixed-order sum, so a
// training run is deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "kifi_icp.hpp"
#include "kifi_internal.hpp"
#include "kifi_synth.h"

namespace kifi {
kifi_status model_from_depth(kifi_model *m, const uint16_t *d_mm, const float e2w[16], const float w2e[16]);
void icp_set_pose(IcpWork *w, const double *e2w, const float *w2e, hipStream_t st);
} // namespace kifi

using namespace kifi;

namespace {

typedef short bf16x8 __attribute__((ext_vector_type(8)));      // bf16 bits
typedef __bf16 mfma_bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// SplitMix64 (Steele, Lea, Flood 2014), the synthetic frames' generator
inline uint64_t splitmix_host(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

constexpr int kInW = 128, kInH = 106, kInC = 2;
constexpr int kNumConv = 5;
constexpr int kConvC[kNumConv + 1] = { 2, 16, 32, 64, 64, 128 };
constexpr int kFc1 = 256, kOut = 6;
constexpr int kMaxBatch = 256;
// twist normalisation: network output x scale = (omega [rad], v [m])
constexpr float kTwistScale[6] = { 0.15f, 0.15f, 0.15f, 0.06f, 0.06f, 0.06f };

__host__ __device__ inline int conv_out(int n) { return (n - 1) / 2 + 1; }   // k 3, stride 2, pad 1

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);     // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

uint16_t f2bf_host(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// ------------------------------------------------------------------ GEMM core
// C[M x N] (+)= A[M x K] B[K x N]; Op supplies a(m, k), b(k, n) as bf16 bits (0
// outside the matrix) and store(m, n, value, split).  blockIdx.z splits K into
// chunks of op.kchunk (a multiple of 32).  Operand layout of
// v_mfma_f32_16x16x32_bf16: lane l holds row (A) / column (B) l % 16 at eight
// consecutive k from 8 (l / 16); A and B share the k assignment.
template <class Op>
__global__ __launch_bounds__(256) void k_gemm(Op op)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m0 = blockIdx.x * 64 + (wave >> 1) * 32, n0 = blockIdx.y * 64 + (wave & 1) * 32;
    const int kb = blockIdx.z * op.kchunk, ke = min(op.K, kb + op.kchunk);
    const int r = lane & 15, kq = 8 * (lane >> 4);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            acc[i][j] = f32x4{ 0.0f, 0.0f, 0.0f, 0.0f };
    for (int k = kb; k < ke; k += 32) {
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            op.a8(m0 + 16 * i + r, k + kq, ke, a[i]);
            op.b8(k + kq, n0 + 16 * i + r, ke, b[i]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mfma_bf16x8, a[i]),
                                                                    __builtin_bit_cast(mfma_bf16x8, b[j]),
                                                                    acc[i][j], 0, 0, 0);
    }
    // D layout of v_mfma_f32_16x16x32_bf16: lane holds rows 4*(lane/16)+e of column lane%16
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + 16 * i + 4 * (lane >> 4) + e, n = n0 + 16 * j + r;
                if (m < op.M && n < op.N)
                    op.store(m, n, acc[i][j][e], blockIdx.z);
            }
}

template <class Op>
void launch_gemm(const Op &op, int splits, hipStream_t st)
{
    dim3 grid((op.M + 63) / 64, (op.N + 63) / 64, splits);
    hipLaunchKernelGGL(k_gemm<Op>, grid, dim3(256), 0, st, op);
}

// Eight consecutive k (k0 .. k0+7) of an operand row or column, 0 at k >= ke.
// The default takes them one element at a time; an operator whose element
// index needs runtime divisions replaces it with one decomposition per call.
template <class D>
struct GemmElems {
    __device__ void a8(int m, int k0, int ke, bf16x8 &o) const
    {
        const D &d = static_cast<const D &>(*this);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            o[j] = (short)(k0 + j < ke ? d.a(m, k0 + j) : 0);
    }
    __device__ void b8(int k0, int n, int ke, bf16x8 &o) const
    {
        const D &d = static_cast<const D &>(*this);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            o[j] = (short)(k0 + j < ke ? d.b(k0 + j, n) : 0);
    }
};

// Split of a reduction dimension K into `splits` chunks of multiples of 32.
inline int chunk_of(int K, int splits) { return ((K + splits - 1) / splits + 31) / 32 * 32; }
inline int chunk_all(int K) { return (K + 31) / 32 * 32; }

struct ConvGeom {
    int B, Ci, Hi, Wi, Co, Ho, Wo;
};

// forward: M = B*Ho*Wo, N = Co, K = Ci*9; epilogue bias + ReLU -> bf16
struct OpConvFwd : GemmElems<OpConvFwd> {
    ConvGeom g;
    const uint16_t *X, *W;
    const float *bias;
    uint16_t *Y;
    int M, N, K, kchunk;
    __device__ uint16_t a(int m, int k) const
    {
        if (m >= M)
            return 0;
        const int hw = g.Ho * g.Wo, b = m / hw, p = m - b * hw, ho = p / g.Wo, wo = p - ho * g.Wo;
        const int ci = k / 9, t = k - ci * 9, ky = t / 3, kx = t - ky * 3;
        const int hi = 2 * ho - 1 + ky, wi = 2 * wo - 1 + kx;
        if (hi < 0 || hi >= g.Hi || wi < 0 || wi >= g.Wi)
            return 0;
        return X[((size_t)(b * g.Ci + ci) * g.Hi + hi) * g.Wi + wi];
    }
    __device__ uint16_t b(int k, int n) const { return n < N ? W[(size_t)n * K + k] : 0; }
    __device__ void store(int m, int n, float v, int) const
    {
        const int hw = g.Ho * g.Wo, b = m / hw, p = m - b * hw;
        v += bias[n];
        Y[((size_t)(b * g.Co + n)) * hw + p] = f2bf(v > 0.0f ? v : 0.0f);
    }
};

// backward data: dX[b][ci][hi][wi] = sum over (co, ky, kx) with hi = 2 ho - 1 + ky
// of dY[b][co][ho][wo] W[co][ci][ky][kx], then x ReLU'(X)
struct OpConvBwdData : GemmElems<OpConvBwdData> {
    ConvGeom g;
    const uint16_t *dY, *W, *X;
    uint16_t *dX;
    int M, N, K, kchunk;
    __device__ uint16_t a(int m, int k) const
    {
        if (m >= M)
            return 0;
        const int hw = g.Hi * g.Wi, b = m / hw, p = m - b * hw, hi = p / g.Wi, wi = p - hi * g.Wi;
        const int co = k / 9, t = k - co * 9, ky = t / 3, kx = t - ky * 3;
        const int hh = hi + 1 - ky, ww = wi + 1 - kx;
        if (hh < 0 || ww < 0 || (hh & 1) || (ww & 1))
            return 0;
        const int ho = hh >> 1, wo = ww >> 1;
        if (ho >= g.Ho || wo >= g.Wo)
            return 0;
        return dY[((size_t)(b * g.Co + co) * g.Ho + ho) * g.Wo + wo];
    }
    __device__ uint16_t b(int k, int n) const
    {
        if (n >= N)
            return 0;
        const int co = k / 9, t = k - co * 9;
        return W[((size_t)co * g.Ci + n) * 9 + t];
    }
    __device__ void store(int m, int n, float v, int) const
    {
        const int hw = g.Hi * g.Wi, b = m / hw, p = m - b * hw;
        const size_t o = ((size_t)(b * g.Ci + n)) * hw + p;
        dX[o] = bf2f(X[o]) > 0.0f ? f2bf(v) : (uint16_t)0;
    }
};

// backward weights: dW[co][ci*9+t] = sum over (b, ho, wo) dY[b][co][ho][wo] X[b][ci][2ho-1+ky][2wo-1+kx]
struct OpConvBwdW : GemmElems<OpConvBwdW> {
    ConvGeom g;
    const uint16_t *dY, *X;
    float *part;     // [splits][M][N]
    int M, N, K, kchunk;
    __device__ uint16_t a(int m, int k) const
    {
        if (m >= M)
            return 0;
        const int hw = g.Ho * g.Wo, b = k / hw, p = k - b * hw;
        return dY[((size_t)(b * g.Co + m)) * hw + p];
    }
    __device__ uint16_t b(int k, int n) const
    {
        if (n >= N)
            return 0;
        const int hw = g.Ho * g.Wo, b = k / hw, p = k - b * hw, ho = p / g.Wo, wo = p - ho * g.Wo;
        const int ci = n / 9, t = n - ci * 9, ky = t / 3, kx = t - ky * 3;
        const int hi = 2 * ho - 1 + ky, wi = 2 * wo - 1 + kx;
        if (hi < 0 || hi >= g.Hi || wi < 0 || wi >= g.Wi)
            return 0;
        return X[((size_t)(b * g.Ci + ci) * g.Hi + hi) * g.Wi + wi];
    }
    // Eight k of a row / column: k = (b, ho, wo) is decomposed once and stepped,
    // and every load is issued (indices clamped into the tensors) before the
    // masks apply.  Same values as a() / b().
    __device__ void a8(int m, int k0, int ke, bf16x8 &o) const
    {
        const int hw = g.Ho * g.Wo, mm = min(m, M - 1);
        int b = k0 / hw, p = k0 - b * hw;
        uint16_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const size_t idx = k0 + j < K ? ((size_t)(b * g.Co + mm)) * hw + p : 0;
            v[j] = dY[idx];
            if (++p == hw) {
                p = 0;
                ++b;
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            o[j] = (short)(m < M && k0 + j < ke ? v[j] : 0);
    }
    __device__ void b8(int k0, int n, int ke, bf16x8 &o) const
    {
        const int hw = g.Ho * g.Wo, nn = min(n, N - 1);
        const int ci = nn / 9, t = nn - ci * 9, ky = t / 3, kx = t - ky * 3;
        int b = k0 / hw, p = k0 - b * hw, ho = p / g.Wo, wo = p - ho * g.Wo;
        uint16_t v[8];
        bool ok[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int hi = 2 * ho - 1 + ky, wi = 2 * wo - 1 + kx;
            ok[j] = n < N && k0 + j < ke && hi >= 0 && hi < g.Hi && wi >= 0 && wi < g.Wi;
            const int bc = min(b, g.B - 1), hc = min(max(hi, 0), g.Hi - 1), wc = min(max(wi, 0), g.Wi - 1);
            v[j] = X[((size_t)(bc * g.Ci + ci) * g.Hi + hc) * g.Wi + wc];
            if (++wo == g.Wo) {
                wo = 0;
                if (++ho == g.Ho) {
                    ho = 0;
                    ++b;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            o[j] = (short)(ok[j] ? v[j] : 0);
    }
    __device__ void store(int m, int n, float v, int z) const { part[((size_t)z * M + m) * N + n] = v; }
};

// fully connected forward: Y[b][n] = sum_k X[b][k] W[n][k] + bias[n] (ReLU -> bf16, or f32 out)
struct OpFcFwd : GemmElems<OpFcFwd> {
    const uint16_t *X, *W;
    const float *bias;
    uint16_t *Y;       // bf16 + ReLU when set
    float *Yf;         // f32, no activation, when set
    int M, N, K, kchunk;
    __device__ uint16_t a(int m, int k) const { return m < M ? X[(size_t)m * K + k] : 0; }
    __device__ uint16_t b(int k, int n) const { return n < N ? W[(size_t)n * K + k] : 0; }
    __device__ void store(int m, int n, float v, int) const
    {
        v += bias[n];
        if (Yf)
            Yf[(size_t)m * N + n] = v;
        else
            Y[(size_t)m * N + n] = f2bf(v > 0.0f ? v : 0.0f);
    }
};

// fully connected backward data: dX[b][k] = sum_n dY[b][n] W[n][k], x ReLU'(X)
struct OpFcBwdData : GemmElems<OpFcBwdData> {
    const uint16_t *dY, *W, *X;
    uint16_t *dX;
    int M, N, K, kchunk;    // M = B, N = in features, K = out features
    __device__ uint16_t a(int m, int k) const { return m < M ? dY[(size_t)m * K + k] : 0; }
    __device__ uint16_t b(int k, int n) const { return n < N ? W[(size_t)k * N + n] : 0; }
    __device__ void store(int m, int n, float v, int) const
    {
        const size_t o = (size_t)m * N + n;
        dX[o] = bf2f(X[o]) > 0.0f ? f2bf(v) : (uint16_t)0;
    }
};

// fully connected backward weights: dW[n][k] = sum_b dY[b][n] X[b][k]
struct OpFcBwdW : GemmElems<OpFcBwdW> {
    const uint16_t *dY, *X;
    float *part;
    int M, N, K, kchunk;    // M = out features, N = in features, K = B
    int ldy;
    __device__ uint16_t a(int m, int k) const { return m < M ? dY[(size_t)k * ldy + m] : 0; }
    __device__ uint16_t b(int k, int n) const { return n < N ? X[(size_t)k * N + n] : 0; }
    __device__ void store(int m, int n, float v, int z) const { part[((size_t)z * M + m) * N + n] = v; }
};

// Fully connected forward for a few rows (the pipeline's single-sample
// prediction): one wave per output, the K products strided over the lanes in
// 16-byte loads and summed in binary32 by a wave reduction.  The MFMA tile path
// would run a handful of workgroups along the whole of K for M = 1.
template <bool OUT_F32>
__global__ __launch_bounds__(256) void k_fc_small(const uint16_t *__restrict__ X, const uint16_t *__restrict__ W,
                                                  const float *__restrict__ bias, int M, int N, int K, int vec,
                                                  uint16_t *__restrict__ Y, float *__restrict__ Yf)
{
    const int wave = (int)((blockIdx.x * 256u + threadIdx.x) / 64u), lane = threadIdx.x % 64;
    if (wave >= M * N)
        return;                                   // whole waves: the index is uniform per wave
    const int m = wave / N, n = wave - m * N;
    const uint16_t *x = X + (size_t)m * K, *w = W + (size_t)n * K;
    float s = 0.0f;
    if (vec) {                                    // K % 8 == 0 and 16-byte aligned rows
        for (int k = lane * 8; k < K; k += 64 * 8) {
            const uint4 xa = *reinterpret_cast<const uint4 *>(x + k), wa = *reinterpret_cast<const uint4 *>(w + k);
            const uint32_t xs[4] = { xa.x, xa.y, xa.z, xa.w }, ws[4] = { wa.x, wa.y, wa.z, wa.w };
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                s += bf2f((uint16_t)(xs[j] & 0xffffu)) * bf2f((uint16_t)(ws[j] & 0xffffu));
                s += bf2f((uint16_t)(xs[j] >> 16)) * bf2f((uint16_t)(ws[j] >> 16));
            }
        }
    } else {
        for (int k = lane; k < K; k += 64)
            s += bf2f(x[k]) * bf2f(w[k]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        s += __shfl_xor(s, o, 64);
    if (lane == 0) {
        const float v = s + bias[n];
        if (OUT_F32)
            Yf[(size_t)m * N + n] = v;
        else
            Y[(size_t)m * N + n] = f2bf(v > 0.0f ? v : 0.0f);
    }
}

// Direct 3x3 / stride-2 / pad-1 convolution for a few inputs (the per-frame
// prediction), bias + ReLU -> bf16 as the GEMM path stores it.  Workgroup =
// (band of R output rows, output channel co, input b): the band's input rows
// (every channel; zero rows outside the image and a zero column left of it)
// and the channel's weights are staged in LDS, every global load of a round
// (NX 16-byte input chunks and NW weight loads per thread, clamped so that none
// is conditional) issued before the first LDS store, so a layer waits for one
// memory round trip.  Thread (pixel p, slice s) then sums the input channels
// s, s + nks, ... of its pixel in binary32, and the nks slices of a pixel are
// added in slice order through LDS.  One channel per workgroup and K split over
// threads keep 64..224 workgroups and at most 18 x Ci / nks products per
// thread; the implicit-GEMM tiles would run M = Ho Wo (16 .. 3392 here) through
// a few workgroups walking the whole of K.  Needs Wi % 8 == 0 (so Wo = Wi / 2
// and column 2 Wo - 1 is in the image) and 16-byte aligned input rows; WV
// (16-byte weight chunks) also K % 8 == 0 and 16-byte aligned weights.  The
// host checks.
struct ConvSmall {
    ConvGeom g;
    const uint16_t *X, *W;
    const float *bias;
    uint16_t *Y;
    int R, nks;
};

constexpr int kXPad = 8;   // LDS row: 8 zero elements (the last is column -1), then Wi

__host__ __device__ constexpr int conv_small_kp(int K, bool wv) { return wv ? K + 8 : K + 2; }

template <bool WV>
__global__ __launch_bounds__(256) void k_conv_small(ConvSmall c)
{
    constexpr int NX = 12, NW = 1;
    extern __shared__ __align__(16) uint16_t smem[];
    const ConvGeom &g = c.g;
    const int r0 = blockIdx.x * c.R, co = blockIdx.y, b = blockIdx.z;
    const int rows_out = min(c.R, g.Ho - r0), rows_in = 2 * rows_out + 1, hi0 = 2 * r0 - 1;
    const int Wp = g.Wi + kXPad, K = g.Ci * 9;
    uint16_t *sx = smem;                                  // [Ci][rows_in][Wp]
    uint16_t *sw = smem + (size_t)g.Ci * rows_in * Wp;    // [Kp]
    float *part = reinterpret_cast<float *>(sw + conv_small_kp(K, WV));   // [nks][R Wo]
    const int cpr = Wp / 8, nch = g.Ci * rows_in * cpr;   // 16-byte input chunks; chunk 0 of a row is the pad
    const int nwl = WV ? K / 8 : K;                       // weight loads: chunks or elements
    const uint16_t *wsrc = c.W + (size_t)co * K;
    for (int r = 0; r * NX * 256 < nch || r * NW * 256 < nwl; ++r) {
        uint4 vx[NX];
        uint32_t inside = 0;   // the zero of a pad chunk is applied at the store: writing it into a
                               // load's destination register would wait for that load
#pragma unroll
        for (int u = 0; u < NX; ++u) {
            const int i = min((r * NX + u) * 256 + (int)threadIdx.x, nch - 1);
            const int row = i / cpr, ch = i - row * cpr, ci = row / rows_in, hi = hi0 + (row - ci * rows_in);
            const int hic = min(max(hi, 0), g.Hi - 1), chc = max(ch - 1, 0);
            vx[u] = *reinterpret_cast<const uint4 *>(c.X + ((size_t)(b * g.Ci + ci) * g.Hi + hic) * g.Wi + chc * 8);
            inside |= (uint32_t)(ch > 0 && hi >= 0 && hi < g.Hi) << u;
        }
        const int iw = min(r * NW * 256 + (int)threadIdx.x, nwl - 1);
        uint4 vw = make_uint4(0u, 0u, 0u, 0u);
        uint16_t ew = 0;
        if constexpr (WV)
            vw = reinterpret_cast<const uint4 *>(wsrc)[iw];
        else
            ew = wsrc[iw];
#pragma unroll
        for (int u = 0; u < NX; ++u) {
            const int i = (r * NX + u) * 256 + (int)threadIdx.x;
            if (i < nch)
                reinterpret_cast<uint4 *>(sx)[i] = inside >> u & 1u ? vx[u] : make_uint4(0u, 0u, 0u, 0u);
        }
        if (r * NW * 256 + (int)threadIdx.x < nwl) {
            if constexpr (WV)
                reinterpret_cast<uint4 *>(sw)[iw] = vw;
            else
                sw[iw] = ew;
        }
    }
    __syncthreads();
    const int pp = c.R * g.Wo, p = threadIdx.x % pp, ks = threadIdx.x / pp;
    const bool live = ks < c.nks && p < rows_out * g.Wo;
    float acc = 0.0f;
    if (live) {
        const int hl = p / g.Wo, wo = p - hl * g.Wo;
        const uint16_t *xr0 = sx + 2 * hl * Wp + kXPad - 1 + 2 * wo;   // channel 0, tap (0, 0): column 2 wo - 1
#pragma unroll 2
        for (int ci = ks; ci < g.Ci; ci += c.nks) {
            const uint16_t *xr = xr0 + ci * rows_in * Wp, *wr = sw + ci * 9;
#pragma unroll
            for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                for (int kx = 0; kx < 3; ++kx)
                    acc += bf2f(xr[ky * Wp + kx]) * bf2f(wr[ky * 3 + kx]);
        }
    }
    if (c.nks > 1) {
        if (live)
            part[ks * pp + p] = acc;
        __syncthreads();
        if (!live || ks != 0)
            return;
        for (int k = 1; k < c.nks; ++k)
            acc += part[k * pp + p];
    } else if (!live) {
        return;
    }
    const int hl = p / g.Wo, wo = p - hl * g.Wo;
    const float v = acc + c.bias[co];
    c.Y[((size_t)(b * g.Co + co) * g.Ho + r0 + hl) * g.Wo + wo] = f2bf(v > 0.0f ? v : 0.0f);
}

// fixed-order sum of split partials into the flat gradient
__global__ void k_sum_splits(const float *__restrict__ part, int splits, int n, float *__restrict__ out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int z = 0; z < splits; ++z)
            s += part[(size_t)z * n + i];
        out[i] = s;
    }
}

// Bias gradient db[c] = sum over b, p of dY[b][c][p], part 1 of 2: workgroup
// (c, s) sums the planes of batch slice b = s, s + S, ... (16-byte loads along
// a plane when HW % 8 == 0) into part[s][c]; k_sum_splits then adds the S
// slices of a channel in slice order, so the gradient is deterministic.  One
// workgroup per channel ran 16 workgroups over 868k values each for conv1.
__global__ __launch_bounds__(256) void k_bias_part(const uint16_t *__restrict__ dY, int B, int C, int HW, int S,
                                                   float *__restrict__ part)
{
    __shared__ float red[256];
    const int c = blockIdx.x, s = blockIdx.y, nb = (B - s + S - 1) / S;
    float acc = 0.0f;
    if (HW % 8 == 0) {
        const int per = HW / 8;
        for (int i = threadIdx.x; i < nb * per; i += 256) {
            const int j = i / per, q = i - j * per;
            const uint4 v = *reinterpret_cast<const uint4 *>(dY + ((size_t)(s + j * S) * C + c) * HW + q * 8);
            const uint32_t w[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                acc += bf2f((uint16_t)(w[k] & 0xffffu));
                acc += bf2f((uint16_t)(w[k] >> 16));
            }
        }
    } else {
        for (int i = threadIdx.x; i < nb * HW; i += 256) {
            const int j = i / HW, p = i - j * HW;
            acc += bf2f(dY[((size_t)(s + j * S) * C + c) * HW + p]);
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[(size_t)s * C + c] = red[0];
}

// loss 0.5/B sum (y - t)^2 and dy = (y - t)/B (bf16); single workgroup
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ y, const float *__restrict__ t, int n,
                                              float invB, uint16_t *__restrict__ dy, float *__restrict__ loss)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) {
        const float d = y[i] - t[i];
        s += d * d;
        dy[i] = f2bf(d * invB);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *loss = 0.5f * red[0] * invB;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, uint16_t *__restrict__ pb, int n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.0f - b1) * gi;
        const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = v;
        const float w = p[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
        p[i] = w;
        pb[i] = f2bf(w);
    }
}

__global__ void k_to_bf16(const float *__restrict__ in, uint16_t *__restrict__ out, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = f2bf(in[i]);
}

__global__ void k_gather_batch(const uint16_t *__restrict__ set_x, const float *__restrict__ set_t,
                               const int *__restrict__ idx, int B, uint16_t *__restrict__ x, float *__restrict__ t)
{
    const size_t per = (size_t)kInC * kInH * kInW;
    const int b = blockIdx.y;
    if (b >= B)
        return;
    const size_t src = (size_t)idx[b] * per;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < per; i += (size_t)gridDim.x * blockDim.x)
        x[(size_t)b * per + i] = set_x[src + i];
    if (blockIdx.x == 0 && threadIdx.x < kOut)
        t[b * kOut + threadIdx.x] = set_t[(size_t)idx[b] * kOut + threadIdx.x];
}

// ------------------------------------------------------------- input frames
// 4x4-block mean of the valid depths of a uint16 mm frame, / 4 m, as bf16 (one channel)
__global__ void k_downsample(const uint16_t *__restrict__ mm, int W, int H, uint16_t *__restrict__ out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= kInW * kInH)
        return;
    const int y = q / kInW, x = q - y * kInW;
    uint16_t v[16];   // all 16 loads in flight together (clamped, the bounds applied after)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            v[4 * j + i] = mm[(size_t)min(4 * y + j, H - 1) * W + min(4 * x + i, W - 1)];
    float s = 0.0f, c = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint16_t d = 4 * y + j < H && 4 * x + i < W ? v[4 * j + i] : (uint16_t)0;
            if (d) {
                s += (float)d * 0.001f;
                c += 1.0f;
            }
        }
    out[q] = f2bf(c > 0.0f ? (s / c) * 0.25f : 0.0f);
}

// --------------------------------------------------------------- rendering
// Ray casting of kifi_scene on the device, double precision, the same geometry
// and noise model as kifi_scene_render (synth.c): Box-Muller noise N(0, (k z^2)^2)
// from SplitMix64 keyed by (seed, pixel), rounding to mm, Bernoulli invalid mask.
__device__ inline uint64_t sm64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
__device__ inline double unit01(uint64_t r) { return ((r >> 11) + 1) * (1.0 / 9007199254740992.0); }

__device__ double hit_sphere_d(const double *o, const double *d, const double *sp)
{
    const double oc[3] = { o[0] - sp[0], o[1] - sp[1], o[2] - sp[2] };
    const double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double b = 2.0 * (d[0] * oc[0] + d[1] * oc[1] + d[2] * oc[2]);
    const double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - sp[3] * sp[3];
    const double disc = b * b - 4.0 * a * c;
    if (disc < 0.0)
        return INFINITY;
    const double sq = sqrt(disc);
    const double t0 = (-b - sq) / (2.0 * a), t1 = (-b + sq) / (2.0 * a);
    if (t0 > 1e-9)
        return t0;
    if (t1 > 1e-9)
        return t1;
    return INFINITY;
}

__device__ double hit_box_d(const double *o, const double *d, const double *bx, bool inside)
{
    double tn = -INFINITY, tf = INFINITY;
    for (int a = 0; a < 3; ++a) {
        if (d[a] == 0.0) {
            if (o[a] < bx[a] || o[a] > bx[a + 3])
                return INFINITY;
            continue;
        }
        double t1 = (bx[a] - o[a]) / d[a], t2 = (bx[a + 3] - o[a]) / d[a];
        if (t1 > t2) {
            const double t = t1;
            t1 = t2;
            t2 = t;
        }
        tn = t1 > tn ? t1 : tn;
        tf = t2 < tf ? t2 : tf;
    }
    if (tn > tf)
        return INFINITY;
    if (inside)
        return tf > 1e-9 ? tf : INFINITY;
    return tn > 1e-9 ? tn : INFINITY;
}

struct RenderArgs {
    kifi_scene sc;
    double e2w[16];
    int W, H;
    float fx, fy, cx, cy;
    uint64_t key;
    double noise_k, invalid;
};

__global__ __launch_bounds__(256) void k_render(RenderArgs ra, uint16_t *__restrict__ out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ra.W * ra.H)
        return;
    const int v = q / ra.W, u = q - v * ra.W;
    const double o[3] = { ra.e2w[12], ra.e2w[13], ra.e2w[14] };
    const double de[3] = { (u + 0.5 - ra.cx) / ra.fx, (ra.cy - v - 0.5) / ra.fy, -1.0 };
    double d[3];
    for (int r = 0; r < 3; ++r)
        d[r] = ra.e2w[r] * de[0] + ra.e2w[r + 4] * de[1] + ra.e2w[r + 8] * de[2];
    double t = INFINITY;
    if (ra.sc.kind == KIFI_SCENE_PLANE_SPHERES && d[2] != 0.0) {
        const double tp = (ra.sc.plane_z - o[2]) / d[2];
        if (tp > 1e-9)
            t = tp;
    }
    if (ra.sc.kind == KIFI_SCENE_ROOM) {
        const double tr = hit_box_d(o, d, ra.sc.room, true);
        t = tr < t ? tr : t;
    }
    for (int i = 0; i < ra.sc.n_spheres; ++i) {
        const double ts = hit_sphere_d(o, d, ra.sc.spheres[i]);
        t = ts < t ? ts : t;
    }
    for (int i = 0; i < ra.sc.n_boxes; ++i) {
        const double tb = hit_box_d(o, d, ra.sc.boxes[i], false);
        t = tb < t ? tb : t;
    }
    const uint64_t c = ra.key + 4 * (uint64_t)q;
    const double u1 = unit01(sm64(c)), u2 = unit01(sm64(c + 1)), u3 = unit01(sm64(c + 2));
    const double g = sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
    uint16_t mm = 0;
    if (isfinite(t)) {
        const double z = t + g * ra.noise_k * t * t;
        const double m = floor(z * 1000.0 + 0.5);
        if (m > 0.0 && m < 65535.0)
            mm = (uint16_t)m;
    }
    if (u3 <= ra.invalid)
        mm = 0;
    out[q] = mm;
}

void render_launch(const kifi_scene &sc, const kifi_intrinsics &cam, const double e2w[16], uint64_t seed,
                   double noise_k, double invalid, uint16_t *out, hipStream_t st)
{
    RenderArgs ra;
    ra.sc = sc;
    for (int i = 0; i < 16; ++i)
        ra.e2w[i] = e2w[i];
    ra.W = cam.width;
    ra.H = cam.height;
    ra.fx = cam.fx;
    ra.fy = cam.fy;
    ra.cx = cam.cx;
    ra.cy = cam.cy;
    ra.key = splitmix_host(seed);
    ra.noise_k = noise_k;
    ra.invalid = invalid;
    const int P = cam.width * cam.height;
    hipLaunchKernelGGL(k_render, dim3((P + 255) / 256), dim3(256), 0, st, ra, out);
}

// ------------------------------------------------------------- pose algebra
// Twist (omega, v) <-> rigid transform [R(omega) | v], column storage, FP64.
__host__ __device__ inline void twist_exp(const double xi[6], double T[16])
{
    const double wx = xi[0], wy = xi[1], wz = xi[2];
    const double th = sqrt(wx * wx + wy * wy + wz * wz);
    double A, Bc;
    if (th < 1e-8) {
        A = 1.0;
        Bc = 0.5;
    } else {
        A = sin(th) / th;
        Bc = (1.0 - cos(th)) / (th * th);
    }
    const double K[9] = { 0, wz, -wy, -wz, 0, wx, wy, -wx, 0 };   // column storage of [w]x
    double K2[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            K2[r + 3 * c] = K[r] * K[3 * c] + K[r + 3] * K[3 * c + 1] + K[r + 6] * K[3 * c + 2];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            T[r + 4 * c] = (r == c ? 1.0 : 0.0) + A * K[r + 3 * c] + Bc * K2[r + 3 * c];
    T[3] = T[7] = T[11] = 0.0;
    T[12] = xi[3];
    T[13] = xi[4];
    T[14] = xi[5];
    T[15] = 1.0;
}

__host__ __device__ inline void twist_log(const double T[16], double xi[6])
{
    const double tr = T[0] + T[5] + T[10];
    const double c = fmin(1.0, fmax(-1.0, 0.5 * (tr - 1.0)));
    const double th = acos(c);
    // skew part: (R - R^T)/2 = sin(th) [axis]x
    const double sx = 0.5 * (T[6] - T[9]), sy = 0.5 * (T[8] - T[2]), sz = 0.5 * (T[1] - T[4]);
    const double s = sin(th);
    const double f = th < 1e-8 ? 1.0 : th / s;
    xi[0] = sx * f;
    xi[1] = sy * f;
    xi[2] = sz * f;
    xi[3] = T[12];
    xi[4] = T[13];
    xi[5] = T[14];
}

__host__ __device__ inline void mul44(const double *A, const double *B, double *C)
{
    double r[16];
    for (int c = 0; c < 4; ++c)
        for (int i = 0; i < 4; ++i)
            r[i + 4 * c] = A[i] * B[4 * c] + A[i + 4] * B[4 * c + 1] + A[i + 8] * B[4 * c + 2] + A[i + 12] * B[4 * c + 3];
    for (int i = 0; i < 16; ++i)
        C[i] = r[i];
}

__host__ __device__ inline void inv_rigid(const double *M, double *out)
{
    double r[16];
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 3; ++i)
            r[i + 4 * c] = M[c + 4 * i];
    for (int i = 0; i < 3; ++i)
        r[12 + i] = -(r[i] * M[12] + r[i + 4] * M[13] + r[i + 8] * M[14]);
    r[3] = r[7] = r[11] = 0.0;
    r[15] = 1.0;
    for (int i = 0; i < 16; ++i)
        out[i] = r[i];
}

// target of pair i: T* = inv(P_A) * P_B(ICP), normalised twist; valid if ICP kept
// enough inliers and stayed within 2 cm / 2 deg of the ground truth
__global__ void k_pair_target(const IcpState *__restrict__ st, const double *__restrict__ pa, const double *__restrict__ gt,
                              float *__restrict__ target, float *__restrict__ truth, int *__restrict__ valid)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    double ia[16], T[16], xi[6];
    inv_rigid(pa, ia);
    mul44(ia, st->pose, T);
    twist_log(T, xi);
    bool ok = !st->failed && st->inliers > 2000;
    double dw = 0.0, dv = 0.0;
    for (int j = 0; j < 3; ++j) {
        dw += (xi[j] - gt[j]) * (xi[j] - gt[j]);
        dv += (xi[3 + j] - gt[3 + j]) * (xi[3 + j] - gt[3 + j]);
    }
    ok = ok && sqrt(dw) < 0.035 && sqrt(dv) < 0.02;
    for (int j = 0; j < 6; ++j) {
        target[j] = (float)(xi[j] / kTwistScale[j]);
        truth[j] = (float)gt[j];
    }
    *valid = ok ? 1 : 0;
}

// guess = prev * exp(twist), twist = network output x scale (pipeline use)
__global__ void k_twist_guess(const float *__restrict__ y, const double *__restrict__ prev, double *__restrict__ guess)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    double xi[6], T[16];
    for (int j = 0; j < 6; ++j)
        xi[j] = (double)y[j] * (double)kTwistScale[j];
    twist_exp(xi, T);
    mul44(prev, T, guess);
}

} // namespace

// =================================================================== host side
namespace kifi {

struct NetLayout {
    size_t w[7], b[7];      // offsets of the 5 conv and 2 fc weight/bias blocks
    size_t total;
    NetLayout()
    {
        size_t o = 0;
        for (int l = 0; l < kNumConv; ++l) {
            w[l] = o;
            o += (size_t)kConvC[l + 1] * kConvC[l] * 9;
            b[l] = o;
            o += kConvC[l + 1];
        }
        const int fin = kConvC[kNumConv] * 4 * 4;
        w[5] = o; o += (size_t)kFc1 * fin;
        b[5] = o; o += kFc1;
        w[6] = o; o += (size_t)kOut * kFc1;
        b[6] = o; o += kOut;
        total = o;
    }
};

} // namespace kifi

struct kifi_twistnet {
    kifi_ctx *ctx = nullptr;
    kifi::NetLayout L;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;   // master weights, grads, Adam
    uint16_t *pb = nullptr;                                           // bf16 weights
    long long step = 0;
    // activations (bf16) and their gradients, sized for kMaxBatch
    uint16_t *act[7] = {};     // 0: input, 1..5 conv outputs, 6: fc1 output
    uint16_t *dact[7] = {};
    float *y = nullptr, *t = nullptr, *loss = nullptr;
    uint16_t *dy = nullptr;
    float *part = nullptr;
    size_t part_floats = 0;
    int *d_idx = nullptr;
    // pipeline state: previous frame's downsampled image, and the per-frame input
    uint16_t *prev_small = nullptr;
    bool have_prev = false;
    float *pipe_y = nullptr;
};

struct kifi_pairset {
    kifi_ctx *ctx = nullptr;
    int n = 0;
    uint16_t *x = nullptr;     // n x 2 x 106 x 128 bf16
    float *t = nullptr;        // n x 6 normalised targets
    float *truth = nullptr;    // n x 6 ground-truth twists (physical units)
    int *valid = nullptr;      // n
};

namespace {

size_t act_elems(int l)
{
    if (l == 0)
        return (size_t)kInC * kInH * kInW;
    if (l == 6)
        return kFc1;
    int h = kInH, w = kInW;
    for (int i = 0; i < l; ++i) {
        h = conv_out(h);
        w = conv_out(w);
    }
    return (size_t)kConvC[l] * h * w;
}

ConvGeom conv_geom(int l, int B)
{
    int h = kInH, w = kInW;
    for (int i = 0; i < l; ++i) {
        h = conv_out(h);
        w = conv_out(w);
    }
    ConvGeom g;
    g.B = B;
    g.Ci = kConvC[l];
    g.Hi = h;
    g.Wi = w;
    g.Co = kConvC[l + 1];
    g.Ho = conv_out(h);
    g.Wo = conv_out(w);
    return g;
}

constexpr int kSmallBatch = 4;   // batch sizes that run the direct convolution and one wave per FC output

// k_conv_small for one layer; false when its shape or alignment does not fit
// (the caller then runs the GEMM path).
bool launch_conv_small(const OpConvFwd &op, hipStream_t st)
{
    const ConvGeom &g = op.g;
    const int K = g.Ci * 9, Wp = g.Wi + kXPad;
    if (g.Wi % 8 != 0 || K % 2 != 0 || (reinterpret_cast<uintptr_t>(op.X) & 15u))
        return false;
    const bool wv = K % 8 == 0 && (reinterpret_cast<uintptr_t>(op.W) & 15u) == 0;
    // a band of output rows fills the 256 threads at most, its input rows stay within 48 KiB
    int R = std::min(g.Ho, std::max(1, 256 / g.Wo));
    while (R > 1 && (size_t)g.Ci * (2 * R + 1) * Wp * 2 > 48 * 1024)
        --R;
    const int nks = std::max(1, std::min(g.Ci, 256 / (R * g.Wo)));
    const ConvSmall c{ g, op.X, op.W, op.bias, op.Y, R, nks };
    const size_t lds = ((size_t)g.Ci * (2 * R + 1) * Wp + conv_small_kp(K, wv)) * 2 + (size_t)256 * 4;
    const int nch = g.Ci * (2 * R + 1) * (Wp / 8), nwl = wv ? K / 8 : K;
    if (R * g.Wo > 256 || lds > 64 * 1024 || nch > 12 * 256 * 4 || nwl > 256 * 4)
        return false;
    const dim3 grid((g.Ho + R - 1) / R, g.Co, g.B);
    if (wv)
        hipLaunchKernelGGL(k_conv_small<true>, grid, dim3(256), lds, st, c);
    else
        hipLaunchKernelGGL(k_conv_small<false>, grid, dim3(256), lds, st, c);
    return true;
}

// forward of B inputs already in net->act[0]; outputs net->y (B x 6, normalised)
void net_forward(kifi_twistnet *n, int B, hipStream_t st)
{
    const kifi::NetLayout &L = n->L;
    for (int l = 0; l < kNumConv; ++l) {
        OpConvFwd op;
        op.g = conv_geom(l, B);
        op.X = n->act[l];
        op.W = n->pb + L.w[l];
        op.bias = n->p + L.b[l];
        op.Y = n->act[l + 1];
        op.M = B * op.g.Ho * op.g.Wo;
        op.N = op.g.Co;
        op.K = op.g.Ci * 9;
        op.kchunk = chunk_all(op.K);
        if (B > kSmallBatch || !launch_conv_small(op, st))
            launch_gemm(op, 1, st);
    }
    const int K5 = (int)act_elems(5);
    if (B <= kSmallBatch) {
        // a few rows: one wave per output instead of MFMA tiles along all of K
        auto vec = [](const void *a, const void *b, int K) {
            return (int)(K % 8 == 0 && ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15u) == 0);
        };
        const uint16_t *w5 = n->pb + L.w[5], *w6 = n->pb + L.w[6];
        hipLaunchKernelGGL(k_fc_small<false>, dim3((B * kFc1 * 64 + 255) / 256), dim3(256), 0, st, n->act[5], w5,
                           n->p + L.b[5], B, kFc1, K5, vec(n->act[5], w5, K5), n->act[6], (float *)nullptr);
        hipLaunchKernelGGL(k_fc_small<true>, dim3((B * kOut * 64 + 255) / 256), dim3(256), 0, st, n->act[6], w6,
                           n->p + L.b[6], B, kOut, kFc1, vec(n->act[6], w6, kFc1), (uint16_t *)nullptr, n->y);
        return;
    }
    OpFcFwd f1;
    f1.X = n->act[5];
    f1.W = n->pb + L.w[5];
    f1.bias = n->p + L.b[5];
    f1.Y = n->act[6];
    f1.Yf = nullptr;
    f1.M = B;
    f1.N = kFc1;
    f1.K = K5;
    f1.kchunk = f1.K;
    launch_gemm(f1, 1, st);
    OpFcFwd f2;
    f2.X = n->act[6];
    f2.W = n->pb + L.w[6];
    f2.bias = n->p + L.b[6];
    f2.Y = nullptr;
    f2.Yf = n->y;
    f2.M = B;
    f2.N = kOut;
    f2.K = kFc1;
    f2.kchunk = f2.K;
    launch_gemm(f2, 1, st);
}

int splits_for(int K) { return std::max(1, std::min(256, K / 2048)); }

void sum_splits(kifi_twistnet *n, int splits, int count, float *out, hipStream_t st);

// db of a layer whose output gradient dY is B x C x HW: about 512 workgroups of
// batch slices, then the fixed-order sum of the slices (uses n->part)
void bias_grad(kifi_twistnet *n, const uint16_t *dY, int B, int C, int HW, float *db, hipStream_t st)
{
    const int S = std::max(1, std::min(B, 512 / C));
    hipLaunchKernelGGL(k_bias_part, dim3(C, S), dim3(256), 0, st, dY, B, C, HW, S, n->part);
    sum_splits(n, S, C, db, st);
}

void sum_splits(kifi_twistnet *n, int splits, int count, float *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_sum_splits, dim3(std::min(1024, (count + 255) / 256)), dim3(256), 0, st, n->part, splits,
                       count, out);
}

// backward from net->t (targets, B x 6): loss -> net->loss, gradients -> net->g
void net_backward(kifi_twistnet *n, int B, hipStream_t st)
{
    const kifi::NetLayout &L = n->L;
    hipLaunchKernelGGL(k_loss, dim3(1), dim3(256), 0, st, n->y, n->t, B * kOut, 1.0f / (float)B, n->dy, n->loss);
    // fc2: dW = dy^T h, db = sum dy, dh = dy W2 x ReLU'(h)
    {
        OpFcBwdW op;
        op.dY = n->dy;
        op.X = n->act[6];
        op.part = n->part;
        op.M = kOut;
        op.N = kFc1;
        op.K = B;
        op.kchunk = chunk_all(B);
        op.ldy = kOut;
        launch_gemm(op, 1, st);
        sum_splits(n, 1, kOut * kFc1, n->g + L.w[6], st);
        bias_grad(n, n->dy, B, kOut, 1, n->g + L.b[6], st);
        OpFcBwdData od;
        od.dY = n->dy;
        od.W = n->pb + L.w[6];
        od.X = n->act[6];
        od.dX = n->dact[6];
        od.M = B;
        od.N = kFc1;
        od.K = kOut;
        od.kchunk = chunk_all(kOut);
        launch_gemm(od, 1, st);
    }
    // fc1
    {
        const int fin = (int)act_elems(5);
        OpFcBwdW op;
        op.dY = n->dact[6];
        op.X = n->act[5];
        op.part = n->part;
        op.M = kFc1;
        op.N = fin;
        op.K = B;
        op.kchunk = chunk_all(B);
        op.ldy = kFc1;
        launch_gemm(op, 1, st);
        sum_splits(n, 1, kFc1 * fin, n->g + L.w[5], st);
        bias_grad(n, n->dact[6], B, kFc1, 1, n->g + L.b[5], st);
        OpFcBwdData od;
        od.dY = n->dact[6];
        od.W = n->pb + L.w[5];
        od.X = n->act[5];
        od.dX = n->dact[5];
        od.M = B;
        od.N = fin;
        od.K = kFc1;
        od.kchunk = chunk_all(kFc1);
        launch_gemm(od, 1, st);
    }
    for (int l = kNumConv - 1; l >= 0; --l) {
        const ConvGeom g = conv_geom(l, B);
        OpConvBwdW ow;
        ow.g = g;
        ow.dY = n->dact[l + 1];
        ow.X = n->act[l];
        ow.part = n->part;
        ow.M = g.Co;
        ow.N = g.Ci * 9;
        ow.K = B * g.Ho * g.Wo;
        const int splits = splits_for(ow.K);
        ow.kchunk = chunk_of(ow.K, splits);
        const int used = (ow.K + ow.kchunk - 1) / ow.kchunk;
        launch_gemm(ow, used, st);
        sum_splits(n, used, ow.M * ow.N, n->g + L.w[l], st);
        bias_grad(n, n->dact[l + 1], B, g.Co, g.Ho * g.Wo, n->g + L.b[l], st);
        if (l == 0)
            break;
        OpConvBwdData od;
        od.g = g;
        od.dY = n->dact[l + 1];
        od.W = n->pb + L.w[l];
        od.X = n->act[l];
        od.dX = n->dact[l];
        od.M = B * g.Hi * g.Wi;
        od.N = g.Ci;
        od.K = g.Co * 9;
        od.kchunk = chunk_all(od.K);
        launch_gemm(od, 1, st);
    }
}

void adam_step(kifi_twistnet *n, float lr, hipStream_t st)
{
    n->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - (float)std::pow((double)b1, (double)n->step);
    const float c2 = 1.0f - (float)std::pow((double)b2, (double)n->step);
    const int cnt = (int)n->L.total;
    hipLaunchKernelGGL(k_adam, dim3(1024), dim3(256), 0, st, n->p, n->g, n->m1, n->m2, n->pb, cnt, lr, b1, b2,
                       1e-8f, c1, c2);
}

kifi_status upload_inputs(kifi_twistnet *n, const float *x, int B)
{
    const size_t cnt = (size_t)B * act_elems(0);
    std::vector<uint16_t> h(cnt);
    for (size_t i = 0; i < cnt; ++i)
        h[i] = f2bf_host(x[i]);
    KIFI_HIP(n->ctx, hipMemcpyAsync(n->act[0], h.data(), cnt * 2, hipMemcpyHostToDevice, n->ctx->stream));
    KIFI_HIP(n->ctx, hipStreamSynchronize(n->ctx->stream));
    return KIFI_OK;
}

double gauss_host(uint64_t &s)
{
    s = splitmix_host(s);
    const double u1 = ((s >> 11) + 1) * (1.0 / 9007199254740992.0);
    s = splitmix_host(s);
    const double u2 = ((s >> 11) + 1) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
}

double uniform_host(uint64_t &s)
{
    s = splitmix_host(s);
    return ((s >> 11) + 1) * (1.0 / 9007199254740992.0);
}

} // namespace

namespace kifi {

void twistnet_observe(kifi_twistnet *n, const uint16_t *d_mm, int W, int H, hipStream_t st);

// Frame-level hook for the pipeline: twist of (previous frame, this frame) and
// the guess prev * exp(twist) into guess_dev.  Returns false (no guess) on the
// first frame it sees.
bool twistnet_guess(kifi_twistnet *n, const uint16_t *d_mm, int W, int H, const double *prev_pose_dev,
                    double *guess_dev, hipStream_t st)
{
    const size_t per = (size_t)kInH * kInW;
    const int blocks = (int)((per + 255) / 256);
    if (!n->have_prev) {
        twistnet_observe(n, d_mm, W, H, st);
        return false;
    }
    // input channels: previous frame, this frame
    (void)hipMemcpyAsync(n->act[0], n->prev_small, per * 2, hipMemcpyDeviceToDevice, st);
    hipLaunchKernelGGL(k_downsample, dim3(blocks), dim3(256), 0, st, d_mm, W, H, n->act[0] + per);
    net_forward(n, 1, st);
    hipLaunchKernelGGL(k_twist_guess, dim3(1), dim3(64), 0, st, n->y, prev_pose_dev, guess_dev);
    (void)hipMemcpyAsync(n->prev_small, n->act[0] + per, per * 2, hipMemcpyDeviceToDevice, st);
    return true;
}

// A frame the network does not predict for (first frame, given pose): keep its
// downsampled image as the previous frame of the next prediction.
void twistnet_observe(kifi_twistnet *n, const uint16_t *d_mm, int W, int H, hipStream_t st)
{
    const int blocks = (int)(((size_t)kInH * kInW + 255) / 256);
    hipLaunchKernelGGL(k_downsample, dim3(blocks), dim3(256), 0, st, d_mm, W, H, n->prev_small);
    n->have_prev = true;
}

void twistnet_reset_frames(kifi_twistnet *n) { n->have_prev = false; }

} // namespace kifi

// ===================================================================== C ABI
extern "C" {

size_t kifi_twistnet_param_count(void) { return kifi::NetLayout().total; }

void kifi_twistnet_destroy(kifi_twistnet *n)
{
    if (!n)
        return;
    (void)hipStreamSynchronize(n->ctx->stream);
    void *bufs[] = { n->p, n->g, n->m1, n->m2, n->pb, n->y, n->t, n->loss, n->dy, n->part, n->d_idx,
                     n->prev_small, n->pipe_y };
    for (void *b : bufs)
        if (b)
            (void)hipFree(b);
    for (int l = 0; l < 7; ++l) {
        if (n->act[l])
            (void)hipFree(n->act[l]);
        if (n->dact[l])
            (void)hipFree(n->dact[l]);
    }
    delete n;
}

kifi_status kifi_twistnet_create(kifi_ctx *ctx, uint64_t seed, kifi_twistnet **out)
{
    if (!ctx || !out)
        return KIFI_ERR_ARG;
    *out = nullptr;
    kifi_twistnet *n = new (std::nothrow) kifi_twistnet;
    if (!n)
        return KIFI_ERR_NOMEM;
    n->ctx = ctx;
    const size_t P = n->L.total;
    hipError_t e = hipSuccess;
    auto A = [&](void **ptr, size_t bytes) {
        if (e == hipSuccess)
            e = hipMalloc(ptr, bytes);
        if (e == hipSuccess)
            e = hipMemsetAsync(*ptr, 0, bytes, ctx->stream);
    };
    A((void **)&n->p, P * 4);
    A((void **)&n->g, P * 4);
    A((void **)&n->m1, P * 4);
    A((void **)&n->m2, P * 4);
    A((void **)&n->pb, P * 2);
    for (int l = 0; l < 7; ++l) {
        A((void **)&n->act[l], (size_t)kMaxBatch * act_elems(l) * 2);
        A((void **)&n->dact[l], (size_t)kMaxBatch * act_elems(l) * 2);
    }
    A((void **)&n->y, kMaxBatch * kOut * 4);
    A((void **)&n->t, kMaxBatch * kOut * 4);
    A((void **)&n->loss, 4);
    A((void **)&n->dy, kMaxBatch * kOut * 2);
    // split partials: the largest of splits x M x N over the weight gradients
    size_t pf = (size_t)kFc1 * act_elems(5);
    for (int l = 0; l < kNumConv; ++l) {
        const ConvGeom g = conv_geom(l, kMaxBatch);
        const int K = kMaxBatch * g.Ho * g.Wo;
        const int s = splits_for(K), used = (K + chunk_of(K, s) - 1) / chunk_of(K, s);
        pf = std::max(pf, (size_t)used * g.Co * g.Ci * 9);
    }
    n->part_floats = pf;
    A((void **)&n->part, pf * 4);
    A((void **)&n->d_idx, kMaxBatch * 4);
    A((void **)&n->prev_small, (size_t)kInH * kInW * 2);
    if (e != hipSuccess) {
        kifi_twistnet_destroy(n);
        return hip_error(ctx, e, "kifi_twistnet_create");
    }
    // seeded He-normal init (variance 2 / fan_in), output layer 0.1x; zero biases
    std::vector<float> h(P, 0.0f);
    uint64_t s = splitmix_host(seed ^ 0x7a15e7ull);
    for (int l = 0; l < 7; ++l) {
        const size_t cnt = n->L.b[l] - n->L.w[l];
        const double fan_in = l < kNumConv ? kConvC[l] * 9.0 : (l == 5 ? (double)act_elems(5) : (double)kFc1);
        const double sd = std::sqrt(2.0 / fan_in) * (l == 6 ? 0.1 : 1.0);
        for (size_t i = 0; i < cnt; ++i)
            h[n->L.w[l] + i] = (float)(gauss_host(s) * sd);
    }
    kifi_status st = kifi_twistnet_set_params(n, h.data());
    if (st != KIFI_OK) {
        kifi_twistnet_destroy(n);
        return st;
    }
    *out = n;
    return KIFI_OK;
}

kifi_status kifi_twistnet_get_params(kifi_twistnet *n, float *out)
{
    if (!n || !out)
        return KIFI_ERR_ARG;
    KIFI_HIP(n->ctx, hipMemcpyAsync(out, n->p, n->L.total * 4, hipMemcpyDeviceToHost, n->ctx->stream));
    KIFI_HIP(n->ctx, hipStreamSynchronize(n->ctx->stream));
    return KIFI_OK;
}

kifi_status kifi_twistnet_set_params(kifi_twistnet *n, const float *in)
{
    if (!n || !in)
        return KIFI_ERR_ARG;
    hipStream_t st = n->ctx->stream;
    KIFI_HIP(n->ctx, hipMemcpyAsync(n->p, in, n->L.total * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_to_bf16, dim3(1024), dim3(256), 0, st, n->p, n->pb, n->L.total);
    KIFI_HIP(n->ctx, hipMemsetAsync(n->m1, 0, n->L.total * 4, st));
    KIFI_HIP(n->ctx, hipMemsetAsync(n->m2, 0, n->L.total * 4, st));
    n->step = 0;
    KIFI_HIP(n->ctx, hipStreamSynchronize(st));
    return KIFI_OK;
}

kifi_status kifi_twistnet_forward(kifi_twistnet *n, const float *x, int B, float *y)
{
    if (!n || !x || !y || B < 1 || B > kMaxBatch)
        return KIFI_ERR_ARG;
    kifi_status s = upload_inputs(n, x, B);
    if (s != KIFI_OK)
        return s;
    net_forward(n, B, n->ctx->stream);
    KIFI_HIP(n->ctx, hipGetLastError());
    KIFI_HIP(n->ctx, hipMemcpyAsync(y, n->y, (size_t)B * kOut * 4, hipMemcpyDeviceToHost, n->ctx->stream));
    KIFI_HIP(n->ctx, hipStreamSynchronize(n->ctx->stream));
    return KIFI_OK;
}

kifi_status kifi_twistnet_loss_grad(kifi_twistnet *n, const float *x, const float *t, int B, float *loss,
                                    float *grad)
{
    if (!n || !x || !t || B < 1 || B > kMaxBatch)
        return KIFI_ERR_ARG;
    kifi_status s = upload_inputs(n, x, B);
    if (s != KIFI_OK)
        return s;
    hipStream_t st = n->ctx->stream;
    KIFI_HIP(n->ctx, hipMemcpyAsync(n->t, t, (size_t)B * kOut * 4, hipMemcpyHostToDevice, st));
    net_forward(n, B, st);
    net_backward(n, B, st);
    KIFI_HIP(n->ctx, hipGetLastError());
    if (loss)
        KIFI_HIP(n->ctx, hipMemcpyAsync(loss, n->loss, 4, hipMemcpyDeviceToHost, st));
    if (grad)
        KIFI_HIP(n->ctx, hipMemcpyAsync(grad, n->g, n->L.total * 4, hipMemcpyDeviceToHost, st));
    KIFI_HIP(n->ctx, hipStreamSynchronize(st));
    return KIFI_OK;
}

kifi_status kifi_twistnet_apply_gradient(kifi_twistnet *n, const float *grad, float lr)
{
    if (!n || !grad || !(lr >= 0.0f))
        return KIFI_ERR_ARG;
    hipStream_t st = n->ctx->stream;
    KIFI_HIP(n->ctx, hipMemcpyAsync(n->g, grad, n->L.total * 4, hipMemcpyHostToDevice, st));
    adam_step(n, lr, st);
    KIFI_HIP(n->ctx, hipGetLastError());
    KIFI_HIP(n->ctx, hipStreamSynchronize(st));
    return KIFI_OK;
}

kifi_status kifi_twistnet_train(kifi_twistnet *n, kifi_pairset *set, int steps, int batch, float lr, uint64_t seed,
                                float *losses)
{
    if (!n || !set || steps < 0 || batch < 1 || batch > kMaxBatch || set->n < 1)
        return KIFI_ERR_ARG;
    hipStream_t st = n->ctx->stream;
    // valid pairs only
    std::vector<int> valid(set->n), pool;
    KIFI_HIP(n->ctx, hipMemcpyAsync(valid.data(), set->valid, set->n * 4, hipMemcpyDeviceToHost, st));
    KIFI_HIP(n->ctx, hipStreamSynchronize(st));
    for (int i = 0; i < set->n; ++i)
        if (valid[i])
            pool.push_back(i);
    if (pool.empty())
        return set_error(n->ctx, KIFI_ERR_STATE, "kifi_twistnet_train: no valid pair in the set");
    std::vector<float> lossv(steps, 0.0f);
    float *d_losses = nullptr;
    if (steps)
        KIFI_HIP(n->ctx, hipMalloc(&d_losses, (size_t)steps * 4));
    // batches: seeded shuffles of the pool, one epoch after another
    uint64_t s = splitmix_host(seed ^ 0xba7c4ull);
    std::vector<int> order(pool), idx(batch);
    size_t pos = order.size();
    std::vector<std::vector<int>> all(steps, std::vector<int>(batch));
    for (int k = 0; k < steps; ++k)
        for (int b = 0; b < batch; ++b) {
            if (pos == order.size()) {
                for (size_t i = order.size() - 1; i > 0; --i) {
                    s = splitmix_host(s);
                    std::swap(order[i], order[s % (i + 1)]);
                }
                pos = 0;
            }
            all[k][b] = order[pos++];
        }
    int *d_all = nullptr;
    hipError_t e = steps ? hipMalloc(&d_all, (size_t)steps * batch * 4) : hipSuccess;
    std::vector<int> flat((size_t)steps * batch);
    for (int k = 0; k < steps; ++k)
        std::copy(all[k].begin(), all[k].end(), flat.begin() + (size_t)k * batch);
    if (e == hipSuccess && steps)
        e = hipMemcpyAsync(d_all, flat.data(), flat.size() * 4, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
        if (d_losses) (void)hipFree(d_losses);
        if (d_all) (void)hipFree(d_all);
        return hip_error(n->ctx, e, "kifi_twistnet_train setup");
    }
    const size_t per = act_elems(0);
    for (int k = 0; k < steps; ++k) {
        hipLaunchKernelGGL(k_gather_batch, dim3((unsigned)((per + 255) / 256 / 4), batch), dim3(256), 0, st,
                           set->x, set->t, d_all + (size_t)k * batch, batch, n->act[0], n->t);
        net_forward(n, batch, st);
        net_backward(n, batch, st);
        adam_step(n, lr, st);
        (void)hipMemcpyAsync(d_losses + k, n->loss, 4, hipMemcpyDeviceToDevice, st);
    }
    e = hipGetLastError();
    if (e == hipSuccess && steps)
        e = hipMemcpyAsync(lossv.data(), d_losses, (size_t)steps * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (d_losses) (void)hipFree(d_losses);
    if (d_all) (void)hipFree(d_all);
    if (e != hipSuccess)
        return hip_error(n->ctx, e, "kifi_twistnet_train");
    if (losses)
        std::copy(lossv.begin(), lossv.end(), losses);
    return KIFI_OK;
}

kifi_status kifi_twistnet_predict_set(kifi_twistnet *n, kifi_pairset *set, int first, int count, float *twists)
{
    if (!n || !set || !twists || first < 0 || count < 0 || first + count > set->n)
        return KIFI_ERR_ARG;
    hipStream_t st = n->ctx->stream;
    const size_t per = act_elems(0);
    std::vector<float> y(kMaxBatch * kOut);
    for (int b0 = 0; b0 < count; b0 += kMaxBatch) {
        const int B = std::min(kMaxBatch, count - b0);
        KIFI_HIP(n->ctx, hipMemcpyAsync(n->act[0], set->x + (size_t)(first + b0) * per, (size_t)B * per * 2,
                                        hipMemcpyDeviceToDevice, st));
        net_forward(n, B, st);
        KIFI_HIP(n->ctx, hipMemcpyAsync(y.data(), n->y, (size_t)B * kOut * 4, hipMemcpyDeviceToHost, st));
        KIFI_HIP(n->ctx, hipStreamSynchronize(st));
        for (int i = 0; i < B * kOut; ++i)
            twists[(size_t)b0 * kOut + i] = y[i] * kTwistScale[i % kOut];
    }
    return KIFI_OK;
}

void kifi_pairset_destroy(kifi_pairset *s)
{
    if (!s)
        return;
    (void)hipStreamSynchronize(s->ctx->stream);
    if (s->x) (void)hipFree(s->x);
    if (s->t) (void)hipFree(s->t);
    if (s->truth) (void)hipFree(s->truth);
    if (s->valid) (void)hipFree(s->valid);
    delete s;
}

static kifi_status pairset_generate(kifi_ctx *ctx, const kifi_intrinsics *cam, const kifi_scene *scenes, int n_scenes,
                                    int n, uint64_t seed, const kifi_motion_steps &ms, kifi_pairset **out);

kifi_status kifi_pairset_generate(kifi_ctx *ctx, const kifi_intrinsics *cam, const kifi_scene *scenes, int n_scenes,
                                  int n, uint64_t seed, float max_trans, float max_rot_deg, kifi_pairset **out)
{
    kifi_motion_steps ms{};
    ms.max_trans = max_trans;
    ms.max_rot_deg = max_rot_deg;
    return pairset_generate(ctx, cam, scenes, n_scenes, n, seed, ms, out);
}

kifi_status kifi_pairset_generate_steps(kifi_ctx *ctx, const kifi_intrinsics *cam, const kifi_scene *scenes,
                                        int n_scenes, int n, uint64_t seed, const kifi_motion_steps *steps,
                                        kifi_pairset **out)
{
    if (!steps || !(steps->min_trans >= 0.0f) || !(steps->max_trans >= steps->min_trans) ||
        !(steps->min_rot_deg >= 0.0f) || !(steps->max_rot_deg >= steps->min_rot_deg) ||
        !(steps->max_pitch_step_deg >= 0.0f))
        return ctx ? set_error(ctx, KIFI_ERR_ARG, "kifi_pairset_generate_steps: invalid step ranges")
                   : KIFI_ERR_ARG;
    return pairset_generate(ctx, cam, scenes, n_scenes, n, seed, *steps, out);
}

static kifi_status pairset_generate(kifi_ctx *ctx, const kifi_intrinsics *cam, const kifi_scene *scenes, int n_scenes,
                                    int n, uint64_t seed, const kifi_motion_steps &ms, kifi_pairset **out)
{
    if (!ctx || !cam || !scenes || n_scenes < 1 || n < 1 || !out || cam->width > 4 * kInW ||
        cam->height > 4 * kInH + 3 || cam->width % 4 || cam->height % 4)
        return KIFI_ERR_ARG;
    *out = nullptr;
    hipStream_t st = ctx->stream;
    kifi_pairset *s = new (std::nothrow) kifi_pairset;
    if (!s)
        return KIFI_ERR_NOMEM;
    s->ctx = ctx;
    s->n = n;
    const size_t per = (size_t)kInC * kInH * kInW;
    const size_t P = (size_t)cam->width * cam->height;
    hipError_t e = hipMalloc(&s->x, (size_t)n * per * 2);
    if (e == hipSuccess) e = hipMalloc(&s->t, (size_t)n * kOut * 4);
    if (e == hipSuccess) e = hipMalloc(&s->truth, (size_t)n * kOut * 4);
    if (e == hipSuccess) e = hipMalloc(&s->valid, (size_t)n * 4);
    uint16_t *dA = nullptr, *dB = nullptr;
    double *d_pose = nullptr;     // per pair: P_A (16), P_B ground truth (16), twist (6)
    if (e == hipSuccess) e = hipMalloc(&dA, P * 2);
    if (e == hipSuccess) e = hipMalloc(&dB, P * 2);
    if (e == hipSuccess) e = hipMalloc(&d_pose, (size_t)n * 38 * 8);
    if (e != hipSuccess) {
        if (dA) (void)hipFree(dA);
        if (dB) (void)hipFree(dB);
        if (d_pose) (void)hipFree(d_pose);
        kifi_pairset_destroy(s);
        return hip_error(ctx, e, "kifi_pairset_generate");
    }
    // host: poses of every pair
    std::vector<double> poses((size_t)n * 38);
    uint64_t r = splitmix_host(seed ^ 0x9a125e7ull);
    const double deg = 3.14159265358979323846 / 180.0;
    for (int i = 0; i < n; ++i) {
        double *PA = &poses[(size_t)i * 38], *PB = PA + 16, *xi = PA + 32;
        // base pose: inside the room, radius <= 1 m around the centre, looking out
        const double rad = std::sqrt(uniform_host(r)) * 1.0, ang = uniform_host(r) * 2.0 * 3.14159265358979323846;
        const double px = rad * std::cos(ang), pz = rad * std::sin(ang), py = -0.1 + 0.4 * uniform_host(r);
        const double yaw = uniform_host(r) * 2.0 * 3.14159265358979323846;
        const double pitch = -(20.0 * uniform_host(r)) * deg, roll = (6.0 * uniform_host(r) - 3.0) * deg;
        // R = Ry(yaw) Rx(pitch) Rz(roll) as twists composed
        double Ty[16], Tx[16], Tz[16], T1[16];
        const double wy[6] = { 0, yaw, 0, 0, 0, 0 }, wx[6] = { pitch, 0, 0, 0, 0, 0 }, wz[6] = { 0, 0, roll, 0, 0, 0 };
        twist_exp(wy, Ty);
        twist_exp(wx, Tx);
        twist_exp(wz, Tz);
        mul44(Ty, Tx, T1);
        mul44(T1, Tz, PA);
        PA[12] = px; PA[13] = py; PA[14] = pz;
        if (!ms.walk) {
            // relative motion: random axis, angle U[min_rot, max_rot]; random
            // direction, length U[min_trans, max_trans]
            double ax[3], dir[3], na = 0, nd = 0;
            for (int j = 0; j < 3; ++j) {
                ax[j] = gauss_host(r);
                dir[j] = gauss_host(r);
                na += ax[j] * ax[j];
                nd += dir[j] * dir[j];
            }
            na = std::sqrt(na);
            nd = std::sqrt(nd);
            const double a = (ms.min_rot_deg + uniform_host(r) * (ms.max_rot_deg - ms.min_rot_deg)) * deg;
            const double l = ms.min_trans + uniform_host(r) * (ms.max_trans - ms.min_trans);
            for (int j = 0; j < 3; ++j) {
                xi[j] = ax[j] / na * a;
                xi[3 + j] = dir[j] / nd * l;
            }
            double T[16];
            twist_exp(xi, T);
            mul44(PA, T, PB);
        } else {
            // a step of a walk: yaw +-U[min_rot, max_rot], pitch +-U[0, max_pitch_step]
            // (kept in [-20, 0] deg), horizontal translation U[min_trans, max_trans]
            // in a random direction; the same roll
            const double ang = (ms.min_rot_deg + uniform_host(r) * (ms.max_rot_deg - ms.min_rot_deg)) * deg *
                               (uniform_host(r) < 0.5 ? -1.0 : 1.0);
            double p2 = pitch + (2.0 * uniform_host(r) - 1.0) * ms.max_pitch_step_deg * deg;
            p2 = std::min(0.0, std::max(-20.0 * deg, p2));
            const double len = ms.min_trans + uniform_host(r) * (ms.max_trans - ms.min_trans);
            const double dir = uniform_host(r) * 2.0 * 3.14159265358979323846;
            const double wy2[6] = { 0, yaw + ang, 0, 0, 0, 0 }, wx2[6] = { p2, 0, 0, 0, 0, 0 };
            double Ty2[16], Tx2[16], T2[16];
            twist_exp(wy2, Ty2);
            twist_exp(wx2, Tx2);
            mul44(Ty2, Tx2, T2);
            mul44(T2, Tz, PB);
            PB[12] = px + len * std::cos(dir);
            PB[13] = py;
            PB[14] = pz + len * std::sin(dir);
            // the twist of PA^-1 PB (eye(B) -> eye(A)) for the truth record
            double Ai[16], D[16];
            inv_rigid(PA, Ai);
            mul44(Ai, PB, D);
            twist_log(D, xi);
        }
    }
    e = hipMemcpyAsync(d_pose, poses.data(), poses.size() * 8, hipMemcpyHostToDevice, st);
    kifi_status status = e == hipSuccess ? KIFI_OK : hip_error(ctx, e, "kifi_pairset_generate upload");
    kifi_model m;
    IcpWork w;
    if (status == KIFI_OK) {
        m.ctx = ctx;
        m.cam = *cam;
        m.mp.W = cam->width;
        m.mp.H = cam->height;
        m.mp.fx = cam->fx;
        m.mp.fy = cam->fy;
        m.mp.cx = cam->cx;
        m.mp.cy = cam->cy;
        m.mp.znear = 0.1f;
        status = model_alloc(&m);
    }
    if (status == KIFI_OK)
        status = icp_alloc(&w, ctx, cam->width, cam->height);
    kifi_icp_params ip;
    kifi_icp_params_default(&ip);
    const int blocks = (int)((kInH * kInW + 255) / 256);
    for (int i = 0; i < n && status == KIFI_OK; ++i) {
        const double *PA = &poses[(size_t)i * 38], *PB = PA + 16;
        const kifi_scene &sc = scenes[i % n_scenes];
        const uint64_t fs = splitmix_host(seed + 2 * (uint64_t)i);
        render_launch(sc, *cam, PA, fs, 1.5e-3, 0.02, dA, st);
        render_launch(sc, *cam, PB, fs + 1, 1.5e-3, 0.02, dB, st);
        uint16_t *xi = s->x + (size_t)i * per;
        hipLaunchKernelGGL(k_downsample, dim3(blocks), dim3(256), 0, st, dA, cam->width, cam->height, xi);
        hipLaunchKernelGGL(k_downsample, dim3(blocks), dim3(256), 0, st, dB, cam->width, cam->height,
                           xi + (size_t)kInH * kInW);
        // target: ICP of frame B against frame A's own surface, started at the truth
        float e2w[16], w2e[16];
        double iA[16];
        inv_rigid(PA, iA);
        for (int j = 0; j < 16; ++j) {
            e2w[j] = (float)PA[j];
            w2e[j] = (float)iA[j];
        }
        status = model_from_depth(&m, dA, e2w, w2e);
        if (status != KIFI_OK)
            break;
        icp_set_pose(&w, PB, nullptr, st);
        status = icp_run(&m, &w, dB, cam, PB, nullptr, &ip);
        if (status != KIFI_OK)
            break;
        hipLaunchKernelGGL(k_pair_target, dim3(1), dim3(64), 0, st, w.d_state, d_pose + (size_t)i * 38,
                           d_pose + (size_t)i * 38 + 32, s->t + (size_t)i * kOut, s->truth + (size_t)i * kOut,
                           s->valid + i);
    }
    if (status == KIFI_OK) {
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        if (e != hipSuccess)
            status = hip_error(ctx, e, "kifi_pairset_generate");
    }
    (void)hipStreamSynchronize(st);
    icp_free(&w);
    model_free(&m);
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(d_pose);
    if (status != KIFI_OK) {
        kifi_pairset_destroy(s);
        return status;
    }
    *out = s;
    return KIFI_OK;
}

kifi_status kifi_pairset_download(kifi_pairset *s, float *inputs, float *targets, float *truth, int *valid)
{
    if (!s)
        return KIFI_ERR_ARG;
    hipStream_t st = s->ctx->stream;
    const size_t per = (size_t)kInC * kInH * kInW;
    if (inputs) {
        std::vector<uint16_t> h((size_t)s->n * per);
        KIFI_HIP(s->ctx, hipMemcpyAsync(h.data(), s->x, h.size() * 2, hipMemcpyDeviceToHost, st));
        KIFI_HIP(s->ctx, hipStreamSynchronize(st));
        for (size_t i = 0; i < h.size(); ++i) {
            const uint32_t u = (uint32_t)h[i] << 16;
            std::memcpy(&inputs[i], &u, 4);
        }
    }
    if (targets)
        KIFI_HIP(s->ctx, hipMemcpyAsync(targets, s->t, (size_t)s->n * kOut * 4, hipMemcpyDeviceToHost, st));
    if (truth)
        KIFI_HIP(s->ctx, hipMemcpyAsync(truth, s->truth, (size_t)s->n * kOut * 4, hipMemcpyDeviceToHost, st));
    if (valid)
        KIFI_HIP(s->ctx, hipMemcpyAsync(valid, s->valid, (size_t)s->n * 4, hipMemcpyDeviceToHost, st));
    KIFI_HIP(s->ctx, hipStreamSynchronize(st));
    return KIFI_OK;
}

int kifi_pairset_size(const kifi_pairset *s) { return s ? s->n : 0; }

kifi_status kifi_render(kifi_ctx *ctx, const kifi_scene *sc, const kifi_intrinsics *cam, const double e2w[16],
                        uint64_t noise_seed, double noise_k, double invalid_frac, uint16_t *out, int out_on_device)
{
    if (!ctx || !sc || !cam || !e2w || !out || cam->width <= 0 || cam->height <= 0)
        return KIFI_ERR_ARG;
    hipStream_t st = ctx->stream;
    const size_t P = (size_t)cam->width * cam->height;
    uint16_t *d = out;
    if (!out_on_device)
        KIFI_HIP(ctx, hipMalloc(&d, P * 2));
    render_launch(*sc, *cam, e2w, noise_seed, noise_k, invalid_frac, d, st);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && !out_on_device)
        e = hipMemcpyAsync(out, d, P * 2, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && !out_on_device)
        e = hipStreamSynchronize(st);
    if (!out_on_device)
        (void)hipFree(d);
    if (e != hipSuccess)
        return hip_error(ctx, e, "kifi_render");
    return KIFI_OK;
}

} // extern "C"
