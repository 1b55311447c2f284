// kifi_icp.hip -- model prediction and projective ICP on the GPU.
//
// Reference: ICP::Align / AlignStep (KinectFission/kifi/ICP.cpp:9-265) runs 4 fixed
// Horn point-to-point steps over a 21x21 window search, against a model point
// buffer from Mesher::Mesh + Renderer::Bin (Pipeline.cpp:48-52).  Here:
//   - model prediction: one pass over the volume splats every near-surface voxel
//     (|D| < dmax) at its estimated surface depth along the viewing ray into a
//     z-buffer (64-bit atomicMin of depth | voxel index), then resolves points and
//     normals (binary32, bit-exact with the oracle);
//   - tracking: projective association on a 3-level depth pyramid, point-to-plane
//     6x6 normal equations (or Horn's cross-covariance) accumulated in binary32 with
//     compensated (TwoSum) reductions, solved on the device in FP64 (Cholesky /
//     cyclic Jacobi for Horn's 4x4), pose composed in FP64.  Every iteration is one
//     kernel: the workgroup that finishes last reduces the partials and solves, and
//     a device-side convergence flag turns the level's remaining launches into
//     early exits, so a frame needs no host sync.  The fallback pass behind a
//     learned start is one launch whose workgroups hand the iterations over among
//     themselves (it rarely runs, and then costs more per iteration).
#include <cfloat>
#include <vector>

#include "kifi_internal.hpp"
#include "kifi_prims.hpp"
#include "kifi_icp.hpp"

namespace kifi {

constexpr int kIT = 256;

// ------------------------------------------------------------ voxel access
struct VoxF32i { float sum, weight; };
__device__ __forceinline__ float vw(const VoxF32i &v) { return v.weight; }
__device__ __forceinline__ float vd(const VoxF32i &v) { return v.sum / v.weight; }
__device__ __forceinline__ float vw(uint32_t v) { return (float)((v >> 16) & 0xffu); }
__device__ __forceinline__ float vd(uint32_t v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu)); }

__device__ __forceinline__ uint32_t orderable_f(float z)
{
    if (z == 0.0f)
        z = 0.0f;
    const uint32_t b = __float_as_uint(z);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Consumes the words of a loaded value in an empty asm statement: the load is
// issued before this point, so loads placed ahead of it wait together instead
// of being sunk, one by one, below the tests that use them.
template <typename T>
__device__ __forceinline__ void loaded_here(const T &v)
{
    static_assert(sizeof(T) % 4 == 0, "whole words");
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k)
        asm volatile("" ::"v"(w[k]));
}

// The model sample of a voxel (key, val): false outside the band or behind
// the near plane.
template <typename K, typename V>
__device__ __forceinline__ bool model_sample_of(K key, const V &val, const ModelParams &mp, const Mat4 &w2e,
                                                float &ex, float &ey, float &z, float &zs, float c[3])
{
    if (!(vw(val) > 0.0f))
        return false;
    const float D = vd(val);
    if (!(fabsf(D) < mp.dmax))
        return false;
    uint32_t x, y, zz;
    unpack_key<K>(key, x, y, zz);
    c[0] = (float)x * mp.vl + mp.c0;
    c[1] = (float)y * mp.vl + mp.c0;
    c[2] = (float)zz * mp.vl + mp.c0;
    const float p[4] = { c[0], c[1], c[2], 1.0f };
    float e[4];
    mat_vec(w2e, p, e);
    const float depth = -e[2];
    if (!(depth > mp.znear))
        return false;
    ex = e[0];
    ey = e[1];
    z = depth;
    zs = depth + D;
    return true;
}

template <typename K, typename V>
__device__ __forceinline__ bool model_sample(const K *keys, const V *vals, uint64_t i, const ModelParams &mp,
                                             const Mat4 &w2e, float &ex, float &ey, float &z, float &zs,
                                             float c[3])
{
    const V val = vals[i];
    const K key = keys[i];      // both loads in flight before the first test
    loaded_here(val);
    loaded_here(key);
    return model_sample_of<K, V>(key, val, mp, w2e, ex, ey, z, zs, c);
}

template <typename K, typename V>
__global__ __launch_bounds__(kIT) void k_model_splat(const K *__restrict__ keys, const V *__restrict__ vals,
                                                     const uint64_t *__restrict__ d_n, ModelParams mp,
                                                     const ModelPose *__restrict__ pose,
                                                     unsigned long long *__restrict__ zbuf)
{
    const uint64_t n = *d_n;
    const Mat4 w2e = pose->w2e;
    const int W = mp.W, H = mp.H;
    for (uint64_t i = (uint64_t)blockIdx.x * kIT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kIT) {
        float ex, ey, z, zs, c[3];
        if (!model_sample<K, V>(keys, vals, i, mp, w2e, ex, ey, z, zs, c))
            continue;
        const float u = mp.fx * ex / z + mp.cx;
        const float v = mp.cy - mp.fy * ey / z;
        if (!(u > -8.0f && u < (float)W + 8.0f && v > -8.0f && v < (float)H + 8.0f))
            continue;
        const float kf = ceilf(mp.vl * mp.fx / z);
        const int k = kf < 1.0f ? 1 : (kf > 3.0f ? 3 : (int)kf);
        const float half = 0.5f * (float)(k - 1);
        const int u0 = (int)floorf(u - half), v0 = (int)floorf(v - half);
        // nearest voxel centre wins (its depth is exact geometry, so the choice
        // does not select on the noise of D); its surface estimate is zs
        const unsigned long long key = (unsigned long long)orderable_f(z) << 32 | (uint32_t)i;
        for (int dv = 0; dv < k; ++dv)
            for (int du = 0; du < k; ++du) {
                const int pu = u0 + du, pv = v0 + dv;
                if (pu < 0 || pu >= W || pv < 0 || pv >= H)
                    continue;
                atomicMin(&zbuf[(size_t)pv * W + pu], key);
            }
    }
}

// The model normal at pixel (x, y) from its point P and the points NB pixels
// right, left, down and up (central differences), facing the camera; zero at
// the border, where a neighbour is missing or further than disc from P.
constexpr int kNrmBase = 4;   // central-difference baseline in pixels
__device__ __forceinline__ float4 model_normal(const float4 &P, const float4 (&nb)[4], int x, int y, int W, int H,
                                               const ModelPose *__restrict__ pose, float disc)
{
    constexpr int NB = kNrmBase;
    float4 N = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (P.w != 0.0f && x >= NB && x < W - NB && y >= NB && y < H - NB) {
        const float disc2 = disc * disc;
        bool ok = true;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const float4 Q = nb[d];
            const float a = Q.x - P.x, b = Q.y - P.y, c = Q.z - P.z;
            ok = ok && Q.w != 0.0f && (a * a + b * b) + c * c < disc2;
        }
        if (ok) {
            const float4 r = nb[0], l = nb[1], dn = nb[2], up = nb[3];
            const float ux = r.x - l.x, uy = r.y - l.y, uz = r.z - l.z;
            const float vx = dn.x - up.x, vy = dn.y - up.y, vz = dn.z - up.z;
            float nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
            const float len = sqrtf((nx * nx + ny * ny) + nz * nz);
            if (len > 0.0f) {
                nx = nx / len; ny = ny / len; nz = nz / len;
                const float wx = P.x - pose->cam[0], wy = P.y - pose->cam[1], wz = P.z - pose->cam[2];
                if ((nx * wx + ny * wy) + nz * wz > 0.0f) {
                    nx = -nx; ny = -ny; nz = -nz;
                }
                N = make_float4(nx, ny, nz, 1.0f);
            }
        }
    }
    return N;
}

__global__ __launch_bounds__(kIT) void k_model_normals(const float4 *__restrict__ pts, int W, int H,
                                                       const ModelPose *__restrict__ pose, float disc,
                                                       float4 *__restrict__ nrm)
{
    const int q = blockIdx.x * kIT + threadIdx.x;
    if (q >= W * H)
        return;
    const int y = q / W, x = q - y * W;
    constexpr int NB = kNrmBase;
    const int dxs[4] = { NB, -NB, 0, 0 }, dys[4] = { 0, 0, NB, -NB };
    // the point and its four neighbours in flight together (clamped to the image;
    // the tests decide whether they are used)
    const float4 P = pts[q];
    float4 nb[4];
#pragma unroll
    for (int d = 0; d < 4; ++d)
        nb[d] = pts[(size_t)min(max(y + dys[d], 0), H - 1) * W + min(max(x + dxs[d], 0), W - 1)];
    nrm[q] = model_normal(P, nb, x, y, W, H, pose, disc);
}

// The model's points and normals in one launch.  A pixel's point is the
// surface estimate of the voxel its z-buffer entry names: the voxel centre
// moved along the viewing ray to depth z + D.  A workgroup resolves a
// kMTile x kMTile tile of pixels and the kNrmBase-pixel ring around it into
// LDS (every z-buffer read, then every voxel read, in flight together), then
// writes the tile's points and takes their normals from LDS (k_model_normals'
// rule).
constexpr int kMTile = 16;
constexpr int kMSpan = kMTile + 2 * kNrmBase;
constexpr int kMRounds = (kMSpan * kMSpan + kIT - 1) / kIT;
static_assert(kMTile * kMTile == kIT, "one tile pixel per thread");

template <typename K, typename V>
__global__ __launch_bounds__(kIT) void k_model_points(const K *__restrict__ keys, const V *__restrict__ vals,
                                                      ModelParams mp, const ModelPose *__restrict__ pose,
                                                      const unsigned long long *__restrict__ zbuf, float disc,
                                                      float4 *__restrict__ pts, float4 *__restrict__ nrm)
{
    __shared__ float4 sp[kMSpan * kMSpan];
    const int W = mp.W, H = mp.H;
    const int x0 = (int)blockIdx.x * kMTile - kNrmBase, y0 = (int)blockIdx.y * kMTile - kNrmBase;
    const Mat4 w2e = pose->w2e;
    unsigned long long zk[kMRounds];
#pragma unroll
    for (int r = 0; r < kMRounds; ++r) {
        const int i = r * kIT + (int)threadIdx.x;
        const int x = x0 + i % kMSpan, y = y0 + i / kMSpan;
        const bool in = i < kMSpan * kMSpan && x >= 0 && x < W && y >= 0 && y < H;
        zk[r] = zbuf[in ? (size_t)y * W + x : 0];
        if (!in)
            zk[r] = ~0ull;
    }
    K kk[kMRounds];
    V vv[kMRounds];
#pragma unroll
    for (int r = 0; r < kMRounds; ++r) {
        const uint32_t idx = zk[r] != ~0ull ? (uint32_t)zk[r] : 0u;    // index 0: read, unused
        kk[r] = keys[idx];
        vv[r] = vals[idx];
    }
#pragma unroll
    for (int r = 0; r < kMRounds; ++r) {
        const int i = r * kIT + (int)threadIdx.x;
        float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (zk[r] != ~0ull) {
            float ex, ey, z, zs, c[3];
            model_sample_of<K, V>(kk[r], vv[r], mp, w2e, ex, ey, z, zs, c);
            const float s = zs / z;
            o.x = pose->cam[0] + (c[0] - pose->cam[0]) * s;
            o.y = pose->cam[1] + (c[1] - pose->cam[1]) * s;
            o.z = pose->cam[2] + (c[2] - pose->cam[2]) * s;
            o.w = 1.0f;
        }
        if (i < kMSpan * kMSpan)
            sp[i] = o;
    }
    __syncthreads();
    const int lx = (int)threadIdx.x % kMTile + kNrmBase, ly = (int)threadIdx.x / kMTile + kNrmBase;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= W || y >= H)
        return;
    constexpr int NB = kNrmBase;
    const float4 P = sp[ly * kMSpan + lx];
    const float4 nb[4] = { sp[ly * kMSpan + lx + NB], sp[ly * kMSpan + lx - NB], sp[(ly + NB) * kMSpan + lx],
                           sp[(ly - NB) * kMSpan + lx] };
    const size_t q = (size_t)y * W + x;
    pts[q] = P;
    nrm[q] = model_normal(P, nb, x, y, W, H, pose, disc);
}

// model pose from a world->eye matrix (host value or device pointer)
__global__ void k_model_pose(Mat4 w2e_val, const Mat4 *__restrict__ w2e_dev, Mat4 clip,
                             ModelPose *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    ModelPose p;
    p.w2e = w2e_dev ? *w2e_dev : w2e_val;
    mat_mul(clip.m, p.w2e.m, p.proj.m);
    Mat4 e2w = p.w2e;
    invert_transform(e2w.m);
    p.cam[0] = e2w.m[12];
    p.cam[1] = e2w.m[13];
    p.cam[2] = e2w.m[14];
    *out = p;
}

// k_model_pose and the z-buffer clear (a fill kernel, never a memset node) in one launch
__global__ void k_model_begin(Mat4 w2e_val, const Mat4 *__restrict__ w2e_dev, Mat4 clip, ModelPose *__restrict__ out,
                              unsigned long long *__restrict__ zbuf, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        zbuf[i] = ~0ull;
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    ModelPose p;
    p.w2e = w2e_dev ? *w2e_dev : w2e_val;
    mat_mul(clip.m, p.w2e.m, p.proj.m);
    Mat4 e2w = p.w2e;
    invert_transform(e2w.m);
    p.cam[0] = e2w.m[12];
    p.cam[1] = e2w.m[13];
    p.cam[2] = e2w.m[14];
    *out = p;
}

// ------------------------------------------------------------ depth pyramid
// A pixel of pyramid level l+1: the mean of the valid (> 0) depths of its 2x2
// pixels at level l, summed in the order (0,0),(1,0),(0,1),(1,1); 0 if none is.
// k_icp_begin builds the levels.
__device__ __forceinline__ float mean_valid4(float a, float b, float c, float d)
{
    float s = 0.0f, n = 0.0f;
    const float v[4] = { a, b, c, d };
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (v[k] > 0.0f) {
            s += v[k];
            n += 1.0f;
        }
    return n > 0.0f ? s / n : 0.0f;
}


// ------------------------------------------------------------ compensated sums
struct CSum {
    float s, c;
};

__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e)
{
    s = a + b;
    const float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

// COMP = false: plain binary32 addition (the naive reduction of the precision
// study); the compensation term stays zero.
template <bool COMP = true>
__device__ __forceinline__ void cadd(CSum &x, float s2, float c2)
{
    if constexpr (COMP) {
        float t, e;
        two_sum(x.s, s2, t, e);
        x.s = t;
        x.c = x.c + c2 + e;
    } else {
        x.s = x.s + s2;
    }
}

// Compensated block reduction through an LDS transpose: every thread stores its
// NV pairs as one row; kSlices x NV threads then each fold a 1/kSlices column
// slice sequentially, and NV threads fold the slices into lds.result.  Every
// step is parallel over the values: a butterfly of cross-lane shuffles would
// chain 6 x 2 x NV dependent permutes per wave (about 9 us for NV = 29), and a
// single thread folding all slices about 5 us.  Result in lds.result.
constexpr int kSlices = 8;
static_assert(kIT / kSlices % 4 == 0, "a slice's rows split over four chains");
template <int NV>
struct CsumLds {
    CSum rows[kIT * NV];
    CSum slices[kSlices * NV];
    CSum result[NV];
};

template <int NV, bool COMP = true>
__device__ __forceinline__ void block_csum(const CSum (&v)[NV], CsumLds<NV> &lds)
{
    static_assert(kSlices * NV <= kIT, "one thread per (slice, value)");
    constexpr int kRows = kIT / kSlices;
#pragma unroll
    for (int j = 0; j < NV; ++j)
        lds.rows[threadIdx.x * NV + j] = v[j];
    __syncthreads();
    if (threadIdx.x < kSlices * NV) {
        // a slice's rows in kChains interleaved chains (row r to chain r % kChains):
        // kChains LDS reads in flight per step instead of one dependent chain
        constexpr int kChains = 4;
        const int j = threadIdx.x % NV, sl = threadIdx.x / NV;
        CSum a[kChains];
#pragma unroll
        for (int c = 0; c < kChains; ++c)
            a[c] = lds.rows[(sl * kRows + c) * NV + j];
#pragma unroll
        for (int r = kChains; r < kRows; r += kChains) {
            CSum b[kChains];
#pragma unroll
            for (int c = 0; c < kChains; ++c)
                b[c] = lds.rows[(sl * kRows + r + c) * NV + j];
#pragma unroll
            for (int c = 0; c < kChains; ++c)
                cadd<COMP>(a[c], b[c].s, b[c].c);
        }
        cadd<COMP>(a[0], a[1].s, a[1].c);
        cadd<COMP>(a[2], a[3].s, a[3].c);
        cadd<COMP>(a[0], a[2].s, a[2].c);
        lds.slices[sl * NV + j] = a[0];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        CSum sv[kSlices];
#pragma unroll
        for (int sl = 0; sl < kSlices; ++sl)
            sv[sl] = lds.slices[sl * NV + threadIdx.x];
        CSum a = sv[0];
#pragma unroll
        for (int sl = 1; sl < kSlices; ++sl)
            cadd<COMP>(a, sv[sl].s, sv[sl].c);
        lds.result[threadIdx.x] = a;
    }
    __syncthreads();
}

// ------------------------------------------------------------ association
// NV = 29 for point-to-plane (21 upper-triangular A, 6 b, count, r^2),
//      17 for Horn (sum p (3), sum q (3), sum p q^T (9), count, r^2) about origin o.
template <int MODE>
struct IcpNV;
template <> struct IcpNV<KIFI_ICP_POINT_TO_PLANE> { static constexpr int nv = 30; };
template <> struct IcpNV<KIFI_ICP_HORN> { static constexpr int nv = 17; };

template <int MODE, bool COMP>
__device__ void icp_finish(const CSum *partials, int nblocks, IcpState *st, int level, const IcpSolveParams &sp,
                           const ModelPose *mpose, CsumLds<IcpNV<MODE>::nv> &lds);

// Block partial -> device memory; the workgroup that takes the last ticket
// reduces all partials and runs the FP64 solve (no spin-wait anywhere).
// Ordering: each workgroup's partial store is released at agent scope before
// its ticket increment (gfx950: L2 write-back, no invalidate); only the last
// workgroup acquires (one L2 invalidate per launch).  A full __threadfence()
// here would invalidate the XCD's L2 once per workgroup and make every other
// workgroup re-fetch the model maps from HBM.
// GEN (the multi-iteration kernel): nblocks workgroups take part (the first
// ones of the grid), and the solving workgroup then releases its state writes
// and sets st->gen = gen, which all workgroups wait for.
template <int MODE, bool COMP, int NV, bool GEN = false>
__device__ __forceinline__ void icp_publish(CSum (&acc)[NV], CsumLds<NV> &lds, CSum *partials, IcpState *st, int level,
                                            const IcpSolveParams &sp, const ModelPose *mpose, unsigned gen = 0,
                                            int nblocks = 0, unsigned *gen_ctr = nullptr)
{
    __shared__ bool last;
    const unsigned nb = GEN ? (unsigned)nblocks : gridDim.x;
    block_csum<NV, COMP>(acc, lds);
    if (threadIdx.x < NV)
        partials[(size_t)blockIdx.x * NV + threadIdx.x] = lds.result[threadIdx.x];
    __syncthreads();     // the stores of threads 0..NV-1 precede thread 0's release
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned t = __hip_atomic_fetch_add(&st->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = t == nb - 1;
    }
    __syncthreads();
    if (!last)
        return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // see every workgroup's partial
    icp_finish<MODE, COMP>(partials, (int)nb, st, level, sp, mpose, lds);
    if (threadIdx.x == 0) {
        st->ticket = 0;
        if (GEN) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(gen_ctr ? gen_ctr : &st->gen, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// One association pass of a level with the live pose `pose`, strided over
// nblocks workgroups (workgroups beyond them add nothing): the correspondence
// sums of this thread in acc.
template <int MODE, bool COMP, int NV>
__device__ __forceinline__ void icp_accumulate(const IcpLevel &lv, const ModelParams &mp,
                                               const ModelPose *__restrict__ mpose, const float4 *__restrict__ mpts,
                                               const float4 *__restrict__ mnrm, const Mat4 &pose, float dist_thresh,
                                               int nblocks, CSum (&acc)[NV])
{
#pragma unroll
    for (int j = 0; j < NV; ++j)
        acc[j] = CSum{ 0.0f, 0.0f };
    const int P = (int)blockIdx.x < nblocks ? lv.W * lv.H : 0;
    for (int q = blockIdx.x * kIT + threadIdx.x; q < P; q += nblocks * kIT) {
        const int v = q / lv.W, u = q - v * lv.W;
        const float d = lv.depth[q];
        if (d > 0.0f) {
            const float pe[4] = { ((float)u + 0.5f - lv.cx) / lv.fx * d, (lv.cy - (float)v - 0.5f) / lv.fy * d, -d,
                                  1.0f };
            float p[4];
            mat_vec(pose, pe, p);
            float e[4];
            mat_vec(mpose->w2e, p, e);
            const float z = -e[2];
            if (z > 0.0f) {
                const float fu = mp.fx * e[0] / z + mp.cx, fv = mp.cy - mp.fy * e[1] / z;
                if (fu >= 0.0f && fu < (float)mp.W && fv >= 0.0f && fv < (float)mp.H) {
                    const size_t mq = (size_t)(int)fv * mp.W + (int)fu;
                    const float4 Q = mpts[mq], N = mnrm[mq];
                    if (Q.w != 0.0f && N.w != 0.0f) {
                        const float dx = Q.x - p[0], dy = Q.y - p[1], dz = Q.z - p[2];
                        if ((dx * dx + dy * dy) + dz * dz <= dist_thresh * dist_thresh) {
                            if (MODE == KIFI_ICP_POINT_TO_PLANE) {
                                const float r = (N.x * dx + N.y * dy) + N.z * dz;
                                const float J[6] = { p[1] * N.z - p[2] * N.y, p[2] * N.x - p[0] * N.z,
                                                     p[0] * N.y - p[1] * N.x, N.x, N.y, N.z };
                                // inverse variance of the sensor noise model sigma ~ z^2
                                const float iz = 1.0f / d, wgt = iz * iz * iz * iz;
                                int t = 0;
#pragma unroll
                                for (int i = 0; i < 6; ++i)
#pragma unroll
                                    for (int j = 0; j <= i; ++j)
                                    {
                                        cadd<COMP>(acc[t], wgt * J[i] * J[j], 0.0f);
                                        ++t;
                                    }
#pragma unroll
                                for (int i = 0; i < 6; ++i)
                                    cadd<COMP>(acc[21 + i], wgt * J[i] * r, 0.0f);
                                cadd<COMP>(acc[27], 1.0f, 0.0f);
                                cadd<COMP>(acc[28], r * r, 0.0f);
                                cadd<COMP>(acc[29], wgt * (r * r), 0.0f);
                            } else {
                                const float a[3] = { p[0] - mpose->cam[0], p[1] - mpose->cam[1], p[2] - mpose->cam[2] };
                                const float b[3] = { Q.x - mpose->cam[0], Q.y - mpose->cam[1], Q.z - mpose->cam[2] };
#pragma unroll
                                for (int i = 0; i < 3; ++i) {
                                    cadd<COMP>(acc[i], a[i], 0.0f);
                                    cadd<COMP>(acc[3 + i], b[i], 0.0f);
#pragma unroll
                                    for (int j = 0; j < 3; ++j)
                                        cadd<COMP>(acc[6 + 3 * i + j], a[i] * b[j], 0.0f);
                                }
                                cadd<COMP>(acc[15], 1.0f, 0.0f);
                                cadd<COMP>(acc[16], (dx * dx + dy * dy) + dz * dz, 0.0f);
                            }
                        }
                    }
                }
            }
        }
    }
}

template <int MODE, bool COMP>
__global__ __launch_bounds__(kIT) void k_icp_assoc(IcpLevel lv, ModelParams mp, const ModelPose *__restrict__ mpose,
                                                   const float4 *__restrict__ mpts,
                                                   const float4 *__restrict__ mnrm,
                                                   IcpState *__restrict__ st, float dist_thresh,
                                                   CSum *__restrict__ partials, int level, IcpSolveParams sp)
{
    constexpr int NV = IcpNV<MODE>::nv;
    constexpr int MODE_OF_KERNEL = MODE;
    __shared__ CsumLds<NV> lds;
    if (st->level_done[level])
        return;
    CSum acc[NV];
    icp_accumulate<MODE, COMP>(lv, mp, mpose, mpts, mnrm, st->pose_f, dist_thresh, (int)gridDim.x, acc);
    icp_publish<MODE_OF_KERNEL, COMP>(acc, lds, partials, st, level, sp, mpose);
}

// ICP::AlignStep association (ICP.cpp:66-131): every 4th column, back-projected
// in the reference's form, projected with EyeToClipRH * dstWorldToEye, nearest
// model point of a 21 x 21 window within 10 cm (strictly smaller wins, scan order).
// The window clamps to the model map (the reference hard-codes 640 x 480).
// Sums about the model camera centre o: sum a, sum b, sum a b^T, count, sum |b-a|^2.
__global__ __launch_bounds__(kIT) void k_icp_assoc_window(IcpLevel lv, ModelParams mp,
                                                          const ModelPose *__restrict__ mpose,
                                                          const float4 *__restrict__ mpts,
                                                          IcpState *__restrict__ st,
                                                          CSum *__restrict__ partials, int level,
                                                          IcpSolveParams sp)
{
    constexpr int NV = IcpNV<KIFI_ICP_HORN>::nv;
    constexpr int MODE_OF_KERNEL = KIFI_ICP_HORN;
    __shared__ CsumLds<NV> lds;
    if (st->level_done[level])
        return;
    CSum acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j)
        acc[j] = CSum{ 0.0f, 0.0f };
    const int cols = (lv.W + 3) / 4;
    for (int q = blockIdx.x * kIT + threadIdx.x; q < cols * lv.H; q += gridDim.x * kIT) {
        const int y = q / cols, x = 4 * (q - y * cols);
        const float d = lv.depth[(size_t)y * lv.W + x];
        if (0.0f != d) {
            const float flInvX = 1.0f / lv.fx, flInvY = 1.0f / lv.fy;     // :50-57
            const float ppX = (0.5f - lv.cx) * flInvX, ppY = (lv.cy - 0.5f) * flInvY;
            float pt[4] = { (float)x, -(float)y, -1.0f, 0.0f };          // :74-77
            pt[0] = (pt[0] * flInvX + ppX) * d;
            pt[1] = (pt[1] * flInvY + ppY) * d;
            pt[2] = (pt[2] * 1.0f + 0.0f) * d;
            pt[3] = 1.0f;
            float p[4];
            mat_vec(st->pose_f, pt, p);                                   // :79
            float uv[4];
            mat_vec(mpose->proj, p, uv);                                  // :81
            const float w = uv[3];
            const float ux = uv[0] / w, uy = uv[1] / w;
            const float halfW = (float)mp.W * 0.5f, halfH = (float)mp.H * 0.5f;
            const float fu = ux * halfW + halfW, fv = uy * halfH + halfH;
            if (fu > -2147483648.0f && fu < 2147483648.0f && fv > -2147483648.0f && fv < 2147483648.0f) {
                int u = (int)fu - 10;                                     // :100-104
                int v = mp.H - 1 - (int)fv - 10;
                u = u < 0 ? 0 : u;
                v = v < 0 ? 0 : v;
                const int yend = min(mp.H, v + 21), xend = min(mp.W, u + 21);
                float mind = 0.01f;
                float4 best = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                for (int y1 = v; y1 < yend; ++y1)                         // :108-125
                    for (int x1 = u; x1 < xend; ++x1) {
                        const float4 Q = mpts[(size_t)y1 * mp.W + x1];
                        if (Q.x == 0.0f && Q.y == 0.0f && Q.z == 0.0f)
                            continue;
                        const float dx = Q.x - p[0], dy = Q.y - p[1], dz = Q.z - p[2], dw = 1.0f - p[3];
                        const float l2 = (dx * dx + dy * dy) + (dz * dz + dw * dw);
                        if (l2 < mind) {
                            mind = l2;
                            best = Q;
                        }
                    }
                if (mind < 0.01f) {                                       // :126-131
                    const float a[3] = { p[0] - mpose->cam[0], p[1] - mpose->cam[1], p[2] - mpose->cam[2] };
                    const float b[3] = { best.x - mpose->cam[0], best.y - mpose->cam[1], best.z - mpose->cam[2] };
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        cadd(acc[i], a[i], 0.0f);
                        cadd(acc[3 + i], b[i], 0.0f);
#pragma unroll
                        for (int j = 0; j < 3; ++j)
                            cadd(acc[6 + 3 * i + j], a[i] * b[j], 0.0f);
                    }
                    cadd(acc[15], 1.0f, 0.0f);
                    cadd(acc[16], mind, 0.0f);
                }
            }
        }
    }
    icp_publish<MODE_OF_KERNEL, true>(acc, lds, partials, st, level, sp, mpose);
}

// ---------------------------------------------------------------- FP64 solve
// One thread solves A x = b: six FP64 divisions (the reciprocals of the
// diagonal) instead of 27; each one is a long dependent sequence on the
// solving thread, which the other workgroups' next launch waits for.
__device__ bool cholesky6(const double *A, const double *b, double *x)
{
    double L[36], inv[6];
    for (int i = 0; i < 36; ++i)
        L[i] = 0.0;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = A[6 * i + j];
            for (int k = 0; k < j; ++k)
                s -= L[6 * i + k] * L[6 * j + k];
            if (i == j) {
                if (!(s > 0.0))
                    return false;
                L[6 * i + i] = sqrt(s);
                inv[i] = 1.0 / L[6 * i + i];
            } else {
                L[6 * i + j] = s * inv[j];
            }
        }
    double y[6];
    for (int i = 0; i < 6; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k)
            s -= L[6 * i + k] * y[k];
        y[i] = s * inv[i];
    }
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < 6; ++k)
            s -= L[6 * k + i] * x[k];
        x[i] = s * inv[i];
    }
    return true;
}

__device__ void twist_to_mat(const double xi[6], double *M)
{
    const double wx = xi[0], wy = xi[1], wz = xi[2];
    const double th = sqrt(wx * wx + wy * wy + wz * wz);
    const double K[9] = { 0, wz, -wy, -wz, 0, wx, wy, -wx, 0 };
    double a, b;
    if (th < 1e-12) {
        a = 1.0;
        b = 0.5;
    } else {
        double s, c;
        sincos(th, &s, &c);
        const double it = 1.0 / th;
        a = s * it;
        b = (1.0 - c) * (it * it);
    }
    for (int i = 0; i < 16; ++i)
        M[i] = 0.0;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            double k2 = 0.0;
            for (int k = 0; k < 3; ++k)
                k2 += K[r + 3 * k] * K[k + 3 * c];
            M[r + 4 * c] = (r == c ? 1.0 : 0.0) + a * K[r + 3 * c] + b * k2;
        }
    M[12] = xi[3];
    M[13] = xi[4];
    M[14] = xi[5];
    M[15] = 1.0;
}

__device__ void mat_mul_d(const double *A, const double *B, double *C)
{
    double R[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k)
                s += A[r + 4 * k] * B[k + 4 * c];
            R[r + 4 * c] = s;
        }
    for (int i = 0; i < 16; ++i)
        C[i] = R[i];
}

// cyclic Jacobi eigen decomposition of a symmetric 4x4 (FP64); V columns = vectors
__device__ void jacobi4(double A[4][4], double V[16], double ev[4])
{
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            V[r + 4 * c] = r == c ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                if (r != c) off += A[r][c] * A[r][c];
                else diag += A[r][c] * A[r][c];
            }
        if (off <= 1e-30 * diag || off == 0.0)
            break;
        for (int p = 0; p < 3; ++p)
            for (int q = p + 1; q < 4; ++q) {
                const double apq = A[p][q];
                if (apq == 0.0)
                    continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 4; ++k) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 4; ++k) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 4; ++k) {
                    const double vkp = V[k + 4 * p], vkq = V[k + 4 * q];
                    V[k + 4 * p] = c * vkp - s * vkq;
                    V[k + 4 * q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < 4; ++i)
        ev[i] = A[i][i];
}

// Horn's closed form (ICP.cpp:199-264) from centred cross-covariance, FP64:
// returns the rigid increment D mapping the src points onto the dst points.
__device__ void horn_increment(const double S[9], const double ms[3], const double md[3], double D[16])
{
    const double Sxx = S[0], Sxy = S[1], Sxz = S[2], Syx = S[3], Syy = S[4], Syz = S[5], Szx = S[6],
                 Szy = S[7], Szz = S[8];
    double N[4][4] = { { Sxx + Syy + Szz, Syz - Szy, Szx - Sxz, Sxy - Syx },
                       { Syz - Szy, Sxx - Syy - Szz, Sxy + Syx, Szx + Sxz },
                       { Szx - Sxz, Sxy + Syx, Syy - Sxx - Szz, Syz + Szy },
                       { Sxy - Syx, Szx + Sxz, Syz + Szy, Szz - Sxx - Syy } };
    double V[16], ev[4];
    jacobi4(N, V, ev);
    int im = 0;
    for (int i = 1; i < 4; ++i)
        if (ev[i] > ev[im])
            im = i;
    double q[4] = { V[4 * im], V[4 * im + 1], V[4 * im + 2], V[4 * im + 3] };
    const double len = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int i = 0; i < 4; ++i)
        q[i] /= len;
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    double R[9];
    R[0] = 1 - 2 * (z * z + w * w); R[3] = 2 * (y * z - x * w); R[6] = 2 * (y * w + x * z);
    R[1] = 2 * (y * z + x * w); R[4] = 1 - 2 * (y * y + w * w); R[7] = 2 * (z * w - x * y);
    R[2] = 2 * (y * w - x * z); R[5] = 2 * (z * w + x * y); R[8] = 1 - 2 * (y * y + z * z);
    for (int i = 0; i < 16; ++i)
        D[i] = 0.0;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            D[r + 4 * c] = R[r + 3 * c];
    for (int r = 0; r < 3; ++r)
        D[12 + r] = md[r] - (R[r] * ms[0] + R[r + 3] * ms[1] + R[r + 6] * ms[2]);
    D[15] = 1.0;
}

// Motion prior (kifi_icp_params.motion_prior).  A world-frame increment
// (w, v) moves the camera centre c by w x c + v to first order; the prior
// measures the motion since the pass's start as m = M xi with
// M = [[I, 0], [-[c]x, I]] (c: the start's centre), accumulated over the
// pass's solves, and each solve minimises |J xi - r|^2 + mu |m_acc + M xi|^2:
// (A + mu M^T M) xi = b - mu M^T m_acc.
__device__ void prior_reset(IcpState *st, const double *start)
{
    for (int i = 0; i < 6; ++i)
        st->prior_acc[i] = 0.0;
    for (int i = 0; i < 3; ++i)
        st->prior_c[i] = start[12 + i];
}

__device__ void prior_matrix(const double c[3], double M[36])
{
    for (int i = 0; i < 36; ++i)
        M[i] = 0.0;
    for (int i = 0; i < 6; ++i)
        M[6 * i + i] = 1.0;
    // rows 3..5, columns 0..2: -[c]x
    M[6 * 3 + 1] = c[2];
    M[6 * 3 + 2] = -c[1];
    M[6 * 4 + 0] = -c[2];
    M[6 * 4 + 2] = c[0];
    M[6 * 5 + 0] = c[1];
    M[6 * 5 + 1] = -c[0];
}

__device__ void prior_apply(const IcpState *st, double mu, double A[36], double b[6])
{
    double M[36];
    prior_matrix(st->prior_c, M);
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k)
            s += M[6 * k + i] * st->prior_acc[k];
        b[i] -= mu * s;
        for (int j = 0; j < 6; ++j) {
            double h = 0.0;
            for (int k = 0; k < 6; ++k)
                h += M[6 * k + i] * M[6 * k + j];
            A[6 * i + j] += mu * h;
        }
    }
}

__device__ void prior_accumulate(IcpState *st, const double xi[6])
{
    double M[36];
    prior_matrix(st->prior_c, M);
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (int j = 0; j < 6; ++j)
            s += M[6 * i + j] * xi[j];
        st->prior_acc[i] += s;
    }
}

// Reduce the block partials (compensated), solve in FP64, compose the pose.
// Run by the whole last workgroup of an association kernel; the partials of the
// other workgroups are read with device-scope loads after the ticket.
template <int MODE, bool COMP>
__device__ void icp_finish(const CSum *partials, int nblocks, IcpState *st, int level, const IcpSolveParams &sp,
                           const ModelPose *mpose, CsumLds<IcpNV<MODE>::nv> &lds)
{
    constexpr int NV = IcpNV<MODE>::nv;
    CSum acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j)
        acc[j] = CSum{ 0.0f, 0.0f };
    // plain loads: the caller's device-scope acquire fence after the ticket has
    // invalidated this CU's L1, so they read the other workgroups' partials from L2
    for (int b = threadIdx.x; b < nblocks; b += kIT)
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const CSum p = partials[(size_t)b * NV + j];
            cadd<COMP>(acc[j], p.s, p.c);
        }
    block_csum<NV, COMP>(acc, lds);
    if (threadIdx.x != 0)
        return;
    double v[NV];
    for (int j = 0; j < NV; ++j) {
        v[j] = (double)lds.result[j].s + (double)lds.result[j].c;
        st->sums[j] = v[j];
    }
    // count, sum r^2 and (point-to-plane) sum r^2 / d^4: the residual scaled to
    // 1 m depth by the sensor noise model sigma ~ z^2
    constexpr int IC = MODE == KIFI_ICP_POINT_TO_PLANE ? 27 : 15;
    const double cnt = v[IC], e2 = v[IC + 1], ew2 = v[MODE == KIFI_ICP_POINT_TO_PLANE ? IC + 2 : IC + 1];
    const double icnt = cnt > 0.0 ? 1.0 / cnt : 0.0;
    st->rms = (float)sqrt(e2 * icnt);
    st->rms_1m = (float)sqrt(ew2 * icnt);
    st->inliers = (int)cnt;
    if (cnt < (double)sp.min_inliers) {
        st->level_done[level] = 1;
        st->failed = 1;
        return;
    }
    double D[16], xi[6];
    bool ok;
    if (MODE == KIFI_ICP_POINT_TO_PLANE) {
        double A[36], b[6];
        int t = 0;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j <= i; ++j) {
                A[6 * i + j] = v[t];
                A[6 * j + i] = v[t];
                ++t;
            }
        for (int i = 0; i < 6; ++i)
            b[i] = v[21 + i];
        if (sp.prior > 0.0)
            prior_apply(st, sp.prior * cnt, A, b);
        ok = cholesky6(A, b, xi);
        if (ok) {
            twist_to_mat(xi, D);
            if (sp.prior > 0.0)
                prior_accumulate(st, xi);
        }
    } else {
        const double ms[3] = { v[0] / cnt, v[1] / cnt, v[2] / cnt }, md[3] = { v[3] / cnt, v[4] / cnt, v[5] / cnt };
        double S[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                S[3 * i + j] = v[6 + 3 * i + j] - cnt * ms[i] * md[j];
        double Dl[16];
        horn_increment(S, ms, md, Dl);
        // increment about the model camera origin o: D = T(o) Dl T(-o)
        const double o[3] = { mpose->cam[0], mpose->cam[1], mpose->cam[2] };
        for (int i = 0; i < 16; ++i)
            D[i] = Dl[i];
        for (int r = 0; r < 3; ++r)
            D[12 + r] = Dl[12 + r] + o[r] - (Dl[r] * o[0] + Dl[r + 4] * o[1] + Dl[r + 8] * o[2]);
        // twist magnitude for the convergence test: rotation angle and translation
        const double tr = Dl[0] + Dl[5] + Dl[10];
        const double ang = acos(fmin(1.0, fmax(-1.0, 0.5 * (tr - 1.0))));
        xi[0] = ang; xi[1] = 0; xi[2] = 0;
        xi[3] = D[12]; xi[4] = D[13]; xi[5] = D[14];
        ok = true;
    }
    if (!ok) {
        st->level_done[level] = 1;
        st->failed = 1;
        return;
    }
    mat_mul_d(D, st->pose, st->pose);
    for (int i = 0; i < 16; ++i)
        st->pose_f.m[i] = (float)st->pose[i];
    st->iters[level] += 1;
    st->total_iters += 1;
    // |w| < eps_rot and |t| < eps_trans, compared as squares
    const double rot2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
    const double tr2 = xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5];
    if (rot2 < sp.eps_rot * sp.eps_rot && tr2 < sp.eps_trans * sp.eps_trans)
        st->level_done[level] = 1;
}

// state control kernels
// Level-1 pixel (x, y) of the pyramid straight from the raw frame: mean_valid4
// of its 2x2 level-0 depths in metres (the values k_icp_begin stores at level 0).
template <typename D>
__device__ __forceinline__ float pyramid_l1(const D *depth, int W, int x, int y)
{
    const size_t r0 = (size_t)(2 * y) * W + 2 * x, r1 = r0 + W;
    return mean_valid4(depth_m(depth, r0), depth_m(depth, r0 + 1), depth_m(depth, r1), depth_m(depth, r1 + 1));
}

// Frame start, in one launch: the state from the guess (one thread), the
// level-0 depth map in metres (DepthStream.cpp:89-90) and, with LV levels, the
// coarser ones (each pixel the mean_valid4 of its 2x2 pixels one level
// finer).  Levels 1 and 2 are computed from the raw
// frame, not from stored level-0 values, so nothing in the launch waits on
// another thread's store: threads [0, n1) write level 1, threads [n1, n1 + n2)
// level 2 (each recomputing its four level-1 values).  The grid has at least
// W*H >= n1 + n2 threads.
template <typename D, int LV>
__global__ void k_icp_begin(IcpState *__restrict__ st, const double *__restrict__ guess_dev, Pose64 guess_val,
                            const D *__restrict__ depth, int W, int H, float *__restrict__ lv0,
                            float *__restrict__ lv1, float *__restrict__ lv2)
{
    const int n = W * H;
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    for (int i = gt; i < n; i += gridDim.x * blockDim.x)
        lv0[i] = depth_m(depth, i);
    if (LV >= 2) {
        const int W1 = W / 2, n1 = W1 * (H / 2);
        if (gt < n1) {
            const int y = gt / W1, x = gt - y * W1;
            lv1[gt] = pyramid_l1(depth, W, x, y);
        }
        if (LV == 3) {
            const int W2 = W / 4, q = gt - n1;
            if (q >= 0 && q < W2 * (H / 4)) {
                const int y = q / W2, x = q - y * W2;
                lv2[q] = mean_valid4(pyramid_l1(depth, W, 2 * x, 2 * y), pyramid_l1(depth, W, 2 * x + 1, 2 * y),
                                     pyramid_l1(depth, W, 2 * x, 2 * y + 1),
                                     pyramid_l1(depth, W, 2 * x + 1, 2 * y + 1));
            }
        }
    }
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    const double *g = guess_dev ? guess_dev : guess_val.m;
    for (int i = 0; i < 16; ++i) {
        st->pose[i] = g[i];
        st->guess[i] = g[i];
        st->pose_f.m[i] = (float)g[i];
    }
    for (int l = 0; l < 4; ++l)
        st->iters[l] = 0;
    st->total_iters = 0;
    for (int l = 0; l < 4; ++l)
        st->level_done[l] = 0;
    st->failed = 0;
    st->fallback = 0;
    st->rms = 0.0f;
    st->rms_1m = 0.0f;
    st->inliers = 0;
    st->ticket = 0;
    for (int i = 0; i < 6; ++i)
        st->tail_gen[i] = 0;   // each level tail of the frame counts its hand-offs from 0
    prior_reset(st, g);
}


// Fallback gate (no host round trip): when the first pass started from an
// external guess and its 1 m-scaled residual or its iteration count exceeds the
// limits, the second pass re-runs plain ICP from the previous frame's pose; the
// first pass's result is kept aside for fallback_pick.
__device__ __forceinline__ bool fallback_needed(const IcpState *st, float max_rms, int max_iters)
{
    return st->failed || (max_rms > 0.0f && st->rms_1m > max_rms) || (max_iters > 0 && st->total_iters >= max_iters);
}

__device__ void fallback_setup(IcpState *st)
{
    prior_reset(st, st->prev);
    st->fallback = 1;
    st->alt_rms = st->failed ? FLT_MAX : st->rms_1m;
    st->alt_inliers = st->inliers;
    st->failed = 0;
    for (int i = 0; i < 16; ++i) {
        st->alt_pose[i] = st->pose[i];
        st->pose[i] = st->prev[i];
        st->pose_f.m[i] = (float)st->prev[i];
    }
    for (int l = 0; l < 4; ++l)
        st->level_done[l] = 0;
}

// ICP judges the two starts: after a fallback pass, the first result stands
// unless it failed or broke the residual limit and the fallback's residual is
// lower (a converged pass with an acceptable residual is never replaced: a
// far-off start can reach a wrong minimum with a smaller residual).
__device__ void fallback_pick(IcpState *st, float max_rms)
{
    const float second = st->failed ? FLT_MAX : st->rms_1m;
    const bool first_bad = st->alt_rms == FLT_MAX || (max_rms > 0.0f && st->alt_rms > max_rms);
    if (!first_bad || st->alt_rms <= second) {
        for (int i = 0; i < 16; ++i) {
            st->pose[i] = st->alt_pose[i];
            st->pose_f.m[i] = (float)st->alt_pose[i];
        }
        st->rms_1m = st->alt_rms;
        st->inliers = st->alt_inliers;
        st->failed = 0;
    }
}

// Gate of the second pass when it runs as one launch per iteration: a good
// first pass marks every level done, which turns those launches into early exits.
__global__ void k_icp_fallback_check(IcpState *__restrict__ st, float max_rms, int max_iters)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    if (fallback_needed(st, max_rms, max_iters)) {
        fallback_setup(st);
    } else {
        for (int l = 0; l < 4; ++l)
            st->level_done[l] = 1;
    }
}

__global__ void k_icp_fallback_pick(IcpState *__restrict__ st, float max_rms)
{
    if (threadIdx.x != 0 || blockIdx.x != 0 || !st->fallback)
        return;
    fallback_pick(st, max_rms);
}

// ------------------------------------------------ fallback pass in one launch
// The whole fallback pass (levels levels-1 down to 0) in one launch, so that a
// frame whose first pass is good pays one early-exit launch for it.  Every
// workgroup evaluates the gate on the first pass's state; when it holds, the
// last workgroup to arrive (so no workgroup can still be reading the gate's
// inputs) sets the state up.  Every workgroup is resident (the host launches
// at most one per CU after checking that a CU holds one), so a workgroup can
// wait for the others: each iteration ends in icp_publish's ticket reduction,
// whose solving workgroup releases the new state and then publishes st->gen;
// the others poll gen with relaxed loads and read the pose and level flag with
// coherent loads once it moves, so a converged level ends the loop for every
// workgroup at the same point.  Each level runs on the workgroup count that
// its launch-per-iteration kernel has (the first nb[l] of the grid; the others
// only wait), so the sums are bit-identical to a pass launched iteration by
// iteration.  Workgroup 0 judges the two passes at the end.  Waits are bounded:
// a workgroup that polls spin_limit times counts a timeout and leaves; the
// others then miss its ticket and time out in turn, so the launch always ends
// (and the host reports the timeout).  In-kernel iterations cost more than
// launched ones here (the hand-off is slower than a kernel boundary), which is
// why only the rarely needed second pass runs this way.
struct IcpFallback {
    IcpLevel lv[3];
    IcpSolveParams sp[3];
    float    gate[3];
    int      nb[3];             // workgroups of a level's association (as one launch per iteration has)
    int      iters[3];
    int      levels;
    float    max_rms;
    int      max_iters;
    unsigned spin_limit;
};

// Waits (thread 0 polling, the workgroup at the barrier) until st->gen reaches
// target; false after spin_limit polls.  No acquire fence follows: the state a
// waiting workgroup needs next (pose and level flag) it reads with coherent
// atomic loads after the wait, which the solving workgroup's release made
// visible before it moved gen.  An agent-scope acquire here would invalidate the
// L2 of the XCD once per wave, evicting the model maps that the other
// workgroups are gathering from.
__device__ bool icp_wait_gen(IcpState *st, unsigned target, unsigned spin_limit, unsigned *gen_ctr = nullptr)
{
    __shared__ int ok;
    if (threadIdx.x == 0) {
        int good = 1;
        unsigned n = 0;
        while ((int)(__hip_atomic_load(gen_ctr ? gen_ctr : &st->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
                     target) < 0) {
            if (++n > spin_limit) {
                good = 0;
                __hip_atomic_fetch_add(&st->spin_timeouts, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

template <bool COMP>
__global__ __launch_bounds__(kIT) void k_icp_fallback_pass(IcpFallback a, ModelParams mp,
                                                           const ModelPose *__restrict__ mpose,
                                                           const float4 *__restrict__ mpts,
                                                           const float4 *__restrict__ mnrm, IcpState *st,
                                                           CSum *__restrict__ partials)
{
    constexpr int MODE = KIFI_ICP_POINT_TO_PLANE;
    constexpr int NV = IcpNV<MODE>::nv;
    __shared__ CsumLds<NV> lds;
    __shared__ float s_pose[16];
    __shared__ int s_done;
    if (!fallback_needed(st, a.max_rms, a.max_iters))
        return;
    unsigned gen = __hip_atomic_load(&st->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0) {
        const unsigned t = __hip_atomic_fetch_add(&st->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t == gridDim.x - 1) {
            fallback_setup(st);
            st->ticket = 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(&st->gen, gen + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (!icp_wait_gen(st, ++gen, a.spin_limit))
        return;
    for (int l = a.levels - 1; l >= 0; --l) {
        for (int it = 0; it < a.iters[l]; ++it) {
            if (threadIdx.x < 16)
                s_pose[threadIdx.x] = __hip_atomic_load(&st->pose_f.m[threadIdx.x], __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            else if (threadIdx.x == 16)
                s_done = __hip_atomic_load(&st->level_done[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __syncthreads();
            const bool done = s_done != 0;
            Mat4 pose;
#pragma unroll
            for (int i = 0; i < 16; ++i)
                pose.m[i] = s_pose[i];
            __syncthreads();     // read by every wave before the next level's loads rewrite them
            if (done)
                break;
            ++gen;
            if ((int)blockIdx.x < a.nb[l]) {   // the rest of the grid has no pixels at this level: it only waits
                CSum acc[NV];
                icp_accumulate<MODE, COMP>(a.lv[l], mp, mpose, mpts, mnrm, pose, a.gate[l], a.nb[l], acc);
                icp_publish<MODE, COMP, NV, true>(acc, lds, partials, st, l, a.sp[l], mpose, gen, a.nb[l]);
            }
            if (!icp_wait_gen(st, gen, a.spin_limit))
                return;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // the last solve's state
        fallback_pick(st, a.max_rms);
    }
}

// ------------------------------------------------ the rest of a level in one launch
// A pass launches the first iterations of a level one by one and the rest of
// the level as one launch: a level usually converges within its first
// iterations, and each launch that only finds the level converged costs a
// dispatch.  The launch has the level's own workgroup count (every workgroup
// takes part, so the sums equal those of launched iterations bit for bit) and
// hands iterations over as the fallback pass does: the solving workgroup of
// icp_publish releases the state and moves st->gen, the others poll it
// (bounded by spin_limit) and read the pose and the level flag with coherent
// loads.  Each tail launch of a frame has a counter of its own
// (st->tail_gen[pass * 3 + level]), which k_icp_begin clears: it starts at 0,
// so no workgroup reads a counter that another may already have moved, and
// nothing needs resetting on the way out.
template <bool COMP>
__global__ __launch_bounds__(kIT) void k_icp_level_tail(IcpLevel lv, ModelParams mp,
                                                        const ModelPose *__restrict__ mpose,
                                                        const float4 *__restrict__ mpts,
                                                        const float4 *__restrict__ mnrm, IcpState *st, float gate,
                                                        CSum *__restrict__ partials, int level, IcpSolveParams sp,
                                                        int n_iters, unsigned spin_limit, int slot)
{
    constexpr int MODE = KIFI_ICP_POINT_TO_PLANE;
    constexpr int NV = IcpNV<MODE>::nv;
    __shared__ CsumLds<NV> lds;
    __shared__ float s_pose[16];
    __shared__ int s_done;
    unsigned gen = 0;
    for (int it = 0; it < n_iters; ++it) {
        if (threadIdx.x < 16)
            s_pose[threadIdx.x] = __hip_atomic_load(&st->pose_f.m[threadIdx.x], __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        else if (threadIdx.x == 16)
            s_done = __hip_atomic_load(&st->level_done[level], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const bool done = s_done != 0;
        Mat4 pose;
#pragma unroll
        for (int i = 0; i < 16; ++i)
            pose.m[i] = s_pose[i];
        __syncthreads();     // read by every wave before the next iteration rewrites them
        if (done)
            break;
        ++gen;
        CSum acc[NV];
        icp_accumulate<MODE, COMP>(lv, mp, mpose, mpts, mnrm, pose, gate, (int)gridDim.x, acc);
        icp_publish<MODE, COMP, NV, true>(acc, lds, partials, st, level, sp, mpose, gen, (int)gridDim.x,
                                          &st->tail_gen[slot]);
        if (!icp_wait_gen(st, gen, spin_limit, &st->tail_gen[slot]))
            break;
    }
}

// The rest of level level_t (as k_icp_level_tail, on that level's own
// workgroup count ntb: the first ntb workgroups of the grid take part, the
// others wait) and then the first iteration of the next finer level (as
// k_icp_assoc, on the whole grid).  A frame whose level level_t converged
// within its launched iterations, the common case, reads the flag and goes
// straight on: one launch for the tail and the next level's first iteration
// instead of two.  The sums of both parts are those of their own launches bit
// for bit.
template <bool COMP>
__global__ __launch_bounds__(kIT) void k_icp_tail_assoc(IcpLevel lt, float gate_t, IcpSolveParams sp_t, int level_t,
                                                        int n_iters, int ntb, int slot, IcpLevel lv, float gate,
                                                        IcpSolveParams sp, int level, ModelParams mp,
                                                        const ModelPose *__restrict__ mpose,
                                                        const float4 *__restrict__ mpts,
                                                        const float4 *__restrict__ mnrm, IcpState *st,
                                                        CSum *__restrict__ partials, unsigned spin_limit)
{
    constexpr int MODE = KIFI_ICP_POINT_TO_PLANE;
    constexpr int NV = IcpNV<MODE>::nv;
    __shared__ CsumLds<NV> lds;
    __shared__ float s_pose[16];
    __shared__ int s_done, s_next_done;
    unsigned gen = 0;
    for (int it = 0;; ++it) {
        // the state after the last hand-off (coherent loads)
        if (threadIdx.x < 16)
            s_pose[threadIdx.x] = __hip_atomic_load(&st->pose_f.m[threadIdx.x], __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        else if (threadIdx.x == 16)
            s_done = __hip_atomic_load(&st->level_done[level_t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else if (threadIdx.x == 17)
            s_next_done = __hip_atomic_load(&st->level_done[level], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const bool tail_done = s_done != 0 || it >= n_iters;
        const bool next_done = s_next_done != 0;
        Mat4 pose;
#pragma unroll
        for (int i = 0; i < 16; ++i)
            pose.m[i] = s_pose[i];
        __syncthreads();     // read by every wave before the next iteration rewrites them
        if (tail_done) {
            if (next_done)
                return;
            CSum acc[NV];
            icp_accumulate<MODE, COMP>(lv, mp, mpose, mpts, mnrm, pose, gate, (int)gridDim.x, acc);
            icp_publish<MODE, COMP>(acc, lds, partials, st, level, sp, mpose);
            return;
        }
        ++gen;
        if ((int)blockIdx.x < ntb) {
            CSum acc[NV];
            icp_accumulate<MODE, COMP>(lt, mp, mpose, mpts, mnrm, pose, gate_t, ntb, acc);
            icp_publish<MODE, COMP, NV, true>(acc, lds, partials, st, level_t, sp_t, mpose, gen, ntb,
                                              &st->tail_gen[slot]);
        }
        if (!icp_wait_gen(st, gen, spin_limit, &st->tail_gen[slot]))
            return;
    }
}

// Frame end; with w2e set, also the binary32 world->eye of the tracked pose
// (as k_icp_w2e writes it).
__global__ void k_icp_frame_end(IcpState *__restrict__ st, Mat4 *__restrict__ w2e)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    icp_frame_end_dev(st, w2e);
}

// Sets the pose without tracking (first frame or a given pose).
__global__ void k_icp_set_pose(IcpState *__restrict__ st, Pose64 p)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    for (int i = 0; i < 16; ++i) {
        st->pose[i] = p.m[i];
        st->prev[i] = p.m[i];
        st->pose_f.m[i] = (float)p.m[i];
    }
}

// world->eye binary32 pose given by the caller -> FP64 eye->world state
__global__ void k_icp_set_pose_w2e(IcpState *__restrict__ st, Mat4 w2e)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    Mat4 e = w2e;
    invert_transform(e.m);
    for (int i = 0; i < 16; ++i) {
        st->pose[i] = e.m[i];
        st->prev[i] = e.m[i];
        st->pose_f.m[i] = e.m[i];
    }
}

// the previous pose only (the start of the next frame), from a binary32
// world->eye; the current pose keeps the tracked result
__global__ void k_icp_set_prev_w2e(IcpState *__restrict__ st, Mat4 w2e)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    Mat4 e = w2e;
    invert_transform(e.m);
    for (int i = 0; i < 16; ++i)
        st->prev[i] = e.m[i];
}

// world->eye (float) of the tracked pose, as Pipeline.cpp:41-42 derives it
__global__ void k_icp_w2e(const IcpState *__restrict__ st, Mat4 *__restrict__ w2e)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    Mat4 m = st->pose_f;
    invert_transform(m.m);
    *w2e = m;
}

// ====================================================================== host
static size_t al256(size_t x) { return (x + 255) & ~size_t(255); }

template <typename K, typename V>
static void launch_model(kifi_model *m, kifi_volume *vol, hipStream_t st, bool splat)
{
    const ModelParams mp = m->mp;
    const int P = mp.W * mp.H;
    const K *keys = (const K *)vol->keys[vol->cur];
    const V *vals = (const V *)vol->vals[vol->cur];
    if (splat) {
        const unsigned grid = (unsigned)std::max(1, vol->ctx->num_cus * 8);
        hipLaunchKernelGGL((k_model_splat<K, V>), dim3(grid), dim3(kIT), 0, st, keys, vals, vol->d_size, mp,
                           m->d_pose, m->d_zbuf);
    }
    const dim3 tiles((unsigned)((mp.W + kMTile - 1) / kMTile), (unsigned)((mp.H + kMTile - 1) / kMTile));
    hipLaunchKernelGGL((k_model_points<K, V>), tiles, dim3(kIT), 0, st, keys, vals, mp, m->d_pose, m->d_zbuf,
                       m->disc, m->d_pts, m->d_nrm);
}

// Model pose of this prediction and a cleared z-buffer.
void model_prepare(kifi_model *m, kifi_volume *vol)
{
    m->mp.vl = vol->consts.vl;
    m->mp.c0 = vol->consts.vlHalfPlusMin;
    m->mp.dmax = vol->consts.vl;                 // near-surface band: |D| < one voxel
}

void model_begin(kifi_model *m, kifi_volume *vol, const float *w2e_host, const Mat4 *w2e_dev)
{
    hipStream_t st = m->ctx->stream;
    Mat4 wv;
    for (int i = 0; i < 16; ++i)
        wv.m[i] = w2e_host ? w2e_host[i] : 0.0f;
    Mat4 clip;
    eye_to_clip_rh(m->cam.width, m->cam.height, m->cam.fx, m->cam.fy, m->cam.cx, m->cam.cy, m->cam.znear,
                   m->cam.zfar, clip.m);
    model_prepare(m, vol);
    const int P = m->mp.W * m->mp.H;
    hipLaunchKernelGGL(k_model_begin, dim3((P + 255) / 256), dim3(256), 0, st, wv, w2e_dev, clip, m->d_pose,
                       m->d_zbuf, P);
}

// Splat target for the fused voxel pass of an integration.
SplatParams model_splat_params(kifi_model *m, const Mat4 *w2e_dev)
{
    SplatParams sp;
    sp.w2e = w2e_dev;
    sp.zbuf = m->d_zbuf;
    sp.W = m->mp.W;
    sp.H = m->mp.H;
    sp.fx = m->mp.fx;
    sp.fy = m->mp.fy;
    sp.cx = m->mp.cx;
    sp.cy = m->mp.cy;
    sp.vl = m->mp.vl;
    sp.c0 = m->mp.c0;
    sp.dmax = m->mp.dmax;
    sp.znear = m->mp.znear;
    sp.pose_out = nullptr;
    sp.icp_end = nullptr;
    return sp;
}

// Resolve + normals (and the splat itself unless it was fused).
void model_end(kifi_model *m, kifi_volume *vol, bool splat)
{
    hipStream_t st = m->ctx->stream;
    if (vol->desc.key_bits == 32) {
        if (vol->desc.storage == KIFI_STORAGE_F32) launch_model<uint32_t, VoxF32i>(m, vol, st, splat);
        else launch_model<uint32_t, uint32_t>(m, vol, st, splat);
    } else {
        if (vol->desc.storage == KIFI_STORAGE_F32) launch_model<uint64_t, VoxF32i>(m, vol, st, splat);
        else launch_model<uint64_t, uint32_t>(m, vol, st, splat);
    }
}

kifi_status model_predict(kifi_model *m, kifi_volume *vol, const float *w2e_host, const Mat4 *w2e_dev)
{
    kifi_ctx *ctx = m->ctx;
    const int t_m = timer_begin(ctx, KIFI_STAGE_MODEL);
    model_begin(m, vol, w2e_host, w2e_dev);
    model_end(m, vol, true);
    timer_end(ctx, KIFI_STAGE_MODEL, t_m);
    KIFI_HIP(ctx, hipGetLastError());
    return KIFI_OK;
}

// Frame-to-frame model: a depth frame back-projected at its pose (binary32,
// p = e2w * ((u + 0.5 - cx)/fx d, (cy - v - 0.5)/fy d, -d, 1)), then normals.
__global__ __launch_bounds__(kIT) void k_model_backproject(const uint16_t *__restrict__ mm, ModelParams mp, Mat4 e2w,
                                                           float4 *__restrict__ pts)
{
    const int q = blockIdx.x * kIT + threadIdx.x;
    if (q >= mp.W * mp.H)
        return;
    const int v = q / mp.W, u = q - v * mp.W;
    const float d = (float)mm[q] * 0.001f;
    float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (d > 0.0f) {
        const float pe[4] = { ((float)u + 0.5f - mp.cx) / mp.fx * d, (mp.cy - (float)v - 0.5f) / mp.fy * d, -d, 1.0f };
        float p[4];
        mat_vec(e2w, pe, p);
        o = make_float4(p[0], p[1], p[2], 1.0f);
    }
    pts[q] = o;
}

kifi_status model_from_depth(kifi_model *m, const uint16_t *d_mm, const float e2w[16], const float w2e[16])
{
    kifi_ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    Mat4 wv, ev, clip;
    for (int i = 0; i < 16; ++i) {
        wv.m[i] = w2e[i];
        ev.m[i] = e2w[i];
    }
    eye_to_clip_rh(m->cam.width, m->cam.height, m->cam.fx, m->cam.fy, m->cam.cx, m->cam.cy, m->cam.znear,
                   m->cam.zfar, clip.m);
    const int P = m->mp.W * m->mp.H;
    hipLaunchKernelGGL(k_model_pose, dim3(1), dim3(64), 0, st, wv, (const Mat4 *)nullptr, clip, m->d_pose);
    hipLaunchKernelGGL(k_model_backproject, dim3((P + kIT - 1) / kIT), dim3(kIT), 0, st, d_mm, m->mp, ev, m->d_pts);
    hipLaunchKernelGGL(k_model_normals, dim3((P + kIT - 1) / kIT), dim3(kIT), 0, st, m->d_pts, m->mp.W, m->mp.H,
                       m->d_pose, m->disc, m->d_nrm);
    KIFI_HIP(ctx, hipGetLastError());
    return KIFI_OK;
}

// A caller's point buffer as the model (ICP::Align's synthDepthBuffer and
// synthEyeToWorld, ICP.cpp:38-48): W*H float3 world points, all-zero = empty
// (ICP.cpp:112-113).  The association projection is EyeToClipRH of the raw
// camera times the buffer's worldToEye (ICP.cpp:48, 81), as in the reference.
kifi_status model_set_points(kifi_model *m, const float *xyz, const float *w2e, const kifi_intrinsics *proj_cam)
{
    kifi_ctx *ctx = m->ctx;
    const size_t P = (size_t)m->mp.W * m->mp.H;
    std::vector<float4> tmp(P);
    for (size_t i = 0; i < P; ++i) {
        const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        tmp[i] = make_float4(x, y, z, (x == 0.0f && y == 0.0f && z == 0.0f) ? 0.0f : 1.0f);
    }
    KIFI_HIP(ctx, hipMemcpyAsync(m->d_pts, tmp.data(), P * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
    Mat4 wv, clip;
    for (int i = 0; i < 16; ++i)
        wv.m[i] = w2e[i];
    eye_to_clip_rh(proj_cam->width, proj_cam->height, proj_cam->fx, proj_cam->fy, proj_cam->cx, proj_cam->cy,
                   proj_cam->znear, proj_cam->zfar, clip.m);
    hipLaunchKernelGGL(k_model_pose, dim3(1), dim3(64), 0, ctx->stream, wv, (const Mat4 *)nullptr, clip, m->d_pose);
    KIFI_HIP(ctx, hipGetLastError());
    KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));     // tmp is released on return
    return KIFI_OK;
}

kifi_status model_alloc(kifi_model *m)
{
    kifi_ctx *ctx = m->ctx;
    const size_t P = (size_t)m->cam.width * m->cam.height;
    KIFI_HIP(ctx, hipMalloc(&m->d_pts, P * sizeof(float4)));
    KIFI_HIP(ctx, hipMalloc(&m->d_nrm, P * sizeof(float4)));
    KIFI_HIP(ctx, hipMalloc(&m->d_zbuf, P * 8));
    KIFI_HIP(ctx, hipMalloc(&m->d_pose, sizeof(ModelPose)));
    KIFI_HIP(ctx, hipMemset(m->d_pts, 0, P * sizeof(float4)));
    KIFI_HIP(ctx, hipMemset(m->d_nrm, 0, P * sizeof(float4)));
    return KIFI_OK;
}

void model_free(kifi_model *m)
{
    if (m->d_pts) (void)hipFree(m->d_pts);
    if (m->d_nrm) (void)hipFree(m->d_nrm);
    if (m->d_zbuf) (void)hipFree(m->d_zbuf);
    if (m->d_pose) (void)hipFree(m->d_pose);
}

// Work buffer of a W x H frame: the three pyramid levels, then the block partials.
size_t icp_work_bytes(int W, int H)
{
    const size_t P = (size_t)W * H;
    const size_t nb = (P + kIT - 1) / kIT;
    return al256(P * 4) + al256(P / 4 * 4 + 4) + al256(P / 16 * 4 + 4) +
           al256(nb * IcpNV<KIFI_ICP_POINT_TO_PLANE>::nv * sizeof(CSum));
}

void icp_work_carve(IcpWork *w, void *buf, int W, int H)
{
    const size_t P = (size_t)W * H;
    w->buf = buf;
    char *p = (char *)buf;
    w->lv[0] = (float *)p; p += al256(P * 4);
    w->lv[1] = (float *)p; p += al256(P / 4 * 4 + 4);
    w->lv[2] = (float *)p; p += al256(P / 16 * 4 + 4);
    w->partials = p;
    w->W = W;
    w->H = H;
}

kifi_status icp_alloc(IcpWork *w, kifi_ctx *ctx, int W, int H)
{
    void *buf = nullptr;
    KIFI_HIP(ctx, hipMalloc(&buf, icp_work_bytes(W, H)));
    icp_work_carve(w, buf, W, H);
    KIFI_HIP(ctx, hipMalloc(&w->d_state, sizeof(IcpState)));
    KIFI_HIP(ctx, hipMemset(w->d_state, 0, sizeof(IcpState)));
    return KIFI_OK;
}

void icp_free(IcpWork *w)
{
    if (w->buf) (void)hipFree(w->buf);
    if (w->d_state) (void)hipFree(w->d_state);
    w->buf = nullptr;
    w->d_state = nullptr;
}

static void icp_pass(kifi_model *m, IcpWork *w, const kifi_intrinsics *cam, const kifi_icp_params *ip, int levels,
                     int pass);
static bool icp_fallback_multi(kifi_model *m, IcpWork *w, const kifi_intrinsics *cam, const kifi_icp_params *ip,
                               int levels);

template <typename D>
static void launch_icp_begin(int levels, int P, hipStream_t st, IcpWork *w, const double *guess_dev, Pose64 gv,
                             const D *depth, int W, int H)
{
    const dim3 grid((P + 255) / 256), block(256);
    if (levels == 3)
        hipLaunchKernelGGL((k_icp_begin<D, 3>), grid, block, 0, st, w->d_state, guess_dev, gv, depth, W, H, w->lv[0],
                           w->lv[1], w->lv[2]);
    else if (levels == 2)
        hipLaunchKernelGGL((k_icp_begin<D, 2>), grid, block, 0, st, w->d_state, guess_dev, gv, depth, W, H, w->lv[0],
                           w->lv[1], w->lv[2]);
    else
        hipLaunchKernelGGL((k_icp_begin<D, 1>), grid, block, 0, st, w->d_state, guess_dev, gv, depth, W, H, w->lv[0],
                           w->lv[1], w->lv[2]);
}

// Runs the whole pyramid schedule asynchronously on the context's stream.
void icp_frame_end(IcpWork *w, Mat4 *w2e_out, hipStream_t st)
{
    hipLaunchKernelGGL(k_icp_frame_end, dim3(1), dim3(64), 0, st, w->d_state, w2e_out);
}

kifi_status icp_run(kifi_model *m, IcpWork *w, DepthIn depth, const kifi_intrinsics *cam,
                    const double *guess_host, const double *guess_dev, const kifi_icp_params *ip, Mat4 *w2e_out,
                    bool defer_end)
{
    kifi_ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    const int W = cam->width, H = cam->height;
    const int P = W * H;
    Pose64 gv;
    for (int i = 0; i < 16; ++i)
        gv.m[i] = guess_host ? guess_host[i] : 0.0;
    const int t_icp = timer_begin(ctx, KIFI_STAGE_ICP);
    const int levels = ip->levels < 1 ? 1 : (ip->levels > 3 ? 3 : ip->levels);
    if (depth.metres)
        launch_icp_begin(levels, P, st, w, guess_dev, gv, (const float *)depth.p, W, H);
    else
        launch_icp_begin(levels, P, st, w, guess_dev, gv, (const uint16_t *)depth.p, W, H);
    // the first pass starts from the predicted pose when a fallback is configured,
    // and may stop its levels earlier than the fallback from the previous pose
    const bool with_fallback = ip->fallback_rms > 0.0f || ip->fallback_iters > 0;
    kifi_icp_params first = *ip;
    for (int l = 0; l < 3; ++l)
        if (with_fallback && ip->first_iters[l] > 0)
            first.iters[l] = ip->first_iters[l];
    icp_pass(m, w, cam, &first, levels, 0);
    if (with_fallback && !icp_fallback_multi(m, w, cam, ip, levels)) {
        hipLaunchKernelGGL(k_icp_fallback_check, dim3(1), dim3(64), 0, st, w->d_state, ip->fallback_rms,
                           ip->fallback_iters);
        icp_pass(m, w, cam, ip, levels, 1);
        hipLaunchKernelGGL(k_icp_fallback_pick, dim3(1), dim3(64), 0, st, w->d_state, ip->fallback_rms);
    }
    if (!defer_end)
        icp_frame_end(w, w2e_out, st);
    timer_end(ctx, KIFI_STAGE_ICP, t_icp);
    KIFI_HIP(ctx, hipGetLastError());
    return KIFI_OK;
}

void icp_set_pose(IcpWork *w, const double *e2w, const float *w2e, hipStream_t st)
{
    if (e2w) {
        Pose64 p;
        for (int i = 0; i < 16; ++i)
            p.m[i] = e2w[i];
        hipLaunchKernelGGL(k_icp_set_pose, dim3(1), dim3(64), 0, st, w->d_state, p);
    } else {
        Mat4 m;
        for (int i = 0; i < 16; ++i)
            m.m[i] = w2e[i];
        hipLaunchKernelGGL(k_icp_set_pose_w2e, dim3(1), dim3(64), 0, st, w->d_state, m);
    }
}

// Pyramid level l of the live frame (the reference's camera scaled by 2^-l) and
// its association gate.
static IcpLevel icp_level(const IcpWork *w, const kifi_intrinsics *cam, const kifi_icp_params *ip, int l, float *gate)
{
    IcpLevel lv;
    const float sc = 1.0f / (float)(1 << l);
    lv.W = cam->width >> l;
    lv.H = cam->height >> l;
    lv.fx = cam->fx * sc;
    lv.fy = cam->fy * sc;
    lv.cx = cam->cx * sc;
    lv.cy = cam->cy * sc;
    lv.depth = w->lv[l];
    float g = ip->dist_thresh;
    for (int k = 0; k < l; ++k)
        g *= ip->level_gate_growth > 0.0f ? ip->level_gate_growth : 1.0f;
    *gate = g;
    return lv;
}

// Convergence test of level l: the increments below eps x eps_level_growth^l.
static IcpSolveParams solve_params(const kifi_icp_params *ip, int l)
{
    double g = 1.0;
    for (int k = 0; k < l; ++k)
        g *= ip->eps_level_growth > 0.0f ? ip->eps_level_growth : 1.0f;
    IcpSolveParams sp;
    sp.eps_rot = ip->eps_rot * g;
    sp.eps_trans = ip->eps_trans * g;
    sp.min_inliers = ip->min_inliers;
    sp.prior = ip->motion_prior > 0.0f ? (double)ip->motion_prior : 0.0;
    return sp;
}

// grid-stride: at most one workgroup per CU, so the last workgroup's reduction
// covers few partials
static int icp_blocks(const kifi_ctx *ctx, const IcpLevel &lv)
{
    return std::min((lv.W * lv.H + kIT - 1) / kIT, std::max(1, ctx->num_cus));
}

// Iterations of a level launched one by one before the rest of the level runs
// as one launch (k_icp_level_tail), per level 512x424, 256x212, 128x106: most
// frames converge within them (1.5, 1.3 and 2.5 solves on average on the room
// trajectory).
constexpr int kLaunchedIters[3] = { 2, 2, 3 };

static bool icp_tail_ok(kifi_ctx *ctx)
{
    if (ctx->tail_ok < 0) {
        int a = 0, b = 0;
        int c = 0, d = 0;
        const hipError_t e0 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_icp_level_tail<true>, kIT, 0);
        const hipError_t e1 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_icp_level_tail<false>, kIT, 0);
        const hipError_t e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_icp_tail_assoc<true>, kIT, 0);
        const hipError_t e3 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&d, k_icp_tail_assoc<false>, kIT, 0);
        (void)hipGetLastError();
        ctx->tail_ok = e0 == hipSuccess && e1 == hipSuccess && e2 == hipSuccess && e3 == hipSuccess && a >= 1 &&
                               b >= 1 && c >= 1 && d >= 1
                           ? 1
                           : 0;
    }
    return ctx->tail_ok != 0;
}

static void icp_pass(kifi_model *m, IcpWork *w, const kifi_intrinsics *cam, const kifi_icp_params *ip, int levels,
                     int pass)
{
    hipStream_t st = m->ctx->stream;
    const bool tails = ip->mode != KIFI_ICP_HORN && icp_tail_ok(m->ctx);
    const unsigned spin_limit = ip->spin_limit > 0 ? (unsigned)ip->spin_limit : 1u << 18;
    // the rest of a coarser level still to run: in the next level's first
    // launch (k_icp_tail_assoc), or on its own before a level with no launches
    struct Tail {
        bool pending = false;
        IcpLevel lv;
        float gate;
        IcpSolveParams sp;
        int level, nb, iters;
    } tail;
    auto launch_tail = [&]() {
        auto k = ip->naive_sums ? k_icp_level_tail<false> : k_icp_level_tail<true>;
        hipLaunchKernelGGL(k, dim3(tail.nb), dim3(kIT), 0, st, tail.lv, m->mp, m->d_pose, m->d_pts, m->d_nrm,
                           w->d_state, tail.gate, (CSum *)w->partials, tail.level, tail.sp, tail.iters, spin_limit,
                           pass * 3 + tail.level);
        tail.pending = false;
    };
    // Horn mode is the reference's ICP::Align: full resolution only
    for (int l = ip->mode == KIFI_ICP_HORN ? 0 : levels - 1; l >= 0; --l) {
        const IcpSolveParams sp = solve_params(ip, l);
        float gate;
        const IcpLevel lv = icp_level(w, cam, ip, l, &gate);
        const int nb = icp_blocks(m->ctx, lv);
        const int launched = tails ? std::min(ip->iters[l], kLaunchedIters[l]) : ip->iters[l];
        if (tail.pending && (launched == 0 || tail.nb > nb))
            launch_tail();
        for (int it = 0; it < launched; ++it) {
            if (ip->mode == KIFI_ICP_HORN) {
                const int nbw = std::min((((lv.W + 3) / 4) * lv.H + kIT - 1) / kIT, std::max(1, m->ctx->num_cus));
                hipLaunchKernelGGL(k_icp_assoc_window, dim3(nbw), dim3(kIT), 0, st, lv, m->mp, m->d_pose, m->d_pts,
                                   w->d_state, (CSum *)w->partials, l, sp);
            } else if (tail.pending) {
                auto k = ip->naive_sums ? k_icp_tail_assoc<false> : k_icp_tail_assoc<true>;
                hipLaunchKernelGGL(k, dim3(nb), dim3(kIT), 0, st, tail.lv, tail.gate, tail.sp, tail.level, tail.iters,
                                   tail.nb, pass * 3 + tail.level, lv, gate, sp, l, m->mp, m->d_pose, m->d_pts,
                                   m->d_nrm, w->d_state, (CSum *)w->partials, spin_limit);
                tail.pending = false;
            } else {
                if (ip->naive_sums)
                    hipLaunchKernelGGL((k_icp_assoc<KIFI_ICP_POINT_TO_PLANE, false>), dim3(nb), dim3(kIT), 0, st,
                                       lv, m->mp, m->d_pose, m->d_pts, m->d_nrm, w->d_state, gate,
                                       (CSum *)w->partials, l, sp);
                else
                    hipLaunchKernelGGL((k_icp_assoc<KIFI_ICP_POINT_TO_PLANE, true>), dim3(nb), dim3(kIT), 0, st,
                                       lv, m->mp, m->d_pose, m->d_pts, m->d_nrm, w->d_state, gate,
                                       (CSum *)w->partials, l, sp);
            }
        }
        if (launched < ip->iters[l]) {
            tail.pending = true;
            tail.lv = lv;
            tail.gate = gate;
            tail.sp = sp;
            tail.level = l;
            tail.nb = nb;
            tail.iters = ip->iters[l] - launched;
            if (l == 0)
                launch_tail();
        }
    }
    if (tail.pending)
        launch_tail();
}

// The point-to-plane fallback pass as one launch of k_icp_fallback_pass: a
// frame whose first pass is good pays one early-exit launch instead of one per
// iteration slot.  Its workgroups wait for each other, so all of them must be
// resident at once: the grid is at most one workgroup per CU, and the occupancy
// query (once per context) confirms that a CU holds one.  An ordinary launch
// then suffices; hipLaunchCooperativeKernel gives the same guarantee at the
// cost of a device-wide synchronisation per launch.  False without that
// confirmation (the caller then enqueues the pass launch by launch).
static bool icp_fallback_multi(kifi_model *m, IcpWork *w, const kifi_intrinsics *cam, const kifi_icp_params *ip,
                               int levels)
{
    kifi_ctx *ctx = m->ctx;
    if (ip->mode != KIFI_ICP_POINT_TO_PLANE)
        return false;
    if (ctx->multi_ok < 0) {
        int a = 0, b = 0;
        const hipError_t e0 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_icp_fallback_pass<true>, kIT, 0);
        const hipError_t e1 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_icp_fallback_pass<false>, kIT, 0);
        (void)hipGetLastError();
        ctx->multi_ok = e0 == hipSuccess && e1 == hipSuccess && a >= 1 && b >= 1 ? 1 : 0;
    }
    if (!ctx->multi_ok)
        return false;
    IcpFallback a{};
    for (int l = 0; l < 3; ++l) {
        a.lv[l] = icp_level(w, cam, ip, l, &a.gate[l]);
        a.sp[l] = solve_params(ip, l);
        a.nb[l] = icp_blocks(ctx, a.lv[l]);
        a.iters[l] = l < levels ? ip->iters[l] : 0;
    }
    a.levels = levels;
    a.max_rms = ip->fallback_rms;
    a.max_iters = ip->fallback_iters;
    // ~0.1 s of polling by default; an iteration takes tens of microseconds
    a.spin_limit = ip->spin_limit > 0 ? (unsigned)ip->spin_limit : 1u << 18;
    const dim3 grid(icp_blocks(ctx, a.lv[0]));   // <= num_cus
    if (ip->naive_sums)
        hipLaunchKernelGGL(k_icp_fallback_pass<false>, grid, dim3(kIT), 0, ctx->stream, a, m->mp, m->d_pose,
                           m->d_pts, m->d_nrm, w->d_state, (CSum *)w->partials);
    else
        hipLaunchKernelGGL(k_icp_fallback_pass<true>, grid, dim3(kIT), 0, ctx->stream, a, m->mp, m->d_pose,
                           m->d_pts, m->d_nrm, w->d_state, (CSum *)w->partials);
    return true;
}

__global__ void k_store_mat(Mat4 v, Mat4 *__restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *out = v;
}

void icp_set_prev(IcpWork *w, const float *w2e, hipStream_t st)
{
    Mat4 m;
    for (int i = 0; i < 16; ++i)
        m.m[i] = w2e[i];
    hipLaunchKernelGGL(k_icp_set_prev_w2e, dim3(1), dim3(64), 0, st, w->d_state, m);
}

void store_w2e(const float *w2e, Mat4 *d_out, hipStream_t st)
{
    Mat4 m;
    for (int i = 0; i < 16; ++i)
        m.m[i] = w2e[i];
    hipLaunchKernelGGL(k_store_mat, dim3(1), dim3(64), 0, st, m, d_out);
}

void icp_w2e(IcpWork *w, Mat4 *d_out, hipStream_t st)
{
    hipLaunchKernelGGL(k_icp_w2e, dim3(1), dim3(64), 0, st, w->d_state, d_out);
}

// ------------------------------------------------ Horn from correspondences
// GPU twin of ICP.cpp:166-264 given explicit correspondences: two compensated
// passes (centroids, then centred cross-covariance), FP64 Horn solve.
__global__ __launch_bounds__(kIT) void k_horn_sums(const float *__restrict__ src, const float *__restrict__ dst,
                                                   uint32_t n, const double *__restrict__ means, CSum *__restrict__ out)
{
    __shared__ CsumLds<9> lds;
    CSum acc[9];
#pragma unroll
    for (int j = 0; j < 9; ++j)
        acc[j] = CSum{ 0.0f, 0.0f };
    float ms[3] = { 0, 0, 0 }, md[3] = { 0, 0, 0 };
    if (means)
        for (int i = 0; i < 3; ++i) {
            ms[i] = (float)means[i];
            md[i] = (float)means[3 + i];
        }
    for (uint32_t i = blockIdx.x * kIT + threadIdx.x; i < n; i += gridDim.x * kIT) {
        float a[3], b[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = src[3 * (size_t)i + k] - ms[k];
            b[k] = dst[3 * (size_t)i + k] - md[k];
        }
        if (!means) {
            for (int k = 0; k < 3; ++k) {
                cadd(acc[k], a[k], 0.0f);
                cadd(acc[3 + k], b[k], 0.0f);
            }
        } else {
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    cadd(acc[3 * r + c], a[r] * b[c], 0.0f);
        }
    }
    block_csum<9>(acc, lds);
    if (threadIdx.x < 9)
        out[(size_t)blockIdx.x * 9 + threadIdx.x] = lds.result[threadIdx.x];
}

__global__ __launch_bounds__(kIT) void k_horn_finish(const CSum *__restrict__ part, int nb, uint32_t n, int phase,
                                                     double *__restrict__ means, float *__restrict__ out)
{
    __shared__ CsumLds<9> lds;
    CSum acc[9];
#pragma unroll
    for (int j = 0; j < 9; ++j)
        acc[j] = CSum{ 0.0f, 0.0f };
    for (int b = threadIdx.x; b < nb; b += kIT)
        for (int j = 0; j < 9; ++j)
            cadd(acc[j], part[(size_t)b * 9 + j].s, part[(size_t)b * 9 + j].c);
    block_csum<9>(acc, lds);
    if (threadIdx.x != 0)
        return;
    double v[9];
    for (int j = 0; j < 9; ++j)
        v[j] = (double)lds.result[j].s + (double)lds.result[j].c;
    if (phase == 0) {
        // centroids (ICP.cpp:166-167), rounded to binary32 as the reference keeps them
        for (int j = 0; j < 6; ++j)
            means[j] = (double)(float)(v[j] / (double)n);
        return;
    }
    const double ms[3] = { means[0], means[1], means[2] }, md[3] = { means[3], means[4], means[5] };
    double D[16];
    horn_increment(v, ms, md, D);
    for (int i = 0; i < 16; ++i)
        out[i] = (float)D[i];
}

kifi_status horn_align(kifi_ctx *ctx, const float *src, const float *dst, size_t n, float out[16])
{
    hipStream_t st = ctx->stream;
    const int nb = (int)std::min<size_t>((n + kIT - 1) / kIT, 1024);
    const size_t bytes = al256(n * 12) * 2 + al256((size_t)nb * 9 * sizeof(CSum)) + al256(64) + al256(64);
    kifi_status s = ensure_scratch(ctx, bytes);
    if (s != KIFI_OK)
        return s;
    char *p = (char *)ctx->scratch;
    float *ds = (float *)p; p += al256(n * 12);
    float *dd = (float *)p; p += al256(n * 12);
    CSum *part = (CSum *)p; p += al256((size_t)nb * 9 * sizeof(CSum));
    double *means = (double *)p; p += al256(64);
    float *dout = (float *)p;
    KIFI_HIP(ctx, hipMemcpyAsync(ds, src, n * 12, hipMemcpyHostToDevice, st));
    KIFI_HIP(ctx, hipMemcpyAsync(dd, dst, n * 12, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_horn_sums, dim3(nb), dim3(kIT), 0, st, ds, dd, (uint32_t)n, (const double *)nullptr, part);
    hipLaunchKernelGGL(k_horn_finish, dim3(1), dim3(kIT), 0, st, part, nb, (uint32_t)n, 0, means, dout);
    hipLaunchKernelGGL(k_horn_sums, dim3(nb), dim3(kIT), 0, st, ds, dd, (uint32_t)n, means, part);
    hipLaunchKernelGGL(k_horn_finish, dim3(1), dim3(kIT), 0, st, part, nb, (uint32_t)n, 1, means, dout);
    KIFI_HIP(ctx, hipGetLastError());
    KIFI_HIP(ctx, hipMemcpyAsync(out, dout, 64, hipMemcpyDeviceToHost, st));
    KIFI_HIP(ctx, hipStreamSynchronize(st));
    return KIFI_OK;
}

} // namespace kifi
