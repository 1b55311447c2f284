// kifi_integrate.hip -- per-frame fusion on the GPU:
//   keygen (+2x2x2 neighbourhood) -> LSD radix sort -> locate (each unique
//   candidate's rank in the volume; drop the maximum and the keys the volume
//   holds) -> compact -> partition into merged tiles -> union with the volume
//   fused with the weighted TSDF update.
//
// Reference path: Integrator::Integrate (KinectFission/kifi/Integrator.cpp:26-71).
// The reference sorts the K pixel keys, uniques them and dilates them with
// three passes (:186-250) whose x pass never emits the largest key's +1
// neighbour.  The dilated set equals the sorted unique set of all eight
// 2x2x2-neighbourhood keys of every pixel key minus its single maximum, so one
// sort of the 8K candidates replaces sort + unique + three unions.
#include <cfloat>

#include "kifi_icp.hpp"
#include "kifi_internal.hpp"
#include "kifi_prims.hpp"

namespace kifi {

constexpr int kThreads = 256;
constexpr int kRsItems = 8;
constexpr int kRsTile = kThreads * kRsItems;   // 2048 keys per radix tile
constexpr int kMgTile = 2048;                  // merged positions per merge tile

// ===================================================================== reset
// The update's cone pre-test (ClipPre).  Row r of clip space at the centre of
// voxel X is sum_j m_rj (X_j vl + c0) + m_r3.  The reference evaluates it in
// binary32 in source order; the pre-test as three fused multiply-adds of
// a_j = m_rj vl and a_3 = m_r3 + sum_j m_rj c0.  With u = 2^-24,
// S = sum_j |m_rj| C + |m_r3| and C = side + |c0| bounding every centre, each
// lies within about 8 u S of the exact value, so within 16 u S of the other;
// eps = 64 u S, maximised over the three rows, keeps a margin of 4.
__device__ void clip_pre(const Mat4 &w2c, const VolumeConsts &vc, ClipPre &p)
{
    const float C = (float)vc.res * vc.vl + fabsf(vc.vlHalfPlusMin);
    const int rows[3] = { 0, 1, 3 };
    float *dst[3] = { p.x, p.y, p.w };
    float smax = 0.0f;
    for (int k = 0; k < 3; ++k) {
        const int r = rows[k];
        const float m0 = w2c.m[r], m1 = w2c.m[r + 4], m2 = w2c.m[r + 8], m3 = w2c.m[r + 12];
        dst[k][0] = m0 * vc.vl;
        dst[k][1] = m1 * vc.vl;
        dst[k][2] = m2 * vc.vl;
        dst[k][3] = ((m0 * vc.vlHalfPlusMin + m1 * vc.vlHalfPlusMin) + m2 * vc.vlHalfPlusMin) + m3;
        const float S = (fabsf(m0) + fabsf(m1) + fabsf(m2)) * C + fabsf(m3);
        smax = fmaxf(smax, S);
    }
    p.eps = 64.0f * 0x1p-24f * smax;
}

// The frame's matrices (Integrator.cpp:41-42: worldToClip = EyeToClipRH() *
// worldToEye, eyeToWorld = invert_transform(worldToEye)), in binary32 on the
// device so that a device-resident pose (tracking) and a host pose give the
// same bits.  worldToEye is the tracker's pose inverted when the frame ends a
// tracking (icp_frame_end_dev writes the same matrix), else the device matrix,
// else the host one.
struct FrameStart {
    Mat4 clip, w2e_val;
    Mat4 *w2e_dev;           // device worldToEye (written here when icp_end is set)
    FrameMats *mats;         // out: this frame's matrices
    ModelPose *model_out;    // out when set: the model pose of the prediction this frame splats
    IcpState *icp_end;       // when set: the tracking this frame ends
    FrameCounters *cnt;
};

__device__ __forceinline__ Mat4 frame_w2e(const FrameStart &fs)
{
    if (fs.icp_end) {
        Mat4 m = fs.icp_end->pose_f;
        invert_transform(m.m);
        return m;
    }
    return fs.w2e_dev ? *fs.w2e_dev : fs.w2e_val;
}

// The frame's start, one thread: the end of the tracking (when icp_end), the
// per-frame counters, the matrices and the model pose.
__device__ void frame_start(const FrameStart &fs, const Mat4 &w2e, const VolumeConsts &vc)
{
    if (fs.icp_end)
        icp_frame_end_dev(fs.icp_end, fs.w2e_dev);
    FrameCounters *c = fs.cnt;
    c->n_cand = 0; c->n_keys = 0; c->n_unique = 0; c->n_new = 0;
    c->v_new = 0; c->n_tiles = 0; c->overflow = 0;
    FrameMats f;
    f.w2e = w2e;
    mat_mul(fs.clip.m, f.w2e.m, f.w2c.m);
    f.e2w = f.w2e;
    invert_transform(f.e2w.m);
    clip_pre(f.w2c, vc, f.pre);
    *fs.mats = f;
    if (fs.model_out) {
        ModelPose mp;
        mp.w2e = f.w2e;
        mp.proj = f.w2c;
        mp.cam[0] = f.e2w.m[12];
        mp.cam[1] = f.e2w.m[13];
        mp.cam[2] = f.e2w.m[14];
        *fs.model_out = mp;
    }
}

// ==================================================================== keygen
struct KeygenParams {
    int   W, H;
    float flInvX, flInvY, ppX, ppY, maxIndex;
    VolumeConsts vc;
    FrameStart fs;           // the frame's pose and its start (workgroup 0)
    unsigned long long *zbuf;   // when set: the model z-buffer of nz pixels, cleared to ~0 (a fill, never a memset)
    int   nz;
};



// Integrator::DepthMap2PointCloud (Integrator.cpp:76-182) per pixel, followed by
// the eight keys k + {0,1}x + {0,1}y + {0,1}z (integer adds, Integrator.cpp:192-194,
// 218, 242).  Output order is irrelevant: the sort that follows fixes it.
// pixel_key: the key of pixel pix at depth d metres (Integrator.cpp:109-178);
// false when it emits none.
template <typename K>
__device__ __forceinline__ bool pixel_key(float d, const KeygenParams &p, const Mat4 &e2w, int pix, K &key)
{
    const int v = pix / p.W, u = pix - v * p.W;
    if (!(d > 0.0f))
        return false;
    float q[4] = { (float)u + 0.0f, 0.0f + -(float)v, 0.0f + -1.0f, 0.0f + 0.0f };
    q[0] = q[0] * p.flInvX + p.ppX;                               // :126-129
    q[1] = q[1] * p.flInvY + p.ppY;
    q[2] = q[2] * 1.0f + 0.0f;
    q[3] = q[3] * 1.0f + 0.0f;
    q[0] = q[0] * d + 0.0f;                                       // :136-139
    q[1] = q[1] * d + 0.0f;
    q[2] = q[2] * d + 0.0f;
    q[3] = q[3] * d + 1.0f;
    mat_vec(e2w, q, q);                                           // :141-144
    const float ix = q[0] * p.vc.vlInv + p.vc.minOverNegVl - 0.5f;   // :146-149
    const float iy = q[1] * p.vc.vlInv + p.vc.minOverNegVl - 0.5f;
    const float iz = q[2] * p.vc.vlInv + p.vc.minOverNegVl - 0.5f;
    const float iw = q[3] * 1.0f + 0.0f - 0.5f;
    const bool valid = ix >= 0.0f && ix < p.maxIndex && iy >= 0.0f && iy < p.maxIndex && iz >= 0.0f &&
                       iz < p.maxIndex && iw >= 0.0f && iw < p.maxIndex;   // :151-154
    if (valid)
        key = pack_key<K>((K)(uint32_t)ix, (K)(uint32_t)iy, (K)(uint32_t)iz);
    return valid;
}

constexpr int kKgItems = 1;   // pixels per thread: a workgroup's 256 pixels give at most 2048 keys, one radix tile
static_assert(kKgItems * kThreads * 8 == kRsTile, "a keygen workgroup fills at most one radix tile");

// dm_out (uint16 frames): receives the frame in metres for the voxel update
// Workgroup t writes its keys to radix tile t (out + t kRsTile; its key count
// | valid pixels << 16 to tile_n[t]) and the tile's counts of the first radix
// digit to counts[d Tb + t], so the sort's first pass needs no upsweep.  No
// counter atomics: the first pass's scan sums tile_n into the frame's counts.
// Every workgroup derives the frame's eyeToWorld itself (frame_w2e, the same
// operations as frame_start), and workgroup 0 also starts the frame, so the
// frame needs no launch of its own before this one.  The four keys of a half-block
// differ only in y and z, above the first digit's bits, so each thread adds to
// at most two digit counters.
template <typename K, typename D, int DB>
__global__ __launch_bounds__(kThreads) void k_keygen(const D *__restrict__ depth,
                                                     KeygenParams p, K *__restrict__ out,
                                                     FrameCounters *__restrict__ cnt, float *__restrict__ dm_out,
                                                     uint32_t *__restrict__ counts, int Tb,
                                                     uint32_t *__restrict__ tile_n)
{
    constexpr int ND = 1 << DB;
    constexpr uint32_t MASK = ND - 1;
    static_assert(KeyTraits<K>::kb >= DB, "the y and z offsets leave the first digit unchanged");
    __shared__ uint32_t tmp[kThreads / kWave + 1];
    __shared__ uint32_t hist[ND];
    __shared__ Mat4 s_e2w;
    const int P = p.W * p.H;
    for (int q = threadIdx.x; q < ND; q += kThreads)
        hist[q] = 0;
    if (threadIdx.x == 0) {
        const Mat4 w2e = frame_w2e(p.fs);
        Mat4 e2w = w2e;
        invert_transform(e2w.m);
        s_e2w = e2w;
        if (blockIdx.x == 0)
            frame_start(p.fs, w2e, p.vc);
    }
    if (p.zbuf)
        for (int i = blockIdx.x * kThreads + threadIdx.x; i < p.nz; i += gridDim.x * kThreads)
            p.zbuf[i] = ~0ull;
    K keys[kKgItems];
    uint32_t flags = 0, c = 0;
    // every depth load in flight before the first use (a load under a branch
    // would be waited for on its own)
    float dm[kKgItems];
#pragma unroll
    for (int j = 0; j < kKgItems; ++j)
        dm[j] = depth_m(depth, (size_t)min((blockIdx.x * kKgItems + j) * kThreads + threadIdx.x, P - 1));
    __syncthreads();     // s_e2w
    const Mat4 e2w = s_e2w;
#pragma unroll
    for (int j = 0; j < kKgItems; ++j) {
        const int pix = (blockIdx.x * kKgItems + j) * kThreads + threadIdx.x;
        if (dm_out && pix < P)
            dm_out[pix] = dm[j];                                  // the metre frame of the update
        K k = 0;
        const bool ok = pix < P && pixel_key<K>(dm[j], p, e2w, pix, k);   // DepthStream.cpp:89-90
        keys[j] = k;
        flags |= (uint32_t)ok << j;
        c += ok;
    }
    // The 2x2x2 block of a key is its x+0 half (key + {0, dy, dz, dz+dy}) and its
    // x+1 half (the same + dx).  A half already in the left lane's block (the
    // previous pixel of the row, same key or one x step away) is not written:
    // the sort and locate that follow see the same key set (every dropped key
    // is written by some lane further left), from fewer candidates.
    constexpr K dx = 1, dy = K(1) << KeyTraits<K>::kb, dz = K(1) << (2 * KeyTraits<K>::kb);
    const int lane = __lane_id();
    uint32_t halves = 0, nw = 0;   // 2 bits per pixel: bit 0 = x+0 half, bit 1 = x+1 half
#pragma unroll
    for (int j = 0; j < kKgItems; ++j) {
        const bool ok = flags >> j & 1u;
        const K lk = __shfl_up(keys[j], 1, kWave);
        const int lok = __shfl_up((int)ok, 1, kWave);
        uint32_t h = ok ? 3u : 0u;
        if (ok && lane > 0 && lok) {
            if (lk == keys[j])
                h = 0u;
            else if (lk == keys[j] + dx)
                h = 1u;            // the left block's x+0 half is this block's x+1 half
            else if (lk + dx == keys[j])
                h = 2u;            // the left block's x+1 half is this block's x+0 half
        }
        halves |= h << (2 * j);
        nw += 4u * ((h & 1u) + (h >> 1));
    }
    // one scan for both counts: written candidates (<= 2048 per block) and valid
    // pixels; its barriers also order the histogram's clearing before the adds
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan<kThreads>(nw | (c << 16), tmp, tot);
#pragma unroll
    for (int j = 0; j < kKgItems; ++j) {
        const uint32_t h = halves >> (2 * j) & 3u;
        if (h & 1u)
            atomicAdd(&hist[(uint32_t)keys[j] & MASK], 4u);
        if (h & 2u)
            atomicAdd(&hist[(uint32_t)(keys[j] + dx) & MASK], 4u);
    }
    if (threadIdx.x == 0)
        tile_n[blockIdx.x] = tot;
    __syncthreads();
    for (int q = threadIdx.x; q < ND; q += kThreads)
        counts[(size_t)q * Tb + blockIdx.x] = hist[q];
    if (!nw)
        return;
    K *o = out + (size_t)blockIdx.x * kRsTile + (ex & 0xffffu);
#pragma unroll
    for (int j = 0; j < kKgItems; ++j) {
        const uint32_t h = halves >> (2 * j) & 3u;
        const K key = keys[j];
        if (h & 1u) {
            o[0] = key;      o[1] = key + dy;
            o[2] = key + dz; o[3] = key + dz + dy;
            o += 4;
        }
        if (h & 2u) {
            o[0] = key + dx;      o[1] = key + dy + dx;
            o[2] = key + dz + dx; o[3] = key + dz + dy + dx;
            o += 4;
        }
    }
}

// ================================================================ radix sort
// LSD radix sort with DB-bit digits (the reference's util::radix_sort,
// algorithm.h:63-91, sorts 8-bit digits with the same stability contract; any
// digit width gives the same order).  Only the bits a key can use are sorted:
// 2 kb + log2(res) (54 for 64-bit keys at res 4096: six 9-bit passes instead
// of eight 8-bit ones).  Per pass: upsweep (tile digit counts, wave-aggregated
// LDS histogram), scan (one workgroup per digit), downsweep (stable wave-match
// ranking, LDS staging, coalesced scatter).

template <typename K, int DB>
__global__ __launch_bounds__(kThreads) void k_rs_upsweep(const K *__restrict__ keys,
                                                         const uint32_t *__restrict__ d_n, int shift,
                                                         uint32_t *__restrict__ counts, int Tb)
{
    constexpr int ND = 1 << DB;
    constexpr uint32_t MASK = ND - 1;
    __shared__ uint32_t hist[ND];
    // the tile's keys are read before the count arrives (the candidate buffers
    // hold whole tiles), so the two reads overlap
    const size_t base = (size_t)blockIdx.x * kRsTile;
    K kk[kRsItems];
#pragma unroll
    for (int j = 0; j < kRsItems; ++j)
        kk[j] = keys[base + (size_t)j * kThreads + threadIdx.x];
    const uint32_t n = *d_n;
    for (int q = threadIdx.x; q < ND; q += kThreads)
        hist[q] = 0;
    __syncthreads();
    if (base < n) {
#pragma unroll
        for (int j = 0; j < kRsItems; ++j) {
            const size_t i = base + (size_t)j * kThreads + threadIdx.x;
            const bool ok = i < n;
            const uint32_t d = ok ? (uint32_t)(kk[j] >> shift) & MASK : 0u;
            const uint64_t act = __ballot(ok);
            const uint64_t m = match_digit(d, DB, act);
            if (ok && (m & lanemask_lt()) == 0)
                atomicAdd(&hist[d], popc64(m));
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < ND; q += kThreads)
        counts[(size_t)q * Tb + blockIdx.x] = hist[q];
}

// Per digit (one workgroup each): exclusive scan of the digit's tile counts,
// in place, and the digit's total (no atomics; the downsweep scans the totals).
// First pass: one more workgroup sums keygen's tile counts into the frame's
// candidate count and valid-pixel count.
__global__ __launch_bounds__(kThreads) void k_rs_scan(uint32_t *__restrict__ counts, int Tb,
                                                      uint32_t *__restrict__ totals,
                                                      const uint32_t *__restrict__ tile_n,
                                                      FrameCounters *__restrict__ cnt)
{
    __shared__ uint32_t tmp[kThreads / kWave + 1];
    const int d = blockIdx.x;
    if (tile_n && d == (int)gridDim.x - 1) {
        uint32_t keys = 0, px = 0;
        for (int t = threadIdx.x; t < Tb; t += kThreads) {
            const uint32_t v = tile_n[t];
            keys += v & 0xffffu;
            px += v >> 16;
        }
        keys = block_reduce_sum<kThreads>(keys, tmp);
        px = block_reduce_sum<kThreads>(px, tmp);
        if (threadIdx.x == 0) {
            cnt->n_cand = keys;
            cnt->n_keys = px;
        }
        return;
    }
    uint32_t carry = 0;
    uint32_t *row = counts + (size_t)d * Tb;
    constexpr int IT = 4;
    for (int start = 0; start < Tb; start += kThreads * IT) {
        uint32_t loc[IT], sum = 0;
        const int mb = start + threadIdx.x * IT;
#pragma unroll
        for (int q = 0; q < IT; ++q) {
            loc[q] = mb + q < Tb ? row[mb + q] : 0u;
            sum += loc[q];
        }
        uint32_t tot;
        uint32_t run = carry + block_exclusive_scan<kThreads>(sum, tmp, tot);
#pragma unroll
        for (int q = 0; q < IT; ++q) {
            if (mb + q < Tb)
                row[mb + q] = run;
            run += loc[q];
        }
        carry += tot;
    }
    if (threadIdx.x == 0)
        totals[d] = carry;
}

// tile_n (the first pass): tile t holds tile_n[t] & 0xffff keys at its start
// (keygen's tiles); otherwise the n keys are dense.
template <typename K, int DB>
__global__ __launch_bounds__(kThreads) void k_rs_downsweep(const K *__restrict__ in,
                                                           K *__restrict__ out,
                                                           const uint32_t *__restrict__ d_n, int shift,
                                                           const uint32_t *__restrict__ offsets, int Tb,
                                                           const uint32_t *__restrict__ totals,
                                                           const uint32_t *__restrict__ tile_n)
{
    constexpr int ND = 1 << DB;
    constexpr uint32_t MASK = ND - 1;
    constexpr int DPT = ND / kThreads;     // digits per thread in the digit scans
    static_assert(ND % kThreads == 0, "digit count a multiple of the workgroup");
    __shared__ uint32_t wcnt[kThreads / kWave][ND];
    __shared__ uint32_t dstart[ND];
    __shared__ uint32_t gbase[ND];
    __shared__ uint32_t tmp[kThreads / kWave + 1];
    __shared__ K stage[kRsTile];
    const size_t tbase = (size_t)blockIdx.x * kRsTile;
    const int w = threadIdx.x / kWave, lane = __lane_id();
    // Each wave ranks its own contiguous 64 x kRsItems-key sub-tile, item by item, so the
    // per-wave counters need no barrier (a wave's LDS operations are in order).  The keys
    // are read before the count arrives (the candidate buffers hold whole tiles).
    K keys[kRsItems];
    uint32_t rank[kRsItems];
    const size_t wbase = tbase + (size_t)w * (kWave * kRsItems);
#pragma unroll
    for (int j = 0; j < kRsItems; ++j)
        keys[j] = in[wbase + (size_t)j * kWave + lane];
    // this thread's digits' totals and tile offsets, read with the keys
    const int d0 = threadIdx.x * DPT;
    uint32_t gtot[DPT], goff[DPT];
#pragma unroll
    for (int e = 0; e < DPT; ++e) {
        gtot[e] = totals[d0 + e];
        goff[e] = offsets[(size_t)(d0 + e) * Tb + blockIdx.x];
    }
    uint32_t tn;          // keys of this tile, from its start
    if (tile_n) {
        tn = tile_n[blockIdx.x] & 0xffffu;
    } else {
        const uint32_t n = *d_n;
        tn = tbase < n ? (uint32_t)min((size_t)kRsTile, (size_t)n - tbase) : 0u;
    }
    if (tn == 0)
        return;
    const uint32_t wloc = (uint32_t)w * (kWave * kRsItems);   // the wave's first key in the tile
    for (int q = lane; q < ND; q += kWave)
        wcnt[w][q] = 0;
#pragma unroll
    for (int j = 0; j < kRsItems; ++j) {
        const bool ok = wloc + (uint32_t)(j * kWave + lane) < tn;
        const K k = keys[j];
        const uint32_t d = (uint32_t)(k >> shift) & MASK;
        const uint64_t act = __ballot(ok);
        const uint64_t m = match_digit(d, DB, act);
        const uint32_t before = popc64(m & lanemask_lt());
        uint32_t old = 0;
        if (ok)
            old = wcnt[w][d];
        rank[j] = old + before;
        if (ok && before == 0)
            wcnt[w][d] = old + popc64(m);
    }
    __syncthreads();
    {
        // thread t owns digits t*DPT .. t*DPT+DPT-1: per digit the waves'
        // exclusive offsets, the tile's digit starts and the global bases
        uint32_t cnt[DPT], tsum = 0, gsum = 0;
#pragma unroll
        for (int e = 0; e < DPT; ++e) {
            uint32_t acc = 0;
#pragma unroll
            for (int q = 0; q < kThreads / kWave; ++q) {
                const uint32_t c = wcnt[q][d0 + e];
                wcnt[q][d0 + e] = acc;
                acc += c;
            }
            cnt[e] = acc;
            tsum += acc;
            gsum += gtot[e];
        }
        uint32_t tot;
        uint32_t ex = block_exclusive_scan<kThreads>(tsum, tmp, tot);
        uint32_t dbase = block_exclusive_scan<kThreads>(gsum, tmp, tot);
#pragma unroll
        for (int e = 0; e < DPT; ++e) {
            const int d = d0 + e;
            dstart[d] = ex;
            gbase[d] = dbase + goff[e] - ex;
            ex += cnt[e];
            dbase += gtot[e];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRsItems; ++j) {
        if (wloc + (uint32_t)(j * kWave + lane) < tn) {
            const uint32_t d = (uint32_t)(keys[j] >> shift) & MASK;
            stage[dstart[d] + wcnt[w][d] + rank[j]] = keys[j];
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tn; i += kThreads) {
        const K k = stage[i];
        const uint32_t d = (uint32_t)(k >> shift) & MASK;
        out[gbase[d] + i] = k;
    }
}

// ======================================================== locate / compact
// The sorted candidates B (with repeats) against the volume A, two launches:
//   k_locate:  per unique candidate b, r = #A <= b (util::set_union puts A
//              first on ties, algorithm.h:138-198).  b is dropped when A holds
//              it already or when it is the largest candidate (the reference's
//              x pass never emits it, Integrator.cpp:231-245).
//   k_compact: the kept keys (the frame's new voxels) in key order with their
//              merged positions (the j-th kept key lands at r + j), and the
//              partition of the merged volume into tiles of kMgTile outputs.
// Every search runs on a sorted array W with two sampled levels beside it,
// S1[i] = W[i << SH1] and S2[i] = W[i << kS2Shift]: a search over S2 (a few KB,
// cache resident) and two short windows, instead of a binary search of the
// whole array whose every step waits for HBM.  The volume's keys keep every
// 16th (kVolS1) and every 4096th key.

constexpr int kVolS1 = 4;
constexpr int kKeyPad = 32;   // keys of padding past a volume key buffer's capacity
constexpr int kS2Shift = 12;

template <int SH1>
__host__ __device__ constexpr uint64_t s1_count(uint64_t n) { return (n + (1u << SH1) - 1) >> SH1; }
__host__ __device__ constexpr uint64_t s2_count(uint64_t n) { return (n + 4095) >> kS2Shift; }

// r[e] = #W[0, n) <= x[e] for n >= 1; NE independent searches interleaved
template <typename K, int NE, int SH1>
__device__ __forceinline__ void rank_le(const K *__restrict__ W, const K *__restrict__ S1, const K *__restrict__ S2,
                                        uint64_t n, const K (&x)[NE], uint64_t (&r)[NE])
{
    constexpr uint64_t G1 = 1u << (kS2Shift - SH1), G0 = 1u << SH1;   // S1 per S2 entry, W per S1 entry
    const uint64_t n2 = s2_count(n), n1 = s1_count<SH1>(n);
    uint64_t top = 1;
    while (top * 2 <= n2)
        top *= 2;
    uint64_t c[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e)
        c[e] = 0;
    // every probe is a load (clamped into the array), the step is a select:
    // the NE searches keep their loads in flight together
    for (uint64_t s = top; s; s >>= 1)
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const K w = S2[min(c[e] + s, n2) - 1];
            if (c[e] + s <= n2 && w <= x[e])
                c[e] += s;
        }
    // S1[G1 (c-1)] = S2[c-1] <= x < S2[c] = S1[G1 c]: the rest of that window
    uint64_t w0[NE];
    uint32_t len[NE], k[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const uint64_t cc = c[e] ? c[e] : 1;
        w0[e] = G1 * (cc - 1) + 1;
        len[e] = (uint32_t)(min(G1 * cc, n1) - w0[e]);
        k[e] = 0;
    }
#pragma unroll
    for (uint32_t s = (uint32_t)G1 / 2; s; s >>= 1)
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const K w = S1[w0[e] + min(k[e] + s, len[e]) - 1];   // S1[w0 - 1] when the window is empty
            if (k[e] + s <= len[e] && w <= x[e])
                k[e] += s;
        }
    // W[G0 (c1-1)] = S1[c1-1] <= x: the rest of that window of W
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const uint64_t c1 = w0[e] + k[e];
        w0[e] = G0 * (c1 - 1) + 1;
        len[e] = (uint32_t)(min(G0 * c1, n) - w0[e]);
        k[e] = 0;
    }
#pragma unroll
    for (uint32_t s = (uint32_t)G0 / 2; s; s >>= 1)
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const K w = W[w0[e] + min(k[e] + s, len[e]) - 1];
            if (k[e] + s <= len[e] && w <= x[e])
                k[e] += s;
        }
#pragma unroll
    for (int e = 0; e < NE; ++e)
        r[e] = c[e] ? w0[e] + k[e] : 0;
}

// lo + #W[lo, hi) <= x, found by the 64 lanes of one wave together (x is
// wave-uniform): each round probes 64 evenly spaced entries at once and keeps
// the gap between the last probe <= x and the first one above it.
template <typename K>
__device__ __forceinline__ uint64_t wave_count_le(const K *__restrict__ W, uint64_t lo, uint64_t hi, K x)
{
    const uint64_t lane = (uint64_t)__lane_id();
    while (hi > lo) {
        const uint64_t step = (hi - lo + 63) >> 6;
        const uint64_t p = lo + (lane + 1) * step - 1;
        const bool le = p < hi && W[p] <= x;
        const uint64_t c = (uint64_t)popc64(__ballot(le));
        if (step == 1)
            return lo + c;
        const uint64_t nlo = lo + c * step;
        hi = min(hi, lo + (c + 1) * step - 1);
        lo = nlo;
    }
    return lo;
}

// #W[0, n) <= x for n >= 1 through the two sampled levels, one wave
template <typename K, int SH1>
__device__ __forceinline__ uint64_t wave_rank_le(const K *__restrict__ W, const K *__restrict__ S1,
                                                 const K *__restrict__ S2, uint64_t n, K x)
{
    constexpr uint64_t G1 = 1u << (kS2Shift - SH1), G0 = 1u << SH1;
    const uint64_t c2 = wave_count_le(S2, 0, s2_count(n), x);
    if (c2 == 0)
        return 0;
    const uint64_t c1 = wave_count_le(S1, G1 * (c2 - 1) + 1, min(G1 * c2, s1_count<SH1>(n)), x);
    return wave_count_le(W, G0 * (c1 - 1) + 1, min(G0 * c1, n), x);
}

// dst[0, len) = src[0, len) for len <= kThreads * R through LDS: every load in
// flight at once (one HBM round trip; a load-store loop would wait for each)
template <typename K, int R>
__device__ __forceinline__ void lds_fill(K *dst, const K *__restrict__ src, uint32_t len)
{
    if (len == 0)
        return;
    K t[R];
#pragma unroll
    for (int q = 0; q < R; ++q)
        if ((uint32_t)q * kThreads < len)   // workgroup-uniform: only the rounds len needs
            t[q] = src[min(q * kThreads + threadIdx.x, len - 1)];
#pragma unroll
    for (int q = 0; q < R; ++q)
        if (q * kThreads + threadIdx.x < len)
            dst[q * kThreads + threadIdx.x] = t[q];
}

// c[e] = #win[0, len) <= x[e] for every e in `mask` (win sorted in LDS;
// power-of-two steps from top, the largest power of two <= len); the NE
// searches interleave
template <typename K, int NE>
__device__ __forceinline__ void lds_count_le(const K *win, uint32_t len, uint32_t top, const K (&x)[NE],
                                             uint32_t mask, uint32_t (&c)[NE])
{
#pragma unroll
    for (int e = 0; e < NE; ++e)
        c[e] = 0;
    if (len == 0)
        return;
    // every probe is a read (clamped into the window) and the step a select:
    // the NE reads of a step are in flight together, where a read under a
    // branch would be waited for on its own
    for (uint32_t st = top; st; st >>= 1) {
        K w[NE];
#pragma unroll
        for (int e = 0; e < NE; ++e)
            w[e] = win[min(c[e] + st, len) - 1];
#pragma unroll
        for (int e = 0; e < NE; ++e)
            c[e] += ((mask >> e & 1u) && c[e] + st <= len && w[e] <= x[e]) ? st : 0u;
    }
}

// (number of keys <= x in the window W[w0, w0 + 16 NW), whether x is one of
// them); keys at nA or beyond do not count.  The window is read as aligned
// 16-byte vectors: w0 is a multiple of 16 keys and the key buffers hold
// kKeyPad keys of padding past their capacity.
template <typename K, int NW, int NE>
__device__ __forceinline__ void window_count(const K *__restrict__ A, uint64_t w0, uint64_t nA, const K (&x)[NE],
                                             uint32_t mask, uint32_t (&cnt)[NE], uint32_t &eqm)
{
    constexpr int NV = NW * 16 * (int)sizeof(K) / 16;
    uint4 v[NV];
    const uint4 *vp = (const uint4 *)(A + w0);
#pragma unroll
    for (int q = 0; q < NV; ++q)
        v[q] = vp[q];
    K wk[NW * 16];
    __builtin_memcpy(wk, v, sizeof(v));
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        uint32_t c = 0;
        bool eq = false;
#pragma unroll
        for (int q = 0; q < NW * 16; ++q) {
            const bool in = w0 + q < nA;
            c += in && wk[q] <= x[e];
            eq |= in && wk[q] == x[e];
        }
        cnt[e] = c;
        if ((mask >> e & 1u) && eq)
            eqm |= 1u << e;
    }
}

// Chunks of kRsTile sorted candidates; candidate e * kThreads + t of a chunk
// belongs to thread t (so at every step of the lookups below the 64 lanes of a
// wave hold neighbouring candidates, whose LDS probes fall on neighbouring
// words).  Two waves bound the volume ranks of the chunk's first and last
// candidate through the sampled levels (c1 = #S1 <= x); every rank lies in
// A[16 (c1_first - 1), 16 c1_last).  When that stretch fits the LDS window it
// is loaded there in one round and each candidate is looked up in LDS.
// Otherwise the stretch's every-16th keys go through LDS (in pages) and each
// candidate reads the 16-key volume window its rank lies in.  The kept keys
// and their ranks are compacted, in key order, within the chunk's own region
// of Bc / Rc.
template <typename K>
__global__ __launch_bounds__(kThreads) void k_locate(const K *__restrict__ B, const uint32_t *__restrict__ d_n,
                                                     const K *__restrict__ A, const K *__restrict__ AS1,
                                                     const K *__restrict__ AS2, const uint64_t *__restrict__ d_size,
                                                     FrameCounters *__restrict__ cnt, K *__restrict__ Bc,
                                                     uint64_t *__restrict__ Rc, uint32_t *__restrict__ chunk_kept,
                                                     uint32_t *__restrict__ chunk_heads)
{
    constexpr uint32_t WIN = 32768 / sizeof(K);
    constexpr int RF = WIN / kThreads;
    constexpr int NWV = kThreads / kWave;
    constexpr uint64_t G1 = 1u << (kS2Shift - kVolS1), G0 = 1u << kVolS1;
    __shared__ K win[WIN];
    __shared__ uint64_t sh_c1[2];
    __shared__ uint32_t sh_off[kRsItems * NWV + 1];
    __shared__ uint32_t tmp[kThreads / kWave + 1];
    const uint32_t n = *d_n;
    const uint64_t nA = *d_size;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        cnt->v_old = nA;
    const size_t base = (size_t)blockIdx.x * kRsTile;
    if (base >= n)
        return;
    const size_t last = (size_t)n - 1;
    const size_t tlast = min(base + kRsTile, (size_t)n) - 1;
    const int w = threadIdx.x / kWave, lane = __lane_id();
    // every load in flight together (clamped, not conditional); the two bound
    // searches start at once, their first loads behind these
    const K bmax = B[last];
    K b[kRsItems], pv[kRsItems];
#pragma unroll
    for (int e = 0; e < kRsItems; ++e) {
        const size_t i = base + (size_t)e * kThreads + threadIdx.x;
        b[e] = B[min(i, last)];
        pv[e] = B[i > 0 ? min(i - 1, last) : 0];
    }
    if (nA > 0 && w < 2) {
        const K x = B[w == 0 ? base : tlast];
        const uint64_t c2 = wave_count_le(AS2, 0, s2_count(nA), x);
        const uint64_t c1 = c2 ? wave_count_le(AS1, G1 * (c2 - 1) + 1, min(G1 * c2, s1_count<kVolS1>(nA)), x) : 0;
        if (lane == 0)
            sh_c1[w] = c1;
    }
    uint32_t look = 0, heads = 0;   // look bit e: a unique key other than the maximum
#pragma unroll
    for (int e = 0; e < kRsItems; ++e) {
        const size_t i = base + (size_t)e * kThreads + threadIdx.x;
        const bool head = i < n && (i == 0 || b[e] != pv[e]);
        heads += head;
        look |= (uint32_t)(head && b[e] != bmax) << e;
    }
    uint64_t r[kRsItems];
#pragma unroll
    for (int e = 0; e < kRsItems; ++e)
        r[e] = 0;
    uint32_t dupm = 0;   // candidates found in the volume (A wins)
    if (nA > 0) {   // block-uniform
        __syncthreads();
        const uint64_t c1f = sh_c1[0], c1l = sh_c1[1];
        const uint64_t ws = c1f ? G0 * (c1f - 1) : 0, we = min(G0 * c1l, nA);
        const uint64_t len = we > ws ? we - ws : 0;   // A[ws] <= every candidate (c1f > 0), A[we] above them
        if (len <= WIN) {   // block-uniform
            lds_fill<K, RF>(win, A + ws, (uint32_t)len);
            __syncthreads();
            uint32_t top = 1;
            while (2 * top <= (uint32_t)len)
                top *= 2;
            uint32_t c[kRsItems];
            lds_count_le<K, kRsItems>(win, (uint32_t)len, top, b, look, c);
            K at[kRsItems];   // the window key just below each rank (win[0] stands in when there is none)
#pragma unroll
            for (int e = 0; e < kRsItems; ++e)
                at[e] = win[c[e] ? c[e] - 1 : 0];
#pragma unroll
            for (int e = 0; e < kRsItems; ++e) {
                r[e] = ws + c[e];
                if ((look >> e & 1u) && c[e] && at[e] == b[e])
                    dupm |= 1u << e;
            }
        } else {
            // S1[i_lo, c1l) through LDS, in pages of WIN: S1[i_lo] <= every
            // candidate (c1f > 0) and S1[c1l] above them; c1 = #S1 <= b sums the
            // pages' counts
            const uint64_t i_lo = c1f ? c1f - 1 : 0, ns = c1l - i_lo;
            if (ns <= 8 * (uint64_t)WIN) {   // block-uniform
                uint32_t c[kRsItems];
#pragma unroll
                for (int e = 0; e < kRsItems; ++e)
                    c[e] = 0;
#pragma unroll 1
                for (uint64_t p0 = 0; p0 < ns; p0 += WIN) {
                    const uint32_t pl = (uint32_t)min(ns - p0, (uint64_t)WIN);
                    if (p0)
                        __syncthreads();   // the previous page's lookups are done
                    // two rounds of half a window: a whole-window round here
                    // would hold the kernel at half the occupancy
                    for (uint32_t k0 = 0; k0 < pl; k0 += RF / 2 * kThreads)
                        lds_fill<K, RF / 2>(win + k0, AS1 + i_lo + p0 + k0, min(pl - k0, (uint32_t)(RF / 2 * kThreads)));
                    __syncthreads();
                    uint32_t top = 1;
                    while (2 * top <= pl)
                        top *= 2;
                    uint32_t cp[kRsItems];
                    lds_count_le<K, kRsItems>(win, pl, top, b, look, cp);
#pragma unroll
                    for (int e = 0; e < kRsItems; ++e)
                        c[e] += cp[e];
                }
                // c1 = #S1 <= b: the rank lies in A[16 (c1-1), 16 c1] (0 below A[0]);
                // one 16-key window per candidate, GE candidates' reads at once
                constexpr int GE = sizeof(K) == 4 ? 2 : 1;
#pragma unroll
                for (int e0 = 0; e0 < kRsItems; e0 += GE) {
                    uint64_t w0[GE];
                    uint4 v[GE][sizeof(K)];
#pragma unroll
                    for (int g = 0; g < GE; ++g) {
                        const uint64_t c1 = i_lo + c[e0 + g];
                        w0[g] = (look >> (e0 + g) & 1u) && c1 ? G0 * (c1 - 1) : 0;
                        const uint4 *vp = (const uint4 *)(A + w0[g]);
#pragma unroll
                        for (int q = 0; q < (int)sizeof(K); ++q)
                            v[g][q] = vp[q];
                    }
#pragma unroll
                    for (int g = 0; g < GE; ++g) {
                        K wk[16];
                        __builtin_memcpy(wk, v[g], sizeof(wk));
                        uint32_t cn = 0;
                        bool eq = false;
#pragma unroll
                        for (int q = 0; q < 16; ++q) {
                            const bool in = w0[g] + q < nA;
                            cn += in && wk[q] <= b[e0 + g];
                            eq |= in && wk[q] == b[e0 + g];
                        }
                        r[e0 + g] = w0[g] + cn;
                        if ((look >> (e0 + g) & 1u) && eq)
                            dupm |= 1u << (e0 + g);
                    }
                    asm volatile("" ::: "memory");   // one group's windows in registers at a time
                }
            } else {
                rank_le<K, kRsItems, kVolS1>(A, AS1, AS2, nA, b, r);
                K at[kRsItems];
#pragma unroll
                for (int e = 0; e < kRsItems; ++e)
                    at[e] = A[r[e] ? r[e] - 1 : 0];
#pragma unroll
                for (int e = 0; e < kRsItems; ++e)
                    if (r[e] && at[e] == b[e])
                        dupm |= 1u << e;
            }
        }
    }
    // compaction in candidate order e * kThreads + t: per (e, wave) counts by
    // ballot, their exclusive scan, then each kept candidate's offset
    const uint32_t keep = look & ~dupm;
    uint64_t bal[kRsItems];
#pragma unroll
    for (int e = 0; e < kRsItems; ++e) {
        bal[e] = __ballot(keep >> e & 1u);
        if (lane == 0)
            sh_off[e * NWV + w] = (uint32_t)popc64(bal[e]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int q = 0; q < kRsItems * NWV; ++q) {
            const uint32_t t = sh_off[q];
            sh_off[q] = acc;
            acc += t;
        }
        sh_off[kRsItems * NWV] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < kRsItems; ++e)
        if (keep >> e & 1u) {
            const uint32_t off = sh_off[e * NWV + w] + (uint32_t)popc64(bal[e] & lanemask_lt());
            Bc[base + off] = b[e];
            Rc[base + off] = r[e];
        }
    const uint32_t hsum = block_reduce_sum<kThreads>(heads, tmp);
    if (threadIdx.x == 0) {
        chunk_kept[blockIdx.x] = sh_off[kRsItems * NWV];
        chunk_heads[blockIdx.x] = hsum;
    }
}

// The kept keys of every chunk, in key order, with their merged positions (the
// j-th kept key lands at its rank + j), and the merge's tile partition: tile t
// starts at old voxel kMgTile t - m and new voxel m, m = #(positions <
// kMgTile t).  New voxel j owns the tile boundaries in (position of j - 1,
// position of j] and writes m = j there; the last non-empty chunk (chunk 0
// when nothing is kept) writes the boundaries after the last new voxel and
// closes the frame's counts: the new size, the capacity checks, the running
// totals and the host ring slot.  A gap of more than kGapDirect boundaries is
// written by the whole workgroup.
constexpr int kGapDirect = 4;
constexpr int kGapList = 256;

template <typename K>
__global__ __launch_bounds__(kThreads) void k_compact(const K *__restrict__ Bc, const uint64_t *__restrict__ Rc,
                                                      const uint32_t *__restrict__ d_n,
                                                      const uint32_t *__restrict__ chunk_kept,
                                                      const uint32_t *__restrict__ chunk_heads,
                                                      FrameCounters *__restrict__ cnt, K *__restrict__ Bn,
                                                      uint64_t *__restrict__ on, uint64_t *__restrict__ part_a,
                                                      uint64_t *__restrict__ part_b, int Tbound,
                                                      uint64_t *__restrict__ d_size, uint64_t cap,
                                                      FrameCounters *__restrict__ host_cnt)
{
    __shared__ uint32_t tmp[kThreads / kWave + 1];
    __shared__ uint64_t gap_lo[kGapList], gap_hi[kGapList];
    __shared__ uint32_t gap_j[kGapList];
    __shared__ uint32_t n_gap;
    const uint32_t n = *d_n;
    const uint32_t nchunks = (n + kRsTile - 1) / kRsTile;
    const uint32_t c = blockIdx.x;
    const uint32_t kc = c < nchunks ? chunk_kept[c] : 0u;
    if (kc == 0 && c != 0)
        return;
    const size_t base = (size_t)c * kRsTile;
    // the chunk's first kept entries are read before the prefix is known
    K kb = K(0);
    uint64_t kr = 0, kp = 0;
    if (threadIdx.x < kc) {
        kb = Bc[base + threadIdx.x];
        kr = Rc[base + threadIdx.x];
        if (threadIdx.x > 0)
            kp = Rc[base + threadIdx.x - 1];
    }
    // kept keys before and after this chunk, and the nearest earlier chunk that kept any
    uint32_t s = 0, sa = 0, lastq = 0;
    for (uint32_t q = threadIdx.x; q < nchunks; q += kThreads) {
        const uint32_t v = chunk_kept[q];
        if (q < c) {
            s += v;
            if (v)
                lastq = max(lastq, q + 1);
        } else if (q > c) {
            sa += v;
        }
    }
    if (threadIdx.x == 0)
        n_gap = 0;
    const uint32_t pre = block_reduce_sum<kThreads>(s, tmp);
    const uint32_t post = block_reduce_sum<kThreads>(sa, tmp);
    lastq = block_reduce_max<kThreads>(lastq, tmp);     // 1 + that chunk, 0 when none
    const uint32_t nk = pre + kc + post;
    // position of the kept key just before this chunk's first (none: ~0)
    uint64_t o_before = ~0ull;
    if (lastq) {
        const uint32_t q = lastq - 1;
        o_before = Rc[(size_t)q * kRsTile + chunk_kept[q] - 1] + (pre - 1);
    }
    for (uint32_t i = threadIdx.x; i < kc; i += kThreads) {
        if (i >= kThreads) {
            kb = Bc[base + i];
            kr = Rc[base + i];
            kp = Rc[base + i - 1];
        }
        const uint32_t j = pre + i;
        const uint64_t o = kr + j;
        Bn[j] = kb;
        on[j] = o;
        const uint64_t o_prev = i ? kp + (j - 1) : o_before;
        const uint64_t t_lo = o_prev == ~0ull ? 0 : o_prev / kMgTile + 1, t_hi = o / kMgTile;
        if (t_lo > t_hi)
            continue;
        if (t_hi - t_lo < kGapDirect) {
            for (uint64_t t = t_lo; t <= t_hi; ++t)
                if (t <= (uint64_t)Tbound) {
                    part_b[t] = j;
                    part_a[t] = t * kMgTile - j;
                }
        } else {
            const uint32_t g = atomicAdd(&n_gap, 1u);
            if (g < kGapList) {
                gap_lo[g] = t_lo;
                gap_hi[g] = t_hi;
                gap_j[g] = j;
            } else {
                for (uint64_t t = t_lo; t <= t_hi && t <= (uint64_t)Tbound; ++t) {
                    part_b[t] = j;
                    part_a[t] = t * kMgTile - j;
                }
            }
        }
    }
    __syncthreads();
    const uint32_t ng = min(n_gap, (uint32_t)kGapList);
    for (uint32_t g = 0; g < ng; ++g) {
        const uint64_t hi = min(gap_hi[g], (uint64_t)Tbound);
        const uint32_t j = gap_j[g];
        for (uint64_t t = gap_lo[g] + threadIdx.x; t <= hi; t += kThreads) {
            part_b[t] = j;
            part_a[t] = t * kMgTile - j;
        }
    }
    // the last new voxel's chunk (chunk 0 when there is none)
    if (!(post == 0 && (kc > 0 || nk == 0)))
        return;
    const uint64_t v0 = cnt->v_old, v1 = v0 + nk;
    const uint64_t T = (v1 + kMgTile - 1) / kMgTile;
    const uint64_t o_last = nk ? Rc[base + kc - 1] + (nk - 1) : 0;
    const uint64_t t0 = nk ? o_last / kMgTile + 1 : 0;
    for (uint64_t t = t0 + threadIdx.x; t <= T && t <= (uint64_t)Tbound; t += kThreads) {
        part_b[t] = nk;
        part_a[t] = min(t * kMgTile, v1) - nk;
    }
    uint32_t hs = 0;
    for (uint32_t q = threadIdx.x; q < nchunks; q += kThreads)
        hs += chunk_heads[q];
    hs = block_reduce_sum<kThreads>(hs, tmp);
    if (threadIdx.x == 0) {
        cnt->n_kept = nk;
        cnt->n_unique = hs;
        cnt->n_new = hs ? hs - 1 : 0;
        cnt->v_new = v1;
        cnt->n_tiles = (uint32_t)min(T, (uint64_t)0xffffffffu);
        if (v1 > cap || T > (uint64_t)Tbound) {
            cnt->overflow = 1;
        } else {
            *d_size = v1;
            cnt->tot_frames += 1;
            cnt->tot_v_old += v0;
            cnt->tot_new += cnt->n_new;
            cnt->tot_v_new += v1;
            cnt->tot_keys += cnt->n_keys;
        }
        *host_cnt = *cnt;   // host-mapped: read after the frame's event
    }
}

// The two sampled levels of the volume's keys, rebuilt after the host wrote
// the volume (upload, load, growth); an integration keeps them current.
template <typename K>
__global__ __launch_bounds__(kThreads) void k_build_samples(const K *__restrict__ A,
                                                            const uint64_t *__restrict__ d_size,
                                                            K *__restrict__ S1, K *__restrict__ S2)
{
    const uint64_t n = *d_size, n1 = s1_count<kVolS1>(n);
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n1; i += (uint64_t)gridDim.x * kThreads) {
        const K k = A[i << kVolS1];
        S1[i] = k;
        if ((i & ((1u << (kS2Shift - kVolS1)) - 1)) == 0)
            S2[i >> (kS2Shift - kVolS1)] = k;
    }
}

struct UpdateParams {
    const FrameMats *mats;   // device: w2c of this frame
    float vl, c0, trunc, ndcU, ndcV, maxU, maxV;
    int   W, H;
    const float *depth;      // the frame in metres
    unsigned long long *upd; // A's running counters (kUpdSlots x kUpdStride)
};

// Integrator::UpdateVoxels per voxel (Integrator.cpp:285-356), in two halves so
// that a lane can issue the depth gathers of all its voxels before it uses one.
// voxel_pixel: the depth pixel the voxel reads (:285-319), or -1 when it is
// outside the image.
template <typename K>
__device__ __forceinline__ float voxel_centre(uint32_t c, const UpdateParams &p)
{
    return (float)c * p.vl + p.c0;                                 // Volume.cpp:75-78
}

// Outside the update's cone for sure: the pre-test's |x| or |y| exceeds
// 1.02 |w| + eps, so the reference-order binary32 values have |x/w| or |y/w|
// above 1.01 (their distance from the pre-test is at most eps / 2), which lands
// outside [0,W) x [0,H) whatever the rounding.  NaN is never outside.
template <typename K>
__device__ __forceinline__ bool voxel_outside(K key, const ClipPre &c)
{
    uint32_t x, y, z;
    unpack_key<K>(key, x, y, z);
    const float fx = (float)x, fy = (float)y, fz = (float)z;
    const float px = fmaf(c.x[0], fx, fmaf(c.x[1], fy, fmaf(c.x[2], fz, c.x[3])));
    const float py = fmaf(c.y[0], fx, fmaf(c.y[1], fy, fmaf(c.y[2], fz, c.y[3])));
    const float pw = fmaf(c.w[0], fx, fmaf(c.w[1], fy, fmaf(c.w[2], fz, c.w[3])));
    const float lim = fmaf(1.02f, fabsf(pw), c.eps);
    return fabsf(px) > lim || fabsf(py) > lim;
}

template <typename K>
__device__ __forceinline__ int voxel_pixel(K key, const UpdateParams &p, const Mat4 &w2c, const ClipPre &pre)
{
    if (voxel_outside<K>(key, pre))   // most of the volume: no projection in the reference's order
        return -1;
    uint32_t x, y, z;
    unpack_key<K>(key, x, y, z);
    float k[4] = { voxel_centre<K>(x, p), voxel_centre<K>(y, p), voxel_centre<K>(z, p), 1.0f };
    mat_vec(w2c, k, k);                                            // :290-293
    const float kw = k[3];
    // :305-315 divides all four lanes by w and keeps the voxel only if
    // (z/w)*0+0 and (w/w)*0+0 are in range, i.e. z/w is finite and w is finite
    // and nonzero.  With w finite and nonzero, RN(z/w) is finite exactly when
    // |z| < |w| (2^128 - 2^103) (the rounding boundary above FLT_MAX; a tie
    // rounds to the even infinity).  Binary32, no division in the common case:
    // |z| < |w| 2^127 is finite (the scaling by a power of two is exact, or
    // overflows to infinity, which only accepts what is finite anyway);
    // |z| >= |w| 2^128 is not; between the two the division decides.  NaN z
    // fails every comparison and the division's finiteness.
    const bool wok = kw != 0.0f && fabsf(kw) <= FLT_MAX;
    const float za = fabsf(k[2]), wa = fabsf(kw);
    bool zok = za < wa * 0x1p127f;
    if (!zok && !(za >= wa * 0x1p127f * 2.0f))   // |w| 2^128, or infinity when that overflows
        zok = __builtin_isfinite(k[2] / kw);
    float u = k[0] / kw, v = k[1] / kw;
    u = u * p.ndcU + p.ndcU;                                       // :310-313
    v = v * p.ndcV + p.ndcV;
    if ((u >= 0.0f) & (u < p.maxU) & (v >= 0.0f) & (v < p.maxV) & zok & wok)
        return (int)((p.H - (unsigned)v - 1) * (unsigned)p.W + (unsigned)u);   // :319
    return -1;
}

// voxel_update: with the gathered depth (m) of the voxel's pixel, true and the
// clamped signed distance when the voxel receives an update (:319-323).  w is
// recomputed: the same products and sums in the same order as mat_vec's row 3.
template <typename K>
__device__ __forceinline__ bool voxel_update(K key, float d, const UpdateParams &p, const Mat4 &w2c,
                                             float &sd_out)
{
    uint32_t x, y, z;
    unpack_key<K>(key, x, y, z);
    const float cx = voxel_centre<K>(x, p), cy = voxel_centre<K>(y, p), cz = voxel_centre<K>(z, p);
    const float kw = (w2c.m[3] * cx + w2c.m[7] * cy) + (w2c.m[11] * cz + w2c.m[15] * 1.0f);
    const float sd = d - kw;                                       // :319-320
    if (d > 0.0f && sd >= -p.trunc) {                              // :322
        sd_out = p.trunc < sd ? p.trunc : sd;                      // std::min, :323
        return true;
    }
    return false;
}

template <typename K>
__device__ __forceinline__ bool voxel_sample(K key, const UpdateParams &p, const Mat4 &w2c, const ClipPre &pre,
                                             float &sd_out)
{
    const int pix = voxel_pixel<K>(key, p, w2c, pre);
    return pix >= 0 && voxel_update<K>(key, p.depth[pix], p, w2c, sd_out);
}

struct VoxF32 { float sum, weight; };

__device__ __forceinline__ void voxel_apply(VoxF32 &v, float s, const float *)
{
    v.sum += s;                                                    // Voxel.h:53-57
    v.weight++;
}

// binary16 distance + uint8 weight; binary32 update d' = (d*w + s) * R[w] with
// R[w] = 1/(w + 1) correctly rounded (the table `recip`, in LDS)
__device__ __forceinline__ void voxel_apply(uint32_t &v, float s, const float *recip)
{
    const uint32_t w = (v >> 16) & 0xffu;
    const float d = (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu));
    const float wf = (float)w;
    float dn = (d * wf + s) * recip[w];
    // the product is rounded to binary32, then to binary16: without this the
    // compiler fuses multiply and conversion into one mixed-precision FMA,
    // which rounds once, straight to binary16
    asm volatile("" : "+v"(dn));
    const uint32_t wn = w < 255u ? w + 1u : 255u;
    v = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)dn) | (wn << 16);
}

// Model prediction folded into the voxel pass of a tracked frame: every
// near-surface voxel (|D| < dmax after its update) is splatted into the model
// z-buffer at the frame's pose, keyed by its index in the merged volume.  Same
// arithmetic as the stand-alone model splat (kifi_icp.hip).
__device__ __forceinline__ float vox_weight(const VoxF32 &v) { return v.weight; }
__device__ __forceinline__ float vox_dist(const VoxF32 &v) { return v.sum / v.weight; }
__device__ __forceinline__ float vox_weight(uint32_t v) { return (float)((v >> 16) & 0xffu); }
__device__ __forceinline__ float vox_dist(uint32_t v)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu));
}

__device__ __forceinline__ uint32_t orderable_z(float z)
{
    if (z == 0.0f)
        z = 0.0f;
    const uint32_t b = __float_as_uint(z);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// |D| < dmax of the stored voxel.  Binary32 sum/weight: a product test decides
// all but the cases within 2^-20 of the bound (the rounded quotient can only
// land on dmax inside that margin), which take the exact division.
__device__ __forceinline__ bool near_surface(const VoxF32 &v, float dmax)
{
    const float a = fabsf(v.sum), t = dmax * v.weight;
    if (a >= t * (1.0f + 0x1p-20f))
        return false;
    if (a <= t * (1.0f - 0x1p-20f))
        return true;
    return fabsf(v.sum / v.weight) < dmax;
}
__device__ __forceinline__ bool near_surface(uint32_t v, float dmax) { return fabsf(vox_dist(v)) < dmax; }

// The voxel enters the model: weighted and within dmax of the surface.
template <typename V>
__device__ __forceinline__ bool splat_candidate(const V &val, const SplatParams &sp)
{
    return vox_weight(val) > 0.0f && near_surface(val, sp.dmax);
}

// The voxel's eye coordinates and whether it can reach the splat's image: in
// front of the near plane and not rejected by the multiply-only test well
// outside the image (2 % + 1 px slack, rounding-safe).
template <typename K>
__device__ __forceinline__ bool splat_visible(K key, const SplatParams &sp, const Mat4 &w2e, float &z, float &fxe,
                                              float &fye)
{
    uint32_t x, y, zz;
    unpack_key<K>(key, x, y, zz);
    const float p[4] = { (float)x * sp.vl + sp.c0, (float)y * sp.vl + sp.c0, (float)zz * sp.vl + sp.c0, 1.0f };
    float e[4];
    mat_vec(w2e, p, e);
    z = -e[2];
    if (!(z > sp.znear))
        return false;
    fxe = sp.fx * e[0];
    fye = sp.fy * e[1];
    return !(fxe > ((float)sp.W + 8.0f - sp.cx) * z * 1.02f + 1.0f || fxe < (-8.0f - sp.cx) * z * 1.02f - 1.0f ||
             -fye > ((float)sp.H + 8.0f - sp.cy) * z * 1.02f + 1.0f || -fye < (-8.0f - sp.cy) * z * 1.02f - 1.0f);
}

// Splat of a candidate voxel with merged index o into the model z-buffer.
template <typename K>
__device__ __forceinline__ void splat_point(K key, uint64_t o, const SplatParams &sp, const Mat4 &w2e)
{
    float z, fxe, fye;
    if (!splat_visible<K>(key, sp, w2e, z, fxe, fye))
        return;
    const float u = fxe / z + sp.cx;
    const float v = sp.cy - fye / z;
    if (!(u > -8.0f && u < (float)sp.W + 8.0f && v > -8.0f && v < (float)sp.H + 8.0f))
        return;
    const float kf = ceilf(sp.vl * sp.fx / z);
    const int k = kf < 1.0f ? 1 : (kf > 3.0f ? 3 : (int)kf);
    const float half = 0.5f * (float)(k - 1);
    const int u0 = (int)floorf(u - half), v0 = (int)floorf(v - half);
    const unsigned long long zk = (unsigned long long)orderable_z(z) << 32 | (uint32_t)o;
    for (int dv = 0; dv < k; ++dv)
        for (int du = 0; du < k; ++du) {
            const int pu = u0 + du, pv = v0 + dv;
            if (pu < 0 || pu >= sp.W || pv < 0 || pv >= sp.H)
                continue;
            // no read of the slot first: a non-returning atomic does not hold the
            // tile up, a read would (measured: merge+update 2.6% faster without)
            atomicMin(&sp.zbuf[(size_t)pv * sp.W + pu], zk);
        }
}

// A tile's splat goes through an LDS table first: pixel -> minimum 64-bit key,
// kZSlots slots, open addressing with at most kZProbe probes; a pixel that
// finds no slot goes straight to the global z-buffer.  The tile then issues one
// global atomic per occupied slot.  The candidates of one tile (consecutive
// keys: neighbouring voxel rows) hit the same pixels often, and the global
// 64-bit atomics are what the splat costs.  The minimum is the same.
constexpr int kZSlots = 512;
constexpr int kZProbe = 8;
constexpr uint32_t kZEmpty = 0xffffffffu;

__device__ __forceinline__ void zslot_min(uint32_t *tkey, unsigned long long *tval, unsigned long long *zbuf,
                                          uint32_t pix, unsigned long long zk)
{
    uint32_t s = (pix * 2654435761u) >> (32 - 9);      // kZSlots = 2^9
    for (int probe = 0; probe < kZProbe; ++probe) {
        const uint32_t prev = atomicCAS(&tkey[s], kZEmpty, pix);
        if (prev == kZEmpty || prev == pix) {
            atomicMin(&tval[s], zk);
            return;
        }
        s = (s + 1) & (kZSlots - 1);
    }
    atomicMin(&zbuf[pix], zk);
}

template <typename K>
__device__ __forceinline__ void splat_point_lds(K key, uint64_t o, const SplatParams &sp, const Mat4 &w2e,
                                                uint32_t *tkey, unsigned long long *tval)
{
    float z, fxe, fye;
    if (!splat_visible<K>(key, sp, w2e, z, fxe, fye))
        return;
    const float u = fxe / z + sp.cx;
    const float v = sp.cy - fye / z;
    if (!(u > -8.0f && u < (float)sp.W + 8.0f && v > -8.0f && v < (float)sp.H + 8.0f))
        return;
    const float kf = ceilf(sp.vl * sp.fx / z);
    const int k = kf < 1.0f ? 1 : (kf > 3.0f ? 3 : (int)kf);
    const float half = 0.5f * (float)(k - 1);
    const int u0 = (int)floorf(u - half), v0 = (int)floorf(v - half);
    const unsigned long long zk = (unsigned long long)orderable_z(z) << 32 | (uint32_t)o;
    for (int dv = 0; dv < k; ++dv)
        for (int du = 0; du < k; ++du) {
            const int pu = u0 + du, pv = v0 + dv;
            if (pu < 0 || pu >= sp.W || pv < 0 || pv >= sp.H)
                continue;
            zslot_min(tkey, tval, sp.zbuf, (uint32_t)(pv * sp.W + pu), zk);
        }
}

template <typename K, typename V>
__device__ __forceinline__ void splat_voxel(K key, const V &val, uint64_t o, const SplatParams &sp, const Mat4 &w2e)
{
    if (splat_candidate(val, sp))
        splat_point<K>(key, o, sp, w2e);
}

// LDS index of the shift histogram: one pad word per 32 keeps a lane's 16
// consecutive entries (the scan) and a wave's consecutive entries (the old-voxel
// pass) on distinct banks
__device__ __forceinline__ uint32_t hpad(uint32_t i) { return i + (i >> 5); }
constexpr int kHistWords = kMgTile + kMgTile / 32;

template <typename K, typename V, bool SPLAT>
__global__ __launch_bounds__(kThreads) void k_merge_update(
    const K *__restrict__ A, const V *__restrict__ Av, const K *__restrict__ B, const uint64_t *__restrict__ on,
    const uint64_t *__restrict__ part_a, const uint64_t *__restrict__ part_b,
    const FrameCounters *__restrict__ cnt, K *__restrict__ O, V *__restrict__ Ov, K *__restrict__ OS1,
    K *__restrict__ OS2, UpdateParams up, SplatParams sp)
{
    __shared__ K sA[kMgTile];
    __shared__ uint32_t h[kHistWords];   // per A position: new keys sorting just below it
    __shared__ uint32_t tmp[kThreads / kWave + 1];
    constexpr bool kF16 = sizeof(V) == 4;
    __shared__ float recip[kF16 ? 256 : 1];   // binary16 storage: R[w] = 1/(w + 1)
    __shared__ uint32_t ztk[SPLAT ? kZSlots : 1];              // the splat's pixel table
    __shared__ unsigned long long ztv[SPLAT ? kZSlots : 1];
    const uint64_t T = cnt->n_tiles;
    if (blockIdx.x >= T || cnt->overflow)
        return;
    if (kF16)
        recip[threadIdx.x] = 1.0f / (float)(threadIdx.x + 1);   // correctly rounded; visible after the barrier below
    // the frame matrix into registers once (the output stores would otherwise
    // force a reload per voxel); the splat's matrix is loaded where it is used,
    // so the two are not live together
    const Mat4 w2c = up.mats->w2c;
    const ClipPre pre = up.mats->pre;
    const uint64_t upd_end = cnt->v_new / 4 * 4;                   // Integrator.cpp:269
    // the tile: merged outputs [obase, obase + kMgTile), old voxels [a0, a1),
    // new voxels [b0, b1) (none of them in the volume, each with its merged
    // position)
    const uint64_t a0 = part_a[blockIdx.x], a1 = part_a[blockIdx.x + 1];
    const uint64_t b0 = part_b[blockIdx.x], b1 = part_b[blockIdx.x + 1];
    const uint64_t obase = (uint64_t)blockIdx.x * kMgTile;
    // ranges that do not add up to the tile (never, unless the partition is
    // broken) are not touched: no access outside the buffers
    if (a1 < a0 || b1 < b0 || a1 > cnt->v_old || b1 > cnt->n_kept ||
        (a1 - a0) + (b1 - b0) != min((uint64_t)kMgTile, cnt->v_new - obase))
        return;
    const uint32_t na = (uint32_t)(a1 - a0), nb = (uint32_t)(b1 - b0);
    // outputs by 32-bit offset from the tile's base (offsets < 2^13); updates
    // stop at offset ulim
    K *const Ot = O + obase;
    V *const Ovt = Ov + obase;
    const uint32_t ulim = upd_end > obase ? (uint32_t)min(upd_end - obase, (uint64_t)0xffffffffu) : 0u;
    // Every global load of the tile is issued up front: old keys, new keys and
    // their positions, then old values, which stay in registers until the
    // old-voxel pass.  Element i of either range belongs to lane i % kThreads, so
    // the loads coalesce and a tile's few new keys spread over many lanes.
    constexpr int PER = kMgTile / kThreads;
    K ak[PER], bk[PER];
    uint32_t bp[PER];   // a new key's offset in the tile
    V av[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = q * kThreads + threadIdx.x;
        if (i < na)
            ak[q] = A[a0 + i];
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = q * kThreads + threadIdx.x;
        if (i < nb) {
            bk[q] = B[b0 + i];
            bp[q] = (uint32_t)(on[b0 + i] - obase);
        }
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = q * kThreads + threadIdx.x;
        if (i < na)
            av[q] = Av[a0 + i];
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = q * kThreads + threadIdx.x;
        if (i < na)
            sA[i] = ak[q];
    }
    __syncthreads();
    // New voxels: Voxel() default, then the update.  New key j of the tile at
    // offset p has p - j old keys of the tile before it.
    uint32_t lbq[PER], nupd = 0;   // nupd: voxels this lane updated (A)
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t j = q * kThreads + threadIdx.x;
        lbq[q] = kMgTile;   // counts for no old voxel
        // (the offset check holds for a correct partition; it keeps a broken
        // one inside the tile)
        if (j < nb && bp[q] < na + nb && bp[q] - j <= na) {
            const uint32_t ol = bp[q];
            lbq[q] = ol - j;
            V val{};
            float s;
            if (ol < ulim && voxel_sample<K>(bk[q], up, w2c, pre, s)) {
                voxel_apply(val, s, recip);
                ++nupd;
            }
            Ot[ol] = bk[q];
            Ovt[ol] = val;
            if ((ol & ((1u << kVolS1) - 1)) == 0)
                OS1[(obase + ol) >> kVolS1] = bk[q];
            if (((obase + ol) & 4095u) == 0)
                OS2[(obase + ol) >> kS2Shift] = bk[q];
            if (SPLAT)
                splat_voxel<K, V>(bk[q], val, obase + ol, sp, *sp.w2e);
        }
    }
    const uint32_t nbk = nb;
    // shift of A[i] = kept new keys below it = kept keys with lb <= i: an LDS
    // histogram of the lbs and an inclusive scan over the tile's A positions.
    // nbk is uniform; a tile that inserts nothing copies its old voxels in place.
    if (nbk > 0) {
        for (uint32_t i = threadIdx.x; i < kHistWords; i += kThreads)
            h[i] = 0;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (lbq[q] < na)
                atomicAdd(&h[hpad(lbq[q])], 1u);
        __syncthreads();
        uint32_t loc[PER], sum = 0;
        const uint32_t base = threadIdx.x * PER;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            loc[q] = h[hpad(base + q)];
            sum += loc[q];
        }
        uint32_t total;
        uint32_t run = block_exclusive_scan<kThreads>(sum, tmp, total);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            run += loc[q];
            h[hpad(base + q)] = run;
        }
        __syncthreads();
    }
    // old voxels keep their value (first range wins), then update: project all
    // of the lane's voxels, gather their depths together, then finish each
    int pix[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = q * kThreads + threadIdx.x;
        pix[q] = -1;
        if (i < na && i + (nbk ? h[hpad(i)] : 0u) < ulim)
            pix[q] = voxel_pixel<K>(sA[i], up, w2c, pre);
    }
    // unconditional loads (pixel 0 stands in for "none"): a load under a branch
    // would be waited for before its value merges, one gather at a time
    float dmm[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q)
        dmm[q] = up.depth[pix[q] >= 0 ? pix[q] : 0];
    uint32_t cand = 0;                 // splat candidates among the lane's voxels
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = q * kThreads + threadIdx.x;
        if (i < na) {
            const K k = sA[i];
            const uint32_t ol = i + (nbk ? h[hpad(i)] : 0u);
            V val = av[q];
            float s;
            if (pix[q] >= 0 && voxel_update<K>(k, dmm[q], up, w2c, s)) {
                voxel_apply(val, s, recip);
                ++nupd;
            }
            Ot[ol] = k;
            Ovt[ol] = val;
            if ((ol & ((1u << kVolS1) - 1)) == 0)
                OS1[(obase + ol) >> kVolS1] = k;
            if (((obase + ol) & 4095u) == 0)
                OS2[(obase + ol) >> kS2Shift] = k;
            if (SPLAT)
                cand |= (uint32_t)splat_candidate(val, sp) << q;
        }
    }
    if (SPLAT) {
        // The candidates are a minority spread over every wave: splatting them
        // inline would run the projection for whole waves.  They are compacted
        // into a list (tile index | output offset << 16, both < 2^16) in the
        // histogram's space and splatted densely.
        uint32_t ent[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const uint32_t i = q * kThreads + threadIdx.x;
            ent[q] = (cand >> q & 1u) ? i | ((i + (nbk ? h[hpad(i)] : 0u)) << 16) : 0u;
        }
        uint32_t ncand;
        uint32_t pos = block_exclusive_scan<kThreads>((uint32_t)__builtin_popcount(cand), tmp, ncand);
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (cand >> q & 1u)
                h[pos++] = ent[q];
        if (ncand) {   // block-uniform
            for (int t = threadIdx.x; t < kZSlots; t += kThreads) {
                ztk[t] = kZEmpty;
                ztv[t] = ~0ull;
            }
            __syncthreads();   // the list in h and the cleared table
            const Mat4 w2e = *sp.w2e;
            for (uint32_t t = threadIdx.x; t < ncand; t += kThreads) {
                const uint32_t e = h[t];
                splat_point_lds<K>(sA[e & 0xffffu], obase + (e >> 16), sp, w2e, ztk, ztv);
            }
            __syncthreads();
            for (int t = threadIdx.x; t < kZSlots; t += kThreads) {
                const uint32_t pix = ztk[t];
                if (pix != kZEmpty)
                    atomicMin(&sp.zbuf[pix], ztv[t]);
            }
        }
    }
    // A, the updated voxels: one atomic per tile on the tile's running counter
    nupd = block_reduce_sum<kThreads>(nupd, tmp);
    if (threadIdx.x == 0 && nupd)
        atomicAdd(&up.upd[(blockIdx.x % kUpdSlots) * kUpdStride], (unsigned long long)nupd);
}

// ====================================================================== host
kifi_status grow_volume(kifi_volume *vol, size_t needed);

// scratch layout for one integration (all offsets 256-byte aligned)
struct IntegrateScratch {
    void     *candA, *candB;
    uint32_t *rs_counts;     // [digits][Tb]
    uint32_t *rs_totals;     // [passes][digits]
    uint32_t *chunk_kept;    // [Tb]
    uint32_t *chunk_heads;   // [Tb]
    uint32_t *tile_n;        // [Tb] keys per keygen tile
    uint64_t *rank;          // [8P] the kept candidates' volume ranks, per chunk
    uint64_t *on;            // [8P] merged positions of the new voxels
    float    *depth_m;       // [P] a uint16 frame in metres
    size_t    bytes;
};

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

static IntegrateScratch layout_scratch(char *base, size_t P, size_t kbytes, int passes, int digits)
{
    IntegrateScratch s{};
    const size_t ncand = 8 * P;
    const size_t Tb = (ncand + kRsTile - 1) / kRsTile;
    size_t off = 0;
    // whole tiles: the sort reads a tile's keys before it knows the count
    s.candA = base + off; off += align256(Tb * kRsTile * kbytes);
    s.candB = base + off; off += align256(Tb * kRsTile * kbytes);
    s.rs_counts = (uint32_t *)(base + off); off += align256((size_t)digits * Tb * 4);
    s.rs_totals = (uint32_t *)(base + off); off += align256((size_t)passes * digits * 4);
    s.chunk_kept = (uint32_t *)(base + off); off += align256(Tb * 4);
    s.chunk_heads = (uint32_t *)(base + off); off += align256(Tb * 4);
    s.tile_n = (uint32_t *)(base + off); off += align256(Tb * 4);
    s.rank = (uint64_t *)(base + off); off += align256(ncand * 8);
    s.on = (uint64_t *)(base + off); off += align256(ncand * 8);
    s.depth_m = (float *)(base + off); off += align256(P * 4);
    s.bytes = off;
    return s;
}

template <typename K, typename V>
static kifi_status integrate_impl(kifi_volume *vol, DepthIn d_depth, const kifi_intrinsics *cam,
                                  const float *w2e, const Mat4 *w2e_dev, const SplatParams *splat)
{
    kifi_ctx *ctx = vol->ctx;
    const int W = cam->width, H = cam->height;
    const size_t P = (size_t)W * H;
    // the bits a key can use: z << 2 kb | y << kb | x with every axis below res
    constexpr int DB = KeyTraits<K>::digit_bits;
    constexpr int ND = 1 << DB;
    int res_bits = 0;
    while ((1 << res_bits) < vol->desc.res)
        ++res_bits;
    const int passes = (2 * KeyTraits<K>::kb + res_bits + DB - 1) / DB;
    const size_t Tb = (8 * P + kRsTile - 1) / kRsTile;

    // Capacity: the volume can grow by at most 8P voxels per frame.  This
    // frame's ring slot is free once the frame that used it kSizeRing frames
    // ago has finished (the wait leaves the frames after it queued).
    constexpr int kRing = kifi_volume::kSizeRing;
    const int slot = (int)(vol->frames_enq % kRing);
    if (vol->ring_frame[slot])
        KIFI_HIP(ctx, hipEventSynchronize(vol->ring_ev[slot]));
    ring_refresh(vol);
    size_t bound = vol->known_size + 8 * P * (size_t)(vol->frames_enq - vol->known_frame);   // V0 <= bound
    if (bound + 8 * P > vol->cap) {
        KIFI_HIP(ctx, hipStreamSynchronize(ctx->stream));
        uint64_t exact = 0;
        KIFI_HIP(ctx, hipMemcpy(&exact, vol->d_size, 8, hipMemcpyDeviceToHost));
        ring_settle(vol, exact);
        bound = exact;
        if (exact + 8 * P * kRing > vol->cap) {
            // room for a full ring of frames in flight, so that the bound does
            // not force another synchronisation before the ring throttles
            // (the capacity grows by factors of 1.5, so such syncs stay rare)
            kifi_status st = grow_volume(vol, exact + 8 * P * (kRing + 1));
            if (st != KIFI_OK)
                return st;
        }
    }
    const size_t Tm = (bound + 8 * P + kMgTile - 1) / kMgTile + 1;

    // scratch: integration buffers + merge partitions
    const IntegrateScratch probe = layout_scratch(nullptr, P, sizeof(K), passes, ND);
    const size_t merge_bytes = 2 * align256((Tm + 2) * 8);
    kifi_status st = ensure_scratch(ctx, probe.bytes + merge_bytes);
    if (st != KIFI_OK)
        return st;
    char *base = (char *)ctx->scratch;
    IntegrateScratch s = layout_scratch(base, P, sizeof(K), passes, ND);
    uint64_t *part_a = (uint64_t *)(base + s.bytes);
    uint64_t *part_b = (uint64_t *)(base + s.bytes + align256((Tm + 2) * 8));

    FrameCounters *cnt = vol->d_counters;
    hipStream_t stream = ctx->stream;

    Mat4 clip, wval;
    eye_to_clip_rh(cam->width, cam->height, cam->fx, cam->fy, cam->cx, cam->cy, cam->znear, cam->zfar, clip.m);
    for (int i = 0; i < 16; ++i)
        wval.m[i] = w2e ? w2e[i] : 0.0f;
    const bool model_start = splat && splat->pose_out;

    KeygenParams kp;
    kp.W = W; kp.H = H;
    kp.flInvX = 1.0f / cam->fx;
    kp.flInvY = 1.0f / cam->fy;
    kp.ppX = (0.5f - cam->cx) * kp.flInvX;                          // Integrator.cpp:95-101
    kp.ppY = (cam->cy - 0.5f) * kp.flInvY;
    kp.maxIndex = (float)(vol->desc.res - 1);
    kp.vc = vol->consts;
    kp.fs.clip = clip;
    kp.fs.w2e_val = wval;
    kp.fs.w2e_dev = const_cast<Mat4 *>(w2e_dev);
    kp.fs.mats = vol->d_mats;
    kp.fs.model_out = model_start ? splat->pose_out : nullptr;
    kp.fs.icp_end = splat ? splat->icp_end : nullptr;
    kp.fs.cnt = cnt;
    kp.zbuf = model_start ? splat->zbuf : nullptr;
    kp.nz = model_start ? splat->W * splat->H : 0;
    const int t_sort = timer_begin(ctx, KIFI_STAGE_SORT);
    // keygen: one workgroup per radix tile, with the first pass's tile counts
    const size_t kg_blocks = (P + kThreads * kKgItems - 1) / (kThreads * kKgItems);
    if (kg_blocks != Tb)
        return set_error(ctx, KIFI_ERR_STATE, "integrate: keygen tiles do not match the radix tiles");
    const float *frame_m = d_depth.metres ? (const float *)d_depth.p : s.depth_m;
    if (d_depth.metres)
        hipLaunchKernelGGL((k_keygen<K, float, DB>), dim3((unsigned)Tb), dim3(kThreads), 0, stream,
                           (const float *)d_depth.p, kp, (K *)s.candA, cnt, (float *)nullptr, s.rs_counts, (int)Tb,
                           s.tile_n);
    else
        hipLaunchKernelGGL((k_keygen<K, uint16_t, DB>), dim3((unsigned)Tb), dim3(kThreads), 0, stream,
                           (const uint16_t *)d_depth.p, kp, (K *)s.candA, cnt, s.depth_m, s.rs_counts, (int)Tb,
                           s.tile_n);

    // radix sort of the candidates (the sorted keys end in src); the first
    // pass sorts keygen's partly filled tiles into dense order
    K *src = (K *)s.candA, *dst = (K *)s.candB;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = DB * pass;
        uint32_t *tot = s.rs_totals + pass * ND;
        if (pass > 0)
            hipLaunchKernelGGL((k_rs_upsweep<K, DB>), dim3((unsigned)Tb), dim3(kThreads), 0, stream, src,
                               &cnt->n_cand, shift, s.rs_counts, (int)Tb);
        hipLaunchKernelGGL(k_rs_scan, dim3(pass == 0 ? ND + 1 : ND), dim3(kThreads), 0, stream, s.rs_counts,
                           (int)Tb, tot, pass == 0 ? s.tile_n : nullptr, cnt);
        hipLaunchKernelGGL((k_rs_downsweep<K, DB>), dim3((unsigned)Tb), dim3(kThreads), 0, stream, src, dst,
                           &cnt->n_cand, shift, s.rs_counts, (int)Tb, tot, pass == 0 ? s.tile_n : nullptr);
        K *t = src; src = dst; dst = t;
    }

    // new voxels, their merged positions and the merge's tile partition
    const int cur = vol->cur, nxt = cur ^ 1;
    if (!vol->samples_valid) {
        const size_t n1 = s1_count<kVolS1>(bound);
        const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((n1 + kThreads - 1) / kThreads, 4096));
        hipLaunchKernelGGL(k_build_samples<K>, dim3(g), dim3(kThreads), 0, stream, (const K *)vol->keys[cur],
                           vol->d_size, (K *)vol->samp1[cur], (K *)vol->samp2[cur]);
        vol->samples_valid = true;
    }
    // the kept keys per chunk go to dst (free after the sort), then compacted
    // into src (the sorted candidates are not needed after k_locate)
    hipLaunchKernelGGL(k_locate<K>, dim3((unsigned)Tb), dim3(kThreads), 0, stream, src, &cnt->n_cand,
                       (const K *)vol->keys[cur], (const K *)vol->samp1[cur], (const K *)vol->samp2[cur],
                       vol->d_size, cnt, dst, s.rank, s.chunk_kept, s.chunk_heads);
    hipLaunchKernelGGL(k_compact<K>, dim3((unsigned)Tb), dim3(kThreads), 0, stream, dst, s.rank, &cnt->n_cand,
                       s.chunk_kept, s.chunk_heads, cnt, src, s.on, part_a, part_b, (int)Tm, vol->d_size,
                       (uint64_t)vol->cap, vol->h_ring_dev + slot);
    timer_end(ctx, KIFI_STAGE_SORT, t_sort);
    const K *Bnew = src;
    UpdateParams up;
    up.mats = vol->d_mats;
    up.vl = vol->consts.vl;
    up.c0 = vol->consts.vlHalfPlusMin;
    up.trunc = vol->consts.trunc;
    up.ndcU = (float)W * 0.5f;                                     // Integrator.cpp:262
    up.ndcV = (float)H * 0.5f;
    up.maxU = (float)W;                                            // :265
    up.maxV = (float)H;
    up.W = W; up.H = H;
    up.depth = frame_m;
    up.upd = (unsigned long long *)vol->d_upd;
    const int t_mu = timer_begin(ctx, KIFI_STAGE_MERGE_UPDATE);
    if (splat)
        hipLaunchKernelGGL((k_merge_update<K, V, true>), dim3((unsigned)Tm), dim3(kThreads), 0, stream,
                           (const K *)vol->keys[cur], (const V *)vol->vals[cur], Bnew, s.on, part_a, part_b,
                           cnt, (K *)vol->keys[nxt], (V *)vol->vals[nxt], (K *)vol->samp1[nxt],
                           (K *)vol->samp2[nxt], up, *splat);
    else
        hipLaunchKernelGGL((k_merge_update<K, V, false>), dim3((unsigned)Tm), dim3(kThreads), 0, stream,
                           (const K *)vol->keys[cur], (const V *)vol->vals[cur], Bnew, s.on, part_a, part_b,
                           cnt, (K *)vol->keys[nxt], (V *)vol->vals[nxt], (K *)vol->samp1[nxt],
                           (K *)vol->samp2[nxt], up, SplatParams{});
    timer_end(ctx, KIFI_STAGE_MERGE_UPDATE, t_mu);
    KIFI_HIP(ctx, hipGetLastError());
    vol->cur = nxt;

    // k_compact wrote the counters into the frame's host-mapped ring slot;
    // the event marks when they may be read to refresh the host size bound
    KIFI_HIP(ctx, hipEventRecord(vol->ring_ev[slot], stream));
    vol->ring_frame[slot] = ++vol->frames_enq;
    return KIFI_OK;
}

kifi_status integrate_dispatch(kifi_volume *vol, DepthIn d_depth, const kifi_intrinsics *cam,
                               const float *w2e, const Mat4 *w2e_dev, const SplatParams *splat)
{
    if (vol->desc.key_bits == 32) {
        if (vol->desc.storage == KIFI_STORAGE_F32)
            return integrate_impl<uint32_t, VoxF32>(vol, d_depth, cam, w2e, w2e_dev, splat);
        return integrate_impl<uint32_t, uint32_t>(vol, d_depth, cam, w2e, w2e_dev, splat);
    }
    if (vol->desc.storage == KIFI_STORAGE_F32)
        return integrate_impl<uint64_t, VoxF32>(vol, d_depth, cam, w2e, w2e_dev, splat);
    return integrate_impl<uint64_t, uint32_t>(vol, d_depth, cam, w2e, w2e_dev, splat);
}

} // namespace kifi
