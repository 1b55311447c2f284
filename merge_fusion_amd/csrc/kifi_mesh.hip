// kifi_mesh.hip -- marching cubes over the sorted sparse volume and the
// z-buffered point splat (model prediction).
//
// Reference: Mesher::Generate<bool> / Mesher::Mesh (KinectFission/kifi/Mesher.cpp:14-45,
// 85-229) and Renderer::Bin (Renderer.cpp:39-80).
//
// The reference sweeps voxels i = V-2 .. 0 with three monotone pointers whose
// value for voxel i is min(V-2, upper_bound(key_i + delta) - 1); here every
// voxel computes them independently with a wave-cooperative windowed search.
// Outputs are placed by scans of per-voxel counts: the generation order of the
// reference is the descending voxel order, and its vertex-ID order (3*key+axis)
// is the ascending voxel order, so the vertex/index join needs no sort.
#include <cfloat>

#include "kifi_internal.hpp"
#include "kifi_prims.hpp"

namespace kifi {

constexpr int kMT = 256;
constexpr int kScanTile = 4096;

// mc_tables.inc: kTriOffsets[257], kTriTable[2460] in __constant__ memory
#include "mc_tables.inc"

// ------------------------------------------------------------- device scan
// Exclusive scan of n uint32 counts into out (may alias), total to *total.
__global__ __launch_bounds__(kMT) void k_scan_reduce(const uint32_t *__restrict__ in, uint64_t n,
                                                     uint32_t *__restrict__ tile_sums)
{
    __shared__ uint32_t tmp[kMT / kWave + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int q = 0; q < kScanTile / kMT; ++q) {
        const uint64_t i = base + (uint64_t)q * kMT + threadIdx.x;
        if (i < n)
            s += in[i];
    }
    s = block_reduce_sum<kMT>(s, tmp);
    if (threadIdx.x == 0)
        tile_sums[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(uint32_t *__restrict__ tile_sums, uint64_t nt,
                                                     uint64_t *__restrict__ total)
{
    block_scan_array<1024>(tile_sums, (size_t)nt, total);
}

__global__ __launch_bounds__(kMT) void k_scan_apply(const uint32_t *__restrict__ in, uint64_t n,
                                                    const uint32_t *__restrict__ tile_sums,
                                                    uint32_t *__restrict__ out)
{
    __shared__ uint32_t tmp[kMT / kWave + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    constexpr int PER = kScanTile / kMT;
    const uint64_t mine = base + (uint64_t)threadIdx.x * PER;
    uint32_t loc[PER], s = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        loc[q] = mine + q < n ? in[mine + q] : 0u;
        s += loc[q];
    }
    uint32_t tot;
    uint32_t run = tile_sums[blockIdx.x] + block_exclusive_scan<kMT>(s, tmp, tot);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        if (mine + q < n)
            out[mine + q] = run;
        run += loc[q];
    }
}

static void device_scan(hipStream_t st, const uint32_t *in, uint32_t *out, uint64_t n,
                        uint32_t *tile_sums, uint64_t *total)
{
    const uint64_t nt = (n + kScanTile - 1) / kScanTile;
    if (nt == 0) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, tile_sums, (uint64_t)0, total);
        return;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nt), dim3(kMT), 0, st, in, n, tile_sums);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, tile_sums, nt, total);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nt), dim3(kMT), 0, st, in, n, tile_sums, out);
}

// ------------------------------------------------------------ voxel access
struct VoxF32m { float sum, weight; };

__device__ __forceinline__ float vweight(const VoxF32m &v) { return v.weight; }
__device__ __forceinline__ float vdist(const VoxF32m &v) { return v.sum / v.weight; } // Voxel.h:39-44
__device__ __forceinline__ float vweight(uint32_t v) { return (float)((v >> 16) & 0xffu); }
__device__ __forceinline__ float vdist(uint32_t v)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu));
}

template <typename K>
__device__ __forceinline__ uint64_t upper_bound_g(const K *keys, uint64_t lo, uint64_t hi, K t)
{
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (keys[mid] <= t)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Neighbour search.  The reference walks pointers to the +y, +z (and +y+z)
// neighbours of each voxel (Mesher.cpp:122-132, 174-175); here that is
// min(V-2, upper_bound(keys, key + offset) - 1).  Voxels go in tiles of
// kMeshTile; for a tile the upper bounds of all its targets lie between those
// of its first voxel and of the next tile's first voxel, which k_mesh_windows
// finds by one binary search per tile boundary.  Each tile loads that key
// window [w_t - 1, w_{t+1}] into LDS and searches there; a window wider than
// kMeshWin keys is searched in global memory instead.
constexpr int kMeshTile = 1024;
constexpr int kMeshWin = 2048;

template <typename K, bool TRIS>
struct MeshOffsets {
    static constexpr int n = TRIS ? 3 : 2;
    __device__ static K get(int o)
    {
        constexpr K d2 = K(1) << KeyTraits<K>::kb, d4 = K(1) << (2 * KeyTraits<K>::kb);
        return o == 0 ? d2 : (o == 1 ? d4 : d4 + d2);
    }
};

// w[o * (T + 1) + t] = upper_bound(keys, keys[min(t * kMeshTile, V - 1)] + offset_o)
template <typename K, bool TRIS>
__global__ __launch_bounds__(kMT) void k_mesh_windows(const K *__restrict__ keys, uint64_t n, uint64_t T,
                                                      uint64_t *__restrict__ w)
{
    constexpr int NO = MeshOffsets<K, TRIS>::n;
    const uint64_t g = (uint64_t)blockIdx.x * kMT + threadIdx.x;
    if (g >= (T + 1) * NO)
        return;
    const int o = (int)(g / (T + 1));
    const uint64_t t = g % (T + 1);
    const uint64_t i = min(t * (uint64_t)kMeshTile, n - 1);
    w[g] = upper_bound_g(keys, 0, n, (K)(keys[i] + MeshOffsets<K, TRIS>::get(o)));
}

template <typename K, bool TRIS>
struct MeshWin {
    static constexpr int NO = MeshOffsets<K, TRIS>::n;
    K s[NO][kMeshWin];
};

// The tile's windows: [lo, hi) per offset, in LDS when lds.  Uniform over the block.
template <typename K, bool TRIS>
struct WinRange {
    uint64_t lo[MeshOffsets<K, TRIS>::n], hi[MeshOffsets<K, TRIS>::n];
    bool lds[MeshOffsets<K, TRIS>::n];
};

template <typename K, bool TRIS>
__device__ __forceinline__ WinRange<K, TRIS> mesh_load_windows(const K *keys, uint64_t n, const uint64_t *w,
                                                               uint64_t T, uint64_t t, MeshWin<K, TRIS> &mw)
{
    WinRange<K, TRIS> r;
#pragma unroll
    for (int o = 0; o < MeshOffsets<K, TRIS>::n; ++o) {
        const uint64_t a = min(w[o * (T + 1) + t], n - 1), b = w[o * (T + 1) + t + 1];
        r.lo[o] = a >= 1 ? a - 1 : 0;
        r.hi[o] = min(b + 1, n);
        r.lds[o] = r.hi[o] - r.lo[o] <= (uint64_t)kMeshWin;
        if (r.lds[o])
            for (uint32_t j = threadIdx.x; j < r.hi[o] - r.lo[o]; j += kMT)
                mw.s[o][j] = keys[r.lo[o] + j];
    }
    __syncthreads();
    return r;
}

// min(V-2, upper_bound(target) - 1) and the keys at that index and the next
template <typename K, bool TRIS>
__device__ __forceinline__ int64_t win_walk(const K *keys, uint64_t n, const MeshWin<K, TRIS> &mw,
                                            const WinRange<K, TRIS> &wr, int o, K target, K &k_at, K &k_next)
{
    uint64_t u;
    if (wr.lds[o]) {
        const K *s = mw.s[o];
        uint32_t a = 0, b = (uint32_t)(wr.hi[o] - wr.lo[o]);
        while (a < b) {
            const uint32_t mid = (a + b) >> 1;
            if (s[mid] <= target)
                a = mid + 1;
            else
                b = mid;
        }
        u = wr.lo[o] + a;
    } else {
        u = upper_bound_g(keys, wr.lo[o], wr.hi[o], target);
    }
    const int64_t cl = (int64_t)n - 2;
    int64_t r = (int64_t)u - 1;
    r = r < cl ? r : cl;
    if (wr.lds[o]) {
        k_at = mw.s[o][r - (int64_t)wr.lo[o]];
        k_next = mw.s[o][r + 1 - (int64_t)wr.lo[o]];
    } else {
        k_at = keys[r];
        k_next = keys[r + 1];
    }
    return r;
}

struct MeshParams {
    float vl, c0;
};

template <typename K>
struct Corners {
    int64_t c[8];
};

// Neighbour search and vertex flags of voxel i (Mesher.cpp:117-168).
template <typename K, typename V, bool TRIS>
__device__ __forceinline__ uint32_t voxel_topology(const K *keys, const V *vals, uint64_t n, uint64_t i,
                                                   const MeshWin<K, TRIS> &mw, const WinRange<K, TRIS> &wr,
                                                   Corners<K> &cn, int &lut)
{
    constexpr K d1 = 1, d2 = K(1) << KeyTraits<K>::kb, d4 = K(1) << (2 * KeyTraits<K>::kb);
    lut = -1;
    // the voxel's key, its successor's key and its value in flight together
    // (i <= V - 2, so i + 1 is a voxel)
    const K k = keys[i], kNext = keys[i + 1];
    const V v0 = vals[i];
    K kTop, kTop1, kFront, kFront1, kFT = 0, kFT1 = 0;
    const int64_t iTop = win_walk<K, TRIS>(keys, n, mw, wr, 0, k + d2, kTop, kTop1);       // :122-123
    const int64_t iFront = win_walk<K, TRIS>(keys, n, mw, wr, 1, k + d4, kFront, kFront1);  // :125-126
    int64_t iFrontTop = 0;
    if (TRIS)
        iFrontTop = win_walk<K, TRIS>(keys, n, mw, wr, 2, k + d4 + d2, kFT, kFT1);          // :174-175
    if (vweight(v0) == 0.0f)                                       // :119-120
        return 0;
    cn.c[0] = (int64_t)i;
    cn.c[1] = kNext == k + d1 ? (int64_t)i + 1 : -1;               // :132
    cn.c[2] = kTop == k + d2 ? iTop : -1;                          // :129
    cn.c[4] = kFront == k + d4 ? iFront : -1;                      // :130
    if (TRIS) {
        cn.c[6] = kFT == k + d4 + d2 ? iFrontTop : -1;             // :177
        cn.c[3] = kTop1 == k + d2 + d1 ? iTop + 1 : -1;            // :179
        cn.c[5] = kFront1 == k + d4 + d1 ? iFront + 1 : -1;
        cn.c[7] = kFT1 == k + d4 + d2 + d1 ? iFrontTop + 1 : -1;
    }
    // every corner value in flight together: absent corners read the voxel's
    // own value (in bounds), and the tests below ignore it
    constexpr int NC = TRIS ? 8 : 5;
    V vc[NC];
#pragma unroll
    for (int c = 1; c < NC; ++c)
        if (TRIS || c == 1 || c == 2 || c == 4)
            vc[c] = vals[cn.c[c] >= 0 ? (uint64_t)cn.c[c] : i];
    const float dself = vdist(v0);
    uint32_t flags = 0;
    const int axis_corner[3] = { 1, 2, 4 };
#pragma unroll
    for (int a = 0; a < 3; ++a) {                                  // :140-168
        const int ci = axis_corner[a];
        if (cn.c[ci] >= 0) {
            const V vn = vc[ci];
            if (vweight(vn) > 0.0f && dself * vdist(vn) <= 0.0f)
                flags |= 1u << a;
        }
    }
    if (TRIS) {
        bool skip = false;
        float d[8];
        d[0] = dself;
#pragma unroll
        for (int c = 1; c < 8; ++c) {                              // :183-188
            if (cn.c[c] < 0) {
                skip = true;
                d[c] = 0.0f;
                continue;
            }
            if (vweight(vc[c]) == 0.0f)
                skip = true;
            d[c] = skip ? 0.0f : vdist(vc[c]);
        }
        if (!skip)                                                 // :191-199
            lut = (d[0] < 0.0f) << 2 | (d[1] < 0.0f) << 3 | (d[2] < 0.0f) << 6 |
                  (d[3] < 0.0f) << 7 | (d[4] < 0.0f) << 1 | (d[5] < 0.0f) << 0 |
                  (d[6] < 0.0f) << 5 | (d[7] < 0.0f) << 4;
    }
    return flags;
}

// The indexed mesh (Mesh(volume, verts, idx)), pass 1: per-voxel vertex flags,
// vertex counts (forward: the vertex IDs' order) and index counts (reverse:
// generation order), a tile of kMeshTile voxels at a time.  TRIS selects the
// cube corners the topology reads (the vertex-only mesh has kernels of its own).
template <typename K, typename V, bool TRIS>
__global__ __launch_bounds__(kMT) void k_mesh_count(const K *__restrict__ keys, const V *__restrict__ vals,
                                                    uint64_t n, const uint64_t *__restrict__ win, uint64_t T,
                                                    uint8_t *__restrict__ vflags,
                                                    uint32_t *__restrict__ cnt_v_fwd,
                                                    uint32_t *__restrict__ cnt_i_rev)
{
    __shared__ MeshWin<K, TRIS> mw;
    const uint64_t m = n - 1;                           // voxels 0 .. V-2
    for (uint64_t t = blockIdx.x; t < T; t += gridDim.x) {
        const WinRange<K, TRIS> wr = mesh_load_windows<K, TRIS>(keys, n, win, T, t, mw);
        const uint64_t i0 = t * kMeshTile, i1 = min(i0 + kMeshTile, m);
        for (uint64_t i = i0 + threadIdx.x; i < i1; i += kMT) {
            Corners<K> cn;
            int lut;
            const uint32_t f = voxel_topology<K, V, TRIS>(keys, vals, n, i, mw, wr, cn, lut);
            const uint32_t nv = __popc(f);
            vflags[i] = (uint8_t)f;
            cnt_v_fwd[i] = nv;
            if (TRIS)
                cnt_i_rev[m - 1 - i] = lut >= 0 ? (uint32_t)(kTriOffsets[lut + 1] - kTriOffsets[lut]) : 0u;
        }
        __syncthreads();                                // the windows are reloaded for the next tile
    }
}

// Edge e of the cube -> (corner, axis), from localToGlobal (Mesher.cpp:206-218).
__constant__ int kEdgeCorner[12] = { 4, 0, 0, 1, 6, 2, 2, 3, 5, 4, 0, 1 };
__constant__ int kEdgeAxis[12] = { 0, 2, 0, 2, 0, 2, 0, 2, 1, 1, 1, 1 };
// the coordinate offsets of that corner (dx, dy, dz)
__constant__ int kCornerOff[8][3] = { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 1, 1, 0 },
                                      { 0, 0, 1 }, { 1, 0, 1 }, { 0, 1, 1 }, { 1, 1, 1 } };

// The indexed mesh, pass 2: vertices (in vertex-ID order) and indices at their
// scanned positions.
template <typename K, typename V, bool TRIS>
__global__ __launch_bounds__(kMT) void k_mesh_emit(const K *__restrict__ keys, const V *__restrict__ vals,
                                                   uint64_t n, const uint64_t *__restrict__ win, uint64_t T,
                                                   MeshParams mp,
                                                   const uint8_t *__restrict__ vflags,
                                                   const uint32_t *__restrict__ voff,   // first vertex per voxel (ID order)
                                                   const uint32_t *__restrict__ ioff_rev,
                                                   float *__restrict__ verts, uint32_t *__restrict__ idx,
                                                   uint64_t *__restrict__ raw_ids,
                                                   unsigned long long *__restrict__ stall_id)
{
    __shared__ MeshWin<K, TRIS> mw;
    const uint64_t m = n - 1;
    constexpr int kb = KeyTraits<K>::kb;
    for (uint64_t t = blockIdx.x; t < T; t += gridDim.x) {
      const WinRange<K, TRIS> wr = mesh_load_windows<K, TRIS>(keys, n, win, T, t, mw);
      const uint64_t i0 = t * kMeshTile, i1 = min(i0 + kMeshTile, m);
      for (uint64_t i = i0 + threadIdx.x; i < i1; i += kMT) {
        Corners<K> cn;
        int lut;
        const uint32_t f = voxel_topology<K, V, TRIS>(keys, vals, n, i, mw, wr, cn, lut);
        if (f) {
            const K k = keys[i];
            uint32_t x0, y0, z0;
            unpack_key<K>(k, x0, y0, z0);
            // Volume::VoxelCenter(int, int, int), Volume.cpp:80-88
            const float vx = (float)(int)x0 * mp.vl + mp.c0, vy = (float)(int)y0 * mp.vl + mp.c0,
                        vz = (float)(int)z0 * mp.vl + mp.c0;
            const V v0 = vals[i];
            const float dself = vdist(v0);
            uint32_t o = voff[i];
            const int axis_corner[3] = { 1, 2, 4 };
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (!(f >> a & 1u))
                    continue;
                const float dn = vdist(vals[cn.c[axis_corner[a]]]);
                const float t = fabsf(dself) / (fabsf(dself) + fabsf(dn)) * mp.vl; // :143/153/163
                verts[3 * (size_t)o + 0] = a == 0 ? vx + t : vx;
                verts[3 * (size_t)o + 1] = a == 1 ? vy + t : vy;
                verts[3 * (size_t)o + 2] = a == 2 ? vz + t : vz;
                ++o;
            }
        }
        if (TRIS && lut >= 0) {
            const K k = keys[i];
            uint32_t x0, y0, z0;
            unpack_key<K>(k, x0, y0, z0);
            uint32_t o = ioff_rev[m - 1 - i];
            for (int t = kTriOffsets[lut]; t < kTriOffsets[lut + 1]; ++t, ++o) {
                const int e = kTriTable[t];
                const int c = kEdgeCorner[e], a = kEdgeAxis[e];
                // the reference's vertex ID: pack(x0+dx, y0+dy, z0+dz) * 3 + axis (OR-pack)
                const K pk = pack_key<K>((K)x0 + kCornerOff[c][0], (K)y0 + kCornerOff[c][1],
                                         (K)z0 + kCornerOff[c][2]);
                int64_t vi = cn.c[c];
                if (keys[vi] != pk) {        // only where x0+1 or y0+1 carries into the next field
                    const uint64_t ub = upper_bound_g(keys, 0, n, pk);
                    vi = (ub > 0 && keys[ub - 1] == pk) ? (int64_t)ub - 1 : -1;
                }
                // a vertex exists only for voxels <= V-2 that emitted along this axis
                const bool have = vi >= 0 && (uint64_t)vi + 2 <= n && (vflags[vi] >> a & 1u);
                const uint64_t rid = 3ull * (uint64_t)pk + (uint64_t)a;
                raw_ids[o] = rid;
                if (have) {
                    const uint32_t fl = vflags[vi];
                    idx[o] = voff[vi] + __popc(fl & ((1u << a) - 1u));
                } else {
                    idx[o] = (uint32_t)rid;
                    atomicMin(stall_id, (unsigned long long)rid);
                }
                (void)kb;
            }
        }
      }
      __syncthreads();                                  // the windows are reloaded for the next tile
    }
}

// The reference's join stalls at the smallest unmatched index ID (Mesher.cpp:35-38):
// every index whose ID is >= that ID keeps its raw ID.
__global__ void k_mesh_stall_fix(const uint64_t *__restrict__ raw_ids, uint64_t ni,
                                 const unsigned long long *__restrict__ stall_id,
                                 uint32_t *__restrict__ idx)
{
    const unsigned long long s = *stall_id;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < ni;
         q += (uint64_t)gridDim.x * blockDim.x)
        if (raw_ids[q] >= s)
            idx[q] = (uint32_t)raw_ids[q];
}

// ------------------------------------------------- vertices only (Mesh(volume, verts))
// The vertex-only mesh needs each voxel's vertex flags and, for the output
// positions, only how many vertices the voxels after it emit.  Pass 1 keeps
// the flags (one byte per voxel) and one vertex count per tile; the tile
// counts are scanned in generation order (descending tiles); pass 2 reads the
// flags, places each voxel by a scan within its tile and revisits only the
// voxels that emit.
//
// Tiles go to XCDs in contiguous runs (workgroup b runs on XCD b % 8): the
// +y and +z key windows of consecutive tiles overlap, and a tile's keys were
// read shortly before as part of an earlier tile's windows, so they meet in
// the XCD's own L2.
constexpr int kVQ = kMeshTile / kMT;          // voxels per thread
constexpr int kWinRounds = kMeshWin / kMT;    // window loads per thread per offset
constexpr int kXcds = 8;

template <typename K>
struct VWin {
    K s[2][kMeshWin];                         // +y and +z windows
};

template <typename K>
struct VRange {
    uint64_t lo[2];
    uint32_t len[2];
    bool lds[2];
};

__device__ __forceinline__ uint64_t xcd_tile(uint64_t T)
{
    const uint64_t per = (T + kXcds - 1) / kXcds;
    return (uint64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

// The +y / +z windows of tile t ([w_t - 1, w_{t+1}] clipped to the volume,
// as in mesh_load_windows) and their fill into LDS: every load of both windows
// is issued before the first LDS store.
template <typename K>
__device__ __forceinline__ VRange<K> vwin_load(const K *keys, uint64_t n, const uint64_t *w, uint64_t T,
                                               uint64_t t, VWin<K> &mw)
{
    VRange<K> r;
    K x[2][kWinRounds];
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        const uint64_t a = min(w[o * (T + 1) + t], n - 1), b = w[o * (T + 1) + t + 1];
        const uint64_t lo = a >= 1 ? a - 1 : 0, hi = min(b + 1, n);
        r.lo[o] = lo;
        r.lds[o] = hi - lo <= (uint64_t)kMeshWin;
        r.len[o] = r.lds[o] ? (uint32_t)(hi - lo) : 0u;
        const uint32_t last = r.len[o] ? r.len[o] - 1 : 0u;
#pragma unroll
        for (int q = 0; q < kWinRounds; ++q)       // clamped: in bounds, the extra ones unused
            x[o][q] = keys[lo + min((uint32_t)(q * kMT) + threadIdx.x, last)];
    }
#pragma unroll
    for (int o = 0; o < 2; ++o)
#pragma unroll
        for (int q = 0; q < kWinRounds; ++q)
            if ((uint32_t)(q * kMT) + threadIdx.x < r.len[o])
                mw.s[o][q * kMT + threadIdx.x] = x[o][q];
    __syncthreads();
    return r;
}

// upper_bound of NC targets per offset (chains interleaved step by step) ->
// min(V-2, upper_bound - 1) in r and the key there in k_at.  Targets of an
// offset whose window did not fit LDS are searched in global memory.
template <typename K, int NC>
__device__ __forceinline__ void vwin_search(const K *keys, uint64_t n, const VWin<K> &mw, const VRange<K> &wr,
                                            const K (&tg)[2][NC], const bool (&need)[2][NC],
                                            int64_t (&r)[2][NC], K (&k_at)[2][NC])
{
    uint32_t a[2][NC];
    uint32_t L = 1;
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        L = max(L, wr.len[o]);
#pragma unroll
        for (int c = 0; c < NC; ++c)
            a[o][c] = 0;
    }
    for (uint32_t st = 1u << (31 - __clz(L)); st; st >>= 1) {
#pragma unroll
        for (int o = 0; o < 2; ++o)
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const uint32_t p = a[o][c] + st;
                const K x = mw.s[o][max(min(p, wr.len[o]), 1u) - 1];
                a[o][c] = (p <= wr.len[o] && x <= tg[o][c]) ? p : a[o][c];
            }
    }
    const int64_t cl = (int64_t)n - 2;
#pragma unroll
    for (int o = 0; o < 2; ++o)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            uint64_t u;
            if (wr.lds[o])
                u = wr.lo[o] + a[o][c];
            else
                u = need[o][c] ? upper_bound_g(keys, wr.lo[o], n, tg[o][c]) : wr.lo[o] + 1;
            const int64_t rr = min((int64_t)u - 1, cl);
            r[o][c] = rr;
            k_at[o][c] = wr.lds[o] ? mw.s[o][rr - (int64_t)wr.lo[o]] : keys[rr];
        }
}

// Pass 1: the vertex flags of voxels i0 + q * kMT + tid (Mesher.cpp:117-168)
// and the tile's vertex count at tile_rev[T - 1 - t].
template <typename K, typename V>
__global__ __launch_bounds__(kMT) void k_mesh_vcount(const K *__restrict__ keys, const V *__restrict__ vals,
                                                     uint64_t n, const uint64_t *__restrict__ win, uint64_t T,
                                                     uint8_t *__restrict__ vflags, uint32_t *__restrict__ tile_rev)
{
    __shared__ VWin<K> mw;
    __shared__ uint32_t red[kMT / kWave + 1];
    constexpr K d1 = 1, d2 = K(1) << KeyTraits<K>::kb, d4 = K(1) << (2 * KeyTraits<K>::kb);
    const uint64_t t = xcd_tile(T);
    if (t >= T)
        return;
    const uint64_t m = n - 1;
    const uint64_t i0 = t * kMeshTile, i1 = min(i0 + kMeshTile, m);
    // own keys, successors and values in flight with the windows
    K k[kVQ], kn[kVQ];
    V v0[kVQ];
#pragma unroll
    for (int q = 0; q < kVQ; ++q) {
        const uint64_t i = min(i0 + (uint64_t)(q * kMT) + threadIdx.x, m - 1);   // i + 1 <= V - 1
        k[q] = keys[i];
        kn[q] = keys[i + 1];
        v0[q] = vals[i];
    }
    const VRange<K> wr = vwin_load<K>(keys, n, win, T, t, mw);
    K tg[2][kVQ], k_at[2][kVQ];
    bool need[2][kVQ];
    int64_t r[2][kVQ];
#pragma unroll
    for (int q = 0; q < kVQ; ++q) {
        tg[0][q] = k[q] + d2;                                        // :122-123
        tg[1][q] = k[q] + d4;                                        // :125-126
        need[0][q] = need[1][q] = i0 + (uint64_t)(q * kMT) + threadIdx.x < i1;
    }
    vwin_search<K, kVQ>(keys, n, mw, wr, tg, need, r, k_at);
    // the corner values: absent corners read the voxel's own value
    V vc[3][kVQ];
    int64_t cn[3][kVQ];
#pragma unroll
    for (int q = 0; q < kVQ; ++q) {
        const int64_t i = (int64_t)min(i0 + (uint64_t)(q * kMT) + threadIdx.x, m - 1);
        cn[0][q] = kn[q] == k[q] + d1 ? i + 1 : -1;                  // :132
        cn[1][q] = k_at[0][q] == k[q] + d2 ? r[0][q] : -1;           // :129
        cn[2][q] = k_at[1][q] == k[q] + d4 ? r[1][q] : -1;           // :130
#pragma unroll
        for (int a = 0; a < 3; ++a)
            vc[a][q] = vals[cn[a][q] >= 0 ? cn[a][q] : i];
    }
    uint32_t cnt = 0;
#pragma unroll
    for (int q = 0; q < kVQ; ++q) {
        const uint64_t i = i0 + (uint64_t)(q * kMT) + threadIdx.x;
        uint32_t f = 0;
        if (vweight(v0[q]) != 0.0f) {                                // :119-120
            const float dself = vdist(v0[q]);
#pragma unroll
            for (int a = 0; a < 3; ++a)                              // :140-168
                if (cn[a][q] >= 0 && vweight(vc[a][q]) > 0.0f && dself * vdist(vc[a][q]) <= 0.0f)
                    f |= 1u << a;
        }
        if (i < i1) {
            vflags[i] = (uint8_t)f;
            cnt += __popc(f);
        }
    }
    cnt = block_reduce_sum<kMT>(cnt, red);
    if (threadIdx.x == 0)
        tile_rev[T - 1 - t] = cnt;
}

// Pass 2: the vertices of the flagged voxels of tile t.  Thread tid reads the
// flags of voxels i0 + 4 tid .. + 3; one block scan gives both each voxel's
// place in generation order (the vertices of the voxels after it, then its
// own in axis order x, y, z) and its slot in an LDS list of flagged voxels,
// which the workgroup then works through.
template <typename K, typename V>
__global__ __launch_bounds__(kMT) void k_mesh_vemit(const K *__restrict__ keys, const V *__restrict__ vals,
                                                    uint64_t n, const uint64_t *__restrict__ win, uint64_t T,
                                                    MeshParams mp, const uint8_t *__restrict__ vflags,
                                                    const uint32_t *__restrict__ tile_off_rev,
                                                    float *__restrict__ verts)
{
    static_assert(kVQ == 4, "four flag bytes per thread");
    __shared__ VWin<K> mw;
    __shared__ uint32_t tmp[kMT / kWave + 1];
    __shared__ uint16_t sh_ent[kMeshTile];    // voxel within the tile | flags << 10
    __shared__ uint32_t sh_off[kMeshTile];    // its first vertex
    constexpr K d2 = K(1) << KeyTraits<K>::kb, d4 = K(1) << (2 * KeyTraits<K>::kb);
    const uint64_t t = xcd_tile(T);
    if (t >= T)
        return;
    const uint64_t m = n - 1;
    const uint64_t i0 = t * kMeshTile, i1 = min(i0 + kMeshTile, m);
    const uint64_t my = i0 + 4 * (uint64_t)threadIdx.x;
    uint32_t f4 = my < i1 ? *(const uint32_t *)(vflags + my) : 0u;  // 4 bytes from a 4-aligned voxel
    if (my + 4 > i1 && my < i1)
        f4 &= 0xffffffffu >> (8 * (uint32_t)(my + 4 - i1));
    const uint32_t base = tile_off_rev[T - 1 - t];
    const VRange<K> wr = vwin_load<K>(keys, n, win, T, t, mw);
    uint32_t nf = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        nf += (f4 >> (8 * q) & 0xffu) != 0u;
    const uint32_t cnt = __popc(f4);
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan<kMT>(cnt << 16 | nf, tmp, tot);
    // vertices of the voxels after mine in this tile
    uint32_t o = base + (tot >> 16) - (ex >> 16) - cnt;
    uint32_t slot = ex & 0xffffu;
#pragma unroll
    for (int q = 3; q >= 0; --q) {
        const uint32_t f = f4 >> (8 * q) & 0xffu;
        if (f) {
            sh_ent[slot] = (uint16_t)((4 * threadIdx.x + q) | f << 10);
            sh_off[slot] = o;
            ++slot;
            o += __popc(f);
        }
    }
    __syncthreads();
    const uint32_t nlist = tot & 0xffffu;
    for (uint32_t e = threadIdx.x; e < nlist; e += kMT) {
        const uint32_t ent = sh_ent[e];
        const uint64_t i = i0 + (ent & 1023u);
        const uint32_t f = ent >> 10;
        const K k = keys[i];
        const V v0 = vals[i];
        K tg[2][1] = { { k + d2 }, { k + d4 } }, k_at[2][1];
        const bool need[2][1] = { { (f & 2u) != 0u }, { (f & 4u) != 0u } };
        int64_t r[2][1];
        vwin_search<K, 1>(keys, n, mw, wr, tg, need, r, k_at);
        // a set flag means the corner exists: i + 1, and the searched voxels
        const V vn[3] = { vals[i + 1], vals[r[0][0]], vals[r[1][0]] };
        uint32_t x0, y0, z0;
        unpack_key<K>(k, x0, y0, z0);
        // Volume::VoxelCenter(int, int, int), Volume.cpp:80-88
        const float vx = (float)(int)x0 * mp.vl + mp.c0, vy = (float)(int)y0 * mp.vl + mp.c0,
                    vz = (float)(int)z0 * mp.vl + mp.c0;
        const float dself = vdist(v0);
        uint32_t oo = sh_off[e];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (!(f >> a & 1u))
                continue;
            const float dn = vdist(vn[a]);
            const float tt = fabsf(dself) / (fabsf(dself) + fabsf(dn)) * mp.vl;   // :143/153/163
            verts[3 * (size_t)oo + 0] = a == 0 ? vx + tt : vx;
            verts[3 * (size_t)oo + 1] = a == 1 ? vy + tt : vy;
            verts[3 * (size_t)oo + 2] = a == 2 ? vz + tt : vz;
            ++oo;
        }
    }
}

// ---------------------------------------------------------------- Bin
__device__ __forceinline__ bool to_int_checked(float f, int &out)
{
    if (!(f > -2147483648.0f && f < 2147483648.0f))
        return false;
    out = (int)f;
    return true;
}

__device__ __forceinline__ uint32_t orderable(float z)
{
    if (z == 0.0f)
        z = 0.0f;                                  // -0 and +0 compare equal in the reference
    const uint32_t b = __float_as_uint(z);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void k_fill_u64(unsigned long long *__restrict__ p, uint64_t n, unsigned long long v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Renderer::Bin (Renderer.cpp:59-79): nearest z (NDC) wins, the earlier point on ties.
__global__ void k_bin_splat(const float *__restrict__ pts, uint64_t n, Mat4 w2c, int W, int H,
                            unsigned long long *__restrict__ zbuf)
{
    const float hw = (float)W * 0.5f, hh = (float)H * 0.5f;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        float p[4] = { pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], 1.0f };
        mat_vec(w2c, p, p);
        const float w = p[3];
        const float x = p[0] / w, y = p[1] / w, z = p[2] / w;     // math.h:414-419
        int u, yi;
        if (!to_int_checked(x * hw + hw, u) || !to_int_checked(y * hh + hh, yi))
            continue;
        const long v = (long)H - 1 - yi;
        if (u >= 0 && u < W && v >= 0 && v < H && z >= -1.0f && z <= 1.0f) {
            const unsigned long long key = (unsigned long long)orderable(z) << 32 | (uint32_t)i;
            atomicMin(&zbuf[(size_t)v * W + u], key);
        }
    }
}

__global__ void k_bin_resolve(const float *__restrict__ pts, const unsigned long long *__restrict__ zbuf,
                              int W, int H, float *__restrict__ out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= W * H)
        return;
    const unsigned long long k = zbuf[q];
    float x = 0.0f, y = 0.0f, z = 0.0f;
    if (k != ~0ull) {
        const uint32_t i = (uint32_t)k;
        x = pts[3 * (size_t)i];
        y = pts[3 * (size_t)i + 1];
        z = pts[3 * (size_t)i + 2];
    }
    out[3 * q] = x;
    out[3 * q + 1] = y;
    out[3 * q + 2] = z;
}

// =================================================================== host
static size_t a256(size_t x) { return (x + 255) & ~size_t(255); }

template <typename K, typename V>
static kifi_status mesh_impl(kifi_volume *vol, bool tris, float *verts, size_t vcap, uint32_t *idx,
                             size_t icap, size_t *nverts, size_t *nidx, int *stalled, bool async_copy)
{
    kifi_ctx *ctx = vol->ctx;
    hipStream_t st = ctx->stream;
    KIFI_HIP(ctx, hipStreamSynchronize(st));
    uint64_t n = 0;
    KIFI_HIP(ctx, hipMemcpy(&n, vol->d_size, 8, hipMemcpyDeviceToHost));
    if (stalled)
        *stalled = 0;
    if (n < 2) {
        *nverts = 0;
        if (nidx)
            *nidx = 0;
        return KIFI_OK;
    }
    const uint64_t m = n - 1;
    const uint64_t nt = (m + kScanTile - 1) / kScanTile + 1;
    const uint64_t T = (m + kMeshTile - 1) / kMeshTile;
    const int NO = tris ? 3 : 2;
    // vertices only: flags (read 4 at a time) and tile counts; with indices:
    // flags, per-voxel counts and their scans
    size_t bytes = a256(m + 4) + a256(16) + a256(16) + a256((T + 1) * NO * 8);
    bytes += tris ? 2 * a256(m * 4) + a256(nt * 4) : a256(T * 4);
    kifi_status s = ensure_scratch(ctx, bytes);
    if (s != KIFI_OK)
        return s;
    char *p = (char *)ctx->scratch;
    uint8_t *vflags = (uint8_t *)p; p += a256(m + 4);
    uint64_t *totals = (uint64_t *)p; p += a256(16);
    unsigned long long *stall = (unsigned long long *)p; p += a256(16);
    uint64_t *win = (uint64_t *)p; p += a256((T + 1) * NO * 8);
    uint32_t *cv_fwd = (uint32_t *)p, *ci_rev = cv_fwd, *tsum = cv_fwd, *tile_rev = cv_fwd;
    if (tris) {
        ci_rev = cv_fwd + a256(m * 4) / 4;
        tsum = ci_rev + a256(m * 4) / 4;
    }

    const K *keys = (const K *)vol->keys[vol->cur];
    const V *vals = (const V *)vol->vals[vol->cur];
    const unsigned grid = (unsigned)std::min<uint64_t>(T, (uint64_t)ctx->num_cus * 8);
    const unsigned xgrid = (unsigned)((T + kXcds - 1) / kXcds * kXcds);   // a whole run per XCD
    const unsigned wgrid = (unsigned)(((T + 1) * NO + kMT - 1) / kMT);
    const int t_mesh = timer_begin(ctx, KIFI_STAGE_MESH);
    if (tris) {
        hipLaunchKernelGGL((k_mesh_windows<K, true>), dim3(wgrid), dim3(kMT), 0, st, keys, n, T, win);
        hipLaunchKernelGGL((k_mesh_count<K, V, true>), dim3(grid), dim3(kMT), 0, st, keys, vals, n, win, T,
                           vflags, cv_fwd, ci_rev);
        device_scan(st, cv_fwd, cv_fwd, m, tsum, totals);
        device_scan(st, ci_rev, ci_rev, m, tsum, totals + 1);
    } else {
        hipLaunchKernelGGL((k_mesh_windows<K, false>), dim3(wgrid), dim3(kMT), 0, st, keys, n, T, win);
        hipLaunchKernelGGL((k_mesh_vcount<K, V>), dim3(xgrid), dim3(kMT), 0, st, keys, vals, n, win, T,
                           vflags, tile_rev);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, tile_rev, T, totals);
    }
    timer_end(ctx, KIFI_STAGE_MESH, t_mesh);
    uint64_t h_tot[2] = { 0, 0 };
    KIFI_HIP(ctx, hipMemcpyAsync(h_tot, totals, tris ? 16 : 8, hipMemcpyDeviceToHost, st));
    KIFI_HIP(ctx, hipStreamSynchronize(st));
    *nverts = h_tot[0];
    if (nidx)
        *nidx = tris ? h_tot[1] : 0;
    // per-voxel output offsets and the indices themselves are 32-bit
    if (h_tot[0] > 0xffffffffull || (tris && h_tot[1] > 0xffffffffull))
        return set_error(ctx, KIFI_ERR_CAPACITY, "mesh: more than 2^32 - 1 vertices or indices (32-bit offsets "
                                                 "and indices)");
    if (h_tot[0] > vcap || (tris && h_tot[1] > icap))
        return KIFI_ERR_CAPACITY;

    // output buffers on the device (after the counts), then copy to the host
    const size_t vb = a256(h_tot[0] * 12 + 4), ib = tris ? a256(h_tot[1] * 4 + 4) : 0,
                 rb = tris ? a256(h_tot[1] * 8 + 8) : 0;
    s = ensure_out(ctx, vb + ib + rb);
    if (s != KIFI_OK)
        return s;
    if (async_copy && !ctx->out_stream) {
        KIFI_HIP(ctx, hipStreamCreateWithFlags(&ctx->out_stream, hipStreamNonBlocking));
        KIFI_HIP(ctx, hipEventCreateWithFlags(&ctx->out_ready, hipEventDisableTiming));
        KIFI_HIP(ctx, hipEventCreateWithFlags(&ctx->out_copied, hipEventDisableTiming));
    }
    if (ctx->out_pending)          // the previous asynchronous copy still reads out_scratch
        KIFI_HIP(ctx, hipStreamWaitEvent(st, ctx->out_copied, 0));
    char *q = (char *)ctx->out_scratch;
    float *d_verts = (float *)q; q += vb;
    uint32_t *d_idx = (uint32_t *)q; q += ib;
    uint64_t *d_raw = (uint64_t *)q;
    MeshParams mp{ vol->consts.vl, vol->consts.vlHalfPlusMin };
    const int t_emit = timer_begin(ctx, KIFI_STAGE_MESH);
    if (tris) {
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, stall, (uint64_t)1, ~0ull);
        hipLaunchKernelGGL((k_mesh_emit<K, V, true>), dim3(grid), dim3(kMT), 0, st, keys, vals, n, win, T, mp,
                           vflags, cv_fwd, ci_rev, d_verts, d_idx, d_raw, stall);
        unsigned long long h_stall = ~0ull;
        KIFI_HIP(ctx, hipMemcpyAsync(&h_stall, stall, 8, hipMemcpyDeviceToHost, st));
        KIFI_HIP(ctx, hipStreamSynchronize(st));
        if (h_stall != ~0ull) {
            hipLaunchKernelGGL(k_mesh_stall_fix, dim3(256), dim3(256), 0, st, d_raw, (uint64_t)h_tot[1],
                               stall, d_idx);
            if (stalled)
                *stalled = 1;
        }
    } else {
        hipLaunchKernelGGL((k_mesh_vemit<K, V>), dim3(xgrid), dim3(kMT), 0, st, keys, vals, n, win, T, mp,
                           vflags, tile_rev, d_verts);
    }
    timer_end(ctx, KIFI_STAGE_MESH, t_emit);
    KIFI_HIP(ctx, hipGetLastError());
    if (async_copy) {
        // the copy runs on its own stream behind the emit; the frames queued
        // next on `stream` overlap it, and out_settle waits for its end
        if (h_tot[0]) {
            KIFI_HIP(ctx, hipEventRecord(ctx->out_ready, st));
            KIFI_HIP(ctx, hipStreamWaitEvent(ctx->out_stream, ctx->out_ready, 0));
            KIFI_HIP(ctx, hipMemcpyAsync(verts, d_verts, h_tot[0] * 12, hipMemcpyDeviceToHost,
                                         ctx->out_stream));
            KIFI_HIP(ctx, hipEventRecord(ctx->out_copied, ctx->out_stream));
            ctx->out_pending = true;
        }
        return KIFI_OK;
    }
    if (h_tot[0])
        KIFI_HIP(ctx, hipMemcpyAsync(verts, d_verts, h_tot[0] * 12, hipMemcpyDeviceToHost, st));
    if (tris && h_tot[1])
        KIFI_HIP(ctx, hipMemcpyAsync(idx, d_idx, h_tot[1] * 4, hipMemcpyDeviceToHost, st));
    KIFI_HIP(ctx, hipStreamSynchronize(st));
    return KIFI_OK;
}

kifi_status mesh_dispatch(kifi_volume *vol, bool tris, float *verts, size_t vcap, uint32_t *idx, size_t icap,
                          size_t *nverts, size_t *nidx, int *stalled, bool async_copy)
{
    auto run = [&](auto impl) { return impl(vol, tris, verts, vcap, idx, icap, nverts, nidx, stalled, async_copy); };
    if (vol->desc.key_bits == 32) {
        if (vol->desc.storage == KIFI_STORAGE_F32)
            return run(mesh_impl<uint32_t, VoxF32m>);
        return run(mesh_impl<uint32_t, uint32_t>);
    }
    if (vol->desc.storage == KIFI_STORAGE_F32)
        return run(mesh_impl<uint64_t, VoxF32m>);
    return run(mesh_impl<uint64_t, uint32_t>);
}

kifi_status bin_host(kifi_ctx *ctx, const float *points, size_t n, const float w2c[16], int W, int H,
                     float *out)
{
    hipStream_t st = ctx->stream;
    const size_t P = (size_t)W * H;
    const size_t bytes = a256(n * 12 + 4) + a256(P * 8) + a256(P * 12);
    kifi_status s = ensure_scratch(ctx, bytes);
    if (s != KIFI_OK)
        return s;
    char *p = (char *)ctx->scratch;
    float *d_pts = (float *)p; p += a256(n * 12 + 4);
    unsigned long long *zb = (unsigned long long *)p; p += a256(P * 8);
    float *d_out = (float *)p;
    if (n)
        KIFI_HIP(ctx, hipMemcpyAsync(d_pts, points, n * 12, hipMemcpyHostToDevice, st));
    Mat4 m;
    for (int i = 0; i < 16; ++i)
        m.m[i] = w2c[i];
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, zb, (uint64_t)P, ~0ull);
    if (n)
        hipLaunchKernelGGL(k_bin_splat, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0,
                           st, d_pts, (uint64_t)n, m, W, H, zb);
    hipLaunchKernelGGL(k_bin_resolve, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, d_pts, zb, W, H,
                       d_out);
    KIFI_HIP(ctx, hipGetLastError());
    KIFI_HIP(ctx, hipMemcpyAsync(out, d_out, P * 12, hipMemcpyDeviceToHost, st));
    KIFI_HIP(ctx, hipStreamSynchronize(st));
    return KIFI_OK;
}

} // namespace kifi
