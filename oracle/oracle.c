/*
 * oracle.c -- CPU restatement of the kifi fusion path (test infrastructure only).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * library, as the checker.  Reference citations are relative to KinectFission/kifi/.
 * Build: oracle/Makefile (gcc -O2 -msse2 -ffp-contract=off, never -ffast-math).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* =========================== util/math.h =========================== */

/* float4x4 * float4, math.h:362-365 (and the SSE twin :381-397):
 * (col0*x + col1*y) + (col2*z + col3*w), every product rounded separately. */
void orc_mat_vec(const float m[16], const float v[4], float out[4])
{
    float r[4];
    for (int i = 0; i < 4; ++i)
        r[i] = (m[i] * v[0] + m[i + 4] * v[1]) + (m[i + 8] * v[2] + m[i + 12] * v[3]);
    memcpy(out, r, sizeof r);
}

/* float4x4 * float4x4, math.h:367-377: result.colj = m * n.colj */
void orc_mat_mul(const float m[16], const float n[16], float out[16])
{
    float r[16];
    for (int c = 0; c < 4; ++c)
        orc_mat_vec(m, n + 4 * c, r + 4 * c);
    memcpy(out, r, sizeof r);
}

/* math.h:615-625 */
void orc_invert_transform(float m[16])
{
    float t[4] = { -m[12], -m[13], -m[14], 1.0f }; /* tInv = -col3; tInv.w = 1 */
    m[12] = 0.0f; m[13] = 0.0f; m[14] = 0.0f; m[15] = 1.0f;
    float s;
    s = m[1]; m[1] = m[4]; m[4] = s;  /* col0.y <-> col1.x */
    s = m[2]; m[2] = m[8]; m[8] = s;  /* col0.z <-> col2.x */
    s = m[6]; m[6] = m[9]; m[9] = s;  /* col1.z <-> col2.y */
    float c3[4];
    orc_mat_vec(m, t, c3);
    memcpy(m + 12, c3, sizeof c3);
}

/* DepthSensorParams.cpp:47-63; element (r,c) lives at m[r + 4c] (math.h:259-265) */
void orc_eye_to_clip_rh(const orc_camera *cam, float out[16])
{
    memset(out, 0, 16 * sizeof(float));
    float W = (float)cam->width, H = (float)cam->height;
    out[0 + 4 * 0] = 2.0f * cam->fx / W;
    out[0 + 4 * 2] = W - 2.0f * cam->cx;
    out[1 + 4 * 1] = 2.0f * cam->fy / H;
    out[1 + 4 * 2] = H - 2.0f * cam->cy;
    out[2 + 4 * 2] = -(cam->zfar + cam->znear) / (cam->zfar - cam->znear);
    out[2 + 4 * 3] = -2.0f * cam->zfar * cam->znear / (cam->zfar - cam->znear);
    out[3 + 4 * 2] = -1.0f;
}

/* Volume.cpp:12-37 */
void orc_volume_desc_init(orc_volume_desc *d, int res, float side, float trunc, int kb)
{
    d->res = res;
    d->kb = kb;
    d->side = side;
    d->trunc = trunc;
    float minx = -0.5f * side;                 /* Volume.cpp:46-49 Minimum() */
    d->vl = side / (float)res;
    d->vlHalfPlusMin = d->vl * 0.5f + minx;
    d->vlInv = 1.0f / d->vl;
    d->minOverNegVl = d->vlInv * -minx;
}

/* math.h:442-445 (OR of shifted fields, kb = 10 is the reference layout) */
uint64_t orc_pack(uint64_t x, uint64_t y, uint64_t z, int kb)
{
    return z << (2 * kb) | y << kb | x;
}

/* math.h:452-457 */
void orc_unpack(uint64_t k, int kb, uint32_t *x, uint32_t *y, uint32_t *z)
{
    uint64_t m = (1ull << kb) - 1;
    *x = (uint32_t)(k & m);
    *y = (uint32_t)(k >> kb & m);
    *z = (uint32_t)(k >> (2 * kb));
}

/* =========================== Voxel =========================== */

/* Voxel.h:53-57: the voxel keeps the sum of distances and the update count */
void orc_voxel_update_f32(float *sw, float d)
{
    sw[0] += d;
    sw[1]++;
}

/* Voxel.h:39-44 */
float orc_voxel_distance_f32(const float *sw)
{
    return sw[0] / sw[1];
}

/* IEEE binary32 -> binary16, round to nearest, ties to even (own implementation). */
uint16_t orc_f32_to_f16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xffu)
        return (uint16_t)(sign | 0x7c00u | (man ? (0x200u | (man >> 13)) : 0u));
    int e = (int)exp - 127 + 15;
    if (e >= 31)
        return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10)
            return (uint16_t)sign;
        man |= 0x800000u;
        int shift = 14 - e;
        uint32_t hm = man >> shift;
        uint32_t rem = man & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (hm & 1u)))
            hm++;
        return (uint16_t)(sign | hm);
    }
    uint32_t h = sign | ((uint32_t)e << 10) | (man >> 13);
    uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u)))
        h++;
    return (uint16_t)h;
}

float orc_f16_to_f32(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1fu;
    uint32_t man = h & 0x3ffu;
    uint32_t x;
    if (exp == 0) {
        if (man == 0) {
            x = sign;
        } else { /* subnormal: renormalise */
            int e = -1;
            do { man <<= 1; ++e; } while (!(man & 0x400u));
            x = sign | ((uint32_t)(127 - 15 - e) << 23) | ((man & 0x3ffu) << 13);
        }
    } else if (exp == 31) {
        x = sign | 0x7f800000u | (man << 13);
    } else {
        x = sign | ((exp + 127 - 15) << 23) | (man << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

/* FP16 distance + uint8 weight storage (no reference counterpart; the unbuilt
 * cuda/Voxel.h:38-64 packs unorm16 + 8-bit weight with a running average).
 * Update in binary32: d' = (d*w + s) * R[w] with R[w] = 1/(w + 1) correctly
 * rounded to binary32 (a 256-entry table; one rounding more than a division,
 * far below the binary16 rounding of the stored distance); the weight
 * saturates at 255, where R[255] = 1/256 exactly. */
float orc_f16u8_recip(uint32_t w)
{
    return 1.0f / (float)(w + 1u);
}

void orc_voxel_update_f16u8(uint32_t *p, float s)
{
    uint32_t w = (*p >> 16) & 0xffu;
    float d = orc_f16_to_f32((uint16_t)(*p & 0xffffu));
    float wf = (float)w;
    float dn = (d * wf + s) * orc_f16u8_recip(w);
    uint32_t wn = w < 255u ? w + 1u : 255u;
    *p = (uint32_t)orc_f32_to_f16(dn) | (wn << 16);
}

float orc_voxel_distance_f16u8(uint32_t p)
{
    return orc_f16_to_f32((uint16_t)(p & 0xffffu));
}

static size_t value_bytes(int storage)
{
    return storage == ORC_STORE_F32 ? 8 : (storage == ORC_STORE_F64 ? 16 : 4);
}

static float vox_weight(const void *values, size_t i, int storage)
{
    if (storage == ORC_STORE_F32)
        return ((const float *)values)[2 * i + 1];
    if (storage == ORC_STORE_F64)
        return (float)((const double *)values)[2 * i + 1];
    return (float)((((const uint32_t *)values)[i] >> 16) & 0xffu);
}

static float vox_distance(const void *values, size_t i, int storage)
{
    if (storage == ORC_STORE_F32)
        return orc_voxel_distance_f32((const float *)values + 2 * i);
    if (storage == ORC_STORE_F64) {
        const double *sw = (const double *)values + 2 * i;
        return (float)(sw[0] / sw[1]);
    }
    return orc_voxel_distance_f16u8(((const uint32_t *)values)[i]);
}

static void vox_update(void *values, size_t i, int storage, float d)
{
    if (storage == ORC_STORE_F32)
        orc_voxel_update_f32((float *)values + 2 * i, d);
    else
        orc_voxel_update_f16u8((uint32_t *)values + i, d);
}

/* =========================== Integrator =========================== */

/* Integrator::DepthMap2PointCloud, Integrator.cpp:76-182.  Each SSE lane is a
 * separate binary32 computation; the lane-wise ops are spelled out here. */
size_t orc_keygen(const float *depth, const orc_camera *cam, const float e2w[16],
                  const orc_volume_desc *vol, uint64_t *out)
{
    const int W = cam->width, H = cam->height;
    const float flInvX = 1.0f / cam->fx, flInvY = 1.0f / cam->fy;
    const float ppX = (0.5f - cam->cx) * flInvX;       /* :95-101 */
    const float ppY = (cam->cy - 0.5f) * flInvY;
    const float ppZ = 0.0f * 1.0f, ppW = 0.0f * 1.0f;
    const float maxIndex = (float)(vol->res - 1);
    size_t n = 0;

    for (int v = 0; v < H; ++v) {
        const float pointY = -(float)v;                 /* :111 point = (0, -v, -1, 0) */
        for (int u = 0; u < W; ++u) {
            float d = depth[(size_t)v * W + u];
            if (!(d > 0.0f))                            /* :117-119, :151-154 lane masks */
                continue;
            /* loadss(u) + point */
            float p[4] = { (float)u + 0.0f, 0.0f + pointY, 0.0f + -1.0f, 0.0f + 0.0f };
            /* point * flInv + ppOverFl  (:126-129) */
            p[0] = p[0] * flInvX + ppX;
            p[1] = p[1] * flInvY + ppY;
            p[2] = p[2] * 1.0f + ppZ;
            p[3] = p[3] * 1.0f + ppW;
            /* point * depth + (0,0,0,1)  (:136-139) */
            p[0] = p[0] * d + 0.0f;
            p[1] = p[1] * d + 0.0f;
            p[2] = p[2] * d + 0.0f;
            p[3] = p[3] * d + 1.0f;
            orc_mat_vec(e2w, p, p);                     /* :141-144 */
            /* Volume::VoxelIndex(p) - half  (:146-149, Volume.cpp:90-93) */
            float ix = p[0] * vol->vlInv + vol->minOverNegVl - 0.5f;
            float iy = p[1] * vol->vlInv + vol->minOverNegVl - 0.5f;
            float iz = p[2] * vol->vlInv + vol->minOverNegVl - 0.5f;
            float iw = p[3] * 1.0f + 0.0f - 0.5f;
            int ok = ix >= 0.0f && ix < maxIndex && iy >= 0.0f && iy < maxIndex &&
                     iz >= 0.0f && iz < maxIndex && iw >= 0.0f && iw < maxIndex;
            if (ok)
                out[n++] = orc_pack((uint64_t)(uint32_t)ix, (uint64_t)(uint32_t)iy,
                                    (uint64_t)(uint32_t)iz, vol->kb);
        }
    }
    return n;
}

/* util::radix_sort (algorithm.h:63-91): LSD, 8-bit digits, counting + exclusive
 * scan + stable scatter, ping-pong.  key_bits = 32 reproduces the reference. */
void orc_radix_sort(uint64_t *keys, size_t n, int key_bits, uint64_t *scratch)
{
    uint64_t *first = keys, *tmp = scratch;
    size_t cnt[256];
    int passes = key_bits / 8;
    for (int pass = 0; pass < passes; ++pass) {
        int shift = 8 * pass;
        memset(cnt, 0, sizeof cnt);
        for (size_t i = 0; i < n; ++i)
            ++cnt[(first[i] >> shift) & 0xff];
        size_t acc = 0;                                   /* numeric.h:90-108 */
        for (int b = 0; b < 256; ++b) {
            size_t t = cnt[b];
            cnt[b] = acc;
            acc += t;
        }
        for (size_t i = 0; i < n; ++i)
            tmp[cnt[(first[i] >> shift) & 0xff]++] = first[i];
        uint64_t *s = first; first = tmp; tmp = s;
    }
    if (first != keys)
        memcpy(keys, first, n * sizeof *keys);
}

/* std::unique, Integrator.cpp:50-55 */
size_t orc_unique(uint64_t *k, size_t n)
{
    if (n == 0)
        return 0;
    size_t d = 0;
    for (size_t i = 1; i < n; ++i)
        if (k[i] != k[d])
            k[++d] = k[i];
    return d + 1;
}

/* std::set_union of a sorted unique range with itself shifted by delta */
static size_t union_shift(const uint64_t *a, size_t n, uint64_t delta, uint64_t *out)
{
    size_t i = 0, j = 0, o = 0;
    while (i < n && j < n) {
        uint64_t x = a[i], y = a[j] + delta;
        if (x < y) { out[o++] = x; ++i; }
        else if (y < x) { out[o++] = y; ++j; }
        else { out[o++] = x; ++i; ++j; }
    }
    while (i < n) out[o++] = a[i++];
    while (j < n) out[o++] = a[j++] + delta;
    return o;
}

/* Integrator::ExpandChunks / ExpandChunksHelper, Integrator.cpp:186-250.
 * z and y passes: set_union(S, S+delta) (:211-225); x pass: the linear loop of
 * :227-248, which never emits the last element's +1 neighbour. */
size_t orc_expand(uint64_t *keys, size_t n, int kb, uint64_t *scratch, size_t cap)
{
    if (n == 0)
        return 0;
    if (cap < 8 * n)
        return (size_t)-1;
    uint64_t dz = orc_pack(0, 0, 1, kb), dy = orc_pack(0, 1, 0, kb);
    n = union_shift(keys, n, dz, scratch);
    memcpy(keys, scratch, n * sizeof *keys);
    n = union_shift(keys, n, dy, scratch);
    memcpy(keys, scratch, n * sizeof *keys);
    /* x pass */
    size_t dst = 0;
    uint64_t prev = keys[0];
    for (size_t i = 0; i < n; ++i) {
        uint64_t x = keys[i];
        scratch[dst] = prev;
        dst += (x != prev);
        scratch[dst++] = x;
        prev = x + 1;
    }
    memcpy(keys, scratch, dst * sizeof *keys);
    return dst;
}

/* util::set_union with values, algorithm.h:138-198 (the first range wins ties).
 * The second range's values are all the same default value (make_const_iterator,
 * Integrator.cpp:60-63).  Returns the output size. */
size_t orc_set_union(const uint64_t *k1, size_t n1, const uint64_t *k2, size_t n2,
                     const void *v1, const void *vdef, size_t vb, uint64_t *ko, void *vo)
{
    const char *a = (const char *)v1;
    char *o = (char *)vo;
    size_t i = 0, j = 0, r = 0;
    while (i < n1 && j < n2) {
        if (k1[i] < k2[j]) {
            ko[r] = k1[i]; memcpy(o + r * vb, a + i * vb, vb); ++r; ++i;
        } else if (k2[j] < k1[i]) {
            ko[r] = k2[j]; memcpy(o + r * vb, vdef, vb); ++r; ++j;
        } else {
            ko[r] = k1[i]; memcpy(o + r * vb, a + i * vb, vb); ++r; ++i; ++j;
        }
    }
    while (i < n1) { ko[r] = k1[i]; memcpy(o + r * vb, a + i * vb, vb); ++r; ++i; }
    while (j < n2) { ko[r] = k2[j]; memcpy(o + r * vb, vdef, vb); ++r; ++j; }
    return r;
}

/* Integrator::UpdateVoxels, Integrator.cpp:253-358 (covers floor(n/4)*4 voxels). */
void orc_update(const uint64_t *keys, void *values, size_t n, int storage, const float *depth,
                const orc_camera *cam, const float w2c[16], const orc_volume_desc *vol)
{
    const int W = cam->width, H = cam->height;
    const float ndcU = (float)W * 0.5f, ndcV = (float)H * 0.5f;   /* :262 */
    const float maxU = (float)W, maxV = (float)H;                  /* :265 */
    const float trunc = vol->trunc;
    size_t end = n / 4 * 4;                                        /* :269 */
    for (size_t i = 0; i < end; ++i) {
        uint32_t x, y, z;
        orc_unpack(keys[i], vol->kb, &x, &y, &z);
        /* Volume::VoxelCenter(vector), Volume.cpp:75-78 */
        float k[4] = { (float)x * vol->vl + vol->vlHalfPlusMin,
                       (float)y * vol->vl + vol->vlHalfPlusMin,
                       (float)z * vol->vl + vol->vlHalfPlusMin,
                       1.0f * 1.0f + 0.0f };
        orc_mat_vec(w2c, k, k);                                    /* :290-293 */
        float kw = k[3];
        float dist = kw;                                           /* :300-303 */
        float u = k[0] / kw, v = k[1] / kw, zz = k[2] / kw, ww = k[3] / kw; /* :305-308 */
        u = u * ndcU + ndcU;                                       /* :310-313 */
        v = v * ndcV + ndcV;
        zz = zz * 0.0f + 0.0f;
        ww = ww * 0.0f + 0.0f;
        if (u >= 0.0f && u < maxU && v >= 0.0f && v < maxV && zz >= -1.0f && zz < 1.0f &&
            ww >= -FLT_MAX && ww < FLT_MAX) {                       /* :315 */
            float d = depth[(size_t)(H - (unsigned)v - 1) * W + (unsigned)u]; /* :319 */
            float sd = d - dist;
            if (d > 0.0f && sd >= -trunc)                          /* :322-323 */
                vox_update(values, i, storage, trunc < sd ? trunc : sd);
        }
    }
}

/* FP64 twin of the update (same voxels, same tests, double arithmetic), used to
 * report the error of the binary32/binary16 paths. sum_weight holds {sum, weight}. */
void orc_update_f64(const uint64_t *keys, double *sw, size_t n, const float *depth,
                    const orc_camera *cam, const double w2c[16], const orc_volume_desc *vol)
{
    const int W = cam->width, H = cam->height;
    const double vl = (double)vol->side / vol->res;
    const double c0 = vl * 0.5 - 0.5 * (double)vol->side;
    size_t end = n / 4 * 4;
    for (size_t i = 0; i < end; ++i) {
        uint32_t x, y, z;
        orc_unpack(keys[i], vol->kb, &x, &y, &z);
        double p[4] = { x * vl + c0, y * vl + c0, z * vl + c0, 1.0 }, k[4];
        for (int r = 0; r < 4; ++r)
            k[r] = w2c[r] * p[0] + w2c[r + 4] * p[1] + w2c[r + 8] * p[2] + w2c[r + 12] * p[3];
        double kw = k[3];
        double u = k[0] / kw * (W * 0.5) + W * 0.5, v = k[1] / kw * (H * 0.5) + H * 0.5;
        double zz = k[2] / kw;
        if (!(kw != 0.0) || !isfinite(zz) || !(u >= 0 && u < W && v >= 0 && v < H))
            continue;
        double d = depth[(size_t)(H - (unsigned)v - 1) * W + (unsigned)u];
        double sd = d - kw;
        if (d > 0.0 && sd >= -(double)vol->trunc) {
            sw[2 * i] += sd < vol->trunc ? sd : vol->trunc;
            sw[2 * i + 1] += 1.0;
        }
    }
}

/* ------------------------- the volume container ------------------------- */

orc_volume *orc_volume_create(int res, float side, float trunc, int kb, int storage)
{
    orc_volume *v = (orc_volume *)calloc(1, sizeof *v);
    if (!v)
        return NULL;
    orc_volume_desc_init(&v->desc, res, side, trunc, kb);
    v->storage = storage;
    return v;
}

void orc_volume_destroy(orc_volume *v)
{
    if (!v)
        return;
    free(v->keys);
    free(v->values);
    free(v);
}

void orc_volume_clear(orc_volume *v) { v->size = 0; }
size_t orc_volume_size(const orc_volume *v) { return v->size; }

static int volume_reserve(orc_volume *v, size_t n)
{
    if (n <= v->cap)
        return 0;
    size_t c = v->cap ? v->cap : 1024;
    while (c < n)
        c *= 2;
    uint64_t *k = (uint64_t *)realloc(v->keys, c * sizeof *k);
    if (!k)
        return -1;
    v->keys = k;
    void *val = realloc(v->values, c * value_bytes(v->storage));
    if (!val)
        return -1;
    v->values = val;
    v->cap = c;
    return 0;
}

size_t orc_volume_read(const orc_volume *v, uint64_t *keys, void *values, size_t cap)
{
    size_t n = v->size < cap ? v->size : cap;
    if (keys)
        memcpy(keys, v->keys, n * sizeof *keys);
    if (values)
        memcpy(values, v->values, n * value_bytes(v->storage));
    return v->size;
}

int orc_volume_write(orc_volume *v, const uint64_t *keys, const void *values, size_t n)
{
    if (volume_reserve(v, n))
        return -1;
    memcpy(v->keys, keys, n * sizeof *keys);
    memcpy(v->values, values, n * value_bytes(v->storage));
    v->size = n;
    return 0;
}

/* Integrator::Integrate, Integrator.cpp:26-71, with the frame given in mm and
 * converted as DepthStream.cpp:89-90. */
static size_t integrate_m(orc_volume *v, const float *depth, const orc_camera *cam, const float worldToEye[16],
                          size_t stats[4]);

size_t orc_integrate(orc_volume *v, const uint16_t *mm, const orc_camera *cam,
                     const float worldToEye[16], size_t stats[4])
{
    const size_t P = (size_t)cam->width * cam->height;
    float *depth = (float *)malloc((P ? P : 1) * sizeof(float));
    if (!depth)
        return (size_t)-1;
    for (size_t i = 0; i < P; ++i)
        depth[i] = (float)mm[i] * 0.001f;
    size_t r = integrate_m(v, depth, cam, worldToEye, stats);
    free(depth);
    return r;
}

/* Integrate of a vector2d<float> frame in metres, used as it is (DepthStream's
 * FLOAT texels, DepthStream.cpp:78-80). */
size_t orc_integrate_f32(orc_volume *v, const float *depth_m, const orc_camera *cam,
                         const float worldToEye[16], size_t stats[4])
{
    return integrate_m(v, depth_m, cam, worldToEye, stats);
}

static size_t integrate_m(orc_volume *v, const float *depth, const orc_camera *cam, const float worldToEye[16],
                          size_t stats[4])
{
    const size_t P = (size_t)cam->width * cam->height;
    if (cam->width % 4 != 0 || v->desc.res < 2)          /* :89-90 asserts */
        return (size_t)-1;
    uint64_t *keys = (uint64_t *)malloc(8 * P * sizeof(uint64_t));
    uint64_t *scr = (uint64_t *)malloc(8 * P * sizeof(uint64_t));
    if (!keys || !scr) {
        free(keys); free(scr);
        return (size_t)-1;
    }

    float clip[16], w2c[16], e2w[16];
    orc_eye_to_clip_rh(cam, clip);
    orc_mat_mul(clip, worldToEye, w2c);                  /* :41 */
    memcpy(e2w, worldToEye, sizeof e2w);
    orc_invert_transform(e2w);                           /* :42 */

    size_t K = orc_keygen(depth, cam, e2w, &v->desc, keys);          /* :46 */
    orc_radix_sort(keys, K, v->desc.kb == 10 ? 32 : 64, scr);         /* :48 */
    size_t U = orc_unique(keys, K);                                   /* :50-55 */
    size_t E = orc_expand(keys, U, v->desc.kb, scr, 8 * P);           /* :57 */

    /* flat_map::insert (flat_map.h:164-197): union into fresh arrays, then swap */
    size_t vb = value_bytes(v->storage);
    size_t cap = v->size + E;
    uint64_t *nk = (uint64_t *)malloc((cap ? cap : 1) * sizeof *nk);
    void *nv = malloc((cap ? cap : 1) * vb);
    uint8_t vdef[16] = { 0 };                             /* Voxel() = {0, 0} */
    size_t V1 = orc_set_union(v->keys, v->size, keys, E, v->values, vdef, vb, nk, nv);
    free(v->keys);
    free(v->values);
    v->keys = nk;
    v->values = nv;
    v->size = V1;
    v->cap = cap;

    if (v->storage == ORC_STORE_F64) {
        /* the full-FP64 run: same keys, projection and update in double from
         * the same binary32 inputs (intrinsics, worldToEye, depth) */
        double clip[16] = { 0 }, w2c64[16];
        const double W = cam->width, H = cam->height;
        clip[0] = 2.0 * cam->fx / W;
        clip[8] = W - 2.0 * cam->cx;
        clip[5] = 2.0 * cam->fy / H;
        clip[9] = H - 2.0 * cam->cy;
        clip[10] = -((double)cam->zfar + cam->znear) / ((double)cam->zfar - cam->znear);
        clip[14] = -2.0 * cam->zfar * (double)cam->znear / ((double)cam->zfar - cam->znear);
        clip[11] = -1.0;
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) {
                double acc = 0.0;
                for (int k = 0; k < 4; ++k)
                    acc += clip[r + 4 * k] * (double)worldToEye[k + 4 * c];
                w2c64[r + 4 * c] = acc;
            }
        orc_update_f64(v->keys, (double *)v->values, v->size, depth, cam, w2c64, &v->desc);
    } else {
        orc_update(v->keys, v->values, v->size, v->storage, depth, cam, w2c, &v->desc); /* :67 */
    }
    if (stats) {
        stats[0] = K; stats[1] = U; stats[2] = E; stats[3] = V1;
    }
    free(keys); free(scr);
    return E;
}

/* =========================== Mesher =========================== */

/* kTriOffsets[257] (Mesher.cpp:239-261) and kTriTable[2460] (Mesher.cpp:264-475), data
 * generated by oracle/gen_mc_tables.py and checked by sha256 fingerprint */
#include "mc_tables.inc"

const int *orc_tri_offsets(void) { return kTriOffsets; }
const int *orc_tri_table(void) { return kTriTable; }

typedef struct {
    uint64_t key;
    uint32_t axis;
} vid_t; /* vertex ID 3*key + axis of Mesher.cpp:147/157/167, held as a pair */

static int vid_lt(vid_t a, vid_t b)
{
    return a.key < b.key || (a.key == b.key && a.axis < b.axis);
}

static int vid_eq(vid_t a, vid_t b) { return a.key == b.key && a.axis == b.axis; }

/* Mesher::Generate<bool>, Mesher.cpp:85-229.  vids/iids may be NULL when
 * triangles are not requested. */
static size_t generate(const orc_volume *vol, int tris, float *verts, size_t vcap, vid_t *vids,
                       vid_t *iids, size_t icap, size_t *n_iid)
{
    const uint64_t *keys = vol->keys;
    const void *val = vol->values;
    const int st = vol->storage, kb = vol->desc.kb;
    const long n = (long)vol->size;
    const float vl = vol->desc.vl, c0 = vol->desc.vlHalfPlusMin;
    size_t nv = 0, ni = 0;
    long iTop = n - 2, iFront = n - 2, iFrontTop = n - 2;              /* :98-100 */
    const uint64_t d1 = orc_pack(1, 0, 0, kb), d2 = orc_pack(0, 1, 0, kb),
                   d3 = orc_pack(1, 1, 0, kb), d4 = orc_pack(0, 0, 1, kb),
                   d5 = orc_pack(1, 0, 1, kb), d6 = orc_pack(0, 1, 1, kb),
                   d7 = orc_pack(1, 1, 1, kb);                          /* :104-114 */

    for (long i = n - 2; i >= 0; --i) {                                  /* :117 */
        if (0.0f == vox_weight(val, (size_t)i, st))                      /* :119-120 */
            continue;
        uint64_t k = keys[i];
        while (keys[iTop] > k + d2) --iTop;                              /* :122-126 */
        while (keys[iFront] > k + d4) --iFront;

        /* corner c present (weight as stored) or the default Voxel() */
        long ci[8];
        ci[0] = i;
        ci[2] = keys[iTop] == k + d2 ? iTop : -1;                        /* :128-132 */
        ci[4] = keys[iFront] == k + d4 ? iFront : -1;
        ci[1] = keys[i + 1] == k + d1 ? i + 1 : -1;
#define W_(c) (ci[c] < 0 ? 0.0f : vox_weight(val, (size_t)ci[c], st))
#define D_(c) (vox_distance(val, (size_t)ci[c], st))
        uint32_t x0, y0, z0;
        orc_unpack(k, kb, &x0, &y0, &z0);
        /* Volume::VoxelCenter(int,int,int), Volume.cpp:80-88 */
        float vx = (float)(int)x0 * vl + c0, vy = (float)(int)y0 * vl + c0,
              vz = (float)(int)z0 * vl + c0;
        float dself = D_(0);
        if (W_(1) > 0.0f && dself * D_(1) <= 0.0f) {                      /* :140-148 */
            if (nv < vcap) {
                verts[3 * nv] = vx + fabsf(dself) / (fabsf(dself) + fabsf(D_(1))) * vl;
                verts[3 * nv + 1] = vy;
                verts[3 * nv + 2] = vz;
                if (tris) { vids[nv].key = k; vids[nv].axis = 0; }
            }
            ++nv;
        }
        if (W_(2) > 0.0f && dself * D_(2) <= 0.0f) {                      /* :150-158 */
            if (nv < vcap) {
                verts[3 * nv] = vx;
                verts[3 * nv + 1] = vy + fabsf(dself) / (fabsf(dself) + fabsf(D_(2))) * vl;
                verts[3 * nv + 2] = vz;
                if (tris) { vids[nv].key = k; vids[nv].axis = 1; }
            }
            ++nv;
        }
        if (W_(4) > 0.0f && dself * D_(4) <= 0.0f) {                      /* :160-168 */
            if (nv < vcap) {
                verts[3 * nv] = vx;
                verts[3 * nv + 1] = vy;
                verts[3 * nv + 2] = vz + fabsf(dself) / (fabsf(dself) + fabsf(D_(4))) * vl;
                if (tris) { vids[nv].key = k; vids[nv].axis = 2; }
            }
            ++nv;
        }
        if (!tris)
            continue;
        while (keys[iFrontTop] > k + d6) --iFrontTop;                    /* :174-175 */
        ci[6] = keys[iFrontTop] == k + d6 ? iFrontTop : -1;              /* :177 */
        ci[3] = keys[iTop + 1] == k + d3 ? iTop + 1 : -1;                /* :179-181 */
        ci[5] = keys[iFront + 1] == k + d5 ? iFront + 1 : -1;
        ci[7] = keys[iFrontTop + 1] == k + d7 ? iFrontTop + 1 : -1;
        int skip = 0;
        for (int c = 1; c < 8; ++c)                                      /* :183-188 */
            skip |= (0.0f == W_(c));
        if (skip)
            continue;
        int lut = (D_(0) < 0.0f) << 2 | (D_(1) < 0.0f) << 3 | (D_(2) < 0.0f) << 6 |
                  (D_(3) < 0.0f) << 7 | (D_(4) < 0.0f) << 1 | (D_(5) < 0.0f) << 0 |
                  (D_(6) < 0.0f) << 5 | (D_(7) < 0.0f) << 4;              /* :191-199 */
        uint64_t x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;                 /* :201-204 */
        vid_t l2g[12] = {                                                /* :206-218 */
            { orc_pack(x0, y0, z1, kb), 0 }, { orc_pack(x0, y0, z0, kb), 2 },
            { orc_pack(x0, y0, z0, kb), 0 }, { orc_pack(x1, y0, z0, kb), 2 },
            { orc_pack(x0, y1, z1, kb), 0 }, { orc_pack(x0, y1, z0, kb), 2 },
            { orc_pack(x0, y1, z0, kb), 0 }, { orc_pack(x1, y1, z0, kb), 2 },
            { orc_pack(x1, y0, z1, kb), 1 }, { orc_pack(x0, y0, z1, kb), 1 },
            { orc_pack(x0, y0, z0, kb), 1 }, { orc_pack(x1, y0, z0, kb), 1 },
        };
        for (int t = kTriOffsets[lut]; t < kTriOffsets[lut + 1]; ++t) {  /* :220-226 */
            if (ni < icap)
                iids[ni] = l2g[kTriTable[t]];
            ++ni;
        }
#undef W_
#undef D_
    }
    if (n_iid)
        *n_iid = ni;
    return nv;
}

size_t orc_mesh(const orc_volume *v, float *verts, size_t cap)
{
    return generate(v, 0, verts, cap, NULL, NULL, 0, NULL);
}

/* stable merge sort of (id, payload) by id */
typedef struct {
    vid_t id;
    uint32_t pay;
} vid_pair;

static void msort(vid_pair *a, vid_pair *t, size_t n)
{
    if (n < 2)
        return;
    size_t h = n / 2;
    msort(a, t, h);
    msort(a + h, t, n - h);
    size_t i = 0, j = h, o = 0;
    while (i < h && j < n)
        t[o++] = vid_lt(a[j].id, a[i].id) ? a[j++] : a[i++];
    while (i < h) t[o++] = a[i++];
    while (j < n) t[o++] = a[j++];
    memcpy(a, t, n * sizeof *a);
}

/* Mesher::Mesh with indices, Mesher.cpp:19-45: sort vertex IDs with their
 * vertices, sort index IDs with their positions (both stable, as the LSD radix
 * sort of algorithm.h:93-134), merge-join ID -> vertex ordinal, scatter back. */
size_t orc_mesh_indexed(const orc_volume *v, float *verts, size_t vcap, uint32_t *idx,
                        size_t icap, size_t *n_idx, int *stalled)
{
    size_t n = v->size;
    vid_t *vids = (vid_t *)malloc((3 * n + 1) * sizeof *vids);
    vid_t *iids = (vid_t *)malloc((15 * n + 1) * sizeof *iids);
    float *gv = (float *)malloc((9 * n + 3) * sizeof(float));
    size_t ni = 0;
    size_t nv = generate(v, 1, gv, 3 * n, vids, iids, 15 * n, &ni);

    vid_pair *pv = (vid_pair *)malloc((nv + 1) * sizeof *pv);
    vid_pair *pi = (vid_pair *)malloc((ni + 1) * sizeof *pi);
    vid_pair *tmp = (vid_pair *)malloc(((nv > ni ? nv : ni) + 1) * sizeof *tmp);
    for (size_t i = 0; i < nv; ++i) { pv[i].id = vids[i]; pv[i].pay = (uint32_t)i; }
    for (size_t i = 0; i < ni; ++i) { pi[i].id = iids[i]; pi[i].pay = (uint32_t)i; }
    msort(pv, tmp, nv);                                              /* :25-26 */
    msort(pi, tmp, ni);                                              /* :28-33 */

    for (size_t i = 0; i < nv && i < vcap; ++i)
        memcpy(verts + 3 * i, gv + 3 * (size_t)pv[i].pay, 3 * sizeof(float));

    /* :35-38; an index ID with no vertex ID stalls the join (all later stay raw) */
    uint32_t *mapped = (uint32_t *)malloc((ni + 1) * sizeof *mapped);
    for (size_t j = 0; j < ni; ++j)   /* raw ID value as the reference keeps it (32-bit) */
        mapped[j] = (uint32_t)(3 * pi[j].id.key + pi[j].id.axis);
    size_t j = 0;
    for (size_t i = 0; i < nv; ++i)
        while (j < ni && vid_eq(pi[j].id, pv[i].id))
            mapped[j++] = (uint32_t)i;
    if (stalled)
        *stalled = j < ni;
    /* :40-44 scatter back to generation order */
    for (size_t q = 0; q < ni; ++q)
        if (pi[q].pay < icap)
            idx[pi[q].pay] = mapped[q];
    if (n_idx)
        *n_idx = ni;
    free(mapped); free(pv); free(pi); free(tmp); free(gv); free(vids); free(iids);
    return nv;
}

/* =========================== Renderer::Bin =========================== */

/* (int) conversion as the x86 reference does it for in-range values; values
 * outside int range (or NaN) can never land inside the image, return INT_MIN. */
static int to_int(float f)
{
    if (!(f > -2147483648.0f && f < 2147483648.0f))
        return (-2147483647 - 1);
    return (int)f;
}

/* Renderer.cpp:39-80 */
void orc_bin(const float *pts, size_t n, const float w2c[16], int W, int H, float *out,
             float *zbuf)
{
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        zbuf[i] = FLT_MAX;
        out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = 0.0f;
    }
    const float hw = (float)W * 0.5f, hh = (float)H * 0.5f;
    for (size_t i = 0; i < n; ++i) {
        float p[4] = { pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], 1.0f };
        orc_mat_vec(w2c, p, p);
        float w = p[3];
        float x = p[0] / w, y = p[1] / w, z = p[2] / w;       /* homogenize, math.h:414-419 */
        int u = to_int(x * hw + hw);
        int yi = to_int(y * hh + hh);
        if (yi == (-2147483647 - 1))
            continue;
        long v = (long)H - 1 - yi;
        if (u >= 0 && u < W && v >= 0 && v < H && z >= -1.0f && z <= 1.0f) {
            size_t o = (size_t)v * W + u;
            if (z < zbuf[o]) {
                zbuf[o] = z;
                memcpy(out + 3 * o, pts + 3 * i, 3 * sizeof(float));
            }
        }
    }
}

/* =========================== ICP (Horn) =========================== */

/* util::kahan_sum (numeric.h:52-74): y = x - c; t = s + y; c = (t - s) - y; s = t */
float orc_kahan_sum_f32(const float *x, size_t n)
{
    float s = 0.0f, c = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        float y = x[i] - c;
        float t = s + y;
        c = t - s - y;
        s = t;
    }
    return s;
}

/* numeric.h:78-108 */
void orc_partial_sum_exclusive_u32(const uint32_t *in, size_t n, uint32_t *out)
{
    uint32_t acc = 0;
    for (size_t i = 0; i < n; ++i) {
        uint32_t t = in[i];
        out[i] = acc;
        acc += t;
    }
}

/* Cyclic Jacobi eigen decomposition of a symmetric 4x4 matrix in FP64 (Jacobi
 * 1846; Golub & Van Loan, Matrix Computations, sec. 8.5).  Eigenvectors are the
 * columns of evecs (column storage). */
void orc_jacobi4(const double in[16], double ev[4], double V[16])
{
    double A[4][4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            A[r][c] = 0.5 * (in[r + 4 * c] + in[c + 4 * r]);
            V[r + 4 * c] = r == c ? 1.0 : 0.0;
        }
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
                double apq = A[p][q];
                if (apq == 0.0)
                    continue;
                double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 4; ++k) { /* A <- A J */
                    double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 4; ++k) { /* A <- J^T A */
                    double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 4; ++k) { /* V <- V J */
                    double vkp = V[k + 4 * p], vkq = V[k + 4 * q];
                    V[k + 4 * p] = c * vkp - s * vkq;
                    V[k + 4 * q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < 4; ++i)
        ev[i] = A[i][i];
}

typedef struct { float s[4], c[4]; } kahan4;

static void kahan4_add(kahan4 *k, const float x[4])
{
    for (int i = 0; i < 4; ++i) {
        float y = x[i] - k->c[i];
        float t = k->s[i] + y;
        k->c[i] = t - k->s[i] - y;
        k->s[i] = t;
    }
}

typedef struct { float s, c; } kahan1;

static void kahan1_add(kahan1 *k, float x)
{
    float y = x - k->c;
    float t = k->s + y;
    k->c = t - k->s - y;
    k->s = t;
}

static void horn_from_sums(const kahan4 *ss, const kahan4 *ds, const float *src, const float *dst, size_t n,
                           float out[16]);

/* ICP::AlignStep, ICP.cpp:35-265, in full: every 4th column of the raw frame
 * (depth in metres, W x H) is back-projected with the guess, projected into the
 * model buffer (mW x mH float3, 0 = invalid) with EyeToClipRH * dstWorldToEye,
 * matched to the nearest model point of a 21 x 21 window within 10 cm (first
 * minimum in scan order), and Horn's closed form aligns the matches.  The
 * window clamps to the model buffer (the reference hard-codes 640 x 480,
 * ICP.cpp:108-109).  Returns the number of associations; out = T*R. */
size_t orc_align_step(const float *depth, const orc_camera *cam, const float guess_e2w[16],
                      const float *model, int mW, int mH, const float model_e2w[16], float out[16],
                      float *scratch_src, float *scratch_dst)
{
    const int W = cam->width, H = cam->height;
    float dstW2E[16];
    memcpy(dstW2E, model_e2w, sizeof dstW2E);
    orc_invert_transform(dstW2E);                                  /* :48 */
    const float flInvX = 1.0f / cam->fx, flInvY = 1.0f / cam->fy;  /* :50-57 */
    const float ppX = (0.5f - cam->cx) * flInvX, ppY = (cam->cy - 0.5f) * flInvY;
    const float halfW = (float)mW * 0.5f, halfH = (float)mH * 0.5f; /* :59-60 */
    float clip[16], proj[16];
    orc_eye_to_clip_rh(cam, clip);
    orc_mat_mul(clip, dstW2E, proj);                               /* :81 */
    kahan4 ss = { { 0 }, { 0 } }, ds = { { 0 }, { 0 } };
    size_t n = 0;
    for (int y = 0; y < H; ++y)                                    /* :66-67 */
        for (int x = 0; x < W; x += 4) {
            float d = depth[(size_t)y * W + x];
            if (0.0f == d)
                continue;
            float pt[4] = { (float)x, -(float)y, -1.0f, 0.0f };    /* :74-77 */
            pt[0] = (pt[0] * flInvX + ppX) * d;
            pt[1] = (pt[1] * flInvY + ppY) * d;
            pt[2] = (pt[2] * 1.0f + 0.0f) * d;
            pt[3] = 1.0f;
            orc_mat_vec(guess_e2w, pt, pt);                        /* :79 */
            float uv[4];
            orc_mat_vec(proj, pt, uv);
            float w = uv[3];
            uv[0] = uv[0] / w; uv[1] = uv[1] / w;                  /* homogenize */
            int u = to_int(uv[0] * halfW + halfW), v0 = to_int(uv[1] * halfH + halfH);
            if (u == (-2147483647 - 1) || v0 == (-2147483647 - 1))
                continue;
            u -= 10;                                               /* :100-104 */
            long v = (long)mH - 1 - v0 - 10;
            if (u < 0) u = 0;
            if (v < 0) v = 0;
            float mind = 0.01f, p2[4] = { 0, 0, 0, 0 };
            long yend = v + 21 < mH ? v + 21 : mH;
            int xend = u + 21 < mW ? u + 21 : mW;
            for (long y1 = v; y1 < yend; ++y1)                     /* :108-125 */
                for (int x1 = u; x1 < xend; ++x1) {
                    const float *q = model + 3 * ((size_t)y1 * mW + x1);
                    if (q[0] == 0.0f && q[1] == 0.0f && q[2] == 0.0f)
                        continue;
                    float dx = q[0] - pt[0], dy = q[1] - pt[1], dz = q[2] - pt[2], dw = 1.0f - pt[3];
                    float l2 = (dx * dx + dy * dy) + (dz * dz + dw * dw);
                    if (l2 < mind) {
                        mind = l2;
                        p2[0] = q[0]; p2[1] = q[1]; p2[2] = q[2]; p2[3] = 1.0f;
                    }
                }
            if (mind < 0.01f) {                                    /* :126-131 */
                memcpy(scratch_src + 3 * n, pt, 3 * sizeof(float));
                memcpy(scratch_dst + 3 * n, p2, 3 * sizeof(float));
                float a[4] = { pt[0], pt[1], pt[2], pt[3] }, b[4] = { p2[0], p2[1], p2[2], 1.0f };
                kahan4_add(&ss, a);
                kahan4_add(&ds, b);
                ++n;
            }
        }
    if (n == 0) {
        float I[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
        memcpy(out, I, sizeof I);
        return 0;
    }
    horn_from_sums(&ss, &ds, scratch_src, scratch_dst, n, out);
    return n;
}

/* ICP::Align, ICP.cpp:9-31: four AlignStep iterations, result = step * result. */
size_t orc_icp_align_ref(const float *depth, const orc_camera *cam, const float guess_e2w[16],
                         const float *model, int mW, int mH, const float model_e2w[16], int iters,
                         float out_e2w[16])
{
    size_t cap = (size_t)cam->height * (cam->width / 4 + 1);
    float *s = (float *)malloc(3 * cap * sizeof(float)), *d = (float *)malloc(3 * cap * sizeof(float));
    float res[16], step[16];
    memcpy(res, guess_e2w, sizeof res);
    size_t n = 0;
    for (int i = 0; i < iters; ++i) {
        n = orc_align_step(depth, cam, res, model, mW, mH, model_e2w, step, s, d);
        orc_mat_mul(step, res, res);
    }
    memcpy(out_e2w, res, sizeof res);
    free(s);
    free(d);
    return n;
}

/* ICP::AlignStep from the correspondences on, ICP.cpp:166-264. */
void orc_horn_step(const float *src, const float *dst, size_t n, float out[16])
{
    kahan4 ss = { { 0 }, { 0 } }, ds = { { 0 }, { 0 } };
    for (size_t i = 0; i < n; ++i) {                         /* :129-130 in assoc order */
        float a[4] = { src[3 * i], src[3 * i + 1], src[3 * i + 2], 1.0f };
        float b[4] = { dst[3 * i], dst[3 * i + 1], dst[3 * i + 2], 1.0f };
        kahan4_add(&ss, a);
        kahan4_add(&ds, b);
    }
    horn_from_sums(&ss, &ds, src, dst, n, out);
}

static void horn_from_sums(const kahan4 *ssp, const kahan4 *dsp, const float *src, const float *dst, size_t n,
                           float out[16])
{
    kahan4 ss = *ssp, ds = *dsp;
    float nf = (float)n, sm[4], dm[4];
    for (int i = 0; i < 4; ++i) {                            /* :166-167 */
        sm[i] = ss.s[i] / nf;
        dm[i] = ds.s[i] / nf;
    }
    kahan1 S[9];
    memset(S, 0, sizeof S);
    for (size_t i = 0; i < n; ++i) {                         /* :175-197 */
        float a[3], b[3];
        for (int j = 0; j < 3; ++j) {
            a[j] = src[3 * i + j] - sm[j];
            b[j] = dst[3 * i + j] - dm[j];
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                kahan1_add(&S[3 * r + c], a[r] * b[c]);
    }
    float Sxx = S[0].s, Sxy = S[1].s, Sxz = S[2].s, Syx = S[3].s, Syy = S[4].s, Syz = S[5].s,
          Szx = S[6].s, Szy = S[7].s, Szz = S[8].s;
    float Nr[4][4] = {                                       /* :199-213 */
        { Sxx + Syy + Szz, Syz - Szy, Szx - Sxz, Sxy - Syx },
        { 0, Sxx - Syy - Szz, Sxy + Syx, Szx + Sxz },
        { 0, 0, Syy - Sxx - Szz, Syz + Szy },
        { 0, 0, 0, Szz - Sxx - Syy },
    };
    double N[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            N[r + 4 * c] = r <= c ? Nr[r][c] : Nr[c][r];
    double ev[4], V[16];
    orc_jacobi4(N, ev, V);                                   /* :215-216 */
    int im = 0;
    for (int i = 1; i < 4; ++i)                              /* :218-221 */
        if (ev[i] > ev[im])
            im = i;
    float q[4] = { (float)V[4 * im], (float)V[4 * im + 1], (float)V[4 * im + 2],
                   (float)V[4 * im + 3] };
    float len = sqrtf((q[0] * q[0] + q[1] * q[1]) + (q[2] * q[2] + q[3] * q[3]));
    for (int i = 0; i < 4; ++i)                              /* :231 normalize */
        q[i] = q[i] / len;
    float xy = q[0] * q[1], xz = q[0] * q[2], xw = q[0] * q[3], yy = q[1] * q[1],
          yz = q[1] * q[2], yw = q[1] * q[3], zz = q[2] * q[2], zw = q[2] * q[3],
          ww = q[3] * q[3];
    float R[16] = { 0 };                                     /* :233-254 */
    R[0] = 1.0f - 2.0f * (zz + ww); R[4] = 2.0f * (yz - xw); R[8] = 2.0f * (yw + xz);
    R[1] = 2.0f * (yz + xw); R[5] = 1.0f - 2.0f * (yy + ww); R[9] = 2.0f * (zw - xy);
    R[2] = 2.0f * (yw - xz); R[6] = 2.0f * (zw + xy); R[10] = 1.0f - 2.0f * (yy + zz);
    R[15] = 1.0f;
    float Rs[4];
    orc_mat_vec(R, sm, Rs);                                  /* :256 */
    float T[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
    for (int i = 0; i < 3; ++i)
        T[12 + i] = dm[i] - Rs[i];
    T[15] = 1.0f;                                            /* :258-261 */
    orc_mat_mul(T, R, out);                                  /* :263 */
}
