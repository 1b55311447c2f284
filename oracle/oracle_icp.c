/*
 * oracle_icp.c -- CPU oracle for model prediction and point-to-plane ICP
 * (test infrastructure only; see oracle.h).
 *
 * No reference counterpart: the reference tracks with 4 fixed Horn steps over
 * a window search (KinectFission/kifi/ICP.cpp:9-265) that the build replaces
 * by projective point-to-plane ICP on a 3-level pyramid.  Two kinds of code:
 *   - model splat / resolve / normals and the depth pyramid are restated in
 *     binary32 in the same order as the HIP kernels (bit-exact parity);
 *   - the ICP solve is FP64 throughout (the FP64 oracle the GPU pose error is
 *     reported against).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------ model splat */

static uint32_t orderable(float z)
{
    if (z == 0.0f)
        z = 0.0f;
    uint32_t b;
    memcpy(&b, &z, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

static float vdist(const orc_volume *v, size_t i)
{
    if (v->storage == ORC_STORE_F32) {
        const float *sw = (const float *)v->values + 2 * i;
        return sw[0] / sw[1];
    }
    return orc_f16_to_f32((uint16_t)(((const uint32_t *)v->values)[i] & 0xffffu));
}

static float vweight(const orc_volume *v, size_t i)
{
    if (v->storage == ORC_STORE_F32)
        return ((const float *)v->values)[2 * i + 1];
    return (float)((((const uint32_t *)v->values)[i] >> 16) & 0xffu);
}

/* Surface sample of voxel i seen from the model camera: returns 0 if the voxel
 * does not contribute, else fills the eye-space centre and the surface depth. */
static int model_sample(const orc_volume *v, size_t i, const float w2e[16], float dmax, float znear,
                        float *ex, float *ey, float *z, float *zs, float c[3])
{
    if (!(vweight(v, i) > 0.0f))
        return 0;
    float D = vdist(v, i);
    if (!(fabsf(D) < dmax))
        return 0;
    uint32_t x, y, zz;
    orc_unpack(v->keys[i], v->desc.kb, &x, &y, &zz);
    c[0] = (float)x * v->desc.vl + v->desc.vlHalfPlusMin;
    c[1] = (float)y * v->desc.vl + v->desc.vlHalfPlusMin;
    c[2] = (float)zz * v->desc.vl + v->desc.vlHalfPlusMin;
    float p[4] = { c[0], c[1], c[2], 1.0f }, e[4];
    orc_mat_vec(w2e, p, e);
    float depth = -e[2];
    if (!(depth > znear))
        return 0;
    *ex = e[0];
    *ey = e[1];
    *z = depth;
    *zs = depth + D;
    return 1;
}

/* Model prediction: z-buffered splat of near-surface voxels (|D| < dmax) at the
 * estimated surface depth along their ray; footprint k x k pixels with
 * k = clamp(ceil(vl*fx/z), 1, 3).  zbuf holds (orderable(zs) << 32 | index). */
void orc_model_splat(const orc_volume *v, const float w2e[16], const orc_camera *cam, float dmax,
                     float znear, uint64_t *zbuf)
{
    const int W = cam->width, H = cam->height;
    for (size_t q = 0; q < (size_t)W * H; ++q)
        zbuf[q] = ~0ull;
    for (size_t i = 0; i < v->size; ++i) {
        float ex, ey, z, zs, c[3];
        if (!model_sample(v, i, w2e, dmax, znear, &ex, &ey, &z, &zs, c))
            continue;
        float u = cam->fx * ex / z + cam->cx;
        float vv = cam->cy - cam->fy * ey / z;
        if (!(u > -8.0f && u < (float)W + 8.0f && vv > -8.0f && vv < (float)H + 8.0f))
            continue;
        float kf = ceilf(v->desc.vl * cam->fx / z);
        int k = kf < 1.0f ? 1 : (kf > 3.0f ? 3 : (int)kf);
        float half = 0.5f * (float)(k - 1);
        int u0 = (int)floorf(u - half), v0 = (int)floorf(vv - half);
        uint64_t key = (uint64_t)orderable(z) << 32 | (uint32_t)i;
        for (int dv = 0; dv < k; ++dv)
            for (int du = 0; du < k; ++du) {
                int pu = u0 + du, pv = v0 + dv;
                if (pu < 0 || pu >= W || pv < 0 || pv >= H)
                    continue;
                uint64_t *slot = &zbuf[(size_t)pv * W + pu];
                if (key < *slot)
                    *slot = key;
            }
    }
}

/* Resolve the splat into a world-space point map: p = cam + (c - cam) * (zs / z). */
void orc_model_resolve(const orc_volume *v, const float w2e[16], const float camPos[3],
                       const uint64_t *zbuf, int W, int H, float dmax, float znear, float *pts)
{
    for (size_t q = 0; q < (size_t)W * H; ++q) {
        float *o = pts + 4 * q;
        o[0] = o[1] = o[2] = o[3] = 0.0f;
        if (zbuf[q] == ~0ull)
            continue;
        size_t i = (uint32_t)zbuf[q];
        float ex, ey, z, zs, c[3];
        model_sample(v, i, w2e, dmax, znear, &ex, &ey, &z, &zs, c);
        float s = zs / z;
        o[0] = camPos[0] + (c[0] - camPos[0]) * s;
        o[1] = camPos[1] + (c[1] - camPos[1]) * s;
        o[2] = camPos[2] + (c[2] - camPos[2]) * s;
        o[3] = 1.0f;
    }
}

/* Normals from the point map by central differences over +-NB pixels (all four
 * neighbours valid and closer than `disc`), oriented towards the camera.
 * A baseline of several pixels keeps the sensor noise of the fused points from
 * tilting the normals.  pts/nrm are W*H*4 (w = validity). */
#define NB 4
void orc_model_normals4(const float *pts, int W, int H, const float camPos[3], float disc, float *nrm)
{
    const float disc2 = disc * disc;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const float *P = pts + 4 * ((size_t)y * W + x);
            float *N = nrm + 4 * ((size_t)y * W + x);
            N[0] = N[1] = N[2] = N[3] = 0.0f;
            if (P[3] == 0.0f || x < NB || x >= W - NB || y < NB || y >= H - NB)
                continue;
            const float *nb[4];
            const int dxs[4] = { NB, -NB, 0, 0 }, dys[4] = { 0, 0, NB, -NB };
            int ok = 1;
            for (int d = 0; d < 4; ++d) {
                const float *Q = pts + 4 * ((size_t)(y + dys[d]) * W + (x + dxs[d]));
                float a = Q[0] - P[0], b = Q[1] - P[1], c = Q[2] - P[2];
                ok &= Q[3] != 0.0f && (a * a + b * b) + c * c < disc2;
                nb[d] = Q;
            }
            if (!ok)
                continue;
            const float *r = nb[0], *l = nb[1], *dn = nb[2], *up = nb[3];
            float ux = r[0] - l[0], uy = r[1] - l[1], uz = r[2] - l[2];
            float vx = dn[0] - up[0], vy = dn[1] - up[1], vz = dn[2] - up[2];
            float nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
            float len = sqrtf((nx * nx + ny * ny) + nz * nz);
            if (!(len > 0.0f))
                continue;
            nx = nx / len; ny = ny / len; nz = nz / len;
            float wx = P[0] - camPos[0], wy = P[1] - camPos[1], wz = P[2] - camPos[2];
            if ((nx * wx + ny * wy) + nz * wz > 0.0f) {
                nx = -nx; ny = -ny; nz = -nz;
            }
            N[0] = nx; N[1] = ny; N[2] = nz; N[3] = 1.0f;
        }
}

/* 2x2 mean of valid depths, in source order ((0,0),(1,0),(0,1),(1,1)); binary32. */
void orc_depth_pyramid_level(const float *src, int W, int H, float *dst)
{
    int w = W / 2, h = H / 2;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float s = 0.0f, c = 0.0f;
            for (int j = 0; j < 2; ++j)
                for (int i = 0; i < 2; ++i) {
                    float d = src[(size_t)(2 * y + j) * W + 2 * x + i];
                    if (d > 0.0f) {
                        s += d;
                        c += 1.0f;
                    }
                }
            dst[(size_t)y * w + x] = c > 0.0f ? s / c : 0.0f;
        }
}

/* ------------------------------------------------------------- FP64 ICP */

static int cholesky6(double A[36], double b[6], double x[6])
{
    double L[36] = { 0 };
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = A[6 * i + j];
            for (int k = 0; k < j; ++k)
                s -= L[6 * i + k] * L[6 * j + k];
            if (i == j) {
                if (!(s > 0.0))
                    return -1;
                L[6 * i + i] = sqrt(s);
            } else {
                L[6 * i + j] = s / L[6 * j + j];
            }
        }
    double y[6];
    for (int i = 0; i < 6; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k)
            s -= L[6 * i + k] * y[k];
        y[i] = s / L[6 * i + i];
    }
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < 6; ++k)
            s -= L[6 * k + i] * x[k];
        x[i] = s / L[6 * i + i];
    }
    return 0;
}

/* Rodrigues rotation of the twist (w, t) as a 4x4 (column storage) */
static void twist_to_mat(const double xi[6], double M[16])
{
    double wx = xi[0], wy = xi[1], wz = xi[2];
    double th = sqrt(wx * wx + wy * wy + wz * wz);
    double K[9] = { 0, wz, -wy, -wz, 0, wx, wy, -wx, 0 }; /* column storage of [w]x */
    double a, b;
    if (th < 1e-12) {
        a = 1.0;
        b = 0.5;
    } else {
        a = sin(th) / th;
        b = (1.0 - cos(th)) / (th * th);
    }
    double K2[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            double s = 0;
            for (int k = 0; k < 3; ++k)
                s += K[r + 3 * k] * K[k + 3 * c];
            K2[r + 3 * c] = s;
        }
    memset(M, 0, 16 * sizeof(double));
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            M[r + 4 * c] = (r == c ? 1.0 : 0.0) + a * K[r + 3 * c] + b * K2[r + 3 * c];
    M[12] = xi[3];
    M[13] = xi[4];
    M[14] = xi[5];
    M[15] = 1.0;
}

static void mat_mul_d(const double *A, const double *B, double *C)
{
    double R[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double s = 0;
            for (int k = 0; k < 4; ++k)
                s += A[r + 4 * k] * B[k + 4 * c];
            R[r + 4 * c] = s;
        }
    memcpy(C, R, sizeof R);
}

/* One projective association in FP64: the point-to-plane normal equations in
 * the layout of the GPU's reduction, out[0..20] = sum w J_i J_j (i >= j, row by
 * row), out[21..26] = sum w J_i r, out[27] = count, out[28] = sum r^2,
 * out[29] = sum w r^2, with w = 1/d^4 and r = n . (q - p). */
static void icp_sums_f64(const float *depth, int Wl, int Hl, double fxl, double fyl, double cxl, double cyl,
                         const float *pts, const float *nrm, const orc_camera *cam, const double W2E[16],
                         const double pose[16], double dist_thresh, double out[30])
{
    memset(out, 0, 30 * sizeof(double));
    for (int v = 0; v < Hl; ++v)
        for (int u = 0; u < Wl; ++u) {
            double d = depth[(size_t)v * Wl + u];
            if (!(d > 0.0))
                continue;
            double pe[3] = { (u + 0.5 - cxl) / fxl * d, (cyl - v - 0.5) / fyl * d, -d };
            double p[3];
            for (int r = 0; r < 3; ++r)
                p[r] = pose[r] * pe[0] + pose[r + 4] * pe[1] + pose[r + 8] * pe[2] + pose[r + 12];
            double e[3];
            for (int r = 0; r < 3; ++r)
                e[r] = W2E[r] * p[0] + W2E[r + 4] * p[1] + W2E[r + 8] * p[2] + W2E[r + 12];
            double z = -e[2];
            if (!(z > 0.0))
                continue;
            double fu = cam->fx * e[0] / z + cam->cx, fv = cam->cy - cam->fy * e[1] / z;
            if (!(fu >= 0.0 && fu < cam->width && fv >= 0.0 && fv < cam->height))
                continue;
            size_t q = (size_t)(int)fv * cam->width + (int)fu;
            const float *Q = pts + 4 * q, *N = nrm + 4 * q;
            if (Q[3] == 0.0f || N[3] == 0.0f)
                continue;
            double df[3] = { Q[0] - p[0], Q[1] - p[1], Q[2] - p[2] };
            if (df[0] * df[0] + df[1] * df[1] + df[2] * df[2] > dist_thresh * dist_thresh)
                continue;
            double n[3] = { N[0], N[1], N[2] };
            double r = n[0] * df[0] + n[1] * df[1] + n[2] * df[2];
            double J[6] = { p[1] * n[2] - p[2] * n[1], p[2] * n[0] - p[0] * n[2],
                            p[0] * n[1] - p[1] * n[0], n[0], n[1], n[2] };
            /* inverse variance of the sensor noise model sigma ~ z^2 */
            double iz = 1.0 / d, wgt = iz * iz * iz * iz;
            int t = 0;
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j <= i; ++j)
                    out[t++] += wgt * J[i] * J[j];
            for (int i = 0; i < 6; ++i)
                out[21 + i] += wgt * J[i] * r;
            out[27] += 1.0;
            out[28] += r * r;
            out[29] += wgt * r * r;
        }
}

void orc_icp_sums_f64(const float *depth, int Wl, int Hl, double fxl, double fyl, double cxl, double cyl,
                      const float *pts, const float *nrm, const orc_camera *cam, const double W2E[16],
                      const double pose[16], double dist_thresh, double out[30])
{
    icp_sums_f64(depth, Wl, Hl, fxl, fyl, cxl, cyl, pts, nrm, cam, W2E, pose, dist_thresh, out);
}

/* The same association in binary32, operation for operation as the HIP
 * association kernel (k_icp_assoc): level intrinsics = full-resolution values
 * x 2^-level, back-projection, pose and model transforms in the order of
 * mat_vec, projection, gate, point-to-plane residual and Jacobian, w = 1/d^4.
 * Each binary32 term is then summed in double, so out[] is the (practically
 * exact) sum of the very terms the GPU reduces: the GPU's difference from it
 * is its summation error alone.  Layout as orc_icp_sums_f64. */
static void mat_vec_f32(const float *m, const float v[4], float out[4])
{
    float r[4];
    for (int i = 0; i < 4; ++i)
        r[i] = (m[i] * v[0] + m[i + 4] * v[1]) + (m[i + 8] * v[2] + m[i + 12] * v[3]);
    memcpy(out, r, sizeof r);
}

void orc_icp_sums_f32terms(const float *depth, int level, const float *pts, const float *nrm,
                           const orc_camera *cam, const float w2e[16], const float pose_f[16], float gate,
                           double out[30])
{
    const float sc = 1.0f / (float)(1 << level);
    const int Wl = cam->width >> level, Hl = cam->height >> level;
    const float fxl = cam->fx * sc, fyl = cam->fy * sc, cxl = cam->cx * sc, cyl = cam->cy * sc;
    const float Wm = (float)cam->width, Hm = (float)cam->height;
    memset(out, 0, 30 * sizeof(double));
    for (int v = 0; v < Hl; ++v)
        for (int u = 0; u < Wl; ++u) {
            const float d = depth[(size_t)v * Wl + u];
            if (!(d > 0.0f))
                continue;
            const float pe[4] = { ((float)u + 0.5f - cxl) / fxl * d, (cyl - (float)v - 0.5f) / fyl * d, -d, 1.0f };
            float p[4], e[4];
            mat_vec_f32(pose_f, pe, p);
            mat_vec_f32(w2e, p, e);
            const float z = -e[2];
            if (!(z > 0.0f))
                continue;
            const float fu = cam->fx * e[0] / z + cam->cx, fv = cam->cy - cam->fy * e[1] / z;
            if (!(fu >= 0.0f && fu < Wm && fv >= 0.0f && fv < Hm))
                continue;
            const size_t q = (size_t)(int)fv * cam->width + (int)fu;
            const float *Q = pts + 4 * q, *N = nrm + 4 * q;
            if (Q[3] == 0.0f || N[3] == 0.0f)
                continue;
            const float dx = Q[0] - p[0], dy = Q[1] - p[1], dz = Q[2] - p[2];
            if (!((dx * dx + dy * dy) + dz * dz <= gate * gate))
                continue;
            const float r = (N[0] * dx + N[1] * dy) + N[2] * dz;
            const float J[6] = { p[1] * N[2] - p[2] * N[1], p[2] * N[0] - p[0] * N[2], p[0] * N[1] - p[1] * N[0],
                                 N[0], N[1], N[2] };
            const float iz = 1.0f / d, wgt = iz * iz * iz * iz;
            int t = 0;
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j <= i; ++j)
                    out[t++] += (double)(wgt * J[i] * J[j]);
            for (int i = 0; i < 6; ++i)
                out[21 + i] += (double)(wgt * J[i] * r);
            out[27] += 1.0;
            out[28] += (double)(r * r);
            out[29] += (double)(wgt * (r * r));
        }
}

/* One pyramid level of projective point-to-plane ICP in FP64.
 * Live level: depth (metres, binary32 as the GPU sees it), Wl x Hl, intrinsics
 * (fxl, fyl, cxl, cyl).  Model: pts/nrm W*H*4 at the full-resolution camera `cam`
 * with world->eye modelW2E.  pose_io: eye->world guess in, result out.
 * Convergence: |w| < eps_rot and |t| < eps_trans.  Returns iterations run;
 * stats[0] = RMS of the last association, stats[1] = its inlier count. */
/* The motion prior of kifi_icp_params.motion_prior, restated: the motion since
 * the pass's start is m = M xi accumulated over the solves, with
 * M = [[I, 0], [-[c]x, I]] for the start's camera centre c (a world-frame
 * increment (w, v) moves c by w x c + v to first order), and each solve is
 * (A + mu M^T M) xi = b - mu M^T m_acc with mu = weight x inliers: the normal
 * equations of |J xi - r|^2 + mu |m_acc + M xi|^2.  No reference counterpart. */
static void prior_matrix(const double c[3], double M[36])
{
    memset(M, 0, 36 * sizeof(double));
    for (int i = 0; i < 6; ++i)
        M[6 * i + i] = 1.0;
    /* rows 3..5, columns 0..2: -[c]x, so that row 3 + k is (w x c)_k */
    M[6 * 3 + 1] = c[2];
    M[6 * 3 + 2] = -c[1];
    M[6 * 4 + 0] = -c[2];
    M[6 * 4 + 2] = c[0];
    M[6 * 5 + 0] = c[1];
    M[6 * 5 + 1] = -c[0];
}

int orc_icp_level_f64_prior(const float *depth, int Wl, int Hl, double fxl, double fyl, double cxl, double cyl,
                            const float *pts, const float *nrm, const orc_camera *cam, const double W2E[16],
                            double pose[16], int max_iters, double dist_thresh, double eps_rot,
                            double eps_trans, double prior, const double prior_c[3], double prior_acc[6],
                            double stats[2])
{
    int it = 0;
    double M[36];
    prior_matrix(prior_c, M);
    stats[0] = 0.0;
    stats[1] = 0.0;
    for (; it < max_iters; ++it) {
        double sums[30];
        icp_sums_f64(depth, Wl, Hl, fxl, fyl, cxl, cyl, pts, nrm, cam, W2E, pose, dist_thresh, sums);
        double A[36], b[6], cnt = sums[27], e2 = sums[28];
        int t = 0;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j <= i; ++j) {
                A[6 * i + j] = sums[t];
                A[6 * j + i] = sums[t];
                ++t;
            }
        for (int i = 0; i < 6; ++i)
            b[i] = sums[21 + i];
        stats[0] = cnt > 0 ? sqrt(e2 / cnt) : 0.0;
        stats[1] = cnt;
        if (cnt < 12.0)
            break;
        if (prior > 0.0) {
            const double mu = prior * cnt;
            for (int i = 0; i < 6; ++i) {
                double s = 0.0;
                for (int k = 0; k < 6; ++k)
                    s += M[6 * k + i] * prior_acc[k];
                b[i] -= mu * s;
                for (int j = 0; j < 6; ++j) {
                    double h = 0.0;
                    for (int k = 0; k < 6; ++k)
                        h += M[6 * k + i] * M[6 * k + j];
                    A[6 * i + j] += mu * h;
                }
            }
        }
        double xi[6];
        if (cholesky6(A, b, xi))
            break;
        if (prior > 0.0)
            for (int i = 0; i < 6; ++i) {
                double s = 0.0;
                for (int j = 0; j < 6; ++j)
                    s += M[6 * i + j] * xi[j];
                prior_acc[i] += s;
            }
        double D[16];
        twist_to_mat(xi, D);
        mat_mul_d(D, pose, pose);
        double rot = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
        double tr = sqrt(xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5]);
        if (rot < eps_rot && tr < eps_trans) {
            ++it;
            break;
        }
    }
    return it;
}

int orc_icp_level_f64(const float *depth, int Wl, int Hl, double fxl, double fyl, double cxl, double cyl,
                      const float *pts, const float *nrm, const orc_camera *cam, const double W2E[16],
                      double pose[16], int max_iters, double dist_thresh, double eps_rot,
                      double eps_trans, double stats[2])
{
    const double c[3] = { 0.0, 0.0, 0.0 };
    double acc[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
    return orc_icp_level_f64_prior(depth, Wl, Hl, fxl, fyl, cxl, cyl, pts, nrm, cam, W2E, pose, max_iters,
                                   dist_thresh, eps_rot, eps_trans, 0.0, c, acc, stats);
}
