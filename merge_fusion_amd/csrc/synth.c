/*
 * synth.c -- seeded procedural depth frames (host C, libkifi_synth.so).
 *
 * Replaces the reference's offline fixture tools (poly2depth/, kinect2depth/)
 * with analytic ray casting of planes, spheres and axis-aligned boxes, so that
 * every workload is reproducible from a seed.  All geometry is double precision.
 */
#include "kifi_synth.h"

#include <math.h>
#include <string.h>

#define PI 3.14159265358979323846

uint64_t kifi_splitmix64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

/* uniform double in (0, 1] from a 64-bit draw */
static double unit(uint64_t r)
{
    return ((r >> 11) + 1) * (1.0 / 9007199254740992.0);
}

typedef struct {
    uint64_t s;
} rng_t;

static double rng_uniform(rng_t *g)
{
    g->s = kifi_splitmix64(g->s);
    return unit(g->s);
}

static double rng_range(rng_t *g, double a, double b)
{
    return a + (b - a) * rng_uniform(g);
}

int kifi_scene_make(int kind, uint64_t seed, kifi_scene *s)
{
    memset(s, 0, sizeof *s);
    s->kind = kind;
    rng_t g = { kifi_splitmix64(seed ^ 0x5ce4e5b9d1ce4e5bull) };
    if (kind == KIFI_SCENE_PLANE_SPHERES) {
        s->plane_z = -2.5;
        s->n_spheres = 3;
        for (int i = 0; i < 3; ++i) {
            double r = rng_range(&g, 0.2, 0.5);
            double dist = rng_range(&g, 1.0, 2.45 - r);           /* in front of the plane */
            double half = dist * 0.55;                           /* inside the frustum    */
            s->spheres[i][0] = rng_range(&g, -half, half);
            s->spheres[i][1] = rng_range(&g, -half * 0.8, half * 0.8);
            s->spheres[i][2] = -dist;
            s->spheres[i][3] = r;
        }
        return 0;
    }
    if (kind == KIFI_SCENE_ROOM) {
        const double room[6] = { -2.0, -1.5, -2.0, 2.0, 1.5, 2.0 };
        memcpy(s->room, room, sizeof room);
        s->n_boxes = 10;
        for (int i = 0; i < 10; ++i) {
            /* boxes stand on the floor or hang on a wall, outside the camera circle */
            double sx = rng_range(&g, 0.2, 0.7), sy = rng_range(&g, 0.2, 1.2),
                   sz = rng_range(&g, 0.2, 0.7);
            double ang = rng_range(&g, 0.0, 2.0 * PI);
            double rad = rng_range(&g, 1.45, 1.75);
            double cx = rad * cos(ang), cz = rad * sin(ang);
            double y0 = room[1];
            if (i % 3 == 2)
                y0 = rng_range(&g, -0.8, 0.6);
            s->boxes[i][0] = cx - 0.5 * sx; s->boxes[i][1] = y0; s->boxes[i][2] = cz - 0.5 * sz;
            s->boxes[i][3] = cx + 0.5 * sx; s->boxes[i][4] = y0 + sy; s->boxes[i][5] = cz + 0.5 * sz;
        }
        s->n_spheres = 6;
        for (int i = 0; i < 6; ++i) {
            double r = rng_range(&g, 0.12, 0.3);
            double ang = rng_range(&g, 0.0, 2.0 * PI);
            double rad = i < 2 ? rng_range(&g, 0.0, 0.5) : rng_range(&g, 1.45, 1.95 - r);
            s->spheres[i][0] = rad * cos(ang);
            s->spheres[i][1] = rng_range(&g, -1.2, 1.0);
            s->spheres[i][2] = rad * sin(ang);
            s->spheres[i][3] = r;
        }
        return 0;
    }
    return -1;
}

static double hit_sphere(const double *o, const double *d, const double *sp)
{
    double oc[3] = { o[0] - sp[0], o[1] - sp[1], o[2] - sp[2] };
    double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double b = 2.0 * (d[0] * oc[0] + d[1] * oc[1] + d[2] * oc[2]);
    double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - sp[3] * sp[3];
    double disc = b * b - 4.0 * a * c;
    if (disc < 0.0)
        return INFINITY;
    double sq = sqrt(disc);
    double t0 = (-b - sq) / (2.0 * a), t1 = (-b + sq) / (2.0 * a);
    if (t0 > 1e-9)
        return t0;
    if (t1 > 1e-9)
        return t1;
    return INFINITY;
}

/* slab test; inside = 1 returns the exit distance (room walls seen from inside) */
static double hit_box(const double *o, const double *d, const double *b, int inside)
{
    double tn = -INFINITY, tf = INFINITY;
    for (int a = 0; a < 3; ++a) {
        if (d[a] == 0.0) {
            if (o[a] < b[a] || o[a] > b[a + 3])
                return INFINITY;
            continue;
        }
        double t1 = (b[a] - o[a]) / d[a], t2 = (b[a + 3] - o[a]) / d[a];
        if (t1 > t2) { double t = t1; t1 = t2; t2 = t; }
        if (t1 > tn) tn = t1;
        if (t2 < tf) tf = t2;
    }
    if (tn > tf)
        return INFINITY;
    if (inside)
        return tf > 1e-9 ? tf : INFINITY;
    if (tn > 1e-9)
        return tn;
    return INFINITY;
}

int kifi_scene_render(const kifi_scene *s, int W, int H, float fx, float fy, float cx, float cy,
                      const double *e2w, uint64_t noise_seed, double noise_k, double invalid_frac,
                      uint16_t *out, float *truth)
{
    if (W <= 0 || H <= 0 || !out)
        return -1;
    const double o[3] = { e2w[12], e2w[13], e2w[14] };
    const uint64_t key = kifi_splitmix64(noise_seed);
#pragma omp parallel for schedule(static)
    for (int v = 0; v < H; ++v) {
        for (int u = 0; u < W; ++u) {
            double de[3] = { (u + 0.5 - cx) / fx, (cy - v - 0.5) / fy, -1.0 };
            double d[3];
            for (int r = 0; r < 3; ++r)
                d[r] = e2w[r] * de[0] + e2w[r + 4] * de[1] + e2w[r + 8] * de[2];
            double t = INFINITY;
            if (s->kind == KIFI_SCENE_PLANE_SPHERES && d[2] != 0.0) {
                double tp = (s->plane_z - o[2]) / d[2];
                if (tp > 1e-9)
                    t = tp;
            }
            if (s->kind == KIFI_SCENE_ROOM) {
                double tr = hit_box(o, d, s->room, 1);
                if (tr < t) t = tr;
            }
            for (int i = 0; i < s->n_spheres; ++i) {
                double ts = hit_sphere(o, d, s->spheres[i]);
                if (ts < t) t = ts;
            }
            for (int i = 0; i < s->n_boxes; ++i) {
                double tb = hit_box(o, d, s->boxes[i], 0);
                if (tb < t) t = tb;
            }
            size_t idx = (size_t)v * W + u;
            /* depth along the optical axis equals t because de.z = -1 */
            uint64_t c = key + 4 * (uint64_t)idx;
            double u1 = unit(kifi_splitmix64(c)), u2 = unit(kifi_splitmix64(c + 1));
            double u3 = unit(kifi_splitmix64(c + 2));
            double g = sqrt(-2.0 * log(u1)) * cos(2.0 * PI * u2); /* Box-Muller */
            uint16_t mm = 0;
            if (isfinite(t)) {
                double z = t + g * noise_k * t * t;
                double m = floor(z * 1000.0 + 0.5);
                if (m > 0.0 && m < 65535.0)
                    mm = (uint16_t)m;
            }
            if (u3 <= invalid_frac)
                mm = 0;
            out[idx] = mm;
            if (truth)
                truth[idx] = isfinite(t) ? (float)t : 0.0f;
        }
    }
    return 0;
}

/* eye->world from yaw (about world +y), pitch (about eye +x) and position */
static void pose_from(double yaw, double pitch, double roll, const double *p, double *m)
{
    double cy = cos(yaw), sy = sin(yaw), cp = cos(pitch), sp = sin(pitch), cr = cos(roll),
           sr = sin(roll);
    /* R = Ry(yaw) * Rx(pitch) * Rz(roll) */
    double Ry[9] = { cy, 0, -sy, 0, 1, 0, sy, 0, cy };      /* column storage 3x3 */
    double Rx[9] = { 1, 0, 0, 0, cp, sp, 0, -sp, cp };
    double Rz[9] = { cr, sr, 0, -sr, cr, 0, 0, 0, 1 };
    double A[9], R[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            A[r + 3 * c] = Ry[r] * Rx[3 * c] + Ry[r + 3] * Rx[3 * c + 1] + Ry[r + 6] * Rx[3 * c + 2];
        }
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            R[r + 3 * c] = A[r] * Rz[3 * c] + A[r + 3] * Rz[3 * c + 1] + A[r + 6] * Rz[3 * c + 2];
    memset(m, 0, 16 * sizeof(double));
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            m[r + 4 * c] = R[r + 3 * c];
    m[12] = p[0]; m[13] = p[1]; m[14] = p[2]; m[15] = 1.0;
}

int kifi_trajectory_pose(int traj, uint64_t seed, int frame, double *m)
{
    if (frame < 0)
        return -1;
    if (traj == KIFI_TRAJ_STATIC) {
        double p[3] = { 0, 0, 0 };
        pose_from(0, 0, 0, p, m);
        return 0;
    }
    if (traj == KIFI_TRAJ_CIRCLE) {
        const double r = 1.2, step = 0.015, deg = PI / 180.0;
        double th0 = 0.25 * PI;
        double th = th0 + frame * step / r;
        rng_t g = { kifi_splitmix64(seed * 0x100000001b3ull + (uint64_t)frame) };
        double jt[3] = { rng_range(&g, -0.002, 0.002), rng_range(&g, -0.002, 0.002),
                         rng_range(&g, -0.002, 0.002) };
        double jy = rng_range(&g, -0.2, 0.2) * deg, jp = rng_range(&g, -0.2, 0.2) * deg,
               jr = rng_range(&g, -0.2, 0.2) * deg;
        double p[3] = { r * cos(th) + jt[0], 0.1 + jt[1], r * sin(th) + jt[2] };
        /* start looking at the room centre; the view turns 1 degree per frame */
        double yaw0 = atan2(cos(th0), sin(th0));
        double yaw = yaw0 + frame * deg + jy;
        pose_from(yaw, -12.0 * deg + jp, jr, p, m);
        return 0;
    }
    if (traj == KIFI_TRAJ_FAST) {
        const double deg = PI / 180.0;
        rng_t g = { kifi_splitmix64(seed ^ 0xfa57fa57fa57ull) };
        double p[3] = { 0.0, 0.1, 0.6 }, yaw = 0.0, pitch = -10.0 * deg;
        for (int f = 1; f <= frame; ++f) {
            double len = rng_range(&g, 0.02, 0.10), dir = rng_range(&g, 0.0, 2.0 * PI);
            double ang = rng_range(&g, 2.0, 15.0) * deg * (rng_uniform(&g) < 0.5 ? -1.0 : 1.0);
            double dp = rng_range(&g, -2.0, 2.0) * deg;
            double nx = p[0] + len * cos(dir), nz = p[2] + len * sin(dir);
            if (nx * nx + nz * nz > 1.0) { /* reflect back towards the centre */
                nx = p[0] - len * cos(dir);
                nz = p[2] - len * sin(dir);
            }
            p[0] = nx; p[2] = nz;
            yaw += ang;
            pitch += dp;
            if (pitch > 0.0) pitch = 0.0;
            if (pitch < -20.0 * deg) pitch = -20.0 * deg;
        }
        pose_from(yaw, pitch, 0.0, p, m);
        return 0;
    }
    if (traj == KIFI_TRAJ_SWAY) {
        /* independent seeded views about the identity pose; frame 0 is the identity */
        const double deg = PI / 180.0;
        double p[3] = { 0, 0, 0 };
        if (frame == 0) {
            pose_from(0, 0, 0, p, m);
            return 0;
        }
        rng_t g = { kifi_splitmix64(seed * 0x9e3779b97f4a7c15ull + (uint64_t)frame) };
        for (int i = 0; i < 3; ++i)
            p[i] = rng_range(&g, -0.05, 0.05);
        const double yaw = rng_range(&g, -2.0, 2.0) * deg, pitch = rng_range(&g, -2.0, 2.0) * deg,
                     roll = rng_range(&g, -2.0, 2.0) * deg;
        pose_from(yaw, pitch, roll, p, m);
        return 0;
    }
    return -1;
}
