/*
 * kifi_synth.h -- seeded procedural depth frames for the fusion workloads.
 *
 * No reference counterpart as a library: the reference renders its fixtures
 * offline with a C#/XNA tool (poly2depth/, output format poly2depth/README.txt:23-55).
 * These frames feed the benchmark and the parity tests identically on the CPU
 * and the GPU.  Host-only C ABI (libkifi_synth.so).
 *
 * Conventions (same as the fusion library): eye space is right-handed and looks
 * along -z; a pixel (u, v) (v counted from the top row) has its centre on the
 * ray ((u + 0.5 - cx)/fx, (cy - v - 0.5)/fy, -1); depth is the distance along
 * the optical axis in millimetres, 0 = invalid; matrices are float/double[16]
 * in column storage (element (r, c) at m[r + 4c]).
 */
#ifndef KIFI_SYNTH_H
#define KIFI_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    KIFI_SCENE_PLANE_SPHERES = 0, /* plane at z = -2.5 m + 3 spheres, identity pose */
    KIFI_SCENE_ROOM = 1           /* 4 x 3 x 4 m room + 10 boxes + 6 spheres        */
};

enum {
    KIFI_TRAJ_STATIC = 0, /* identity pose                                         */
    KIFI_TRAJ_CIRCLE = 1, /* r = 1.2 m circle, 1.5 cm and 1 deg per frame, jitter  */
    KIFI_TRAJ_FAST = 2,   /* random walk, steps U[2,10] cm and U[2,15] deg         */
    KIFI_TRAJ_SWAY = 3    /* independent views about the identity: U[-5,5] cm per  */
                          /* axis, yaw/pitch/roll U[-2,2] deg; frame 0 = identity  */
};

typedef struct kifi_scene_s {
    int    kind;
    int    n_spheres, n_boxes;
    double spheres[8][4];   /* centre xyz, radius                    */
    double boxes[16][6];    /* min xyz, max xyz                      */
    double plane_z;         /* KIFI_SCENE_PLANE_SPHERES: plane z = .. */
    double room[6];         /* KIFI_SCENE_ROOM: inner box min, max   */
} kifi_scene;

/* Builds the scene for `kind` from `seed`.  Returns 0 on success. */
int kifi_scene_make(int kind, uint64_t seed, kifi_scene *out);

/* Renders one frame: exact ray casting in double precision, then additive
 * Gaussian noise N(0, (noise_k * z^2)^2) metres, rounding to mm, and a
 * Bernoulli(invalid_frac) mask of zeroed pixels, all drawn from a counter-based
 * generator keyed by (noise_seed, pixel).  true_depth_m may be NULL. */
int kifi_scene_render(const kifi_scene *s, int width, int height, float fx, float fy, float cx,
                      float cy, const double eyeToWorld[16], uint64_t noise_seed, double noise_k,
                      double invalid_frac, uint16_t *out_mm, float *true_depth_m);

/* Ground-truth eye->world pose of frame `frame` for trajectory `traj`.
 * KIFI_TRAJ_FAST is a seeded walk: poses depend on all earlier frames. */
int kifi_trajectory_pose(int traj, uint64_t seed, int frame, double eyeToWorld[16]);

/* SplitMix64 step (Steele, Lea, Flood 2014), exposed for seeding tests. */
uint64_t kifi_splitmix64(uint64_t x);

#ifdef __cplusplus
}
#endif
#endif
