/* sanitize_driver.c -- TEST INFRASTRUCTURE: runs the CPU oracle and the seeded
 * scene renderer (merge_fusion_amd/csrc/synth.c) through every entry point a
 * test uses, for a build with -fsanitize=address,undefined (oracle/Makefile,
 * target `sanitize`).  Exits 0 when every step ran; the sanitizers abort on
 * the first finding.  Not part of any product path. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kifi_synth.h"
#include "oracle.h"

#define W 256
#define H 212
#define P (W * H)

static void render(const kifi_scene *sc, int traj, int frame, uint16_t *mm, float w2e[16], double e2w[16])
{
    kifi_trajectory_pose(traj, 2, frame, e2w);
    kifi_scene_render(sc, W, H, 182.5f, 182.5f, W * 0.5f, H * 0.5f, e2w, 1000u + (unsigned)frame, 1.5e-3, 0.02,
                      mm, NULL);
    /* world->eye = inverse of eye->world */
    float m[16];
    for (int i = 0; i < 16; ++i)
        m[i] = (float)e2w[i];
    orc_invert_transform(m);
    memcpy(w2e, m, sizeof m);
}

static int run_volume(const kifi_scene *sc, int res, float side, float trunc, int kb, int storage, int frames)
{
    orc_camera cam = { W, H, 182.5f, 182.5f, W * 0.5f, H * 0.5f, 0.4f, 8.0f };
    orc_volume *v = orc_volume_create(res, side, trunc, kb, storage);
    uint16_t *mm = malloc(P * sizeof *mm);
    float w2e[16];
    double e2w[16];
    for (int f = 0; f < frames; ++f) {
        render(sc, KIFI_TRAJ_CIRCLE, f, mm, w2e, e2w);
        size_t st[4];
        if (orc_integrate(v, mm, &cam, w2e, st) == (size_t)-1)
            return 1;
    }
    const size_t n = orc_volume_size(v);
    float *verts = malloc((3 * n + 1) * 3 * sizeof *verts);
    uint32_t *idx = malloc((15 * n + 1) * sizeof *idx);
    size_t ni = 0;
    int stalled = 0;
    const size_t nv = orc_mesh(v, verts, 3 * n + 1);
    const size_t nvi = orc_mesh_indexed(v, verts, 3 * n + 1, idx, 15 * n + 1, &ni, &stalled);
    /* model prediction at the last pose, then FP64 ICP and the sums against it */
    uint64_t *zbuf = malloc(P * sizeof *zbuf);
    float *pts = calloc(P * 4, sizeof *pts), *nrm = calloc(P * 4, sizeof *nrm);
    for (int i = 0; i < P; ++i)
        zbuf[i] = ~0ull;
    orc_model_splat(v, w2e, &cam, v->desc.vl, 0.1f, zbuf);
    const float campos[3] = { (float)e2w[12], (float)e2w[13], (float)e2w[14] };
    orc_model_resolve(v, w2e, campos, zbuf, W, H, v->desc.vl, 0.1f, pts);
    orc_model_normals4(pts, W, H, campos, 0.15f, nrm);
    float *d0 = malloc(P * sizeof *d0), *d1 = malloc(P / 4 * sizeof *d1);
    render(sc, KIFI_TRAJ_CIRCLE, frames, mm, w2e, e2w);
    for (int i = 0; i < P; ++i)
        d0[i] = (float)mm[i] * 0.001f;
    orc_depth_pyramid_level(d0, W, H, d1);
    double W2E[16], pose[16], stats[2], sums[30];
    for (int i = 0; i < 16; ++i) {
        W2E[i] = w2e[i];
        pose[i] = e2w[i];
    }
    const int it = orc_icp_level_f64(d1, W / 2, H / 2, 91.25, 91.25, 64.0, 53.0, pts, nrm, &cam, W2E, pose, 5, 0.2,
                                     8e-4, 8e-4, stats);
    orc_icp_sums_f64(d0, W, H, 182.5, 182.5, 128.0, 106.0, pts, nrm, &cam, W2E, pose, 0.1, sums);
    float posef[16];
    for (int i = 0; i < 16; ++i)
        posef[i] = (float)pose[i];
    double sums32[30];
    orc_icp_sums_f32terms(d0, 0, pts, nrm, &cam, w2e, posef, 0.1f, sums32);
    /* the reference's window-search ICP against the model points (float3) */
    float *m3 = malloc(P * 3 * sizeof *m3);
    for (int i = 0; i < P; ++i)
        for (int c = 0; c < 3; ++c)
            m3[3 * i + c] = pts[4 * i + 3] != 0.0f ? pts[4 * i + c] : 0.0f;
    float out[16], ge[16];
    for (int i = 0; i < 16; ++i)
        ge[i] = (float)e2w[i];
    orc_icp_align_ref(d0, &cam, ge, m3, W, H, ge, 4, out);
    float w2c[16], clip[16], *bin = malloc(P * 3 * sizeof *bin), *bz = malloc(P * sizeof *bz);
    orc_eye_to_clip_rh(&cam, clip);
    orc_mat_mul(clip, w2e, w2c);
    orc_bin(verts, nv, w2c, W, H, bin, bz);
    printf("res %d kb %d storage %d: %zu voxels, %zu + %zu vertices, %zu indices, icp %d iterations\n", res, kb,
           storage, n, nv, nvi, ni, it);
    free(bin); free(bz); free(m3); free(d0); free(d1); free(zbuf); free(pts); free(nrm); free(verts); free(idx);
    free(mm);
    orc_volume_destroy(v);
    return 0;
}

int main(void)
{
    kifi_scene sc;
    if (kifi_scene_make(KIFI_SCENE_ROOM, 2, &sc))
        return 1;
    int bad = run_volume(&sc, 1024, 4.096f, 0.02f, 10, 0, 4);         /* binary32, 32-bit keys */
    bad |= run_volume(&sc, 1024, 4.096f, 0.02f, 10, 1, 3);            /* binary16 + u8 */
    bad |= run_volume(&sc, 2048, 4.096f, 0.02f, 21, 0, 2);            /* 64-bit keys */
    bad |= run_volume(&sc, 1024, 4.096f, 0.02f, 10, 2, 2);            /* FP64 storage */
    double a[16] = { 4, 1, 0, 0.5, 1, 3, 0.2, 0, 0, 0.2, 2, 0.1, 0.5, 0, 0.1, 1 }, ev[4], vec[16];
    orc_jacobi4(a, ev, vec);
    float src[300], dst[300], hs[16];
    for (int i = 0; i < 300; ++i) {
        src[i] = (float)((i * 37) % 101) * 0.01f;
        dst[i] = src[i] + 0.01f;
    }
    orc_horn_step(src, dst, 100, hs);
    uint32_t in[1000], ex[1000];
    for (int i = 0; i < 1000; ++i)
        in[i] = (uint32_t)i;
    orc_partial_sum_exclusive_u32(in, 1000, ex);
    float ks[1000];
    for (int i = 0; i < 1000; ++i)
        ks[i] = 1e-3f;
    printf("kahan %.6f, jacobi %.4f\n", orc_kahan_sum_f32(ks, 1000), ev[0]);
    return bad;
}
