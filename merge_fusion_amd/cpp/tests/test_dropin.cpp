// test_dropin.cpp -- the reference's class API (KinectFission/kifi/*.h) used the
// way its own callers use it (UnitTests/Integrator.cpp:15-55, Pipeline.cpp:26-68),
// compiled as ordinary C++ against kifi/kifi.hpp and libkifi_hip.so.
//
// Writes what each call produced into the output directory given as argv[1]
// (raw little-endian arrays, one file per result); the pytest driver
// tests/test_gpu_cpp_dropin.py compares them with the oracle.  Checks that need
// no oracle (sizes, invariants) are asserted here and reported as
// "CHECK <name> ok|FAIL"; the exit status is the number of failures.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kifi/kifi.hpp"
#include "kifi_synth.h"

using namespace kifi;
using util::float3;
using util::float4x4;

static int g_fail = 0;
static void check(bool ok, const char *name)
{
    std::printf("CHECK %s %s\n", name, ok ? "ok" : "FAIL");
    std::fflush(stdout);
    g_fail += ok ? 0 : 1;
}

static std::string g_dir;
template <typename T>
static void dump(const char *name, const T *data, std::size_t n)
{
    const std::string p = g_dir + "/" + name;
    FILE *f = std::fopen(p.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", p.c_str());
        std::exit(100);
    }
    if (n)
        std::fwrite(data, sizeof(T), n, f);
    std::fclose(f);
}

// a frame in metres exactly as DepthStream delivers it: (float)mm * 0.001f
static util::vector2d<float> render_m(const kifi_scene &sc, const DepthSensorParams &cam, const double e2w[16],
                                      uint64_t seed)
{
    const int W = cam.ResolutionPixels().x, H = cam.ResolutionPixels().y;
    std::vector<uint16_t> mm((std::size_t)W * H);
    kifi_scene_render(&sc, W, H, cam.FocalLengthPixels().x, cam.FocalLengthPixels().y, cam.PrincipalPointPixels().x,
                      cam.PrincipalPointPixels().y, e2w, seed, 1.5e-3, 0.02, mm.data(), nullptr);
    util::vector2d<float> m(W, H);
    for (std::size_t i = 0; i < mm.size(); ++i)
        m[i] = (float)mm[i] * 0.001f;
    return m;
}

static float4x4 to_f(const double m[16])
{
    float4x4 r;
    for (int i = 0; i < 16; ++i)
        r.data()[i] = (float)m[i];
    return r;
}

static float4x4 world_to_eye(const double e2w[16])
{
    float4x4 r = to_f(e2w);
    util::invert_transform(r);
    return r;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s OUT_DIR\n", argv[0]);
        return 100;
    }
    g_dir = argv[1];
    const DepthSensorParams cam(util::int2(512, 424), util::float2(365.0f), util::float2(256.0f, 212.0f),
                                util::float2(0.4f, 8.0f));

    // ---- Integrator + Volume::Data: one frame of the plane + spheres scene, 8 mm voxels
    {
        kifi_scene sc;
        kifi_scene_make(KIFI_SCENE_PLANE_SPHERES, 1, &sc);
        double I[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
        util::vector2d<float> frame = render_m(sc, cam, I, 7);
        dump("c1_depth_m.f32", frame.data(), frame.size());
        Volume vol(1024, 8.192f, 0.04f);
        Integrator integrator;
        integrator.Integrate(vol, frame, cam, float4x4::identity());
        VolumeData data = vol.Data();
        check(data.size() > 100000, "integrator_volume_nonempty");
        bool sorted = true;
        for (std::size_t i = 1; i < data.size(); ++i)
            sorted = sorted && data.keys_cbegin()[i - 1] < data.keys_cbegin()[i];
        check(sorted, "volume_keys_strictly_increasing");
        dump("c1_keys.u32", data.keys_cbegin(), data.size());
        dump("c1_values.f32x2", data.values_cbegin(), data.size());

        // ---- Mesher: both overloads, and the OBJ export
        Mesher mesher;
        std::vector<float3> verts, iverts;
        std::vector<unsigned> idx;
        mesher.Mesh(vol, verts);
        mesher.Mesh(vol, iverts, idx);
        check(!verts.empty() && verts.size() == iverts.size(), "mesher_vertex_counts_agree");
        bool in_range = idx.size() % 3 == 0;
        for (unsigned i : idx)
            in_range = in_range && i < iverts.size();
        check(in_range, "mesher_indices_in_range");
        dump("c1_verts.f32x3", &verts[0].x, verts.size() * 3);
        dump("c1_iverts.f32x3", &iverts[0].x, iverts.size() * 3);
        dump("c1_idx.u32", idx.data(), idx.size());
        const std::string obj = g_dir + "/c1_mesh.obj";
        Mesher::Mesh2Obj(iverts, idx, obj.c_str());

        // ---- checkpoint / resume: the volume to a file and back
        const std::string vf = g_dir + "/c1.kvol";
        vol.Save(vf.c_str());
        Volume back = Volume::Load(vf.c_str());
        VolumeData d2 = back.Data();
        check(back.Resolution() == 1024 && d2.size() == data.size() &&
                  std::memcmp(d2.keys_cbegin(), data.keys_cbegin(), data.size() * 4) == 0 &&
                  std::memcmp(d2.values_cbegin(), data.values_cbegin(), data.size() * sizeof(Voxel)) == 0,
              "volume_save_load_round_trip");
    }

    // ---- Renderer::Bin + ICP::Align: the reference's tracking chain (Pipeline.cpp:37-52)
    {
        kifi_scene sc;
        kifi_scene_make(KIFI_SCENE_ROOM, 2, &sc);
        Volume vol(1024, 4.096f, 0.02f);
        Integrator integrator;
        double e2w[16];
        for (int f = 0; f < 3; ++f) {
            kifi_trajectory_pose(KIFI_TRAJ_CIRCLE, 2, f, e2w);
            integrator.Integrate(vol, render_m(sc, cam, e2w, 10 + f), cam, world_to_eye(e2w));
        }
        Mesher mesher;
        std::vector<float3> cloud;
        mesher.Mesh(vol, cloud);
        const float4x4 e2w2 = to_f(e2w);
        util::vector2d<float3> synth(1024, 768);
        Renderer renderer;
        renderer.Bin(cloud, cam.EyeToClipRH() * world_to_eye(e2w), synth);
        std::size_t filled = 0;
        for (const float3 &p : synth)
            filled += (p.x != 0.0f || p.y != 0.0f || p.z != 0.0f) ? 1 : 0;
        check(filled > synth.size() / 4, "bin_fills_buffer");
        double e2w3[16];
        kifi_trajectory_pose(KIFI_TRAJ_CIRCLE, 2, 3, e2w3);
        util::vector2d<float> raw = render_m(sc, cam, e2w3, 11);
        ICP icp;
        const float4x4 aligned = icp.Align(raw, e2w2, synth, e2w2, cam);
        check(icp.LastStats().iterations == 4 && !icp.LastStats().failed, "icp_align_ran_4_steps");
        dump("icp_synth.f32x3", &synth[0].x, synth.size() * 3);
        dump("icp_raw_m.f32", raw.data(), raw.size());
        dump("icp_guess.f32", e2w2.data(), 16);
        dump("icp_aligned.f32", aligned.data(), 16);
    }

    // ---- Pipeline: given poses (the reference's active branch), then tracking
    {
        kifi_scene sc;
        kifi_scene_make(KIFI_SCENE_ROOM, 2, &sc);
        Pipeline pipe(cam, 1024, 4.096f, 0.02f);
        double e2w[16];
        for (int f = 0; f < 3; ++f) {
            kifi_trajectory_pose(KIFI_TRAJ_CIRCLE, 2, f, e2w);
            pipe.Integrate(render_m(sc, cam, e2w, 10 + f), world_to_eye(e2w));
        }
        const float4x4 last = pipe.EyeToWorld();
        float err = 0.0f;
        for (int i = 0; i < 16; ++i)
            err = std::fmax(err, std::fabs(last.data()[i] - (float)e2w[i]));
        check(err < 1e-5f, "pipeline_eye_to_world_is_given_pose");
        VolumeData data = pipe.Volume().Data();
        dump("pipe_keys.u32", data.keys_cbegin(), data.size());
        dump("pipe_values.f32x2", data.values_cbegin(), data.size());
        for (int f = 3; f < 8; ++f) {
            kifi_trajectory_pose(KIFI_TRAJ_CIRCLE, 2, f, e2w);
            pipe.Integrate(render_m(sc, cam, e2w, 300 + f));
        }
        const float4x4 tracked = pipe.EyeToWorld();
        double dt = 0.0;
        for (int r = 0; r < 3; ++r)
            dt += (tracked(r, 3) - e2w[12 + r]) * (tracked(r, 3) - e2w[12 + r]);
        check(std::sqrt(dt) < 0.02, "pipeline_tracking_within_2cm");
        check(pipe.Stats().frames == 5, "pipeline_tracked_5_frames");
        std::vector<float3> verts;
        pipe.Mesh(verts);
        check(!verts.empty(), "pipeline_mesh_nonempty");
    }
    // ---- DepthStream + Integrator + Mesher on a .depth file (the reference's
    // UnitTests/Integrator.cpp:21-52 scenario: Volume(256, 2, 0.02), KinectParams
    // 640x480 far), when a file is given as argv[2]
    if (argc >= 3) {
        DepthStream ds(argv[2]);
        check(ds.Good() && ds.FrameWidth() == 640 && ds.FrameHeight() == 480, "depthstream_header");
        util::vector2d<float> depth;
        float4x4 view;
        const bool got = ds.NextFrame(depth, view);
        check(got, "depthstream_first_frame");
        if (got) {
            dump("ds_depth_m.f32", depth.data(), depth.size());
            dump("ds_view.f32", view.data(), 16);
            Volume vol(256, 2.0f, 0.02f);
            Integrator integrator;
            integrator.Integrate(vol, depth,
                                 DepthSensorParams::KinectParams(KinectDepthSensorResolution640x480,
                                                                 KinectDepthSensorModeFar),
                                 view);
            VolumeData data = vol.Data();
            check(data.size() > 1000, "depthstream_volume_nonempty");
            dump("ds_keys.u32", data.keys_cbegin(), data.size());
            dump("ds_values.f32x2", data.values_cbegin(), data.size());
            Mesher mesher;
            std::vector<float3> verts;
            mesher.Mesh(vol, verts);
            dump("ds_verts.f32x3", &verts[0].x, verts.size() * 3);
        }
        // the file holds no more frames than its header says, and none past its end
        int n = got ? 1 : 0;
        while (ds.NextFrame(depth, view))
            ++n;
        check(n <= ds.FrameCount(), "depthstream_frames_within_header");
        std::printf("depthstream frames read %d of %d\n", n, ds.FrameCount());
    }
    std::printf("DONE failures=%d\n", g_fail);
    return g_fail;
}
