"""GPU parity on the reference's own .depth fixtures and for frames in metres.

- imrod_v2.depth (one frame of binary32 metres, not on the millimetre grid)
  integrated as the reference's UnitTests/Integrator.cpp:21-52 does
  (Volume(256, 2, 0.02), KinectParams 640x480 far, the file's view matrix):
  keys, values and mesh bit-exact against the oracle;
- imrod_v1.depth (uint16 mm) the same way;
- a metre frame equal to mm * 0.001f gives the millimetre path's volume bit for
  bit, for kifi_integrate_f32 and for the pipeline (given pose and tracking).
"""
import numpy as np
import pytest

import merge_fusion_amd as mf
import oracle as orc
from test_depthstream import fixture

pytestmark = pytest.mark.gpu

KINECT = mf.Intrinsics(640, 480, 571.26, 571.26)     # DepthSensorParams.cpp:68-82, 640x480
KINECT.cx, KINECT.cy, KINECT.znear, KINECT.zfar = 320.0, 240.0, 0.8, 4.0
CAM = mf.Intrinsics(512, 424, 365.0, 365.0)


def _same_volume(gv, ov):
    gk, gvals = gv.download()
    ok, ovals = ov.read()
    assert np.array_equal(gk, ok)
    assert np.array_equal(np.asarray(gvals).view(np.uint32).ravel(), np.asarray(ovals).view(np.uint32).ravel())
    return len(gk)


@pytest.mark.parametrize("name", ["imrod_v2.depth", "imrod_v1.depth"])
def test_reference_fixture_integrates_bit_exact(ctx, tmp_path, name):
    with mf.DepthStream(fixture(tmp_path, name)) as ds:
        depth, view = ds.next_frame(metres=ds.texel == 1)
    gv = mf.Volume(ctx, 256, 2.0, 0.02)
    ov = orc.Volume(256, 2.0, 0.02)
    gv.integrate(depth, KINECT, view)
    ov.integrate(depth, KINECT, view)
    assert _same_volume(gv, ov) > 10000
    g, o = gv.mesh(), ov.mesh()
    assert len(g) > 1000 and np.array_equal(g.view(np.uint32), o.view(np.uint32))


def test_metre_frames_match_millimetre_frames(ctx):
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    frames, poses = [], []
    for f in range(6):
        pose = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 2 * f)
        frames.append(mf.render(sc, CAM, pose, noise_seed=40 + f))
        poses.append(pose)
    metres = [(d.astype(np.float32) * np.float32(0.001)).astype(np.float32) for d in frames]
    # Integrator
    a, b = mf.Volume(ctx, 1024, 4.096, 0.02), mf.Volume(ctx, 1024, 4.096, 0.02)
    for d, m, p in zip(frames, metres, poses):
        w2e = mf.invert_rigid(p).astype(np.float32)
        a.integrate(d, CAM, w2e)
        b.integrate(m, CAM, w2e)
    ka, va = a.download()
    kb, vb = b.download()
    assert np.array_equal(ka, kb) and np.array_equal(va.view(np.uint32), vb.view(np.uint32))
    # Pipeline: the first frame at its pose, the rest tracked
    out = []
    for fr in (frames, metres):
        pipe = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02)
        for i, d in enumerate(fr):
            pipe.integrate(d, mf.invert_rigid(poses[0]).astype(np.float32) if i == 0 else None)
        out.append((pipe.pose(), pipe.volume.download(), pipe.stats()["sum_iters"]))
        pipe.close()
    (p0, (k0, v0), it0), (p1, (k1, v1), it1) = out
    assert np.array_equal(p0, p1) and it0 == it1
    assert np.array_equal(k0, k1) and np.array_equal(v0.view(np.uint32), v1.view(np.uint32))
    assert mf.pose_error(p0, poses[-1])[0] < 0.02
