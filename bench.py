#!/usr/bin/env python3
"""Fused frames/s of the kifi fusion pipeline on one MI355X.

Workload (the trajectory configuration): 200 seeded 512x424 uint16 frames of a
procedural room (4 x 3 x 4 m box, 10 boxes, 6 spheres), camera on a circle of
r = 1.2 m at 1.5 cm / 1 deg per frame with jitter, noise N(0, (1.5e-3 z^2)^2) m
and 2 % invalid pixels; 4 mm voxels (res 1024, side 4.096 m, truncation 2 cm,
binary32 sum + weight, 32-bit keys); per frame: 3-level point-to-plane ICP
(10/5/4 iterations) against the model of the previous frame, integration at
the tracked pose, model prediction; a mesh (Mesher::Mesh, vertices copied to
page-locked host memory while the next frames run) at frames 49, 99, 149 and
199, all on the host before the step ends.

One step = the whole 200-frame trajectory from an empty volume (frame 0 at
its true pose, frames 1..199 tracked, the four meshes), so every step is the
workload as stated.  Frames are rendered on the host and copied to HBM before
timing; `value` = 200 x steps / the timed wall time.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]

Prints one JSON line.  Besides the headline it reports: the fused-frame
algorithmic GB/s against HBM peak and against a measured copy kernel; the
single-frame plain configuration (8 mm voxels, integrate + mesh); the host-frame
(PCIe-inclusive) rate; the CPU leg (the single-threaded oracle timed on one
pinned host core, which is also the checker of the GPU's errors against the
FP64 oracle on sampled frames and voxels); the learned-vs-plain comparison on
the 1000-frame fast trajectory (network trained here on this GPU); and the
2 mm / 64-bit-key precision variant at about 50 M voxels.
"""
import argparse
import csv
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fused frames/s at 512x424, 4 mm voxels; ICP iters/frame, learned vs plain init"
HBM_PEAK_GBS = 8000.0
W, H, F = 512, 424, 365.0
P = W * H
SEED = 2
RES, SIDE, TRUNC = 1024, 4.096, 0.02
S_K, S_V = 4, 8            # key and value bytes (32-bit keys, binary32 sum + weight)
TRAJ_FRAMES, MESH_EVERY = 200, 50
LEVEL_PIXELS = (P, P // 4, P // 16)   # ICP pyramid 512x424, 256x212, 128x106


def log(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print("[bench %s] %s" % (time.strftime("%H:%M:%S"), msg), file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed trajectories of 200 frames")
    ap.add_argument("--warmup", type=int, default=2, help="untimed trajectories before the timed ones")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU leg (baseline timing and FP64 checks)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-initialisation comparison")
    ap.add_argument("--fast-frames", type=int, default=1000, help="frames of the fast trajectory")
    ap.add_argument("--train-pairs", type=int, default=200000)
    ap.add_argument("--train-epochs", type=int, default=32)
    ap.add_argument("--open-loop-frames", type=int, default=300, help="fast-trajectory frames of the open loop")
    ap.add_argument("--no-precision", action="store_true", help="skip the 2 mm / 64-bit-key variant at scale")
    ap.add_argument("--precision-voxels", type=int, default=50_000_000,
                    help="active voxels the 2 mm volume is grown to before its timed frames")
    ap.add_argument("--precision-steps", type=int, default=20)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a plain copy kernel of this size 20 times (HBM counter calibration)")
    ap.add_argument("--pmc-window", metavar="COUNTER_CSV",
                    help="no GPU work: from a rocprofv3 --pmc counter collection of a run with the same --steps and "
                         "--warmup, each counter's mean over the roofline kernel's headline-pass timed launches, "
                         "and over the copy kernel's launches of a --calibrate-copy-mib run")
    ap.add_argument("--trace-window", metavar="KERNEL_TRACE_CSV",
                    help="no GPU work: from a rocprofv3 kernel trace of a run with the same --steps and --warmup, "
                         "the roofline kernel's launches of that run's roofline pass, timed window only")
    return ap.parse_args()


def render_frames(mf, cam, n):
    scene = mf.make_scene(mf.SCENE_ROOM, SEED)
    poses = [mf.trajectory_pose(mf.TRAJ_CIRCLE, SEED, f) for f in range(n)]
    frames = np.stack([mf.render(scene, cam, poses[f], noise_seed=1000 + f) for f in range(n)])
    return frames, poses


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


class PinnedCore:
    """Runs the enclosed block on one host core (sched_setaffinity of this
    process, restored afterwards): the reference path is single-threaded."""

    def __enter__(self):
        self.saved = os.sched_getaffinity(0)
        self.core = min(self.saved)
        os.sched_setaffinity(0, {self.core})
        return self

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.saved)
        return False


def cpu_baseline(mf, cam, frames, poses, budget_s):
    """The oracle pipeline (FP64 point-to-plane ICP 10/5/4, binary32 integration,
    model prediction) on the first frames of the trajectory, on one pinned
    host core, until the time budget is spent."""
    import oracle as orc
    with PinnedCore() as pin:
        vol = orc.Volume(RES, SIDE, TRUNC)
        w2e0 = mf.invert_rigid(poses[0]).astype(np.float32)
        vol.integrate(frames[0], cam, w2e0)
        pts, nrm = orc.model_predict(vol, w2e0, cam)
        model_w2e, pose = w2e0, poses[0]
        times = []
        t_start = time.perf_counter()
        f = 1
        while f < len(frames) and time.perf_counter() - t_start < budget_s:
            t0 = time.perf_counter()
            pose, _, _ = orc.icp_f64(frames[f], cam, pts, nrm, model_w2e.astype(np.float64), pose)
            w2e = mf.invert_rigid(pose).astype(np.float32)
            vol.integrate(frames[f], cam, w2e)
            pts, nrm = orc.model_predict(vol, w2e, cam)
            model_w2e = w2e
            times.append(time.perf_counter() - t0)
            f += 1
        core = pin.core
    ts = np.array(times)
    med = float(np.median(ts[3:] if len(ts) > 5 else ts))
    return {
        "value": round(1.0 / med, 4),
        "unit": "frames/s",
        "cores": 1,
        "kind": "port",
        "sample": "oracle pipeline (FP64 point-to-plane ICP 10/5/4, integration, model prediction) on frames "
                  "1..%d of the same trajectory (volume 0 -> %d voxels), median frame time %.3f s after 3 "
                  "warm-up frames; pinned to host core %d of %d (%s)"
                  % (f - 1, vol.size(), med, core, os.cpu_count() or 0, cpu_model()),
    }


def copy_rate_gbs(mf, ctx, mib=1024, launches=20):
    """Read + write GB/s of the library's plain 16-byte copy kernel on this GPU."""
    nbytes = mib << 20
    a = ctx.device_upload(np.zeros(nbytes // 4, np.uint32))
    b = ctx.device_upload(np.zeros(nbytes // 4, np.uint32))
    ctx.device_copy(b.address(), a.address(), nbytes)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(launches):
        ctx.device_copy(b.address(), a.address(), nbytes)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    a.free()
    b.free()
    return 2.0 * nbytes * launches / dt / 1e9


def int_bytes(d, s_k, s_v, frames):
    """Algorithmic bytes of `frames` integrations from the running-total deltas d:
    2P + 4 S_k K + 2 S_k E + (S_k + S_v)(V0 + 2 V1) + S_v A per frame."""
    return (2 * P * frames + 4 * s_k * d["keys"] + 2 * s_k * d["new"] +
            (s_k + s_v) * (d["v_old"] + 2 * d["v_new"]) + s_v * d["updated"])


def icp_bytes(level_iters, frames):
    """36 P_l per iteration at level l, plus 2P + sum_l 24 P_l per tracked frame
    (pyramid and normals)."""
    return (sum(36 * LEVEL_PIXELS[l] * level_iters[l] for l in range(3)) +
            frames * (2 * P + sum(24 * p for p in LEVEL_PIXELS)))


def headline(mf, ctx, cam, args, frames, dframes, poses):
    """Config-2 trajectories.  Passes: the headline (no events), the roofline
    kernel's launch durations (only its events; the profiler window), every
    stage timer, no meshing, and host frames (PCIe copy in each frame)."""
    fbytes = P * 2
    mesh_buf = [ctx.host_empty((1 << 22, 3), np.float32)]
    w2e0 = mf.invert_rigid(poses[0]).astype(np.float32)

    def trajectory(pipe, mesh=True, host=None, record=None):
        pipe.reset()
        meshes = []
        for f in range(TRAJ_FRAMES):
            w2e = w2e0 if f == 0 else None
            if host is not None:
                pipe.integrate(host[f], w2e)
            else:
                pipe.integrate_device(dframes.address(f * fbytes), w2e)
            if record is not None:
                record(f, pipe)
            if mesh and (f + 1) % MESH_EVERY == 0:
                # the vertex copy overlaps the next frames; every mesh is on the
                # host before the trajectory returns
                nv = pipe.volume.mesh_into_async(mesh_buf[0])
                if nv > len(mesh_buf[0]):
                    ctx.mesh_wait()
                    mesh_buf[0] = ctx.host_empty((int(nv * 1.5), 3), np.float32)
                    nv = pipe.volume.mesh_into_async(mesh_buf[0])
                meshes.append(int(nv))
        ctx.mesh_wait()
        return meshes

    def run_pass(stages, mesh=True, host=None):
        pipe = mf.Pipeline(ctx, cam, RES, SIDE, TRUNC)
        for _ in range(args.warmup):
            trajectory(pipe, mesh, host)
        ctx.synchronize()
        st0, tot0 = pipe.stats(), pipe.volume.totals()
        if stages != []:
            ctx.timing(True, stages)
        t0 = time.perf_counter()
        meshes = []
        for _ in range(args.steps):
            meshes = trajectory(pipe, mesh, host)
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
        ctx.timing(False)
        st1, tot1 = pipe.stats(), pipe.volume.totals()
        stage = {name: ctx.timing_read(s) for name, s in (("merge_update", mf.STAGE_MERGE_UPDATE),
                                                          ("sort", mf.STAGE_SORT), ("icp", mf.STAGE_ICP),
                                                          ("model", mf.STAGE_MODEL), ("mesh", mf.STAGE_MESH))}
        pipe.close()
        d = {k: tot1[k] - tot0[k] for k in tot1}
        icp = {"frames": st1["frames"] - st0["frames"], "sum_iters": st1["sum_iters"] - st0["sum_iters"],
               "fallbacks": st1["fallbacks"] - st0["fallbacks"],
               "level_iters": [st1["sum_level_iters"][l] - st0["sum_level_iters"][l] for l in range(3)]}
        return dict(elapsed=elapsed, tot=d, icp=icp, stage=stage, meshes=meshes)

    main = run_pass([])
    roof = run_pass([mf.STAGE_MERGE_UPDATE])
    diag = run_pass(None)
    nomesh = run_pass([], mesh=False)
    hframes = ctx.host_empty((TRAJ_FRAMES, H, W), np.uint16)
    hframes[:] = frames
    h2d = run_pass([], mesh=False, host=hframes)

    # one more trajectory, untimed: every frame's pose, the volume and the
    # model at sampled frames (for the FP64 checks of the CPU leg)
    quality = {"poses": [], "volume_at_mesh": [], "samples": {}}
    sample_frames = (50, 100, 150, 199)

    def record(f, pipe):
        quality["poses"].append(pipe.pose())
        if (f + 1) % MESH_EVERY == 0:
            quality["volume_at_mesh"].append(int(pipe.volume.size()))
        if f + 1 in sample_frames:
            pts, nrm = pipe.model.download()       # the model the next frame is tracked against
            quality["samples"][f + 1] = {"pts": pts, "nrm": nrm, "prev_pose": pipe.pose()}
        if f in sample_frames:
            quality["samples"][f]["pose"] = pipe.pose()
            quality["samples"][f]["start_pose"] = pipe.stats()["start_pose"]

    qpipe = mf.Pipeline(ctx, cam, RES, SIDE, TRUNC)
    qmeshes = trajectory(qpipe, True, None, record)
    qpipe.close()
    return dict(main=main, roof=roof, diag=diag, nomesh=nomesh, h2d=h2d, quality=quality, qmeshes=qmeshes)


def twistnet_flops_per_sample():
    """Multiply-adds x 2 of one forward pass (3x3 stride-2 convolutions, two FC layers)."""
    ch, h, w, fl = [2, 16, 32, 64, 64, 128], 106, 128, 0
    for l in range(5):
        h, w = (h - 1) // 2 + 1, (w - 1) // 2 + 1
        fl += 2 * h * w * ch[l + 1] * ch[l] * 9
    return fl + 2 * (2048 * 256 + 256 * 6)


# The fast trajectory's tracker settings, the same for the learned and the plain
# start: pyramid 10/6/4 iterations (20 in all, so the fallback's iteration
# trigger is the cap), gates 10/20/40 cm, the motion prior, and for the learned
# start the fallback to plain ICP from the previous pose.
C5_ITERS = (10, 6, 4)
C5_GATE_GROWTH = 2.0
C5_PRIOR = 1e-4
C5_FALLBACK_RMS = 0.005
C5_SCENE_SEED, C5_TRAJ_SEED = SEED, 3


def c5_params(mf, learned):
    prm = mf.IcpParams.default()
    prm.iters[0], prm.iters[1], prm.iters[2] = C5_ITERS
    prm.level_gate_growth = C5_GATE_GROWTH
    prm.motion_prior = C5_PRIOR
    if learned:
        prm.fallback_rms = C5_FALLBACK_RMS
        prm.fallback_iters = sum(C5_ITERS)
    return prm


def trajectory_errors(est, truth):
    """Per-frame errors of eye->world poses against the truth: absolute
    (m, deg), the absolute trajectory error after the rigid alignment of the
    positions (Horn / Umeyama without scale, the TUM RGB-D definition) and the
    relative pose error of consecutive frames."""
    E = [np.asarray(p, np.float64).reshape(4, 4).T for p in est]
    G = [np.asarray(p, np.float64).reshape(4, 4).T for p in truth]
    ab = np.array([pose_error_mats(a, g) for a, g in zip(E, G)])
    pe, pg = np.array([a[:3, 3] for a in E]), np.array([g[:3, 3] for g in G])
    me, mg = pe.mean(0), pg.mean(0)
    U, _, Vt = np.linalg.svd((pg - mg).T @ (pe - me))
    D = np.eye(3)
    D[2, 2] = np.sign(np.linalg.det(U @ Vt))
    R = U @ D @ Vt
    al = (R @ (pe - me).T).T + mg
    rel = np.array([pose_error_mats(np.linalg.inv(G[f - 1]) @ G[f], np.linalg.inv(E[f - 1]) @ E[f])
                    for f in range(1, len(E))])
    return ab, float(np.sqrt(np.mean(np.sum((al - pg) ** 2, 1)))), rel


def pose_error_mats(a, b):
    d = np.linalg.inv(b) @ a
    c = np.clip((np.trace(d[:3, :3]) - 1.0) * 0.5, -1.0, 1.0)
    return np.linalg.norm(d[:3, 3]), float(np.degrees(np.arccos(c)))


def net_input(d_mm):
    """The network's input channel of a frame: 4x4 block mean of the valid
    depths (m) / 4 m, as the pipeline's downsampling kernel forms it."""
    b = (d_mm.astype(np.float32) * np.float32(0.001)).reshape(H // 4, 4, W // 4, 4).transpose(0, 2, 1, 3)
    b = b.reshape(H // 4, W // 4, 16)
    c = (b > 0).sum(-1)
    return np.where(c > 0, b.sum(-1) / np.maximum(c, 1) * np.float32(0.25), 0.0).astype(np.float32)


def compose_twist(prev_e2w, tw):
    """prev * [R(omega) | v] in column storage: the start the pipeline derives
    from a predicted twist."""
    w, v = np.asarray(tw[:3], np.float64), np.asarray(tw[3:], np.float64)
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    R = np.eye(3) + (np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K if th > 1e-12 else K)
    T = np.eye(4)
    T[:3, :3], T[:3, 3] = R, v
    return (np.asarray(prev_e2w, np.float64).reshape(4, 4).T @ T).T.reshape(16)


def learned_section(mf, ctx, cam, args):
    """Learned vs plain initial pose on the fast trajectory (per-frame steps of
    U[2,10] cm horizontal translation and U[2,15] deg yaw, 1000 frames): the
    network is trained here, on this GPU, by the library's own forward/backward
    kernels on pairs drawn from the same motion law (seeded, bounded budget);
    then every comparison runs both starts under the same ICP schedule, gates
    and motion prior, on the same device-resident frames:
    - closed loop: the pipelines as they are (plain loses the camera);
    - fused at the true poses: every frame is tracked as in the closed loop
      but fused at its true pose (kifi_pipeline_integrate_at), and the next
      frame starts from that pose, so that both trackers work on equal volumes
      (the frames/s ratio on equal volumes);
    - open loop: ICP of frame f against a model integrated at the true poses
      of frames 0..f-1, also from the true pose (the floor of any start)."""
    scenes = [mf.make_scene(mf.SCENE_ROOM, 100 + k) for k in range(64)]
    law = dict(min_trans=0.02, max_trans=0.10, min_rot_deg=2.0, max_rot_deg=15.0, walk=True,
               max_pitch_step_deg=2.0)
    t0 = time.perf_counter()
    pairs = mf.PairSet(ctx, cam, scenes, args.train_pairs, seed=7, **law)
    ctx.synchronize()
    t_gen = time.perf_counter() - t0
    _, _, _, pvalid = pairs.download(inputs=False)
    net = mf.TwistNet(ctx, 1)
    spe = args.train_pairs // 256
    losses = []
    t0 = time.perf_counter()
    log("fast trajectory: %d training pairs generated in %.1f s; training %d epochs" % (
        args.train_pairs, t_gen, args.train_epochs))
    for ep in range(args.train_epochs):
        if ep % 8 == 0:
            log("  epoch %d" % ep)
        lr = 1e-3 * 0.5 * (1.0 + np.cos(np.pi * ep / args.train_epochs))
        losses.append(net.train(pairs, spe, 256, lr, seed=100 + ep))
    ctx.synchronize()
    t_train = time.perf_counter() - t0
    pairs.close()
    hold = mf.PairSet(ctx, cam, [mf.make_scene(mf.SCENE_ROOM, 500 + k) for k in range(8)], 1024, seed=2, **law)
    _, _, truth, valid = hold.download(inputs=False)
    pred = net.predict(hold)
    hold.close()
    rot_err = np.degrees(np.linalg.norm(pred[valid, :3] - truth[valid, :3], axis=1))
    tr_err = np.linalg.norm(pred[valid, 3:] - truth[valid, 3:], axis=1)
    samples = spe * 256 * args.train_epochs
    train_tflops = 3.0 * twistnet_flops_per_sample() * samples / t_train / 1e12

    nf = args.fast_frames
    scene = mf.make_scene(mf.SCENE_ROOM, C5_SCENE_SEED)
    poses = [mf.trajectory_pose(mf.TRAJ_FAST, C5_TRAJ_SEED, f) for f in range(nf)]
    fbytes = W * H * 2
    buf = ctx.device_upload(np.zeros(nf * W * H, np.uint16))
    for f in range(nf):
        mf.render_device(ctx, scene, cam, poses[f], 5000 + f, out_ptr=buf.address(f * fbytes))
    ctx.synchronize()
    w2e0 = mf.invert_rigid(poses[0]).astype(np.float32)
    samp_frames = set(range(100, nf, 200))
    c5_samples = {}

    def make(storage, learned):
        pipe = mf.Pipeline(ctx, cam, RES, SIDE, TRUNC, storage=storage, params=c5_params(mf, learned))
        if learned:
            pipe.set_twistnet(net)
        return pipe

    w2e_true = [mf.invert_rigid(p).astype(np.float32) for p in poses]

    def closed(storage, learned, at_truth, track):
        pipe = make(storage, learned)
        est, fb_frames = [], []
        ctx.synchronize()
        t0 = time.perf_counter()
        for f in range(nf):
            sample = track and learned and not at_truth and f in samp_frames
            if sample:
                pts, nrm = pipe.model.download()
                prev = pipe.pose()
            if at_truth:
                pipe.integrate_at(buf.address(f * fbytes), w2e_true[f], on_device=True)
            else:
                pipe.integrate_device(buf.address(f * fbytes), w2e0 if f == 0 else None)
            if track:
                est.append(pipe.pose())
                if f > 0:
                    st = pipe.stats()
                    if st["fallback"]:
                        fb_frames.append(f)
                    if sample:
                        c5_samples[f] = {"pts": pts, "nrm": nrm, "prev_pose": prev, "start_pose": st["start_pose"],
                                         "pose": est[-1], "fallback": st["fallback"]}
        ctx.synchronize()
        dt = time.perf_counter() - t0
        st = pipe.stats()
        pipe.close()
        r = {"icp_iters_per_frame": round(st["sum_iters"] / max(st["frames"], 1), 3),
             "icp_level_iters_per_frame": [round(x / max(st["frames"], 1), 3) for x in st["sum_level_iters"]],
             "fallback_rate": round(st["fallbacks"] / max(st["frames"], 1), 4)}
        if not track:
            r["frames_per_s"] = round((nf - 1) / dt, 2)
            return r
        ab, ate_al, rel = trajectory_errors(est, poses)
        over = np.nonzero(ab[:, 0] > 0.10)[0]
        r.update({
            "ate_m": round(float(np.sqrt(np.mean(ab[:, 0] ** 2))), 4),
            "ate_aligned_m": round(ate_al, 4),
            "max_error_m_deg": [round(float(ab[:, 0].max()), 4), round(float(ab[:, 1].max()), 3)],
            "first_frame_error_over_10cm": int(over[0]) if len(over) else None,
            "frames_error_over_10cm": int(len(over)),
            "rpe_rmse_mm_deg": [round(float(np.sqrt(np.mean(rel[:, 0] ** 2))) * 1e3, 3),
                                round(float(np.sqrt(np.mean(rel[:, 1] ** 2))), 4)],
            "frames_rpe_over_1cm_or_1deg": int(((rel[:, 0] > 0.01) | (rel[:, 1] > 1.0)).sum()),
        })
        if at_truth:
            r["frames_within_2cm_1deg"] = round(float(np.mean((ab[1:, 0] < 0.02) & (ab[1:, 1] < 1.0))), 4)
        if learned and not at_truth:
            r["_est"] = est
        return r

    res = {"closed_loop": {}, "fused_at_true_poses": {}}
    for name, storage, learned in (("plain_f32", mf.STORAGE_F32, False), ("plain_f16", mf.STORAGE_F16U8, False),
                                   ("learned_f16", mf.STORAGE_F16U8, True)):
        for mode, at_truth in (("closed_loop", False), ("fused_at_true_poses", True)):
            timed = closed(storage, learned, at_truth, False)
            tracked = closed(storage, learned, at_truth, True)
            res[mode][name] = dict(tracked, frames_per_s=timed["frames_per_s"],
                                   deterministic=timed["icp_iters_per_frame"] == tracked["icp_iters_per_frame"])
    est_learned = res["closed_loop"]["learned_f16"].pop("_est")

    # open loop: the model integrated at the true poses of frames 0..f-1
    n_ol = min(nf, args.open_loop_frames)
    host = [mf.render_device(ctx, scene, cam, poses[f], 5000 + f) for f in range(n_ol)]
    prm = c5_params(mf, False)
    small = [net_input(h) for h in host]
    x = np.stack([np.stack([small[f - 1], small[f]]) for f in range(1, n_ol)])
    tw = np.concatenate([net.forward(x[k:k + 256]) for k in range(0, len(x), 256)]) * mf.TWIST_SCALE
    pipe = mf.Pipeline(ctx, cam, RES, SIDE, TRUNC, storage=mf.STORAGE_F16U8, params=prm)
    out = {k: ([], []) for k in ("true_pose", "learned", "previous_pose")}
    pred_err = []
    for f in range(1, n_ol):
        pipe.integrate(host[f - 1], mf.invert_rigid(poses[f - 1]).astype(np.float32))
        lg = compose_twist(poses[f - 1], tw[f - 1])
        pred_err.append(mf.pose_error(lg, poses[f]))
        for name, guess in (("true_pose", poses[f]), ("learned", lg), ("previous_pose", poses[f - 1])):
            pose, st = pipe.model.icp_align(host[f], guess, prm)
            out[name][0].append(st["iterations"])
            out[name][1].append(mf.pose_error(pose, poses[f]))
    pipe.close()
    ol = {"frames": n_ol - 1}
    for name, (its, errs) in out.items():
        its, errs = np.array(its), np.array(errs)
        ol[name] = {"icp_iters_per_frame": round(float(its.mean()), 3),
                    "frames_within_2cm_1deg": round(float(np.mean((errs[:, 0] < 0.02) & (errs[:, 1] < 1.0))), 4),
                    "median_error_cm_deg": [round(float(np.median(errs[:, 0])) * 100, 3),
                                            round(float(np.median(errs[:, 1])), 3)]}
    pe = np.array(pred_err)
    ol["prediction_error"] = {"mean_cm_deg": [round(float(pe[:, 0].mean()) * 100, 3), round(float(pe[:, 1].mean()), 3)],
                              "p90_cm_deg": [round(float(np.percentile(pe[:, 0], 90)) * 100, 3),
                                             round(float(np.percentile(pe[:, 1], 90)), 3)]}
    ol["icp_iters_reduction"] = round(1.0 - ol["learned"]["icp_iters_per_frame"] /
                                      ol["previous_pose"]["icp_iters_per_frame"], 4)
    ol["icp_iters_reduction_bound"] = round(1.0 - ol["true_pose"]["icp_iters_per_frame"] /
                                            ol["previous_pose"]["icp_iters_per_frame"], 4)
    buf.free()
    net.close()
    cl, ra = res["closed_loop"], res["fused_at_true_poses"]

    def ratio(d, a, b, k):
        return round(d[a][k] / d[b][k], 3)

    return {
        "workload": "fast trajectory (per-frame steps U[2,10] cm horizontal, U[2,15] deg yaw, pitch +-2 deg), "
                    "%d frames of the room scene, 512x424, 4 mm voxels; both starts: point-to-plane ICP "
                    "%d/%d/%d (20 iterations) with %d/%d/%d cm gates and motion prior %g per inlier; learned: CNN "
                    "twist as the start, fallback to plain ICP from the previous pose when the 1 m-scaled "
                    "residual > %g mm or the first pass used all 20 iterations" % (
                        (nf,) + C5_ITERS + tuple(int(10 * C5_GATE_GROWTH ** l) for l in range(3)) +
                        (C5_PRIOR, C5_FALLBACK_RMS * 1e3)),
        "training": {
            "pairs": args.train_pairs, "pairs_valid_frac": round(float(pvalid.mean()), 4),
            "motion_law": "walk steps: translation U[2,10] cm horizontal, yaw +-U[2,15] deg, pitch +-U[0,2] deg",
            "epochs": args.train_epochs, "batch": 256, "optimizer": "adam, cosine lr 1e-3",
            "pair_generation_s": round(t_gen, 2), "train_s": round(t_train, 2),
            "samples_per_s": round(samples / t_train, 1), "tflops": round(train_tflops, 3),
            "mfma_bf16_peak_tflops": 2500.0, "mfma_frac": round(train_tflops / 2500.0, 6),
            "loss_first_last": [round(float(losses[0][:20].mean()), 4), round(float(losses[-1][-20:].mean()), 4)],
            "heldout_rot_err_deg_mean_p90": [round(float(rot_err.mean()), 3), round(float(np.percentile(rot_err, 90)), 3)],
            "heldout_trans_err_cm_mean_p90": [round(float(tr_err.mean() * 100), 3),
                                              round(float(np.percentile(tr_err, 90) * 100), 3)],
            "heldout_prev_pose_rot_err_deg_mean": round(float(np.degrees(np.linalg.norm(truth[valid, :3],
                                                                                          axis=1)).mean()), 3),
            "heldout_prev_pose_trans_err_cm_mean": round(float(np.linalg.norm(truth[valid, 3:],
                                                                              axis=1).mean() * 100), 3),
        },
        "closed_loop": cl,
        "fused_at_true_poses": ra,
        "open_loop": ol,
        "fps_ratio_learned_f16_vs_plain_f32": {"closed_loop": ratio(cl, "learned_f16", "plain_f32", "frames_per_s"),
                                               "fused_at_true_poses": ratio(ra, "learned_f16", "plain_f32",
                                                                            "frames_per_s")},
        "icp_iters_reduction_vs_plain_f32": {
            "closed_loop": round(1.0 - ratio(cl, "learned_f16", "plain_f32", "icp_iters_per_frame"), 4),
            "fused_at_true_poses": round(1.0 - ratio(ra, "learned_f16", "plain_f32", "icp_iters_per_frame"), 4),
            "open_loop": ol["icp_iters_reduction"]},
        "_samples": c5_samples, "_est": est_learned, "_scene_traj": (C5_SCENE_SEED, C5_TRAJ_SEED),
    }


def precision_section(mf, ctx, cam, args):
    """The precision variant at scale: the plane + spheres scene at 2 mm voxels
    (res 4096, side 8.192 m, truncation 4 cm, 64-bit keys), binary16 + u8 and
    binary32 storage.  Each volume is first grown to args.precision_voxels
    active voxels by integrating seeded views about the identity pose (rendered
    on the GPU); the timed steps then integrate the identity view, fresh noise
    each step, into it: merge + TSDF update over tens of millions of voxels.
    Frames are resident in HBM before timing."""
    sc = mf.make_scene(mf.SCENE_PLANE_SPHERES, 1)
    res, side, trunc, sk = 4096, 8.192, 0.04, 8
    fbytes = W * H * 2
    K = args.precision_steps
    ident = np.eye(4, dtype=np.float32).reshape(16)
    timed = ctx.device_upload(np.zeros((K, H, W), np.uint16))
    grow = ctx.device_upload(np.zeros((H, W), np.uint16))
    for k in range(K):
        mf.render_device(ctx, sc, cam, np.eye(4), 7000 + k, out_ptr=timed.address(k * fbytes))
    out = {"workload": "plane + 3 spheres, 512x424, 2 mm voxels (res 4096, side 8.192 m, trunc 4 cm), 64-bit keys; "
                       "volume grown by seeded views about the identity, then %d identity views integrated" % K}
    for name, storage, sv in (("f16_u8", mf.STORAGE_F16U8, 4), ("f32", mf.STORAGE_F32, 8)):
        vol = mf.Volume(ctx, res, side, trunc, 64, storage)
        f = 0
        while f < 400:
            pose = mf.trajectory_pose(mf.TRAJ_SWAY, 1, f)
            mf.render_device(ctx, sc, cam, pose, 100 + f, out_ptr=grow.address(0))
            vol.integrate_device(grow.address(0), cam, mf.invert_rigid(pose).astype(np.float32))
            f += 1
            if f % 8 == 0 and vol.size() >= args.precision_voxels:
                break
        ctx.synchronize()
        v_start = int(vol.size())
        tot0 = vol.totals()
        ctx.timing_read(mf.STAGE_MERGE_UPDATE)
        ctx.timing(True, [mf.STAGE_MERGE_UPDATE])
        t0 = time.perf_counter()
        for k in range(K):
            vol.integrate_device(timed.address(k * fbytes), cam, ident)
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
        ctx.timing(False)
        tot1 = vol.totals()
        mu_ms, mu_n = ctx.timing_read(mf.STAGE_MERGE_UPDATE)
        launches = tot1["frames"] - tot0["frames"]
        dv0, de, dv1 = (tot1[k] - tot0[k] for k in ("v_old", "new", "v_new"))
        alg = ((sk + sv) * (dv0 + dv1) + sk * de + 2 * W * H * launches) / max(launches, 1)
        gbs = alg / (mu_ms / mu_n * 1e-3) / 1e9 if mu_n else 0.0
        out[name] = {
            "voxels_before": v_start, "voxels_after": int(vol.size()), "growth_frames": f,
            "ms_per_frame": round(1000.0 * elapsed / K, 4),
            "merge_update_ms": round(mu_ms / mu_n, 4) if mu_n else None,
            "alg_bytes_per_launch": int(alg),
            "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
        }
        vol.close()
    timed.free()
    grow.free()
    return out


ROOF_KERNEL = "k_merge_update<unsigned int, kifi::VoxF32, true>"


def trace_window(args):
    """The headline's launches in a rocprofv3 kernel trace of a bench run with
    the same --steps and --warmup.  Every frame of the headline passes launches
    the roofline kernel's tracked binary32 instance once, and nothing before
    them does: pass 1 (no events) is launches [0, n), pass 2 (the roofline
    pass) is [n, 2n) with n = 200 (warmup + steps), and their timed windows are
    the last 200 x steps of each.  Prints the roofline kernel's mean duration
    over pass 2's window (the figure the bench's avg_launch_ms measures with
    HIP events) and, for pass 1's window (the headline value's own frames),
    every kernel's calls and mean duration per frame, the kernel time per frame
    and the span per frame (from the end of the roofline launch before the
    window to the end of the window's last one)."""
    rows = []
    with open(args.trace_window, newline="") as fh:
        for r in csv.DictReader(fh):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    roof = [(s_, e) for s_, e, k in rows if k.startswith("void kifi::" + ROOF_KERNEL) or k.startswith(ROOF_KERNEL)]
    n = TRAJ_FRAMES * (args.warmup + args.steps)
    lo = TRAJ_FRAMES * args.warmup
    win = roof[n + lo: 2 * n]
    dur = np.array([e - s_ for s_, e in win], np.float64) * 1e-6
    out = {"kernel": ROOF_KERNEL, "launches_in_trace": len(roof), "window_launches": len(win),
           "window": "[%d, %d)" % (n + lo, 2 * n),
           "mean_ms": round(float(dur.mean()), 5) if len(dur) else None,
           "median_ms": round(float(np.median(dur)), 5) if len(dur) else None}
    if len(roof) >= n and lo > 0:
        t0, t1 = roof[lo - 1][1], roof[n - 1][1]
        frames = n - lo
        per, busy, nk = {}, 0, 0
        for s_, e, k in rows:
            if t0 < e <= t1:
                name = k.split("(")[0].replace("void ", "")
                c = per.setdefault(name, [0, 0])
                c[0] += 1
                c[1] += e - s_
                busy += e - s_
                nk += 1
        out["headline_pass_window"] = {
            "frames": frames, "span_us_per_frame": round((t1 - t0) * 1e-3 / frames, 2),
            "kernel_us_per_frame": round(busy * 1e-3 / frames, 2),
            "launches_per_frame": round(nk / frames, 2),
            "kernels": {k: {"calls_per_frame": round(c / frames, 3), "avg_us": round(t * 1e-3 / c, 2),
                            "us_per_frame": round(t * 1e-3 / frames, 2)}
                        for k, (c, t) in sorted(per.items(), key=lambda kv: -kv[1][1])}}
    # the 2 mm variant's timed launches: the last --precision-steps of each 64-bit-key instance
    tails = {}
    for s_, e, k in rows:
        if "k_merge_update<unsigned long" in k:
            tails.setdefault(k.split("(")[0].replace("void ", ""), []).append(e - s_)
    if tails:
        out["precision_timed_merge_update_us"] = {
            k: round(float(np.mean(v[-args.precision_steps:])) * 1e-3, 2) for k, v in sorted(tails.items())}
    print(json.dumps(out))


def pmc_window(args):
    """Counter means of a rocprofv3 --pmc collection: the roofline kernel's
    launches of the headline pass's timed trajectories (as trace_window picks
    them: pass 1 is the first TRAJ_FRAMES x (warmup + steps) launches), and the
    device copy kernel's launches (a --calibrate-copy-mib run)."""
    vals = {}
    with open(args.pmc_window, newline="") as fh:
        for r in csv.DictReader(fh):
            name = r["Kernel_Name"]
            kind = "roof" if ROOF_KERNEL in name else "copy" if "k_copy16" in name else None
            if kind is None:
                continue
            key = (kind, r["Counter_Name"])
            vals.setdefault(key, {})
            d = int(r["Dispatch_Id"])
            vals[key][d] = vals[key].get(d, 0.0) + float(r["Counter_Value"])
    out = {}
    for (kind, counter), per in sorted(vals.items()):
        v = np.array([per[d] for d in sorted(per)], np.float64)
        if kind == "roof":
            lo, hi = TRAJ_FRAMES * args.warmup, TRAJ_FRAMES * (args.warmup + args.steps)
            v = v[lo:hi]
        out["%s_%s" % (kind, counter)] = {"launches": int(len(v)), "mean": float(v.mean()) if len(v) else None}
    print(json.dumps(out))


def calibrate_copy(mf, mib):
    ctx = mf.Context(0)
    nbytes = mib << 20
    a = ctx.device_upload(np.zeros(nbytes // 4, np.uint32))
    b = ctx.device_upload(np.zeros(nbytes // 4, np.uint32))
    for _ in range(20):
        ctx.device_copy(b.address(), a.address(), nbytes)
    ctx.synchronize()
    print(json.dumps({"calibrate_copy_bytes": nbytes, "launches": 20}))


C1_RES, C1_SIDE, C1_TRUNC = 1024, 8.192, 0.04


def config1_frame(mf, cam):
    """The plain single-frame configuration: plane at 2.5 m + 3 spheres, identity pose."""
    scene = mf.make_scene(mf.SCENE_PLANE_SPHERES, 1)
    pose = mf.trajectory_pose(mf.TRAJ_STATIC, 1, 0)
    return mf.render(scene, cam, pose, noise_seed=7), mf.invert_rigid(pose).astype(np.float32)


def config1_section(mf, ctx, cam, steps=50):
    """One frame integrated into an empty 8 mm volume (res 1024, side 8.192 m,
    truncation 4 cm, binary32) and meshed (Mesher::Mesh, vertices to page-locked
    host memory): Integrator.cpp:26-71 + Mesher.cpp:14-17.  A step clears the
    volume, integrates the HBM-resident frame and meshes."""
    depth, w2e = config1_frame(mf, cam)
    dframe = ctx.device_upload(depth)
    vol = mf.Volume(ctx, C1_RES, C1_SIDE, C1_TRUNC)
    buf = ctx.host_empty((1 << 21, 3), np.float32)

    def step(mesh):
        vol.clear()
        vol.integrate_device(dframe.address(), cam, w2e)
        return vol.mesh_into(buf) if mesh else 0

    for _ in range(3):
        nv = step(True)
    res = {}
    for name, mesh in (("integrate_mesh", True), ("integrate", False)):
        ctx.synchronize()
        tot0 = vol.totals()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(mesh)
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / steps
        d = {k: v - tot0[k] for k, v in vol.totals().items()}
        nvox = int(vol.size())
        b = int_bytes(d, S_K, S_V, steps) / steps + ((S_K + S_V) * nvox + 12 * nv if mesh else 0)
        res[name] = {"frames_per_s": round(1.0 / dt, 2), "ms_per_frame": round(1e3 * dt, 4),
                     "alg_bytes": int(b), "GBs": round(b / dt / 1e9, 2),
                     "frac_hbm_peak": round(b / dt / 1e9 / HBM_PEAK_GBS, 5)}
    keys, vals = vol.download()
    verts = buf[:nv].copy()
    vol.close()
    dframe.free()
    return {"workload": "one 512x424 frame of plane + 3 spheres at the identity pose, 8 mm voxels (res 1024, "
                        "side 8.192 m, trunc 4 cm), binary32 sum + weight, integrate into an empty volume + "
                        "Mesher::Mesh", "voxels": len(keys), "mesh_vertices": int(nv), **res,
            "_result": (depth, w2e, keys, vals, verts)}


def tsdf_error(gk, gd, gw, ok, ov):
    """Distance error of a GPU volume (keys gk, distances gd, weights gw) against
    an FP64 oracle volume (keys ok, values ov = [sum, weight] doubles), on the
    voxels where both have the same positive weight.  A voxel whose centre
    projects within rounding of a pixel edge reads the neighbouring pixel in one
    precision ("pixel choice": the depth difference of two pixels, an error
    above 0.1 mm, ten binary16 ulps at the truncation distance); the others
    give the arithmetic error."""
    if not np.array_equal(gk.astype(np.uint64), ok):
        return {"keys_equal": False}
    w64 = ov[:, 1]
    m = (w64 > 0) & (gw == w64)
    e = np.abs(gd[m] - ov[m, 0] / w64[m])
    choice = e > 1e-4
    a = e[~choice]
    return {"keys_equal": True, "voxels": int(len(gk)), "weight_equal_frac": round(float(np.mean(gw == w64)), 6),
            "pixel_choice_frac": round(float(choice.mean()), 6), "max_m": float(a.max()),
            "p99_m": float(np.quantile(a, 0.99)), "mean_m": float(a.mean()),
            "all_voxels_p9999_m": float(np.quantile(e, 0.9999))}


def cpu_leg(mf, ctx, cam, args, frames, poses, head, c1, learned, precision):
    """The CPU side of the bench, on one pinned host core: the oracle timed
    (the CPU baseline), and the oracle as the checker of the GPU's results:
    config 1 bit-exactness, and errors against FP64 (poses of sampled tracked
    frames against FP64 ICP from the same start and model; TSDF distances
    against an FP64 volume fed the same frames and poses)."""
    import oracle as orc
    out = {"cpu_baseline": cpu_baseline(mf, cam, frames, poses, args.cpu_seconds)}
    with PinnedCore():
        # config 1: the oracle's integrate + mesh, timed and compared bit for bit
        depth, w2e, gk, gv, gverts = c1["_result"]
        t0 = time.perf_counter()
        ov = orc.Volume(C1_RES, C1_SIDE, C1_TRUNC)
        ov.integrate(depth, cam, w2e)
        om = ov.mesh()
        dt = time.perf_counter() - t0
        ok, ovals = ov.read()
        g = np.stack([gv["sum"], gv["weight"]], 1)
        c1["cpu_oracle"] = {"frames_per_s": round(1.0 / dt, 3), "cores": 1}
        c1["bit_exact_vs_oracle"] = {"keys": bool(np.array_equal(gk.astype(np.uint64), ok)),
                                     "values": bool(np.array_equal(g.view(np.uint32), ovals.view(np.uint32))),
                                     "mesh": bool(np.array_equal(gverts.view(np.uint32), om.view(np.uint32)))}
        # config 2: sampled tracked frames against FP64 ICP from the same start and model
        q = head["quality"]
        pose_checks = []
        for f, smp in sorted(q["samples"].items()):
            if "pose" not in smp:
                continue
            model_w2e = orc.invert_transform(np.asarray(smp["prev_pose"], np.float64).astype(np.float32))
            opose, oit, _ = orc.icp_f64(frames[f], cam, smp["pts"], smp["nrm"], model_w2e.astype(np.float64),
                                        smp["start_pose"])
            dt_m, dr = mf.pose_error(smp["pose"], opose)
            pose_checks.append([f, round(dt_m * 1e3, 5), round(dr, 6), int(oit)])
        # config 2: TSDF of the first frames at the GPU-tracked poses, binary32 GPU vs FP64 oracle
        n_t = 20
        gvol = mf.Volume(ctx, RES, SIDE, TRUNC)
        o64 = orc.Volume(RES, SIDE, TRUNC, 10, orc.STORE_F64)
        for f in range(n_t):
            w = (mf.invert_rigid(poses[0]).astype(np.float32) if f == 0 else
                 orc.invert_transform(np.asarray(q["poses"][f], np.float64).astype(np.float32)))
            gvol.integrate(frames[f], cam, w)
            o64.integrate(frames[f], cam, w)
        gk2, gv2 = gvol.download()
        gvol.close()
        ok2, ov2 = o64.read()
        del o64
        with np.errstate(invalid="ignore", divide="ignore"):
            gd2 = gv2["sum"].astype(np.float64) / gv2["weight"]
        err2 = tsdf_error(gk2, gd2, gv2["weight"].astype(np.float64), ok2, ov2)
        c5 = c3 = None
        if learned is not None:
            c5 = config5_fp64(mf, ctx, cam, orc, learned)
        if precision is not None:
            c3 = config3_fp64(mf, ctx, cam, orc)
    out["error_vs_fp64"] = {
        "config2_pose_mm_deg": {"frames": pose_checks,
                                "note": "[frame, mm, deg, FP64 iterations]: GPU pose of the tracked frame against "
                                        "FP64 point-to-plane ICP (oracle) from the GPU's start pose against the "
                                        "GPU's model maps; tolerance 1 mm / 0.1 deg"},
        "config2_tsdf_f32": dict(err2, frames=n_t, note="frames 0..%d at the GPU-tracked poses into a binary32 "
                                                         "GPU volume and an FP64 oracle volume" % (n_t - 1)),
    }
    if c5 is not None:
        out["error_vs_fp64"].update(c5)
    if c3 is not None:
        out["error_vs_fp64"].update(c3)
    return out


def f16u8_split(v):
    """(distance, weight) of binary16 + u8 voxels packed as half | weight << 16."""
    v = np.asarray(v, np.uint32)
    return (v & 0xffff).astype(np.uint16).view(np.float16).astype(np.float64), ((v >> 16) & 0xff).astype(np.float64)


def config5_fp64(mf, ctx, cam, orc, learned):
    """Fast trajectory, learned start: sampled tracked frames against FP64 ICP
    from the same start, model and settings (prior included); the binary16 + u8
    TSDF of the first frames at the tracked poses against an FP64 volume."""
    scene_seed, traj_seed = learned["_scene_traj"]
    scene = mf.make_scene(mf.SCENE_ROOM, scene_seed)
    checks = []
    for f, smp in sorted(learned["_samples"].items()):
        if smp["fallback"]:
            continue
        depth = mf.render_device(ctx, scene, cam, mf.trajectory_pose(mf.TRAJ_FAST, traj_seed, f), 5000 + f)
        model_w2e = orc.invert_transform(np.asarray(smp["prev_pose"], np.float64).astype(np.float32))
        opose, oit, _ = orc.icp_f64(depth, cam, smp["pts"], smp["nrm"], model_w2e.astype(np.float64),
                                    smp["start_pose"], iters=C5_ITERS, gate_growth=C5_GATE_GROWTH,
                                    motion_prior=C5_PRIOR)
        dt_m, dr = mf.pose_error(smp["pose"], opose)
        checks.append([f, round(dt_m * 1e3, 5), round(dr, 6), int(oit)])
    n_t = 20
    est = learned["_est"]
    gvol = mf.Volume(ctx, RES, SIDE, TRUNC, 32, mf.STORAGE_F16U8)
    o64 = orc.Volume(RES, SIDE, TRUNC, 10, orc.STORE_F64)
    for f in range(n_t):
        pose = mf.trajectory_pose(mf.TRAJ_FAST, traj_seed, f) if f == 0 else est[f]
        depth = mf.render_device(ctx, scene, cam, mf.trajectory_pose(mf.TRAJ_FAST, traj_seed, f), 5000 + f)
        w = orc.invert_transform(np.asarray(pose, np.float64).astype(np.float32))
        gvol.integrate(depth, cam, w)
        o64.integrate(depth, cam, w)
    gk, gv = gvol.download()
    gvol.close()
    ok, ov = o64.read()
    del o64
    gd, gw = f16u8_split(gv)
    return {
        "config5_pose_mm_deg": {"frames": checks,
                                "note": "[frame, mm, deg, FP64 iterations]: learned-start GPU pose against FP64 "
                                        "point-to-plane ICP (oracle, same motion prior, schedule and gates) from "
                                        "the GPU's start pose against the GPU's model maps; tolerance 1 mm / "
                                        "0.1 deg; frames that fell back are skipped"},
        "config5_tsdf_f16u8": dict(tsdf_error(gk, gd, gw, ok, ov), frames=n_t,
                                   note="fast-trajectory frames 0..%d at the learned pipeline's tracked poses into "
                                        "a binary16 + u8 GPU volume and an FP64 oracle volume" % (n_t - 1)),
    }


def config3_fp64(mf, ctx, cam, orc, n_t=4):
    """Precision variant (2 mm voxels, 64-bit keys): the first seeded views of
    its growth phase into binary32 and binary16 + u8 GPU volumes and an FP64
    oracle volume."""
    sc = mf.make_scene(mf.SCENE_PLANE_SPHERES, 1)
    res, side, trunc = 4096, 8.192, 0.04
    frames = []
    for f in range(n_t):
        pose = mf.trajectory_pose(mf.TRAJ_SWAY, 1, f)
        frames.append((mf.render_device(ctx, sc, cam, pose, 100 + f), mf.invert_rigid(pose).astype(np.float32)))
    o64 = orc.Volume(res, side, trunc, 21, orc.STORE_F64)
    for d, w in frames:
        o64.integrate(d, cam, w)
    ok, ov = o64.read()
    del o64
    out = {}
    for name, storage in (("f32", mf.STORAGE_F32), ("f16u8", mf.STORAGE_F16U8)):
        vol = mf.Volume(ctx, res, side, trunc, 64, storage)
        for d, w in frames:
            vol.integrate(d, cam, w)
        gk, gv = vol.download()
        vol.close()
        if storage == mf.STORAGE_F32:
            with np.errstate(invalid="ignore", divide="ignore"):
                gd = gv["sum"].astype(np.float64) / gv["weight"]
            gw = gv["weight"].astype(np.float64)
        else:
            gd, gw = f16u8_split(gv)
        out["config3_tsdf_" + name] = dict(tsdf_error(gk, gd, gw, ok, ov), frames=n_t,
                                           note="plane + spheres, 2 mm voxels, 64-bit keys: the first %d seeded "
                                                "views of the growth phase, GPU vs FP64 oracle volume" % n_t)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: the fusion path runs on one GPU (it is not sharded across GPUs); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.trace_window:
        trace_window(args)
        return
    if args.pmc_window:
        pmc_window(args)
        return
    import merge_fusion_amd as mf
    if args.calibrate_copy_mib:
        calibrate_copy(mf, args.calibrate_copy_mib)
        return
    cam = mf.Intrinsics(W, H, F, F)
    frames, poses = render_frames(mf, cam, TRAJ_FRAMES)

    ctx = mf.Context(0)
    dframes = ctx.device_upload(frames)
    log("headline: %d + %d trajectories of %d frames, 5 passes" % (args.warmup, args.steps, TRAJ_FRAMES))
    head = headline(mf, ctx, cam, args, frames, dframes, poses)
    log("headline done")
    copy_gbs = copy_rate_gbs(mf, ctx)
    c1 = config1_section(mf, ctx, cam)
    log("config 1 done")
    learned = None if args.no_learned else learned_section(mf, ctx, cam, args)
    log("fast trajectory done")
    precision = None if args.no_precision else precision_section(mf, ctx, cam, args)
    log("2 mm variant done")
    cpu = None if args.no_cpu else cpu_leg(mf, ctx, cam, args, frames, poses, head, c1, learned, precision)
    log("CPU leg done")
    for k in [k for k in (learned or {}) if k.startswith("_")]:
        del learned[k]
    c1.pop("_result")

    main_p, roof, diag = head["main"], head["roof"], head["diag"]
    frames_timed = TRAJ_FRAMES * args.steps
    fps = frames_timed / main_p["elapsed"]
    stage = dict(diag["stage"])
    stage["merge_update"] = roof["stage"]["merge_update"]

    # roofline of the dominant kernel: fused merge + TSDF update.  Algorithmic
    # bytes per launch = read V0 keys+values + E new keys, write V1 keys+values,
    # read the depth frame once.
    d = roof["tot"]
    launches = d["frames"]
    alg_bytes = (S_K + S_V) * (d["v_old"] + d["v_new"]) + S_K * d["new"] + 2 * P * launches
    mu_ms, mu_n = stage["merge_update"]
    achieved = alg_bytes / launches / (mu_ms / mu_n * 1e-3) / 1e9 if mu_n else 0.0
    # HBM traffic per launch: the measured traffic / algorithmic ratio of this
    # kernel on this workload (rocprofv3 FETCH_SIZE / WRITE_SIZE, copy-kernel
    # calibrated) x this run's algorithmic bytes
    traffic, traffic_note = None, None
    pmc_path = os.path.join(ROOT, "profiles", "merge_update_pmc.json")
    if os.path.exists(pmc_path) and launches:
        with open(pmc_path) as fh:
            pmc = json.load(fh)
        if pmc.get("traffic_over_alg"):
            traffic = int(pmc["traffic_over_alg"] * alg_bytes / launches)
            traffic_note = "measured traffic/algorithmic ratio %.3f (%s) x this run's algorithmic bytes" % (
                pmc["traffic_over_alg"], pmc.get("source"))

    # fused-frame algorithmic bytes of the headline pass: integration + ICP + meshes
    dm, im = main_p["tot"], main_p["icp"]
    b_int = int_bytes(dm, S_K, S_V, dm["frames"])
    b_icp = icp_bytes(im["level_iters"], im["frames"])
    q = head["quality"]
    b_mesh = args.steps * sum((S_K + S_V) * v + 12 * nv for v, nv in zip(q["volume_at_mesh"], head["qmeshes"]))
    fused_gbs = (b_int + b_icp + b_mesh) / main_p["elapsed"] / 1e9

    poses_gpu = q["poses"]
    errs = np.array([mf.pose_error(poses_gpu[f], poses[f]) for f in range(TRAJ_FRAMES)])
    icp_frames = max(im["frames"], 1)
    out = {
        "metric": METRIC,
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * main_p["elapsed"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "room trajectory, 512x424, 4 mm voxels (res 1024, side 4.096 m, trunc 2 cm), "
                        "3-level point-to-plane ICP 10/5/4, mesh every 50 frames, plain init; one step = the "
                        "whole 200-frame trajectory from an empty volume",
            "frames": TRAJ_FRAMES,
            "frames_per_step": TRAJ_FRAMES,
            "meshes_per_step": TRAJ_FRAMES // MESH_EVERY,
            "key_bits": 32,
            "storage": "f32 sum + f32 weight",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_merge_update",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_note": traffic_note,
            "alg_bytes_per_launch": int(alg_bytes / max(launches, 1)),
            "avg_launch_ms": round(mu_ms / mu_n, 5) if mu_n else None,
            "launches": mu_n,
        },
        "cpu_baseline": cpu["cpu_baseline"] if cpu else None,
        "fused_frame": {
            "alg_bytes_per_frame": int((b_int + b_icp + b_mesh) / frames_timed),
            "alg_bytes_split": {"integrate": int(b_int / frames_timed), "icp": int(b_icp / frames_timed),
                                "mesh": int(b_mesh / frames_timed)},
            "GBs": round(fused_gbs, 2),
            "frac_hbm_peak": round(fused_gbs / HBM_PEAK_GBS, 4),
            "copy_kernel_GBs": round(copy_gbs, 1),
            "frac_copy_kernel": round(fused_gbs / copy_gbs, 4),
        },
        "icp_iters_per_frame": round(im["sum_iters"] / icp_frames, 3),
        "icp_level_iters_per_frame": [round(x / icp_frames, 3) for x in im["level_iters"]],
        "icp_fallback_rate": round(im["fallbacks"] / icp_frames, 4),
        "pose_error_vs_truth": {
            "ate_m": round(float(np.sqrt(np.mean(errs[:, 0] ** 2))), 5),
            "at_mesh_frames_m_deg": [[f, round(float(errs[f, 0]), 5), round(float(errs[f, 1]), 4)]
                                     for f in range(MESH_EVERY - 1, TRAJ_FRAMES, MESH_EVERY)],
        },
        "error_vs_fp64": cpu["error_vs_fp64"] if cpu else None,
        "volume_voxels_at_mesh": q["volume_at_mesh"],
        "mesh_vertices": head["qmeshes"],
        "deterministic": head["qmeshes"] == main_p["meshes"] == roof["meshes"] == diag["meshes"],
        "stage_ms_per_frame": {k: round(v[0] / frames_timed, 4) for k, v in stage.items()},
        "stage_timing_note": "value from a pass with no events; merge_update from a second pass over the same "
                             "trajectories with only its events (%.1f frames/s), the other stages from a third "
                             "pass with every stage timer on (%.1f frames/s)"
                             % (frames_timed / roof["elapsed"], frames_timed / diag["elapsed"]),
        "fps_no_mesh": round(frames_timed / head["nomesh"]["elapsed"], 3),
        "fps_with_h2d_no_mesh": round(frames_timed / head["h2d"]["elapsed"], 3),
        "config1_plain_8mm": c1,
        "learned_vs_plain": learned,
        "precision_variant_2mm": precision,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
