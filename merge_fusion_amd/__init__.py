"""MI355X-native KinectFission (kifi) fusion hot path.

Python binding of the C ABI in include/kifi.h (libkifi_hip.so) and of the
procedural frame generator in include/kifi_synth.h (libkifi_synth.so).  The
binding mirrors the reference's classes (KinectFission/kifi/Volume.h,
Integrator.h, Mesher.h, Renderer.h): a Volume is created from
(resolution, side, truncation), Integrator.integrate(volume, frame, camera,
worldToEye), Mesher.mesh(volume[, with indices]).

There is no CPU fallback: if the HIP library is not built or no GPU is
present, every call raises.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

from .depthstream import DepthStream, write_depth  # noqa: E402,F401

_PKG = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(_PKG, "libkifi_hip.so")
SYNTH_LIB_PATH = os.path.join(_PKG, "libkifi_synth.so")

KIFI_OK, KIFI_ERR_ARG, KIFI_ERR_HIP, KIFI_ERR_NOMEM, KIFI_ERR_CAPACITY, KIFI_ERR_STATE = range(6)
STORAGE_F32, STORAGE_F16U8 = 0, 1
SCENE_PLANE_SPHERES, SCENE_ROOM = 0, 1
TRAJ_STATIC, TRAJ_CIRCLE, TRAJ_FAST, TRAJ_SWAY = 0, 1, 2, 3

HIP_EXPORTS = [
    "kifi_ctx_create", "kifi_ctx_destroy", "kifi_last_error", "kifi_ctx_synchronize", "kifi_ctx_stream",
    "kifi_volume_create", "kifi_volume_destroy", "kifi_volume_clear", "kifi_volume_size",
    "kifi_volume_download", "kifi_volume_upload", "kifi_integrate", "kifi_integrate_device",
    "kifi_volume_save", "kifi_volume_load", "kifi_integrate_f32", "kifi_integrate_device_f32", "kifi_pipeline_integrate_f32", "kifi_icp_align_points_f32",
    "kifi_integrate_stats", "kifi_mesh", "kifi_mesh_indexed", "kifi_mesh_async", "kifi_mesh_wait", "kifi_bin",
    "kifi_icp_params_default", "kifi_model_create", "kifi_model_destroy", "kifi_model_predict",
    "kifi_model_download", "kifi_icp_align", "kifi_icp_normal_equations", "kifi_horn_align", "kifi_pipeline_create",
    "kifi_pipeline_destroy", "kifi_pipeline_volume", "kifi_pipeline_model", "kifi_pipeline_integrate",
    "kifi_pipeline_set_guess", "kifi_pipeline_integrate_at", "kifi_pipeline_set_pose", "kifi_pipeline_pose", "kifi_pipeline_stats",
    "kifi_pipeline_reset", "kifi_icp_align_device_noalloc",
    "kifi_timing_enable", "kifi_timing_read", "kifi_device_alloc", "kifi_device_free", "kifi_host_alloc", "kifi_host_free", "kifi_memcpy_to_device",
    "kifi_device_copy", "kifi_integrate_totals", "kifi_icp_align_points",
    "kifi_twistnet_param_count", "kifi_twistnet_create", "kifi_twistnet_destroy", "kifi_twistnet_get_params",
    "kifi_twistnet_set_params", "kifi_twistnet_forward", "kifi_twistnet_loss_grad", "kifi_twistnet_train",
    "kifi_twistnet_predict_set", "kifi_pairset_generate", "kifi_pairset_generate_steps", "kifi_pairset_destroy", "kifi_pairset_size",
    "kifi_pairset_download", "kifi_render", "kifi_pipeline_set_twistnet", "kifi_twistnet_apply_gradient",
    "kifi_segmented_inclusive_scan", "kifi_segmented_exclusive_scan", "kifi_segmented_reduce",
]
STAGE_MERGE_UPDATE, STAGE_SORT, STAGE_ICP, STAGE_MODEL, STAGE_MESH = range(5)
ICP_POINT_TO_PLANE, ICP_HORN = 0, 1


class KifiError(RuntimeError):
    pass


class Intrinsics(ctypes.Structure):
    """DepthSensorParams (DepthSensorParams.h:23-45)."""
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("fx", ctypes.c_float),
                ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("znear", ctypes.c_float), ("zfar", ctypes.c_float)]

    def __init__(self, width=512, height=424, fx=365.0, fy=365.0, cx=None, cy=None, znear=0.4, zfar=8.0):
        super().__init__(width, height, fx, fy, width * 0.5 if cx is None else cx,
                         height * 0.5 if cy is None else cy, znear, zfar)


class VolumeDesc(ctypes.Structure):
    _fields_ = [("res", ctypes.c_int), ("side", ctypes.c_float), ("trunc", ctypes.c_float),
                ("key_bits", ctypes.c_int), ("storage", ctypes.c_int)]


class IcpParams(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("levels", ctypes.c_int), ("iters", ctypes.c_int * 3),
                ("dist_thresh", ctypes.c_float), ("eps_rot", ctypes.c_float), ("eps_trans", ctypes.c_float),
                ("min_inliers", ctypes.c_int), ("fallback_rms", ctypes.c_float), ("fallback_iters", ctypes.c_int),
                ("level_gate_growth", ctypes.c_float), ("naive_sums", ctypes.c_int),
                ("eps_level_growth", ctypes.c_float), ("spin_limit", ctypes.c_int),
                ("first_iters", ctypes.c_int * 3), ("motion_prior", ctypes.c_float)]

    @classmethod
    def default(cls):
        p = cls()
        load_hip().kifi_icp_params_default(ctypes.byref(p))
        return p


class IcpStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("iters", ctypes.c_int * 3), ("rms", ctypes.c_float),
                ("inliers", ctypes.c_int), ("failed", ctypes.c_int), ("fallback", ctypes.c_int),
                ("sum_iters", ctypes.c_longlong), ("frames", ctypes.c_longlong), ("fallbacks", ctypes.c_longlong),
                ("rms_1m", ctypes.c_float), ("sum_level_iters", ctypes.c_longlong * 3),
                ("start_pose", ctypes.c_double * 16)]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = list(v) if k in ("iters", "sum_level_iters") else (
                np.array(v, np.float64) if k == "start_pose" else v)
        return out


class Scene(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n_spheres", ctypes.c_int), ("n_boxes", ctypes.c_int),
                ("spheres", (ctypes.c_double * 4) * 8), ("boxes", (ctypes.c_double * 6) * 16),
                ("plane_z", ctypes.c_double), ("room", ctypes.c_double * 6)]


_hip = None
_synth = None


def _f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def load_hip():
    """Load libkifi_hip.so (raises if it was not built)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise KifiError("libkifi_hip.so is not built; run python -m merge_fusion_amd.build")
        lib = ctypes.CDLL(HIP_LIB_PATH)
        vp, sz, i32, f32p = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_float)
        szp = ctypes.POINTER(ctypes.c_size_t)
        f64p = ctypes.POINTER(ctypes.c_double)
        sigs = {
            "kifi_ctx_create": ([i32, ctypes.POINTER(vp)], i32),
            "kifi_ctx_destroy": ([vp], None),
            "kifi_last_error": ([vp], ctypes.c_char_p),
            "kifi_ctx_synchronize": ([vp], i32),
            "kifi_ctx_stream": ([vp], vp),
            "kifi_volume_create": ([vp, ctypes.POINTER(VolumeDesc), ctypes.POINTER(vp)], i32),
            "kifi_volume_destroy": ([vp], None),
            "kifi_volume_clear": ([vp], i32),
            "kifi_volume_size": ([vp, szp], i32),
            "kifi_volume_download": ([vp, vp, vp, sz, szp], i32),
            "kifi_volume_upload": ([vp, vp, vp, sz], i32),
            "kifi_integrate": ([vp, vp, ctypes.POINTER(Intrinsics), f32p], i32),
            "kifi_integrate_device": ([vp, vp, ctypes.POINTER(Intrinsics), f32p], i32),
            "kifi_integrate_f32": ([vp, vp, ctypes.POINTER(Intrinsics), f32p], i32),
            "kifi_volume_save": ([vp, ctypes.c_char_p], i32),
            "kifi_volume_load": ([vp, ctypes.c_char_p, ctypes.POINTER(vp)], i32),
            "kifi_integrate_device_f32": ([vp, vp, ctypes.POINTER(Intrinsics), f32p], i32),
            "kifi_integrate_stats": ([vp, szp], i32),
            "kifi_mesh": ([vp, f32p, sz, szp], i32),
            "kifi_mesh_async": ([vp, f32p, sz, szp], i32),
            "kifi_mesh_wait": ([vp], i32),
            "kifi_mesh_indexed": ([vp, f32p, sz, ctypes.POINTER(ctypes.c_uint32), sz, szp, szp,
                                   ctypes.POINTER(ctypes.c_int)], i32),
            "kifi_bin": ([vp, f32p, sz, f32p, i32, i32, f32p], i32),
            "kifi_icp_params_default": ([ctypes.POINTER(IcpParams)], i32),
            "kifi_model_create": ([vp, ctypes.POINTER(Intrinsics), ctypes.POINTER(vp)], i32),
            "kifi_model_destroy": ([vp], None),
            "kifi_model_predict": ([vp, vp, f32p], i32),
            "kifi_model_download": ([vp, f32p, f32p], i32),
            "kifi_icp_align": ([vp, vp, ctypes.POINTER(Intrinsics), f64p, ctypes.POINTER(IcpParams), f64p,
                                ctypes.POINTER(IcpStats)], i32),
            "kifi_icp_normal_equations": ([vp, vp, ctypes.POINTER(Intrinsics), f64p, i32, ctypes.c_float, i32,
                                           f64p], i32),
            "kifi_icp_align_device_noalloc": ([vp, vp, ctypes.POINTER(Intrinsics), f64p, ctypes.POINTER(IcpParams),
                                               f64p, ctypes.POINTER(IcpStats)], i32),
            "kifi_horn_align": ([vp, f32p, f32p, sz, f32p], i32),
            "kifi_pipeline_create": ([vp, ctypes.POINTER(VolumeDesc), ctypes.POINTER(Intrinsics),
                                      ctypes.POINTER(IcpParams), ctypes.POINTER(vp)], i32),
            "kifi_pipeline_destroy": ([vp], None),
            "kifi_pipeline_volume": ([vp], vp),
            "kifi_pipeline_model": ([vp], vp),
            "kifi_pipeline_integrate": ([vp, vp, i32, f32p], i32),
            "kifi_pipeline_integrate_f32": ([vp, vp, i32, f32p], i32),
            "kifi_pipeline_set_guess": ([vp, f64p], i32),
            "kifi_pipeline_reset": ([vp], i32),
            "kifi_pipeline_set_pose": ([vp, f64p], i32),
            "kifi_pipeline_pose": ([vp, f64p], i32),
            "kifi_pipeline_stats": ([vp, ctypes.POINTER(IcpStats)], i32),
            "kifi_timing_enable": ([vp, i32], i32),
            "kifi_timing_read": ([vp, i32, f64p, ctypes.POINTER(ctypes.c_longlong)], i32),
            "kifi_device_alloc": ([vp, sz, ctypes.POINTER(vp)], i32),
            "kifi_device_free": ([vp, vp], i32),
            "kifi_host_alloc": ([vp, sz, ctypes.POINTER(vp)], i32),
            "kifi_host_free": ([vp, vp], i32),
            "kifi_memcpy_to_device": ([vp, vp, vp, sz], i32),
            "kifi_device_copy": ([vp, vp, vp, sz], i32),
            "kifi_integrate_totals": ([vp, ctypes.POINTER(ctypes.c_uint64)], i32),
            "kifi_icp_align_points": ([vp, vp, ctypes.POINTER(Intrinsics), f32p, f32p, i32, i32, f32p, i32, f32p,
                                       ctypes.POINTER(IcpStats)], i32),
            "kifi_icp_align_points_f32": ([vp, vp, ctypes.POINTER(Intrinsics), f32p, f32p, i32, i32, f32p, i32, f32p,
                                       ctypes.POINTER(IcpStats)], i32),
            "kifi_twistnet_param_count": ([], sz),
            "kifi_twistnet_create": ([vp, ctypes.c_uint64, ctypes.POINTER(vp)], i32),
            "kifi_twistnet_destroy": ([vp], None),
            "kifi_twistnet_get_params": ([vp, f32p], i32),
            "kifi_twistnet_set_params": ([vp, f32p], i32),
            "kifi_twistnet_forward": ([vp, f32p, i32, f32p], i32),
            "kifi_twistnet_loss_grad": ([vp, f32p, f32p, i32, f32p, f32p], i32),
            "kifi_twistnet_train": ([vp, vp, i32, i32, ctypes.c_float, ctypes.c_uint64, f32p], i32),
            "kifi_twistnet_apply_gradient": ([vp, f32p, ctypes.c_float], i32),
            "kifi_segmented_inclusive_scan": ([vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp], i32),
            "kifi_segmented_exclusive_scan": ([vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp], i32),
            "kifi_segmented_reduce": ([vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp], i32),
            "kifi_twistnet_predict_set": ([vp, vp, i32, i32, f32p], i32),
            "kifi_pairset_generate": ([vp, ctypes.POINTER(Intrinsics), vp, i32, i32, ctypes.c_uint64,
                                       ctypes.c_float, ctypes.c_float, ctypes.POINTER(vp)], i32),
            "kifi_pairset_generate_steps": ([vp, ctypes.POINTER(Intrinsics), vp, i32, i32, ctypes.c_uint64,
                                             ctypes.POINTER(MotionSteps), ctypes.POINTER(vp)], i32),
            "kifi_pairset_destroy": ([vp], None),
            "kifi_pairset_size": ([vp], i32),
            "kifi_pairset_download": ([vp, f32p, f32p, f32p, ctypes.POINTER(ctypes.c_int)], i32),
            "kifi_render": ([vp, vp, ctypes.POINTER(Intrinsics), f64p, ctypes.c_uint64, ctypes.c_double,
                             ctypes.c_double, vp, i32], i32),
            "kifi_pipeline_set_twistnet": ([vp, vp], i32),
            "kifi_pipeline_integrate_at": ([vp, vp, i32, f32p], i32),
        }
        for name, (args, res) in sigs.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _hip = lib
    return _hip


def load_synth():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_LIB_PATH):
            raise KifiError("libkifi_synth.so is not built; run python -m merge_fusion_amd.build")
        lib = ctypes.CDLL(SYNTH_LIB_PATH)
        lib.kifi_scene_make.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(Scene)]
        lib.kifi_scene_make.restype = ctypes.c_int
        lib.kifi_scene_render.argtypes = [ctypes.POINTER(Scene), ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                          ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                          ctypes.POINTER(ctypes.c_double), ctypes.c_uint64, ctypes.c_double,
                                          ctypes.c_double, ctypes.POINTER(ctypes.c_uint16),
                                          ctypes.POINTER(ctypes.c_float)]
        lib.kifi_scene_render.restype = ctypes.c_int
        lib.kifi_trajectory_pose.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_double)]
        lib.kifi_trajectory_pose.restype = ctypes.c_int
        lib.kifi_splitmix64.argtypes = [ctypes.c_uint64]
        lib.kifi_splitmix64.restype = ctypes.c_uint64
        _synth = lib
    return _synth


# ---------------------------------------------------------------- synthetic data
def make_scene(kind, seed):
    s = Scene()
    if load_synth().kifi_scene_make(kind, seed, ctypes.byref(s)) != 0:
        raise KifiError("unknown scene kind %d" % kind)
    return s


def trajectory_pose(traj, seed, frame):
    """Ground-truth eye->world pose (float64[16], column storage)."""
    m = np.zeros(16, np.float64)
    if load_synth().kifi_trajectory_pose(traj, seed, frame, m.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
        raise KifiError("bad trajectory request")
    return m


def render(scene, cam, eye_to_world, noise_seed, noise_k=1.5e-3, invalid_frac=0.02, with_truth=False):
    """One uint16 mm frame (H, W) of `scene` seen from eye_to_world (float64[16])."""
    out = np.zeros((cam.height, cam.width), np.uint16)
    truth = np.zeros((cam.height, cam.width), np.float32) if with_truth else None
    e2w = np.ascontiguousarray(eye_to_world, np.float64)
    rc = load_synth().kifi_scene_render(
        ctypes.byref(scene), cam.width, cam.height, cam.fx, cam.fy, cam.cx, cam.cy,
        e2w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), noise_seed, noise_k, invalid_frac,
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
        truth.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if truth is not None else None)
    if rc:
        raise KifiError("render failed")
    return (out, truth) if with_truth else out


def invert_rigid(m):
    """Inverse of a rigid eye->world / world->eye matrix (column storage, float64)."""
    M = np.asarray(m, np.float64).reshape(4, 4).T     # row-major view
    R, t = M[:3, :3], M[:3, 3]
    out = np.eye(4)
    out[:3, :3] = R.T
    out[:3, 3] = -R.T @ t
    return out.T.reshape(16).copy()


# ---------------------------------------------------------------- HIP objects
# Contexts still open when the interpreter exits are closed by an exit hook, in
# order (children, host buffers, then the context), before module teardown
# starts dropping objects in an arbitrary order.
_live_contexts = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    for c in list(_live_contexts):
        try:
            c.close()
        except Exception:
            pass


class Context:
    """kifi_ctx: one HIP stream on one device."""

    def __init__(self, device=0):
        lib = load_hip()
        h = ctypes.c_void_p()
        rc = lib.kifi_ctx_create(device, ctypes.byref(h))
        if rc != KIFI_OK:
            raise KifiError("kifi_ctx_create failed (%d): no usable GPU?" % rc)
        self._h = h
        self._lib = lib
        # objects created on this context; closed before the context is destroyed
        self._children = weakref.WeakSet()
        # page-locked host buffers by address, freed when their arrays are gone
        # or when the context closes
        self._host = {}
        _live_contexts.add(self)

    def _adopt(self, obj):
        self._children.add(obj)

    @property
    def handle(self):
        return self._h

    def check(self, rc, what):
        if rc != KIFI_OK:
            msg = self._lib.kifi_last_error(self._h)
            raise KifiError("%s: status %d: %s" % (what, rc, msg.decode() if msg else ""))

    def synchronize(self):
        self.check(self._lib.kifi_ctx_synchronize(self._h), "synchronize")

    def mesh_wait(self):
        """Waits for the vertex copies queued by Volume.mesh_into_async."""
        self.check(self._lib.kifi_mesh_wait(self._h), "kifi_mesh_wait")

    def close(self):
        if self._h:
            # pipelines first (they own a volume and a model), then the rest
            kids = sorted(list(self._children), key=lambda o: not isinstance(o, Pipeline))
            for obj in kids:
                try:
                    obj.close()
                except Exception:
                    pass
            for buf in list(self._host.values()):
                buf.free()
            self._host.clear()
            self._lib.kifi_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def timing(self, enable, stages=None):
        """Stage timers on (all stages, or the STAGE_* listed) or off."""
        mask = 0
        if enable:
            mask = (1 << 5) - 1 if stages is None else sum(1 << s for s in stages)
        self.check(self._lib.kifi_timing_enable(self._h, mask), "kifi_timing_enable")

    def timing_read(self, stage):
        """(total ms, launches) of a stage since the last read."""
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        self.check(self._lib.kifi_timing_read(self._h, stage, ctypes.byref(ms), ctypes.byref(n)), "kifi_timing_read")
        return ms.value, n.value

    def host_empty(self, shape, dtype):
        """A numpy array in page-locked host memory (kifi_host_alloc), freed when the
        array and its views are gone, or when the context closes (the array must
        not be used after that)."""
        dt = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dt.itemsize
        p = ctypes.c_void_p()
        self.check(self._lib.kifi_host_alloc(self._h, nbytes, ctypes.byref(p)), "kifi_host_alloc")
        self._host[p.value] = HostBuffer(self, p, nbytes)
        raw = (ctypes.c_char * max(nbytes, 1)).from_address(p.value)
        arr = np.frombuffer(raw, dtype=dt, count=int(np.prod(shape))).reshape(shape)
        # every view of arr keeps raw alive: free the buffer once the last one goes
        weakref.finalize(raw, self._release_host, p.value)
        return arr

    def _release_host(self, addr):
        buf = self._host.pop(addr, None)
        if buf is not None:
            buf.free()

    def device_upload(self, array):
        """Copy a numpy array into a new device buffer; returns DeviceBuffer."""
        a = np.ascontiguousarray(array)
        p = ctypes.c_void_p()
        self.check(self._lib.kifi_device_alloc(self._h, a.nbytes, ctypes.byref(p)), "kifi_device_alloc")
        buf = DeviceBuffer(self, p, a.nbytes)
        self._adopt(buf)
        self.check(self._lib.kifi_memcpy_to_device(self._h, p, a.ctypes.data, a.nbytes), "kifi_memcpy_to_device")
        return buf

    def device_copy(self, dst, src, nbytes):
        self.check(self._lib.kifi_device_copy(self._h, ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes),
                   "kifi_device_copy")

    def segmented(self, op, data, segment):
        """op: "inclusive" / "exclusive" scan per segment, or "reduce" (sum per
        segment) of uint32 data, on the device's workgroup primitives."""
        d = np.ascontiguousarray(data, np.uint32)
        n = d.size
        out = np.zeros(n if op != "reduce" else (n + segment - 1) // segment, np.uint32)
        fn = {"inclusive": self._lib.kifi_segmented_inclusive_scan, "exclusive": self._lib.kifi_segmented_exclusive_scan,
              "reduce": self._lib.kifi_segmented_reduce}[op]
        self.check(fn(self._h, d.ctypes.data, n, segment, out.ctypes.data), "kifi_segmented_" + op)
        return out

    def bin(self, points, world_to_clip, width, height):
        """Renderer::Bin: z-buffered splat into a (H, W, 3) float32 map."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        m = np.ascontiguousarray(world_to_clip, np.float32).reshape(16)
        out = np.zeros((height, width, 3), np.float32)
        self.check(self._lib.kifi_bin(self._h, _f32p(pts), pts.shape[0], _f32p(m), width, height, _f32p(out)),
                   "kifi_bin")
        return out


class DeviceBuffer:
    def __init__(self, ctx, ptr, nbytes):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes

    def address(self, offset=0):
        return self.ptr.value + offset

    def free(self):
        if self.ptr and self.ctx.handle:
            self.ctx._lib.kifi_device_free(self.ctx.handle, self.ptr)
        self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def load_volume(ctx, path):
    """A volume from a kifi_volume_save file (description, keys and values)."""
    with open(path, "rb") as f:
        head = f.read(36)
    if len(head) < 36 or head[:8] != b"KIFIVOL\n":
        raise KifiError("not a volume file: %s" % path)
    res, = np.frombuffer(head[12:16], "<i4")
    side, trunc = np.frombuffer(head[16:24], "<f4")
    kb, storage = np.frombuffer(head[24:32], "<i4")
    return Volume(ctx, int(res), float(side), float(trunc), int(kb), int(storage), _path=path)


class HostBuffer:
    def __init__(self, ctx, ptr, nbytes):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes

    def free(self):
        if self.ptr and self.ctx.handle:
            self.ctx._lib.kifi_host_free(self.ctx.handle, self.ptr)
        self.ptr = None

    close = free


class Volume:
    """kifi::Volume (Volume.h:15-31) resident on the GPU."""

    def __init__(self, ctx, res, side, trunc, key_bits=32, storage=STORAGE_F32, _path=None):
        self.ctx = ctx
        self.desc = VolumeDesc(res, side, trunc, key_bits, storage)
        h = ctypes.c_void_p()
        if _path is None:
            ctx.check(ctx._lib.kifi_volume_create(ctx.handle, ctypes.byref(self.desc), ctypes.byref(h)),
                      "kifi_volume_create")
        else:
            ctx.check(ctx._lib.kifi_volume_load(ctx.handle, os.fsencode(_path), ctypes.byref(h)), "kifi_volume_load")
        self._h = h
        ctx._adopt(self)
        self.key_dtype = np.uint32 if key_bits == 32 else np.uint64
        self.val_dtype = np.dtype([("sum", np.float32), ("weight", np.float32)]) if storage == STORAGE_F32 \
            else np.uint32

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            self.ctx._lib.kifi_volume_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self):
        n = ctypes.c_size_t()
        self.ctx.check(self.ctx._lib.kifi_volume_size(self._h, ctypes.byref(n)), "kifi_volume_size")
        return n.value

    def clear(self):
        self.ctx.check(self.ctx._lib.kifi_volume_clear(self._h), "kifi_volume_clear")

    def save(self, path):
        """Checkpoint to a file (kifi_volume_save)."""
        self.ctx.check(self.ctx._lib.kifi_volume_save(self._h, os.fsencode(path)), "kifi_volume_save")

    def download(self):
        """(keys, values) of the sorted flat map."""
        n = self.size()
        keys = np.zeros(n, self.key_dtype)
        vals = np.zeros(n, self.val_dtype)
        got = ctypes.c_size_t()
        self.ctx.check(self.ctx._lib.kifi_volume_download(self._h, keys.ctypes.data, vals.ctypes.data, n,
                                                          ctypes.byref(got)), "kifi_volume_download")
        return keys, vals

    def upload(self, keys, values):
        k = np.ascontiguousarray(keys, self.key_dtype)
        v = np.ascontiguousarray(values, self.val_dtype)
        self.ctx.check(self.ctx._lib.kifi_volume_upload(self._h, k.ctypes.data, v.ctypes.data, k.shape[0]),
                       "kifi_volume_upload")

    def integrate(self, depth, cam, world_to_eye):
        """Integrator::Integrate (Integrator.cpp:26-71): depth (H, W) uint16 mm, or
        float32 metres used as they are (a vector2d<float> frame)."""
        metres = np.asarray(depth).dtype == np.float32
        d = np.ascontiguousarray(depth, np.float32 if metres else np.uint16)
        if d.shape != (cam.height, cam.width):
            raise KifiError("frame shape does not match the intrinsics")
        m = np.ascontiguousarray(world_to_eye, np.float32).reshape(16)
        fn = self.ctx._lib.kifi_integrate_f32 if metres else self.ctx._lib.kifi_integrate
        self.ctx.check(fn(self._h, d.ctypes.data, ctypes.byref(cam), _f32p(m)), "kifi_integrate")

    def integrate_device(self, depth_ptr, cam, world_to_eye, metres=False):
        m = np.ascontiguousarray(world_to_eye, np.float32).reshape(16)
        fn = self.ctx._lib.kifi_integrate_device_f32 if metres else self.ctx._lib.kifi_integrate_device
        self.ctx.check(fn(self._h, ctypes.c_void_p(depth_ptr), ctypes.byref(cam), _f32p(m)), "kifi_integrate_device")

    def totals(self):
        """Running sums since creation: integrations, V0, E, V1, K, A (updated voxels)."""
        t = (ctypes.c_uint64 * 6)()
        self.ctx.check(self.ctx._lib.kifi_integrate_totals(self._h, t), "kifi_integrate_totals")
        return {"frames": t[0], "v_old": t[1], "new": t[2], "v_new": t[3], "keys": t[4], "updated": t[5]}

    def stats(self):
        s = (ctypes.c_size_t * 4)()
        self.ctx.check(self.ctx._lib.kifi_integrate_stats(self._h, s), "kifi_integrate_stats")
        return {"keys": s[0], "candidates": s[1], "new": s[2], "size": s[3]}

    def mesh(self):
        """Mesher::Mesh(volume, verts): (N, 3) float32 in generation order."""
        lib = self.ctx._lib
        n = ctypes.c_size_t()
        rc = lib.kifi_mesh(self._h, None, 0, ctypes.byref(n))
        if rc not in (KIFI_OK, KIFI_ERR_CAPACITY):
            self.ctx.check(rc, "kifi_mesh")
        verts = np.zeros((n.value, 3), np.float32)
        self.ctx.check(lib.kifi_mesh(self._h, _f32p(verts), n.value, ctypes.byref(n)), "kifi_mesh")
        return verts

    def mesh_into(self, buf):
        """Mesher::Mesh into a preallocated (N, 3) float32 buffer; returns the vertex count
        (larger than len(buf) if the buffer was too small and nothing was written)."""
        n = ctypes.c_size_t()
        rc = self.ctx._lib.kifi_mesh(self._h, _f32p(buf), buf.shape[0], ctypes.byref(n))
        if rc not in (KIFI_OK, KIFI_ERR_CAPACITY):
            self.ctx.check(rc, "kifi_mesh")
        return n.value

    def mesh_into_async(self, buf):
        """mesh_into whose copy to buf overlaps the work queued after it: returns the
        vertex count once the kernels and the copy are queued (nothing is queued if
        it exceeds len(buf)); buf holds the vertices after Context.mesh_wait (or
        synchronize) and must be kept alive until then.  Overlaps only into
        page-locked buffers (Context.host_empty)."""
        n = ctypes.c_size_t()
        rc = self.ctx._lib.kifi_mesh_async(self._h, _f32p(buf), buf.shape[0], ctypes.byref(n))
        if rc not in (KIFI_OK, KIFI_ERR_CAPACITY):
            self.ctx.check(rc, "kifi_mesh_async")
        return n.value

    def mesh_indexed(self):
        """Mesher::Mesh(volume, verts, idx): (verts sorted by vertex ID, indices, stalled)."""
        lib = self.ctx._lib
        nv, ni, st = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_int()
        rc = lib.kifi_mesh_indexed(self._h, None, 0, None, 0, ctypes.byref(nv), ctypes.byref(ni),
                                   ctypes.byref(st))
        if rc not in (KIFI_OK, KIFI_ERR_CAPACITY):
            self.ctx.check(rc, "kifi_mesh_indexed")
        verts = np.zeros((nv.value, 3), np.float32)
        idx = np.zeros(ni.value, np.uint32)
        self.ctx.check(lib.kifi_mesh_indexed(self._h, _f32p(verts), nv.value,
                                             idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ni.value,
                                             ctypes.byref(nv), ctypes.byref(ni), ctypes.byref(st)),
                       "kifi_mesh_indexed")
        return verts, idx, bool(st.value)


class Model:
    """Device-resident model point/normal maps (ICP's model input)."""

    def __init__(self, ctx, cam, handle=None):
        self.ctx, self.cam = ctx, cam
        self._owned = handle is None
        if handle is None:
            h = ctypes.c_void_p()
            ctx.check(ctx._lib.kifi_model_create(ctx.handle, ctypes.byref(cam), ctypes.byref(h)),
                      "kifi_model_create")
            handle = h
        self._h = handle
        ctx._adopt(self)

    def close(self):
        if self._h and self._owned:
            self.ctx._lib.kifi_model_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def predict(self, volume, world_to_eye):
        m = np.ascontiguousarray(world_to_eye, np.float32).reshape(16)
        self.ctx.check(self.ctx._lib.kifi_model_predict(self._h, volume.handle, _f32p(m)), "kifi_model_predict")

    def download(self):
        P = self.cam.width * self.cam.height
        pts = np.zeros((self.cam.height, self.cam.width, 4), np.float32)
        nrm = np.zeros_like(pts)
        self.ctx.check(self.ctx._lib.kifi_model_download(self._h, _f32p(pts), _f32p(nrm)), "kifi_model_download")
        return pts, nrm

    def icp_align(self, depth_mm, guess_e2w, params=None):
        """ICP::Align: returns (eye->world float64[16], stats dict)."""
        p = params if params is not None else IcpParams.default()
        d = np.ascontiguousarray(depth_mm, np.uint16)
        if d.shape != (self.cam.height, self.cam.width):
            raise KifiError("frame shape does not match the intrinsics")
        g = np.ascontiguousarray(guess_e2w, np.float64).reshape(16)
        out = np.zeros(16, np.float64)
        st = IcpStats()
        self.ctx.check(self.ctx._lib.kifi_icp_align(self._h, d.ctypes.data, ctypes.byref(self.cam), _f64p(g),
                                                    ctypes.byref(p), _f64p(out), ctypes.byref(st)),
                       "kifi_icp_align")
        return out, st.as_dict()

    def icp_align_device_noalloc(self, depth_ptr, guess_e2w, params=None):
        """kifi_icp_align_device_noalloc: the same alignment of a frame already in
        device memory (uint16 mm at depth_ptr), allocating nothing."""
        p = params if params is not None else IcpParams.default()
        g = np.ascontiguousarray(guess_e2w, np.float64).reshape(16)
        out = np.zeros(16, np.float64)
        st = IcpStats()
        self.ctx.check(self.ctx._lib.kifi_icp_align_device_noalloc(self._h, ctypes.c_void_p(depth_ptr),
                                                                   ctypes.byref(self.cam), _f64p(g), ctypes.byref(p),
                                                                   _f64p(out), ctypes.byref(st)),
                       "kifi_icp_align_device_noalloc")
        return out, st.as_dict()

    def normal_equations(self, depth_mm, e2w, level=0, dist_thresh=0.1, naive=False):
        """One association at the eye->world pose e2w on `level`: the 30 reduced sums
        (21 sum w J_i J_j for i >= j row by row, 6 sum w J_i r, count, sum r^2,
        sum w r^2) as the tracker sums them, in float64."""
        d = np.ascontiguousarray(depth_mm, np.uint16)
        if d.shape != (self.cam.height, self.cam.width):
            raise KifiError("frame shape does not match the intrinsics")
        g = np.ascontiguousarray(e2w, np.float64).reshape(16)
        out = np.zeros(30, np.float64)
        self.ctx.check(self.ctx._lib.kifi_icp_normal_equations(self._h, d.ctypes.data, ctypes.byref(self.cam), _f64p(g),
                                                               int(level), float(dist_thresh), int(bool(naive)),
                                                               _f64p(out)), "kifi_icp_normal_equations")
        return out


def horn_align(ctx, src, dst):
    s = np.ascontiguousarray(src, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dst, np.float32).reshape(-1, 3)
    out = np.zeros(16, np.float32)
    ctx.check(ctx._lib.kifi_horn_align(ctx.handle, _f32p(s), _f32p(d), s.shape[0], _f32p(out)), "kifi_horn_align")
    return out


def icp_align_points(ctx, depth, cam, guess_e2w, synth_xyz, synth_e2w, iterations=4):
    """ICP::Align (ICP.cpp:9-265) against a host point buffer: (pose float32[16], stats).
    depth: uint16 mm or float32 metres."""
    metres = np.asarray(depth).dtype == np.float32
    d = np.ascontiguousarray(depth, np.float32 if metres else np.uint16)
    if d.shape != (cam.height, cam.width):
        raise KifiError("depth frame must be (height, width) of the camera")
    syn = np.ascontiguousarray(np.asarray(synth_xyz, np.float32)[..., :3])
    g = np.ascontiguousarray(guess_e2w, np.float32).reshape(16)
    se = np.ascontiguousarray(synth_e2w, np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    st = IcpStats()
    fn = ctx._lib.kifi_icp_align_points_f32 if metres else ctx._lib.kifi_icp_align_points
    ctx.check(fn(ctx.handle, d.ctypes.data, ctypes.byref(cam), _f32p(g), _f32p(syn), syn.shape[1], syn.shape[0],
                 _f32p(se), iterations, _f32p(out), ctypes.byref(st)), "kifi_icp_align_points")
    return out, st.as_dict()


class _BorrowedVolume(Volume):
    def __init__(self, ctx, handle, desc):
        self.ctx = ctx
        self.desc = desc
        self._h = handle
        self.key_dtype = np.uint32 if desc.key_bits == 32 else np.uint64
        self.val_dtype = np.dtype([("sum", np.float32), ("weight", np.float32)]) \
            if desc.storage == STORAGE_F32 else np.uint32

    def close(self):
        self._h = None


class Pipeline:
    """kifi::Pipeline (Pipeline.h:19-41): volume + model + tracking, one call per frame."""

    def __init__(self, ctx, cam, res=1024, side=4.096, trunc=0.02, key_bits=32, storage=STORAGE_F32,
                 params=None):
        self.ctx, self.cam = ctx, cam
        self.params = params if params is not None else IcpParams.default()
        desc = VolumeDesc(res, side, trunc, key_bits, storage)
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.kifi_pipeline_create(ctx.handle, ctypes.byref(desc), ctypes.byref(cam),
                                                ctypes.byref(self.params), ctypes.byref(h)), "kifi_pipeline_create")
        self._h = h
        ctx._adopt(self)
        self.volume = _BorrowedVolume(ctx, ctypes.c_void_p(ctx._lib.kifi_pipeline_volume(h)), desc)
        self.model = Model(ctx, cam, ctypes.c_void_p(ctx._lib.kifi_pipeline_model(h)))
        self.model._owned = False

    def close(self):
        if self._h:
            self.ctx._lib.kifi_pipeline_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def integrate(self, depth, world_to_eye=None):
        """Integrate at world_to_eye, or track with ICP when it is None.  depth: uint16
        mm or float32 metres."""
        metres = np.asarray(depth).dtype == np.float32
        d = np.ascontiguousarray(depth, np.float32 if metres else np.uint16)
        if d.shape != (self.cam.height, self.cam.width):
            raise KifiError("frame shape does not match the intrinsics")
        m = None if world_to_eye is None else _f32p(np.ascontiguousarray(world_to_eye, np.float32).reshape(16))
        fn = self.ctx._lib.kifi_pipeline_integrate_f32 if metres else self.ctx._lib.kifi_pipeline_integrate
        self.ctx.check(fn(self._h, d.ctypes.data, 0, m), "kifi_pipeline_integrate")

    def integrate_device(self, depth_ptr, world_to_eye=None, metres=False):
        m = None if world_to_eye is None else _f32p(np.ascontiguousarray(world_to_eye, np.float32).reshape(16))
        fn = self.ctx._lib.kifi_pipeline_integrate_f32 if metres else self.ctx._lib.kifi_pipeline_integrate
        self.ctx.check(fn(self._h, ctypes.c_void_p(depth_ptr), 1, m), "kifi_pipeline_integrate")

    def integrate_at(self, depth, fuse_world_to_eye, on_device=False):
        """kifi_pipeline_integrate_at: track the mm frame (host array, or a device
        pointer with on_device), fuse it at fuse_world_to_eye."""
        m = _f32p(np.ascontiguousarray(fuse_world_to_eye, np.float32).reshape(16))
        if on_device:
            ptr = ctypes.c_void_p(depth)
        else:
            d = np.ascontiguousarray(depth, np.uint16)
            if d.shape != (self.cam.height, self.cam.width):
                raise KifiError("frame shape does not match the intrinsics")
            ptr = d.ctypes.data
        self.ctx.check(self.ctx._lib.kifi_pipeline_integrate_at(self._h, ptr, int(bool(on_device)), m),
                       "kifi_pipeline_integrate_at")

    def set_twistnet(self, net):
        """Learned initial pose for tracked frames (None detaches)."""
        self.ctx.check(self.ctx._lib.kifi_pipeline_set_twistnet(self._h, net._h if net is not None else None),
                       "kifi_pipeline_set_twistnet")
        self._net = net

    def reset(self):
        """Empty volume, identity pose, next frame is the first; allocations and running totals kept."""
        self.ctx.check(self.ctx._lib.kifi_pipeline_reset(self._h), "kifi_pipeline_reset")

    def set_guess(self, e2w):
        g = np.ascontiguousarray(e2w, np.float64).reshape(16)
        self.ctx.check(self.ctx._lib.kifi_pipeline_set_guess(self._h, _f64p(g)), "kifi_pipeline_set_guess")

    def set_pose(self, e2w):
        g = np.ascontiguousarray(e2w, np.float64).reshape(16)
        self.ctx.check(self.ctx._lib.kifi_pipeline_set_pose(self._h, _f64p(g)), "kifi_pipeline_set_pose")

    def pose(self):
        out = np.zeros(16, np.float64)
        self.ctx.check(self.ctx._lib.kifi_pipeline_pose(self._h, _f64p(out)), "kifi_pipeline_pose")
        return out

    def stats(self):
        st = IcpStats()
        self.ctx.check(self.ctx._lib.kifi_pipeline_stats(self._h, ctypes.byref(st)), "kifi_pipeline_stats")
        return st.as_dict()


def pose_error(a, b):
    """(translation error m, rotation error deg) between two eye->world poses."""
    A = np.asarray(a, np.float64).reshape(4, 4).T
    B = np.asarray(b, np.float64).reshape(4, 4).T
    dt = np.linalg.norm(A[:3, 3] - B[:3, 3])
    R = A[:3, :3].T @ B[:3, :3]
    # angle from the skew part (sin) and the trace (cos): a rotation that began as
    # a binary32 matrix is orthonormal only to ~1e-8, which moves the trace of
    # R (a symmetric error) but not its skew part; arccos of the trace alone
    # would read that as ~0.01 deg
    s = 0.5 * np.linalg.norm([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    ang = np.degrees(np.arctan2(s, (np.trace(R) - 1) / 2))
    return dt, ang


# ------------------------------------------------------- learned initial pose
TWIST_IN_SHAPE = (2, 106, 128)
TWIST_SCALE = np.array([0.15, 0.15, 0.15, 0.06, 0.06, 0.06], np.float32)


class TwistNet:
    """BF16 CNN twist predictor (kifi_twistnet): 2 x 106 x 128 -> 6 raw outputs."""

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.kifi_twistnet_create(ctx.handle, seed, ctypes.byref(h)), "kifi_twistnet_create")
        self._h = h
        self.n_params = ctx._lib.kifi_twistnet_param_count()
        ctx._adopt(self)

    def close(self):
        if self._h:
            self.ctx._lib.kifi_twistnet_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        out = np.zeros(self.n_params, np.float32)
        self.ctx.check(self.ctx._lib.kifi_twistnet_get_params(self._h, _f32p(out)), "kifi_twistnet_get_params")
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.n_params:
            raise KifiError("expected %d parameters" % self.n_params)
        self.ctx.check(self.ctx._lib.kifi_twistnet_set_params(self._h, _f32p(p)), "kifi_twistnet_set_params")

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape((-1,) + TWIST_IN_SHAPE)
        y = np.zeros((x.shape[0], 6), np.float32)
        self.ctx.check(self.ctx._lib.kifi_twistnet_forward(self._h, _f32p(x), x.shape[0], _f32p(y)),
                       "kifi_twistnet_forward")
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape((-1,) + TWIST_IN_SHAPE)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 6)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.n_params, np.float32)
        self.ctx.check(self.ctx._lib.kifi_twistnet_loss_grad(self._h, _f32p(x), _f32p(t), x.shape[0], _f32p(loss),
                                                             _f32p(g)), "kifi_twistnet_loss_grad")
        return float(loss[0]), g

    def train(self, pairs, steps, batch=256, lr=1e-3, seed=1):
        losses = np.zeros(max(steps, 1), np.float32)
        self.ctx.check(self.ctx._lib.kifi_twistnet_train(self._h, pairs._h, steps, batch, lr, seed, _f32p(losses)),
                       "kifi_twistnet_train")
        return losses[:steps]

    def apply_gradient(self, grad, lr):
        """One Adam step with a caller's gradient (parameter order)."""
        g = np.ascontiguousarray(grad, np.float32).reshape(-1)
        if g.size != self.n_params:
            raise KifiError("gradient size does not match the network")
        self.ctx.check(self.ctx._lib.kifi_twistnet_apply_gradient(self._h, _f32p(g), lr),
                       "kifi_twistnet_apply_gradient")

    def predict(self, pairs, first=0, n=None):
        n = pairs.size - first if n is None else n
        out = np.zeros((n, 6), np.float32)
        self.ctx.check(self.ctx._lib.kifi_twistnet_predict_set(self._h, pairs._h, first, n, _f32p(out)),
                       "kifi_twistnet_predict_set")
        return out


class MotionSteps(ctypes.Structure):
    """kifi_motion_steps: the ranges of a training pair's relative motion."""
    _fields_ = [("min_trans", ctypes.c_float), ("max_trans", ctypes.c_float),
                ("min_rot_deg", ctypes.c_float), ("max_rot_deg", ctypes.c_float),
                ("max_pitch_step_deg", ctypes.c_float), ("walk", ctypes.c_int)]


class PairSet:
    """Device-resident training pairs (kifi_pairset) with ICP-converged targets.

    The relative motion has a random axis and direction, angle
    U[min_rot_deg, max_rot_deg], length U[min_trans, max_trans]; walk=True draws
    it as a walking camera's step instead (yaw +-U[min_rot_deg, max_rot_deg],
    pitch +-U[0, max_pitch_step_deg], horizontal translation)."""

    def __init__(self, ctx, cam, scenes, n, seed=1, max_trans=0.10, max_rot_deg=15.0, min_trans=0.0,
                 min_rot_deg=0.0, walk=False, max_pitch_step_deg=2.0):
        self.ctx = ctx
        arr = (Scene * len(scenes))(*scenes)
        h = ctypes.c_void_p()
        if min_trans == 0.0 and min_rot_deg == 0.0 and not walk:
            ctx.check(ctx._lib.kifi_pairset_generate(ctx.handle, ctypes.byref(cam), ctypes.cast(arr, ctypes.c_void_p),
                                                     len(scenes), n, seed, max_trans, max_rot_deg, ctypes.byref(h)),
                      "kifi_pairset_generate")
        else:
            ms = MotionSteps(min_trans, max_trans, min_rot_deg, max_rot_deg, max_pitch_step_deg, int(bool(walk)))
            ctx.check(ctx._lib.kifi_pairset_generate_steps(ctx.handle, ctypes.byref(cam),
                                                           ctypes.cast(arr, ctypes.c_void_p), len(scenes), n, seed,
                                                           ctypes.byref(ms), ctypes.byref(h)),
                      "kifi_pairset_generate_steps")
        self._h = h
        self.size = n
        ctx._adopt(self)

    def close(self):
        if self._h:
            self.ctx._lib.kifi_pairset_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def download(self, inputs=True):
        x = np.zeros((self.size,) + TWIST_IN_SHAPE, np.float32) if inputs else None
        t = np.zeros((self.size, 6), np.float32)
        g = np.zeros((self.size, 6), np.float32)
        v = np.zeros(self.size, np.int32)
        self.ctx.check(self.ctx._lib.kifi_pairset_download(
            self._h, _f32p(x) if inputs else None, _f32p(t), _f32p(g), v.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
            "kifi_pairset_download")
        return x, t, g, v.astype(bool)


def render_device(ctx, scene, cam, eye_to_world, noise_seed, noise_k=1.5e-3, invalid_frac=0.02, out_ptr=None):
    """kifi_render: the renderer on the GPU; host (H, W) uint16, or into a device pointer."""
    e2w = np.ascontiguousarray(eye_to_world, np.float64)
    if out_ptr is not None:
        ctx.check(ctx._lib.kifi_render(ctx.handle, ctypes.byref(scene), ctypes.byref(cam), _f64p(e2w), noise_seed,
                                       noise_k, invalid_frac, ctypes.c_void_p(out_ptr), 1), "kifi_render")
        return None
    out = np.zeros((cam.height, cam.width), np.uint16)
    ctx.check(ctx._lib.kifi_render(ctx.handle, ctypes.byref(scene), ctypes.byref(cam), _f64p(e2w), noise_seed,
                                   noise_k, invalid_frac, out.ctypes.data, 0), "kifi_render")
    return out
