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
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(_PKG, "libkifi_hip.so")
SYNTH_LIB_PATH = os.path.join(_PKG, "libkifi_synth.so")

KIFI_OK, KIFI_ERR_ARG, KIFI_ERR_HIP, KIFI_ERR_NOMEM, KIFI_ERR_CAPACITY, KIFI_ERR_STATE = range(6)
STORAGE_F32, STORAGE_F16U8 = 0, 1
SCENE_PLANE_SPHERES, SCENE_ROOM = 0, 1
TRAJ_STATIC, TRAJ_CIRCLE, TRAJ_FAST = 0, 1, 2

HIP_EXPORTS = [
    "kifi_ctx_create", "kifi_ctx_destroy", "kifi_last_error", "kifi_ctx_synchronize", "kifi_ctx_stream",
    "kifi_volume_create", "kifi_volume_destroy", "kifi_volume_clear", "kifi_volume_size",
    "kifi_volume_download", "kifi_volume_upload", "kifi_integrate", "kifi_integrate_device",
    "kifi_integrate_stats", "kifi_mesh", "kifi_mesh_indexed", "kifi_bin",
]


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


class Scene(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n_spheres", ctypes.c_int), ("n_boxes", ctypes.c_int),
                ("spheres", (ctypes.c_double * 4) * 8), ("boxes", (ctypes.c_double * 6) * 16),
                ("plane_z", ctypes.c_double), ("room", ctypes.c_double * 6)]


_hip = None
_synth = None


def _f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def load_hip():
    """Load libkifi_hip.so (raises if it was not built)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise KifiError("libkifi_hip.so is not built; run python -m merge_fusion_amd.build")
        lib = ctypes.CDLL(HIP_LIB_PATH)
        vp, sz, i32, f32p = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_float)
        szp = ctypes.POINTER(ctypes.c_size_t)
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
            "kifi_integrate_stats": ([vp, szp], i32),
            "kifi_mesh": ([vp, f32p, sz, szp], i32),
            "kifi_mesh_indexed": ([vp, f32p, sz, ctypes.POINTER(ctypes.c_uint32), sz, szp, szp,
                                   ctypes.POINTER(ctypes.c_int)], i32),
            "kifi_bin": ([vp, f32p, sz, f32p, i32, i32, f32p], i32),
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

    @property
    def handle(self):
        return self._h

    def check(self, rc, what):
        if rc != KIFI_OK:
            msg = self._lib.kifi_last_error(self._h)
            raise KifiError("%s: status %d: %s" % (what, rc, msg.decode() if msg else ""))

    def synchronize(self):
        self.check(self._lib.kifi_ctx_synchronize(self._h), "synchronize")

    def close(self):
        if self._h:
            self._lib.kifi_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bin(self, points, world_to_clip, width, height):
        """Renderer::Bin: z-buffered splat into a (H, W, 3) float32 map."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        m = np.ascontiguousarray(world_to_clip, np.float32).reshape(16)
        out = np.zeros((height, width, 3), np.float32)
        self.check(self._lib.kifi_bin(self._h, _f32p(pts), pts.shape[0], _f32p(m), width, height, _f32p(out)),
                   "kifi_bin")
        return out


class Volume:
    """kifi::Volume (Volume.h:15-31) resident on the GPU."""

    def __init__(self, ctx, res, side, trunc, key_bits=32, storage=STORAGE_F32):
        self.ctx = ctx
        self.desc = VolumeDesc(res, side, trunc, key_bits, storage)
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.kifi_volume_create(ctx.handle, ctypes.byref(self.desc), ctypes.byref(h)),
                  "kifi_volume_create")
        self._h = h
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

    def integrate(self, depth_mm, cam, world_to_eye):
        """Integrator::Integrate (Integrator.cpp:26-71); depth in mm (H, W) uint16."""
        d = np.ascontiguousarray(depth_mm, np.uint16)
        if d.shape != (cam.height, cam.width):
            raise KifiError("frame shape does not match the intrinsics")
        m = np.ascontiguousarray(world_to_eye, np.float32).reshape(16)
        self.ctx.check(self.ctx._lib.kifi_integrate(self._h, d.ctypes.data, ctypes.byref(cam), _f32p(m)),
                       "kifi_integrate")

    def integrate_device(self, depth_ptr, cam, world_to_eye):
        m = np.ascontiguousarray(world_to_eye, np.float32).reshape(16)
        self.ctx.check(self.ctx._lib.kifi_integrate_device(self._h, ctypes.c_void_p(depth_ptr),
                                                           ctypes.byref(cam), _f32p(m)),
                       "kifi_integrate_device")

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
