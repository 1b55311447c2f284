"""ctypes binding of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU oracle restates the reference's fusion path (KinectFission/kifi) in
plain C.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import this module, and only as the checker; the product path in
merge_fusion_amd/ never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
STORE_F32, STORE_F16U8, STORE_F64 = 0, 1, 2


class Camera(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("fx", ctypes.c_float),
                ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("znear", ctypes.c_float), ("zfar", ctypes.c_float)]


class VolumeDesc(ctypes.Structure):
    _fields_ = [("res", ctypes.c_int), ("kb", ctypes.c_int), ("side", ctypes.c_float),
                ("trunc", ctypes.c_float), ("vl", ctypes.c_float), ("vlHalfPlusMin", ctypes.c_float),
                ("vlInv", ctypes.c_float), ("minOverNegVl", ctypes.c_float)]


class OVolume(ctypes.Structure):
    _fields_ = [("desc", VolumeDesc), ("storage", ctypes.c_int), ("size", ctypes.c_size_t),
                ("cap", ctypes.c_size_t), ("keys", ctypes.POINTER(ctypes.c_uint64)), ("values", ctypes.c_void_p)]


_lib = None
F32P = ctypes.POINTER(ctypes.c_float)
F64P = ctypes.POINTER(ctypes.c_double)
U64P = ctypes.POINTER(ctypes.c_uint64)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        sig = {
            "orc_mat_mul": ([F32P, F32P, F32P], None),
            "orc_mat_vec": ([F32P, F32P, F32P], None),
            "orc_invert_transform": ([F32P], None),
            "orc_eye_to_clip_rh": ([ctypes.POINTER(Camera), F32P], None),
            "orc_volume_desc_init": ([ctypes.POINTER(VolumeDesc), i32, ctypes.c_float, ctypes.c_float, i32], None),
            "orc_pack": ([ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, i32], ctypes.c_uint64),
            "orc_keygen": ([F32P, ctypes.POINTER(Camera), F32P, ctypes.POINTER(VolumeDesc), U64P], sz),
            "orc_radix_sort": ([U64P, sz, i32, U64P], None),
            "orc_unique": ([U64P, sz], sz),
            "orc_expand": ([U64P, sz, i32, U64P, sz], sz),
            "orc_set_union": ([U64P, sz, U64P, sz, vp, vp, sz, U64P, vp], sz),
            "orc_update": ([U64P, vp, sz, i32, F32P, ctypes.POINTER(Camera), F32P, ctypes.POINTER(VolumeDesc)], None),
            "orc_update_f64": ([U64P, F64P, sz, F32P, ctypes.POINTER(Camera), F64P, ctypes.POINTER(VolumeDesc)], None),
            "orc_volume_create": ([i32, ctypes.c_float, ctypes.c_float, i32, i32], ctypes.POINTER(OVolume)),
            "orc_volume_destroy": ([ctypes.POINTER(OVolume)], None),
            "orc_volume_clear": ([ctypes.POINTER(OVolume)], None),
            "orc_volume_size": ([ctypes.POINTER(OVolume)], sz),
            "orc_volume_read": ([ctypes.POINTER(OVolume), U64P, vp, sz], sz),
            "orc_volume_write": ([ctypes.POINTER(OVolume), U64P, vp, sz], i32),
            "orc_integrate": ([ctypes.POINTER(OVolume), ctypes.POINTER(ctypes.c_uint16), ctypes.POINTER(Camera),
                               F32P, ctypes.POINTER(ctypes.c_size_t)], sz),
            "orc_integrate_f32": ([ctypes.POINTER(OVolume), F32P, ctypes.POINTER(Camera),
                                   F32P, ctypes.POINTER(ctypes.c_size_t)], sz),
            "orc_voxel_update_f32": ([F32P, ctypes.c_float], None),
            "orc_voxel_distance_f32": ([F32P], ctypes.c_float),
            "orc_voxel_update_f16u8": ([ctypes.POINTER(ctypes.c_uint32), ctypes.c_float], None),
            "orc_voxel_distance_f16u8": ([ctypes.c_uint32], ctypes.c_float),
            "orc_f32_to_f16": ([ctypes.c_float], ctypes.c_uint16),
            "orc_f16_to_f32": ([ctypes.c_uint16], ctypes.c_float),
            "orc_mesh": ([ctypes.POINTER(OVolume), F32P, sz], sz),
            "orc_mesh_indexed": ([ctypes.POINTER(OVolume), F32P, sz, ctypes.POINTER(ctypes.c_uint32), sz,
                                  ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int)], sz),
            "orc_tri_offsets": ([], ctypes.POINTER(ctypes.c_int)),
            "orc_tri_table": ([], ctypes.POINTER(ctypes.c_int)),
            "orc_bin": ([F32P, sz, F32P, i32, i32, F32P, F32P], None),
            "orc_horn_step": ([F32P, F32P, sz, F32P], None),
            "orc_jacobi4": ([F64P, F64P, F64P], None),
            "orc_kahan_sum_f32": ([F32P, sz], ctypes.c_float),
            "orc_partial_sum_exclusive_u32": ([ctypes.POINTER(ctypes.c_uint32), sz,
                                               ctypes.POINTER(ctypes.c_uint32)], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def f32p(a):
    return a.ctypes.data_as(F32P)


def u64p(a):
    return a.ctypes.data_as(U64P)


def camera(cam):
    """Oracle camera from any object with the DepthSensorParams fields."""
    return Camera(cam.width, cam.height, cam.fx, cam.fy, cam.cx, cam.cy, cam.znear, cam.zfar)


def volume_desc(res, side, trunc, kb=10):
    d = VolumeDesc()
    lib().orc_volume_desc_init(ctypes.byref(d), res, side, trunc, kb)
    return d


def mat_mul(m, n):
    m = np.ascontiguousarray(m, np.float32); n = np.ascontiguousarray(n, np.float32)
    out = np.zeros(16, np.float32)
    lib().orc_mat_mul(f32p(m), f32p(n), f32p(out))
    return out


def invert_transform(m):
    out = np.array(m, np.float32).reshape(16).copy()
    lib().orc_invert_transform(f32p(out))
    return out


def eye_to_clip(cam):
    out = np.zeros(16, np.float32)
    lib().orc_eye_to_clip_rh(ctypes.byref(camera(cam)), f32p(out))
    return out


class Volume:
    """Oracle flat_map volume (keys uint64, values float {sum, weight} or packed uint32)."""

    def __init__(self, res, side, trunc, kb=10, storage=STORE_F32):
        self.h = lib().orc_volume_create(res, side, trunc, kb, storage)
        self.storage = storage
        self.kb = kb

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_volume_destroy(self.h)
            self.h = None

    @property
    def desc(self):
        return self.h.contents.desc

    def size(self):
        return lib().orc_volume_size(self.h)

    def read(self):
        n = self.size()
        keys = np.zeros(n, np.uint64)
        vals = (np.zeros((n, 2), np.float32) if self.storage == STORE_F32 else
                np.zeros((n, 2), np.float64) if self.storage == STORE_F64 else np.zeros(n, np.uint32))
        lib().orc_volume_read(self.h, u64p(keys), vals.ctypes.data, n)
        return keys, vals

    def write(self, keys, vals):
        k = np.ascontiguousarray(keys, np.uint64)
        v = np.ascontiguousarray(vals, {STORE_F32: np.float32, STORE_F64: np.float64}.get(self.storage, np.uint32))
        if lib().orc_volume_write(self.h, u64p(k), v.ctypes.data, k.shape[0]):
            raise MemoryError

    def integrate(self, depth, cam, world_to_eye):
        """uint16 mm frames convert as DepthStream.cpp:89-90; float32 metres are used as they are."""
        m = np.ascontiguousarray(world_to_eye, np.float32).reshape(16)
        st = (ctypes.c_size_t * 4)()
        if np.asarray(depth).dtype == np.float32:
            d = np.ascontiguousarray(depth, np.float32)
            r = lib().orc_integrate_f32(self.h, f32p(d), ctypes.byref(camera(cam)), f32p(m), st)
        else:
            d = np.ascontiguousarray(depth, np.uint16)
            r = lib().orc_integrate(self.h, d.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                    ctypes.byref(camera(cam)), f32p(m), st)
        if r == ctypes.c_size_t(-1).value:
            raise ValueError("orc_integrate rejected its input")
        return {"keys": st[0], "unique": st[1], "new": st[2], "size": st[3]}

    def mesh(self):
        n = self.size()
        verts = np.zeros((3 * n + 1, 3), np.float32)
        nv = lib().orc_mesh(self.h, f32p(verts), 3 * n + 1)
        return verts[:nv].copy()

    def mesh_indexed(self):
        n = self.size()
        verts = np.zeros((3 * n + 1, 3), np.float32)
        idx = np.zeros(15 * n + 1, np.uint32)
        ni = ctypes.c_size_t()
        st = ctypes.c_int()
        nv = lib().orc_mesh_indexed(self.h, f32p(verts), 3 * n + 1,
                                    idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 15 * n + 1,
                                    ctypes.byref(ni), ctypes.byref(st))
        return verts[:nv].copy(), idx[:ni.value].copy(), bool(st.value)


def keygen(depth_m, cam, eye_to_world, desc):
    d = np.ascontiguousarray(depth_m, np.float32)
    out = np.zeros(d.size, np.uint64)
    m = np.ascontiguousarray(eye_to_world, np.float32).reshape(16)
    n = lib().orc_keygen(f32p(d), ctypes.byref(camera(cam)), f32p(m), ctypes.byref(desc), u64p(out))
    return out[:n].copy()


def radix_sort(keys, key_bits=32):
    k = np.ascontiguousarray(keys, np.uint64).copy()
    scr = np.zeros_like(k)
    lib().orc_radix_sort(u64p(k), k.shape[0], key_bits, u64p(scr))
    return k


def unique(keys):
    k = np.ascontiguousarray(keys, np.uint64).copy()
    n = lib().orc_unique(u64p(k), k.shape[0])
    return k[:n].copy()


def expand(keys, kb=10):
    k = np.zeros(max(8 * len(keys), 1), np.uint64)
    k[:len(keys)] = keys
    scr = np.zeros_like(k)
    n = lib().orc_expand(u64p(k), len(keys), kb, u64p(scr), k.shape[0])
    return k[:n].copy()


def set_union_u32(k1, v1, k2, vdef):
    """set_union with int32 values (the tie keeps range 1)."""
    k1 = np.ascontiguousarray(k1, np.uint64); k2 = np.ascontiguousarray(k2, np.uint64)
    v1 = np.ascontiguousarray(v1, np.int32)
    d = np.array([vdef], np.int32)
    ko = np.zeros(len(k1) + len(k2) + 1, np.uint64)
    vo = np.zeros(len(k1) + len(k2) + 1, np.int32)
    n = lib().orc_set_union(u64p(k1), len(k1), u64p(k2), len(k2), v1.ctypes.data, d.ctypes.data, 4,
                            u64p(ko), vo.ctypes.data)
    return ko[:n].copy(), vo[:n].copy()


def bin_points(points, world_to_clip, W, H):
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    m = np.ascontiguousarray(world_to_clip, np.float32).reshape(16)
    out = np.zeros((H, W, 3), np.float32)
    zb = np.zeros((H, W), np.float32)
    lib().orc_bin(f32p(p), p.shape[0], f32p(m), W, H, f32p(out), f32p(zb))
    return out


def horn_step(src, dst):
    s = np.ascontiguousarray(src, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dst, np.float32).reshape(-1, 3)
    out = np.zeros(16, np.float32)
    lib().orc_horn_step(f32p(s), f32p(d), s.shape[0], f32p(out))
    return out


def jacobi4(a):
    a = np.ascontiguousarray(a, np.float64).reshape(16)
    ev = np.zeros(4); V = np.zeros(16)
    lib().orc_jacobi4(a.ctypes.data_as(F64P), ev.ctypes.data_as(F64P), V.ctypes.data_as(F64P))
    return ev, V.reshape(4, 4).T  # columns are eigenvectors (row-major view)


def kahan_sum_f32(x):
    x = np.ascontiguousarray(x, np.float32)
    return lib().orc_kahan_sum_f32(f32p(x), x.shape[0])


def _icp_sigs(L):
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    L.orc_model_splat.argtypes = [ctypes.POINTER(OVolume), F32P, ctypes.POINTER(Camera), ctypes.c_float,
                                  ctypes.c_float, U64P]
    L.orc_model_splat.restype = None
    L.orc_model_resolve.argtypes = [ctypes.POINTER(OVolume), F32P, F32P, U64P, i32, i32, ctypes.c_float,
                                    ctypes.c_float, F32P]
    L.orc_model_resolve.restype = None
    L.orc_model_normals4.argtypes = [F32P, i32, i32, F32P, ctypes.c_float, F32P]
    L.orc_model_normals4.restype = None
    L.orc_depth_pyramid_level.argtypes = [F32P, i32, i32, F32P]
    L.orc_depth_pyramid_level.restype = None
    d = ctypes.c_double
    L.orc_icp_level_f64.argtypes = [F32P, i32, i32, d, d, d, d, F32P, F32P, ctypes.POINTER(Camera), F64P, F64P,
                                    i32, d, d, d, F64P]
    L.orc_icp_level_f64.restype = i32
    L.orc_icp_level_f64_prior.argtypes = [F32P, i32, i32, d, d, d, d, F32P, F32P, ctypes.POINTER(Camera), F64P,
                                          F64P, i32, d, d, d, d, F64P, F64P, F64P]
    L.orc_icp_level_f64_prior.restype = i32
    L.orc_icp_sums_f64.argtypes = [F32P, i32, i32, d, d, d, d, F32P, F32P, ctypes.POINTER(Camera), F64P, F64P, d,
                                   F64P]
    L.orc_icp_sums_f64.restype = None
    L.orc_icp_sums_f32terms.argtypes = [F32P, i32, F32P, F32P, ctypes.POINTER(Camera), F32P, F32P, ctypes.c_float,
                                        F64P]
    L.orc_icp_sums_f32terms.restype = None


def model_predict(vol, world_to_eye, cam, dmax=None, znear=0.1, disc=0.15):
    """Model point/normal maps (H, W, 4) of an oracle Volume seen from world_to_eye."""
    L = lib()
    if not hasattr(L, "_icp_ready"):
        _icp_sigs(L)
        L._icp_ready = True
    w2e = np.ascontiguousarray(world_to_eye, np.float32).reshape(16)
    c = camera(cam)
    W, H = cam.width, cam.height
    dm = vol.desc.vl if dmax is None else dmax
    zb = np.zeros(W * H, np.uint64)
    L.orc_model_splat(vol.h, f32p(w2e), ctypes.byref(c), dm, znear, u64p(zb))
    e2w = invert_transform(w2e)
    campos = np.ascontiguousarray(e2w[12:15], np.float32)
    pts = np.zeros((H, W, 4), np.float32)
    L.orc_model_resolve(vol.h, f32p(w2e), f32p(campos), u64p(zb), W, H, dm, znear, f32p(pts))
    nrm = np.zeros_like(pts)
    L.orc_model_normals4(f32p(pts), W, H, f32p(campos), disc, f32p(nrm))
    return pts, nrm


def depth_pyramid(depth_mm, levels=3):
    """Level maps in metres, binary32, as the GPU builds them."""
    L = lib()
    if not hasattr(L, "_icp_ready"):
        _icp_sigs(L)
        L._icp_ready = True
    d0 = (np.ascontiguousarray(depth_mm, np.uint16).astype(np.float32) * np.float32(0.001)).astype(np.float32)
    out = [d0]
    for _ in range(1, levels):
        src = out[-1]
        H, W = src.shape
        dst = np.zeros((H // 2, W // 2), np.float32)
        L.orc_depth_pyramid_level(f32p(np.ascontiguousarray(src)), W, H, f32p(dst))
        out.append(dst)
    return out


def icp_f64(depth_mm, cam, pts, nrm, model_w2e, guess_e2w, iters=(10, 5, 4), dist_thresh=0.1,
            eps_rot=2e-4, eps_trans=2e-4, gate_growth=1.0, eps_growth=4.0, motion_prior=0.0):
    """FP64 point-to-plane ICP over the pyramid; returns (pose, total iterations, last rms).
    Level l gates correspondences at dist_thresh x gate_growth^l and converges
    at eps_rot, eps_trans x eps_growth^l; motion_prior > 0 adds the prior on the
    motion from guess_e2w (kifi_icp_params.motion_prior), carried across levels."""
    L = lib()
    if not hasattr(L, "_icp_ready"):
        _icp_sigs(L)
        L._icp_ready = True
    pyr = depth_pyramid(depth_mm, len(iters))
    pose = np.ascontiguousarray(guess_e2w, np.float64).reshape(16).copy()
    W2E = np.ascontiguousarray(model_w2e, np.float64).reshape(16)
    c = camera(cam)
    P = np.ascontiguousarray(pts, np.float32)
    N = np.ascontiguousarray(nrm, np.float32)
    total, st = 0, np.zeros(2)
    centre = pose[12:15].copy()
    acc = np.zeros(6)
    for l in range(len(iters) - 1, -1, -1):
        d = pyr[l]
        s = 1.0 / (1 << l)
        n = L.orc_icp_level_f64_prior(f32p(d), d.shape[1], d.shape[0], cam.fx * s, cam.fy * s, cam.cx * s,
                                      cam.cy * s, f32p(P), f32p(N), ctypes.byref(c), W2E.ctypes.data_as(F64P),
                                      pose.ctypes.data_as(F64P), iters[l], dist_thresh * gate_growth ** l,
                                      eps_rot * eps_growth ** l, eps_trans * eps_growth ** l, float(motion_prior),
                                      centre.ctypes.data_as(F64P), acc.ctypes.data_as(F64P),
                                      st.ctypes.data_as(F64P))
        total += n
    return pose, total, st[0]


def icp_sums_f64(depth_mm, cam, pts, nrm, model_w2e, e2w, level=0, dist_thresh=0.1):
    """FP64 normal equations of one association of pyramid level `level` at the
    eye->world pose e2w: 30 sums in the layout of kifi_icp_normal_equations."""
    L = lib()
    if not hasattr(L, "_icp_ready"):
        _icp_sigs(L)
        L._icp_ready = True
    d = depth_pyramid(depth_mm, level + 1)[level]
    s = 1.0 / (1 << level)
    W2E = np.ascontiguousarray(model_w2e, np.float64).reshape(16)
    pose = np.ascontiguousarray(e2w, np.float64).reshape(16)
    out = np.zeros(30, np.float64)
    c = camera(cam)
    L.orc_icp_sums_f64(f32p(np.ascontiguousarray(d)), d.shape[1], d.shape[0], cam.fx * s, cam.fy * s, cam.cx * s,
                       cam.cy * s, f32p(np.ascontiguousarray(pts, np.float32)),
                       f32p(np.ascontiguousarray(nrm, np.float32)), ctypes.byref(c), W2E.ctypes.data_as(F64P),
                       pose.ctypes.data_as(F64P), dist_thresh, out.ctypes.data_as(F64P))
    return out


def icp_sums_f32terms(depth_mm, cam, pts, nrm, model_w2e, e2w, level=0, dist_thresh=0.1):
    """The binary32 terms of the GPU's association (pose rounded to binary32 as the
    GPU does), summed in double: the reference for the GPU's summation error."""
    L = lib()
    if not hasattr(L, "_icp_ready"):
        _icp_sigs(L)
        L._icp_ready = True
    d = depth_pyramid(depth_mm, level + 1)[level]
    w2e = np.ascontiguousarray(model_w2e, np.float32).reshape(16)
    pose_f = np.ascontiguousarray(np.asarray(e2w, np.float64).reshape(16).astype(np.float32))
    out = np.zeros(30, np.float64)
    c = camera(cam)
    L.orc_icp_sums_f32terms(f32p(np.ascontiguousarray(d)), level, f32p(np.ascontiguousarray(pts, np.float32)),
                            f32p(np.ascontiguousarray(nrm, np.float32)), ctypes.byref(c), f32p(w2e), f32p(pose_f),
                            dist_thresh, out.ctypes.data_as(F64P))
    return out


def icp_align_ref(depth_mm, cam, model_pts4, model_e2w, guess_e2w, iters=4):
    """ICP::Align restated (ICP.cpp:9-265): window search + Horn, binary32."""
    L = lib()
    if not hasattr(L, "_ref_ready"):
        L.orc_icp_align_ref.argtypes = [F32P, ctypes.POINTER(Camera), F32P, F32P, ctypes.c_int, ctypes.c_int,
                                        F32P, ctypes.c_int, F32P]
        L.orc_icp_align_ref.restype = ctypes.c_size_t
        L._ref_ready = True
    d = (np.ascontiguousarray(depth_mm, np.uint16).astype(np.float32) * np.float32(0.001)).astype(np.float32)
    m3 = np.ascontiguousarray(np.asarray(model_pts4, np.float32)[..., :3])
    H, W = m3.shape[:2]
    me = np.ascontiguousarray(model_e2w, np.float32).reshape(16)
    g = np.ascontiguousarray(guess_e2w, np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    n = L.orc_icp_align_ref(f32p(d), ctypes.byref(camera(cam)), f32p(g), f32p(m3), W, H, f32p(me), iters,
                            f32p(out))
    return out, n


def partial_sum_exclusive(x):
    x = np.ascontiguousarray(x, np.uint32)
    out = np.zeros_like(x)
    lib().orc_partial_sum_exclusive_u32(x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), x.shape[0],
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out
