"""ctypes access to the oracle (oracle/liboracle.so) and to the reference's
own host math compiled into oracle/_ref/libref_math.so.

Test infrastructure only: the oracle is the checker the HIP pipeline is
compared against.
"""
import contextlib
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libref_math.so")

P = C.c_void_p


class OrVolume(C.Structure):
    _fields_ = [("dim", C.c_int * 3), ("origin", C.c_float * 3), ("voxel", C.c_float), ("mu", C.c_float),
                ("max_w", C.c_float)]


class OrSessionDesc(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int),
        ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
        ("levels", C.c_int), ("iters", C.c_int * 3),
        ("icp_dist2", C.c_float), ("icp_min_valid", C.c_int),
        ("max_motion_m", C.c_float), ("max_motion_deg", C.c_float),
        ("vol", OrVolume),
        ("bil_sigma_s", C.c_float), ("bil_sigma_r", C.c_float), ("bil_radius", C.c_int),
        ("fp32", C.c_int), ("icp_rel_pivot", C.c_float),
    ]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


_or = None


def oracle():
    global _or
    if _or is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError("oracle not built: make -C oracle")
        _or = _load_oracle(ORACLE_PATH)
    return _or


@contextlib.contextmanager
def using(path):
    """Route every oracle call of this module through the build at `path`
    (one of the variant builds of oracle/Makefile) inside the block."""
    global _or
    saved = oracle()
    _or = _load_oracle(path)
    try:
        yield
    finally:
        _or = saved


def _load_oracle(path):
    L = C.CDLL(path)
    L.or_icp_system_d.restype = C.c_longlong
    L.or_icp_system_d.argtypes = [P, C.c_int, C.c_int, P, P, P, P, C.c_float, P, P, P]
    L.or_bilateral_d.argtypes = [P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, P]
    L.or_bilateral_f.argtypes = [P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, P]
    L.or_gaussian_weights.argtypes = [C.c_float, C.c_int, P]
    L.or_pyr_down_d.argtypes = [P, C.c_int, C.c_int, P]
    L.or_pyr_down_f.argtypes = [P, C.c_int, C.c_int, P]
    L.or_intrinsics_pyramid.argtypes = [P, C.c_int, P]
    L.or_vnmap_d.argtypes = [P, C.c_int, C.c_int, P, C.c_int, P, P]
    L.or_cholesky_solve.argtypes = [P, P, P]
    L.or_cholesky_solve.restype = C.c_int
    L.or_expmap.argtypes = [P, P]
    L.or_rigid_inverse.argtypes = [P, P]
    L.or_matmul4.argtypes = [P, P, P]
    L.or_icp_track_d.argtypes = [P, C.c_int, C.c_int, P, C.c_int, P, P, P, C.c_float, C.c_int, C.c_float,
                                 C.c_float, C.c_float, P, P, P]
    L.or_cholesky_status.argtypes = [P, P, C.c_double, P]
    L.or_cholesky_status.restype = C.c_int
    L.or_icp_track_d.restype = C.c_int
    L.or_integrate_d.argtypes = [P, C.c_int, C.c_int, P, P, P, C.POINTER(OrVolume), P, P, P, P, P]
    L.or_integrate_f.argtypes = [P, C.c_int, C.c_int, P, P, P, C.POINTER(OrVolume), P, P, P, P, P]
    L.or_raycast_d.argtypes = [P, P, C.POINTER(OrVolume), C.c_int, C.c_int, P, P, P, P, P, P, P, P]
    L.or_raycast_f.argtypes = [P, P, C.POINTER(OrVolume), C.c_int, C.c_int, P, P, P, P, P, P, P, P]
    L.or_session_create.argtypes = [C.POINTER(OrSessionDesc), P]
    L.or_session_create.restype = P
    L.or_session_destroy.argtypes = [P]
    L.or_fuse_frame.argtypes = [P, P, P]
    L.or_fuse_frame.restype = C.c_int
    L.or_session_pose.argtypes = [P, P]
    L.or_session_maps.argtypes = [P, C.c_int, P, P, P]
    L.or_session_volume.argtypes = [P, P, P]
    L.or_session_filtered.argtypes = [P, C.c_int, P]
    L.or_depth_to_m.argtypes = [P, C.c_longlong, P]
    L.or_num_threads.restype = C.c_int
    return L


_ref = None


def ref_math():
    """The reference's own _yooji_math_inlines.h, or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        L = C.CDLL(REF_PATH)
        L.ref_expmap.argtypes = [P, P]
        L.ref_inverse4x4.argtypes = [P, P]
        L.ref_inverse4x4.restype = C.c_int
        L.ref_matmul4.argtypes = [P, P, P]
        L.ref_norm_inf.argtypes = [P]
        L.ref_norm_inf.restype = C.c_float
        _ref = L
    return _ref


def volume(desc):
    v = OrVolume()
    for k in range(3):
        v.dim[k] = desc.vol_dim[k]
        v.origin[k] = desc.origin[k]
    v.voxel = desc.voxel
    v.mu = desc.mu
    v.max_w = desc.max_weight
    return v


def K_levels(desc, levels=3):
    K0 = np.array([desc.fx, desc.fy, desc.cx, desc.cy], np.float32)
    out = np.zeros(4 * levels, np.float32)
    oracle().or_intrinsics_pyramid(_p(K0), levels, _p(out))
    return out.reshape(levels, 4)


def depth_to_m(mm):
    mm = np.ascontiguousarray(mm, np.uint16)
    out = np.zeros(mm.shape, np.float32)
    oracle().or_depth_to_m(_p(mm), mm.size, _p(out))
    return out


def bilateral(mm, sigma_s, sigma_r, radius, fp32=False):
    mm = np.ascontiguousarray(mm, np.uint16)
    H, W = mm.shape
    out = np.zeros((H, W), np.float32 if fp32 else np.float64)
    f = oracle().or_bilateral_f if fp32 else oracle().or_bilateral_d
    f(_p(mm), W, H, sigma_s, sigma_r, radius, _p(out))
    return out


def gaussian_weights(sigma_s, radius):
    g = np.zeros(2 * radius + 1, np.float32)
    oracle().or_gaussian_weights(sigma_s, radius, _p(g))
    return g


def pyr_down(img):
    img = np.ascontiguousarray(img, np.float64)
    H, W = img.shape
    out = np.zeros((H // 2, W // 2), np.float64)
    oracle().or_pyr_down_d(_p(img), W, H, _p(out))
    return out


def vnmap(depth, K, inplace=False):
    depth = np.ascontiguousarray(depth, np.float32)
    H, W = depth.shape
    K = np.ascontiguousarray(K, np.float32)
    v = np.zeros((H, W, 3), np.float64)
    n = np.zeros((H, W, 3), np.float64)
    oracle().or_vnmap_d(_p(depth), W, H, _p(K), 1 if inplace else 0, _p(v), _p(n))
    return v, n


def icp_system(depth1, K, mvtx, mnrm, T21, dist2):
    depth1 = np.ascontiguousarray(depth1, np.float32)
    H, W = depth1.shape
    ATA = np.zeros(36)
    ATb = np.zeros(6)
    mask = np.zeros((H, W), np.uint8)
    n = oracle().or_icp_system_d(_p(depth1), W, H, _p(np.ascontiguousarray(K, np.float32)),
                                 _p(np.ascontiguousarray(mvtx, np.float32)),
                                 _p(np.ascontiguousarray(mnrm, np.float32)),
                                 _p(np.ascontiguousarray(T21, np.float32).reshape(16)), dist2, _p(ATA), _p(ATb),
                                 _p(mask))
    return ATA.reshape(6, 6), ATb, n, mask


def cholesky_status(A, b, rel):
    A = np.ascontiguousarray(A, np.float64).reshape(36)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(6)
    st = oracle().or_cholesky_status(_p(A), _p(b), float(rel), _p(x))
    return st, x


def cholesky_solve(A, b):
    A = np.ascontiguousarray(A, np.float64).reshape(36)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(6)
    ok = oracle().or_cholesky_solve(_p(A), _p(b), _p(x))
    return bool(ok), x


def expmap(x):
    x = np.ascontiguousarray(x, np.float64)
    T = np.zeros(16)
    oracle().or_expmap(_p(x), _p(T))
    return T.reshape(4, 4)


def rigid_inverse(T):
    T = np.ascontiguousarray(T, np.float64).reshape(16)
    o = np.zeros(16)
    oracle().or_rigid_inverse(_p(T), _p(o))
    return o.reshape(4, 4)


def icp_track(depths, K, levels, iters, mvtx, mnrm, dist2, T_gc, min_valid=0, max_m=0.0, max_deg=0.0, T_init=None,
              rel_pivot=0.0):
    d = [np.ascontiguousarray(x, np.float32) for x in depths]
    mv = [np.ascontiguousarray(x, np.float32) for x in mvtx]
    mn = [np.ascontiguousarray(x, np.float32) for x in mnrm]
    H0, W0 = d[0].shape
    dp = (P * levels)(*[_p(x) for x in d])
    mvp = (P * levels)(*[_p(x) for x in mv])
    mnp = (P * levels)(*[_p(x) for x in mn])
    it = (C.c_int * 3)(*iters)
    T = np.ascontiguousarray(T_gc, np.float64).reshape(16).copy()
    Ti = None if T_init is None else np.ascontiguousarray(T_init, np.float64).reshape(16)
    nd = C.c_int()
    st = oracle().or_icp_track_d(dp, W0, H0, _p(np.ascontiguousarray(K, np.float32)), levels, it, mvp, mnp, dist2,
                                 min_valid, max_m, max_deg, rel_pivot, _p(Ti), _p(T), C.byref(nd))
    return st, T.reshape(4, 4), nd.value


def integrate(mm, K, Tgc_f, Tcg_f, vol, F, Wt, fp32=False):
    mm = np.ascontiguousarray(mm, np.uint16)
    H, W = mm.shape
    nb = (vol.dim[0] // 8) * (vol.dim[1] // 8) * (vol.dim[2] // 8)
    bm = np.zeros(nb, np.uint8)
    upd = np.zeros(F.shape, np.uint8)
    counts = np.zeros(2, np.int64)
    f = oracle().or_integrate_f if fp32 else oracle().or_integrate_d
    f(_p(mm), W, H, _p(np.ascontiguousarray(K, np.float32)), _p(np.ascontiguousarray(Tgc_f, np.float32)),
      _p(np.ascontiguousarray(Tcg_f, np.float32)), C.byref(vol), _p(F), _p(Wt), _p(bm), _p(upd), _p(counts))
    return bm, upd, counts


def integrate_color(mm, rgb, K, Tgc_f, Tcg_f, vol, F, Wt, Cvol):
    """Production integrate (g_ctvc) with a registered colour frame; rgb is
    (3, H, W) uint8, Cvol (3, N) uint8 updated in place."""
    mm = np.ascontiguousarray(mm, np.uint16)
    rgb = np.ascontiguousarray(rgb, np.uint8)
    H, W = mm.shape
    assert rgb.shape == (3, H, W) and Cvol.dtype == np.uint8 and Cvol.flags.c_contiguous
    nb = (vol.dim[0] // 8) * (vol.dim[1] // 8) * (vol.dim[2] // 8)
    bm = np.zeros(nb, np.uint8)
    upd = np.zeros(F.shape, np.uint8)
    counts = np.zeros(2, np.int64)
    oracle().or_integrate_color_d(_p(mm), _p(rgb), W, H, _p(np.ascontiguousarray(K, np.float32)),
                                  _p(np.ascontiguousarray(Tgc_f, np.float32)),
                                  _p(np.ascontiguousarray(Tcg_f, np.float32)), C.byref(vol), _p(F), _p(Wt),
                                  _p(Cvol), _p(bm), _p(upd), _p(counts))
    return bm, upd, counts


def depth_reproject(src, K_src, T_dst_src, dst_shape, K_dst):
    src = np.ascontiguousarray(src, np.uint16)
    sh, sw = src.shape
    out = np.zeros(dst_shape, np.uint16)
    oracle().or_depth_reproject(_p(src), sw, sh, _p(np.ascontiguousarray(K_src, np.float32)),
                                _p(np.ascontiguousarray(T_dst_src, np.float32)), _p(out), dst_shape[1], dst_shape[0],
                                _p(np.ascontiguousarray(K_dst, np.float32)))
    return out


def volume_carve(F, Wt, Cvol, bmask, valid, vol):
    """g_uvv on FP64 F/Wt (updated in place), optional (3, N) colour and the
    brick mask (updated in place); returns the number of bricks reset."""
    oracle().or_volume_carve.restype = C.c_longlong
    valid = np.ascontiguousarray(valid, np.uint8)
    return oracle().or_volume_carve(_p(F), _p(Wt), _p(Cvol), _p(bmask), _p(valid), C.byref(vol))


def depth_crop(mm, K, T_cg_f, vol):
    out = np.ascontiguousarray(mm, np.uint16).copy()
    H, W = out.shape
    oracle().or_depth_crop(_p(out), W, H, _p(np.ascontiguousarray(K, np.float32)),
                           _p(np.ascontiguousarray(T_cg_f, np.float32)), C.byref(vol))
    return out


def raycast_color(F, Wt, Cvol, vol, W, H, K, Tgc_f, Tcg_f):
    """Raycast (FP64 values) with the texture map: returns depth, vtx, nrm,
    img and tex (3, H, W) uint8."""
    F = np.ascontiguousarray(F, np.float32)
    Wt = np.ascontiguousarray(Wt, np.float32)
    Cvol = np.ascontiguousarray(Cvol, np.uint8)
    depth = np.zeros((H, W))
    vtx = np.zeros((H, W, 3))
    nrm = np.zeros((H, W, 3))
    img = np.zeros((H, W), np.uint8)
    tex = np.zeros((3, H, W), np.uint8)
    oracle().or_raycast_color_d(_p(F), _p(Wt), _p(Cvol), C.byref(vol), W, H, _p(np.ascontiguousarray(K, np.float32)),
                                _p(np.ascontiguousarray(Tgc_f, np.float32)),
                                _p(np.ascontiguousarray(Tcg_f, np.float32)), _p(depth), _p(vtx), _p(nrm), _p(img),
                                _p(tex))
    return depth, vtx, nrm, img, tex


def raycast(F, Wt, vol, W, H, K, Tgc_f, Tcg_f, fp32=False):
    F = np.ascontiguousarray(F, np.float32)
    Wt = np.ascontiguousarray(Wt, np.float32)
    dt = np.float32 if fp32 else np.float64
    depth = np.zeros((H, W), dt)
    vtx = np.zeros((H, W, 3), dt)
    nrm = np.zeros((H, W, 3), dt)
    img = np.zeros((H, W), np.uint8)
    stats = np.zeros(4, np.int64)
    f = oracle().or_raycast_f if fp32 else oracle().or_raycast_d
    f(_p(F), _p(Wt), C.byref(vol), W, H, _p(np.ascontiguousarray(K, np.float32)),
      _p(np.ascontiguousarray(Tgc_f, np.float32)), _p(np.ascontiguousarray(Tcg_f, np.float32)), _p(depth), _p(vtx),
      _p(nrm), _p(img), _p(stats))
    return depth, vtx, nrm, img, stats


class OracleSession:
    """The oracle's whole per-frame pipeline (FP64 values, or FP32 with fp32=True)."""

    def __init__(self, desc, T_gc0, fp32=False):
        d = OrSessionDesc()
        for f in ["width", "height", "fx", "fy", "cx", "cy", "levels", "icp_dist2", "icp_min_valid",
                  "max_motion_m", "max_motion_deg", "bil_sigma_s", "bil_sigma_r", "bil_radius", "icp_rel_pivot"]:
            setattr(d, f, getattr(desc, f))
        for k in range(3):
            d.iters[k] = desc.iters[k]
        d.vol = volume(desc)
        d.fp32 = 1 if fp32 else 0
        self.desc = desc
        self.d = d
        T = np.ascontiguousarray(T_gc0, np.float64).reshape(16)
        self.h = oracle().or_session_create(C.byref(d), _p(T))

    def close(self):
        if self.h:
            oracle().or_session_destroy(self.h)
            self.h = None

    def fuse(self, mm):
        mm = np.ascontiguousarray(mm, np.uint16)
        ms = np.zeros(5)
        st = oracle().or_fuse_frame(self.h, _p(mm), _p(ms))
        return st, ms

    def pose(self):
        T = np.zeros(16)
        oracle().or_session_pose(self.h, _p(T))
        return T.reshape(4, 4)

    def maps(self, level):
        H, W = self.desc.height >> level, self.desc.width >> level
        d = np.zeros((H, W), np.float32)
        v = np.zeros((H, W, 3), np.float32)
        n = np.zeros((H, W, 3), np.float32)
        oracle().or_session_maps(self.h, level, _p(d), _p(v), _p(n))
        return d, v, n

    def volume(self):
        dd = self.desc.vol_dim
        F = np.zeros((dd[2], dd[1], dd[0]), np.float32)
        Wt = np.zeros_like(F)
        oracle().or_session_volume(self.h, _p(F), _p(Wt))
        return F, Wt

    def filtered(self, level):
        H, W = self.desc.height >> level, self.desc.width >> level
        o = np.zeros((H, W), np.float32)
        oracle().or_session_filtered(self.h, level, _p(o))
        return o


def pose_error(T_a, T_b):
    """Translation (m) and geodesic rotation angle (deg) between two poses
    (the metric of z_yooji_evaluation.cpp:153-160)."""
    Ra, Rb = T_a[:3, :3], T_b[:3, :3]
    c = (np.trace(Ra @ Rb.T) - 1.0) / 2.0
    ang = np.degrees(np.arccos(np.clip(c, -1.0, 1.0)))
    # compare camera centres (translation of the inverse poses)
    ca = -Ra.T @ T_a[:3, 3]
    cb = -Rb.T @ T_b[:3, 3]
    return float(np.linalg.norm(ca - cb)), float(ang)
