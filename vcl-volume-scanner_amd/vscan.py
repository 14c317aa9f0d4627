"""ctypes binding of the fusion library's C ABI (include/vscan.h) and of the
synthetic depth-stream generator (include/vscan_scene.h), plus the workload
presets the tests and the benchmark share.

The product is libvscan.so (HIP kernels for gfx950 behind the C ABI). This
module only marshals arguments; it never computes a stage itself, and it
raises if the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VSCAN_LIB may name another build of the library (build.py --variant)
LIB_PATH = os.environ.get("VSCAN_LIB") or os.path.join(HERE, "libvscan.so")
SCENE_PATH = os.path.join(HERE, "libvscan_scene.so")

(VS_OK, VS_EINVAL, VS_ENOMEM, VS_EHIP, VS_TRACK_LOST, VS_MOTION_REJECTED, VS_BUDGET, VS_TOO_FEW_POINTS,
 VS_SINGULAR) = range(9)
VS_TSDF_F32, VS_TSDF_F16 = 0, 1
VS_WEIGHT_F32, VS_WEIGHT_U8 = 0, 1
VS_INTEGRATE_DEPTH, VS_INTEGRATE_COLOR = 0, 1
VS_DEPTH_MM_U16, VS_DEPTH_M_F32 = 0, 1
VS_TUNE_OFF, VS_TUNE_ONLINE, VS_TUNE_CACHED = 0, 1, 2


class SessionDesc(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int),
        ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
        ("levels", C.c_int), ("iters", C.c_int * 3),
        ("icp_dist2", C.c_float), ("icp_min_valid", C.c_int),
        ("max_motion_m", C.c_float), ("max_motion_deg", C.c_float),
        ("vol_dim", C.c_int * 3), ("origin", C.c_float * 3),
        ("voxel", C.c_float), ("mu", C.c_float), ("max_weight", C.c_float),
        ("tsdf_type", C.c_int), ("weight_type", C.c_int),
        ("bil_sigma_s", C.c_float), ("bil_sigma_r", C.c_float), ("bil_radius", C.c_int),
        ("integrate_mode", C.c_int), ("depth_units", C.c_int), ("icp_rel_pivot", C.c_float),
    ]


class SnapshotHeader(C.Structure):
    _fields_ = [("magic", C.c_char * 8), ("version", C.c_uint32), ("header_bytes", C.c_uint32),
                ("desc", SessionDesc), ("T_gc", C.c_double * 16), ("frames", C.c_int32),
                ("bricks_total", C.c_uint32), ("bricks_stored", C.c_uint32), ("value_bytes", C.c_uint32),
                ("weight_bytes", C.c_uint32), ("color_bytes", C.c_uint32)]


class ServerDesc(C.Structure):
    _fields_ = [("max_slots", C.c_int), ("tsdf_type", C.c_int), ("weight_type", C.c_int),
                ("hbm_budget", C.c_ulonglong)]


class Prim(C.Structure):
    _fields_ = [("kind", C.c_int), ("p", C.c_double * 6)]


class Scene(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int),
        ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
        ("n_prims", C.c_int), ("prims", Prim * 64),
        ("radius0", C.c_double), ("radius1", C.c_double), ("height0", C.c_double), ("height_amp", C.c_double),
        ("deg_per_frame", C.c_double), ("phase0_deg", C.c_double),
        ("target", C.c_double * 3), ("n_frames", C.c_int),
        ("seed", C.c_uint64), ("dropout", C.c_double), ("noise", C.c_int), ("max_range", C.c_double),
    ]


_lib = None
_scene = None

P = C.c_void_p
dptr = C.POINTER(C.c_double)
fptr = C.POINTER(C.c_float)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib():
    """The fusion library; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run vcl-volume-scanner_amd/build.py")
        L = C.CDLL(LIB_PATH)
        L.vs_status_string.restype = C.c_char_p
        L.vs_last_error.restype = C.c_char_p
        L.vs_session_bytes.restype = C.c_size_t
        L.vs_session_bytes.argtypes = [C.POINTER(SessionDesc)]
        L.vs_session_desc_default.argtypes = [C.POINTER(SessionDesc), C.c_int, C.c_int, C.c_int, C.c_float]
        L.vs_session_desc_default.restype = None
        for name in ["vs_ctx_create", "vs_ctx_destroy", "vs_session_create", "vs_session_destroy",
                     "vs_session_reset", "vs_session_desc_get", "vs_session_sync", "vs_set_depth", "vs_bilateral",
                     "vs_pyramid", "vs_vertex_normal_maps", "vs_track", "vs_integrate", "vs_raycast",
                     "vs_fuse_frame", "vs_fuse_batch", "vs_get_frame_status", "vs_get_pose", "vs_set_pose",
                     "vs_get_filtered", "vs_set_filtered", "vs_get_maps", "vs_set_maps", "vs_volume_download",
                     "vs_volume_upload", "vs_icp_system", "vs_session_set_timing", "vs_get_stage_ms",
                     "vs_get_integrate_counts", "vs_raycast_counts", "vs_ctx_set_timing", "vs_ctx_stage_ms",
                     "vs_server_create", "vs_server_destroy", "vs_server_open", "vs_server_close",
                     "vs_server_step", "vs_server_usage", "vs_set_color", "vs_get_texture",
                     "vs_volume_download_color", "vs_session_snapshot_bytes", "vs_session_snapshot",
                     "vs_session_restore", "vs_server_evict", "vs_server_open_snapshot", "vs_depth_reproject",
                     "vs_depth_crop_to_volume", "vs_volume_carve", "vs_ctx_set_tuning", "vs_ctx_tuning",
                     "vs_bilateral_filter_depth", "vs_integrate_depth", "vs_render_maps", "vs_icp_system_noalloc",
                     "vs_vertex_normal_maps_inplace"]:
            getattr(L, name).restype = C.c_int
        L.vs_ctx_set_tuning.argtypes = [P, C.c_int, C.c_char_p]
        L.vs_ctx_tuning.argtypes = [P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.vs_ctx_create.argtypes = [C.c_int, C.POINTER(P)]
        L.vs_session_create.argtypes = [P, C.POINTER(SessionDesc), P, C.POINTER(P)]
        L.vs_session_reset.argtypes = [P, P, P]
        L.vs_set_depth.argtypes = [P, P, C.c_int, P]
        L.vs_bilateral.argtypes = [P, P]
        L.vs_bilateral_filter_depth.argtypes = [P, P, C.c_int, P, P]
        L.vs_integrate_depth.argtypes = [P, P, P]
        L.vs_render_maps.argtypes = [P, C.c_int, P, P, P, P, P]
        L.vs_pyramid.argtypes = [P, P]
        L.vs_vertex_normal_maps.argtypes = [P, C.c_int, P, P, P]
        L.vs_vertex_normal_maps_inplace.argtypes = [P, C.c_int, P, P, P]
        L.vs_track.argtypes = [P, P, P]
        L.vs_integrate.argtypes = [P, P]
        L.vs_raycast.argtypes = [P, C.c_int, P]
        L.vs_fuse_frame.argtypes = [P, P, C.c_int, P]
        L.vs_fuse_batch.argtypes = [P, P, P, C.c_int, C.c_int, P]
        L.vs_get_frame_status.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.vs_get_pose.argtypes = [P, P]
        L.vs_set_pose.argtypes = [P, P, P]
        L.vs_get_filtered.argtypes = [P, C.c_int, P]
        L.vs_set_filtered.argtypes = [P, C.c_int, P]
        L.vs_get_maps.argtypes = [P, C.c_int, P, P, P, P]
        L.vs_set_maps.argtypes = [P, C.c_int, P, P]
        L.vs_volume_download.argtypes = [P, P, P, P]
        L.vs_volume_upload.argtypes = [P, P, P]
        L.vs_icp_system.argtypes = [P, C.c_int, P, P, P, C.POINTER(C.c_longlong), P]
        L.vs_icp_system_noalloc.argtypes = [P, C.c_int, P, P, P, P]
        L.vs_session_set_timing.argtypes = [P, C.c_int]
        L.vs_get_stage_ms.argtypes = [P, P]
        L.vs_get_integrate_counts.argtypes = [P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.vs_raycast_counts.argtypes = [P, C.c_int, P]
        L.vs_ctx_set_timing.argtypes = [P, C.c_int]
        L.vs_ctx_stage_ms.argtypes = [P, P]
        L.vs_set_color.argtypes = [P, P, C.c_int, P]
        L.vs_depth_reproject.argtypes = [P, P, C.c_int, C.c_int, P, P, C.c_int, P, P]
        L.vs_volume_carve.argtypes = [P, P, C.c_int, C.POINTER(C.c_uint)]
        L.vs_depth_crop_to_volume.argtypes = [P, P, C.c_int, C.c_int, P, P, C.c_int, P]
        L.vs_session_snapshot_bytes.argtypes = [P, C.POINTER(C.c_size_t)]
        L.vs_session_snapshot.argtypes = [P, P, C.c_size_t, C.POINTER(C.c_size_t)]
        L.vs_session_restore.argtypes = [P, P, C.c_size_t, C.POINTER(P)]
        L.vs_server_evict.argtypes = [P, C.c_int, P, C.c_size_t, C.POINTER(C.c_size_t)]
        L.vs_server_open_snapshot.argtypes = [P, P, C.c_size_t, C.POINTER(C.c_int)]
        L.vs_get_texture.argtypes = [P, C.c_int, P]
        L.vs_volume_download_color.argtypes = [P, P]
        L.vs_server_create.argtypes = [P, C.POINTER(ServerDesc), C.POINTER(P)]
        L.vs_server_destroy.argtypes = [P]
        L.vs_server_open.argtypes = [P, C.POINTER(SessionDesc), P, C.POINTER(C.c_int)]
        L.vs_server_close.argtypes = [P, C.c_int]
        L.vs_server_session.argtypes = [P, C.c_int]
        L.vs_server_session.restype = P
        L.vs_server_step.argtypes = [P, P, P, C.c_int, C.c_int, P]
        L.vs_server_usage.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_ulonglong)]
        _lib = L
    return _lib


def scene_lib():
    global _scene
    if _scene is None:
        if not os.path.exists(SCENE_PATH):
            raise RuntimeError(f"{SCENE_PATH} is not built: run vcl-volume-scanner_amd/build.py")
        S = C.CDLL(SCENE_PATH)
        S.vs_scene_preset.argtypes = [C.c_int, C.c_uint64, C.c_int, C.c_int, C.POINTER(Scene)]
        S.vs_scene_pose.argtypes = [C.POINTER(Scene), C.c_int, P]
        S.vs_scene_render.argtypes = [C.POINTER(Scene), C.c_int, P]
        S.vs_scene_render_pose.argtypes = [C.POINTER(Scene), P, C.c_int, P]
        S.vs_scene_render_exact.argtypes = [C.POINTER(Scene), P, P]
        S.vs_scene_sdf.argtypes = [C.POINTER(Scene), P]
        S.vs_scene_sdf.restype = C.c_double
        for f in ["vs_scene_preset", "vs_scene_pose", "vs_scene_render", "vs_scene_render_pose",
                  "vs_scene_render_exact"]:
            getattr(S, f).restype = None
        _scene = S
    return _scene


class VsError(RuntimeError):
    def __init__(self, code, where):
        L = lib()
        super().__init__(f"{where}: {L.vs_status_string(code).decode()} ({code}): {L.vs_last_error().decode()}")
        self.code = code


def check(code, where):
    if code != VS_OK:
        raise VsError(code, where)


# ---------------------------------------------------------------------------
# synthetic scenes
# ---------------------------------------------------------------------------
class SceneStream:
    """Seeded depth stream: uint16 mm frames and ground-truth poses."""

    def __init__(self, config, seed=1, width=0, height=0, noise=True, dropout=None):
        self.sc = Scene()
        scene_lib().vs_scene_preset(config, seed, width, height, C.byref(self.sc))
        if not noise:
            self.sc.noise = 0
        if dropout is not None:
            self.sc.dropout = dropout
        self.W, self.H = self.sc.width, self.sc.height

    @property
    def K(self):
        return (self.sc.fx, self.sc.fy, self.sc.cx, self.sc.cy)

    def pose(self, frame):
        T = np.zeros(16, np.float64)
        scene_lib().vs_scene_pose(C.byref(self.sc), frame, _ptr(T))
        return T.reshape(4, 4)

    def frame(self, frame):
        d = np.zeros((self.H, self.W), np.uint16)
        scene_lib().vs_scene_render(C.byref(self.sc), frame, _ptr(d))
        return d

    def frame_at(self, T, frame=0):
        d = np.zeros((self.H, self.W), np.uint16)
        T = np.ascontiguousarray(T, np.float64)
        scene_lib().vs_scene_render_pose(C.byref(self.sc), _ptr(T), frame, _ptr(d))
        return d

    def exact_depth(self, T):
        d = np.zeros((self.H, self.W), np.float64)
        T = np.ascontiguousarray(T, np.float64)
        scene_lib().vs_scene_render_exact(C.byref(self.sc), _ptr(T), _ptr(d))
        return d

    def sdf(self, pts):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        out = np.empty(len(pts))
        S = scene_lib()
        for i, p in enumerate(pts):
            out[i] = S.vs_scene_sdf(C.byref(self.sc), _ptr(p))
        return out


# ---------------------------------------------------------------------------
# workloads (the five configurations the measurements are quoted on)
# ---------------------------------------------------------------------------
WORKLOADS = {
    # plain, CPU-runnable: 30 x 320x240, 128^3 FP32 over a 1 m cube, mu = 3 voxels
    "plain_320x240_128": dict(config=0, W=320, H=240, n=128, cube=1.0, origin=(-0.5, -0.5, -0.3), mu_vox=3.0,
                              tsdf=VS_TSDF_F32, weight=VS_WEIGHT_F32, max_w=100.0, frames=30, iters=(10, 5, 4)),
    # plain on GPU: 640x480, 256^3 FP32 over 1.2 m, helical orbit, 300 frames
    "plain_640x480_256": dict(config=1, W=640, H=480, n=256, cube=1.2, origin=(-0.6, -0.6, -0.35), mu_vox=3.0,
                              tsdf=VS_TSDF_F32, weight=VS_WEIGHT_F32, max_w=100.0, frames=300, iters=(10, 5, 4)),
    # precision form: 640x480, 512^3 FP16 + u8 (cap 128) over 1.5 m, 20 random primitives
    "fp16_640x480_512": dict(config=2, W=640, H=480, n=512, cube=1.5, origin=(-0.75, -0.75, -0.4), mu_vox=3.0,
                             tsdf=VS_TSDF_F16, weight=VS_WEIGHT_U8, max_w=128.0, frames=300, iters=(10, 5, 4)),
}


def make_desc(w, tsdf=None, weight=None, n=None, W=None, H=None, K=None, integrate_mode=VS_INTEGRATE_DEPTH,
              depth_units=VS_DEPTH_MM_U16):
    d = SessionDesc()
    W = W or w["W"]
    H = H or w["H"]
    n = n or w["n"]
    lib().vs_session_desc_default(C.byref(d), W, H, n, w["cube"])
    if K is not None:
        d.fx, d.fy, d.cx, d.cy = K
    vox = w["cube"] / n
    d.voxel = vox
    d.mu = w["mu_vox"] * vox
    for k in range(3):
        d.origin[k] = w["origin"][k]
    d.max_weight = w["max_w"]
    d.tsdf_type = w["tsdf"] if tsdf is None else tsdf
    d.weight_type = w["weight"] if weight is None else weight
    for k in range(3):
        d.iters[k] = w["iters"][k]
    d.bil_sigma_s = 4.0
    d.bil_sigma_r = 0.030
    d.bil_radius = 2
    d.icp_dist2 = (0.055 * w["cube"]) ** 2
    d.integrate_mode = integrate_mode
    d.depth_units = depth_units
    d.icp_rel_pivot = 0.0
    return d


# ---------------------------------------------------------------------------
# handles
# ---------------------------------------------------------------------------
class Context:
    def __init__(self, device=0):
        self.h = P()
        check(lib().vs_ctx_create(device, C.byref(self.h)), "vs_ctx_create")

    def close(self):
        if self.h:
            lib().vs_ctx_destroy(self.h)
            self.h = P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_timing(self, on):
        check(lib().vs_ctx_set_timing(self.h, 1 if on else 0), "vs_ctx_set_timing")

    # integrate workgroup counts the tuner tries, in the order of vs_ctx_tuning's means
    INTEG_GRIDS = (1024, 2048, 4096, 8192, 16384, 32768)
    # shading candidates (ray block, w = waves per SIMD the registers aim for; none = 4)
    SHADE_OPTS = ("8x8", "32x2", "8x8w5", "32x2w5")

    def set_tuning(self, mode=VS_TUNE_ONLINE, cache_path=None):
        """Launch-shape autotuning (vs_ctx_set_tuning)."""
        path = cache_path.encode() if cache_path else None
        check(lib().vs_ctx_set_tuning(self.h, int(mode), path), "vs_ctx_set_tuning")

    def tuning(self):
        """Settled launch shapes per size class (vs_ctx_tuning): {class: {"march", "shade",
        "integrate", "ms"}}, ms = {"march": {block: ms}, "shade": {...}, "integrate": {grid: ms}}
        for classes tuned in this context, None for ones read from a cache file."""
        need = C.c_size_t()
        check(lib().vs_ctx_tuning(self.h, None, 0, C.byref(need)), "vs_ctx_tuning")
        buf = C.create_string_buffer(need.value)
        check(lib().vs_ctx_tuning(self.h, buf, need.value, None), "vs_ctx_tuning")
        out = {}
        for line in buf.value.decode().splitlines():
            f = line.split("|")
            ms = None
            if f[4] != "cached":
                m, sh, ig = ([float(x) for x in g.split(",")] for g in f[4:7])
                ms = {"march": dict(zip(["8x8", "16x4", "32x2"], m)), "shade": dict(zip(self.SHADE_OPTS, sh)),
                      "integrate": dict(zip(self.INTEG_GRIDS, ig))}
            out[f[0]] = {"march": f[1], "shade": f[2], "integrate": int(f[3]), "ms": ms}
        return out

    def stage_ms(self):
        ms = np.zeros(8, np.float32)
        check(lib().vs_ctx_stage_ms(self.h, _ptr(ms)), "vs_ctx_stage_ms")
        return ms

    def fuse_batch(self, sessions, depths, is_host=True, stream=None):
        n = len(sessions)
        hs = (P * n)(*[s.h for s in sessions])
        keep = [np.ascontiguousarray(d, np.uint16) if is_host else d for d in depths]
        ds = (P * n)(*[(_ptr(k) if is_host else P(k)) for k in keep])
        check(lib().vs_fuse_batch(self.h, hs, ds, n, 1 if is_host else 0, stream), "vs_fuse_batch")


class Session:
    def __init__(self, ctx, desc, T_gc0=None, handle=None):
        self.ctx = ctx
        self.desc = desc
        self.W, self.H = desc.width, desc.height
        self.owned = handle is None
        if handle is not None:  # a session owned by a server slot
            self.h = P(handle)
            return
        self.h = P()
        T = None if T_gc0 is None else np.ascontiguousarray(T_gc0, np.float64).reshape(16)
        check(lib().vs_session_create(ctx.h, C.byref(desc), _ptr(T), C.byref(self.h)), "vs_session_create")

    def close(self):
        if self.h and self.owned:
            lib().vs_session_destroy(self.h)
        self.h = P()

    def level_shape(self, l):
        return (self.H >> l, self.W >> l)

    @property
    def nvox(self):
        d = self.desc
        return d.vol_dim[0] * d.vol_dim[1] * d.vol_dim[2]

    def snapshot(self):
        """The session's snapshot (the open sparse format of vscan.h) as bytes."""
        n = C.c_size_t()
        check(lib().vs_session_snapshot_bytes(self.h, C.byref(n)), "vs_session_snapshot_bytes")
        buf = np.zeros(n.value, np.uint8)
        check(lib().vs_session_snapshot(self.h, _ptr(buf), n.value, C.byref(n)), "vs_session_snapshot")
        return buf[:n.value].tobytes()

    def carve(self, valid):
        """g_uvv with a host validity volume (D, H, W) of bytes; returns the
        number of bricks reset."""
        valid = np.ascontiguousarray(valid, np.uint8)
        assert valid.size == self.nvox
        n = C.c_uint()
        check(lib().vs_volume_carve(self.h, _ptr(valid), 1, C.byref(n)), "vs_volume_carve")
        return n.value

    def set_color(self, rgb):
        """rgb: (3, H, W) uint8 planes registered to the depth frame."""
        rgb = np.ascontiguousarray(rgb, np.uint8)
        assert rgb.shape == (3, self.H, self.W)
        check(lib().vs_set_color(self.h, _ptr(rgb), 1, None), "vs_set_color")

    def texture(self, level):
        h, w = self.level_shape(level)
        out = np.zeros((3, h, w), np.uint8)
        check(lib().vs_get_texture(self.h, level, _ptr(out)), "vs_get_texture")
        return out

    def volume_color(self):
        d = self.desc
        out = np.zeros((3, d.vol_dim[2], d.vol_dim[1], d.vol_dim[0]), np.uint8)
        check(lib().vs_volume_download_color(self.h, _ptr(out)), "vs_volume_download_color")
        return out

    def frame_dtype(self):
        return np.float32 if self.desc.depth_units == VS_DEPTH_M_F32 else np.uint16

    def set_depth(self, depth):
        depth = np.ascontiguousarray(depth, self.frame_dtype())
        check(lib().vs_set_depth(self.h, _ptr(depth), 1, None), "vs_set_depth")

    def bilateral(self):
        check(lib().vs_bilateral(self.h, None), "vs_bilateral")

    def pyramid(self):
        check(lib().vs_pyramid(self.h, None), "vs_pyramid")

    def track(self, T_init=None):
        T = None if T_init is None else np.ascontiguousarray(T_init, np.float64).reshape(16)
        check(lib().vs_track(self.h, _ptr(T), None), "vs_track")

    def integrate(self):
        check(lib().vs_integrate(self.h, None), "vs_integrate")

    def raycast(self, level):
        check(lib().vs_raycast(self.h, level, None), "vs_raycast")

    def fuse(self, depth):
        """Fuse a host uint16 frame; None fuses the frame already in the
        session's raw buffer (vs_set_depth or a pre-step)."""
        if depth is None:
            check(lib().vs_fuse_frame(self.h, None, 1, None), "vs_fuse_frame")
            return
        depth = np.ascontiguousarray(depth, self.frame_dtype())
        check(lib().vs_fuse_frame(self.h, _ptr(depth), 1, None), "vs_fuse_frame")

    def depth_reproject(self, src_ptr, w, h, K_src, T_dst_src, dst_ptr=None, is_float=False):
        """z_tdrcc on device buffers (pointers as ints); dst None = the raw buffer."""
        K = np.ascontiguousarray(K_src, np.float32)
        T = np.ascontiguousarray(T_dst_src, np.float32).reshape(16)
        check(lib().vs_depth_reproject(self.h, P(src_ptr), w, h, _ptr(K), _ptr(T), 1 if is_float else 0,
                                       P(dst_ptr) if dst_ptr else None, None), "vs_depth_reproject")

    def depth_crop(self, depth_ptr=None, w=0, h=0, K=None, T_cg=None, is_float=False):
        """god's volume crop in place; depth_ptr None = the raw buffer."""
        Kp = None if K is None else np.ascontiguousarray(K, np.float32)
        Tp = None if T_cg is None else np.ascontiguousarray(T_cg, np.float32).reshape(16)
        check(lib().vs_depth_crop_to_volume(self.h, P(depth_ptr) if depth_ptr else None, w, h, _ptr(Kp), _ptr(Tp),
                                            1 if is_float else 0, None), "vs_depth_crop_to_volume")

    def sync(self):
        check(lib().vs_session_sync(self.h), "vs_session_sync")

    def status(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().vs_get_frame_status(self.h, C.byref(a), C.byref(b), C.byref(c)), "vs_get_frame_status")
        return a.value, b.value, c.value

    def pose(self):
        T = np.zeros(16, np.float64)
        check(lib().vs_get_pose(self.h, _ptr(T)), "vs_get_pose")
        return T.reshape(4, 4)

    def set_pose(self, T):
        T = np.ascontiguousarray(T, np.float64).reshape(16)
        check(lib().vs_set_pose(self.h, _ptr(T), None), "vs_set_pose")

    def filtered(self, level):
        out = np.zeros(self.level_shape(level), np.float32)
        check(lib().vs_get_filtered(self.h, level, _ptr(out)), "vs_get_filtered")
        return out

    def set_filtered(self, level, img):
        img = np.ascontiguousarray(img, np.float32)
        check(lib().vs_set_filtered(self.h, level, _ptr(img)), "vs_set_filtered")

    def maps(self, level):
        h, w = self.level_shape(level)
        d = np.zeros((h, w), np.float32)
        v = np.zeros((h, w, 3), np.float32)
        n = np.zeros((h, w, 3), np.float32)
        img = np.zeros((h, w), np.uint8)
        check(lib().vs_get_maps(self.h, level, _ptr(d), _ptr(v), _ptr(n), _ptr(img)), "vs_get_maps")
        return d, v, n, img

    def set_maps(self, level, vtx, nrm):
        vtx = np.ascontiguousarray(vtx, np.float32)
        nrm = np.ascontiguousarray(nrm, np.float32)
        check(lib().vs_set_maps(self.h, level, _ptr(vtx), _ptr(nrm)), "vs_set_maps")

    def volume(self):
        d = self.desc
        shp = (d.vol_dim[2], d.vol_dim[1], d.vol_dim[0])
        F = np.zeros(shp, np.float32)
        Wt = np.zeros(shp, np.float32)
        nb = (d.vol_dim[0] // 8) * (d.vol_dim[1] // 8) * (d.vol_dim[2] // 8)
        bm = np.zeros(nb, np.uint8)
        check(lib().vs_volume_download(self.h, _ptr(F), _ptr(Wt), _ptr(bm)), "vs_volume_download")
        return F, Wt, bm

    def set_volume(self, F, Wt):
        F = np.ascontiguousarray(F, np.float32)
        Wt = np.ascontiguousarray(Wt, np.float32)
        check(lib().vs_volume_upload(self.h, _ptr(F), _ptr(Wt)), "vs_volume_upload")

    def icp_system(self, level, T21, want_mask=False):
        T21 = np.ascontiguousarray(T21, np.float32).reshape(16)
        ATA = np.zeros(36, np.float64)
        ATb = np.zeros(6, np.float64)
        n = C.c_longlong()
        mask = np.zeros(self.level_shape(level), np.uint8) if want_mask else None
        check(lib().vs_icp_system(self.h, level, _ptr(T21), _ptr(ATA), _ptr(ATb), C.byref(n), _ptr(mask)),
              "vs_icp_system")
        return ATA.reshape(6, 6), ATb, n.value, mask

    def set_timing(self, on):
        check(lib().vs_session_set_timing(self.h, 1 if on else 0), "vs_session_set_timing")

    def stage_ms(self):
        ms = np.zeros(5, np.float32)
        check(lib().vs_get_stage_ms(self.h, _ptr(ms)), "vs_get_stage_ms")
        return ms

    def integrate_counts(self):
        a, b = C.c_longlong(), C.c_longlong()
        check(lib().vs_get_integrate_counts(self.h, C.byref(a), C.byref(b)), "vs_get_integrate_counts")
        return a.value, b.value

    def raycast_counts(self, level):
        """vs_raycast_counts: [nearest taps, trilinear evaluations, hits, nearest taps outside the grid,
        nearest taps without weight, nearest-tap bytes, trilinear bytes, normal bytes]."""
        c = np.zeros(8, np.uint64)
        check(lib().vs_raycast_counts(self.h, level, _ptr(c)), "vs_raycast_counts")
        return c


def restore(ctx, data):
    """A new Session from snapshot bytes."""
    buf = np.frombuffer(data, np.uint8)
    h = P()
    check(lib().vs_session_restore(ctx.h, _ptr(buf), len(data), C.byref(h)), "vs_session_restore")
    hdr = SnapshotHeader.from_buffer_copy(data[:C.sizeof(SnapshotHeader)])
    s = Session(ctx, hdr.desc, handle=h.value)
    s.owned = True
    return s


def session_bytes(desc):
    return lib().vs_session_bytes(C.byref(desc))


class Server:
    """Slot table of the serving form (vs_server_*): sessions opened into
    slots under an HBM budget, one frame per listed slot per step."""

    def __init__(self, ctx, max_slots, tsdf=0, weight=0, hbm_budget=0):
        self.ctx = ctx
        self.h = P()
        d = ServerDesc(max_slots, tsdf, weight, hbm_budget)
        check(lib().vs_server_create(ctx.h, C.byref(d), C.byref(self.h)), "vs_server_create")
        self.sessions = {}

    def close(self):
        if self.h:
            lib().vs_server_destroy(self.h)
            self.h = P()
            self.sessions.clear()

    def open(self, desc, T_gc0=None):
        """Slot index of the new session, or None when the budget or the
        slots are exhausted (VS_BUDGET)."""
        T = None if T_gc0 is None else np.ascontiguousarray(T_gc0, np.float64).reshape(16)
        slot = C.c_int(-1)
        code = lib().vs_server_open(self.h, C.byref(desc), _ptr(T), C.byref(slot))
        if code == VS_BUDGET:
            return None
        check(code, "vs_server_open")
        self.sessions[slot.value] = Session(self.ctx, desc, handle=lib().vs_server_session(self.h, slot.value))
        return slot.value

    def close_slot(self, slot):
        s = self.sessions.pop(slot)
        s.h = P()
        check(lib().vs_server_close(self.h, slot), "vs_server_close")

    def session(self, slot):
        return self.sessions[slot]

    def step(self, slots, depths, is_host=True, stream=None):
        n = len(slots)
        sl = np.ascontiguousarray(slots, np.int32)
        keep = [np.ascontiguousarray(d, np.uint16) if is_host else d for d in depths]
        ds = (P * n)(*[(_ptr(k) if is_host else P(k)) for k in keep])
        check(lib().vs_server_step(self.h, _ptr(sl), ds, n, 1 if is_host else 0, stream), "vs_server_step")

    def evict(self, slot):
        """Snapshot bytes of the slot's session; the slot is closed."""
        n = C.c_size_t()
        check(lib().vs_server_evict(self.h, slot, None, 0, C.byref(n)), "vs_server_evict")
        buf = np.zeros(n.value, np.uint8)
        s = self.sessions.pop(slot)
        s.h = P()
        check(lib().vs_server_evict(self.h, slot, _ptr(buf), n.value, C.byref(n)), "vs_server_evict")
        return buf[:n.value].tobytes()

    def open_snapshot(self, data):
        """Slot of the session restored from snapshot bytes, or None (VS_BUDGET)."""
        buf = np.frombuffer(data, np.uint8)
        slot = C.c_int(-1)
        code = lib().vs_server_open_snapshot(self.h, _ptr(buf), len(data), C.byref(slot))
        if code == VS_BUDGET:
            return None
        check(code, "vs_server_open_snapshot")
        hdr = SnapshotHeader.from_buffer_copy(data[:C.sizeof(SnapshotHeader)])
        self.sessions[slot.value] = Session(self.ctx, hdr.desc, handle=lib().vs_server_session(self.h, slot.value))
        return slot.value

    def usage(self):
        n, b = C.c_int(), C.c_ulonglong()
        check(lib().vs_server_usage(self.h, C.byref(n), C.byref(b)), "vs_server_usage")
        return n.value, b.value
