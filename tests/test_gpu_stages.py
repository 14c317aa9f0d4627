"""Stage-by-stage parity of the HIP pipeline with the oracle, on identical
seeded inputs (the plain 320x240 / 128^3 workload unless stated).

Bars (from the workload specification): filtered depth within 1e-3 mm, TSDF
within 1e-4 (normalised), raycast points within 0.5 voxel, pixel-validity and
voxel/brick masks bit-exact.
"""
import numpy as np
import pytest

import oracle_lib as orc
import vscan

pytestmark = pytest.mark.gpu

W0 = vscan.WORKLOADS["plain_320x240_128"]


def f32pose(T):
    Tcg = orc.rigid_inverse(T)
    return np.float32(T), np.float32(Tcg)


@pytest.fixture(scope="module")
def stream():
    return vscan.SceneStream(0, seed=1)


@pytest.fixture()
def sess(ctx, stream):
    d = vscan.make_desc(W0, K=stream.K)
    s = vscan.Session(ctx, d, T_gc0=stream.pose(0))
    yield s
    s.close()


@pytest.mark.parametrize("radius,sigma_s", [(1, 1.5), (2, 4.0), (3, 2.5)])
def test_bilateral_parity(ctx, stream, radius, sigma_s):
    print(f"case: bilateral 320x240 frames 0,7, radius {radius}", flush=True)
    d = vscan.make_desc(W0, K=stream.K)
    d.bil_radius = radius
    d.bil_sigma_s = sigma_s
    sess = vscan.Session(ctx, d, T_gc0=stream.pose(0))
    try:
        for fr in (0, 7):
            mm = stream.frame(fr)
            sess.set_depth(mm)
            sess.bilateral()
            got = sess.filtered(0)
            ref = orc.bilateral(mm, d.bil_sigma_s, d.bil_sigma_r, d.bil_radius)
            assert np.array_equal(got > 0, ref > 0), "validity mask differs"
            err = np.abs(got.astype(np.float64) - ref)
            assert err.max() <= 1e-6, f"max filtered-depth error {err.max() * 1e3:.3g} mm > 1e-3 mm"
    finally:
        sess.close()


def test_bilateral_640(ctx):
    print("case: bilateral 640x480", flush=True)
    st = vscan.SceneStream(1, seed=2)
    d = vscan.make_desc(vscan.WORKLOADS["plain_640x480_256"], K=st.K, n=64)
    s = vscan.Session(ctx, d, T_gc0=st.pose(0))
    try:
        mm = st.frame(3)
        s.set_depth(mm)
        s.bilateral()
        got = s.filtered(0)
        ref = orc.bilateral(mm, d.bil_sigma_s, d.bil_sigma_r, d.bil_radius)
        assert np.array_equal(got > 0, ref > 0)
        assert np.abs(got - ref).max() <= 1e-6
    finally:
        s.close()


def test_pyramid_parity(sess, stream):
    print("case: pyramid", flush=True)
    sess.set_depth(stream.frame(2))
    sess.bilateral()
    sess.pyramid()
    for l in (1, 2):
        src = sess.filtered(l - 1)
        got = sess.filtered(l)
        ref = orc.pyr_down(src.astype(np.float64))
        assert np.array_equal(got > 0, ref > 0), f"level {l} mask"
        assert np.abs(got - ref).max() <= 1e-6, f"level {l}"


def test_vertex_normal_parity(sess, stream):
    import hiprt

    print("case: vertex/normal maps", flush=True)
    sess.set_depth(stream.frame(4))
    sess.bilateral()
    sess.pyramid()
    K = orc.K_levels(sess.desc)
    for l in range(3):
        h, w = sess.level_shape(l)
        v = hiprt.DeviceBuffer(h * w * 12)
        n = hiprt.DeviceBuffer(h * w * 12)
        vscan.check(vscan.lib().vs_vertex_normal_maps(sess.h, l, v.ptr, n.ptr, None), "vnmap")
        sess.sync()
        gv = v.download(np.zeros((h, w, 3), np.float32))
        gn = n.download(np.zeros((h, w, 3), np.float32))
        v.free()
        n.free()
        rv, rn = orc.vnmap(sess.filtered(l), K[l])
        assert np.array_equal(gv[..., 0] == -100, rv[..., 0] == -100), f"vertex mask level {l}"
        assert np.array_equal(gn[..., 0] == -100, rn[..., 0] == -100), f"normal mask level {l}"
        ok = rn[..., 0] != -100
        assert np.abs(gv - rv)[rv[..., 0] != -100].max() <= 1e-6
        assert np.abs(gn - rn)[ok].max() <= 1e-5


@pytest.mark.parametrize("cfg,frame", [(0, 4), (1, 6)])
def test_vertex_normal_inplace_parity(ctx, cfg, frame):
    """The reference loop's in-place order against the oracle's in-place
    form: masks bit-exact at every level, and different from the
    order-independent form next to grazing pixels."""
    import hiprt

    print(f"case: in-place vertex/normal maps, scene {cfg} frame {frame}", flush=True)
    st = vscan.SceneStream(cfg, seed=1)
    w = W0 if cfg == 0 else vscan.WORKLOADS["plain_640x480_256"]
    d = vscan.make_desc(w, K=st.K, n=64)
    s = vscan.Session(ctx, d, T_gc0=st.pose(0))
    try:
        s.set_depth(st.frame(frame))
        s.bilateral()
        s.pyramid()
        K = orc.K_levels(d)
        differ = 0
        for l in range(3):
            h, wl = s.level_shape(l)
            v = hiprt.DeviceBuffer(h * wl * 12)
            n = hiprt.DeviceBuffer(h * wl * 12)
            try:
                vscan.check(vscan.lib().vs_vertex_normal_maps_inplace(s.h, l, v.ptr, n.ptr, None), "vnmap inplace")
                s.sync()
                gv = v.download(np.zeros((h, wl, 3), np.float32))
                gn = n.download(np.zeros((h, wl, 3), np.float32))
            finally:
                v.free()
                n.free()
            rv, rn = orc.vnmap(s.filtered(l), K[l], inplace=True)
            ov, on = orc.vnmap(s.filtered(l), K[l], inplace=False)
            assert np.array_equal(gv[..., 0] == -100, rv[..., 0] == -100), f"vertex mask level {l}"
            assert np.array_equal(gn[..., 0] == -100, rn[..., 0] == -100), f"normal mask level {l}"
            ok = rn[..., 0] != -100
            assert np.abs(gv - rv)[rv[..., 0] != -100].max() <= 1e-6
            assert np.abs(gn - rn)[ok].max() <= 1e-5
            differ += int(np.sum((rn[..., 0] == -100) != (on[..., 0] == -100)))
        assert differ > 0, "the scene has no grazing pixel next to a valid normal"
    finally:
        s.close()


def _oracle_volume(stream, desc, frames):
    """FP64 oracle TSDF built from ground-truth poses, returned as float32."""
    vol = orc.volume(desc)
    N = desc.vol_dim[0] * desc.vol_dim[1] * desc.vol_dim[2]
    F = np.ones(N)
    Wt = np.zeros(N)
    K = orc.K_levels(desc)[0]
    for fr in frames:
        Tg, Tc = f32pose(stream.pose(fr))
        orc.integrate(stream.frame(fr), K, Tg, Tc, vol, F, Wt)
    shp = (desc.vol_dim[2], desc.vol_dim[1], desc.vol_dim[0])
    return F.reshape(shp), Wt.reshape(shp)


def test_integrate_parity(sess, stream):
    print("case: integrate 3 frames at ground-truth poses", flush=True)
    d = sess.desc
    vol = orc.volume(d)
    N = sess.nvox
    F = np.ones(N)
    Wt = np.zeros(N)
    K = orc.K_levels(d)[0]
    for fr in (0, 5, 11):
        T = stream.pose(fr)
        mm = stream.frame(fr)
        sess.set_pose(T)
        sess.set_depth(mm)
        sess.integrate()
        Tg, Tc = f32pose(T)
        bm, upd, counts = orc.integrate(mm, K, Tg, Tc, vol, F, Wt)
        gF, gW, gbm = sess.volume()
        assert np.array_equal(gbm, bm), f"brick mask differs at frame {fr}"
        vb, nu = sess.integrate_counts()
        assert vb == counts[0] and nu == counts[1], (vb, nu, counts)
        assert np.array_equal(gW.reshape(-1), Wt.astype(np.float32)), "weights differ"
        err = np.abs(gF.reshape(-1).astype(np.float64) - F)
        assert err.max() <= 1e-4, f"TSDF error {err.max()}"


def test_raycast_parity(sess, stream):
    print("case: raycast 3 levels on an oracle volume", flush=True)
    d = sess.desc
    F, Wt = _oracle_volume(stream, d, range(0, 12, 3))
    sess.set_volume(F.astype(np.float32), Wt.astype(np.float32))
    T = stream.pose(12)
    sess.set_pose(T)
    Tg, Tc = f32pose(T)
    K = orc.K_levels(d)
    vol = orc.volume(d)
    vox = d.voxel
    for l in range(3):
        sess.raycast(l)
        gd, gv, gn, gimg = sess.maps(l)
        h, w = sess.level_shape(l)
        rd, rv, rn, rimg, stats = orc.raycast(F.astype(np.float32), Wt.astype(np.float32), vol, w, h, K[l], Tg, Tc)
        gm = gv[..., 0] != -100
        rm = rv[..., 0] != -100
        assert np.array_equal(gm, rm), f"level {l}: {np.sum(gm != rm)} rays differ in hit/miss"
        assert gm.sum() > 0.2 * gm.size
        assert np.abs(gv - rv)[rm].max() <= 0.5 * vox
        assert np.abs(gn - rn)[rm].max() <= 1e-3
        assert np.abs(gd - rd)[rm].max() <= 0.5 * vox
        assert np.abs(gimg.astype(int) - rimg.astype(int)).max() <= 1


def test_icp_system_parity(sess, stream):
    print("case: one ICP system per level", flush=True)
    d = sess.desc
    F, Wt = _oracle_volume(stream, d, range(0, 9, 3))
    sess.set_volume(F.astype(np.float32), Wt.astype(np.float32))
    sess.set_pose(stream.pose(9))
    for l in range(3):
        sess.raycast(l)
    sess.set_depth(stream.frame(10))
    sess.bilateral()
    sess.pyramid()
    K = orc.K_levels(d)
    T21 = np.eye(4, dtype=np.float32)
    for l in range(3):
        _, mv, mn, _ = sess.maps(l)
        A, b, n, mask = sess.icp_system(l, T21, want_mask=True)
        rA, rb, rn, rmask = orc.icp_system(sess.filtered(l), K[l], mv, mn, T21, d.icp_dist2)
        assert np.array_equal(mask, rmask), f"level {l}: correspondence mask differs ({np.sum(mask != rmask)})"
        assert n == rn
        scale = np.abs(rA).max()
        assert np.abs(A - rA).max() <= 1e-5 * scale
        assert np.abs(b - rb).max() <= 1e-5 * max(np.abs(rb).max(), 1e-12) + 1e-7 * scale


def test_track_parity(ctx):
    """Full coarse-to-fine track on the 20-primitive stream (on the plain
    sphere+box scene the normal equations are too ill-conditioned for FP32
    and FP64 sums to agree on the pose; see test_gpu_pipeline.py)."""
    print("case: full coarse-to-fine track", flush=True)
    stream = vscan.SceneStream(4, seed=1)
    d = vscan.make_desc(W0, K=stream.K)
    sess = vscan.Session(ctx, d, T_gc0=stream.pose(0))
    try:
        _track_parity(sess, stream)
    finally:
        sess.close()


def _track_parity(sess, stream):
    d = sess.desc
    F, Wt = _oracle_volume(stream, d, range(0, 9, 2))
    sess.set_volume(F.astype(np.float32), Wt.astype(np.float32))
    T0 = stream.pose(9)
    sess.set_pose(T0)
    for l in range(3):
        sess.raycast(l)
    maps = [sess.maps(l) for l in range(3)]
    sess.set_depth(stream.frame(11))
    sess.bilateral()
    sess.pyramid()
    depths = [sess.filtered(l) for l in range(3)]
    sess.track()
    st, _, iters = sess.status()
    assert st == 0
    T_gpu = sess.pose()
    K = orc.K_levels(d).reshape(-1)
    rst, T_or, nsolve = orc.icp_track(depths, K, 3, list(d.iters), [m[1] for m in maps], [m[2] for m in maps],
                                      d.icp_dist2, T0)
    assert rst == 0
    dt, da = orc.pose_error(T_gpu, T_or)
    assert dt <= 0.5e-3 and da <= 0.05, (dt, da)
    # and both land near the ground truth of frame 11
    gt_t, gt_a = orc.pose_error(T_or, stream.pose(11))
    assert gt_t < 0.01 and gt_a < 1.0, (gt_t, gt_a)


def _grid_parity(ctx, st, d, shapes):
    """Integrate frames 0, 4, 8 at ground-truth poses and compare with the
    oracle (brick mask, visible and updated counts and weights bit-exact,
    TSDF within 1e-4), then raycast the oracle's volume at pose 10 (hit masks
    bit-exact, points within 0.5 voxel) at every level; `shapes` are the
    expected (width, height) of the levels."""
    nx, ny, nz = d.vol_dim[0], d.vol_dim[1], d.vol_dim[2]
    s = vscan.Session(ctx, d, T_gc0=st.pose(0))
    try:
        vol = orc.volume(d)
        N = nx * ny * nz
        assert s.nvox == N
        F = np.ones(N)
        Wt = np.zeros(N)
        K = orc.K_levels(d)
        for fr in (0, 4, 8):
            T = st.pose(fr)
            mm = st.frame(fr)
            s.set_pose(T)
            s.set_depth(mm)
            s.integrate()
            Tg, Tc = f32pose(T)
            bm, upd, counts = orc.integrate(mm, K[0], Tg, Tc, vol, F, Wt)
            gF, gW, gbm = s.volume()
            assert gF.shape == (nz, ny, nx)
            assert np.array_equal(gbm, bm), f"brick mask differs at frame {fr}"
            assert s.integrate_counts() == (counts[0], counts[1]), (s.integrate_counts(), counts)
            assert np.array_equal(gW.reshape(-1), Wt.astype(np.float32)), "weights differ"
            assert np.abs(gF.reshape(-1).astype(np.float64) - F).max() <= 1e-4
        assert counts[1] > 10000
        F32, W32 = F.astype(np.float32), Wt.astype(np.float32)
        s.set_volume(F32, W32)
        T = st.pose(10)
        s.set_pose(T)
        Tg, Tc = f32pose(T)
        for l in range(3):
            s.raycast(l)
            _, gv, _, _ = s.maps(l)
            h, w = s.level_shape(l)
            assert (w, h) == shapes[l]
            _, rv, _, _, _ = orc.raycast(F32, W32, vol, w, h, K[l], Tg, Tc)
            gm, rm = gv[..., 0] != -100, rv[..., 0] != -100
            assert np.array_equal(gm, rm), f"level {l}: {int(np.sum(gm != rm))} rays differ in hit/miss"
            assert gm.sum() > 0.1 * gm.size
            assert np.abs(gv - rv)[rm].max() <= 0.5 * d.voxel
    finally:
        s.close()


def test_noncubic_grid_odd_frame(ctx):
    """A grid of 120 x 104 x 72 voxels (sides not multiples of the 16-voxel
    superbrick, so the last superbrick of each axis is partly outside) and a
    200x150 frame (levels 100x75 and 50x37, odd halvings). Cubic grids and
    even frames cannot catch an axis mix-up in the brick or superbrick
    indexing."""
    print("case: non-cubic 120x104x72 grid, 200x150 frame", flush=True)
    st = vscan.SceneStream(0, seed=2, width=200, height=150)
    d = vscan.make_desc(W0, K=st.K, W=200, H=150)
    d.vol_dim[0], d.vol_dim[1], d.vol_dim[2] = 120, 104, 72
    _grid_parity(ctx, st, d, ((200, 150), (100, 75), (50, 37)))


def test_longest_grid_side(ctx):
    """A grid at the largest side the ABI accepts, 2048 voxels along x (256
    bricks, 128 superbricks; 2048 x 64 x 64, a 16 m strip through the scene):
    the largest brick coordinates and a ray start 7 m behind the camera (the
    cube's bounding sphere), with the same bars as above."""
    print("case: 2048x64x64 grid, 320x240 frame", flush=True)
    st = vscan.SceneStream(0, seed=3)
    d = vscan.make_desc(W0, K=st.K)
    d.vol_dim[0], d.vol_dim[1], d.vol_dim[2] = 2048, 64, 64
    d.origin[0], d.origin[1], d.origin[2] = -1024 * d.voxel, -32 * d.voxel, -0.05
    _grid_parity(ctx, st, d, ((320, 240), (160, 120), (80, 60)))
