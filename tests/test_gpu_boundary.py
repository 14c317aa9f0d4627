"""The reference operators' own buffer conventions at the C ABI, and the
tracker's failure paths, on the GPU.

* Float-metre sessions (VS_DEPTH_M_F32, the units of bfd_Bilateral_Filter_Depth
  and cdtoc, z_yooji_cuda_img_proc.cu:394-442, z_yooji_cuda_volume_integrator.cu:936-972)
  give bit-identical results to uint16-millimetre sessions fed mm * 0.001f,
  with invalid pixels written as 0, negative, NaN or infinite metres; also
  inside one batch with both kinds.
* vs_bilateral_filter_depth, vs_integrate_depth and vs_render_maps write and
  read caller-owned device buffers (rmi_Render_Maps_for_ICP's io_map_* outputs,
  z_yooji_cuda_tsdf_renderer.cu:1642-1656).
* Failure paths (so_Scan_Object keeps the pose and skips fusion,
  z_yooji_object_scanner.cpp:553-557): a planar model is rank 3 and trips the
  relative pivot guard (VS_SINGULAR); the motion gate
  (z_yooji_camera_tracker.cpp:357-371) and the minimum-correspondence gate
  (:274) forced from the host; the initial-pose single-level track
  (z_yooji_cuda_pose_tracker.cu:830-839). Each against the oracle.
"""
import numpy as np
import pytest

import hiprt
import oracle_lib as orc
import vscan

pytestmark = pytest.mark.gpu

W0 = vscan.WORKLOADS["plain_320x240_128"]


def say(*a):
    print(*a, flush=True)


def float_frame(mm, seed):
    """mm * 0.001f, with the invalid (0 mm) pixels written as a mix of 0,
    -1, NaN, +inf and -inf metres."""
    f = orc.depth_to_m(mm)
    bad = np.flatnonzero(mm.reshape(-1) == 0)
    codes = np.random.default_rng(seed).integers(0, 5, bad.size)
    vals = np.array([0.0, -1.0, np.nan, np.inf, -np.inf], np.float32)
    f.reshape(-1)[bad] = vals[codes]
    return f


def _same(a, b, what):
    assert a.status() == b.status(), what
    assert np.array_equal(a.pose(), b.pose()), f"{what}: pose differs"
    Fa, Wa, Ba = a.volume()
    Fb, Wb, Bb = b.volume()
    assert np.array_equal(Fa, Fb) and np.array_equal(Wa, Wb) and np.array_equal(Ba, Bb), f"{what}: volume differs"
    for l in range(3):
        assert all(np.array_equal(x, y) for x, y in zip(a.maps(l), b.maps(l))), f"{what}: maps L{l} differ"
    for l in range(3):
        assert np.array_equal(a.filtered(l), b.filtered(l)), f"{what}: pyramid L{l} differs"


def test_float_metre_sessions_equal_millimetre_sessions(ctx):
    say("case: float-metre vs uint16-mm sessions, alone and in one batch, 6 frames")
    st = vscan.SceneStream(4, seed=1)
    du = vscan.make_desc(W0, K=st.K)
    df = vscan.make_desc(W0, K=st.K, depth_units=vscan.VS_DEPTH_M_F32)
    assert vscan.session_bytes(df) == vscan.session_bytes(du) + du.width * du.height * 2
    a = vscan.Session(ctx, du, T_gc0=st.pose(0))
    b = vscan.Session(ctx, df, T_gc0=st.pose(0))
    c = vscan.Session(ctx, du, T_gc0=st.pose(0))
    e = vscan.Session(ctx, df, T_gc0=st.pose(0))
    try:
        for fr in range(6):
            mm = st.frame(fr)
            ff = float_frame(mm, fr)
            a.fuse(mm)
            b.fuse(ff)
            ctx.fuse_batch([c, e], [mm, ff.view(np.uint16)])
        _same(a, b, "float alone")
        _same(a, c, "mm in batch")
        _same(a, e, "float in batch")
        assert a.status()[0] == 0 and a.status()[1] == 6
    finally:
        for s in (a, b, c, e):
            s.close()


@pytest.mark.parametrize("units", [vscan.VS_DEPTH_MM_U16, vscan.VS_DEPTH_M_F32])
def test_bilateral_filter_depth_into_caller_buffer(ctx, units):
    say(f"case: vs_bilateral_filter_depth, units {units}, host and device input")
    st = vscan.SceneStream(0, seed=2)
    d = vscan.make_desc(W0, K=st.K, depth_units=units)
    s = vscan.Session(ctx, d, T_gc0=st.pose(0))
    out = hiprt.DeviceBuffer(d.width * d.height * 4)
    dev_in = hiprt.DeviceBuffer(d.width * d.height * 4)
    L = vscan.lib()
    try:
        for fr, host in ((3, True), (5, False)):
            mm = st.frame(fr)
            frame = float_frame(mm, fr) if units == vscan.VS_DEPTH_M_F32 else mm
            frame = np.ascontiguousarray(frame)
            if host:
                src = frame.ctypes.data_as(vscan.P)
            else:
                dev_in.upload(frame)
                src = dev_in.ptr
            vscan.check(L.vs_bilateral_filter_depth(s.h, src, 1 if host else 0, out.ptr, None), "bfd")
            s.sync()
            got = out.download(np.zeros((d.height, d.width), np.float32))
            assert np.array_equal(got, s.filtered(0))
            ref = orc.bilateral(mm, d.bil_sigma_s, d.bil_sigma_r, d.bil_radius)
            assert np.array_equal(got > 0, ref > 0)
            assert np.abs(got - ref).max() <= 1e-6
    finally:
        out.free()
        dev_in.free()
        s.close()


def test_integrate_depth_and_render_maps_device_buffers(ctx):
    say("case: vs_integrate_depth from a device float frame, vs_render_maps into device buffers")
    st = vscan.SceneStream(4, seed=3)
    du = vscan.make_desc(W0, K=st.K)
    df = vscan.make_desc(W0, K=st.K, depth_units=vscan.VS_DEPTH_M_F32)
    a = vscan.Session(ctx, du, T_gc0=st.pose(0))
    b = vscan.Session(ctx, df, T_gc0=st.pose(0))
    L = vscan.lib()
    frame = hiprt.DeviceBuffer(du.width * du.height * 4)
    bufs = []
    try:
        for fr in (0, 4, 8):
            T = st.pose(fr)
            mm = st.frame(fr)
            a.set_pose(T)
            b.set_pose(T)
            a.set_depth(mm)
            a.integrate()
            frame.upload(np.ascontiguousarray(float_frame(mm, fr)))
            vscan.check(L.vs_integrate_depth(b.h, frame.ptr, None), "vs_integrate_depth")
        Fa, Wa, Ba = a.volume()
        Fb, Wb, Bb = b.volume()
        assert np.array_equal(Fa, Fb) and np.array_equal(Wa, Wb) and np.array_equal(Ba, Bb)
        assert Wa.max() > 0
        b.set_pose(st.pose(9))
        for l in range(3):
            h, w = b.level_shape(l)
            bd, bv, bn, bi = (hiprt.DeviceBuffer(h * w * 4), hiprt.DeviceBuffer(h * w * 12),
                              hiprt.DeviceBuffer(h * w * 12), hiprt.DeviceBuffer(h * w))
            bufs += [bd, bv, bn, bi]
            vscan.check(L.vs_render_maps(b.h, l, bd.ptr, bv.ptr, bn.ptr, bi.ptr, None), "vs_render_maps")
            b.sync()
            gd, gv, gn, gi = b.maps(l)
            dd = bd.download(np.zeros((h, w), np.float32))
            vv = bv.download(np.zeros((h, w, 3), np.float32))
            nn = bn.download(np.zeros((h, w, 3), np.float32))
            ii = bi.download(np.zeros((h, w), np.uint8))
            assert np.array_equal(dd, gd) and np.array_equal(vv, gv) and np.array_equal(nn, gn)
            assert np.array_equal(ii, gi)
            hit = vv[..., 0] != -100
            assert hit.sum() > 0.2 * hit.size and np.all(dd[~hit] == 0) and np.array_equal(dd[hit], vv[..., 2][hit])
        # a NULL output skips that buffer; the model maps still update
        vscan.check(L.vs_render_maps(b.h, 0, None, None, None, None, None), "vs_render_maps")
        b.sync()
    finally:
        for x in bufs + [frame]:
            x.free()
        a.close()
        b.close()


# ---------------------------------------------------------------- failure paths
def _analytic_plane(d, st):
    n = d.vol_dim[0]
    c = np.arange(n) + 0.5
    z = d.origin[2] + c * d.voxel
    F = np.clip(z / d.mu, -1.0, 1.0).astype(np.float32)  # the ground plane z = 0
    F = np.broadcast_to(F[:, None, None], (n, n, n)).copy()
    return F, np.ones_like(F)


def _prepared(ctx, d, st, T_model, mm, F, Wt):
    """A session holding volume (F, Wt), model maps raycast at T_model and
    the pyramid of frame mm; plus the same inputs for the oracle."""
    s = vscan.Session(ctx, d, T_gc0=T_model)
    s.set_volume(F, Wt)
    for l in range(3):
        s.raycast(l)
    maps = [s.maps(l) for l in range(3)]
    s.set_depth(mm)
    s.bilateral()
    s.pyramid()
    depths = [s.filtered(l) for l in range(3)]
    return s, depths, [m[1] for m in maps], [m[2] for m in maps]


def _oracle_volume(st, d, frames):
    vol = orc.volume(d)
    N = d.vol_dim[0] * d.vol_dim[1] * d.vol_dim[2]
    F = np.ones(N)
    Wt = np.zeros(N)
    K = orc.K_levels(d)
    for fr in frames:
        T = st.pose(fr)
        orc.integrate(st.frame(fr), K[0], np.float32(T), np.float32(orc.rigid_inverse(T)), vol, F, Wt)
    sh = (d.vol_dim[2], d.vol_dim[1], d.vol_dim[0])
    return F.reshape(sh).astype(np.float32), Wt.reshape(sh).astype(np.float32)


def _track_both(ctx, st, d, F, Wt, T0, frame):
    s, depths, mv, mn = _prepared(ctx, d, st, T0, st.frame(frame), F, Wt)
    try:
        s.track()
        status, _, solves = s.status()
        pose = s.pose()
    finally:
        s.close()
    ost, T_or, nsolve = orc.icp_track(depths, orc.K_levels(d).reshape(-1), 3, list(d.iters), mv, mn, d.icp_dist2,
                                      T0, rel_pivot=d.icp_rel_pivot)
    return status, solves, pose, ost, nsolve, T_or


def test_exact_plane_model_fails(ctx):
    """An exactly planar model (the analytic distance field of the ground
    plane): the normal equations are rank 3. With the relative guard on, the
    solve fails on the GPU and in the oracle (the oracle's FP64 Jacobian
    already trips the reference's absolute 1e-8 pivot, VS_TRACK_LOST; the
    GPU's FP32 Jacobian leaves a pivot above 1e-8 that the relative guard
    catches, VS_SINGULAR) and the pose is kept."""
    say("case: exactly planar model")
    st = vscan.SceneStream(0, seed=1, noise=False, dropout=0.0)
    st.sc.n_prims = 1
    d = vscan.make_desc(W0, K=st.K, n=64)
    d.icp_rel_pivot = 1e-4
    F, Wt = _analytic_plane(d, st)
    T0 = st.pose(0)
    status, solves, pose, ost, nsolve, _ = _track_both(ctx, st, d, F, Wt, T0, 1)
    say(f"  GPU status {status} after {solves} solves, oracle {ost} after {nsolve}")
    assert status in (vscan.VS_TRACK_LOST, vscan.VS_SINGULAR) and ost in (vscan.VS_TRACK_LOST, vscan.VS_SINGULAR)
    assert solves == 1 and nsolve == 1
    assert np.array_equal(pose, T0)


def test_relative_pivot_guard_on_a_plane_scene(ctx):
    """The ground plane alone, integrated from noisy frames: the smallest
    squared Cholesky pivot is below 2e-2 of its diagonal entry (about 8e-3 on
    this stream), while the scene with 20 primitives stays near 1e-1. With
    icp_rel_pivot = 2e-2 the plane's first solve is VS_SINGULAR on the GPU and
    in the oracle, and the primitive scene tracks as without the guard."""
    say("case: relative pivot guard, plane-only scene vs primitive scene")
    for cfg, nprims, want in ((0, 1, vscan.VS_SINGULAR), (4, None, vscan.VS_OK)):
        st = vscan.SceneStream(cfg, seed=1)
        if nprims:
            st.sc.n_prims = nprims
        d = vscan.make_desc(W0, K=st.K)
        F, Wt = _oracle_volume(st, d, (0, 3, 6))
        T0 = st.pose(6)
        d.icp_rel_pivot = 2e-2
        status, solves, pose, ost, nsolve, T_or = _track_both(ctx, st, d, F, Wt, T0, 7)
        say(f"  scene {cfg}: GPU status {status} after {solves} solves, oracle {ost} after {nsolve}")
        assert status == want and ost == want and solves == nsolve
        if want == vscan.VS_OK:
            dt, da = orc.pose_error(pose, T_or)
            assert dt <= 0.5e-3 and da <= 0.05
        else:
            assert solves == 1 and np.array_equal(pose, T0)


def test_motion_and_correspondence_gates(ctx):
    """Gates forced from the host: a 0.1 mm / 0.01 deg motion limit rejects
    the orbit's 1 deg step (VS_MOTION_REJECTED), and a correspondence
    minimum above the pixel count stops the first solve (VS_TOO_FEW_POINTS);
    the GPU and the oracle pipeline agree, the pose is kept and the frame is
    not fused."""
    say("case: motion gate and minimum-correspondence gate")
    st = vscan.SceneStream(4, seed=1)
    cases = [("motion", dict(max_motion_m=1e-4, max_motion_deg=0.01), vscan.VS_MOTION_REJECTED),
             ("min_valid", dict(icp_min_valid=10 ** 7), vscan.VS_TOO_FEW_POINTS),
             ("loose", dict(max_motion_m=0.1, max_motion_deg=30.0), vscan.VS_OK)]
    for name, fields, want in cases:
        d = vscan.make_desc(W0, K=st.K)
        for k, v in fields.items():
            setattr(d, k, v)
        s = vscan.Session(ctx, d, T_gc0=st.pose(0))
        o = orc.OracleSession(d, st.pose(0))
        try:
            s.fuse(st.frame(0))
            o.fuse(st.frame(0))
            T1 = s.pose()
            F1, W1, _ = s.volume()
            s.fuse(st.frame(1))
            ost, _ = o.fuse(st.frame(1))
            status, frames, solves = s.status()
            say(f"  {name}: GPU status {status} ({solves} solves), oracle {ost}")
            assert status == want and ost == want
            if want != vscan.VS_OK:
                assert frames == 1 and np.array_equal(s.pose(), T1)
                F2, W2, _ = s.volume()
                assert np.array_equal(F1, F2) and np.array_equal(W1, W2)
            else:
                assert frames == 2
                dt, da = orc.pose_error(s.pose(), o.pose())
                assert dt <= 0.5e-3 and da <= 0.05
        finally:
            s.close()
            o.close()


def test_track_from_initial_pose_single_level(ctx):
    """vs_track with T_01_init runs level 0 only, starting from T_init
    (z_yooji_cuda_pose_tracker.cu:830-839), as the oracle's tracker."""
    say("case: initial-pose single-level track")
    st = vscan.SceneStream(4, seed=1)
    d = vscan.make_desc(W0, K=st.K)
    K = orc.K_levels(d)
    F, Wt = _oracle_volume(st, d, range(0, 9, 2))
    T0 = st.pose(9)
    s, depths, mv, mn = _prepared(ctx, d, st, T0, st.frame(11), F, Wt)
    try:
        rel = st.pose(11) @ orc.rigid_inverse(T0)  # camera t0 -> camera t1, ground truth
        T_init = orc.expmap(np.array([0.002, -0.001, 0.0015, 0.003, -0.002, 0.001])) @ rel
        s.track(T_init)
        status, _, solves = s.status()
        ost, T_or, nsolve = orc.icp_track(depths, K.reshape(-1), 3, list(d.iters), mv, mn, d.icp_dist2, T0,
                                          T_init=T_init)
        assert status == 0 and ost == 0
        assert solves == nsolve and solves <= d.iters[0]
        dt, da = orc.pose_error(s.pose(), T_or)
        gt_t, gt_a = orc.pose_error(T_or, st.pose(11))
        say(f"  {solves} level-0 solves; GPU vs oracle {dt * 1e3:.4f} mm {da:.5f} deg; oracle vs truth "
            f"{gt_t * 1e3:.3f} mm {gt_a:.4f} deg")
        assert dt <= 0.5e-3 and da <= 0.05
        assert gt_t < 0.01 and gt_a < 1.0
    finally:
        s.close()
