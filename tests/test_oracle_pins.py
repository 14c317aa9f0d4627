"""Pins of the oracle: every stage is checked against something the oracle's
own formula does not restate — closed forms, invariants, special cases, the
reference's own compiled math (oracle/_ref), or brute-force geometry.

Runs on the CPU (no GPU marker).
"""
import numpy as np
import pytest

import oracle_lib as orc
import vscan

RNG = np.random.default_rng(7)


# ---------------------------------------------------------------- A1 ingest
def test_depth_to_metres_every_value():
    mm = np.arange(65536, dtype=np.uint16)
    got = orc.depth_to_m(mm)
    want = mm.astype(np.float32) * np.float32(0.001)  # one IEEE FP32 product
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


# ---------------------------------------------------------------- A2-A3 bilateral
def test_gaussian_weights_closed_form():
    for sigma, r in [(4.0, 2), (1.5, 3), (0.7, 1)]:
        g = orc.gaussian_weights(sigma, r)
        x = np.arange(-r, r + 1)
        assert g[r] == 1.0
        assert np.allclose(g, np.exp(-x * x / (2 * sigma * sigma)), rtol=2e-7, atol=0)
        assert np.array_equal(g, g[::-1])


def test_bilateral_constant_image_is_identity():
    mm = np.full((20, 24), 1234, np.uint16)
    out = orc.bilateral(mm, 4.0, 0.03, 2)
    assert np.allclose(out, np.float32(1234) * np.float32(0.001), rtol=0, atol=1e-15)


def test_bilateral_isolated_zero_mask():
    mm = np.full((16, 16), 900, np.uint16)
    mm[7, 9] = 0
    out = orc.bilateral(mm, 4.0, 0.03, 2)
    zero = out == 0
    want = np.zeros_like(zero)
    want[5:10, 7:12] = True  # the 5x5 window around the dropout
    assert np.array_equal(zero, want)


def test_bilateral_linear_ramp_is_identity_in_interior():
    x = np.arange(40)
    mm = np.tile((800 + 3 * x).astype(np.uint16), (30, 1))
    out = orc.bilateral(mm, 4.0, 0.03, 2)
    d = mm.astype(np.float32) * np.float32(0.001)
    # symmetric taps cancel on a ramp; the residue is the FP32 rounding of the
    # mm -> m inputs (not exactly linear), far below a misplaced tap (~1e-4 m)
    assert np.abs(out[:, 2:-2] - d[:, 2:-2]).max() < 1e-7


def test_bilateral_preserves_step_edge_with_small_range_sigma():
    mm = np.full((12, 20), 700, np.uint16)
    mm[:, 10:] = 800
    out = orc.bilateral(mm, 4.0, 0.001, 2)  # sigma_r = 1 mm vs a 100 mm step
    a, b = np.float32(700) * np.float32(0.001), np.float32(800) * np.float32(0.001)
    assert np.abs(out[:, :10] - a).max() < 1e-9
    assert np.abs(out[:, 10:] - b).max() < 1e-9


def test_bilateral_mirror_symmetry_and_clamp():
    mm = (600 + RNG.integers(0, 80, (17, 23))).astype(np.uint16)
    out = orc.bilateral(mm, 4.0, 0.03, 2)
    flip = orc.bilateral(mm[:, ::-1].copy(), 4.0, 0.03, 2)[:, ::-1]
    assert np.abs(out - flip).max() < 1e-12
    # clamp-to-edge: a corner pixel equals the filter of the padded image
    pad = np.pad(mm, 2, mode="edge")
    big = orc.bilateral(pad, 4.0, 0.03, 2)
    assert np.abs(big[2:-2, 2:-2] - out)[2:-2, 2:-2].max() < 1e-12


# ---------------------------------------------------------------- A4-A5 pyramid
def test_pyramid_constant_interior_and_zero_border():
    img = np.full((32, 40), 1.25)
    out = orc.pyr_down(img)
    H2, W2 = out.shape
    i = 2 * np.arange(W2) + 1
    j = 2 * np.arange(H2) + 1
    inside = ((j[:, None] >= 2) & (j[:, None] <= 32 - 3)) & ((i[None, :] >= 2) & (i[None, :] <= 40 - 3))
    assert np.all(out[inside] == 1.25)
    assert np.all(out[~inside] == 0.0)


def test_pyramid_impulse_response_is_binomial_at_odd_pixels():
    w = np.array([1, 4, 6, 4, 1], float)
    for (y0, x0) in [(9, 12), (10, 13), (15, 15)]:
        img = np.zeros((32, 32))
        img[y0, x0] = 256.0
        out = orc.pyr_down(img)
        want = np.zeros_like(out)
        for j2 in range(16):
            for i2 in range(16):
                dy, dx = y0 - (2 * j2 + 1), x0 - (2 * i2 + 1)
                if abs(dy) <= 2 and abs(dx) <= 2 and 2 <= 2 * i2 + 1 <= 29 and 2 <= 2 * j2 + 1 <= 29:
                    want[j2, i2] = w[dy + 2] * w[dx + 2]
        assert np.array_equal(out, want)


def test_intrinsics_pyramid_closed_form():
    class D:
        fx, fy, cx, cy = 525.0, 525.0, 319.5, 239.5

    K = orc.K_levels(D)
    for l in range(3):
        assert K[l, 0] == 525.0 / 2 ** l
        assert K[l, 2] == (319.5 + 0.5) / 2 ** l - 0.5
        assert K[l, 3] == (239.5 + 0.5) / 2 ** l - 0.5


# ---------------------------------------------------------------- A6-A7 vertex/normal maps
def _plane_depth(n_cam, dist, K, W, H):
    """Exact depth of the plane n.X = dist seen by a pinhole camera."""
    u, v = np.meshgrid(np.arange(W), np.arange(H))
    ray = np.stack([(u - K[2]) / K[0], (v - K[3]) / K[1], np.ones_like(u, float)], -1)
    return (dist / (ray @ n_cam)).astype(np.float32)


def test_backproject_project_round_trip():
    K = np.array([262.5, 262.5, 159.5, 119.5], np.float32)
    uu, vv = np.meshgrid(np.arange(32), np.arange(24))
    d = (0.6 + 0.002 * uu + 0.001 * vv + 0.0005 * RNG.random((24, 32))).astype(np.float32)
    v, _ = orc.vnmap(d, K)
    ok = v[..., 0] != -100  # some noisy pixels fail the grazing test
    assert ok.mean() > 0.7
    u_ = K[0] * v[..., 0] / v[..., 2] + K[2]
    v_ = K[1] * v[..., 1] / v[..., 2] + K[3]
    uu, vv = np.meshgrid(np.arange(32), np.arange(24))
    assert np.abs(u_ - uu)[ok].max() < 1e-4 and np.abs(v_ - vv)[ok].max() < 1e-4


def test_normals_of_analytic_planes():
    K = np.array([262.5, 262.5, 159.5, 119.5], np.float32)
    for tilt in (0.0, 20.0, 45.0):
        t = np.radians(tilt)
        n = np.array([np.sin(t), 0.0, -np.cos(t)])  # facing the camera
        d = _plane_depth(n, -0.8, K, 64, 48)
        _, nrm = orc.vnmap(d, K)
        ok = nrm[..., 0] != -100
        assert ok[1:-1, 1:-1].all()
        assert np.abs(nrm[ok] - n).max() < 2e-3
    # grazing plane (70 deg): every normal rejected, vertices invalidated too
    t = np.radians(70.0)
    n = np.array([np.sin(t), 0.0, -np.cos(t)])
    d = _plane_depth(n, -0.8, K, 64, 48)
    d[d <= 0] = 0
    v, nrm = orc.vnmap(d, K)
    assert np.all(nrm[..., 0] == -100)
    assert np.all(v[1:-1, 1:-1][(d[1:-1, 1:-1] > 0)][..., 0] == -100)


def test_inplace_vertex_invalidation_only_differs_next_to_grazing_pixels():
    st = vscan.SceneStream(0, seed=1, noise=False, dropout=0.0)
    d = st.exact_depth(st.pose(0)).astype(np.float32)
    K = np.array(st.K, np.float32)
    v0, n0 = orc.vnmap(d, K, inplace=False)
    v1, n1 = orc.vnmap(d, K, inplace=True)
    diff = np.any(n0 != n1, axis=-1)
    graz = (v0[..., 0] == -100) & (d > 0)
    near = np.zeros_like(graz)
    near[1:-1, 1:-1] = graz[1:-1, :-2] | graz[:-2, 1:-1]
    assert np.all(~diff | near)


# ---------------------------------------------------------------- A10-A11 solve and pose algebra
def test_cholesky_matches_dense_solve():
    for _ in range(20):
        M = RNG.normal(size=(12, 6))
        A = M.T @ M + 1e-3 * np.eye(6)
        b = RNG.normal(size=6)
        ok, x = orc.cholesky_solve(A, b)
        assert ok
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)


def test_cholesky_fails_on_singular_zero_and_nan():
    M = RNG.normal(size=(12, 3))
    A = np.zeros((6, 6))
    A[:3, :3] = M.T @ M  # rank 3
    assert not orc.cholesky_solve(A, np.ones(6))[0]
    assert not orc.cholesky_solve(np.zeros((6, 6)), np.zeros(6))[0]
    B = np.eye(6)
    B[2, 2] = np.nan
    assert not orc.cholesky_solve(B, np.ones(6))[0]


def _expm(x):
    from scipy.linalg import expm

    w, v = x[:3], x[3:]
    X = np.zeros((4, 4))
    X[:3, :3] = [[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]
    X[:3, 3] = v
    return expm(X)


def test_expmap_matches_matrix_exponential():
    for scale in (1e-9, 1e-5, 1e-3, 0.05, 0.4, 1.5):
        x = RNG.normal(size=6) * scale
        assert np.allclose(orc.expmap(x), _expm(x), rtol=0, atol=1e-13)
    T = orc.expmap(np.array([0, 0, 0, 0.1, -0.2, 0.3]))
    assert np.array_equal(T[:3, :3], np.eye(3)) and np.allclose(T[:3, 3], [0.1, -0.2, 0.3], atol=0)


def test_expmap_and_inverse_match_reference_build():
    ref = orc.ref_math()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for _ in range(50):
        x = (RNG.normal(size=6) * 0.05).astype(np.float32)
        Tr = np.zeros(16, np.float32)
        ref.ref_expmap(x.ctypes.data_as(orc.C.c_void_p), Tr.ctypes.data_as(orc.C.c_void_p))
        assert np.allclose(orc.expmap(x.astype(np.float64)), Tr.reshape(4, 4), rtol=0, atol=2e-6)
        T = orc.expmap(RNG.normal(size=6) * 0.3)
        Ti = np.zeros(16, np.float32)
        assert ref.ref_inverse4x4(np.float32(T).ctypes.data_as(orc.C.c_void_p), Ti.ctypes.data_as(orc.C.c_void_p))
        assert np.allclose(orc.rigid_inverse(T), Ti.reshape(4, 4), rtol=0, atol=2e-6)


def test_rigid_inverse():
    for _ in range(10):
        T = orc.expmap(RNG.normal(size=6))
        assert np.allclose(orc.rigid_inverse(T) @ T, np.eye(4), atol=1e-13)


# ---------------------------------------------------------------- A8-A9, A12 tracking
def _analytic_volume(stream, desc):
    n = desc.vol_dim[0]
    vox = desc.voxel
    org = np.array(list(desc.origin))
    i = np.arange(n)
    z, y, x = np.meshgrid(i, i, i, indexing="ij")
    pts = np.stack([org[0] + (x + 0.5) * vox, org[1] + (y + 0.5) * vox, org[2] + (z + 0.5) * vox], -1)
    sdf = stream.sdf(pts.reshape(-1, 3)).reshape(n, n, n)
    return np.clip(sdf / desc.mu, -1, 1).astype(np.float32), np.ones((n, n, n), np.float32)


@pytest.fixture(scope="module")
def prim_setup():
    st = vscan.SceneStream(4, seed=1, noise=False, dropout=0.0)
    d = vscan.make_desc(vscan.WORKLOADS["plain_320x240_128"], K=st.K, n=96)
    F, Wt = _analytic_volume(st, d)
    return st, d, F, Wt


def _model_maps(F, Wt, d, T):
    K = orc.K_levels(d)
    Tg, Tc = np.float32(T), np.float32(orc.rigid_inverse(T))
    out = []
    for l in range(3):
        out.append(orc.raycast(F, Wt, orc.volume(d), d.width >> l, d.height >> l, K[l], Tg, Tc))
    return out


def test_icp_identity_on_its_own_raycast(prim_setup):
    st, d, F, Wt = prim_setup
    T = st.pose(0)
    maps = _model_maps(F, Wt, d, T)
    K = orc.K_levels(d)
    depth, vtx, nrm = maps[0][0].astype(np.float32), maps[0][1], maps[0][2]
    A, b, n, mask = orc.icp_system(depth, K[0], vtx, nrm, np.eye(4), d.icp_dist2)
    # every model pixel finds itself, except border pixels whose round trip
    # lands a hair outside [0, W-1]
    assert int((depth > 0).sum()) - 16 <= n <= int((depth > 0).sum())
    ok, x = orc.cholesky_solve(A, b)
    assert ok and np.abs(x).max() < 5e-6
    ev = np.linalg.eigvalsh(A)
    assert np.allclose(A, A.T, atol=0) and ev.min() > -1e-9 * ev.max()


def test_icp_planar_scene_is_rank_three():
    st = vscan.SceneStream(0, seed=1, noise=False, dropout=0.0)
    st.sc.n_prims = 1  # the ground plane alone
    d = vscan.make_desc(vscan.WORKLOADS["plain_320x240_128"], K=st.K, n=64)
    F, Wt = _analytic_volume(st, d)
    T = st.pose(0)
    maps = _model_maps(F, Wt, d, T)
    K = orc.K_levels(d)
    A, b, n, _ = orc.icp_system(maps[0][0].astype(np.float32), K[0], maps[0][1], maps[0][2], np.eye(4),
                                d.icp_dist2)
    ev = np.linalg.eigvalsh(A)
    assert n > 1000
    assert ev[2] < 1e-6 * ev[5] and ev[3] > 1e-3 * ev[5]


def test_icp_all_zero_depth_fails_exactly():
    K = np.array([[262.5, 262.5, 159.5, 119.5], [131.25, 131.25, 79.5, 59.5], [65.625, 65.625, 39.5, 29.5]],
                 np.float32)
    dz = [np.zeros((240 >> l, 320 >> l), np.float32) for l in range(3)]
    mv = [np.full((240 >> l, 320 >> l, 3), -100, np.float32) for l in range(3)]
    st, T, ns = orc.icp_track(dz, K.reshape(-1), 3, [10, 5, 4], mv, mv, 0.003, np.eye(4))
    assert st == 4 and ns == 1 and np.array_equal(T, np.eye(4))


def test_warp_and_recover(prim_setup):
    """Known rigid motion <= 3 deg / 2 cm recovered from an analytic model
    (pattern of opencv rgbd test_odometry.cpp:212-332)."""
    st, d, F, Wt = prim_setup
    K = orc.K_levels(d).reshape(-1)
    T0 = st.pose(0)
    maps = _model_maps(F, Wt, d, T0)
    mv = [m[1].astype(np.float32) for m in maps]
    mn = [m[2].astype(np.float32) for m in maps]
    rng = np.random.default_rng(11)
    for _ in range(4):
        w = rng.normal(size=3)
        w *= np.radians(rng.uniform(0.5, 3.0)) / np.linalg.norm(w)
        t = rng.normal(size=3)
        t *= rng.uniform(0.005, 0.02) / np.linalg.norm(t)
        T1 = orc.expmap(np.r_[w, t]) @ T0
        depth = st.exact_depth(T1).astype(np.float32)
        pyr = [depth]
        for l in (1, 2):
            pyr.append(orc.pyr_down(pyr[-1].astype(np.float64)).astype(np.float32))
        status, T, _ = orc.icp_track(pyr, K, 3, [10, 5, 4], mv, mn, d.icp_dist2, T0)
        assert status == 0
        dt, da = orc.pose_error(T, T1)
        assert dt < 2e-3 and da < 0.2, (dt, da)


# ---------------------------------------------------------------- A13-A15 integration
def _front_camera():
    # camera at the origin looking along +z; the volume lies in front of it
    d = vscan.make_desc(vscan.WORKLOADS["plain_320x240_128"], n=32)
    d.voxel = 0.02
    d.mu = 0.06
    d.origin[0], d.origin[1], d.origin[2] = -0.32, -0.32, 0.5
    d.max_weight = 3.0
    return d


def test_integrate_fronto_parallel_plane_closed_form():
    d = _front_camera()
    vol = orc.volume(d)
    K = orc.K_levels(d)[0]
    Z = 0.83
    mm = np.full((d.height, d.width), int(Z * 1000), np.uint16)
    N = 32 ** 3
    F = np.ones(N)
    Wt = np.zeros(N)
    I4 = np.eye(4, dtype=np.float32)
    for k in range(5):
        bm, upd, cnt = orc.integrate(mm, K, I4, I4, vol, F, Wt)
    i = np.arange(32)
    zc = (np.float32(i) + np.float32(0.5)) / np.float32(1.0 / np.float32(0.02)) + np.float32(0.5)
    zc = zc.astype(np.float32)
    depth = np.float32(int(Z * 1000)) * np.float32(0.001)
    td = (depth - zc).astype(np.float32)
    Fz = np.minimum(1.0, td.astype(np.float64) / np.float64(np.float32(0.06)))
    F3, W3, U3 = F.reshape(32, 32, 32), Wt.reshape(32, 32, 32), upd.reshape(32, 32, 32)
    seen = W3 > 0
    # every column inside the image footprint is updated exactly on z where td >= -mu
    for z in range(32):
        if td[z] < -np.float32(0.06):
            assert not seen[z].any()
        elif seen[z].any():
            assert np.allclose(F3[z][seen[z]], Fz[z], atol=1e-12)
            assert np.all(W3[z][seen[z]] == 3.0)  # capped at max_weight after 5 frames
    assert seen.any()


def test_integrate_running_average_of_two_frames():
    d = _front_camera()
    d.max_weight = 100.0
    vol = orc.volume(d)
    K = orc.K_levels(d)[0]
    N = 32 ** 3
    F = np.ones(N)
    Wt = np.zeros(N)
    I4 = np.eye(4, dtype=np.float32)
    zs = []
    for Zmm in (830, 850):
        mm = np.full((d.height, d.width), Zmm, np.uint16)
        orc.integrate(mm, K, I4, I4, vol, F, Wt)
        zs.append(np.float32(Zmm) * np.float32(0.001))
    i = np.arange(32)
    zc = ((np.float32(i) + np.float32(0.5)) / np.float32(1.0 / np.float32(0.02)) + np.float32(0.5)).astype(np.float32)
    F3, W3 = F.reshape(32, 32, 32), Wt.reshape(32, 32, 32)
    mu = np.float64(np.float32(0.06))
    for z in range(32):
        both = W3[z] == 2
        if both.any():
            t1 = min(1.0, np.float32(zs[0] - zc[z]) / mu)
            t2 = min(1.0, np.float32(zs[1] - zc[z]) / mu)
            assert np.allclose(F3[z][both], (t1 + t2) / 2, atol=1e-12)


# ------------------------------------------------- production integrate with colour
def _color_frame(H, W, rgb):
    return np.stack([np.full((H, W), c, np.uint8) for c in rgb])


def test_integrate_color_constant_frame_and_frozen_voxels():
    """A constant colour frame leaves that colour in every updated voxel (to
    within the truncation of the 1/newW product); at the weight cap the voxel
    freezes: neither its TSDF nor its colour changes again (g_ctvc
    :391-393), unlike g_ctv, which keeps averaging at the cap."""
    d = _front_camera()
    vol = orc.volume(d)
    K = orc.K_levels(d)[0]
    N = 32 ** 3
    F, Wt, Cv = np.ones(N), np.zeros(N), np.zeros((3, N), np.uint8)
    I4 = np.eye(4, dtype=np.float32)
    rgb = _color_frame(d.height, d.width, (200, 100, 50))
    for Zmm in (830, 830, 830):
        orc.integrate_color(np.full((d.height, d.width), Zmm, np.uint16), rgb, K, I4, I4, vol, F, Wt, Cv)
    seen = Wt > 0
    assert seen.sum() > 1000 and np.all(Wt[seen] == 3.0)
    for k, c in enumerate((200, 100, 50)):
        # 0 where the unrounded projection leaves [0, W-1] x [0, H-1]: the
        # colour tap fails there although the rounded depth pixel is inside
        v = Cv[k][seen]
        assert np.all((v == c) | (v == c - 1) | (v == 0)) and ((v == c) | (v == c - 1)).mean() > 0.95
        assert np.all(Cv[k][~seen] == 0)
    F3, W3, C3 = F.copy(), Wt.copy(), Cv.copy()
    # two more frames of another plane and colour: frozen at max_weight = 3
    rgb2 = _color_frame(d.height, d.width, (10, 20, 30))
    for _ in range(2):
        orc.integrate_color(np.full((d.height, d.width), 850, np.uint16), rgb2, K, I4, I4, vol, F, Wt, Cv)
    assert np.array_equal(F[seen], F3[seen]) and np.array_equal(Cv[:, seen], C3[:, seen])
    assert np.all(Wt <= 3.0)


def test_integrate_color_ray_distance_sdf_on_the_optical_axis():
    """On the pixel ray through the principal point lambda = 1, so the
    production SDF (depth - |g - C| / lambda) of a voxel centred on the
    optical axis is depth - |g|. Elsewhere |g - C| / lambda is the voxel's
    depth along the ray of its rounded pixel, within the half-pixel angle of
    g_ctv's depth - z (< 1.2 mm here, 0.02 of mu, at 19 deg off axis)."""
    d = _front_camera()
    d.cx, d.cy = 160.0, 120.0  # principal point on a pixel centre
    d.origin[0] = d.origin[1] = -0.33  # voxel column (16, 16) centred on x = y = 0
    vol = orc.volume(d)
    K = orc.K_levels(d)[0]
    N = 32 ** 3
    F, Wt, Cv = np.ones(N), np.zeros(N), np.zeros((3, N), np.uint8)
    Fg, Wg = np.ones(N), np.zeros(N)
    I4 = np.eye(4, dtype=np.float32)
    mm = np.full((d.height, d.width), 830, np.uint16)
    orc.integrate_color(mm, _color_frame(d.height, d.width, (9, 9, 9)), K, I4, I4, vol, F, Wt, Cv)
    orc.integrate(mm, K, I4, I4, vol, Fg, Wg)
    F3, W3, Fg3 = F.reshape(32, 32, 32), Wt.reshape(32, 32, 32), Fg.reshape(32, 32, 32)
    depth = float(np.float32(830) * np.float32(0.001))
    vinv = np.float32(1.0) / np.float32(0.02)
    c = np.float32((np.float32(16) + np.float32(0.5)) / vinv + np.float32(-0.33))
    checked = 0
    for z in range(32):
        if W3[z, 16, 16] == 0:
            continue
        zc = np.float32((np.float32(z) + np.float32(0.5)) / vinv + np.float32(0.5))
        g = np.array([c, c, zc], np.float64)
        want = min(1.0, (depth - np.linalg.norm(g)) / float(np.float32(0.06)))
        assert abs(F3[z, 16, 16] - want) < 2e-6
        checked += 1
    assert checked > 3
    both = (W3 > 0) & (Wg.reshape(32, 32, 32) > 0) & (Fg3 < 1.0) & (F3 < 1.0)
    assert both.sum() > 1000 and np.abs(F3[both] - Fg3[both]).max() < 0.02


def test_integrate_color_failed_channel_keeps_its_value():
    """Q17: a channel whose bilinear tap reads an intensity of 0 is not
    updated; the other channels are."""
    d = _front_camera()
    vol = orc.volume(d)
    K = orc.K_levels(d)[0]
    N = 32 ** 3
    F, Wt, Cv = np.ones(N), np.zeros(N), np.zeros((3, N), np.uint8)
    I4 = np.eye(4, dtype=np.float32)
    mm = np.full((d.height, d.width), 830, np.uint16)
    orc.integrate_color(mm, _color_frame(d.height, d.width, (80, 0, 160)), K, I4, I4, vol, F, Wt, Cv)
    seen = Wt > 0
    inside = seen & (Cv[0] == 80)  # the red tap succeeded: the projection is inside the image
    assert inside.sum() > 0.95 * seen.sum()
    assert np.all(Cv[1][seen] == 0) and np.all(Cv[2][inside] == 160)


def test_raycast_texture_of_a_uniform_colour_volume(prim_setup):
    """Trilinear colour of eight equal corners is that colour: every hit of an
    analytic volume painted (12, 34, 56) renders exactly that; misses are 0."""
    st, d, F, Wt = prim_setup
    T = st.pose(2)
    K = orc.K_levels(d)
    N = F.size
    Cv = np.empty((3, N), np.uint8)
    Cv[0], Cv[1], Cv[2] = 12, 34, 56
    Tg, Tc = np.float32(T), np.float32(orc.rigid_inverse(T))
    depth, vtx, nrm, img, tex = orc.raycast_color(F, Wt, Cv, orc.volume(d), d.width, d.height, K[0], Tg, Tc)
    d2, v2, n2, i2, _ = orc.raycast(F, Wt, orc.volume(d), d.width, d.height, K[0], Tg, Tc)
    assert np.array_equal(vtx, v2) and np.array_equal(img, i2)
    hit = vtx[..., 0] != -100
    assert hit.mean() > 0.5
    full = np.ones(hit.shape, bool)
    for k, c in enumerate((12, 34, 56)):
        # a hit whose 2x2x2 colour stencil touches an unobserved voxel renders 0 (d_grvi failure)
        assert np.all((tex[k][hit] == c) | (tex[k][hit] == 0))
        assert np.all(tex[k][~hit] == 0)
        full &= tex[k] == c
    assert full[hit].mean() > 0.9


# ------------------------------------------------------------------- input pre-steps
def test_depth_reproject_identity_and_integer_shift():
    """Same camera: the frame comes back unchanged. A translation of the
    destination camera by t along x moves a fronto-parallel plane at depth Z
    by exactly fx * t / Z pixels when that is an integer."""
    K = np.array([262.5, 262.5, 159.5, 119.5], np.float32)
    rng = np.random.default_rng(5)
    mm = rng.integers(400, 2000, size=(240, 320)).astype(np.uint16)
    mm[rng.random((240, 320)) < 0.05] = 0
    I4 = np.eye(4, dtype=np.float32)
    assert np.array_equal(orc.depth_reproject(mm, K, I4, (240, 320), K), mm)
    Z = 1050  # mm; fx * t / Z = 262.5 * 0.02 / 1.05 = 5 pixels
    plane = np.full((240, 320), Z, np.uint16)
    T = np.eye(4, dtype=np.float32)
    T[0, 3] = np.float32(-0.02)  # the destination camera sits 2 cm to the right
    out = orc.depth_reproject(plane, K, T, (240, 320), K)
    assert np.all(out[:, :315] == Z) and np.all(out[:, 315:] == 0)


def test_depth_reproject_keeps_the_nearest_depth():
    """Two source pixels landing on one destination pixel: the nearer depth is
    kept whichever is visited first."""
    K = np.array([100.0, 100.0, 8.0, 8.0], np.float32)
    src = np.zeros((17, 17), np.uint16)
    src[8, 8] = 1000
    src[8, 9] = 800  # at 0.8 m, 1 px right: x = 0.008
    T = np.eye(4, dtype=np.float32)
    T[0, 3] = np.float32(-0.008)  # moves the 0.8 m point onto the principal point
    out = orc.depth_reproject(src, K, T, (17, 17), K)
    assert out[8, 8] == 800 and out[8, 0] == 0
    src2 = src.copy()
    src2[8, 8], src2[8, 9] = 0, 800
    assert orc.depth_reproject(src2, K, T, (17, 17), K)[8, 8] == 800


def test_depth_crop_closed_form():
    """A fronto-parallel plane: depths whose point lies inside the volume box
    survive, the rest are zeroed (columns outside the box's x/y extent, and
    every pixel when the plane lies beyond the box)."""
    d = _front_camera()
    vol = orc.volume(d)
    K = orc.K_levels(d)[0]
    I4 = np.eye(4, dtype=np.float32)
    mm = np.full((d.height, d.width), 830, np.uint16)
    out = orc.depth_crop(mm, K, I4, vol)
    # box: x, y in [-0.32, 0.32], z in [0.5, 1.14]; X = 0.83 (u - cx) / fx
    u = np.arange(d.width, dtype=np.float32)
    X = np.float32(0.83) * (u - np.float32(K[2])) / np.float32(K[0])
    inside_u = (X >= np.float32(-0.32)) & (X <= np.float32(-0.32) + np.float32(32 * 0.02))
    v = np.arange(d.height, dtype=np.float32)
    Y = np.float32(0.83) * (v - np.float32(K[3])) / np.float32(K[1])
    inside_v = (Y >= np.float32(-0.32)) & (Y <= np.float32(-0.32) + np.float32(32 * 0.02))
    want = np.where(inside_v[:, None] & inside_u[None, :], 830, 0)
    assert np.array_equal(out, want) and 0 < (out > 0).sum() < out.size
    far = orc.depth_crop(np.full((d.height, d.width), 1500, np.uint16), K, I4, vol)
    assert not far.any()


# ------------------------------------------------------------------- volume carving
def test_volume_carve_closed_form():
    """A masked brick with a single valid voxel survives untouched; masked
    bricks with none are reset and leave the mask; unmasked bricks are never
    touched, valid voxels or not."""
    d = _front_camera()
    vol = orc.volume(d)
    N = 32 ** 3
    rng = np.random.default_rng(3)
    F = rng.uniform(-1, 1, N)
    Wt = rng.integers(1, 4, N).astype(np.float64)
    Cv = rng.integers(1, 255, (3, N)).astype(np.uint8)
    bm = np.zeros(64, np.uint8)
    bm[[0, 1, 2, 9]] = 1
    valid = np.zeros((32, 32, 32), np.uint8)
    valid[0, 0, 8 + 7] = 1  # one voxel of brick 1 (bx = 1)
    valid[8:16, 8:16, 8:16] = 1  # all of brick (1,1,1) = 21, not in the mask
    F0, W0_, C0 = F.copy(), Wt.copy(), Cv.copy()
    n = orc.volume_carve(F, Wt, Cv, bm, valid, vol)
    assert n == 3 and list(np.nonzero(bm)[0]) == [1]
    F3, W3 = F.reshape(32, 32, 32), Wt.reshape(32, 32, 32)
    C3 = Cv.reshape(3, 32, 32, 32)
    for b in (0, 2, 9):
        bz, by, bx = b // 16, (b // 4) % 4, b % 4
        sl = (slice(bz * 8, bz * 8 + 8), slice(by * 8, by * 8 + 8), slice(bx * 8, bx * 8 + 8))
        assert np.all(F3[sl] == 1.0) and np.all(W3[sl] == 0.0) and np.all(C3[(slice(None),) + sl] == 0)
    keep = np.ones(N, bool).reshape(32, 32, 32)
    for b in (0, 2, 9):
        bz, by, bx = b // 16, (b // 4) % 4, b % 4
        keep[bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] = False
    keep = keep.reshape(-1)
    assert np.array_equal(F[keep], F0[keep]) and np.array_equal(Wt[keep], W0_[keep])
    assert np.array_equal(Cv[:, keep], C0[:, keep])


# ---------------------------------------------------------------- A16-A18 raycast
def test_raycast_analytic_scene_within_half_voxel(prim_setup):
    st, d, F, Wt = prim_setup
    T = st.pose(3)
    K = orc.K_levels(d)
    Tg, Tc = np.float32(T), np.float32(orc.rigid_inverse(T))
    depth, vtx, nrm, img, stats = orc.raycast(F, Wt, orc.volume(d), d.width, d.height, K[0], Tg, Tc)
    exact = st.exact_depth(T)
    hit = vtx[..., 0] != -100
    assert hit.mean() > 0.5
    err = np.abs(depth - exact)[hit & (exact > 0)]
    assert np.percentile(err, 95) < 0.5 * d.voxel
    assert np.array_equal(depth[hit], vtx[..., 2][hit])
    assert np.allclose(np.linalg.norm(nrm[hit], axis=-1), 1.0, atol=1e-9)


def test_raycast_empty_volume_and_looking_away():
    d = vscan.make_desc(vscan.WORKLOADS["plain_320x240_128"], n=32)
    F = np.ones((32, 32, 32), np.float32)
    W = np.zeros_like(F)
    K = orc.K_levels(d)[0]
    st = vscan.SceneStream(0, seed=1)
    T = st.pose(0)
    depth, vtx, nrm, img, _ = orc.raycast(F, W, orc.volume(d), d.width, d.height, K, np.float32(T),
                                          np.float32(orc.rigid_inverse(T)))
    assert np.all(depth == 0) and np.all(vtx == -100) and np.all(nrm == -100) and np.all(img == 0)
    # camera 3 m away, turned away from the cube: every ray fails the cone
    # test, whatever the volume holds
    Fs, Ws = np.full_like(F, -0.5), np.ones_like(W)
    C = np.array([3.0, 0.0, 0.45])
    R = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, -1.0], [1.0, 0.0, 0.0]])  # looks along +x, away from the cube
    Ta = np.eye(4)
    Ta[:3, :3] = R
    Ta[:3, 3] = -R @ C
    _, vtx, _, _, stats = orc.raycast(Fs, Ws, orc.volume(d), d.width, d.height, K, np.float32(Ta),
                                      np.float32(orc.rigid_inverse(Ta)))
    assert np.all(vtx == -100) and stats[0] == 0


# ---------------------------------------------------------------- A19 pipeline
def test_oracle_pipeline_tracks_ground_truth():
    st = vscan.SceneStream(4, seed=1)
    d = vscan.make_desc(vscan.WORKLOADS["plain_320x240_128"], K=st.K)
    o = orc.OracleSession(d, st.pose(0))
    try:
        for fr in range(12):
            s, _ = o.fuse(st.frame(fr))
            assert s == 0
        dt, da = orc.pose_error(o.pose(), st.pose(11))
        assert dt < 5e-3 and da < 0.5, (dt, da)
    finally:
        o.close()
