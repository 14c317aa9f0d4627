"""Pins of the oracle's visible-brick set (A13), surface normal (A17) and
shading (A18) against geometry the oracle's own code does not restate:
hand-derived brick sets of exactly representable rays, an FP64 recomputation
of the sampled points of a fronto-parallel plane, and the analytic normals
of planar and spherical distance fields. Each `pin_*` function raises
AssertionError when the oracle is wrong; tests/test_oracle_mutants.py runs
them against oracle builds with one planted mistake each.

Reference: g_fvsc_Find_Visible_Sub_Cubes (z_yooji_cuda_volume_integrator.cu:84-152)
with d_gpv_Get_Position_in_Voxel (_yooji_cuda_math.h:265-282) and ROUNDF
(_yooji_cuda_defines.h:53); d_csnt_Compute_Surface_Normal_from_TSDF
(z_yooji_cuda_tsdf_renderer.cu:288-464); the normal image of g_rmi
(:1163-1164) with light (0, 0, -1) (z_yooji_object_scanner.cpp:1313).

Runs on the CPU (no GPU marker).
"""
import numpy as np

import oracle_lib as orc
import vscan

I4 = np.eye(4, dtype=np.float32)


# ---------------------------------------------------------------- A13
def _one_ray_mask(oz):
    """Brick mask of a frame whose only valid pixel is the principal point
    (160, 120) at 36 m, with a 32^3 grid of 1/64 m voxels, mu = 3 voxels.
    The ray is the optical axis: its direction (0,0,64)/64 and every sample
    z = 36 + n*3/64 are exact in FP32, so the voxel coordinates are exact:
    x = y = 15 (brick 1) and z = (36 + 3n/64 - oz)*64 - 0.5."""
    W, H = 320, 240
    K = np.array([262.5, 262.5, 160.0, 120.0], np.float32)
    mm = np.zeros((H, W), np.uint16)
    mm[120, 160] = 36000  # 36000 * 0.001f == 36.0f exactly
    vol = orc.OrVolume()
    for k in range(3):
        vol.dim[k] = 32
    vol.origin[0] = vol.origin[1] = -15.5 / 64
    vol.origin[2] = oz
    vol.voxel = 1.0 / 64
    vol.mu = 3.0 / 64
    vol.max_w = 10.0
    F = np.ones(32 ** 3)
    Wt = np.zeros(32 ** 3)
    bm, _, counts = orc.integrate(mm, K, I4, I4, vol, F, Wt)
    assert counts[0] == bm.sum()
    return bm.reshape(4, 4, 4)  # (bz, by, bx)


def _bricks(*bz):
    m = np.zeros((4, 4, 4), np.uint8)
    for z in bz:
        m[z, 1, 1] = 1
    return m


def pin_a13_single_ray_brick_sets():
    # z(n) = 12.5 + 3n for n = -4..1: 0.5, 3.5, 6.5, 9.5, 12.5, 15.5, which
    # round half away from zero to 1, 4, 7, 10, 13, 16: bricks 0, 0, 0, 1, 1, 2
    got = _one_ray_mask(36.0 - 4 * 3 / 64 - 1 / 64)
    assert np.array_equal(got, _bricks(0, 1, 2)), np.argwhere(got)
    # z(-4) = -0.5 exactly: inside the grid (d_gpv accepts -0.5), ROUNDF
    # gives -1, and the macro's missing parentheses leave -1 undivided, so the
    # brick index is out of range and the loop breaks: no brick at all
    got = _one_ray_mask(36.0 - 4 * 3 / 64)
    assert not got.any(), np.argwhere(got)
    # sample n = -4 lies below the grid (skipped), n = -3 hits z = -0.5 and
    # breaks before n = -2..1 are reached
    got = _one_ray_mask(36.0 - 3 * 3 / 64)
    assert not got.any(), np.argwhere(got)


def pin_a13_plane_brick_set():
    """Fronto-parallel plane at 0.83 m filling a 320x240 frame, identity
    pose, 64^3 grid of 1 cm voxels from (-0.32, -0.24, 0.60), mu = 3 cm. The
    oracle's brick set is compared with one computed in FP64 from the ray
    geometry: bricks of samples farther than 1e-3 voxel from a rounding or
    grid boundary must be marked, and every marked brick must be one that a
    sample can reach with either rounding."""
    W, H = 320, 240
    K = np.array([262.5, 262.5, 159.5, 119.5], np.float32)
    n, vox, mu = 64, 0.01, 0.03
    o = np.array([-0.32, -0.24, 0.60])
    vol = orc.OrVolume()
    for k in range(3):
        vol.dim[k] = n
        vol.origin[k] = o[k]
    vol.voxel, vol.mu, vol.max_w = vox, mu, 10.0
    mm = np.full((H, W), 830, np.uint16)
    F = np.ones(n ** 3)
    Wt = np.zeros(n ** 3)
    bm, _, _ = orc.integrate(mm, K, I4, I4, vol, F, Wt)
    got = set(map(tuple, np.argwhere(bm.reshape(8, 8, 8))))  # (bz, by, bx)

    Z = np.float64(np.float32(830) * np.float32(0.001))
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    p = np.stack([Z * (xs - 159.5) / 262.5, Z * (ys - 119.5) / 262.5, np.full_like(xs, Z)], -1).reshape(-1, 3)
    dirn = p / np.linalg.norm(p, axis=1, keepdims=True)  # the pixel ray leaves the camera centre
    sure, maybe = set(), set()
    tol = 1e-3
    for k in range(-4, 2):
        v = (p + k * mu * dirn - o) / vox - 0.5
        inside = np.all((v >= -0.5 + tol) & (v < n - 0.5 - tol), axis=1)
        edge = np.all((v >= -0.5 - tol) & (v < n - 0.5 + tol), axis=1) & ~inside
        frac = v + 0.5 - np.floor(v + 0.5)
        amb = np.any((frac < tol) | (frac > 1 - tol), axis=1)
        lo = np.floor(v + 0.5 - tol).astype(int) // 8
        hi = np.floor(v + 0.5 + tol).astype(int) // 8
        for b in np.unique(lo[inside & ~amb][:, ::-1], axis=0):
            sure.add(tuple(b))
        cand = inside | edge
        for arr in (lo, hi):
            sel = arr[cand][:, ::-1]
            ok = np.all((sel >= 0) & (sel < 8), axis=1)
            for b in np.unique(sel[ok], axis=0):
                maybe.add(tuple(b))
    assert len(sure) > 20
    assert sure <= got, f"bricks the geometry requires but the oracle misses: {sorted(sure - got)[:8]}"
    assert got <= maybe, f"bricks no sample reaches: {sorted(got - maybe)[:8]}"


# ---------------------------------------------------------------- A17-A18
def _look_at(C, target=(0.0, 0.0, 0.0)):
    """World -> camera pose of a camera at C looking at target, z up
    (x right, y down, z forward)."""
    C = np.asarray(C, float)
    f = np.asarray(target, float) - C
    f /= np.linalg.norm(f)
    r = np.array([f[1], -f[0], 0.0])
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    R = np.stack([r, d, f])
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = -R @ C
    return T


def _field_volume(sdf, n=48, voxel=0.01, mu_vox=5.0):
    """Truncated signed distance sdf/mu clipped to [-1, 1] at the voxel
    centres of an n^3 grid centred on the origin, weight 1 everywhere."""
    d = vscan.make_desc(vscan.WORKLOADS["plain_320x240_128"], n=n)
    d.voxel = voxel
    d.mu = mu_vox * voxel
    for k in range(3):
        d.origin[k] = -0.5 * n * voxel
    c = -0.5 * n * voxel + (np.arange(n) + 0.5) * voxel
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    F = np.clip(sdf(x, y, z) / d.mu, -1.0, 1.0).astype(np.float32)
    return d, F, np.ones_like(F)


def _raycast(d, F, Wt, T):
    K = orc.K_levels(d)[0]
    return orc.raycast(F, Wt, orc.volume(d), d.width, d.height, K, np.float32(T),
                       np.float32(orc.rigid_inverse(T)))


PLANE_N = np.array([0.3, -0.2, 1.0]) / np.linalg.norm([0.3, -0.2, 1.0])
PLANE_CAM = (0.25, 0.45, 0.5)


def _plane_scene():
    d, F, Wt = _field_volume(lambda x, y, z: PLANE_N[0] * x + PLANE_N[1] * y + PLANE_N[2] * z)
    T = _look_at(PLANE_CAM)
    return d, F, Wt, T


def _world(vtx, T):
    R, t = T[:3, :3], T[:3, 3]
    return (vtx - t) @ R  # camera -> world: R^T (p - t)


def pin_a17_plane_normals():
    """The 32-tap gradient of a linear field is exact: every hit on the plane
    has the plane normal in the camera frame. (A few rays also stop where the
    field meets the grid's faces, as the reference's do: voxels outside the
    grid read as free space there. Those hits are not on the plane.)"""
    d, F, Wt, T = _plane_scene()
    _, vtx, nrm, _, _ = _raycast(d, F, Wt, T)
    hit = vtx[..., 0] != -100
    on = hit & (np.abs(_world(vtx, T) @ PLANE_N) < 0.5 * d.voxel)
    assert on.sum() > 2000 and on.sum() > 0.95 * hit.sum()
    want = T[:3, :3] @ PLANE_N
    err = np.abs(nrm[on] - want).max()
    assert err < 1e-5, f"plane normal off by {err:.3g}"


def pin_a17_sphere_normals():
    """Normals of a sphere of radius 0.12 m against the radial direction at
    the hit point, rotated into the camera frame."""
    c0 = np.array([0.02, -0.01, 0.0])
    d, F, Wt = _field_volume(lambda x, y, z: np.sqrt((x - c0[0]) ** 2 + (y - c0[1]) ** 2 + (z - c0[2]) ** 2) - 0.12)
    T = _look_at((0.45, 0.2, 0.3))
    _, vtx, nrm, _, _ = _raycast(d, F, Wt, T)
    hit = vtx[..., 0] != -100
    radial = _world(vtx, T) - c0
    on = hit & (np.abs(np.linalg.norm(radial, axis=-1) - 0.12) < 0.5 * d.voxel)
    assert on.sum() > 2000 and on.sum() > 0.95 * hit.sum()
    want = (radial[on] / np.linalg.norm(radial[on], axis=1, keepdims=True)) @ T[:3, :3].T
    ang = np.degrees(np.arccos(np.clip(np.sum(nrm[on] * want, axis=1), -1, 1)))
    assert np.median(ang) < 0.2 and ang.max() < 1.5, (np.median(ang), ang.max())


def pin_a18_shade():
    """Normal image = (uchar)clamp((0.8 n.l + 0.2) 255) with l = (0, 0, -1),
    from the analytic normal (one value for every hit of the plane)."""
    d, F, Wt, T = _plane_scene()
    _, vtx, _, img, _ = _raycast(d, F, Wt, T)
    hit = vtx[..., 0] != -100
    on = hit & (np.abs(_world(vtx, T) @ PLANE_N) < 0.5 * d.voxel)
    nc = T[:3, :3] @ PLANE_N
    v = (0.8 * (-nc[2]) + 0.2) * 255.0
    assert 0.2 < v - np.floor(v) < 0.8  # the scene keeps the value clear of a rounding boundary
    want = np.uint8(np.clip(np.floor(v), 0, 255))
    assert on.sum() > 2000
    assert np.all(img[on] == want), (np.unique(img[on]), want)
    assert np.all(img[~hit] == 0)


def test_a13_single_ray_brick_sets():
    pin_a13_single_ray_brick_sets()


def test_a13_plane_brick_set():
    pin_a13_plane_brick_set()


def test_a17_plane_normals():
    pin_a17_plane_normals()


def test_a17_sphere_normals():
    pin_a17_sphere_normals()


def test_a18_shade():
    pin_a18_shade()
