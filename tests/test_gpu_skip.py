"""The raycast's empty-space skipping on adversarial volumes.

The march skips the steps that stay inside regions whose voxels all read 1.0:
superbricks at distance d from the nearest band superbrick (the (2d-1)^3 box
around them), bricks without band bit, and the space outside the grid. The
skips must leave every hit/miss decision and sample position exactly as the
reference's step-by-step march (tsdf_renderer.cu:466-635) makes them. This
test uploads volumes built to exercise every kind of region: a sphere whose
free space in front of the surface holds weight with value exactly 1.0
(occupied bricks without band bit), 150 small spheres of 1.5-4 voxels
scattered far from it (each a weighted 9^3 box, so rays hit, graze or pass
just outside them), and single weighted voxels of exactly -1.0 or 1.0, so the
distance field has boxes of many sizes around small band regions. The GPU
raycast's hit mask must equal the oracle's bit for bit and its points lie
within 0.5 voxel, at several poses, for FP32 and FP16 storage; the oracle's
own answer without the scattered features must differ on more than 100 rays,
which shows the rays do meet them.
"""
import numpy as np
import pytest

import oracle_lib as orc
import vscan

pytestmark = pytest.mark.gpu

W0 = vscan.WORKLOADS["plain_320x240_128"]


def _grid(d):
    n = d.vol_dim[0]
    org = np.array(d.origin, np.float64)
    idx = np.arange(n, dtype=np.float64)
    z, y, x = np.meshgrid(idx, idx, idx, indexing="ij")  # linear order z, y, x (x fastest)
    return org[0] + (x + 0.5) * d.voxel, org[1] + (y + 0.5) * d.voxel, org[2] + (z + 0.5) * d.voxel


def adversarial_volume(d, seed, blobs=True):
    """A sphere (r 0.18 m) whose free space in front of the surface holds
    weight 1 and value exactly 1.0 up to 6 mu, plus (blobs) small spheres of
    1.5-4 voxels scattered far from it, each weighted over a 9^3 box, and
    single voxels of value exactly -1.0 or 1.0."""
    n = d.vol_dim[0]
    px, py, pz = _grid(d)
    sdf = np.sqrt(px ** 2 + py ** 2 + (pz - 0.2) ** 2) - 0.18
    F = np.ones(n ** 3, np.float64)
    W = np.zeros(n ** 3, np.float64)
    near = ((sdf > -d.mu) & (sdf < 6 * d.mu)).reshape(-1)
    F[near] = np.clip(sdf.reshape(-1)[near] / d.mu, -1.0, 1.0)
    W[near] = 1.0
    if not blobs:
        return F.astype(np.float32), W.astype(np.float32)
    rng = np.random.default_rng(seed)
    F3, W3 = F.reshape(n, n, n), W.reshape(n, n, n)
    far = np.argwhere(sdf > 12 * d.mu)
    far = far[(far.min(1) >= 5) & (far.max(1) <= n - 6)]
    for c in far[rng.choice(len(far), 150, replace=False)]:
        r = rng.uniform(1.5, 4.0) * d.voxel
        zs, ys, xs = (slice(c[k] - 4, c[k] + 5) for k in range(3))
        dist = np.sqrt((px[zs, ys, xs] - px[tuple(c)]) ** 2 + (py[zs, ys, xs] - py[tuple(c)]) ** 2 +
                       (pz[zs, ys, xs] - pz[tuple(c)]) ** 2) - r
        F3[zs, ys, xs] = np.clip(dist / d.mu, -1.0, 1.0)
        W3[zs, ys, xs] = 1.0
    for c in far[rng.choice(len(far), 200, replace=False)]:
        if W3[tuple(c)] == 0.0:
            F3[tuple(c)] = rng.choice([-1.0, 1.0])
            W3[tuple(c)] = 1.0
    return F.astype(np.float32), W.astype(np.float32)


@pytest.mark.parametrize("tsdf", [vscan.VS_TSDF_F32, vscan.VS_TSDF_F16])
def test_skips_keep_every_decision(ctx, tsdf):
    print(f"case: adversarial volume (sphere, free space at 1.0, small spheres, speckles), storage {tsdf}", flush=True)
    st = vscan.SceneStream(0, seed=1)
    d = vscan.make_desc(W0, K=st.K, tsdf=tsdf, weight=vscan.VS_WEIGHT_F32)
    s = vscan.Session(ctx, d, T_gc0=st.pose(0))
    try:
        F, W = adversarial_volume(d, seed=11)
        if tsdf == vscan.VS_TSDF_F16:
            F = F.astype(np.float16).astype(np.float32)  # what the volume holds
        s.set_volume(F, W)
        gF, gW, _ = s.volume()
        assert np.array_equal(gF.reshape(-1), F) and np.array_equal(gW.reshape(-1), W)
        F0, W0v = adversarial_volume(d, seed=11, blobs=False)
        K = orc.K_levels(d)
        vol = orc.volume(d)
        hits = changed = 0
        for fr in (0, 9, 21):
            T = st.pose(fr)
            s.set_pose(T)
            Tg, Tc = np.float32(T), np.float32(orc.rigid_inverse(T))
            for l in range(3):
                s.raycast(l)
                _, gv, _, _ = s.maps(l)
                h, w = s.level_shape(l)
                _, rv, _, _, _ = orc.raycast(F, W, vol, w, h, K[l], Tg, Tc)
                gm, rm = gv[..., 0] != -100, rv[..., 0] != -100
                assert np.array_equal(gm, rm), f"pose {fr} level {l}: {int(np.sum(gm != rm))} rays differ in hit/miss"
                if rm.any():
                    assert np.abs(gv - rv)[rm].max() <= 0.5 * d.voxel
                hits += int(rm.sum())
                # the scattered features decide some rays (without them the
                # oracle's answer differs), so they are exercised
                _, rv0, _, _, _ = orc.raycast(F0, W0v, vol, w, h, K[l], Tg, Tc)
                changed += int(np.sum(rm != (rv0[..., 0] != -100)))
        print(f"hits {hits}, rays decided by the scattered features {changed}", flush=True)
        assert hits > 10000 and changed > 100
    finally:
        s.close()
