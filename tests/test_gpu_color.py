"""Colour form (VS_INTEGRATE_COLOR): the production integrate g_ctvc with a
registered colour frame, and the raycast's texture map, against the oracle.

Bars: brick mask, updated-voxel set, weights and every colour byte bit-exact
(colour arithmetic is FP32 in the reference's order on both sides); TSDF
within 1e-4; texture bytes bit-exact where both sides hit.
"""
import numpy as np
import pytest

import oracle_lib as orc
import vscan

pytestmark = pytest.mark.gpu

W0 = vscan.WORKLOADS["plain_320x240_128"]


def f32pose(T):
    return np.float32(T), np.float32(orc.rigid_inverse(T))


def color_frame(fr, H, W):
    """Seeded smooth colour pattern with a few zero intensities (which make
    the bilinear colour tap fail, decision Q17)."""
    y, x = np.mgrid[0:H, 0:W]
    r = 30 + (3 * x + 7 * fr) % 220
    g = 30 + (5 * y + 11 * fr) % 220
    b = (x + 2 * y + 13 * fr) % 256
    return np.stack([r, g, b]).astype(np.uint8)


@pytest.fixture(scope="module")
def stream():
    return vscan.SceneStream(4, seed=1)


@pytest.mark.parametrize("types", [(vscan.VS_TSDF_F32, vscan.VS_WEIGHT_F32), (vscan.VS_TSDF_F32, vscan.VS_WEIGHT_U8)])
def test_integrate_color_parity(ctx, stream, types):
    print(f"case: production colour integrate, 3 frames at ground-truth poses, types {types}", flush=True)
    w = dict(W0)
    w["max_w"] = 4.0  # reached within the test: frozen voxels are exercised
    d = vscan.make_desc(w, tsdf=types[0], weight=types[1], K=stream.K, integrate_mode=vscan.VS_INTEGRATE_COLOR)
    s = vscan.Session(ctx, d, T_gc0=stream.pose(0))
    try:
        vol = orc.volume(d)
        N = s.nvox
        F, Wt, Cv = np.ones(N), np.zeros(N), np.zeros((3, N), np.uint8)
        K = orc.K_levels(d)[0]
        for k, fr in enumerate((0, 2, 4, 6, 8, 10)):
            T = stream.pose(fr)
            mm = stream.frame(fr)
            rgb = color_frame(fr, d.height, d.width)
            s.set_pose(T)
            s.set_depth(mm)
            s.set_color(rgb)
            s.integrate()
            Tg, Tc = f32pose(T)
            bm, upd, counts = orc.integrate_color(mm, rgb, K, Tg, Tc, vol, F, Wt, Cv)
            gF, gW, gbm = s.volume()
            gC = s.volume_color().reshape(3, -1)
            assert np.array_equal(gbm, bm), f"brick mask differs at frame {fr}"
            vb, nu = s.integrate_counts()
            assert vb == counts[0] and nu == counts[1], (vb, nu, counts)
            assert np.array_equal(gW.reshape(-1), Wt.astype(np.float32)), f"weights differ at frame {fr}"
            err = np.abs(gF.reshape(-1).astype(np.float64) - F)
            assert err.max() <= 1e-4, f"TSDF error {err.max()} at frame {fr}"
            diff = gC != Cv
            assert not diff.any(), f"frame {fr}: {int(diff.sum())} colour bytes differ"
        assert (Wt == 4.0).sum() > 1000  # some voxels froze
        assert (Cv[0] > 0).sum() > 1000
    finally:
        s.close()


def test_texture_parity(ctx, stream):
    print("case: texture maps of 3 levels vs the oracle", flush=True)
    d = vscan.make_desc(W0, K=stream.K, integrate_mode=vscan.VS_INTEGRATE_COLOR)
    s = vscan.Session(ctx, d, T_gc0=stream.pose(0))
    try:
        for fr in range(0, 12, 2):
            s.set_pose(stream.pose(fr))
            s.set_depth(stream.frame(fr))
            s.set_color(color_frame(fr, d.height, d.width))
            s.integrate()
        gF, gW, _ = s.volume()
        gC = s.volume_color().reshape(3, -1)
        T = stream.pose(12)
        s.set_pose(T)
        Tg, Tc = f32pose(T)
        K = orc.K_levels(d)
        vol = orc.volume(d)
        for l in range(3):
            s.raycast(l)
            _, gv, _, _ = s.maps(l)
            gt = s.texture(l)
            h, w = s.level_shape(l)
            _, rv, _, _, rt = orc.raycast_color(gF.reshape(-1), gW.reshape(-1), gC, vol, w, h, K[l], Tg, Tc)
            gm, rm = gv[..., 0] != -100, rv[..., 0] != -100
            assert np.array_equal(gm, rm), f"level {l}: hit masks differ"
            assert np.all(gt[:, ~gm] == 0)
            both = gm & rm
            bad = (gt != rt)[:, both]
            assert not bad.any(), f"level {l}: {int(bad.sum())} texture bytes differ"
            assert (gt[0][gm] > 0).mean() > 0.8
    finally:
        s.close()


def test_color_batch_equals_alone(ctx):
    print("case: colour sessions fused alone and in a batch", flush=True)
    made = []
    for cfg, seed, Wd, Hd, n in [(4, 1, 320, 240, 128), (3, 1003, 320, 240, 96)]:
        st = vscan.SceneStream(cfg, seed=seed, width=Wd, height=Hd)
        made.append((st, vscan.make_desc(W0, K=st.K, n=n, W=Wd, H=Hd, integrate_mode=vscan.VS_INTEGRATE_COLOR)))
    alone = [vscan.Session(ctx, d, T_gc0=st.pose(0)) for st, d in made]
    mixed = [vscan.Session(ctx, d, T_gc0=st.pose(0)) for st, d in made]
    try:
        for fr in range(5):
            frames = [st.frame(fr) for st, _ in made]
            for s, (st, d), f in zip(alone, made, frames):
                s.set_color(color_frame(fr, d.height, d.width))
                s.fuse(f)
            for s, (st, d) in zip(mixed, made):
                s.set_color(color_frame(fr, d.height, d.width))
                s.sync()
            ctx.fuse_batch(mixed, frames)
        for a, m in zip(alone, mixed):
            assert np.array_equal(a.pose(), m.pose())
            assert np.array_equal(a.volume_color(), m.volume_color())
            for l in range(3):
                assert np.array_equal(a.texture(l), m.texture(l))
    finally:
        for s in alone + mixed:
            s.close()
