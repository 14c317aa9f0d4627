"""Launch-shape autotuning (vs_ctx_set_tuning / vs_ctx_tuning).

Every candidate shape gives the same results, so a batch fused while the
tuner times its candidates, and after it settles on one, must be
bit-identical to the same batch fused with the fixed shapes; likewise a
session's own frames. The settled shapes are reported per size class with
the mean time of every candidate, and a second context reads them back from
the cache file without tuning again.
"""
import numpy as np
import pytest

import vscan

pytestmark = pytest.mark.gpu

W0 = vscan.WORKLOADS["plain_320x240_128"]
FRAMES = 16  # 12 tuning steps, the settling step, then graph replays


def _made():
    made = []
    for cfg, seed, Wd, Hd, n in [(4, 1, 320, 240, 128), (3, 1003, 320, 240, 96), (3, 1004, 160, 120, 64)]:
        st = vscan.SceneStream(cfg, seed=seed, width=Wd, height=Hd)
        made.append((st, vscan.make_desc(W0, K=st.K, n=n, W=Wd, H=Hd)))
    return made


def _same(a, b, what):
    assert np.array_equal(a.pose(), b.pose()), f"{what}: pose differs"
    Fa, Wa, _ = a.volume()
    Fb, Wb, _ = b.volume()
    assert np.array_equal(Fa, Fb) and np.array_equal(Wa, Wb), f"{what}: volume differs"
    for l in range(3):
        assert all(np.array_equal(x, y) for x, y in zip(a.maps(l), b.maps(l))), f"{what}: maps L{l} differ"


def test_tuned_steps_equal_fixed_shapes(ctx, tmp_path):
    print("case: batch of 3 and a lone session, tuned context vs fixed shapes, 16 frames", flush=True)
    made = _made()
    frames = [[st.frame(fr) for st, _ in made] for fr in range(FRAMES)]
    cache = str(tmp_path / "shapes.txt")
    fixed = [vscan.Session(ctx, d, T_gc0=st.pose(0)) for st, d in made]
    lone_fixed = vscan.Session(ctx, made[1][1], T_gc0=made[1][0].pose(0))
    tctx = vscan.Context(0)
    tuned, lone_tuned = [], None
    try:
        tctx.set_tuning(vscan.VS_TUNE_ONLINE, cache)
        tuned = [vscan.Session(tctx, d, T_gc0=st.pose(0)) for st, d in made]
        lone_tuned = vscan.Session(tctx, made[1][1], T_gc0=made[1][0].pose(0))
        for fr in range(FRAMES):
            ctx.fuse_batch(fixed, frames[fr])
            tctx.fuse_batch(tuned, frames[fr])
            lone_fixed.fuse(frames[fr][1])
            lone_tuned.fuse(frames[fr][1])
        for a, b in zip(fixed, tuned):
            a.sync()
            b.sync()
            _same(a, b, "batch member")
        lone_fixed.sync()
        lone_tuned.sync()
        _same(lone_fixed, lone_tuned, "lone session")
        rep = tctx.tuning()
        assert set(rep) == {"320x240 v128 f32+f32 n4", "320x240 v128 f32+f32 n1"}, rep
        for cls, r in rep.items():
            print(cls, r, flush=True)
            assert r["march"] in ("8x8", "16x4", "32x2") and r["shade"] in vscan.Context.SHADE_OPTS
            ms = r["ms"]
            assert all(v > 0 for v in ms["march"].values()) and all(v > 0 for v in ms["shade"].values())
            assert all(v > 0 for v in ms["integrate"].values())
            # ray blocks: the fastest, unless the first candidate (8x8) is
            # within 3% of it; integrate grids: the fastest (the reported
            # means are rounded to 1e-4 ms: a case that close to the 3%
            # boundary is not decided here)
            for stage in ("march", "shade"):
                best = min(ms[stage].values())
                first = ms[stage]["8x8"]
                if abs(first - best / (1.0 - 0.03)) < 2e-4:
                    continue
                want = "8x8" if first < best / (1.0 - 0.03) else min(ms[stage], key=ms[stage].get)
                assert r[stage] == want or ms[stage][r[stage]] == ms[stage][want], (stage, r[stage], ms[stage])
            assert ms["integrate"][r["integrate"]] == min(ms["integrate"].values())
    finally:
        for s in tuned + [lone_tuned]:
            if s is not None:
                s.close()
        tctx.close()
        for s in fixed + [lone_fixed]:
            s.close()
    # a new context reads the settled shapes back
    c2 = vscan.Context(0)
    try:
        c2.set_tuning(vscan.VS_TUNE_ONLINE, cache)
        back = c2.tuning()
        assert set(back) == set(rep)
        for cls in rep:
            assert back[cls]["ms"] is None
            for k in ("march", "shade", "integrate"):
                assert back[cls][k] == rep[cls][k]
    finally:
        c2.close()
