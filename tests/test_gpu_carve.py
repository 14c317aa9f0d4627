"""Volume carving (g_uvv) on the device against the oracle: the bricks of
the last frame's visibility mask with no valid voxel are reset, bit for bit
(TSDF, weight, colour, mask, count)."""
import numpy as np
import pytest

import oracle_lib as orc
import vscan

pytestmark = pytest.mark.gpu

W0 = vscan.WORKLOADS["plain_320x240_128"]


def test_carve_parity(ctx):
    print("case: carving with a spherical valid volume, colour session", flush=True)
    st = vscan.SceneStream(4, seed=1)
    d = vscan.make_desc(W0, K=st.K, integrate_mode=vscan.VS_INTEGRATE_COLOR)
    s = vscan.Session(ctx, d, T_gc0=st.pose(0))
    try:
        rgb = np.stack([np.full((d.height, d.width), v, np.uint8) for v in (50, 100, 150)])
        for fr in range(5):
            s.set_color(rgb)
            s.fuse(st.frame(fr))
        F, Wt, bm = s.volume()
        Cv = s.volume_color().reshape(3, -1).copy()
        F, Wt = F.reshape(-1).astype(np.float64), Wt.reshape(-1).astype(np.float64)
        n = d.vol_dim[0]
        z, y, x = np.mgrid[0:n, 0:n, 0:n]
        c = (n - 1) / 2.0
        valid = (((x - c) ** 2 + (y - c) ** 2 + (z - c * 0.6) ** 2) < (0.3 * n) ** 2).astype(np.uint8)
        want = orc.volume_carve(F, Wt, Cv, bm, valid, orc.volume(d))
        got = s.carve(valid)
        assert got == want and 0 < got < int(bm.sum() + got), (got, want)
        gF, gW, gbm = s.volume()
        assert np.array_equal(gbm, bm)
        assert np.array_equal(gF.reshape(-1), F.astype(np.float32)) and np.array_equal(gW.reshape(-1), Wt.astype(np.float32))
        assert np.array_equal(s.volume_color().reshape(3, -1), Cv)
        # the carved session still fuses
        s.set_color(rgb)
        s.fuse(st.frame(5))
        assert s.status()[0] == 0
    finally:
        s.close()
