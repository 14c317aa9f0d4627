"""vs_icp_system_noalloc: the ICP system entry point that allocates nothing
(its tile partials live in the library's fixed device array) gives the same
bits as vs_icp_system, on two calls in a row (the second on a smaller input,
so the first call's partials are still in the array), on a second session,
and refuses a level larger than the array with VS_EINVAL."""
import numpy as np
import pytest

import hiprt
import oracle_lib as orc
import vscan

pytestmark = pytest.mark.gpu

W0 = vscan.WORKLOADS["plain_320x240_128"]


def _ready_session(ctx, W, H, seed):
    st = vscan.SceneStream(4, seed=seed, width=W, height=H)
    d = vscan.make_desc(W0, K=st.K, W=W, H=H)
    s = vscan.Session(ctx, d, T_gc0=st.pose(0))
    vol = orc.volume(d)
    N = d.vol_dim[0] * d.vol_dim[1] * d.vol_dim[2]
    F, Wt = np.ones(N), np.zeros(N)
    K = orc.K_levels(d)
    for fr in (0, 3, 6):
        T = st.pose(fr)
        orc.integrate(st.frame(fr), K[0], np.float32(T), np.float32(orc.rigid_inverse(T)), vol, F, Wt)
    sh = (d.vol_dim[2], d.vol_dim[1], d.vol_dim[0])
    s.set_volume(F.reshape(sh).astype(np.float32), Wt.reshape(sh).astype(np.float32))
    s.set_pose(st.pose(6))
    for l in range(3):
        s.raycast(l)
    s.set_depth(st.frame(7))
    s.bilateral()
    s.pyramid()
    return s


def _noalloc(s, level, T21):
    h, w = s.level_shape(level)
    sums = hiprt.DeviceBuffer(29 * 8)
    mask = hiprt.DeviceBuffer(h * w)
    try:
        T = np.ascontiguousarray(T21, np.float32).reshape(16)
        vscan.check(vscan.lib().vs_icp_system_noalloc(s.h, level, T.ctypes.data_as(vscan.P), sums.ptr, mask.ptr,
                                                      None), "vs_icp_system_noalloc")
        s.sync()
        out = sums.download(np.zeros(29, np.float64))
        m = mask.download(np.zeros((h, w), np.uint8))
    finally:
        sums.free()
        mask.free()
    A = np.zeros((6, 6))
    k = 0
    for i in range(6):
        for j in range(i, 6):
            A[i, j] = A[j, i] = out[k]
            k += 1
    return A, out[21:27], out[27], m


def _check(s, level, T21, what):
    A, b, n, mask = s.icp_system(level, T21, want_mask=True)
    A2, b2, n2, mask2 = _noalloc(s, level, T21)
    assert np.array_equal(A, A2) and np.array_equal(b, b2), f"{what}: sums differ"
    assert n2 == n and n > 100, (what, n, n2)
    assert np.array_equal(mask, mask2), f"{what}: masks differ"
    print(f"  {what}: {n} correspondences, bit-identical", flush=True)


def test_noalloc_icp_system_matches_allocating_entry_point(ctx):
    print("case: vs_icp_system_noalloc vs vs_icp_system", flush=True)
    T21 = orc.expmap(np.array([0.002, -0.001, 0.001, 0.002, 0.001, -0.003]))
    s = _ready_session(ctx, 320, 240, 1)
    try:
        _check(s, 0, T21, "320x240 level 0 (300 tiles)")
        _check(s, 2, T21, "320x240 level 2 (20 tiles, after the level-0 call)")
        _check(s, 1, np.eye(4), "320x240 level 1 (75 tiles)")
    finally:
        s.close()
    s = _ready_session(ctx, 160, 120, 2)
    try:
        _check(s, 0, T21, "160x120 level 0 (75 tiles, another session)")
    finally:
        s.close()


def test_noalloc_icp_system_refuses_levels_beyond_its_array(ctx):
    print("case: vs_icp_system_noalloc size limit", flush=True)
    d = vscan.make_desc(W0, n=32, W=2048, H=1040)  # 64 x 130 = 8320 tiles at level 0, 2080 at level 1
    s = vscan.Session(ctx, d)
    sums = hiprt.DeviceBuffer(29 * 8)
    try:
        T = np.eye(4, dtype=np.float32).reshape(16)
        code = vscan.lib().vs_icp_system_noalloc(s.h, 0, T.ctypes.data_as(vscan.P), sums.ptr, None, None)
        assert code == vscan.VS_EINVAL
        assert b"8192" in vscan.lib().vs_last_error()
        code = vscan.lib().vs_icp_system_noalloc(s.h, 1, T.ctypes.data_as(vscan.P), sums.ptr, None, None)
        assert code == vscan.VS_OK
        s.sync()
    finally:
        sums.free()
        s.close()
