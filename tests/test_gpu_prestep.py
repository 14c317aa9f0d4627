"""Input pre-steps on the device against the oracle: the depth reprojection
into another camera (z_tdrcc) and the volume crop (god), both bit-exact on
uint16 millimetres; and fusing the session's raw buffer after a pre-step
(NULL frame) equals fusing the same frame passed in."""
import numpy as np
import pytest

import hiprt
import oracle_lib as orc
import vscan

pytestmark = pytest.mark.gpu

W0 = vscan.WORKLOADS["plain_320x240_128"]


def _extrinsic():
    """Depth camera -> colour camera of a typical RGB-D pair: 2.5 cm baseline,
    a small rotation."""
    a = np.radians(0.7)
    R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = [-0.025, 0.001, 0.0004]
    return np.float32(T)


def _dev(arr):
    b = hiprt.DeviceBuffer(arr.nbytes)
    b.upload(np.ascontiguousarray(arr))
    return b


def test_depth_reproject_parity(ctx):
    print("case: depth reprojection 320x240 -> 320x240, device vs oracle", flush=True)
    st = vscan.SceneStream(4, seed=1)
    d = vscan.make_desc(W0, K=st.K)
    s = vscan.Session(ctx, d, T_gc0=st.pose(0))
    bufs = []
    try:
        K_src = np.array([258.0, 259.5, 161.2, 118.4], np.float32)  # a different depth-camera model
        T = _extrinsic()
        for fr in (0, 7):
            mm = st.frame(fr)
            src, dst = _dev(mm), hiprt.DeviceBuffer(mm.nbytes)
            bufs += [src, dst]
            s.depth_reproject(src.ptr.value, 320, 240, K_src, T, dst_ptr=dst.ptr.value)
            s.sync()
            got = dst.download(np.zeros_like(mm))
            want = orc.depth_reproject(mm, K_src, T, mm.shape, np.float32(st.K))
            assert np.array_equal(got, want), f"frame {fr}: {int((got != want).sum())} pixels differ"
            assert (want > 0).mean() > 0.5
    finally:
        s.close()
        for b in bufs:
            b.free()


def test_depth_crop_parity_and_raw_buffer_fusion(ctx):
    print("case: volume crop vs oracle; fusing the raw buffer after pre-steps", flush=True)
    st = vscan.SceneStream(4, seed=1)
    w = dict(W0)
    w["cube"] = 0.7  # a box smaller than the view, so the crop removes depths
    w["origin"] = (-0.35, -0.35, -0.05)
    d = vscan.make_desc(w, K=st.K)
    a = vscan.Session(ctx, d, T_gc0=st.pose(0))
    b = vscan.Session(ctx, d, T_gc0=st.pose(0))
    bufs = []
    try:
        vol = orc.volume(d)
        K = np.float32(st.K)
        for fr in range(4):
            mm = st.frame(fr)
            T_cg = np.float32(orc.rigid_inverse(a.pose()))
            want = orc.depth_crop(mm, K, T_cg, vol)
            assert 0.05 < (want == 0).mean() - (mm == 0).mean() < 0.95
            # session a: set the raw buffer, crop it with the session's own pose, fuse NULL
            a.set_depth(mm)
            a.depth_crop()
            a.sync()
            buf = _dev(mm)
            bufs.append(buf)
            b.depth_crop(buf.ptr.value, 320, 240, K, T_cg)
            b.sync()
            got = buf.download(np.zeros_like(mm))
            assert np.array_equal(got, want), f"frame {fr}: {int((got != want).sum())} pixels differ"
            a.fuse(None)
            b.fuse(want)
            assert np.array_equal(a.pose(), b.pose())
        Fa, Wa, _ = a.volume()
        Fb, Wb, _ = b.volume()
        assert np.array_equal(Fa, Fb) and np.array_equal(Wa, Wb)
    finally:
        a.close()
        b.close()
        for x in bufs:
            x.free()
