"""Snapshots (the open sparse format of vscan.h): save, restore, and the
server's evict / resume.

* A restored session holds the same volume, pose and frame count, and fusing
  further frames gives bit-identical results to the session that was never
  saved (its model maps are re-rendered at the saved pose).
* Only occupied bricks are stored.
* A slot evicted from the server and opened again from its snapshot
  continues exactly like a twin that stayed resident.
* Damaged snapshots are refused.
"""
import ctypes as C

import numpy as np
import pytest

import vscan

pytestmark = pytest.mark.gpu

W0 = vscan.WORKLOADS["plain_320x240_128"]


def _same(a, b, what):
    assert a.status()[:2] == b.status()[:2], what
    assert np.array_equal(a.pose(), b.pose()), f"{what}: pose differs"
    Fa, Wa, _ = a.volume()
    Fb, Wb, _ = b.volume()
    assert np.array_equal(Fa, Fb) and np.array_equal(Wa, Wb), f"{what}: volume differs"
    for l in range(3):
        assert all(np.array_equal(x, y) for x, y in zip(a.maps(l), b.maps(l))), f"{what}: maps L{l} differ"


@pytest.mark.parametrize("form", ["f32", "f16u8", "color"])
def test_restore_continues_bit_exact(ctx, form):
    print(f"case: snapshot / restore / continue, form {form}", flush=True)
    st = vscan.SceneStream(4, seed=1)
    w = dict(W0)
    kw = {}
    if form == "f16u8":
        w["max_w"] = 128.0
        kw = dict(tsdf=vscan.VS_TSDF_F16, weight=vscan.VS_WEIGHT_U8)
    if form == "color":
        kw = dict(integrate_mode=vscan.VS_INTEGRATE_COLOR)
    d = vscan.make_desc(w, K=st.K, **kw)
    rgb = np.stack([np.full((d.height, d.width), v, np.uint8) for v in (90, 140, 200)])
    a = vscan.Session(ctx, d, T_gc0=st.pose(0))
    b = None
    try:
        for fr in range(6):
            if form == "color":
                a.set_color(rgb)
            a.fuse(st.frame(fr))
        data = a.snapshot()
        hdr = vscan.SnapshotHeader.from_buffer_copy(data[:C.sizeof(vscan.SnapshotHeader)])
        assert hdr.magic == b"VSCANSNP" and hdr.frames == 6 and 0 < hdr.bricks_stored < hdr.bricks_total
        per = 512 * (hdr.value_bytes + hdr.weight_bytes + hdr.color_bytes) + 4
        assert len(data) == C.sizeof(vscan.SnapshotHeader) + hdr.bricks_stored * per
        b = vscan.restore(ctx, data)
        _same(a, b, "right after restore")
        if form == "color":
            assert np.array_equal(a.volume_color(), b.volume_color())
        for fr in range(6, 10):
            for s in (a, b):
                if form == "color":
                    s.set_color(rgb)
                s.fuse(st.frame(fr))
        _same(a, b, "4 frames after restore")
        if form == "color":
            assert np.array_equal(a.volume_color(), b.volume_color())
            assert np.array_equal(a.texture(0), b.texture(0))
    finally:
        a.close()
        if b is not None:
            b.close()


def test_server_evict_and_resume(ctx):
    print("case: server evict / resume vs a resident twin", flush=True)
    st = vscan.SceneStream(3, seed=1004, width=320, height=240)
    d = vscan.make_desc(W0, K=st.K, n=96, W=320, H=240)
    srv = vscan.Server(ctx, 2)
    twin = vscan.Session(ctx, d, T_gc0=st.pose(0))
    try:
        slot = srv.open(d, T_gc0=st.pose(0))
        for fr in range(5):
            srv.step([slot], [st.frame(fr)])
            twin.fuse(st.frame(fr))
        data = srv.evict(slot)
        assert srv.usage() == (0, 0)
        slot2 = srv.open_snapshot(data)
        assert slot2 is not None
        for fr in range(5, 9):
            srv.step([slot2], [st.frame(fr)])
            twin.fuse(st.frame(fr))
        srv.session(slot2).sync()
        _same(twin, srv.session(slot2), "resumed slot")
    finally:
        srv.close()
        twin.close()


def test_damaged_snapshots_are_refused(ctx):
    print("case: damaged snapshots", flush=True)
    st = vscan.SceneStream(4, seed=1)
    d = vscan.make_desc(W0, K=st.K, n=64)
    a = vscan.Session(ctx, d, T_gc0=st.pose(0))
    try:
        a.fuse(st.frame(0))
        data = bytearray(a.snapshot())
    finally:
        a.close()
    with pytest.raises(vscan.VsError):
        vscan.restore(ctx, bytes(data[:-1]))  # truncated
    bad = bytearray(data)
    bad[0:8] = b"NOTASNAP"
    with pytest.raises(vscan.VsError):
        vscan.restore(ctx, bytes(bad))
    hs = C.sizeof(vscan.SnapshotHeader)
    bad = bytearray(data)
    bad[hs:hs + 4] = (10 ** 9).to_bytes(4, "little")  # brick index out of range
    with pytest.raises(vscan.VsError):
        vscan.restore(ctx, bytes(bad))
