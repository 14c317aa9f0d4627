"""Serving form: the slot table (vs_server_*).

* Sessions opened while others are running join the next step; a step fuses
  whichever slots have a frame; closing a slot leaves the others untouched.
  Every session's pose, status, volume and model maps are bit-identical to the
  same frames fused alone with vs_fuse_frame.
* Admission: a session whose bytes would exceed the HBM budget, or a full
  table, is refused with VS_BUDGET and allocates nothing.
"""
import numpy as np
import pytest

import hiprt
import vscan

pytestmark = pytest.mark.gpu


def _spec(tsdf, weight, cfg, seed, W, H, n):
    st = vscan.SceneStream(cfg, seed=seed, width=W, height=H)
    w = dict(vscan.WORKLOADS["plain_320x240_128"])
    if tsdf == vscan.VS_TSDF_F16:
        w["max_w"] = 128.0
    return st, vscan.make_desc(w, tsdf=tsdf, weight=weight, n=n, W=W, H=H, K=st.K)


def _same(a, m, what):
    assert a.status() == m.status(), what
    assert np.array_equal(a.pose(), m.pose()), f"{what}: pose differs"
    Fa, Wa, Ba = a.volume()
    Fm, Wm, Bm = m.volume()
    assert np.array_equal(Fa, Fm) and np.array_equal(Wa, Wm) and np.array_equal(Ba, Bm), f"{what}: volume differs"
    for l in range(3):
        assert all(np.array_equal(x, y) for x, y in zip(a.maps(l), m.maps(l))), f"{what}: maps L{l} differ"


@pytest.mark.parametrize("types", [(vscan.VS_TSDF_F32, vscan.VS_WEIGHT_F32), (vscan.VS_TSDF_F16, vscan.VS_WEIGHT_U8)])
def test_server_mix_equals_alone(ctx, types):
    print(f"case: server mix vs alone, types {types}", flush=True)
    specs = {"A": (4, 1, 320, 240, 128), "B": (3, 1001, 640, 480, 64), "C": (3, 1002, 320, 240, 96)}
    made = {k: _spec(types[0], types[1], *v) for k, v in specs.items()}
    # (opens, closes, frames this step); a device-frame step checks the gather path
    plan = [(["A"], [], ["A"]), ([], [], ["A"]), (["B"], [], ["A", "B"]), ([], [], ["B"]),
            (["C"], [], ["C", "A", "B"]), ([], [], ["C"]), ([], ["B"], ["A", "C"]), ([], [], ["A", "C"])]
    srv = vscan.Server(ctx, 8, tsdf=types[0], weight=types[1])
    alone = {k: vscan.Session(ctx, d, T_gc0=st.pose(0)) for k, (st, d) in made.items()}
    slot, count, bufs = {}, {k: 0 for k in made}, []
    try:
        for step, (opens, closes, who) in enumerate(plan):
            for k in opens:
                st, d = made[k]
                slot[k] = srv.open(d, T_gc0=st.pose(0))
                assert slot[k] is not None
            for k in closes:
                _same(alone[k], srv.session(slot[k]), f"{k} before close")
                srv.close_slot(slot.pop(k))
            frames = []
            for k in who:
                frames.append(made[k][0].frame(count[k]))
                alone[k].fuse(frames[-1])
                count[k] += 1
            if step == 4:
                dev = []
                for f in frames:
                    b = hiprt.DeviceBuffer(f.nbytes)
                    b.upload(np.ascontiguousarray(f))
                    dev.append(b)
                bufs.extend(dev)
                srv.step([slot[k] for k in who], [b.ptr.value for b in dev], is_host=False)
            else:
                srv.step([slot[k] for k in who], frames)
            for k in who:
                srv.session(slot[k]).sync()
        for k in slot:
            _same(alone[k], srv.session(slot[k]), f"{k} at end ({count[k]} frames)")
        n_open, used = srv.usage()
        assert n_open == 2 and used == sum(vscan.session_bytes(made[k][1]) for k in slot)
    finally:
        srv.close()
        for s in alone.values():
            s.close()
        for b in bufs:
            b.free()


def test_server_admission(ctx):
    print("case: server admission by budget and slots", flush=True)
    st, d = _spec(vscan.VS_TSDF_F16, vscan.VS_WEIGHT_U8, 3, 5, 320, 240, 64)
    need = vscan.session_bytes(d)
    srv = vscan.Server(ctx, 3, tsdf=vscan.VS_TSDF_F16, weight=vscan.VS_WEIGHT_U8, hbm_budget=2 * need + need // 2)
    try:
        a = srv.open(d, T_gc0=st.pose(0))
        b = srv.open(d, T_gc0=st.pose(0))
        assert a is not None and b is not None and a != b
        assert srv.open(d, T_gc0=st.pose(0)) is None  # over budget
        assert srv.usage() == (2, 2 * need)
        srv.close_slot(a)
        c = srv.open(d, T_gc0=st.pose(0))
        assert c == a  # freed slot and bytes are reused
        srv.step([b, c], [st.frame(0), st.frame(0)])
        srv.session(b).sync()
        assert srv.session(b).status()[1] == 1 and srv.session(c).status()[1] == 1
        with pytest.raises(vscan.VsError):
            srv.step([b, b], [st.frame(1), st.frame(1)])  # a slot twice in one step
        with pytest.raises(vscan.VsError):
            srv.open(vscan.make_desc(dict(vscan.WORKLOADS["plain_320x240_128"]), n=64, W=320, H=240, K=st.K))
    finally:
        srv.close()
    full = vscan.Server(ctx, 1, tsdf=vscan.VS_TSDF_F16, weight=vscan.VS_WEIGHT_U8)
    try:
        assert full.open(d, T_gc0=st.pose(0)) == 0
        assert full.open(d, T_gc0=st.pose(0)) is None  # no free slot
    finally:
        full.close()
