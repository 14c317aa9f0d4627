"""Batched execution and the precision form.

* A session gives bit-identical results alone (vs_fuse_frame) and inside a
  batch with other sessions of different frame and volume sizes
  (vs_fuse_batch): per-session reductions are summed in a fixed tile order and
  every brick and pixel is computed independently of its neighbours in the
  batch.
* The FP16-value / u8-weight volume tracks the FP64 oracle within
  0.5 mm / 0.05 deg and keeps the surface within 1 mm.
"""
import numpy as np
import pytest

import oracle_lib as orc
import vscan

pytestmark = pytest.mark.gpu


def _sessions(ctx, tsdf, weight, specs):
    out = []
    for cfg, seed, W, H, n in specs:
        st = vscan.SceneStream(cfg, seed=seed, width=W, height=H)
        w = dict(vscan.WORKLOADS["plain_320x240_128"])
        if tsdf == vscan.VS_TSDF_F16:
            w["max_w"] = 128.0
        d = vscan.make_desc(w, tsdf=tsdf, weight=weight, n=n, W=W, H=H, K=st.K)
        out.append((st, d))
    return out


@pytest.mark.parametrize("types", [(vscan.VS_TSDF_F32, vscan.VS_WEIGHT_F32), (vscan.VS_TSDF_F16, vscan.VS_WEIGHT_U8)])
def test_batch_equals_alone(ctx, types):
    print(f"case: batch vs alone, types {types}", flush=True)
    specs = [(4, 1, 320, 240, 128), (3, 1001, 640, 480, 64), (3, 1002, 320, 240, 96)]
    made = _sessions(ctx, types[0], types[1], specs)
    alone = [vscan.Session(ctx, d, T_gc0=st.pose(0)) for st, d in made]
    mixed = [vscan.Session(ctx, d, T_gc0=st.pose(0)) for st, d in made]
    try:
        for fr in range(6):
            frames = [st.frame(fr) for st, _ in made]
            for s, f in zip(alone, frames):
                s.fuse(f)
            ctx.fuse_batch(mixed, frames)
            for a, m in zip(alone, mixed):
                assert a.status() == m.status()
                assert np.array_equal(a.pose(), m.pose()), f"frame {fr}: pose differs"
        for a, m in zip(alone, mixed):
            Fa, Wa, Ba = a.volume()
            Fm, Wm, Bm = m.volume()
            assert np.array_equal(Fa, Fm) and np.array_equal(Wa, Wm) and np.array_equal(Ba, Bm)
            for l in range(3):
                assert all(np.array_equal(x, y) for x, y in zip(a.maps(l), m.maps(l)))
    finally:
        for s in alone + mixed:
            s.close()


def test_split_batch_equals_alone(ctx):
    """A batch of 16 or more sessions runs in parts, independent branches
    of the captured step graph; every session still matches its lone run bit
    for bit."""
    print("case: 18-session batch (graph branches) vs alone, 4 frames", flush=True)
    specs = [(3, 2000 + i, 160 if i % 3 else 320, 120 if i % 3 else 240, 64 if i % 2 else 96) for i in range(18)]
    made = _sessions(ctx, vscan.VS_TSDF_F16, vscan.VS_WEIGHT_U8, specs)
    alone = [vscan.Session(ctx, d, T_gc0=st.pose(0)) for st, d in made]
    mixed = [vscan.Session(ctx, d, T_gc0=st.pose(0)) for st, d in made]
    try:
        for fr in range(4):
            frames = [st.frame(fr) for st, _ in made]
            for s, f in zip(alone, frames):
                s.fuse(f)
            ctx.fuse_batch(mixed, frames)
        for a, m in zip(alone, mixed):
            assert a.status() == m.status()
            assert np.array_equal(a.pose(), m.pose())
            Fa, Wa, _ = a.volume()
            Fm, Wm, _ = m.volume()
            assert np.array_equal(Fa, Fm) and np.array_equal(Wa, Wm)
            for l in range(3):
                assert all(np.array_equal(x, y) for x, y in zip(a.maps(l), m.maps(l)))
    finally:
        for s in alone + mixed:
            s.close()


def test_fp16_form_tracks_fp64_oracle(ctx):
    print("case: FP16 + u8 volume vs FP64 oracle, 320x240 / 128^3, 20 frames", flush=True)
    st = vscan.SceneStream(4, seed=1)
    w = dict(vscan.WORKLOADS["plain_320x240_128"])
    w["max_w"] = 128.0
    d = vscan.make_desc(w, tsdf=vscan.VS_TSDF_F16, weight=vscan.VS_WEIGHT_U8, K=st.K)
    d64 = vscan.make_desc(w, K=st.K)
    s = vscan.Session(ctx, d, T_gc0=st.pose(0))
    o = orc.OracleSession(d64, st.pose(0))
    worst = (0.0, 0.0)
    try:
        for fr in range(20):
            mm = st.frame(fr)
            s.fuse(mm)
            o.fuse(mm)
            assert s.status()[0] == 0
            dt, da = orc.pose_error(s.pose(), o.pose())
            worst = (max(worst[0], dt), max(worst[1], da))
        print(f"FP16 form worst pose gap: {worst[0] * 1e3:.4f} mm, {worst[1]:.5f} deg")
        assert worst[0] <= 0.5e-3 and worst[1] <= 0.05
        gF, gW, _ = s.volume()
        oF, oW = o.volume()
        seen = (gW > 0) & (oW > 0)
        err = np.abs(gF - oF)[seen]
        print(f"FP16 TSDF error vs FP64: median {np.median(err):.2e}, p99 {np.percentile(err, 99):.2e}")
        _, gv, _, _ = s.maps(0)
        _, ov, _ = o.maps(0)
        both = (gv[..., 0] != -100) & (ov[..., 0] != -100)
        surf = np.linalg.norm(gv - ov, axis=-1)[both]
        print(f"FP16 surface gap: p50 {np.median(surf) * 1e3:.4f} mm, p99 {np.percentile(surf, 99) * 1e3:.4f} mm")
        assert np.percentile(surf, 99) <= 1e-3
    finally:
        s.close()
        o.close()


def test_failure_keeps_pose_and_skips_fusion(ctx):
    """An all-zero frame after the first one: the solve fails (J^T J = 0), the
    pose is kept and the volume is untouched (so_Scan_Object's failure path)."""
    print("case: tracking failure on an empty frame", flush=True)
    st = vscan.SceneStream(4, seed=2)
    d = vscan.make_desc(vscan.WORKLOADS["plain_320x240_128"], K=st.K)
    s = vscan.Session(ctx, d, T_gc0=st.pose(0))
    try:
        s.fuse(st.frame(0))
        T1 = s.pose()
        F1, W1, _ = s.volume()
        s.fuse(np.zeros((d.height, d.width), np.uint16))
        status, frames, solves = s.status()
        assert status == vscan.VS_TRACK_LOST and frames == 1 and solves == 1
        assert np.array_equal(s.pose(), T1)
        F2, W2, _ = s.volume()
        assert np.array_equal(F1, F2) and np.array_equal(W1, W2)
        s.fuse(st.frame(1))  # recovers on the next real frame
        assert s.status()[0] == 0
    finally:
        s.close()


def test_wide_offsets_equal_32bit_offsets(ctx):
    """The march and the shading address the volume with 32-bit offsets
    when every session of a launch has its voxel records below 4 GB, and
    with 64-bit addresses otherwise. A session batched with a 1024^3 FP32
    session (8 GB of records) takes the 64-bit path and must give the same
    bits as the same session alone (32-bit path)."""
    print("case: 128^3 session alone vs batched with a 1024^3 FP32 session, 4 frames", flush=True)
    (st, d), (stb, db) = _sessions(ctx, vscan.VS_TSDF_F32, vscan.VS_WEIGHT_F32,
                                   [(4, 1, 320, 240, 128), (3, 7, 320, 240, 1024)])
    assert vscan.session_bytes(db) > 8 << 30  # 4 GiB of values and 4 GiB of weights
    alone = vscan.Session(ctx, d, T_gc0=st.pose(0))
    mixed = vscan.Session(ctx, d, T_gc0=st.pose(0))
    big = vscan.Session(ctx, db, T_gc0=stb.pose(0))
    try:
        for fr in range(4):
            alone.fuse(st.frame(fr))
            ctx.fuse_batch([mixed, big], [st.frame(fr), stb.frame(fr)])
            assert alone.status() == mixed.status()
            assert np.array_equal(alone.pose(), mixed.pose()), f"frame {fr}: pose differs"
        for l in range(3):
            assert all(np.array_equal(x, y) for x, y in zip(alone.maps(l), mixed.maps(l))), f"maps L{l} differ"
    finally:
        for s in (alone, mixed, big):
            s.close()


def test_host_frames_equal_device_frames(ctx):
    """Frames handed to vs_fuse_batch in host memory (ordinary numpy arrays or
    page-locked buffers) go through the batch's two device staging areas in
    turn; the results equal those of the same frames passed in device memory,
    over consecutive steps (both areas, reused)."""
    import hiprt
    print("case: host frames (pageable, pinned) vs device frames", flush=True)
    specs = [(4, 11, 320, 240, 96), (3, 12, 640, 480, 64), (3, 13, 320, 240, 64)]
    made = _sessions(ctx, vscan.VS_TSDF_F16, vscan.VS_WEIGHT_U8, specs)
    runs = {k: [vscan.Session(ctx, d, T_gc0=st.pose(0)) for st, d in made] for k in ("page", "pin", "dev")}
    bufs = []
    try:
        for fr in range(5):
            frames = [np.ascontiguousarray(st.frame(fr), np.uint16) for st, _ in made]
            ctx.fuse_batch(runs["page"], frames)
            pins = []
            for f in frames:
                pb = hiprt.PinnedBuffer(f.shape, np.uint16)
                pb.array[...] = f
                pins.append(pb)
            ctx.fuse_batch(runs["pin"], [pb.array for pb in pins])
            devs = []
            for f in frames:
                db = hiprt.DeviceBuffer(f.nbytes)
                db.upload(f)
                devs.append(db)
            ctx.fuse_batch(runs["dev"], [db.ptr.value for db in devs], is_host=False)
            for s in runs["dev"]:
                s.pose()  # waits for the step
            bufs += pins + devs
            for a, b, c in zip(runs["page"], runs["pin"], runs["dev"]):
                assert a.status() == b.status() == c.status()
                assert np.array_equal(a.pose(), c.pose()) and np.array_equal(b.pose(), c.pose()), f"frame {fr}"
        for a, b, c in zip(runs["page"], runs["pin"], runs["dev"]):
            Fc, Wc, _ = c.volume()
            for x in (a, b):
                Fx, Wx, _ = x.volume()
                assert np.array_equal(Fx, Fc) and np.array_equal(Wx, Wc)
                for l in range(3):
                    assert all(np.array_equal(u, v) for u, v in zip(x.maps(l), c.maps(l)))
    finally:
        for ss in runs.values():
            for s in ss:
                s.close()
        for b in bufs:
            b.free()
