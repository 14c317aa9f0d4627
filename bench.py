#!/usr/bin/env python3
"""Fused-frame throughput of the batched fusion step on one MI355X.

Workload (the headline metric's configuration): 64 independent scan sessions,
each a 640x480 uint16 depth stream of its own seeded random-primitive scene and
orbit, each fusing into its own 512^3 TSDF (FP16 value + u8 weight, cap 128)
over a 1.5 m cube. One step = one fused frame (bilateral -> pyramid -> ICP
(3 levels, 10/5/4 iterations) -> brick-culled integration -> raycast of 3
levels) for every session, launched as one batch. Depth frames are rendered on
the host beforehand and resident in HBM when the timed region starts.

Prints one JSON line (stdout). Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vcl-volume-scanner_amd"))

import hiprt  # noqa: E402
import vscan  # noqa: E402

METRIC = "fused frames/s at 640x480/512^3 TSDF; p50/p99 frame latency, 64 sessions"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_sessions(ctx, n_sess, frames, W, H, vol_n, host_frames=0):
    """Sessions of the benchmark set with their first `frames` frames resident
    in HBM, and the `host_frames` frames after those kept in host memory."""
    w = vscan.WORKLOADS["fp16_640x480_512"]
    sessions, streams, bufs, host = [], [], [], []
    for i in range(n_sess):
        st = vscan.SceneStream(3, seed=1000 + i, width=W, height=H)
        d = vscan.make_desc(w, K=st.K, n=vol_n, W=W, H=H)
        s = vscan.Session(ctx, d, T_gc0=st.pose(0))
        fr = np.stack([st.frame(k) for k in range(frames)])
        buf = hiprt.DeviceBuffer(fr.nbytes)
        buf.upload(np.ascontiguousarray(fr))
        if host_frames:
            host.append(np.stack([st.frame(frames + k) for k in range(host_frames)]))
        sessions.append(s)
        streams.append(st)
        bufs.append(buf)
        if (i + 1) % 16 == 0:
            log(f"  sessions ready: {i + 1}/{n_sess}")
    return sessions, streams, bufs, host


def plain_serial(ctx, streams, bufs, W, H, vol_n, warm, steps, stream):
    """The plain form: FP32 value + FP32 weight volumes, each session fused on
    its own (vs_fuse_frame, one graph per session), one after the other, on
    the same resident frames. Returns fused frames/s."""
    w = vscan.WORKLOADS["fp16_640x480_512"]
    sess = []
    for st in streams:
        d = vscan.make_desc(w, tsdf=vscan.VS_TSDF_F32, weight=vscan.VS_WEIGHT_F32, K=st.K, n=vol_n, W=W, H=H)
        d.max_weight = 100.0
        sess.append(vscan.Session(ctx, d, T_gc0=st.pose(0)))
    L = vscan.lib()

    def run(k):
        for s, b in zip(sess, bufs):
            vscan.check(L.vs_fuse_frame(s.h, vscan.P(b.ptr.value + k * W * H * 2), 0, stream.h), "vs_fuse_frame")

    for k in range(warm):
        run(k)
    stream.sync()
    t0 = time.perf_counter()
    for k in range(warm, warm + steps):
        run(k)
    stream.sync()
    dt = time.perf_counter() - t0
    for s in sess:
        s.close()
    return len(streams) * steps / dt


def frame_ptrs(bufs, k, W, H):
    off = k * W * H * 2
    return [b.ptr.value + off for b in bufs]


def integrate_bytes(sessions, sv_sw):
    """Algorithmic bytes of one integrate launch over the batch:
    N_vis*(s_v+s_w) + N_upd*(s_v+s_w) + W*H*4 per session."""
    tot = 0
    for s in sessions:
        vb, nu = s.integrate_counts()
        tot += (vb * 512 + nu) * sv_sw + s.W * s.H * 4
    return tot


def raycast_bytes_level(sessions, level, acc=None):
    """Algorithmic bytes of the raycast of `level`: the bytes the reference's
    lookups request (d_gtvu / d_gtvi / d_csnt: nothing for a tap outside the
    grid, the weight alone for a voxel without weight, weight + value
    otherwise, stencils up to their first failing tap; counted by the
    instrumented march) + W_l*H_l*29 written per session (depth 4, vertex 12,
    normal 12, shade 1). acc (dict, optional) sums the tap counts."""
    tot = 0
    for s in sessions:
        c = [int(x) for x in s.raycast_counts(level)]
        h, w = s.level_shape(level)
        tot += c[5] + c[6] + c[7] + w * h * 29
        if acc is not None:
            for k, v in zip(["nearest_taps", "trilinear_evals", "hits", "nearest_outside_grid",
                             "nearest_no_weight", "nearest_bytes", "trilinear_bytes", "normal_bytes"], c):
                acc[k] = acc.get(k, 0) + v
            acc["map_bytes"] = acc.get("map_bytes", 0) + w * h * 29
    return tot


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_batch64_640x480_512_fp16.json")


def measured_traffic(kernel, args):
    """Memory bytes per launch of `kernel` on the default workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by a copy of known
    size), and the file they come from; (None, None) for other workloads or
    when no such measurement exists."""
    if args.sessions != 64 or args.volume != 512 or not os.path.exists(TRAFFIC_FILE):
        return None, None
    with open(TRAFFIC_FILE) as f:
        return json.load(f).get("traffic_bytes", {}).get(kernel), os.path.relpath(TRAFFIC_FILE, ROOT)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def stage_stats(ms_rows, names):
    a = np.asarray(ms_rows, float)
    return {n: {"median": round(float(np.median(a[:, i])), 3), "p99": round(float(np.percentile(a[:, i], 99)), 3)}
            for i, n in enumerate(names)}


def tracking_quality(sessions, streams, frame):
    """Pose of each session against its stream's ground truth at `frame`:
    sessions off by more than 1 cm or 1 degree (mis-registered, whatever
    their status), and the median / max camera-centre and rotation errors."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc

    errs = [orc.pose_error(s.pose(), st.pose(frame)) for s, st in zip(sessions, streams)]
    t = np.array([e[0] for e in errs]) * 1e3
    a = np.array([e[1] for e in errs])
    return {"frame": frame, "sessions_off_by_1cm_or_1deg": int(np.sum((t > 10.0) | (a > 1.0))),
            "translation_mm": {"median": round(float(np.median(t)), 3), "max": round(float(t.max()), 3)},
            "rotation_deg": {"median": round(float(np.median(a)), 4), "max": round(float(a.max()), 4)},
            "definition": "camera centre distance and geodesic rotation angle of each session's tracked pose "
                          "to the synthetic stream's ground truth"}


def _cpu_run(orc, desc, stream, frames, warm):
    """The oracle's FP32/OpenMP pipeline (the reference algorithm's CPU path)
    over `frames` frames of one stream; returns (fused frames, seconds, per-
    frame stage ms after `warm` frames)."""
    o = orc.OracleSession(desc, stream.pose(0), fp32=True)
    rows, fused, t_all = [], 0, 0.0
    try:
        for k in range(frames):
            mm = stream.frame(k)
            t0 = time.perf_counter()
            st, ms = o.fuse(mm)
            dt = time.perf_counter() - t0
            if k >= warm:
                rows.append(list(ms) + [dt * 1e3])
                t_all += dt
                fused += 1
    finally:
        o.close()
    return fused, t_all, rows


CPU_STAGES = ["bilateral", "pyramid", "track", "integrate", "raycast", "fused_frame"]


def cpu_baseline(sessions=4, frames=31, warm=1):
    """The reference algorithm's CPU path (the oracle instantiated in FP32 with
    OpenMP over all host threads) on a bounded sample of the benchmark's
    workload: the first `sessions` sessions of the benchmark set (640x480,
    512^3 volume held in FP32, the CPU path having no FP16 storage), `frames`
    frames each, the first `warm` of each untimed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc

    w = vscan.WORKLOADS["fp16_640x480_512"]
    cores = orc.oracle().or_num_threads()
    fused, secs, rows = 0, 0.0, []
    for i in range(sessions):
        st = vscan.SceneStream(3, seed=1000 + i)
        d = vscan.make_desc(w, tsdf=vscan.VS_TSDF_F32, weight=vscan.VS_WEIGHT_F32, K=st.K)
        f, t, r = _cpu_run(orc, d, st, frames, warm)
        fused += f
        secs += t
        rows += r
    return {"value": round(fused / secs, 3), "unit": "fused frames/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{sessions} sessions of the benchmark set x {frames - warm} frames of 640x480 after {warm} "
                      f"untimed frame each, 512^3 volume in FP32, oracle FP32/OpenMP pipeline ({cores} threads)",
            "seconds": round(secs, 2), "stage_ms": stage_stats(rows, CPU_STAGES)}


def cpu_plain_configs():
    """The CPU-baseline plan for the plain configurations: 320x240 / 128^3
    FP32, 30 frames, and 640x480 / 256^3 FP32 (first 65 of its 300 frames),
    per stage median and p99 after 5 warm-up frames."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc

    out = {}
    for name, cfg, frames in (("plain_320x240_128", 0, 30), ("plain_640x480_256", 1, 65)):
        st = vscan.SceneStream(cfg, seed=1)
        d = vscan.make_desc(vscan.WORKLOADS[name], K=st.K)
        fused, secs, rows = _cpu_run(orc, d, st, frames, 5)
        out[name] = {"fused_frames_per_s": round(fused / secs, 3), "frames_timed": fused,
                     "stage_ms": stage_stats(rows, CPU_STAGES)}
    out["cores"] = orc.oracle().or_num_threads()
    out["cpu_model"] = cpu_model()
    out["definition"] = ("oracle FP32/OpenMP pipeline, steady_clock per stage, median/p99 over the frames after "
                         "5 warm-up frames")
    return out


def gpu_plain_config2(ctx, frames=35, warm=5):
    """The plain FP32 pipeline on the GPU at 640x480 / 256^3 over 1.2 m (one
    session, vs_fuse_frame, events around each stage): per-stage median/p99 ms,
    GB/s of algorithmic bytes, and the error against the FP64 oracle on
    identical inputs (frame 0: filter, integrate at the ground-truth pose,
    raycast of that volume)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc

    w = vscan.WORKLOADS["plain_640x480_256"]
    st = vscan.SceneStream(1, seed=1)
    d = vscan.make_desc(w, K=st.K)
    s = vscan.Session(ctx, d, T_gc0=st.pose(0))
    s.set_timing(True)
    rows, byts = [], []
    W, H = d.width, d.height
    try:
        for k in range(frames):
            s.fuse(st.frame(k))
            s.sync()
            if k < warm:
                continue
            ms = s.stage_ms()
            rows.append(list(ms) + [float(ms.sum())])
            vb, nu = s.integrate_counts()
            # bytes per stage: filter W*H*(2+4); pyramid sum_l (read l-1 + write l) * 4; ICP W_l*H_l*28 per
            # configured iteration; integrate N_vis*8 + N_upd*8 + W*H*4; raycast as the headline's
            pyr = sum(((W >> (l - 1)) * (H >> (l - 1)) + (W >> l) * (H >> l)) * 4 for l in (1, 2))
            icp = sum((W >> l) * (H >> l) * 28 * d.iters[l] for l in range(3))
            byts.append([W * H * 6, pyr, icp, (vb * 512 + nu) * 8 + W * H * 4])
        ray = sum(raycast_bytes_level([s], l) for l in range(3))
        stage = stage_stats(rows, ["bilateral", "pyramid", "track", "integrate", "raycast", "fused_frame"])
        med = {k: v["median"] for k, v in stage.items()}
        b = np.mean(np.asarray(byts, float), axis=0)
        gbs = {"bilateral": b[0], "pyramid": b[1], "track": b[2], "integrate": b[3], "raycast": ray}
        gbs = {k: round(v / (med[k] * 1e-3) / 1e9, 1) for k, v in gbs.items() if med[k] > 0}
    finally:
        s.close()
    # identical-input errors against the FP64 oracle
    mm = st.frame(0)
    g = vscan.Session(ctx, d, T_gc0=st.pose(0))
    try:
        g.set_depth(mm)
        g.bilateral()
        filt = g.filtered(0)
        ref = orc.bilateral(mm, d.bil_sigma_s, d.bil_sigma_r, d.bil_radius)
        g.integrate()
        F, Wt, _ = g.volume()
        vol = orc.volume(d)
        N = F.size
        oF, oW = np.ones(N), np.zeros(N)
        T = st.pose(0)
        orc.integrate(mm, orc.K_levels(d)[0], np.float32(T), np.float32(orc.rigid_inverse(T)), vol, oF, oW)
        seen = oW > 0
        g.raycast(0)
        _, gv, _, _ = g.maps(0)
        _, rv, _, _, _ = orc.raycast(F, Wt, vol, W, H, orc.K_levels(d)[0], np.float32(T),
                                     np.float32(orc.rigid_inverse(T)))
        err = {"filtered_depth_max_mm": float(np.abs(filt - ref).max() * 1e3),
               "filter_mask_mismatches": int(np.sum((filt > 0) != (ref > 0))),
               "tsdf_max": float(np.abs(F.reshape(-1) - oF)[seen].max()),
               "weight_mismatches": int(np.sum(Wt.reshape(-1) != oW.astype(np.float32))),
               "raycast_hit_mask_mismatches": int(np.sum((gv[..., 0] != -100) != (rv[..., 0] != -100))),
               "raycast_point_max_voxels": float(np.abs(gv - rv)[rv[..., 0] != -100].max() / d.voxel)}
    finally:
        g.close()
    return {"workload": "one session, 640x480 uint16 depth, 256^3 FP32 TSDF + FP32 weight over 1.2 m, helical orbit "
                        "(sphere + box + plane)", "frames_timed": len(rows), "stage_ms": stage, "GB_per_s": gbs,
            "error_vs_fp64_oracle_frame0": err,
            "definition": "events around each stage of vs_fuse_frame; GB/s of algorithmic bytes (bilateral W*H*6, "
                          "pyramid reads+writes, ICP W_l*H_l*28 per configured iteration, integrate "
                          "N_vis*8+N_upd*8+W*H*4, raycast reference-requested bytes + maps); the FP32 256^3 volume "
                          "(134 MB) fits the 256 MB Infinity Cache, so these are not HBM fractions"}


def serve(ctx, streams, bufs, frames, W, H, seconds, with_plain):
    """Serving form: Poisson session arrivals (2/s after 64 at start),
    lifetimes 150-900 frames at 30 fps +-5 ms, volumes drawn from
    {128,256,384,512}^3 FP16 + u8, admitted under a 200 GB budget into a
    128-slot table. Frames stream from the resident pools (forward/backward
    walk over each scene's frames). The same trace is then run through the
    plain pipeline (FP32 volumes, one vs_fuse_frame per frame)."""
    import serving

    w = vscan.WORKLOADS["fp16_640x480_512"]
    K = streams[0].K
    reqs = serving.make_requests(7, seconds, len(bufs), initial=64)
    ptr = lambda scene, k: bufs[scene].ptr.value + k * W * H * 2  # noqa: E731
    pose0 = lambda scene: streams[scene].pose(0)  # noqa: E731

    def desc16(n):
        return vscan.make_desc(w, K=K, n=n, W=W, H=H)

    def desc32(n):
        d = vscan.make_desc(w, tsdf=vscan.VS_TSDF_F32, weight=vscan.VS_WEIGHT_F32, K=K, n=n, W=W, H=H)
        d.max_weight = 100.0
        return d

    log(f"serving: {len(reqs)} sessions over {seconds:.0f} s, batched slot table")
    loop = serving.ServeLoop(serving.BatchedBackend(ctx, 128, 200 << 30), desc16, ptr, frames, pose0)
    res = loop.run(reqs, seconds, log=log)
    loop.close()
    res["definition"] = ("latency = completion of the step that fused the frame (GPU event) - frame arrival; "
                         "sessions arrive Poisson 2/s after 64 at t=0, 30 fps +-5 ms, 150-900 frames, volumes "
                         "{128,256,384,512}^3 FP16+u8, 200 GB budget, 128 slots")
    log(f"  batched: {res['fused_frames_per_s']} fused frames/s, p50 {res['latency_ms']['p50']} ms, "
        f"p99 {res['latency_ms']['p99']} ms")
    if with_plain:
        reqs = serving.make_requests(7, seconds, len(bufs), initial=64)
        log("serving: same trace, plain serial pipeline")
        loop = serving.ServeLoop(serving.SerialBackend(ctx, 128, 200 << 30), desc32, ptr, frames, pose0,
                                 max_batch=1)
        plain = loop.run(reqs, seconds, log=log)
        loop.close()
        log(f"  plain: {plain['fused_frames_per_s']} fused frames/s, p50 {plain['latency_ms']['p50']} ms, "
            f"p99 {plain['latency_ms']['p99']} ms")
        res["plain_serial_fp32"] = plain
    return res


TUNE_STEPS = 13  # the tuner's 12 candidate steps + the step that settles the shapes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--sessions", type=int, default=64)
    ap.add_argument("--volume", type=int, default=512)
    ap.add_argument("--cpu-sessions", type=int, default=4)
    ap.add_argument("--cpu-frames", type=int, default=31)
    ap.add_argument("--no-host-frames", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="length of the serving-form run (0 = skip)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="end with one device-to-device copy of this size (a known-size kernel for counter scaling)")
    ap.add_argument("--serve-plain", type=int, default=1,
                    help="also run the same request trace through the plain serial pipeline")
    ap.add_argument("--no-tune", action="store_true",
                    help="fixed launch shapes instead of the online autotuner")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build shards sessions across GPUs only by running one process per "
              "GPU, and multi-GPU runs are not part of this measurement (use --gpus 1).", file=sys.stderr)
        sys.exit(2)
    W, H = 640, 480
    K, Wm = args.steps, args.warmup
    T = 0 if args.no_tune else TUNE_STEPS
    frames = T + Wm + K + 2
    log(f"bench: {args.sessions} sessions, {W}x{H}, {args.volume}^3 FP16+u8, warmup {Wm}, steps {K}")
    ctx = vscan.Context(0)
    sessions, streams, bufs, host = make_sessions(ctx, args.sessions, frames, W, H, args.volume,
                                                  0 if args.no_host_frames else 2 * K)
    stream = hiprt.Stream()

    def step(k):
        ctx.fuse_batch(sessions, frame_ptrs(bufs, T + k, W, H), is_host=False, stream=stream.h)

    tune_s = 0.0
    if T:
        # launch-shape autotuning: the first steps of the batch's size class
        # time every candidate shape (setup, before the warmup)
        log(f"tuning launch shapes ({T} steps)")
        ctx.set_tuning(vscan.VS_TUNE_ONLINE)
        t0 = time.perf_counter()
        for k in range(T):
            ctx.fuse_batch(sessions, frame_ptrs(bufs, k, W, H), is_host=False, stream=stream.h)
        stream.sync()
        tune_s = time.perf_counter() - t0
        # the settled shapes stay; the later runs (plain pipeline, serving)
        # tune nothing inside their timed regions
        ctx.set_tuning(vscan.VS_TUNE_CACHED)
    log("warmup")
    for k in range(Wm):
        step(k)
    stream.sync()
    hiprt.device_sync()
    ev = [hiprt.Event() for _ in range(K + 1)]
    log("timed region")
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(K):
        step(Wm + i)
        ev[i + 1].record(stream)
    stream.sync()
    hiprt.device_sync()
    wall = time.perf_counter() - t0
    lat = np.array([ev[i].elapsed_ms(ev[i + 1]) for i in range(K)])
    gpu_ms = float(lat.sum())
    n = args.sessions
    value = n * K / wall
    statuses = [s.status()[0] for s in sessions]
    quality = tracking_quality(sessions, streams, T + Wm + K - 1)
    log(f"timed: {wall * 1e3:.1f} ms wall, {gpu_ms:.1f} ms GPU, {value:.1f} fused frames/s, "
        f"lost sessions: {sum(1 for x in statuses if x != 0)}, off ground truth by >1 cm/1 deg: "
        f"{quality['sessions_off_by_1cm_or_1deg']}")

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "fused frames/s", "n_gpus": 1, "steps": K,
        "warmup": Wm, "ms_per_step": round(wall * 1e3 / K, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n} sessions x 640x480 uint16 depth, {args.volume}^3 FP16 TSDF + u8 weight "
                               f"(cap 128) over 1.5 m, seeded random-primitive scenes and orbits",
                   "sessions": n, "width": W, "height": H, "volume": args.volume, "tsdf": "f16", "weight": "u8",
                   "icp_iters": [10, 5, 4]},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                       "definition": "GPU time of the batched step that fuses the frame (all frames of a step "
                                     "arrive when it starts)"},
        "gpu_ms_per_step": round(gpu_ms / K, 3),
        "sessions_failed_status": int(sum(1 for x in statuses if x != 0)),
        "tracking_vs_ground_truth": quality,
    }
    if T:
        out["launch_shapes"] = {"tuning_steps": T, "tuning_s": round(tune_s, 3), "classes": ctx.tuning(),
                                "definition": "shapes the online autotuner settled on per size class; ms = mean "
                                              "stage time of each candidate over its tuning steps"}

    if not args.no_roofline:
        # measurement pass: eager launches with events between kernels
        ctx.set_timing(True)
        reps = 3
        acc = np.zeros(8)
        integ_b = ray_b = 0
        for r in range(reps):
            step(Wm + K + (r % 2))
            acc += ctx.stage_ms()
        acc /= reps
        ctx.set_timing(False)
        sv_sw = 3
        integ_b = integrate_bytes(sessions, sv_sw)
        taps = {}
        ray_b = sum(raycast_bytes_level(sessions, l, taps) for l in range(3))
        stage = dict(zip(["reset_bilateral", "pyramid", "icp", "brick_visibility", "integrate", "raycast_march",
                          "raycast_shade"], [round(float(x), 4) for x in acc[:7]]))
        out["stage_ms"] = stage
        cand = {"k_integrate": (integ_b, acc[4]),
                "k_raycast + k_raycast_shade (levels 0-2, one launch each)": (ray_b, acc[5] + acc[6])}
        name = max(cand, key=lambda k: cand[k][1])
        b, ms = cand[name]
        ach = b / (ms * 1e-3) / 1e9
        traffic, traffic_src = measured_traffic(name, args)
        out["roofline"] = {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                           "traffic": traffic, "algorithmic_bytes_per_launch": int(b),
                           "avg_launch_ms": round(float(ms), 4),
                           "traffic_source": (f"{traffic_src}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this "
                                              f"workload, copied (not measured in this run)") if traffic else None,
                           "physical_frac": (round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic
                                             else None),
                           "definition": "achieved = bytes the reference's lookups request (instrumented march, "
                                         "raycast_taps below) / average launch time of this run; physical_frac = "
                                         "measured memory traffic / the same time / peak"}
        out["raycast_taps"] = taps
        out["other_kernels"] = {k: {"GB/s": round(v[0] / (v[1] * 1e-3) / 1e9, 1), "ms": round(float(v[1]), 4),
                                    "bytes": int(v[0])} for k, v in cand.items()}
        upd = sum(s.integrate_counts()[1] for s in sessions)
        out["voxel_updates_per_s"] = round(upd / (acc.sum() * 1e-3), 1)
    if host:
        # the same steps with frames handed over in host memory (the PCIe
        # upload of each session's 614 KB frame inside the timed region), on
        # the frames that follow the ones fused so far: K steps from ordinary
        # (pageable) numpy arrays, then K from page-locked buffers
        def host_run(frames_of):
            stream.sync()
            t0 = time.perf_counter()
            for i in range(K):
                ctx.fuse_batch(sessions, frames_of(i), is_host=True, stream=stream.h)
            stream.sync()
            return time.perf_counter() - t0

        hw = host_run(lambda i: [hf[i] for hf in host])
        pinned = []
        for hf in host:
            pb = hiprt.PinnedBuffer(hf[K:].shape, hf.dtype)
            pb.array[...] = hf[K:]
            pinned.append(pb)
        pw = host_run(lambda i: [pb.array[i] for pb in pinned])
        for pb in pinned:
            pb.free()
        out["host_frames"] = {"value": round(n * K / hw, 2), "unit": "fused frames/s",
                              "ms_per_step": round(hw * 1e3 / K, 3),
                              "pinned_value": round(n * K / pw, 2), "pinned_ms_per_step": round(pw * 1e3 / K, 3),
                              "definition": "same batched step with each session's uint16 frame passed in host "
                                            "memory (vs_fuse_batch is_host=1: one 614 KB upload per session inside "
                                            "the timed region); value: ordinary numpy arrays, pinned_value: "
                                            "page-locked buffers (hipHostMalloc)"}
        log(f"host frames: {out['host_frames']['value']} fused frames/s, "
            f"pinned {out['host_frames']['pinned_value']}")
        del host
    if args.serve_seconds > 0:
        for s in sessions:  # the serving run opens its own sessions
            s.close()
        sessions = []
        out["serving"] = serve(ctx, streams, bufs, frames, W, H, args.serve_seconds, args.serve_plain)
    if not args.no_plain:
        log("plain FP32 one-session-at-a-time pipeline")
        for s in sessions:  # free the FP16 volumes first
            s.close()
        sessions = []
        pf = plain_serial(ctx, streams, bufs, W, H, args.volume, 2, min(K, 5), stream)
        out["plain_serial_fp32"] = {"value": round(pf, 2), "unit": "fused frames/s",
                                    "definition": "same sessions and frames, FP32 value + FP32 weight volumes, "
                                                  "each session fused alone (vs_fuse_frame) one after another"}
        out["speedup_vs_plain_serial"] = round(value / pf, 3)
    if not args.no_plain:
        log("plain FP32 pipeline at 640x480 / 256^3, per stage")
        out["plain_640x480_256_gpu"] = gpu_plain_config2(ctx)
    if not args.no_cpu_baseline:
        log("cpu baseline (benchmark workload sample)")
        out["cpu_baseline"] = cpu_baseline(args.cpu_sessions, args.cpu_frames)
        log(f"  {out['cpu_baseline']['value']} fused frames/s in {out['cpu_baseline']['seconds']} s")
        log("cpu baseline (plain configurations)")
        out["cpu_plain_configs"] = cpu_plain_configs()
    for s in sessions:
        s.close()
    for b in bufs:
        b.free()
    if args.calib_copy_mib > 0:
        nb = args.calib_copy_mib << 20
        a, b = hiprt.DeviceBuffer(nb), hiprt.DeviceBuffer(nb)
        hiprt.copy_d2d(b, a, nb)
        hiprt.device_sync()
        log(f"calibration copy: {nb} bytes device to device")
        a.free()
        b.free()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
