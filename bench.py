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


def make_sessions(ctx, n_sess, frames, W, H, vol_n):
    w = vscan.WORKLOADS["fp16_640x480_512"]
    sessions, streams, bufs = [], [], []
    for i in range(n_sess):
        st = vscan.SceneStream(3, seed=1000 + i, width=W, height=H)
        d = vscan.make_desc(w, K=st.K, n=vol_n, W=W, H=H)
        s = vscan.Session(ctx, d, T_gc0=st.pose(0))
        fr = np.stack([st.frame(k) for k in range(frames)])
        buf = hiprt.DeviceBuffer(fr.nbytes)
        buf.upload(np.ascontiguousarray(fr))
        sessions.append(s)
        streams.append(st)
        bufs.append(buf)
        if (i + 1) % 16 == 0:
            log(f"  sessions ready: {i + 1}/{n_sess}")
    return sessions, streams, bufs


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


def raycast_bytes_level(sessions, level, sv_sw):
    """Algorithmic bytes of the raycast of `level`:
    sum_rays (n_near + 8 n_tri + 32 hit)(s_v+s_w) + W_l*H_l*29 per session."""
    tot = 0
    for s in sessions:
        n_near, n_tri, hit = (int(x) for x in s.raycast_counts(level))
        h, w = s.level_shape(level)
        tot += (n_near + 8 * n_tri + 32 * hit) * sv_sw + w * h * 29
    return tot


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_batch64_640x480_512_fp16.json")


def measured_traffic(kernel, args):
    """Memory bytes per launch of `kernel` on the default workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by a copy of known
    size); None for other workloads or when no such measurement exists."""
    if args.sessions != 64 or args.volume != 512 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        return json.load(f).get("traffic_bytes", {}).get(kernel)


def cpu_baseline(frames):
    """The reference algorithm's CPU path (the oracle instantiated in FP32 with
    OpenMP) on one session of the same kind: 640x480, 512^3 volume held in FP32
    (the CPU path has no FP16 storage)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc

    w = vscan.WORKLOADS["fp16_640x480_512"]
    st = vscan.SceneStream(3, seed=1000)
    d = vscan.make_desc(w, K=st.K)
    o = orc.OracleSession(d, st.pose(0), fp32=True)
    cores = orc.oracle().or_num_threads()
    mm = [st.frame(k) for k in range(frames)]
    o.fuse(mm[0])  # first frame (no tracking) is warm-up
    t0 = time.perf_counter()
    for k in range(1, frames):
        o.fuse(mm[k])
    dt = time.perf_counter() - t0
    o.close()
    return {"value": (frames - 1) / dt, "unit": "fused frames/s", "cores": cores, "kind": "port",
            "sample": f"1 session x {frames - 1} tracked frames of 640x480 after 1 warm-up frame, 512^3 volume "
                      f"in FP32 (same scene family), OpenMP {cores} threads"}


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
    ap.add_argument("--cpu-frames", type=int, default=6)
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
    sessions, streams, bufs = make_sessions(ctx, args.sessions, frames, W, H, args.volume)
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
    log(f"timed: {wall * 1e3:.1f} ms wall, {gpu_ms:.1f} ms GPU, {value:.1f} fused frames/s, "
        f"lost sessions: {sum(1 for x in statuses if x != 0)}")

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
        ray_b = sum(raycast_bytes_level(sessions, l, sv_sw) for l in range(3))
        stage = dict(zip(["reset_bilateral", "pyramid", "icp", "brick_visibility", "integrate", "raycast_march",
                          "raycast_shade"], [round(float(x), 4) for x in acc[:7]]))
        out["stage_ms"] = stage
        cand = {"k_integrate": (integ_b, acc[4]),
                "k_raycast + k_raycast_shade (levels 0-2, one launch each)": (ray_b, acc[5] + acc[6])}
        name = max(cand, key=lambda k: cand[k][1])
        b, ms = cand[name]
        ach = b / (ms * 1e-3) / 1e9
        out["roofline"] = {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                           "traffic": measured_traffic(name, args), "algorithmic_bytes_per_launch": int(b),
                           "avg_launch_ms": round(float(ms), 4)}
        out["other_kernels"] = {k: {"GB/s": round(v[0] / (v[1] * 1e-3) / 1e9, 1), "ms": round(float(v[1]), 4),
                                    "bytes": int(v[0])} for k, v in cand.items()}
        upd = sum(s.integrate_counts()[1] for s in sessions)
        out["voxel_updates_per_s"] = round(upd / (acc.sum() * 1e-3), 1)
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
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_frames)
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
