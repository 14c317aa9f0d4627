"""Request loop of the serving form over the slot table (vs_server_*).

Scan sessions arrive as a Poisson process; each streams depth frames at a
fixed rate with jitter for its lifetime and is then closed, freeing its slot
and HBM. The loop admits sessions when a slot and budget are free (a refused
session waits and retries), and keeps at most `max_inflight` batched steps
queued on one HIP stream. A step fuses, for every open session, its oldest
frame that has arrived by the time the step is formed; frames that arrive
while steps run join the next step (continuous batching). Frames of a session
are fused in order, one per step.

Latency of a frame = completion of the step that fused it (HIP event, GPU
clock, referenced to an event synchronised when the loop starts) minus its
arrival time on the host clock. A session's frames start arriving when it is
admitted; the time it waited for admission (a full table or budget) is
reported separately, as are the host's time in the open, close and step calls.
"""
import collections
import dataclasses
import time

import numpy as np

import hiprt
import vscan


@dataclasses.dataclass
class Request:
    rid: int
    t_open: float      # seconds after the loop starts
    n_frames: int      # lifetime in frames
    scene: int         # index of the frame pool it streams from
    vol_n: int         # volume edge in voxels
    arrivals: np.ndarray = None  # frame arrival times (s)


def make_requests(seed, duration, n_scenes, rate=2.0, initial=16, life=(150, 900), vols=(128, 256, 384, 512),
                  fps=30.0, jitter=0.005):
    """`initial` sessions open at t=0, then Poisson arrivals at `rate`/s until
    `duration`. Frame j of a session arrives at t_open + j/fps + U(-jitter,
    jitter), clipped so arrivals stay ordered."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    opens = [0.0] * initial
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        opens.append(t)
    for i, t0 in enumerate(opens):
        n = int(rng.integers(life[0], life[1] + 1))
        arr = t0 + np.arange(n) / fps + rng.uniform(-jitter, jitter, n)
        arr[0] = max(arr[0], t0)
        arr = np.maximum.accumulate(arr)
        out.append(Request(i, t0, n, i % n_scenes, int(rng.choice(vols)), arr))
    return out


def pingpong(j, n):
    """Frame j of an endless forward/backward walk over a pool of n frames."""
    if n == 1:
        return 0
    p = j % (2 * n - 2)
    return p if p < n else 2 * n - 2 - p


class BatchedBackend:
    """Sessions in one vs_server slot table; a step is one vs_server_step."""

    def __init__(self, ctx, max_slots, hbm_budget, tsdf=vscan.VS_TSDF_F16, weight=vscan.VS_WEIGHT_U8):
        self.srv = vscan.Server(ctx, max_slots, tsdf=tsdf, weight=weight, hbm_budget=hbm_budget)

    def open(self, desc, T_gc0):
        return self.srv.open(desc, T_gc0=T_gc0)

    def close_slot(self, slot):
        self.srv.close_slot(slot)

    def session(self, slot):
        return self.srv.session(slot)

    def step(self, slots, ptrs, stream):
        self.srv.step(slots, ptrs, is_host=False, stream=stream.h)

    def usage(self):
        return self.srv.usage()

    def close(self):
        self.srv.close()


class SerialBackend:
    """The plain pipeline: each session its own vs_session, each frame its own
    vs_fuse_frame, one after another; the same slot count and HBM budget."""

    def __init__(self, ctx, max_slots, hbm_budget):
        self.ctx, self.max_slots, self.budget = ctx, max_slots, hbm_budget
        self.sess, self.bytes = {}, 0

    def open(self, desc, T_gc0):
        need = vscan.session_bytes(desc)
        if len(self.sess) >= self.max_slots or (self.budget and self.bytes + need > self.budget):
            return None
        slot = next(i for i in range(self.max_slots) if i not in self.sess)
        self.sess[slot] = vscan.Session(self.ctx, desc, T_gc0=T_gc0)
        self.bytes += need
        return slot

    def close_slot(self, slot):
        s = self.sess.pop(slot)
        self.bytes -= vscan.session_bytes(s.desc)
        s.close()

    def session(self, slot):
        return self.sess[slot]

    def step(self, slots, ptrs, stream):
        L = vscan.lib()
        for slot, p in zip(slots, ptrs):
            vscan.check(L.vs_fuse_frame(self.sess[slot].h, vscan.P(p), 0, stream.h), "vs_fuse_frame")

    def usage(self):
        return len(self.sess), self.bytes

    def close(self):
        for slot in list(self.sess):
            self.close_slot(slot)


class ServeLoop:
    """Runs requests against a backend (BatchedBackend or SerialBackend).
    `frame_ptr(scene, k)` gives the device address of frame k of a scene's
    pool (uint16, W x H), `pool_len` its frame count, `pose0(scene)` the pose
    of frame 0, `desc_for(vol_n)` the session descriptor. `max_batch` caps the
    frames per step (1 = one frame at a time, oldest arrival first)."""

    def __init__(self, backend, desc_for, frame_ptr, pool_len, pose0, max_inflight=2, max_batch=None,
                 admit_per_pass=2):
        self.backend = backend
        # sessions admitted per pass of the loop: opening one takes ~1-3 ms of
        # host time (allocation, volume clear), so admitting a whole burst
        # before stepping would hold back the frames of the open sessions
        self.admit_per_pass = admit_per_pass
        self.desc_for, self.frame_ptr, self.pool_len, self.pose0 = desc_for, frame_ptr, pool_len, pose0
        self.max_inflight = max_inflight
        self.max_batch = max_batch
        self.stream = hiprt.Stream()

    def close(self):
        self.backend.close()
        self.stream.destroy()

    def run(self, requests, duration, log=None):
        srv, stream = self.backend, self.stream
        waiting = collections.deque(sorted(requests, key=lambda r: r.t_open))
        live = {}          # slot -> [request, next frame to submit, frames completed]
        inflight = collections.deque()  # (event, [(slot, arrival)], t_submit)
        lat, sizes, refused, opened, closed = [], [], 0, 0, 0
        admit_wait, t_open_call, t_close_call, t_step_call = [], [], [], []
        peak_open, peak_bytes, lost = 0, 0, 0
        ev0 = hiprt.Event()
        ev0.record(stream)
        ev0.sync()
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0  # noqa: E731
        last_log = 0.0

        def retire(block):
            while inflight:
                ev, members, _ = inflight[0]
                if not block and not ev.query():
                    return
                ev.sync()
                done = ev0.elapsed_ms(ev) * 1e-3
                for slot, arr in members:
                    lat.append(done - arr)
                    live[slot][2] += 1
                inflight.popleft()
                ev.destroy()
                block = False

        while True:
            now = clock()
            if log and now - last_log > 5.0:
                last_log = now
                log(f"  serve t={now:.1f}s open={len(live)} fused={len(lat)}")
            retire(False)
            # sessions whose every frame has completed leave the table
            for slot in [s for s, v in live.items() if v[2] == v[0].n_frames]:
                st = srv.session(slot).status()
                lost += st[0] != 0
                tc = time.perf_counter()
                srv.close_slot(slot)
                t_close_call.append(time.perf_counter() - tc)
                del live[slot]
                closed += 1
            # admissions (in arrival order; a refused session waits), a few
            # per pass so the open sessions' frames keep being stepped
            admitted_now = 0
            while waiting and waiting[0].t_open <= now and now < duration and admitted_now < self.admit_per_pass:
                r = waiting[0]
                tc = time.perf_counter()
                slot = srv.open(self.desc_for(r.vol_n), T_gc0=self.pose0(r.scene))
                t_open_call.append(time.perf_counter() - tc)
                if slot is None:
                    refused += 1
                    break
                waiting.popleft()
                # the client starts streaming once its session is admitted
                wait = max(clock() - r.t_open, 0.0)
                admit_wait.append(wait)
                r.arrivals = r.arrivals + wait
                live[slot] = [r, 0, 0]
                opened += 1
                admitted_now += 1
            n_open, used = srv.usage()
            peak_open, peak_bytes = max(peak_open, n_open), max(peak_bytes, used)
            now = clock()
            if now >= duration:
                # stop submitting; drain what is queued
                retire(True)
                break
            ready = [(slot, v) for slot, v in live.items() if v[1] < v[0].n_frames and v[0].arrivals[v[1]] <= now]
            if self.max_batch and len(ready) > self.max_batch:
                ready = sorted(ready, key=lambda sv: sv[1][0].arrivals[sv[1][1]])[:self.max_batch]
            if ready and len(inflight) < self.max_inflight:
                slots = [s for s, _ in ready]
                ptrs = [self.frame_ptr(v[0].scene, pingpong(v[1], self.pool_len)) for _, v in ready]
                members = [(s, float(v[0].arrivals[v[1]])) for s, v in ready]
                for _, v in ready:
                    v[1] += 1
                tc = time.perf_counter()
                srv.step(slots, ptrs, stream)
                t_step_call.append(time.perf_counter() - tc)
                ev = hiprt.Event()
                ev.record(stream)
                inflight.append((ev, members, now))
                sizes.append(len(slots))
                continue
            if len(inflight) >= self.max_inflight:
                retire(True)
                continue
            # wait for the next frame or session arrival, or the oldest step
            nxt = [v[0].arrivals[v[1]] for v in live.values() if v[1] < v[0].n_frames]
            if waiting:
                nxt.append(waiting[0].t_open)
            t_next = min(nxt + [duration])
            while clock() < t_next and not (inflight and inflight[0][0].query()):
                time.sleep(2e-4)
        wall = clock()
        ev0.destroy()
        n_fused = len(lat)
        lat = np.array(lat) if lat else np.zeros(1)

        def dist_ms(xs):
            a = np.array(xs) * 1e3 if xs else np.zeros(1)
            return {"p50": round(float(np.percentile(a, 50)), 3), "p99": round(float(np.percentile(a, 99)), 3),
                    "max": round(float(a.max()), 3), "n": len(xs)}

        return {
            "duration_s": round(wall, 3),
            "fused_frames": n_fused,
            "fused_frames_per_s": round(n_fused / wall, 2),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                           "p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
                           "max": round(float(lat.max()) * 1e3, 3),
                           "frames_over_33ms": int(np.sum(lat > 0.033))},
            "admission_wait_ms": dist_ms(admit_wait),
            "host_call_ms": {"open": dist_ms(t_open_call), "close": dist_ms(t_close_call),
                             "step": dist_ms(t_step_call)},
            "sessions_opened": opened, "sessions_closed": closed, "open_at_end": len(live),
            "admissions_refused": refused, "peak_open": peak_open, "peak_session_bytes": int(peak_bytes),
            "sessions_lost_tracking": int(lost),
            "mean_step_frames": round(float(np.mean(sizes)) if sizes else 0.0, 2),
        }
