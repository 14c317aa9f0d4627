"""Host logic of the serving request loop (no GPU): the request trace and the
frame-pool walk."""
import numpy as np

import serving


def test_requests_trace():
    reqs = serving.make_requests(7, 60.0, 64, rate=2.0, initial=16)
    assert sum(1 for r in reqs if r.t_open == 0.0) >= 16
    later = np.array([r.t_open for r in reqs if r.t_open > 0])
    assert np.all(np.diff(later) > 0) and later.max() < 60.0
    # Poisson count over 60 s at 2/s: mean 120, sd ~11
    assert 80 < len(later) < 160
    for r in reqs:
        assert 150 <= r.n_frames <= 900 and r.vol_n in (128, 256, 384, 512)
        a = r.arrivals
        assert len(a) == r.n_frames and np.all(np.diff(a) >= 0) and a[0] >= r.t_open
        nominal = r.t_open + np.arange(r.n_frames) / 30.0
        assert np.all(np.abs(a[1:] - nominal[1:]) <= 0.005 + 1e-9 + 1.0 / 30.0)
    again = serving.make_requests(7, 60.0, 64, rate=2.0, initial=16)
    assert [(r.t_open, r.n_frames, r.vol_n) for r in again] == [(r.t_open, r.n_frames, r.vol_n) for r in reqs]


def test_pingpong_walk_is_continuous():
    n = 5
    walk = [serving.pingpong(j, n) for j in range(40)]
    assert walk[:10] == [0, 1, 2, 3, 4, 3, 2, 1, 0, 1]
    assert all(abs(a - b) == 1 for a, b in zip(walk, walk[1:]))
    assert all(0 <= k < n for k in walk)
    assert [serving.pingpong(j, 1) for j in range(3)] == [0, 0, 0]
