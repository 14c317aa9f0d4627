"""The pins catch the mistakes they exist for: each pin of
tests/test_oracle_pins_geometry.py is run against an oracle build with one
planted mistake (oracle/Makefile target `variants`, macro OR_MUTATION) and
must fail there, while passing on the real oracle.

Also: the march's iteration cap (decision Q13: the forward-only step bound
+ 64) decides no ray of the test scenes: an oracle built with a cap 16x
longer renders identical maps, on FP32 volumes and on FP16-rounded ones
(where a fine-search ray can keep stepping back behind a -1.0 voxel).

Runs on the CPU (no GPU marker).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as orc
import test_oracle_pins_geometry as pins
import vscan

VAR = os.path.join(orc.ROOT, "oracle", "_var")


@pytest.fixture(scope="module")
def variants():
    subprocess.run(["make", "-s", "-j5", "-C", os.path.join(orc.ROOT, "oracle"), "variants"], check=True)
    return VAR


@pytest.mark.parametrize("mutant,pin", [
    (1, pins.pin_a13_single_ray_brick_sets),  # samples n = -1..1
    (1, pins.pin_a13_plane_brick_set),
    (2, pins.pin_a13_single_ray_brick_sets),  # ROUNDF(v)/8 without the precedence quirk
    (3, pins.pin_a17_plane_normals),          # one gradient tap one voxel off
    (3, pins.pin_a17_sphere_normals),
    (4, pins.pin_a18_shade),                  # shading factor 0.75
])
def test_pin_fails_on_its_mutant(variants, mutant, pin):
    pin()  # the real oracle passes
    with orc.using(os.path.join(variants, f"liboracle_m{mutant}.so")):
        with pytest.raises(AssertionError):
            pin()


def _volume(cfg, w, frames, W=None, H=None):
    st = vscan.SceneStream(cfg, seed=1, width=W or 0, height=H or 0)
    d = vscan.make_desc(w, K=st.K, W=W, H=H)
    vol = orc.volume(d)
    N = d.vol_dim[0] * d.vol_dim[1] * d.vol_dim[2]
    F = np.ones(N)
    Wt = np.zeros(N)
    K = orc.K_levels(d)
    for fr in frames:
        T = st.pose(fr)
        orc.integrate(st.frame(fr), K[0], np.float32(T), np.float32(orc.rigid_inverse(T)), vol, F, Wt)
    return st, d, F, Wt


def _maps(d, F, Wt, T):
    K = orc.K_levels(d)
    Tg, Tc = np.float32(T), np.float32(orc.rigid_inverse(T))
    out = []
    for l in range(3):
        depth, vtx, nrm, img, stats = orc.raycast(F, Wt, orc.volume(d), d.width >> l, d.height >> l, K[l], Tg, Tc)
        out.append((depth, vtx, nrm, img, stats[:3]))
    return out


@pytest.mark.parametrize("case", ["plain_320x240_128", "primitives_320x240_128", "fp16_640x480_512"])
def test_march_cap_decides_no_ray(variants, case):
    if case == "plain_320x240_128":
        st, d, F, Wt = _volume(0, vscan.WORKLOADS[case], range(0, 12, 3))
        Fv = F.astype(np.float32)
    elif case == "primitives_320x240_128":
        st, d, F, Wt = _volume(4, vscan.WORKLOADS["plain_320x240_128"], range(0, 12, 3))
        Fv = F.astype(np.float32)
    else:
        st, d, F, Wt = _volume(2, vscan.WORKLOADS[case], range(0, 12, 3))
        Fv = F.astype(np.float16).astype(np.float32)  # the FP16 storage's values
    Wv = Wt.astype(np.float32)
    T = st.pose(13)
    base = _maps(d, Fv, Wv, T)
    with orc.using(os.path.join(variants, "liboracle_cap16.so")):
        long_cap = _maps(d, Fv, Wv, T)
    for l, (a, b) in enumerate(zip(base, long_cap)):
        hits = int((a[1][..., 0] != -100).sum())
        assert hits > 0.1 * a[0].size, (case, l, hits)
        for x, y in zip(a[:4], b[:4]):
            assert np.array_equal(x, y), f"{case} level {l}: the cap changed a ray"
