"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: a
driver runs every stage (the FP64 and FP32 pipelines over six frames, and
each stand-alone operator) on a seeded synthetic stream; any out-of-bounds
access, use after free, leak or undefined behaviour aborts it. Runs on the
CPU (no GPU marker)."""
import os
import subprocess

import oracle_lib as orc


def test_oracle_stages_clean_under_asan_and_ubsan():
    odir = os.path.join(orc.ROOT, "oracle")
    subprocess.run(["make", "-s", "-C", odir, "sanitize"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(odir, "_var", "oracle_sanitize")], capture_output=True, text=True, env=env,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-3000:]
    assert "sanitized oracle run complete" in r.stdout
