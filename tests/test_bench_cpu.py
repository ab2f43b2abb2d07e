"""CPU tests of bench.py's bookkeeping: the counter-traffic lookup keyed by
the kernel configuration, the kernel named in the roofline, the refusal of
more than one GPU, and the arithmetic of the committed traffic record."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

DEFAULT = {"kernel": 2, "cols_per_thread": 2, "strip": 64, "segment": 128, "prefetch_rows": 1,
           "steps_per_launch": 6, "col_block": 2}


def record():
    kv = {}
    for line in open(bench.TRAFFIC_FILE):
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    return kv


def test_traffic_record_is_the_scaled_counters():
    """traffic = fetch_scale x FETCH_SIZE + write_scale x WRITE_SIZE (KiB),
    and the scales are what the same counters read for the 1 GiB copy."""
    kv = record()
    fetch, write = float(kv["fetch_size_kib_avg"]), float(kv["write_size_kib_avg"])
    fs, ws = float(kv["fetch_scale"]), float(kv["write_scale"])
    assert fs == pytest.approx(1048576.0 / float(kv["copy_fetch_size_kib_for_1gib_read"]), rel=1e-4)
    assert ws == pytest.approx(1048576.0 / float(kv["copy_write_size_kib_for_1gib_write"]), rel=1e-4)
    assert float(kv["traffic_bytes_per_launch"]) == pytest.approx((fs * fetch + ws * write) * 1024, rel=1e-6)


def test_traffic_lookup_matches_only_its_configuration():
    kv = record()
    assert bench.measured_traffic(DEFAULT) == float(kv["traffic_bytes_per_launch"])
    for key, other in (("segment", 64), ("steps_per_launch", 4), ("col_block", 8), ("prefetch_rows", 2)):
        assert bench.measured_traffic({**DEFAULT, key: other}) is None


def test_profile_summary_names_the_recorded_kernel():
    kv = record()
    md = open(os.path.join(ROOT, "profiles", "fom_fusedW_config2_fp64.md")).read()
    assert kv["kernel"] in md
    assert f"{float(kv['fetch_size_kib_avg']):,.1f}" in md and f"{float(kv['write_size_kib_avg']):,.1f}" in md


@pytest.mark.parametrize("cfg,name", [
    (DEFAULT, "k_fusedW<double,2,6>"),
    ({**DEFAULT, "steps_per_launch": 4, "col_block": 8, "segment": 64}, "k_fusedN<double,2,4>"),
    ({**DEFAULT, "steps_per_launch": 3, "cols_per_thread": 1, "col_block": 1}, "k_fusedW<double,1,3>"),
    ({**DEFAULT, "steps_per_launch": 2}, "k_fused2<double,2>"),
    ({**DEFAULT, "steps_per_launch": 1, "prefetch_rows": 2}, "k_fused<double,2,2>"),
    ({**DEFAULT, "strip": 32, "steps_per_launch": 4}, "k_fusedN<double,2,4>"),
])
def test_kernel_name_follows_the_dispatch(cfg, name):
    """k_fusedW runs 3-6 steps when a strip is one wavefront (cols_per_thread
    = col_block <= 2, 64-point strips); k_fusedN other 3-4 step shapes."""
    assert bench.fused_kernel_name(cfg) == name


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
