"""bench.py's host logic, on the CPU: the one-GPU contract, the float time
loop it counts substeps with, the committed traffic lookup, kernel labels
and the CPU-baseline leg."""
import os
import subprocess
import sys
from types import SimpleNamespace

import numpy as np

import mercury_amd as mh
import oracle
from conftest import REPO

sys.path.insert(0, REPO)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=REPO)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_substep_count_follows_the_float_time_loop():
    # Simulator::Simulate's `remaining -= dt` recurrence, as the oracle runs it
    for duration, dt in ((1.0, 1e-4), (0.1, 1e-4), (1.0, 1e-3), (0.05, 3e-4)):
        assert bench.float_loop_steps(duration, dt) == len(oracle.step_times(0.0, dt, duration))
    assert bench.float_loop_steps(1.0, 1e-4) == 10000


def test_traffic_lookup_reads_the_committed_counter_summary():
    for workload, kernel in (("voxel", "mass"), ("nn", "stream (chunked)"), ("nn", "mass-stream")):
        t = bench.measured_traffic(workload, kernel, 1000)
        assert t is not None and t["bytes_per_launch"] > 0
        assert t["source"].startswith("profiles/traffic.json: ")
    # streaming kernels read their records every substep; the mass-parallel
    # kernel reads its state once per launch
    assert bench.measured_traffic("nn", "mass-stream", 2000)["bytes_per_launch"] == \
        2 * bench.measured_traffic("nn", "mass-stream", 1000)["bytes_per_launch"]
    assert bench.measured_traffic("voxel", "mass", 2000)["bytes_per_launch"] == \
        bench.measured_traffic("voxel", "mass", 1000)["bytes_per_launch"]
    assert bench.measured_traffic("voxel", "no-such-kernel", 1000) is None


def test_kernel_labels():
    def label(kernel, chunk=0):
        return bench.kernel_label(SimpleNamespace(kernel=kernel, stream_chunk=chunk))
    assert label(mh.MH_KERNEL_MASS) == "mass"
    assert label(mh.MH_KERNEL_MASS_GATHER | mh.MH_KERNEL_STREAM) == "mass-gather+stream"
    assert label(mh.MH_KERNEL_STREAM, chunk=5120) == "stream (chunked)"
    assert label(mh.MH_KERNEL_LDS | mh.MH_KERNEL_STREAM, chunk=512) == "lds+stream (chunked)"
    assert label(mh.MH_KERNEL_MASS_STREAM) == "mass-stream"
    # every --kernel choice names a family the library knows, plus a
    # streaming one
    for name, bits in bench.KERNEL_FAMILIES.items():
        assert bits == 0 or bits & (mh.MH_KERNEL_STREAM | mh.MH_KERNEL_MASS_STREAM)


def test_cpu_baseline_leg_runs_the_mirror():
    ms = mh.workload_materials()
    els = mh.voxel_population(ms, 2, voxels_per_axis=3, seed=1, mass=0.001)
    rate, secs = bench.cpu_baseline(els, mh.MH_SOLVER_FORCE, 1e-4, 40, 2, bench.LAND)
    assert rate > 0 and secs > 0
    rate64, _ = bench.cpu_baseline(els, mh.MH_SOLVER_XPBD_REF, 1e-3, 20, 1, {}, f64=True)
    assert rate64 > 0
    assert np.isfinite(rate) and np.isfinite(rate64)


def test_roofline_head_takes_the_larger_fraction():
    valu = {"achieved": 5.9e11, "peak": 6.144e11, "unit": "wave64 instr/s", "frac": 5.9e11 / 6.144e11}
    head = bench.roofline_head("lds", 3.3e4, 1.4e5, "GB/s", {"valu_issue": valu})
    assert head["bound"] == "valu" and head["frac"] == valu["frac"]
    assert head["byte_model_bound"] == "lds" and abs(head["byte_model_frac"] - 3.3e4 / 1.4e5) < 1e-12
    low = dict(valu, achieved=4.0e11, frac=4.0e11 / 6.144e11)
    head = bench.roofline_head("hbm", 6.5e3, 8.0e3, "GB/s", {"valu_issue": low})
    assert head["bound"] == "hbm" and abs(head["frac"] - 6.5e3 / 8.0e3) < 1e-12
    assert head["valu_frac"] == low["frac"]
    head = bench.roofline_head("hbm", 6.5e3, 8.0e3, "GB/s", {})
    assert head["bound"] == "hbm" and "valu_frac" not in head
