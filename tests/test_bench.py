"""CPU tests of bench.py's host logic: the one-GPU contract, the rocprof kernel names of the
workloads' kernels and the roofline arithmetic (no device is touched)."""
import os
import subprocess
import sys

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("mixed,learned,expect", [
    (True, True, ["k_face_rmhd_weno_rows<2, 0>", "k_face_rmhd_weno_rows<2, 1>", "k_face_rmhd_weno_lanes<2>",
                  "k_update<9, float, true>"]),
    (False, False, ["k_face_rmhd<2, 2, 0, double>", "k_face_rmhd<2, 2, 1, double>", "k_face_rmhd<2, 2, 2, double>",
                    "k_update<9, double, false>"]),
])
def test_kernel_names_of_the_rmhd_workload(mixed, learned, expect):
    s = bench.load_workload("rmhd_jet_256")
    got = [bench.kernel_name(s, w, mixed, learned) for w in ("face_x", "face_y", "face_z", "stage_update_c2p")]
    # the FP32 WENO5 faces are the shared-edge kernels (rows for x/y, lanes for z)
    assert got == expect


def test_kernel_names_of_the_srhd_workload():
    s = bench.load_workload("srhd_jet_256")
    assert bench.kernel_name(s, "face_y", True, False) == "k_face_srhd<0, 1, 1, float>"
    assert bench.kernel_name(s, "stage_update_c2p", True, False) == "k_update<5, float, false>"


def test_roofline_picks_the_slowest_kernel_and_counts_its_bytes():
    """RMHD WENO5 at 256^3, mixed: x faces move 9 x 8 B of primitives, two 9 x 4 B fluxes and two
    4 B carried speeds per face over 257 x 256 x 256 faces; the update 4 x 9 x 8 B of state, 8 B
    of psi and four 9 x 4 B flux reads per cell over 262^3 cells."""
    s = bench.load_workload("rmhd_jet_256")
    kern = {"launches_face": [3, 3, 3], "ms_face": [3.0, 2.4, 2.1], "launches_update": 3, "ms_update": 6.3}
    r = bench.kernel_roofline(s, kern, True, True)
    assert r["kernel"] == "stage_update_c2p" and r["avg_ms"] == pytest.approx(2.1)
    assert r["bytes_per_launch"] == (4 * 9 * 8 + 8 + 4 * 9 * 4) * 262 ** 3
    assert r["frac"] == pytest.approx(r["achieved"] / bench.HBM_PEAK_GBS, rel=1e-4)
    kern["ms_face"][0] = 9.0  # x faces slowest now
    r = bench.kernel_roofline(s, kern, True, True)
    assert r["kernel"] == "face_x"
    assert r["bytes_per_launch"] == (9 * 8 + 2 * 9 * 4 + 2 * 4) * 257 * 256 * 256
