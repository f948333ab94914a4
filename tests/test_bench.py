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


@pytest.mark.parametrize("mixed,learned,face_path,expect", [
    (True, True, 1, ["k_face_rmhd<2, 2, 0, float>", "k_face_rmhd<2, 2, 1, float>", "k_face_rmhd<2, 2, 2, float>",
                     "k_update<9, float, true>"]),
    (True, False, 2, ["k_face_rmhd_weno_rows<2, 0>", "k_face_rmhd_weno_rows<2, 1>", "k_face_rmhd_weno_lanes<2>",
                      "k_update<9, float, false>"]),
    (True, True, 3, ["k_face_rmhd_weno_march<2, 0>", "k_face_rmhd_weno_march<2, 1>", "k_face_rmhd_weno_lanes_div<2>",
                     "k_update<9, float, true>"]),
    (True, True, 4, ["k_face_rmhd_weno_march<2, 0>", "k_face_rmhd_weno_march<2, 1>", "k_face_rmhd_weno_lanes_div<2>",
                     "k_face_march_update<2, true>"]),
    (False, False, 3, ["k_face_rmhd<2, 2, 0, double>", "k_face_rmhd<2, 2, 1, double>", "k_face_rmhd<2, 2, 2, double>",
                       "k_update<9, double, false>"]),
])
def test_kernel_names_of_the_rmhd_workload(mixed, learned, face_path, expect):
    """The face-kernel variants (afd_params.face_path) of the mixed WENO5 faces; FP64 has one."""
    s = bench.load_workload("rmhd_jet_256")
    got = [bench.kernel_name(s, w, mixed, learned, face_path) for w in ("face_x", "face_y", "face_z", "stage_update_c2p")]
    assert got == expect


def test_face_variant():
    """face_path 0 takes the header's default; the variants apply only to mixed RMHD WENO5 below
    2^29 cells (the 1024^3 grid takes one thread per face); other workloads have one face kernel."""
    from astrofd_amd import _lib as L
    s = bench.load_workload("rmhd_jet_256")
    assert bench.face_variant(s, True, 0) == L.AFD_FACE_PATH_DEFAULT
    assert [bench.face_variant(s, True, v) for v in (1, 2, 3, 4)] == [1, 2, 3, 4]
    assert bench.face_variant(s, False, 3) == 0
    assert bench.face_variant(dict(s, NX=1024, NY=1024, NZ=1024), True, 3) == 1
    assert bench.face_variant(bench.load_workload("srhd_jet_256"), True, 3) == 0


def test_kernel_names_of_the_srhd_workload():
    s = bench.load_workload("srhd_jet_256")
    assert bench.kernel_name(s, "face_y", True, False) == "k_face_srhd<0, 1, 1, float>"
    assert bench.kernel_name(s, "stage_update_c2p", True, False) == "k_update<5, float, false>"


def test_algorithmic_update_bytes_are_the_metrics():
    """13 state arrays of nv x 8 B per zone-update for SSPRK3 (8 for SSPRK2), spread over the stages:
    936 / 3 = 312 B a cell per stage for RMHD, 1144 / 3 for the 512^3 jet with Te/Ti, 520 / 3 for
    SRHD, 320 / 2 for the SSPRK2 SRHD jet."""
    assert bench.algorithmic_update_bytes(bench.load_workload("rmhd_jet_256")) == 312
    assert bench.algorithmic_update_bytes(bench.load_workload("rmhd_jet_512")) == pytest.approx(1144 / 3)
    assert bench.algorithmic_update_bytes(bench.load_workload("srhd_jet_256")) == pytest.approx(520 / 3)
    assert bench.algorithmic_update_bytes(bench.load_workload("srhd_jet_32")) == 160


def test_impl_update_bytes():
    """RMHD SSPRK3, mixed, flux-divergence path: stage 0 reads P0, writes U1 and P1; stage 1 also
    reads U1; stage 2 reads U2 and writes no U; each reads psi and the 9 x 4 B divergence."""
    s = bench.load_workload("rmhd_jet_256")
    stage = [72 + 72 + 72 + 36 + 8, 72 + 72 + 72 + 72 + 36 + 8, 72 + 72 + 72 + 36 + 8]
    assert bench.impl_update_bytes(s, True, 3) == pytest.approx(sum(stage) / 3)
    # stored FP32 face fluxes: four 9 x 4 B flux reads instead of the divergence
    assert bench.impl_update_bytes(s, True, 2) == pytest.approx(sum(b - 36 + 4 * 36 for b in stage) / 3)
    # FP64: four 9 x 8 B flux reads
    stage = [b - 36 + 4 * 72 for b in stage]
    assert bench.impl_update_bytes(s, False) == pytest.approx(sum(stage) / 3)


def test_roofline_picks_the_slowest_kernel_and_counts_its_bytes():
    """RMHD WENO5 at 256^3, mixed: the update's bytes per launch are the metric's 312 B a cell per
    stage over 262^3 cells; the x pencils move 9 x 8 B of primitives, 2 x 4 B of carried speeds and
    read and write 9 x 4 B of divergence per cell over 256^3 cells."""
    s = bench.load_workload("rmhd_jet_256")
    kern = {"launches_face": [3, 3, 3], "ms_face": [3.0, 2.4, 2.1], "launches_update": 3, "ms_update": 6.3}
    r = bench.kernel_roofline(s, kern, True, True, face_path=3)
    assert r["kernel"] == "stage_update_c2p" and r["avg_ms"] == pytest.approx(2.1)
    assert r["bytes_per_launch"] == 312 * 262 ** 3
    assert r["impl_bytes_per_launch"] == int(bench.impl_update_bytes(s, True, 3) * 262 ** 3)
    assert r["frac"] == pytest.approx(r["achieved"] / bench.HBM_PEAK_GBS, rel=1e-4)
    assert r["achieved"] == pytest.approx(312 * 262 ** 3 / 2.1e-3 / 1e9, rel=1e-4)
    kern["ms_face"][0] = 9.0  # x faces slowest now
    r = bench.kernel_roofline(s, kern, True, True, face_path=3)
    assert r["kernel"] == "face_x"
    assert r["bytes_per_launch"] == (9 * 8 + 2 * 4 + 2 * 9 * 4) * 256 ** 3
    r = bench.kernel_roofline(s, kern, False, False)  # FP64: per face, two fluxes and the speeds
    assert r["bytes_per_launch"] == (9 * 8 + 2 * 9 * 8 + 2 * 8) * 257 * 256 * 256


@pytest.mark.parametrize("workload", ["rmhd_jet_256", "rmhd_jet_512", "srhd_jet_256"])
def test_traffic_names_its_source(workload):
    """Every workload the bench line reports has a committed traffic profile holding the stage
    update of both runs (mixed with its default start, FP64 plain) under the names the roofline
    looks up, and the roofline names the file it read."""
    s = bench.load_workload(workload)
    rm = s["PHYSICS"] == "rmhd"
    kern = {"launches_face": [3, 3, 3], "ms_face": [3.0, 2.4, 2.1], "launches_update": 3, "ms_update": 6.3}
    table, _ = bench.measured_traffic(workload)
    for mixed in (True, False):
        learned = mixed and rm
        r = bench.kernel_roofline(s, kern, mixed, learned, workload)
        assert r["kernel"] == "stage_update_c2p"
        assert r["rocprof_name"] in table
        assert r["traffic"] == round(table[r["rocprof_name"]]) and r["traffic"] > 0
        assert r["traffic_source"].startswith(f"profiles/{workload}_traffic.txt")
