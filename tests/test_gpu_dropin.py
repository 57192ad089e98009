"""Drop-in test: the reference's own CLI driver (NRPyElliptic_Ccodes/main.c, unmodified),
linked once against the reference objects (CPU) and once against libnrpyell_hip.so (GPU),
run with the same arguments; its output files must agree. Both binaries are built from the
reference sources by oracle/Makefile (target `drivers`) where /root/reference exists and
travel with the tree; without them the test is skipped."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
REF = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref")
MAIN_REF, MAIN_HIP = os.path.join(REF, "main_ref"), os.path.join(REF, "main_hip")


@pytest.mark.skipif(not (os.path.exists(MAIN_REF) and os.path.exists(MAIN_HIP)),
                    reason="reference driver binaries not built in this tree")
def test_reference_driver_runs_on_the_gpu_library(tmp_path):
    args = ["32", "32", "8", "200"]
    outs = {}
    for name, exe in (("cpu", MAIN_REF), ("gpu", MAIN_HIP)):
        d = tmp_path / name
        d.mkdir()
        r = subprocess.run([os.path.abspath(exe), *args], cwd=d, capture_output=True, text=True, timeout=300,
                           env=dict(os.environ, OMP_NUM_THREADS="4"))
        assert r.returncode == 0, r.stderr[-2000:]
        outs[name] = d / "output_NRPyElliptic"
    # residual history (main.c:153-155): step, time, log10 L2 at every step
    a = np.loadtxt(outs["cpu"] / "residual_l2_norm.txt")
    b = np.loadtxt(outs["gpu"] / "residual_l2_norm.txt")
    assert a.shape == b.shape == (201, 3)
    assert np.array_equal(a[:, 0], b[:, 0])
    assert np.max(np.abs(a[:, 2] - b[:, 2])) <= 1e-9
    # final u along the z axis, and the set-up fields (gridfunction_z_axis text format)
    for f in ("uu_final_1D.txt", "wavespeed_gf_1D.txt", "psi_background_gf_1D.txt", "ADD_times_AUU_gf_1D.txt"):
        x, y = np.loadtxt(outs["cpu"] / f), np.loadtxt(outs["gpu"] / f)
        assert np.array_equal(x[:, 0], y[:, 0]), f  # coordinates printed identically
        assert np.max(np.abs(x[:, 1] - y[:, 1])) <= 1e-12 * max(1.0, np.max(np.abs(x[:, 1]))), f
    # a 2-D dump (gridfunction_xz_plane) of u at the last output step
    last = sorted(os.listdir(outs["cpu"] / "2D"))[-1]
    x, y = np.loadtxt(outs["cpu"] / "2D" / last), np.loadtxt(outs["gpu"] / "2D" / last)
    assert np.array_equal(x[:, :2], y[:, :2])
    assert np.max(np.abs(x[:, 2] - y[:, 2])) <= 1e-11
