"""The library's host-side compute_TRC / set_optimum_eta_damping (the Einstein Toolkit driver's
eta_damping = -1 policy, conformally_flat_BBH_set_optimum_eta_damping.c:3-27) against the
oracle's restatement, which tests/test_oracle_adm.py pins. Host code only: no GPU needed."""
import ctypes as C
import math

import pytest

import oracle
from nrpyelliptic_amd import abi, config, runtime


def test_compute_TRC_matches_oracle():
    L = runtime.lib()
    O = oracle.lib()
    for A, b, C0, dt, N1, N2 in ((1e6, 5.0, 0.7, 1.7e-3, 128, 16), (1e6, 5.0, 0.5, 3e-3, 64, 4),
                                 (1e4, 3.0, 0.7, 2.2e-3, 96, 8)):
        d1, d2 = math.pi / N1, 2 * math.pi / N2
        assert L.compute_TRC(A, b, C0, dt, d1, d2) == O.orc_compute_TRC(A, b, C0, dt, d1, d2)


def test_set_optimum_eta_damping_reads_grid_parameters():
    L = runtime.lib()
    p = abi.paramstruct()
    L.set_Cparameters_to_default(C.byref(p))
    cfg = config.equal_mass_boosted()
    cfg.fill_params(p)
    p.dxx1, p.dxx2 = math.pi / 128, 2 * math.pi / 16
    L.set_optimum_eta_damping(1.7e-3, 0.7, C.byref(p))
    want = oracle.lib().orc_optimum_eta(p.AMAX, p.bScale, p.dxx1, p.dxx2, 1.7e-3, 0.7)
    assert p.eta_damping == want and 0 < want < 100


def test_optimum_eta_policy_requires_shipped_grid():
    """The thorn computes the optimum only for domain_size 1e6, sinh_width 0.07, foci 5."""
    class Fake:
        pass
    P = Fake()
    P.gd = abi.griddata_struct()
    runtime.lib().set_Cparameters_to_default(C.byref(P.gd.params))
    P.gd.params.AMAX, P.gd.params.SINHWAA, P.gd.params.bScale = 1e6, 0.07, 4.0
    with pytest.raises(runtime.NrpyellError, match="foci_position"):
        runtime.Problem.set_optimum_eta_damping(P)
