"""nrpyell_L2_norm_of_gf_noalloc (the allocation-free L2 norm, every device buffer in one fixed
per-GPU scratch array) against the allocating L2_norm_of_gf on the same inputs, through the
C ABI. Two calls in a row, the second on a smaller grid, so the second call runs on scratch the
first one filled. Bar: the L2 values agree to 1e-13 relative (the volume weights come from the
device's exp/sin/cos instead of the host's, and the volume is summed in another order), and
both agree with the oracle's L2."""
import ctypes as C

import numpy as np
import pytest

import oracle
from nrpyelliptic_amd import abi, config, runtime

pytestmark = pytest.mark.gpu

GRIDS = [config.equal_mass_boosted((128, 128, 16)), config.unequal_mass_q4((32, 32, 8))]


def test_noalloc_l2_matches_allocating_l2_twice_in_a_row():
    for k, cfg in enumerate(GRIDS):
        print(f"noalloc L2: call {k + 1} on {tuple(cfg.Nxx)}", flush=True)
        P = runtime.Problem(cfg)
        try:
            O = oracle.OracleProblem(cfg)
            rng = np.random.default_rng(11 + k)
            for radius in (100.0, 1e300):
                gfs = rng.uniform(-1, 1, 2 * P.ntot)
                for gf_index in (0, 1):
                    want = P.l2(gfs, radius=radius, gf_index=gf_index)
                    got = P.l2_noalloc(gfs, radius=radius, gf_index=gf_index)
                    assert abs(10 ** got / 10 ** want - 1) <= 1e-13, (cfg.Nxx, radius, gf_index, got, want)
                    assert abs(got - O.l2(gfs[gf_index * P.ntot:(gf_index + 1) * P.ntot], radius)) <= 1e-13
        finally:
            P.close()


def test_noalloc_l2_refuses_a_grid_beyond_its_scratch():
    cfg = config.unequal_mass_q4((32, 32, 8))
    P = runtime.Problem(cfg)
    try:
        big = abi.paramstruct.from_buffer_copy(P.gd.params)
        big.Nxx0 = big.Nxx1 = 600
        big.Nxx_plus_2NGHOSTS0 = big.Nxx_plus_2NGHOSTS1 = 610
        out = C.c_double(123.0)
        d = runtime.DeviceArray(2 * P.ntot).upload(np.zeros(2 * P.ntot))
        rc = runtime.lib().nrpyell_L2_norm_of_gf_noalloc(C.byref(big), 0, 100.0, P.gd.xx, d.as_real_p(),
                                                          C.byref(out))
        assert rc == -1
        assert "scratch" in runtime.lib().nrpyell_last_error().decode()
        assert out.value == 123.0
        # a refused call leaves the entry point usable
        gf = np.random.default_rng(2).uniform(-1, 1, P.ntot)
        assert abs(10 ** P.l2_noalloc(gf) / 10 ** P.l2(gf) - 1) <= 1e-13
    finally:
        P.close()
