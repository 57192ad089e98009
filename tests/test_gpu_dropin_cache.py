"""MoL_step_forward_in_time keeps a solver per griddata (device tables built once). These tests
change what that solver was built from and check that the next step follows the change, as the
reference's MoL step (which reads auxevol and rfm on every call, MoL_step_forward_in_time.c:26-31)
does: new punctures written into the same auxevol buffer, by the library's own
auxevol_gfs_all_points (main.c:97-105) or by any other writer, and a closed Problem replaced by
another one on the same grid whose buffers may land at the same addresses. Bar: the oracle's
RK4 step on the same inputs, 1e-12 relative (identical inputs) or 1e-10 (auxevol computed on
each side)."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

import oracle
from nrpyelliptic_amd import config, runtime

pytestmark = pytest.mark.gpu

SHAPE = (16, 16, 6)
CFG_A = config.shipped_free_parameters(SHAPE)
# same grid (bScale, AMAX, SINHWAA), other punctures
CFG_B = dataclasses.replace(CFG_A, bare_mass=(0.3, 0.6), P0=(0.0, 0.2, 0.05), P1=(0.0, -0.2, -0.05),
                            S0=(0.1, 0.0, 0.2), S1=(0.0, 0.0, -0.15), name="other_punctures")


def rel_maxnorm(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def use_oracle_inputs(P, O):
    for orc_name, ref_name in zip(oracle.ORC_RFM, oracle.abi.RFM_FIELDS):
        runtime.upload_ptr(getattr(P.gd.rfmstruct, ref_name), O.rfm_arrays[orc_name])
    P.set("auxevol_gfs", O.aux)


def step_both(P, O, y0, nsteps):
    P.set("y_n_gfs", y0)
    st = O.new_state()
    st["y_n"][:] = y0
    for _ in range(nsteps):
        P.mol_step(O.dt)
        O.mol_step(st)
    return P.get("y_n_gfs"), st["y_n"]


@pytest.mark.parametrize("writer", ["auxevol_gfs_all_points", "upload"])
def test_auxevol_rewritten_in_place(writer):
    OA, OB = oracle.OracleProblem(CFG_A), oracle.OracleProblem(CFG_B)
    assert OA.dt == OB.dt
    P = runtime.Problem(CFG_A)
    try:
        use_oracle_inputs(P, OA)
        got, want = step_both(P, OA, np.zeros(2 * P.ntot), 3)  # builds the cached solver
        assert rel_maxnorm(got, want) <= 1e-12
        y0 = got.copy()
        if writer == "upload":
            P.set("auxevol_gfs", OB.aux)
            tol = 1e-12
        else:
            CFG_B.fill_params(P.gd.params)
            runtime.lib().auxevol_gfs_all_points(C.byref(P.gd.params), P.gd.xx, P.gd.gridfuncs.auxevol_gfs)
            tol = 1e-10
        got, want = step_both(P, OB, y0, 2)
        assert rel_maxnorm(got, want) <= tol
        # and the tables of the first punctures would have been visibly wrong
        stale = OA.new_state()
        stale["y_n"][:] = y0
        OA.mol_step(stale)
        OA.mol_step(stale)
        assert rel_maxnorm(stale["y_n"], want) > 1e3 * tol
    finally:
        P.close()


def test_problem_replaced_on_the_same_grid():
    P = runtime.Problem(CFG_A)
    OA = oracle.OracleProblem(CFG_A)
    use_oracle_inputs(P, OA)
    step_both(P, OA, np.zeros(2 * P.ntot), 2)
    P.close()
    OB = oracle.OracleProblem(CFG_B)
    for k in range(3):  # reopen a few times: allocations are likely to reuse addresses
        Q = runtime.Problem(CFG_B)
        try:
            use_oracle_inputs(Q, OB)
            got, want = step_both(Q, OB, np.zeros(2 * Q.ntot), 3)
            assert rel_maxnorm(got, want) <= 1e-12, k
        finally:
            Q.close()


def test_steps_from_two_threads():
    """MoL_step_forward_in_time from two host threads at once, on two griddata of different
    grids (each call blocks until its step is done; the solver cache is shared): each state
    equals the same steps run from one thread."""
    import threading
    cfgs = [config.shipped_free_parameters((16, 16, 6)), config.shipped_free_parameters((32, 16, 8))]
    probs = [runtime.Problem(c) for c in cfgs]
    try:
        y0 = [np.random.default_rng(20 + k).uniform(-1e-2, 1e-2, 2 * P.ntot) for k, P in enumerate(probs)]
        want = []
        for P, y in zip(probs, y0):
            P.set("y_n_gfs", y)
            for _ in range(12):
                P.mol_step()
            want.append(P.get("y_n_gfs"))
        for P, y in zip(probs, y0):
            P.set("y_n_gfs", y)
        errors = []

        def run(P):
            try:
                for _ in range(12):
                    P.mol_step()
            except Exception as e:  # reported below
                errors.append(e)

        threads = [threading.Thread(target=run, args=(P,)) for P in probs]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=60)
        assert not errors and not any(t.is_alive() for t in threads)
        for P, w in zip(probs, want):
            assert np.array_equal(P.get("y_n_gfs"), w)
    finally:
        for P in probs:
            P.close()
