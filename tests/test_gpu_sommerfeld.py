"""GPU parity of the Einstein Toolkit thorn's Sommerfeld outer BC against the oracle
restatement (itself pinned in test_oracle_sommerfeld.py): the drop-in entry point
nrpyell_apply_bcs_sommerfeld on every test grid, the fused solver with outer_bc="sommerfeld6"
along a 200-step trajectory for each tile shape, the FP32 phase and the batched solver; and
the solution property that the outer BC changes how the relaxation gets there, not where
it ends: converged u under both BCs agree.

Tolerances as for the radiation BC: one BC application <= 1e-13 relative max-norm, 200
relaxation steps <= 1e-10 max-abs.
"""
import numpy as np
import pytest

import oracle
from nrpyelliptic_amd import config, runtime
from test_gpu_parity import IDS, SMALL, interior_mask, rel_maxnorm, use_oracle_inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=range(len(SMALL)), ids=IDS)
def pair(request):
    cfg = SMALL[request.param]
    P = runtime.Problem(cfg)
    O = oracle.OracleProblem(cfg, outer_bc="sommerfeld6")
    yield P, O
    P.close()


@pytest.mark.parametrize("variant", ["thorn", "finf_nofalloff"])
def test_dropin_bc_matches_oracle(pair, variant):
    P, O = pair
    print(f"sommerfeld drop-in: grid {P.Ng} variant {variant}", flush=True)
    rng = np.random.default_rng(11)
    st = rng.uniform(-0.1, 0.1, 2 * P.ntot)
    rhs = rng.uniform(-1, 1, 2 * P.ntot)
    if variant == "thorn":
        kw = dict(speed=(runtime.ET_SOMMERFELD_SPEED,) * 2, f_infinity=(0.0, 0.0), falloff=(3.0, 3.0))
    else:
        kw = dict(speed=(2.0e6, 3.0e6), f_infinity=(0.25, -0.5), falloff=(0.0, 2.0))
    want = rhs.copy()
    O.apply_sommerfeld(st, want, speed=kw["speed"], finf=kw["f_infinity"], falloff=kw["falloff"])
    got = P.apply_sommerfeld(st, rhs, **kw)
    assert rel_maxnorm(got, want) <= 1e-13
    # it wrote the ghost points and left the interior alone
    inner = np.tile(interior_mask(P.Ng), 2)
    assert np.array_equal(got[inner], rhs[inner]) and not np.array_equal(got, rhs)


@pytest.mark.parametrize("tile", [(32, 1), (16, 1), (64, 2), (32, 2)], ids=lambda t: f"tx{t[0]}_pl{t[1]}")
def test_solver_trajectory_matches_oracle(pair, tile):
    P, O = pair
    print(f"sommerfeld solver trajectory: grid {P.Ng} tile {tile}", flush=True)
    use_oracle_inputs(P, O)
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    S = P.solver(check_every=10, steps_per_graph=50, tile_i0=tile[0], planes=tile[1], outer_bc="sommerfeld6")
    S.steps(200)
    got = P.get("y_n_gfs")
    st = O.new_state()
    for _ in range(200):
        O.mol_step(st)
    assert np.max(np.abs(got - st["y_n"])) <= 1e-10
    S.close()


def test_outer_bc_option_changes_the_trajectory():
    """The option reaches the fused kernel: 100 steps with each BC give different states."""
    cfg = SMALL[1]
    P = runtime.Problem(cfg)
    outs = []
    for bc in ("radiation", "sommerfeld6"):
        P.set("y_n_gfs", np.zeros(2 * P.ntot))
        S = P.solver(check_every=10, steps_per_graph=50, outer_bc=bc)
        S.steps(100)
        outs.append(P.get("y_n_gfs"))
        S.close()
    assert not np.array_equal(outs[0], outs[1])
    with pytest.raises(runtime.NrpyellError):
        P.solver(outer_bc="neumann")
    with pytest.raises(runtime.NrpyellError, match="EINVAL"):
        P.solver(outer_bc=7)
    P.close()


@pytest.fixture(scope="module")
def converged():
    """q = 4 configuration on a 96x64x8 grid relaxed to log10 L2 <= -9 under each BC."""
    cfg = config.unequal_mass_q4(Nxx=(96, 64, 8))
    P = runtime.Problem(cfg)
    out = {}
    for bc in ("radiation", "sommerfeld6"):
        P.set("y_n_gfs", np.zeros(2 * P.ntot))
        S = P.solver(log10_target=-9.0, max_steps=20000, outer_bc=bc)
        r = S.solve()
        print(f"converged solve outer_bc={bc}: steps {r.steps} log10 {r.log10_l2:.2f}", flush=True)
        out[bc] = (r, P.get("y_n_gfs")[: P.ntot].copy())
        S.close()
    yield P, out
    P.close()


def test_both_bcs_converge_to_the_same_solution(converged):
    P, out = converged
    (ra, ua), (rs, us) = out["radiation"], out["sommerfeld6"]
    assert ra.converged and rs.converged
    m = np.zeros(P.Ng[::-1], dtype=bool)
    m[5:-5, 5:-5, 5:-5] = True
    m = m.ravel()
    # L2 <= 1e-9 bounds u to ~1e-7 of its scale; the two runs take different paths there
    assert np.max(np.abs(ua[m] - us[m])) <= 1e-6 * np.max(np.abs(ua[m]))


def test_fp32_phase_and_batch_with_the_thorn_bc(converged):
    """Mixed precision and the batched solver accept the option and reach the same u."""
    P, out = converged
    _, us = out["sommerfeld6"]
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    S = P.solver(log10_target=-9.0, max_steps=20000, precision=32, outer_bc="sommerfeld6")
    r = S.solve()
    print(f"fp32+fp64 sommerfeld: steps {r.steps} fp32 {r.fp32_steps}", flush=True)
    assert r.converged and r.fp32_steps > 0
    u32 = P.get("y_n_gfs")[: P.ntot]
    S.close()
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    B = runtime.BatchSolver([P], log10_target=-9.0, max_steps=20000, outer_bc="sommerfeld6")
    res, _ = B.solve()
    ub = P.get("y_n_gfs")[: P.ntot]
    B.close()
    assert res[0].converged and res[0].steps == out["sommerfeld6"][0].steps
    assert np.array_equal(ub, us)
    scale = np.max(np.abs(us))
    assert np.max(np.abs(u32 - us)) <= 1e-6 * scale


def test_bc_geometry_follows_the_grid_values():
    """The outer-BC point lists are rebuilt when the geometry changes behind the same xx
    pointers (AMAX changed in place): both drop-in BCs then match the oracle on the new grid."""
    cfg = SMALL[0]
    P = runtime.Problem(cfg)
    O = oracle.OracleProblem(cfg, outer_bc="sommerfeld6")
    rng = np.random.default_rng(5)
    st = rng.uniform(-0.1, 0.1, 2 * P.ntot)
    rhs = rng.uniform(-1, 1, 2 * P.ntot)
    c = P.gd.params.uu_wavespeed_at_OB
    for amax in (P.gd.params.AMAX, 0.5 * P.gd.params.AMAX):
        print(f"bc geometry with AMAX {amax}", flush=True)
        P.gd.params.AMAX = amax
        O.g.AMAX = amax
        O.g.cOB[0] = O.g.cOB[1] = c
        want = rhs.copy()
        O.apply_radiation_and_inner(st, want)
        assert rel_maxnorm(P.apply_radiation_and_inner(st, rhs), want) <= 1e-13
        want = rhs.copy()
        O.apply_sommerfeld(st, want)
        got = P.apply_sommerfeld(st, rhs, speed=(c, c))
        assert rel_maxnorm(got, want) <= 1e-13
    P.close()
