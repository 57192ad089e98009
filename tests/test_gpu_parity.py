"""GPU parity of the drop-in entry points of libnrpyell_hip.so against the oracle
(oracle/liboracle.so, itself pinned to the compiled reference in test_oracle.py).

Tolerances: a single RHS evaluation agrees to <= 1e-12 relative max-norm
(rhs_eval and residual_all_points are in fact bit-exact: same FMA tree, IEEE division);
bcstruct tables bit-exact; converged u to <= 1e-10 max-abs.
"""
import numpy as np
import pytest

import oracle
from nrpyelliptic_amd import config, runtime

pytestmark = pytest.mark.gpu

# Square and non-square grids (Nxx0 != Nxx1 catches any i0/i1 mix-up), ragged and full tiles,
# generic spins and momenta.
SMALL = [config.shipped_free_parameters((16, 16, 6)), config.unequal_mass_q4((32, 32, 8)),
         config.random_batch(1, seed=42, grids=False, Nxx=(24, 16, 6))[0],
         config.random_batch(1, seed=5, grids=False, Nxx=(16, 40, 8))[0],
         config.random_batch(1, seed=9, grids=False, Nxx=(33, 16, 6))[0],
         config.random_batch(1, seed=42, grids=False, Nxx=(96, 128, 16))[0],
         config.random_batch(1, seed=7, grids=False, Nxx=(160, 128, 16))[0],
         config.random_batch(1, seed=13, grids=False, Nxx=(70, 18, 8))[0]]
# 33: the tile holding the last interior i0 holds only that column (the one-sided BC stencil
# reaches into the previous tile); 160x128x16: the largest grid of the batched workloads;
# 70x18x8: Nxx1 divides neither the drop-in kernels' 4-row tiles, the fused kernels' 8-row tiles
# nor the 16-row BC blocks
IDS = ["shipped_16x16x6", "q4_32x32x8", "rand_24x16x6", "rand_16x40x8", "rand_33x16x6", "rand_96x128x16",
       "max_160x128x16", "rand_70x18x8"]


def interior_mask(Ng):
    m = np.zeros(Ng[::-1], dtype=bool)
    m[5:-5, 5:-5, 5:-5] = True
    return m.ravel()


def rel_maxnorm(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module", params=range(len(SMALL)), ids=IDS)
def pair(request):
    cfg = SMALL[request.param]
    P = runtime.Problem(cfg)
    O = oracle.OracleProblem(cfg)
    yield P, O
    P.close()


def use_oracle_inputs(P, O):
    """Load the oracle's rfm factors and auxevol into the device griddata so both sides see
    identical inputs."""
    for orc_name, ref_name in zip(oracle.ORC_RFM, oracle.abi.RFM_FIELDS):
        runtime.upload_ptr(getattr(P.gd.rfmstruct, ref_name), O.rfm_arrays[orc_name])
    P.set("auxevol_gfs", O.aux)


def test_setup_matches_oracle(pair):
    P, O = pair
    assert P.dt == O.dt
    for orc_name, ref_name in zip(oracle.ORC_RFM, oracle.abi.RFM_FIELDS):
        got, want = P.rfm(ref_name), O.rfm_arrays[orc_name]
        assert rel_maxnorm(got, want) <= 1e-14, ref_name
    aux, nt = P.get("auxevol_gfs"), P.ntot
    assert rel_maxnorm(aux[:nt], O.aux[:nt]) <= 1e-13  # AA^2, scaled by its max
    assert rel_maxnorm(aux[nt:2 * nt], O.aux[nt:2 * nt]) <= 1e-14  # psi_bg
    m = interior_mask(P.Ng)
    assert rel_maxnorm(aux[2 * nt:][m], O.aux[2 * nt:][m]) <= 1e-14  # wavespeed (interior)
    assert P.gd.params.uu_wavespeed_at_OB == pytest.approx(O.g.cOB[0], rel=1e-14)


def test_bcstruct_bit_exact(pair):
    P, O = pair
    rec, outer = P.bc_tables()
    dst, src, par = O.inner_arrays()
    assert np.array_equal(rec["dst"], dst) and np.array_equal(rec["src"], src) and np.array_equal(rec["par"], par)
    assert outer == O.outer_lists()


def test_rhs_eval_bit_exact(pair):
    P, O = pair
    use_oracle_inputs(P, O)
    st = np.random.default_rng(1234).uniform(-0.1, 0.1, 2 * P.ntot)
    got, want = P.rhs(st), O.rhs(st)
    m = np.concatenate([interior_mask(P.Ng)] * 2)
    assert np.array_equal(got[m], want[m]), f"max diff {np.max(np.abs(got[m]-want[m]))}"


def test_residual_bit_exact(pair):
    P, O = pair
    use_oracle_inputs(P, O)
    st = np.random.default_rng(99).uniform(-0.1, 0.1, 2 * P.ntot)
    got, want = P.residual(st), O.residual(st)
    m = interior_mask(P.Ng)
    assert np.array_equal(got[m], want[m])


def test_rhs_and_residual_bit_exact_on_a_large_ragged_grid():
    """rhs_eval and residual_all_points on 200x130x16 (2 x 64 + 72 columns, 32 x 4 + 2 rows of the
    drop-in kernels' tiles): bit-exact against the oracle at every interior point."""
    cfg = config.random_batch(1, seed=17, grids=False, Nxx=(200, 130, 16))[0]
    P, O = runtime.Problem(cfg), oracle.OracleProblem(cfg)
    try:
        use_oracle_inputs(P, O)
        st = np.random.default_rng(7).uniform(-0.1, 0.1, 2 * P.ntot)
        m = interior_mask(P.Ng)
        got, want = P.rhs(st), O.rhs(st)
        assert np.array_equal(got[np.concatenate([m, m])], want[np.concatenate([m, m])])
        got, want = P.residual(st), O.residual(st)
        assert np.array_equal(got[m], want[m])
    finally:
        P.close()


def test_radiation_and_inner_bc(pair):
    P, O = pair
    use_oracle_inputs(P, O)
    rng = np.random.default_rng(7)
    st = rng.uniform(-0.1, 0.1, 2 * P.ntot)
    rhs = rng.uniform(-1, 1, 2 * P.ntot)
    O.g.cOB[0] = O.g.cOB[1] = P.gd.params.uu_wavespeed_at_OB
    want = rhs.copy()
    O.apply_radiation_and_inner(st, want)
    got = P.apply_radiation_and_inner(st, rhs)
    assert rel_maxnorm(got, want) <= 1e-13


def test_inner_gather_exact(pair):
    P, O = pair
    g = np.random.default_rng(3).uniform(-1, 1, 2 * P.ntot)
    want = g.copy()
    O.apply_inner_only(want)
    assert np.array_equal(P.apply_inner_only(g), want)


def test_l2_norm(pair):
    P, O = pair
    gf = np.random.default_rng(5).uniform(-1, 1, P.ntot)
    assert abs(P.l2(gf) - O.l2(gf)) <= 1e-13


def test_l2_norm_back_to_back_calls(pair):
    """L2_norm_of_gf's result is written by the device into host memory and read after the
    stream synchronises: back-to-back calls on different device gridfunctions and radii (no
    upload, so no other synchronisation, in between) each return their own value."""
    import ctypes as C
    from nrpyelliptic_amd import runtime
    P, O = pair
    rng = np.random.default_rng(11)
    gfs = [rng.uniform(-1, 1, P.ntot) * (k + 1) for k in range(4)]
    dev = [runtime.DeviceArray(P.ntot).upload(gf) for gf in gfs]
    L = runtime.lib()
    try:
        for radius in (100.0, 3.0):
            got = [L.L2_norm_of_gf(C.byref(P.gd.params), 0, radius, P.gd.xx, d.as_real_p()) for d in dev]
            for gf, v in zip(gfs, got):
                assert abs(v - O.l2(gf, radius)) <= 1e-13 * max(1.0, abs(v)), (radius, v)
    finally:
        for d in dev:
            d.free()


@pytest.mark.parametrize("nsteps", [1, 20])
def test_mol_step_matches_oracle(pair, nsteps):
    """MoL_step_forward_in_time (the fused kernels: two launches of two substeps where the grid
    allows, else four) vs the oracle's reference-order RK4; y_n at every point including all ghost
    zones."""
    P, O = pair
    use_oracle_inputs(P, O)
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    st = O.new_state()
    for _ in range(nsteps):
        P.mol_step(O.dt)
        O.mol_step(st)
    got = P.get("y_n_gfs")
    assert rel_maxnorm(got, st["y_n"]) <= 1e-12


def test_mol_step_follows_dt_and_eta_changes(pair):
    """MoL_step_forward_in_time replays one captured graph per cached solver; a call with another
    dt, or after the damping in the params changed, must step with the new values (the graph is
    recaptured), as the reference's step reads them on every call."""
    P, O = pair
    use_oracle_inputs(P, O)
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    st = O.new_state()
    eta0 = P.gd.params.eta_damping
    try:
        for k, (dt, eta) in enumerate([(O.dt, eta0), (0.5 * O.dt, eta0), (0.5 * O.dt, eta0), (0.5 * O.dt, 2.0 * eta0),
                                       (O.dt, 2.0 * eta0)]):
            P.gd.params.eta_damping = eta
            O.g.eta = eta
            P.mol_step(dt)
            O.mol_step(st, dt=dt)
            assert rel_maxnorm(P.get("y_n_gfs"), st["y_n"]) <= 1e-12, k
    finally:
        P.gd.params.eta_damping = eta0
        O.g.eta = eta0


@pytest.mark.parametrize("tile", [(32, 1), (16, 1), (64, 2), (32, 2)], ids=lambda t: f"tx{t[0]}_pl{t[1]}")
def test_solver_matches_oracle_trajectory(pair, tile):
    """The graph-captured solver for 200 steps (every tile shape the kernel has, one and two
    i2 planes per thread) equals the oracle's reference-order RK4 at every point."""
    P, O = pair
    use_oracle_inputs(P, O)
    P.set("y_n_gfs", np.zeros(2 * P.ntot))
    S = P.solver(check_every=10, steps_per_graph=50, tile_i0=tile[0], planes=tile[1])
    S.steps(200)
    got = P.get("y_n_gfs")
    st = O.new_state()
    for _ in range(200):
        O.mol_step(st)
    assert np.max(np.abs(got - st["y_n"])) <= 1e-10
    l2 = S.log10_residual()
    assert abs(l2 - O.log10_residual(st["y_n"])) <= 1e-9
    S.close()
