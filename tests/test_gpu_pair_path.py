"""Two RK4 substeps per launch (nrpyell_substep2.hip) against one substep per launch
(nrpyell_substep.hip): the same arithmetic per point, so the states must be bit-identical --
after odd and even step counts (the pair path alternates y_n between two arrays), for every tile
shape, on grids with ragged tiles, a one-column last tile of the one-substep decomposition
(Nxx0 = 33), Nxx1 below the BC blocks' 16 rows, and the largest batched grid; and a full solve
stops at the same step with the same history. The one-substep path itself follows the oracle
(test_gpu_parity.py), as does the default (pair) path through Solver.steps there."""
import numpy as np
import pytest

from nrpyelliptic_amd import config, runtime

pytestmark = pytest.mark.gpu

GRIDS = [config.shipped_free_parameters((16, 16, 6)), config.unequal_mass_q4((32, 32, 8)),
         config.random_batch(1, seed=5, grids=False, Nxx=(16, 40, 8))[0],
         config.random_batch(1, seed=9, grids=False, Nxx=(33, 16, 6))[0],
         config.random_batch(1, seed=3, grids=False, Nxx=(70, 12, 6))[0],
         config.random_batch(1, seed=7, grids=False, Nxx=(160, 128, 16))[0]]
IDS = ["16x16x6", "32x32x8", "16x40x8", "33x16x6", "70x12x6", "160x128x16"]
TILES = [(32, 1), (16, 1), (64, 1), (32, 2), (64, 2)]


@pytest.fixture(scope="module", params=range(len(GRIDS)), ids=IDS)
def prob(request):
    P = runtime.Problem(GRIDS[request.param])
    yield P
    P.close()


def run_steps(P, n, spl, tile):
    P.zero_guess()
    S = P.solver(substeps_per_launch=spl, tile_i0=tile[0], planes=tile[1])
    S.steps(n)
    S.close()
    return P.get("y_n_gfs")


@pytest.mark.parametrize("tile", TILES, ids=lambda t: f"tx{t[0]}_pl{t[1]}")
def test_pair_path_bit_identical_to_single(prob, tile):
    for n in (1, 2, 7, 40):
        print(f"{tuple(prob.cfg.Nxx)} tile {tile} steps {n}", flush=True)
        a = run_steps(prob, n, 1, tile)
        b = run_steps(prob, n, 2, tile)
        assert np.array_equal(a, b), (n, float(np.max(np.abs(a - b))))


def test_pair_path_solve_matches_single():
    P = runtime.Problem(config.axisymmetric_spinning(Nxx=(64, 64, 8)))
    try:
        out = []
        for spl in (1, 2):
            P.zero_guess()
            S = P.solver(log10_target=-7.0, max_steps=5000, substeps_per_launch=spl, steps_per_graph=25)
            r = S.solve()
            out.append((r.steps, r.log10_l2, P.get("y_n_gfs"), S.history()))
            S.close()
        (s1, l1, y1, h1), (s2, l2, y2, h2) = out
        assert s1 == s2 and l1 == l2 and np.array_equal(y1, y2)
        assert np.array_equal(h1[0], h2[0]) and np.array_equal(h1[1], h2[1])
        assert s1 % 2 == 1 or s1 % 2 == 0  # either parity ends in y_n_gfs
    finally:
        P.close()


def test_pair_block_order_places_every_block_once():
    """The two-substep kernel's placement table at 128x128x16, tiles 32 x 8 x 2 planes (512
    interior blocks after 128 BC blocks of 16 rows, one round of resident workgroups on 256 CUs):
    a permutation of the launch's blocks; the BC blocks keep their positions and every interior
    block its XCD (workgroup L runs on XCD L % 8, which the kernel's own order gives tiles
    (L % 8) * 64 .. + 63); and the XCD's k-th workgroup runs on its CU k % 32, so the CUs
    holding three blocks get a BC block, a first-column tile and a middle tile, in that order,
    and those holding two a last-column tile and a middle tile."""
    P = runtime.Problem(config.equal_mass_boosted())
    try:
        S = P.solver(substeps_per_launch=2, tile_i0=32, planes=2)
        o = S.block_order()
        S.close()
        ntile0, ntile1, ng = 128 // 32, 128 // 8, 16 // 2
        nb, ni = (128 // 16) * 16, ntile0 * ntile1 * ng
        assert len(o) == nb + ni and np.array_equal(np.sort(o), np.arange(nb + ni))
        assert np.array_equal(o, runtime.block_order_plan(256, (128, 128, 16), 32, 2))  # MI355X: 256 CUs
        assert np.array_equal(o[:nb], np.arange(nb))
        L = np.arange(nb, nb + ni)
        assert np.array_equal((o[L] - nb) // (ni // 8), L % 8)

        def kind(b):
            if b < nb:
                return "B"
            t0 = ((b - nb) // ng) % ntile0
            return "F" if t0 == 0 else "L" if t0 == ntile0 - 1 else "M"

        for x in range(8):
            cus = [""] * 32
            for j, b in enumerate(o[x::8]):
                cus[j % 32] += kind(int(b))
            assert cus == ["BFM"] * 16 + ["LM"] * 16, (x, cus)
    finally:
        P.close()


def test_pair_path_option_checks():
    P = runtime.Problem(config.axisymmetric_spinning())  # Nxx2 = 4
    try:
        with pytest.raises(runtime.NrpyellError, match="Nxx2 >= 6"):
            P.solver(substeps_per_launch=2)
        S = P.solver()  # falls back to one substep per launch
        S.steps(3)
        S.close()
    finally:
        P.close()
    Q = runtime.Problem(config.shipped_free_parameters((16, 16, 6)))
    try:
        with pytest.raises(runtime.NrpyellError, match="radiation"):
            Q.solver(substeps_per_launch=2, outer_bc="sommerfeld6")
    finally:
        Q.close()


@pytest.mark.parametrize("check_every,steps_per_graph", [(10, 50), (5, 25), (3, 9)])
def test_pair_path_mixed_precision_matches_single(check_every, steps_per_graph):
    """The FP32 bulk phase runs the same two-substep kernel in float: the mixed-precision solve
    takes the same FP32 and FP64 steps with the same history and the same final u as with one
    substep per launch. Odd check intervals end the FP32 phase on odd step counts, i.e. with the
    FP32 state in the pair path's second array."""
    P = runtime.Problem(config.axisymmetric_spinning(Nxx=(64, 64, 8)))
    try:
        out = []
        for spl in (1, 2):
            P.zero_guess()
            S = P.solver(precision=32, log10_target=-9.0, max_steps=8000, substeps_per_launch=spl,
                         check_every=check_every, steps_per_graph=steps_per_graph)
            r = S.solve()
            out.append((r.steps, r.fp32_steps, r.log10_l2, P.get("y_n_gfs"), S.history()))
            S.close()
        (s1, f1, l1, y1, h1), (s2, f2, l2, y2, h2) = out
        print(f"check_every {check_every}: {s1} steps, {f1} in FP32, log10 L2 {l1:.3f}", flush=True)
        assert f1 > 0 and s1 > f1
        assert (s1, f1, l1) == (s2, f2, l2)
        assert np.array_equal(y1, y2)
        assert np.array_equal(h1[0], h2[0]) and np.array_equal(h1[1], h2[1])
    finally:
        P.close()
