"""The two-substep kernel's block placement (nrpyell_block_order_plan, host code, no GPU): a
permutation of the launch's blocks that keeps the BC blocks first and every interior block on
the XCD the kernel's own order gives it (workgroup L runs on XCD L % 8; that order gives XCD x
the interior blocks x*per .. x*per + per - 1, whole tile columns), never puts two BC blocks on
one CU (an XCD's k-th workgroup runs on its CU k % P, P = CUs per XCD) nor more than three
blocks on a CU, and at 128x128x16 on 256 CUs gives the split measured fastest: BC block, first-
column tile, middle tile (in that dispatch order) on the CUs holding three blocks, last-column
tile and middle tile on those holding two. Launches that do not fit one round of resident
workgroups keep the kernel's own order (empty plan)."""
import numpy as np
import pytest

from nrpyelliptic_amd import runtime

BC_ROWS = 16  # rows per BC block of the two-substep kernel


def counts(Nxx, tile_i0, planes):
    ty = 256 // tile_i0
    nb = -(-Nxx[1] // BC_ROWS) * Nxx[2]
    ntile0, ntile1, ng = -(-Nxx[0] // tile_i0), -(-Nxx[1] // ty), Nxx[2] // planes
    return nb, ntile0, ntile1, ng


def per_cu(o, nb, ncu, kind):
    P = ncu // 8
    out = []
    for x in range(8):
        cus = [""] * P
        for j, b in enumerate(o[x::8]):
            cus[j % P] += kind(int(b))
        out.append(cus)
    return out


@pytest.mark.parametrize("ncu,Nxx,tile", [
    (256, (128, 128, 16), (32, 2)),
    (256, (160, 128, 16), (32, 2)),
    (256, (96, 96, 16), (32, 2)),
    (256, (64, 64, 16), (32, 2)),
    (256, (128, 128, 8), (32, 1)),
    (256, (64, 128, 16), (64, 2)),
    (304, (128, 128, 16), (32, 2)),
    (256, (128, 100, 16), (32, 2)),   # ragged last tile row
])
def test_plan_is_a_placement(ncu, Nxx, tile):
    o = runtime.block_order_plan(ncu, Nxx, *tile)
    nb, ntile0, ntile1, ng = counts(Nxx, *tile)
    n = nb + ntile0 * ntile1 * ng
    assert len(o) == n, "expected a plan for a launch of one round"
    assert np.array_equal(np.sort(o), np.arange(n))
    assert np.array_equal(o[:nb], np.arange(nb))
    ni, L = n - nb, np.arange(nb, n)
    assert np.array_equal((o[L] - nb) // (ni // 8), L % 8)
    cus = per_cu(o, nb, ncu, lambda b: "B" if b < nb else "I")
    for x in range(8):
        for c in cus[x]:
            assert c.count("B") <= 1 and len(c) <= 3, (x, c)


def test_bench_grid_split():
    Nxx, tile = (128, 128, 16), (32, 2)
    o = runtime.block_order_plan(256, Nxx, *tile)
    nb, ntile0, _, ng = counts(Nxx, *tile)

    def kind(b):
        if b < nb:
            return "B"
        t0 = ((b - nb) // ng) % ntile0
        return "F" if t0 == 0 else "L" if t0 == ntile0 - 1 else "M"

    for x, cus in enumerate(per_cu(o, nb, 256, kind)):
        assert cus == ["BFM"] * 16 + ["LM"] * 16, (x, cus)


@pytest.mark.parametrize("ncu,Nxx,tile", [
    (256, (128, 128, 16), (16, 1)),   # 1152 blocks: more than one round
    (256, (256, 256, 32), (32, 2)),   # larger than a round
    (64, (128, 128, 16), (32, 2)),    # too few CUs
    (250, (128, 128, 16), (32, 2)),   # CUs not a multiple of 8 XCDs
    (256, (96, 40, 6), (32, 2)),      # 45 interior and 18 BC blocks: not multiples of 8
])
def test_plan_keeps_kernel_order_when_it_cannot_place(ncu, Nxx, tile):
    assert len(runtime.block_order_plan(ncu, Nxx, *tile)) == 0


@pytest.mark.parametrize("args", [(0, (128, 128, 16), 32, 2), (256, (128, 128, 16), 48, 2),
                                  (256, (128, 128, 16), 16, 2), (256, (128, 128, 15), 32, 2)])
def test_plan_rejects_invalid_arguments(args):
    with pytest.raises(runtime.NrpyellError):
        runtime.block_order_plan(*args)
