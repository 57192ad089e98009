"""CPU checks of the C-ABI library (no GPU needed): it loads, exports every symbol the
public header declares, and its host-side set-up functions reproduce the reference's
(set_Cparameters_to_default, set_Nxx_dxx_invdx_params__and__xx, bcstruct_set_up,
find_timestep) byte for byte. Also the ghost-map check that guards the fused kernels.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from nrpyelliptic_amd import abi, config, runtime

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HEADER = os.path.join(ROOT, "include", "nrpyelliptic_amd.h")
HAVE_REF = oracle.ref_lib() is not None


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?[\s\*]+([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, re.M)
    return sorted(set(n for n in names if n not in ("if", "defined")))


def test_library_exports_every_declared_symbol():
    L = runtime.lib()
    names = header_functions()
    assert "rhs_eval" in names and "MoL_step_forward_in_time" in names and "nrpyell_solver_solve" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_surface_matches_prototypes():
    """Every reference-driver function of the hot path is exported under its own name."""
    L = runtime.lib()
    for n in ("rhs_eval", "MoL_step_forward_in_time", "apply_bcs_outerradiation_and_inner", "apply_bcs_inner_only",
              "residual_all_points", "L2_norm_of_gf", "bcstruct_set_up", "set_Nxx_dxx_invdx_params__and__xx",
              "find_timestep", "MoL_malloc_y_n_gfs", "MoL_malloc_non_y_n_gfs", "rfm_precompute_rfmstruct_define",
              "auxevol_gfs_all_points", "wavespeed_gf_all_points", "compute_wavespeed_at_OB",
              "initial_guess_all_points", "gridfunction_z_axis", "gridfunction_xz_plane"):
        assert hasattr(L, n), n


def test_struct_layouts():
    # member order is the ABI (NRPy_basic_defines.h)
    assert C.sizeof(abi.paramstruct) == 8 * 71 + 4 * 8  # 71 REAL members, then 8 ints
    assert C.sizeof(abi.innerpt_bc_struct) == 20 and C.sizeof(abi.outerpt_bc_struct) == 10
    assert abi.paramstruct.Nxx0.offset == 8 * 71
    assert C.sizeof(abi.griddata_struct) == C.sizeof(abi.paramstruct) + 3 * 8 + C.sizeof(abi.bc_struct) + \
        C.sizeof(abi.rfm_struct) + C.sizeof(abi.MoL_gridfunctions_struct)


@pytest.mark.skipif(not HAVE_REF, reason="compiled reference not built here")
def test_defaults_and_grid_match_reference_bytes():
    L, R = runtime.lib(), oracle.ref_lib()
    p1, p2 = abi.paramstruct(), abi.paramstruct()
    L.set_Cparameters_to_default(C.byref(p1))
    R.set_Cparameters_to_default(C.byref(p2))
    assert bytes(p1) == bytes(p2)
    for eigen in (0, 1):
        for shape in [(16, 16, 6), (128, 128, 16)]:
            Nxx = (C.c_int * 3)(*shape)
            xx1, xx2 = abi.XX3(), abi.XX3()
            L.set_Nxx_dxx_invdx_params__and__xx(eigen, C.byref(Nxx), C.byref(p1), xx1)
            R.set_Nxx_dxx_invdx_params__and__xx(eigen, C.byref(Nxx), C.byref(p2), xx2)
            assert bytes(p1) == bytes(p2)
            for d in range(3):
                n = shape[d] + 10
                assert np.array_equal(np.ctypeslib.as_array(xx1[d], (n,)), np.ctypeslib.as_array(xx2[d], (n,)))


def _host_tables(cfg):
    L = runtime.lib()
    gd = abi.griddata_struct()
    p = gd.params
    L.set_Cparameters_to_default(C.byref(p))
    cfg.fill_params(p)
    Nxx = (C.c_int * 3)(*cfg.Nxx)
    L.set_Nxx_dxx_invdx_params__and__xx(1, C.byref(Nxx), C.byref(p), gd.xx)
    runtime.check(L.nrpyell_bcstruct_set_up_checked(C.byref(p), gd.xx, C.byref(gd.bcstruct)), "bcstruct")
    L.set_Nxx_dxx_invdx_params__and__xx(0, C.byref(Nxx), C.byref(p), gd.xx)
    dt = L.find_timestep(C.byref(p), gd.xx, cfg.CFL)
    return gd, dt


@pytest.mark.parametrize("shape", [(16, 16, 6), (64, 64, 4), (128, 128, 16), (160, 96, 8), (96, 128, 16)])
def test_bcstruct_and_timestep_match_oracle(shape):
    cfg = config.shipped_free_parameters(shape)
    gd, dt = _host_tables(cfg)
    O = oracle.OracleProblem(cfg)
    bc = gd.bcstruct
    n = bc.bc_info.num_inner_boundary_points
    raw = C.string_at(C.cast(bc.inner_bc_array, C.c_void_p), n * 20)
    rec = np.frombuffer(raw, dtype=np.dtype([("dst", "<i4"), ("src", "<i4"), ("par", "i1", 10), ("pad", "V2")]))
    dst, src, par = O.inner_arrays()
    assert np.array_equal(rec["dst"], dst) and np.array_equal(rec["src"], src) and np.array_equal(rec["par"], par)
    outer = []
    for gz in range(5):
        for dirn in range(3):
            m = bc.bc_info.num_pure_outer_boundary_points[gz][dirn]
            arr = bc.pure_outer_bc_array[dirn + 3 * gz]
            outer.append([(arr[k].i0, arr[k].i1, arr[k].i2, arr[k].FACEX0, arr[k].FACEX1, arr[k].FACEX2)
                          for k in range(m)])
    assert outer == O.outer_lists()
    assert dt == O.dt
    assert runtime.lib().nrpyell_ghostmap_check(C.byref(gd.params), C.byref(gd.bcstruct)) == 0


@pytest.mark.skipif(not HAVE_REF, reason="compiled reference not built here")
@pytest.mark.parametrize("amax", [1.0e6, 1.0e4])
def test_bcstruct_loop_bounds_match_reference(amax):
    cfg = config.shipped_free_parameters((32, 32, 8))
    cfg.AMAX = amax
    gd, _ = _host_tables(cfg)
    R = oracle.RefProblem(cfg)
    assert bytes(gd.bcstruct.bc_info) == bytes(R.gd.bcstruct.bc_info)


def _inner_bytes(gd):
    bc = gd.bcstruct
    n = bc.bc_info.num_inner_boundary_points
    return C.string_at(C.cast(bc.inner_bc_array, C.c_void_p), n * C.sizeof(abi.innerpt_bc_struct))


def test_bcstruct_cache_hands_out_equal_independent_copies():
    """A grid seen before gets the same table bytes from the per-grid cache, in fresh
    allocations the caller owns; another AMAX (which changes the parities) is built anew."""
    cfg = config.shipped_free_parameters((24, 16, 8))
    a, _ = _host_tables(cfg)
    b, _ = _host_tables(cfg)
    assert _inner_bytes(a) == _inner_bytes(b) and bytes(a.bcstruct.bc_info) == bytes(b.bcstruct.bc_info)
    assert C.addressof(a.bcstruct.inner_bc_array.contents) != C.addressof(b.bcstruct.inner_bc_array.contents)
    b.bcstruct.inner_bc_array[3].srcpt += 1  # a caller's change stays in its own copy
    c, _ = _host_tables(cfg)
    assert _inner_bytes(c) == _inner_bytes(a)
    cfg2 = config.shipped_free_parameters((24, 16, 8))
    cfg2.AMAX = 1.0e4
    d, _ = _host_tables(cfg2)
    O = oracle.OracleProblem(cfg2)
    dst, src, par = O.inner_arrays()
    raw = np.frombuffer(_inner_bytes(d), dtype=np.dtype([("dst", "<i4"), ("src", "<i4"), ("par", "i1", 10),
                                                           ("pad", "V2")]))
    assert np.array_equal(raw["dst"], dst) and np.array_equal(raw["src"], src) and np.array_equal(raw["par"], par)
    for gd in (a, b, c, d):
        runtime.lib().nrpyell_bcstruct_free(C.byref(gd.bcstruct))


def test_ghostmap_check_rejects_a_corrupted_table():
    """A planted mistake in the table (one source index off by one) must be refused, so the
    fused kernels never run on a grid whose ghost map they do not reproduce."""
    cfg = config.shipped_free_parameters((16, 16, 6))
    gd, _ = _host_tables(cfg)
    L = runtime.lib()
    assert L.nrpyell_ghostmap_check(C.byref(gd.params), C.byref(gd.bcstruct)) == 0
    gd.bcstruct.inner_bc_array[100].srcpt += 1
    assert L.nrpyell_ghostmap_check(C.byref(gd.params), C.byref(gd.bcstruct)) == -3
    assert b"maps to" in L.nrpyell_last_error()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(runtime, "_LIB", None)
    monkeypatch.setattr(runtime, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(runtime.NrpyellError):
        runtime.lib()


_NO_DEVICE_PROBE = r"""
import ctypes as C, sys
L = C.CDLL(sys.argv[1])
L.nrpyell_last_error.restype = C.c_char_p
p = C.c_void_p()
if L.nrpyell_malloc_device(C.byref(p), C.c_size_t(64)) == 0:
    print("device-present"); sys.exit(0)
h = C.c_void_p()
rc = L.nrpyell_mlp_create(C.c_ulonglong(1), C.byref(h))
print("mlp_create", rc, L.nrpyell_last_error().decode())
sys.stdout.flush()
if sys.argv[2] == "reference":
    # a void reference-surface entry point: prints and exits(1) like the reference
    sys.path.insert(0, sys.argv[3])
    from nrpyelliptic_amd import abi
    prm = abi.paramstruct()
    L.set_Cparameters_to_default(C.byref(prm))
    L.rhs_eval(C.byref(prm), None, None, None, None)
    print("rhs_eval returned")
"""


def _run_no_device(mode):
    import subprocess
    import sys
    r = subprocess.run([sys.executable, "-c", _NO_DEVICE_PROBE, runtime.LIB_PATH, mode, ROOT], capture_output=True,
                       text=True, timeout=300)
    if "device-present" in r.stdout:
        pytest.skip("a GPU is visible: the HIP failure cannot be forced from the host")
    return r


def test_hip_failures_return_ehip_from_the_extension_api():
    """A failing HIP call inside an nrpyell_* entry point (here: no device, so hipMalloc fails)
    comes back as NRPYELL_EHIP with the message, and the process goes on."""
    r = _run_no_device("api")
    assert r.returncode == 0, r.stderr
    assert "mlp_create -2 hipMalloc" in r.stdout and "failed" in r.stdout


def test_hip_failures_exit_from_the_reference_surface():
    """The reference's void functions keep the reference's error convention: print, exit(1)."""
    r = _run_no_device("reference")
    assert "mlp_create -2" in r.stdout
    assert r.returncode == 1 and "rhs_eval returned" not in r.stdout
    assert "libnrpyell_hip: rhs_eval failed" in r.stderr
