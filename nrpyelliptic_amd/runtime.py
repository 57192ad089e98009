"""Python host layer over libnrpyell_hip.so.

`Problem` mirrors the reference driver's set-up sequence (NRPyElliptic_Ccodes/main.c:12-105)
call for call, through this library's C ABI, so the device holds the same griddata the
reference would build on the host.  `Solver` wraps the resident hipGraph solver.

The product path never falls back to CPU code: if the HIP library is missing or no GPU is
visible, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import abi
from .config import PunctureConfig

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("NRPYELL_LIB", os.path.join(HERE, "libnrpyell_hip.so"))

NRPYELL_OK = 0
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "EGHOSTMAP", -4: "ENOTCONVERGED", -5: "ENAN"}


class solve_options(C.Structure):
    _fields_ = [("precision", C.c_int), ("check_every", C.c_int), ("steps_per_graph", C.c_int),
                ("log10_target", C.c_double), ("max_steps", C.c_int), ("fp32_switch_log10", C.c_double),
                ("fp32_plateau_steps", C.c_int), ("integration_radius", C.c_double),
                ("tile_i0", C.c_int), ("planes", C.c_int), ("outer_bc", C.c_int), ("outer_bc_speed", C.c_double),
                ("stop_rule", C.c_int), ("substeps_per_launch", C.c_int)]


# stopping rules of the resident solver (nrpyell_solve_options.stop_rule)
STOP_RULE = {"every_step": 0, "et_thorn": 1}


# outer boundary conditions of the solvers (nrpyell_solve_options.outer_bc)
OUTER_BC = {"radiation": 0, "sommerfeld6": 1}
# the Einstein Toolkit thorn's Sommerfeld wave speed (NRPYELL_ET_SOMMERFELD_SPEED)
ET_SOMMERFELD_SPEED = 4.4793122676309114e+06


class admbase_options(C.Structure):
    """nrpyell_admbase_options: the thorn's Initialize_ADMBase parameters."""
    _fields_ = [("orbital_plane", C.c_int), ("position_shift", C.c_double * 3), ("initial_lapse", C.c_int),
                ("lapse_exponent_n", C.c_double), ("NRPy_epsilon", C.c_double), ("interpolation_order", C.c_int)]


ADMBASE_NAMES = ["uuGF", "alp", "betax", "betay", "betaz", "gxx", "gxy", "gxz", "gyy", "gyz", "gzz",
                 "kxx", "kxy", "kxz", "kyy", "kyz", "kzz"]


class solve_result(C.Structure):
    _fields_ = [("steps", C.c_int), ("fp32_steps", C.c_int), ("log10_l2", C.c_double),
                ("log10_l2_initial", C.c_double), ("seconds", C.c_double), ("converged", C.c_int),
                ("status", C.c_int)]


# nrpyell_solve_result.status
STATUS = {0: "max_steps", 1: "converged", 2: "nan", 3: "diverged"}


class NrpyellError(RuntimeError):
    pass


_LIB = None


def lib() -> C.CDLL:
    """Load libnrpyell_hip.so (built by nrpyelliptic_amd.build / __graft_entry__.build)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise NrpyellError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    L = abi.bind(C.CDLL(LIB_PATH))
    vp = C.c_void_p
    L.nrpyell_device_synchronize.restype = C.c_int
    L.nrpyell_malloc_device.argtypes = [C.POINTER(vp), C.c_size_t]
    L.nrpyell_free_device.argtypes = [vp]
    L.nrpyell_memcpy_h2d.argtypes = [vp, vp, C.c_size_t]
    L.nrpyell_memcpy_d2h.argtypes = [vp, vp, C.c_size_t]
    L.nrpyell_memset_device.argtypes = [vp, C.c_int, C.c_size_t]
    L.nrpyell_last_error.restype = C.c_char_p
    L.nrpyell_bcstruct_set_up_checked.argtypes = [C.POINTER(abi.paramstruct), abi.XX3, C.POINTER(abi.bc_struct)]
    L.nrpyell_bcstruct_free.argtypes = [C.POINTER(abi.bc_struct)]
    L.nrpyell_bcstruct_free.restype = None
    L.nrpyell_ghostmap_check.argtypes = [C.POINTER(abi.paramstruct), C.POINTER(abi.bc_struct)]
    L.nrpyell_default_options.argtypes = [C.POINTER(solve_options)]
    L.nrpyell_default_options.restype = None
    L.nrpyell_solver_create.argtypes = [C.POINTER(abi.griddata_struct), C.c_double, C.POINTER(solve_options),
                                        C.POINTER(vp)]
    L.nrpyell_solver_destroy.argtypes = [vp]
    L.nrpyell_solver_solve.argtypes = [vp, C.POINTER(solve_result)]
    L.nrpyell_solver_steps.argtypes = [vp, C.c_int]
    L.nrpyell_solver_residual_l2.argtypes = [vp, C.POINTER(C.c_double)]
    L.nrpyell_solver_time_substep.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.nrpyell_solver_timed_steps.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.nrpyell_solver_history.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.nrpyell_solver_substeps_per_launch.argtypes = [vp]
    L.nrpyell_solver_block_order.argtypes = [vp, C.POINTER(C.c_ushort), C.c_int]
    L.nrpyell_block_order_plan.argtypes = [C.c_int] * 6 + [C.POINTER(C.c_ushort), C.c_int]
    L.nrpyell_set_eta_damping.argtypes = [C.POINTER(abi.griddata_struct), C.c_double]
    L.nrpyell_batch_create.argtypes = [C.POINTER(C.POINTER(abi.griddata_struct)), C.POINTER(C.c_double),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.POINTER(solve_options),
                                       C.POINTER(vp)]
    L.nrpyell_batch_solve.argtypes = [vp, C.POINTER(solve_result), C.POINTER(C.c_double)]
    L.nrpyell_batch_destroy.argtypes = [vp]
    L.nrpyell_Initialize_ADMBase.argtypes = ([C.POINTER(abi.griddata_struct), C.POINTER(admbase_options),
                                              C.POINTER(C.c_int)] + [abi.REALP] * 20)
    L.nrpyell_apply_bcs_sommerfeld.argtypes = [C.POINTER(abi.paramstruct), C.POINTER(abi.bc_struct), abi.XX3,
                                               C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                               abi.REALP, abi.REALP]
    L.nrpyell_L2_norm_of_gf_noalloc.argtypes = [C.POINTER(abi.paramstruct), C.c_int, C.c_double, abi.XX3, abi.REALP,
                                                C.POINTER(C.c_double)]
    L.compute_TRC.argtypes = [C.c_double] * 6
    L.compute_TRC.restype = C.c_double
    L.set_optimum_eta_damping.argtypes = [C.c_double, C.c_double, C.POINTER(abi.paramstruct)]
    L.set_optimum_eta_damping.restype = None
    _LIB = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != NRPYELL_OK:
        msg = lib().nrpyell_last_error().decode(errors="replace")
        raise NrpyellError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")


def default_options(**kw) -> solve_options:
    """nrpyell_solve_options with the library defaults, overridden by keyword; outer_bc takes
    "radiation" or "sommerfeld6" and stop_rule "every_step" or "et_thorn" as well as the
    integers."""
    o = solve_options()
    lib().nrpyell_default_options(C.byref(o))
    names = {"outer_bc": OUTER_BC, "stop_rule": STOP_RULE}
    for k, v in kw.items():
        if k in names and isinstance(v, str):
            if v not in names[k]:
                raise NrpyellError(f"{k} {v!r}: expected one of {sorted(names[k])}")
            v = names[k][v]
        setattr(o, k, v)
    return o


def block_order_plan(ncu: int, Nxx, tile_i0: int = 32, planes: int = 2) -> np.ndarray:
    """Placement of the two-substep kernel's blocks that a solver plans for a grid of Nxx
    interior points on a GPU of ncu CUs (entry L: the block workgroup L computes, BC blocks
    first); empty when the kernel keeps its own order. Host computation only."""
    buf = (C.c_ushort * 1024)()
    n = lib().nrpyell_block_order_plan(int(ncu), int(Nxx[0]), int(Nxx[1]), int(Nxx[2]), int(tile_i0), int(planes),
                                       buf, 1024)
    if n < 0:
        check(n, "block_order_plan")
    return np.array(buf[:n], dtype=np.int64)


class DeviceArray:
    """A device allocation of float64 (or another dtype) made through the library."""

    def __init__(self, n: int, dtype=np.float64):
        self.n, self.dtype = int(n), np.dtype(dtype)
        self.ptr = C.c_void_p()
        check(lib().nrpyell_malloc_device(C.byref(self.ptr), self.nbytes), "malloc")

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(lib().nrpyell_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes), "h2d")
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        check(lib().nrpyell_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), "d2h")
        return out

    def as_real_p(self):
        return C.cast(self.ptr, abi.REALP)

    def free(self):
        if self.ptr:
            lib().nrpyell_free_device(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download_ptr(p, n: int, dtype=np.float64) -> np.ndarray:
    out = np.empty(n, dtype=dtype)
    check(lib().nrpyell_memcpy_d2h(out.ctypes.data, C.cast(p, C.c_void_p), out.nbytes), "d2h")
    return out


def upload_ptr(p, a: np.ndarray) -> None:
    a = np.ascontiguousarray(a)
    check(lib().nrpyell_memcpy_h2d(C.cast(p, C.c_void_p), a.ctypes.data, a.nbytes), "h2d")


class Problem:
    """One puncture configuration set up on the GPU exactly as main.c:12-105 sets it up on
    the host: eigen grid -> bcstruct, SinhSymTP grid -> dt, MoL buffers, rfm factors, zero
    initial guess, auxevol (psi_bg, AA^2), wavespeed, wavespeed at the outer boundary."""

    def __init__(self, cfg: PunctureConfig):
        L = lib()
        self.cfg = cfg
        gd = abi.griddata_struct()
        self.gd = gd
        p = gd.params
        L.set_Cparameters_to_default(C.byref(p))
        p.outer_bc_type = abi.RADIATION_OUTER_BCS
        cfg.fill_params(p)
        Nxx = (C.c_int * 3)(*[int(n) for n in cfg.Nxx])
        L.set_Nxx_dxx_invdx_params__and__xx(1, C.byref(Nxx), C.byref(p), gd.xx)
        check(L.nrpyell_bcstruct_set_up_checked(C.byref(p), gd.xx, C.byref(gd.bcstruct)), "bcstruct_set_up")
        self._free_xx()
        L.set_Nxx_dxx_invdx_params__and__xx(0, C.byref(Nxx), C.byref(p), gd.xx)
        self.Ng = [p.Nxx_plus_2NGHOSTS0, p.Nxx_plus_2NGHOSTS1, p.Nxx_plus_2NGHOSTS2]
        self.ntot = self.Ng[0] * self.Ng[1] * self.Ng[2]
        self.xx = [np.ctypeslib.as_array(gd.xx[d], shape=(self.Ng[d],)) for d in range(3)]
        self.dt = L.find_timestep(C.byref(p), gd.xx, cfg.CFL)
        L.MoL_malloc_y_n_gfs(C.byref(p), C.byref(gd.gridfuncs))
        L.rfm_precompute_rfmstruct_malloc(C.byref(p), C.byref(gd.rfmstruct))
        L.rfm_precompute_rfmstruct_define(C.byref(p), gd.xx, C.byref(gd.rfmstruct))
        L.initial_guess_all_points(C.byref(p), gd.xx, gd.gridfuncs.y_n_gfs)
        L.MoL_malloc_non_y_n_gfs(C.byref(p), C.byref(gd.gridfuncs))
        L.auxevol_gfs_all_points(C.byref(p), gd.xx, gd.gridfuncs.auxevol_gfs)
        L.wavespeed_gf_all_points(C.byref(p), cfg.CFL, self.dt, gd.xx, gd.gridfuncs.auxevol_gfs)
        c = L.compute_wavespeed_at_OB(C.byref(p), gd.gridfuncs.auxevol_gfs)
        p.uu_wavespeed_at_OB = p.vv_wavespeed_at_OB = c
        self.cfl = cfg.CFL
        self._solver: Optional[Solver] = None
        self._closed = False

    def _free_xx(self):
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        for d in range(3):
            libc.free(C.cast(self.gd.xx[d], C.c_void_p))
            self.gd.xx[d] = abi.REALP()

    def set_wave_speed_cfl(self, cfl: float) -> None:
        """Recompute the wave speed c = cfl * (local minimum proper spacing) / dt
        (wavespeed_gf_all_points.c:40-41) and the outer-boundary speed from it
        (compute_wavespeed_at_OB.c:10). dt stays: find_timestep returns the smallest proper
        spacing whatever the CFL factor (find_timestep.c:45). The relaxation's steady state,
        c^2 (lap u + AA^2/(8 (psi+u)^7)) = 0 with the radiation BC, does not depend on c; the
        factor sets how fast the relaxation moves and whether RK4 stays stable. Solvers built
        before the call keep the old speed."""
        L = lib()
        p = self.gd.params
        L.wavespeed_gf_all_points(C.byref(p), float(cfl), self.dt, self.gd.xx, self.gd.gridfuncs.auxevol_gfs)
        c = L.compute_wavespeed_at_OB(C.byref(p), self.gd.gridfuncs.auxevol_gfs)
        p.uu_wavespeed_at_OB = p.vv_wavespeed_at_OB = c
        self.cfl = float(cfl)

    def set_eta_damping(self, eta: float) -> None:
        """Change the damping mid-relaxation keeping u and u_t = v - eta u (nrpyell_set_eta_damping)."""
        check(lib().nrpyell_set_eta_damping(C.byref(self.gd), float(eta)), "set_eta_damping")

    def zero_guess(self) -> None:
        """u = v = 0 at every point (initial_guess_all_points.c:7), the reference's start."""
        lib().initial_guess_all_points(C.byref(self.gd.params), self.gd.xx, self.gd.gridfuncs.y_n_gfs)

    # --- device arrays of the griddata ----------------------------------------------------
    def get(self, name: str) -> np.ndarray:
        n = (3 if name == "auxevol_gfs" else 2) * self.ntot
        return download_ptr(getattr(self.gd.gridfuncs, name), n)

    def set(self, name: str, a: np.ndarray) -> None:
        upload_ptr(getattr(self.gd.gridfuncs, name), np.ascontiguousarray(a, dtype=np.float64))

    def rfm(self, field: str) -> np.ndarray:
        n = self.Ng[0] if field.startswith(("f0", "f3")) else self.Ng[1] if field.startswith("f1") \
            else self.Ng[0] * self.Ng[1]
        return download_ptr(getattr(self.gd.rfmstruct, field), n)

    # --- the reference-surface operators ----------------------------------------------------
    def rhs(self, state: np.ndarray) -> np.ndarray:
        L = lib()
        din = DeviceArray(2 * self.ntot).upload(state)
        dout = DeviceArray(2 * self.ntot).upload(np.zeros(2 * self.ntot))
        L.rhs_eval(C.byref(self.gd.params), C.byref(self.gd.rfmstruct), self.gd.gridfuncs.auxevol_gfs,
                   din.as_real_p(), dout.as_real_p())
        return dout.download()

    def residual(self, state: np.ndarray) -> np.ndarray:
        din = DeviceArray(2 * self.ntot).upload(state)
        dout = DeviceArray(self.ntot).upload(np.zeros(self.ntot))
        lib().residual_all_points(C.byref(self.gd.params), C.byref(self.gd.rfmstruct), self.gd.gridfuncs.auxevol_gfs,
                                  din.as_real_p(), dout.as_real_p())
        return dout.download()

    def apply_radiation_and_inner(self, gfs: np.ndarray, rhs: np.ndarray) -> np.ndarray:
        p = self.gd.params
        dg = DeviceArray(2 * self.ntot).upload(gfs)
        dr = DeviceArray(2 * self.ntot).upload(rhs)
        ws = (C.c_double * 2)(p.uu_wavespeed_at_OB, p.vv_wavespeed_at_OB)
        finf = (C.c_double * 2)(0.0, 0.0)
        lib().apply_bcs_outerradiation_and_inner(C.byref(p), C.byref(self.gd.bcstruct), self.gd.xx, C.byref(ws),
                                                 C.byref(finf), dg.as_real_p(), dr.as_real_p())
        return dr.download()

    def apply_sommerfeld(self, gfs: np.ndarray, rhs: np.ndarray, speed=(ET_SOMMERFELD_SPEED,) * 2,
                         f_infinity=(0.0, 0.0), falloff=(3.0, 3.0)) -> np.ndarray:
        """The Einstein Toolkit thorn's outer BC (nrpyell_apply_bcs_sommerfeld) on a copy of rhs;
        defaults are the thorn's constants."""
        dg = DeviceArray(2 * self.ntot).upload(gfs)
        dr = DeviceArray(2 * self.ntot).upload(rhs)
        arr = [(C.c_double * 2)(*map(float, a)) for a in (speed, f_infinity, falloff)]
        check(lib().nrpyell_apply_bcs_sommerfeld(C.byref(self.gd.params), C.byref(self.gd.bcstruct), self.gd.xx,
                                                 arr[0], arr[1], arr[2], dg.as_real_p(), dr.as_real_p()),
              "apply_bcs_sommerfeld")
        return dr.download()

    def apply_inner_only(self, gfs: np.ndarray) -> np.ndarray:
        dg = DeviceArray(2 * self.ntot).upload(gfs)
        lib().apply_bcs_inner_only(C.byref(self.gd.params), C.byref(self.gd.bcstruct), dg.as_real_p())
        return dg.download()

    def mol_step(self, dt: Optional[float] = None) -> None:
        lib().MoL_step_forward_in_time(C.byref(self.gd), self.dt if dt is None else dt)

    def l2(self, gf: np.ndarray, radius: float = 100.0, gf_index: int = 0) -> float:
        d = DeviceArray(gf.size).upload(gf)
        return lib().L2_norm_of_gf(C.byref(self.gd.params), gf_index, radius, self.gd.xx, d.as_real_p())

    def l2_noalloc(self, gf: np.ndarray, radius: float = 100.0, gf_index: int = 0) -> float:
        """L2_norm_of_gf through nrpyell_L2_norm_of_gf_noalloc (no device allocation inside the
        call; the upload of gf here is the caller's)."""
        d = DeviceArray(gf.size).upload(gf)
        out = C.c_double()
        check(lib().nrpyell_L2_norm_of_gf_noalloc(C.byref(self.gd.params), gf_index, radius, self.gd.xx,
                                                  d.as_real_p(), C.byref(out)), "L2_norm_of_gf_noalloc")
        return out.value

    def log10_residual(self, radius: float = 100.0) -> float:
        """residual_all_points + L2_norm_of_gf on y_n, as main.c:146-151 does each step."""
        L = lib()
        p = self.gd.params
        gf = self.gd.gridfuncs
        L.residual_all_points(C.byref(p), C.byref(self.gd.rfmstruct), gf.auxevol_gfs, gf.y_n_gfs,
                              gf.diagnostic_output_gfs)
        return L.L2_norm_of_gf(C.byref(p), abi.UUGF, radius, self.gd.xx, gf.diagnostic_output_gfs)

    def bc_tables(self):
        bc = self.gd.bcstruct
        n = bc.bc_info.num_inner_boundary_points
        raw = C.string_at(C.cast(bc.inner_bc_array, C.c_void_p), n * C.sizeof(abi.innerpt_bc_struct))
        rec = np.frombuffer(raw, dtype=np.dtype([("dst", "<i4"), ("src", "<i4"), ("par", "i1", 10), ("pad", "V2")]))
        outer = []
        for gz in range(abi.NGHOSTS):
            for dirn in range(3):
                m = bc.bc_info.num_pure_outer_boundary_points[gz][dirn]
                arr = bc.pure_outer_bc_array[dirn + 3 * gz]
                outer.append([(arr[k].i0, arr[k].i1, arr[k].i2, arr[k].FACEX0, arr[k].FACEX1, arr[k].FACEX2)
                              for k in range(m)])
        return rec, outer

    # --- checkpoint / resume ------------------------------------------------------------------
    _CKPT_PARAMS = ("AMAX", "SINHWAA", "bScale", "eta_damping", "bare_mass_0", "bare_mass_1",
                    "puncture_0_x", "puncture_0_y", "puncture_0_z", "puncture_1_x", "puncture_1_y", "puncture_1_z",
                    "P0_x", "P0_y", "P0_z", "P1_x", "P1_y", "P1_z", "S0_x", "S0_y", "S0_z", "S1_x", "S1_y", "S1_z",
                    "dxx0", "dxx1", "dxx2")

    def save_state(self, path: str, steps: int = 0) -> None:
        """y_n (u and v, ghosts included), the steps taken and the grid and puncture parameters,
        to an .npz file of plain arrays. A solve resumed from it continues bit for bit: an RK4
        step reads nothing but y_n."""
        p = self.gd.params
        np.savez(path, y_n=self.get("y_n_gfs"), steps=np.int64(steps),
                 grid=np.array([p.Nxx0, p.Nxx1, p.Nxx2], dtype=np.int64),
                 params=np.array([getattr(p, k) for k in self._CKPT_PARAMS]), dt=np.float64(self.dt))

    def load_state(self, path: str) -> int:
        """Restore y_n from save_state's file (refused unless grid, parameters and dt are those of
        this problem); returns the steps recorded with it."""
        d = np.load(path)
        p = self.gd.params
        if (not np.array_equal(d["grid"], [p.Nxx0, p.Nxx1, p.Nxx2])
                or not np.array_equal(d["params"], [getattr(p, k) for k in self._CKPT_PARAMS])
                or float(d["dt"]) != self.dt):
            raise NrpyellError(f"{path}: checkpoint of a different grid, puncture set or time step")
        self.set("y_n_gfs", d["y_n"])
        return int(d["steps"])

    # --- Einstein Toolkit consumer path ----------------------------------------------------
    def set_optimum_eta_damping(self) -> float:
        """The thorn's eta_damping = -1 policy (NRPyEllipticET_conformally_flat_BBH_Hyperbolic_Relaxation.c:146-156):
        the fitted optimum, available only for domain_size 1e6, sinh_width 0.07, foci_position 5."""
        p = self.gd.params
        if not (p.AMAX == 1e6 and p.SINHWAA == 0.07 and p.bScale == 5):
            raise NrpyellError("Automatic computation of the optimum eta_damping parameter is only available if "
                               f"domain_size = 1e6 (got {p.AMAX:e}), sinh_width = 0.07 (got {p.SINHWAA:e}) and "
                               f"foci_position = 5 (got {p.bScale:e})")
        lib().set_optimum_eta_damping(self.dt, self.cfg.CFL, C.byref(p))
        self._solver = None
        return p.eta_damping

    def initialize_admbase(self, x, y, z, orbital_plane: str = "xy", position_shift=(0.0, 0.0, 0.0),
                           initial_lapse: str = "NRPyEllipticET-psi^n", lapse_exponent_n: float = -2.0,
                           NRPy_epsilon: float = 1e-8, interpolation_order: int = 4) -> dict:
        """ADMBase fields at the Cartesian points (x, y, z) from the current y_n (the thorn's
        Initialize_ADMBase with its parameter defaults); returns the 17 arrays by name."""
        if orbital_plane not in ("xy", "yx", "xz", "zx"):
            raise NrpyellError('Unsupported orbital plane. Available options are "xy" ("yx") or "xz" ("zx").')
        if initial_lapse not in ("NRPyEllipticET-psi^n", "NRPyEllipticET-averaged"):
            raise NrpyellError(f"NRPyEllipticET does not support the lapse initial data: {initial_lapse}")
        o = admbase_options()
        o.orbital_plane = 0 if orbital_plane in ("xy", "yx") else 1
        for d in range(3):
            o.position_shift[d] = float(position_shift[d])
        o.initial_lapse = 0 if initial_lapse == "NRPyEllipticET-psi^n" else 1
        o.lapse_exponent_n, o.NRPy_epsilon, o.interpolation_order = lapse_exponent_n, NRPy_epsilon, interpolation_order
        pts = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (x, y, z)]
        n = pts[0].size
        if any(a.size != n for a in pts):
            raise NrpyellError("x, y, z must have the same size")
        lsh = (C.c_int * 3)(n, 1, 1)
        dev = [DeviceArray(max(n, 1)).upload(a) if n else DeviceArray(1) for a in pts]
        outs = [DeviceArray(max(n, 1)) for _ in ADMBASE_NAMES]
        check(lib().nrpyell_Initialize_ADMBase(C.byref(self.gd), C.byref(o), lsh, *[d.as_real_p() for d in dev],
                                               *[d.as_real_p() for d in outs]), "Initialize_ADMBase")
        res = {k: d.download()[:n] for k, d in zip(ADMBASE_NAMES, outs)}
        for d in dev + outs:
            d.free()
        return res

    # --- resident solver ----------------------------------------------------------------------
    def solver(self, **opts) -> "Solver":
        return Solver(self, **opts)

    def close(self):
        if self._closed:
            return
        L = lib()
        p = self.gd.params
        if self.gd.gridfuncs.y_n_gfs:
            L.MoL_free_memory_y_n_gfs(C.byref(p), C.byref(self.gd.gridfuncs))
        if self.gd.gridfuncs.k_odd_gfs:
            L.MoL_free_memory_non_y_n_gfs(C.byref(p), C.byref(self.gd.gridfuncs))
        if self.gd.rfmstruct.f0_of_xx0:
            L.rfm_precompute_rfmstruct_freemem(C.byref(p), C.byref(self.gd.rfmstruct))
        L.nrpyell_bcstruct_free(C.byref(self.gd.bcstruct))
        self._closed = True


class Solver:
    """Resident solver: hipGraph-captured RK4 steps with an on-device residual check every
    `check_every` steps (FP64, or FP32 bulk then FP64 finish with precision=32)."""

    def __init__(self, prob: Problem, **opts):
        self.prob = prob
        self.opts = default_options(**opts)
        self.h = C.c_void_p()
        check(lib().nrpyell_solver_create(C.byref(prob.gd), prob.dt, C.byref(self.opts), C.byref(self.h)),
              "solver_create")

    def solve(self) -> solve_result:
        r = solve_result()
        rc = lib().nrpyell_solver_solve(self.h, C.byref(r))
        self.last = r
        if rc not in (NRPYELL_OK, -4):
            check(rc, "solve")
        return r

    def steps(self, n: int) -> None:
        check(lib().nrpyell_solver_steps(self.h, n), "steps")

    def log10_residual(self) -> float:
        v = C.c_double()
        check(lib().nrpyell_solver_residual_l2(self.h, C.byref(v)), "residual_l2")
        return v.value

    def timed_steps(self, n: int):
        """Exactly n production RK4 steps timed with HIP events: (milliseconds, final log10 L2)."""
        ms, l2 = C.c_double(), C.c_double()
        check(lib().nrpyell_solver_timed_steps(self.h, n, C.byref(ms), C.byref(l2)), "timed_steps")
        return ms.value, l2.value

    def write_residual_log(self, path: str, max_entries: int = 4096) -> int:
        """The residual history in main.c's residual_l2_norm.txt format (main.c:153-155):
        "%6d %10.4e %.17e" = step, time = step * dt, log10 L2 of the residual; returns the
        number of lines."""
        steps, vals = self.history(max_entries)
        dt = self.prob.dt
        with open(path, "w") as f:
            for n, l in zip(steps, vals):
                f.write("%6d %10.4e %.17e\n" % (int(n), float(n) * dt, float(l)))
        return len(steps)

    def history(self, max_entries: int = 4096):
        steps = (C.c_int * max_entries)()
        vals = (C.c_double * max_entries)()
        n = lib().nrpyell_solver_history(self.h, max_entries, steps, vals)
        return np.array(steps[:n]), np.array(vals[:n])

    @property
    def substeps_per_launch(self) -> int:
        """1 or 2: RK4 substeps per launch of the FP64 steps."""
        return lib().nrpyell_solver_substeps_per_launch(self.h)

    def block_order(self) -> np.ndarray:
        """Placement of the two-substep kernel's blocks: entry L is the block (BC blocks first,
        then interior blocks in tile order) that workgroup L of a launch computes; empty when the
        kernel runs its own order."""
        buf = (C.c_ushort * 1024)()
        n = lib().nrpyell_solver_block_order(self.h, buf, 1024)
        return np.array(buf[:n], dtype=np.int64)

    def time_substep(self, stage: int, n: int) -> float:
        v = C.c_double()
        check(lib().nrpyell_solver_time_substep(self.h, stage, n, C.byref(v)), "time_substep")
        return v.value

    def close(self):
        if self.h:
            lib().nrpyell_solver_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchSolver:
    """Relaxes many configurations (uneven grids) together on one GPU: one launch per RK4
    substep per tile width covers the whole batch; each configuration stops at its own
    residual target. `tiles` gives each configuration's tile as an i0 width or a
    (width, planes) pair (from the autotuner)."""

    def __init__(self, problems, tiles=None, **opts):
        self.problems = list(problems)
        n = len(self.problems)
        self.opts = default_options(**opts)
        gds = (C.POINTER(abi.griddata_struct) * n)(*[C.pointer(p.gd) for p in self.problems])
        dts = (C.c_double * n)(*[p.dt for p in self.problems])
        pairs = [t if isinstance(t, (tuple, list)) else (t, 0) for t in (tiles or [0] * n)]
        tl = (C.c_int * n)(*[int(t[0]) for t in pairs])
        pl = (C.c_int * n)(*[int(t[1]) for t in pairs])
        self.h = C.c_void_p()
        check(lib().nrpyell_batch_create(gds, dts, tl, pl, n, C.byref(self.opts), C.byref(self.h)), "batch_create")

    def solve(self):
        n = len(self.problems)
        res = (solve_result * n)()
        sec = C.c_double()
        rc = lib().nrpyell_batch_solve(self.h, res, C.byref(sec))
        # not converged / NaN in some configurations: reported per configuration (converged = 0)
        if rc not in (NRPYELL_OK, -4, -5):
            check(rc, "batch_solve")
        return list(res), sec.value

    def close(self):
        if self.h:
            lib().nrpyell_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# A configuration whose relaxation diverges at the reference's CFL factor 0.7 is solved again
# from the zero guess at this factor. Frozen-coefficient bound: the 10th-order second
# difference has symbol at most 6.8267 (at the Nyquist mode), and with the local wave speed
# c = CFL ds_min/dt the RK4 step sees |lambda dt| <= CFL sqrt(6.8267 (1 + r1^2 + r2^2)), r the
# ratios of the smallest proper spacing to the other two; RK4 is stable on the imaginary axis up
# to 2 sqrt(2). At 0.7 every grid of the batched workloads exceeds that somewhere near the
# origin (2.6-3.2; most configurations relax anyway); at 0.5 every one stays below it (<= 2.26).
FALLBACK_CFL = 0.5


def relax_batch(problems, tiles=None, fallback_cfl: Optional[float] = FALLBACK_CFL, **opts):
    """Relax `problems` together (BatchSolver); then every configuration whose relaxation
    diverged (status "diverged": its residual rose 4 decades above its lowest) is relaxed again,
    from the zero guess, with the wave speed set at `fallback_cfl` (set_wave_speed_cfl; the
    converged u does not depend on it). Returns (results, seconds, info): results[k] is the
    configuration's final solve, info[k] {"fallback": bool, "first_steps": steps of the diverged
    attempt}. Problems that fell back keep the fallback wave speed."""
    tiles = list(tiles) if tiles is not None else None
    B = BatchSolver(problems, tiles=tiles, **opts)
    try:
        res, sec = B.solve()
    finally:
        B.close()
    info = [{"fallback": False, "first_steps": 0} for _ in problems]
    redo = [k for k, r in enumerate(res) if r.status == 3]
    if redo and fallback_cfl:
        sub = [problems[k] for k in redo]
        for P in sub:
            P.set_wave_speed_cfl(fallback_cfl)
            P.zero_guess()
        B2 = BatchSolver(sub, tiles=[tiles[k] for k in redo] if tiles is not None else None, **opts)
        try:
            res2, sec2 = B2.solve()
        finally:
            B2.close()
        sec += sec2
        for k, r in zip(redo, res2):
            info[k] = {"fallback": True, "first_steps": int(res[k].steps)}
            res[k] = r
    return res, sec, info
