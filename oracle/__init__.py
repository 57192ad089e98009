"""Checker bindings: the C restatement (liboracle.so) and, when it was built, the compiled
reference C (oracle/_ref).

TEST INFRASTRUCTURE. Imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg. Nothing in nrpyelliptic_amd/ imports this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from nrpyelliptic_amd import abi
from nrpyelliptic_amd.config import PunctureConfig

HERE = os.path.dirname(os.path.abspath(__file__))
NG = 5
DP = C.POINTER(C.c_double)


class OrcGrid(C.Structure):
    _fields_ = [("N", C.c_int * 3), ("Ng", C.c_int * 3), ("dxx", C.c_double * 3), ("invdx", C.c_double * 3),
                ("AMAX", C.c_double), ("SINHWAA", C.c_double), ("bScale", C.c_double), ("eta", C.c_double),
                ("mass", C.c_double * 2), ("pos", (C.c_double * 3) * 2), ("P", (C.c_double * 3) * 2),
                ("S", (C.c_double * 3) * 2), ("cOB", C.c_double * 2), ("outer_bc", C.c_int)]


ORC_BC_RADIATION, ORC_BC_SOMMERFELD6 = 0, 1
# the wave speed the Einstein Toolkit thorn's Sommerfeld BC uses on every grid
# (conformally_flat_BBH_apply_bcs_sommerfeld.h:12)
ET_SOMMERFELD_SPEED = 4.4793122676309114e+06


ORC_RFM = ["f0", "f0_D0", "f0_DD00", "f0_DDD000", "f1", "f1_D1", "f1_DD11", "f2", "f2_D0", "f2_D1",
           "f2_DD00", "f2_DD11", "f3", "f3_D0", "f3_DD00"]


class OrcRfm(C.Structure):
    _fields_ = [(n, DP) for n in ORC_RFM]


class OrcBc(C.Structure):
    _fields_ = [("num_inner", C.c_int), ("inner", C.POINTER(abi.innerpt_bc_struct)),
                ("num_outer", C.c_int * 15), ("outer", C.POINTER(abi.outerpt_bc_struct) * 15)]


class OrcAdmOpts(C.Structure):
    _fields_ = [("orbital_plane", C.c_int), ("position_shift", C.c_double * 3), ("initial_lapse", C.c_int),
                ("lapse_exponent", C.c_double), ("nrpy_epsilon", C.c_double), ("interp_order", C.c_int)]


ADM_NAMES = ["uuGF", "alp", "betax", "betay", "betaz", "gxx", "gxy", "gxz", "gyy", "gyz", "gzz",
             "kxx", "kxy", "kxz", "kyy", "kyz", "kzz"]


def adm_opts(orbital_plane="xy", position_shift=(0.0, 0.0, 0.0), initial_lapse="psi^n", lapse_exponent=-2.0,
             nrpy_epsilon=1e-8, interp_order=4) -> "OrcAdmOpts":
    o = OrcAdmOpts()
    o.orbital_plane = 0 if orbital_plane in ("xy", "yx") else 1
    for d in range(3):
        o.position_shift[d] = position_shift[d]
    o.initial_lapse = 0 if initial_lapse in ("psi^n", "NRPyEllipticET-psi^n") else 1
    o.lapse_exponent, o.nrpy_epsilon, o.interp_order = lapse_exponent, nrpy_epsilon, interp_order
    return o


def _p(a):
    return a.ctypes.data_as(DP)


_LIB = None


def build():
    """Compile liboracle.so (and the reference build when /root/reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if os.path.isdir("/root/reference/NRPyElliptic_Ccodes"):
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref"], check=True)
        if os.path.exists(os.path.join(HERE, "..", "nrpyelliptic_amd", "libnrpyell_hip.so")):
            subprocess.run(["make", "-s", "-C", HERE, "drivers"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        G = C.POINTER(OrcGrid)
        L.orc_set_grid.argtypes = [G, C.c_int, DP, DP, DP]
        L.orc_rfm_define.argtypes = [G, DP, DP, C.POINTER(OrcRfm)]
        L.orc_auxevol.argtypes = [G, DP, DP, DP, DP]
        L.orc_find_timestep.argtypes = [G, DP, DP, DP]
        L.orc_find_timestep.restype = C.c_double
        L.orc_wavespeed.argtypes = [G, C.c_double, C.c_double, DP, DP, DP, DP]
        L.orc_wavespeed_at_OB.argtypes = [G, DP]
        L.orc_wavespeed_at_OB.restype = C.c_double
        L.orc_bcstruct.argtypes = [G, DP, DP, DP, C.POINTER(OrcBc)]
        L.orc_rhs_eval.argtypes = [G, C.POINTER(OrcRfm), DP, DP, DP]
        L.orc_residual.argtypes = [G, C.POINTER(OrcRfm), DP, DP, DP]
        L.orc_apply_bcs_inner_only.argtypes = [G, C.POINTER(OrcBc), DP]
        L.orc_apply_bcs_radiation_and_inner.argtypes = [G, C.POINTER(OrcBc), DP, DP, DP, DP, DP]
        L.orc_apply_bcs_sommerfeld.argtypes = [G, C.POINTER(OrcBc), DP, DP, DP, DP, DP, DP, DP, DP]
        L.orc_sommerfeld_fd1.argtypes = [G, DP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_sommerfeld_fd1.restype = C.c_double
        L.orc_mol_step.argtypes = [G, C.POINTER(OrcRfm), C.POINTER(OrcBc), DP, DP, DP, DP, DP, DP, DP, DP,
                                   C.c_double]
        L.orc_l2_norm.argtypes = [G, C.c_double, DP, DP, DP, DP]
        L.orc_l2_norm.restype = C.c_double
        L.orc_fd1_weight.argtypes = [C.c_int, C.c_int]
        L.orc_fd1_weight.restype = C.c_double
        L.orc_get_xx_from_xyz.argtypes = [C.c_double, C.c_double, C.c_double, DP, DP]
        L.orc_get_xx_from_xyz.restype = None
        L.orc_lagrange_interp3.argtypes = [C.POINTER(C.c_int), DP, DP, DP, C.c_int, DP, DP]
        L.orc_adm_point.argtypes = [G, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, DP]
        L.orc_adm_point.restype = None
        L.orc_initialize_admbase.argtypes = [G, DP, DP, DP, DP, C.POINTER(OrcAdmOpts), C.c_long, DP, DP, DP,
                                             C.POINTER(DP)]
        L.orc_compute_TRC.argtypes = [C.c_double] * 6
        L.orc_compute_TRC.restype = C.c_double
        L.orc_optimum_eta.argtypes = [C.c_double] * 6
        L.orc_optimum_eta.restype = C.c_double
        _LIB = L
    return _LIB


class OracleProblem:
    """One configuration set up entirely by the C restatement, arrays held in numpy."""

    def __init__(self, cfg: PunctureConfig, outer_bc: str = "radiation", outer_bc_speed: float = 0.0):
        """outer_bc "radiation" (NRPyElliptic) or "sommerfeld6" (the Einstein Toolkit thorn's
        BC, with its hard-coded wave speed unless outer_bc_speed > 0)."""
        L = lib()
        self.cfg = cfg
        g = OrcGrid()
        for d in range(3):
            g.N[d] = int(cfg.Nxx[d])
        g.AMAX, g.SINHWAA, g.bScale, g.eta = cfg.AMAX, cfg.SINHWAA, cfg.bScale, cfg.eta
        g.mass[0], g.mass[1] = cfg.bare_mass
        for d in range(3):
            g.pos[0][d], g.pos[1][d] = cfg.pos0[d], cfg.pos1[d]
            g.P[0][d], g.P[1][d] = cfg.P0[d], cfg.P1[d]
            g.S[0][d], g.S[1][d] = cfg.S0[d], cfg.S1[d]
        self.g = g
        Ng = [int(n) + 2 * NG for n in cfg.Nxx]
        self.Ng = Ng
        self.ntot = Ng[0] * Ng[1] * Ng[2]
        # eigen grid -> bcstruct (main.c:46-49), then the SinhSymTP grid (main.c:56-57)
        self.xxe = [np.zeros(n) for n in Ng]
        L.orc_set_grid(C.byref(g), 1, *[_p(a) for a in self.xxe])
        self._bc_storage = []
        self.bc = OrcBc()
        self.bc.inner = self._alloc(abi.innerpt_bc_struct, self.ntot)
        for s in range(15):
            self.bc.outer[s] = self._alloc(abi.outerpt_bc_struct, self.ntot)
        if L.orc_bcstruct(C.byref(g), *[_p(a) for a in self.xxe], C.byref(self.bc)) != 0:
            raise RuntimeError("oracle bcstruct: Cartesian round-trip disagreement")
        self.xx = [np.zeros(n) for n in Ng]
        L.orc_set_grid(C.byref(g), 0, *[_p(a) for a in self.xx])
        self.dt = L.orc_find_timestep(C.byref(g), *[_p(a) for a in self.xx])
        self.rfm_arrays = {}
        self.rfm = OrcRfm()
        for n in ORC_RFM:
            size = Ng[0] if n.startswith(("f0", "f3")) else Ng[1] if n.startswith("f1") else Ng[0] * Ng[1]
            a = np.zeros(size)
            self.rfm_arrays[n] = a
            setattr(self.rfm, n, _p(a))
        L.orc_rfm_define(C.byref(g), _p(self.xx[0]), _p(self.xx[1]), C.byref(self.rfm))
        self.aux = np.zeros(3 * self.ntot)
        L.orc_auxevol(C.byref(g), *[_p(a) for a in self.xx], _p(self.aux))
        L.orc_wavespeed(C.byref(g), cfg.CFL, self.dt, *[_p(a) for a in self.xx], _p(self.aux))
        c = L.orc_wavespeed_at_OB(C.byref(g), _p(self.aux))
        if outer_bc == "sommerfeld6":
            g.outer_bc = ORC_BC_SOMMERFELD6
            c = outer_bc_speed if outer_bc_speed > 0 else ET_SOMMERFELD_SPEED
        elif outer_bc != "radiation":
            raise ValueError(f"outer_bc {outer_bc!r}")
        g.cOB[0] = g.cOB[1] = c

    def initialize_admbase(self, u, x, y, z, **opts):
        """The Einstein Toolkit consumer path on the oracle's grid: u (gridfunction 0 with
        ghosts) -> dict of the 17 ADMBase arrays at the points (x, y, z)."""
        o = adm_opts(**opts)
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        u = np.ascontiguousarray(u[: self.ntot], dtype=np.float64)
        outs = [np.zeros(len(x)) for _ in ADM_NAMES]
        optr = (DP * len(outs))(*[_p(a) for a in outs])
        rc = lib().orc_initialize_admbase(C.byref(self.g), *[_p(a) for a in self.xx], _p(u), C.byref(o), len(x),
                                          _p(x), _p(y), _p(z), optr)
        if rc != 0:
            raise RuntimeError(f"oracle initialize_admbase: rc {rc}")
        return dict(zip(ADM_NAMES, outs))

    def _alloc(self, T, n):
        arr = (T * n)()
        self._bc_storage.append(arr)
        return C.cast(arr, C.POINTER(T))

    # --- tables as numpy -----------------------------------------------------------------
    def inner_table(self):
        n = self.bc.num_inner
        raw = np.ctypeslib.as_array(C.cast(self.bc.inner, C.POINTER(C.c_uint8 * C.sizeof(abi.innerpt_bc_struct))),
                                    shape=(n,))
        return bytes(memoryview(np.ascontiguousarray(raw)))

    def inner_arrays(self):
        n = self.bc.num_inner
        dst = np.array([self.bc.inner[k].dstpt for k in range(n)], dtype=np.int64)
        src = np.array([self.bc.inner[k].srcpt for k in range(n)], dtype=np.int64)
        par = np.array([list(self.bc.inner[k].parity) for k in range(n)], dtype=np.int8)
        return dst, src, par

    def outer_lists(self):
        out = []
        for s in range(15):
            n = self.bc.num_outer[s]
            out.append([(o.i0, o.i1, o.i2, o.FACEX0, o.FACEX1, o.FACEX2)
                        for o in (self.bc.outer[s][k] for k in range(n))])
        return out

    # --- operators ------------------------------------------------------------------------
    def rhs(self, state, rfm=None, aux=None):
        out = np.zeros(2 * self.ntot)
        lib().orc_rhs_eval(C.byref(self.g), C.byref(rfm or self.rfm), _p(self.aux if aux is None else aux),
                           _p(state), _p(out))
        return out

    def residual(self, u_gfs, aux=None):
        out = np.zeros(self.ntot)
        lib().orc_residual(C.byref(self.g), C.byref(self.rfm), _p(self.aux if aux is None else aux), _p(u_gfs),
                           _p(out))
        return out

    def apply_radiation_and_inner(self, gfs, rhs):
        lib().orc_apply_bcs_radiation_and_inner(C.byref(self.g), C.byref(self.bc), *[_p(a) for a in self.xx],
                                                _p(gfs), _p(rhs))

    def apply_sommerfeld(self, gfs, rhs, speed=None, finf=(0.0, 0.0), falloff=(3.0, 3.0)):
        sp = np.array(speed if speed is not None else self.g.cOB[:], dtype=np.float64)
        fi, fo = np.array(finf, dtype=np.float64), np.array(falloff, dtype=np.float64)
        lib().orc_apply_bcs_sommerfeld(C.byref(self.g), C.byref(self.bc), *[_p(a) for a in self.xx], _p(sp),
                                       _p(fi), _p(fo), _p(gfs), _p(rhs))

    def sommerfeld_fd1(self, f, d, i0, i1, i2, face):
        return lib().orc_sommerfeld_fd1(C.byref(self.g), _p(f), d, i0, i1, i2, face)

    def apply_inner_only(self, gfs):
        lib().orc_apply_bcs_inner_only(C.byref(self.g), C.byref(self.bc), _p(gfs))

    def new_state(self):
        return {k: np.zeros(2 * self.ntot) for k in ("y_n", "rt", "k_odd", "k_even")}

    def mol_step(self, st, dt=None, aux=None):
        lib().orc_mol_step(C.byref(self.g), C.byref(self.rfm), C.byref(self.bc), *[_p(a) for a in self.xx],
                           _p(self.aux if aux is None else aux), _p(st["y_n"]), _p(st["rt"]), _p(st["k_odd"]),
                           _p(st["k_even"]), self.dt if dt is None else dt)

    def l2(self, gf, radius=100.0):
        return lib().orc_l2_norm(C.byref(self.g), radius, *[_p(a) for a in self.xx], _p(gf))

    def log10_residual(self, y_n, radius=100.0):
        return self.l2(self.residual(y_n), radius)

    def relax_every_step(self, st, target, max_steps):
        """The residual of y_n before every step (main.c:136-151: residual_all_points +
        L2_norm_of_gf at the top of each iteration, then MoL_step_forward_in_time at :185);
        stop at the first log10 L2 <= target or after max_steps steps. Returns (steps,
        history of (steps, log10 L2))."""
        hist = []
        n = 0
        while True:
            l2 = self.log10_residual(st["y_n"])
            hist.append((n, l2))
            if l2 <= target or n >= max_steps:
                return n, hist
            self.mol_step(st)
            n += 1

    def relax_et_thorn(self, st, target, max_iterations, info_output_freq):
        """The Einstein Toolkit thorn's relaxation loop
        (NRPyEllipticET_conformally_flat_BBH_Hyperbolic_Relaxation.c:181-229): k =
        info_output_freq (1 when 0); for n = 0..max_iterations: one MoL step, then, when
        n % k == 0 or n == max_iterations, the residual of the state after n + 1 steps; stop when
        log10 L2 < target. If no check passes the loop ends after max_iterations + 1 steps (the
        error branch after the loop tests n == N_final, which the loop never leaves n at).
        Returns (steps, converged, history of (steps, log10 L2))."""
        k = info_output_freq if info_output_freq != 0 else 1
        hist = []
        for n in range(max_iterations + 1):
            self.mol_step(st)
            if n % k == 0 or n == max_iterations:
                l2 = self.log10_residual(st["y_n"])
                hist.append((n + 1, l2))
                if l2 < target:
                    return n + 1, True, hist
        return max_iterations + 1, False, hist


# ---------------------------------------------------------------------------------------
# The compiled reference (oracle/_ref), present where it could be built.
# ---------------------------------------------------------------------------------------
def _cpu_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            return " avx512f " in f.read().replace("\n", " ")
    except OSError:
        return False


_REF = None


def ref_lib():
    """The reference C build, or None when oracle/_ref was not built."""
    global _REF
    if _REF is None:
        name = "libnrpyell_ref_v4.so" if _cpu_has_avx512() else "libnrpyell_ref_v3.so"
        path = os.path.join(HERE, "_ref", name)
        if not os.path.exists(path):
            return None
        _REF = abi.bind(C.CDLL(path))
    return _REF


class RefProblem:
    """One configuration set up by the compiled reference C, driven through its own ABI
    exactly as main.c:12-105 does (host arrays)."""

    def __init__(self, cfg: PunctureConfig, lib_=None):
        R = lib_ or ref_lib()
        if R is None:
            raise RuntimeError("reference build not available")
        self.R = R
        self.cfg = cfg
        gd = abi.griddata_struct()
        self.gd = gd
        R.set_Cparameters_to_default(C.byref(gd.params))
        gd.params.outer_bc_type = abi.RADIATION_OUTER_BCS
        cfg.fill_params(gd.params)
        Nxx = (C.c_int * 3)(*[int(n) for n in cfg.Nxx])
        self._keep = []
        R.set_Nxx_dxx_invdx_params__and__xx(1, C.byref(Nxx), C.byref(gd.params), gd.xx)
        self.xxe = [np.ctypeslib.as_array(gd.xx[d], shape=(int(cfg.Nxx[d]) + 2 * NG,)).copy() for d in range(3)]
        R.bcstruct_set_up(C.byref(gd.params), gd.xx, C.byref(gd.bcstruct))
        R.set_Nxx_dxx_invdx_params__and__xx(0, C.byref(Nxx), C.byref(gd.params), gd.xx)
        p = gd.params
        self.Ng = [p.Nxx_plus_2NGHOSTS0, p.Nxx_plus_2NGHOSTS1, p.Nxx_plus_2NGHOSTS2]
        self.ntot = self.Ng[0] * self.Ng[1] * self.Ng[2]
        self.xx = [np.ctypeslib.as_array(gd.xx[d], shape=(self.Ng[d],)) for d in range(3)]
        self.dt = R.find_timestep(C.byref(p), gd.xx, cfg.CFL)
        R.MoL_malloc_y_n_gfs(C.byref(p), C.byref(gd.gridfuncs))
        R.rfm_precompute_rfmstruct_malloc(C.byref(p), C.byref(gd.rfmstruct))
        R.rfm_precompute_rfmstruct_define(C.byref(p), gd.xx, C.byref(gd.rfmstruct))
        R.initial_guess_all_points(C.byref(p), gd.xx, gd.gridfuncs.y_n_gfs)
        R.MoL_malloc_non_y_n_gfs(C.byref(p), C.byref(gd.gridfuncs))
        R.auxevol_gfs_all_points(C.byref(p), gd.xx, gd.gridfuncs.auxevol_gfs)
        R.wavespeed_gf_all_points(C.byref(p), cfg.CFL, self.dt, gd.xx, gd.gridfuncs.auxevol_gfs)
        c = R.compute_wavespeed_at_OB(C.byref(p), gd.gridfuncs.auxevol_gfs)
        p.uu_wavespeed_at_OB = p.vv_wavespeed_at_OB = c
        # The stage loops of the SIMD build run over all points, so every array is initialised
        # (auxevol wavespeed ghosts are never written by the reference; zero them).
        nt = self.ntot
        for name in ("y_nplus1_running_total_gfs", "k_odd_gfs", "k_even_gfs"):
            self.arr(name)[:] = 0.0
        aux = self.arr("auxevol_gfs")
        ws = np.ctypeslib.as_array(gd.gridfuncs.auxevol_gfs, shape=(3 * nt,))
        mask = np.ones(self.Ng[::-1], dtype=bool)
        mask[NG:-NG, NG:-NG, NG:-NG] = False
        ws[2 * nt:][mask.ravel()] = 0.0
        del aux

    def arr(self, name):
        n = 3 * self.ntot if name == "auxevol_gfs" else 2 * self.ntot
        return np.ctypeslib.as_array(getattr(self.gd.gridfuncs, name), shape=(n,))

    def rfm_arr(self, field):
        """rfm array by oracle name (f0, f2_D0, ...) or reference member name."""
        if field in ORC_RFM:
            field = abi.RFM_FIELDS[ORC_RFM.index(field)]
        n = self.Ng[0] if field.startswith(("f0", "f3")) else self.Ng[1] if field.startswith("f1") \
            else self.Ng[0] * self.Ng[1]
        return np.ctypeslib.as_array(getattr(self.gd.rfmstruct, field), shape=(n,))

    def inner_arrays(self):
        bc = self.gd.bcstruct
        n = bc.bc_info.num_inner_boundary_points
        dst = np.array([bc.inner_bc_array[k].dstpt for k in range(n)], dtype=np.int64)
        src = np.array([bc.inner_bc_array[k].srcpt for k in range(n)], dtype=np.int64)
        par = np.array([list(bc.inner_bc_array[k].parity) for k in range(n)], dtype=np.int8)
        return dst, src, par

    def outer_lists(self):
        bc = self.gd.bcstruct
        out = []
        for gz in range(NG):
            for dirn in range(3):
                n = bc.bc_info.num_pure_outer_boundary_points[gz][dirn]
                arr = bc.pure_outer_bc_array[dirn + 3 * gz]
                out.append([(arr[k].i0, arr[k].i1, arr[k].i2, arr[k].FACEX0, arr[k].FACEX1, arr[k].FACEX2)
                            for k in range(n)])
        return out

    def rhs(self, state):
        out = np.zeros(2 * self.ntot)
        self.R.rhs_eval(C.byref(self.gd.params), C.byref(self.gd.rfmstruct), self.gd.gridfuncs.auxevol_gfs,
                        _p(state), _p(out))
        return out

    def residual(self, u_gfs):
        out = np.zeros(2 * self.ntot)
        self.R.residual_all_points(C.byref(self.gd.params), C.byref(self.gd.rfmstruct),
                                   self.gd.gridfuncs.auxevol_gfs, _p(u_gfs), _p(out))
        return out[: self.ntot]

    def apply_radiation_and_inner(self, gfs, rhs):
        p = self.gd.params
        ws = (C.c_double * 2)(p.uu_wavespeed_at_OB, p.vv_wavespeed_at_OB)
        finf = (C.c_double * 2)(0.0, 0.0)
        self.R.apply_bcs_outerradiation_and_inner(C.byref(p), C.byref(self.gd.bcstruct), self.gd.xx, C.byref(ws),
                                                  C.byref(finf), _p(gfs), _p(rhs))

    def mol_step(self):
        self.R.MoL_step_forward_in_time(C.byref(self.gd), self.dt)

    def l2(self, gf, radius=100.0):
        return self.R.L2_norm_of_gf(C.byref(self.gd.params), 0, radius, self.gd.xx, _p(gf))
