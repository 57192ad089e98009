"""ctypes mirror of the reference driver's C ABI (NRPyElliptic_Ccodes/NRPy_basic_defines.h).

The member order of every struct is the ABI (NRPy_basic_defines.h:40-120, 135-151,
164-195, 206-213, 263-269); include/nrpyelliptic_amd.h declares the same layout for C
callers.  Because libnrpyell_hip.so keeps this ABI, the same Python structs drive both
it and a build of the reference C (which the tests use as a checker).
"""
import ctypes as C

REAL = C.c_double
REALP = C.POINTER(C.c_double)
NGHOSTS = 5
NUM_EVOL_GFS = 2
UUGF, VVGF = 0, 1
NUM_AUXEVOL_GFS = 3
ADD_TIMES_AUUGF, PSI_BACKGROUNDGF, WAVESPEEDGF = 0, 1, 2
EXTRAPOLATION_OUTER_BCS, RADIATION_OUTER_BCS = 0, 1

_PARAM_REALS = (
    "AMAX Cart_CoM_offsetx Cart_CoM_offsety Cart_CoM_offsetz Cart_originx Cart_originy Cart_originz "
    "P0_x P0_y P0_z P1_x P1_y P1_z S0_x S0_y S0_z S1_x S1_y S1_z SINHWAA bScale bare_mass_0 bare_mass_1 "
    "dxx0 dxx1 dxx2 eta_damping f0_of_xx0 f0_of_xx0__D0 f0_of_xx0__DD00 f0_of_xx0__DDD000 f1_of_xx1 "
    "f1_of_xx1__D1 f1_of_xx1__DD11 f1_of_xx1__DDD111 f2_of_xx0_xx1 f2_of_xx0_xx1__D0 f2_of_xx0_xx1__D1 "
    "f2_of_xx0_xx1__DD00 f2_of_xx0_xx1__DD11 f2_of_xx1 f2_of_xx1__D1 f2_of_xx1__DD11 f2_of_xx1__DDD111 "
    "f3_of_xx0 f3_of_xx0__D0 f3_of_xx0__DD00 f4_of_xx2 f4_of_xx2__D2 f4_of_xx2__DD22 invdx0 invdx1 invdx2 "
    "puncture_0_x puncture_0_y puncture_0_z puncture_1_x puncture_1_y puncture_1_z time "
    "uu_wavespeed_at_OB vv_wavespeed_at_OB xx0 xx1 xx2 xxmax0 xxmax1 xxmax2 xxmin0 xxmin1 xxmin2"
).split()
_PARAM_INTS = ("Nxx0 Nxx1 Nxx2 Nxx_plus_2NGHOSTS0 Nxx_plus_2NGHOSTS1 Nxx_plus_2NGHOSTS2 "
               "has_outer_boundary outer_bc_type").split()


class paramstruct(C.Structure):
    _fields_ = [(n, REAL) for n in _PARAM_REALS] + [(n, C.c_int) for n in _PARAM_INTS]


RFM_FIELDS = ("f0_of_xx0 f0_of_xx0__D0 f0_of_xx0__DD00 f0_of_xx0__DDD000 f1_of_xx1 f1_of_xx1__D1 "
              "f1_of_xx1__DD11 f2_of_xx0_xx1 f2_of_xx0_xx1__D0 f2_of_xx0_xx1__D1 f2_of_xx0_xx1__DD00 "
              "f2_of_xx0_xx1__DD11 f3_of_xx0 f3_of_xx0__D0 f3_of_xx0__DD00").split()


class rfm_struct(C.Structure):
    _fields_ = [(n, REALP) for n in RFM_FIELDS]


class innerpt_bc_struct(C.Structure):
    _fields_ = [("dstpt", C.c_int), ("srcpt", C.c_int), ("parity", C.c_int8 * 10)]


class outerpt_bc_struct(C.Structure):
    _fields_ = [("i0", C.c_short), ("i1", C.c_short), ("i2", C.c_short),
                ("FACEX0", C.c_int8), ("FACEX1", C.c_int8), ("FACEX2", C.c_int8)]


class bc_info_struct(C.Structure):
    _fields_ = [("num_inner_boundary_points", C.c_int),
                ("num_pure_outer_boundary_points", (C.c_int * 3) * NGHOSTS),
                ("bc_loop_bounds", ((C.c_int * 6) * 6) * NGHOSTS)]


class bc_struct(C.Structure):
    _fields_ = [("inner_bc_array", C.POINTER(innerpt_bc_struct)),
                ("pure_outer_bc_array", C.POINTER(outerpt_bc_struct) * (NGHOSTS * 3)),
                ("bc_info", bc_info_struct)]


class MoL_gridfunctions_struct(C.Structure):
    _fields_ = [(n, REALP) for n in ("y_n_gfs y_nplus1_running_total_gfs k_odd_gfs k_even_gfs "
                                     "auxevol_gfs diagnostic_output_gfs").split()]


class griddata_struct(C.Structure):
    _fields_ = [("params", paramstruct), ("xx", REALP * 3), ("bcstruct", bc_struct),
                ("rfmstruct", rfm_struct), ("gridfuncs", MoL_gridfunctions_struct)]


XX3 = REALP * 3

# Function prototypes shared by the reference build and libnrpyell_hip.so
# (NRPy_function_prototypes.h).
PROTOTYPES = {
    "set_Cparameters_to_default": (None, [C.POINTER(paramstruct)]),
    "set_Nxx_dxx_invdx_params__and__xx": (None, [C.c_int, C.POINTER(C.c_int * 3), C.POINTER(paramstruct), XX3]),
    "bcstruct_set_up": (None, [C.POINTER(paramstruct), XX3, C.POINTER(bc_struct)]),
    "find_timestep": (REAL, [C.POINTER(paramstruct), XX3, REAL]),
    "MoL_malloc_y_n_gfs": (None, [C.POINTER(paramstruct), C.POINTER(MoL_gridfunctions_struct)]),
    "MoL_malloc_non_y_n_gfs": (None, [C.POINTER(paramstruct), C.POINTER(MoL_gridfunctions_struct)]),
    "MoL_free_memory_y_n_gfs": (None, [C.POINTER(paramstruct), C.POINTER(MoL_gridfunctions_struct)]),
    "MoL_free_memory_non_y_n_gfs": (None, [C.POINTER(paramstruct), C.POINTER(MoL_gridfunctions_struct)]),
    "rfm_precompute_rfmstruct_malloc": (None, [C.POINTER(paramstruct), C.POINTER(rfm_struct)]),
    "rfm_precompute_rfmstruct_define": (None, [C.POINTER(paramstruct), XX3, C.POINTER(rfm_struct)]),
    "rfm_precompute_rfmstruct_freemem": (None, [C.POINTER(paramstruct), C.POINTER(rfm_struct)]),
    "initial_guess_all_points": (None, [C.POINTER(paramstruct), XX3, REALP]),
    "auxevol_gfs_all_points": (None, [C.POINTER(paramstruct), XX3, REALP]),
    "wavespeed_gf_all_points": (None, [C.POINTER(paramstruct), REAL, REAL, XX3, REALP]),
    "compute_wavespeed_at_OB": (REAL, [C.POINTER(paramstruct), REALP]),
    "rhs_eval": (None, [C.POINTER(paramstruct), C.POINTER(rfm_struct), REALP, REALP, REALP]),
    "apply_bcs_outerradiation_and_inner": (None, [C.POINTER(paramstruct), C.POINTER(bc_struct), XX3,
                                                  C.POINTER(REAL * 2), C.POINTER(REAL * 2), REALP, REALP]),
    "apply_bcs_inner_only": (None, [C.POINTER(paramstruct), C.POINTER(bc_struct), REALP]),
    "MoL_step_forward_in_time": (None, [C.POINTER(griddata_struct), REAL]),
    "residual_all_points": (None, [C.POINTER(paramstruct), C.POINTER(rfm_struct), REALP, REALP, REALP]),
    "L2_norm_of_gf": (REAL, [C.POINTER(paramstruct), C.c_int, REAL, XX3, REALP]),
}


def bind(lib):
    """Attach argtypes/restype for every reference-surface symbol the library exports."""
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return lib
