"""ctypes binding of libfluid_gfx950.so (C ABI in include/fluid_gfx950.h).

The library is loaded after torch so both share torch's HIP runtime.  There
is no CPU fallback: a missing or unloadable library raises.
"""

import ctypes
from pathlib import Path

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

LIB_PATH = Path(__file__).resolve().parent / "libfluid_gfx950.so"

F64, F32 = 0, 1
MODE_RESIDUAL, MODE_RHS = 0, 1
TERM_EULER, TERM_VISCOUS = 1, 2
BND_WALL, BND_INLET, BND_NOSLIP, BND_SLIP = 1, 2, 4, 8


class Params(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("gas_constant", ctypes.c_double),
                ("dynamic_viscosity", ctypes.c_double),
                ("thermal_conductivity", ctypes.c_double), ("br2_penalty", ctypes.c_double),
                ("inviscid_flux", ctypes.c_int32), ("nvars", ctypes.c_int32),
                ("degree", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("use_wale", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("wale_constant", ctypes.c_double), ("turbulent_prandtl", ctypes.c_double)]


class BlockDesc(ctypes.Structure):
    _fields_ = [("element_type", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("count", ctypes.c_int64)] + [
        (k, ctypes.c_void_p) for k in ("Bt", "Dt", "D", "wq", "minv", "jinv", "scale", "detw",
                                       "rec_ptr", "rec_off", "rec_sup", "rec_nf", "minvT",
                                       "sf1d", "width")]


class GroupDesc(ctypes.Structure):
    _fields_ = [("face_type", ctypes.c_int32), ("owner_block", ctypes.c_int32),
                ("neighbor_block", ctypes.c_int32), ("sub", ctypes.c_int32),
                ("count", ctypes.c_int64)] + [
        (k, ctypes.c_void_p) for k in ("owner_elem", "neighbor_elem", "geom", "owner_table",
                                       "neighbor_table", "owner_tables", "neighbor_tables",
                                       "owner_sup", "neighbor_sup", "face_weights", "bflags")] + [
        ("out_owner", ctypes.c_int64), ("out_neighbor", ctypes.c_int64)] + [
        (k, ctypes.c_void_p) for k in ("owner_sf", "neighbor_sf", "face_sf1d")]


class MoodType(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in ("N", "Q", "NS", "NI", "NB", "reserved")] + [
        (k, ctypes.c_void_p) for k in ("p1proj", "samples", "wq", "P", "PT", "RT", "subvol", "int_o",
                                       "int_n", "int_area", "int_dc", "int_ptr", "int_list",
                                       "bnd_face", "bnd_gather", "tgt_ptr", "tgt_list")]


class MoodBlock(ctypes.Structure):
    _fields_ = [("max_faces", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("elem_faces", ctypes.c_void_p), ("flags", ctypes.c_void_p),
                ("sensor", ctypes.c_void_p)]


class MoodGroup(ctypes.Structure):
    _fields_ = [("K", ctypes.c_int32), ("reserved", ctypes.c_int32)] + [
        (k, ctypes.c_void_p) for k in ("pair_table", "pair_ids", "pair_geom", "basis_o",
                                       "basis_n")]


class MoodParams(ctypes.Structure):
    _fields_ = [("threshold", ctypes.c_double), ("density_floor", ctypes.c_double),
                ("pressure_floor", ctypes.c_double), ("relative_floor", ctypes.c_double),
                ("reconstruction", ctypes.c_int32), ("sensor_on", ctypes.c_int32)]


class ChemParams(ctypes.Structure):
    _fields_ = [("newton_iters", ctypes.c_int32), ("dimensionless", ctypes.c_int32),
                ("gamma", ctypes.c_double), ("T_ref", ctypes.c_double),
                ("rho_ref", ctypes.c_double), ("e_ref", ctypes.c_double),
                ("t_ref", ctypes.c_double)]


class DtBlock(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("length", "sliver", "detw", "node_weights")]


class DtConn(ctypes.Structure):
    _fields_ = [("owner_block", ctypes.c_int32), ("neighbor_block", ctypes.c_int32),
                ("count", ctypes.c_int64), ("owner_cells", ctypes.c_void_p),
                ("neighbor_cells", ctypes.c_void_p)]


class StepState(ctypes.Structure):
    """fsg_step_state of include/fluid_gfx950.h (held in device memory)."""
    _fields_ = [(k, ctypes.c_double) for k in ("cfl_dt", "retry", "time", "target_time", "dt",
                                               "factor", "min_dt", "last_rho_min",
                                               "last_p_min")] + [
        (k, ctypes.c_int64) for k in ("step", "maximum_steps", "attempts", "rejections")] + [
        (k, ctypes.c_int32) for k in ("consecutive", "maximum_rejections", "status", "active",
                                      "accept", "have_dt", "refresh_interval", "reserved")]


class XlaRhsConfig(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("boundary", ctypes.c_void_p), ("mode", ctypes.c_int32),
                ("terms", ctypes.c_int32), ("nblocks", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class Boundary(ctypes.Structure):
    _fields_ = [("has_boundary", ctypes.c_int32), ("has_wall_temperature", ctypes.c_int32),
                ("wall_temperature", ctypes.c_double), ("farfield", ctypes.c_double * 16)]


_lib = None


# FSG_ABI_VERSION of include/fluid_gfx950.h that these structures mirror
ABI_VERSION = 2


def lib():
    """The loaded library; raises when it has not been built or was built
    against another ABI version."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with jax-fluid-sim_amd/build.py")
        L = ctypes.CDLL(str(LIB_PATH))
        vp, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        L.fsg_create.argtypes = [ctypes.POINTER(Params), i32, ctypes.POINTER(BlockDesc), i32,
                                 ctypes.POINTER(GroupDesc), vp, vp, ctypes.POINTER(vp)]
        L.fsg_create.restype = ctypes.c_int
        L.fsg_residual.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), i32, i32,
                                   ctypes.POINTER(Boundary), vp]
        L.fsg_residual.restype = ctypes.c_int
        L.fsg_residual_masked.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), i32, i32,
                                          ctypes.POINTER(vp), ctypes.POINTER(Boundary), vp]
        L.fsg_residual_masked.restype = ctypes.c_int
        L.fsg_rk_stage.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp),
                                   f64, ctypes.POINTER(vp), f64, f64, i32,
                                   ctypes.POINTER(Boundary), vp, vp]
        L.fsg_rk_stage.restype = ctypes.c_int
        L.fsg_diag_reset.argtypes = [vp, vp]
        L.fsg_diag_reset.restype = ctypes.c_int
        L.fsg_timing.argtypes = [vp, i32]
        L.fsg_timing.restype = ctypes.c_int
        L.fsg_timing_read.argtypes = [vp, vp, vp, vp, vp]
        L.fsg_timing_read.restype = ctypes.c_int
        L.fsg_destroy.argtypes = [vp]
        L.fsg_destroy.restype = None
        L.fsg_last_error.argtypes = []
        L.fsg_last_error.restype = ctypes.c_char_p
        L.fsg_abi_version.restype = ctypes.c_int32
        if L.fsg_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI version {L.fsg_abi_version()}, these bindings "
                               f"{ABI_VERSION}: rebuild it with jax-fluid-sim_amd/build.py")
        L.fsg_chem_table_size.restype = ctypes.c_int32
        L.fsg_chem_integrate.argtypes = [vp, vp, vp, vp, f64, i32, i32, vp, vp, i64, vp]
        L.fsg_chem_integrate.restype = ctypes.c_int
        L.fsg_chem_react.argtypes = [vp, vp, i64, i64, vp, f64, ctypes.POINTER(ChemParams), vp]
        L.fsg_chem_react.restype = ctypes.c_int
        L.fsg_chem_last_error.restype = ctypes.c_char_p
        L.fsg_chem_integrate_flagged.argtypes = [vp, vp, vp, vp, ctypes.c_double, i32, i32, vp,
                                                 vp, vp, vp, vp, i64, vp]
        L.fsg_chem_integrate_flagged.restype = ctypes.c_int
        L.fsg_chem_integrate_flagged_noalloc.argtypes = [vp, vp, vp, vp, ctypes.c_double, i32,
                                                         i32, vp, vp, vp, i64, vp]
        L.fsg_chem_integrate_flagged_noalloc.restype = ctypes.c_int
        L.fsg_chem_check.argtypes = [vp, vp, vp, vp, f64, i32, vp, f64, f64, vp, vp, i64, vp]
        L.fsg_chem_check.restype = ctypes.c_int
        L.fsg_chem_prepare.argtypes = [vp, vp, i32, i64, ctypes.POINTER(ChemParams), vp, vp, vp]
        L.fsg_chem_prepare.restype = ctypes.c_int
        L.fsg_chem_finish.argtypes = [vp, i32, i64, ctypes.POINTER(ChemParams), vp, vp]
        L.fsg_chem_finish.restype = ctypes.c_int
        pp = ctypes.POINTER(vp)
        L.fsg_mfma_selftest.argtypes = [i32, vp, vp, vp, vp]
        L.fsg_set_option.argtypes = [vp, i32, i32]
        L.fsg_set_option.restype = ctypes.c_int
        L.fsg_mfma_selftest.restype = ctypes.c_int
        L.fsg_mood_create.argtypes = [vp, ctypes.POINTER(MoodParams), ctypes.POINTER(MoodType),
                                      ctypes.POINTER(MoodBlock), ctypes.POINTER(MoodGroup), vp, vp,
                                      ctypes.POINTER(vp)]
        L.fsg_mood_destroy.argtypes = [vp]
        L.fsg_mood_destroy.restype = None
        L.fsg_mood_troubled.argtypes = [vp, pp, vp]
        L.fsg_mood_positivity.argtypes = [vp, pp, pp, vp]
        L.fsg_mood_recompute.argtypes = [vp, pp, pp, pp, pp, f64, pp, i32, vp, pp, vp]
        L.fsg_mood_stage.argtypes = [vp, pp, pp, pp, pp, pp, pp, f64, pp, f64, f64, vp, vp, vp]
        L.fsg_mood_stage_fused.argtypes = [vp, pp, pp, pp, pp, pp, pp, f64, pp, f64, f64, vp, vp,
                                           vp]
        L.fsg_mood_count.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), vp]
        L.fsg_mood_modal.argtypes = [vp, pp, pp, f64, f64, vp, vp]
        L.fsg_xla_parse.argtypes = [ctypes.c_char_p, ctypes.c_size_t,
                                    ctypes.POINTER(XlaRhsConfig)]
        L.fsg_xla_parse.restype = ctypes.c_int
        L.fsg_xla_rhs.argtypes = [vp, pp, ctypes.c_char_p, ctypes.c_size_t]
        L.fsg_xla_rhs.restype = None
        L.fsg_mood_modal.restype = ctypes.c_int
        for name in ("fsg_mood_create", "fsg_mood_troubled", "fsg_mood_positivity",
                     "fsg_mood_recompute", "fsg_mood_stage", "fsg_mood_stage_fused",
                     "fsg_mood_count"):
            getattr(L, name).restype = ctypes.c_int
        pp = ctypes.POINTER(vp)
        L.fsg_dt_work_size.argtypes = [vp, i32, ctypes.POINTER(DtConn)]
        L.fsg_dt_work_size.restype = ctypes.c_int64
        L.fsg_dt_create.argtypes = [vp, ctypes.POINTER(DtBlock), i32, ctypes.POINTER(DtConn), vp,
                                    i64, ctypes.POINTER(vp)]
        L.fsg_dt_destroy.argtypes = [vp]
        L.fsg_dt_destroy.restype = None
        L.fsg_stable_dt.argtypes = [vp, pp, f64, vp, vp, vp]
        L.fsg_local_dt.argtypes = [vp, pp, f64, pp, pp, vp, vp]
        L.fsg_conserved_totals.argtypes = [vp, pp, vp, vp]
        L.fsg_step_begin.argtypes = [vp, vp, pp, pp, i32, pp, f64, vp]
        L.fsg_step_commit.argtypes = [vp, vp, vp, f64, f64, pp, pp, vp]
        for name in ("fsg_dt_create", "fsg_stable_dt", "fsg_local_dt", "fsg_conserved_totals",
                     "fsg_step_begin", "fsg_step_commit"):
            getattr(L, name).restype = ctypes.c_int
        L.fsg_dt_last_error.restype = ctypes.c_char_p
        f32 = ctypes.c_float
        L.fsg_mlp_linear_forward.argtypes = [vp, vp, vp, vp, vp, i32, i32, i32, i32, vp]
        L.fsg_mlp_linear_backward.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32,
                                              vp]
        L.fsg_mlp_backward_work_size.argtypes = [i32, i32, i32]
        L.fsg_mlp_backward_work_size.restype = ctypes.c_int64
        L.fsg_mlp_mse_grad.argtypes = [vp, vp, vp, i64, f32, vp]
        L.fsg_mlp_adam.argtypes = [vp, vp, vp, vp, i64, f32, f32, f32, f32, i32, vp]
        L.fsg_mlp_infer.argtypes = [vp, vp, vp, vp, vp, vp, f64, vp, vp, i64, i32, vp]
        for name in ("fsg_mlp_linear_forward", "fsg_mlp_linear_backward", "fsg_mlp_mse_grad",
                     "fsg_mlp_adam", "fsg_mlp_infer"):
            getattr(L, name).restype = ctypes.c_int
        L.fsg_mlp_packed_weight_count.restype = ctypes.c_int32
        L.fsg_mlp_consts_size.restype = ctypes.c_int32
        L.fsg_mlp_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RuntimeError(f"libfluid_gfx950 error {rc}: {lib().fsg_last_error().decode()}")


def check_dt(rc):
    if rc != 0:
        raise RuntimeError(f"libfluid_gfx950 time-step error {rc}: "
                           f"{lib().fsg_dt_last_error().decode()}")


def check_chem(rc):
    if rc != 0:
        raise RuntimeError(f"libfluid_gfx950 chemistry error {rc}: "
                           f"{lib().fsg_chem_last_error().decode()}")


def check_mlp(rc):
    if rc != 0:
        raise RuntimeError(f"libfluid_gfx950 surrogate error {rc}: "
                           f"{lib().fsg_mlp_last_error().decode()}")


def ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def ptr_array(tensors):
    arr = (ctypes.c_void_p * len(tensors))()
    for i, t in enumerate(tensors):
        arr[i] = 0 if t is None else t.data_ptr()
    return arr


def stream_handle(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
