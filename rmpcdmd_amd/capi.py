"""ctypes binding of librmpcd_amd.so (the C ABI declared in include/rmpcd_capi.h).

Loading fails loudly when the library has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# RMPCD_AMD_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("RMPCD_AMD_LIB") or os.path.join(HERE, "librmpcd_amd.so")

RMPCD_OK = 0
RMPCD_ERR_ARG = -1
RMPCD_ERR_HIP = -2
RMPCD_ERR_OVERFLOW = -3
RMPCD_ERR_DOMAIN = -4
FP64, FP32 = 0, 1
REAC_BIT, MD_BIT, WALL_BIT, PAST_MD_BIT, CATALYZED_BIT = 1, 2, 3, 4, 6
PERIODIC_BC, BOUNCE_BACK_BC, SPECULAR_BC = 1, 2, 3


class RmpcdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code
        self.msg = msg


class Box(C.Structure):
    _fields_ = [("L", C.c_int32 * 3), ("M", C.c_int32 * 3), ("ncells", C.c_int32),
                ("has_walls", C.c_int32), ("a", C.c_double), ("edges", C.c_double * 3),
                ("origin", C.c_double * 3), ("bc", C.c_int32 * 3), ("max_v", C.c_double)]


class Particles(C.Structure):
    _fields_ = [("n", C.c_int64), ("form", C.c_int32), ("pad_", C.c_int32),
                ("pos", C.c_void_p), ("vel", C.c_void_p), ("force", C.c_void_p),
                ("force_old", C.c_void_p), ("image", C.c_void_p), ("id", C.c_void_p),
                ("species", C.c_void_p), ("flags", C.c_void_p)]


class Cells(C.Structure):
    _fields_ = [("cell_count", C.c_void_p), ("cell_start", C.c_void_p),
                ("is_md", C.c_void_p), ("is_reac", C.c_void_p)]


class LJ(C.Structure):
    _fields_ = [("n1", C.c_int32), ("n2", C.c_int32), ("eps", C.c_double * 16),
                ("sigma", C.c_double * 16), ("cut", C.c_double * 16),
                ("cut_sq", C.c_double * 16), ("cut_energy", C.c_double * 16)]


SIM_REFERENCE, SIM_FUSED = 0, 1


class DimerConfig(C.Structure):
    _fields_ = [("L", C.c_int32 * 3), ("rho", C.c_int32), ("n_md", C.c_int32), ("form", C.c_int32),
                ("mode", C.c_int32), ("bulk_rmpcd", C.c_int32), ("T", C.c_double), ("tau", C.c_double),
                ("sigma_C", C.c_double), ("sigma_N", C.c_double), ("d", C.c_double),
                ("epsilon_C", C.c_double * 2), ("epsilon_N", C.c_double * 2),
                ("epsilon_CC", C.c_double), ("epsilon_NC", C.c_double), ("epsilon_NN", C.c_double),
                ("prob", C.c_double), ("bulk_rate", C.c_double), ("skin", C.c_double),
                ("nmax_list", C.c_int32), ("n_replicas", C.c_int32), ("seed", C.c_uint64),
                ("replica_L", C.c_void_p), ("stream_blocks", C.c_int32), ("sort_blocks", C.c_int32),
                ("lazy", C.c_int32), ("extra_sort_blocks", C.c_int32), ("sort_order", C.c_int32),
                ("groups", C.c_int32), ("collide_cells", C.c_int32)]


class DimerStats(C.Structure):
    _fields_ = [("n_solvent", C.c_int64), ("md_steps", C.c_int64), ("cycles", C.c_int64),
                ("n_extra_sorting", C.c_int64), ("n_replicas", C.c_int64),
                ("n_solvent_total", C.c_int64), ("e1", C.c_double), ("e2", C.c_double),
                ("dt", C.c_double), ("mass", C.c_double * 2), ("nmax", C.c_int32), ("error", C.c_int32),
                ("list", C.c_void_p), ("nlist", C.c_void_p), ("pos_old", C.c_void_p),
                ("colloid_pos_old", C.c_void_p)]


_P = C.c_void_p
_D = C.c_double
_I = C.c_int32
_I64 = C.c_int64
_U64 = C.c_uint64
_BOXP = C.POINTER(Box)
_PARTP = C.POINTER(Particles)
_CELLP = C.POINTER(Cells)
_LJP = C.POINTER(LJ)
PROG_FLUID, PROG_SPHERE, PROG_COLLOIDS = 0, 1, 2


class RigidBody(C.Structure):
    """rmpcd_rigid_body_t: rigid_body_t (md.f90:39-57), zero-based inclusive range;
    lives in device memory, pos_body is a device pointer"""
    _fields_ = [("i_start", C.c_int32), ("i_stop", C.c_int32), ("mass", C.c_double),
                ("pos", C.c_double * 3), ("vel", C.c_double * 3), ("force", C.c_double * 3),
                ("q", C.c_double * 4), ("L", C.c_double * 3), ("torque", C.c_double * 3),
                ("I_body", C.c_double * 3), ("omega_body", C.c_double * 3), ("pos_body", C.c_void_p)]


class ProgramConfig(C.Structure):
    _fields_ = [("kind", C.c_int32), ("L", C.c_int32 * 3), ("rho", C.c_int32), ("n_md", C.c_int32),
                ("form", C.c_int32), ("n_colloids", C.c_int32), ("nmax_list", C.c_int32), ("n_thermo", C.c_int32),
                ("T", C.c_double), ("tau", C.c_double), ("sigma", C.c_double), ("epsilon", C.c_double),
                ("epsilon_colloids", C.c_double), ("prob", C.c_double), ("bulk_rate", C.c_double),
                ("skin", C.c_double), ("seed", C.c_uint64), ("T_final", C.c_double), ("n_loop", C.c_int32),
                ("thermostat", C.c_int32), ("hydro", C.c_int32), ("graph", C.c_int32)]


class ProgramStats(C.Structure):
    _fields_ = [("n_solvent", C.c_int64), ("n_colloids", C.c_int64), ("md_steps", C.c_int64),
                ("cycles", C.c_int64), ("n_extra_sorting", C.c_int64), ("e1", C.c_double), ("e2", C.c_double),
                ("temperature", C.c_double), ("dt", C.c_double), ("mass", C.c_double), ("nmax", C.c_int32),
                ("last_sort", C.c_int32), ("list", C.c_void_p), ("nlist", C.c_void_p), ("pos_old", C.c_void_p),
                ("colloid_pos_old", C.c_void_p)]


_D3 = C.POINTER(C.c_double)

# name -> argtypes; every function returns int status unless listed in _RESTYPE
SIGNATURES = {
    "rmpcd_last_error": [],
    "rmpcd_init": [_I],
    "rmpcd_set_stream": [_P],
    "rmpcd_synchronize": [],
    "rmpcd_version": [],
    "rmpcd_malloc": [C.POINTER(C.c_void_p), C.c_size_t],
    "rmpcd_free": [_P],
    "rmpcd_memcpy": [_P, _P, C.c_size_t, _I],
    "rmpcd_compact_p_to_h": [C.POINTER(_I), C.POINTER(_I)],
    "rmpcd_compact_h_to_p": [_I, C.POINTER(_I), C.POINTER(_I)],
    "rmpcd_hilbert_table": [_BOXP, _P],
    "rmpcd_box_init": [_BOXP, C.POINTER(_I), _D, _I],
    "rmpcd_random_shift": [_BOXP, _D3],
    "rmpcd_count_particles": [_BOXP, _PARTP, _CELLP],
    "rmpcd_cartesian_indices": [_BOXP, _PARTP, _P],
    "rmpcd_sort": [_BOXP, _PARTP, _PARTP, _P, _CELLP, _I],
    "rmpcd_maximum_displacement": [_PARTP, _P, _D3],
    "rmpcd_cell_maximum_displacement": [_BOXP, _CELLP, _PARTP, _P, _D, _D3],
    "rmpcd_rand_sphere": [_P, _I64, _I64, _P, _P],
    "rmpcd_simple_mpcd_step": [_BOXP, _CELLP, _PARTP, _P, _U64, _U64, _I, _D, _I, _D, _I],
    "rmpcd_compute_temperature": [_BOXP, _CELLP, _PARTP, _D3],
    "rmpcd_compute_temperature_noalloc": [_BOXP, _CELLP, _PARTP, _P],
    "rmpcd_compute_temperature_profile": [_BOXP, _CELLP, _PARTP, _D3, C.c_double, C.c_double, _I, _P, _P],
    "rmpcd_mpcd_stream_periodic": [_BOXP, _PARTP, _D],
    "rmpcd_bulk_reaction": [_BOXP, _CELLP, _PARTP, _I, _I, _D, _D, _P, _U64, _U64],
    "rmpcd_mpcat_step": [_BOXP, _CELLP, _PARTP, _P, _U64, _U64, _D, _I],
    "rmpcd_wall_mpcd_step": [_BOXP, _CELLP, _PARTP, _I, _P, _P, _I, _D, _I, _I, _D, _P, _P, _P, _U64, _U64],
    "rmpcd_mpcd_stream_xforce_yzwall": [_BOXP, _PARTP, _D, _D],
    "rmpcd_rescale_cells": [_BOXP, _CELLP, _PARTP, _D],
    "rmpcd_bulk_reaction_endothermic": [_BOXP, _CELLP, _PARTP, _I, _I, _D, _D, _D, _P, _U64, _U64],
    "rmpcd_mpcd_stream_nogravity_zwall": [_BOXP, _PARTP, _D],
    "rmpcd_md_pos": [_PARTP, _D],
    "rmpcd_md_vel": [_PARTP, _D],
    "rmpcd_apply_pbc": [_PARTP, _D3],
    "rmpcd_cell_md_pos": [_BOXP, _CELLP, _PARTP, _D, _I],
    "rmpcd_cell_md_vel": [_BOXP, _CELLP, _PARTP, _D, _I],
    "rmpcd_cell_md_pos_wall": [_BOXP, _CELLP, _PARTP, _D, _I, _I],
    "rmpcd_rattle_dimer_pos": [_PARTP, _P, _D3, _I, _D, _D, _D3],
    "rmpcd_rattle_dimer_vel": [_PARTP, _D3, _I, _D, _D, _D3],
    "rmpcd_rattle_body_pos": [_PARTP, _P, _P, _P, _I, _D3, _I, _D, _D3, _D, C.POINTER(_I)],
    "rmpcd_rattle_body_vel": [_PARTP, _P, _P, _I, _D3, _I, _D, _D3, _D, C.POINTER(_I)],
    "rmpcd_elastic_network": [_PARTP, _P, _P, _I, _D, _D3, _D3],
    "rmpcd_rigid_body_init": [_P, _PARTP, _D3, _I, _D3, _I, _I, _P],
    "rmpcd_rigid_body_compute_force_torque": [_P, _PARTP],
    "rmpcd_rigid_body_vv1": [_P, _PARTP, _D, _D, _D3, C.POINTER(_I)],
    "rmpcd_rigid_body_vv2": [_P, _PARTP, _D],
    "rmpcd_colloid_md_pos": [_PARTP, _P, _D],
    "rmpcd_colloid_md_vel": [_PARTP, _P, _D],
    "rmpcd_lj_params_init": [_LJP, _I, _I, _D3, _D3, _D3],
    "rmpcd_make_stencil": [_D, _D, C.POINTER(_I), _I],
    "rmpcd_update_list": [_BOXP, _CELLP, _PARTP, _PARTP, _D, _P, _I, _LJP, _I, _P, _P],
    "rmpcd_compute_force": [_PARTP, _PARTP, _P, _P, _I, _D3, _LJP, _D3],
    "rmpcd_compute_force_n2": [_PARTP, _D3, _LJP, _D3],
    "rmpcd_flag_particles_nl": [_PARTP, _PARTP, _P, _P, _I, _D3, _D],
    "rmpcd_change_species": [_PARTP, _PARTP, _D3, _D, _D, _D, _P, _U64, _U64],
    "rmpcd_refuel": [_PARTP, _PARTP, _D3, _I],
    "rmpcd_flag_particles_nl_prob": [_PARTP, _PARTP, _P, _P, _I, _D3, _D, _D, _P, _U64, _U64],
    "rmpcd_change_species_sphere": [_PARTP, _PARTP, _D3, _D],
    "rmpcd_program_create": [C.POINTER(ProgramConfig), C.POINTER(C.c_void_p)],
    "rmpcd_program_run": [_P, _I],
    "rmpcd_program_state": [_P, _PARTP, _PARTP, _BOXP, _CELLP, C.POINTER(ProgramStats)],
    "rmpcd_program_destroy": [_P],
    "rmpcd_dimer_create": [C.POINTER(DimerConfig), C.POINTER(C.c_void_p)],
    "rmpcd_dimer_run": [_P, _I],
    "rmpcd_dimer_state": [_P, _I, _PARTP, _PARTP, _BOXP, _CELLP, C.POINTER(DimerStats)],
    "rmpcd_dimer_destroy": [_P],
    "rmpcd_dimer_profile": [_P, _I],
    "rmpcd_dimer_stream_time": [_P, _D3, C.POINTER(_I)],
    "rmpcd_threefry_uniform": [_U64, _U64, _U64, _I64, _P],
    "rmpcd_threefry2x64_20": [C.POINTER(_U64), C.POINTER(_U64), C.POINTER(_U64)],
}
_RESTYPE = {"rmpcd_last_error": C.c_char_p, "rmpcd_version": C.c_char_p,
            "rmpcd_compact_p_to_h": _I, "rmpcd_compact_h_to_p": None,
            "rmpcd_make_stencil": _I, "rmpcd_threefry2x64_20": None}

_lib = None


def lib():
    """the loaded library; raises if it is not built (no fallback exists)"""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("librmpcd_amd.so is not built: run __graft_entry__.build() or "
                               "python -m rmpcdmd_amd.build")
        lib_ = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib_, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = lib_
    return _lib


def call(name, *args):
    """call an entry point that returns a status; raise RmpcdError on failure"""
    rc = getattr(lib(), name)(*args)
    if rc != RMPCD_OK:
        raise RmpcdError(rc, lib().rmpcd_last_error().decode())
    return rc


def d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def i3(v):
    return (C.c_int32 * 3)(*[int(x) for x in v])
