"""ctypes bindings of ``include/uvlm_gpu.h`` and ``include/uvlm_host.h``.

The libraries are built in-tree (``uvlm_amd/lib``) by ``uvlm_amd.build``.
Loading fails loudly when they are missing: there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
GPU_LIB = os.path.join(LIBDIR, "libuvlm_gpu.so")
HOST_LIB = os.path.join(LIBDIR, "libuvlm.so")

# include/uvlm_gpu.h
OK, ERR_ARG, ERR_HIP, ERR_OOM, ERR_SINGULAR, ERR_STATE, ERR_CAPACITY, ERR_NOCONV = (
    0, -1, -2, -3, -4, -5, -6, -7)
PREC_F64_EXACT, PREC_F64, PREC_F32C = 0x0, 0x1, 0x2
SCALE_PER_RING, CHORDWISE_ONLY, SEQUENTIAL = 0x10, 0x20, 0x100
SOLVE_F64, SOLVE_F32_IR, SOLVE_BF16_IR, SOLVE_F64_FULLPIV, SOLVE_F64_INV = 0, 1, 2, 3, 4
ADVECT_EULER, ADVECT_RK2, ADVECT_RK4, ADVECT_AB2, ADVECT_RK4_CLASSIC = 0, 1, 2, 3, 4
BATCH_FORCE_PIVOTED = 1
MLP_IN, MLP_OUT, MLP_PARAMS, MLP_FP32, MLP_BF16 = 32, 30, 41120, 0, 1
SCHEMES = {"euler": ADVECT_EULER, "RK2": ADVECT_RK2, "RK4": ADVECT_RK4,
           "AB2": ADVECT_AB2, "RK4classic": ADVECT_RK4_CLASSIC}
GEN = {"rect": 0, "naca4": 1, "naca00": 2}
MAX_WINGS = 16

# Every exported symbol of the two headers (checked by the CPU test suite).
GPU_SYMBOLS = [
    "uvlm_create", "uvlm_destroy", "uvlm_last_error", "uvlm_sync", "uvlm_set_cutoff", "uvlm_source_cache_stats",
    "uvlm_biot_savart", "uvlm_induced_velocity", "uvlm_induced_velocity_noalloc", "uvlm_ring_induced_velocity",
    "uvlm_influence_matrix", "uvlm_ring_influence_matrix", "uvlm_solve",
    "uvlm_sim_init", "uvlm_sim_set_wing",
    "uvlm_sim_shed", "uvlm_sim_solve", "uvlm_sim_loads", "uvlm_sim_advect",
    "uvlm_sim_erase_oldest_wake", "uvlm_sim_wake_shape", "uvlm_sim_get_wake",
    "uvlm_sim_set_wake", "uvlm_sim_get_gamma", "uvlm_sim_get_loads",
    "uvlm_sim_pair_counts", "uvlm_profile_enable", "uvlm_profile_read",
    "uvlm_profile_read_near", "uvlm_batch_check_plan",
    "uvlm_farfield_velocity", "uvlm_sim_set_farfield", "uvlm_sim_set_farfield_focus",
    "uvlm_sim_farfield_stats",
    "uvlm_sim_inverse_stats", "uvlm_sim_refine_stats",
    "uvlm_mlp_forward", "uvlm_mlp_grad", "uvlm_mlp_train", "uvlm_farfield_set_mlp",
    "uvlm_farfield_mlp_stats", "uvlm_farfield_samples", "uvlm_sim_farfield_samples",
    "uvlm_farfield_target_order",
    "uvlm_sim_farfield_learn",
    "uvlm_batch_init", "uvlm_batch_run", "uvlm_batch_autotune", "uvlm_batch_steps",
    "uvlm_batch_get_gamma", "uvlm_batch_get_wing", "uvlm_batch_get_wake",
    "uvlm_batch_get_loads", "uvlm_batch_pairs", "uvlm_batch_set_option", "uvlm_batch_free",
]
HOST_SYMBOLS = [
    "uvlm_host_sim_create", "uvlm_host_sim_destroy", "uvlm_host_last_error",
    "uvlm_host_sim_step", "uvlm_host_sim_sync", "uvlm_host_sim_wing_nodes",
    "uvlm_host_sim_gamma", "uvlm_host_sim_wake_shape", "uvlm_host_sim_wake",
    "uvlm_host_sim_set_wake", "uvlm_host_sim_loads", "uvlm_host_sim_pairs",
    "uvlm_host_sim_context", "uvlm_host_build_wing", "uvlm_host_legacy_run",
    "uvlm_host_sim_set_outputs", "uvlm_host_legacy_set_outputs", "uvlm_host_velocity_at",
    "uvlm_host_snapshot_velocity", "uvlm_host_plane_velocity", "uvlm_host_grid_velocity",
    "uvlm_host_calc_impulse", "uvlm_host_lamb_vector", "uvlm_host_encode_panels",
    "uvlm_host_encode_sheets",
]


class WingSpec(C.Structure):
    _fields_ = [
        ("generator", C.c_int), ("digit", C.c_int),
        ("chord", C.c_double), ("span", C.c_double),
        ("origin", C.c_double * 3), ("alpha", C.c_double),
        ("plug_amp", C.c_double), ("plug_omega", C.c_double), ("plug_phase", C.c_double),
        ("flap_amp", C.c_double), ("flap_omega", C.c_double), ("flap_phase", C.c_double),
        ("fourier_omega", C.c_double),
        ("twist_amp", (C.c_double * 2) * 4), ("twist_phase", (C.c_double * 2) * 4),
        ("bend_amp", (C.c_double * 2) * 4), ("bend_phase", (C.c_double * 2) * 4),
    ]


class CaseSpec(C.Structure):
    _fields_ = [
        ("num_wings", C.c_int), ("wings", WingSpec * MAX_WINGS),
        ("rows", C.c_size_t), ("cols", C.c_size_t),
        ("ff", C.c_double * 3), ("dt", C.c_double), ("scheme", C.c_int),
        ("ignore_steps", C.c_size_t), ("erase_wake_after", C.c_size_t),
        ("loads", C.c_int), ("precision", C.c_uint), ("solve_mode", C.c_int),
        ("cutoff", C.c_double),
    ]


class BatchCase(C.Structure):
    """uvlm_batch_case (include/uvlm_gpu.h)."""
    _fields_ = [
        ("rows", C.c_size_t), ("cols", C.c_size_t),
        ("chord", C.c_double), ("span", C.c_double),
        ("dt", C.c_double), ("alpha", C.c_double),
        ("origin", C.c_double * 3), ("ff", C.c_double * 3),
        ("plug", C.c_double * 3), ("flap", C.c_double * 3),
    ]


_gpu = None
_host = None
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p
_sz = C.c_size_t


class MissingLibrary(RuntimeError):
    pass


def _load(path):
    if not os.path.exists(path):
        raise MissingLibrary(
            f"{path} is not built; run `python -m uvlm_amd.build` (or "
            "__graft_entry__.build()) first -- the UVLM hot path has no CPU fallback")
    return C.CDLL(path)


def gpu_lib():
    """The gfx950 library (include/uvlm_gpu.h)."""
    global _gpu
    if _gpu is None:
        L = _load(GPU_LIB)
        L.uvlm_create.argtypes = [C.c_int, C.POINTER(_vp)]
        L.uvlm_destroy.argtypes = [_vp]
        L.uvlm_last_error.argtypes = [_vp]
        L.uvlm_last_error.restype = C.c_char_p
        L.uvlm_sync.argtypes = [_vp]
        L.uvlm_set_cutoff.argtypes = [_vp, C.c_double]
        L.uvlm_source_cache_stats.argtypes = [_vp, C.POINTER(C.c_longlong),
                                              C.POINTER(C.c_longlong)]
        L.uvlm_induced_velocity_noalloc.argtypes = [_vp, _vp, _sz, _vp, _vp, _sz, _sz, _sz,
                                                    C.c_uint, _vp, _vp]
        L.uvlm_biot_savart.argtypes = [_vp, _dp, _dp, _dp, _sz, C.c_uint, _dp]
        L.uvlm_induced_velocity.argtypes = [_vp, _dp, _sz, _dp, _dp, _sz, _sz, _sz,
                                            C.c_uint, _dp]
        L.uvlm_ring_induced_velocity.argtypes = [_vp, _dp, _sz, _dp, _dp, _sz, C.c_uint, _dp]
        L.uvlm_influence_matrix.argtypes = [_vp, _dp, _sz, _sz, _sz, _dp, _dp, C.c_uint, _dp]
        L.uvlm_ring_influence_matrix.argtypes = [_vp, _dp, _sz, _dp, _dp, _dp]
        L.uvlm_solve.argtypes = [_vp, _sz, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int,
                                 C.POINTER(C.c_int)]
        L.uvlm_sim_init.argtypes = [_vp, _sz, _sz, _sz, _sz, C.c_uint]
        L.uvlm_sim_set_wing.argtypes = [_vp, _dp]
        L.uvlm_sim_shed.argtypes = [_vp, _sz]
        L.uvlm_sim_solve.argtypes = [_vp, _dp, _dp, C.c_int, _dp]
        L.uvlm_sim_loads.argtypes = [_vp, _dp, _dp, C.c_double, _dp]
        L.uvlm_sim_advect.argtypes = [_vp, C.c_int, _sz, _dp, C.c_double]
        L.uvlm_sim_erase_oldest_wake.argtypes = [_vp]
        L.uvlm_sim_wake_shape.argtypes = [_vp, C.POINTER(_sz), C.POINTER(_sz)]
        L.uvlm_sim_get_wake.argtypes = [_vp, _dp, _dp]
        L.uvlm_sim_set_wake.argtypes = [_vp, _sz, _sz, _dp, _dp, _dp]
        L.uvlm_sim_get_gamma.argtypes = [_vp, _dp]
        L.uvlm_sim_get_loads.argtypes = [_vp, _dp, C.POINTER(_sz)]
        L.uvlm_sim_pair_counts.argtypes = [_vp, _dp, _dp]
        L.uvlm_profile_enable.argtypes = [_vp, C.c_int]
        L.uvlm_profile_read.argtypes = [_vp, C.POINTER(_sz), _dp, _dp]
        L.uvlm_profile_read_near.argtypes = [_vp, C.POINTER(_sz), _dp, _dp]
        L.uvlm_farfield_velocity.argtypes = [_vp, _dp, _sz, _dp, _dp, _sz, _sz, _sz, C.c_double,
                                             _dp]
        L.uvlm_sim_set_farfield.argtypes = [_vp, C.c_double, C.c_int, C.c_double]
        L.uvlm_sim_set_farfield_focus.argtypes = [_vp, C.c_double, C.c_double]
        L.uvlm_sim_farfield_stats.argtypes = [_vp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        _ll = C.POINTER(C.c_longlong)
        L.uvlm_sim_inverse_stats.argtypes = [_vp, _ll, _ll, _ll, C.POINTER(C.c_int)]
        L.uvlm_sim_refine_stats.argtypes = [_vp, _ll, _ll, C.POINTER(C.c_int)]
        L.uvlm_mlp_forward.argtypes = [_vp, _fp, _fp, _sz, _fp, C.c_int]
        L.uvlm_mlp_grad.argtypes = [_vp, _fp, _fp, _fp, _sz, _fp, _fp]
        L.uvlm_mlp_train.argtypes = [_vp, _fp, _fp, _fp, _sz, _sz, _sz, C.c_float, C.c_uint,
                                     _fp]
        L.uvlm_farfield_set_mlp.argtypes = [_vp, _fp, C.c_double]
        _ull = C.POINTER(C.c_ulonglong)
        L.uvlm_farfield_mlp_stats.argtypes = [_vp, _ull, _ull]
        L.uvlm_farfield_target_order.argtypes = [_vp, _dp, _sz, C.POINTER(C.c_uint),
                                                 C.POINTER(C.c_int)]
        L.uvlm_farfield_samples.argtypes = [_vp, _dp, _sz, _dp, _dp, _sz, _sz, _sz, C.c_double,
                                            C.c_uint, _fp, _fp, C.POINTER(_sz)]
        L.uvlm_sim_farfield_samples.argtypes = [_vp, C.c_double, C.c_uint, _fp, _fp,
                                                C.POINTER(_sz)]
        L.uvlm_sim_farfield_learn.argtypes = [_vp, C.c_double, _sz, _sz, _sz, C.c_float,
                                              C.c_uint, _fp, _fp]
        L.uvlm_batch_init.argtypes = [_vp, C.POINTER(BatchCase), _sz, _dp, _sz, C.c_int, _sz]
        L.uvlm_batch_check_plan.argtypes = [C.POINTER(BatchCase), _sz, _sz, C.c_char_p, _sz]
        L.uvlm_batch_run.argtypes = [_vp, _sz, C.c_int]
        L.uvlm_batch_autotune.argtypes = [_vp, C.POINTER(C.c_int), _dp, C.c_int]
        L.uvlm_batch_steps.argtypes = [_vp, C.POINTER(_sz)]
        L.uvlm_batch_get_gamma.argtypes = [_vp, _sz, _dp]
        L.uvlm_batch_get_wing.argtypes = [_vp, _sz, _dp]
        L.uvlm_batch_get_wake.argtypes = [_vp, _sz, _dp, _dp]
        L.uvlm_batch_get_loads.argtypes = [_vp, _sz, _dp, C.POINTER(_sz)]
        L.uvlm_batch_pairs.argtypes = [_vp, _dp]
        L.uvlm_batch_set_option.argtypes = [_vp, C.c_int, C.c_int]
        L.uvlm_batch_free.argtypes = [_vp]
        _gpu = L
    return _gpu


def host_lib():
    """The host library (include/uvlm_host.h), which links the GPU library."""
    global _host
    if _host is None:
        gpu_lib()
        L = _load(HOST_LIB)
        L.uvlm_host_sim_create.argtypes = [C.POINTER(CaseSpec), _sz, C.POINTER(_vp)]
        L.uvlm_host_sim_destroy.argtypes = [_vp]
        L.uvlm_host_last_error.restype = C.c_char_p
        L.uvlm_host_sim_step.argtypes = [_vp, _sz]
        L.uvlm_host_sim_sync.argtypes = [_vp]
        L.uvlm_host_sim_wing_nodes.argtypes = [_vp, _dp]
        L.uvlm_host_sim_gamma.argtypes = [_vp, _dp]
        L.uvlm_host_sim_wake_shape.argtypes = [_vp, C.POINTER(_sz), C.POINTER(_sz)]
        L.uvlm_host_sim_wake.argtypes = [_vp, _dp, _dp]
        L.uvlm_host_sim_set_wake.argtypes = [_vp, _sz, _sz, _dp, _dp, _dp]
        L.uvlm_host_sim_loads.argtypes = [_vp, _dp, C.POINTER(_sz)]
        L.uvlm_host_sim_pairs.argtypes = [_vp, _dp, _dp]
        L.uvlm_host_sim_context.argtypes = [_vp]
        L.uvlm_host_build_wing.argtypes = [C.POINTER(WingSpec), _sz, _sz, _dp]
        L.uvlm_host_legacy_run.argtypes = [C.POINTER(CaseSpec), _sz, _dp, _dp, _dp, _dp,
                                           C.POINTER(_sz)]
        L.uvlm_host_sim_context.restype = _vp
        _cp = C.c_char_p
        L.uvlm_host_sim_set_outputs.argtypes = [_vp, _cp, _cp]
        L.uvlm_host_legacy_set_outputs.argtypes = [_cp, _cp, C.c_int]
        L.uvlm_host_velocity_at.argtypes = [_cp, _dp, C.c_int, C.c_int, C.c_int, _cp,
                                            C.POINTER(_sz)]
        L.uvlm_host_snapshot_velocity.argtypes = [_cp, C.c_int, _sz, _dp, _sz, _dp]
        L.uvlm_host_plane_velocity.argtypes = [_cp, C.c_int, _sz, _cp, _dp, C.c_double, _sz,
                                               _cp, _cp]
        L.uvlm_host_grid_velocity.argtypes = [_cp, C.c_int, _sz, _dp, _dp, _sz, _cp,
                                              C.POINTER(_sz)]
        L.uvlm_host_calc_impulse.argtypes = [_cp, _dp, _dp, _dp, _dp]
        L.uvlm_host_lamb_vector.argtypes = [_dp, C.c_double, _dp, _dp, _dp, _dp, _sz, _dp]
        L.uvlm_host_encode_panels.argtypes = [_dp, _dp, _sz, _sz, _sz, _dp, _dp, _sz, _vp,
                                              C.POINTER(_sz)]
        L.uvlm_host_encode_sheets.argtypes = [_dp, _dp, _sz, _sz, _sz, _dp, _dp, _sz,
                                              C.c_double, _vp, C.POINTER(_sz)]
        _host = L
    return _host
