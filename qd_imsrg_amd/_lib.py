"""ctypes binding of libqdimsrg.so (the C ABI in include/qdimsrg.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libqdimsrg.so")

vp = C.c_void_p
dp = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)
ip = C.POINTER(C.c_int)
fp = C.POINTER(C.c_float)


class Info(C.Structure):
    _fields_ = [("nbody", C.c_int64), ("nshell", C.c_int64), ("msp", C.c_int64), ("nblock", C.c_int64),
                ("ncc", C.c_int64), ("neq", C.c_int64), ("cc_size", C.c_int64), ("n2b", C.c_int64)]


class FlowParams(C.Structure):
    _fields_ = [("stp", C.c_double), ("conv", C.c_double), ("max_iter", C.c_int), ("pandya", C.c_int),
                ("verbose", C.c_int), ("warm_start", C.c_int), ("diis", C.c_int), ("stop_rule", C.c_int),
                ("conv_deds", C.c_double), ("tda_ml", C.c_int), ("tda_ms", C.c_int), ("tda_cut", C.c_int),
                ("spectrum_path", C.c_char_p)]


class FlowResult(C.Structure):
    _fields_ = [("s", C.c_double), ("e0", C.c_double), ("crit", C.c_double), ("iterations", C.c_int),
                ("rejected", C.c_int), ("bch_terms", C.c_int64), ("converged", C.c_int),
                ("trace_rows", C.c_int), ("mbpt2", C.c_double), ("deds", C.c_double)]


class GsParams(C.Structure):
    _fields_ = [("relerr", C.c_double), ("abserr", C.c_double), ("stp", C.c_double), ("conv", C.c_double),
                ("max_outputs", C.c_int), ("verbose", C.c_int), ("maxnum", C.c_int)]


class GsResult(C.Structure):
    _fields_ = [("s", C.c_double), ("e0", C.c_double), ("crit", C.c_double), ("outputs", C.c_int),
                ("nsteps", C.c_int), ("nfe", C.c_int64), ("iflag", C.c_int), ("converged", C.c_int)]


# name -> (restype, argtypes); the exported surface of include/qdimsrg.h
SIGNATURES = {
    "qd_last_error": (C.c_char_p, []),
    "qd_context_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(vp)]),
    "qd_context_destroy": (C.c_int, [vp]),
    "qd_context_info": (C.c_int, [vp, C.POINTER(Info)]),
    "qd_context_table": (C.c_int64, [vp, C.c_int, i64p, C.c_int64]),
    "qd_layout_table": (C.c_int64, [C.c_int, C.c_int, C.c_int, i64p, C.c_int64]),
    "qd_context_sync": (C.c_int, [vp]),
    "qd_op_create": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
    "qd_op_destroy": (C.c_int, [vp]),
    "qd_op_set_herm": (C.c_int, [vp, C.c_int]),
    "qd_op_get_herm": (C.c_int, [vp, ip]),
    "qd_op_upload": (C.c_int, [vp, dp, C.c_int64]),
    "qd_op_download": (C.c_int, [vp, dp, C.c_int64]),
    "qd_op_device_ptr": (vp, [vp]),
    "qd_cc_create": (C.c_int, [vp, C.c_double, C.POINTER(vp)]),
    "qd_cc_destroy": (C.c_int, [vp]),
    "qd_cc_set_herm": (C.c_int, [vp, C.c_double]),
    "qd_cc_upload": (C.c_int, [vp, dp, C.c_int64]),
    "qd_cc_download": (C.c_int, [vp, dp, C.c_int64]),
    "qd_cc_zero": (C.c_int, [vp, vp]),
    "qd_commutator_110": (C.c_int, [vp, vp, vp, dp]),
    "qd_commutator_220": (C.c_int, [vp, vp, vp, dp]),
    "qd_commutator_111": (C.c_int, [vp, vp, vp, vp]),
    "qd_commutator_121": (C.c_int, [vp, vp, vp, vp]),
    "qd_commutator_221": (C.c_int, [vp, vp, vp, vp]),
    "qd_commutator_122": (C.c_int, [vp, vp, vp, vp]),
    "qd_commutator_222": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "qd_calc_cc": (C.c_int, [vp, vp, vp]),
    "qd_build_white": (C.c_int, [vp, vp, vp]),
    "qd_add_arrays": (C.c_int, [vp, vp, C.c_double, vp, C.c_double, vp]),
    "qd_copy_arrays": (C.c_int, [vp, vp, vp]),
    "qd_set_to_zero": (C.c_int, [vp, vp]),
    "qd_mat_2_norm": (C.c_int, [vp, vp, dp]),
    "qd_mbpt2": (C.c_int, [vp, vp, dp]),
    "qd_mbpt2_noalloc": (C.c_int, [vp, vp, dp]),
    "qd_commutator": (C.c_int, [vp, vp, vp, vp, vp, vp, dp]),
    "qd_commutator_cc": (C.c_int, [vp, vp, vp, vp, vp, vp, dp]),
    "qd_bch_expand": (C.c_int, [vp, vp, vp, vp, vp, vp, ip]),
    "qd_magnus_expand": (C.c_int, [vp, vp, vp, vp]),
    "qd_euler_step": (C.c_int, [vp, vp, vp, dp, C.c_double]),
    "qd_dgam": (C.c_int, [vp, vp, vp, vp, vp]),
    "qd_flow_params_default": (None, [C.POINTER(FlowParams)]),
    "qd_magnus_flow": (C.c_int, [vp, vp, vp, vp, C.POINTER(FlowParams), C.POINTER(FlowResult), dp, C.c_int]),
    "qd_timing_enable": (C.c_int, [vp, C.c_int]),
    "qd_timing_read": (C.c_int, [vp, dp, i64p, C.c_int]),
    "qd_gemm_flops": (C.c_int, [vp, dp, dp, dp]),
    "qd_set_launch_mode": (C.c_int, [vp, C.c_int]),
    "qd_set_graphs": (C.c_int, [vp, C.c_int]),
    "qd_graph_count": (C.c_int64, [vp]),
    "qd_autotune": (C.c_int, [vp, C.c_int, dp, dp]),
    "qd_set_precision": (C.c_int, [vp, C.c_int]),
    "qd_ode_create": (C.c_int, [vp, C.POINTER(vp)]),
    "qd_ode_destroy": (C.c_int, [vp]),
    "qd_ode_integrate": (C.c_int, [vp, vp, dp, C.c_double, dp, dp, ip]),
    "qd_ode_stats": (C.c_int, [vp, ip, i64p, ip, dp]),
    "qd_gs_params_default": (None, [C.POINTER(GsParams)]),
    "qd_gs_flow": (C.c_int, [vp, vp, vp, C.POINTER(GsParams), C.POINTER(GsResult), dp, C.c_int]),
    "qd_mlp_create": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(vp)]),
    "qd_mlp_destroy": (C.c_int, [vp]),
    "qd_mlp_num_params": (C.c_int64, [vp]),
    "qd_mlp_get_params": (C.c_int, [vp, fp, C.c_int64]),
    "qd_mlp_set_params": (C.c_int, [vp, fp, C.c_int64]),
    "qd_mlp_forward": (C.c_int, [vp, fp, C.c_int, fp]),
    "qd_mlp_loss_grad": (C.c_int, [vp, fp, fp, C.c_int, fp, fp]),
    "qd_mlp_train": (C.c_int, [vp, fp, fp, C.c_int, C.c_int, C.c_float, fp]),
    "qd_basis_create": (C.c_int, [vp, C.c_int, dp, dp, C.POINTER(vp)]),
    "qd_basis_destroy": (C.c_int, [vp]),
    "qd_basis_expand": (C.c_int, [vp, dp, vp]),
    "qd_get_precision": (C.c_int, [vp, ip]),
    "qd_gemm_shapes": (C.c_int, [vp, ip, C.c_int64]),
    "qd_set_gemm_shapes": (C.c_int, [vp, ip, C.c_int64]),
    "qd_gemm_mfma_flops": (C.c_int, [vp, dp, dp, dp]),
    "qd_v_int": (C.c_double, [ip, ip, ip, ip, C.c_double]),
    "qd_hf_hamiltonian": (C.c_int, [C.c_int, C.c_int, C.c_double, dp, C.c_int64, dp, ip]),
    "qd_hf_hamiltonian_me": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_char_p, dp, C.c_int64, dp, ip]),
    "qd_hf_hamiltonian_device": (C.c_int, [vp, C.c_double, vp, dp, ip]),
    "qd_me_bin_write": (C.c_int, [C.c_int, C.c_char_p, i64p]),
    "qd_tda_spectrum": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, dp, C.c_int, ip]),
    "qd_eom_dim": (C.c_int, [vp, C.c_int, C.c_int, i64p, i64p]),
    "qd_eom_matvec": (C.c_int, [vp, vp, C.c_int, C.c_int, dp, dp, C.c_int64]),
    "qd_eom_excited_states": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, dp, dp, ip]),
    "qd_eom13_dim": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, i64p, i64p]),
    "qd_eom13_matvec": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, dp, dp, C.c_int64]),
    "qd_eom_attached_removed": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, ip]),
    "qd_export_hamiltonian": (C.c_int, [vp, C.c_char_p]),
    "qd_import_hamiltonian": (C.c_int, [vp, C.c_char_p]),
    "qd_export_hamiltonian_host": (C.c_int, [C.c_int, C.c_int, dp, C.c_int64, C.c_char_p]),
    "qd_import_hamiltonian_host": (C.c_int, [C.c_int, C.c_int, dp, C.c_int64, C.c_char_p]),
}

_LIB = None


def load(path: str | None = None):
    """Load the HIP library once; raises if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = path or LIBPATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m qd_imsrg_amd.build` "
                           "(there is no CPU fallback)")
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


class QDError(RuntimeError):
    pass


def check(rc: int) -> None:
    if rc != 0:
        msg = _LIB.qd_last_error().decode() if _LIB is not None else ""
        raise QDError(f"qdimsrg error {rc}: {msg}")
