"""ctypes binding of libaem.so (declarations in include/aem.h).

The library is built in-tree by ``python -m alpineem_amd.build``.  Importing
this module never falls back to anything else: when libaem.so is missing the
load raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaem.so")

AEM_OK = 0
STATUS_NAMES = {0: "AEM_OK", 1: "AEM_ERR_INPUT", 2: "AEM_ERR_IO", 3: "AEM_ERR_DEVICE", 4: "AEM_ERR_STATE",
                5: "AEM_ERR_UNSUPPORTED", 6: "AEM_ERR_NUMERIC"}

SERIES = {"port_v": 0, "port_i": 1, "incident": 2, "e_theta": 3, "e_phi": 4, "reflected": 5,
          "transmitted": 6, "ff_pc": 7, "spice_v": 8, "spice_i": 9, "reflected_tm": 10, "transmitted_tm": 11}
FIELDS = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}
IMAG = 0x100    # AEM_IMAG: the imaginary part of a kmax simulation's series or field

# (name, restype, argtypes) for every entry point of include/aem.h
SIGNATURES = [
    ("aem_last_error", C.c_char_p, []),
    ("aem_version", C.c_char_p, []),
    ("aem_plan_from_inputs", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("aem_plan_from_text", C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    ("aem_plan_options_init", None, [C.c_void_p]),
    ("aem_plan_from_text_ex", C.c_int, [C.c_char_p, C.c_char_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("aem_plan_from_inputs_ex", C.c_int, [C.c_char_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("aem_plan_view_get", C.c_void_p, [C.c_void_p]),
    ("aem_plan_write_geometry", C.c_int, [C.c_void_p, C.c_char_p]),
    ("aem_plan_destroy", None, [C.c_void_p]),
    ("aem_sim_create", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("aem_sim_create_ex", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("aem_batch_create", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    ("aem_batch_run", C.c_int, [C.c_void_p, C.c_int64, C.c_int32]),
    ("aem_batch_last_run_ms", C.c_float, [C.c_void_p]),
    ("aem_batch_size", C.c_int32, [C.c_void_p]),
    ("aem_batch_destroy", None, [C.c_void_p]),
    ("aem_create_from_inputs", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("aem_destroy", None, [C.c_void_p]),
    ("aem_run", C.c_int, [C.c_void_p, C.c_int64, C.c_int32]),
    ("aem_steps_done", C.c_int64, [C.c_void_p]),
    ("aem_sync", C.c_int, [C.c_void_p]),
    ("aem_finalize", C.c_int, [C.c_void_p]),
    ("aem_port_exchange", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("aem_get_series", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int64]),
    ("aem_get_field", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]),
    ("aem_set_field", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]),
    ("aem_get_video", C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_int64, C.c_void_p, C.c_int64]),
    ("aem_get_ntff_raw", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]),
    ("aem_get_ff_time", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("aem_get_dft", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]),
    ("aem_get_dft_noalloc", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]),
    ("aem_objective_set", C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_double, C.c_int32]),
    ("aem_objective_get", C.c_int, [C.c_void_p, C.c_void_p]),
    ("aem_ffdft_set", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int32]),
    ("aem_ffdft_get", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("aem_adjoint_record", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("aem_adjoint_eps_gradient", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    ("aem_design_box", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("aem_set_design_eps", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("aem_adjoint_design_gradient", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("aem_adjoint", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    ("aem_adjoint_series", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    ("aem_adjoint_checkpoint", C.c_int, [C.c_void_p, C.c_int32]),
    ("aem_adjoint_step", C.c_int, [C.c_void_p, C.c_int64]),
    ("aem_adjoint_get_field", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]),
    ("aem_adjoint_set_field", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]),
    ("aem_write_outputs", C.c_int, [C.c_void_p, C.c_char_p]),
    ("aem_last_run_ms", C.c_float, [C.c_void_p]),
    ("aem_time_phases", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_float), C.c_int32]),
    ("aem_core_cells", C.c_int64, [C.c_void_p]),
    ("aem_single_pass", C.c_int32, [C.c_void_p]),
]

_lib = None


class AemError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def lib() -> C.CDLL:
    """Load libaem.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with 'python -m alpineem_amd.build'")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != AEM_OK:
        raise AemError(status, lib().aem_last_error().decode())


_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)

NTFF_BINNING = {"ordered": 0, "atomic": 1, "histogram": 2, "off": 3}


class PlanOptions(C.Structure):
    """aem_plan_options (include/aem_plan.h)."""
    _fields_ = [("cpml_cells", C.c_int32), ("cpml_grading", C.c_int32), ("cpml_kappa_max", C.c_float),
                ("cpml_alpha_max", C.c_float), ("cpml_reflection", C.c_float), ("tfsf_inset", C.c_int32),
                ("ntff_gap", C.c_int32), ("kmax", C.c_int32)]


class SimOptions(C.Structure):
    """aem_sim_options (include/aem.h)."""
    _fields_ = [("ntff_binning", C.c_int32), ("planes", C.c_int32), ("no_cpml_fusion", C.c_int32),
                ("no_single_pass", C.c_int32), ("ntff_row_scan", C.c_int32)]


class PortView(C.Structure):
    """aem_port_view (include/aem_plan.h)."""
    _fields_ = [("dir", C.c_int32), ("loc", C.c_int32 * 3), ("size", C.c_int32 * 3),
                ("L", C.c_float), ("C", C.c_float), ("Z", C.c_float), ("dl", C.c_float),
                ("kI", C.c_float), ("kV", C.c_float), ("kS", C.c_float),
                ("cur_num", C.c_float), ("cur_den", C.c_float), ("gap_div", C.c_float),
                ("excited", C.c_int32)]


class SpiceView(C.Structure):
    """aem_spice_view (include/aem_plan.h)."""
    _fields_ = [("dir", C.c_int32), ("loc", C.c_int32 * 3), ("size", C.c_int32 * 3), ("v_bias", C.c_float),
                ("cap", C.c_float), ("cur_num", C.c_float), ("cur_den", C.c_float), ("gap_div", C.c_float)]


class PlanView(C.Structure):
    """aem_plan_view (include/aem_plan.h), field for field."""
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("NX", C.c_int32), ("NY", C.c_int32), ("NZ", C.c_int32),
                ("PX", C.c_int32), ("pbc", C.c_int32 * 3),
                ("plane", C.c_int64), ("ncell", C.c_int64),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("dt", C.c_float), ("DB", C.c_float), ("time_steps", C.c_int32),
                ("pulse_type", C.c_int32), ("spread", C.c_float), ("t_spread", C.c_float),
                ("f_center", C.c_float),
                ("den_ex", _FP), ("den_ey", _FP), ("den_ez", _FP),
                ("den_hx", _FP), ("den_hy", _FP), ("den_hz", _FP),
                ("n_tuple", C.c_int32), ("idx_bytes", C.c_int32), ("tuples", _FP),
                ("tuple_idx8", C.POINTER(C.c_uint8)), ("tuple_idx16", C.POINTER(C.c_uint16)),
                ("npml", C.c_int32 * 6),
                ("be", _FP * 6), ("ce", _FP * 6), ("bh", _FP * 6), ("ch", _FP * 6),
                ("n_sheet_planes", C.c_int32 * 3), ("sheet_pos", _IP * 3),
                ("sheet_sig", _FP * 3), ("sheet_ep", _FP * 3), ("sheet_dim", (C.c_int32 * 2) * 3),
                ("sheet_thickness", C.c_float),
                ("n_poles", C.c_int32), ("has_drude", C.c_int32), ("dbox", (C.c_int32 * 6) * 3),
                ("Ja", _FP * 3), ("Jb", _FP * 3),
                ("has_tfsf", C.c_int32), ("tf_lo", C.c_int32 * 3), ("tf_hi", C.c_int32 * 3),
                ("delay", C.c_int32 * 3), ("WE", C.c_float * 3), ("WH", C.c_float * 3),
                ("s_th", C.c_float), ("c_th", C.c_float), ("s_ph", C.c_float), ("c_ph", C.c_float),
                ("s_pol", C.c_float), ("c_pol", C.c_float),
                ("theta", C.c_float), ("phi", C.c_float), ("pol", C.c_float),
                ("kHinc", C.c_float * 3), ("kEinc", C.c_float * 3), ("inv_c", C.c_float),
                ("plane_wave_amp", C.c_float),
                ("n_ports", C.c_int32), ("M", C.c_int32), ("ports", C.POINTER(PortView)),
                ("antenna_amp", C.c_float), ("excitation_port", C.c_int32),
                ("n_angles", C.c_int32), ("ang", _FP), ("ic", C.c_int32), ("jc", C.c_int32),
                ("kc", C.c_int32), ("r_relay", C.c_float), ("len_ff", C.c_int32),
                ("ff_lo", C.c_int32 * 3), ("ff_hi", C.c_int32 * 3), ("ff_K", C.c_float * 3),
                ("plane_sum_axis", C.c_int32), ("ps_c1", C.c_float), ("ps_c2", C.c_float),
                ("excitation", C.c_int32), ("min_steps", C.c_float),
                ("video_on", C.c_int32), ("slice", C.c_int32), ("slice_location", C.c_int32),
                ("vid_size1", C.c_int32), ("vid_size2", C.c_int32),
                ("src_table", _FP), ("exp_half", C.c_float),
                ("n_spice", C.c_int32), ("n_spice_extra", C.c_int32), ("spice", C.POINTER(SpiceView)),
                ("spice_names", C.c_char_p),
                ("kmax", C.c_int32), ("mode_type", C.c_int32), ("k_direction", C.c_int32), ("k_axis", C.c_int32),
                ("kpl_H", C.c_int32), ("kpl_E", C.c_int32), ("k_vec", C.c_float * 3), ("k_exc", C.c_float * 3),
                ("f_adj", C.c_float), ("bloch_c", C.c_float * 3), ("bloch_s", C.c_float * 3),
                ("src_table_im", C.POINTER(C.c_float)), ("ksrc", C.POINTER(C.c_float)),
                ("mirror_type", C.c_int32), ("wall", C.c_float * 6), ("delay_m", C.c_int32 * 3),
                ("WE_m", C.c_float * 3), ("WH_m", C.c_float * 3), ("s_thm", C.c_float), ("c_thm", C.c_float),
                ("s_phm", C.c_float), ("c_phm", C.c_float), ("use_m", C.c_float * 3), ("off_m", C.c_float * 3),
                ("sgn_M", C.c_float * 3), ("sgn_J", C.c_float * 3)]
