"""ctypes binding of libpct.so (the C ABI declared in include/pct.h and
include/pct_synth_api.h).

The library is loaded from this package directory. There is no fallback: if
libpct.so is missing the import of the GPU entry points fails loudly.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpct.so")

c_i32, c_i64, c_u64, c_f32, c_f64, c_vp = (ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float,
                                          ctypes.c_double, ctypes.c_void_p)


class PctParams(ctypes.Structure):
    _fields_ = [
        ("columns", c_i32), ("rows", c_i32), ("slices", c_i32),
        ("t_bins", c_i32), ("v_bins", c_i32), ("angular_bins", c_i32),
        ("recon_cyl_radius", c_f64), ("recon_cyl_height", c_f64), ("slice_thickness", c_f64),
        ("t_bin_size", c_f64), ("v_bin_size", c_f64), ("angular_bin_size", c_f64),
        ("ssd_t_size", c_f64), ("ssd_v_size", c_f64), ("source_radius", c_f64),
        ("gantry_angle_interval", c_f64), ("sigmas_to_keep", c_f64), ("fbp_threshold", c_f64),
        ("rsp_air", c_f64), ("sc_threshold", c_f64), ("msc_threshold", c_f64),
        ("hull_filter_threshold", c_f64), ("hull_filter_radius", c_i32), ("msc_diff_thresh", c_i32),
        ("e_0", c_f64), ("x_0", c_f64), ("a", c_f64 * 6),
        ("max_intersections", c_i32), ("drop_block_size", c_i32), ("lambda_", c_f32), ("iterations", c_i32),
        ("eta", c_f64), ("psi_sign", c_i32), ("sm_lower_threshold", c_f64), ("sm_scale_threshold", c_f64),
    ]

    def copy(self):
        c = PctParams()
        ctypes.pointer(c)[0] = self
        return c


class PctRawSoa(ctypes.Structure):
    _fields_ = [("gantry_angle", c_vp)] + [(n, c_vp) for n in (
        "t_in_1", "t_in_2", "t_out_1", "t_out_2", "u_in_1", "u_in_2", "u_out_1", "u_out_2",
        "v_in_1", "v_in_2", "v_out_1", "v_out_2", "wepl")] + [("n", c_i64)]


HIST_FLOAT_FIELDS = ("x_entry", "y_entry", "z_entry", "x_exit", "y_exit", "z_exit",
                     "xy_entry_angle", "xz_entry_angle", "xy_exit_angle", "xz_exit_angle", "wepl")
RAW_FLOAT_FIELDS = ("t_in_1", "t_in_2", "t_out_1", "t_out_2", "u_in_1", "u_in_2", "u_out_1", "u_out_2",
                    "v_in_1", "v_in_2", "v_out_1", "v_out_2", "wepl")


class PctHistSoa(ctypes.Structure):
    _fields_ = [(n, c_vp) for n in HIST_FLOAT_FIELDS] + [("bin_num", c_vp), ("first_voxel", c_vp), ("n", c_i64)]


class PctBinStats(ctypes.Structure):
    _fields_ = [(n, c_vp) for n in ("counts", "mean_wepl", "mean_rel_ut", "mean_rel_uv", "stddev_wepl",
                                    "stddev_rel_ut", "stddev_rel_uv")]


BIN_STAT_FIELDS = ("counts", "mean_wepl", "mean_rel_ut", "mean_rel_uv", "stddev_wepl", "stddev_rel_ut",
                   "stddev_rel_uv")

SYNTH_MAX_CYL = 16


class PctSynthCfg(ctypes.Structure):
    _fields_ = [
        ("seed", c_u64), ("n_histories", c_i64), ("n_angles", c_i32), ("angle_step_deg", c_f64),
        ("u_in_1", c_f64), ("u_in_2", c_f64), ("u_out_1", c_f64), ("u_out_2", c_f64),
        ("t_half", c_f64), ("v_half", c_f64), ("source_radius", c_f64), ("beta_cp", c_f64),
        ("x0", c_f64), ("straggle", c_f64), ("n_cyl", c_i32),
        ("cyl_x", c_f64 * SYNTH_MAX_CYL), ("cyl_y", c_f64 * SYNTH_MAX_CYL),
        ("cyl_r", c_f64 * SYNTH_MAX_CYL), ("cyl_rsp", c_f64 * SYNTH_MAX_CYL),
        ("n_ell", c_i32),
        ("ell_x", c_f64 * SYNTH_MAX_CYL), ("ell_y", c_f64 * SYNTH_MAX_CYL), ("ell_z", c_f64 * SYNTH_MAX_CYL),
        ("ell_a", c_f64 * SYNTH_MAX_CYL), ("ell_b", c_f64 * SYNTH_MAX_CYL), ("ell_c", c_f64 * SYNTH_MAX_CYL),
        ("ell_phi", c_f64 * SYNTH_MAX_CYL), ("ell_rsp", c_f64 * SYNTH_MAX_CYL),
    ]


class PctSynthSoa(ctypes.Structure):
    _fields_ = [("gantry_angle", c_vp)] + [(n, c_vp) for n in RAW_FLOAT_FIELDS]


class PctNetTrainCfg(ctypes.Structure):
    _fields_ = [("epochs", c_i32), ("batch", c_i32), ("lr", c_f32), ("lr_min", c_f32), ("beta1", c_f32),
                ("beta2", c_f32), ("eps", c_f32), ("seed", c_u64)]


NET_CELLS = 64


class PctNetCalib(ctypes.Structure):
    _fields_ = [("samples", c_i64), ("list_mismatch", c_i64), ("in_range", c_i64), ("surrogate", c_i64),
                ("surrogate_mismatch", c_i64), ("dev_median", c_f64),
                ("dev_p99", c_f64), ("dev_max", c_f64), ("cells_fallback", c_i32),
                ("cell_samples", c_i32 * NET_CELLS), ("cell_p99", c_f32 * NET_CELLS),
                ("cell_fallback", ctypes.c_uint8 * NET_CELLS), ("cell_list_mismatch", c_f32 * NET_CELLS)]


PCTD_STR = 256


class PctPctdHeader(ctypes.Structure):
    _fields_ = [("version", c_i32), ("n_histories", c_i32), ("projection_angle", c_f32), ("beam_energy", c_f32),
                ("generation_date", c_i32), ("preprocess_date", c_i32),
                ("phantom_name", ctypes.c_char * PCTD_STR), ("data_source", ctypes.c_char * PCTD_STR),
                ("prepared_by", ctypes.c_char * PCTD_STR), ("data_offset", c_i64)]


PCT_TXT_F32, PCT_TXT_F64, PCT_TXT_I32, PCT_TXT_U32, PCT_TXT_BOOL = 0, 1, 2, 3, 4

# name -> (restype, argtypes) of every exported entry point of include/pct.h,
# include/pct_net.h, include/pct_synth_api.h and include/pct_io.h
P = ctypes.POINTER
SIGNATURES = {
    "pct_create": (c_vp, [P(PctParams), c_vp]),
    "pct_destroy": (None, [c_vp]),
    "pct_sync": (ctypes.c_int, [c_vp]),
    "pct_last_error": (ctypes.c_char_p, [c_vp]),
    "pct_default_params": (None, [P(PctParams)]),
    "pct_recon_volume_intersections": (ctypes.c_int, [c_vp, P(PctRawSoa), P(PctHistSoa), c_vp]),
    "pct_hull_init": (ctypes.c_int, [c_vp, c_vp]),
    "pct_space_carve": (ctypes.c_int, [c_vp, P(PctHistSoa), c_vp, c_vp]),
    "pct_msc_init": (ctypes.c_int, [c_vp, c_vp]),
    "pct_msc_count": (ctypes.c_int, [c_vp, P(PctHistSoa), c_vp, c_vp]),
    "pct_msc_edge": (ctypes.c_int, [c_vp, c_vp, c_vp]),
    "pct_sm_count": (ctypes.c_int, [c_vp, P(PctHistSoa), c_vp, c_vp]),
    "pct_sm_edge": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "pct_hull_select": (ctypes.c_int, [c_vp, c_vp, c_vp]),
    "pct_fbp_filter": (ctypes.c_int, [c_vp, c_vp, c_vp]),
    "pct_backproject": (ctypes.c_int, [c_vp, c_vp, c_vp]),
    "pct_fbp_threshold": (ctypes.c_int, [c_vp, c_vp, c_vp]),
    "pct_initial_iterate": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "pct_binning": (ctypes.c_int, [c_vp, P(PctHistSoa), c_vp, c_vp]),
    "pct_bin_statistics": (ctypes.c_int, [c_vp, P(PctHistSoa), c_vp, c_vp, P(PctBinStats)]),
    "pct_statistical_cuts": (ctypes.c_int, [c_vp, P(PctHistSoa), c_vp, P(PctBinStats), c_vp, c_vp]),
    "pct_mlp_endpoints": (ctypes.c_int, [c_vp, P(PctHistSoa), c_vp, c_vp]),
    "pct_compact": (ctypes.c_int, [c_vp, P(PctHistSoa), c_vp, P(PctHistSoa), P(c_i64)]),
    "pct_mlp_capacity": (ctypes.c_int, [c_vp, P(PctHistSoa), c_i64, c_i64, c_vp, P(c_i64)]),
    "pct_mlp_paths": (ctypes.c_int, [c_vp, P(PctHistSoa), c_i64, c_i64, c_vp, c_vp, c_vp, ctypes.c_int]),
    "pct_drop_block_update": (ctypes.c_int, [c_vp, P(PctHistSoa), c_i64, c_i64, c_vp, c_vp, c_vp, c_f32, ctypes.c_int]),
    "pct_drop_block_update_noalloc": (ctypes.c_int, [c_vp, P(PctHistSoa), c_i64, c_i64, c_vp, c_vp, c_vp, c_f32,
                                                     ctypes.c_int]),
    "pct_drop_image_update": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "pct_drop_iteration": (ctypes.c_int, [c_vp, P(PctHistSoa), c_vp, ctypes.c_int, c_f32, ctypes.c_int]),
    "pct_drop_invalidate": (ctypes.c_int, [c_vp]),
    "pct_drop_plan_bytes": (c_i64, [c_i64, c_i64, c_i64, c_i64, ctypes.c_int]),
    "pct_drop_plan_choice": (ctypes.c_int, [c_i64, c_i64, c_i64, c_i64, c_i64, ctypes.c_int, c_i64, ctypes.c_int,
                                            ctypes.POINTER(c_i64)]),
    "pct_drop_prepare": (ctypes.c_int, [c_vp, P(PctHistSoa), ctypes.c_int, ctypes.c_int]),
    "pct_set_profiling": (ctypes.c_int, [c_vp, ctypes.c_int]),
    "pct_last_timings": (ctypes.c_int, [c_vp, c_vp]),
    "pct_drop_stats": (ctypes.c_int, [c_vp, P(PctHistSoa), ctypes.c_int, P(c_i64), P(c_i64)]),
    "pct_set_tuning": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int]),
    "pct_batch_create": (c_vp, [P(c_vp), ctypes.c_int]),
    "pct_batch_destroy": (None, [c_vp]),
    "pct_batch_prepare": (ctypes.c_int, [c_vp, P(PctHistSoa), P(ctypes.c_int), ctypes.c_int]),
    "pct_batch_iteration": (ctypes.c_int, [c_vp, P(PctHistSoa), P(c_vp), P(ctypes.c_int), P(c_f32), ctypes.c_int]),
    "pct_selftest_math": (ctypes.c_int, [c_vp, ctypes.c_int, c_vp, c_vp, c_vp, c_i64]),
    "pct_net_features": (ctypes.c_int, [c_vp, P(PctHistSoa), c_i64, c_i64, c_vp, c_vp]),
    "pct_net_gradient": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, P(c_f64)]),
    "pct_net_train": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, P(PctNetTrainCfg), c_vp]),
    "pct_net_set": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "pct_net_set_hermite": (ctypes.c_int, [c_vp, c_vp, c_vp]),
    "pct_net_forward": (ctypes.c_int, [c_vp, c_vp, c_i64, c_vp]),
    "pct_net_calibrate": (ctypes.c_int, [c_vp, P(PctHistSoa), c_i64, c_i64, c_i64, c_f32, c_i32, P(PctNetCalib)]),
    "pct_net_calibrate_rule": (ctypes.c_int, [c_vp, P(PctHistSoa), c_i64, c_i64, c_i64, ctypes.c_int, c_f32, c_i32,
                                              P(PctNetCalib)]),
    "pct_net_fallback_stats": (ctypes.c_int, [c_vp, P(c_i64), P(c_i64)]),
    "pct_synth_generate": (ctypes.c_int, [c_vp, P(PctSynthCfg), c_i64, c_i64, P(PctSynthSoa)]),
    "pct_synth_generate_host": (ctypes.c_int, [P(PctSynthCfg), c_i64, c_i64, P(PctSynthSoa)]),
    "pct_pctd_read_header": (ctypes.c_int, [ctypes.c_char_p, P(PctPctdHeader)]),
    "pct_pctd_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, c_i64, P(PctSynthSoa), P(c_i64)]),
    "pct_pctd_write": (ctypes.c_int, [ctypes.c_char_p, P(PctPctdHeader), P(PctRawSoa)]),
    "pct_array_to_disk": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, c_vp, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, c_i64, ctypes.c_int]),
    "pct_io_last_error": (ctypes.c_char_p, []),
    "pct_mlp_paths_file_info": (ctypes.c_int, [ctypes.c_char_p, P(c_i64), P(c_i64), P(c_i64)]),
    "pct_mlp_paths_file_read": (ctypes.c_int, [ctypes.c_char_p, c_i64, c_i64, c_vp, c_vp, P(c_i64)]),
    "pct_mlp_paths_file_write": (ctypes.c_int, [ctypes.c_char_p, c_i64, c_vp, c_vp]),
    "pct_path_errors_read": (ctypes.c_int, [ctypes.c_char_p, c_i64, c_vp]),
    "pct_path_errors_write": (ctypes.c_int, [ctypes.c_char_p, c_i64, c_vp]),
    "pct_chord_lengths": (ctypes.c_int, [c_vp, P(PctHistSoa), ctypes.c_int, c_vp]),
    "pct_drop_lists_iteration": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                                ctypes.c_int, c_f32, ctypes.c_int]),
}

PCT_NOALLOC_MAX_HISTORIES = 8192
PCT_NET_RULE_DEVIATION = 0
PCT_NET_RULE_LISTS = 1
PCT_NOALLOC_MAX_SLOTS = 512
PCT_MLP_FP64 = 0
PCT_MLP_TABLE = 1
PCT_MLP_SURROGATE = 2
NET_IN, NET_OUT, NET_PARAMS = 12, 8, 5512
PCT_DROP_CACHED = 0x100
PCT_DROP_ATOMIC = 0x200
PCT_DROP_NO_GRAPH = 0x400
PCT_DROP_F32 = 0x800
PCT_DROP_ROBUST = 0x1000
PCT_LISTS_DROP, PCT_LISTS_DROP_RED, PCT_LISTS_ERROR_RED, PCT_LISTS_ERROR, PCT_LISTS_ERROR_RED_SP = 0, 1, 2, 3, 4
PCT_LISTS_REUSE = 0x100
PCT_DROP_ROBUST2 = 0x2000
PCT_DROP_PERSISTENT = 0x4000
# pct_set_tuning keys
PCT_TUNE_MLP_WAVES = 1
PCT_TUNE_CHUNK_FIRST = 2
PCT_TUNE_CHUNK_MAX = 3
PCT_TUNE_OVERLAP = 4
PCT_TUNE_BLOCK_FORM = 5
PCT_TUNE_PLAN_LIMIT_MB = 6
PCT_TUNE_WAIT_GIVE_UP = 7
# the precision form: tabulated MLP with FP32 stepping + FP32/Kahan update
PCT_PRECISION = PCT_MLP_TABLE | PCT_DROP_F32
# the learned-path form: surrogate MLP (exact fallback) + FP32/Kahan update
PCT_LEARNED = PCT_MLP_SURROGATE | PCT_DROP_F32

_lib = None


def lib():
    """The loaded libpct.so (raises if it was never built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libpct.so not found at {LIB_PATH}: run __graft_entry__.build() / python -m pct_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def default_params() -> PctParams:
    p = PctParams()
    lib().pct_default_params(ctypes.byref(p))
    return p
