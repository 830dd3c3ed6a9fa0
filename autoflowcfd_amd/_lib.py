"""ctypes binding of libafr.so (the C ABI declared in include/afr.h)."""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libafr.so")
ABI_VERSION = 6

c_vp = ctypes.c_void_p


class AfrMeshDesc(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("order", ctypes.c_int32), ("dtype", ctypes.c_int32),
        ("n_rows", ctypes.c_int32), ("n_cells", ctypes.c_int64), ("n_prism", ctypes.c_int64),
        ("det", c_vp), ("adj", c_vp), ("adj_fine", c_vp),
        ("slot_nb", c_vp), ("slot_blk", c_vp), ("slot_mask", c_vp), ("slot_bc", c_vp), ("slot_fallback", c_vp),
        ("slot_tn", c_vp), ("slot_adjrow", c_vp), ("rows", c_vp),
        ("rows_nnz", c_vp), ("rows_idx", c_vp), ("rows_w", c_vp), ("rows_knz", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("c2f", c_vp * 2), ("op", c_vp * 2), ("E", c_vp * 2), ("D", c_vp * 2), ("filt", c_vp * 2),
        ("gL", c_vp), ("gR", c_vp), ("fp_of_sp", c_vp), ("coord_of_sp", c_vp), ("prism_tab", c_vp),
        ("prism_schedule", c_vp),
        ("gcl", c_vp), ("bc_params", c_vp), ("cell_case", c_vp), ("n_cases", ctypes.c_int32),
        ("pad2_", ctypes.c_int32), ("case_prism_off", c_vp), ("case_tet_off", c_vp), ("case_wall_off", c_vp), ("ref_state", ctypes.c_double * 5),
        ("work_q", c_vp), ("work_grad", c_vp), ("work_umean", c_vp),
        ("work_rvol", c_vp), ("work_h", c_vp),
        ("work_vpack", c_vp), ("dsp_w", c_vp * 2), ("esp_w", c_vp * 2), ("dsp_i", c_vp * 2), ("esp_i", c_vp * 2),
        ("metric_flux_scale", c_vp), ("volume", c_vp),
        ("work_r", c_vp), ("work_s1", c_vp), ("work_s2", c_vp), ("work_dt", c_vp), ("work_part", c_vp),
        ("n_records", ctypes.c_int64), ("rec_owner", c_vp), ("rec_oslot", c_vp), ("rec_flags", c_vp),
        ("rec_tn", c_vp), ("cell_inc", c_vp),
        ("work_pack", c_vp), ("work_nbp", c_vp), ("work_tvol", c_vp), ("work_tsrc", c_vp), ("work_crec", c_vp),
        ("n_wall_rec", ctypes.c_int64), ("wrec_owner", c_vp), ("wrec_oslot", c_vp), ("wrec_geom", c_vp),
        ("w1d", c_vp), ("work_fpart", c_vp), ("grid_cap", ctypes.c_int32 * 8),
        ("variant", ctypes.c_int32 * 8),
        ("state_ref", c_vp),
    ]


class AfrGeomDesc(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("order", ctypes.c_int32), ("dtype", ctypes.c_int32),
        ("pad_", ctypes.c_int32), ("n_cells", ctypes.c_int64), ("n_prism", ctypes.c_int64),
        ("nodes", c_vp), ("prism_conn", c_vp), ("tet_conn", c_vp),
        ("sps", c_vp), ("fine", c_vp), ("w_sps", c_vp), ("op64", c_vp * 2), ("E64", c_vp * 2),
        ("slot_tn64", c_vp),
        ("det", c_vp), ("adj", c_vp), ("adj_fine", c_vp), ("gcl", c_vp), ("metric_flux_scale", c_vp),
        ("slot_adjrow", c_vp), ("volume", c_vp), ("det_mean", c_vp), ("slot_fallback", c_vp), ("bad_det", c_vp),
        ("sps_xyz", c_vp),
    ]


class AfrP0Desc(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("dtype", ctypes.c_int32), ("n_cells", ctypes.c_int64),
        ("n_records", ctypes.c_int64),
        ("det", c_vp), ("adj", c_vp), ("mfs", c_vp), ("vol1", c_vp), ("cell_volume", c_vp),
        ("rec_owner", c_vp), ("rec_nb", c_vp), ("rec_bc", c_vp), ("rec_flags", c_vp), ("rec_geom", c_vp),
        ("cell_inc", c_vp), ("cell_ghost_rec", c_vp), ("bc_params", c_vp),
        ("work_r", c_vp), ("work_s1", c_vp), ("work_s2", c_vp), ("work_dt", c_vp), ("work_part", c_vp),
        ("rec_src", c_vp), ("rec_w", c_vp), ("d_wall", c_vp), ("work_sst", c_vp),
        ("state_ref", c_vp),
    ]


class AfrTopoDesc(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("pad_", ctypes.c_int32), ("n_nodes", ctypes.c_int64),
        ("n_prism", ctypes.c_int64), ("n_tet", ctypes.c_int64),
        ("nodes", c_vp), ("prism_conn", c_vp), ("tet_conn", c_vp), ("work", c_vp), ("work_bytes", ctypes.c_size_t),
        ("owner", c_vp), ("neighbor", c_vp), ("owner_code", c_vp), ("neighbor_code", c_vp), ("node_ids", c_vp),
        ("normal", c_vp), ("area", c_vp), ("center", c_vp), ("n_faces", c_vp), ("status", c_vp),
    ]


class AfrSlotDesc(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("pad_", ctypes.c_int32), ("n_prism", ctypes.c_int64),
        ("n_tet", ctypes.c_int64), ("n_records", ctypes.c_int64),
        ("nodes", c_vp), ("prism_conn", c_vp), ("tet_conn", c_vp),
        ("owner", c_vp), ("neighbor", c_vp), ("owner_code", c_vp), ("neighbor_code", c_vp), ("bc", c_vp),
        ("bc_row", c_vp), ("normal", c_vp), ("translation", c_vp), ("work", c_vp), ("work_bytes", ctypes.c_size_t),
        ("nb", c_vp), ("blk", c_vp), ("slot_bc", c_vp), ("slot_bc_row", c_vp), ("tn", c_vp), ("trans", c_vp),
        ("other", c_vp), ("rec", c_vp), ("slot_flags", c_vp), ("rep", c_vp), ("run_start", c_vp),
        ("sorted_slot", c_vp), ("counts", c_vp), ("status", c_vp),
    ]


# every function declared in include/afr.h: name -> (restype, argtypes)
SIGNATURES = {
    "afr_version": (ctypes.c_int, []),
    "afr_last_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "afr_mesh_create": (ctypes.c_int, [ctypes.POINTER(AfrMeshDesc), ctypes.POINTER(c_vp)]),
    "afr_mesh_destroy": (ctypes.c_int, [c_vp]),
    "afr_metrics": (ctypes.c_int, [ctypes.POINTER(AfrGeomDesc), c_vp]),
    "afr_primitives": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int64, ctypes.c_int32,
                                      ctypes.POINTER(ctypes.c_double), c_vp]),
    "afr_prepare_state": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "afr_inviscid_residual": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_int32, c_vp]),
    "afr_inviscid_from_primitives": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int32, c_vp]),
    "afr_viscous_residual": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_int32, c_vp]),
    "afr_residual": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int32, c_vp]),
    "afr_viscous_volume": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          c_vp]),
    "afr_viscous_interface": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int32, ctypes.c_int32, c_vp]),
    "afr_local_time_step": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_double, c_vp]),
    "afr_rk_stage": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, c_vp]),
    "afr_residual_norm_f64": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int32, c_vp, c_vp]),
    "afr_residual_norm_f64_noalloc": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int32, c_vp, c_vp]),
    "afr_p0_residual": (ctypes.c_int, [ctypes.POINTER(AfrP0Desc), c_vp, c_vp, ctypes.c_double, c_vp, c_vp]),
    "afr_p0_ssp_rk3_step": (ctypes.c_int, [ctypes.POINTER(AfrP0Desc), c_vp, c_vp, ctypes.c_double, ctypes.c_int32,
                                           c_vp, c_vp]),
    "afr_face_topology_workspace": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    "afr_face_topology": (ctypes.c_int, [ctypes.POINTER(AfrTopoDesc), c_vp]),
    "afr_face_slots_workspace": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    "afr_face_slots": (ctypes.c_int, [ctypes.POINTER(AfrSlotDesc), c_vp]),
    "afr_p0_sst_step": (ctypes.c_int, [ctypes.POINTER(AfrP0Desc), c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_double,
                                       ctypes.c_int32, c_vp, c_vp]),
    "afr_interpolate_order": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_int32, c_vp]),
    "afr_ssp_rk3_step": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int32, c_vp, c_vp]),
    "afr_ssp_rk3_step_dt": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int32, c_vp, c_vp]),
    "afr_wall_distance": (ctypes.c_int, [c_vp, ctypes.c_int64, c_vp, ctypes.c_int32, c_vp, ctypes.c_int32, c_vp]),
    "afr_sst_step": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_double, c_vp]),
    "afr_sst_rk3_step": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int32, c_vp, c_vp]),
    "afr_local_time_step_nut": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_double, c_vp]),
    "afr_forces_f64": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_double, c_vp, c_vp]),
    "afr_mesh_set_grid": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.c_int32]),
    "afr_mesh_get_grid": (ctypes.c_int, [c_vp, ctypes.c_int32]),
    "afr_mesh_set_variant": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.c_int32]),
    "afr_mesh_get_variant": (ctypes.c_int, [c_vp, ctypes.c_int32]),
    "afr_graph_begin": (ctypes.c_int, [c_vp]),
    "afr_graph_end": (ctypes.c_int, [c_vp, ctypes.POINTER(c_vp)]),
    "afr_graph_launch": (ctypes.c_int, [c_vp, c_vp]),
    "afr_graph_destroy": (ctypes.c_int, [c_vp]),
    "afr_residual_norm_cases": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int32, c_vp, c_vp]),
    "afr_forces_cases_f64": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_double, c_vp, c_vp]),
    "afr_suppress_residual_outliers": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                                      ctypes.c_int32, ctypes.c_int32, c_vp]),
}

_lib = None


def load():
    """Load libafr.so (or the build named by AFR_LIBRARY); raises if it has not been built
    (there is no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("AFR_LIBRARY", LIB_PATH)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m autoflowcfd_amd.build`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.afr_version() != ABI_VERSION:
        raise RuntimeError("libafr.so ABI version mismatch; rebuild it")
    _lib = lib
    return lib


def check(status: int, what: str):
    if status != 0:
        buf = ctypes.create_string_buffer(512)
        load().afr_last_error(buf, 512)
        raise RuntimeError(f"{what} failed: {buf.value.decode(errors='replace')}")
