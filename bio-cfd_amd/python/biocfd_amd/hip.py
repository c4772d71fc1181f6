"""ctypes binding of libbiocfd_hip.so (include/biocfd_hip.h).

Only plain pointers and sizes cross the boundary.  Every call raises
BiocfdError with the library's own message when it returns non-zero; there is
no fallback path: if the library is missing the import of this module fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "..", "lib", "libbiocfd_hip.so"))

# enums of include/biocfd_hip.h
BF = dict(u=0, v=1, w=2, ut=3, vt=4, wt=5, p=6, b=7, pc=8, pco=9)
BT = dict(cell=0, cell2=1, nodeIdTag=2, cell_n=3, cell_pr=4)
BG_NAMES = ["x1", "y1", "z1", "deltax", "deltay", "deltaz", "xu", "yu", "zu", "xv", "yv", "zv",
            "xw", "yw", "zw", "xp", "yp", "zp"]
BL = dict(fluidIndexPtr=0, redCellIndexPtr=1, blackCellIndexPtr=2, solidIndexPtr=3,
          interceptedIndexPtr=4, TSIndexPtr=5)
BN_NAMES = ["nelp", "nelu1", "nelu2", "nelv1", "nelv2", "nelw1", "nelw2"]
BND_NAMES = ["pNormDis", "u1NormDis", "u2NormDis", "v1NormDis", "v2NormDis", "w1NormDis",
             "w2NormDis"]
BE_NAMES = ["xcent", "ycent", "zcent", "cosAlpha", "cosBeta", "cosGamma", "element_length"]
BGH_NAMES = ["p_ghost", "pt_ghost", "u1_ghost", "u1t_ghost", "u2_ghost", "u2t_ghost", "v1_ghost",
             "v1t_ghost", "v2_ghost", "v2t_ghost", "w1_ghost", "w1t_ghost", "w2_ghost",
             "w2t_ghost"]
BC = dict(fluid=0, solid=1, ib=2, red=3, black=4, ts=5)
BTB_NAMES = ["uu", "vv", "ww", "uv", "uw", "vu", "vw", "wu", "wv"]
BK_N = 11
BKS_N = 35

# declared C entry points: name -> argtypes
_P = C.c_void_p
_I = C.c_int
_L = C.c_int64
_D = C.c_double
SIGNATURES = {
    "biocfd_init": [_I],
    "biocfd_finalize": [],
    "biocfd_timing_enable": [_I],
    "biocfd_timing_read": [_P, _P, _P],
    "biocfd_copy_probe": [_L, _I],
    "biocfd_synchronize": [],
    "biocfd_block_create": [_I, _L, _L, _L, _D, _L, _D, _D, _D],
    "biocfd_block_destroy": [_I],
    "biocfd_grid_upload": [_I, _I, _P, _L],
    "biocfd_field_upload": [_I, _I, _P],
    "biocfd_field_download": [_I, _I, _P],
    "biocfd_tag_upload": [_I, _I, _P],
    "biocfd_tag_download": [_I, _I, _P],
    "biocfd_counts": [_I, _P],
    "biocfd_list_download": [_I, _I, _P, _L],
    "biocfd_nel_download": [_I, _I, _P, _L],
    "biocfd_normdis_download": [_I, _I, _P, _L],
    "biocfd_ghost_download": [_I, _I, _P, _L],
    "biocfd_elem_download": [_I, _I, _P, _L],
    "biocfd_index_ts_download": [_I, _P, _L],
    "biocfd_coef_download": [_I, _I, _P, _L],
    "biocfd_table_download": [_I, _I, _I, _P, _L],
    "biocfd_sor_state": [_I, _P, _P, _P],
    "biocfd_sor_exhausted": [_I, _P],
    "biocfd_health_check": [_I, _P, _P, _P],
    "biocfd_health_config": [_I, _D, _I],
    "biocfd_inject_alloc_failure": [_I],
    "biocfd_body_create": [_I, _L, _L, _P, _P, _P, _P, _P],
    "biocfd_coefficient_matrix": [_I, _I],
    "biocfd_non_uni_coeff": [_I],
    "biocfd_ns_momentum": [_I, _D, _D],
    "biocfd_ns_momentum_set": [_I, _I, _D, _D],
    "biocfd_velocity_bc": [_I, _D, _D],
    "biocfd_solid_cell_bc": [_I],
    "biocfd_solid_cell_bc_move": [_I, _L],
    "biocfd_update_velocity_newv": [_I, _L],
    "biocfd_poisson_solver": [_I, _I, _L, _D, _D, _D],
    "biocfd_compute_div": [_I],
    "biocfd_redblack_sor": [_I, _L, _D, _D],
    "biocfd_correct": [_I, _D],
    "biocfd_sor_config": [_I, _I],
    "biocfd_sor_block_config": [_I, _I, _I],
    "biocfd_sor_stop_rule": [_I, _I],
    "biocfd_sor_variant_info": [_I, _P],
    "biocfd_sor_time": [_I, _I, _I, _I, _D, _P],
    "biocfd_set_precision": [_I],
    "biocfd_mixed_config": [_I, _I, _I, _D],
    "biocfd_mixed_variant_info": [_I, _P],
    "biocfd_mixed_time": [_I, _I, _I, _I, _D, _P],
    "biocfd_residual_norm": [_I, _D, _P],
    "biocfd_projection": [_I, _L, _D],
    "biocfd_projection_adjoint": [_I, _L, _D, _P, _P, _P, _P, _P, _P, _P],
    "biocfd_coarse_update_adjoint": [_I, _I, _P, _P],
    "biocfd_fine_update_bd_adjoint": [_I, _I, _P, _P],
    "biocfd_compute_surface_norm": [_I, _P, _P, _P, _L],
    "biocfd_tagging": [_I],
    "biocfd_selective_retagging": [_I, _L],
    "biocfd_cell_count_solid": [_I],
    "biocfd_cell_count_solid_coarse": [_I],
    "biocfd_fine_block_cell": [_I],
    "biocfd_compute_norm_distance": [_I],
    "biocfd_find_ts_cells": [_I],
    "biocfd_velocity_forcing1": [_I, _P],
    "biocfd_pressure_forcing1": [_I, _P],
    "biocfd_velocity_forcing_field": [_I, _P],
    "biocfd_pressure_forcing_field": [_I, _P],
    "biocfd_velocity_forcing_ghost": [_I, _P],
    "biocfd_pressure_forcing_ghost": [_I, _P],
    "biocfd_stress_calculation": [_I, _D, _D, _P, _P],
    "biocfd_stress_calculation_noalloc": [_I, _D, _D, _P, _P],
    "biocfd_interface_create": [_I, _I, _I, _L, _L, _P],
    "biocfd_coarse_update": [_I],
    "biocfd_fine_update_bd": [_I],
    "biocfd_fine_update_bd_one": [_I, _I],
    "biocfd_interface_table": [_I, _I, _I, _P, _P, _L],
    "biocfd_shift_surface_nodes_initial": [_L, _P, _P, _P, _P, _D, _D, _L, _D, _D, _D, _P, _P, _P, _P],
    "biocfd_compute_surface_variables": [_L, _P, _P, _P, _P, _D, _D, _L, _D, _D, _D, _D, _P, _P, _P, _P],
    "biocfd_find_dist_node": [_L, _P, _P, _P, _P],
    "biocfd_block_move_check": [_I, _P, _P, _P, _P, _P, _P],
    "biocfd_change_block_coords": [_I, _L],
    "biocfd_change_block_coords_interfaces": [_I, _L],
    "biocfd_change_block_interface": [_I, _L],
    "biocfd_cell_count_solid_coarse_mv": [_I, _P],
    "biocfd_fine_update_mv": [_I],
    "biocfd_tagging_move": [_I, _L],
    "biocfd_grid_download": [_I, _I, _P, _L],
    "biocfd_output_fields": [_I, _P, _P, _P, _P, _P, _P, _P],
    "biocfd_sor_batching": [_I],
    "biocfd_sor_graphs": [_I],
    "biocfd_interp_batching": [_I],
    "biocfd_sor_graph_stats": [_P, _P],
    "biocfd_momentum_config": [_I, _I],
    "biocfd_momentum_precision": [_I],
    "biocfd_momentum_divisions": [_I],
    "biocfd_ts_pass_config": [_I],
    "biocfd_ad_reset": [_I],
    "biocfd_ad_upload": [_I, _I, _P],
    "biocfd_ad_download": [_I, _I, _P],
    "biocfd_ad_kin": [_I, _P, _I],
    "biocfd_ad_ghost_upload": [_I, _I, _P, _L],
    "biocfd_ad_ghost_download": [_I, _I, _P, _L],
    "biocfd_tlm_momentum": [_I, _D, _D],
    "biocfd_adj_momentum": [_I, _D, _D],
    "biocfd_tlm_velocity_bc": [_I, _D, _D],
    "biocfd_adj_velocity_bc": [_I, _D, _D],
    "biocfd_tlm_solid_cell_bc": [_I],
    "biocfd_adj_solid_cell_bc": [_I],
    "biocfd_tlm_forcing": [_I, _I, _P, _P, _I],
    "biocfd_adj_forcing": [_I, _I, _P, _I],
    "biocfd_tlm_poisson": [_I, _L, _D, _D],
    "biocfd_adj_poisson": [_I, _L, _D, _D],
    "biocfd_stress_sums": [_I, _D, _D, _P, _P],
    "biocfd_tlm_stress": [_I, _D, _D, _P, _P, _I, _P],
    "biocfd_adj_stress": [_I, _D, _D, _P, _P, _I],
    "biocfd_tlm_step_start": [_I, _L, _D, _D, _D, _P, _P, _I],
    "biocfd_tlm_step_end": [_I, _P, _P, _I],
    "biocfd_adj_step_end": [_I, _P, _I],
    "biocfd_adj_step_start": [_I, _L, _D, _D, _D, _P, _I],
    "biocfd_ad_record": [_I, _I],
    "biocfd_ad_geometry_tangent": [_I, _D, _P, _P, _P],
    "biocfd_ad_geometry_adjoint": [_I, _D, _P, _P, _P, _I],
    "biocfd_snapshot_save": [_I, _I],
    "biocfd_snapshot_restore": [_I, _I],
    "biocfd_snapshot_free": [_I, _I],
}


class BiocfdError(RuntimeError):
    """A non-zero return of the library; code is that return value (-1
    for most failures, -3 / -4 for the failure checks of the pressure
    solve)."""

    def __init__(self, msg: str, code: int | None = None):
        super().__init__(msg)
        self.code = code


def load(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise BiocfdError(f"HIP library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.biocfd_last_error.argtypes = []
    lib.biocfd_last_error.restype = C.c_char_p
    return lib


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def ptr(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def call(name: str, *args) -> int:
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().biocfd_last_error().decode(errors="replace")
        raise BiocfdError(f"{name} failed ({rc}): {msg}", rc)
    return rc
