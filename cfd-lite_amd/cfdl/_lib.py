"""ctypes binding of libcfdl_hip.so (declarations in include/cfdl.h)."""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)

PI = C.POINTER(C.c_int)
PD = C.POINTER(C.c_double)
VP = C.c_void_p


class CfdlError(RuntimeError):
    """A libcfdl_hip call returned a non-zero status."""


def library_path():
    return os.environ.get("CFDL_LIBRARY", os.path.join(_PKG, "libcfdl_hip.so"))


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "cfdl_last_error": (C.c_char_p, []),
    "cfdl_version": (C.c_int, []),
    "cfdl_mesh_box": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_uint64, C.POINTER(VP)]),
    "cfdl_mesh_from_elements": (C.c_int, [C.c_int, PD, PD, PD, C.c_int, PI, PI, PI, C.POINTER(VP)]),
    "cfdl_mesh_partition": (C.c_int, [VP, C.c_int]),
    "cfdl_mesh_renumber": (C.c_int, [VP, C.c_uint64]),
    "cfdl_mesh_info": (C.c_int, [VP, PI, PI, PI, PI, PI, PI]),
    "cfdl_mesh_get_i32": (C.c_int, [VP, C.c_char_p, PI, C.c_int64]),
    "cfdl_mesh_get_f64": (C.c_int, [VP, C.c_char_p, PD, C.c_int64]),
    "cfdl_mesh_free": (None, [VP]),
    "cfdl_calc_residual": (C.c_int, [C.c_int, VP, VP, VP, VP, VP, VP, PD, PD, VP]),
    "cfdl_calc_residual_noalloc": (C.c_int, [C.c_int, VP, VP, VP, VP, VP, VP, PD, PD, VP]),
    "cfdl_mesh_mg_get": (C.c_int, [VP, C.c_int, C.c_char_p, PI, C.c_int64, C.POINTER(C.c_int64)]),
    "cfdl_mesh_set_section_names": (C.c_int, [VP, C.c_int, C.POINTER(C.c_char_p)]),
    "cfdl_mesh_write_vtu": (C.c_int, [VP, C.c_char_p, C.c_int, C.POINTER(C.c_char_p), PI,
                                      C.POINTER(PD)]),
    "cfdl_ctx_create": (C.c_int, [C.c_int, C.c_int, C.c_int, PI, PI, PI, PI, PD, PD, PD, PD, PD, PD,
                                  C.c_int, PI, PI, C.c_int, PI, PI, PI, PD, PD, C.c_double, C.c_int,
                                  C.POINTER(VP)]),
    "cfdl_ctx_create_cavity": (C.c_int, [VP, C.c_double, C.c_double, C.c_double, C.c_int,
                                         C.POINTER(VP)]),
    "cfdl_ctx_create_channel": (C.c_int, [VP, C.c_double, C.c_double, C.c_double, C.c_int,
                                          C.c_double, C.POINTER(VP)]),
    "cfdl_set_inlet_profile": (C.c_int, [VP, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_double]),
    "cfdl_ctx_destroy": (None, [VP]),
    "cfdl_ctx_reset": (C.c_int, [VP]),
    "cfdl_set_relaxation": (C.c_int, [VP, C.c_double, C.c_double]),
    "cfdl_set_pc_order": (C.c_int, [VP, C.c_int]),
    "cfdl_set_assembly_precision": (C.c_int, [VP, C.c_int]),
    "cfdl_set_gs_order": (C.c_int, [VP, C.c_int]),
    "cfdl_set_properties": (C.c_int, [VP, PD, PD]),
    "cfdl_set_dt": (C.c_int, [VP, C.c_double]),
    "cfdl_set_bc_kind": (C.c_int, [VP, C.c_int, C.c_int, C.c_int]),
    "cfdl_set_field": (C.c_int, [VP, C.c_int, PD, C.c_int64]),
    "cfdl_get_field": (C.c_int, [VP, C.c_int, PD, C.c_int64]),
    "cfdl_synchronize": (C.c_int, [VP]),
    "cfdl_update_boundaries": (C.c_int, [VP]),
    "cfdl_calc_coef_uvw": (C.c_int, [VP]),
    "cfdl_solve_gs": (C.c_int, [VP, C.c_char_p, C.c_int, C.c_int]),
    "cfdl_solve_gs_uvw": (C.c_int, [VP]),
    "cfdl_calc_grad": (C.c_int, [VP, C.c_int, C.c_int]),
    "cfdl_calc_mip": (C.c_int, [VP, C.c_int]),
    "cfdl_calc_coef_p": (C.c_int, [VP]),
    "cfdl_set_a_0": (C.c_int, [VP, C.c_int]),
    "cfdl_solve": (C.c_int, [VP, C.c_char_p]),
    "cfdl_adjust_pc": (C.c_int, [VP]),
    "cfdl_update_uvwp": (C.c_int, [VP]),
    "cfdl_solve_uvwp": (C.c_int, [VP]),
    "cfdl_update_time": (C.c_int, [VP]),
    "cfdl_calc_mass_imbalance": (C.c_int, [VP, C.c_int]),
    "cfdl_set_thermal": (C.c_int, [VP, PD, PD]),
    "cfdl_construct_energy": (C.c_int, [VP]),
    "cfdl_update_boundaries_energy": (C.c_int, [VP]),
    "cfdl_calc_coef_energy": (C.c_int, [VP]),
    "cfdl_solve_energy": (C.c_int, [VP]),
    "cfdl_run_energy": (C.c_int, [VP, C.c_int, C.c_int]),
    "cfdl_run": (C.c_int, [VP, C.c_int, C.c_int]),
    "cfdl_batch_create": (C.c_int, [C.POINTER(VP)]),
    "cfdl_batch_attach": (C.c_int, [VP, VP]),
    "cfdl_batch_detach": (C.c_int, [VP, VP]),
    "cfdl_batch_size": (C.c_int, [VP, PI]),
    "cfdl_batch_set_block": (C.c_int, [VP, C.c_int]),
    "cfdl_batch_block": (C.c_int, [VP, PI]),
    "cfdl_batch_run": (C.c_int, [VP, C.c_int, C.c_int]),
    "cfdl_batch_set_graphs": (C.c_int, [VP, C.c_int]),
    "cfdl_batch_graph_stats": (C.c_int, [VP, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "cfdl_batch_synchronize": (C.c_int, [VP]),
    "cfdl_batch_destroy": (None, [VP]),
    "cfdl_log_count": (C.c_int, [VP, PI]),
    "cfdl_log_get": (C.c_int, [VP, C.c_int, C.c_char_p, PI, PD, PD, PD]),
    "cfdl_log_clear": (C.c_int, [VP]),
    "cfdl_stats": (C.c_int, [VP, PD, C.c_int]),
    "cfdl_set_profile": (C.c_int, [VP, C.c_int]),
    "cfdl_stats_reset": (C.c_int, [VP]),
}


def exported_names():
    return list(_SIGS)


def load_library(path=None):
    """Load libcfdl_hip.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or library_path()
    if not os.path.exists(path):
        raise CfdlError("libcfdl_hip.so not found at %s: build it first "
                        "(python -c 'import __graft_entry__ as g; g.build()')" % path)
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


_cgns = None

# libcfdl_cgns.so (include/cfdl_cgns.h): CGNS mesh files over HDF5
_CGNS_SIGS = {
    "cfdl_cgns_last_error": (C.c_char_p, []),
    "cfdl_mesh_read_cgns": (C.c_int, [C.c_char_p, C.POINTER(VP)]),
    "cfdl_mesh_write_cgns": (C.c_int, [VP, C.c_char_p]),
    "cfdl_cgns_section_name": (C.c_int, [VP, C.c_int, C.c_char_p, C.c_int]),
}


def cgns_exported_names():
    return list(_CGNS_SIGS)


def cgns_library_path():
    return os.path.join(_PKG, "libcfdl_cgns.so")


def load_cgns_library():
    """Load libcfdl_cgns.so once (after libcfdl_hip.so); raise if it is not built."""
    global _cgns
    if _cgns is not None:
        return _cgns
    load_library()
    path = cgns_library_path()
    if not os.path.exists(path):
        raise CfdlError("libcfdl_cgns.so not found at %s (built when HDF5 headers are present)"
                        % path)
    lib = C.CDLL(path)
    for name, (res, args) in _CGNS_SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _cgns = lib
    return lib


def check_cgns(rc):
    if rc != 0:
        msg = _cgns.cfdl_cgns_last_error().decode(errors="replace")
        raise CfdlError("cfdl cgns error %d: %s" % (rc, msg))


def check(rc):
    if rc != 0:
        msg = _lib.cfdl_last_error().decode(errors="replace")
        raise CfdlError("cfdl error %d: %s" % (rc, msg))
