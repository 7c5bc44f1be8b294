"""ctypes binding of libtmrx.so (the C-ABI declared in include/tmrx.h).

The library is loaded from this package directory.  There is no fallback:
if the native library is missing, every entry point raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# TMRX_LIB selects another build of the same library (the sanitizer build of
# tmr_amd.build.build_sanitized, for host-code tests)
LIB_PATH = os.environ.get("TMRX_LIB") or os.path.join(_HERE, "libtmrx.so")


class Octant(C.Structure):
    """Same layout as TMROctant (src/TMROctant.h:48-53)."""
    _fields_ = [("block", C.c_int32), ("x", C.c_int32), ("y", C.c_int32),
                ("z", C.c_int32), ("tag", C.c_int32), ("level", C.c_int16),
                ("info", C.c_int16)]


class Material(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("ys", C.c_double), ("q", C.c_double),
                ("k0", C.c_double), ("eps", C.c_double), ("penalty_type", C.c_int)]


class SolveOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("max_iters", C.c_int),
                ("precision", C.c_int), ("cheb_degree", C.c_int),
                ("cheb_lower", C.c_double), ("cheb_upper", C.c_double),
                ("eig_iters", C.c_int), ("monitor", C.c_int), ("zero_guess", C.c_int),
                ("fp32_valu", C.c_int), ("no_graph", C.c_int),
                ("cheb_pre_degree", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("rnorm0", C.c_double),
                ("rnorm", C.c_double), ("setup_ms", C.c_double), ("solve_ms", C.c_double)]


P = C.c_void_p
I = C.c_int
D = C.c_double
IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)
CIP = C.POINTER(IP)
CDP = C.POINTER(DP)
S = C.c_char_p

# name: (restype, argtypes)
_SIGS = {
    "tmrx_last_error": (S, []),
    "tmrx_version": (S, []),
    "tmrx_forest_create": (I, [I, I, C.POINTER(P)]),
    "tmrx_forest_free": (None, [P]),
    "tmrx_forest_set_connectivity": (I, [P, I, IP, I]),
    "tmrx_forest_get_connectivity": (I, [P, IP, IP, IP, IP, CIP, CIP, CIP, CIP]),
    "tmrx_forest_set_node_points": (I, [P, DP]),
    "tmrx_forest_set_name": (I, [P, I, I, S]),
    "tmrx_forest_create_trees": (I, [P, I]),
    "tmrx_forest_create_trees_per_block": (I, [P, IP, I]),
    "tmrx_forest_set_octants": (I, [P, P, I]),
    "tmrx_forest_refine": (I, [P, IP, I, I]),
    "tmrx_forest_balance": (I, [P, I]),
    "tmrx_forest_write_vtk": (I, [P, C.c_char_p, P, I, C.c_char_p, P, I, C.c_char_p]),
    "tmrx_check_tile_schedule": (I, [I, P, P, P, C.c_char_p, I]),
    "tmrx_forest_balance_device": (I, [P, I, I, DP]),
    "tmrx_forest_coarsen": (I, [P, C.POINTER(P)]),
    "tmrx_forest_duplicate": (I, [P, C.POINTER(P)]),
    "tmrx_forest_repartition": (I, [P]),
    "tmrx_forest_create_nodes": (I, [P]),
    "tmrx_forest_create_nodes_device": (I, [P, I, DP]),
    "tmrx_forest_get_octants": (I, [P, C.POINTER(C.POINTER(Octant)), IP]),
    "tmrx_forest_get_node_conn": (I, [P, CIP, IP, IP, IP]),
    "tmrx_forest_set_mesh_order": (I, [P, I, I]),
    "tmrx_forest_get_dep_node_conn": (I, [P, CIP, CIP, CDP]),
    "tmrx_forest_get_owned_node_range": (I, [P, IP]),
    "tmrx_forest_get_node_numbers": (I, [P, CIP, IP]),
    "tmrx_forest_get_points": (I, [P, CDP, IP]),
    "tmrx_forest_get_nodes_with_name": (I, [P, S, IP, I]),
    "tmrx_forest_get_octs_with_name": (I, [P, S, C.POINTER(Octant), I]),
    "tmrx_forest_create_interpolation": (I, [P, P, IP, IP, CIP, CIP, CDP]),
    "tmrx_forest_create_interpolation_device": (I, [P, P, I, IP, IP, CIP, CIP, CDP, DP]),
    "tmrx_forest_create_interpolation_geometric": (I, [P, P, IP, IP, CIP, CIP, CDP]),
    "tmrx_forest_element_center_weights": (I, [P, P, IP, IP, CIP, CIP, CDP]),
    "tmrx_forest_interpolate_vec": (I, [P, P, I, P, P, I]),
    "tmrx_forest_generate_bin_file": (I, [P, S, P, I, I, D, IP]),
    "tmrx_convert_bin_to_stl": (I, [S, S]),
    "tmrx_filter_create": (I, [P, I, D, I, I, C.POINTER(P)]),
    "tmrx_filter_create_custom": (I, [I, I, P, P, P, P, I, P]),
    "tmrx_filter_free": (None, [P]),
    "tmrx_filter_apply": (I, [P, I, P, P, I]),
    "tmrx_filter_apply_transpose": (I, [P, I, P, P, I]),
    "tmrx_filter_info": (I, [P, IP, IP]),
    "tmrx_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
    "tmrx_problem_create": (I, [I, C.POINTER(P), C.POINTER(Material), I, C.POINTER(S), IP, I,
                                C.POINTER(P)]),
    "tmrx_problem_free": (None, [P]),
    "tmrx_mg_create": (I, [P, D, I, I, I]),
    "tmrx_problem_smoother_order": (I, [P, I, IP, IP, IP, IP]),
    "tmrx_problem_level_matrix": (I, [P, I, IP, IP, P, P, P]),
    "tmrx_problem_get_sizes": (I, [P, IP, IP, IP]),
    "tmrx_problem_set_stream": (I, [P, P]),
    "tmrx_problem_set_design": (I, [P, P, I, I]),
    "tmrx_problem_set_design_nodal": (I, [P, P, I, I]),
    "tmrx_problem_set_bc_values": (I, [P, P, I, I]),
    "tmrx_problem_design_info": (I, [P, IP, IP]),
    "tmrx_problem_level_design": (I, [P, I, DP]),
    "tmrx_problem_setup": (I, [P, C.POINTER(SolveOpts)]),
    "tmrx_problem_solve": (I, [P, P, P, I, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
    "tmrx_problem_solve_multi": (I, [P, I, P, P, I, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
    "tmrx_problem_num_components": (I, [P, IP]),
    "tmrx_problem_component_info": (I, [P, I, I, IP, IP, C.POINTER(C.c_double)]),
    "tmrx_problem_autotune": (I, [P, IP, I, I, IP, C.POINTER(C.c_double)]),
    "tmrx_problem_tuning": (I, [P, IP, IP]),
    "tmrx_problem_set_fp32_form": (I, [P, I]),
    "tmrx_problem_tile_info": (I, [P, I, IP, IP]),
    "tmrx_problem_set_tiling": (I, [P, I]),
    "tmrx_problem_true_residual": (I, [P, DP]),
    "tmrx_problem_set_state": (I, [P, P, P, I]),
    "tmrx_problem_compliance": (I, [P, DP]),
    "tmrx_problem_compliance_grad": (I, [P, P, I]),
    "tmrx_problem_apply": (I, [P, P, P, I, I]),
    "tmrx_problem_mg_op": (I, [P, I, I, P, P, I, P, P]),
    "tmrx_problem_traction_load": (I, [P, S, DP, DP]),
    "tmrx_problem_stress_pnorm": (I, [P, D, DP]),
    "tmrx_problem_stress_pnorm_noalloc": (I, [P, D, DP]),
    "tmrx_problem_stress_pnorm_grad": (I, [P, D, P, I, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
    "tmrx_problem_kernel_timing": (I, [P, I]),
    "tmrx_problem_kernel_timing_result": (I, [P, DP, IP, IP]),
}

_lib = None


def lib():
    """Load libtmrx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtmrx.so not found at %s; run tmr_amd.build.build()" % LIB_PATH)
        # One HIP runtime per process: torch ships its own libamdhip64 under
        # the same soname, and a second copy loaded after ours cannot open
        # the GPU.  Loading torch first makes this library bind to torch's.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def register(name, restype, argtypes):
    fn = getattr(lib(), name)
    fn.restype = restype
    fn.argtypes = argtypes
    _SIGS[name] = (restype, argtypes)
    return fn


class TmrxError(RuntimeError):
    pass


def check(status, what=""):
    if status < 0:
        msg = lib().tmrx_last_error().decode()
        raise TmrxError("%s: %s" % (what, msg) if what else msg)
    return status


def declared_symbols():
    """Names of every function declared in include/tmrx.h."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "tmrx.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(tmrx_[a-z0-9_]+)\s*\(", txt)))
