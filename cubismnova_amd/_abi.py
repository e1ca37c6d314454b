"""ctypes declarations of the C ABI in include/cubismnova_amd.h.

The shared library is loaded from this package directory.  There is no
fallback: if the library is missing the import of the compute API raises, so
that nothing silently runs on the CPU.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcubismnova_amd.so")

# status codes (enum cnv_status)
CNV_OK = 0
CNV_ERR_INVALID = 1
CNV_ERR_ALLOC = 2
CNV_ERR_HIP = 3
CNV_ERR_UNSUPPORTED = 4
CNV_ERR_NOT_ALLOCATED = 5
CNV_ERR_SIZE = 6

CNV_F32 = 0
CNV_F64 = 1
CNV_CELL, CNV_NODE, CNV_FACE = 0, 1, 2

CNV_BC_BASE = 0
CNV_BC_ABSORBING = 1
CNV_BC_DIRICHLET = 2
CNV_BC_SYMMETRY = 3


class cnv_bc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("dir", ctypes.c_int), ("side", ctypes.c_int),
                ("reserved", ctypes.c_int), ("value", ctypes.c_double)]


class cnv_stencil(ctypes.Structure):
    _fields_ = [("begin", ctypes.c_int * 3), ("end", ctypes.c_int * 3), ("tensorial", ctypes.c_int)]


class cnv_grid_desc(ctypes.Structure):
    _fields_ = [("nblocks", ctypes.c_int * 3), ("block_cells", ctypes.c_int * 3),
                ("ncomp", ctypes.c_int), ("dtype", ctypes.c_int),
                ("begin", ctypes.c_double * 3), ("end", ctypes.c_double * 3),
                ("entity", ctypes.c_int), ("reserved", ctypes.c_int)]


class cnv_solve_opts(ctypes.Structure):
    _fields_ = [("refine", ctypes.c_int), ("sweeps_per_cycle", ctypes.c_int), ("max_cycles", ctypes.c_int),
                ("omega", ctypes.c_double), ("tol", ctypes.c_double)]


class cnv_solve_info(ctypes.Structure):
    _fields_ = [("cycles", ctypes.c_int), ("sweeps", ctypes.c_longlong), ("rel_residual", ctypes.c_double)]


class cnv_batch_item(ctypes.Structure):
    _fields_ = [("grid_in", ctypes.c_void_p), ("grid_out", ctypes.c_void_p), ("op", ctypes.c_int),
                ("ghost", ctypes.c_int), ("nu_dt", ctypes.c_double)]


CNV_OP_DIFFUSION = 0
CNV_OP_WENO5 = 1


c_int_p = ctypes.POINTER(ctypes.c_int)
vp = ctypes.c_void_p
i32 = ctypes.c_int
sz = ctypes.c_size_t
f64 = ctypes.c_double

# name -> (restype, argtypes); every entry point of the header
SIGNATURES = {
    "cnv_strerror": (ctypes.c_char_p, [i32]),
    "cnv_last_error": (ctypes.c_char_p, []),
    "cnv_device_count": (i32, [c_int_p]),
    "cnv_malloc": (i32, [ctypes.POINTER(vp), sz]),
    "cnv_free": (i32, [vp]),
    "cnv_memcpy_h2d": (i32, [vp, vp, sz, vp]),
    "cnv_memcpy_d2h": (i32, [vp, vp, sz, vp]),
    "cnv_memcpy_d2d": (i32, [vp, vp, sz, vp]),
    "cnv_fill": (i32, [vp, i32, f64, sz, vp]),
    "cnv_memcpy_stream": (i32, [vp, vp, sz, vp]),
    "cnv_memcpy_fanout": (i32, [vp, sz, vp, sz, i32, vp]),
    "cnv_stream_create": (i32, [ctypes.POINTER(vp)]),
    "cnv_stream_destroy": (i32, [vp]),
    "cnv_stream_synchronize": (i32, [vp]),
    "cnv_grid_create": (i32, [ctypes.POINTER(cnv_grid_desc), ctypes.POINTER(vp)]),
    "cnv_grid_wrap": (i32, [ctypes.POINTER(cnv_grid_desc), vp, sz, ctypes.POINTER(vp)]),
    "cnv_grid_write_hdf": (i32, [vp, ctypes.c_char_p, ctypes.c_char_p, i32, f64, i32, i32, c_int_p, c_int_p, vp]),
    "cnv_grid_read_hdf": (i32, [vp, ctypes.c_char_p, i32, i32, c_int_p, c_int_p, vp]),
    "cnv_grid_destroy": (i32, [vp]),
    "cnv_grid_info": (i32, [vp, ctypes.POINTER(cnv_grid_desc), ctypes.POINTER(sz), ctypes.POINTER(sz)]),
    "cnv_grid_data": (i32, [vp, i32, ctypes.POINTER(vp)]),
    "cnv_grid_upload": (i32, [vp, i32, vp, sz, vp]),
    "cnv_grid_download": (i32, [vp, i32, vp, sz, vp]),
    "cnv_grid_set_block_bcs": (i32, [vp, i32, c_int_p, ctypes.POINTER(cnv_bc), i32]),
    "cnv_grid_add_face_bc": (i32, [vp, i32, ctypes.POINTER(cnv_bc)]),
    "cnv_grid_clear_bcs": (i32, [vp, i32]),
    "cnv_lab_extent": (i32, [vp, ctypes.POINTER(cnv_stencil), c_int_p]),
    "cnv_lab_geometry": (i32, [vp, i32, c_int_p, ctypes.POINTER(cnv_stencil), c_int_p, c_int_p]),
    "cnv_lab_load": (i32, [vp, i32, c_int_p, ctypes.POINTER(cnv_stencil), i32, ctypes.POINTER(cnv_bc), i32,
                           vp, vp]),
    "cnv_lab_load_all": (i32, [vp, i32, ctypes.POINTER(cnv_stencil), vp, vp]),
    "cnv_laplacian": (i32, [vp, i32, vp, i32, vp]),
    "cnv_diffusion_step": (i32, [vp, vp, f64, vp, vp, vp]),
    "cnv_diffusion_step_noalloc": (i32, [ctypes.POINTER(cnv_grid_desc), vp, vp, f64, vp, vp]),
    "cnv_diffusion_run": (i32, [vp, vp, i32, f64, vp, vp]),
    "cnv_grid_sums": (i32, [vp, i32, vp, vp]),
    "cnv_autotune_stencil7": (i32, [vp, vp, i32, c_int_p, c_int_p, vp, i32, c_int_p]),
    "cnv_tune_set_stencil7": (i32, [i32, i32, i32, i32]),
    "cnv_tune_get_stencil7": (i32, [i32, i32, c_int_p, c_int_p]),
    "cnv_autotune_diffusion_run": (i32, [vp, vp, i32, c_int_p, c_int_p, vp, i32, c_int_p]),
    "cnv_tune_set_diffusion_run": (i32, [i32, i32, i32, i32]),
    "cnv_tune_get_diffusion_run": (i32, [i32, i32, c_int_p, c_int_p]),
    "cnv_grid_compare": (i32, [vp, i32, vp, i32, vp, vp]),
    "cnv_grid_set_zfaces": (i32, [vp, vp, i32]),
    "cnv_jacobi_sweep": (i32, [vp, i32, vp, i32, vp, i32, f64, vp]),
    "cnv_jacobi_defect": (i32, [vp, i32, vp, i32, vp, i32, vp, vp]),
    "cnv_jacobi_sweep2": (i32, [vp, i32, vp, i32, vp, i32, f64, vp]),
    "cnv_tune_set_jacobi2": (i32, [i32, i32, i32, i32, i32]),
    "cnv_tune_get_jacobi2": (i32, [i32, i32, c_int_p, c_int_p, c_int_p]),
    "cnv_add_correction": (i32, [vp, i32, vp, i32, vp]),
    "cnv_autotune_jacobi": (i32, [vp, vp, vp, i32, c_int_p, c_int_p, c_int_p, vp, i32, c_int_p]),
    "cnv_tune_set_jacobi": (i32, [i32, i32, i32, i32, i32]),
    "cnv_tune_get_jacobi": (i32, [i32, i32, c_int_p, c_int_p, c_int_p]),
    "cnv_jacobi_solve": (i32, [vp, vp, vp, vp, vp, vp, i32, vp]),
    "cnv_weno5": (i32, [vp, vp, vp]),
    "cnv_autotune_weno5": (i32, [vp, vp, i32, c_int_p, c_int_p, vp, i32, c_int_p]),
    "cnv_tune_set_weno5": (i32, [i32, i32, i32, i32]),
    "cnv_tune_get_weno5": (i32, [i32, i32, c_int_p, c_int_p]),
    "cnv_batch_create": (i32, [vp, i32, ctypes.POINTER(vp)]),
    "cnv_batch_autotune": (i32, [vp, i32, vp, i32, c_int_p]),
    "cnv_batch_run": (i32, [vp, vp]),
    "cnv_batch_info": (i32, [vp, c_int_p, c_int_p, ctypes.POINTER(sz)]),
    "cnv_batch_class_info": (i32, [vp, i32, c_int_p, c_int_p]),
    "cnv_batch_form": (i32, [vp, c_int_p, c_int_p]),
    "cnv_batch_set_form": (i32, [vp, i32]),
    "cnv_batch_destroy": (i32, [vp]),
    "cnv_field_op2": (i32, [i32, i32, vp, vp, vp, sz, vp]),
    "cnv_field_op1": (i32, [i32, i32, vp, f64, vp, sz, vp]),
    "cnv_field_rcp": (i32, [i32, vp, f64, vp, sz, vp]),
    "cnv_grid_field_op": (i32, [i32, vp, i32, vp, i32, vp, i32, ctypes.c_double, vp]),
    "cnv_grid_prepare": (i32, [vp]),
    "cnv_graph_begin": (i32, [vp]),
    "cnv_graph_end": (i32, [vp, ctypes.POINTER(vp)]),
    "cnv_graph_launch": (i32, [vp, vp]),
    "cnv_graph_destroy": (i32, [vp]),
}

# the reference's field-op hooks (FieldOperator.h:149-227): void, blocking
for _op in ("Add", "Sub", "Mul", "Div"):
    for _t, _ct in (("double", f64), ("float", ctypes.c_float)):
        SIGNATURES["field%s2_%s" % (_op, _t)] = (None, [vp, vp, vp, sz])
        SIGNATURES["field%s1_%s" % (_op, _t)] = (None, [vp, _ct, vp, sz])
SIGNATURES["fieldRcp_double"] = (None, [vp, f64, vp, sz])
SIGNATURES["fieldRcp_float"] = (None, [vp, ctypes.c_float, vp, sz])

CNV_FIELD_ADD, CNV_FIELD_SUB, CNV_FIELD_MUL, CNV_FIELD_DIV, CNV_FIELD_RCP, CNV_FIELD_NEG = 0, 1, 2, 3, 4, 5

_lib = None


class CnvError(RuntimeError):
    def __init__(self, status, fn, msg):
        self.status = status
        super().__init__("%s: %s (status %d)" % (fn, msg, status))


def declared_symbols():
    """Names of all entry points declared in include/cubismnova_amd.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "cubismnova_amd.h")
    txt = open(hdr).read()
    names = set(re.findall(r"\b(cnv_[a-z0-9_]+)\s*\(", txt))
    names |= set(re.findall(r"\b(field(?:Add|Sub|Mul|Div)[12]_(?:double|float)|fieldRcp_(?:double|float))\s*\(", txt))
    return sorted(names)


def lib():
    """Load libcubismnova_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libcubismnova_amd.so is not built: run `python -m cubismnova_amd.build` "
                              "(expected at %s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    """Call an entry point; raise CnvError on a non-zero status."""
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != CNV_OK:
        msg = L.cnv_last_error().decode() or L.cnv_strerror(rc).decode()
        raise CnvError(rc, name, msg)
    return rc
