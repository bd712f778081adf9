"""ctypes binding of libpyprop_amd.so, the C-ABI of the HIP path.

The library is built in-tree by ``pyprop_amd.build`` (hipcc, gfx950).  Every
entry point returns an int status: 0 ok, <0 argument / unsupported / HIP
errors, >0 numerical failures; :func:`check` turns them into exceptions the
way the reference turns C++ ``runtime_error`` into Python ``RuntimeError``.
There is no CPU fallback: if the library is missing, importing the compute
path raises.
"""
import atexit
import ctypes
import itertools
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpyprop_amd.so")
if os.environ.get("PYPROP_AMD_LIB"):   # an alternative in-tree build (launch-shape tuning)
    LIB_PATH = os.path.abspath(os.environ["PYPROP_AMD_LIB"])

PP_OK, PP_EARG, PP_EUNSUP, PP_EHIP, PP_ETIMEOUT = 0, -1, -2, -3, -4

# storage ids (tensorpotentialmultiply_generator.py:1336)
STORAGE_IDS = {"Ident": 0, "Diag": 1, "Simp": 2, "Herm": 3, "BandNH": 4, "Band": 5}

# overlap operations
OV_SOLVE_SQRT, OV_SOLVE_SQRT_CONJ, OV_MUL_SQRT, OV_MUL_SQRT_CONJ, OV_SOLVE, OV_MUL = range(6)

EXPORTS = [
    "pp_band_cholesky_upper", "pp_overlap_create", "pp_overlap_create_chunked", "pp_overlap_factor", "pp_overlap_destroy",
    "pp_overlap_apply", "pp_tp_multiply", "pp_pamp_create", "pp_pamp_add_term", "pp_pamp_add_factored_term",
    "pp_lowrank_factor", "pp_pamp_propagate", "pp_pamp_diagnostics", "pp_pamp_set_mixed", "pp_pamp_set_fiber_split", "pp_pamp_set_pade_scaling", "pp_expokit_propagate", "pp_expokit_diagnostics", "pp_pamp_batch_propagate", "pp_pamp_batch_graph_create", "pp_pamp_batch_graph_launch", "pp_pamp_batch_graph_destroy",
    "pp_gram_schmidt", "pp_gram_schmidt_noalloc", "pp_inner_product", "pp_basis_combine", "pp_band_lu_solve", "pp_pamp_lds_estimate", "pp_pamp_set_timing", "pp_pamp_timings", "pp_pamp_destroy",
    "pp_pamp3_create", "pp_pamp3_add_kron_term", "pp_pamp3_add_multipole_term", "pp_pamp3_apply",
    "pp_pamp3_propagate", "pp_pamp3_diagnostics", "pp_pamp3_set_timing", "pp_pamp3_kernel_time",
    "pp_pamp3_set_radial_supports", "pp_pamp3_set_mixed",
    "pp_pamp3_multipole_flops", "pp_pamp3_destroy",
    "pp_radial_prec_create", "pp_radial_prec_solve", "pp_radial_prec_destroy",
    "pp_pamp_set_exponentiation", "pp_pamp3_set_exponentiation",
    "pp_projector_create", "pp_projector_apply", "pp_projector_destroy",
]

_lib = None


class NativeError(RuntimeError):
    pass


def check(rc, what):
    if rc == PP_OK:
        return
    names = {PP_EARG: "invalid argument", PP_EUNSUP: "unsupported on the HIP path", PP_EHIP: "HIP runtime error",
             PP_ETIMEOUT: "a grid barrier of the persistent launch timed out (its results are invalid)"}
    raise NativeError("%s failed: %s (status %d)" % (what, names.get(rc, "numerical failure"), rc))


def lib():
    """Load libpyprop_amd.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeError("libpyprop_amd.so not built; run pyprop_amd.build.build() (hipcc, gfx950)")
    # torch ships its own libamdhip64.so.7; importing it first makes the
    # library bind to that same HIP runtime (one runtime per process).  Loaded
    # the other way round, /opt/rocm's runtime comes first and the library's
    # own calls then report "no ROCm-capable device is detected".
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    D = ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.pp_band_cholesky_upper.argtypes = [I, I, dp]
    L.pp_overlap_create.argtypes = [I, I, dp, ctypes.POINTER(P), ip]
    L.pp_overlap_create_chunked.argtypes = [I, I, dp, I, ctypes.POINTER(P), ip]
    L.pp_overlap_factor.argtypes = [P, dp]
    L.pp_overlap_destroy.argtypes = [P]
    L.pp_overlap_apply.argtypes = [P, I, I, I, ip, P, P]
    L.pp_tp_multiply.argtypes = [I, ip, ip, P, D, ip, P, P, ctypes.POINTER(ip), P]
    L.pp_pamp_create.argtypes = [I, I, I, I, P, ctypes.POINTER(P)]
    L.pp_pamp_add_term.argtypes = [P, I, I, ip, I, I, P]
    L.pp_pamp_add_factored_term.argtypes = [P, I, I, ip, I, I, I, dp, dp]
    L.pp_lowrank_factor.argtypes = [I, ctypes.c_long, dp, I, ip, dp, dp, dp]
    L.pp_pamp_propagate.argtypes = [P, P, I, D, D, dp, P]
    L.pp_pamp_diagnostics.argtypes = [P, I, dp, P]
    L.pp_pamp_set_mixed.argtypes = [P, ctypes.c_uint]
    L.pp_pamp_set_fiber_split.argtypes = [P, I]
    L.pp_pamp_set_pade_scaling.argtypes = [P, I]
    L.pp_expokit_propagate.argtypes = [P, P, I, D, D, dp, D, D, P]
    L.pp_expokit_diagnostics.argtypes = [P, I, dp, P]
    L.pp_pamp_batch_propagate.argtypes = [I, ctypes.POINTER(P), ctypes.POINTER(P), I, dp, ctypes.POINTER(dp), P]
    L.pp_pamp_batch_graph_create.argtypes = [I, ctypes.POINTER(P), ctypes.POINTER(P), I, dp, ctypes.POINTER(P)]
    L.pp_pamp_batch_graph_launch.argtypes = [P, ctypes.POINTER(dp), P]
    L.pp_pamp_batch_graph_destroy.argtypes = [P]
    L.pp_gram_schmidt.argtypes = [P, ctypes.c_long, I, P, I, dp, dp, P]
    L.pp_gram_schmidt_noalloc.argtypes = [P, ctypes.c_long, I, P, I, dp, dp, P]
    L.pp_inner_product.argtypes = [P, P, ctypes.c_long, dp, P]
    L.pp_basis_combine.argtypes = [P, ctypes.c_long, I, dp, I, P, P]
    L.pp_band_lu_solve.argtypes = [I, I, I, I, P, ctypes.POINTER(ctypes.c_int), P, P]
    L.pp_pamp_set_timing.argtypes = [P, I]
    L.pp_pamp_timings.argtypes = [P, ctypes.POINTER(ctypes.c_longlong), I, ip]
    L.pp_pamp_destroy.argtypes = [P]
    L.pp_pamp_lds_estimate.argtypes = [I, I, I, I, I, I]
    L.pp_pamp3_create.argtypes = [I, I, I, I, P, ctypes.POINTER(P)]
    L.pp_pamp3_add_kron_term.argtypes = [P, I, ip, dp, dp, dp, I]
    L.pp_pamp3_add_multipole_term.argtypes = [P, I, ip, dp, dp, dp, dp, I]
    L.pp_pamp3_set_radial_supports.argtypes = [P, ip, ip]
    L.pp_pamp3_apply.argtypes = [P, P, P, dp, I, P]
    L.pp_pamp3_propagate.argtypes = [P, P, I, D, D, dp, P]
    L.pp_pamp3_diagnostics.argtypes = [P, I, dp, P]
    L.pp_pamp3_destroy.argtypes = [P]
    L.pp_pamp3_set_timing.argtypes = [P, I]
    L.pp_pamp3_set_mixed.argtypes = [P, I]
    L.pp_pamp3_kernel_time.argtypes = [P, dp, ctypes.POINTER(ctypes.c_long)]
    L.pp_pamp3_multipole_flops.argtypes = [P, dp]
    L.pp_radial_prec_create.argtypes = [I, I, I, ip, ip, dp, dp, ctypes.POINTER(P)]
    L.pp_radial_prec_solve.argtypes = [P, P, D, D, D, D, P]
    L.pp_radial_prec_destroy.argtypes = [P]
    L.pp_pamp_set_exponentiation.argtypes = [P, I]
    L.pp_pamp3_set_exponentiation.argtypes = [P, I]
    L.pp_projector_create.argtypes = [I, I, ip, ip, dp, ctypes.POINTER(P)]
    L.pp_projector_apply.argtypes = [P, ctypes.c_long, P, P, P]
    L.pp_projector_destroy.argtypes = [P]
    for name in EXPORTS:
        getattr(L, name).restype = ctypes.c_int
    L.pp_pamp_lds_estimate.restype = ctypes.c_long
    _lib = L
    return L


# Native objects (overlaps, engines, graphs) owned by Python wrappers.  They
# are destroyed when their wrapper is collected, or, for those still alive
# when the interpreter exits, by an exit hook that runs while the HIP runtime
# is still up, newest first (engines and graphs before the overlaps they
# use).  After that hook no wrapper calls into the library: a finalizer that
# runs during interpreter teardown would otherwise free device memory after
# the runtime (or a profiler attached to it) has shut down.
_live = {}
_serial = itertools.count()
_closed = False


def track(destroy, handle):
    """Register a native handle with its destroy entry point; returns the
    token to release it with."""
    token = next(_serial)
    _live[token] = (destroy, handle)
    return token


def release(token):
    """Destroy a registered handle now; a no-op once the exit hook ran (or
    while the interpreter tears this module down)."""
    if _closed is not False or _live is None or _lib is None:
        return
    item = _live.pop(token, None)
    if item is not None:
        getattr(_lib, item[0])(item[1])


@atexit.register
def _release_all():
    global _closed
    for token in sorted(_live, reverse=True):
        destroy, handle = _live.pop(token)
        if _lib is not None:
            getattr(_lib, destroy)(handle)
    _closed = True


def dptr(a):
    """ctypes double* of a contiguous numpy array."""
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
