"""ctypes bindings of ``libsqdft_amd.so`` (C-ABI declared in include/*.h).

The product path is the HIP library; there is no CPU fallback.  Importing
this module raises if the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsqdft_amd.so")

c_int_p = C.POINTER(C.c_int)
c_dbl_p = C.POINTER(C.c_double)
c_long_p = C.POINTER(C.c_long)


class TypeSpec(C.Structure):
    _fields_ = [("symbol", C.c_char_p), ("Z", C.c_int), ("lloc", C.c_int),
                ("psd_path", C.c_char_p), ("natm", C.c_int), ("xyz", c_dbl_p), ("rb", C.c_double)]


class ScfParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxiter", C.c_int), ("beta", C.c_double),
                ("m", C.c_int), ("p", C.c_int), ("scf_tol", C.c_double), ("scf_miniter", C.c_int),
                ("scf_maxiter", C.c_int), ("beta_scf", C.c_double), ("m_scf", C.c_int), ("p_scf", C.c_int)]


class LocalSpec(C.Structure):
    _fields_ = [("n_typ", C.c_int), ("size", c_int_p), ("r", C.POINTER(c_dbl_p)), ("rvloc", C.POINTER(c_dbl_p)),
                ("dvloc", C.POINTER(c_dbl_p)), ("Z", c_dbl_p), ("n_img", C.c_int), ("img_xyz", c_dbl_p),
                ("img_type", c_int_p), ("img_atom", c_int_p), ("img_box", c_int_p), ("n_atm", C.c_int),
                ("uu", C.POINTER(c_dbl_p)), ("duu", C.POINTER(c_dbl_p))]


class Spec(C.Structure):
    _fields_ = [("n", C.c_int), ("L", C.c_double), ("h", C.c_double), ("fd_order", C.c_int),
                ("Rcut", C.c_double), ("npl", C.c_int), ("T", C.c_double),
                ("lanczos_tol", C.c_double), ("fermi_tol", C.c_double), ("fermi_alg", C.c_int),
                ("n_typ", C.c_int), ("types", C.POINTER(TypeSpec))]


class Info(C.Structure):
    _fields_ = [("n", C.c_int), ("FDn", C.c_int), ("nloc", C.c_int), ("npl", C.c_int),
                ("fermi_alg", C.c_int), ("n_typ", C.c_int), ("n_atm", C.c_int), ("n_img", C.c_int),
                ("L", C.c_double), ("h", C.c_double), ("Rcut", C.c_double), ("bet", C.c_double),
                ("lanczos_tol", C.c_double), ("fermi_tol", C.c_double),
                ("coeff_lap", C.c_double * 16), ("coeff_grad", C.c_double * 16),
                ("udv_len", C.c_long), ("max_nproj", C.c_int)]


class TypeInfo(C.Structure):
    _fields_ = [("Z", C.c_int), ("lloc", C.c_int), ("lmax", C.c_int), ("natm", C.c_int),
                ("nlocc", C.c_int), ("nproj", C.c_int), ("rc", C.c_double), ("rz", C.c_double),
                ("denom", C.c_double * 4), ("size", C.c_int), ("rb", C.c_double)]


class Params(C.Structure):
    _fields_ = [("n", C.c_int), ("h", C.c_double), ("FDn", C.c_int),
                ("coeff_lap", C.c_double * 16), ("coeff_grad", C.c_double * 16),
                ("nloc", C.c_int), ("npl", C.c_int), ("bet", C.c_double),
                ("lanczos_tol", C.c_double), ("fermi_tol", C.c_double), ("fermi_alg", C.c_int),
                ("precision", C.c_int), ("batch_nodes", C.c_int), ("ckpt_interval", C.c_int),
                ("device", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("t_lanczos", C.c_double), ("t_cheb", C.c_double), ("t_fermi", C.c_double),
                ("t_rho", C.c_double), ("t_total", C.c_double),
                ("lanczos_iters_node0", C.c_int), ("lanczos_iters_min", C.c_int),
                ("lanczos_iters_max", C.c_int), ("lanczos_passes", C.c_long),
                ("fermi_iters", C.c_int), ("lambda_min", C.c_double), ("lambda_max", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


# name -> (restype, argtypes); mirrors include/sqdft_system.h and include/sqdft_gpu.h
SIGNATURES = {
    # sqdft_system.h
    "sqdft_system_create": (C.c_int, [C.POINTER(Spec), C.POINTER(C.c_void_p)]),
    "sqdft_system_from_files": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "sqdft_system_destroy": (None, [C.c_void_p]),
    "sqdft_system_info": (C.c_int, [C.c_void_p, C.POINTER(Info)]),
    "sqdft_system_type_info": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(TypeInfo)]),
    "sqdft_system_radial": (C.c_int, [C.c_void_p, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p]),
    "sqdft_system_atoms": (C.c_int, [C.c_void_p, c_dbl_p, c_int_p]),
    "sqdft_system_images": (C.c_int, [C.c_void_p, c_dbl_p, c_int_p, c_int_p, c_int_p, c_long_p]),
    "sqdft_system_udv": (C.c_int, [C.c_void_p, c_dbl_p]),
    "sqg_poisson_aaj": (C.c_int, [C.c_void_p, c_dbl_p, c_dbl_p, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int,
                                  c_dbl_p, c_int_p, c_dbl_p]),
    "sqg_vxc_lda": (C.c_int, [C.c_void_p, c_dbl_p, c_dbl_p]),
    "sqg_total_energy": (C.c_int, [C.c_void_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_int, c_dbl_p]),
    "sqdft_system_pseudocharge": (C.c_int, [C.c_void_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p,
                                            c_dbl_p]),
    "sqg_forces_local": (C.c_int, [C.c_void_p, C.POINTER(LocalSpec), c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p,
                                   c_dbl_p]),
    "sqg_pseudocharge": (C.c_int, [C.c_void_p, C.POINTER(LocalSpec), c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p,
                                   c_dbl_p, c_dbl_p, c_dbl_p]),
    "sqdft_system_uu_deriv": (C.c_int, [C.c_void_p, C.c_int, c_dbl_p]),
    "sqdft_system_set_positions": (C.c_int, [C.c_void_p, c_dbl_p]),
    "sqdft_system_rb_images": (C.c_int, [C.c_void_p, c_int_p, c_dbl_p, c_int_p, c_int_p, c_int_p]),
    "sqdft_system_vloc_deriv": (C.c_int, [C.c_void_p, C.c_int, c_dbl_p]),
    "sqdft_system_scf_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sqdft_spline_eval": (C.c_int, [c_dbl_p, c_dbl_p, c_dbl_p, C.c_int, c_dbl_p, c_dbl_p, C.c_int]),
    "sqdft_spline_derivs": (C.c_int, [c_dbl_p, c_dbl_p, c_dbl_p, C.c_int]),
    # sqdft_gpu.h
    "sqg_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "sqg_destroy": (C.c_int, [C.c_void_p]),
    "sqg_set_geometry": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p, c_dbl_p, C.c_int, C.c_int,
                                   c_int_p, c_int_p, c_int_p, c_dbl_p]),
    "sqg_density": (C.c_int, [C.c_void_p, c_dbl_p, C.c_int, C.c_double, c_dbl_p, c_dbl_p, c_dbl_p,
                              c_dbl_p, C.POINTER(Stats)]),
    "sqg_forces_nonlocal": (C.c_int, [C.c_void_p, c_dbl_p]),
    "sqg_get_node_data": (C.c_int, [C.c_void_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, c_int_p]),
    "sqg_node_moments": (C.c_int, [C.c_void_p, c_dbl_p, c_int_p, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p,
                                   c_int_p]),
    "sqg_hsub_apply_batch": (C.c_int, [C.c_void_p, c_dbl_p, c_int_p, C.c_int, c_dbl_p, c_dbl_p]),
    # veff, v, hv are device addresses
    "sqg_hsub_apply_batch_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, c_int_p, C.c_int, C.c_void_p,
                                               C.c_void_p]),
    "sqg_last_kernel_timing": (C.c_int, [C.c_void_p, C.c_int, c_long_p, c_dbl_p, c_dbl_p]),
    "sqg_density_fixed": (C.c_int, [C.c_void_p, c_dbl_p, C.c_double, C.c_double, C.c_double, c_dbl_p,
                                    c_dbl_p]),
    "sqg_adjoint_veff": (C.c_int, [C.c_void_p, c_dbl_p, C.c_int, c_dbl_p]),
    "sqg_tangent_veff": (C.c_int, [C.c_void_p, c_dbl_p, C.c_int, c_dbl_p]),
}

STATUS = {
    0: "SQG_OK", -1: "SQG_E_ARG", -2: "SQG_E_HIP", -3: "SQG_E_NOGEOM", -4: "SQG_E_NPL",
    -5: "SQG_E_NEG_RHO", -6: "SQG_E_QL", -7: "SQG_E_BRACKET", -8: "SQG_E_BRENT_ITMAX",
    -9: "SQG_E_DENOM", -10: "SQG_E_CUBE", -11: "SQG_E_OVERFLOW", -12: "SQG_E_NOSTATE",
    -13: "SQG_E_FDN", -14: "SQG_E_SPLINE", 1: "SQG_W_LANCZOS_MAXIT", 2: "SQG_W_NEWTON_MAXIT",
}

_lib = None


def load() -> C.CDLL:
    """Loads the HIP library; raises ImportError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m sqdft_amd.build` first")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class SQError(RuntimeError):
    def __init__(self, where: str, code: int):
        super().__init__(f"{where} returned {code} ({STATUS.get(code, '?')})")
        self.code = code


def check(where: str, code: int, allow_warn: bool = True) -> int:
    if code < 0 or (code > 0 and not allow_warn):
        raise SQError(where, code)
    return code


def dptr(a) -> c_dbl_p:
    return a.ctypes.data_as(c_dbl_p)


def iptr(a) -> c_int_p:
    return a.ctypes.data_as(c_int_p)


def lptr(a) -> c_long_p:
    return a.ctypes.data_as(c_long_p)
