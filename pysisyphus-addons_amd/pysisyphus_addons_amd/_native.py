"""ctypes binding of libdfhip.so (see include/dfhip.h)."""

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(os.environ.get("DFHIP_LIB", Path(__file__).resolve().parent / "libdfhip.so"))

_lib = None


class DfhipError(RuntimeError):
    """Raised when a libdfhip.so call reports a non-zero status."""

    def __init__(self, status, msg):
        super().__init__(f"dfhip status {status}: {msg}")
        self.status = status


class Basis(ctypes.Structure):
    _fields_ = [
        ("nshells", ctypes.c_int32),
        ("ndata", ctypes.c_int32),
        ("bas_centers", ctypes.c_void_p),
        ("bas_spec", ctypes.c_void_p),
        ("bas_data", ctypes.c_void_p),
    ]


class Options(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("screening", ctypes.c_int32),
        ("thresh", ctypes.c_double),
        ("dens_layout", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
        ("max_tile_bytes", ctypes.c_int64),
        ("dens_on_device", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("fp32_below", ctypes.c_double),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("ntriplets", ctypes.c_int64),
        ("ncomputed", ctypes.c_int64),
        ("nskipped", ctypes.c_int64),
        ("nbf", ctypes.c_int64),
        ("naux", ctypes.c_int64),
        ("npair_functions", ctypes.c_int64),
        ("ncolumn_blocks", ctypes.c_int32),
    ] + [(n, ctypes.c_double) for n in (
        "ms_upload", "ms_schwarz", "ms_metric", "ms_cholesky", "ms_int3c2e", "ms_pack",
        "ms_contract", "ms_solve", "ms_assemble", "ms_download", "ms_total",
        "ms_h2d_densities", "ms_host_plan")] + [
        ("class_triplets", ctypes.c_int64 * 90),
        ("class_prim_triplets", ctypes.c_int64 * 90),
        ("ncomputed_fp32", ctypes.c_int64),
        ("npairs_zero_q", ctypes.c_int64),
    ]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name.startswith("class_") else v
        return d


FP64, MIXED = 0, 1
SCREEN_SCHWARZ, SCREEN_NONE, SCREEN_QVL = 0, 1, 2
DENS_FORTRAN, DENS_C = 0, 1

EXPORTED = (
    "f_init", "f_contract_coulomb_densities_2d", "f_contract_coulomb_densities_4d",
    "f_int_schwarz_aux", "f_int_schwarz_bra", "dfhip_last_status", "dfhip_last_error",
    "dfhip_default_options", "dfhip_contract_coulomb_densities", "dfhip_int3c2e_dense",
    "dfhip_metric", "dfhip_boys", "dfhip_release_workspace", "dfhip_contract_coulomb_densities_batch",
    "dfhip_qvl_records", "dfhip_int_schwarz_aux_noalloc", "f_intor_df1c", "f_intor_eri2c", "f_intor_df2c_mo",
    "f_benchmark_int3c2e", "dfhip_metric_powerh", "dfhip_int3c2e_estimates", "f_eval_densities",
    "dfhip_autotune", "dfhip_load_tuning", "dfhip_tuning_choice", "dfhip_tuning_contract", "dfhip_schwarz_bounds",
)


def lib():
    """Load libdfhip.so; raises ImportError when it is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} not found: build it with __graft_entry__.build()")
        L = ctypes.CDLL(str(LIB_PATH))
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.dfhip_last_status.restype = ctypes.c_int
        L.dfhip_last_error.restype = ctypes.c_char_p
        L.dfhip_default_options.argtypes = [ctypes.POINTER(Options)]
        L.dfhip_contract_coulomb_densities.restype = ctypes.c_int
        L.dfhip_contract_coulomb_densities.argtypes = [
            ctypes.POINTER(Basis), ctypes.POINTER(Basis), i32, i32, vp, ctypes.POINTER(Options),
            vp, i32, vp, ctypes.POINTER(Stats)]
        L.dfhip_contract_coulomb_densities_batch.restype = ctypes.c_int
        L.dfhip_contract_coulomb_densities_batch.argtypes = [
            i32, ctypes.POINTER(Basis), ctypes.POINTER(Basis), vp, vp, vp, ctypes.POINTER(Options), vp, vp,
            ctypes.POINTER(Stats), i32]
        L.dfhip_int3c2e_dense.restype = ctypes.c_int
        L.dfhip_int3c2e_dense.argtypes = [ctypes.POINTER(Basis), ctypes.POINTER(Basis),
                                          ctypes.POINTER(Options), vp, ctypes.POINTER(Stats)]
        L.dfhip_int_schwarz_aux_noalloc.restype = ctypes.c_int
        L.dfhip_int_schwarz_aux_noalloc.argtypes = [ctypes.POINTER(Basis), vp]
        L.dfhip_qvl_records.restype = ctypes.c_int
        L.dfhip_qvl_records.argtypes = [ctypes.POINTER(Basis), ctypes.POINTER(Basis), vp, vp]
        L.dfhip_metric.restype = ctypes.c_int
        L.dfhip_metric.argtypes = [ctypes.POINTER(Basis), vp]
        L.dfhip_schwarz_bounds.restype = ctypes.c_int
        L.dfhip_schwarz_bounds.argtypes = [ctypes.POINTER(Basis), ctypes.POINTER(Basis), vp, vp]
        L.dfhip_boys.restype = ctypes.c_int
        L.dfhip_boys.argtypes = [i64, vp, vp, vp]
        args13 = [i32, i32, vp, vp, vp, i32, i32, vp, vp, vp, i32, i32, vp]
        L.f_contract_coulomb_densities_2d.argtypes = args13 + [i32, vp]
        L.f_contract_coulomb_densities_4d.argtypes = args13 + [i32, vp]
        L.f_intor_df1c.argtypes = args13[:10] + [i32, i32, vp]
        L.f_intor_eri2c.argtypes = args13[:10] + [i32, vp]
        L.f_intor_df2c_mo.argtypes = args13[:10] + [i32, i32, vp, i32, vp]
        L.f_benchmark_int3c2e.argtypes = args13[:10]
        L.dfhip_metric_powerh.restype = ctypes.c_int
        L.dfhip_metric_powerh.argtypes = [ctypes.POINTER(Basis), ctypes.c_double, vp]
        L.dfhip_int3c2e_estimates.restype = ctypes.c_int
        L.dfhip_int3c2e_estimates.argtypes = [ctypes.POINTER(Basis), ctypes.POINTER(Basis), vp, i64, vp]
        L.f_int_schwarz_aux.argtypes = [i32, i32, vp, vp, vp, vp]
        L.f_int_schwarz_bra.argtypes = [i32, i32, vp, vp, vp, vp]
        L.f_eval_densities.argtypes = [i32, i32, vp, vp, vp, i32, vp, i32, i32, vp, vp, i32, ctypes.c_double]
        L.dfhip_autotune.restype = ctypes.c_int
        L.dfhip_autotune.argtypes = [ctypes.POINTER(Basis), ctypes.POINTER(Basis), i32, i32, vp,
                                     ctypes.POINTER(Options), ctypes.c_char_p]
        L.dfhip_load_tuning.restype = ctypes.c_int
        L.dfhip_load_tuning.argtypes = [ctypes.c_char_p]
        L.dfhip_tuning_choice.restype = i32
        L.dfhip_tuning_choice.argtypes = [i32, i32, i32, i32]
        L.dfhip_tuning_contract.restype = i32
        L.dfhip_tuning_contract.argtypes = [i32, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def check(status=None):
    L = lib()
    st = L.dfhip_last_status() if status is None else status
    if st != 0:
        raise DfhipError(st, L.dfhip_last_error().decode(errors="replace"))


def default_options(**kw):
    o = Options()
    lib().dfhip_default_options(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class BasisArrays:
    """Keeps Fortran-ordered int32/float64 copies of the basis arrays alive for
    the duration of a call (the reference's Cython layer makes the same
    np.asfortranarray copies, diabatization/_intor.pyx:133-145)."""

    def __init__(self, bas_centers, bas_spec, bas_data):
        self.centers = np.asfortranarray(bas_centers, dtype=np.int32)
        self.spec = np.asfortranarray(bas_spec, dtype=np.int32)
        self.data = np.ascontiguousarray(bas_data, dtype=np.float64)
        if self.centers.ndim != 2 or self.centers.shape[1] != 3:
            raise ValueError("bas_centers must have shape (nshells, 3)")
        if self.spec.ndim != 2 or self.spec.shape[1] != 5 or len(self.spec) != len(self.centers):
            raise ValueError("bas_spec must have shape (nshells, 5)")
        self.struct = Basis(len(self.centers), self.data.size, self.centers.ctypes.data,
                            self.spec.ctypes.data, self.data.ctypes.data)

    @property
    def nsph(self):
        return int(np.sum(2 * self.spec[:, 1] + 1))

    def args(self):
        return (len(self.centers), self.data.size, self.centers.ctypes.data, self.spec.ctypes.data,
                self.data.ctypes.data)
