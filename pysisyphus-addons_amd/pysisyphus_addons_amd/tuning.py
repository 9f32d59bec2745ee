"""The library's autotuner: per (la, lb, lc) class and precision window the
faster of the generated VALU and matrix-core 3c2e kernels, and the
contraction's output-tile width and split-K depth, measured on a caller's
workload (include/dfhip.h, dfhip_autotune).  The table lives in the process
and in a text file that a later process loads with load() or through the
DFHIP_TUNING environment variable."""

import ctypes
import os

import numpy as np

from . import _native as N


def autotune(shells, shells_aux, densities, path=None, **options):
    """Tune on this basis pair and C-ordered densities (ndens, nbfs, nbfs);
    options as for _native.default_options (precision, screening, thresh).
    Returns the table as {"int3c2e": {(la, lb, lc, window): family},
    "contract": (ipt, split_scale)}."""
    b = N.BasisArrays(*shells.as_arrays(fortran=True))
    x = N.BasisArrays(*shells_aux.as_arrays(fortran=True))
    D = np.ascontiguousarray(densities, dtype=np.float64)
    opt = N.default_options(dens_layout=N.DENS_C, **options)
    p = os.fsencode(path) if path is not None else None
    N.check(N.lib().dfhip_autotune(ctypes.byref(b.struct), ctypes.byref(x.struct), D.shape[0], D.shape[1],
                                   D.ctypes.data, ctypes.byref(opt), p))
    return table(D.shape[0])


def load(path=None):
    """Replace the process's table by the file's (None clears it)."""
    N.check(N.lib().dfhip_load_tuning(os.fsencode(path) if path is not None else None))


def table(ndens=None):
    L = N.lib()
    fam = {}
    for la in range(5):
        for lb in range(la + 1):
            for lc in range(6):
                for w in range(2):
                    f = L.dfhip_tuning_choice(la, lb, lc, w)
                    if f > 0:
                        fam[(la, lb, lc, w)] = f
    out = {"int3c2e": fam, "contract": None}
    if ndens is not None:
        ipt, sc = ctypes.c_int32(0), ctypes.c_double(0.0)
        if L.dfhip_tuning_contract(ndens, ctypes.byref(ipt), ctypes.byref(sc)):
            out["contract"] = (ipt.value, sc.value)
    return out
