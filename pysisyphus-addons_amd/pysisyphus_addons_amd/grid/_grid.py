"""Native layer of the grid densities; same function, arguments, defaults and
return layout as the reference's grid/_grid.pyx:31-75 (grid3d of shape
(npoints, 3), Cartesian densities of shape (ndens, nbfs, nbfs), result of
shape (ndens, npoints) in Fortran order).  Failures raise DfhipError."""

import numpy as np

from .. import _native as N


def eval_densities(bas_centers, bas_spec, bas_data, grid3d, densities, blk_size=100, thresh=1e-8):
    b = N.BasisArrays(bas_centers, bas_spec, bas_data)
    grid3d = np.ascontiguousarray(grid3d, dtype=np.float64)
    if grid3d.ndim != 2 or grid3d.shape[1] != 3:
        raise ValueError("grid3d must have shape (npoints, 3)")
    densities = np.asfortranarray(densities, dtype=np.float64)
    if densities.ndim != 3 or densities.shape[1] != densities.shape[2]:
        raise ValueError("densities must have shape (ndens, nbfs, nbfs)")
    ndens, nbfs = densities.shape[:2]
    npoints = grid3d.shape[0]
    # grid3d (npoints, 3) in C order is the (3, npoints) Fortran array of the reference
    out = np.zeros((ndens, npoints), order="F")
    N.lib().f_eval_densities(*b.args(), npoints, grid3d.ctypes.data, ndens, nbfs, densities.ctypes.data,
                             out.ctypes.data, int(blk_size), float(thresh))
    N.check()
    return out
