"""Drop-in for pysisyphus_addons.grid.grid (grid/grid.py:4-22)."""

from . import _grid


def eval_density(shells, grid3d, density, **kwargs):
    """One density matrix of shape (nbfs, nbfs); returns shape (npoints,)."""
    assert density.ndim == 2
    return eval_densities(shells, grid3d, density[None, ...], **kwargs)[0]


def eval_densities(shells, grid3d, densities, **kwargs):
    bas_centers, bas_spec, bas_data = shells.as_arrays(fortran=True)
    return _grid.eval_densities(bas_centers, bas_spec, bas_data, grid3d, densities, **kwargs)
