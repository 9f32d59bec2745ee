"""Drop-in for pysisyphus_addons.diabatization.intor (diabatization/intor.py:4-36).

``shells`` only needs ``as_arrays(fortran=True)`` returning
(bas_centers int32 (n,3), bas_spec int32 (n,5), bas_data float64) with
1-based pointers, as pysisyphus' Shells provide.
"""

from . import _intor


def contract_coulomb_densities_2d(shells, shells_aux, densities):
    bas_centers, bas_spec, bas_data = shells.as_arrays(fortran=True)
    bas_centers_aux, bas_spec_aux, bas_data_aux = shells_aux.as_arrays(fortran=True)
    return _intor.contract_coulomb_densities_2d(
        bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux, bas_data_aux, densities
    )


def contract_coulomb_densities_4d(shells, shells_aux, densities):
    bas_centers, bas_spec, bas_data = shells.as_arrays(fortran=True)
    bas_centers_aux, bas_spec_aux, bas_data_aux = shells_aux.as_arrays(fortran=True)
    return _intor.contract_coulomb_densities_4d(
        bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux, bas_data_aux, densities
    )
