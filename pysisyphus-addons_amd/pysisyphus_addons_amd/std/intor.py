"""Drop-in for pysisyphus_addons.std.intor (std/intor.py:9-66).  The
reference's Wavefunction overload of int_eri2c, which loads the
def2-universal-jfit aux basis through pysisyphus, has no counterpart here:
pass both Shells."""

from . import _intor


def int_eri2c(shells, shells_aux):
    bas_centers, bas_spec, bas_data = shells.as_arrays(fortran=True)
    bas_centers_aux, bas_spec_aux, bas_data_aux = shells_aux.as_arrays(fortran=True)
    return _intor.int_eri2c(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux, bas_data_aux)


def int_df1c(shells, shells_aux):
    bas_centers, bas_spec, bas_data = shells.as_arrays(fortran=True)
    bas_centers_aux, bas_spec_aux, bas_data_aux = shells_aux.as_arrays(fortran=True)
    return _intor.int_df1c(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux, bas_data_aux)


def int_df2c_mo(shells, shells_aux, mo_tensor):
    bas_centers, bas_spec, bas_data = shells.as_arrays(fortran=True)
    bas_centers_aux, bas_spec_aux, bas_data_aux = shells_aux.as_arrays(fortran=True)
    return _intor.int_df2c_mo(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux, bas_data_aux,
                              mo_tensor)
