"""Drop-in for pysisyphus_addons.intor.benchmark (intor/benchmark.py:4-17),
plus the estimate records in memory and the metric power of the eigen route."""

from . import _benchmark


def benchmark_int3c2e(shells, shells_aux):
    bas_centers, bas_spec, bas_data = shells.as_arrays(fortran=True)
    bas_centers_aux, bas_spec_aux, bas_data_aux = shells_aux.as_arrays(fortran=True)
    return _benchmark.benchmark_int3c2e(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux,
                                        bas_data_aux)


def int3c2e_estimates(shells, shells_aux):
    bas_centers, bas_spec, bas_data = shells.as_arrays(fortran=True)
    bas_centers_aux, bas_spec_aux, bas_data_aux = shells_aux.as_arrays(fortran=True)
    return _benchmark.int3c2e_estimates(bas_centers, bas_spec, bas_data, bas_centers_aux, bas_spec_aux,
                                        bas_data_aux)


def metric_powerh(shells_aux, power=-0.5):
    bas_centers, bas_spec, bas_data = shells_aux.as_arrays(fortran=True)
    return _benchmark.metric_powerh(bas_centers, bas_spec, bas_data, power)
