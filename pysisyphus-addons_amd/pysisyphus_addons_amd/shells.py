"""Minimal shell container producing the reference's basis-array ABI.

pysisyphus' ``Shells.as_arrays(fortran=True)`` is the only thing the
reference's Python entry points need from a basis
(``diabatization/intor.py:4-36``).  This module builds the same three arrays
from plain shell descriptions, so that callers without pysisyphus (tests,
benchmarks) can drive the drop-in:

* ``bas_centers`` int32 (nshells, 3): center index, atomic number, 1-based
  pointer to the centre's x, y, z (bohr) in ``bas_data``;
* ``bas_spec`` int32 (nshells, 5): shell index, L, number of primitives,
  number of contractions (1), 1-based pointer to the contraction
  coefficients in ``bas_data`` (the exponents follow them);
* ``bas_data`` float64.

The column meaning follows ``common/shells.f90:100-110`` and the read at
``common/shells.f90:149-167``.  Contraction coefficients refer to normalized
primitives; the library normalizes the contracted shell itself
(``common/normalization.f90:10-38``).
"""

from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np


@dataclass
class Shell:
    L: int
    center: Sequence[float]
    exps: Sequence[float]
    coeffs: Sequence[float]
    center_ind: int = 0
    atomic_num: int = 1


@dataclass
class Shells:
    shells: List[Shell] = field(default_factory=list)

    def __len__(self):
        return len(self.shells)

    @property
    def nbf(self):
        """Number of spherical basis functions."""
        return sum(2 * s.L + 1 for s in self.shells)

    @property
    def lmax(self):
        return max(s.L for s in self.shells)

    def as_arrays(self, fortran=True):
        """Return (bas_centers, bas_spec, bas_data); C-contiguous, pointers
        1-based when fortran is True (0-based otherwise)."""
        off = 1 if fortran else 0
        n = len(self.shells)
        bas_centers = np.zeros((n, 3), dtype=np.int32)
        bas_spec = np.zeros((n, 5), dtype=np.int32)
        data = []
        center_ptr = {}
        for i, sh in enumerate(self.shells):
            key = (sh.center_ind, tuple(float(x) for x in sh.center))
            if key not in center_ptr:
                center_ptr[key] = len(data)
                data.extend(float(x) for x in sh.center)
            nprim = len(sh.exps)
            assert len(sh.coeffs) == nprim
            bas_centers[i] = (sh.center_ind, sh.atomic_num, center_ptr[key] + off)
            bas_spec[i] = (i, sh.L, nprim, 1, len(data) + off)
            data.extend(float(c) for c in sh.coeffs)
            data.extend(float(e) for e in sh.exps)
        return bas_centers, bas_spec, np.array(data, dtype=np.float64)

    @staticmethod
    def from_arrays(bas_centers, bas_spec, bas_data, fortran=True):
        off = 1 if fortran else 0
        shells = []
        for (cind, Z, cptr), (_, L, nprim, _, pptr) in zip(bas_centers, bas_spec):
            c0 = cptr - off
            p0 = pptr - off
            shells.append(
                Shell(
                    L=int(L),
                    center=bas_data[c0 : c0 + 3].copy(),
                    coeffs=bas_data[p0 : p0 + nprim].copy(),
                    exps=bas_data[p0 + nprim : p0 + 2 * nprim].copy(),
                    center_ind=int(cind),
                    atomic_num=int(Z),
                )
            )
        return Shells(shells)
