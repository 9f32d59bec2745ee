"""Extract the reference's Coulomb-tensor golden case into tests/golden/bena2.npz.

Source (read-only reference test data, tests/test_coulomb_tensor/):
  bena2_dcat.bson                        ORCA 5.0.4 JSON wavefunction (BeNa2 2+)
  bena2_dcat_densities_relaxed.npy       (2, 2, 50, 50) state/transition densities
  test_coulomb_tensor.py:13-31           density stacking and expected J values

The basis is converted to the reference's bas_* array ABI (shell order as in
the file, one centre per atom).  Coordinates are converted from Angstrom
with CODATA 2018's 0.529177210903 Angstrom/bohr, the factor with which the
oracle reproduces the expected J values to ~2e-11 (ORCA's own 0.5291772083
gives 1.4e-7 and CODATA 2010 gives 9e-10); ``bas_data_orca`` holds the same
basis with ORCA's factor, which reproduces ORCA's overlap matrix.  ORCA's overlap matrix is stored
permuted from ORCA's within-shell order (p: z,x,y; d: z2,xz,yz,x2-y2,xy) to
m = -l..l (p: y,z,x; d: xy,yz,z2,xz,x2-y2), the order the densities use.

Run once in a container that has the reference mounted:
    python tests/golden/make_bena2.py /root/reference
"""

import sys
from pathlib import Path

import bson
import numpy as np

HERE = Path(__file__).resolve().parent
ANG_PER_BOHR_ORCA = 0.5291772083
ANG_PER_BOHR_CODATA2018 = 0.529177210903

# J values of test_coulomb_tensor.py:17-31 (nstates = 2)
J_REF = np.array(
    [
        [[[344.547007476478, 0.00000000011], [0.00000000011, 344.802879544756]],
         [[0.00000000011, 0.014478048022], [0.014478048022, 0.000000000096]]],
        [[[0.00000000011, 0.014478048022], [0.014478048022, 0.000000000096]],
         [[344.802879544756, 0.000000000096], [0.000000000096, 345.075649169291]]],
    ]
)

LMAP = {"s": 0, "p": 1, "d": 2, "f": 3, "g": 4}
# position in ORCA's shell block of each m = -l..l component
ORCA_TO_M = {0: [0], 1: [2, 0, 1], 2: [4, 2, 0, 1, 3]}


def main(ref_root):
    tdir = Path(ref_root) / "tests" / "test_coulomb_tensor"
    doc = bson.decode_all((tdir / "bena2_dcat.bson").read_bytes())[0]["Molecule"]
    assert doc["CoordinateUnits"] == "Angs"
    centers, specs, data = [], [], []
    perm = []
    bf = 0
    cptrs = []
    for atom in doc["Atoms"]:
        xyz = np.array(atom["Coords"]) / ANG_PER_BOHR_CODATA2018
        cptr = len(data) + 1
        cptrs.append(cptr)
        data.extend(xyz.tolist())
        for sh in atom["BasisFunctions"]:
            L = LMAP[sh["Shell"]]
            n = len(sh["Exponents"])
            centers.append((atom["Idx"], atom["ElementNumber"], cptr))
            specs.append((len(specs), L, n, 1, len(data) + 1))
            data.extend(sh["Coefficients"])
            data.extend(sh["Exponents"])
            perm.extend(bf + k for k in ORCA_TO_M[L])
            bf += 2 * L + 1
    data_orca = np.array(data, dtype=np.float64)
    for atom, cptr in zip(doc["Atoms"], cptrs):
        data_orca[cptr - 1 : cptr + 2] = np.array(atom["Coords"]) / ANG_PER_BOHR_ORCA
    S_orca = np.array(doc["S-Matrix"])
    perm = np.array(perm)
    S = S_orca[np.ix_(perm, perm)]
    D = np.load(tdir / "bena2_dcat_densities_relaxed.npy", allow_pickle=False)
    dens = np.stack((D[0, 0], D[0, 1], D[1, 1]), axis=0)
    np.savez_compressed(
        HERE / "bena2.npz",
        bas_centers=np.array(centers, dtype=np.int32),
        bas_spec=np.array(specs, dtype=np.int32),
        bas_data=np.array(data, dtype=np.float64),
        bas_data_orca=data_orca,
        densities=dens,
        S=S,
        J_ref=J_REF,
        ang_per_bohr=ANG_PER_BOHR_CODATA2018,
        atom_coords_ang=np.array([a["Coords"] for a in doc["Atoms"]]),
    )
    print("wrote", HERE / "bena2.npz", "nbf", bf, "nshells", len(specs))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
