"""Synthetic molecules, basis sets and densities of the five benchmark
workloads (random atoms in a cube, even-tempered contracted orbital shells,
log-uniform uncontracted auxiliary shells, random symmetric trace-1
densities).  Everything is seeded with numpy's PCG64.

Workload ids:
  0: 8 atoms, 10 bohr cube, orbital 3s2p1d, aux 4s3p2d1f, 3 states
  1: 60 atoms, 22 bohr cube, orbital 3s2p1d, aux 4s3p2d1f, 8 states
  2: as 1 with orbital +1f and aux +1g
  3: 256 molecules of 5-40 atoms, cube side 4 N^(1/3), as 0, 4 states each
  4: 200 atoms, 34 bohr cube, orbital 4s3p2d1f1g, aux 5s4p3d2f1g1h, 20 states
"""

import numpy as np

from .shells import Shell, Shells

SEED_BASE = 20260000
MOL_SEED_BASE = 20263000

_ORB = {
    "spd": (3, 2, 1),
    "spdf": (3, 2, 1, 1),
    "spdfg": (4, 3, 2, 1, 1),
}
_AUX = {
    "spdf": (4, 3, 2, 1),
    "spdfg": (4, 3, 2, 1, 1),
    "spdfgh": (5, 4, 3, 2, 1, 1),
}

WORKLOADS = {
    0: dict(natoms=8, side=10.0, orb="spd", aux="spdf", nstates=3, alpha0="grid"),
    1: dict(natoms=60, side=22.0, orb="spd", aux="spdf", nstates=8, alpha0="grid"),
    2: dict(natoms=60, side=22.0, orb="spdf", aux="spdfg", nstates=8, alpha0="grid"),
    4: dict(natoms=200, side=34.0, orb="spdfg", aux="spdfgh", nstates=20, alpha0="loguniform"),
}
_AUX_RANGE = {"spdf": (0.05, 200.0), "spdfg": (0.05, 200.0), "spdfgh": (0.03, 500.0)}


def random_geometry(rng, natoms, side, min_dist=2.0, max_tries=100000):
    """Uniform points in [0, side]^3 with pairwise distance >= min_dist."""
    coords = []
    tries = 0
    while len(coords) < natoms:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place atoms")
        x = rng.uniform(0.0, side, size=3)
        if all(np.linalg.norm(x - c) >= min_dist for c in coords):
            coords.append(x)
    return np.array(coords)


def _orbital_shells(rng, coords, orb, alpha0):
    shells = []
    for ci, xyz in enumerate(coords):
        for L, count in enumerate(_ORB[orb]):
            for _ in range(count):
                if alpha0 == "grid":
                    a0 = 0.1 * 3.0 ** rng.integers(0, 6)
                else:
                    a0 = np.exp(rng.uniform(np.log(0.05), np.log(2000.0 / 9.0)))
                exps = a0 * 3.0 ** np.arange(3)
                coeffs = rng.uniform(0.2, 1.0, size=3)
                shells.append(Shell(L=L, center=xyz, exps=exps, coeffs=coeffs, center_ind=ci))
    return Shells(shells)


def _aux_shells(rng, coords, aux):
    lo, hi = _AUX_RANGE[aux]
    shells = []
    for ci, xyz in enumerate(coords):
        for L, count in enumerate(_AUX[aux]):
            for _ in range(count):
                e = np.exp(rng.uniform(np.log(lo), np.log(hi)))
                shells.append(Shell(L=L, center=xyz, exps=[e], coeffs=[1.0], center_ind=ci))
    return Shells(shells)


def random_densities(rng, nstates, nbf):
    """ndens = nstates(nstates+1)/2 symmetric matrices (A + A^T)/2, A ~ U(0,1),
    each scaled to unit trace; C-ordered (ndens, nbf, nbf)."""
    ndens = nstates * (nstates + 1) // 2
    D = np.empty((ndens, nbf, nbf))
    for i in range(ndens):
        A = rng.uniform(0.0, 1.0, size=(nbf, nbf))
        A = 0.5 * (A + A.T)
        D[i] = A / np.trace(A)
    return D


def molecule(natoms, side, orb="spd", aux="spdf", nstates=3, alpha0="grid", seed=0, with_densities=True):
    rng = np.random.Generator(np.random.PCG64(seed))
    coords = random_geometry(rng, natoms, side)
    shells = _orbital_shells(rng, coords, orb, alpha0)
    shells_aux = _aux_shells(rng, coords, aux)
    dens = random_densities(rng, nstates, shells.nbf) if with_densities else None
    return shells, shells_aux, dens


def workload(wid, with_densities=True):
    """(shells, shells_aux, densities) of workload 0, 1, 2 or 4; densities None
    with with_densities=False (workload 4 holds 102 GB of them)."""
    w = WORKLOADS[wid]
    return molecule(w["natoms"], w["side"], w["orb"], w["aux"], w["nstates"], w["alpha0"],
                    seed=SEED_BASE + wid, with_densities=with_densities)


def workload4_reduced(natoms=10):
    """The 200-atom recipe (orbital 4s3p2d1f1g, aux 5s4p3d2f1g1h, 20 states)
    on natoms atoms at the same atom density (cube side 34 (natoms/200)^(1/3)
    bohr): small enough for the oracle, with every class of the large run."""
    w = WORKLOADS[4]
    side = w["side"] * (natoms / w["natoms"]) ** (1.0 / 3.0)
    return molecule(natoms, side, w["orb"], w["aux"], w["nstates"], w["alpha0"],
                    seed=SEED_BASE + 4000 + natoms)


def batch_molecule(i):
    """Molecule i (0..255) of the batched workload 3."""
    rng = np.random.Generator(np.random.PCG64(MOL_SEED_BASE + i))
    natoms = int(rng.integers(5, 41))
    side = 4.0 * natoms ** (1.0 / 3.0)
    return molecule(natoms, side, "spd", "spdf", 4, "grid", seed=MOL_SEED_BASE + 100000 + i)
