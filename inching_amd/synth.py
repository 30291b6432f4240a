"""Seeded synthetic structures (coordinates in Angstrom).

* ``globule``    random sequential addition (dart throwing) in a sphere with a minimum
                 pair distance; the radius follows from the atom count and density.
* ``icosahedral_group``  the 60 proper rotations of the icosahedron.
* ``assembly``   subunit globules placed on an icosahedral orbit (each subunit is
                 rotated by its group element).
* ``batch_sizes`` log-uniform atom counts for batched solves.

Poisson-disk sampling at a 2.8 A minimum separation saturates near 0.029 atoms/A^3, so
the default minimum separation is 2.2 A, which reaches 0.058 atoms/A^3.
"""
import math

import numpy as np

from . import ops

DENSITY = 0.058
DMIN = 2.2


def radius_for(n_atoms, density=DENSITY):
    return (3.0 * n_atoms / (4.0 * math.pi * density)) ** (1.0 / 3.0)


def globule(n_atoms, seed, density=DENSITY, dmin=DMIN):
    return ops.synth_globule(n_atoms, radius_for(n_atoms, density), dmin, seed)


def _axis_rot(axis, ang):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * (K @ K)


def icosahedral_group():
    phi = (1 + 5 ** 0.5) / 2
    g5 = _axis_rot((0.0, 1.0, phi), 2 * math.pi / 5)
    g3 = _axis_rot((1.0, 1.0, 1.0), 2 * math.pi / 3)
    elems = [np.eye(3)]
    frontier = [np.eye(3)]

    def known(M):
        return any(np.allclose(M, E, atol=1e-9) for E in elems)

    while frontier:
        nxt = []
        for M in frontier:
            for g in (g5, g3):
                P = g @ M
                if not known(P):
                    elems.append(P)
                    nxt.append(P)
        frontier = nxt
    if len(elems) != 60:
        raise RuntimeError("icosahedral closure produced %d elements" % len(elems))
    return np.stack(elems)


def assembly(n_sub=60, atoms_per_sub=40000, shell_radius=450.0, seed0=100, orbit_seed=7,
             density=DENSITY, dmin=DMIN):
    """Subunits k = 0..n_sub-1 (globule seed seed0 + k) rotated by the k-th group element
    and centred on the orbit of a seeded generic unit vector scaled to shell_radius.
    Returns (xyz [n_sub*atoms_per_sub, 3], component id per atom)."""
    G = icosahedral_group()[:n_sub]
    rng = np.random.default_rng(orbit_seed)
    u = rng.standard_normal(3)
    u /= np.linalg.norm(u)
    parts = []
    for k in range(n_sub):
        sub = globule(atoms_per_sub, seed0 + k, density, dmin)
        parts.append(sub @ G[k].T + shell_radius * (G[k] @ u))
    xyz = np.concatenate(parts)
    comp = np.repeat(np.arange(n_sub, dtype=np.int32), atoms_per_sub)
    return xyz, comp


def batch_sizes(count=512, lo=500, hi=20000, seed=2):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=count)).astype(np.int64)
