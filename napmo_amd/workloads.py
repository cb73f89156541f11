"""Seeded synthetic workloads of the grid hot path (geometry, basis, density matrix),
shared by the benchmark and the tests. All coordinates in bohr; seeds use
numpy.random.default_rng (PCG64).

* h2_system: H2 at (0, 0, +-0.7), 2 x 100 x 110 Chebyshev/Lebedev, cc-pVDZ (n = 10), lmax 10
* chon16_system: 16 atoms {C,H,N,O} in a 12 bohr box (min separation 1.8), 50 x 590 per atom
  (472,000 points), 300 Cartesian s/p/d functions (30 cycles of s, p, d shells, cycle c on
  atom c % 16, K ~ U{1,2,3} primitives, zeta log-uniform [0.1, 1e4], c ~ U(0.2, 1)),
  P = 2 C C^T from 40 orthonormal occupied orbitals, lmax 12
"""
import json
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data')
with open(os.path.join(_DATA, 'basis_sets.json')) as _f:
    BASIS = json.load(_f)

H2_ATOMS = [('H', (0.0, 0.0, 0.7)), ('H', (0.0, 0.0, -0.7))]


def random_molecule(rng, natom, box, elements=('C', 'H', 'N', 'O'), min_sep=1.8):
    """Elements uniform over ``elements``; positions uniform in [0, box]^3, redrawn until
    every pair is at least ``min_sep`` apart."""
    syms = [elements[k] for k in rng.integers(0, len(elements), natom)]
    while True:
        pos = rng.uniform(0.0, box, (natom, 3))
        d = np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1))
        d[np.diag_indices(natom)] = np.inf
        if d.min() >= min_sep:
            return list(zip(syms, [tuple(float(c) for c in p) for p in pos]))


def random_shells(rng, natom, ncycle, lmax_shell=2):
    """``ncycle`` cycles of shells l = 0..lmax_shell; cycle c goes to atom c % natom."""
    per_atom = [[] for _ in range(natom)]
    for c in range(ncycle):
        for l in range(lmax_shell + 1):
            k = int(rng.integers(1, 4))
            zeta = np.exp(rng.uniform(np.log(0.1), np.log(1e4), k))
            coef = rng.uniform(0.2, 1.0, k)
            per_atom[c % natom].append({'l': l, 'prim': [float(z) for z in zeta],
                                        'cont': [float(x) for x in coef]})
    return per_atom


def named_shells(atoms, basis_name):
    return [BASIS[basis_name][s] for s, _ in atoms]


def h2_system():
    atoms = H2_ATOMS
    return atoms, named_shells(atoms, 'CC-PVDZ'), 100, 110, 10


def chon16_system(seed=2, nrad=50, nang=590, natom=16, ncycle=30):
    rng = np.random.default_rng(seed)
    atoms = random_molecule(rng, natom, 12.0)
    shells = random_shells(rng, natom, ncycle)
    return atoms, shells, nrad, nang, 12


def even_tempered_shells(nexp=8, zeta0=4.0, ratio=10.0 ** (1.0 / 7.0), lmax_shell=2):
    """Nuclear basis of one quantum proton: zeta_k = zeta0 ratio^k (k < nexp), one
    uncontracted shell per (k, l <= lmax_shell); 8 x (1 + 3 + 6) = 80 functions."""
    return [{'l': l, 'prim': [float(zeta0 * ratio ** k)], 'cont': [1.0]}
            for k in range(nexp) for l in range(lmax_shell + 1)]


def apmo_system(seed=3, nheavy=4, nh=4, box=12.0, ncycle=12):
    """Electrons + quantum protons: ``nheavy`` atoms from {C, N, O} and ``nh`` H atoms in a
    ``box`` bohr cube (min separation 1.8). Electrons: one 100 x 302 grid per atom
    (241,600 points), ``ncycle`` cycles of s/p/d random shells over all atoms (120
    functions). Protons (species H_1, charge +1, eta 1): one 60 x 194 grid per H (46,560
    points) and the even-tempered 80-function basis on each (320). lmax = half the
    Lebedev degree (14 and 11). Returns a dict of both species."""
    rng = np.random.default_rng(seed)
    heavy = [('C', 'N', 'O')[k] for k in rng.integers(0, 3, nheavy)]
    syms = heavy + ['H'] * nh
    while True:
        pos = rng.uniform(0.0, box, (len(syms), 3))
        d = np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1))
        d[np.diag_indices(len(syms))] = np.inf
        if d.min() >= 1.8:
            break
    atoms = list(zip(syms, [tuple(float(c) for c in p) for p in pos]))
    e_shells = random_shells(rng, len(atoms), ncycle)
    protons = [a for a in atoms if a[0] == 'H']
    n_shells = [even_tempered_shells() for _ in protons]
    nelec = sum({'H': 1, 'C': 6, 'N': 7, 'O': 8}[s] for s in syms)
    return {'atoms': atoms, 'electron': {'centres': atoms, 'shells': e_shells, 'nrad': 100,
                                         'nang': 302, 'lmax': 14, 'charge': -1.0, 'eta': 2.0,
                                         'nocc': nelec // 2},
            'proton': {'centres': protons, 'shells': n_shells, 'nrad': 60, 'nang': 194, 'lmax': 11,
                       'charge': 1.0, 'eta': 1.0, 'nocc': len(protons)}}


LEBEDEV_CHOICES = (110, 146, 170, 194, 230, 266, 302, 350, 434, 590)
ZNUC = {'H': 1, 'C': 6, 'N': 7, 'O': 8}


def batch_molecules(seed=5, count=64):
    """``count`` small molecules: n_atoms ~ U{2..12} of {C,H,N,O} in a box of side
    2.5 n^(1/3) + 2 bohr (min separation 1.8); per atom nr ~ U{35..110} Chebyshev points
    and na uniform over LEBEDEV_CHOICES; U{2..20} cycles of s/p/d shells (n = 20..200);
    nocc = ceil(n_elec / 2) clipped to n; lmax = min(12, Lebedev degree of the smallest
    angular grid / 2)."""
    from napmo_amd.grids.lebedev import lebedev_get_order
    rng = np.random.default_rng(seed)
    mols = []
    for _ in range(count):
        natom = int(rng.integers(2, 13))
        atoms = random_molecule(rng, natom, 2.5 * natom ** (1.0 / 3.0) + 2.0)
        sizes = [(int(rng.integers(35, 111)), int(rng.choice(LEBEDEV_CHOICES))) for _ in range(natom)]
        ncycle = int(rng.integers(2, 21))
        shells = random_shells(rng, natom, ncycle)
        n = 10 * ncycle
        nelec = sum(ZNUC[s] for s, _ in atoms)
        lmax = min(12, lebedev_get_order(min(na for _, na in sizes)) // 2)
        mols.append({'atoms': atoms, 'sizes': sizes, 'shells': shells, 'nbasis': n,
                     'nocc': min((nelec + 1) // 2, n), 'lmax': lmax})
    return mols


def occupied_orbitals(rng, S, nocc):
    """C = C0 (C0^T S C0)^(-1/2), C0 ~ N(0, 1): nocc S-orthonormal orbitals [n x nocc]."""
    C0 = rng.normal(size=(S.shape[0], nocc))
    w, V = np.linalg.eigh(C0.T @ S @ C0)
    return C0 @ (V @ np.diag(w ** -0.5) @ V.T)


def density_matrix(rng, S, nocc):
    """P = 2 C C^T from ``occupied_orbitals`` (same random stream)."""
    C = occupied_orbitals(rng, S, nocc)
    return 2.0 * C @ C.T
