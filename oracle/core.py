"""CPU float64 oracle for the nAPMO grid hot path. TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker. The product
(``napmo_amd``) never imports it and has no CPU fallback.

Every function restates one step of the reference algorithm and cites the
reference file:line it follows. Heavy O(N*M^2) / O(N*M*L) / O(N*n^2) loops also
exist in C (``oracle/napmo_oracle.c``, through ``oracle/cext.py``); the numpy
versions here are the readable statement used at small sizes and to check the C
ones.

Pinning (what fixes each function besides the reference text):

* lebedev               -- oracle/_ref (reference generator), exactness of the rules
* transforms, radial    -- oracle/_ref, He 50-point radii golden vector, Gaussian
                           radial integral closed form
* becke_weights         -- partition of unity, nuclear limit, permutation symmetry,
                           brute-force per-point Python loop (reference becke.cpp
                           needs Eigen: no _ref vector)
* spherical_expansion   -- oracle/_ref dot_multi_moments, 1s/1pz closed forms
* splines / build_ode2  -- oracle/_ref build_ode2, solve_cubic_spline_system,
                           CubicSpline_eval; closed-form Poisson of a Gaussian
* eval_decomposition    -- oracle/_ref eval_decomposition_grid; re-synthesis of rho
* basis_compute         -- oracle/_ref BasisSet_compute, reference test constants
* contractions, energy  -- textbook identities (einsum), trace identities
* compute_dpsi          -- exact solution of its radial equations on one centre
                           (fourth-order convergence), planted-mistake checks
* common_index          -- constructed grids with known shared points, swap and order
                           rules, signed zero
* epc17_2               -- the published closed form at hand-computed points; both
                           potential forms equal the derivatives of rhoE c_ene
* read_lowdin2_grid     -- a file written with known content
"""
from fractions import Fraction
import json
import math
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), '..', 'napmo_amd', 'data')

PI_4 = 4.0 * math.pi
INV_SQRT_4_PI = 0.28209479177387814347  # evaluate.cpp:127


# ----------------------------------------------------------------------------------
# Tabulated data
# ----------------------------------------------------------------------------------

_LEB = None
_ELEM = None


def _lebedev_table():
    global _LEB
    if _LEB is None:
        with open(os.path.join(_DATA, 'lebedev_orbits.json')) as f:
            _LEB = json.load(f)
    return _LEB


def elements():
    global _ELEM
    if _ELEM is None:
        with open(os.path.join(_DATA, 'elements.json')) as f:
            _ELEM = json.load(f)
    return _ELEM


LEBEDEV_ORDER = {6: 3, 14: 5, 26: 7, 38: 9, 50: 11, 74: 13, 86: 15, 110: 17, 146: 19,
                 170: 21, 194: 23, 230: 25, 266: 27, 302: 29, 350: 31, 434: 35, 590: 41,
                 770: 47, 974: 53, 1202: 59, 1454: 65, 1730: 71, 2030: 77, 2354: 83,
                 2702: 89, 3074: 95, 3470: 101, 3890: 107, 4334: 113, 4802: 119,
                 5294: 125, 5810: 131}


def lebedev_get_order(n):
    """napmo/grids/lebedev.py:10-27 (KeyError for an unknown point count)."""
    return LEBEDEV_ORDER[n]


def lebedev(npoint):
    """Unit vectors [n,3] and weights [n] (sum 1) of the ``npoint`` Lebedev rule.

    Follows src/grid/sphere_lebedev_rule.cpp:61 (gen_oh) and :639 (ld_by_order) via
    the orbit table: symbol magnitudes are formed in the same arithmetic order
    (gen_oh: b = sqrt(1 - 2a^2) for code 4, sqrt(1 - a^2) for code 5,
    c = sqrt(1 - a^2 - b^2) for code 6).
    """
    tab = _lebedev_table()
    orbits = tab['rules'][str(npoint)]
    pts, wts = [], []
    for code, a, b, v in orbits:
        if code == 1:
            sym = (1.0, 0.0, 0.0)
        elif code == 2:
            sym = (1.0 / math.sqrt(2.0), 0.0, 0.0)
        elif code == 3:
            sym = (1.0 / math.sqrt(3.0), 0.0, 0.0)
        elif code == 4:
            sym = (a, math.sqrt(1.0 - 2.0 * a * a), 0.0)
        elif code == 5:
            sym = (a, math.sqrt(1.0 - a * a), 0.0)
        else:
            sym = (a, b, math.sqrt(1.0 - a * a - b * b))
        for row in tab['templates'][str(code)]:
            pts.append([0.0 if s == 0 else math.copysign(sym[abs(s) - 1], s) for s in row])
            wts.append(v)
    return np.array(pts, dtype=np.float64), np.array(wts, dtype=np.float64)


def atomic_radius(symbol):
    """Radius choice of napmo/grids/radial.py:20-48 (H 'atomic_radii_2' fallback,
    isotope 'X_k' takes X's 'atomic_radii', known symbols their own)."""
    db = elements()
    radii = db.get('H', {}).get('atomic_radii_2', 1.0)
    k = symbol.find('_')
    if k > 0:
        radii = db.get(symbol[:k], {}).get('atomic_radii', 1.0)
    if symbol in db:
        radii = db[symbol].get('atomic_radii', 1.0)
    return radii


# ----------------------------------------------------------------------------------
# Radial transforms and grids
# ----------------------------------------------------------------------------------

class ChebyshevRTransform:
    """src/horton/rtransform.cpp:308-341: t in [0, n-1], t' = n - t, z = pi/(n+1),
    r = R (1 + cos(t' z)) / (1 - cos(t' z))."""
    kind = 'chebyshev'

    def __init__(self, radii, npoint):
        if npoint < 2:
            raise ValueError('A radial grid consists of at least two points.')
        self.radii = float(radii)
        self.npoint = int(npoint)
        self.z = math.pi / (npoint + 1)

    def radius(self, t):
        t = self.npoint - np.asarray(t, dtype=np.float64)
        aux = np.cos(t * self.z)
        return self.radii * (1 + aux) / (1 - aux)

    def deriv(self, t):
        t = self.npoint - np.asarray(t, dtype=np.float64)
        aux = t * self.z
        return (2 * self.radii * self.z * np.sin(aux)) / (np.cos(aux) - 1) ** 2

    def deriv2(self, t):
        t = self.npoint - np.asarray(t, dtype=np.float64)
        aux = t * self.z
        return 0.5 * self.z * self.z * self.radii * (2 + np.cos(aux)) * (1 / np.sin(aux * 0.5)) ** 4

    def deriv3(self, t):
        t = self.npoint - np.asarray(t, dtype=np.float64)
        aux = t * self.z
        aux2 = aux * 0.5
        return (0.25 * self.z * self.z * self.z * self.radii * (11 * np.cos(aux2) + np.cos(1.5 * aux))
                * (1 / np.sin(aux2)) ** 5)

    def inv(self, r):
        r = np.asarray(r, dtype=np.float64)
        return -(np.arccos((r - self.radii) / (r + self.radii)) / self.z) + self.npoint


class PowerRTransform:
    """src/horton/rtransform.cpp:257-283: r = rmin (t+1)^p, p = ln(rmax/rmin)/ln n."""
    kind = 'power'

    def __init__(self, rmin, rmax, npoint):
        if npoint < 2:
            raise ValueError('A radial grid consists of at least two points.')
        if rmin >= rmax:
            raise ValueError('rmin must be below rmax.')
        self.rmin, self.rmax, self.npoint = float(rmin), float(rmax), int(npoint)
        self.power = (math.log(rmax) - math.log(rmin)) / math.log(npoint)
        if self.power < 2.0:
            raise ValueError('Power must be at least two for a decent intgration.')
        self.radii = self.rmax  # napmo/grids/radial_transform.py:199-200

    def radius(self, t):
        return self.rmin * np.power(np.asarray(t, dtype=np.float64) + 1, self.power)

    def deriv(self, t):
        return self.power * self.rmin * np.power(np.asarray(t, dtype=np.float64) + 1, self.power - 1)

    def deriv2(self, t):
        return self.power * (self.power - 1) * self.rmin * np.power(np.asarray(t, dtype=np.float64) + 1,
                                                                   self.power - 2)

    def deriv3(self, t):
        p = self.power
        return p * (p - 1) * (p - 2) * self.rmin * np.power(np.asarray(t, dtype=np.float64) + 1, p - 3)

    def inv(self, r):
        return np.power(np.asarray(r, dtype=np.float64) / self.rmin, 1.0 / self.power) - 1


class IdentityRTransform:
    """src/horton/rtransform.cpp (IdentityRTransform): r = t."""
    kind = 'identity'

    def __init__(self, npoint):
        self.npoint = int(npoint)
        self.radii = 0.0

    def radius(self, t):
        return np.asarray(t, dtype=np.float64) * 1.0

    def deriv(self, t):
        return np.ones_like(np.asarray(t, dtype=np.float64))

    def deriv2(self, t):
        return np.zeros_like(np.asarray(t, dtype=np.float64))

    def deriv3(self, t):
        return np.zeros_like(np.asarray(t, dtype=np.float64))

    def inv(self, r):
        return np.asarray(r, dtype=np.float64) * 1.0


def radial_grid(rtf):
    """src/grid/radial.cpp:14-29: r_i = r(i), rw_i = ((4 pi dr/dt) r_i) r_i."""
    t = np.arange(rtf.npoint, dtype=np.float64)
    r = rtf.radius(t)
    rw = PI_4 * rtf.deriv(t) * r * r
    return r, rw


class AtomicGrid:
    """src/grid/atomic_grid.cpp:12-51: radial-major points p = u_j r_i + O,
    weights aw_j rw_i, index i*nang + j."""

    def __init__(self, nrad, nang, origin, symbol='H', rtransform=None):
        self.symbol = symbol
        self.origin = np.asarray(origin, dtype=np.float64).copy()
        self.rtf = rtransform if rtransform is not None else ChebyshevRTransform(atomic_radius(symbol), nrad)
        self.radii = self.rtf.radii
        self.nrad = self.rtf.npoint
        self.nang = nang
        self.r, self.rw = radial_grid(self.rtf)
        self.u, self.aw = lebedev(nang)
        pts = self.u[None, :, :] * self.r[:, None, None] + self.origin[None, None, :]
        self.points = pts.reshape(-1, 3)
        self.weights = (self.aw[None, :] * self.rw[:, None]).reshape(-1)
        self.size = self.points.shape[0]


class BeckeGrid:
    """napmo/grids/becke.py:33-69 + src/grid/becke.cpp:12-54. ``centers`` is a list of
    (symbol, origin). Atomic grids may differ in size (a generalisation of the
    reference, which assumes equal sizes, becke.cpp:121: both agree when equal)."""

    def __init__(self, centers, nrad, nang, rtransform=None, sizes=None):
        self.atgrids = []
        for k, (sym, origin) in enumerate(centers):
            nr, na = (sizes[k] if sizes is not None else (nrad, nang))
            self.atgrids.append(AtomicGrid(nr, na, origin, sym, rtransform))
        self.ncenter = len(self.atgrids)
        self.origin = np.array([g.origin for g in self.atgrids])
        self.radii = np.array([g.radii for g in self.atgrids])
        self.points = np.concatenate([g.points for g in self.atgrids])
        self.weights = np.concatenate([g.weights for g in self.atgrids])
        self.offsets = np.cumsum([0] + [g.size for g in self.atgrids])
        self.size = int(self.offsets[-1])
        self.nrad, self.nang = nrad, nang
        self.becke_weights = becke_weights(self.points, self.origin, self.radii, self.offsets)

    def integrate(self, f):
        return grid_integrate(f, self.becke_weights, self.weights)


# ----------------------------------------------------------------------------------
# Becke partition (src/grid/becke.cpp:79-176)
# ----------------------------------------------------------------------------------

def becke_pair_tables(origin, radii):
    """R_ij and clamped a_ij (becke.cpp:83-115). Returns full matrices with the
    antisymmetric sign applied as the evaluation does (sign -1 when i < j)."""
    m = len(radii)
    R = np.zeros((m, m))
    A = np.zeros((m, m))
    for i in range(m):
        for j in range(i + 1):
            d = origin[i] - origin[j]
            R[i, j] = R[j, i] = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
            u = (radii[i] - radii[j]) / (radii[i] + radii[j])  # eq. A6
            u = u / (u * u - 1.0)                               # eq. A5
            u = min(max(u, -0.45), 0.45)                          # eq. A3
            A[i, j] = u
            A[j, i] = u
    return R, A


def becke_weights(points, origin, radii, offsets):
    """bw_p = P_owner(p) / sum_i P_i(p), P_i = prod_{j != i} s(mu_ij + a_ij (1 - mu_ij^2))
    with three smoothing passes s <- s(3 - s^2)/2 then s <- (1 - s)/2
    (becke.cpp:119-172)."""
    m = len(radii)
    R, A = becke_pair_tables(origin, radii)
    dist = np.empty((m, points.shape[0]))
    for i in range(m):
        d = points - origin[i]
        dist[i] = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    owner = np.repeat(np.arange(m), np.diff(offsets))
    total = np.zeros(points.shape[0])
    mine = np.zeros(points.shape[0])
    for i in range(m):
        p = np.ones(points.shape[0])
        for j in range(m):
            if i == j:
                continue
            mu = (dist[i] - dist[j]) / R[i, j]
            sign = -1.0 if i < j else 1.0
            s = mu + A[i, j] * (sign * (1.0 - mu * mu))
            for _ in range(3):
                s = 0.5 * s * (3 - s * s)
            s = 0.5 * (1.0 - s)
            p = p * s
        mine = np.where(owner == i, p, mine)
        total = total + p
    return mine / total


def becke_weights_pointwise(point, owner, origin, radii):
    """Brute-force single-point Becke weight written straight from Becke (1988)
    eqs. 11, 13, 19-22, A2-A6, used only to pin ``becke_weights``."""
    m = len(radii)
    cell = []
    for i in range(m):
        prod = 1.0
        ri = math.dist(point, origin[i])
        for j in range(m):
            if j == i:
                continue
            rj = math.dist(point, origin[j])
            mu = (ri - rj) / math.dist(origin[i], origin[j])
            chi = radii[i] / radii[j]
            u = (chi - 1) / (chi + 1)
            a = min(max(u / (u * u - 1), -0.45), 0.45)
            nu = mu + a * (1 - mu * mu)
            f = nu
            for _ in range(3):
                f = 1.5 * f - 0.5 * f ** 3
            prod *= 0.5 * (1 - f)
        cell.append(prod)
    return cell[owner] / sum(cell)


def grid_integrate(f, bw, w):
    """src/grid/becke.cpp:178-184: sum_p f_p bw_p w_p."""
    return float(np.sum(f * bw * w))


# ----------------------------------------------------------------------------------
# Real solid harmonics (src/horton/moments.cpp:59-125) and spherical expansion
# ----------------------------------------------------------------------------------

def pure_polynomials(z, x, y, lmax):
    """Racah-normalised regular solid harmonics for l = 1..lmax in the per-l order
    [C_l0, C_l1, S_l1, ..., C_ll, S_ll], from inputs (z, x, y)
    (moments.cpp:59-125). Returns [(lmax+1)^2 - 1, npts]."""
    z = np.asarray(z, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    out = [z, x, y]
    if lmax <= 1:
        return np.array(out[:max(0, (lmax + 1) ** 2 - 1)])
    r2 = x * x + y * y + z * z
    pi_old = [np.ones_like(z)] + [None] * lmax
    pi_new = [z, np.ones_like(z)] + [None] * (lmax - 1)
    a = [None, x] + [None] * (lmax - 1)
    b = [None, y] + [None] * (lmax - 1)
    for l in range(2, lmax + 1):
        factor = 2 * l - 1
        for m in range(0, l - 1):
            tmp = pi_old[m]
            pi_old[m] = pi_new[m]
            pi_new[m] = (z * factor * pi_old[m] - r2 * (l + m - 1) * tmp) / (l - m)
        pi_old[l - 1] = pi_new[l - 1]
        pi_new[l] = factor * pi_old[l - 1]
        pi_new[l - 1] = z * pi_new[l]
        a[l] = x * a[l - 1] - y * b[l - 1]
        b[l] = x * b[l - 1] + y * a[l - 1]
        out.append(pi_new[0])
        fac = math.sqrt(2.0)
        for m in range(1, l + 1):
            fac /= math.sqrt((l + m) * (l - m + 1))
            out.append(fac * a[m] * pi_new[m])
            out.append(fac * b[m] * pi_new[m])
    return np.array(out)


def spherical_expansion(atgrid, f, lmax):
    """rho_lm(r_i) of src/grid/atomic_grid.cpp:82-136 with dot_multi_moments
    mtype=4 (src/horton/dot_multi.cpp:148-186): per radial shell,
    sum_j f w S_lm(u) / sum_j w, times sqrt(4 pi (2l+1)). Returns [nrad, L]."""
    if lmax < 0:
        raise ValueError('lmax can not be negative.')
    L = (lmax + 1) ** 2
    ff = f * atgrid.weights
    d = atgrid.points - atgrid.origin
    r = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    moments = np.ones((L, atgrid.size))
    if lmax > 0:
        moments[1:] = pure_polynomials(d[:, 2] / r, d[:, 0] / r, d[:, 1] / r, lmax)
    nr, na = atgrid.nrad, atgrid.nang
    out = np.einsum('lij,ij->il', moments.reshape(L, nr, na), ff.reshape(nr, na))
    sigma = atgrid.weights.reshape(nr, na).sum(axis=1)
    out /= sigma[:, None]
    lvec = np.repeat(np.arange(lmax + 1), 2 * np.arange(lmax + 1) + 1)
    out *= np.sqrt(4.0 * math.pi * (2.0 * lvec + 1))[None, :]
    return out


# ----------------------------------------------------------------------------------
# Cubic splines on the t-axis (src/horton/cubic_spline.cpp)
# ----------------------------------------------------------------------------------

def tridiagsym_solve(diag_mid, diag_up, right):
    """src/horton/cubic_spline.cpp:54-79 (same elimination order)."""
    n = len(diag_mid)
    diag_up = np.array(diag_up, dtype=np.float64)
    right = np.array(right, dtype=np.float64)
    tmp1 = diag_up[0]
    tmp2 = 0.0
    diag_up[0] /= diag_mid[0]
    right[0] /= diag_mid[0]
    for i in range(1, n):
        den = diag_mid[i] - diag_up[i - 1] * tmp1
        if i < n - 1:
            tmp2 = diag_up[i]
            diag_up[i] /= den
        right[i] = (right[i] - right[i - 1] * tmp1) / den
        tmp1 = tmp2
    sol = np.empty(n)
    sol[n - 1] = right[n - 1]
    for i in range(n - 2, -1, -1):
        sol[i] = right[i] - diag_up[i] * sol[i + 1]
    return sol


def spline_dt(y):
    """solve_cubic_spline_system, src/horton/cubic_spline.cpp:81-103: the (2,4,...,4,2|1)
    system for the t-derivatives of the C2 interpolating spline."""
    y = np.asarray(y, dtype=np.float64)
    n = len(y)
    mid = np.full(n, 4.0)
    mid[0] = mid[-1] = 2.0
    up = np.ones(n - 1)
    right = np.empty(n)
    right[0] = 3.0 * (y[1] - y[0])
    right[1:-1] = 3.0 * (y[2:] - y[:-2])
    right[-1] = 3.0 * (y[-1] - y[-2])
    return tridiagsym_solve(mid, up, right)


def spline_dt_matrix(n):
    """Dense T with dt = T y, the linear map of ``spline_dt``."""
    T = np.empty((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        T[:, j] = spline_dt(e)
    return T


class PotentialExtrapolation:
    """src/horton/extrapolation.cpp:149-183: left amp r^l, right amp / r^(l+1)."""

    def __init__(self, l):
        if l < 0:
            raise ValueError('The argument l cannot be negative.')
        self.l = int(l)

    def prepare(self, y, first_x, last_x):
        self.amp_left = y[0] / first_x ** self.l
        self.amp_right = y[-1] * last_x ** (self.l + 1)

    def eval_left(self, x):
        return self.amp_left * x ** self.l

    def eval_right(self, x):
        return self.amp_right / x ** (self.l + 1)


class ZeroExtrapolation:
    def prepare(self, y, first_x, last_x):
        pass

    def eval_left(self, x):
        return 0.0 * x

    def eval_right(self, x):
        return 0.0 * x


class Spline:
    """CubicSpline of src/horton/cubic_spline.cpp:138-181 (y and dt on the t-grid)."""

    def __init__(self, y, dt, rtf, extrapolation=None):
        self.y = np.asarray(y, dtype=np.float64)
        self.dt = np.asarray(dt, dtype=np.float64)
        self.rtf = rtf
        self.n = len(self.y)
        self.first_x = float(rtf.radius(0))
        self.last_x = float(rtf.radius(self.n - 1))
        self.extrapolation = extrapolation if extrapolation is not None else ZeroExtrapolation()
        self.extrapolation.prepare(self.y, self.first_x, self.last_x)

    def __call__(self, x):
        x = np.asarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        left = x < self.first_x
        right = x > self.last_x
        mid = ~(left | right)
        if left.any():
            out[left] = self.extrapolation.eval_left(x[left])
        if right.any():
            out[right] = self.extrapolation.eval_right(x[right])
        if mid.any():
            t = self.rtf.inv(x[mid])
            j = np.floor(t).astype(np.int64)
            j = np.where(j == self.n - 1, self.n - 2, j)
            j = np.where(j < 0, 0, j)
            u = t - j
            y0, y1 = self.y[j], self.y[j + 1]
            d0, d1 = self.dt[j], self.dt[j + 1]
            z = y1 - y0
            out[mid] = y0 + u * (d0 + u * (3 * z - 2 * d0 - d1 + u * (-2 * z + d0 + d1)))
        return out


# ----------------------------------------------------------------------------------
# Hermite finite elements (src/horton/ode2.cpp:31-321), integrals from the polynomials
# ----------------------------------------------------------------------------------

# Cubic Hermite shape functions on one unit segment u in [0,1], as integer
# coefficient lists [c0, c1, c2, c3]: node at the segment's left end (value, slope)
# and at its right end (value, slope).
_SHAPE = {
    ('left', 0): [1, 0, -3, 2],
    ('left', 1): [0, 1, -2, 1],
    ('right', 0): [0, 0, 3, -2],
    ('right', 1): [0, 0, -1, 1],
}


def _pmul(p, q):
    out = [Fraction(0)] * (len(p) + len(q) - 1)
    for i, a in enumerate(p):
        for j, b in enumerate(q):
            out[i + j] += Fraction(a) * Fraction(b)
    return out


def _pder(p):
    return [Fraction(k) * Fraction(c) for k, c in enumerate(p)][1:] or [Fraction(0)]


def _pint01(p):
    return sum((Fraction(c) / (k + 1) for k, c in enumerate(p)), Fraction(0))


def _piece(center, kind, seg, deriv):
    """Polynomial of basis function (center, kind) on segment [seg, seg+1], or None."""
    if center == seg:
        p = _SHAPE[('left', kind)]
    elif center == seg + 1:
        p = _SHAPE[('right', kind)]
    else:
        return None
    p = [Fraction(c) for c in p]
    return _pder(p) if deriv else p


_O_CACHE = {}


def _seg_overlap(funcs, seg):
    """Exact integral over [seg, seg+1] of the product of (center, kind, deriv) funcs,
    rounded to float64 once (as a rational table entry is)."""
    key = (tuple((c - seg, k, d) for c, k, d in funcs))
    if key in _O_CACHE:
        return _O_CACHE[key]
    prod = [Fraction(1)]
    val = 0.0
    for c, k, d in funcs:
        p = _piece(c, k, seg, d)
        if p is None:
            prod = None
            break
        prod = _pmul(prod, p)
    if prod is not None:
        val = float(_pint01(prod))
    _O_CACHE[key] = val
    return val


def hermite_overlap(xmax, funcs):
    """Sum over the (up to two) segments around the first function's node inside
    [0, xmax]: first the left segment, then the right one (ode2.cpp:54-76, 164-199).
    ``funcs`` = [(index, deriv), ...] with index = 2*center + kind."""
    parsed = [(i // 2, i % 2, bool(d)) for i, d in funcs]
    c0 = parsed[0][0]
    for c, _, _ in parsed[1:]:
        if abs(c - c0) >= 2:
            return 0.0
    res = 0.0
    if 1 <= c0 <= xmax:
        res += _seg_overlap(parsed, c0 - 1)
    if 0 <= c0 <= xmax - 1:
        res += _seg_overlap(parsed, c0)
    return res


def _node(x, center, kind, deriv):
    return float(x == center) if kind == deriv else 0.0


def build_ode2(b, a, f, bcs, npoint):
    """Galerkin matrix and right-hand side for u'' + b u' + a u = f on the t-axis with
    Hermite cubic elements; b, a, f given as merged [value, slope] node coefficients
    (src/horton/ode2.cpp:232-321). NaN in ``bcs`` = free. Returns (coeffs, rhs)."""
    nfn = 2 * npoint
    coeffs = np.zeros((nfn, nfn))
    rhs = np.zeros(nfn)
    skip = []
    for k, row in ((0, 0), (1, 1), (2, nfn - 2), (3, nfn - 1)):
        if bcs[k] == bcs[k]:
            coeffs[row, row] = 1.0
            rhs[row] = bcs[k]
            skip.append(row)
    skip = skip[:2]
    xmax = npoint - 1
    for irow in range(nfn):
        if irow in skip:
            continue
        begin = max(2 * (irow // 2 - 1), 0)
        end = min(2 * (irow // 2 + 2), nfn)
        for icol in range(begin, end):
            c = coeffs[irow, icol]
            c -= hermite_overlap(xmax, [(irow, True), (icol, True)])
            for ifoo in range(begin, end):
                if b[ifoo] != 0:
                    c += b[ifoo] * hermite_overlap(xmax, [(irow, False), (icol, True), (ifoo, False)])
                if a[ifoo] != 0:
                    c += a[ifoo] * hermite_overlap(xmax, [(irow, False), (icol, False), (ifoo, False)])
            c += _node(xmax, irow // 2, irow % 2, 0) * _node(xmax, icol // 2, icol % 2, 1)
            c -= _node(0, irow // 2, irow % 2, 0) * _node(0, icol // 2, icol % 2, 1)
            coeffs[irow, icol] = c
        acc = rhs[irow]
        for ifoo in range(begin, end):
            if f[ifoo] != 0:
                acc += f[ifoo] * hermite_overlap(xmax, [(irow, False), (ifoo, False)])
        rhs[irow] = acc
    return coeffs, rhs


def _merge(y, d):
    return np.array([y, d]).T.ravel()


def solve_ode2(b, a, f, bcs, rtf, j1, j2, j3, extrapolation, linsolve=None):
    """napmo/utilities/ode2.py:35-118: transform b, a, f (splines as (y, dx)) to the
    t-axis, build the FE system, solve, return the solution spline (y, dt)."""
    by, bd = b
    ay, ad = a
    fy, fd = f
    j1sq = j1 * j1
    by_new = j1 * by - j2 / j1
    bd_new = j2 * by + j1sq * bd + (j2 * j2 - j1 * j3) / j1sq
    ay_new = ay * j1sq
    ad_new = (ad * j1sq + 2 * ay * j2) * j1
    fy_new = fy * j1sq
    fd_new = (fd * j1sq + 2 * fy * j2) * j1
    new_bcs = (bcs[0], np.nan if bcs[1] is None else bcs[1] * j1[0],
               bcs[2], np.nan if bcs[3] is None else bcs[3] * j1[-1])
    new_bcs = [np.nan if v is None else v for v in new_bcs]
    npoint = len(by_new)
    builder = build_ode2
    try:
        from oracle import cext  # same assembly in C (bit-identical), much faster
        builder = cext.build_ode2
    except (ImportError, OSError):
        pass
    coeffs, rhs = builder(_merge(by_new, bd_new), _merge(ay_new, ad_new),
                          _merge(fy_new, fd_new), new_bcs, npoint)
    if linsolve is None:
        from scipy.sparse import csc_matrix
        from scipy.sparse.linalg import spsolve
        sol = spsolve(csc_matrix(coeffs), rhs)
    else:
        sol = linsolve(coeffs, rhs)
    uy = sol[::2].copy()
    ud = sol[1::2] / j1
    return Spline(uy, ud * j1, rtf, extrapolation)


# ----------------------------------------------------------------------------------
# Poisson solver (napmo/grids/poisson_solver.py:18-118) and synthesis
# ----------------------------------------------------------------------------------

def poisson_solver(grid, dens, lmax, linsolve=None):
    """U[atom][lm] = potential spline of each (atom, l, m), HORTON lm order."""
    dens = dens * grid.becke_weights
    U = []
    for k, atgrid in enumerate(grid.atgrids):
        p = dens[grid.offsets[k]:grid.offsets[k + 1]].copy()
        sph = spherical_expansion(atgrid, p, lmax)
        U.append(poisson_radial(atgrid, sph, lmax, linsolve))
    return U


def poisson_radial(atgrid, sph, lmax, linsolve=None):
    rtf = atgrid.rtf
    t = np.arange(rtf.npoint, dtype=np.float64)
    radii = rtf.radius(t)
    rw = atgrid.rw
    d1 = rtf.deriv(t)
    d2 = rtf.deriv2(t)
    d3 = rtf.deriv3(t)
    radii2 = radii * radii
    radii3 = radii2 * radii
    b = (2 / radii, -2 / radii2)
    result = []
    idx = 0
    for l in range(lmax + 1):
        radiiL = radii ** l
        radiiLm1 = radii ** (-l - 1)
        l2l = l * (l + 1)
        vmax_f = 1.0 / radii[-1] ** (l + 1) / (2 * l + 1)
        vmin_f = radii[0] ** l / (2 * l + 1)
        a = (-l2l / radii2, 2 * l2l / radii3)
        for _m in range(-l, l + 1):
            aux = np.array(sph[:, idx])
            dt = spline_dt(aux)
            dx = dt / d1
            f = (-PI_4 * aux, -PI_4 * dx)
            v_rmax = float(np.sum(aux * radiiL * rw)) * vmax_f
            v_rmin = float(np.sum(aux * radiiLm1 * rw)) * vmin_f
            bcs = (v_rmin, None, v_rmax, None)
            result.append(solve_ode2(b, a, f, bcs, rtf, d1, d2, d3, PotentialExtrapolation(l), linsolve))
            idx += 1
    return result


def eval_decomposition(splines, center, points):
    """src/horton/evaluate.cpp:129-223 without periodic images: sum over lm of
    U_lm(d) sqrt(2l+1)/sqrt(4 pi) d^-l S_lm(delta) at every point."""
    nspline = len(splines)
    lmax = int(round(math.sqrt(nspline))) - 1
    if (lmax + 1) ** 2 != nspline:
        raise ValueError('The number of splines does not match a well-defined lmax.')
    delta = points - center
    x, y, z = delta[:, 0], delta[:, 1], delta[:, 2]
    d = np.sqrt(x * x + y * y + z * z)
    out = splines[0](d) * INV_SQRT_4_PI
    if lmax > 0:
        work = pure_polynomials(z, x, y, lmax)
        counter = 0
        dpowl = np.ones_like(d)
        for l in range(1, lmax + 1):
            dpowl = dpowl / d
            factor = math.sqrt(2 * l + 1)
            for _m in range(-l, l + 1):
                s = splines[counter + 1](d)
                out = out + s * factor * INV_SQRT_4_PI * dpowl * work[counter]
                counter += 1
    return out


def compute_coulomb(grid, dens, lmax, linsolve=None):
    """napmo/wavefunction/ntwobody.py:15-46."""
    U = poisson_solver(grid, dens, lmax, linsolve)
    J = np.zeros(grid.size)
    for k in range(grid.ncenter):
        J += eval_decomposition(U[k], grid.origin[k], grid.points)
    return J


# Finite-difference stencils of RadialGrid::deriv2 (src/grid/radial.cpp:53-192) on the
# uniform omega = (i + 1) pi / (n + 1) axis. Rows 0..3 act on FF(size-1), ..., FF(size-7);
# rows size-4..size-1 on FF(6), ..., FF(0) (both in the order the reference adds them);
# interior rows on FF(idx+4), ..., FF(idx-4), idx = size - i - 1.
_D1_BOT = [(-1764, 4320, -5400, 4800, -2700, 864, -120), (-120, -924, 1800, -1200, 600, -180, 24),
           (24, -288, -420, 960, -360, 96, -12), (-12, 108, -540, 0, 540, -108, 12)]
_D1_TOP = [(-12, 108, -540, 0, 540, -108, 12), (12, -96, 360, -960, 420, 288, -24),
           (-24, 180, -600, 1200, -1800, 924, 120), (120, -864, 2700, -4800, 5400, -4320, 1764)]
_D2_BOT = [(1624, -6264, 10530, -10160, 5940, -1944, 274), (274, -294, -510, 940, -570, 186, -26),
           (-26, 456, -840, 400, 30, -24, 4), (4, -54, 540, -980, 540, -54, 4)]
_D2_TOP = [(4, -54, 540, -980, 540, -54, 4), (4, -24, 30, 400, -840, 456, -26),
           (-26, 186, -570, 940, -510, -294, 274), (274, -1944, 5940, -10160, 10530, -6264, 1624)]
_D1_MID = (144, -1536, 8064, -32256, 0, 32256, -8064, 1536, -144)
_D2_MID = (-36, 512, -4032, 32256, -57400, 32256, -4032, 512, -36)


def _stencil(coeffs, vals):
    acc = 0.0
    for c, v in zip(coeffs, vals):
        if c:
            acc += c * v
    return acc


def radial_deriv2(radii, points, f):
    """RadialGrid::deriv2 (src/grid/radial.cpp:53-192) for the Chebyshev grid
    r = R (1 + x) / (1 - x), x = cos omega: (2/r) f' + f'' from 7-point (ends) and 9-point
    finite differences in omega. Needs at least 8 points."""
    f = np.asarray(f, dtype=np.float64)
    n = len(f)
    if n < 8:
        raise ValueError('deriv2 needs at least 8 radial points')
    H = math.pi / (n + 1.0)
    FF = lambda k: f[k]  # noqa: E731
    F1 = np.zeros(n)
    F2 = np.zeros(n)
    bot = [FF(n - 1 - k) for k in range(7)]
    top = [FF(6 - k) for k in range(7)]
    for i in range(4):
        F1[i] = _stencil(_D1_BOT[i], bot) / (720.0 * H)
        F2[i] = _stencil(_D2_BOT[i], bot) / (360.0 * H * H)
        F1[n - 4 + i] = _stencil(_D1_TOP[i], top) / (720.0 * H)
        F2[n - 4 + i] = _stencil(_D2_TOP[i], top) / (360.0 * H * H)
    for i in range(4, n - 4):
        idx = n - i - 1
        vals = [FF(idx + 4 - k) for k in range(9)]
        F1[i] = _stencil(_D1_MID, vals) / (40320.0 * H)
        F2[i] = _stencil(_D2_MID, vals) / (20160.0 * H * H)
    D = np.zeros(n)
    R2 = radii * radii
    for i in range(n):
        idx = n - i - 1
        om = (i + 1) * H
        X, Y = math.cos(om), math.sin(om)
        Y2 = Y * Y * R2
        Y3 = Y2 * Y
        pre2 = (1.0 - X) * (1.0 - X)
        pre3 = pre2 * (1.0 - X)
        pre4 = pre3 * (1.0 - X)
        D[idx] = ((F1[i] * ((0.5 * pre3 / (R2 * Y)) - (0.25 * pre4 * X / Y3)))
                  - (F1[i] * pre2 / (radii * points[idx] * Y)) + (0.25 * F2[i] * pre4 / Y2))
    return D


def compute_kinetic(grid, psi, lmax):
    """napmo/wavefunction/nkinetic.py:15-73: -1/2 laplacian of psi, per atom from the
    spherical expansion of psi * bw, (deriv2 - l(l+1)/r^2) per (l, m), a cubic spline with
    PotentialExtrapolation(l), synthesised at every grid point."""
    U = np.zeros(grid.size)
    for k, ag in enumerate(grid.atgrids):
        f = psi[grid.offsets[k]:grid.offsets[k + 1]] * grid.becke_weights[grid.offsets[k]:grid.offsets[k + 1]]
        sph = spherical_expansion(ag, f, lmax)
        r = ag.r
        splines = []
        idx = 0
        for l in range(lmax + 1):
            pref = l * (l + 1) / (r * r)
            for _m in range(-l, l + 1):
                aux = np.array(sph[:, idx])
                y = radial_deriv2(ag.rtf.radii, r, aux) - pref * aux
                splines.append(Spline(y, spline_dt(y), ag.rtf, PotentialExtrapolation(l)))
                idx += 1
        U += eval_decomposition(splines, grid.origin[k], grid.points)
    return -0.5 * U


def compute_exchange_potential(grid, psi, D, lmax, linsolve=None):
    """PSIN.compute_exchange_potential (napmo/wavefunction/psi_numeric.py:186-215):
    aux[INDEX(s, v)] = V[psi_s psi_v] for s >= v (INDEX(i, j) = i(i+1)/2 + j, i >= j,
    psi_numeric.py:16-20), Kpot[j] = sum_k sum_l psi_k aux[INDEX(l, j)] D[k, l].
    psi [ndim, N]; returns Kpot [ndim, N]."""
    ndim = psi.shape[0]
    aux = {}
    for s in range(ndim):
        for v in range(s + 1):
            aux[(s, v)] = compute_coulomb(grid, psi[s] * psi[v], lmax, linsolve)
    K = np.zeros_like(psi)
    for j in range(ndim):
        acc = np.zeros(psi.shape[1])
        for k in range(ndim):
            for l in range(ndim):
                acc = acc + psi[k] * aux[(max(l, j), min(l, j))] * D[k, l]
        K[j] = acc
    return K


def spherical_average(atgrid, f):
    """AtomicGrid.spherical_average (napmo/grids/atomic.py:77-90): per radial shell,
    sum_j f w / sum_j w, both sums in point order (src/grid/atomic_grid.cpp:53-80)."""
    nr, na = atgrid.nrad, atgrid.nang
    fw = (np.asarray(f, dtype=np.float64) * atgrid.weights).reshape(nr, na)
    w = atgrid.weights.reshape(nr, na)
    num = np.zeros(nr)
    den = np.zeros(nr)
    for j in range(na):
        num += fw[:, j]
        den += w[:, j]
    return num / den


def compute_dpsi(grid, lmax, phi, doi, oi, ri, V_tot, mass_inv, linsolve=None, evaluate=None):
    """compute_dpsi (napmo/wavefunction/ndpsi.py:15-89), the Becke-Dickson orbital
    correction: per atom, the spherical expansion mu_lm of (phi doi - ri) bw and, per
    (l, m), the radial solution of

        u'' + (2/r) u' + a u = f,   a = -2/m_inv (l(l+1)/r^2 + V_sph - oi - doi),
        f = -2/m_inv mu_lm,   u(r_0) = u(r_max) = 0,

    with V_sph the spherical average of V_tot bw; the (atom, l, m) solutions are splined
    with PotentialExtrapolation(l) and synthesised at every grid point. ``a`` is a spline
    built from its values (slopes from the spline system), as the reference builds it;
    ``evaluate(splines, centre, points)`` may replace eval_decomposition."""
    evaluate = evaluate or eval_decomposition
    bw = grid.becke_weights
    src = (phi * doi - ri) * bw
    c = -2.0 / mass_inv
    out = np.zeros(grid.size)
    for k, ag in enumerate(grid.atgrids):
        sl = slice(grid.offsets[k], grid.offsets[k + 1])
        sph = spherical_expansion(ag, src[sl].copy(), lmax)
        V_sph = spherical_average(ag, V_tot[sl] * bw[sl])
        rtf = ag.rtf
        t = np.arange(rtf.npoint, dtype=np.float64)
        r = ag.r
        j1, j2, j3 = rtf.deriv(t), rtf.deriv2(t), rtf.deriv3(t)
        r2 = r * r
        rinv = 1.0 / r
        r2inv = 1.0 / r2
        b = (2.0 * rinv, -2.0 * r2inv)
        splines = []
        idx = 0
        for l in range(lmax + 1):
            ay = c * ((l * (l + 1.0) * r2inv) + (V_sph - oi - doi))
            a = (ay, spline_dt(ay) / j1)
            for _m in range(-l, l + 1):
                aux = np.array(sph[:, idx])
                f = (c * aux, c * (spline_dt(aux) / j1))
                splines.append(solve_ode2(b, a, f, (0.0, None, 0.0, None), rtf, j1, j2, j3,
                                          PotentialExtrapolation(l), linsolve))
                idx += 1
        out += evaluate(splines, grid.origin[k], grid.points)
    return out


def becke_dickson_step(grid, lmax, psi, T, V_tot, e_prev, mass_inv, linsolve=None, evaluate=None):
    """The per-orbital part of PSIO_BECKE.optimize (napmo/wavefunction/psi_optimization.py
    :175-203): residual R_i = T_i + V_i - e_i psi_i, energy correction
    de_i = integrate(R_i psi_i), correction dpsi_i = compute_dpsi(psi_i, de_i, e_i, R_i, V_i).
    psi, T, V_tot [norb, N] (T and V_tot already applied to psi); returns (dpsi, de, R)."""
    R = np.array([Ti + Vi - (e * p) for Ti, Vi, e, p in zip(T, V_tot, e_prev, psi)])
    de = np.array([grid.integrate(r * p) for r, p in zip(R, psi)])
    dpsi = np.array([compute_dpsi(grid, lmax, p, d, e, r, V, mass_inv, linsolve, evaluate)
                     for p, d, e, r, V in zip(psi, de, e_prev, R, V_tot)])
    return dpsi, de, R


def compute_nuclear(points, charges):
    """napmo/wavefunction/nnuclear.py:15-37: V(p) = sum_A Z_A / |p - R_A|."""
    V = np.zeros(len(points))
    for z, o in charges:
        V += z / np.sqrt(((np.asarray(o) - points) ** 2).sum(1))
    return V


def coulomb_at(grid, dens, lmax, points, linsolve=None):
    """Potential of ``dens`` (given on ``grid``) at arbitrary points: the Becke-Poisson
    expansion of ``grid``'s atoms synthesised at ``points``. The reference evaluates the
    other species' compute_coulomb on the grid both species share
    (napmo/wavefunction/psi_hybrid.py:81-110, napmo/solver/hf_solver.py:95-98); with one
    grid per species the same expansion is synthesised at the other species' points."""
    U = poisson_solver(grid, dens, lmax, linsolve)
    J = np.zeros(len(points))
    for k in range(grid.ncenter):
        J += eval_decomposition(U[k], grid.origin[k], points)
    return J


def apmo_coulomb_build(species, evaluate=None, linsolve=None):
    """Coulomb part of the multi-species (APMO) Fock build. ``species``: list of dicts
    with grid, phi [N_s, n_s], P, charge, lmax. Per species s:

        v_eff = q_s^2 V[rho_s] + q_s sum_{t != s} q_t V[rho_t]   (both at s's points)
        F_s = Phi_s^T diag(w bw v_eff) Phi_s   (= q_s^2 J_self + the coupling matrix of
              napmo/wavefunction/psi_numeric.py:217-241 / psi_hybrid.py:81-110)

    Returns per species (F_s, E_trace = 1/2 sum P o F, E_grid = 1/2 sum w bw rho v_eff,
    E_self = 1/2 q^2 sum w bw rho V[rho_s]). ``evaluate(grid, dens, lmax, points)`` may
    replace coulomb_at (e.g. with the C oracle)."""
    evaluate = evaluate or (lambda g, d, l, pts: coulomb_at(g, d, l, pts, linsolve))
    rhos = [density_from_dm(s['phi'], s['P']) for s in species]
    out = []
    for a, sa in enumerate(species):
        g = sa['grid']
        q = sa['charge']
        v_self = evaluate(g, rhos[a], sa['lmax'], g.points)
        v_cross = np.zeros(g.size)
        for b, sb in enumerate(species):
            if b != a:
                v_cross += sb['charge'] * evaluate(sb['grid'], rhos[b], sb['lmax'], g.points)
        v_eff = q * q * v_self + q * v_cross
        F = coupling_matrix(sa['phi'], v_eff, g.becke_weights, g.weights)
        wbw = g.becke_weights * g.weights
        out.append((F, 0.5 * float(np.sum(sa['P'] * F)), 0.5 * float(np.sum(wbw * rhos[a] * v_eff)),
                    0.5 * q * q * float(np.sum(wbw * rhos[a] * v_self))))
    return out


# ----------------------------------------------------------------------------------
# Gaussian basis (src/system/gto.cpp, src/ints/overlap.cpp, napmo/gto/basis_set.py)
# ----------------------------------------------------------------------------------

def factorial2(n):
    """src/utils/utils.cpp:12-22 (0 for n < 1)."""
    if n < 1:
        return 0
    out = 1
    while n > 0:
        out *= n
        n -= 2
    return out


def cartesian_powers(l):
    """Component order of napmo/gto/basis_set.py:122-143 (xx, xy, xz, yy, yz, zz)."""
    return [(i, k - j, j) for k, i in enumerate(reversed(range(l + 1))) for j in range(k + 1)]


def primitive_norm(zeta, lxyz):
    """src/system/gto.cpp:29-39."""
    ltot = sum(lxyz)
    return (math.pow(2.0 * zeta / math.pi, 0.75)) / math.sqrt(
        factorial2(abs(2 * lxyz[0] - 1)) * factorial2(abs(2 * lxyz[1] - 1))
        * factorial2(abs(2 * lxyz[2] - 1)) / math.pow(4.0 * zeta, ltot))


def same_center_overlap(za, zb, lxyz):
    """<g_a|g_b> of two unnormalised Cartesian primitives with equal centre and
    powers: (pi/g)^(3/2) prod (2l-1)!!/(2g)^l, g = za + zb (closed form of the
    Obara-Saika recursion of src/ints/overlap.cpp:12-135 at PA = PB = 0)."""
    g = za + zb
    val = math.sqrt(math.pi / g) * (math.pi / g)
    for l in lxyz:
        val *= (factorial2(2 * l - 1) if l > 0 else 1) / (2.0 * g) ** l
    return val


def contraction_norm(zetas, coeffs, lxyz):
    """src/system/gto.cpp:180: 1/sqrt(<phi|phi>) over coefficient-scaled normalised
    primitives."""
    s = 0.0
    for za, ca in zip(zetas, coeffs):
        na = primitive_norm(za, lxyz)
        for zb, cb in zip(zetas, coeffs):
            nb = primitive_norm(zb, lxyz)
            s += ca * na * cb * nb * same_center_overlap(za, zb, lxyz)
    return 1.0 / math.sqrt(s)


class Basis:
    """Flat list of contracted Cartesian functions: each entry
    (origin[3], lxyz, zetas, coeffs*prim_norm, contraction_norm)."""

    def __init__(self):
        self.functions = []

    def add_shell(self, origin, l, zetas, coeffs):
        for lxyz in cartesian_powers(l):
            cn = [c * primitive_norm(z, lxyz) for z, c in zip(zetas, coeffs)]
            self.functions.append((np.asarray(origin, dtype=np.float64), lxyz, list(zetas), cn,
                                   contraction_norm(zetas, coeffs, lxyz)))

    @property
    def nbasis(self):
        return len(self.functions)


def basis_compute(basis, points):
    """Phi[p, mu] (src/system/basis_set.cpp:69-82, gto.cpp:41-51, 182-192):
    norm_c * sum_k c_k N_k dx^lx dy^ly dz^lz exp(-zeta_k r^2). [N, n]."""
    out = np.empty((points.shape[0], basis.nbasis))
    for mu, (o, lxyz, zetas, cn, norm) in enumerate(basis.functions):
        d = points - o
        r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        ang = np.power(d[:, 0], lxyz[0]) * np.power(d[:, 1], lxyz[1]) * np.power(d[:, 2], lxyz[2])
        acc = np.zeros(points.shape[0])
        for z, c in zip(zetas, cn):
            acc += c * ang * np.exp(-z * r2)
        out[:, mu] = acc * norm
    return out


def basis_gradient(basis, points):
    """grad_r Phi [3, N, n]: for phi = N_c x^a y^b z^c sum_k c_k e^{-zeta_k r^2},
    d phi/dx = N_c y^b z^c (a x^{a-1} g + x^{a+1} h), g = sum_k c_k e_k,
    h = -2 sum_k c_k zeta_k e_k (differentiation of src/system/gto.cpp:41-51)."""
    out = np.empty((3, points.shape[0], basis.nbasis))
    for mu, (o, lxyz, zetas, cn, norm) in enumerate(basis.functions):
        d = points - o
        r2 = (d * d).sum(1)
        g = np.zeros(len(d))
        h = np.zeros(len(d))
        for z, c in zip(zetas, cn):
            e = c * np.exp(-z * r2)
            g += e
            h -= 2.0 * z * e
        pw = [np.power(d[:, k], lxyz[k]) for k in range(3)]
        for k in range(3):
            a = lxyz[k]
            dk = (a * np.power(d[:, k], a - 1) if a else 0.0) * g + pw[k] * d[:, k] * h
            others = pw[(k + 1) % 3] * pw[(k + 2) % 3]
            out[k, :, mu] = norm * dk * others
    return out


def centre_gradient(basis, points, P, s):
    """dE/dR_c [ncentre, 3] of E = 1/2 sum_p s_p rho_p / ... with s = w bw v held fixed:
    sum_p s_p d rho_p/dR_c = -2 sum_p s_p sum_{mu on c} grad phi_mu(p) (Phi P)_p mu; centres
    are the distinct function origins in first-appearance order."""
    centres, cof = [], []
    for o, *_ in basis.functions:
        for c, x in enumerate(centres):
            if np.array_equal(x, o):
                cof.append(c)
                break
        else:
            centres.append(o)
            cof.append(len(centres) - 1)
    X = basis_compute(basis, points) @ P
    dphi = basis_gradient(basis, points)
    per_fn = np.einsum('p,kpm,pm->mk', s, dphi, X)
    grad = np.zeros((len(centres), 3))
    for mu, c in enumerate(cof):
        grad[c] += per_fn[mu]
    return -2.0 * grad, np.array(centres)


# ----------------------------------------------------------------------------------
# Contractions and energies
# ----------------------------------------------------------------------------------

def density_from_dm(phi, P):
    """rho_p = phi_p^T P phi_p (napmo/wavefunction/psi_hybrid.py:144-154). phi [N, n]."""
    return np.einsum('pm,mn,pn->p', phi, P, phi)


def _butterfly16(v):
    """Sum of 16 values in the order of a 16-lane xor butterfly (offsets 8, 4, 2, 1) read at
    lane 0: the order in which the GPU forms a group's sum."""
    v = list(v)
    for o in (8, 4, 2, 1):
        v = [v[i] + v[i ^ o] for i in range(16)]
    return v[0]


def density_orbitals_screened(phi, C, eta, S, ld=None):
    """rho = eta sum_c X_pc^2, X = Phi C, with the screening of the streamed orbital-form
    density (napmo_amd/csrc/density.hip, k_density_orb_stream with SCR; the order and block
    exponents of FockPlan::density_screen_setup): per 16-row tile t and function f the
    exponent e[t][f] of max |Phi| over the tile (clamped to [-127, 127], -128 if zero); the
    functions ordered by the number of tiles where e reaches the largest one less 30 (fewest
    first, ties by index); per 16-function group g of that order eT[t][g] = max e and
    eC[g] = exponent of the butterfly sum over the group of max_c |C|; the (t, g) products
    with eT + eC <= -S are left out. Returns (rho, order, kept fraction). phi [N, n], C [n, nocc]."""
    N, n = phi.shape
    ld = ld or -(-n // 64) * 64
    nt = -(-N // 16)
    P = np.zeros((nt * 16, ld))
    P[:N, :n] = phi
    m = np.abs(P).reshape(nt, 16, ld).max(axis=1)
    e = np.where(m > 0.0, np.clip(_pow2_ceil_exp(m), -127, 127), -128)
    en = e[:, :n]
    top = en.max()
    reach = ((en != -128) & (en >= top - 30)).sum(axis=0)
    order = np.concatenate([np.argsort(reach, kind='stable'), np.arange(n, ld)])
    G = -(-n // 16)
    eT = e[:, order].reshape(nt, ld // 16, 16).max(axis=2)[:, :G]
    cmax = np.zeros(16 * G)
    cmax[:n] = np.abs(C[order[:n]]).max(axis=1)
    csum = np.array([_butterfly16(cmax[16 * g:16 * g + 16]) for g in range(G)])
    eC = np.where(csum > 0.0, _pow2_ceil_exp(csum), -4000)
    keep = eT + eC[None, :] > -S                                     # [nt, G]
    Pp = P[:, order[:16 * G]] if 16 * G <= ld else None
    Cp = np.zeros((16 * G, C.shape[1]))
    Cp[:n] = C[order[:n]]
    X = np.zeros((nt * 16, C.shape[1]))
    for g in range(G):
        rows = np.repeat(keep[:, g], 16)
        X[rows] += Pp[rows, 16 * g:16 * g + 16] @ Cp[16 * g:16 * g + 16]
    rho = eta * (X[:N] ** 2).sum(axis=1)
    return rho, order, float(keep.mean())


def density_from_dm_screened(phi, dm, epsilon, dmmaxrow, out):
    """nwavefunction_compute_density_from_dm (src/wavefunction/nwavefunction.cpp:109-159)
    with its sequential screening: one running epsilon over the points in order; at the
    first point whose bound falls below it the routine returns and leaves the rest of
    ``out`` untouched; otherwise epsilon /= absmax * n * n before that point's screening.
    Only the lower triangle of dm is read. phi [N, n]; out [N] updated in place."""
    N, n = phi.shape
    for p in range(N):
        row = phi[p]
        if epsilon > 0:
            absmax = 0.0
            for i in range(n):
                absmax = max(abs(row[i]), absmax)
            upper = 0.0
            for i in range(n):
                upper += abs(row[i]) * dmmaxrow[i]
            upper *= n * absmax
            if upper < epsilon:
                return out
            epsilon /= absmax * n * n
        rho = 0.0
        for i in range(n):
            if epsilon > 0 and abs(row[i]) * dmmaxrow[i] < epsilon:
                continue
            tmp = 0.0
            for j in range(i - 1, -1, -1):
                tmp += row[j] * dm[i, j]
            rho += (2.0 * tmp + dm[i, i] * row[i]) * row[i]
        out[p] = rho
    return out


def moment_polynomials(delta, lmax, mtype):
    """Polynomial rows of fill_polynomials_wrapper (src/horton/dot_multi.cpp:55-80):
    mtype 1 Cartesian (moments.cpp:24-57), 2 pure of delta, 3 powers of |delta|
    (moments.cpp:138-155), 4 pure of the unit vector. delta [npt, 3] -> [nmoment-1, npt]."""
    dx, dy, dz = delta[:, 0], delta[:, 1], delta[:, 2]
    if mtype == 1:
        out = [dx, dy, dz]
        old_off, old_n = 0, 3
        for l in range(2, lmax + 1):
            new_n, new_off = old_n + l + 1, old_off + old_n
            for i in range(old_n):
                out.append(out[0] * out[old_off + i])
            for i in range(l):
                out.append(out[1] * out[new_off - l + i])
            out.append(out[2] * out[new_off - 1])
            old_n, old_off = new_n, new_off
        return np.array(out)
    if mtype == 3:
        r = np.sqrt(dx * dx + dy * dy + dz * dz)
        out = [r]
        for _ in range(lmax - 1):
            out.append(r * out[-1])
        return np.array(out)
    if mtype == 4:
        r = np.sqrt(dx * dx + dy * dy + dz * dz)
        dx, dy, dz = dx / r, dy / r, dz / r
    return pure_polynomials(dz, dx, dy, lmax)


def moment_count(lmax, mtype):
    if lmax <= 0:
        return 1
    return {1: (lmax + 1) * (lmax + 2) * (lmax + 3) // 6, 3: lmax + 1}.get(mtype, (lmax + 1) ** 2)


def dot_multi_moments(data, points, center, lmax, mtype, segments, output):
    """src/horton/dot_multi.cpp:148-186: output[seg, m] += sum over the segment's points
    of prod_v data[v, p] * poly_m(p - center), accumulated point by point in order; the
    segment walk is shift_segment's (dot_multi.cpp:46-52). data [nvector, npt]; output
    [nseg, nmoment] updated in place."""
    if lmax < 0:
        raise ValueError('lmax can not be negative.')
    if mtype < 1 or mtype > 4:
        raise ValueError('mtype should be 1, 2, 3 or 4.')
    data = np.atleast_2d(data)
    nvector, npt = data.shape
    term = data[nvector - 1].copy()
    for v in range(nvector - 2, -1, -1):
        term = term * data[v]
    nm = output.shape[1]
    rows = np.zeros((npt, nm))
    rows[:, 0] = term
    if lmax > 0:
        rows[:, 1:] = term[:, None] * moment_polynomials(points - center, lmax, mtype)[:nm - 1].T
    seg_of = np.empty(npt, dtype=np.int64)
    k, end = 0, int(segments[0])
    for p in range(npt):
        if p == end:
            k += 1
            end += int(segments[k])
        seg_of[p] = k
    for k in np.unique(seg_of):
        sel = rows[seg_of == k]
        # sequential left-to-right sums, as the reference's running accumulation
        output[k] = np.cumsum(np.vstack([output[k][None, :], sel]), axis=0)[-1]
    return output


def coupling_matrix(phi, v, bw, w):
    """M_ij = sum_p phi_i phi_j v bw w, upper triangle then mirrored
    (src/wavefunction/nwavefunction.cpp:77-107). phi [N, n]."""
    s = v * bw * w
    return phi.T @ (phi * s[:, None])


def two_body_matrix(phi, J, K, x_factor, bw, w):
    """G = Coulomb + x * Exchange (nwavefunction.cpp:15-75); the exchange block takes
    its upper triangle from sum_p phi_i K_j bw w and mirrors it. phi, K [N, n]."""
    G = coupling_matrix(phi, J, bw, w)
    if x_factor != 0.0:
        E = phi.T @ (K * (bw * w)[:, None])
        E = np.triu(E) + np.triu(E, 1).T
        G = G + x_factor * E
    return G


def coulomb_energy(P, Jmat):
    """E_J = 1/2 sum P o J (the 1/2 D o G term of napmo/scf/scf.py:494-520)."""
    return 0.5 * float(np.sum(P * Jmat))


def grid_fock_build(grid, phi, P, lmax, linsolve=None):
    """One grid Fock build: rho from P, Poisson, J = Phi^T diag(w bw v) Phi, E_J."""
    rho = density_from_dm(phi, P)
    v = compute_coulomb(grid, rho, lmax, linsolve)
    J = coupling_matrix(phi, v, grid.becke_weights, grid.weights)
    return rho, v, J, coulomb_energy(P, J)


# ----------------------------------------------------------------------------------
# Cross-species glue: Lowdin2 grids, common points, epc17-2
# ----------------------------------------------------------------------------------

def read_lowdin2_grid(path):
    """napmo/grids/becke.py:42-50: int64 record (4 x npoint), float64 record reshaped
    [npoint x 4] column-major; rows (x, y, z, Becke weight), src/grid/becke.cpp:57-77."""
    from scipy.io import FortranFile
    with FortranFile(path, 'r') as f:
        size = f.read_ints(dtype=np.int64)[0]
        pw = f.read_reals(dtype=np.float64)
    pw = pw.reshape(int(size / 4), 4, order='F')
    return np.ascontiguousarray(pw)


def common_index(points_a, points_b):
    """MultiGrid::get_common_index (src/grid/multi_grid.cpp:37-86): the grids are swapped
    when b is larger, then every (i, j) with equal coordinates in i-major, j-minor order,
    i in the (possibly swapped) first grid. Returns int array [n x 2]."""
    A, B = np.asarray(points_a), np.asarray(points_b)
    if B.shape[0] > A.shape[0]:
        A, B = B, A
    rows = []
    for i in range(A.shape[0]):
        eq = (A[i, 0] == B[:, 0]) & (A[i, 1] == B[:, 1]) & (A[i, 2] == B[:, 2])
        for j in np.flatnonzero(eq):
            rows.append((i, int(j)))
    return np.array(rows, dtype=np.int64).reshape(-1, 2)


def epc17_2(rhoE, rhoN):
    """napmo/functional/interspecies_correlation.py:36-66 (same expressions, same branch
    rule: the simplified potentials for the whole array when any density < 1e-200)."""
    a, b, c = 2.35, 2.4, 6.6
    denominator = a - b * np.sqrt(rhoE * rhoN) + c * rhoE * rhoN
    c_ene = -rhoN / denominator
    if np.any(np.abs(rhoE) < 1.0e-200) or np.any(np.abs(rhoN) < 1.0e-200):
        c_pot_A = (b * np.sqrt(rhoE) * rhoN ** (1.5) - 2.0 * a * rhoN) / denominator ** 2 / 2.0
        c_pot_B = (b * np.sqrt(rhoN) * rhoE ** (1.5) - 2.0 * a * rhoE) / denominator ** 2 / 2.0
    else:
        c_pot_A = (rhoE * rhoN * (c * rhoN - b * rhoN / (2.0 * np.sqrt(rhoE * rhoN)))
                   - rhoN * denominator) / denominator ** 2
        c_pot_B = (rhoN * rhoE * (c * rhoE - b * rhoE / (2.0 * np.sqrt(rhoE * rhoN)))
                   - rhoE * denominator) / denominator ** 2
    return c_ene, c_pot_A, c_pot_B


def c_2species_grid(index, rho_a, rho_b, bw_b, w_b, vrho_a, vrho_b):
    """The functional step of compute_c_2species_grid (napmo/wavefunction/psi_analytic.py
    :241-258): densities taken at the common points, epc17-2, energy
    (bw_b w_b c_zk rho).sum(), potentials added at the common points (in place)."""
    rho = np.take(rho_a, index[:, 0], axis=0)
    other_rho = np.take(rho_b, index[:, 1], axis=0)
    c_zk, c_vrho, c_other_vrho = epc17_2(rho, other_rho)
    aux = (bw_b[index[:, 1]] * w_b[index[:, 1]] * c_zk * rho).sum()
    vrho_a[index[:, 0]] += c_vrho
    vrho_b[index[:, 1]] += c_other_vrho
    return aux


# ----------------------------------------------------------------------------------
# Digit-sliced (int8) grid contraction: an exact restatement of the product's integer
# arithmetic (TEST INFRASTRUCTURE: it checks the GPU kernel bit for bit; its own accuracy is
# checked against the FP64 contraction coupling_matrix).
# ----------------------------------------------------------------------------------

def _pow2_ceil_exp(m):
    """Smallest e with 2^e >= m elementwise (0 where m == 0)."""
    m = np.asarray(m, dtype=np.float64)
    mant, ex = np.frexp(m)
    e = np.where(mant == 0.5, ex - 1, ex)
    return np.where(m > 0, e, 0).astype(np.int64)


def _split5(J):
    """J in [-2^34, 2^34] (int64) -> five balanced base-128 digits, J = sum d_k 2^(7 (5 - k)),
    split as the kernel does: h = floor((J + 2^13) / 2^14), lo = J - 2^14 h in [-2^13, 2^13);
    d1 = floor((h + 2^13) / 2^14), r = h - 2^14 d1, d2 = floor((r + 64) / 128), d3 = r - 128 d2,
    d4 = floor((lo + 64) / 128), d5 = lo - 128 d4 (each in [-64, 64])."""
    h = (J + (1 << 13)) >> 14
    lo = J - (h << 14)
    d1 = (h + (1 << 13)) >> 14
    r = h - (d1 << 14)
    d2 = (r + 64) >> 7
    d3 = r - (d2 << 7)
    d4 = (lo + 64) >> 7
    d5 = lo - (d4 << 7)
    return [d1, d2, d3, d4, d5]


OZ_ZERO_EXP = -2000


def int8_compact_order(phi, wbw, chunk):
    """The screening order of the int8 contraction (napmo_amd/csrc/ozaki.hip,
    ozaki_compact_order): functions by the number of chunks in which the exponent of their
    chunk maximum of |sqrt|wbw| phi| reaches the largest such exponent less 30, fewest first,
    ties by index."""
    N, n = phi.shape
    nchunk = -(-N // chunk)
    G = np.zeros((nchunk * chunk, n))
    G[:N] = np.sqrt(np.abs(wbw))[:, None] * phi
    m = np.abs(G).reshape(nchunk, chunk, n).max(axis=1)            # [nchunk, n]
    e = np.where(m > 0.0, _pow2_ceil_exp(m), OZ_ZERO_EXP)
    reach = (e >= e.max() - 30).sum(axis=0)
    return np.argsort(reach, kind='stable')


def int8_fock(phi, wbw, v, chunk, stage=64, screen=0, order=None, stats=None):
    """F = Phi^T diag(wbw v) Phi as the int8 MFMA kernel computes it (napmo_amd/csrc/ozaki.hip).

    At = sqrt|wbw| Phi is written per (function, chunk) as a 35-bit integer at the chunk's
    power-of-two maximum (J = rint(At 2^(34 - eA)), |J| <= 2^34); B = sign(wbw) v At as
    JB = rint(JA vs 2^-(eV + eB)) with eV the exponent of the chunk's max|vs| and eB that of the
    bound 2^-6 max_stages(max|vs| 2^-eV (max|d1| + 1)), so |JB| <= 2^34. Both split into five
    balanced base-128 digits; the digit products of level L = i + j (1-based) <= 6 are summed
    exactly over the chunk, one integer S_L per level (|S_L| < 2^27 for up to 63 stages), and the
    chunk's block is 2^(e_row + e_col) ((((S2 2^-12 + S3 2^-19) + S4 2^-26) + S5 2^-33) + S6 2^-40);
    chunks are summed in order. Products of level >= 7 (below 2^-35 of the leading one) are the
    only ones dropped. phi [N, n].

    Screening (screen = S > 0, with ``order`` the operand order of the functions): in each
    chunk the 16 x 16 block (I, J), I <= J in that order, is left out when
    eI + eJ + eV + lg <= -S, eI the largest chunk exponent eA of the block's functions (functions
    that are zero in the chunk count as OZ_ZERO_EXP) and 2^lg >= chunk; every element of the
    left-out chunk sum is below 2^-S. ``stats`` (a dict) receives the counts of blocks kept and
    left out."""
    if order is not None:
        order = np.asarray(order)
        phi = phi[:, order]
    N, n = phi.shape
    nb = -(-n // 16)
    lg = int(np.ceil(np.log2(chunk)))
    nchunk = -(-N // chunk)
    Nq = nchunk * chunk
    G = np.zeros((Nq, n))
    G[:N] = np.sqrt(np.abs(wbw))[:, None] * phi
    vs = np.zeros(Nq)
    vs[:N] = np.where(wbw < 0.0, -v, v)
    F = np.zeros((n, n))
    for c in range(nchunk):
        sl = slice(c * chunk, (c + 1) * chunk)
        Gc = G[sl]
        eA = _pow2_ceil_exp(np.abs(Gc).max(axis=0))
        JA = np.rint(np.ldexp(Gc, (34 - eA)[None, :])).astype(np.int64)
        dA = _split5(JA)
        vc = vs[sl]
        ns = chunk // stage
        maxv = np.abs(vc).reshape(ns, stage).max(axis=1)
        eV = int(_pow2_ceil_exp(maxv.max()))
        md1 = np.abs(dA[0]).reshape(ns, stage, n).max(axis=1)       # [ns, n]
        bound = np.zeros(n)
        for t in range(ns):
            bound = np.fmax(bound, np.ldexp(maxv[t], -eV) * (md1[t] + 1).astype(np.float64))
        bound = np.ldexp(bound, -6)
        eB = _pow2_ceil_exp(bound)
        scB = np.ldexp(1.0, -eV - eB)
        y = vc[:, None] * scB[None, :]
        JB = np.rint(JA.astype(np.float64) * y).astype(np.int64)
        dB = _split5(JB)
        S = {L: np.zeros((n, n), dtype=np.int64) for L in range(2, 7)}
        for t in range(ns):
            st = slice(t * stage, (t + 1) * stage)
            r = [d[st].astype(np.float64) for d in dB]   # rows: B digits
            a = [d[st].astype(np.float64) for d in dA]   # columns: A digits
            for L in range(2, 7):
                T = sum(r[i].T @ a[L - 2 - i] for i in range(L - 1))
                S[L] += np.rint(T).astype(np.int64)       # exact integers
        val = (((S[2] * 2.0 ** -12 + S[3] * 2.0 ** -19) + S[4] * 2.0 ** -26) + S[5] * 2.0 ** -33) + S[6] * 2.0 ** -40
        blockF = np.ldexp(val, (eA + eV + eB)[:, None] + eA[None, :])
        if screen > 0:
            ez = np.full(16 * nb, OZ_ZERO_EXP)
            ez[:n] = np.where(np.abs(Gc).max(axis=0) > 0.0, eA, OZ_ZERO_EXP)
            eblk = ez.reshape(nb, 16).max(axis=1)
            keep = np.repeat(np.repeat(eblk[:, None] + eblk[None, :] + eV + lg > -screen, 16, 0), 16, 1)[:n, :n]
            if stats is not None:
                upper = np.triu(np.ones((nb, nb), dtype=bool))
                kb = (eblk[:, None] + eblk[None, :] + eV + lg > -screen) & upper
                stats['kept'] = stats.get('kept', 0) + int(kb.sum())
                stats['skipped'] = stats.get('skipped', 0) + int((upper & ~kb).sum())
            blockF = np.where(keep, blockF, 0.0)
        F = F + blockF
    U = np.triu(F)
    F = U + np.triu(U, 1).T
    if order is not None:
        out = np.empty_like(F)
        out[np.ix_(order, order)] = F
        F = out
    return F
