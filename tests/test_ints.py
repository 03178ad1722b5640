"""Pins of the integral engine (libcint is absent, so nothing here is compared
with the reference's own integrals: parity against libcint is unpinned).

* Boys function: closed form F_0 = sqrt(pi/T) erf(sqrt T)/2 and the series.
* s-type overlap, kinetic, nuclear attraction, (a|b) and (c|ab): closed forms.
* p, d, f functions: derivative relations.  A Cartesian Gaussian of degree l
  on centre A is a derivative of an s Gaussian with respect to A, e.g.
  x_A exp(-a r_A^2) = (1/2a) d/dA_x exp(-a r_A^2); integrals over p/d/f
  shells are checked against central finite differences of s-shell integrals
  (so a wrong index, sign, Hermite term or spherical coefficient fails).
* normalisation of every shell of the three basis sets (S diagonal = 1 and
  orthogonality inside a shell).
* four-centre (ab|cd) (host, for the SAD guess's atomic SCFs): s closed form,
  8-fold symmetry and positive semidefiniteness, and the limit of a vanishing
  ket exponent, which turns (ab|cd) into the three-centre (ab|c) for every l.
"""
import math

import numpy as np
import pytest
from scipy.special import erf

from chem_amd import _lib, desc, ints


def boys(m, T):
    F = np.zeros(m + 1)
    _lib.check(_lib.lib().mchip_boys(m, T, _lib.dptr(F)))
    return F


@pytest.mark.parametrize("T", [0.0, 1e-8, 0.3, 1.7, 7.9, 19.99, 25.0, 39.97, 40.0, 41.3, 80.0, 300.0])
def test_boys_function(T):
    F = boys(12, T)
    # F_0 closed form
    f0 = 1.0 if T == 0 else 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
    assert abs(F[0] - f0) <= 2e-15 * max(f0, 1e-300) + 1e-300
    # every order against 50-digit quadrature of int_0^1 t^2m exp(-T t^2) dt
    import mpmath
    mpmath.mp.dps = 50
    for m in range(13):
        ref = float(mpmath.quad(lambda t: t ** (2 * m) * mpmath.exp(-T * t * t), [0, 1]))
        assert abs(F[m] - ref) <= 4e-15 * ref + 1e-300, (m, F[m], ref)


def _shell(l, alpha, center, atom=0):
    return desc.Shell(l, np.array(center, dtype=float), atom, np.array([alpha]), np.array([1.0]))


class _Mol:
    """Minimal molecule stand-in for single-primitive test shells."""

    def __init__(self, shells, dfshells=None, atoms=None):
        self.shells = shells
        self.dfshells = dfshells or [_shell(0, 1.0, [0, 0, 0])]
        self.atoms = atoms or [desc.Atom(1, 0.0, 0.0, 0.0)]
        self.basis_handle = desc.BasisHandle(self.shells)
        self.dfbasis_handle = desc.BasisHandle(self.dfshells)
        self.nbf = self.basis_handle.nbf
        self.naux = self.dfbasis_handle.nbf

    def charges(self):
        Z = np.array([a.Z for a in self.atoms], dtype=np.float64)
        return Z, np.ascontiguousarray(np.array([a.xyz for a in self.atoms]).reshape(-1))


def nrm_s(a):
    """Normalisation of exp(-a r^2) (including the 1/sqrt(4 pi) angular factor)."""
    return (2 * a / math.pi) ** 0.75


def test_s_closed_forms():
    a, b, g = 0.8, 1.3, 0.6
    A = np.array([0.1, -0.2, 0.3])
    Bc = np.array([0.9, 0.4, -0.5])
    Cc = np.array([-0.7, 0.2, 1.1])
    mol = _Mol([_shell(0, a, A), _shell(0, b, Bc)], [_shell(0, g, Cc)],
               [desc.Atom(1, *Cc)])
    S, T, V = ints.int1e(mol)
    p = a + b
    mu = a * b / p
    R2 = float(((A - Bc) ** 2).sum())
    P = (a * A + b * Bc) / p
    Kab = math.exp(-mu * R2)
    N = nrm_s(a) * nrm_s(b)
    s_ref = N * (math.pi / p) ** 1.5 * Kab
    t_ref = N * mu * (3 - 2 * mu * R2) * (math.pi / p) ** 1.5 * Kab
    RPC2 = float(((P - Cc) ** 2).sum())
    v_ref = -N * 2 * math.pi / p * Kab * boys(0, p * RPC2)[0]
    assert abs(S[0, 1] - s_ref) < 1e-15
    assert abs(T[0, 1] - t_ref) < 1e-14
    assert abs(V[0, 1] - v_ref) < 1e-14
    # three-centre (c|ab) and two-centre (a|b)
    out = ints.int3c2e_shell(mol, 0, 1, 0)[0, 0, 0]
    q = g
    ref3 = N * nrm_s(g) * 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * boys(0, p * q / (p + q) * RPC2)[0]
    assert abs(out - ref3) < 1e-14 * abs(ref3)
    mol2 = _Mol([_shell(0, a, A)], [_shell(0, a, A), _shell(0, b, Bc)])
    M = ints.int2c2e(mol2)
    ref2 = nrm_s(a) * nrm_s(b) * 2 * math.pi ** 2.5 / (a * b * math.sqrt(a + b)) * boys(0, a * b / (a + b) * R2)[0]
    assert abs(M[0, 1] - ref2) < 1e-14 * abs(ref2)


# real solid harmonics in libcint order as polynomials (unnormalised), l <= 3
SHAPES = {
    1: [{(1, 0, 0): 1}, {(0, 1, 0): 1}, {(0, 0, 1): 1}],
    2: [{(1, 1, 0): 1}, {(0, 1, 1): 1}, {(0, 0, 2): 2, (2, 0, 0): -1, (0, 2, 0): -1},
        {(1, 0, 1): 1}, {(2, 0, 0): 1, (0, 2, 0): -1}],
    3: [{(2, 1, 0): 3, (0, 3, 0): -1}, {(1, 1, 1): 1}, {(0, 1, 2): 4, (2, 1, 0): -1, (0, 3, 0): -1},
        {(0, 0, 3): 2, (2, 0, 1): -3, (0, 2, 1): -3}, {(1, 0, 2): 4, (3, 0, 0): -1, (1, 2, 0): -1},
        {(2, 0, 1): 1, (0, 2, 1): -1}, {(3, 0, 0): 1, (1, 2, 0): -3}],
}


def _sphere_int(a, b, c):
    if a % 2 or b % 2 or c % 2:
        return 0.0
    return 2 * math.gamma((a + 1) / 2) * math.gamma((b + 1) / 2) * math.gamma((c + 1) / 2) / math.gamma((a + b + c + 3) / 2)


def _shape_norm(poly):
    s = 0.0
    for (e1, c1) in poly.items():
        for (e2, c2) in poly.items():
            s += c1 * c2 * _sphere_int(e1[0] + e2[0], e1[1] + e2[1], e1[2] + e2[2])
    return 1 / math.sqrt(s)


def _radial_norm(l, a):
    # 1/sqrt(int r^{2l+2} exp(-2 a r^2) dr)
    n1 = (2 * l + 3) / 2
    return 1 / math.sqrt(math.gamma(n1) / (2 * (2 * a) ** n1))


# Hermite-style derivative: monomial (x-Ax)^i exp(-a r^2) from derivatives of the s Gaussian
# d/dAx exp(-a|r-A|^2) = 2a (x-Ax) exp(...); second: (4a^2 (x-Ax)^2 - 2a) exp(...)
# third: (8a^3 (x-Ax)^3 - 12 a^2 (x-Ax)) exp(...)
def _mono_from_derivs(i):
    """Coefficients c_k with (x-A)^i = sum_k c_k d^k/dA^k (unit-amplitude Gaussian), per power of (2a)."""
    # returns list of (k, coefficient as function of a)
    if i == 0:
        return [(0, lambda a: 1.0)]
    if i == 1:
        return [(1, lambda a: 1 / (2 * a))]
    if i == 2:
        return [(2, lambda a: 1 / (4 * a * a)), (0, lambda a: 1 / (2 * a))]
    if i == 3:
        return [(3, lambda a: 1 / (8 * a ** 3)), (1, lambda a: 12 * a * a / (8 * a ** 3) / (2 * a))]
    raise ValueError


def _fd(f, A, order, h=2e-3):
    """Mixed central finite difference of f(A) with derivative orders (kx,ky,kz)."""
    # 1D stencils (error O(h^4) for 1st/2nd, O(h^2) for 3rd)
    st = {0: [(0, 1.0)],
          1: [(-2, 1 / 12), (-1, -2 / 3), (1, 2 / 3), (2, -1 / 12)],
          2: [(-2, -1 / 12), (-1, 4 / 3), (0, -5 / 2), (1, 4 / 3), (2, -1 / 12)],
          3: [(-2, -1 / 2), (-1, 1.0), (1, -1.0), (2, 1 / 2)]}
    tot = 0.0
    for sx, wx in st[order[0]]:
        for sy, wy in st[order[1]]:
            for sz, wz in st[order[2]]:
                d = np.array([sx, sy, sz]) * h
                tot += wx * wy * wz * f(A + d)
    return tot / h ** sum(order)


def _cart_from_s(f_of_A, A, a, mono):
    """Integral with (x-Ax)^i (y-Ay)^j (z-Az)^k exp(-a r_A^2) in place of the s Gaussian."""
    val = 0.0
    for kx, cx in _mono_from_derivs(mono[0]):
        for ky, cy in _mono_from_derivs(mono[1]):
            for kz, cz in _mono_from_derivs(mono[2]):
                val += cx(a) * cy(a) * cz(a) * _fd(f_of_A, A, (kx, ky, kz))
    return val


@pytest.mark.parametrize("l", [1, 2, 3])
def test_3c_aux_shell_by_derivatives(l):
    """(c_l | s s) with c a p/d/f aux function, against derivatives of (c_s|ss)."""
    a, b, g = 0.9, 0.5, 0.7
    A = np.array([0.2, -0.1, 0.4])
    Bc = np.array([-0.6, 0.5, 0.1])
    C = np.array([0.3, 0.8, -0.9])

    def s_int(Cpos):
        mol = _Mol([_shell(0, a, A), _shell(0, b, Bc)], [_shell(0, g, Cpos)])
        return ints.int3c2e_shell(mol, 0, 1, 0)[0, 0, 0] / nrm_s(g)   # unit-amplitude aux s

    mol = _Mol([_shell(0, a, A), _shell(0, b, Bc)], [_shell(l, g, C)])
    got = ints.int3c2e_shell(mol, 0, 1, 0)[:, 0, 0]
    for m, poly in enumerate(SHAPES[l]):
        ref = 0.0
        for mono, c in poly.items():
            ref += c * _cart_from_s(s_int, C, g, mono)
        ref *= _shape_norm(poly) * _radial_norm(l, g)
        assert abs(got[m] - ref) < 2e-6 * max(1.0, abs(ref)), (l, m, got[m], ref)


@pytest.mark.parametrize("la,lb", [(1, 0), (2, 0), (0, 2), (1, 1), (2, 1)])
def test_3c_ao_pair_by_derivatives(la, lb):
    """(s|ab) with p/d functions on the AO pair, against derivatives with respect to A and B."""
    a, b, g = 1.1, 0.6, 0.5
    A = np.array([0.2, -0.1, 0.4])
    Bc = np.array([-0.6, 0.5, 0.1])
    C = np.array([0.3, 0.8, -0.9])
    mol = _Mol([_shell(la, a, A), _shell(lb, b, Bc)], [_shell(0, g, C)])
    got = ints.int3c2e_shell(mol, 0, 1, 0)[0]
    polys_a = SHAPES[la] if la else [{(0, 0, 0): 1}]
    polys_b = SHAPES[lb] if lb else [{(0, 0, 0): 1}]

    def s_int(Apos, Bpos):
        m2 = _Mol([_shell(0, a, Apos), _shell(0, b, Bpos)], [_shell(0, g, C)])
        return ints.int3c2e_shell(m2, 0, 1, 0)[0, 0, 0] / (nrm_s(a) * nrm_s(b))

    for ia, pa in enumerate(polys_a):
        for ib, pb in enumerate(polys_b):
            ref = 0.0
            for ma, ca in pa.items():
                for mb, cb in pb.items():
                    # derivative w.r.t. A then B (independent centres)
                    f_b = lambda Bpos, ma=ma: _cart_from_s(lambda Ap: s_int(Ap, Bpos), A, a, ma)
                    ref += ca * cb * _cart_from_s(f_b, Bc, b, mb)
            nA = _shape_norm(pa) * _radial_norm(la, a) if la else nrm_s(a)
            nB = _shape_norm(pb) * _radial_norm(lb, b) if lb else nrm_s(b)
            ref *= nA * nB
            assert abs(got[ia, ib] - ref) < 5e-6 * max(1.0, abs(ref)), (la, lb, ia, ib, got[ia, ib], ref)


@pytest.mark.parametrize("l", [1, 2])
def test_1e_by_derivatives(l):
    """S, T, V with a p/d function on the bra against derivatives of s-s integrals."""
    a, b = 0.9, 0.6
    A = np.array([0.2, -0.1, 0.4])
    Bc = np.array([-0.6, 0.5, 0.1])
    Cn = np.array([0.1, 0.3, -0.2])

    def s_ints(Apos):
        m = _Mol([_shell(0, a, Apos), _shell(0, b, Bc)], atoms=[desc.Atom(1, *Cn)])
        S, T, V = ints.int1e(m)
        return np.array([S[0, 1], T[0, 1], V[0, 1]]) / (nrm_s(a) * nrm_s(b))

    mol = _Mol([_shell(l, a, A), _shell(0, b, Bc)], atoms=[desc.Atom(1, *Cn)])
    S, T, V = ints.int1e(mol)
    n = 2 * l + 1
    for m, poly in enumerate(SHAPES[l]):
        ref = np.zeros(3)
        for mono, c in poly.items():
            ref += c * _cart_from_s(s_ints, A, a, mono)
        ref *= _shape_norm(poly) * _radial_norm(l, a) * nrm_s(b)
        got = np.array([S[m, n], T[m, n], V[m, n]])
        assert np.abs(got - ref).max() < 2e-7 * max(1.0, np.abs(ref).max()), (l, m, got, ref)


@pytest.mark.parametrize("name", ["cc-pvdz", "cc-pvdz-jkfit", "cc-pvdz-ri"])
def test_basis_normalisation(name):
    atoms = [desc.Atom(8, 0, 0, 0), desc.Atom(1, 0, 0, 3.0)]
    shells = desc.read_basis(name, atoms)
    m = _Mol(shells, atoms=atoms)
    S, _, _ = ints.int1e(m)
    assert np.abs(np.diag(S) - 1).max() < 1e-13
    off = 0
    for s in shells:
        n = 2 * s.l + 1
        blk = S[off:off + n, off:off + n]
        assert np.abs(blk - np.eye(n)).max() < 1e-13
        off += n


def test_schwarz_bounds_every_block():
    """|(X|mn)| <= Q_X Q_mn for every element of B (Cauchy-Schwarz)."""
    mol = desc.Molecule(desc.water_cluster(2, 5))
    Qao, Qx = ints.schwarz(mol)
    Bh = ints.BTensorHost.build(mol, 0.0)
    Bd = Bh.dense()
    # expand shell factors to functions
    fa = np.concatenate([[i] * s.size for i, s in enumerate(mol.shells)])
    fx = np.concatenate([[i] * s.size for i, s in enumerate(mol.dfshells)])
    bound = Qx[fx][:, None, None] * Qao[np.ix_(fa, fa)][None]
    assert (np.abs(Bd) <= bound * (1 + 1e-12) + 1e-300).all()


def test_schwarz_mask_rule():
    """keep(X,M,N) iff float(|Q_MN| |Q_X|) > eps, compared bit-exactly."""
    nmn = np.array([[1e-5, 3.3e-6], [3.3e-6, 2.0]])
    nx = np.array([1e-5, 1.0000001e-5, 0.5])
    eps = 1e-10
    m = desc.schwarz_block_mask(nmn, nx, eps)
    for x in range(3):
        for i in range(2):
            for j in range(2):
                f = float(np.float32(nx[x] * nmn[i, j]))
                assert m[x, i, j] == (1 if f > eps else 0)
    # the float rounding matters at the boundary: 1e-5 * 1e-5 rounds above 1e-10
    assert float(np.float32(1e-10)) > 1e-10 and m[0, 0, 0] == 1


def test_library_schwarz_mask_matches_rule():
    """The library's block mask (mchip_schwarz_mask, the one the host and GPU
    tensor builds and the C++ layer use) is the float-cast rule applied to its
    block norms, and those norms are the Frobenius norms of the shell factors."""
    for n, seed in ((1, 0), (4, 2)):
        mol = desc.Molecule(desc.water_cluster(n, seed))
        mask, nmn, nx = ints.schwarz_mask(mol, 1e-10)
        assert np.array_equal(mask, desc.schwarz_block_mask(nmn, nx, 1e-10))
        Qao, Qx = ints.schwarz(mol)
        assert np.allclose(nmn, desc.block_norms(Qao, mol.blocks), rtol=1e-14, atol=0)
        assert np.allclose(nx, desc.block_norms(Qx, mol.dfblocks), rtol=1e-14, atol=0)
        assert mask.sum() > 0


def test_4c_s_closed_form():
    """(ab|cd) over four s Gaussians: N 2 pi^(5/2) / (p q sqrt(p+q)) K_ab K_cd F_0(rho R_PQ^2)."""
    ex = [0.8, 1.3, 0.6, 2.1]
    cs = [np.array([0.1, -0.2, 0.3]), np.array([0.9, 0.4, -0.5]), np.array([-0.7, 0.2, 1.1]),
          np.array([0.4, -0.8, -0.3])]
    mol = _Mol([_shell(0, a, c) for a, c in zip(ex, cs)])
    g = ints.int4c2e(mol)
    a, b, c, d = ex
    p, q = a + b, c + d
    P = (a * cs[0] + b * cs[1]) / p
    Q = (c * cs[2] + d * cs[3]) / q
    Kab = math.exp(-a * b / p * float(((cs[0] - cs[1]) ** 2).sum()))
    Kcd = math.exp(-c * d / q * float(((cs[2] - cs[3]) ** 2).sum()))
    N = nrm_s(a) * nrm_s(b) * nrm_s(c) * nrm_s(d)
    ref = N * 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kcd * \
        boys(0, p * q / (p + q) * float(((P - Q) ** 2).sum()))[0]
    assert abs(g[0, 1, 2, 3] - ref) < 1e-14 * abs(ref)


def test_4c_permutational_symmetry():
    """(ij|kl) = (ji|kl) = (ij|lk) = (kl|ij) over mixed s/p/d/f shells, and
    (ij|ij) >= 0."""
    sh = [_shell(2, 0.9, [0.2, -0.1, 0.4]), _shell(1, 0.5, [-0.6, 0.5, 0.1]), _shell(0, 1.4, [0.3, 0.8, -0.9]),
          _shell(3, 0.7, [0.0, 0.1, 0.2]), _shell(1, 2.2, [0.2, -0.1, 0.4])]
    g = ints.int4c2e(_Mol(sh))
    tol = 1e-14 * np.abs(g).max()
    assert np.abs(g - g.transpose(1, 0, 2, 3)).max() < tol
    assert np.abs(g - g.transpose(0, 1, 3, 2)).max() < tol
    assert np.abs(g - g.transpose(2, 3, 0, 1)).max() < tol
    n = g.shape[0]
    d = np.einsum("ijij->ij", g)
    assert (d >= -tol).all()
    # the pair matrix (ij|kl) is positive semidefinite
    w = np.linalg.eigvalsh(g.reshape(n * n, n * n))
    assert w.min() > -1e-12 * w.max()


@pytest.mark.parametrize("la,lb,lc", [(0, 0, 0), (1, 0, 0), (2, 1, 0), (0, 0, 2), (1, 1, 1), (2, 2, 3), (3, 0, 1)])
def test_4c_against_3c_limit(la, lb, lc):
    """(ab|c d) with d an s Gaussian of vanishing exponent on the centre of c
    tends to N_d (ab|c), the three-centre integral with c as auxiliary
    function: the four-centre engine against the separately pinned
    three-centre one, for every l on the bra and on the ket."""
    A = [0.2, -0.1, 0.4]
    Bc = [-0.6, 0.5, 0.1]
    C = [0.3, 0.8, -0.9]
    eta = 1e-10
    a, b, g = 1.1, 0.6, 0.5
    mol4 = _Mol([_shell(la, a, A), _shell(lb, b, Bc), _shell(lc, g, C), _shell(0, eta, C)])
    g4 = ints.int4c2e(mol4)
    na, nb, nc = 2 * la + 1, 2 * lb + 1, 2 * lc + 1
    got = g4[:na, na:na + nb, na + nb:na + nb + nc, -1] / nrm_s(eta)
    mol3 = _Mol([_shell(la, a, A), _shell(lb, b, Bc)], [_shell(lc, g, C)])
    ref = ints.int3c2e_shell(mol3, 0, 1, 0)       # [c][a][b]
    ref = ref.transpose(1, 2, 0)
    assert np.abs(got - ref).max() < 1e-8 * max(1.0, np.abs(ref).max()), (la, lb, lc)
