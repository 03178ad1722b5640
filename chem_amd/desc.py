"""Descriptors: atoms, basis sets, atomic blocking, water clusters, screening.

Reference counterparts (the input side of the DF Fock path):
  * basis reader            src/desc/basis.cpp:278-360 (BSE JSON, drop |c|<=eps)
  * atomic clustering       src/desc/basis.cpp:52-80 (one block per centre)
  * block_to_atom           src/desc/basis.cpp:709-738
  * occupied/virtual blocks src/desc/molecule.hpp:53-193 (split_range)
  * batch bounds            src/dbcsrx/dbcsr_btensor.hpp:35-95 (make_blk_bounds)
  * RCM atom reordering     src/math/other/rcm.hpp:20-230, driver :380-399
  * Schwarz block screening src/ints/screening.cpp:8-30
  * bohr conversion         src/utils/constants.hpp:4
"""
import ctypes
import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib

BOHR_RADIUS = 0.529177210671212
DBL_EPSILON = np.finfo(np.float64).eps
BASIS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")

SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne"]


def atomic_number(sym):
    return SYMBOLS.index(sym.capitalize())


@dataclass
class Atom:
    Z: int
    x: float
    y: float
    z: float

    @property
    def xyz(self):
        return np.array([self.x, self.y, self.z])


@dataclass
class Shell:
    l: int
    center: np.ndarray
    atom: int
    alpha: np.ndarray
    coeff: np.ndarray

    @property
    def size(self):
        return 2 * self.l + 1


def load_basis_json(name):
    path = name if os.path.exists(name) else os.path.join(BASIS_DIR, name.lower() + ".json")
    if not os.path.exists(path):
        raise RuntimeError("Could not find basis " + name)
    with open(path) as f:
        return json.load(f)


def read_basis(name, atoms):
    """Shells for every atom in order (basis.cpp:278-360).

    A single-l entry yields one shell per coefficient column, keeping only the
    primitives with |c| > DBL_EPSILON; a multi-l (sp) entry yields one shell
    per angular momentum with all primitives."""
    data = load_basis_json(name)
    elements = data["elements"]
    shells = []
    for ia, atom in enumerate(atoms):
        entry = elements[str(atom.Z)]["electron_shells"]
        for sh in entry:
            alpha = np.array([float(s) for s in sh["exponents"]])
            ams = sh["angular_momentum"]
            coefs = sh["coefficients"]
            if len(ams) == 1:
                for col in coefs:
                    c = np.array([float(s) for s in col])
                    keep = np.abs(c) > DBL_EPSILON
                    shells.append(Shell(ams[0], atom.xyz.copy(), ia, alpha[keep].copy(), c[keep].copy()))
            else:
                for l, col in zip(ams, coefs):
                    c = np.array([float(s) for s in col])
                    shells.append(Shell(l, atom.xyz.copy(), ia, alpha.copy(), c))
    return shells


class BasisHandle:
    """Owner of an mchip_basis (normalised shells on the C++ side)."""

    def __init__(self, shells):
        L = _lib.lib()
        self.shells = shells
        n = len(shells)
        ls = np.array([s.l for s in shells], dtype=np.int32)
        npr = np.array([len(s.alpha) for s in shells], dtype=np.int32)
        ex = np.ascontiguousarray(np.concatenate([s.alpha for s in shells]), dtype=np.float64)
        cf = np.ascontiguousarray(np.concatenate([s.coeff for s in shells]), dtype=np.float64)
        ce = np.ascontiguousarray(np.array([s.center for s in shells]).reshape(-1), dtype=np.float64)
        at = np.array([s.atom for s in shells], dtype=np.int32)
        h = ctypes.c_void_p()
        _lib.check(L.mchip_basis_create(n, _lib.iptr(ls), _lib.iptr(npr), _lib.dptr(ex),
                                        _lib.dptr(cf), _lib.dptr(ce), _lib.iptr(at),
                                        ctypes.byref(h)))
        self.h = h
        nbf = ctypes.c_int()
        _lib.check(L.mchip_basis_nbf(h, ctypes.byref(nbf)))
        self.nbf = nbf.value
        self._nprim_total = int(npr.sum())

    def norm_coefs(self):
        out = np.zeros(self._nprim_total)
        _lib.check(_lib.lib().mchip_basis_norm_coefs(self.h, _lib.dptr(out)))
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.lib().mchip_basis_destroy(self.h)
                self.h = None
        except Exception:
            pass


def cluster_atomic(shells):
    """Block = all shells on one centre, blocks in first-appearance order
    (basis.cpp:52-80).  Returns list of shell-index lists."""
    blocks = []
    seen = {}
    for i, s in enumerate(shells):
        key = tuple(np.round(s.center, 14))
        if key not in seen:
            seen[key] = len(blocks)
            blocks.append([])
        blocks[seen[key]].append(i)
    return blocks


def split_range(n, split):
    """dbcsr_common.hpp:216-230."""
    if n <= 0:
        return []
    nblock = n // split if n % split == 0 else n // split + 1
    out = [split] * nblock
    if n % split != 0:
        out[-1] = n % split
    return out


def make_blk_bounds(blksizes, nbatches, blkmap=None):
    """Batch bounds in block indices, split at centre boundaries
    (dbcsr_btensor.hpp:35-95)."""
    nblks = len(blksizes)
    nele = sum(blksizes)
    if nblks < nbatches:
        nbatches = nblks
    per = nele / nbatches
    if blkmap is not None and len(blkmap) != nblks:
        raise RuntimeError("Block sizes and block map do not have the same dimension")
    out = []
    cur = 0
    ibatch = 1
    first = 0
    for i in range(nblks):
        cur += blksizes[i]
        cc = blkmap[i] if blkmap is not None else i
        if i != nblks - 1:
            nc = blkmap[i + 1] if blkmap is not None else i + 1
        else:
            nc = -1 if blkmap is not None else i + 1
        if cur >= ibatch * per and nc != cc:
            out.append([first, i])
            first = i + 1
            ibatch += 1
    if not out:
        raise RuntimeError("Block bounds are zero.")
    return out


def rcm_order(atoms, cutoff=5.0):
    """Reverse Cuthill-McKee ordering of atoms (rcm.hpp:20-230): connectivity by
    distance < cutoff (bohr); start from the lower-degree end of the farthest
    pair not yet placed; neighbours queued by decreasing degree.  The
    reference sorts each neighbour list with std::sort (rcm.hpp:121), which for
    lists of at most 16 entries (libstdc++'s insertion-sort threshold; a water
    cluster's 5-bohr lists are far shorter) keeps equal degrees in index
    order: the stable sort here gives the same order for such lists.  Returns
    the new order as indices into atoms (rcm::reordered_idx)."""
    n = len(atoms)
    xyz = np.array([a.xyz for a in atoms])
    dist = np.sqrt(((xyz[:, None, :] - xyz[None, :, :]) ** 2).sum(-1))
    con = (np.abs(dist) < cutoff).astype(int)
    deg = con.sum(1) - 1

    def min_degree(R):
        inR = set(R)
        maxd = 0.0
        n1 = n2 = 0
        for i in range(n):
            for j in range(i + 1, n):
                if dist[i, j] > maxd:
                    if i in inR and j in inR:
                        continue
                    maxd = dist[i, j]
                    n1, n2 = i, j
        if (deg[n1] > deg[n2] or n1 in inR) and n2 not in inR:
            return n2
        return n1

    def neighbours(P):
        q = [i for i in range(n) if con[P, i] == 1 and i != P]
        q.sort(key=lambda i: -deg[i])  # stable, decreasing degree
        return q

    R = []
    Rset = set()
    from collections import deque
    Q = deque()
    while len(R) != n:
        P = min_degree(R)
        R.append(P)
        Rset.add(P)
        Q.extend(neighbours(P))
        while Q:
            q0 = Q[0]
            if q0 in Rset:
                Q.popleft()
                continue
            R.append(q0)
            Rset.add(q0)
            Q.extend(neighbours(q0))
            Q.popleft()
        if len(R) == n:
            break
    R.reverse()
    return R


# ---------------------------------------------------------------------------
# water clusters
# ---------------------------------------------------------------------------
class MT19937_64:
    """64-bit Mersenne Twister (the std::mt19937_64 recurrence and tempering)."""
    NN, MM = 312, 156
    MATRIX_A = 0xB5026F5AA96619E9
    UM, LM = 0xFFFFFFFF80000000, 0x7FFFFFFF
    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.mt = [0] * self.NN
        self.mt[0] = seed & self.MASK
        for i in range(1, self.NN):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & self.MASK
        self.mti = self.NN

    def next_u64(self):
        if self.mti >= self.NN:
            mt = self.mt
            for i in range(self.NN):
                x = (mt[i] & self.UM) | (mt[(i + 1) % self.NN] & self.LM)
                xa = x >> 1
                if x & 1:
                    xa ^= self.MATRIX_A
                mt[i] = mt[(i + self.MM) % self.NN] ^ xa
            self.mti = 0
        x = self.mt[self.mti]
        self.mti += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & self.MASK

    def uniform(self):
        """Uniform double in [0,1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self):
        """Standard normal by Box-Muller (one value per call)."""
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)


WATER_ROH = 0.9572      # angstrom
WATER_HOH = 104.52      # degrees
WATER_VOLUME = 29.9     # angstrom^3 per molecule (liquid density)


def water_monomer():
    """O at origin, H at (+-r sin(a/2), 0, r cos(a/2)) angstrom."""
    half = math.radians(WATER_HOH / 2.0)
    h1 = (WATER_ROH * math.sin(half), 0.0, WATER_ROH * math.cos(half))
    h2 = (-WATER_ROH * math.sin(half), 0.0, WATER_ROH * math.cos(half))
    return [(8, (0.0, 0.0, 0.0)), (1, h1), (1, h2)]


def _quat_rot(q):
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(n, seed, min_oo=2.8, box=None, max_tries=1000000):
    """n waters, uniform random orientation (normalised Gaussian quaternion) and
    uniform centre in a cube of edge (n*29.9)^(1/3) angstrom, rejecting any
    candidate with an O-O distance < min_oo.  Returns atoms in bohr."""
    if box is None:
        box = (n * WATER_VOLUME) ** (1.0 / 3.0)
    rng = MT19937_64(seed)
    mono = water_monomer()
    oxy = []
    atoms = []
    tries = 0
    while len(oxy) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("water_cluster: rejection sampling did not finish")
        q = np.array([rng.normal() for _ in range(4)])
        q /= np.linalg.norm(q)
        Rm = _quat_rot(q)
        c = np.array([rng.uniform() * box for _ in range(3)])
        if n == 1:
            c = np.zeros(3)
            Rm = np.eye(3)
        if any(np.linalg.norm(c - o) < min_oo for o in oxy):
            continue
        oxy.append(c)
        for Z, xyz in mono:
            p = c + Rm @ np.array(xyz)
            atoms.append(Atom(Z, *(p / BOHR_RADIUS)))
    return atoms


# ---------------------------------------------------------------------------
# molecule
# ---------------------------------------------------------------------------
class Molecule:
    """Atoms + orbital/aux basis with atomic blocks (molecule.hpp:23-318).

    reorder (default on, as the reference's "reorder" key,
    megalochem_driver.cpp:380-399): atoms in reverse Cuthill-McKee order of
    their 5-bohr connectivity (rcm_order) before the basis is built, which fixes
    the block order of every tensor."""

    def __init__(self, atoms, basis="cc-pvdz", dfbasis="cc-pvdz-jkfit", charge=0,
                 mult=1, mo_split=5, reorder=True, name="mol"):
        if reorder:
            order = rcm_order(atoms)
            atoms = [atoms[i] for i in order]
        self.atoms = atoms
        self.name = name
        self.charge = charge
        self.mult = mult
        self.mo_split = mo_split
        self.basis_name = basis
        self.dfbasis_name = dfbasis
        self.shells = read_basis(basis, atoms)
        self.dfshells = read_basis(dfbasis, atoms) if dfbasis else None
        self.blocks = cluster_atomic(self.shells)
        self.dfblocks = cluster_atomic(self.dfshells) if dfbasis else None
        self.b = [sum(self.shells[i].size for i in blk) for blk in self.blocks]
        self.x = [sum(self.dfshells[i].size for i in blk) for blk in self.dfblocks] if dfbasis else None
        self.nbf = sum(self.b)
        self.naux = sum(self.x) if dfbasis else 0
        nele = sum(a.Z for a in atoms) - charge
        if (nele + mult - 1) % 2:
            raise RuntimeError("charge/multiplicity inconsistent")
        self.nocc_alpha = (nele + mult - 1) // 2
        self.nocc_beta = (nele - mult + 1) // 2
        self.nocc = self.nocc_alpha
        self.nvir = self.nbf - self.nocc
        self._bh = None
        self._xh = None

    # block dims (molecule.hpp:53-193)
    def oa(self):
        return split_range(self.nocc_alpha, self.mo_split)

    def va(self):
        return split_range(self.nbf - self.nocc_alpha, self.mo_split)

    def ma(self):
        return self.oa() + self.va()

    def block_to_atom(self, which="b"):
        blocks = self.blocks if which == "b" else self.dfblocks
        shells = self.shells if which == "b" else self.dfshells
        return [shells[blk[0]].atom for blk in blocks]

    @property
    def basis_handle(self):
        if self._bh is None:
            self._bh = BasisHandle(self.shells)
        return self._bh

    @property
    def dfbasis_handle(self):
        if self._xh is None:
            self._xh = BasisHandle(self.dfshells)
        return self._xh

    def nuclear_repulsion(self):
        e = 0.0
        for i in range(len(self.atoms)):
            for j in range(i + 1, len(self.atoms)):
                a, b = self.atoms[i], self.atoms[j]
                e += a.Z * b.Z / np.linalg.norm(a.xyz - b.xyz)
        return e

    def charges(self):
        Z = np.array([a.Z for a in self.atoms], dtype=np.float64)
        xyz = np.ascontiguousarray(np.array([a.xyz for a in self.atoms]).reshape(-1))
        return Z, xyz


def block_norms(Q, blocks):
    """Frobenius block norms of a shell-level matrix/vector (dbcsr block_norms)."""
    nb = len(blocks)
    if Q.ndim == 1:
        return np.array([math.sqrt(sum(Q[i] ** 2 for i in blk)) for blk in blocks])
    out = np.zeros((nb, nb))
    for i, bi in enumerate(blocks):
        for j, bj in enumerate(blocks):
            out[i, j] = math.sqrt(float((Q[np.ix_(bi, bj)] ** 2).sum()))
    return out


def schwarz_block_mask(norm_mn, norm_x, eps):
    """keep(X,M,N) iff (float)(|Q_MN| * |Q_X|) > eps  (screening.cpp:23-30).

    The product is formed in double, rounded to float, compared in double."""
    prod = norm_x[:, None, None] * norm_mn[None, :, :]
    f = prod.astype(np.float32).astype(np.float64)
    return (f > eps).astype(np.uint8)
