"""CPU tests: the C ABI library loads and exports every declared symbol; basic
integral identities; oracle self-consistency on water."""
import os
import re

import numpy as np

from chem_amd import _lib, desc, ints

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    text = open(os.path.join(ROOT, "include", "mchip.h")).read()
    names = set(re.findall(r"\b(mchip_[a-z0-9_]+)\s*\(", text))
    assert names, "no declarations found"
    for n in sorted(names):
        assert hasattr(L, n), n
    # every binding in _lib is declared in the header
    assert set(_lib.exported_symbols()) <= names


def test_water_geometry_and_dims():
    atoms = desc.water_cluster(1, 0)
    mol = desc.Molecule(atoms)
    assert (mol.nbf, mol.naux, mol.nocc) == (24, 116, 5)
    # RCM (on by default, as in the reference): all three atoms are within 5
    # bohr of each other with degree 2; the farthest pair is H1-H2, the start
    # H1, then O and H2 in index order: R = [1, 0, 2], reversed H2, O, H1
    assert mol.b == [5, 14, 5] and mol.x == [23, 70, 23]
    assert desc.Molecule(atoms, reorder=False).b == [14, 5, 5]
    d = np.linalg.norm(atoms[1].xyz - atoms[0].xyz) * desc.BOHR_RADIUS
    assert abs(d - 0.9572) < 1e-12
    v1, v2 = atoms[1].xyz - atoms[0].xyz, atoms[2].xyz - atoms[0].xyz
    ang = np.degrees(np.arccos(v1 @ v2 / np.linalg.norm(v1) / np.linalg.norm(v2)))
    assert abs(ang - 104.52) < 1e-10


def test_water_cluster_min_distance():
    atoms = desc.water_cluster(8, 3)
    O = np.array([a.xyz for a in atoms if a.Z == 8]) * desc.BOHR_RADIUS
    d = np.sqrt(((O[:, None] - O[None]) ** 2).sum(-1)) + np.eye(len(O)) * 99
    assert d.min() >= 2.8
    mol = desc.Molecule(atoms)
    assert mol.nbf == 8 * 24 and mol.naux == 8 * 116 and mol.nocc == 40


def test_overlap_normalised_and_metric_spd():
    mol = desc.Molecule(desc.water_cluster(2, 1))
    S, T, V = ints.int1e(mol)
    assert np.abs(np.diag(S) - 1).max() < 1e-13
    assert np.array_equal(S, S.T) and np.array_equal(T, T.T)
    M = ints.int2c2e(mol)
    assert np.linalg.eigvalsh(M).min() > 0
    Bh = ints.BTensorHost.build(mol, 1e-10)
    Bd = Bh.dense()
    assert np.array_equal(Bd, Bd.transpose(0, 2, 1))


def test_cpp_builder_header(tmp_path):
    """include/mchip_fock.hpp compiles as C++17 against the C ABI, links with
    libmchip.so, and reports a missing required factory argument (the
    reference's CHECK_REQUIRED) as std::runtime_error naming it, without a
    device call; unsupported methods and metrics are refused the same way."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "b.cpp"
    src.write_text(r'''
#include <cstdio>
#include <cstring>
#include "mchip_fock.hpp"
static int expect(const char* what, const std::runtime_error& e) { return std::strstr(e.what(), what) ? 0 : 1; }
int main() {
  using namespace mchip;
  int bad = 0;
  try { fock::create_k().method(fock::kmethod::dfmo).occ_nbatches(2).build(); return 2; }
  catch (const std::runtime_error& e) { bad += expect("set_world", e); }
  try { fock::create_j().build(); return 3; }
  catch (const std::runtime_error& e) { bad += expect("method", e); }
  try { mp::create_z().method(mp::zmethod::ll_full).build(); return 4; }
  catch (const std::runtime_error& e) { bad += expect("set_world", e); }
  try { ints::aoloader::create().build(); return 5; }
  catch (const std::runtime_error& e) { bad += expect("set_world", e); }
  try { fock::str_to_kmethod("nope"); return 6; }
  catch (const std::runtime_error& e) { bad += expect("Invalid kmethod", e); }
  if (fock::str_to_kmethod("dflmo") != fock::kmethod::dflmo || fock::str_to_jmethod("dfao") != fock::jmethod::dfao ||
      mp::str_to_zmethod("llmp_mem") != mp::zmethod::llmp_mem) return 7;
  if (desc::split_range(12, 5) != std::vector<int>({5, 5, 2})) return 8;
  if (bad) return 9;
  std::printf("ok\n");
  return 0;
}
''')
    exe = tmp_path / "b"
    lib = os.path.join(root, "chem_amd", "lib")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(root, "include"),
                        str(src), "-L", lib, "-lmchip", "-Wl,-rpath," + lib, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says so,
    before touching the library or a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_bench_help_lists_workloads():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 0
    for c in ("c0", "c1", "c2", "c3", "c4"):
        assert c in r.stdout
    for opt in ("--steps", "--warmup", "--form", "--escalate", "--stall"):
        assert opt in r.stdout


def test_split_range_partitions():
    """split_range (dbcsr_common.hpp:216-230): full blocks of `split` and one
    remainder block, summing to n; brute force over small n and split."""
    for n in range(0, 60):
        for split in range(1, 20):
            r = desc.split_range(n, split)
            assert sum(r) == n
            assert len(r) == -(-n // split)
            assert all(x == split for x in r[:-1])
            if r:
                assert 0 < r[-1] <= split


def test_make_blk_bounds_invariants():
    """make_blk_bounds (dbcsr_btensor.hpp:35-95) on random block sizes and
    centre maps: batches are consecutive from block 0, never split a centre,
    number at most min(nbatches, nblocks), and cover every block whenever the
    reference's floating-point test can close batch min(nbatches, nblocks)."""
    rng = np.random.default_rng(5)
    dropped = 0
    for _ in range(3000):
        use_map = rng.random() < 0.5
        if use_map:
            ncen = int(rng.integers(1, 12))
            blkmap = sorted(int(c) for c in rng.integers(0, ncen, size=int(rng.integers(ncen, 3 * ncen + 1))))
            nblk = len(blkmap)
        else:
            blkmap = None
            nblk = int(rng.integers(1, 30))
        bs = [int(x) for x in rng.integers(1, 15, size=nblk)]
        nb = int(rng.integers(1, 12))
        out = desc.make_blk_bounds(bs, nb, blkmap)
        assert out[0][0] == 0
        for (a0, a1), (b0, _) in zip(out, out[1:]):
            assert a0 <= a1 and b0 == a1 + 1
        assert len(out) <= min(nb, nblk)
        if blkmap is not None:
            assert all(e == nblk - 1 or blkmap[e] != blkmap[e + 1] for _, e in out)
        k = min(nb, nblk)
        closes = float(sum(bs)) >= k * (sum(bs) / k)
        covers = out[-1][1] == nblk - 1
        # a run that ends below batch k closes at the last block anyway; only
        # a threshold k * (nele / k) rounded above nele can leave blocks out
        if closes:
            assert covers
        dropped += not covers
    assert dropped > 0   # the rounding case occurs in this sample


def test_make_blk_bounds_known_cases():
    """Hand-checked splits: equal blocks into equal batches; a centre map that
    forces a batch to run to the centre's end; 211 elements in 11 batches,
    where 11 * (211 / 11) rounds to 211.00000000000003 and the last batch
    is never closed (the reference's rule, kept)."""
    assert desc.make_blk_bounds([2] * 6, 3) == [[0, 1], [2, 3], [4, 5]]
    assert desc.make_blk_bounds([1, 1, 1, 1], 2, [0, 0, 0, 1]) == [[0, 2], [3, 3]]
    assert desc.make_blk_bounds([5, 5], 4) == [[0, 0], [1, 1]]
    assert 11 * (211 / 11) > 211
    bs = [10] * 20 + [11]
    out = desc.make_blk_bounds(bs, 11)
    assert out[-1][1] < len(bs) - 1


def test_product_code_does_not_use_the_oracle():
    """The oracle is the checker only: nothing under chem_amd/ or include/
    imports, links or names the oracle package, and bench.py touches it only
    inside its CPU-baseline function."""
    pat = re.compile(r"^\s*(from\s+oracle\b|import\s+oracle\b)|[\"']oracle[/.]", re.M)
    for top in ("chem_amd", "include"):
        for dirpath, _, files in os.walk(os.path.join(ROOT, top)):
            for f in files:
                if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h", ".c")):
                    text = open(os.path.join(dirpath, f), encoding="utf-8", errors="replace").read()
                    assert not pat.search(text), os.path.join(dirpath, f)
    src = open(os.path.join(ROOT, "bench.py")).read()
    body = src.split("\ndef cpu_baseline(", 1)
    assert len(body) == 2
    outside = body[0] + "\n".join(l for l in body[1].split("\ndef ", 1)[1:])
    assert not re.search(r"^\s*(from\s+oracle\b|import\s+oracle\b)", outside, re.M)


def _atoms_at(xs):
    from chem_amd.desc import Atom
    return [Atom(1, x, 0.0, 0.0) for x in xs]


def test_rcm_order_hand_checked_chain():
    """Four atoms on a line 3 bohr apart (5-bohr cutoff: nearest neighbours
    only), given in the scrambled order x = 6, 0, 9, 3.  By hand
    (rcm.hpp:137-210): the farthest pair is (1, 2) (x = 0 and 9), both of
    degree 1, so the start is node 1; its neighbour 3; then 3's neighbours by
    decreasing degree [0 (deg 2), 1]; then 0's [3, 2]; R = [1, 3, 0, 2],
    reversed [2, 0, 3, 1]: the atoms come out sorted by decreasing x."""
    from chem_amd.desc import rcm_order
    atoms = _atoms_at([6.0, 0.0, 9.0, 3.0])
    assert rcm_order(atoms) == [2, 0, 3, 1]


def test_rcm_order_ties_keep_index_order():
    """A centre with four arms (4 bohr from the centre, 5.7 bohr or more from
    each other) and a tail atom linked to arm 4 only: three arms tie in degree
    and enter the queue in index order (std::sort on at most 16 entries keeps
    ties), so the order is fixed by the tie rule."""
    from chem_amd.desc import Atom, rcm_order
    atoms = [Atom(1, 0.0, 0.0, 0.0),          # 0 centre
             Atom(1, 4.0, 0.0, 0.0),          # 1 arm +x
             Atom(1, -4.0, 0.0, 0.0),         # 2 arm -x
             Atom(1, 0.0, 4.0, 0.0),          # 3 arm +y
             Atom(1, 0.0, -4.0, 0.0),         # 4 arm -y
             Atom(1, 0.0, -8.5, 0.0)]         # 5 tail of arm 4
    # degrees: centre 4, arms 1 except arm 4 (2), tail 1.  Farthest pair not
    # placed: (3, 5) at 12.5 bohr, both degree 1 -> start at 3; its neighbour 0;
    # 0's neighbours by decreasing degree: 4 (deg 2), then 1, 2, 3 in index
    # order; 4's neighbours [0, 5]: R = [3, 0, 4, 1, 2, 5], reversed below.
    assert rcm_order(atoms) == [5, 2, 1, 4, 0, 3]


def test_rcm_reduces_water_cluster_bandwidth():
    """On seeded water clusters the reordered 5-bohr connectivity has a
    bandwidth no larger than in the generator's order, and the order is a
    permutation that the Molecule applies by default (the reference's
    "reorder": true)."""
    from chem_amd import desc
    for n, seed in ((8, 1), (16, 2), (32, 3)):
        atoms = desc.water_cluster(n, seed)
        order = desc.rcm_order(atoms)
        assert sorted(order) == list(range(len(atoms)))

        def bandwidth(idx):
            xyz = np.array([atoms[i].xyz for i in idx])
            d = np.sqrt(((xyz[:, None] - xyz[None]) ** 2).sum(-1))
            i, j = np.nonzero(d < 5.0)
            return int(np.abs(i - j).max())
        assert bandwidth(order) <= bandwidth(range(len(atoms)))
        mol = desc.Molecule(atoms)
        assert [a.xyz.tolist() for a in mol.atoms] == [atoms[i].xyz.tolist() for i in order]
