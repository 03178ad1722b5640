"""HDF5 wavefunction files (chem_amd.wfn, libmchip_wfn.so) against the layout
the reference writes: filio::data_handler (src/io/data_handler.hpp), write_molecule
(src/desc/molecule.hpp:322-352), write_hfwfn / read_hfwfn
(src/desc/wfn.hpp:188-326).

The layout is checked with h5dump, the HDF5 project's own reader, so the pin
does not rest on this package's reader: group per tag, float64 little-endian
row-major matrices nbf x nocc / nbf x nvir, vectors, 1-element energies,
int32 scalars and a 128-byte string.  Values round-trip bit for bit; the
reference's error messages for missing datasets and wrong dimensions are kept.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from chem_amd import desc, wfn

H5DUMP = "/opt/conda/bin/h5dump"

pytestmark = pytest.mark.skipif(not os.path.exists(wfn.WFN_LIB), reason="libmchip_wfn.so not built (no HDF5)")


def _wfn(nbf=24, nocc=5, seed=0, beta=False):
    rng = np.random.default_rng(seed)
    w = wfn.HfWavefunction(rng.standard_normal((nbf, nocc)), rng.standard_normal((nbf, nbf - nocc)),
                           np.sort(rng.standard_normal(nocc)), np.sort(rng.standard_normal(nbf - nocc)),
                           -85.1234567891234, 9.1876543219876, -85.1234567891234 + 9.1876543219876)
    if beta:
        w.c_bo_B = rng.standard_normal((nbf, nocc - 1))
        w.c_bv_B = rng.standard_normal((nbf, nbf - nocc + 1))
        w.eps_occ_B = rng.standard_normal(nocc - 1)
        w.eps_vir_B = rng.standard_normal(nbf - nocc + 1)
    return w


def _dump(path):
    r = subprocess.run([H5DUMP, "-H", str(path)], capture_output=True, text=True, check=True)
    out = {}
    for m in re.finditer(r'DATASET "([^"]+)" \{\s*DATATYPE\s+(.*?)\s*DATASPACE\s+SIMPLE \{ \( ([^)]*) \)', r.stdout, re.S):
        out[m.group(1)] = (" ".join(m.group(2).split()), tuple(int(x) for x in m.group(3).split(",")))
    groups = re.findall(r'GROUP "([^"]+)"', r.stdout)
    return out, groups


def test_round_trip_bit_exact(tmp_path):
    f = tmp_path / "out.h5"
    wfn.create(f)
    w = _wfn()
    wfn.write_hfwfn(f, "hf", w)
    r = wfn.read_hfwfn(f, "hf", 24, 5)
    for k in ("c_bo_A", "c_bv_A", "eps_occ_A", "eps_vir_A"):
        assert np.array_equal(getattr(r, k), getattr(w, k)), k
    assert (r.scf_energy, r.nuc_energy, r.wfn_energy) == (w.scf_energy, w.nuc_energy, w.wfn_energy)
    assert r.c_bo_B is None and r.eps_vir_B is None


def test_unrestricted_parts_optional(tmp_path):
    f = tmp_path / "out.h5"
    wfn.create(f)
    w = _wfn(beta=True)
    wfn.write_hfwfn(f, "uhf", w)
    r = wfn.read_hfwfn(f, "uhf", 24, 5, 4)
    for k in ("c_bo_B", "c_bv_B", "eps_occ_B", "eps_vir_B"):
        assert np.array_equal(getattr(r, k), getattr(w, k)), k


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump absent")
def test_layout_matches_reference_writer(tmp_path):
    f = tmp_path / "out.h5"
    wfn.create(f)
    mol = desc.Molecule(desc.water_cluster(2, 3))
    wfn.write_molecule(f, "mol", mol, name="water dimer")
    w = _wfn(nbf=mol.nbf, nocc=mol.nocc)
    wfn.write_hfwfn(f, "wfn", w)
    ds, groups = _dump(f)
    assert set(groups) >= {"/", "mol", "wfn"}
    f64 = "H5T_IEEE_F64LE"
    n, o = mol.nbf, mol.nocc
    expect = {
        "c_bo_A": (f64, (n, o)), "c_bv_A": (f64, (n, n - o)), "eps_occ_A": (f64, (o,)), "eps_vir_A": (f64, (n - o,)),
        "scf_energy": (f64, (1,)), "nuc_energy": (f64, (1,)), "wfn_energy": (f64, (1,)),
        "charge": ("H5T_STD_I32LE", (1,)), "mult": ("H5T_STD_I32LE", (1,)), "mo_split": ("H5T_STD_I32LE", (1,)),
        "atoms": (f64, (len(mol.atoms), 4)),
    }
    for k, v in expect.items():
        assert ds[k] == v, (k, ds.get(k))
    assert ds["name"][1] == (1,) and "STRSIZE 128" in ds["name"][0]
    assert "c_bo_B" not in ds          # restricted: no beta datasets
    # values through h5dump: the atoms table and the charge
    r = subprocess.run([H5DUMP, "-d", "/mol/atoms", "-m", "%.17g", str(f)], capture_output=True, text=True, check=True)
    body = r.stdout.split("DATA {", 1)[1]
    vals = [float(x) for x in re.findall(r"-?\d+\.?\d*(?:[eE][-+]?\d+)?", re.sub(r"\(\d+,\d+\):", "", body))]
    Z, xyz = mol.charges()
    ref = np.column_stack([Z, xyz.reshape(-1, 3)]).ravel()
    assert np.array_equal(np.array(vals), ref)
    r = subprocess.run([H5DUMP, "-d", "/mol/charge", str(f)], capture_output=True, text=True, check=True)
    assert re.search(r"\(0\): 0\b", r.stdout)


def test_reference_error_messages(tmp_path):
    f = tmp_path / "out.h5"
    with pytest.raises(RuntimeError, match="does not exist!"):
        wfn.write_hfwfn(f, "wfn", _wfn())
    wfn.create(f)
    w = _wfn()
    w.c_bv_A = np.zeros((24, 0))       # no virtual columns: the dataset is not written
    wfn.write_hfwfn(f, "wfn", w)
    with pytest.raises(RuntimeError, match=r"Read hfwfn: could not find wfn/c_bv_A\."):
        wfn.read_hfwfn(f, "wfn", 24, 5)
    wfn.write_hfwfn(f, "wfn2", _wfn())
    with pytest.raises(RuntimeError, match="Read hfwfn: incompatible matrix dimensions."):
        wfn.read_hfwfn(f, "wfn2", 24, 6)
    with pytest.raises(RuntimeError, match="exists already"):
        wfn.write_hfwfn(f, "wfn2", _wfn())


def test_cpp_layer_writes_reference_layout(tmp_path):
    """The C++ layer (include/mchip_wfn.hpp: filio::data_handler,
    desc::write_molecule / write_hfwfn / read_hfwfn) through its compiled test
    driver; the file it writes is then read here and by h5dump."""
    from chem_amd.build import WFN_DRIVER_EXE
    if not os.path.exists(WFN_DRIVER_EXE):
        pytest.skip("wfn_driver not built")
    f = tmp_path / "cpp.h5"
    r = subprocess.run([WFN_DRIVER_EXE, str(f)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    w = wfn.read_hfwfn(f, "hf", 24, 5)
    i = np.arange(24)[:, None]
    j = np.arange(24)[None, :]
    ref = np.sin(0.37 * i + 1.91 * j) / (1 + i + j)
    assert np.array_equal(w.c_bo_A, ref[:, :5]) and np.array_equal(w.c_bv_A, ref[:, 5:])
    assert np.array_equal(w.eps(), -1.0 + 0.1 * np.arange(24))
    assert (w.scf_energy, w.nuc_energy, w.wfn_energy) == (-85.25, 9.125, -76.125)
    if os.path.exists(H5DUMP):
        ds, groups = _dump(f)
        assert {"mol", "hf"} <= set(groups)
        assert ds["atoms"] == ("H5T_IEEE_F64LE", (3, 4)) and ds["mult"] == ("H5T_STD_I32LE", (1,))
