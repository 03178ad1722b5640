"""An SCF on the GPU written to the reference's wavefunction file and read back
by the MP2 step, as the reference's driver chains its hfwfn and mpwfn jobs
(megalochem_driver.cpp:495-515, 600-640; wfn.hpp:266-326): the MP2 energy from
the file equals the one from the SCF in memory bit for bit, and the file's
energies are the SCF's."""
import os

import numpy as np
import pytest

from chem_amd import desc, ints, wfn

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(wfn.WFN_LIB), reason="libmchip_wfn.so not built (no HDF5)")]


def test_scf_to_file_to_mp2(tmp_path):
    from chem_amd.engine import Context, DeviceB, Metric
    from chem_amd.mp2 import sos_mp2
    from chem_amd.scf import run_rhf
    print("case: SCF -> hfwfn file -> MP2, 2 waters", flush=True)
    ctx = Context(0)
    geo = desc.water_cluster(2, 4)
    mol = desc.Molecule(geo)
    S, T, V = ints.int1e(mol)
    dB = DeviceB(ctx, ints.BTensorHost.build(mol, 1e-10))
    met = Metric(ctx, ints.int2c2e(mol))
    res = run_rhf(mol, dB, met, S, T + V, threshold=1e-8, guess="sad")
    assert res.converged
    f = tmp_path / "run.h5"
    wfn.create(f)
    wfn.write_molecule(f, "mol", mol)
    wfn.write_hfwfn(f, "hf", wfn.HfWavefunction.from_scf(mol, res))
    back = wfn.read_hfwfn(f, "hf", mol.nbf, mol.nocc)
    assert back.wfn_energy == res.energy
    assert abs(back.scf_energy + back.nuc_energy - res.energy) < 1e-12
    molri = desc.Molecule(geo, dfbasis="cc-pvdz-ri")
    dBr = DeviceB(ctx, ints.BTensorHost.build(molri, 1e-10))
    metr = Metric(ctx, ints.int2c2e(molri))
    e_mem = sos_mp2(molri, dBr, metr, S, res.C, res.eps, nlap=6)
    e_file = sos_mp2(molri, dBr, metr, S, back.coefficients(), back.eps(), nlap=6)
    print("E_mp2 memory %.14f file %.14f" % (e_mem["energy"], e_file["energy"]))
    assert e_file["energy"] == e_mem["energy"]
    assert np.array_equal(e_file["terms"], e_mem["terms"])
