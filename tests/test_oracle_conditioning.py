"""The FP64 noise floor of the reference's DF-J energy.

E_J = 1/2 gamma^T V^-1 gamma with an explicit V^-1 (jkdf.cpp:137-141): for one
water with cc-pVDZ-jkfit, relative perturbations of gamma at the level of one
rounding (1e-16) change E_J by about 1e-10 Eh.  This is the floor any
cross-implementation energy comparison runs into; K (DFMO, applied through
L^-1 per build) does not have it."""
import numpy as np

from chem_amd import desc, ints
from oracle import dfock


def test_dfj_energy_noise_floor():
    mol = desc.Molecule(desc.water_cluster(1, 7))
    B = ints.BTensorHost.build(mol, 1e-10).dense()
    V = ints.int2c2e(mol)
    Vinv, Linv, _ = dfock.metric_factors(V)
    rng = np.random.default_rng(0)
    C = rng.standard_normal((mol.nbf, mol.nocc)) * 0.3
    P = C @ C.T
    g = np.einsum("xmn,mn->x", B, 2 * P)
    e0 = 0.5 * g @ Vinv @ g
    spread = []
    for _ in range(8):
        gn = g * (1 + rng.standard_normal(g.shape) * 1e-16)
        spread.append(0.5 * gn @ Vinv @ gn - e0)
    spread = np.abs(spread)
    assert np.linalg.cond(V) > 1e6
    assert spread.max() > 1e-11          # the floor is real
    assert spread.max() < 1e-8           # and bounded
