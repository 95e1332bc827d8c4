"""Host-side pieces of the 3D path (basis normalisation, molecule builders, grid base tables)
against the oracle; no device needed."""

import numpy as np
import pytest

from oracle import td as otd
from qex_amd.td import dft, gto


def _check_same(mol, omol):
    np.testing.assert_allclose(mol.coords, omol.coords, rtol=0, atol=1e-14)
    assert mol.nao == omol.nao and mol.nelectron == omol.nelectron
    assert abs(mol.energy_nuc() - omol.energy_nuc()) < 1e-12
    for (a, l, e, c), sh in zip(mol.shells, omol.shells):
        assert a == sh.center and l == sh.l
        np.testing.assert_array_equal(e, sh.exps)
        np.testing.assert_allclose(c, sh.coefs * np.array([otd.prim_norm(x, l) for x in sh.exps]), rtol=1e-14)


@pytest.mark.parametrize("n", [2, 5, 16])
def test_synthetic_chain_matches_oracle(n):
    _check_same(gto.synthetic_chain(n, 3), otd.synthetic_chain(n, 3))


def test_h2_631g_matches_oracle():
    _check_same(gto.Mole([("H", (0, 0, 0)), ("H", (0, 0, 0.74))], basis="6-31g"), otd.h2_molecule(0.74))


def test_grid_base_tables_match_oracle():
    for a, b in zip(dft.radial_treutler(50), otd.radial_treutler(50)):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(dft.angular_product(10, 11), otd.angular_product(10, 11)):
        np.testing.assert_array_equal(a, b)
    _, w = dft.angular_product()
    assert abs(w.sum() - 4 * np.pi) < 1e-13


def test_ao_layout_order():
    mol = gto.synthetic_chain(2, 1)
    assert mol.nao == 10
    assert [l for _, l, _, _ in mol.shells] == [0, 0, 1, 0, 0, 1]


def test_eval_xc_checks_the_network_kind_before_the_device():
    """is_global_xc must name the network's kind (a per-point network is local, the amplitude-encoding
    one global); a mismatch is refused on the host."""
    from qex_amd.td import xc

    W, b = xc.init_local_mlp(0, hidden=(4, 4))
    local = xc.LocalNNXC(W, b)
    Wg, bg = xc.init_global_mlp(0, 6, hidden=(4, 4))
    glob = xc.GlobalNNXC(Wg, bg)
    with pytest.raises(ValueError, match="is_global_xc"):
        xc.eval_xc("NN", np.ones(6), network=local, is_global_xc=True)
    with pytest.raises(ValueError, match="is_global_xc"):
        xc.eval_xc("NN-AmplitudeEncoding", np.ones(6), network=glob, is_global_xc=False)
