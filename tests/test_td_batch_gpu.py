"""Batched 3D SCF (qex3d_scf_batch): molecules of different sizes and methods in one launch
sequence give exactly (bit for bit) what each gives alone.  This is the property the reference's
padded/masked batch is built for (scf_functions_masked.py:970-1113 compare_padded_vs_non_padded);
each solo result is itself pinned to the oracle by test_td_gpu.py."""

import numpy as np
import pytest
import torch

from oracle import td as otd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qtd():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from qex_amd import td
    return td


def _mixed(qtd):
    from qex_amd.td.gto import Mole
    from qex_amd.td.xc import GlobalNNXC

    mfs = []
    mf = qtd.scf.RHF(Mole([("H", (0, 0, 0)), ("H", (0, 0, 0.74))], basis="6-31g"))
    mfs.append(mf)
    mfs.append(qtd.scf.RKS(qtd.gto.synthetic_chain(6, 11)))
    mf = qtd.scf.RKS(qtd.gto.synthetic_chain(4, 7))
    mf.grids = qtd.dft.Grids(mf.mol).build()
    W, b = otd.init_global_mlp(9, mf.grids.size, hidden=(64, 32))
    mf.define_xc_(GlobalNNXC(W, b, scale=0.05))
    mfs.append(mf)
    mf = qtd.scf.RKS(qtd.gto.synthetic_chain(8, 5))
    mf.max_cycle = 5  # stops unconverged while the others go on
    mfs.append(mf)
    mf = qtd.scf.RKS(qtd.gto.synthetic_chain(16, 5))
    mf.max_cycle = 60
    mfs.append(mf)
    mfs.append(qtd.scf.RHF(qtd.gto.synthetic_chain(2, 3)))
    for mf in mfs:
        mf.conv_tol = 1e-9
    return mfs


def _snapshot(mf):
    return dict(e=mf.energies.cpu().numpy().copy(), mo=mf.mo_energy.cpu().numpy().copy(),
                C=mf.mo_coeff.cpu().numpy().copy(), D=mf.dm.cpu().numpy().copy(), cycles=mf.cycles,
                converged=mf.converged, e_tot=mf.e_tot)


def _assert_same(a, b, tag):
    for k in ("e", "mo", "C", "D"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{tag}: {k}")
    assert (a["cycles"], a["converged"], a["e_tot"]) == (b["cycles"], b["converged"], b["e_tot"]), tag


@pytest.mark.parametrize("sync_every", [0, 3])
def test_batch_equals_solo_bitwise(qtd, sync_every):
    mfs = _mixed(qtd)
    solo = []
    for i, mf in enumerate(mfs):
        mf.sync_every = sync_every
        print(f"solo {i}: nao={mf.mol.nao} method={mf.method}", flush=True)
        mf.kernel()
        solo.append(_snapshot(mf))
    assert not solo[3]["converged"] and solo[3]["cycles"] == 5
    assert all(s["converged"] for i, s in enumerate(solo) if i != 3)
    print(f"batch of {len(mfs)}", flush=True)
    qtd.scf.kernel_batch(mfs)
    for i, mf in enumerate(mfs):
        _assert_same(_snapshot(mf), solo[i], f"molecule {i}")
    # order does not matter
    print("reversed batch", flush=True)
    qtd.scf.kernel_batch(mfs[::-1])
    for i, mf in enumerate(mfs):
        _assert_same(_snapshot(mf), solo[i], f"molecule {i} (reversed)")


def test_batch_of_copies_and_a_second_smaller_batch(qtd):
    """Eight copies of one chain in one batch, then a batch of two: every copy equals the solo run."""
    mol = qtd.gto.synthetic_chain(6, 21)
    ref = qtd.scf.RKS(mol)
    ref.kernel()
    r = _snapshot(ref)
    mfs = []
    for _ in range(8):
        mf = qtd.scf.RKS(mol)
        mf._ints, mf.grids, mf._ao = ref._ints, ref.grids, ref._ao
        mfs.append(mf)
    qtd.scf.kernel_batch(mfs)
    for i, mf in enumerate(mfs):
        _assert_same(_snapshot(mf), r, f"copy {i}")
    qtd.scf.kernel_batch(mfs[:2])
    for i, mf in enumerate(mfs[:2]):
        _assert_same(_snapshot(mf), r, f"second batch copy {i}")


def test_eri_batch_equals_solo_bitwise(qtd):
    """qex3d_eri_batch over molecules of every size class (wave-split small ones, 16-lane large ones,
    Schwarz-screened H16), repeated and in mixed order: each pair matrix equals the molecule's own
    qex3d_eri bit for bit; a second, smaller batch reusing the objects gives the same again."""
    from qex_amd.td.gto import Mole
    mols = [qtd.gto.synthetic_chain(16, 116), Mole([("H", (0, 0, 0)), ("H", (0, 0, 0.74))], basis="6-31g"),
            qtd.gto.synthetic_chain(8, 108), qtd.gto.synthetic_chain(4, 104), qtd.gto.synthetic_chain(2, 102),
            qtd.gto.synthetic_chain(16, 116), qtd.gto.synthetic_chain(6, 11)]
    solo = [m.intor_eri_pairs() for m in mols]
    for batch in (mols, mols[2:5]):
        got = qtd.gto.intor_eri_pairs_batch(batch)
        for m, g in zip(batch, got):
            ref = solo[mols.index(m)]
            assert g.shape == ref.shape
            assert torch.equal(g, ref), f"nao {m.nao}: max diff {(g - ref).abs().max().item():.3e}"


def test_eri_batch_argument_checks(qtd):
    import ctypes
    from qex_amd import _lib
    lib = _lib.load()
    assert lib.qex3d_eri_batch_workspace_bytes(0, None) == 0
    m = qtd.gto.synthetic_chain(2, 1)
    bases = (_lib.Qex3dBasis * 1)(m.device_basis())
    assert lib.qex3d_eri_batch(0, bases, None, None, 0, None) == _lib.QEX_E_BADARG
    ptrs = (ctypes.c_void_p * 1)(None)
    assert lib.qex3d_eri_batch(1, bases, ptrs, None, 0, None) == _lib.QEX_E_BADARG
