"""The FP32 grid path of the 3D SCF (precision "fp32_grid": AO values stored in FP32, rho from them on
FP64 matrix cores, the Vxc contraction on FP32 matrix cores with FP64 sums across 16-point units;
the XC function, energies, Coulomb, Fock, DIIS and eigensolve in FP64) against the FP64 path and the
oracle's golden energies.

Stated bounds (per result): |E_fp32 - E_golden| <= 1e-6 Ha (the metric's bar), |E_fp32 - E_fp64|
<= 1e-7 Ha, density L2 on the grid sqrt(sum_g w_g (rho32 - rho64)^2) <= 1e-6.  The measured
values are printed."""

import json
import os

import numpy as np
import pytest
import torch

from oracle import td as otd

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
E_GOLDEN_BOUND = 1e-6
E_FP64_BOUND = 1e-7
RHO_L2_BOUND = 1e-6


@pytest.fixture(scope="module")
def qtd():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from qex_amd import td
    return td


def _solver(qtd, mol, precision, xc="LDA,VWN"):
    mf = qtd.scf.RKS(mol, xc=xc)
    mf.conv_tol, mf.max_cycle, mf.diis_space = 1e-8, 30, 8
    mf.precision = precision
    return mf


def _rho_l2(qtd, mf, D32, D64):
    ao = qtd.numint.eval_ao(mf.mol, mf.grids.coords)
    r32 = qtd.numint.eval_rho(mf.mol, ao, D32)
    r64 = qtd.numint.eval_rho(mf.mol, ao, D64)
    return float(torch.sqrt((mf.grids.weights * (r32 - r64) ** 2).sum()))


def test_eval_ao_f32_is_the_rounded_fp64_values_in_the_padded_layout(qtd):
    mol = qtd.gto.synthetic_chain(5, 3)  # nao 25 -> ld 32
    grids = qtd.dft.Grids(mol).build()
    ao = qtd.numint.eval_ao(mol, grids.coords)
    ao32 = qtd.numint.eval_ao_f32(mol, grids.coords)
    ng, n = ao.shape
    assert ao32.shape == ((ng + 63) // 64 * 64, 32) and ao32.dtype == torch.float32
    ref = ao.float()
    got = ao32[:ng, :n]
    # same FP64 evaluation, one rounding: within half an FP32 ulp of the FP64 value
    assert torch.all((got.double() - ao).abs() <= ao.abs() * 2.0 ** -24 + 1e-45)
    assert int((got != ref).sum()) <= ng * n // 1000  # at most rare last-bit differences
    assert torch.count_nonzero(ao32[ng:]) == 0 and torch.count_nonzero(ao32[:, n:]) == 0


def test_fp32_grid_chains_within_stated_bounds(qtd):
    """H2, H4, H8, H16 of the 3D config (seeds 102..116, n_ao 10..80, conv_tol 1e-8): converged, within
    1e-6 Ha of the oracle's end-to-end golden energies and 1e-7 Ha of the FP64 path.  The density
    comparison runs both paths to conv_tol 1e-11, so that the SCF's own convergence error (about
    sqrt(conv_tol) in the density) is below what the FP32 grid work changes."""
    ref = {c["n_atoms"]: c for c in json.load(open(os.path.join(GOLDEN, "c4_chain_reference.json")))["chains"]}
    for n in (2, 4, 8, 16):
        mol = qtd.gto.synthetic_chain(n, 100 + n)
        print(f"fp32_grid: H{n}", flush=True)
        m64 = _solver(qtd, mol, "fp64")
        e64 = m64.kernel()
        m32 = _solver(qtd, mol, "fp32_grid")
        m32.grids = m64.grids
        e32 = m32.kernel()
        assert m32.converged
        print(f"  H{n}: |dE| vs golden {abs(e32 - ref[n]['e_tot']):.2e}, vs fp64 {abs(e32 - e64):.2e}, "
              f"cycles {m32.cycles} (fp64 {m64.cycles})", flush=True)
        assert abs(e32 - ref[n]["e_tot"]) <= E_GOLDEN_BOUND
        assert abs(e32 - e64) <= E_FP64_BOUND
        tight = []
        for mf in (m64, m32):
            mf.conv_tol, mf.max_cycle = 1e-11, 80
            mf.kernel()
            assert mf.converged
            tight.append(mf)
        l2 = _rho_l2(qtd, m64, m32.dm, m64.dm)
        de = abs(m32.e_tot - m64.e_tot)
        print(f"  H{n} at conv_tol 1e-11: |dE| vs fp64 {de:.2e}, rho L2 {l2:.2e}, cycles {m32.cycles} "
              f"(fp64 {m64.cycles})", flush=True)
        assert de <= E_FP64_BOUND
        assert l2 <= RHO_L2_BOUND


def test_fp32_grid_batch_equals_solo_and_mixes_with_fp64(qtd):
    """A batch holding FP32-grid and FP64 molecules of several widths (ld 16..96) and both LDA codes:
    each result bitwise equal to the molecule solved alone."""
    specs = [(4, "fp32_grid", "LDA,VWN"), (8, "fp64", "LDA,VWN"), (2, "fp32_grid", "lda"), (16, "fp32_grid", "LDA,VWN"),
             (12, "fp32_grid", "LDA,VWN"), (18, "fp32_grid", "lda"), (6, "fp32_grid", "LDA,VWN"),
             (8, "fp32_grid", "LDA,VWN")]
    mols = [qtd.gto.synthetic_chain(n, 40 + n) for n, _, _ in specs]
    assert {(m.nao + 15) // 16 for m in mols} == {1, 2, 3, 4, 5, 6}
    solo = []
    for m, (n, prec, xc) in zip(mols, specs):
        mf = _solver(qtd, m, prec, xc)
        mf.kernel()
        solo.append((mf.energies.cpu().numpy().copy(), mf.dm.cpu().numpy().copy(), mf.cycles))
    mfs = [_solver(qtd, m, prec, xc) for m, (n, prec, xc) in zip(mols, specs)]
    qtd.scf.kernel_batch(mfs)
    for (n, prec, xc), mf, (e, D, cyc) in zip(specs, mfs, solo):
        np.testing.assert_array_equal(mf.energies.cpu().numpy(), e, err_msg=f"H{n} {prec}")
        np.testing.assert_array_equal(mf.dm.cpu().numpy(), D, err_msg=f"H{n} {prec}")
        assert mf.cycles == cyc


def test_fp32_grid_network_methods_within_bound_of_fp64(qtd):
    """The per-point ("NN") and amplitude-encoded network XC with the FP32 grid work: rho from the FP32
    AO values, the network in FP64, the Vxc contraction on FP32 matrix cores."""
    from qex_amd.td.xc import GlobalNNXC, LocalNNXC
    mol = qtd.gto.synthetic_chain(4, 7)
    W, b = otd.init_global_mlp(5, 1, hidden=(16, 8))
    rng = np.random.default_rng(7)
    b = [rng.standard_normal(x.shape) * 0.1 for x in b]
    for kind in ("local", "global"):
        es = {}
        for prec in ("fp64", "fp32_grid"):
            mf = _solver(qtd, mol, prec)
            mf.grids = qtd.dft.Grids(mol).build()
            if kind == "local":
                mf.define_xc_(LocalNNXC(W, b, scale=0.2), xctype="NN")
            else:
                Wg, bg = otd.init_global_mlp(9, mf.grids.size, hidden=(64, 32))
                mf.define_xc_(GlobalNNXC(Wg, bg, scale=0.05))
            mf.max_cycle = 40
            es[prec] = (mf.kernel(), mf.converged)
        print(f"fp32_grid {kind} network: |dE| vs fp64 {abs(es['fp32_grid'][0] - es['fp64'][0]):.2e}", flush=True)
        assert es["fp32_grid"][1] == es["fp64"][1]
        assert abs(es["fp32_grid"][0] - es["fp64"][0]) <= E_FP64_BOUND


def test_fp32_grid_argument_checks(qtd):
    mol = qtd.gto.synthetic_chain(2, 1)
    mf = _solver(qtd, mol, "fp16")
    with pytest.raises(ValueError, match="precision"):
        mf.kernel()
