"""C ABI of libqex.so on the CPU: the library loads, exports every function include/qex.h declares,
and rejects bad arguments with an error code and message before touching the device."""

import ctypes
import os
import re

import pytest

from qex_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "qex.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(qex[0-9a-z_]*)\s*\(", text, re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared_functions()
    assert len(names) >= 8
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert lib.qex_abi_version() == 1


def _problem(**kw):
    p = _lib.Qex1dProblem()
    p.batch, p.num_grids, p.dx, p.num_iterations, p.max_electrons = 2, 513, 0.1, 15, 4
    p.alpha, p.alpha_decay, p.num_mixing_iterations = 0.3, 1.0, 1
    p.density_mse_converge_tolerance = -1.0
    p.loop, p.reflect, p.xc_kind = _lib.QEX1D_LOOP_SCAN, 0, _lib.QEX1D_XC_LDA_EXP
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_workspace_query():
    lib = _lib.load()
    assert lib.qex1d_workspace_bytes(ctypes.byref(_problem())) > 0
    assert lib.qex1d_workspace_bytes(ctypes.byref(_problem(num_grids=4))) == 0
    assert lib.qex1d_workspace_bytes(ctypes.byref(_problem(num_grids=_lib.QEX1D_MAX_GRIDS + 1))) == 0
    assert lib.qex1d_workspace_bytes(ctypes.byref(_problem(max_electrons=_lib.QEX1D_MAX_ELECTRONS + 1))) == 0
    assert lib.qex1d_noninteracting_workspace_bytes(4, 513, 8) > 0
    assert lib.qex1d_noninteracting_workspace_bytes(0, 513, 8) == 0


@pytest.mark.parametrize("field,value,msg", [
    ("batch", 0, "batch"), ("num_grids", 3, "num_grids"), ("dx", -0.1, "dx"), ("num_iterations", 0, "num_iterations"),
    ("num_mixing_iterations", 0, "num_mixing"), ("loop", 7, "loop"), ("xc_kind", 99, "xc_kind"),
])
def test_kohn_sham_rejects_bad_problem(field, value, msg):
    lib = _lib.load()
    ins, outs = _lib.Qex1dInputs(), _lib.Qex1dOutputs()
    rc = lib.qex1d_kohn_sham(ctypes.byref(_problem(**{field: value})), ctypes.byref(ins), ctypes.byref(outs), None, 0,
                             None)
    assert rc == -1
    assert msg in lib.qex_last_error().decode()


def test_kohn_sham_rejects_missing_arrays_and_small_workspace():
    lib = _lib.load()
    p = _problem()
    ins, outs = _lib.Qex1dInputs(), _lib.Qex1dOutputs()
    assert lib.qex1d_kohn_sham(ctypes.byref(p), ctypes.byref(ins), ctypes.byref(outs), None, 0, None) == -1
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails before any launch
    ins = _lib.Qex1dInputs(fake, fake, fake, fake, None, None)
    outs = _lib.Qex1dOutputs(fake, fake, fake, fake, fake, fake, fake, None, None)
    rc = lib.qex1d_kohn_sham(ctypes.byref(p), ctypes.byref(ins), ctypes.byref(outs), fake, 16, None)
    assert rc == -2 and "workspace" in lib.qex_last_error().decode()


def test_mlp_validation():
    lib = _lib.load()
    p = _problem(xc_kind=_lib.QEX1D_XC_LOCAL_MLP, mlp_layers=3, mlp_activation=0, mlp_norm=2.0)
    for i, w in enumerate([17, 64, 64, 1]):
        p.mlp_widths[i] = w
    assert lib.qex1d_workspace_bytes(ctypes.byref(p)) > 0
    p.mlp_widths[0] = 16  # even window
    assert lib.qex1d_workspace_bytes(ctypes.byref(p)) == 0
    p.mlp_widths[0] = 17
    p.mlp_widths[3] = 2  # output width must be 1
    assert lib.qex1d_workspace_bytes(ctypes.byref(p)) == 0


def test_global_mlp_validation():
    lib = _lib.load()
    p = _problem(xc_kind=_lib.QEX1D_XC_GLOBAL_MLP, mlp_layers=3, mlp_activation=1, mlp_norm=2.0)
    for i, w in enumerate([p.num_grids, 32, 16, 1]):
        p.mlp_widths[i] = w
    assert lib.qex1d_workspace_bytes(ctypes.byref(p)) > 0
    p.mlp_widths[0] = p.num_grids - 1  # input width must be the grid size
    assert lib.qex1d_workspace_bytes(ctypes.byref(p)) == 0
    p.mlp_widths[0] = p.num_grids
    p.mlp_widths[1] = 513  # the reference notebook's global network width
    assert lib.qex1d_workspace_bytes(ctypes.byref(p)) > 0
    p.mlp_widths[1] = _lib.QEX1D_MAX_MLP_WIDTH + 1  # hidden widths at most QEX1D_MAX_MLP_WIDTH
    assert lib.qex1d_workspace_bytes(ctypes.byref(p)) == 0
    assert "hidden" in lib.qex_last_error().decode()


def test_engine_refuses_to_run_without_a_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from qex_amd.od import engine

    with pytest.raises(_lib.QexError):
        engine.device()


def _scf_problem(**kw):
    p = _lib.Qex3dScfProblem(10, 2, _lib.QEX3D_RHF, 30, 8, 0, 1e-9, 0.7)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_scf3d_workspace_query():
    lib = _lib.load()
    rhf = lib.qex3d_scf_workspace_bytes(ctypes.byref(_scf_problem()))
    assert rhf > 0
    # Hartree-Fock ignores ngrid: the workspace does not depend on it
    assert lib.qex3d_scf_workspace_bytes(ctypes.byref(_scf_problem(ngrid=5000))) == rhf
    rks = lib.qex3d_scf_workspace_bytes(ctypes.byref(_scf_problem(method=_lib.QEX3D_RKS_LDA, ngrid=5000)))
    assert rks > rhf
    assert lib.qex3d_scf_workspace_bytes(ctypes.byref(_scf_problem(method=_lib.QEX3D_RKS_LDA, ngrid=0))) == 0
    assert lib.qex3d_scf_workspace_bytes(ctypes.byref(_scf_problem(nao=_lib.QEX3D_MAX_NAO + 1))) == 0
    assert lib.qex3d_nr_rks_workspace_bytes(5000, 10) > 0
    assert lib.qex3d_nr_rks_workspace_bytes(0, 10) == 0


@pytest.mark.parametrize("field,value,code,msg", [
    ("nao", 0, -4, "nao"), ("nao", _lib.QEX3D_MAX_NAO + 1, -4, "nao"), ("nelectron", 3, -1, "closed-shell"),
    ("nelectron", 24, -1, "closed-shell"), ("method", 5, -1, "method"), ("max_cycle", 0, -1, "cycle"),
    ("diis_space", 13, -1, "diis"),
])
def test_scf3d_rejects_bad_problem(field, value, code, msg):
    lib = _lib.load()
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails before any launch
    ins = _lib.Qex3dScfInputs(fake, fake, fake, None, None)
    outs = _lib.Qex3dScfOutputs(fake, fake, fake, fake, fake)
    rc = lib.qex3d_scf(ctypes.byref(_scf_problem(**{field: value})), ctypes.byref(ins), ctypes.byref(outs), fake,
                       1 << 30, None)
    assert rc == code
    assert msg in lib.qex3d_last_error().decode()


def test_scf3d_rejects_missing_grid_and_small_workspace():
    lib = _lib.load()
    fake = ctypes.c_void_p(0x1000)
    ins = _lib.Qex3dScfInputs(fake, fake, fake, None, None)
    outs = _lib.Qex3dScfOutputs(fake, fake, fake, fake, fake)
    p = _scf_problem(method=_lib.QEX3D_RKS_LDA, ngrid=100)
    assert lib.qex3d_scf(ctypes.byref(p), ctypes.byref(ins), ctypes.byref(outs), fake, 1 << 30, None) == -1
    assert "grid" in lib.qex3d_last_error().decode()
    rc = lib.qex3d_scf(ctypes.byref(_scf_problem()), ctypes.byref(ins), ctypes.byref(outs), fake, 64, None)
    assert rc == -2 and "workspace" in lib.qex3d_last_error().decode()


def test_scf3d_batch_workspace_and_validation():
    lib = _lib.load()
    P = (_lib.Qex3dScfProblem * 3)(_scf_problem(), _scf_problem(method=_lib.QEX3D_RKS_LDA, ngrid=5000),
                                   _scf_problem(nao=40, nelectron=16))
    total = lib.qex3d_scf_batch_workspace_bytes(3, P)
    solo = sum(lib.qex3d_scf_workspace_bytes(ctypes.byref(P[i])) for i in range(3))
    # one record block for the whole batch instead of one per molecule
    assert 0 < total <= solo
    assert lib.qex3d_scf_batch_workspace_bytes(0, P) == 0
    assert lib.qex3d_scf_batch_workspace_bytes(_lib.QEX3D_MAX_SCF_BATCH + 1, P) == 0
    P[1].ngrid = 0  # RKS without a grid
    assert lib.qex3d_scf_batch_workspace_bytes(3, P) == 0
    P[1].ngrid = 5000
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails before any launch
    I = (_lib.Qex3dScfInputs * 3)(*[_lib.Qex3dScfInputs(fake, fake, fake, fake, fake) for _ in range(3)])
    O = (_lib.Qex3dScfOutputs * 3)(*[_lib.Qex3dScfOutputs(fake, fake, fake, fake, fake) for _ in range(3)])
    assert lib.qex3d_scf_batch(0, P, I, O, fake, 1 << 30, None) == -1
    assert "batch" in lib.qex3d_last_error().decode()
    P[2].nelectron = 7
    assert lib.qex3d_scf_batch(3, P, I, O, fake, 1 << 30, None) == -1
    assert "molecule 2: closed-shell" in lib.qex3d_last_error().decode()
    P[2].nelectron = 16
    assert lib.qex3d_scf_batch(3, P, I, O, fake, 64, None) == -2
    assert "workspace" in lib.qex3d_last_error().decode()


def test_3d_entry_points_reject_null_basis():
    lib = _lib.load()
    fake = ctypes.c_void_p(0x1000)
    assert lib.qex3d_int1e(None, fake, fake, fake, None) == -1
    assert lib.qex3d_eri(None, fake, fake, 1 << 20, None) == -1
    b = _lib.Qex3dBasis(1, 1, _lib.QEX3D_MAX_NAO + 1)
    b.max_nprim = 3
    assert lib.qex3d_eval_ao(ctypes.byref(b), 10, fake, fake, None) == -4
    assert lib.qex3d_eval_rho(10, 0, fake, fake, fake, None) == -1
