"""ctypes binding of libqex.so (the C ABI declared in include/qex.h).

The library is required: importing the engine without it raises immediately; there is no
fallback computation path.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqex.so")

QEX_OK = 0
QEX_E_BADARG = -1
QEX_E_WORKSPACE = -2
QEX1D_NOALLOC_SCRATCH_BYTES = 16 << 20
QEX1D_MAX_GRIDS = 1024
QEX1D_MAX_ELECTRONS = 62
QEX1D_MAX_MLP_LAYERS = 6
QEX1D_MAX_MLP_WIDTH = 1024

QEX1D_LOOP_PYTHON = 0
QEX1D_LOOP_SCAN = 1

QEX1D_XC_LDA_EXP = 0
QEX1D_XC_POWER_LOCAL = 1
QEX1D_XC_POWER_GLOBAL = 2
QEX1D_XC_LOCAL_MLP = 3
QEX1D_XC_NONE = 4
QEX1D_XC_GLOBAL_MLP = 5
QEX1D_XC_EXTERNAL_LOCAL = 6
QEX1D_XC_EXTERNAL_GLOBAL = 7

QEX1D_MLP_OUTPUT_LINEAR = 0
QEX1D_MLP_OUTPUT_NEGATIVITY = 1
OUTPUTS = {"linear": QEX1D_MLP_OUTPUT_LINEAR, "negativity": QEX1D_MLP_OUTPUT_NEGATIVITY}

ACTIVATIONS = {"silu": 0, "swish": 0, "tanh": 1, "softplus": 2, "sigmoid": 3, "relu": 4, "gelu": 5}

QEX_INTERACTION_EXPONENTIAL = 0
QEX_INTERACTION_SOFT = 1


class Qex1dProblem(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_int32),
        ("num_grids", ctypes.c_int32),
        ("dx", ctypes.c_double),
        ("num_iterations", ctypes.c_int32),
        ("max_electrons", ctypes.c_int32),
        ("alpha", ctypes.c_double),
        ("alpha_decay", ctypes.c_double),
        ("num_mixing_iterations", ctypes.c_int32),
        ("density_mse_converge_tolerance", ctypes.c_double),
        ("loop", ctypes.c_int32),
        ("reflect", ctypes.c_int32),
        ("xc_kind", ctypes.c_int32),
        ("xc_c", ctypes.c_double),
        ("xc_p", ctypes.c_double),
        ("xc_eps", ctypes.c_double),
        ("mlp_layers", ctypes.c_int32),
        ("mlp_widths", ctypes.c_int32 * (QEX1D_MAX_MLP_LAYERS + 1)),
        ("mlp_activation", ctypes.c_int32),
        ("mlp_norm", ctypes.c_double),
        ("mlp_output", ctypes.c_int32),
        ("stop_gradient_iterations", ctypes.c_int32),
    ]


class Qex1dInputs(ctypes.Structure):
    _fields_ = [
        ("interaction_table", ctypes.c_void_p),
        ("external_potential", ctypes.c_void_p),
        ("initial_density", ctypes.c_void_p),
        ("num_electrons", ctypes.c_void_p),
        ("reflect_window", ctypes.c_void_p),
        ("xc_params", ctypes.c_void_p),
    ]


class Qex1dOutputs(ctypes.Structure):
    _fields_ = [
        ("density", ctypes.c_void_p),
        ("total_energy", ctypes.c_void_p),
        ("hartree_potential", ctypes.c_void_p),
        ("xc_potential", ctypes.c_void_p),
        ("xc_energy_density", ctypes.c_void_p),
        ("gap", ctypes.c_void_p),
        ("converged", ctypes.c_void_p),
        ("eigenvalues", ctypes.c_void_p),
        ("stats", ctypes.c_void_p),
        ("orbitals", ctypes.c_void_p),
    ]


QEX3D_RKS_LDA = 0
QEX3D_RHF = 1
QEX3D_RKS_NN_GLOBAL = 2
QEX3D_RKS_LDA_X = 3
QEX3D_RKS_NN_LOCAL = 4
QEX3D_MAX_NN_LOCAL_HIDDEN = 512
QEX3D_LDA_SVWN5 = 0
QEX3D_LDA_SLATER = 1
QEX3D_MAX_NN_HIDDEN = 1024
QEX3D_PRECISION_FP64 = 0
QEX3D_PRECISION_FP32_GRID = 1


def ao32_ld(nao):
    """Row length of the FP32 AO layout (QEX3D_AO32_LD)."""
    return (nao + 15) // 16 * 16


def ao32_rows(ngrid):
    """Row count of the FP32 AO layout (QEX3D_AO32_ROWS)."""
    return (ngrid + 63) // 64 * 64
QEX3D_MAX_NAO = 96
QEX3D_PARTITION_BECKE = 0
QEX3D_PARTITION_STRATMANN = 1
QEX3D_MAX_NPRIM = 8
QEX3D_MAX_SCF_BATCH = 1024
QEX3D_MAX_ERI_BATCH = 1024


class Qex3dBasis(ctypes.Structure):
    _fields_ = [("natm", ctypes.c_int32), ("nshell", ctypes.c_int32), ("nao", ctypes.c_int32)] + [
        (name, ctypes.c_void_p) for name in ("atom_coords", "atom_charges", "shell_atom", "shell_l", "shell_nprim",
                                             "shell_offset", "exponents", "coefficients", "ao_shell", "ao_lmn")] + [
        ("max_nprim", ctypes.c_int32)]


class Qex3dScfProblem(ctypes.Structure):
    _fields_ = [("nao", ctypes.c_int32), ("nelectron", ctypes.c_int32), ("method", ctypes.c_int32),
                ("max_cycle", ctypes.c_int32), ("diis_space", ctypes.c_int32), ("ngrid", ctypes.c_int64),
                ("conv_tol", ctypes.c_double), ("energy_nuc", ctypes.c_double), ("sync_every", ctypes.c_int32), ("nn_h1", ctypes.c_int32),
                ("nn_h2", ctypes.c_int32), ("nn_scale", ctypes.c_double), ("precision", ctypes.c_int32)]


class Qex3dScfInputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("S", "H", "eri", "ao", "weights", "nn_params", "ao_f32")]


class Qex3dScfOutputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("e_tot", "mo_energy", "mo_coeff", "dm", "status")]


# Every exported symbol and its ctypes signature (restype, argtypes).
_P = ctypes.c_void_p
_I = ctypes.c_int32
_D = ctypes.c_double
_S = ctypes.c_size_t
SIGNATURES = {
    "qex_last_error": (ctypes.c_char_p, []),
    "qex_abi_version": (_I, []),
    "qex1d_workspace_bytes": (_S, [ctypes.POINTER(Qex1dProblem)]),
    "qex1d_kohn_sham": (_I, [ctypes.POINTER(Qex1dProblem), ctypes.POINTER(Qex1dInputs),
                             ctypes.POINTER(Qex1dOutputs), _P, _S, _P]),
    "qex1d_slots_step": (_I, [ctypes.POINTER(Qex1dProblem), ctypes.POINTER(Qex1dInputs),
                              ctypes.POINTER(Qex1dOutputs), _P, _P, _P, _S, _P]),
    "qex1d_kohn_sham_step": (_I, [ctypes.POINTER(Qex1dProblem), ctypes.POINTER(Qex1dInputs),
                                  ctypes.POINTER(Qex1dOutputs), _I, _P, _S, _P]),
    "qex1d_external_xc_densities": (_I, [ctypes.POINTER(Qex1dProblem), ctypes.POINTER(Qex1dInputs),
                                         ctypes.POINTER(Qex1dOutputs), _I, _P, _S, _P, _P, _P]),
    "qex1d_noninteracting_workspace_bytes": (_S, [_I, _I, _I]),
    "qex1d_solve_noninteracting": (_I, [_I, _I, _D, _I, _P, _P, _P, _P, _P, _P, _P, _P, _S, _P]),
    "qex1d_solve_noninteracting_noalloc": (_I, [_I, _I, _D, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "qex1d_external_potential": (_I, [_I, _I, _P, _I, _P, _P, _I, _D, _D, _P, _P]),
    "qex1d_hartree_potential": (_I, [_I, _I, _D, _P, _P, _P, _P]),
    "qex1d_xc_workspace_bytes": (_S, [ctypes.POINTER(Qex1dProblem)]),
    "qex1d_xc_evaluate": (_I, [ctypes.POINTER(Qex1dProblem), _P, _P, _P, _P, _P, _P, _S, _P]),
    "qex_profile_enable": (_I, [_I, _I]),
    "qex_profile_read": (_I, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32), _I]),
    "qex_profile_kinds": (_I, [ctypes.c_uint32]),
    "qex1d_kohn_sham_vjp_workspace_bytes": (_S, [ctypes.POINTER(Qex1dProblem)]),
    "qex1d_kohn_sham_vjp": (_I, [ctypes.POINTER(Qex1dProblem), ctypes.POINTER(Qex1dInputs),
                                 ctypes.POINTER(Qex1dOutputs), _P, _P, _P, _P, _S, _P]),
    "qex1d_vjp_last_error": (ctypes.c_char_p, []),
    "qex3d_last_error": (ctypes.c_char_p, []),
    "qex3d_int1e": (_I, [ctypes.POINTER(Qex3dBasis), _P, _P, _P, _P]),
    "qex3d_eri_workspace_bytes": (_S, [ctypes.POINTER(Qex3dBasis)]),
    "qex3d_eri": (_I, [ctypes.POINTER(Qex3dBasis), _P, _P, _S, _P]),
    "qex3d_eri_batch_workspace_bytes": (_S, [_I, ctypes.POINTER(Qex3dBasis)]),
    "qex3d_eri_batch": (_I, [_I, ctypes.POINTER(Qex3dBasis), _P, _P, _S, _P]),
    "qex3d_grid": (_I, [ctypes.POINTER(Qex3dBasis), _I, _P, _P, _I, _P, _P, _P, _P, _P]),
    "qex3d_grid_partition": (_I, [ctypes.POINTER(Qex3dBasis), _I, _I, _P, _P, _I, _P, _P, _P, _P, _P]),
    "qex3d_eval_ao": (_I, [ctypes.POINTER(Qex3dBasis), ctypes.c_int64, _P, _P, _P]),
    "qex3d_eval_ao_f32": (_I, [ctypes.POINTER(Qex3dBasis), ctypes.c_int64, _P, _P, _P]),
    "qex3d_eval_rho": (_I, [ctypes.c_int64, _I, _P, _P, _P, _P]),
    "qex3d_nr_rks_workspace_bytes": (_S, [ctypes.c_int64, _I]),
    "qex3d_nr_rks": (_I, [ctypes.c_int64, _I, _P, _P, _P, _P, _P, _P, _S, _P]),
    "qex3d_nr_rks_lda": (_I, [_I, ctypes.c_int64, _I, _P, _P, _P, _P, _P, _P, _S, _P]),
    "qex3d_nr_rks_nn_workspace_bytes": (_S, [ctypes.c_int64, _I, _I, _I]),
    "qex3d_nr_rks_nn": (_I, [ctypes.c_int64, _I, _P, _P, _P, _P, _I, _I, ctypes.c_double, _P, _P, _P, _S, _P]),
    "qex3d_nr_rks_nn_local": (_I, [ctypes.c_int64, _I, _P, _P, _P, _P, _I, _I, ctypes.c_double, _P, _P, _P, _S,
                                    _P]),
    "qex3d_eval_xc_nn_workspace_bytes": (_S, [ctypes.c_int64, _I, _I, _I]),
    "qex3d_eval_xc_nn": (_I, [_I, ctypes.c_int64, _P, _P, _I, _I, ctypes.c_double, _P, _P, _P, _S, _P]),
    "qex3d_gen_eigh_workspace_bytes": (_S, [_I, _I]),
    "qex3d_gen_eigh": (_I, [_I, _I, _P, _P, _P, _P, _P, _S, _P]),
    "qex3d_eigh_vjp_workspace_bytes": (_S, [_I, _I]),
    "qex3d_eigh_vjp": (_I, [_I, _I, _P, _P, _P, _P, _I, _P, _P, _S, _P]),
    "qex3d_scf_workspace_bytes": (_S, [ctypes.POINTER(Qex3dScfProblem)]),
    "qex3d_scf": (_I, [ctypes.POINTER(Qex3dScfProblem), ctypes.POINTER(Qex3dScfInputs),
                       ctypes.POINTER(Qex3dScfOutputs), _P, _S, _P]),
    "qex3d_scf_batch_workspace_bytes": (_S, [_I, ctypes.POINTER(Qex3dScfProblem)]),
    "qex3d_scf_batch": (_I, [_I, ctypes.POINTER(Qex3dScfProblem), ctypes.POINTER(Qex3dScfInputs),
                             ctypes.POINTER(Qex3dScfOutputs), _P, _S, _P]),
}

KERNEL_KINDS = ["mlp", "potential", "eigen", "density", "finalize", "int1e", "eri", "grid", "eval_ao", "rho", "vxc",
                "coulomb", "scf_fock", "scf_eigh", "scf_energy", "scf_other", "nn_global", "adjoint", "adj_mlp",
                "adj_eig"]

_lib = None


class QexError(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Loads libqex.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise QexError(f"{LIB_PATH} is missing: build it with `python -m qex_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check_vjp(rc: int, what: str) -> None:
    if rc != QEX_OK:
        msg = load().qex1d_vjp_last_error()
        raise QexError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def check3(rc: int, what: str) -> None:
    if rc != QEX_OK:
        msg = load().qex3d_last_error()
        raise QexError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def check(rc: int, what: str) -> None:
    if rc != QEX_OK:
        msg = load().qex_last_error()
        raise QexError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
