/*
 * qex.h -- C ABI of the MI355X Kohn-Sham SCF engine (libqex.so, gfx950).
 *
 * Conventions shared by every entry point:
 *  - All array arguments are DEVICE pointers owned by the caller (row-major, float64 unless
 *    stated).  The library allocates nothing: each entry point that needs scratch takes a
 *    caller-provided workspace whose size is returned by the matching *_workspace_bytes().
 *  - Every call takes a hipStream_t (passed as void*) and only enqueues work; nothing
 *    synchronises the host.  Results are ready when the stream is.
 *  - Return value: QEX_OK (0) or a negative QEX_E_* code; qex_last_error() returns a
 *    thread-local message for the last failing call on this thread.  The library never
 *    aborts or exits.
 *  - Kernels are deterministic: reductions run in a fixed order, with no floating-point
 *    atomics, so a system's result does not depend on what else is in the batch.
 */
#ifndef QEX_H
#define QEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QEX_OK 0
#define QEX_E_BADARG (-1)
#define QEX_E_WORKSPACE (-2)
#define QEX_E_LAUNCH (-3)
#define QEX_E_UNSUPPORTED (-4)

#define QEX1D_MAX_GRIDS 1024
#define QEX1D_MAX_ELECTRONS 62
#define QEX1D_MAX_MLP_LAYERS 6
#define QEX1D_MAX_MLP_WIDTH 1024 /* hidden widths; also bounded by LDS (the problem is refused if it does not fit) */
#define QEX1D_MAX_MLP_WINDOW 33

/* SCF loop semantics.
 * PYTHON: jax_dft scf.kohn_sham / qedft od/scf.py:434-560 -- convergence on MSE(n_out - n_in)
 *         of the same iteration; reflection about the centre of the nuclei (window per system).
 * SCAN:   jax_dft jit_scf.kohn_sham / qedft od/jit_scf.py:143-244 -- convergence on
 *         MSE(n_mixed - n_in), visible from the next iteration; reflection about grid centre. */
#define QEX1D_LOOP_PYTHON 0
#define QEX1D_LOOP_SCAN 1

/* Exchange-correlation kinds.  LOCAL kinds give an energy density per point
 * (E = dx e(n).n, v = e + J_e^T n; jax_dft scf.get_xc_energy/get_xc_potential);
 * GLOBAL kinds give f summed to a scalar (E = dx sum f, v = dx df/dn; od/scf.py:32-76). */
#define QEX1D_XC_LDA_EXP 0      /* jax_dft xc LDA for the exponential interaction (local) */
#define QEX1D_XC_POWER_LOCAL 1  /* e = c (n + eps)^p (local) */
#define QEX1D_XC_POWER_GLOBAL 2 /* f = c (n + eps)^p summed (global) */
#define QEX1D_XC_LOCAL_MLP 3    /* windowed per-point MLP (local); window 1 = classical_models.py:123 */
#define QEX1D_XC_NONE 4         /* no XC (v_xc = 0) */
#define QEX1D_XC_GLOBAL_MLP 5   /* MLP on the whole density (global, amplitude encoded):
                                   classical_models.py:177-224 build_global_mlp */
/* Any other functional (an opaque xc_energy_density_fn), evaluated by the caller between
 * iterations (qex1d_kohn_sham_step): xc_params is then an exchange buffer [3][B][N] the caller fills
 * before each step: [0] the energy density at the input density of the iteration (GLOBAL: the
 * scalar f at [0][b][0]), [1] the potential there (LOCAL: d(sum_j e_j n_j)/dn_i; GLOBAL:
 * dE/dn_i), [2] the energy density at the pending output density (GLOBAL: scalar at [2][b][0]).
 * The densities to evaluate at come from qex1d_external_xc_densities. */
#define QEX1D_XC_EXTERNAL_LOCAL 6
#define QEX1D_XC_EXTERNAL_GLOBAL 7

/* Transform applied to the network output of LOCAL_MLP / GLOBAL_MLP (the jax_dft neural_xc layer
 * that qedft/models/wrappers.py:201-203 appends with wrap_with_negative_transform). */
#define QEX1D_MLP_OUTPUT_LINEAR 0
#define QEX1D_MLP_OUTPUT_NEGATIVITY 1 /* e -> -swish(e) = -e sigmoid(e), range (-inf, 0.278] */

#define QEX_ACT_SILU 0
#define QEX_ACT_TANH 1
#define QEX_ACT_SOFTPLUS 2
#define QEX_ACT_SIGMOID 3
#define QEX_ACT_RELU 4
#define QEX_ACT_GELU_TANH 5

#define QEX_INTERACTION_EXPONENTIAL 0 /* A exp(-kappa |r|)   (p0 = A, p1 = kappa) */
#define QEX_INTERACTION_SOFT 1        /* 1 / sqrt(r^2 + p0)  (p0 = soften factor)  */

typedef struct qex1d_problem {
  int32_t batch;          /* B systems sharing one uniform grid */
  int32_t num_grids;      /* N, 5 <= N <= QEX1D_MAX_GRIDS */
  double dx;              /* grid spacing */
  int32_t num_iterations; /* T */
  int32_t max_electrons;  /* >= every num_electrons[b]; <= QEX1D_MAX_ELECTRONS */
  double alpha;           /* linear mixing factor of iteration 0 */
  double alpha_decay;     /* alpha_{k+1} = alpha_k * alpha_decay */
  int32_t num_mixing_iterations;
  double density_mse_converge_tolerance; /* negative disables early stopping */
  int32_t loop;                          /* QEX1D_LOOP_* */
  int32_t reflect;                       /* enforce_reflection_symmetry */
  int32_t xc_kind;                       /* QEX1D_XC_* */
  double xc_c, xc_p, xc_eps;             /* POWER_* functionals */
  /* LOCAL_MLP: mlp_layers Dense layers; widths[0] = window, widths[mlp_layers] = 1. */
  int32_t mlp_layers;
  int32_t mlp_widths[QEX1D_MAX_MLP_LAYERS + 1];
  int32_t mlp_activation; /* QEX_ACT_* */
  double mlp_norm;        /* density_normalization_factor */
  int32_t mlp_output;     /* QEX1D_MLP_OUTPUT_* */
  /* Reverse pass only: the states of the first stop_gradient_iterations iterations are constants
     (jax.lax.stop_gradient for idx <= stop_gradient_step in jax_dft jit_scf.kohn_sham, called with the
     config's value at qedft/train/od/train.py:113: stop_gradient_iterations = stop_gradient_step + 1);
     0 differentiates every iteration.  The forward ignores it. */
  int32_t stop_gradient_iterations;
} qex1d_problem;

typedef struct qex1d_inputs {
  const double* interaction_table;  /* [N]   u(k dx), k = 0..N-1 (even interaction) */
  const double* external_potential; /* [B][N] */
  const double* initial_density;    /* [B][N] */
  const int32_t* num_electrons;     /* [B]   1 <= N_e <= max_electrons */
  const int32_t* reflect_window;    /* [B][2] inclusive index window (lo, hi); used if reflect */
  const double* xc_params;          /* LOCAL_MLP: W1[in][out], b1, W2, b2, ... packed, float64;
                                       EXTERNAL kinds: the exchange buffer [3][B][N] */
} qex1d_inputs;

typedef struct qex1d_outputs {
  double* density;           /* [B][T][N]  mixed density stored in the state */
  double* total_energy;      /* [B][T] */
  double* hartree_potential; /* [B][T][N] */
  double* xc_potential;      /* [B][T][N] */
  double* xc_energy_density; /* [B][T][N] for LOCAL kinds, [B][T] for GLOBAL kinds */
  double* gap;               /* [B][T] */
  uint8_t* converged;        /* [B][T] */
  double* eigenvalues;       /* optional [B][T][max_electrons/2+1], may be NULL */
  uint64_t* stats;           /* optional [2], accumulated (caller zeroes): Sturm multisection rounds
                                executed over all eigenpairs, eigenpairs solved; may be NULL */
  double* orbitals;          /* optional [B][T][K][N], K = max_electrons/2+1: the orbitals of each
                                iteration (2-norm 1; slots beyond N_e/2+1 untouched); may be NULL.
                                Needed by qex1d_kohn_sham_vjp. */
} qex1d_outputs;

/* Workspace bytes needed by qex1d_kohn_sham for this problem (0 on bad arguments). */
size_t qex1d_workspace_bytes(const qex1d_problem* p);

/* Batched Kohn-Sham SCF: T iterations for B systems, every iteration's state stored.
 * Replaces jax_dft scf.kohn_sham / jit_scf.kohn_sham (called from qedft od/eval.py:67-89,
 * od/train.py:108-121) and qedft od/scf.py:434 kohn_sham_amplitude_encoded /
 * od/jit_scf.py:247 (GLOBAL xc kinds). */
int qex1d_kohn_sham(const qex1d_problem* p, const qex1d_inputs* in, const qex1d_outputs* out,
                    void* workspace, size_t workspace_bytes, void* stream);

/* Continuous batching (the serving path): p->batch is a table of B slots, each holding one request
 * that runs its own SCF iteration, so requests join and leave the running batch between iterations.
 * One call is one step: every busy slot advances by one iteration (a request takes
 * num_iterations + 1 steps: the last one only evaluates the final energies, then the slot's
 * trajectory is finalised).  Arrays:
 *  - slot_iteration [B] (device int32): the caller sets every entry to -1 once; the library keeps
 *    it: 0..T while a request runs, -2 once its outputs [s][0..T-1] are final;
 *  - admit [B] (device int32): 1 for each slot in which the caller has placed a new request since
 *    the previous step (its rows of external_potential, initial_density, num_electrons and, with
 *    reflection, reflect_window); only slots at -1 or -2 may be admitted; the caller clears it for
 *    the next step.
 * in/out/workspace are those of qex1d_kohn_sham for B systems (qex1d_workspace_bytes); the
 * workspace carries the slots' state between steps.  A request's results equal those of
 * qex1d_kohn_sham on it alone, bit for bit.  Replaces the reference's one-batch-per-call
 * vmap(kohn_sham) (qedft/train/od/train.py:124-126) for a stream of requests; no counterpart
 * in the reference beyond that. */
int qex1d_slots_step(const qex1d_problem* p, const qex1d_inputs* in, const qex1d_outputs* out, int32_t* slot_iteration,
                     const int32_t* admit, void* workspace, size_t workspace_bytes, void* stream);

/* qex1d_kohn_sham one pass at a time, for functionals the caller evaluates between passes (the
 * EXTERNAL xc kinds; any kind is accepted).  Call with iteration = 0, 1, ..., T in order on one
 * stream: pass k < T runs SCF iteration k, pass T only evaluates the last energies and finalises the
 * outputs.  Pass 0 resets the state kept in the workspace.  The T + 1 passes give the results of
 * qex1d_kohn_sham, bit for bit, when the functional's values are the same.  Replaces the python
 * loop of jax_dft scf.kohn_sham (and qedft od/scf.py:517-560) with an opaque
 * xc_energy_density_fn. */
int qex1d_kohn_sham_step(const qex1d_problem* p, const qex1d_inputs* in, const qex1d_outputs* out,
                         int32_t iteration, void* workspace, size_t workspace_bytes, void* stream);

/* The densities an EXTERNAL functional is evaluated at before pass `iteration` of
 * qex1d_kohn_sham_step, [B][N] each: density_in, the input density of that iteration (flipped and
 * averaged about the system's reflection window for LOCAL kinds when p->reflect), and
 * density_out, the output density whose energy the pass completes (that of iteration - 1; a copy
 * of density_in at iteration 0, where no energy is pending). */
int qex1d_external_xc_densities(const qex1d_problem* p, const qex1d_inputs* in, const qex1d_outputs* out,
                                int32_t iteration, void* workspace, size_t workspace_bytes, double* density_in,
                                double* density_out, void* stream);

/* Workspace bytes for qex1d_solve_noninteracting. */
size_t qex1d_noninteracting_workspace_bytes(int32_t batch, int32_t num_grids, int32_t max_electrons);

/* Batched non-interacting solve: lowest eigenpairs of T + diag(v) (5-point kinetic stencil),
 * density of the first N_e spin-orbitals, sum of their eigenvalues and the gap.
 * Replaces jax_dft scf.solve_noninteracting_system / qedft od/scf.py:233-268 and the
 * "noninteracting" initial density of od/trainer.py:72-75.
 * eigenvalues [B][K] and wavefunctions [B][K][N] (K = max_electrons/2+1, dx-normalised)
 * are optional (NULL). */
int qex1d_solve_noninteracting(int32_t batch, int32_t num_grids, double dx, int32_t max_electrons,
                               const double* potential, const int32_t* num_electrons, double* density,
                               double* total_eigen_energies, double* gap, double* eigenvalues,
                               double* wavefunctions, void* workspace, size_t workspace_bytes,
                               void* stream);

/* Allocation-free variant of qex1d_solve_noninteracting (same arguments, no workspace).  No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops that build initial densities per request).  The
 * workspace lives in a fixed device array of QEX1D_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; a problem whose qex1d_noninteracting_workspace_bytes() exceeds it is refused with
 * QEX_E_WORKSPACE (for N = 513 and N_e <= 8 that is up to 800 systems per call).  Calls on one GPU
 * share that array: calls that overlap in time on the same GPU (different streams or host threads)
 * race and must be serialised by the caller. */
#define QEX1D_NOALLOC_SCRATCH_BYTES (16u << 20)
int qex1d_solve_noninteracting_noalloc(int32_t batch, int32_t num_grids, double dx, int32_t max_electrons,
                                       const double* potential, const int32_t* num_electrons, double* density,
                                       double* total_eigen_energies, double* gap, double* eigenvalues,
                                       double* wavefunctions, void* stream);

/* v_ext[b][i] = -sum_a Z[b][a] u(x_i - R[b][a]) on grids[N].
 * Replaces jax_dft utils.get_atomic_chain_potential (od/scf.py:490-495). */
int qex1d_external_potential(int32_t batch, int32_t num_grids, const double* grids, int32_t num_nuclei,
                             const double* locations, const double* nuclear_charges, int32_t interaction,
                             double p0, double p1, double* out, void* stream);

/* v_H[b] = dx K n[b], K_ij = table[|i-j|].  Replaces od/scf.py:303-325 get_hartree_potential. */
int qex1d_hartree_potential(int32_t batch, int32_t num_grids, double dx, const double* table,
                            const double* density, double* out, void* stream);

/* Workspace bytes for qex1d_xc_evaluate (0 on bad arguments). */
size_t qex1d_xc_workspace_bytes(const qex1d_problem* p);

/* The XC functional p describes (xc_kind, xc_c/xc_p/xc_eps, network widths, activation, norm and
 * output transform; batch, num_grids, dx) on p->batch given densities [B][N], with the network
 * kernels the SCF runs; the SCF-only fields of p are not read, and no reflection is applied.
 * Outputs (each may be NULL):
 *  - energy_density: [B][N] for LOCAL kinds (e_xc per electron at each point), [B] for GLOBAL
 *    kinds (POWER_GLOBAL: sum_i f(n_i); GLOBAL_MLP: the network's f(n));
 *  - xc_energy [B]: LOCAL kinds dx sum_i e_i n_i, GLOBAL kinds dx times the value above;
 *  - xc_potential [B][N]: LOCAL kinds (dE/dn_i)/dx = e_i + sum_j n_j de_j/dn_i, GLOBAL kinds dE/dn_i.
 * Replaces qedft od/scf.py:32-76 get_xc_energy_amplitude_encoded / get_xc_potential_amplitude_encoded
 * (GLOBAL kinds) and jax_dft scf.get_xc_energy / get_xc_potential (LOCAL kinds), and the
 * functional's own evaluation xc_energy_density_fn(density). */
int qex1d_xc_evaluate(const qex1d_problem* p, const double* xc_params, const double* density,
                      double* energy_density, double* xc_energy, double* xc_potential, void* workspace,
                      size_t workspace_bytes, void* stream);

/* Opt-in kernel timing for measurement.  While enabled, qex1d_kohn_sham records a hipEvent pair
 * on its stream around each kernel launch (at most max_launches).  qex_profile_read waits for
 * the recorded events and returns, per kernel kind (0 MLP, 1 potential, 2 eigen, 3 density,
 * 4 finalize), the summed milliseconds and the number of launches, then clears the records.
 * No counterpart in the reference (which has no profiler hooks). */
int qex_profile_enable(int32_t enable, int32_t max_launches);
int qex_profile_read(double* ms_per_kind, int32_t* launches_per_kind, int32_t nkinds);
/* Kernel kinds that get events while profiling is enabled: bit k for kind k (default all).
 * Timing only the kinds of interest keeps the event records out of the other launches. */
int qex_profile_kinds(uint32_t mask);

const char* qex_last_error(void);
int32_t qex_abi_version(void);


/* Reverse pass of qex1d_kohn_sham for the scan loop with a LOCAL_MLP functional (1 or 2 hidden
 * layers, widths <= 64, either output transform), with or without reflection symmetry (the flip
 * inside reflect_window), or a GLOBAL_MLP functional (the same widths; the amplitude-encoded loop
 * of qedft od/jit_scf.py:15-78, 247-330), density_mse_converge_tolerance < 0 and the
 * problem's stop_gradient_iterations: given the forward's
 * inputs, its outputs density / hartree_potential / xc_potential / orbitals, and cotangents of the
 * stored energies d_total_energy [B][T] and densities d_density [B][T][N] (either may be NULL), writes
 * d_xc_params [P] = the gradient with respect to the packed MLP parameters, summed over the batch.
 * Replaces jax.grad through jit_scf.kohn_sham (qedft/train/od/train.py:250-323, the loss of
 * od/trainer.py:139-178). */
size_t qex1d_kohn_sham_vjp_workspace_bytes(const qex1d_problem* p);
int qex1d_kohn_sham_vjp(const qex1d_problem* p, const qex1d_inputs* in, const qex1d_outputs* forward,
                        const double* d_total_energy, const double* d_density, double* d_xc_params, void* workspace,
                        size_t workspace_bytes, void* stream);
const char* qex1d_vjp_last_error(void);

/* ============================================================================================
 * 3D restricted Kohn-Sham (LDA) / Hartree-Fock on an atom-centred grid, Cartesian s/p basis.
 * ========================================================================================== */
#define QEX3D_MAX_NAO 96
#define QEX3D_MAX_DIIS 12
#define QEX3D_MAX_SCF_BATCH 1024
#define QEX3D_MAX_NPRIM 8
#define QEX3D_RKS_LDA 0 /* Slater exchange + VWN5 correlation */
#define QEX3D_RHF 1
#define QEX3D_RKS_NN_GLOBAL 2 /* amplitude-encoded network on the whole grid density */
#define QEX3D_RKS_NN_LOCAL 4  /* per-point network ("NN" xctype): exc_g = net(rho_g), vrho_g = d exc_g / d rho_g */
#define QEX3D_MAX_NN_LOCAL_HIDDEN 512
#define QEX3D_RKS_LDA_X 3     /* Slater exchange alone: pyscf's xc code "lda" (mf.xc = "lda" in
                                 qedft/train/td/trainer_legacy_no_jit.py:635 and
                                 qedft/data_io/td/dataset_generation.py:392) */
/* LDA functionals of qex3d_nr_rks_lda */
#define QEX3D_LDA_SVWN5 0  /* Slater exchange + VWN5 correlation (pyscf "lda,vwn") */
#define QEX3D_LDA_SLATER 1 /* Slater exchange alone (pyscf "lda") */
#define QEX3D_MAX_NN_HIDDEN 1024
/* Precision of the SCF's grid work (qex3d_scf_problem.precision) */
#define QEX3D_PRECISION_FP64 0     /* every stage in FP64 */
#define QEX3D_PRECISION_FP32_GRID 1 /* AO values stored in FP32; rho from them on FP64 matrix cores (FP64
                                      sums); the Vxc contraction on FP32 matrix cores with FP32 sums over
                                      16-point units and FP64 sums across units; the XC function,
                                      energies, Coulomb, Fock, DIIS and eigensolve in FP64 */

/* Basis on the device.  Shells are listed in AO order (an s shell is 1 AO, a p shell 3: x, y, z);
 * coefficients include the primitive normalisation and a normalised contraction. */
typedef struct qex3d_basis {
  int32_t natm, nshell, nao;
  const double* atom_coords;  /* [natm][3] bohr */
  const double* atom_charges; /* [natm] */
  const int32_t* shell_atom;  /* [nshell] */
  const int32_t* shell_l;     /* [nshell] 0 or 1 */
  const int32_t* shell_nprim; /* [nshell] */
  const int32_t* shell_offset; /* [nshell] into exponents / coefficients */
  const double* exponents;
  const double* coefficients;
  const int32_t* ao_shell; /* [nao] */
  const int32_t* ao_lmn;   /* [nao] lx | ly << 2 | lz << 4 */
  int32_t max_nprim;       /* >= every shell_nprim, <= QEX3D_MAX_NPRIM */
} qex3d_basis;

/* S, T, V (nuclear attraction) [nao][nao].  Replaces pyscf int1e_ovlp/kin/nuc (libcint), called at
 * qedft/train/td/hf_legacy.py:452-471 (get_hcore, get_ovlp). */
int qex3d_int1e(const qex3d_basis* b, double* S, double* T, double* V, void* stream);

/* ERIs (ij|kl) as the [npair][npair] matrix over i>=j, k>=l pairs (npair = nao(nao+1)/2, pair index
 * i(i+1)/2 + j), computed per shell quartet (McMurchie-Davidson).  The workspace holds primitive-pair
 * records (nshell(nshell+1)/2 * max_nprim^2 * 128 bytes) and the work lists.
 * Schwarz screening (bases of >= 2048 AO pairs): an entry whose bound max sqrt|(ab|ab)| over its two
 * shell pairs multiplies to below 1e-15 is written as 0 (so every skipped integral is below 1e-15 in
 * magnitude); primitive quartets whose s-type magnitude bound c_ab c_cd 2 pi^(5/2) / (p q sqrt(p + q)) is
 * below 1e-17 (1e-20 with a p function) are skipped.
 * Replaces pyscf int2e (libcint) used by get_jk, qedft/train/td/hf_legacy.py:452-471. */
size_t qex3d_eri_workspace_bytes(const qex3d_basis* b);
int qex3d_eri(const qex3d_basis* b, double* eri, void* workspace, size_t workspace_bytes, void* stream);

/* The ERIs of B molecules in one launch sequence (eri: B device pointers, each [npair_b][npair_b]):
 * every result equals qex3d_eri on that molecule alone, bit for bit; the molecules' work overlaps
 * instead of paying each launch's latency in turn.  Supplies the per-molecule eri
 * inputs of the padded batch (qedft/train/td/scf_functions_masked.py:917-967, eri argument at :920),
 * which the reference builds molecule by molecule with pyscf int2e; no batched counterpart there. */
#define QEX3D_MAX_ERI_BATCH 1024
size_t qex3d_eri_batch_workspace_bytes(int32_t B, const qex3d_basis* bases);
int qex3d_eri_batch(int32_t B, const qex3d_basis* bases, double* const* eri, void* workspace, size_t workspace_bytes,
                    void* stream);

/* Atom-centred grid: natm * nrad * nang points from radial (r, w r^2) and angular (unit vectors,
 * weights summing to 4 pi) base tables, Becke partition.  Replaces pyscf dft.Grids.build
 * (qedft/train/td/trainer_legacy_no_jit.py:248-251). */
int qex3d_grid(const qex3d_basis* b, int32_t nrad, const double* rad_r, const double* rad_w, int32_t nang,
               const double* ang_xyz, const double* ang_w, double* coords, double* weights, void* stream);

/* Fuzzy-cell partition of the atom-centred grid: cell functions P_A = prod_B (1 - g(mu_AB)) / 2 with
 * Becke's g (three iterations of 1.5 mu - 0.5 mu^3) or Stratmann-Scuseria-Frisch's
 * (CPL 257, 213 (1996), eq. 14; pyscf gen_grid.stratmann, which the reference selects with
 * grids.becke_scheme at trainer_legacy_no_jit.py:250).  No atomic-size adjustment (homonuclear
 * chains need none). */
#define QEX3D_PARTITION_BECKE 0
#define QEX3D_PARTITION_STRATMANN 1
/* qex3d_grid with the partition scheme chosen. */
int qex3d_grid_partition(const qex3d_basis* b, int32_t scheme, int32_t nrad, const double* rad_r, const double* rad_w,
                         int32_t nang, const double* ang_xyz, const double* ang_w, double* coords, double* weights,
                         void* stream);

/* ao[g][mu] on ngrid points.  Replaces pyscf numint.eval_ao (td/eval_gto.py:20-70). */
int qex3d_eval_ao(const qex3d_basis* b, int64_t ngrid, const double* coords, double* ao, void* stream);

/* The same AO values rounded to FP32 (computed in FP64) in the layout of the FP32 grid path:
 * ao32[QEX3D_AO32_ROWS(ngrid)][QEX3D_AO32_LD(nao)], columns past nao and rows past ngrid zero.
 * Input of qex3d_scf / qex3d_scf_batch with precision QEX3D_PRECISION_FP32_GRID (the mixed-precision
 * form of the eval_ao call above). */
#define QEX3D_AO32_LD(nao) ((((nao) + 15) / 16) * 16)
#define QEX3D_AO32_ROWS(ngrid) ((((ngrid) + 63) / 64) * 64)
int qex3d_eval_ao_f32(const qex3d_basis* b, int64_t ngrid, const double* coords, float* ao32, void* stream);

/* rho_g = sum_{mu nu} ao[g][mu] dm[mu][nu] ao[g][nu].  Replaces eval_rho
 * (qedft/train/td/numint_legacy.py:351-373). */
int qex3d_eval_rho(int64_t ngrid, int32_t nao, const double* ao, const double* dm, double* rho, void* stream);

/* LDA nr_rks: nelec_excsum[0] = sum rho w, [1] = sum rho w e_xc; vmat = V + V^T with
 * V = ao^T diag(w v_xc / 2) ao.  Replaces nr_rks (qedft/train/td/numint.py:16-273, LDA branch
 * :42-64, symmetrisation :264-265). */
size_t qex3d_nr_rks_workspace_bytes(int64_t ngrid, int32_t nao);
int qex3d_nr_rks(int64_t ngrid, int32_t nao, const double* ao, const double* weights, const double* dm,
                 double* nelec_excsum, double* vmat, void* workspace, size_t workspace_bytes, void* stream);
/* The same for a chosen LDA functional (QEX3D_LDA_*): libxc's eval_xc as nr_rks calls it for the
 * xc_code (qedft/train/td/numint.py:42-64, eval_xc :313-336).  qex3d_nr_rks is functional
 * QEX3D_LDA_SVWN5.  Same workspace as qex3d_nr_rks. */
int qex3d_nr_rks_lda(int32_t functional, int64_t ngrid, int32_t nao, const double* ao, const double* weights,
                     const double* dm, double* nelec_excsum, double* vmat, void* workspace, size_t workspace_bytes,
                     void* stream);

/* Global (amplitude-encoded) neural XC: x = rho[ngrid] -> Dense(h1) -> gelu -> Dense(h2) -> gelu ->
 * Dense(1) -> -scale * swish (gelu in its tanh form); excsum = the output, vrho = d excsum / d rho.
 * Parameters packed as W1[ngrid][h1], b1[h1], W2[h1][h2], b2[h2], W3[h2], b3[1] (flax Dense
 * kernels are (in, out)).  nr_rks "NN-AmplitudeEncoding": nelec_excsum = {sum rho w, excsum},
 * vmat = V + V^T with V = ao^T diag(w vrho / 2) ao.  Replaces exc_and_vrho_global
 * (qedft/train/td/trainer_legacy_no_jit.py:46-53, MLP :96-107) inside nr_rks
 * (qedft/train/td/numint.py:206-231, 264-265). */
size_t qex3d_nr_rks_nn_workspace_bytes(int64_t ngrid, int32_t nao, int32_t h1, int32_t h2);
int qex3d_nr_rks_nn(int64_t ngrid, int32_t nao, const double* ao, const double* weights, const double* dm,
                    const double* nn_params, int32_t h1, int32_t h2, double scale, double* nelec_excsum,
                    double* vmat, void* workspace, size_t workspace_bytes, void* stream);

/* Per-point ("NN", local) neural XC: exc_g = -scale swish(MLP(rho_g)) with MLP = Dense(h1) -> gelu ->
 * Dense(h2) -> gelu -> Dense(1) on the scalar rho_g (gelu in its tanh form), vrho_g = d exc_g / d rho_g
 * (the VJP with a unit cotangent of exc_and_vrho_local, qedft/train/td/trainer_legacy_no_jit.py:56-63,
 * MLP :96-107).  Parameters packed as W1[1][h1], b1, W2[h1][h2], b2, W3[h2], b3 (flax kernels (in, out));
 * h1, h2 <= QEX3D_MAX_NN_LOCAL_HIDDEN.  nr_rks "NN" branch (qedft/train/td/numint.py:179-204):
 * nelec_excsum = {sum rho w, sum rho w exc}, vmat = V + V^T with V = ao^T diag(w vrho / 2) ao.
 * Same workspace as qex3d_nr_rks. */
int qex3d_nr_rks_nn_local(int64_t ngrid, int32_t nao, const double* ao, const double* weights, const double* dm,
                          const double* nn_params, int32_t h1, int32_t h2, double scale, double* nelec_excsum,
                          double* vmat, void* workspace, size_t workspace_bytes, void* stream);

/* The network XC evaluated on a density, as the custom eval_xc installed by define_xc_ returns it
 * (qedft/train/td/trainer_legacy_no_jit.py:76-93 with exc_and_vrho_global / exc_and_vrho_local):
 * is_global != 0: exc[0] = the network's scalar on rho[ngrid] (packing of qex3d_nr_rks_nn),
 * vrho[g] = d exc / d rho_g; is_global == 0: exc[g], vrho[g] per point (packing of
 * qex3d_nr_rks_nn_local). */
size_t qex3d_eval_xc_nn_workspace_bytes(int64_t ngrid, int32_t h1, int32_t h2, int32_t is_global);
int qex3d_eval_xc_nn(int32_t is_global, int64_t ngrid, const double* rho, const double* nn_params, int32_t h1,
                     int32_t h2, double scale, double* exc, double* vrho, void* workspace, size_t workspace_bytes,
                     void* stream);

/* Batched generalized symmetric eigenproblem F C = S C diag(e), S positive definite, n <= 96:
 * e ascending, C^T S C = 1 (columns are eigenvectors).  Cholesky S = L L^T, X = L^-T, then the
 * symmetric eigenproblem of X^T F X in one workgroup (Householder tridiagonalisation, Sturm
 * multisection, inverse iteration, back-transformation).  Replaces generalized_eigh
 * (qedft/train/td/generalized_eigensolver.py:264-330; pyscfad scipy.linalg.eigh at td/hf.py:490-494). */
size_t qex3d_gen_eigh_workspace_bytes(int32_t batch, int32_t n);
int qex3d_gen_eigh(int32_t batch, int32_t n, const double* F, const double* S, double* mo_energy, double* mo_coeff,
                   void* workspace, size_t workspace_bytes, void* stream);

/* Reverse pass of the batched symmetric eigendecomposition A = V diag(w) V^T (w [batch][n] ascending,
 * v [batch][n][n] row-major with eigenvectors in the columns) for cotangents gw [batch][n] and gv
 * [batch][n][n] (either may be NULL): gA = sym(V (F o (U gv) + diag(gw)) U), F_ij = 1/(w_j - w_i)
 * and 0 where |w_j - w_i| < eps^0.6 (degenerate pairs, whose numerator is taken as zero).
 * exact = 0 reproduces qedft/train/td/generalized_eigensolver.py:158-219 (degen_eigh_bwd) as written:
 * its eivect = jnp.transpose(eivec, axes=(-2, -1)) is the identity on a 2-D array, so U = V;
 * exact = 1 uses U = V^T, the derivative of the decomposition. */
size_t qex3d_eigh_vjp_workspace_bytes(int32_t batch, int32_t n);
int qex3d_eigh_vjp(int32_t batch, int32_t n, const double* w, const double* v, const double* gw, const double* gv,
                   int32_t exact, double* gA, void* workspace, size_t workspace_bytes, void* stream);

typedef struct qex3d_scf_problem {
  int32_t nao, nelectron, method, max_cycle, diis_space;
  int64_t ngrid;
  double conv_tol;   /* |dE| < conv_tol and |g|/sqrt(nvir nocc) < sqrt(conv_tol) */
  double energy_nuc;
  int32_t sync_every; /* 0: enqueue max_cycle cycles without host synchronisation (converged cycles are
                         no-ops on the device); k > 0: the host waits on the stream every k cycles and
                         stops enqueueing once converged */
  int32_t nn_h1, nn_h2; /* QEX3D_RKS_NN_GLOBAL: hidden widths */
  double nn_scale;
  int32_t precision; /* QEX3D_PRECISION_FP64 (0) or QEX3D_PRECISION_FP32_GRID (RKS methods) */
} qex3d_scf_problem;

typedef struct qex3d_scf_inputs {
  const double* S;       /* [nao][nao] */
  const double* H;       /* core Hamiltonian T + V */
  const double* eri;     /* pair matrix from qex3d_eri */
  const double* ao;      /* [ngrid][nao] (RKS) */
  const double* weights; /* [ngrid] (RKS) */
  const double* nn_params; /* packed network (QEX3D_RKS_NN_GLOBAL) */
  const float* ao_f32;     /* QEX3D_PRECISION_FP32_GRID: qex3d_eval_ao_f32 layout (ao may then be NULL) */
} qex3d_scf_inputs;

typedef struct qex3d_scf_outputs {
  double* e_tot;     /* [max_cycle + 1]: guess energy, then after each cycle (repeated once converged) */
  double* mo_energy; /* [nao] ascending */
  double* mo_coeff;  /* [nao][nao], columns are orbitals, C^T S C = 1 */
  double* dm;        /* [nao][nao] */
  int32_t* status;   /* [8]: [1] cycles run, [4] converged */
} qex3d_scf_outputs;

/* Closed-shell SCF with CDIIS from cycle 1 and a core-Hamiltonian guess.  Replaces
 * qedft/train/td/hf_legacy.py:51-272 (_scf, kernel) with get_fock (:379-427), generalized eigh
 * (td/generalized_eigensolver.py:264-330), make_rdm1 (td/hf.py:303-308) and get_veff
 * (td/rks_legacy.py:36-127). */
size_t qex3d_scf_workspace_bytes(const qex3d_scf_problem* p);
int qex3d_scf(const qex3d_scf_problem* p, const qex3d_scf_inputs* in, const qex3d_scf_outputs* out, void* workspace,
              size_t workspace_bytes, void* stream);

/* The same SCF for `batch` molecules of any sizes and methods in one sequence of launches (each
 * stage is one launch over the batch).  Replaces the padded/masked batch of
 * qedft/train/td/scf_functions_masked.py:199-303 and :917-967 (_scf_test_padded): every molecule
 * keeps its own nao and grid, so its results equal those of qex3d_scf on it alone, bit for bit.
 * problems/inputs/outputs are host arrays of `batch` entries.  A molecule stops at convergence or
 * at its own max_cycle; the host polls every min(sync_every > 0) cycles (0 for all: no polling). */
size_t qex3d_scf_batch_workspace_bytes(int32_t batch, const qex3d_scf_problem* problems);
int qex3d_scf_batch(int32_t batch, const qex3d_scf_problem* problems, const qex3d_scf_inputs* inputs,
                    const qex3d_scf_outputs* outputs, void* workspace, size_t workspace_bytes, void* stream);

const char* qex3d_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* QEX_H */
