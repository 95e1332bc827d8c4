"""Continuous batching of 1D solves (qex1d_slots_step through engine.SlotRunner): requests admitted
at different steps into a slot table, slots reused after their requests finish; every result is
bitwise equal to the same system solved alone by qex1d_kohn_sham, and within 1e-6 Ha of the oracle."""

import numpy as np
import pytest
import torch

from oracle import od as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from qex_amd.od import engine, utils, xc
    return engine, utils, xc


@pytest.mark.parametrize("decay,slots,per_step", [(1.0, 8, 1), (0.9, 5, 1), (1.0, 6, 3)])
def test_slot_results_equal_solo_solves(mods, decay, slots, per_step):
    engine, utils, xc = mods
    grids = np.linspace(-12.8, 12.8, 257)
    T, alpha = 6, 0.3
    mols = O.random_molecules(31, 14, grids=grids)
    W, b = O.glorot_normal_mlp(3, [9, 16, 16, 1])
    n0 = [O.gaussian_initial_density(grids, l, z, ne, weighted=True) for l, z, ne in mols]
    f = xc.LocalMLPXC(list(zip(W, b)), window=9, activation="silu", norm=2.0)
    run = engine.SlotRunner(grids, utils.exponential_coulomb, f, T, slots=slots, alpha=alpha, alpha_decay=decay,
                            num_mixing_iterations=1, loop="scan", depth=2, iterations_per_step=per_step)
    # the workspace is not cleared by the library: idle slots must stay idle whatever it holds
    run.ws.fill_(0x5A)
    # staggered admissions: 3 now, more after two steps, the rest as slots free up
    pending = list(range(len(mols)))
    results = {}
    plan = {0: 3, 2: 4}
    step = 0
    while len(results) < len(mols):
        want = plan.get(step, len(pending))
        take = []
        while pending and len(take) < min(want, run.free_slots()):
            i = pending.pop(0)
            take.append((i, mols[i][0], mols[i][1], mols[i][2], n0[i]))
        if take:
            run.admit(take)
        if run.busy_slots():
            if not run.can_step():
                for tag, e, conv in run.poll(wait=True):
                    results[tag] = e
            run.step()
        for tag, e, conv in run.poll():
            results[tag] = e
        step += 1
        assert step < 40 * len(mols)
    print(f"slots {slots}: {run.steps} steps for {len(mols)} requests of {T} iterations", flush=True)
    assert run.steps < len(mols) * (T + 1)  # requests overlapped
    xo = O.LocalMLP(W, b, "silu", 9, 2.0)
    for i, (l, z, ne) in enumerate(mols):
        solo = engine.Problem1D(grids, [l], [z], [ne], utils.exponential_coulomb, f, T, alpha=alpha,
                                alpha_decay=decay, num_mixing_iterations=1, loop="scan").run(n0[i][None])
        e_solo = float(solo.total_energy[0, -1].item())
        assert results[i] == e_solo, (i, results[i], e_solo)
        ref = O.jit_kohn_sham(l, z, ne, T, grids, xo, O.exponential_coulomb, n0[i], alpha=alpha, alpha_decay=decay,
                              num_mixing_iterations=1)
        assert abs(results[i] - ref.total_energy[-1]) < 1e-6
