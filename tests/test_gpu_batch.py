"""Batched Chebyshev filter plans (one launch per step for many channels,
graph-captured) against the per-channel filter, and the autotuner.

The batched kernel runs the same per-point arithmetic as the single-channel
kernel, but it is a different instantiation, and the compiler may contract a
multiply-add differently in it, so results agree to FP64 rounding (1e-12 of
the block's largest entry), not bit for bit."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(torch.max(torch.abs(a - b)) / torch.max(torch.abs(b)))


def channels(device):
    from cyclix_msparc_amd.api import Hamiltonian
    from cyclix_msparc_amd.system import small_system
    specs = [(45, 16, 14, (0.0, 1.0, 0.6), 5), (30, 16, 12, (0.0, 2.0, 0.1), 3), (45, 16, 14, (0.0, 0.0, 0.3), 7)]
    hams, blocks = [], []
    for k, (nr, nt, nz, kpt, ns) in enumerate(specs):
        sysm = small_system(nr=nr, nt=nt, nz=nz, twist=0.05, kpt=kpt, nstates=ns, npl=8)
        g = sysm.grid
        r = np.tile(g.r, g.nt * g.nz)
        veff = -1.5 * np.exp(-((r - sysm.atoms[:, 0].max()) / 1.2) ** 2)
        hams.append(Hamiltonian(device, sysm, kpt, veff, sysm.projectors()))
        rng = np.random.default_rng(10 + k)
        X = rng.standard_normal((ns, g.N)) + 1j * rng.standard_normal((ns, g.N))
        blocks.append(torch.from_numpy(X).to(device.torch_device))
    return hams, blocks


@pytest.mark.parametrize("m", [1, 6, 9])
def test_batched_plan_matches_single_filters(device, m):
    from cyclix_msparc_amd.api import FilterPlan
    hams, blocks = channels(device)
    bounds = ([2.0, 3.0, 2.5], [60.0, 55.0, 70.0], [-1.0, -0.5, -2.0])
    refs = [h.chebyshev_filter(X, m, bounds[0][c], bounds[1][c], bounds[2][c]) for c, (h, X) in
            enumerate(zip(hams, blocks))]
    work = [X.clone() for X in blocks]
    plan = FilterPlan(device, hams, work, m)
    for rep in range(2):          # the second pass replays the captured graph
        for X, X0 in zip(work, blocks):
            X.copy_(X0)
        plan.run(*bounds)
        for Y, R in zip(work, refs):
            assert rel(Y, R) < 1e-12, f"pass {rep}: {rel(Y, R):.3e}"
    plan.close()


def test_autotuned_plan_gives_the_same_result(device):
    from cyclix_msparc_amd.api import FilterPlan
    hams, blocks = channels(device)
    bounds = ([2.0, 3.0, 2.5], [60.0, 55.0, 70.0], [-1.0, -0.5, -2.0])
    refs = [h.chebyshev_filter(X, 5, bounds[0][c], bounds[1][c], bounds[2][c]) for c, (h, X) in
            enumerate(zip(hams, blocks))]
    work = [X.clone() for X in blocks]
    plan = FilterPlan(device, hams, work, 5, autotune=True, bounds=bounds)
    t = plan.tuning()
    assert t["tr"] >= 12 and t["tr"] * 16 <= 512
    for X, X0 in zip(work, blocks):
        X.copy_(X0)
    plan.run(*bounds)
    for Y, R in zip(work, refs):
        assert rel(Y, R) < 1e-12
    plan.close()


def test_plan_rejects_mixed_theta_counts(device):
    from cyclix_msparc_amd._native import CyxError
    from cyclix_msparc_amd.api import FilterPlan, Hamiltonian
    from cyclix_msparc_amd.system import small_system
    s1 = small_system(nr=30, nt=16, nz=12, twist=0.05, kpt=(0, 1.0, 0.2))
    s2 = small_system(nr=30, nt=13, nz=12, twist=0.05, kpt=(0, 1.0, 0.2))
    hs = [Hamiltonian(device, s, s.kpt, np.zeros(s.grid.N), s.projectors()) for s in (s1, s2)]
    bl = [torch.zeros((2, s.grid.N), dtype=torch.complex128, device=device.torch_device) for s in (s1, s2)]
    with pytest.raises(CyxError):
        FilterPlan(device, hs, bl, 3)


def test_plan_buffers_survive_handle_growth_and_shared_handles(device):
    """The plan owns its KB coefficients and gather partials: a larger
    single-channel filter on one of its Hamiltonians afterwards (which grows
    the handle's own scratch) leaves the plan's next pass exact, and one
    Hamiltonian may serve two channels of a plan."""
    from cyclix_msparc_amd.api import FilterPlan
    hams, blocks = channels(device)
    hams = [hams[0], hams[1], hams[0]]
    X2 = torch.from_numpy(np.random.default_rng(21).standard_normal((4, blocks[0].shape[1]))
                          + 1j * np.random.default_rng(22).standard_normal((4, blocks[0].shape[1])))
    blocks = [blocks[0], blocks[1], X2.to(device.torch_device)]
    bounds = ([2.0, 3.0, 2.2], [60.0, 55.0, 65.0], [-1.0, -0.5, -1.5])
    refs = [h.chebyshev_filter(X, 6, bounds[0][c], bounds[1][c], bounds[2][c]) for c, (h, X) in
            enumerate(zip(hams, blocks))]
    work = [X.clone() for X in blocks]
    plan = FilterPlan(device, hams, work, 6)
    plan.run(*bounds)
    for Y, R in zip(work, refs):
        assert rel(Y, R) < 1e-12
    big = torch.randn((40, blocks[0].shape[1]), dtype=torch.complex128, device=device.torch_device)
    hams[0].chebyshev_filter(big, 6, 2.0, 60.0, -1.0)
    for X, X0 in zip(work, blocks):
        X.copy_(X0)
    plan.run(*bounds)
    for Y, R in zip(work, refs):
        assert rel(Y, R) < 1e-12
    plan.close()
