"""Device-resident VA13 minimization (needs a GPU).

The state machine is pinned to the oracle's VA13 bit for bit on the CPU
(test_va13_state_machine.py), and the adjoint it calls equals the oracle's
FP32 adjoint in the kernel's summation order bit for bit
(test_gpu_adjoint.py).  So the device minimizer must follow the oracle's
trajectories exactly: for every pose the same final DOFs, last and best
energies and evaluation counts as the oracle's VA13 on that adjoint
(orc_minimize_kernel_order), after 10, 80 and 500 evaluations.
Also: no pose ends above its start, the reported DOFs reproduce the reported
energy exactly, every pose stays within its evaluation limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from gpuattract_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("which", ["rigid", "modes"])
def test_minimize_against_oracle(ctx, W, O, which):
    if which == "rigid":
        wl = W.make_workload("min rigid", 81, 300, 80, 40)
    else:
        wl = W.make_workload("min modes", 82, 300, 80, 40, 3, 2, with_lig_grid=False)
    hs = ctx.upload_workload(wl)
    e0 = ctx.adjoint(hs, wl.dofs)[:, 0]
    for ivmax in (10, 80, 500):
        d, el, eb, nf = ctx.minimize(hs, wl.dofs, ivmax)
        od, oel, oeb, onf = O.minimize_kernel_order(wl, ivmax=ivmax)
        od = np.ascontiguousarray(od[:, :d.shape[1]])  # rigid: 6 DOFs per pose
        assert np.all(nf <= ivmax)
        assert np.array_equal(nf, onf)
        assert np.array_equal(d.view(np.uint32), od.view(np.uint32))
        assert np.array_equal(el.astype(np.float64), oel)
        assert np.array_equal(eb.astype(np.float64), oeb)
        assert np.all(eb <= e0 + 1e-6 * np.abs(e0))
        # the reported DOFs reproduce the reported energy
        assert np.array_equal(ctx.adjoint(hs, d)[:, 0], el)
        print(f"{which}, {ivmax} evaluations: {len(nf)} poses equal to the oracle; mean "
              f"evaluations {nf.mean():.1f}")
    assert np.mean(eb < e0 - 1e-3) > 0.8


def test_minimize_limits_and_errors(ctx, W):
    from gpuattract_amd import hip
    wl = W.make_workload("min small", 83, 200, 60, 16)
    hs = ctx.upload_workload(wl)
    d, el, eb, nf = ctx.minimize(hs, wl.dofs, 0)
    # no line-search evaluation allowed: the start point is the result
    assert np.all(nf == 0) and np.array_equal(d, wl.dofs)
    assert np.array_equal(el, ctx.adjoint(hs, wl.dofs)[:, 0])
    with pytest.raises(hip.AttError):
        ctx.minimize(hs, wl.dofs, -1)
    d2, el2, eb2, nf2 = ctx.minimize(hs, wl.dofs, 25)
    d3, el3, eb3, nf3 = ctx.minimize(hs, wl.dofs, 25)
    assert np.array_equal(d2, d3) and np.array_equal(el2, el3)  # deterministic
    assert np.all(nf2 <= 25)
