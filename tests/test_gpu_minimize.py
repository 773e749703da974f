"""Device-resident VA13 minimization (needs a GPU).

The state machine is pinned to the oracle's VA13 bit for bit on the CPU
(test_va13_state_machine.py); here the whole loop runs on the device with the
GPU adjoint.  GPU and oracle scores differ in the last bits (summation order),
which can move a line-search decision, so trajectories are compared by their
outcome: final energies agree (within 1e-3 relative) for most poses, no pose
ends above its start, the reported DOFs reproduce the reported energy exactly,
and every pose stays within its evaluation limit."""
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
    ivmax = 80
    d, el, eb, nf = ctx.minimize(hs, wl.dofs, ivmax)
    od, oel, oeb, onf = O.minimize(wl, ivmax=ivmax)
    e0 = ctx.adjoint(hs, wl.dofs)[:, 0]
    assert np.all(nf <= ivmax) and np.all(onf <= ivmax)
    assert np.all(eb <= e0 + 1e-6 * np.abs(e0))
    assert np.mean(eb < e0 - 1e-3) > 0.8
    # the reported DOFs reproduce the reported energy
    assert np.array_equal(ctx.adjoint(hs, d)[:, 0], el)
    rel = np.abs(eb.astype(np.float64) - oeb) / (np.abs(oeb) + 1.0)
    print(f"{which}: |E_gpu - E_oracle| / (|E| + 1): median {np.median(rel):.2e}, "
          f"share within 1e-3: {np.mean(rel <= 1e-3):.2f}; mean evaluations {nf.mean():.1f} "
          f"(oracle {onf.mean():.1f})")
    assert np.mean(rel <= 1e-3) >= 0.75


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
