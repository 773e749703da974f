"""Device-resident VA13 minimization (needs a GPU).

The state machine is pinned to the oracle's VA13 bit for bit on the CPU
(test_va13_state_machine.py); here the whole loop runs on the device with the
GPU adjoint.  GPU and oracle scores differ in the last bits (summation order);
on a rugged landscape that moves line-search decisions and, over many
evaluations, sends trajectories to different minima.  The oracle shows the same
sensitivity against itself (FP32 vs FP64 scoring), so it is the yardstick:
  - after 10 evaluations, GPU and FP32 oracle final energies agree (1e-3
    relative) at least as often as FP32 and FP64 oracle runs do, minus 0.1;
  - after 80, likewise minus 0.15, with no systematic bias of the GPU's final
    energies (median difference within the spread);
  - no pose ends above its start, the reported DOFs reproduce the reported
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
    for ivmax, slack in ((10, 0.1), (80, 0.15)):
        d, el, eb, nf = ctx.minimize(hs, wl.dofs, ivmax)
        _, _, oeb, onf = O.minimize(wl, ivmax=ivmax)
        _, _, deb, _ = O.minimize(wl, ivmax=ivmax, f64=True)
        assert np.all(nf <= ivmax) and np.all(onf <= ivmax)
        assert np.all(eb <= e0 + 1e-6 * np.abs(e0))
        # the reported DOFs reproduce the reported energy
        assert np.array_equal(ctx.adjoint(hs, d)[:, 0], el)
        agree = lambda a, b: np.mean(np.abs(a - b) / (np.abs(b) + 1.0) <= 1e-3)
        gpu, self_ = agree(eb.astype(np.float64), oeb), agree(deb, oeb)
        diff = eb.astype(np.float64) - oeb
        print(f"{which}, {ivmax} evaluations: GPU vs FP32 oracle agree {gpu:.2f}, FP64 vs FP32 "
              f"oracle {self_:.2f}; median E difference {np.median(diff):.3g}; mean evaluations "
              f"{nf.mean():.1f} (oracle {onf.mean():.1f})")
        assert gpu >= self_ - slack
        assert abs(np.median(diff)) <= np.std(oeb) + 1e-6
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
