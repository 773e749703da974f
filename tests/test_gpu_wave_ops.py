"""The adjoint pair loop's wave primitives on the device (needs a GPU), against
brute force on the host: segmented run sums (DPP row shifts, then the row
heads by readlane) and the per-lane count of slot ends (a 64-bit mask built
with DPP row rotations).  Forces are small integers, so every sum is exact in
any order and the comparison is exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from gpuattract_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def _runs(rng, kind):
    if kind == "one run":
        lengths = [64]
    elif kind == "single lanes":
        lengths = [1] * 64
    elif kind == "row edges":
        lengths = [16, 15, 2, 16, 15]
    else:
        lengths = []
        while sum(lengths) < 64:
            lengths.append(int(rng.choice([1, 2, 3, 7, 16, 17, 33, 80])))
    seg, k = [], int(rng.integers(0, 100))
    for n in lengths:
        seg += [k] * n
        k += int(rng.integers(1, 3))
    seg = np.array(seg[:64], np.uint32)
    if kind == "ragged tail":
        seg[40:] = 0xFFFFFFFF  # lanes past the last pair
    return seg


@pytest.mark.parametrize("kind", ["one run", "single lanes", "row edges", "random", "ragged tail"])
def test_segmented_run_sums(ctx, kind):
    rng = np.random.default_rng(abs(hash(kind)) % 2**32)
    for _ in range(20):
        seg = _runs(rng, kind)
        F = rng.integers(-50, 50, size=(3, 64)).astype(np.float32)
        ends = np.full(64, 0xFFFFFFFF, np.uint32)
        sums, _, _ = ctx.wave_ops(F, seg, ends, 0)
        for lane in range(64):
            if lane == 0 or seg[lane] != seg[lane - 1]:
                run = seg == seg[lane]
                assert np.array_equal(sums[:, lane], F[:, run].sum(axis=1)), (kind, lane)


def test_slot_end_counts(ctx):
    rng = np.random.default_rng(7)
    for trial in range(200):
        B = int(rng.integers(0, 1 << 20))
        nslots = int(rng.integers(1, 65))
        gaps = rng.choice([1, 1, 2, 5, 40, 100], size=nslots)
        ends = (B + np.cumsum(gaps)).astype(np.uint64)
        e = np.full(64, 0xFFFFFFFF, np.uint64)
        e[:nslots] = ends
        _, counts, starts = ctx.wave_ops(np.zeros((3, 64), np.float32), np.zeros(64, np.uint32),
                                         e.astype(np.uint32), B)
        want = [int(np.sum(e <= B + lane)) for lane in range(64)]
        assert np.array_equal(counts, np.array(want, np.uint32)), trial
        mask = 0
        for v in e:
            if B < v < B + 64:
                mask |= 1 << int(v - B)
        assert starts == mask, trial
