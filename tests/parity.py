"""Parity criteria used by the tests.

Per-bead quantities and indices are compared bit for bit.  Per-pose sums
(energy, force, torque gradient, mode gradients) are compared with a
condition-aware relative error:

    err = |x - y| / S,    S = sum over the terms of |term|   (oracle FP64)

S is the magnitude of the computation behind the component; two correct FP32
implementations that add the same per-bead values in different orders differ
by at most ~n*eps*S.  The bound asserted is err <= 1e-5.  The plain relative
error |x - y| / |y| is reported alongside where a test prints statistics.
"""
import numpy as np

TOL = 1e-5


def cond_err(x, y, scale):
    """Identical values (including equal infinities, which the reference's
    mode correction g + 4 k g^4 produces on clashing poses) count as zero
    error; an infinity on one side only is an infinite error."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    s = np.maximum(np.asarray(scale, np.float64), 1e-30)
    with np.errstate(invalid="ignore", over="ignore"):
        e = np.abs(x - y) / s
    e[x == y] = 0.0
    e[np.isnan(e)] = np.inf
    return e


def forcelim_safe(res64):
    """Poses whose force magnitude is not within 1e-4 (relative) of a
    reducePotForce clamp threshold (|F|^2 = 1e18 after 0, 1, 2 or 3 scalings)."""
    F = np.asarray(res64, np.float64)[:, 1:4]
    m2 = np.sum(F * F, axis=1)
    ok = np.ones(len(m2), bool)
    for j in range(4):
        r = m2 * (1e4 ** j) / 1e18
        ok &= np.abs(r - 1.0) > 1e-4
    return ok


def summarize(name, x, y, scale):
    e = cond_err(x, y, scale)
    fin = np.isfinite(np.asarray(y, np.float64))
    with np.errstate(invalid="ignore", over="ignore"):
        rel = np.abs(np.asarray(x, np.float64) - y)[fin] / np.maximum(np.abs(np.asarray(y)[fin]), 1e-30)
    return (f"{name}: cond err max {e.max():.2e} p99 {np.percentile(e, 99):.2e}; "
            f"plain rel err median {np.median(rel):.2e} p99 {np.percentile(rel, 99):.2e}")
