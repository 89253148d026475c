"""The inside test's fast edge decision (dr_traverse.h edge_fast) never contradicts the exact
one (edge_exact: correctly rounded division and square root, Triangle.impl.h:299-301).

Emulated in float32 with numpy: the hardware reciprocal square root is modelled as the exact
value perturbed by up to 2 ulp either way (its specified error is 1 ulp); the fast decision must
equal the exact one wherever it decides (1 or 0), on random edges and on edges placed at the
threshold, where it must defer (2) instead."""
import numpy as np

TOL = np.float32(32 * 0.000005)
MARGIN = np.float32(4e-6)


def exact(D, S, thr):
    return (D / np.sqrt(S) + TOL) >= thr


def fast(D, S, thr, ulps):
    rs = (np.float32(1.0) / np.sqrt(S.astype(np.float64))).astype(np.float32)
    rs = rs + ulps.astype(np.float32) * np.spacing(rs)          # up to +-2 ulp of error
    q = D * rs
    r = q + TOL
    m = MARGIN * (np.abs(q) + TOL)
    out = np.full(D.shape, 2, np.int8)
    out[(r - m) >= thr] = 1
    out[(r + m) < thr] = 0
    out[~((S >= np.float32(1e-30)) & (S <= np.float32(1e30)))] = 2
    return out


def test_fast_edge_decision_agrees_with_exact():
    rng = np.random.default_rng(0)
    n = 2_000_000
    with np.errstate(all="ignore"):
        D = rng.normal(0, 1, n).astype(np.float32) * np.float32(10) ** rng.integers(-6, 6, n).astype(np.float32)
        S = np.abs(rng.normal(0, 1, n)).astype(np.float32) * np.float32(10) ** rng.integers(-12, 12, n).astype(np.float32)
        thr = rng.normal(0, 1, n).astype(np.float32)
        # half the edges sit on the threshold: thr = the exact left side, nudged by a few ulp
        near = rng.random(n) < 0.5
        lhs = D / np.sqrt(S) + TOL
        thr[near] = lhs[near] + rng.integers(-40, 41, near.sum()).astype(np.float32) * np.spacing(lhs[near])
        ulps = rng.integers(-2, 3, n)
        e = exact(D, S, thr)
        f = fast(D, S, thr, ulps)
    decided = f != 2
    assert decided.mean() > 0.4                       # it decides most edges away from the threshold
    assert np.array_equal(f[decided] == 1, e[decided])
    # close to the threshold it defers
    close = near & (np.abs(thr - lhs) <= 4 * np.spacing(np.abs(lhs) + TOL))
    assert close.sum() > 1000 and np.all(f[close] == 2)
