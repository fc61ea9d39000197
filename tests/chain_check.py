"""Comparison of two Metropolis chains run on the same move list, one of them
the reference (FP64, or the plain chain).

Two chains of exact arithmetic take identical decisions.  A mixed-precision
chain may differ from its reference only where the reference's decision
variable lies within the arithmetic's energy error of the threshold: the
first differing decision must be such a near tie, every energy before it
must agree to that error bound, and after it the chains legitimately follow
different states, so nothing later is compared."""
import numpy as np

# Mixed precision (FP32 bulk, FP64 phases and sums) against FP64 on the same
# state: dE within 7.5e-7 at N = 1e5 (|E| ~ 1.3e3) and within 2e-6 on the
# test systems; chain tests assert 2e-5.
MIXED_DE_BOUND = 2e-5


def thresholds(u, beta=1.0):
    """The decision threshold -ln(u) / beta of each move (dE < threshold accepts)."""
    u = np.asarray(u, dtype=float)
    with np.errstate(divide="ignore"):
        return np.where(u > 0, -np.log(u) / beta, np.inf)


def assert_same_chain(dE_ref, acc_ref, dE, acc, thr, tol, score_ref=None):
    """score_ref: the reference's decision variable (dE - H / beta for a
    learned proposal; dE by default).  Returns the index of the first
    differing decision (len(acc) when there is none)."""
    dE_ref, dE = np.asarray(dE_ref, dtype=float), np.asarray(dE, dtype=float)
    acc_ref, acc = np.asarray(acc_ref), np.asarray(acc)
    s = dE_ref if score_ref is None else np.asarray(score_ref, dtype=float)
    diff = np.nonzero(acc_ref != acc)[0]
    k = int(diff[0]) if diff.size else len(acc)
    a, b = dE_ref[:k], dE[:k]
    fin = np.isfinite(a)
    assert (np.isfinite(b) == fin).all(), "hard-sphere overlaps differ before the first differing decision"
    if fin.any():
        err = np.abs(a[fin] - b[fin]).max()
        assert err < tol, f"dE differs by {err} (bound {tol}) before move {k}"
    if k < len(acc):
        assert np.isfinite(s[k]) and abs(s[k] - thr[k]) < tol, \
            f"decision {k} differs away from the threshold: score {s[k]}, threshold {thr[k]}, bound {tol}"
    return k
