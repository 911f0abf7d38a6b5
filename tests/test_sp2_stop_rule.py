"""The SP2 stop rule of k_sp2_decide (pyseqm_amd/csrc/sp2.hip) in exact arithmetic.

SP2 on a diagonal matrix is SP2 on its eigenvalues, so float64 runs on eigenvalue lists show what
the rule does when rounding is negligible.  The rule: stop at iteration k when |tr X - N| < eps,
e_{k-2} < 3/16 and e_k >= 8 e_{k-2}^2, with e = tr X - tr X^2 = sum x (1 - x).  Its claim is that
without rounding it never fires before the iteration has converged (the kernel then fires at its
first iterate on the FP32 rounding floor).  Checked here on random spectra with gaps from 1e-3 to
1 eV, threefold-degenerate levels and loose Gershgorin-like bounds: at the stop every eigenvalue is
within 1e-12 of 0 or 1.  A check that nothing can fail pins nothing, so variants with a planted
mistake (bound 2 instead of 8, no 3/16 guard, no trace condition) must stop early on some of the
same spectra, and so must the reference's reduced-precision rule (trace error below 1e-2 and not
below its value two iterations earlier, SP2.py:74-78), whose trace error the occupied and virtual
deviations cancel in.
"""
import numpy as np

TEPS = 1e-2  # SP2.py:4, the reduced-precision trace threshold


def _spectrum(rng, degenerate):
    n = int(rng.integers(2, 300))
    N = int(rng.integers(1, n))
    gap = 10 ** rng.uniform(-3, 0)
    occ = np.sort(rng.uniform(rng.uniform(-60, -5), -gap / 2, N))
    vir = np.sort(rng.uniform(gap / 2, 20, n - N))
    if degenerate:
        occ = np.repeat(occ[:(N + 2) // 3], 3)[:N]
        vir = np.repeat(vir[:(n - N + 2) // 3], 3)[:n - N]
    w = np.concatenate([occ, vir])
    h1, hN = w.min() - rng.uniform(0, 5), w.max() + rng.uniform(0, 5)
    return (hN - w) / (hN - h1), N  # X0 = (hN I - F) / (hN - h1): occupied levels near 1


def _run(x, N, rule, maxit=200):
    """Trace-correcting SP2 on eigenvalues (the branch test of SP2.py:55-68); returns the
    eigenvalues at the iteration where `rule` stops."""
    e = [float(np.sum(x * (1 - x)))]
    etr = [abs(float(x.sum()) - N)]
    for k in range(1, maxit + 1):
        tr, fro = x.sum(), (x * x).sum()
        x = x * x if abs(fro - N) < abs(2 * tr - fro - N) else 2 * x - x * x
        e.append(abs(float(x.sum() - (x * x).sum())))
        etr.append(abs(float(x.sum()) - N))
        if k >= 3 and rule(k, e, etr):
            return x
    return x


def kernel_rule(C=8.0, guard=0.1875, trace=True):
    return lambda k, e, etr: (not trace or etr[k] < TEPS) and e[k - 2] < guard and e[k] >= C * e[k - 2] ** 2


def reference_rule(k, e, etr):
    return etr[k] < TEPS and etr[k] >= etr[k - 2]


def _early_stops(rule, trials, seed):
    rng = np.random.default_rng(seed)
    worst, early = 0.0, 0
    for t in range(trials):
        x, N = _spectrum(rng, degenerate=t % 3 == 0)
        xf = _run(x, N, rule)
        dev = float(np.minimum(xf, 1 - xf).max())
        worst = max(worst, dev)
        early += dev > 1e-6
    return worst, early


def test_rule_never_stops_before_convergence_in_exact_arithmetic():
    worst, early = _early_stops(kernel_rule(), 1500, seed=0)
    assert early == 0 and worst < 1e-12, (worst, early)


def test_planted_mistakes_stop_early():
    for name, rule in [("bound 2", kernel_rule(C=2.0)), ("no guard", kernel_rule(guard=10.0)),
                       ("no trace condition", kernel_rule(trace=False)), ("reference trace rule", reference_rule)]:
        _, early = _early_stops(rule, 500, seed=1)
        assert early > 0, name
