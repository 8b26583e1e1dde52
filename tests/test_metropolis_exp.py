"""The Metropolis test's exp (crates/cpm-core/src/monte_carlo.rs:158:
accept iff u < exp(-dH)) on the CPU: the double-double exp the kernels fall
back to near a tie is correctly rounded (checked against 60-digit decimal
arithmetic), and the decision rule built on it (metropolis.hpp, compiled
here for the host) decides every constructed near-tie as the exact value
of exp(-dH) does.  The reference's f64::exp is glibc's, which is itself not
correctly rounded on a small fraction of arguments; there a decision can
differ only if u equals the lower of the two neighbouring doubles."""
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np

from cpm import _native as N
from oracle import oracle as O

getcontext().prec = 60


def exact_rounded_exp_neg(x: float) -> float:
    """exp(-x) rounded to the nearest double, from 60-digit arithmetic."""
    return float(Fraction((-Decimal(float(x))).exp()))


def test_exp_cr_is_correctly_rounded():
    rng = np.random.default_rng(11)
    dh = np.concatenate([rng.uniform(0, 40, 3000), 10 ** rng.uniform(-14, 0, 1000),
                         rng.uniform(40, 708, 1000), [1e-300, 0.5, 1.0, np.log(2.0), 36.7]])
    _, cr, _ = N.metropolis_check(dh, device=False)
    want = np.array([exact_rounded_exp_neg(x) for x in dh])
    assert np.array_equal(cr, want)


def test_glibc_misrounds_where_exp_cr_and_glibc_differ():
    """Over 2e6 arguments exp_cr and glibc differ only where glibc is not
    correctly rounded (each such case checked with decimal arithmetic)."""
    rng = np.random.default_rng(12)
    dh = np.concatenate([rng.uniform(0, 40, 1_500_000), 10 ** rng.uniform(-12, 0, 500_000)])
    libm, cr, _ = N.metropolis_check(dh, device=False)
    glibc = O.exp_neg(dh)
    assert np.array_equal(libm, glibc)  # the host path uses the host libm
    diff = np.nonzero(cr != glibc)[0]
    assert 0 < len(diff) < 2e-3 * len(dh)
    for i in diff[:300]:
        assert cr[i] == exact_rounded_exp_neg(dh[i]) != glibc[i]
        # one ulp apart
        assert np.nextafter(cr[i], glibc[i]) == glibc[i]


def test_near_tie_decisions_follow_the_exact_value():
    """u placed on, just above and just below the libm value of exp(-dH):
    every decision equals u < RN(exp(-dH)), the comparison the reference
    makes wherever its libm rounds correctly."""
    rng = np.random.default_rng(13)
    dh = rng.uniform(0, 36, 20000)
    p = O.exp_neg(dh)
    for u in (p, np.nextafter(p, 2.0), np.nextafter(p, -1.0)):
        _, cr, acc = N.metropolis_check(dh, u, device=False)
        assert np.array_equal(acc, u < cr)
        for i in rng.choice(len(dh), 200, replace=False):
            assert bool(acc[i]) == (u[i] < exact_rounded_exp_neg(dh[i]))
        # where glibc rounds correctly, the decision is glibc's
        ok = cr == p
        assert np.array_equal(acc[ok], (u < p)[ok])


def test_far_from_ties_the_libm_value_decides():
    rng = np.random.default_rng(14)
    dh = rng.uniform(0, 30, 100000)
    u = rng.random(100000)
    libm, _, acc = N.metropolis_check(dh, u, device=False)
    assert np.array_equal(acc, u < libm)
