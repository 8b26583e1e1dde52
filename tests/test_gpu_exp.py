"""The device's Metropolis exp against the host's glibc (the reference's
f64::exp, crates/cpm-core/src/monte_carlo.rs:158) over 1e8 arguments, and
the exact-replay kernels' decisions on every uphill Metropolis draw the
oracle makes in a set of parity runs.

Reported: how often the device libm exp differs from glibc, how often the
correctly rounded exp (the kernels' near-tie fallback) differs from glibc
(only where glibc itself is not correctly rounded), and the number of
decisions that differ (must be 0)."""
import numpy as np
import pytest

from cpm import _native as N
from oracle import oracle as O
from tests.helpers import quickstart_oracle, sorting_pair

pytestmark = pytest.mark.gpu

CHUNK = 10_000_000


def _chunks():
    rng = np.random.default_rng(2027)
    for _ in range(6):
        yield "uniform(0,40]", 40.0 - rng.uniform(0.0, 40.0, CHUNK)
    for _ in range(2):
        yield "loguniform[1e-12,1)", 10.0 ** rng.uniform(-12.0, 0.0, CHUNK)
    for _ in range(2):
        yield "uniform(0,745]", 745.0 - rng.uniform(0.0, 745.0, CHUNK)


def test_device_exp_against_glibc_over_1e8_arguments():
    total = libm_diff = cr_diff = cr_host_diff = 0
    for name, dh in _chunks():
        dev_libm, dev_cr, _ = N.metropolis_check(dh)
        glibc = O.exp_neg(dh)
        _, host_cr, _ = N.metropolis_check(dh, device=False)
        normal = glibc >= 2.0 ** -1022
        total += dh.size
        libm_diff += int((dev_libm != glibc).sum())
        cr_diff += int((dev_cr != glibc)[normal].sum())
        # the device's double-double arithmetic is the host's, bit for bit
        cr_host_diff += int((dev_cr != host_cr)[normal].sum())
        print(f"[exp] {name}: device libm != glibc {int((dev_libm != glibc).sum())}, "
              f"correctly rounded != glibc {int((dev_cr != glibc)[normal].sum())}", flush=True)
    print(f"[exp] total {total} arguments: device libm != glibc {libm_diff} "
          f"({libm_diff / total:.2e}); correctly rounded != glibc {cr_diff} "
          f"({cr_diff / total:.2e}); device vs host correctly rounded {cr_host_diff}")
    assert total >= 100_000_000
    assert cr_host_diff == 0


def _harvest():
    """(dH, u) of every uphill Metropolis test in a set of parity runs."""
    runs = []
    m = O.Model(quickstart_oracle(O, seed=1))
    m.record_metropolis(40_000_000)
    m.run_mcs(10_000)
    runs.append(("quickstart 30x30, 10000 MCS", m.metropolis_draws()))
    for seed, kw, mcs in [(3, dict(grid=(48, 48), n_types=3), 400),
                          (4, dict(grid=(64, 64), n_types=2, boundary="fixed"), 300),
                          (5, dict(grid=(20, 20, 20), n_types=2, tv=40, ti=60, phi=0.5), 60)]:
        m = O.Model(_oracle_cfg(seed, kw))
        m.record_metropolis(40_000_000)
        m.run_mcs(mcs)
        runs.append((f"sorting {kw['grid']}, {mcs} MCS", m.metropolis_draws()))
    return runs


def _oracle_cfg(seed, kw):
    import cpm
    return sorting_pair(cpm, O, seed, **kw)[1]


def test_kernel_decisions_on_harvested_draws_equal_the_oracles():
    total = 0
    for name, draws in _harvest():
        dh, u = draws[:, 0], draws[:, 1]
        glibc = O.exp_neg(dh)
        _, _, acc = N.metropolis_check(dh, u)
        differ = int((acc != (u < glibc)).sum())
        near = int((np.abs(u - glibc) <= glibc * 2.0 ** -40).sum())
        total += len(dh)
        print(f"[exp] {name}: {len(dh)} uphill decisions, {near} near-ties, "
              f"{differ} differ from the oracle", flush=True)
        assert differ == 0
    assert total > 1_000_000
