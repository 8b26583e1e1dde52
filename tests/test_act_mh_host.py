"""Metropolis-Hastings and the Act term on the host side of the product
(CPU only): configuration, hashes, state import, and the two pieces of
arithmetic the exact-replay kernels add for them, compiled for the host
from the same source (cpm-rust_amd/csrc/metropolis.hpp):

* the Act geometric-mean root (energy.rs:443, product.powf(1.0 / count as
  f64)), correctly rounded: checked against 60-digit decimal arithmetic; it
  differs from glibc's pow (the reference's and the oracle's) only where
  glibc is not correctly rounded, and not at all on the powers m^n, m <= 228;
* the Metropolis-Hastings decision (monte_carlo.rs:168-196): far from ties
  it is u < min((n_A / n_B) exp(-dH), 1) with the libm exp; at ties, the
  exact value decides."""
import re
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np
import pytest

import cpm
from cpm import _native as N
from oracle import oracle as O

getcontext().prec = 60


def _sim(acceptance="metropolis", act=False, proposal=None, lambda_act=(0.0, 0.0),
         max_act=(0, 0)):
    kw = {} if proposal is None else {"proposal": proposal}
    sim = cpm.CPM(grid=(24, 24), boundary="periodic", seed=9, acceptance=acceptance,
                  active_terms={"act": act}, **kw)
    for i, nm in enumerate(("a", "b")):
        sim.register_cell_type(nm, 20, 18, 2.0, 0.5, lambda_act=lambda_act[i],
                               max_act=max_act[i])
        sim.add_cells(nm, 4)
    sim.set_adhesion([[0.0, 8.0, 8.0], [8.0, 2.0, 5.0], [8.0, 5.0, 3.0]])
    sim.initialize(warn_on_default_init=False)
    return sim


def _cfg(acceptance="metropolis", act=False, lambda_act=(0.0, 0.0), max_act=(0, 0)):
    return O.Config(grid=(24, 24), boundary="periodic", seed=9,
                    cell_types=[O.CellTypeSpec(nm, 20, 18, 2.0, 0.5, lambda_act[i], max_act[i])
                                for i, nm in enumerate(("a", "b"))],
                    cell_counts=[4, 4], adhesion=[[0.0, 8.0, 8.0], [8.0, 2.0, 5.0],
                                                  [8.0, 5.0, 3.0]],
                    warn_on_default_init=False, acceptance=acceptance, act=act)


@pytest.mark.parametrize("acceptance,act", [("metropolis_hastings", False), ("metropolis", True),
                                            ("metropolis_hastings", True)])
def test_mh_and_act_configs_build_and_hash_like_the_oracle(acceptance, act):
    la, ma = (30.0, 20.0), (20, 40)
    rep = cpm.Replica(_sim(acceptance, act, lambda_act=la, max_act=ma))
    cfg = _cfg(acceptance, act, la, ma)
    assert rep.hashes() == (O.convention_hash(cfg), O.Model(cfg).model_hash())
    # config.rs:777-790: enabling act changes the convention hash
    assert O.convention_hash(cfg) != O.convention_hash(_cfg(acceptance, not act, la, ma))


def test_mh_with_edge_list_keeps_the_reference_message():
    """config.rs:394-411"""
    with pytest.raises(ValueError, match=re.escape(
            "proposal = EdgeList is not valid with acceptance = MetropolisHastings: the "
            "exact-Boltzmann validation's reversibility argument depends on the uniform "
            "proposal's forward/reverse symmetry, which EdgeList does not establish")):
        cpm.Replica(_sim("metropolis_hastings", proposal="edge_list"))
    cpm.Replica(_sim(proposal="edge_list"))


def test_birth_mcs_imports_and_reads_back():
    rep = cpm.Replica(_sim(act=True, lambda_act=(10.0, 10.0), max_act=(20, 20)))
    snap = rep.snapshot()
    assert not rep.birth_mcs().any()
    b = np.random.default_rng(1).integers(0, 1000, (24, 24)).astype(np.uint32)
    rep.set_state(snap["lattice"].ravel(), snap["type_index"], rng=snap["rng"], mcs=1000,
                  birth_mcs=b)
    assert np.array_equal(rep.birth_mcs(), b)
    # without the term births read as zeros and are not kept
    plain = cpm.Replica(_sim())
    plain.set_state(snap["lattice"].ravel(), snap["type_index"], birth_mcs=b)
    assert not plain.birth_mcs().any()


def exact_root(x: float, n: int) -> float:
    """x^(RN(1/n)) rounded to the nearest double, from 60-digit arithmetic."""
    if x == 0.0:
        return 0.0
    y = Decimal(1.0 / n)
    return float(Fraction((y * Decimal(x).ln()).exp()))


def _activity_products(rng, k):
    """Products of 1..7 integer activities in [0, M] multiplied in order,
    as the geometric mean forms them (energy.rs:426-441)."""
    xs, ns = [], []
    for _ in range(k):
        M = int(rng.choice([5, 10, 20, 37, 50, 100, 200, 1000, 100000]))
        n = int(rng.integers(1, 8))
        a = np.where(rng.random(n) < 0.4, M, rng.integers(0, M + 1, n))
        p = 1.0
        for v in a:
            p *= float(v)
        xs.append(p)
        ns.append(n)
    return np.array(xs), np.array(ns, dtype=np.uint32)


def test_act_root_is_correctly_rounded():
    x, n = _activity_products(np.random.default_rng(21), 6000)
    got = N.act_root_check(x, n, device=False)
    want = np.array([exact_root(a, int(b)) for a, b in zip(x, n)])
    assert np.array_equal(got, want)


def test_act_root_differs_from_glibc_only_where_glibc_misrounds():
    x, n = _activity_products(np.random.default_rng(22), 200_000)
    got = N.act_root_check(x, n, device=False)
    glibc = O.root(x, n)
    diff = np.nonzero(got != glibc)[0]
    print(f"[act root] {len(diff)} of {len(x)} differ from glibc pow")
    assert len(diff) < 5e-3 * len(x)
    for i in diff[:200]:
        assert got[i] == exact_root(x[i], int(n[i])) != glibc[i]
        assert np.nextafter(got[i], glibc[i]) == glibc[i]


def test_act_root_equals_glibc_on_every_power_up_to_228():
    m = np.repeat(np.arange(0, 229, dtype=np.float64), 7)
    n = np.tile(np.arange(1, 8, dtype=np.uint32), 229)
    x = np.array([mm ** int(nn) for mm, nn in zip(m, n)])
    got = N.act_root_check(x, n, device=False)
    assert np.array_equal(got, O.root(x, n))
    # the exact powers give the integer back
    assert np.array_equal(got[(n & (n - 1)) == 0], m[(n & (n - 1)) == 0])


def test_mh_decisions_far_from_ties_use_the_libm_exp():
    rng = np.random.default_rng(23)
    k = 200_000
    na = rng.integers(1, 5, k).astype(np.uint32)
    nb = rng.integers(1, 5, k).astype(np.uint32)
    dh = rng.uniform(-8.0, 30.0, k)
    u = rng.random(k)
    got = N.mh_check(na, nb, dh, u, device=False)
    want = u < np.minimum((na / nb) * O.exp_neg(dh), 1.0)
    assert np.array_equal(got, want)


def test_mh_near_ties_follow_the_exact_value():
    """u placed on, just above and just below the libm ratio: every decision
    equals u < RN(q * RN(exp(-dH))) with the exactly rounded exp."""
    rng = np.random.default_rng(24)
    k = 3000
    na = rng.integers(1, 5, k).astype(np.uint32)
    nb = rng.integers(1, 5, k).astype(np.uint32)
    dh = rng.uniform(-1.0, 30.0, k)
    q = na / nb
    ratio = np.minimum(q * O.exp_neg(dh), 1.0)
    for u in (ratio, np.nextafter(ratio, 2.0), np.nextafter(ratio, -1.0)):
        u = np.minimum(u, np.nextafter(1.0, 0.0))
        got = N.mh_check(na, nb, dh, u, device=False)
        for i in range(0, k, 10):
            e = float(Fraction((-Decimal(float(dh[i]))).exp()))
            assert bool(got[i]) == (u[i] < min(q[i] * e, 1.0)), i
