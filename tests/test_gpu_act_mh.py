"""Metropolis-Hastings acceptance and the Act term on the GPU (exact replay,
lane path) against the CPU oracle.  Bar: bit-exact — lattice, bookkeeping,
Act births, FP64 energy bits, RNG state — on 2D and 3D, periodic and
fixed lattices, alone and together, in batches mixed with wave-path
models, through state import and through run_with; and the device's Act
root and MH decisions equal the host build of the same code (whose
rounding tests/test_act_mh_host.py pins)."""
import struct

import numpy as np
import pytest

import cpm
from cpm import _native as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

KEYS = ["lattice", "volume", "interface", "sum_position", "sum_outer", "contact", "rng", "mcs"]


def pair(seed, grid, boundary="periodic", acceptance="metropolis", act=False, n_types=2,
         tv=30, ti=24, phi=0.6, lam_act=(40.0, 25.0, 60.0), max_act=(20, 30, 12), energy=None):
    """The same configuration for the product and the oracle (J drawn from
    default_rng(seed))."""
    rng = np.random.default_rng(seed)
    T = n_types
    J = np.zeros((T + 1, T + 1))
    J[0, 1:] = J[1:, 0] = 16.0
    for a in range(1, T + 1):
        for b in range(a, T + 1):
            J[a, b] = J[b, a] = float(rng.uniform(2, 14))
    n = max(T, round(phi * int(np.prod(grid)) / tv))
    counts = [n // T + (1 if i < n % T else 0) for i in range(T)]
    names = [f"t{i}" for i in range(T)]
    sim = cpm.CPM(grid=grid, boundary=boundary, seed=seed, energy_neighborhood=energy,
                  acceptance=acceptance, active_terms={"act": act})
    for i, nm in enumerate(names):
        sim.register_cell_type(nm, target_volume=tv, target_interface=ti, lambda_volume=5.0,
                               lambda_interface=1.0, lambda_act=lam_act[i], max_act=max_act[i])
    for nm, c in zip(names, counts):
        sim.add_cells(nm, c)
    sim.set_adhesion(J.tolist())
    sim.initialize(warn_on_default_init=False)
    cfg = O.Config(grid=grid, boundary=boundary, seed=seed,
                   cell_types=[O.CellTypeSpec(nm, tv, ti, 5.0, 1.0, lam_act[i], max_act[i])
                               for i, nm in enumerate(names)],
                   cell_counts=counts, adhesion=J.tolist(), energy_neighborhood=energy,
                   warn_on_default_init=False, acceptance=acceptance, act=act)
    return sim, cfg


def assert_same(rep, ref, label):
    dev, want = rep.snapshot(), ref.snapshot()
    for k in KEYS:
        assert np.array_equal(np.asarray(dev[k]), np.asarray(want[k])), f"{label}: {k} differs"
    assert struct.pack("<d", dev["energy"]) == struct.pack("<d", want["energy"]), (
        f"{label}: energy {dev['energy']!r} vs {want['energy']!r}")
    assert np.array_equal(rep.birth_mcs(), ref.birth_mcs()), f"{label}: births differ"


CASES = [
    # label, pair kwargs, MCS
    ("mh_2d_periodic", dict(seed=31, grid=(48, 48), acceptance="metropolis_hastings"), 60),
    ("mh_2d_fixed", dict(seed=32, grid=(40, 36), boundary="fixed",
                         acceptance="metropolis_hastings", n_types=3), 50),
    ("mh_3d", dict(seed=33, grid=(14, 14, 14), acceptance="metropolis_hastings", tv=60, ti=90,
                   phi=0.4), 8),
    ("act_2d_periodic", dict(seed=34, grid=(48, 48), act=True), 60),
    ("act_2d_fixed", dict(seed=35, grid=(36, 44), boundary="fixed", act=True, n_types=3), 50),
    ("act_3d_fixed", dict(seed=36, grid=(12, 14, 12), boundary="fixed", act=True, tv=60,
                          ti=90, phi=0.4), 8),
    ("act_3d_periodic_26", dict(seed=23, grid=(12, 12, 12), act=True, tv=60, ti=120, phi=0.4,
                                energy="twenty_six"), 8),
    ("mh_act_2d", dict(seed=38, grid=(40, 40), acceptance="metropolis_hastings", act=True), 60),
    # strong Act: lambda_Act 300 (Max_Act 20), the protrusive regime
    ("act_strong", dict(seed=39, grid=(48, 48), act=True, lam_act=(300.0, 300.0, 0.0)), 80),
]


@pytest.mark.parametrize("label,kw,mcs", CASES, ids=[c[0] for c in CASES])
def test_exact_replay_parity(label, kw, mcs):
    sim, cfg = pair(**kw)
    print(f"[act/mh] {label}: grid {cfg.grid}, {mcs} MCS", flush=True)
    rep = cpm.Replica(sim)
    ref = O.Model(cfg)
    assert_same(rep, ref, label + " (init)")
    # two calls: the second starts from a device-resident state mid-run
    rep.run_mcs(mcs // 2)
    ref.run_mcs(mcs // 2)
    assert_same(rep, ref, label + " (half)")
    cnt = rep.run_mcs(mcs - mcs // 2)
    ref.run_mcs(mcs - mcs // 2)
    assert_same(rep, ref, label)
    assert cnt["accepted"] > 0


def test_nan_act_prices_follow_the_reference():
    """A type with lambda_Act = 0 and Max_Act = 0 in an Act model prices
    every copy it gains at -(0/0) * 0 = NaN: Metropolis then rejects it
    (NaN <= 0 is false, u < exp(NaN) is false) and Metropolis-Hastings
    accepts it (min(NaN, 1) = 1).  Both as the oracle does."""
    for acc in ("metropolis", "metropolis_hastings"):
        sim, cfg = pair(seed=41, grid=(32, 32), act=True, acceptance=acc,
                        lam_act=(0.0, 30.0), max_act=(0, 20))
        rep, ref = cpm.Replica(sim), O.Model(cfg)
        rep.run_mcs(30)
        ref.run_mcs(30)
        assert_same(rep, ref, f"nan {acc}")


def test_mixed_batch_with_wave_path_models():
    """One run_mcs call over Act / MH models (lane path) and plain
    Metropolis models (wave path): each equals its oracle."""
    kws = [dict(seed=51, grid=(32, 32), act=True), dict(seed=52, grid=(40, 40)),
           dict(seed=53, grid=(36, 36), acceptance="metropolis_hastings"),
           dict(seed=54, grid=(24, 40), boundary="fixed", n_types=3),
           dict(seed=55, grid=(32, 32), acceptance="metropolis_hastings", act=True)]
    pairs = [pair(**kw) for kw in kws]
    reps = [cpm.Replica(s) for s, _ in pairs]
    refs = [O.Model(c) for _, c in pairs]
    cpm.Replica.run_mcs_many(reps, 25)
    for i, (rep, ref) in enumerate(zip(reps, refs)):
        ref.run_mcs(25)
        assert_same(rep, ref, f"batch member {i}")


def test_state_import_with_births_continues_the_oracle():
    sim, cfg = pair(seed=61, grid=(40, 40), act=True, acceptance="metropolis_hastings")
    ref = O.Model(cfg)
    ref.run_mcs(33)
    s = ref.snapshot()
    rep = cpm.Replica(sim, initialize=False)
    rep.set_state(s["lattice"].ravel(), ref.type_index(), volume=s["volume"],
                  interface=s["interface"], sum_position=s["sum_position"],
                  sum_outer=s["sum_outer"], contact=s["contact"], energy=s["energy"],
                  rng=s["rng"], mcs=33, birth_mcs=s["birth_mcs"])
    rep.run_mcs(40)
    ref.run_mcs(40)
    assert_same(rep, ref, "imported")


def test_run_with_samples_equal_the_oracles():
    sim, cfg = pair(seed=71, grid=(40, 40), act=True)
    cfg.burn_in_mcs, cfg.readout_mcs, cfg.sampling_interval_mcs = 20, 40, 10
    got = sim.run(burn_in_mcs=20, readout_mcs=40, sampling_interval_mcs=10)
    want = O.run_with(O.Model(cfg))
    assert len(got.samples) == len(want.samples) == 4
    for a, b in zip(got.samples, want.samples):
        assert a.mcs == b.mcs
        assert np.array_equal(a.volume, b.volume)
        assert np.array_equal(a.interface, b.interface)
        assert np.array_equal(a.centroid, b.centroid)
        assert a.conservative_energy == b.energy
    assert got.status.kind == want.status[0]


def test_sublattice_mode_refuses_mh_and_act():
    for kw in (dict(act=True), dict(acceptance="metropolis_hastings")):
        sim, _ = pair(seed=81, grid=(32, 32), **kw)
        rep = cpm.Replica(sim)
        with pytest.raises(ValueError, match="sublattice mode runs the Metropolis rule"):
            rep.run_mcs(2, mode="sublattice")


def test_device_act_root_equals_the_host_build():
    rng = np.random.default_rng(91)
    k = 2_000_000
    M = rng.choice([5, 10, 20, 37, 50, 100, 200, 1000, 100000], k)
    n = rng.integers(1, 8, k).astype(np.uint32)
    x = np.ones(k)
    for j in range(7):
        a = np.where(rng.random(k) < 0.4, M, rng.integers(0, M + 1))
        x = np.where(j < n, x * a, x)
    dev = N.act_root_check(x, n)
    host = N.act_root_check(x, n, device=False)
    glibc = O.root(x, n)
    print(f"[act root] device vs host differ {int((dev != host).sum())}; "
          f"device vs glibc differ {int((dev != glibc).sum())} of {k}")
    assert np.array_equal(dev, host)


def test_device_mh_decisions_equal_the_host_build():
    rng = np.random.default_rng(92)
    k = 1_000_000
    na = rng.integers(1, 7, k).astype(np.uint32)
    nb = rng.integers(1, 7, k).astype(np.uint32)
    dh = rng.uniform(-10.0, 40.0, k)
    u = rng.random(k)
    # a third of the draws on or next to the libm ratio
    ratio = np.minimum((na / nb) * O.exp_neg(dh), 1.0)
    tie = rng.random(k) < 0.33
    near = np.minimum(np.nextafter(ratio, rng.choice([-1.0, 2.0], k)), np.nextafter(1.0, 0.0))
    u = np.where(tie, near, u)
    assert np.array_equal(N.mh_check(na, nb, dh, u), N.mh_check(na, nb, dh, u, device=False))
