"""The EdgeList proposal on the GPU (exact replay, lane path) against the CPU
oracle: bit-exact lattice, bookkeeping, energy bits, RNG state and edge
list order, on sparse and dense, 2D and 3D, periodic and fixed lattices,
with the Act term, in a batch beside wave-path models, through state
import; the notebook calibration run (notebooks/3d_organoid_demo.ipynb
cell 3) through run_with on the device; and the device's Binomial sampler
against the oracle's, draw for draw."""
import struct

import numpy as np
import pytest

import cpm
from cpm import _native as N
from oracle import oracle as O
from tests.test_gpu_act_mh import pair as mh_pair

pytestmark = pytest.mark.gpu

KEYS = ["lattice", "volume", "interface", "sum_position", "sum_outer", "contact", "rng", "mcs"]


def pair(seed, grid, boundary="periodic", phi=0.15, tv=30, ti=60, act=False, energy=None,
         n_types=2):
    rng = np.random.default_rng(seed)
    T = n_types
    J = np.zeros((T + 1, T + 1))
    J[0, 1:] = J[1:, 0] = 3.0
    for a in range(1, T + 1):
        for b in range(a, T + 1):
            J[a, b] = J[b, a] = float(rng.uniform(1, 8))
    n = max(T, round(phi * int(np.prod(grid)) / tv))
    counts = [n // T + (1 if i < n % T else 0) for i in range(T)]
    names = [f"t{i}" for i in range(T)]
    sim = cpm.CPM(grid=grid, boundary=boundary, seed=seed, proposal="edge_list",
                  energy_neighborhood=energy, active_terms={"act": act})
    for nm in names:
        sim.register_cell_type(nm, tv, ti, 2.0, 0.5, lambda_act=30.0, max_act=20)
    for nm, c in zip(names, counts):
        sim.add_cells(nm, c)
    sim.set_adhesion(J.tolist())
    sim.initialize(warn_on_default_init=False)
    cfg = O.Config(grid=grid, boundary=boundary, seed=seed,
                   cell_types=[O.CellTypeSpec(nm, tv, ti, 2.0, 0.5, 30.0, 20) for nm in names],
                   cell_counts=counts, adhesion=J.tolist(), energy_neighborhood=energy,
                   warn_on_default_init=False, proposal="edge_list", act=act)
    return sim, cfg


def assert_same(rep, ref, label):
    dev, want = rep.snapshot(), ref.snapshot()
    for k in KEYS:
        assert np.array_equal(np.asarray(dev[k]), np.asarray(want[k])), f"{label}: {k} differs"
    assert struct.pack("<d", dev["energy"]) == struct.pack("<d", want["energy"]), label
    assert np.array_equal(rep.edge_sites(), ref.edge_sites()), f"{label}: edge list differs"
    assert np.array_equal(rep.birth_mcs(), ref.birth_mcs()), f"{label}: births differ"


CASES = [
    ("sparse_2d_periodic", dict(seed=101, grid=(64, 64), phi=0.1), 200),
    ("dense_2d_fixed", dict(seed=102, grid=(40, 48), boundary="fixed", phi=0.7, n_types=3), 60),
    ("sparse_3d_fixed", dict(seed=103, grid=(24, 24, 24), boundary="fixed", phi=0.08, tv=60,
                             ti=90), 40),
    ("3d_periodic_26", dict(seed=104, grid=(16, 16, 16), phi=0.3, tv=60, ti=120,
                            energy="twenty_six"), 20),
    ("act_2d", dict(seed=105, grid=(48, 48), phi=0.3, act=True), 80),
    # a lone small cell: n p < 10, the inversion branch of the sampler
    ("binv_2d", dict(seed=106, grid=(60, 60), phi=0.001, tv=3, ti=8, n_types=1), 300),
]


@pytest.mark.parametrize("label,kw,mcs", CASES, ids=[c[0] for c in CASES])
def test_exact_replay_parity(label, kw, mcs):
    sim, cfg = pair(**kw)
    print(f"[edge list] {label}: grid {cfg.grid}, {mcs} MCS", flush=True)
    rep, ref = cpm.Replica(sim), O.Model(cfg)
    assert_same(rep, ref, label + " (init)")
    rep.run_mcs(mcs // 3)
    ref.run_mcs(mcs // 3)
    assert_same(rep, ref, label + " (first call)")
    rep.run_mcs(mcs - mcs // 3)
    ref.run_mcs(mcs - mcs // 3)
    assert_same(rep, ref, label)


def test_batch_beside_wave_path_and_mh_models():
    pairs = [pair(seed=111, grid=(40, 40)), mh_pair(seed=112, grid=(36, 36)),
             pair(seed=113, grid=(20, 20, 20), boundary="fixed", tv=60, ti=90, phi=0.1),
             mh_pair(seed=114, grid=(32, 32), acceptance="metropolis_hastings")]
    reps = [cpm.Replica(s) for s, _ in pairs]
    refs = [O.Model(c) for _, c in pairs]
    cpm.Replica.run_mcs_many(reps, 30)
    for i, (rep, ref) in enumerate(zip(reps, refs)):
        ref.run_mcs(30)
        assert_same(rep, ref, f"batch member {i}")


def test_state_import_with_edge_order_continues_the_oracle():
    sim, cfg = pair(seed=121, grid=(48, 48), phi=0.2)
    ref = O.Model(cfg)
    ref.run_mcs(70)
    s = ref.snapshot()
    rep = cpm.Replica(sim, initialize=False)
    rep.set_state(s["lattice"].ravel(), ref.type_index(), volume=s["volume"],
                  interface=s["interface"], sum_position=s["sum_position"],
                  sum_outer=s["sum_outer"], contact=s["contact"], energy=s["energy"],
                  rng=s["rng"], mcs=70, edge_sites=ref.edge_sites())
    rep.run_mcs(50)
    ref.run_mcs(50)
    assert_same(rep, ref, "imported")


def test_notebook_calibration_on_the_device():
    """notebooks/3d_organoid_demo.ipynb cell 3 through run_with on the
    device: "measured interface: mean=2496.8 std=1.33"."""
    sim = cpm.CPM(grid=(56, 56, 56), boundary="fixed", seed=7, energy_neighborhood="eighteen",
                  proposal="edge_list")
    sim.register_cell_type("probe", 900, 0, 4.0, 0.0)
    sim.add_cells("probe", 1)
    sim.set_adhesion([[0.0, 2.5], [2.5, 0.0]])
    sim.initialize(warn_on_default_init=False)
    r = sim.run(burn_in_mcs=20000, readout_mcs=20000, sampling_interval_mcs=2000)
    itf = np.array([s.interface[0] for s in r.samples], dtype=np.float64)
    assert f"mean={itf.mean():.1f} std={itf.std():.2f}" == "mean=2496.8 std=1.33"


def test_device_binomial_equals_the_oracle():
    rng = np.random.default_rng(131)
    count = 400_000
    n = rng.choice([3, 50, 900, 10000, 262144, 175616, 2 ** 24], count).astype(np.uint64)
    p = np.where(rng.random(count) < 0.5, rng.random(count) * 0.02, rng.random(count))
    p[:100] = 0.0
    p[100:200] = 1.0
    states = rng.integers(1, 2 ** 63, (count, 4), dtype=np.uint64)
    k_dev, st_dev = N.binomial_check(states, n, p)
    k_host, st_host = N.binomial_check(states, n, p, device=False)
    assert np.array_equal(k_dev, k_host)
    assert np.array_equal(st_dev, st_host)
    for i in range(0, count, 97):
        r = O.Rng(1)
        r.s[:] = states[i]
        assert r.binomial(int(n[i]), float(p[i])) == int(k_dev[i]), i
        assert np.array_equal(r.s, st_dev[i]), i


def test_sublattice_mode_runs_edge_list_models_and_rebuilds_the_list():
    """The sublattice mode takes an EdgeList model with its own (uniform)
    proposal, which the reference defines as the same chain with its null
    attempts kept (config.rs:42-70); after the call the edge list is the
    rebuild of the new lattice, in EdgeList::build's order."""
    sim, cfg = pair(seed=141, grid=(24, 24, 24), boundary="fixed", phi=0.1, tv=60, ti=90)
    rep = cpm.Replica(sim)
    rep.run_mcs(12, mode="sublattice")
    snap = rep.snapshot()
    ref = O.Model(cfg)
    ref.load_state(snap["lattice"].ravel(), snap["type_index"])
    assert np.array_equal(rep.edge_sites(), ref.edge_sites())
    itf, sp, so, ct = ref.recompute()
    assert np.array_equal(snap["volume"], ref.volume())
    assert np.array_equal(snap["interface"], itf)
    assert np.array_equal(snap["contact"], ct)
    assert abs(snap["energy"] - ref.global_energy()) <= 1e-9 * abs(ref.global_energy())
    # exact replay carries on from the rebuilt list, keeping it consistent
    rep.run_mcs(5)
    s2 = rep.snapshot()
    assert s2["mcs"] == 17
    ref.load_state(s2["lattice"].ravel(), s2["type_index"])
    assert np.array_equal(np.sort(rep.edge_sites()), np.sort(ref.edge_sites()))


def test_sublattice_edge_list_ks_against_exact_edge_list():
    """256 seeds of a sparse 32^3 fixed box (three types, phi = 0.2), every
    replica first 20 MCS of exact EdgeList replay, then 40 MCS either of
    exact EdgeList replay (lane path) or of sublattice mode: every summary
    statistic passes a two-sample KS test at p > 0.01."""
    from tests.test_gpu_sublattice import SEEDS, _ks, organoid3d_sim, trajectory_stats

    seeds = [cpm.derive_seed_u64(2031, i) for i in range(SEEDS)]
    out = {}
    for mode in ("exact_replay", "sublattice"):
        reps = [cpm.Replica(organoid3d_sim(sd, phi=0.2, proposal="edge_list")) for sd in seeds]
        cpm.Replica.run_mcs_many(reps, 20)
        first = [r.sample() for r in reps]
        cpm.Replica.run_mcs_many(reps, 40, mode=mode)
        out[mode] = []
        for r, f in zip(reps, first):
            last = r.sample()
            last.samples.insert(0, f.samples[0])
            out[mode].append(trajectory_stats(last, 50.0, box=(32, 32, 32)))
    _ks(out["exact_replay"], out["sublattice"], "32^3 fixed EdgeList, 20 exact + 40 MCS")
