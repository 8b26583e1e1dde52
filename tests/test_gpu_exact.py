"""Exact-replay parity on the GPU: the HIP engine against the CPU oracle and
the reference's own frozen values.  Bar: bit-exact (integer lattice and
bookkeeping, FP64 energy bits, RNG state)."""
import hashlib
import struct

import numpy as np
import pytest

import cpm
from oracle import oracle as O
from tests.helpers import (PINS, frozen_oracle_config, frozen_sim, quickstart,
                           quickstart_oracle, sorting_pair)

pytestmark = pytest.mark.gpu

KEYS = ["lattice", "volume", "interface", "sum_position", "sum_outer", "contact", "rng", "mcs"]


def assert_same_state(dev: dict, ref: dict, label: str = ""):
    for k in KEYS:
        a, b = np.asarray(dev[k]), np.asarray(ref[k])
        assert a.shape == b.shape, f"{label} {k}: shape {a.shape} vs {b.shape}"
        assert np.array_equal(a, b), f"{label} {k} differs"
    assert struct.pack("<d", dev["energy"]) == struct.pack("<d", ref["energy"]), (
        f"{label} energy {dev['energy']!r} vs {ref['energy']!r}")


def test_frozen_trajectory_through_the_engine():
    f = PINS["frozen_trajectory"]
    rep = cpm.Replica(frozen_sim(cpm), *f["windows"])
    rep.run_mcs(f["mcs"])
    snap = rep.snapshot()
    assert snap["mcs"] == 500
    assert snap["volume"].tolist() == f["volume"]
    assert snap["interface"].tolist() == f["interface"]
    assert struct.unpack("<Q", struct.pack("<d", snap["energy"]))[0] == \
        f["conservative_energy_bits"]
    cfg = frozen_oracle_config(O)
    assert O.model_hash(cfg, snap["lattice"].ravel(), snap["type_index"]) == f["model_hash"]


def test_frozen_trajectory_through_python_run():
    """tests/python/test_reproducibility.py:72-83"""
    f = PINS["frozen_trajectory"]
    result = frozen_sim(cpm).run(burn_in_mcs=0, readout_mcs=500, sampling_interval_mcs=500)
    assert len(result.samples) == 1
    s = result.samples[0]
    assert s.mcs == 500
    assert list(s.ids) == [1, 2]
    assert list(s.volume) == f["volume"]
    assert list(s.interface) == f["interface"]
    assert struct.unpack("<Q", struct.pack("<d", s.conservative_energy))[0] == \
        f["conservative_energy_bits"]


def test_different_seed_diverges():
    a = frozen_sim(cpm).run(burn_in_mcs=0, readout_mcs=500, sampling_interval_mcs=500)
    b = frozen_sim(cpm, seed=1).run(burn_in_mcs=0, readout_mcs=500, sampling_interval_mcs=500)
    assert a.samples[0].conservative_energy != b.samples[0].conservative_energy


CASES = [
    # (label, builder args, mcs)
    ("quickstart30", dict(kind="quick", grid=(30, 30), seed=1), 200),
    ("quickstart_rect", dict(kind="quick", grid=(40, 24), seed=5), 100),
    ("sorting48_periodic", dict(kind="sort", grid=(48, 48), seed=11, n_types=3), 60),
    ("sorting_fixed_rect", dict(kind="sort", grid=(36, 50), seed=12, n_types=2,
                                boundary="fixed"), 60),
    ("sorting64_4types", dict(kind="sort", grid=(64, 64), seed=13, n_types=4, tv=40, ti=30,
                              phi=0.8), 30),
    ("sorting3d_periodic", dict(kind="sort", grid=(16, 16, 16), seed=21, n_types=2, tv=100,
                                ti=150, phi=0.25), 8),
    ("sorting3d_fixed", dict(kind="sort", grid=(14, 18, 12), seed=22, n_types=3, tv=60,
                             ti=100, phi=0.5, boundary="fixed"), 8),
    ("sorting3d_26", dict(kind="sort", grid=(12, 12, 12), seed=23, n_types=2, tv=60, ti=120,
                          phi=0.4, energy="twenty_six"), 8),
    # the 2D sorting configuration: 256x256 periodic, 4 types, V*=40, I*=30,
    # lambda 5 / 1, phi = 0.8 (1311 cells), J medium row 16 and U[2,14]
    ("sorting256_config", dict(kind="sort", grid=(256, 256), seed=1, n_types=4, tv=40, ti=30,
                               phi=0.8), 5),
    # lattices whose haloed u16 copy exceeds LDS keep sigma in HBM
    ("hbm_sigma_2d", dict(kind="sort", grid=(300, 296), seed=31, n_types=3, tv=60, ti=40,
                          phi=0.6), 2),
    ("hbm_sigma_3d", dict(kind="sort", grid=(48, 48, 48), seed=32, n_types=2, tv=300,
                          ti=300, phi=0.2, boundary="fixed"), 1),
    # more cells than 12-bit packed ids hold: one lane per replica, state in HBM
    ("lane_path_4352_cells", dict(kind="sort", grid=(160, 160), seed=33, n_types=2, tv=5,
                                  ti=10, phi=0.85), 2),
    ("long_run_many_rounds", dict(kind="quick", grid=(30, 30), seed=9), 2000),
    # the edges of the packed 16-bit words: 4095 cells and 15 types stay on
    # the wave path, 4096 cells go to the lane path
    ("wave_path_4095_cells", dict(kind="sort", grid=(160, 160), seed=34, n_types=2, tv=5,
                                  ti=10, phi=4095 * 5 / 25600), 2),
    ("lane_path_4096_cells", dict(kind="sort", grid=(160, 160), seed=35, n_types=2, tv=5,
                                  ti=10, phi=4096 * 5 / 25600), 2),
    ("wave_path_15_types", dict(kind="sort", grid=(64, 64), seed=36, n_types=15, tv=20, ti=18,
                                phi=0.7), 10),
]


def test_sixteen_cell_types_are_refused():
    """The device J table holds (T+1)^2 <= 256 entries: a configuration with
    16 cell types is refused at configuration time with a ValueError."""
    with pytest.raises(ValueError, match="at most 15 cell types"):
        sim, _ = sorting_pair(cpm, O, 37, grid=(64, 64), n_types=16, tv=20, ti=18, phi=0.7)
        cpm.Replica(sim)


def build_case(args):
    args = dict(args)
    kind = args.pop("kind")
    if kind == "quick":
        sim = quickstart(cpm, seed=args["seed"], grid=args["grid"])
        sim.initialize(warn_on_default_init=False)
        cfg = quickstart_oracle(O, seed=args["seed"], grid=args["grid"])
        return sim, cfg
    seed = args.pop("seed")
    return sorting_pair(cpm, O, seed, **args)


@pytest.mark.parametrize("label,args,mcs", CASES, ids=[c[0] for c in CASES])
def test_exact_replay_matches_oracle(label, args, mcs):
    sim, cfg = build_case(args)
    print(f"[case] {label} grid={args['grid']} mcs={mcs}", flush=True)
    rep = cpm.Replica(sim)
    ref = O.Model(cfg)
    assert_same_state(rep.snapshot(), ref.snapshot(), label + "@init")
    cnt = rep.run_mcs(mcs)
    ref.run_mcs(mcs)
    assert_same_state(rep.snapshot(), ref.snapshot(), label)
    oc = ref.counters()
    assert [cnt["attempts"], cnt["non_null"], cnt["topology"], cnt["energy"],
            cnt["accepted"]] == [int(x) for x in oc]
    assert cnt["accepted"] > 0


def test_batch_composition_does_not_change_results():
    """Replicas advanced together in one launch equal each replica run alone
    (batch.rs:251-264 proves the same for thread counts)."""
    pairs = [build_case(c[1]) for c in CASES[:5]]
    reps = [cpm.Replica(s) for s, _ in pairs]
    cpm.Replica.run_mcs_many(reps, 25)
    for (sim, cfg), rep in zip(pairs, reps):
        ref = O.Model(cfg)
        ref.run_mcs(25)
        assert_same_state(rep.snapshot(), ref.snapshot())


def test_tuned_placement_slices_match_oracle():
    """A batch call long enough for the sigma placement autotuner (128 MCS,
    two slices per candidate placement: sigma in HBM for every replica, in
    LDS smallest lattices first or longest chains first) equals the
    uninterrupted reference loop for every replica."""
    pairs = [build_case(c[1]) for c in CASES[:5]]
    reps = [cpm.Replica(s) for s, _ in pairs]
    print("[case] tuner batch of 5 replicas, 128 MCS", flush=True)
    cpm.Replica.run_mcs_many(reps, 128)
    for (sim, cfg), rep in zip(pairs, reps):
        ref = O.Model(cfg)
        ref.run_mcs(128)
        assert_same_state(rep.snapshot(), ref.snapshot())


def test_plain_launches_match_batched():
    sim, _ = build_case(CASES[0][1])
    a, b = cpm.Replica(sim), cpm.Replica(sim)
    a.run_mcs_plain(20)
    b.run_mcs(20)
    assert_same_state(a.snapshot(), b.snapshot())


def test_run_samples_match_oracle_run_with():
    sim = quickstart(cpm, seed=7)
    sim.initialize(warn_on_default_init=False)
    res = sim.run(burn_in_mcs=200, readout_mcs=300, sampling_interval_mcs=100,
                  include_lattice=True)
    cfg = quickstart_oracle(O, seed=7, windows=(200, 300, 100))
    cfg.warn_on_default_init = False
    ref = O.run_with(O.Model(cfg), include_lattice=True)
    assert len(res.samples) == len(ref.samples) == 3
    for s, r in zip(res.samples, ref.samples):
        assert s.mcs == r.mcs
        assert s.conservative_energy == r.energy
        assert np.array_equal(s.volume, r.volume)
        assert np.array_equal(s.interface, r.interface)
        assert np.array_equal(s.contact, r.contact)
        assert np.array_equal(s.lattice, r.lattice)
        assert np.array_equal(s.centroid, r.centroid)
        assert np.allclose(s.eigenvalues, r.eigenvalues, rtol=0, atol=1e-9)
        assert s.fragmented_cells == r.fragmented
        assert s.medium_component_volumes == r.medium_volumes
        assert s.medium_component_touches_boundary == r.medium_touch
    assert res.status.kind == ref.status[0]
    assert res.metadata.model_hash == ref.model_hash
    assert res.metadata.convention_hash == ref.convention_hash
    assert res.metadata.derived_phi == ref.derived_phi


def test_readme_quickstart_full_length_matches_oracle():
    """The CPU-runnable configuration at its full length: README quickstart
    (30x30 periodic, seed 1, 5 cells), burn-in 5000 + readout 5000 MCS,
    a sample every 100 MCS — every sample, the trace-derived status and
    the final hashes bit-identical to the reference loop."""
    sim = quickstart(cpm, seed=1)
    sim.initialize(warn_on_default_init=False)
    print("[case] quickstart 30x30, 5000 + 5000 MCS", flush=True)
    res = sim.run(burn_in_mcs=5000, readout_mcs=5000, sampling_interval_mcs=100)
    cfg = quickstart_oracle(O, seed=1, windows=(5000, 5000, 100))
    cfg.warn_on_default_init = False
    ref = O.run_with(O.Model(cfg))
    assert len(res.samples) == len(ref.samples) == 50
    for s, r in zip(res.samples, ref.samples):
        assert s.mcs == r.mcs
        assert s.conservative_energy == r.energy
        assert np.array_equal(s.volume, r.volume)
        assert np.array_equal(s.interface, r.interface)
        assert np.array_equal(s.contact, r.contact)
        assert np.array_equal(s.centroid, r.centroid)
    assert res.status.kind == ref.status[0]
    assert res.metadata.model_hash == ref.model_hash
    assert res.metadata.derived_phi == ref.derived_phi


def test_organoid_64cubed_matches_oracle():
    """The 3D organoid configuration itself: 64^3 fixed boundary, 3 types in
    nested shells placed explicitly, 18-neighbour energy — sigma kept in
    HBM, the pipelined kernel, four windows per block — for 3 MCS, against
    the reference loop on the same explicit lattice."""
    import bench

    sim = bench.organoid_64(cpm, seed=1)
    sp = sim._spec
    types = [O.CellTypeSpec(t["name"], t["target_volume"], t["target_interface"],
                            t["lambda_volume"], t["lambda_interface"]) for t in sp["cell_types"]]
    type_of = sp["init"]["cell_type_of"]
    cfg = O.Config(grid=(64, 64, 64), boundary="fixed", seed=1, cell_types=types,
                   cell_counts=[type_of.count(k) for k in range(len(types))],
                   adhesion=sp["adhesion"], init_lattice=sp["init"]["lattice"],
                   init_cell_type_of=type_of)
    print("[case] organoid 64^3, 3 MCS", flush=True)
    rep = cpm.Replica(sim)
    ref = O.Model(cfg)
    assert_same_state(rep.snapshot(), ref.snapshot(), "organoid@init")
    cnt = rep.run_mcs(3)
    ref.run_mcs(3)
    assert_same_state(rep.snapshot(), ref.snapshot(), "organoid")
    assert cnt["accepted"] > 0


def test_run_batch_matches_oracle_per_replica():
    """batch.rs:110-153: theta injected per replica, seed derive_seed_u64."""
    sim = cpm.CPM(grid=(20, 20), boundary="periodic", seed=1)
    sim.register_cell_type("a", 9, 20, 1.0, 0.2)
    sim.register_cell_type("b", 9, 20, 1.5, 0.3)
    sim.add_cells("a", 2)
    sim.add_cells("b", 2)
    sim.set_adhesion([[0.0, 3.0, 3.0], [3.0, 1.0, 5.0], [3.0, 5.0, 1.0]])
    sim.initialize(warn_on_default_init=False)
    base = sim.extract_theta()
    thetas = [base, [v * 1.1 for v in base], [v * 0.9 for v in base]]
    results = sim.run_batch(thetas, burn_in_mcs=20, readout_mcs=40, sampling_interval_mcs=20,
                            master_seed=42)
    cfg0 = O.Config(grid=(20, 20), boundary="periodic", seed=0,
                    cell_types=[O.CellTypeSpec("a", 9, 20, 1.0, 0.2),
                                O.CellTypeSpec("b", 9, 20, 1.5, 0.3)],
                    cell_counts=[2, 2], adhesion=[[0.0, 3.0, 3.0], [3.0, 1.0, 5.0],
                                                  [3.0, 5.0, 1.0]],
                    warn_on_default_init=False, burn_in_mcs=20, readout_mcs=40,
                    sampling_interval_mcs=20)
    for i, (th, res) in enumerate(zip(thetas, results)):
        cfg = O.theta_inject(cfg0, th)
        cfg.seed = O.derive_seed_u64(42, i)
        ref = O.run_with(O.Model(cfg))
        assert res.metadata.seed == cfg.seed == PINS["derived_seeds"]["values"][i]
        for s, r in zip(res.samples, ref.samples):
            assert s.conservative_energy == r.energy
            assert np.array_equal(s.volume, r.volume)
        assert res.metadata.model_hash == ref.model_hash


def test_run_batch_configs_prior_box_matches_oracle():
    """The batched workload's own shape at small count: 16 configurations
    from the seeded prior box (sides 48-128, 2 types, phi = 0.7, seeds
    derive_seed_u64(42, i)) through run_batch_configs — uneven grids in one
    batch, the placement tuner's slices, readouts overlapped with the next
    interval — each equal to the reference loop run alone."""
    import bench

    params = bench.draw_batch(16)
    sims, cfgs = [], []
    for p in params:
        seed = cpm.derive_seed_u64(42, p["index"])
        sims.append(bench.make_sim(cpm, p, seed))
        cfgs.append(O.Config(grid=(p["side"], p["side"]), boundary="periodic", seed=seed,
                             cell_types=[O.CellTypeSpec("a", p["tv"], p["ti"], p["lv"][0],
                                                        p["li"][0]),
                                         O.CellTypeSpec("b", p["tv"], p["ti"], p["lv"][1],
                                                        p["li"][1])],
                             cell_counts=p["counts"], adhesion=p["J"],
                             warn_on_default_init=False, burn_in_mcs=0, readout_mcs=400,
                             sampling_interval_mcs=200))
    print("[case] run_batch_configs, 16 prior-box configurations, 400 MCS", flush=True)
    results = cpm.run_batch_configs(sims, 0, 400, 200)
    for res, cfg in zip(results, cfgs):
        ref = O.run_with(O.Model(cfg))
        assert len(res.samples) == len(ref.samples) == 2
        for s, r in zip(res.samples, ref.samples):
            assert s.mcs == r.mcs
            assert s.conservative_energy == r.energy
            assert np.array_equal(s.volume, r.volume)
            assert np.array_equal(s.contact, r.contact)
        assert res.metadata.model_hash == ref.model_hash


def test_device_simple_point_tables_match_the_definition():
    import ctypes as C
    from cpm import _native as N
    t2 = np.zeros(256, dtype=np.uint8)
    t3 = np.zeros(1 << 20, dtype=np.uint64)
    N.check(N.lib().cpm_engine_sp_tables(N.engine(), t2.ctypes.data_as(N.u8p),
                                         t3.ctypes.data_as(N.u64p)))
    assert "".join(str(int(b)) for b in t2) == PINS["sp_table_2d_bits"]
    assert hashlib.sha256(t3.astype("<u8").tobytes()).hexdigest() == PINS["sp_table_3d_sha256"]
    del C


def test_piped_stragglers_match_the_single_wave_kernel():
    """The batched workload (4096 replicas of the prior box, sides 48-128):
    with the longest-chains-first placement at 75 % of the LDS, the
    128x128 lattices that the LDS budget leaves in HBM run on the pipelined
    kernel (sigma in HBM, two waves), beside single-wave chains.  Every
    replica equals its twin run with pipelining switched off, on two calls
    in a row."""
    import bench

    params = bench.draw_batch(4096)
    sims = [bench.make_sim(cpm, p, cpm.derive_seed_u64(42, p["index"])) for p in params]
    out = {}
    try:
        cpm._native.set_sigma_placement("longest", 0.75)
        for piped in (True, False):
            cpm._native.set_pipelining(piped)
            reps = [cpm.Replica(s, initialize=False) for s in sims]
            cpm.Replica.initialize_many(reps)
            n_piped = 0
            for _ in range(2):
                n_piped += cpm.Replica.run_mcs_many(reps, 60)["piped_stragglers"]
            print(f"[case] pipelining {piped}: {n_piped} piped stragglers", flush=True)
            assert (n_piped > 0) == piped
            out[piped] = [r.snapshot() for r in reps]
            del reps
    finally:
        cpm._native.set_sigma_placement("auto", 0.0)
        cpm._native.set_pipelining(True)
    for i, (a, b) in enumerate(zip(out[True], out[False])):
        for key in ("lattice", "volume", "interface", "sum_position", "sum_outer", "contact", "rng"):
            assert np.array_equal(a[key], b[key]), (i, key)
        assert a["energy"] == b["energy"], i
