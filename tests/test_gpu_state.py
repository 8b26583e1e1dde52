"""State import on the GPU: a model handed over mid-trajectory
(cpm_model_set_state with the oracle's state after scatter_and_grow and 37
MCS, RNG position and unwrapped position sums included) continues the
reference's trajectory bit for bit on the device; and sublattice-mode
samples taken in a batch equal those of the same replica run alone (the
readouts of one interval never see the next interval's state)."""
import struct

import numpy as np
import pytest

import cpm
from oracle import oracle as O
from tests.helpers import sorting_pair

pytestmark = pytest.mark.gpu

IMPORT_CASES = [
    # label, sorting_pair kwargs, seed
    ("2d_periodic_lds", dict(grid=(48, 48), n_types=3), 11),
    ("2d_fixed_moore", dict(grid=(40, 40), n_types=2, boundary="fixed"), 12),
    ("3d_periodic", dict(grid=(20, 20, 20), n_types=2, tv=40, ti=60, phi=0.5), 13),
    ("3d_fixed_26", dict(grid=(18, 18, 18), n_types=2, boundary="fixed", tv=40, ti=60,
                         energy="twenty_six"), 14),
    # more than 4095 cells: the lane path
    ("2d_lane_path", dict(grid=(150, 150), n_types=2, tv=4, ti=8, phi=0.8), 15),
]


def _bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


@pytest.mark.parametrize("label,kw,seed", IMPORT_CASES, ids=[c[0] for c in IMPORT_CASES])
def test_imported_state_continues_the_oracle_trajectory(label, kw, seed):
    sim, cfg = sorting_pair(cpm, O, seed, **kw)
    ref = O.Model(cfg)
    ref.run_mcs(37)
    s = ref.snapshot()
    print(f"[state-import] {label}: {ref.n_cells} cells, grid {cfg.grid}", flush=True)
    rep = cpm.Replica(sim, initialize=False)
    rep.set_state(s["lattice"].ravel(), ref.type_index(), volume=s["volume"],
                  interface=s["interface"], sum_position=s["sum_position"],
                  sum_outer=s["sum_outer"], contact=s["contact"], energy=s["energy"],
                  rng=s["rng"], mcs=37)
    rep.run_mcs(50)
    ref.run_mcs(50)
    got, want = rep.snapshot(), ref.snapshot()
    assert got["mcs"] == want["mcs"] == 87
    for k in ("lattice", "volume", "interface", "sum_position", "sum_outer", "contact", "rng"):
        assert np.array_equal(got[k], want[k]), f"{label}: {k} differs"
    assert _bits(got["energy"]) == _bits(want["energy"])
    assert rep.hashes()[1] == ref.model_hash()
    # the readout of the continued state equals the oracle's take_sample
    smp = rep.sample().samples[0]
    osm = O.take_sample(ref)
    assert np.array_equal(smp.volume, osm.volume)
    assert np.array_equal(smp.centroid, osm.centroid)
    assert np.array_equal(smp.eigenvalues, osm.eigenvalues)
    assert np.array_equal(smp.kappa_sq, osm.kappa)
    assert smp.fragmented_cells == osm.fragmented
    assert smp.medium_component_volumes == osm.medium_volumes


def test_reimport_replaces_a_device_resident_state():
    """A model that already ran on the device takes a new state: the device
    copy of the old one is dropped, and the new trajectory is the oracle's."""
    sim, cfg = sorting_pair(cpm, O, 21, grid=(32, 32), n_types=2)
    rep = cpm.Replica(sim)
    rep.run_mcs(5)
    ref = O.Model(cfg)
    ref.run_mcs(60)
    s = ref.snapshot()
    rep.set_state(s["lattice"].ravel(), ref.type_index(), volume=s["volume"],
                  interface=s["interface"], sum_position=s["sum_position"],
                  sum_outer=s["sum_outer"], contact=s["contact"], energy=s["energy"],
                  rng=s["rng"], mcs=60)
    rep.run_mcs(20)
    ref.run_mcs(20)
    got, want = rep.snapshot(), ref.snapshot()
    for k in ("lattice", "volume", "interface", "contact", "rng"):
        assert np.array_equal(got[k], want[k]), k
    assert _bits(got["energy"]) == _bits(want["energy"])


def _sublattice_sim(seed, side=48):
    J = [[0.0, 16.0, 16.0], [16.0, 3.0, 11.0], [16.0, 11.0, 5.0]]
    sim = cpm.CPM(grid=(side, side), boundary="periodic", seed=seed)
    n = round(0.7 * side * side / 30)
    for t in range(2):
        sim.register_cell_type(f"t{t}", target_volume=30, target_interface=24,
                               lambda_volume=5.0, lambda_interface=1.0)
        sim.add_cells(f"t{t}", n // 2 + (1 if t < n % 2 else 0))
    sim.set_adhesion(J)
    sim.initialize(warn_on_default_init=False)
    return sim


def test_sublattice_batch_samples_equal_each_replica_run_alone():
    """Several samples per run over many replicas: every sample of the
    batched run_with equals the same replica advanced alone by run_mcs and
    read at that MCS."""
    seeds = [cpm.derive_seed_u64(7, i) for i in range(48)]
    sims = [_sublattice_sim(sd) for sd in seeds]
    res = cpm.run_batch_configs(sims, burn_in_mcs=10, readout_mcs=40, sampling_interval_mcs=10,
                                mode="sublattice")
    for i in (0, 17, 47):
        rep = cpm.Replica(_sublattice_sim(seeds[i]), 10, 40, 10)
        rep.run_mcs(10, mode="sublattice")
        assert len(res[i].samples) == 4
        for k in range(4):
            rep.run_mcs(10, mode="sublattice")
            alone = rep.sample().samples[0]
            batch = res[i].samples[k]
            assert batch.mcs == alone.mcs == 20 + 10 * k
            assert np.array_equal(batch.volume, alone.volume), (i, k)
            assert np.array_equal(batch.interface, alone.interface), (i, k)
            assert np.array_equal(batch.contact, alone.contact), (i, k)
            assert np.array_equal(batch.centroid, alone.centroid), (i, k)
            assert batch.conservative_energy == alone.conservative_energy, (i, k)
