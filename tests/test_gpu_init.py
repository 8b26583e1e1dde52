"""scatter_and_grow on the device (initialization.rs:202-330, one wave per
model): every model initialised in one batched call must end in exactly the
state the host initialisation gives it (itself pinned against the
reference's 2D organoid notebook, site for site) — lattice, volumes,
interfaces, position sums, contacts, energy bits, RNG position and the
default-init flag — including models the device leaves to the host
(explicit placements, more than 4096 cells) mixed into the same call, and
the reference's error for a configuration too crowded to seed."""
import struct
import warnings

import numpy as np
import pytest

import cpm
from oracle import oracle as O
from tests.helpers import quickstart, quickstart_oracle, sorting_pair

pytestmark = pytest.mark.gpu


def _bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _sims():
    """(label, sim) pairs covering both dimensions and boundaries, types of
    different target volumes and the benchmark shapes; seeds vary within
    each family."""
    out = []
    for s in range(4):
        out.append((f"quickstart_{s}", quickstart(cpm, seed=100 + s)))
    for s, (grid, bnd, nt, tv, phi) in enumerate([((48, 48), "periodic", 3, 30, 0.6),
                                                 ((40, 56), "fixed", 2, 24, 0.6),
                                                 ((64, 64), "periodic", 2, 55, 0.6),
                                                 ((20, 20, 20), "periodic", 2, 40, 0.25),
                                                 ((18, 22, 16), "fixed", 3, 30, 0.2)]):
        sim, _ = sorting_pair(cpm, O, 200 + s, grid=grid, n_types=nt, boundary=bnd, tv=tv,
                              ti=2 * tv, phi=phi)
        out.append((f"sorting_{grid}_{bnd}", sim))
    for s, grid in enumerate([(36, 36), (16, 16, 16), (30, 30), (14, 18, 16)]):
        sim = cpm.CPM(grid=grid, boundary="periodic" if s % 2 else "fixed", seed=300 + s)
        # three types of different target volumes: each cell grows to its own
        for k, tv in enumerate((12, 25, 40)):
            sim.register_cell_type(f"t{k}", target_volume=tv, target_interface=2 * tv,
                                   lambda_volume=4.0, lambda_interface=1.0)
            sim.add_cells(f"t{k}", 3 + 2 * k)
        sim.set_adhesion([[0.0, 9.0, 9.0, 9.0], [9.0, 3.0, 6.0, 6.0], [9.0, 6.0, 3.0, 6.0],
                          [9.0, 6.0, 6.0, 3.0]])
        out.append((f"three_volumes_{grid}", sim))
    # the 256x256 four-type sorting configuration (1311 cells)
    from tests.test_gpu_sublattice import c2_sim
    out.append(("sorting_256x256", c2_sim(7)))
    return out


def _state(rep):
    snap = rep.snapshot()
    snap["energy_bits"] = _bits(snap.pop("energy"))
    snap["used_default"] = rep.used_default
    return snap


def _compare(label, got, want):
    for k in want:
        a, b = got[k], want[k]
        same = np.array_equal(a, b) if isinstance(a, np.ndarray) else a == b
        assert same, f"{label}: {k} differs between device and host initialisation"


def test_device_scatter_and_grow_equals_the_host_initialisation():
    sims = _sims()
    host = [_state(cpm.Replica(s)) for _, s in sims]
    reps = [cpm.Replica(s, initialize=False) for _, s in sims]
    cpm.Replica.initialize_many(reps)
    for (label, _), r, want in zip(sims, reps, host):
        print(f"[init] {label}: {len(want['volume'])} cells", flush=True)
        _compare(label, _state(r), want)


def test_device_initialised_models_run_as_the_host_initialised_ones():
    """The device-initialised state (derived tables included) is what the
    first run starts from; every replica then equals its host-initialised
    twin after 30 MCS, in one batched call."""
    sims = [s for _, s in _sims()]
    host = [cpm.Replica(s) for s in sims]
    dev = [cpm.Replica(s, initialize=False) for s in sims]
    cpm.Replica.initialize_many(dev)
    cpm.Replica.run_mcs_many(host, 30)
    cpm.Replica.run_mcs_many(dev, 30)
    for i, (a, b) in enumerate(zip(dev, host)):
        _compare(f"replica {i}", _state(a), _state(b))


def test_device_initialisation_matches_the_oracle_and_runs_on():
    """The oracle's scatter_and_grow (pinned by the notebook) gives the same
    state, and the device-initialised model then replays the oracle's run."""
    sim = quickstart(cpm, seed=41)
    rep = cpm.Replica(sim, initialize=False)
    cpm.Replica.initialize_many([rep])
    ref = O.Model(quickstart_oracle(O, seed=41))
    got, want = rep.snapshot(), ref.snapshot()
    for k in ("lattice", "volume", "interface", "sum_position", "sum_outer", "contact", "rng"):
        assert np.array_equal(got[k], want[k]), k
    # the device-initialised state is resident: the run starts from it as is
    rep.run_mcs(200)
    ref.run_mcs(200)
    got, want = rep.snapshot(), ref.snapshot()
    for k in ("lattice", "volume", "interface", "sum_position", "sum_outer", "contact", "rng"):
        assert np.array_equal(got[k], want[k]), k
    assert _bits(got["energy"]) == _bits(want["energy"])


def test_models_left_to_the_host_in_the_same_call():
    """An explicit placement and a model of more than 4096 cells initialise
    on the host inside the batched call; the others on the device."""
    from tests.helpers import frozen_sim

    big = cpm.CPM(grid=(150, 150), boundary="periodic", seed=9)
    big.register_cell_type("a", target_volume=4, target_interface=8, lambda_volume=2.0,
                           lambda_interface=1.0)
    big.add_cells("a", 4500)
    big.set_adhesion([[0.0, 6.0], [6.0, 2.0]])
    sims = [("frozen_explicit", frozen_sim(cpm)), ("quickstart", quickstart(cpm, seed=5)),
            ("4500_cells", big)]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        host = [_state(cpm.Replica(s)) for _, s in sims]
        reps = [cpm.Replica(s, initialize=False) for _, s in sims]
    cpm.Replica.initialize_many(reps)
    for (label, _), r, want in zip(sims, reps, host):
        _compare(label, _state(r), want)
    assert [r.used_default for r in reps] == [False, True, True]


def test_too_crowded_to_seed_raises_the_references_error():
    """initialization.rs:242-248: the same message as the host; the lowest
    failing index is reported."""
    crowded = cpm.CPM(grid=(12, 12), boundary="fixed", seed=3)
    crowded.register_cell_type("a", target_volume=60, target_interface=40, lambda_volume=1.0,
                               lambda_interface=1.0)
    crowded.add_cells("a", 20)
    crowded.set_adhesion([[0.0, 5.0], [5.0, 2.0]])
    with pytest.raises(ValueError) as host_err:
        cpm.Replica(crowded)
    reps = [cpm.Replica(quickstart(cpm, seed=2), initialize=False),
            cpm.Replica(crowded, initialize=False)]
    with pytest.raises(ValueError) as dev_err:
        cpm.Replica.initialize_many(reps)
    assert str(dev_err.value) == str(host_err.value)
    assert "too crowded" in str(dev_err.value)


def test_run_batch_configs_same_with_device_and_host_initialisation():
    """The batched driver initialises on the device by default; switching
    it to the host changes nothing in the results."""
    import bench

    sims = [bench.make_sim(cpm, p, cpm.derive_seed_u64(42, p["index"]))
            for p in bench.draw_batch(12)]
    out = {}
    try:
        for on in (True, False):
            cpm._native.set_device_init(on)
            res = cpm.run_batch_configs(sims, burn_in_mcs=0, readout_mcs=40,
                                        sampling_interval_mcs=20, include_lattice=True)
            out[on] = res
    finally:
        cpm._native.set_device_init(True)
    for a, b in zip(out[True], out[False]):
        assert a.metadata.model_hash == b.metadata.model_hash
        for sa, sb in zip(a.samples, b.samples):
            assert np.array_equal(sa.lattice, sb.lattice)
            assert np.array_equal(sa.volume, sb.volume)
