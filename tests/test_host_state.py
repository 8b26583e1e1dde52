"""State import (cpm_model_set_state), one-sample readouts (cpm_model_sample)
and the status rules of the product's host library, checked on the CPU
against the reference's hand-built cases and the oracle.  No GPU: a model
that has never run has no device copy."""
import numpy as np
import pytest

import cpm
from cpm.core import not_equilibrated
from oracle import oracle as O
from tests.test_oracle_pins import CASES, lumens


def _sim(grid, boundary, n_types=1):
    sim = cpm.CPM(grid=grid, boundary=boundary, seed=1)
    names = [f"t{i}" for i in range(n_types)]
    for nm in names:
        sim.register_cell_type(nm, target_volume=10, target_interface=10, lambda_volume=1.0,
                               lambda_interface=1.0)
    for nm in names:
        sim.add_cells(nm, 1)
    T = n_types + 1
    sim.set_adhesion([[0.0 if a == b == 0 else 1.0 for b in range(T)] for a in range(T)])
    return sim


def _replica(grid, boundary, lattice, type_index, **kw):
    rep = cpm.Replica(_sim(grid, boundary), initialize=False)
    rep.set_state(np.asarray(lattice).ravel(), type_index, **kw)
    return rep


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_product_labelling_matches_the_reference_cases(case):
    """labelling.rs:174-336 through cpm_model_set_state + cpm_model_sample."""
    name, grid, boundary, lat, cto, frag, lumen_vols, n_lumens, n_medium = case
    r = _replica(grid, boundary, lat, cto).sample()
    s = r.samples[0]
    if frag is not None:
        assert s.fragmented_cells == frag
    got = lumens(s.medium_component_volumes, s.medium_component_touches_boundary)
    if lumen_vols is not None:
        assert sorted(got) == sorted(lumen_vols)
    if n_lumens is not None:
        assert len(got) == n_lumens
    if n_medium is not None:
        assert len(s.medium_component_volumes) == n_medium
        assert not any(s.medium_component_touches_boundary)
    # and the oracle's labelling of the same lattice agrees component by component
    from tests.test_oracle_pins import _model_with
    of, omv, omt = _model_with(grid, boundary, lat, cto).label()
    assert s.fragmented_cells == of
    assert s.medium_component_volumes == omv
    assert s.medium_component_touches_boundary == omt


def test_product_status_rules_on_hand_built_states():
    """run_with's rules (output.rs:470-509) applied to one sample."""
    pieces = np.zeros((9, 9), dtype=np.uint32)
    pieces[1, 1] = pieces[6, 6] = 1
    r = _replica((9, 9), "fixed", pieces, [0]).sample()
    assert (r.status.kind, r.status.cell_id, r.status.at_mcs) == ("Fragmented", 1, 0)
    lost = np.zeros((9, 9), dtype=np.uint32)
    lost[2:5, 2:5] = 1
    lost[7, 7] = 2
    r = _replica((9, 9), "fixed", lost, [0, 0], mcs=12).sample()
    assert (r.status.kind, r.status.cell_id, r.status.at_mcs) == ("CellLost", 2, 12)
    touching = np.zeros((9, 9), dtype=np.uint32)
    touching[2:5, 2:5] = 1
    touching[5:7, 2:5] = 2
    r = _replica((9, 9), "fixed", touching, [0, 0]).sample()
    assert r.status.kind == "Degenerate"
    apart = np.zeros((9, 9), dtype=np.uint32)
    apart[1:3, 1:3] = 1
    apart[6:8, 6:8] = 2
    r = _replica((9, 9), "fixed", apart, [0, 0]).sample()
    assert r.status.kind == "Ok"


def test_product_not_equilibrated_matches_reference_cases():
    """output.rs:614-627 and the oracle on random traces."""
    assert not_equilibrated([i * 10.0 for i in range(20)])
    assert not not_equilibrated([100.0, 101.0, 99.5, 100.3, 99.8, 100.1, 99.9, 100.2, 100.0,
                                 99.7])
    assert not not_equilibrated([])
    rng = np.random.default_rng(3)
    for _ in range(200):
        n = int(rng.integers(0, 30))
        tr = (rng.normal(size=n) + rng.uniform(-0.3, 0.3) * np.arange(n)).tolist()
        assert not_equilibrated(tr) == O.not_equilibrated(tr)


@pytest.mark.parametrize("boundary", ["fixed", "periodic"])
def test_set_state_derives_missing_fields_like_the_oracle(boundary):
    """With only the lattice given, volume / interface / sums / contact /
    energy are those recompute and global_energy give (the oracle's
    load_state derives them the same way, initialization.rs:396-436)."""
    sim, cfg = _sorting_pair(boundary)
    ref = O.Model(cfg)
    ref.run_mcs(15)
    lat, typ = ref.lattice(), ref.type_index()
    rep = cpm.Replica(sim, initialize=False)
    rep.set_state(lat.ravel(), typ, rng=ref.rng_state(), mcs=15)
    snap = rep.snapshot()
    chk = O.Model(cfg)
    chk.load_state(lat.ravel(), typ)
    assert np.array_equal(snap["lattice"], lat)
    assert np.array_equal(snap["volume"], chk.volume())
    assert np.array_equal(snap["interface"], chk.interface())
    assert np.array_equal(snap["sum_position"], chk.sum_position())
    assert np.array_equal(snap["sum_outer"], chk.sum_outer())
    assert np.array_equal(snap["contact"], chk.contact())
    assert snap["energy"] == chk.global_energy()
    assert np.array_equal(snap["rng"], ref.rng_state())
    assert snap["mcs"] == 15


def test_set_state_keeps_given_fields_bit_for_bit():
    sim, cfg = _sorting_pair("periodic")
    ref = O.Model(cfg)
    ref.run_mcs(7)
    s = ref.snapshot()
    rep = cpm.Replica(sim, initialize=False)
    rep.set_state(s["lattice"].ravel(), ref.type_index(), volume=s["volume"],
                  interface=s["interface"], sum_position=s["sum_position"],
                  sum_outer=s["sum_outer"], contact=s["contact"], energy=s["energy"],
                  rng=s["rng"], mcs=s["mcs"])
    got = rep.snapshot()
    for k in ("lattice", "volume", "interface", "sum_position", "sum_outer", "contact", "rng"):
        assert np.array_equal(got[k], s[k]), k
    assert got["energy"] == s["energy"] and got["mcs"] == 7
    assert rep.hashes()[1] == ref.model_hash()


def test_set_state_rejects_what_the_device_cannot_hold_and_keeps_the_old_state():
    sim, cfg = _sorting_pair("periodic")
    ref = O.Model(cfg)
    rep = cpm.Replica(sim)
    before = rep.snapshot()
    lat = ref.lattice().ravel().copy()
    typ = ref.type_index()
    bad_id = lat.copy()
    bad_id[0] = len(typ) + 1
    with pytest.raises(ValueError, match="lattice references cell id"):
        rep.set_state(bad_id, typ)
    with pytest.raises(ValueError, match="type_index"):
        rep.set_state(lat, np.full(len(typ), 9))
    with pytest.raises(ValueError, match="fixed point"):
        rep.set_state(lat, typ, rng=[0, 0, 0, 0])
    with pytest.raises(ValueError, match="n_cells"):
        rep.set_state(np.zeros_like(lat), [])
    after = rep.snapshot()
    for k in before:
        assert np.array_equal(np.asarray(after[k]), np.asarray(before[k])), k


def _sorting_pair(boundary, seed=5):
    from tests.helpers import sorting_pair
    return sorting_pair(cpm, O, seed, grid=(24, 24), n_types=2, boundary=boundary, tv=20,
                        ti=16)
