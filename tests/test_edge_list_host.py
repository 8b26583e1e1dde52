"""The EdgeList proposal on the host side of the product (CPU only): the
edge list built at initialisation equals the oracle's in order
(EdgeList::build, edge_list.rs:121-131, called by initialization.rs:
427-435), and state import takes a given list order or rebuilds it."""
import re

import numpy as np
import pytest

import cpm
from oracle import oracle as O
from tests.test_oracle_edge_list import calibration_config


def _pair(grid, boundary, seed=4):
    tv = 30 if len(grid) == 2 else 60
    ti = 24 if len(grid) == 2 else 90
    n = max(2, int(0.2 * np.prod(grid) / tv))
    sim = cpm.CPM(grid=grid, boundary=boundary, seed=seed, proposal="edge_list")
    sim.register_cell_type("a", tv, ti, 2.0, 0.5)
    sim.add_cells("a", n)
    sim.set_adhesion([[0.0, 6.0], [6.0, 2.0]])
    sim.initialize(warn_on_default_init=False)
    cfg = O.Config(grid=grid, boundary=boundary, seed=seed,
                   cell_types=[O.CellTypeSpec("a", tv, ti, 2.0, 0.5)], cell_counts=[n],
                   adhesion=[[0.0, 6.0], [6.0, 2.0]], warn_on_default_init=False,
                   proposal="edge_list")
    return sim, cfg


@pytest.mark.parametrize("grid,boundary", [((40, 40), "periodic"), ((30, 44), "fixed"),
                                           ((16, 16, 16), "fixed")])
def test_initial_edge_list_equals_the_oracles(grid, boundary):
    sim, cfg = _pair(grid, boundary)
    rep = cpm.Replica(sim)
    assert np.array_equal(rep.edge_sites(), O.Model(cfg).edge_sites())
    assert rep.hashes()[0] == O.convention_hash(cfg)


def test_notebook_calibration_config_hash_and_edges():
    cfg = calibration_config()
    sim = cpm.CPM(grid=(56, 56, 56), boundary="fixed", seed=7, energy_neighborhood="eighteen",
                  proposal="edge_list")
    sim.register_cell_type("probe", 900, 0, 4.0, 0.0)
    sim.add_cells("probe", 1)
    sim.set_adhesion([[0.0, 2.5], [2.5, 0.0]])
    rep = cpm.Replica(sim, 20000, 20000, 2000)
    ref = O.Model(cfg)
    assert rep.hashes() == (O.convention_hash(cfg), ref.model_hash())
    assert np.array_equal(rep.edge_sites(), ref.edge_sites())


def test_state_import_takes_or_rebuilds_the_edge_list():
    sim, cfg = _pair((32, 32), "periodic")
    rep = cpm.Replica(sim)
    snap = rep.snapshot()
    built = rep.edge_sites()
    order = np.random.default_rng(3).permutation(built)
    rep.set_state(snap["lattice"].ravel(), snap["type_index"], edge_sites=order)
    assert np.array_equal(rep.edge_sites(), order)
    rep.set_state(snap["lattice"].ravel(), snap["type_index"])
    assert np.array_equal(rep.edge_sites(), built)
    with pytest.raises(ValueError, match=re.escape(f"edge site {order[0]} is listed twice")):
        rep.set_state(snap["lattice"].ravel(), snap["type_index"],
                      edge_sites=np.append(order, order[0]))
    with pytest.raises(ValueError, match="is not a site"):
        rep.set_state(snap["lattice"].ravel(), snap["type_index"], edge_sites=[32 * 32])
    # the failed imports left the model as it was
    assert np.array_equal(rep.edge_sites(), built)
