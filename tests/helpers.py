"""Shared fixtures: the reference's test configurations, built for both the
product (`cpm`) and the oracle."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PINS = json.load(open(os.path.join(HERE, "golden", "reference_pins.json")))


def frozen_lattice():
    f = PINS["frozen_trajectory"]
    g = f["grid"][1]
    lat = [0] * (f["grid"][0] * f["grid"][1])
    for cid, cells in f["blocks"].items():
        for r, c in cells:
            lat[r * g + c] = int(cid)
    return lat


def frozen_sim(cpm, seed=None):
    """crates/cpm-core/tests/dimension_consistency.rs:53-100 through the
    Python builder (tests/python/test_reproducibility.py:37-54)."""
    f = PINS["frozen_trajectory"]
    sim = cpm.CPM(grid=tuple(f["grid"]), boundary="periodic",
                  seed=f["seed"] if seed is None else seed,
                  copy_neighborhood="von_neumann", connectivity_neighborhood="von_neumann",
                  acceptance="metropolis")
    for name, tv, ti, lv, li in f["cell_types"]:
        sim.register_cell_type(name, target_volume=tv, target_interface=ti,
                               lambda_volume=lv, lambda_interface=li)
    sim.add_cells("a", 1)
    sim.add_cells("b", 1)
    sim.set_adhesion(f["adhesion"])
    sim.initialize(lattice=frozen_lattice(), cell_type_of=f["cell_type_of"])
    return sim


def frozen_oracle_config(O, seed=None):
    f = PINS["frozen_trajectory"]
    return O.Config(grid=tuple(f["grid"]), boundary="periodic",
                    seed=f["seed"] if seed is None else seed,
                    cell_types=[O.CellTypeSpec(*t) for t in f["cell_types"]],
                    cell_counts=[1, 1], adhesion=f["adhesion"], init_lattice=frozen_lattice(),
                    init_cell_type_of=f["cell_type_of"], burn_in_mcs=0, readout_mcs=1,
                    sampling_interval_mcs=1)


def quickstart(cpm, seed=1, grid=(30, 30)):
    """README quickstart (README.md:47-83; tests/python/test_builder.py:11-28)."""
    sim = cpm.CPM(grid=grid, boundary="periodic", seed=seed,
                  copy_neighborhood="von_neumann", connectivity_neighborhood="von_neumann")
    sim.register_cell_type(name="epithelial", target_volume=50, target_interface=75,
                           lambda_volume=10.0, lambda_interface=2.0)
    sim.add_cells(cell_type="epithelial", n=5)
    sim.set_adhesion([[0.0, 5.0], [5.0, 2.0]])
    return sim


def quickstart_oracle(O, seed=1, grid=(30, 30), windows=(0, 1, 1)):
    return O.Config(grid=grid, boundary="periodic", seed=seed,
                    cell_types=[O.CellTypeSpec("epithelial", 50, 75, 10.0, 2.0)],
                    cell_counts=[5], adhesion=[[0.0, 5.0], [5.0, 2.0]],
                    burn_in_mcs=windows[0], readout_mcs=windows[1],
                    sampling_interval_mcs=windows[2])


def sorting_pair(cpm, O, seed, grid=(48, 48), n_types=3, boundary="periodic", phi=0.7,
                 tv=30, ti=24, energy=None):
    """A multi-type differential-adhesion configuration built identically
    for the product and the oracle (J drawn from numpy's default_rng(seed))."""
    rng = np.random.default_rng(seed)
    T = n_types
    J = np.zeros((T + 1, T + 1))
    J[0, 1:] = J[1:, 0] = 16.0
    for a in range(1, T + 1):
        for b in range(a, T + 1):
            J[a, b] = J[b, a] = float(rng.uniform(2, 14))
    n_sites = int(np.prod(grid))
    n = max(T, round(phi * n_sites / tv))
    counts = [n // T + (1 if i < n % T else 0) for i in range(T)]
    names = [f"t{i}" for i in range(T)]
    sim = cpm.CPM(grid=grid, boundary=boundary, seed=seed, energy_neighborhood=energy)
    for nm in names:
        sim.register_cell_type(nm, target_volume=tv, target_interface=ti, lambda_volume=5.0,
                               lambda_interface=1.0)
    for nm, c in zip(names, counts):
        sim.add_cells(nm, c)
    sim.set_adhesion(J.tolist())
    sim.initialize(warn_on_default_init=False)
    cfg = O.Config(grid=grid, boundary=boundary, seed=seed,
                   cell_types=[O.CellTypeSpec(nm, tv, ti, 5.0, 1.0) for nm in names],
                   cell_counts=counts, adhesion=J.tolist(), energy_neighborhood=energy,
                   warn_on_default_init=False)
    return sim, cfg
