"""Host-side behaviour of the product library (CPU only): the C ABI loads
and exports what include/cpm_engine.h declares, and the ctypes mirrors
of its structs have the header's layout; configuration errors,
initial placement, theta and run-record hashes match the reference /
oracle.  Nothing here launches a kernel."""
import ctypes as C
import os
import re
import subprocess
import warnings

import numpy as np
import pytest

import cpm
from cpm import _native as N
from oracle import oracle as O
from tests.helpers import PINS, frozen_oracle_config, frozen_sim, quickstart, quickstart_oracle, \
    sorting_pair

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "cpm_engine.h")).read()
    declared = set(re.findall(r"\b(cpm_[a-z0-9_]+)\s*\(", header))
    lib = C.CDLL(N.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert len(declared) >= 20


_MIRRORS = {"cpm_config": N.CpmConfig, "cpm_counters": N.CpmCounters,
            "cpm_model_info": N.CpmModelInfo, "cpm_model_state": N.CpmModelState,
            "cpm_result_info": N.CpmResultInfo, "cpm_sample_view": N.CpmSampleView,
            "cpm_samples_shape": N.CpmSamplesShape, "cpm_samples_buffers": N.CpmSamplesBuffers,
            "cpm_server_options": N.CpmServerOptions, "cpm_request_info": N.CpmRequestInfo,
            "cpm_server_stats": N.CpmServerStats}


def test_ctypes_structs_match_the_header(tmp_path):
    """Every struct of include/cpm_engine.h has a ctypes mirror with the same
    fields at the same offsets and the same size, as the C compiler lays the
    header out (a mirror that drifts corrupts memory silently)."""
    header = open(os.path.join(ROOT, "include", "cpm_engine.h")).read()
    declared = set(re.findall(r"typedef struct \{[^}]*\}\s*(cpm_[a-z_]+)\s*;", header))
    assert declared == set(_MIRRORS), declared ^ set(_MIRRORS)
    lines = ['#include <stdio.h>', '#include "cpm_engine.h"', "int main(void) {"]
    for name, mirror in _MIRRORS.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for field, _ in mirror._fields_:
            lines.append(f'printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ["return 0;", "}"]
    src, exe = tmp_path / "layout.c", tmp_path / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in
               subprocess.run([str(exe)], check=True, capture_output=True,
                              text=True).stdout.splitlines())
    for name, mirror in _MIRRORS.items():
        assert int(got[f"{name} size"]) == C.sizeof(mirror), name
        for field, _ in mirror._fields_:
            assert int(got[f"{name} {field}"]) == getattr(mirror, field).offset, (name, field)


def _same_init(sim, cfg):
    rep = cpm.Replica(sim)
    ref = O.Model(cfg)
    s, o = rep.snapshot(), ref.snapshot()
    for k in ("lattice", "volume", "interface", "sum_position", "sum_outer", "contact", "rng"):
        assert np.array_equal(np.asarray(s[k]), np.asarray(o[k])), k
    assert s["energy"] == o["energy"]
    assert rep.global_energy() == ref.global_energy()


def test_explicit_init_matches_oracle():
    _same_init(frozen_sim(cpm), frozen_oracle_config(O))


@pytest.mark.parametrize("seed", [1, 2, 99])
def test_scatter_and_grow_matches_oracle(seed):
    sim = quickstart(cpm, seed=seed)
    sim.initialize(warn_on_default_init=False)
    _same_init(sim, quickstart_oracle(O, seed=seed))


@pytest.mark.parametrize("args", [
    dict(grid=(48, 32), n_types=3),
    dict(grid=(40, 40), n_types=2, boundary="fixed"),
    dict(grid=(12, 14, 10), n_types=2, tv=40, ti=80, phi=0.5),
    dict(grid=(12, 12, 12), n_types=3, tv=30, ti=60, phi=0.6, boundary="fixed"),
])
def test_multitype_init_matches_oracle(args):
    sim, cfg = sorting_pair(cpm, O, 5, **args)
    _same_init(sim, cfg)


def test_derived_seeds_from_the_product():
    d = PINS["derived_seeds"]
    assert [cpm.derive_seed_u64(d["master_seed"], i) for i in range(5)] == d["values"]


def test_theta_names_and_extract_match_reference_layout():
    """tests/python/test_batch.py:29-42 and batch_execution_demo.ipynb:124-138"""
    sim = cpm.CPM(grid=(20, 20), boundary="periodic", seed=1)
    sim.register_cell_type("a", target_volume=9, target_interface=20, lambda_volume=1.0,
                           lambda_interface=1.0)
    sim.register_cell_type("b", target_volume=9, target_interface=20, lambda_volume=1.0,
                           lambda_interface=1.0)
    sim.set_adhesion([[0.0, 1.0, 1.0], [1.0, 0.2, 1.0], [1.0, 1.0, 0.2]])
    assert sim.theta_names() == [
        "lambda_volume[a]", "lambda_volume[b]", "lambda_interface[a]", "lambda_interface[b]",
        "J[medium][medium]", "J[medium][a]", "J[medium][b]", "J[a][a]", "J[a][b]", "J[b][b]",
        "lambda_act[a]", "lambda_act[b]", "max_act[a]", "max_act[b]"]
    assert sim.extract_theta() == [1.0, 1.0, 1.0, 1.0, 0.0, 1.0, 1.0, 0.2, 1.0, 0.2, 0.0, 0.0,
                                   0.0, 0.0]


# --- tests/python/test_errors.py:8-51 ---------------------------------------

def test_unsupported_grid_length_raises_value_error():
    with pytest.raises(ValueError, match="length 2 or 3"):
        cpm.CPM(grid=(1, 2, 3, 4), boundary="periodic", seed=1)


def test_unknown_neighborhood_string_raises_value_error():
    with pytest.raises(ValueError, match="unknown neighbourhood"):
        cpm.CPM(grid=(10, 10), boundary="periodic", seed=1, copy_neighborhood="hexagonal")


def test_unknown_boundary_string_raises_value_error():
    with pytest.raises(ValueError, match="unknown boundary"):
        cpm.CPM(grid=(10, 10), boundary="squishy", seed=1)


def test_non_square_adhesion_matrix_raises_before_any_device_work():
    sim = cpm.CPM(grid=(10, 10), boundary="periodic", seed=1)
    sim.register_cell_type(name="a", target_volume=20, target_interface=50, lambda_volume=1.0,
                           lambda_interface=1.0)
    sim.add_cells(cell_type="a", n=1)
    sim.set_adhesion([[0.0, 1.0], [1.0, 0.0], [0.0, 0.0]])
    with pytest.raises(ValueError, match=r"adhesion matrix must be 2x2 \(n_types \+ 1 for "
                                         r"medium\), got 3x2"):
        sim.run(burn_in_mcs=1, readout_mcs=1, sampling_interval_mcs=1)


def test_positive_lambda_act_with_zero_max_act_raises_immediately():
    sim = cpm.CPM(grid=(10, 10), boundary="periodic", seed=1)
    with pytest.raises(ValueError, match="max_act"):
        sim.register_cell_type(name="a", target_volume=20, target_interface=50,
                               lambda_volume=1.0, lambda_interface=1.0, lambda_act=1.0, max_act=0)


def test_unknown_cell_type_in_add_cells_raises_value_error():
    sim = cpm.CPM(grid=(10, 10), boundary="periodic", seed=1)
    with pytest.raises(ValueError, match="unknown cell type"):
        sim.add_cells(cell_type="ghost", n=1)


@pytest.mark.parametrize("kw,msg", [
    (dict(copy_neighborhood="moore"),
     "copy_neighborhood Moore must be a subset of connectivity_neighborhood VonNeumann"),
    (dict(connectivity_neighborhood="moore"),
     "connectivity_neighborhood must be VonNeumann in this version"),
    (dict(energy_neighborhood="eighteen"), "Eighteen-neighbour is 3D only, got D = 2"),
])
def test_validation_messages_match_config_rs(kw, msg):
    """config.rs:287-413 and neighborhood.rs:179-263"""
    sim = cpm.CPM(grid=(10, 10), boundary="periodic", seed=1, **kw)
    sim.register_cell_type("a", 20, 50, 1.0, 1.0)
    sim.add_cells("a", 1)
    sim.set_adhesion([[0.0, 1.0], [1.0, 0.0]])
    with pytest.raises(ValueError, match=re.escape(msg)):
        cpm.Replica(sim)


def test_asymmetric_adhesion_message():
    sim = cpm.CPM(grid=(10, 10), boundary="periodic", seed=1)
    sim.register_cell_type("a", 20, 50, 1.0, 1.0)
    sim.add_cells("a", 1)
    sim.set_adhesion([[0.0, 1.0], [2.5, 0.0]])
    with pytest.raises(ValueError, match=re.escape(
            "adhesion matrix must be symmetric in v1: [0][1] = 1 but [1][0] = 2.5")):
        cpm.Replica(sim)


def test_init_errors_match_initialization_rs():
    sim = cpm.CPM(grid=(10, 10), boundary="periodic", seed=1)
    sim.register_cell_type("a", 20, 50, 1.0, 1.0)
    sim.set_adhesion([[0.0, 1.0], [1.0, 0.0]])
    with pytest.raises(ValueError, match="cell_counts sums to zero cells"):
        cpm.Replica(sim)
    lat = np.zeros(100, dtype=int)
    lat[0] = 1
    lat[55] = 1
    sim.initialize(lattice=lat, cell_type_of=[0])
    with pytest.raises(ValueError, match="cell 1 is not connected under connectivity_neighborhood"):
        cpm.Replica(sim)
    sim.add_cells("a", 200)
    sim.initialize(warn_on_default_init=False)
    with pytest.raises(ValueError, match=re.escape(
            "could not place all 200 seed points with minimum separation 2.52 on a grid of "
            "[10, 10]")):
        cpm.Replica(sim)


def test_wrong_length_theta_raises_before_device_work():
    sim = cpm.CPM(grid=(20, 20), boundary="periodic", seed=1)
    sim.register_cell_type("a", 9, 20, 1.0, 0.2)
    sim.add_cells("a", 2)
    sim.set_adhesion([[0.0, 3.0], [3.0, 1.0]])
    base = sim.extract_theta()
    with pytest.raises(ValueError, match="theta at index 1 has length 6, expected 7"):
        sim.run_batch([base, base[:-1]], burn_in_mcs=5, readout_mcs=5, sampling_interval_mcs=5,
                      master_seed=1)


def test_convention_and_model_hash_match_oracle():
    """config.rs:423-556: the product's serde_json bytes and BLAKE3 agree
    with the oracle's, whose model_hash is pinned by the frozen trajectory."""
    f = PINS["frozen_trajectory"]
    rep = cpm.Replica(frozen_sim(cpm), *f["windows"])
    cfg = frozen_oracle_config(O)
    conv, model = rep.hashes()
    assert conv == O.convention_hash(cfg)
    snap = rep.snapshot()
    assert model == O.model_hash(cfg, snap["lattice"].ravel(), snap["type_index"])
    sim = quickstart(cpm, seed=4)
    sim.initialize(warn_on_default_init=False)
    rep2 = cpm.Replica(sim, 10, 20, 5)
    cfg2 = quickstart_oracle(O, seed=4, windows=(10, 20, 5))
    cfg2.warn_on_default_init = False
    assert rep2.hashes() == (O.convention_hash(cfg2), O.Model(cfg2).model_hash())


def test_cells_for_density():
    """tests/python/test_density_helper.py"""
    n = cpm.cells_for_density(grid=(64, 64), target_volume=100, phi=0.8)
    assert n * 100 / 4096 == pytest.approx(0.8, abs=0.02)
    with pytest.warns(cpm.CPMLowResolutionWarning):
        cpm.cells_for_density(grid=(64, 64), target_volume=4, phi=0.8)
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        cpm.cells_for_density(grid=(64, 64), target_volume=1000, phi=0.8)
        assert not any(issubclass(w.category, cpm.CPMLowResolutionWarning) for w in caught)
    assert cpm.cells_for_density((256, 256), 40, 0.8) == 1311
