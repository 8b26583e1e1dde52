"""Readouts on the device (the labelling, per-cell centroid and gyration
tensor, 2D shape descriptors; output.rs:147-337, labelling.rs:70-154)
against the oracle's take_sample, bit for bit: the reference's hand-built
labelling lattices (labelling.rs:174-336) imported with
cpm_model_set_state and made resident, and lattices after runs of 2D/3D,
fixed/periodic configurations, including many-cell and lane-path ones."""
import numpy as np
import pytest

import cpm
from oracle import oracle as O
from tests.helpers import sorting_pair
from tests.test_host_state import _sim
from tests.test_oracle_pins import CASES, _model_with

pytestmark = pytest.mark.gpu


def _same_sample(s, o):
    assert np.array_equal(s.volume, o.volume)
    assert np.array_equal(s.interface, o.interface)
    assert np.array_equal(s.contact, o.contact)
    # bit for bit (both sides compute the same IEEE operations)
    assert s.centroid.tobytes() == np.asarray(o.centroid, dtype=np.float64).tobytes()
    assert s.eigenvalues.tobytes() == np.asarray(o.eigenvalues, dtype=np.float64).tobytes()
    assert s.radius_of_gyration_sq.tobytes() == np.asarray(o.rg2, dtype=np.float64).tobytes()
    assert s.kappa_sq.tobytes() == np.asarray(o.kappa, dtype=np.float64).tobytes()
    assert s.fragmented_cells == list(o.fragmented)
    assert s.medium_component_volumes == list(o.medium_volumes)
    assert s.medium_component_touches_boundary == [bool(x) for x in o.medium_touch]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_device_labelling_of_the_reference_cases(case):
    name, grid, boundary, lat, cto, *_ = case
    rep = cpm.Replica(_sim(grid, boundary), initialize=False)
    rep.set_state(np.asarray(lat).ravel(), cto)
    rep.run_mcs(0)  # resident on the device, state unchanged
    s = rep.sample().samples[0]
    of, omv, omt = _model_with(grid, boundary, lat, cto).label()
    print(f"[readout] {name}: fragmented {s.fragmented_cells}, medium {s.medium_component_volumes}",
          flush=True)
    assert s.fragmented_cells == of
    assert s.medium_component_volumes == omv
    assert s.medium_component_touches_boundary == omt


RUNS = [
    ("2d_periodic", dict(grid=(48, 48), n_types=3), 20),
    ("2d_fixed_sparse", dict(grid=(64, 64), n_types=2, boundary="fixed", phi=0.3), 20),
    ("2d_rect", dict(grid=(40, 72), n_types=2, phi=0.5), 15),
    ("3d_periodic", dict(grid=(20, 20, 20), n_types=2, tv=40, ti=60, phi=0.5), 8),
    ("3d_fixed", dict(grid=(18, 22, 16), n_types=2, boundary="fixed", tv=40, ti=60, phi=0.4), 8),
    ("2d_lane_path", dict(grid=(150, 150), n_types=2, tv=4, ti=8, phi=0.8), 3),
]


@pytest.mark.parametrize("label,kw,mcs", RUNS, ids=[r[0] for r in RUNS])
def test_device_readouts_equal_the_oracles_take_sample(label, kw, mcs):
    sim, cfg = sorting_pair(cpm, O, 61, **kw)
    rep = cpm.Replica(sim)
    rep.run_mcs(mcs)
    ref = O.Model(cfg)
    ref.run_mcs(mcs)
    s = rep.sample(include_lattice=True).samples[0]
    o = O.take_sample(ref, include_lattice=True)
    print(f"[readout] {label}: {len(s.volume)} cells, {len(s.medium_component_volumes)} medium "
          f"components, fragmented {s.fragmented_cells}", flush=True)
    assert np.array_equal(s.lattice, o.lattice)
    _same_sample(s, o)
