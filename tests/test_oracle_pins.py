"""Pins of the oracle's initialisation, labelling and status rules against
results the reference repository itself holds (CPU only):

* the 160 lattices of notebooks/2d_organoid_demo.ipynb (scatter-and-grow
  initialisation followed by 400000 MCS; decoded into
  tests/golden/organoid2d_frames.npz by tests/golden/make_organoid_frames.py)
  and the results that notebook prints (cell 5);
* the hand-built lattices of labelling.rs:174-336;
* the not_equilibrated cases of output.rs:614-627 and the end-to-end run of
  output.rs:699-810;
* hand-built sample sequences for the precedence of the status rules
  (output.rs:470-509).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FRAMES = np.load(os.path.join(HERE, "golden", "organoid2d_frames.npz"))

ORGANOID_J = [
    [0.0, 2.5, 2.5, 2.5, 2.5],
    [2.5, 0.05, 4.0, 4.0, 4.0],
    [2.5, 4.0, 0.05, 4.0, 4.0],
    [2.5, 4.0, 4.0, 0.05, 4.0],
    [2.5, 4.0, 4.0, 4.0, 0.05],
]
ORGANOID_TYPES = ["stem", "progenitor", "differentiated", "surface"]


def organoid_config(readout_mcs=400000):
    """notebooks/2d_organoid_demo.ipynb cells 3 and 5."""
    return O.Config(grid=(70, 70), boundary="fixed", seed=11,
                    cell_types=[O.CellTypeSpec(n, 110, 135, 1.2, 0.7) for n in ORGANOID_TYPES],
                    cell_counts=[6, 6, 6, 6], adhesion=ORGANOID_J, burn_in_mcs=0,
                    readout_mcs=readout_mcs, sampling_interval_mcs=2500)


def test_organoid_notebook_first_frames_site_for_site():
    """scatter_and_grow (initialization.rs:202-330) and the first 20000 MCS:
    every site of the first eight sampled lattices equals the notebook's."""
    m = O.Model(organoid_config())
    assert m.used_default
    for k in range(8):
        m.run_mcs(2500)
        assert m.mcs == int(FRAMES["mcs"][k])
        diff = int((m.lattice() != FRAMES["lattices"][k]).sum())
        assert diff == 0, f"frame {k}: {diff} sites differ"


@pytest.mark.slow
def test_organoid_notebook_full_run_matches_every_frame_and_the_printed_results():
    """The notebook's whole run_with (400000 MCS, ~2e9 attempts): all 160
    lattices site for site, and cell 5's printed status, derived_phi and
    frame count."""
    m = O.Model(organoid_config())
    r = O.run_with(m, include_lattice=True)
    printed = str(FRAMES["printed"])
    assert "status: Degenerate" in printed and "derived_phi: 0.537" in printed
    assert "frames captured: 160" in printed
    assert r.status[0] == "Degenerate"
    assert round(r.derived_phi, 3) == 0.537
    assert len(r.samples) == 160 == len(FRAMES["lattices"])
    for k, s in enumerate(r.samples):
        assert s.mcs == int(FRAMES["mcs"][k])
        assert np.array_equal(s.lattice, FRAMES["lattices"][k]), f"frame {k} differs"


# ------------------------------------------------------------ labelling

def _model_with(grid, boundary, lattice, cell_type_of):
    """A model holding an arbitrary lattice (labelling needs no dynamics)."""
    n = int(np.prod(grid))
    seed_lat = [0] * n
    seed_lat[0] = 1
    cfg = O.Config(grid=grid, boundary=boundary, seed=1,
                   cell_types=[O.CellTypeSpec("a", 10, 10, 1.0, 1.0)], cell_counts=[1],
                   adhesion=[[0.0, 1.0], [1.0, 1.0]], init_lattice=seed_lat,
                   init_cell_type_of=[0])
    m = O.Model(cfg)
    m.load_state(np.asarray(lattice, dtype=np.uint32).ravel(), cell_type_of)
    return m


def ring_2d(grid=(9, 9), lo=2, hi=6, cid=1, lat=None):
    lat = np.zeros(grid, dtype=np.uint32) if lat is None else lat
    for x in range(lo, hi + 1):
        for y in range(lo, hi + 1):
            if x in (lo, hi) or y in (lo, hi):
                lat[x, y] = cid
    return lat


def lumens(medium_volumes, touches):
    return [v for v, t in zip(medium_volumes, touches) if not t]


def labelling_cases():
    """labelling.rs:174-336: (name, grid, boundary, lattice, cell_type_of,
    expected fragmented, expected lumen volumes (None: count only),
    expected lumen count, expected medium component count or None)."""
    cases = []
    blob = np.zeros((7, 7), dtype=np.uint32)
    blob[2:5, 2:5] = 1
    cases.append(("solid_blob", (7, 7), "fixed", blob, [0], [], [], 0, None))
    cases.append(("hollow_ring", (9, 9), "fixed", ring_2d(), [0], [], [9], 1, None))
    breach = ring_2d()
    breach[4, 2] = 0
    cases.append(("breached_ring", (9, 9), "fixed", breach, [0], None, [], 0, None))
    two = np.zeros((15, 9), dtype=np.uint32)
    for ox, oy, cid in ((1, 2, 1), (8, 2, 2)):
        for x in range(ox, ox + 5):
            for y in range(oy, oy + 5):
                if x in (ox, ox + 4) or y in (oy, oy + 4):
                    two[x, y] = cid
    cases.append(("two_shells", (15, 9), "fixed", two, [0, 0], [], None, 2, None))
    cases.append(("periodic_ring", (9, 9), "periodic", ring_2d(), [0], None, None, 2, 2))
    pieces = np.zeros((9, 9), dtype=np.uint32)
    pieces[1, 1] = pieces[6, 6] = 1
    cases.append(("two_pieces", (9, 9), "fixed", pieces, [0], [1], None, None, None))
    cases.append(("ring_cell", (9, 9), "fixed", ring_2d(), [0], [], None, None, None))
    shell = np.zeros((9, 9, 9), dtype=np.uint32)
    for x in range(2, 7):
        for y in range(2, 7):
            for z in range(2, 7):
                if x in (2, 6) or y in (2, 6) or z in (2, 6):
                    shell[x, y, z] = 1
    cases.append(("hollow_shell_3d", (9, 9, 9), "fixed", shell, [0], [], [27], 1, None))
    return cases


CASES = labelling_cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_labelling_hand_built_cases(case):
    name, grid, boundary, lat, cto, frag, lumen_vols, n_lumens, n_medium = case
    m = _model_with(grid, boundary, lat, cto)
    got_frag, mv, mt = m.label()
    if frag is not None:
        assert got_frag == frag
    got_lumens = lumens(mv, mt)
    if lumen_vols is not None:
        assert sorted(got_lumens) == sorted(lumen_vols)
    if n_lumens is not None:
        assert len(got_lumens) == n_lumens
    if n_medium is not None:
        assert len(mv) == n_medium
        assert not any(mt), "no medium component touches a periodic boundary"


# ------------------------------------------------------------ status rules

def test_not_equilibrated_flags_a_clear_trend():
    """output.rs:614-618"""
    assert O.not_equilibrated([i * 10.0 for i in range(20)])


def test_not_equilibrated_does_not_flag_a_noisy_plateau():
    """output.rs:620-627"""
    assert not O.not_equilibrated([100.0, 101.0, 99.5, 100.3, 99.8, 100.1, 99.9, 100.2, 100.0,
                                   99.7])


def test_not_equilibrated_needs_three_points_and_a_flat_trace_is_fine():
    assert not O.not_equilibrated([1.0, 100.0])
    assert not O.not_equilibrated([5.0] * 10)


def _sample(mcs, volumes, fragmented=(), contact_pairs=()):
    n = len(volumes)
    contact = np.zeros((n + 1, n + 1), dtype=np.uint32)
    for a, b in contact_pairs:
        contact[min(a, b), max(a, b)] = 1
    z = np.zeros(n)
    return O.Sample(mcs, 0.0, np.array(volumes, dtype=np.uint32), z, z, z, z, z, contact,
                    list(fragmented), [], [])


def test_status_rule_order_on_hand_built_samples():
    """output.rs:470-509: NotEquilibrated wins over everything; then the first
    fragmented sample; then a cell at the volume floor over the trailing
    half; then a last sample whose contact graph joins every cell."""
    trend = [i * 10.0 for i in range(20)]
    joined = [(1, 2), (2, 3)]
    ok = [_sample(10, [5, 5, 5]), _sample(20, [5, 5, 5])]
    assert O.status_of(ok, [], 3) == ("Ok", None, None)
    frag = [_sample(10, [5, 5, 5]), _sample(20, [5, 5, 5], [3, 1]), _sample(30, [5, 5, 5], [2])]
    assert O.status_of(frag, [], 3) == ("Fragmented", 3, 20)
    assert O.status_of(frag, trend, 3) == ("NotEquilibrated", None, None)
    # 4 samples: the trailing window is ceil(4 * 0.5) = 2 samples
    lost = [_sample(10, [5, 1, 5]), _sample(20, [5, 9, 5]), _sample(30, [5, 1, 1]),
            _sample(40, [5, 0, 1])]
    assert O.status_of(lost, [], 3) == ("CellLost", 2, 30)
    # 3 samples: ceil(1.5) = 2 trailing samples; cell 2 recovered in between
    assert O.status_of(lost[:3], [], 3) == ("Ok", None, None)
    lost3 = [_sample(10, [5, 5, 5]), _sample(20, [5, 5, 1]), _sample(30, [5, 5, 0])]
    assert O.status_of(lost3, [], 3) == ("CellLost", 3, 20)
    degenerate = [_sample(10, [5, 5, 5]), _sample(20, [5, 5, 5], (), joined)]
    assert O.status_of(degenerate, [], 3) == ("Degenerate", None, None)
    not_joined = [_sample(20, [5, 5, 5], (), [(1, 2)])]
    assert O.status_of(not_joined, [], 3) == ("Ok", None, None)
    assert O.status_of([_sample(20, [5, 1, 5], (), joined)], [], 3)[0] == "CellLost"
    assert O.status_of([_sample(20, [5], (), [])], [], 1) == ("Ok", None, None)


def test_run_end_to_end_status_and_sample_count():
    """output.rs:699-810: 24x24 fixed, seed 7, two types x two cells,
    burn-in 20 / readout 20 / interval 5 -> status Ok, 4 samples of 4 cells,
    no lattice unless asked for."""
    cfg = O.Config(grid=(24, 24), boundary="fixed", seed=7,
                   cell_types=[O.CellTypeSpec("a", 12, 24, 1.0, 0.2),
                               O.CellTypeSpec("b", 12, 24, 1.0, 0.2)],
                   cell_counts=[2, 2],
                   adhesion=[[0.0, 1.0, 1.0], [1.0, 0.5, 1.0], [1.0, 1.0, 0.5]],
                   burn_in_mcs=20, readout_mcs=20, sampling_interval_mcs=5)
    m = O.Model(cfg)
    r = O.run_with(m)
    assert r.status == ("Ok", None, None)
    assert len(r.samples) == 4
    assert all(len(s.volume) == 4 for s in r.samples)
    assert r.samples[0].lattice is None
    assert r.model_hash == m.model_hash()
    assert r.convention_hash == O.convention_hash(cfg)


def test_derived_phi_matches_hand_computation():
    """output.rs:630-660"""
    cfg = O.Config(grid=(10, 10), boundary="fixed", seed=1,
                   cell_types=[O.CellTypeSpec("a", 9, 20, 1.0, 0.1)], cell_counts=[1],
                   adhesion=[[0.0, 0.0], [0.0, 0.0]], burn_in_mcs=0, readout_mcs=1,
                   sampling_interval_mcs=1)
    m = O.Model(cfg)
    expected_before = float(m.volume()[0]) / 100.0
    assert expected_before == 0.09
    r = O.run_with(m)
    assert abs(r.derived_phi - float(m.volume()[0]) / 100.0) < 1e-12


def test_include_lattice_snapshot_matches_tracked_volumes():
    """output.rs:812-870: 16x16 fixed, seed 3, two cells."""
    cfg = O.Config(grid=(16, 16), boundary="fixed", seed=3,
                   cell_types=[O.CellTypeSpec("a", 10, 20, 1.0, 0.2)], cell_counts=[2],
                   adhesion=[[0.0, 1.0], [1.0, 0.5]], burn_in_mcs=10, readout_mcs=10,
                   sampling_interval_mcs=5)
    r = O.run_with(O.Model(cfg), include_lattice=True)
    for s in r.samples:
        assert s.lattice.size == 256
        counts = np.bincount(s.lattice.ravel(), minlength=3)
        assert counts[1:].tolist() == s.volume.tolist()
