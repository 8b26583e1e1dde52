"""Every instantiation of the exact-replay wave kernel against the oracle.

The kernel is instantiated per (dimension, energy stencil) x sigma in LDS or
in HBM x one wave or pipelined (two waves) x untimed or time-limited
(serving rounds): 40 variants.  Each energy stencil runs here in a batch
(a long chain, which runs pipelined, beside short ones, which run on one
wave) with sigma placed in LDS and then in HBM, and in serving rounds of
0.5 ms (time-limited; 2D and 3D requests mixed) with pipelining on and off.
Every result is compared with the oracle bit for bit, and the last test
checks, from the engine's own launch counts, that all 40 variants ran."""
import struct

import numpy as np
import pytest

import cpm
from cpm import _native as N
from oracle import oracle as O
from tests.helpers import sorting_pair

pytestmark = pytest.mark.gpu

# (dim, energy stencil, energy-stencil size, big grid, small grid, V*, I*, phi)
STENCILS = [
    (2, "von_neumann", 4, (40, 40), (20, 20), 20, 16, 0.7),
    (2, "moore", 8, (40, 40), (20, 20), 20, 16, 0.7),
    (3, "von_neumann", 6, (16, 16, 16), (10, 10, 10), 27, 30, 0.5),
    (3, "eighteen", 18, (16, 16, 16), (10, 10, 10), 27, 60, 0.5),
    (3, "twenty_six", 26, (16, 16, 16), (10, 10, 10), 27, 70, 0.5),
]
PLACEMENTS = {"lds": ("smallest", 8.0), "hbm": ("smallest", 1e-6)}


@pytest.fixture(autouse=True)
def _engine_defaults():
    yield
    N.set_sigma_placement("auto")
    N.set_pipelining(True)


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def _pair(st, big, seed, boundary="periodic"):
    dim, energy, ne, g_big, g_small, tv, ti, phi = st
    return sorting_pair(cpm, O, seed, grid=g_big if big else g_small, n_types=2,
                        boundary=boundary, phi=phi, tv=tv, ti=ti, energy=energy)


@pytest.mark.parametrize("place", list(PLACEMENTS))
@pytest.mark.parametrize("st", STENCILS, ids=[f"{s[0]}d_{s[1]}" for s in STENCILS])
def test_batch_of_uneven_chains_matches_oracle(st, place):
    """One long chain (pipelined) and three short ones (one wave each)."""
    N.set_sigma_placement(*PLACEMENTS[place])
    pairs = [_pair(st, i == 0, 500 + 10 * i, "fixed" if i == 3 else "periodic")
             for i in range(4)]
    print(f"[case] batch {st[0]}D {st[1]} sigma in {place}", flush=True)
    reps = [cpm.Replica(sim) for sim, _ in pairs]
    cpm.Replica.run_mcs_many(reps, 25)
    for (sim, cfg), rep in zip(pairs, reps):
        ref = O.Model(cfg)
        ref.run_mcs(25)
        got, want = rep.snapshot(), ref.snapshot()
        for k in ("lattice", "volume", "interface", "sum_position", "sum_outer", "contact",
                  "rng"):
            assert np.array_equal(got[k], want[k]), (cfg.grid, k)
        assert _bits(got["energy"]) == _bits(want["energy"])


SERVE_WINDOWS = dict(burn_in_mcs=6, readout_mcs=12, sampling_interval_mcs=6)


@pytest.mark.parametrize("pipelining", [True, False], ids=["pipelined", "one_wave"])
@pytest.mark.parametrize("place", list(PLACEMENTS))
def test_time_limited_serving_mix_of_2d_and_3d_matches_oracle(place, pipelining):
    """2D and 3D requests of every energy stencil served together in 0.5 ms
    rounds (chains stop mid-MCS and resume), sigma in LDS or HBM: each
    result equals the oracle's run_with."""
    N.set_sigma_placement(*PLACEMENTS[place])
    N.set_pipelining(pipelining)
    pairs = [_pair(st, big, 900 + 7 * k + (1 if big else 0))
             for k, st in enumerate(STENCILS) for big in (True, False)]
    print(f"[case] serving mix, sigma in {place}, pipelining {pipelining}", flush=True)
    with cpm.Server(round_ms=0.5) as srv:
        reqs = [srv.submit(sim, **SERVE_WINDOWS, include_lattice=True) for sim, _ in pairs]
        results = [r.result(300) for r in reqs]
        st = srv.stats()
    assert st["failed"] == 0 and st["rounds"] > 10
    for (sim, cfg), res in zip(pairs, results):
        cfg.burn_in_mcs = SERVE_WINDOWS["burn_in_mcs"]
        cfg.readout_mcs = SERVE_WINDOWS["readout_mcs"]
        cfg.sampling_interval_mcs = SERVE_WINDOWS["sampling_interval_mcs"]
        ref = O.run_with(O.Model(cfg), include_lattice=True)
        assert res.metadata.model_hash == ref.model_hash, cfg.grid
        assert res.status.kind == ref.status[0]
        assert len(res.samples) == len(ref.samples)
        for a, b in zip(res.samples, ref.samples):
            assert a.mcs == b.mcs
            assert np.array_equal(a.lattice, b.lattice)
            assert np.array_equal(a.volume, b.volume)
            assert np.array_equal(a.contact, b.contact)
            assert _bits(a.conservative_energy) == _bits(b.energy)


def test_3d_lane_path_matches_oracle():
    """A 3D replica with more cells than the packed sigma words hold (one
    lane per replica, state in HBM)."""
    sim, cfg = sorting_pair(cpm, O, 77, grid=(26, 26, 26), n_types=2, phi=0.8, tv=3, ti=9)
    ref = O.Model(cfg)
    assert ref.n_cells > 4095
    print(f"[case] 3D lane path, {ref.n_cells} cells", flush=True)
    rep = cpm.Replica(sim)
    rep.run_mcs(6)
    ref.run_mcs(6)
    got, want = rep.snapshot(), ref.snapshot()
    for k in ("lattice", "volume", "interface", "sum_position", "sum_outer", "contact", "rng"):
        assert np.array_equal(got[k], want[k]), k
    assert _bits(got["energy"]) == _bits(want["energy"])


def _variant_name(key):
    variant, base = divmod(key, 2000)
    d, ne = (base % 1000) // 100, base % 100
    return (f"k_lds_engine<{d},{ne},{'hbm' if base >= 1000 else 'lds'},"
            f"{'pipelined' if variant & 1 else 'one-wave'},"
            f"{'time-limited' if variant & 2 else 'untimed'}>")


def test_every_wave_kernel_variant_ran():
    """Runs last in this module: the engine's launch counts show that each
    of the 40 instantiations ran in the tests above (or earlier ones)."""
    counts = N.kernel_launches()
    expected = {v * 2000 + hbm * 1000 + d * 100 + ne
                for v in range(4) for hbm in (0, 1) for d, _, ne, *_ in STENCILS}
    assert len(expected) == 40
    for k in sorted(expected):
        print(f"[variants] {_variant_name(k)}: {counts.get(k, 0)} launches")
    missing = sorted(_variant_name(k) for k in expected if not counts.get(k))
    assert not missing, f"never launched: {missing}"
