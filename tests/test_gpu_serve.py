"""Serving loop: requests admitted by HBM footprint and advanced together in
rounds (continuous batching).  A request's result must be the same alone
and in any mix of other requests — bit for bit, in both modes — and the
same as `CPM.run` gives outside the server."""
import threading
import time

import numpy as np
import pytest

import cpm

pytestmark = pytest.mark.gpu


def sim_of(seed, side=40, n_types=2, boundary="periodic", tv=30):
    J = [[0.0, 16.0, 16.0], [16.0, 3.0, 11.0], [16.0, 11.0, 5.0]][: n_types + 1]
    J = [row[: n_types + 1] for row in J]
    s = cpm.CPM(grid=(side, side), boundary=boundary, seed=seed)
    n = max(2, round(0.6 * side * side / tv))
    for t in range(n_types):
        s.register_cell_type(f"t{t}", target_volume=tv, target_interface=24,
                             lambda_volume=5.0, lambda_interface=1.0)
        s.add_cells(f"t{t}", n // n_types + (1 if t < n % n_types else 0))
    s.set_adhesion(J)
    return s


WINDOWS = dict(burn_in_mcs=40, readout_mcs=60, sampling_interval_mcs=20)


def same_result(a, b):
    assert a.metadata.model_hash == b.metadata.model_hash
    assert a.status.kind == b.status.kind
    assert len(a.samples) == len(b.samples)
    for x, y in zip(a.samples, b.samples):
        assert x.mcs == y.mcs
        assert np.array_equal(x.lattice, y.lattice)
        assert np.array_equal(x.volume, y.volume)
        assert np.array_equal(x.interface, y.interface)
        assert x.conservative_energy == y.conservative_energy


@pytest.mark.parametrize("mode", ["exact_replay", "sublattice"])
def test_request_result_is_the_same_alone_and_in_a_mix(mode):
    target = sim_of(1234, side=40)
    direct = target.run(**WINDOWS, include_lattice=True, mode=mode)
    with cpm.Server() as srv:
        alone = srv.submit(target, **WINDOWS, include_lattice=True, mode=mode).result(120)
    same_result(direct, alone)

    # the mix: requests of other sizes and both modes, a small quantum and
    # round length so every request is split into many rounds (of
    # different lengths once its speed is measured), and a budget that
    # makes later arrivals wait for memory
    others = [(sim_of(100 + i, side=s), m) for i, (s, m) in enumerate(
        [(48, "exact_replay"), (64, "sublattice"), (32, "exact_replay"), (56, "sublattice"),
         (40, "exact_replay"), (72, "exact_replay"), (36, "sublattice"), (48, "exact_replay")])]
    with cpm.Server(quantum_attempts=20000, round_ms=0.5, hbm_budget=6 << 20) as srv:
        reqs = [srv.submit(s, **WINDOWS, mode=m) for s, m in others[:4]]
        time.sleep(0.05)  # the target arrives while the first ones run
        mine = srv.submit(target, **WINDOWS, include_lattice=True, mode=mode)
        reqs += [srv.submit(s, **WINDOWS, mode=m) for s, m in others[4:]]
        in_mix = mine.result(300)
        for r in reqs:
            assert r.result(300).status.kind is not None
        st = srv.stats()
    same_result(direct, in_mix)
    assert st["completed"] == len(others) + 1 and st["failed"] == 0
    assert st["rounds"] > len(others)


def test_lane_path_request_in_time_limited_rounds():
    """A request with more cells than the packed words hold (4352 cells on
    160x160, the lane path: one MCS per time-limited round) served beside
    wave-path requests gives the same result as CPM.run."""
    big = cpm.CPM(grid=(160, 160), boundary="periodic", seed=33)
    big.register_cell_type("a", target_volume=5, target_interface=10, lambda_volume=5.0,
                           lambda_interface=1.0)
    big.register_cell_type("b", target_volume=5, target_interface=10, lambda_volume=5.0,
                           lambda_interface=1.0)
    big.add_cells("a", 2176)
    big.add_cells("b", 2176)
    big.set_adhesion([[0.0, 16.0, 16.0], [16.0, 3.0, 11.0], [16.0, 11.0, 5.0]])
    windows = dict(burn_in_mcs=0, readout_mcs=6, sampling_interval_mcs=3)
    direct = big.run(**windows, include_lattice=True)
    small = [sim_of(300 + i, side=48) for i in range(3)]
    with cpm.Server(round_ms=2.0) as srv:
        print("[case] lane-path request among wave-path requests", flush=True)
        reqs = [srv.submit(s, **WINDOWS) for s in small]
        mine = srv.submit(big, **windows, include_lattice=True)
        served = mine.result(300)
        for r in reqs:
            assert r.result(300).status.kind is not None
        st = srv.stats()
    same_result(direct, served)
    assert st["failed"] == 0


def test_admission_respects_the_hbm_budget_and_reports_latency():
    sims = [sim_of(7 + i, side=48) for i in range(12)]
    budget = 4 << 20
    peak = [0]
    stop = threading.Event()
    with cpm.Server(hbm_budget=budget, quantum_attempts=50000) as srv:
        def watch():
            while not stop.is_set():
                peak[0] = max(peak[0], srv.stats()["hbm_in_use_bytes"])
                time.sleep(0.002)
        w = threading.Thread(target=watch)
        w.start()
        reqs = [srv.submit(s, burn_in_mcs=0, readout_mcs=40, sampling_interval_mcs=20)
                for s in sims]
        for r in reqs:
            r.result(300)
        stop.set()
        w.join()
        st = srv.stats()
    assert 0 < peak[0] <= budget
    assert st["completed"] == 12
    lat = np.array([r.latency_s for r in reqs])
    assert np.all(lat > 0)
    for r in reqs:
        assert r.info.submit_s <= r.info.admit_s <= r.info.done_s
        assert r.info.mcs_done == r.info.mcs_total == 40
    # arrival-order admission: a later request is never admitted earlier
    admits = [r.info.admit_s for r in reqs]
    assert admits == sorted(admits)
    print("latency p50/p99 (s):", float(np.percentile(lat, 50)), float(np.percentile(lat, 99)))


def test_requests_that_cannot_run_are_refused_at_submit():
    with cpm.Server(hbm_budget=1 << 16) as srv:
        with pytest.raises(Exception, match="budget"):
            srv.submit(sim_of(1, side=64), **WINDOWS)
    odd = sim_of(2, side=39)
    with cpm.Server() as srv:
        with pytest.raises(ValueError, match="sublattice"):
            srv.submit(odd, **WINDOWS, mode="sublattice")
        bad = cpm.CPM(grid=(30, 30), boundary="periodic", seed=1)
        with pytest.raises(ValueError):
            srv.submit(bad, **WINDOWS)  # no cell types
