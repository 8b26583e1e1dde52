"""Benchmark: copy attempts per second of the exact-replay engine on one
MI355X, on the batched run_batch workload (4096 theta-replicas, 2D periodic
lattices with side drawn from {48, 64, 96, 128}, 2 cell types, phi = 0.7,
summaries every 500 MCS).

A step is one sampling interval (500 MCS) of every replica in the batch,
advanced on the device by the batched exact-replay kernel with all state
resident in HBM.  Beside it the script measures the plain
one-replica-per-launch kernel on a 64x64 replica, 4096 64x64 replicas
batched (the batched-vs-plain ratio on that shape), the run_batch workload
end to end at its full 10k MCS (measured sims/hour), the dominant kernel's
HBM and LDS roofline fractions from its algorithmic bytes (the LDS peak from
a microbenchmark run first), single replicas, 256 seeds of the 256x256
sorting configuration, serving traces at 80 and 200 requests/s, and the CPU
oracle on the host cores (a bounded sample of the same workload).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--replicas N]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cpm-rust_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DOMINANT_KERNEL = "k_lds_engine<2, 8, *, *, false>"  # exact replay, 2D, Moore stencil, any sigma placement and pipelining, no time budget
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_lds_engine_2d_run_batch4096_pmc.txt")
LDS_TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_lds_engine_2d_run_batch4096_lds_pmc.txt")
SIDES = (48, 64, 96, 128)
INTERVAL = 500
MASTER_SEED = 42


def draw_batch(n: int, seed: int = 2026, sides=SIDES):
    """The seeded prior box of the batched workload: J entries U[1,20]
    (J[medium][medium] = 0), lambda_V U[1,20], lambda_I U[0.5,5],
    V* U[30,80], I* U[20,60] (rounded), side uniform over `sides`."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(rng.choice(sides))
        j = rng.uniform(1, 20, size=6)
        J = [[0.0, j[0], j[1]], [j[0], j[2], j[3]], [j[1], j[3], j[4]]]
        lv, li = rng.uniform(1, 20, 2), rng.uniform(0.5, 5, 2)
        tv = int(round(rng.uniform(30, 80)))
        ti = int(round(rng.uniform(20, 60)))
        ncell = max(2, int(round(0.7 * side * side / tv)))
        out.append(dict(side=side, J=J, lv=lv.tolist(), li=li.tolist(), tv=tv, ti=ti,
                        counts=[ncell - ncell // 2, ncell // 2], seed=None, index=i))
    return out


def make_sim(cpm, p, seed):
    sim = cpm.CPM(grid=(p["side"], p["side"]), boundary="periodic", seed=seed)
    for t, name in enumerate(("a", "b")):
        sim.register_cell_type(name, p["tv"], p["ti"], p["lv"][t], p["li"][t])
        sim.add_cells(name, p["counts"][t])
    sim.set_adhesion(p["J"])
    sim.initialize(warn_on_default_init=False)
    return sim


def algorithmic_bytes(c: dict) -> float:
    """2D bytes per launch: 8 per attempt (sigma_t, sigma_s), 28 per
    topology read (the other 7 neighbours), 56 per energy evaluation
    (types of 8 neighbours, V and I of both cells, types of L and G) and
    356 per accepted copy (sigma write, V/I, position sums, 16 contact
    read-modify-writes, energy)."""
    return 8.0 * c["attempts"] + 28.0 * c["topology"] + 56.0 * c["energy"] + 356.0 * c["accepted"]


def _summary_value(path: str, key: str):
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def measured_traffic():
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE passes, each scaled
    by the same counter's reading on a device copy of known size); None if
    the summary is absent."""
    return _summary_value(TRAFFIC_SUMMARY, "traffic_bytes_per_launch")


def measured_lds_traffic():
    """LDS bytes per launch of the dominant kernel (SQ_INSTS_LDS_LOAD_ and
    _STORE_BANDWIDTH, in 64-byte units fixed on the LDS microbenchmark),
    from the committed rocprofv3 --pmc summary; None if it is absent."""
    return _summary_value(LDS_TRAFFIC_SUMMARY, "lds_traffic_bytes_per_launch")


def lds_peak_tbps():
    """Aggregate LDS read bandwidth of this GPU in TB/s, measured by the
    microbenchmark built beside the engine library (every thread of 8
    workgroups per CU streaming 16-byte LDS reads); None if it is absent."""
    import subprocess

    exe = os.path.join(ROOT, "cpm-rust_amd", "cpm", "lds_bandwidth")
    if not os.path.exists(exe):
        return None
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, check=True).stdout
    for line in out.splitlines():
        if line.startswith("lds_read_tbps:"):
            return float(line.split(":", 1)[1])
    return None


def batch_64(cpm, n: int, interval: int, plain_rate: float, threads: int):
    """The batched-vs-plain comparison on its stated shape: `n` replicas of
    2D periodic 64x64 lattices (same prior box, side fixed at 64), one
    sampling interval warm and one timed through run_batch's engine."""
    params = draw_batch(n, seed=2027, sides=(64,))
    from concurrent.futures import ThreadPoolExecutor

    sims = [make_sim(cpm, p, cpm.derive_seed_u64(MASTER_SEED + 1, p["index"])) for p in params]
    with ThreadPoolExecutor(threads) as ex:
        reps = list(ex.map(lambda s: cpm.Replica(s, 0, 10000, INTERVAL), sims))
    cpm.Replica.run_mcs_many(reps, interval)
    t0 = time.perf_counter()
    c = cpm.Replica.run_mcs_many(reps, interval)
    rate = c["attempts"] / (time.perf_counter() - t0)
    print(f"[bench] {n} x 64x64 batched: {rate / 1e9:.2f} G attempts/s, "
          f"{rate / plain_rate:.0f}x the plain kernel", flush=True)
    return {"replicas": n, "interval_mcs": interval, "attempts_per_s": rate,
            "plain_attempts_per_s": plain_rate, "batched_vs_plain": rate / plain_rate}


def run_batch_full(cpm, n: int, mcs: int, interval: int, threads: int):
    """The run_batch workload end to end: `n` configurations of the prior box
    (seeds derive_seed_u64(42, i)), each initialised on the host, run for
    `mcs` MCS with a readout sample every `interval` MCS, and returned as
    RunResults — configuration, initialisation, device run and readouts all
    inside the timed region."""
    params = draw_batch(n)
    sims = [make_sim(cpm, p, cpm.derive_seed_u64(MASTER_SEED, p["index"])) for p in params]
    cnt = {}
    t0 = time.perf_counter()
    results = cpm.run_batch_configs(sims, 0, mcs, interval, counters=cnt)
    wall = time.perf_counter() - t0
    samples = sum(len(r.samples) for r in results)
    print(f"[bench] run_batch end to end: {n} simulations x {mcs} MCS in {wall:.1f} s "
          f"({n * 3600.0 / wall:.0f} sims/hour)", flush=True)
    return {"simulations": n, "mcs": mcs, "sampling_interval_mcs": interval, "wall_s": wall,
            "sims_per_hour": n * 3600.0 / wall, "attempts_per_s": cnt["attempts"] / wall,
            "samples": samples}


def copy_calibration(mib: int = 256, reps: int = 4):
    """A plain device-to-device copy of known size (reads and writes `mib`
    MiB per launch, HIP's own copy kernel), run beside the engine so a PMC
    pass can scale FETCH_SIZE and WRITE_SIZE to bytes.  Calls the HIP
    runtime directly, so no other framework's kernels enter the profile."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed in the copy calibration")
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(reps):
        if hip.hipMemcpy(b, a, C.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed in the copy calibration")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] copy calibration: {reps} device copies of {mib} MiB", flush=True)


def serving_trace(cpm, n_requests: int = 1500, rate: float = 80.0, seed: int = 7,
                  quantum: int = 1 << 17, max_active: int = 4096, round_ms: float = 30.0,
                  side3d=(32, 40), mcs_range=(40, 160)):
    """The serving workload: a Poisson stream (mean `rate` requests/s) of
    mixed run requests — 70 % 2D periodic lattices of side U{64..256}, 30 %
    3D periodic lattices of side U{side3d} (by default 32..40, the
    LDS-resident range; 32..64 is the stated mix, whose larger lattices keep
    sigma in HBM) — with 2-4 cell types, V* = 40, I* = 30 (2D) / 120 (3D),
    lambda_V = 5, lambda_I = 1, J: medium row 16, type pairs U[2,14],
    density 0.8 (2D) / 0.5 (3D), and U[mcs_range] MCS each, rounded down to
    a multiple of 20 (readout, one sample per 20 MCS).  Requests are
    initialised on 16 host threads as they arrive and queued on one server;
    reported: completion latency (server queue + device + readouts)
    p50/p99, sims/hour and copy attempts/s over the trace."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)
    specs, t = [], 0.0
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate)
        if rng.random() < 0.7:
            side = int(rng.integers(64, 257))
            grid, it, phi = (side, side), 30, 0.8
        else:
            side = int(rng.integers(side3d[0], side3d[1] + 1))
            grid, it, phi = (side, side, side), 120, 0.5
        T = int(rng.integers(2, 5))
        J = np.zeros((T + 1, T + 1))
        J[0, 1:] = J[1:, 0] = 16.0
        for a in range(1, T + 1):
            for b in range(a, T + 1):
                J[a, b] = J[b, a] = float(rng.uniform(2, 14))
        mcs = max(20, int(rng.integers(mcs_range[0], mcs_range[1] + 1)) // 20 * 20)
        specs.append(dict(at=t, grid=grid, T=T, J=J.tolist(), it=it, phi=phi, mcs=mcs,
                          seed=int(cpm.derive_seed_u64(seed, i))))

    def build(sp):
        sim = cpm.CPM(grid=sp["grid"], boundary="periodic", seed=sp["seed"])
        n = cpm.cells_for_density(sp["grid"], 40, sp["phi"])
        for k in range(sp["T"]):
            sim.register_cell_type(f"t{k}", 40, sp["it"], 5.0, 1.0)
            sim.add_cells(f"t{k}", n // sp["T"] + (1 if k < n % sp["T"] else 0))
        sim.set_adhesion(sp["J"])
        return sim

    import warnings as _w
    with _w.catch_warnings():
        _w.simplefilter("ignore")
        sims = [build(sp) for sp in specs]
    with cpm.Server(quantum_attempts=quantum, max_active=max_active, round_ms=round_ms) as srv, \
            ThreadPoolExecutor(16) as pool:
        t0 = time.perf_counter()
        futs = []
        for sp, sim in zip(specs, sims):
            delay = sp["at"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            futs.append(pool.submit(lambda s=sim, m=sp["mcs"]: srv.submit(
                s, burn_in_mcs=0, readout_mcs=m, sampling_interval_mcs=20)))
        reqs = [f.result() for f in futs]
        for r in reqs:
            r.result()
        wall = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array([r.latency_s for r in reqs])
    return {"requests": n_requests, "offered_rate_per_s": rate, "wall_s": wall,
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "sims_per_hour": n_requests * 3600.0 / wall,
            "attempts_per_s": st["attempts"] / wall, "rounds": st["rounds"],
            "device_s": st["device_s"], "host_s": st["busy_s"] - st["device_s"],
            "mode": "exact_replay",
            "mix": f"70% 2D side U{{64..256}}, 30% 3D side U{{{side3d[0]}..{side3d[1]}}}, "
                   f"2-4 types, U[{mcs_range[0]},{mcs_range[1]}] MCS per request"}


def sorting_256_parts(seed: int = 1):
    """The single-replica sorting workload: 256x256 periodic, 4 cell types
    scattered at random, phi = 0.8 via cells_for_density with V* = 40,
    I* = 30, lambda_V = 5, lambda_I = 1, J medium row 16 and type pairs
    U[2,14] drawn from the seed.  Returns (cell counts per type, J)."""
    import warnings as _w

    from cpm.parameters import cells_for_density

    rng = np.random.default_rng(seed)
    T = 4
    J = np.zeros((T + 1, T + 1))
    J[0, 1:] = J[1:, 0] = 16.0
    for a in range(1, T + 1):
        for b in range(a, T + 1):
            J[a, b] = J[b, a] = float(rng.uniform(2, 14))
    with _w.catch_warnings():
        _w.simplefilter("ignore")  # the low-resolution warning of V* = 40
        n = cells_for_density((256, 256), 40, 0.8)
    return [n // T + (1 if t < n % T else 0) for t in range(T)], J.tolist()


def sorting_256(cpm, seed: int = 1):
    counts, J = sorting_256_parts(seed)
    sim = cpm.CPM(grid=(256, 256), boundary="periodic", seed=seed)
    for t, c in enumerate(counts):
        sim.register_cell_type(f"t{t}", 40, 30, 5.0, 1.0)
        sim.add_cells(f"t{t}", c)
    sim.set_adhesion(J)
    sim.initialize(warn_on_default_init=False)
    return sim


def organoid_64_parts(seed: int = 1):
    """The 3D organoid workload: 64^3 fixed boundary, 3 cell types in nested
    spherical shells (radii 10 / 18 / 24 sites) cut into cells of about
    V* = 1048 sites (Voronoi within each shell, each cell reduced to its
    largest face-connected piece), I* = 2760, lambda_V = 10, lambda_I = 2,
    J medium row 16 and type pairs U[2,16] drawn from the seed.  Returns
    (lattice, cell type of each cell, J)."""
    from scipy import ndimage

    rng = np.random.default_rng(seed)
    L = 64
    c = (L - 1) / 2.0
    z, y, x = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    lat = np.zeros((L, L, L), dtype=np.int64)
    type_of, nid = [], 0
    for t, (r0, r1) in enumerate(((0, 10), (10, 18), (18, 24))):
        pts = np.argwhere((r >= r0) & (r < r1))
        k = max(1, int(round(len(pts) / 1048)))
        seeds = pts[rng.choice(len(pts), k, replace=False)]
        own = ((pts[:, None, :] - seeds[None, :, :]) ** 2).sum(-1).argmin(1)
        lat[tuple(pts.T)] = own + 1 + nid
        type_of += [t] * k
        nid += k
    for cid in range(1, nid + 1):
        m = lat == cid
        lab, nl = ndimage.label(m)
        if nl > 1:
            sizes = ndimage.sum(m, lab, range(1, nl + 1))
            lat[m & (lab != 1 + int(np.argmax(sizes)))] = 0
    T = 3
    J = np.zeros((T + 1, T + 1))
    J[0, 1:] = J[1:, 0] = 16.0
    for a in range(1, T + 1):
        for b in range(a, T + 1):
            J[a, b] = J[b, a] = float(rng.uniform(2, 16))
    return lat, type_of, J.tolist()


def organoid_64(cpm, seed: int = 1):
    lat, type_of, J = organoid_64_parts(seed)
    sim = cpm.CPM(grid=lat.shape, boundary="fixed", seed=seed)
    for t in range(3):
        sim.register_cell_type(f"t{t}", 1048, 2760, 10.0, 2.0)
    sim.set_adhesion(J)
    sim.initialize(lattice=lat.ravel().tolist(), cell_type_of=type_of)
    return sim


def many_seeds_256(cpm, n: int = 256):
    """The 256x256 sorting configuration over n seeds (its KS workload):
    copy attempts/s of the whole batch in exact replay and in sublattice
    mode (10 MCS in one call after 2 warm ones)."""
    import warnings as _w

    with _w.catch_warnings():
        _w.simplefilter("ignore")
        reps = [cpm.Replica(sorting_256(cpm, seed=s)) for s in range(n)]
    out = {"replicas": n}
    for mode, mcs in (("exact_replay", 10), ("sublattice", 10)):
        cpm.Replica.run_mcs_many(reps, 2, mode=mode)
        t0 = time.perf_counter()
        c = cpm.Replica.run_mcs_many(reps, mcs, mode=mode)
        out[mode + "_attempts_per_s"] = c["attempts"] / (time.perf_counter() - t0)
    print(f"[bench] {n} seeds of 256x256 sorting: exact {out['exact_replay_attempts_per_s'] / 1e9:.2f}"
          f" G/s, sublattice {out['sublattice_attempts_per_s'] / 1e9:.3f} G/s", flush=True)
    return out


def single_replica_rates(cpm):
    """Copy attempts/s of one replica of each single-lattice workload, in
    exact replay and in sublattice mode (short runs after one warm MCS)."""
    out = {}
    for name, make, mcs in (("sorting_256x256_4types", sorting_256, 20),
                            ("organoid_64cubed_3types", organoid_64, 4)):
        row = {}
        for mode in ("exact_replay", "sublattice"):
            rep = cpm.Replica(make(cpm))
            rep.run_mcs(1, mode=mode)
            t0 = time.perf_counter()
            c = rep.run_mcs(mcs, mode=mode)
            row[mode + "_attempts_per_s"] = c["attempts"] / (time.perf_counter() - t0)
        row["mcs_timed"] = mcs
        out[name] = row
    return out


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# the host threads a one-GPU share of the machine may keep busy
HOST_THREADS = 16


def cpu_baseline(params, seconds: float = 12.0):
    """The oracle (C restatement of the reference loop) on the host cores:
    replicas of the same workload advanced by one sampling interval each,
    one replica per thread (ctypes releases the GIL), as the reference's
    run_batch spreads replicas over a thread pool.  Uses HOST_THREADS
    threads (a one-GPU share of the host), fewer if the host has fewer."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O

    cores = min(os.cpu_count() or 1, HOST_THREADS)
    models = []
    for p in params[: cores * 2]:
        cfg = O.Config(grid=(p["side"], p["side"]), boundary="periodic", seed=p["seed"],
                       cell_types=[O.CellTypeSpec("a", p["tv"], p["ti"], p["lv"][0], p["li"][0]),
                                   O.CellTypeSpec("b", p["tv"], p["ti"], p["lv"][1], p["li"][1])],
                       cell_counts=p["counts"], adhesion=p["J"], warn_on_default_init=False)
        m = O.Model(cfg)
        m.cfg_side = p["side"]
        models.append(m)
    # calibrate MCS per replica so the sample takes about `seconds`
    probe = models[0]
    t0 = time.perf_counter()
    probe.run_mcs(2)
    per_site = (time.perf_counter() - t0) / (2 * probe.n_sites)
    total_sites = sum(m.n_sites for m in models)
    mcs = max(1, int(seconds * cores / (per_site * total_sites)))
    before = sum(int(m.counters()[0]) for m in models)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        list(ex.map(lambda m: m.run_mcs(mcs), models))
    dt = time.perf_counter() - t0
    attempts = sum(int(m.counters()[0]) for m in models) - before
    return {"value": attempts / dt, "unit": "copy attempts/s", "cores": cores, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "sample": f"{len(models)} replicas of the batched workload x {mcs} MCS "
                      f"({attempts} attempts in {dt:.1f} s on {cores} threads)",
            "single_core": cpu_single_core(O)}


def _time_oracle(m, seconds: float) -> dict:
    """run_mcs on one core for about `seconds` (after one warm MCS)."""
    m.run_mcs(1)
    t0 = time.perf_counter()
    m.run_mcs(1)
    one = max(time.perf_counter() - t0, 1e-6)
    mcs = max(1, int(seconds / one))
    a0 = int(m.counters()[0])
    t0 = time.perf_counter()
    m.run_mcs(mcs)
    dt = time.perf_counter() - t0
    att = int(m.counters()[0]) - a0
    return {"mcs_per_s": mcs / dt, "attempts_per_s": att / dt, "mcs_timed": mcs}


def cpu_single_core(O, seconds: float = 3.0) -> dict:
    """One-core oracle rates of the single-replica workloads (the reference's
    run is single-threaded): the README quickstart (30x30, 5 cells), the
    256x256 four-type sorting lattice and the 64^3 organoid, plus the
    reference's published 30x30 / 10-cell fixture
    (notebooks/performance_demo.ipynb:106: 68,579.0 MCS/s, hardware not
    stated) as a check on the oracle's speed."""
    out = {}
    qs = O.Config(grid=(30, 30), boundary="periodic", seed=1,
                  cell_types=[O.CellTypeSpec("epithelial", 50, 75, 10.0, 2.0)], cell_counts=[5],
                  adhesion=[[0.0, 5.0], [5.0, 2.0]], warn_on_default_init=False)
    out["quickstart_30x30"] = _time_oracle(O.Model(qs), seconds)
    counts, J = sorting_256_parts(1)
    srt = O.Config(grid=(256, 256), boundary="periodic", seed=1,
                   cell_types=[O.CellTypeSpec(f"t{t}", 40, 30, 5.0, 1.0) for t in range(4)],
                   cell_counts=counts, adhesion=J, warn_on_default_init=False)
    out["sorting_256x256_4types"] = _time_oracle(O.Model(srt), seconds)
    lat, type_of, J = organoid_64_parts(1)
    org = O.Config(grid=lat.shape, boundary="fixed", seed=1,
                   cell_types=[O.CellTypeSpec(f"t{t}", 1048, 2760, 10.0, 2.0) for t in range(3)],
                   cell_counts=[0, 0, 0], adhesion=J, init_lattice=lat.ravel().tolist(),
                   init_cell_type_of=type_of)
    out["organoid_64cubed_3types"] = _time_oracle(O.Model(org), seconds)
    fx = O.Config(grid=(30, 30), boundary="periodic", seed=1,
                  cell_types=[O.CellTypeSpec("a", 30, 40, 1.0, 0.2)], cell_counts=[10],
                  adhesion=[[0.0, 3.0], [3.0, 1.0]], warn_on_default_init=False)
    m = O.Model(fx)
    t0 = time.perf_counter()
    m.run_mcs(20000)
    out["published_fixture_30x30_10cells"] = {
        "oracle_mcs_per_s": 20000 / (time.perf_counter() - t0),
        "reference_published_mcs_per_s": 68579.0,
        "source": "notebooks/performance_demo.ipynb:106 (hardware not stated)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--replicas", type=int, default=4096)
    ap.add_argument("--interval", type=int, default=INTERVAL)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-mcs", type=int, default=20)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-batch64", action="store_true",
                    help="skip the 64x64 batched-vs-plain comparison")
    ap.add_argument("--no-many-seeds", action="store_true",
                    help="skip the 256-seed 256x256 sorting batch")
    ap.add_argument("--no-run-batch", action="store_true",
                    help="skip the end-to-end run_batch (10k MCS per simulation)")
    ap.add_argument("--run-batch-mcs", type=int, default=10000)
    ap.add_argument("--no-lds-peak", action="store_true",
                    help="skip the LDS microbenchmark (a child process; leave it out under a "
                         "profiler that has already opened the GPU)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device copy (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine shards replicas across GPUs only by "
              "index range with no exchange, which this build does not run; use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    lds_peak = None if args.no_lds_peak else lds_peak_tbps()
    if lds_peak is not None:
        print(f"[bench] LDS read peak (microbenchmark): {lds_peak:.1f} TB/s", flush=True)

    import cpm

    t_init0 = time.perf_counter()
    params = draw_batch(args.replicas)
    for p in params:
        p["seed"] = cpm.derive_seed_u64(MASTER_SEED, p["index"])
    from concurrent.futures import ThreadPoolExecutor

    sims = [make_sim(cpm, p, p["seed"]) for p in params]
    threads = min(os.cpu_count() or 1, HOST_THREADS)
    # CPM::new of every replica on the host threads, then scatter_and_grow
    # of all of them on the device (one wave per replica)
    with ThreadPoolExecutor(threads) as ex:
        reps = list(ex.map(lambda s: cpm.Replica(s, 0, 10000, INTERVAL, initialize=False), sims))
    t_create = time.perf_counter() - t_init0
    t_d0 = time.perf_counter()
    cpm.Replica.initialize_many(reps)
    t_dev_init = time.perf_counter() - t_d0
    t_init = time.perf_counter() - t_init0
    # initialisation and the first 1-MCS call (which uploads the state of
    # host-initialised replicas; device-initialised ones are resident) on
    # the device and on the host threads, on copies of the same replicas
    from cpm import _native

    def init_and_first_call(on_device):
        with ThreadPoolExecutor(threads) as ex:
            rs = list(ex.map(lambda s: cpm.Replica(s, 0, 10000, INTERVAL, initialize=False),
                             sims))
        _native.set_device_init(on_device)
        try:
            ta = time.perf_counter()
            cpm.Replica.initialize_many(rs)
            tb = time.perf_counter()
            cpm.Replica.run_mcs_many(rs, 1)
            tc = time.perf_counter()
        finally:
            _native.set_device_init(True)
        return tb - ta, tc - tb

    dev_i, dev_c = init_and_first_call(True)
    host_i, host_c = init_and_first_call(False)
    init_info = {"replicas": len(reps), "create_s": t_create,
                 "device_scatter_and_grow_s": t_dev_init,
                 "device_init_s": dev_i, "device_first_call_s": dev_c,
                 "host_threads_init_s": host_i, "host_threads_first_call_s": host_c,
                 "host_threads": min(os.cpu_count() or 1, 16)}
    print(f"[bench] {len(reps)} replicas initialised in {t_init:.2f} s (create {t_create:.2f} s, "
          f"scatter_and_grow on the device {t_dev_init:.3f} s); init + first 1-MCS call: "
          f"device {dev_i:.3f} + {dev_c:.3f} s, {init_info['host_threads']} host threads "
          f"{host_i:.3f} + {host_c:.3f} s", flush=True)

    for w in range(args.warmup):
        c = cpm.Replica.run_mcs_many(reps, args.interval)
        print(f"[bench] warmup {w}: {c['kernel_ms']:.0f} ms", flush=True)

    totals = {k: 0.0 for k in ("attempts", "non_null", "topology", "energy", "accepted",
                               "exp_near_ties", "kernel_ms", "replay_launches",
                               "replay_launch_ms", "piped_stragglers")}
    t0 = time.perf_counter()
    for s in range(args.steps):
        c = cpm.Replica.run_mcs_many(reps, args.interval)
        for k in totals:
            totals[k] += c[k]
        print(f"[bench] step {s}: {c['kernel_ms']:.0f} ms, {c['attempts'] / c['kernel_ms'] / 1e6:.2f}"
              f" G attempts/s", flush=True)
    wall = time.perf_counter() - t0
    from cpm import _native as _nat
    placement = "%s-first, %.0f %% of the LDS" % (lambda o, sh: (o, 100 * sh))(
        *_nat.sigma_placement())
    value = totals["attempts"] / wall
    # roofline of the dominant kernel: algorithmic bytes per launch over its
    # average launch duration (HIP events around each launch, on its stream)
    launches = max(1.0, totals["replay_launches"])
    avg_launch_s = totals["replay_launch_ms"] / 1e3 / launches
    bytes_per_launch = algorithmic_bytes(totals) / launches
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0

    # plain one-replica-per-launch kernel on a 64x64 replica
    p64 = next(p for p in params if p["side"] == 64)
    plain = cpm.Replica(make_sim(cpm, p64, p64["seed"]), 0, 10000, INTERVAL)
    plain.run_mcs_plain(1)
    t1 = time.perf_counter()
    pc = plain.run_mcs_plain(args.plain_mcs)
    plain_rate = pc["attempts"] / (time.perf_counter() - t1)
    print(f"[bench] plain: {plain_rate / 1e6:.2f} M attempts/s", flush=True)

    # run_batch sims/hour projected from the measured step (20 intervals per
    # simulation) plus the measured host initialisation
    step_s = wall / args.steps
    n_reps = len(reps)
    sims_per_hour = n_reps * 3600.0 / (t_init + 20 * step_s)

    b64 = None if args.no_batch64 else batch_64(cpm, args.replicas, args.interval, plain_rate,
                                                 min(os.cpu_count() or 1, HOST_THREADS))

    full = None
    if not args.no_run_batch:
        del reps  # the step replicas' device memory
        full = run_batch_full(cpm, args.replicas, args.run_batch_mcs, args.interval,
                              min(os.cpu_count() or 1, HOST_THREADS))

    single = single_replica_rates(cpm)
    seeds = None if args.no_many_seeds else many_seeds_256(cpm)
    print("[bench] single replicas: " + ", ".join(
        f"{k} exact {v['exact_replay_attempts_per_s'] / 1e6:.1f} M/s, sublattice "
        f"{v['sublattice_attempts_per_s'] / 1e6:.1f} M/s" for k, v in single.items()), flush=True)
    serving = None
    if not args.no_serving:
        serving = [serving_trace(cpm, rate=r) for r in (80.0, 200.0)]
        # the stated mix's shapes (3D sides up to 64, sigma in HBM for the
        # larger ones) with its 5k-20k MCS scaled by 1/100
        serving.append(serving_trace(cpm, n_requests=400, rate=20.0, side3d=(32, 64),
                                     mcs_range=(50, 200)))
        for sv in serving:
            print(f"[bench] serving at {sv['offered_rate_per_s']:.0f} requests/s: "
                  f"{sv['requests']} requests in {sv['wall_s']:.1f} s, "
                  f"p50 {sv['latency_p50_s']:.3f} s, p99 {sv['latency_p99_s']:.3f} s", flush=True)
    if args.copy_calibration:
        copy_calibration()
    cpu = None if args.no_cpu else cpu_baseline(params)
    line = {
        "metric": "copy attempts/s on 1 GPU (batched exact replay: run_mcs steps over the "
                  "4096-replica run_batch workload)",
        "value": value,
        "unit": "copy attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+f64",
        "data": "synthetic",
        "config": {"workload": f"run_batch: {n_reps} theta-replicas, 2D periodic sides "
                               f"{{48,64,96,128}}, 2 types, phi=0.7, {args.interval} MCS per step",
                   "replicas": n_reps, "interval_mcs": args.interval,
                   "parallelism": "1 GPU, replicas batched in one launch",
                   "sigma_placement": placement},
        "plain_attempts_per_s": plain_rate,
        "batched_vs_plain": value / plain_rate,
        "sims_per_hour": None if full is None else full["sims_per_hour"],
        "sims_per_hour_projected_from_steps": sims_per_hour,
        "run_batch": full,
        "initialization": init_info,
        "counters": {k: totals[k] for k in totals},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": DOMINANT_KERNEL, "launches": int(launches),
                     "avg_launch_ms": avg_launch_s * 1e3,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        # the same algorithmic bytes against the LDS read peak measured on
        # this GPU: every lattice and cell table is LDS-resident in the
        # batched kernel, so its reads and writes are LDS traffic
        "lds_roofline": None if lds_peak is None else {
            "bound": "lds", "achieved": achieved, "peak": lds_peak * 1e3, "unit": "GB/s",
            "frac": achieved / (lds_peak * 1e3), "traffic": measured_lds_traffic(),
            "kernel": DOMINANT_KERNEL,
            "peak_source": "LDS read microbenchmark on this GPU"},
        "batch_64x64": b64,
        "cpu_baseline": cpu,
        # below capacity (latency = service time) and at the configured
        # 200 requests/s (above capacity for this mix: latency includes queueing)
        "serving": serving,
        "single_replica": single,
        "many_seeds_256x256": seeds,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
