"""cpm_run_mcs_noalloc, the entry point that allocates nothing on the device,
called through the C ABI and checked bit for bit against cpm_run_mcs on
twin models: two calls in a row, the second with fewer models, plus its
refusals (a model not yet resident, too many models)."""
import ctypes as C
import struct

import numpy as np
import pytest

import cpm
from cpm import _native as N
from oracle import oracle as O
from tests.helpers import sorting_pair

pytestmark = pytest.mark.gpu

CONFIGS = [
    (31, dict(grid=(48, 48), n_types=3)),
    (32, dict(grid=(64, 64), n_types=2, boundary="fixed")),
    (33, dict(grid=(96, 96), n_types=2)),
    (34, dict(grid=(16, 16, 16), n_types=2, tv=27, ti=30, phi=0.5)),
    (35, dict(grid=(30, 30), n_types=1)),
    # more than 4095 cells: the lane path
    (36, dict(grid=(150, 150), n_types=2, tv=4, ti=8, phi=0.8)),
]


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def _handles(reps):
    return (C.c_void_p * max(len(reps), 1))(*[r.h for r in reps])


def _run_noalloc(reps, n_mcs):
    cnt = N.CpmCounters()
    N.check(N.lib().cpm_run_mcs_noalloc(N.engine(), _handles(reps), len(reps), n_mcs,
                                        C.byref(cnt)))
    return cnt


def _run_reference(reps, n_mcs):
    cnt = N.CpmCounters()
    N.check(N.lib().cpm_run_mcs(N.engine(), _handles(reps), len(reps), n_mcs, N.MODE["exact_replay"],
                                C.byref(cnt)))
    return cnt


def _same(a, b):
    x, y = a.snapshot(), b.snapshot()
    for k in ("lattice", "volume", "interface", "sum_position", "sum_outer", "contact", "rng"):
        assert np.array_equal(x[k], y[k]), k
    assert _bits(x["energy"]) == _bits(y["energy"]) and x["mcs"] == y["mcs"]


def test_noalloc_matches_run_mcs_on_two_calls_the_second_smaller():
    sims = [sorting_pair(cpm, O, seed, **kw)[0] for seed, kw in CONFIGS]
    a = [cpm.Replica(s) for s in sims]
    b = [cpm.Replica(s) for s in sims]
    # residency: one ordinary call each
    _run_reference(a, 1)
    _run_reference(b, 1)
    print("[noalloc] call 1: 6 models x 24 MCS", flush=True)
    ca = _run_noalloc(a, 24)
    cb = _run_reference(b, 24)
    for x, y in zip(a, b):
        _same(x, y)
    assert (ca.attempts, ca.accepted) == (cb.attempts, cb.accepted)
    print("[noalloc] call 2: 3 models x 11 MCS", flush=True)
    _run_noalloc(a[1:4], 11)
    _run_reference(b[1:4], 11)
    for x, y in zip(a, b):
        _same(x, y)
    assert a[0].mcs == 25 and a[1].mcs == 36


def test_noalloc_refuses_what_it_cannot_run():
    sim = sorting_pair(cpm, O, 41, grid=(32, 32), n_types=2)[0]
    fresh = cpm.Replica(sim)
    with pytest.raises(ValueError, match="not resident"):
        _run_noalloc([fresh], 3)
    arr = (C.c_void_p * 1)(fresh.h)
    st = N.lib().cpm_run_mcs_noalloc(N.engine(), arr, 4097, 1, None)
    assert st == N.CPM_ERR_ARG
    assert "exceed the maximum of 4096" in N.lib().cpm_last_error().decode()
    # the refused model is untouched and still runs through the ordinary path
    assert fresh.mcs == 0
    _run_reference([fresh], 2)
    assert fresh.mcs == 2
