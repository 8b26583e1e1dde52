"""Serving loop: simulation requests (one `run` each, output::run_with of
crates/cpm-core/src/output.rs:434-529) submitted at any time, admitted to
the GPU while their device footprint fits an HBM budget, and advanced
together in rounds (continuous batching).  The reference has no serving
layer; a request returns the same `RunResult` `CPM.run` does, and the same
result alone as in any mix of other requests.
"""

from __future__ import annotations

import ctypes as C
import threading

from . import _native as N

STATES = {N.CPM_REQ_QUEUED: "queued", N.CPM_REQ_ACTIVE: "active", N.CPM_REQ_DONE: "done",
          N.CPM_REQ_FAILED: "failed"}


class Request:
    """Handle of a submitted simulation."""

    def __init__(self, server: "Server", rid: int, spec: dict, used_default: bool):
        self.server = server
        self.id = rid
        self.spec = spec
        self.used_default = used_default
        self.info = None
        self._result = None

    def poll(self, timeout: float = 0.0):
        """The `RunResult` once done (None before); waits up to `timeout`
        seconds.  Raises the engine error of a failed request."""
        if self._result is None:
            self._result = self.server._poll(self, timeout)
        return self._result

    def result(self, timeout: float | None = None):
        """Blocks until done (or `timeout` seconds) and returns the result."""
        while self.poll(1.0 if timeout is None else timeout) is None:
            if timeout is not None:
                raise TimeoutError(f"request {self.id} not finished after {timeout} s")
        return self._result

    @property
    def latency_s(self) -> float | None:
        """Completion time minus submission time (None before completion)."""
        if self.info is None or self.info.state != N.CPM_REQ_DONE:
            return None
        return self.info.done_s - self.info.submit_s


class Server:
    """A serving loop on the process-wide engine.

    hbm_budget: device bytes requests may hold at once (None: 80 % of the
    free device memory); max_active: requests on the device at once;
    round_ms: round length — every exact-replay chain runs that long on
    the device per round (stopping mid-MCS if need be; the next round
    resumes it), never past its next sample; quantum_attempts: copy
    attempts per sublattice-mode request per round.
    """

    def __init__(self, hbm_budget: int | None = None, max_active: int = 4096,
                 quantum_attempts: int = 1 << 18, round_ms: float = 30.0):
        opt = N.CpmServerOptions(int(hbm_budget or 0), int(max_active), int(quantum_attempts),
                                 float(round_ms))
        self.h = C.c_void_p()
        N.check(N.lib().cpm_serve_create(N.engine(), C.byref(opt), C.byref(self.h)))
        self._lock = threading.Lock()
        N._servers.add(self)

    def close(self) -> None:
        if getattr(self, "h", None):
            N.lib().cpm_serve_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, sim, burn_in_mcs: int, readout_mcs: int, sampling_interval_mcs: int,
               include_lattice: bool = False, mode: str = "exact_replay") -> Request:
        """Validates and initialises `sim` (a configured `cpm.CPM`) and
        queues one run of it; configuration errors raise `ValueError` here."""
        spec = sim._windowed(burn_in_mcs, readout_mcs, sampling_interval_mcs)
        holder = N.ConfigHolder(spec)
        rid, used = C.c_uint64(), C.c_int32(0)
        N.check(N.lib().cpm_serve_submit(self.h, holder.ref(), 1 if include_lattice else 0,
                                         N.MODE[mode], C.byref(rid), C.byref(used)))
        return Request(self, int(rid.value), spec, bool(used.value))

    def _poll(self, req: Request, timeout: float):
        info, res = N.CpmRequestInfo(), C.c_void_p()
        N.check(N.lib().cpm_serve_poll(self.h, req.id, float(timeout), C.byref(info),
                                       C.byref(res)))
        req.info = info
        if not res.value:
            return None
        return N.RunResult(res, req.spec["grid"], [t["name"] for t in req.spec["cell_types"]])

    def stats(self) -> dict:
        out = N.CpmServerStats()
        N.check(N.lib().cpm_serve_stats(self.h, C.byref(out)))
        return out.as_dict()
