"""Engine access at the granularity of cpm-core's Rust API.

`Replica` is the counterpart of `cpm_core::model::CPM<D>` driven by
`initialization::initialize` and `monte_carlo::run_mcs`
(crates/cpm-core/src/model.rs:44-76, initialization.rs:39-55,
monte_carlo.rs:327-357): one model whose state lives on the GPU between
calls.  `Replica.run_mcs_many` advances many replicas in one launch.
"""

from __future__ import annotations

import ctypes as C
import warnings

import numpy as np

from . import _native as N
from .warnings import CPMInitializationWarning


def derive_seed_u64(master_seed: int, index: int) -> int:
    """crates/cpm-core/src/rng.rs:37-40"""
    return int(N.lib().cpm_derive_seed_u64(int(master_seed), int(index)))


class Replica:
    """A validated, initialised model (CPM::new + initialize)."""

    def __init__(self, sim, burn_in_mcs: int = 0, readout_mcs: int = 1,
                 sampling_interval_mcs: int = 1, initialize: bool = True):
        spec = sim._windowed(burn_in_mcs, readout_mcs, sampling_interval_mcs)
        self.spec = spec
        self.dim = len(spec["grid"])
        L = N.lib()
        self._holder = N.ConfigHolder(spec)
        self.h = C.c_void_p()
        N.check(L.cpm_model_create(self._holder.ref(), C.byref(self.h)))
        self.used_default = False
        if not initialize:  # the state comes from set_state
            return
        used = C.c_int32(0)
        try:
            N.check(L.cpm_model_initialize(self.h, C.byref(used)))
        except Exception:
            L.cpm_model_destroy(self.h)
            self.h = None
            raise
        self.used_default = bool(used.value)

    def __del__(self):
        if getattr(self, "h", None):
            N.lib().cpm_model_destroy(self.h)
            self.h = None

    @staticmethod
    def initialize_many(replicas) -> None:
        """initialization::initialize of replicas created with
        initialize=False, in one call (as run_batch's workers do):
        scatter_and_grow runs on the device, one wave per replica."""
        arr = (C.c_void_p * max(len(replicas), 1))(*[r.h for r in replicas])
        used = (C.c_int32 * max(len(replicas), 1))()
        N.check(N.lib().cpm_model_initialize_batch(N.engine(), arr, len(replicas), used))
        for r, u in zip(replicas, used):
            r.used_default = bool(u)

    # ------------------------------------------------------------ dynamics
    def run_mcs(self, n: int, mode: str = "exact_replay") -> dict:
        return Replica.run_mcs_many([self], n, mode)

    @staticmethod
    def run_mcs_many(replicas, n: int, mode: str = "exact_replay") -> dict:
        arr = (C.c_void_p * max(len(replicas), 1))(*[r.h for r in replicas])
        cnt = N.CpmCounters()
        N.check(N.lib().cpm_run_mcs(N.engine(), arr, len(replicas), int(n), N.MODE[mode],
                                    C.byref(cnt)))
        return cnt.as_dict()

    def run_mcs_plain(self, n: int) -> dict:
        cnt = N.CpmCounters()
        N.check(N.lib().cpm_run_mcs_plain(N.engine(), self.h, int(n), C.byref(cnt)))
        return cnt.as_dict()

    # ------------------------------------------------------------ state
    def info(self) -> N.CpmModelInfo:
        out = N.CpmModelInfo()
        N.check(N.lib().cpm_model_get_info(self.h, C.byref(out)))
        return out

    @property
    def mcs(self) -> int:
        return int(self.info().mcs)

    @property
    def energy(self) -> float:
        return float(self.info().conservative_energy)

    def chain_rate(self) -> float:
        """Copy attempts/s of this replica's own chain in its last
        exact-replay call on the wave path (device clock; 0 before one)."""
        out = C.c_double()
        N.check(N.lib().cpm_model_chain_rate(self.h, C.byref(out)))
        return float(out.value)

    def rng_state(self) -> np.ndarray:
        return np.array(list(self.info().rng_state), dtype=np.uint64)

    def snapshot(self) -> dict:
        info = self.info()
        n, D = int(info.n_cells), self.dim
        lat = np.zeros(int(info.n_sites), dtype=np.uint32)
        vol = np.zeros(n, dtype=np.uint32)
        itf = np.zeros(n, dtype=np.uint32)
        typ = np.zeros(n, dtype=np.uint32)
        sp = np.zeros(n * D, dtype=np.int64)
        so = np.zeros(n * D * D, dtype=np.int64)
        ct = np.zeros((n + 1) * (n + 1), dtype=np.uint32)
        N.check(N.lib().cpm_model_get_state(
            self.h, lat.ctypes.data_as(N.u32p), vol.ctypes.data_as(N.u32p),
            itf.ctypes.data_as(N.u32p), typ.ctypes.data_as(N.u32p), sp.ctypes.data_as(N.i64p),
            so.ctypes.data_as(N.i64p), ct.ctypes.data_as(N.u32p)))
        return {
            "lattice": lat.reshape(self.spec["grid"]), "volume": vol, "interface": itf,
            "type_index": typ, "sum_position": sp.reshape(n, D),
            "sum_outer": so.reshape(n, D, D), "contact": ct.reshape(n + 1, n + 1),
            "energy": float(info.conservative_energy),
            "rng": np.array(list(info.rng_state), dtype=np.uint64), "mcs": int(info.mcs),
        }

    def hashes(self) -> tuple[str, str]:
        """(convention_hash, model_hash) of the current state."""
        a, b = C.create_string_buffer(65), C.create_string_buffer(65)
        N.check(N.lib().cpm_model_hashes(self.h, a, b))
        return a.value.decode(), b.value.decode()

    def global_energy(self) -> float:
        out = C.c_double()
        N.check(N.lib().cpm_model_global_energy(self.h, C.byref(out)))
        return float(out.value)

    def set_state(self, lattice, type_index, volume=None, interface=None, sum_position=None,
                  sum_outer=None, contact=None, energy=None, rng=None, mcs: int = 0,
                  birth_mcs=None, edge_sites=None) -> None:
        """Replaces the whole state by the given one (the fields of CPM<D> and
        State<D>, crates/cpm-core/src/model.rs:15-38, state.rs:168-203, and
        the Act term's per-site birth MCS, energy.rs:337-344, and the
        EdgeList proposal's edge list in its order, edge_list.rs:42-45, as
        interior row-major indices); omitted per-cell fields are derived
        from the lattice as initialization does, an omitted energy by
        global_energy, an omitted rng keeps the current stream, omitted
        births are 0, an omitted edge list is rebuilt from the lattice."""
        keep = []

        def arr(x, dtype, ptype):
            if x is None:
                return None
            a = np.ascontiguousarray(np.asarray(x, dtype=dtype).ravel())
            keep.append(a)
            return a.ctypes.data_as(ptype)

        typ = np.asarray(type_index, dtype=np.uint32).ravel()
        st = N.CpmModelState()
        st.lattice = arr(lattice, np.uint32, N.u32p)
        st.n_cells = len(typ)
        st.type_index = arr(typ, np.uint32, N.u32p)
        st.volume = arr(volume, np.uint32, N.u32p)
        st.interface = arr(interface, np.uint32, N.u32p)
        st.sum_position = arr(sum_position, np.int64, N.i64p)
        st.sum_outer = arr(sum_outer, np.int64, N.i64p)
        st.contact = arr(contact, np.uint32, N.u32p)
        st.has_energy = 0 if energy is None else 1
        st.conservative_energy = 0.0 if energy is None else float(energy)
        st.rng_state = arr(rng, np.uint64, N.u64p)
        st.mcs = int(mcs)
        st.birth_mcs = arr(birth_mcs, np.uint32, N.u32p)
        st.edge_sites = arr(edge_sites, np.uint64, N.u64p)
        st.n_edge_sites = 0 if edge_sites is None else int(np.asarray(edge_sites).size)
        if birth_mcs is not None and np.asarray(birth_mcs).size != np.asarray(lattice).size:
            raise ValueError("birth_mcs must hold one MCS per site")
        if np.asarray(lattice).size != int(np.prod(self.spec["grid"])):
            raise ValueError("lattice must hold one id per site")
        N.check(N.lib().cpm_model_set_state(self.h, C.byref(st)))

    def birth_mcs(self) -> np.ndarray:
        """The Act term's per-site birth MCS (energy.rs:337-344), shaped like
        the grid; zeros when the term is off."""
        grid = tuple(self.spec["grid"])
        out = np.zeros(int(np.prod(grid)), dtype=np.uint32)
        N.check(N.lib().cpm_model_birth_mcs(self.h, out.ctypes.data_as(N.u32p), out.size))
        return out.reshape(grid)

    def edge_sites(self) -> np.ndarray:
        """The EdgeList proposal's edge list in its order (interior row-major
        site indices); empty under the uniform proposal."""
        n = C.c_size_t(0)
        cap = int(np.prod(self.spec["grid"]))
        out = np.zeros(cap, dtype=np.uint64)
        N.check(N.lib().cpm_model_edge_sites(self.h, out.ctypes.data_as(N.u64p), cap,
                                             C.byref(n)))
        return out[:n.value]

    def sample(self, include_lattice: bool = False):
        """take_sample of the current state (crates/cpm-core/src/
        output.rs:289-337) as a RunResult holding that one sample, its
        status from run_with's rules applied to it."""
        h = C.c_void_p()
        N.check(N.lib().cpm_model_sample(self.h, 1 if include_lattice else 0, C.byref(h)))
        return N.RunResult(h, self.spec["grid"], [t["name"] for t in self.spec["cell_types"]])


def not_equilibrated(trace) -> bool:
    """crates/cpm-core/src/output.rs:388-420 through the engine library."""
    a = np.ascontiguousarray(np.asarray(trace, dtype=np.float64))
    out = C.c_int32(0)
    N.check(N.lib().cpm_not_equilibrated(a.ctypes.data_as(N.f64p), len(a), C.byref(out)))
    return bool(out.value)


def run_batch_configs(sims, burn_in_mcs: int, readout_mcs: int, sampling_interval_mcs: int,
                      include_lattice: bool = False, mode: str = "exact_replay",
                      counters: dict | None = None) -> list:
    """Runs every configured `CPM` builder in `sims` to completion (its own
    grid, cell types, theta and seed), all resident on the GPU together: the
    per-replica `CPM::new` + `initialize` + `run_with` of
    crates/cpm-core/src/batch.rs:130-145 without a shared grid."""
    specs = [s._windowed(burn_in_mcs, readout_mcs, sampling_interval_mcs) for s in sims]
    holders = [N.ConfigHolder(sp) for sp in specs]
    n = len(holders)
    ptrs = (C.POINTER(N.CpmConfig) * max(n, 1))(*[C.pointer(h.c) for h in holders])
    handles = (C.c_void_p * max(n, 1))()
    any_used = C.c_int32(0)
    cnt = N.CpmCounters()
    N.check(N.lib().cpm_run_batch_configs(N.engine(), ptrs, n, 1 if include_lattice else 0,
                                          N.MODE[mode], handles, C.byref(any_used),
                                          C.byref(cnt)))
    if counters is not None:
        counters.update(cnt.as_dict())
    out = [N.RunResult(C.c_void_p(handles[i]), specs[i]["grid"],
                       [t["name"] for t in specs[i]["cell_types"]]) for i in range(n)]
    if any_used.value and any(s._warn_on_default_init for s in sims):
        warnings.warn("default (scatter-and-grow) initial placement used in at least one "
                      "batch member", CPMInitializationWarning, stacklevel=2)
    return out
