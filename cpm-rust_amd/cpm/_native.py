"""ctypes binding of the engine's C ABI (include/cpm_engine.h).

This module plays the role of the reference's PyO3 extension `cpm._cpm`
(crates/cpm-py/src/): it marshals builder state into `cpm_config`, maps
every CONFIG/INIT failure to `ValueError` with the engine's message
(crates/cpm-py/src/errors.rs:16-33) and converts results into NumPy-backed
objects (crates/cpm-py/src/convert.rs).  The library is loaded from this
package directory; there is no fallback implementation — importing the
package without the built library raises.
"""

from __future__ import annotations

import atexit
import ctypes as C
import os
import threading
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcpm_engine.so")

CPM_OK, CPM_ERR_CONFIG, CPM_ERR_INIT, CPM_ERR_DEVICE, CPM_ERR_OOM, CPM_ERR_ARG = range(6)
BOUNDARY = {"fixed": 0, "periodic": 1}
NB = {"von_neumann": 0, "moore": 1, "eighteen": 2, "twenty_six": 3, "twenty_six_weighted": 4}
NB_DEFAULT = -1
ACCEPT = {"metropolis": 0, "metropolis_hastings": 1}
PROPOSAL = {"uniform": 0, "edge_list": 1}
MODE = {"exact_replay": 0, "sublattice": 1}
CPM_REQ_QUEUED, CPM_REQ_ACTIVE, CPM_REQ_DONE, CPM_REQ_FAILED = range(4)

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)


class CpmConfig(C.Structure):
    _fields_ = [
        ("dim", C.c_uint32), ("grid", C.c_uint64 * 3), ("boundary", C.c_int32),
        ("copy_neighborhood", C.c_int32), ("energy_neighborhood", C.c_int32),
        ("connectivity_neighborhood", C.c_int32), ("acceptance", C.c_int32),
        ("proposal", C.c_int32), ("act_active", C.c_int32), ("n_types", C.c_uint32),
        ("type_names", C.POINTER(C.c_char_p)), ("target_volume", u32p),
        ("target_interface", u32p), ("lambda_volume", f64p), ("lambda_interface", f64p),
        ("lambda_act", f64p), ("max_act", u32p), ("cell_counts", u64p),
        ("adhesion_rows", C.c_uint32), ("adhesion_row_len", u32p), ("adhesion", f64p),
        ("init_explicit", C.c_int32), ("init_lattice", u32p), ("init_lattice_len", C.c_uint64),
        ("init_cell_type_of", u64p), ("init_n_cells", C.c_uint64),
        ("warn_on_default_init", C.c_int32), ("burn_in_mcs", C.c_uint64),
        ("readout_mcs", C.c_uint64), ("sampling_interval_mcs", C.c_uint64), ("seed", C.c_uint64),
    ]


class CpmCounters(C.Structure):
    _fields_ = [("attempts", C.c_uint64), ("non_null", C.c_uint64), ("topology", C.c_uint64),
                ("energy", C.c_uint64), ("accepted", C.c_uint64),
                ("exp_near_ties", C.c_uint64), ("kernel_ms", C.c_double),
                ("replay_launches", C.c_uint64), ("replay_launch_ms", C.c_double),
                ("piped_stragglers", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class CpmServerOptions(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_active", C.c_uint32),
                ("quantum_attempts", C.c_uint64), ("round_ms", C.c_double)]


class CpmRequestInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("submit_s", C.c_double), ("admit_s", C.c_double),
                ("done_s", C.c_double), ("footprint_bytes", C.c_uint64),
                ("mcs_done", C.c_uint64), ("mcs_total", C.c_uint64)]


class CpmServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("queued", C.c_uint64), ("active", C.c_uint64),
                ("hbm_budget_bytes", C.c_uint64), ("hbm_in_use_bytes", C.c_uint64),
                ("rounds", C.c_uint64), ("attempts", C.c_uint64), ("busy_s", C.c_double),
                ("device_s", C.c_double), ("uptime_s", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class CpmModelInfo(C.Structure):
    _fields_ = [("n_sites", C.c_uint64), ("n_cells", C.c_uint64), ("mcs", C.c_uint64),
                ("conservative_energy", C.c_double), ("rng_state", C.c_uint64 * 4)]


class CpmModelState(C.Structure):
    _fields_ = [("lattice", u32p), ("n_cells", C.c_uint64), ("type_index", u32p),
                ("volume", u32p), ("interface", u32p), ("sum_position", i64p),
                ("sum_outer", i64p), ("contact", u32p), ("has_energy", C.c_int32),
                ("conservative_energy", C.c_double), ("rng_state", u64p), ("mcs", C.c_uint64),
                ("birth_mcs", u32p), ("edge_sites", u64p), ("n_edge_sites", C.c_uint64)]


class CpmResultInfo(C.Structure):
    _fields_ = [
        ("cpm_version", C.c_char_p), ("convention_version", C.c_uint32),
        ("convention_hash", C.c_char_p), ("model_hash", C.c_char_p), ("seed", C.c_uint64),
        ("seed_policy", C.c_char_p), ("derived_phi", C.c_double),
        ("theta_layout_version", C.c_uint32), ("status_kind", C.c_char_p),
        ("status_has_cell", C.c_int32), ("status_cell_id", C.c_uint32),
        ("status_at_mcs", C.c_uint64), ("n_samples", C.c_uint64), ("n_cells", C.c_uint64),
        ("dim", C.c_uint32), ("cell_type_index", u32p), ("used_default", C.c_int32),
    ]


class CpmSampleView(C.Structure):
    _fields_ = [
        ("mcs", C.c_uint64), ("conservative_energy", C.c_double), ("n_cells", C.c_uint64),
        ("ids", u32p), ("volume", u32p), ("interface", u32p), ("centroid", f64p),
        ("eigenvalues", f64p), ("radius_of_gyration_sq", f64p), ("kappa_sq", f64p),
        ("contact", u32p), ("n_fragmented", C.c_uint64), ("fragmented_cells", u32p),
        ("n_medium_components", C.c_uint64), ("medium_component_volumes", u32p),
        ("medium_component_touches_boundary", u8p), ("lattice", u32p),
    ]


class CpmSamplesShape(C.Structure):
    _fields_ = [
        ("n_samples", C.c_uint64), ("n_cells", C.c_uint64), ("dim", C.c_uint32),
        ("has_lattice", C.c_int32), ("lattice_sites", C.c_uint64),
        ("n_fragmented_total", C.c_uint64), ("n_medium_total", C.c_uint64),
    ]


class CpmSamplesBuffers(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in (
        "mcs", "conservative_energy", "volume", "interface", "centroid", "eigenvalues",
        "radius_of_gyration_sq", "kappa_sq", "contact", "fragmented_offsets",
        "fragmented_cells", "medium_offsets", "medium_component_volumes",
        "medium_component_touches_boundary", "lattice")]


_lib = None
_lib_lock = threading.Lock()


def lib():
    """The loaded engine library.  Raises if it has not been built."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"cpm engine library not built: {LIB_PATH} is missing "
                "(run `python cpm-rust_amd/build.py`)")
        L = C.CDLL(LIB_PATH)
        st = C.c_int
        vp = C.c_void_p
        L.cpm_last_error.restype = C.c_char_p
        L.cpm_version.restype = C.c_char_p
        L.cpm_engine_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.cpm_engine_create.restype = st
        L.cpm_engine_destroy.argtypes = [vp]
        L.cpm_model_create.argtypes = [C.POINTER(CpmConfig), C.POINTER(vp)]
        L.cpm_model_create.restype = st
        L.cpm_model_destroy.argtypes = [vp]
        L.cpm_model_initialize.argtypes = [vp, C.POINTER(C.c_int32)]
        L.cpm_model_initialize.restype = st
        L.cpm_model_initialize_batch.argtypes = [vp, C.POINTER(vp), C.c_size_t,
                                                 C.POINTER(C.c_int32)]
        L.cpm_model_initialize_batch.restype = st
        L.cpm_run_mcs.argtypes = [vp, C.POINTER(vp), C.c_size_t, C.c_uint64, C.c_int32,
                                  C.POINTER(CpmCounters)]
        L.cpm_run_mcs.restype = st
        L.cpm_run_mcs_noalloc.argtypes = [vp, C.POINTER(vp), C.c_size_t, C.c_uint64,
                                          C.POINTER(CpmCounters)]
        L.cpm_run_mcs_noalloc.restype = st
        L.cpm_run_mcs_plain.argtypes = [vp, vp, C.c_uint64, C.POINTER(CpmCounters)]
        L.cpm_run_mcs_plain.restype = st
        L.cpm_model_get_info.argtypes = [vp, C.POINTER(CpmModelInfo)]
        L.cpm_model_get_info.restype = st
        L.cpm_model_get_state.argtypes = [vp, u32p, u32p, u32p, u32p, i64p, i64p, u32p]
        L.cpm_model_get_state.restype = st
        L.cpm_model_hashes.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.cpm_model_hashes.restype = st
        L.cpm_model_global_energy.argtypes = [vp, f64p]
        L.cpm_model_global_energy.restype = st
        L.cpm_model_set_state.argtypes = [vp, C.POINTER(CpmModelState)]
        L.cpm_model_set_state.restype = st
        L.cpm_model_sample.argtypes = [vp, C.c_int32, C.POINTER(vp)]
        L.cpm_model_sample.restype = st
        L.cpm_not_equilibrated.argtypes = [f64p, C.c_size_t, C.POINTER(C.c_int32)]
        L.cpm_not_equilibrated.restype = st
        L.cpm_run.argtypes = [vp, vp, C.c_int32, C.c_int32, C.POINTER(vp)]
        L.cpm_run.restype = st
        L.cpm_run_batch.argtypes = [vp, C.POINTER(CpmConfig), f64p, C.c_size_t, C.c_size_t,
                                    C.c_uint64, C.c_int32, C.c_int32, C.POINTER(vp),
                                    C.POINTER(C.c_int32)]
        L.cpm_run_batch.restype = st
        L.cpm_run_batch_configs.argtypes = [vp, C.POINTER(C.POINTER(CpmConfig)), C.c_size_t,
                                            C.c_int32, C.c_int32, C.POINTER(vp),
                                            C.POINTER(C.c_int32), C.POINTER(CpmCounters)]
        L.cpm_run_batch_configs.restype = st
        L.cpm_result_get_info.argtypes = [vp, C.POINTER(CpmResultInfo)]
        L.cpm_result_get_info.restype = st
        L.cpm_result_get_sample.argtypes = [vp, C.c_uint64, C.POINTER(CpmSampleView)]
        L.cpm_result_get_sample.restype = st
        L.cpm_result_get_samples_shape.argtypes = [vp, C.POINTER(CpmSamplesShape)]
        L.cpm_result_get_samples_shape.restype = st
        L.cpm_result_copy_samples.argtypes = [vp, C.POINTER(CpmSamplesBuffers)]
        L.cpm_result_copy_samples.restype = st
        L.cpm_result_destroy.argtypes = [vp]
        L.cpm_theta_len.argtypes = [C.c_uint32]
        L.cpm_theta_len.restype = C.c_size_t
        L.cpm_theta_extract.argtypes = [C.POINTER(CpmConfig), f64p, C.c_size_t]
        L.cpm_theta_extract.restype = st
        L.cpm_theta_name.argtypes = [C.POINTER(CpmConfig), C.c_size_t, C.c_char_p, C.c_size_t]
        L.cpm_theta_name.restype = st
        L.cpm_derive_seed_u64.argtypes = [C.c_uint64, C.c_uint64]
        L.cpm_derive_seed_u64.restype = C.c_uint64
        L.cpm_engine_sp_tables.argtypes = [vp, u8p, u64p]
        L.cpm_engine_sp_tables.restype = st
        L.cpm_engine_set_sigma_placement.argtypes = [vp, C.c_int32, C.c_double]
        L.cpm_engine_set_sigma_placement.restype = st
        L.cpm_metropolis_check.argtypes = [vp, f64p, f64p, C.c_size_t, f64p, f64p, u8p]
        L.cpm_metropolis_check.restype = st
        L.cpm_act_root_check.argtypes = [vp, f64p, u32p, C.c_size_t, f64p]
        L.cpm_act_root_check.restype = st
        L.cpm_mh_check.argtypes = [vp, u32p, u32p, f64p, f64p, C.c_size_t, u8p]
        L.cpm_mh_check.restype = st
        L.cpm_model_birth_mcs.argtypes = [vp, u32p, C.c_size_t]
        L.cpm_model_birth_mcs.restype = st
        L.cpm_model_edge_sites.argtypes = [vp, u64p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.cpm_model_edge_sites.restype = st
        L.cpm_binomial_check.argtypes = [vp, u64p, u64p, f64p, C.c_size_t, u64p]
        L.cpm_binomial_check.restype = st
        L.cpm_engine_set_pipelining.argtypes = [vp, C.c_int32]
        L.cpm_engine_set_pipelining.restype = st
        L.cpm_engine_sigma_placement.argtypes = [vp, C.POINTER(C.c_int32),
                                                 C.POINTER(C.c_double)]
        L.cpm_engine_sigma_placement.restype = st
        L.cpm_model_chain_rate.argtypes = [vp, C.POINTER(C.c_double)]
        L.cpm_model_chain_rate.restype = st
        L.cpm_engine_set_device_init.argtypes = [vp, C.c_int32]
        L.cpm_engine_set_device_init.restype = st
        L.cpm_engine_kernel_launches.argtypes = [vp, C.POINTER(C.c_int32), u64p, C.c_size_t,
                                                 C.POINTER(C.c_size_t)]
        L.cpm_engine_kernel_launches.restype = st
        L.cpm_serve_create.argtypes = [vp, C.POINTER(CpmServerOptions), C.POINTER(vp)]
        L.cpm_serve_create.restype = st
        L.cpm_serve_destroy.argtypes = [vp]
        L.cpm_serve_submit.argtypes = [vp, C.POINTER(CpmConfig), C.c_int32, C.c_int32,
                                       C.POINTER(C.c_uint64), C.POINTER(C.c_int32)]
        L.cpm_serve_submit.restype = st
        L.cpm_serve_poll.argtypes = [vp, C.c_uint64, C.c_double, C.POINTER(CpmRequestInfo),
                                     C.POINTER(vp)]
        L.cpm_serve_poll.restype = st
        L.cpm_serve_stats.argtypes = [vp, C.POINTER(CpmServerStats)]
        L.cpm_serve_stats.restype = st
        _lib = L
        return L


class EngineError(RuntimeError):
    """A device-side failure (no GPU, HIP error, out of memory)."""


def check(status: int) -> None:
    if status == CPM_OK:
        return
    msg = lib().cpm_last_error().decode("utf-8", "replace")
    if status in (CPM_ERR_CONFIG, CPM_ERR_INIT, CPM_ERR_ARG):
        raise ValueError(msg)
    raise EngineError(msg)


_engine = None
_engine_lock = threading.Lock()
# servers still open, closed before the engine at interpreter exit
_servers = weakref.WeakSet()


def engine() -> C.c_void_p:
    """The process-wide engine on HIP device 0 (created on first use and
    destroyed at interpreter exit, after any server still open)."""
    global _engine
    with _engine_lock:
        if _engine is None:
            h = C.c_void_p()
            check(lib().cpm_engine_create(0, C.byref(h)))
            _engine = h
            atexit.register(_destroy_engine)
        return _engine


def set_sigma_placement(order: str = "auto", share: float = 0.0) -> None:
    """Fixes the exact-replay sigma placement of the process-wide engine
    ("smallest" or "longest" first, filling `share` of the chip's LDS) or
    returns it to the autotuner ("auto")."""
    code = {"auto": 0, "smallest": 1, "longest": 2}[order]
    check(lib().cpm_engine_set_sigma_placement(engine(), code, float(share)))


def metropolis_check(delta_h, u=None, device: bool = True):
    """(exp_libm, exp_cr, accept) of cpm_metropolis_check: the device's (or,
    with device=False, the host's) libm exp(-dH), the correctly rounded
    exp(-dH), and the exact-replay decision for u (None without u)."""
    dh = np.ascontiguousarray(np.asarray(delta_h, dtype=np.float64))
    pl = np.empty_like(dh)
    pc = np.empty_like(dh)
    uu = None if u is None else np.ascontiguousarray(np.asarray(u, dtype=np.float64))
    acc = None if u is None else np.empty(dh.size, dtype=np.uint8)
    check(lib().cpm_metropolis_check(
        engine() if device else None, dh.ctypes.data_as(f64p),
        None if uu is None else uu.ctypes.data_as(f64p), dh.size, pl.ctypes.data_as(f64p),
        pc.ctypes.data_as(f64p), None if acc is None else acc.ctypes.data_as(u8p)))
    return pl, pc, (None if acc is None else acc.astype(bool))


def act_root_check(product, count, device: bool = True):
    """cpm_act_root_check: the Act term's geometric-mean root
    product^(1/count) as the exact-replay kernels compute it on the device
    (or, with device=False, the same code on the host)."""
    x = np.ascontiguousarray(np.asarray(product, dtype=np.float64))
    n = np.ascontiguousarray(np.asarray(count, dtype=np.uint32))
    if x.shape != n.shape:
        raise ValueError("product and count must have the same shape")
    out = np.empty_like(x)
    check(lib().cpm_act_root_check(engine() if device else None, x.ctypes.data_as(f64p),
                                   n.ctypes.data_as(u32p), x.size, out.ctypes.data_as(f64p)))
    return out


def mh_check(n_a, n_b, delta_h, u, device: bool = True):
    """cpm_mh_check: the exact-replay Metropolis-Hastings decisions."""
    a = np.ascontiguousarray(np.asarray(n_a, dtype=np.uint32))
    b = np.ascontiguousarray(np.asarray(n_b, dtype=np.uint32))
    dh = np.ascontiguousarray(np.asarray(delta_h, dtype=np.float64))
    uu = np.ascontiguousarray(np.asarray(u, dtype=np.float64))
    if not (a.shape == b.shape == dh.shape == uu.shape):
        raise ValueError("n_a, n_b, delta_h and u must have the same shape")
    acc = np.empty(dh.size, dtype=np.uint8)
    check(lib().cpm_mh_check(engine() if device else None, a.ctypes.data_as(u32p),
                             b.ctypes.data_as(u32p), dh.ctypes.data_as(f64p),
                             uu.ctypes.data_as(f64p), dh.size, acc.ctypes.data_as(u8p)))
    return acc.astype(bool)


def binomial_check(states, n, p, device: bool = True):
    """cpm_binomial_check: one Binomial(n[i], p[i]) sample per Xoshiro256++
    state (rows of `states`, shape (count, 4)) as the exact-replay kernels
    draw it; returns (k, advanced states)."""
    st = np.ascontiguousarray(np.array(states, dtype=np.uint64).reshape(-1, 4))
    nn = np.ascontiguousarray(np.asarray(n, dtype=np.uint64))
    pp = np.ascontiguousarray(np.asarray(p, dtype=np.float64))
    if not (len(st) == nn.size == pp.size):
        raise ValueError("states, n and p must have the same length")
    k = np.zeros(nn.size, dtype=np.uint64)
    check(lib().cpm_binomial_check(engine() if device else None, st.ctypes.data_as(u64p),
                                   nn.ctypes.data_as(u64p), pp.ctypes.data_as(f64p), nn.size,
                                   k.ctypes.data_as(u64p)))
    return k, st


def sigma_placement() -> tuple[str, float]:
    """The placement ("smallest"/"longest" first, LDS share) of the last
    exact-replay call ("none", 0.0 before any)."""
    o, sh = C.c_int32(0), C.c_double(0.0)
    check(lib().cpm_engine_sigma_placement(engine(), C.byref(o), C.byref(sh)))
    return {0: "none", 1: "smallest", 2: "longest"}[o.value], float(sh.value)


def set_pipelining(on: bool) -> None:
    """Allows (default) or forbids the two-wave pipelined exact-replay kernel."""
    check(lib().cpm_engine_set_pipelining(engine(), 1 if on else 0))


def set_device_init(on: bool) -> None:
    """Batches run scatter_and_grow on the device (default) or on the host
    threads; the initial states are the same."""
    check(lib().cpm_engine_set_device_init(engine(), 1 if on else 0))


def kernel_launches() -> dict:
    """{instantiation key: launches} of the exact-replay wave kernels
    (key layout in cpm_engine_kernel_launches)."""
    n = C.c_size_t(0)
    check(lib().cpm_engine_kernel_launches(engine(), None, None, 0, C.byref(n)))
    cap = max(int(n.value), 1)
    keys = (C.c_int32 * cap)()
    counts = (C.c_uint64 * cap)()
    check(lib().cpm_engine_kernel_launches(engine(), keys, counts, cap, C.byref(n)))
    return {int(keys[i]): int(counts[i]) for i in range(int(n.value))}


def _destroy_engine() -> None:
    """Releases the engine's streams, events and device buffers while the
    HIP runtime is still up (models and results hold their own memory and
    may be freed later)."""
    global _engine
    for srv in list(_servers):
        srv.close()
    with _engine_lock:
        if _engine is not None:
            lib().cpm_engine_destroy(_engine)
            _engine = None


def _arr(values, dtype):
    return np.ascontiguousarray(np.asarray(values, dtype=dtype))


class ConfigHolder:
    """A filled `cpm_config` plus the buffers it points into."""

    def __init__(self, spec: dict):
        self.keep = []
        c = CpmConfig()
        grid = list(spec["grid"])
        c.dim = len(grid)
        for i, g in enumerate(grid[:3]):
            c.grid[i] = int(g)
        c.boundary = spec["boundary"]
        c.copy_neighborhood = spec["copy_neighborhood"]
        c.energy_neighborhood = spec["energy_neighborhood"]
        c.connectivity_neighborhood = spec["connectivity_neighborhood"]
        c.acceptance = spec["acceptance"]
        c.proposal = spec["proposal"]
        c.act_active = 1 if spec["act"] else 0
        types = spec["cell_types"]
        c.n_types = len(types)
        names = (C.c_char_p * max(len(types), 1))(*[t["name"].encode("utf-8") for t in types])
        self.keep.append(names)
        c.type_names = C.cast(names, C.POINTER(C.c_char_p))

        def buf(values, dtype, ptype):
            a = _arr(values if len(values) else [0], dtype)
            self.keep.append(a)
            return a.ctypes.data_as(ptype)

        c.target_volume = buf([t["target_volume"] for t in types], np.uint32, u32p)
        c.target_interface = buf([t["target_interface"] for t in types], np.uint32, u32p)
        c.lambda_volume = buf([t["lambda_volume"] for t in types], np.float64, f64p)
        c.lambda_interface = buf([t["lambda_interface"] for t in types], np.float64, f64p)
        c.lambda_act = buf([t["lambda_act"] for t in types], np.float64, f64p)
        c.max_act = buf([t["max_act"] for t in types], np.uint32, u32p)
        c.cell_counts = buf(spec["cell_counts"], np.uint64, u64p)
        adh = spec["adhesion"]
        if adh is None:
            c.adhesion_rows = 0
        else:
            c.adhesion_rows = len(adh)
            c.adhesion_row_len = buf([len(r) for r in adh], np.uint32, u32p)
            flat = [float(v) for r in adh for v in r]
            c.adhesion = buf(flat, np.float64, f64p)
        init = spec["init"]
        if init is not None and init["lattice"] is not None:
            c.init_explicit = 1
            lat = init["lattice"]
            c.init_lattice = buf(lat, np.uint32, u32p)
            c.init_lattice_len = len(lat)
            cto = init["cell_type_of"]
            c.init_cell_type_of = buf(cto, np.uint64, u64p)
            c.init_n_cells = len(cto)
        c.warn_on_default_init = 1 if (init is None or init["warn"]) else 0
        c.burn_in_mcs = spec.get("burn_in_mcs", 1)
        c.readout_mcs = spec.get("readout_mcs", 1)
        c.sampling_interval_mcs = spec.get("sampling_interval_mcs", 1)
        c.seed = spec["seed"]
        self.c = c

    def ref(self):
        return C.byref(self.c)


# ------------------------------------------------------------------ results

class RunMetadata:
    """crates/cpm-py/src/convert.rs:27-38"""

    def __init__(self, info: CpmResultInfo):
        self.cpm_version = info.cpm_version.decode()
        self.convention_version = int(info.convention_version)
        self.convention_hash = info.convention_hash.decode()
        self.model_hash = info.model_hash.decode()
        self.seed = int(info.seed)
        self.seed_policy = info.seed_policy.decode()
        self.derived_phi = float(info.derived_phi)
        self.theta_layout_version = int(info.theta_layout_version)


class Status:
    """crates/cpm-py/src/convert.rs:40-51"""

    def __init__(self, info: CpmResultInfo):
        self.kind = info.status_kind.decode()
        self.cell_id = int(info.status_cell_id) if info.status_has_cell else None
        self.at_mcs = int(info.status_at_mcs) if info.status_has_cell else None


def _np(ptr, n, dtype, shape=None):
    """A NumPy copy of n elements at a C pointer (one C-level copy: batch
    results hold tens of thousands of samples, so the per-call cost counts)."""
    dtype = np.dtype(dtype)
    if n == 0:
        out = np.zeros(0, dtype=dtype)
    else:
        out = np.frombuffer(C.string_at(ptr, n * dtype.itemsize), dtype=dtype).copy()
    return out.reshape(shape) if shape is not None else out


class Sample:
    """crates/cpm-py/src/convert.rs:53-93"""

    @classmethod
    def _of_packed(cls, k, a, grid):
        """Sample k of the field-by-field arrays `a` (each field one NumPy
        array over all samples; the per-sample arrays are rows of it)."""
        self = cls.__new__(cls)
        self.mcs = int(a["mcs"][k])
        self.conservative_energy = float(a["energy"][k])
        self.ids = a["ids"].copy()
        self.volume = a["volume"][k]
        self.interface = a["interface"][k]
        self.centroid = a["centroid"][k]
        self.eigenvalues = a["eigenvalues"][k]
        self.radius_of_gyration_sq = a["rg2"][k]
        self.kappa_sq = a["kappa"][k]
        self.contact = a["contact"][k]
        f0, f1 = a["frag_off"][k], a["frag_off"][k + 1]
        self.fragmented_cells = a["frag"][f0:f1].tolist()
        m0, m1 = a["med_off"][k], a["med_off"][k + 1]
        self.medium_component_volumes = a["med_vol"][m0:m1].tolist()
        self.medium_component_touches_boundary = [bool(x) for x in a["med_touch"][m0:m1]]
        self.lattice = a["lattice"][k] if a["lattice"] is not None else None
        return self

    def __init__(self, v: CpmSampleView, dim: int, grid):
        n = int(v.n_cells)
        self.mcs = int(v.mcs)
        self.conservative_energy = float(v.conservative_energy)
        self.ids = _np(v.ids, n, np.uint32)
        self.volume = _np(v.volume, n, np.uint32)
        self.interface = _np(v.interface, n, np.uint32)
        self.centroid = _np(v.centroid, n * dim, np.float64, (n, dim))
        self.eigenvalues = _np(v.eigenvalues, n * dim, np.float64, (n, dim))
        self.radius_of_gyration_sq = _np(v.radius_of_gyration_sq, n, np.float64)
        self.kappa_sq = _np(v.kappa_sq, n, np.float64)
        self.contact = _np(v.contact, (n + 1) * (n + 1), np.uint32, (n + 1, n + 1))
        self.fragmented_cells = _np(v.fragmented_cells, int(v.n_fragmented), np.uint32).tolist()
        k = int(v.n_medium_components)
        self.medium_component_volumes = _np(v.medium_component_volumes, k, np.uint32).tolist()
        self.medium_component_touches_boundary = [
            bool(x) for x in _np(v.medium_component_touches_boundary, k, np.uint8)]
        if v.lattice:
            self.lattice = _np(v.lattice, int(np.prod(grid)), np.uint32, tuple(grid))
        else:
            self.lattice = None


class RunResult:
    """crates/cpm-py/src/convert.rs:95-115"""

    def __init__(self, handle: C.c_void_p, grid, cell_type_names):
        L = lib()
        info = CpmResultInfo()
        check(L.cpm_result_get_info(handle, C.byref(info)))
        self.metadata = RunMetadata(info)
        self.status = Status(info)
        dim = int(info.dim)
        sh = CpmSamplesShape()
        check(L.cpm_result_get_samples_shape(handle, C.byref(sh)))
        ns, n = int(sh.n_samples), int(sh.n_cells)
        a = {
            "mcs": np.empty(ns, np.uint64), "energy": np.empty(ns, np.float64),
            "ids": np.arange(1, n + 1, dtype=np.uint32),
            "volume": np.empty((ns, n), np.uint32), "interface": np.empty((ns, n), np.uint32),
            "centroid": np.empty((ns, n, dim)), "eigenvalues": np.empty((ns, n, dim)),
            "rg2": np.empty((ns, n)), "kappa": np.empty((ns, n)),
            "contact": np.empty((ns, n + 1, n + 1), np.uint32),
            "frag_off": np.empty(ns + 1, np.uint64), "med_off": np.empty(ns + 1, np.uint64),
            "frag": np.empty(int(sh.n_fragmented_total), np.uint32),
            "med_vol": np.empty(int(sh.n_medium_total), np.uint32),
            "med_touch": np.empty(int(sh.n_medium_total), np.uint8),
            "lattice": (np.empty((ns,) + tuple(grid), np.uint32) if sh.has_lattice else None),
        }
        def ptr(x):
            return x.ctypes.data if x is not None else None
        buf = CpmSamplesBuffers(
            ptr(a["mcs"]), ptr(a["energy"]), ptr(a["volume"]), ptr(a["interface"]),
            ptr(a["centroid"]), ptr(a["eigenvalues"]), ptr(a["rg2"]), ptr(a["kappa"]),
            ptr(a["contact"]), ptr(a["frag_off"]), ptr(a["frag"]), ptr(a["med_off"]),
            ptr(a["med_vol"]), ptr(a["med_touch"]), ptr(a["lattice"]))
        check(L.cpm_result_copy_samples(handle, C.byref(buf)))
        self.samples = [Sample._of_packed(k, a, grid) for k in range(ns)]
        self.cell_type_index = _np(info.cell_type_index, int(info.n_cells), np.uint32)
        self.cell_type_names = list(cell_type_names)
        self.used_default = bool(info.used_default)
        L.cpm_result_destroy(handle)
