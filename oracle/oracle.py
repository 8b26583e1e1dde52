"""ctypes front-end of the CPU oracle — TEST INFRASTRUCTURE ONLY.

Only `tests/`, `__graft_entry__.smoke()` and the CPU-baseline leg of
`bench.py` import this module, and only as the checker.  It loads
`oracle/_build/libcpm_oracle.so` (built by `make -C oracle`) and adds the
pieces of the reference that are plain data plumbing:

* the `run_with` driver loop (crates/cpm-core/src/output.rs:434-468) and the
  status rules (output.rs:388-559),
* per-cell readouts (output.rs:147-231, 289-302),
* `convention_hash` / `model_hash` (crates/cpm-core/src/config.rs:423-556):
  the serde_json compact byte layout of the two hashed views, hashed with
  the oracle's BLAKE3,
* `theta` layout and injection (crates/cpm-core/src/theta.rs:43-178).

Pins: see cpm_oracle.h.  `model_hash` is pinned by the frozen trajectory's
hash (crates/cpm-core/tests/dimension_consistency.rs:127-130); the readout
formulas by closed forms in tests/test_oracle.py.
"""

from __future__ import annotations

import ctypes as C
import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libcpm_oracle.so")

NB = {"von_neumann": 0, "moore": 1, "eighteen": 2, "twenty_six": 3}
NB_SERDE = {
    "von_neumann": "VonNeumann",
    "moore": "Moore",
    "eighteen": "Eighteen",
    "twenty_six": "TwentySix",
    "twenty_six_weighted": "TwentySixWeighted",
}
FIXED, PERIODIC = 0, 1
CONVENTION_VERSION = 1
THETA_LAYOUT_VERSION = 1

_lib = None


def build() -> None:
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    u64p = C.POINTER(C.c_uint64)
    u32p = C.POINTER(C.c_uint32)
    i64p = C.POINTER(C.c_int64)
    f64p = C.POINTER(C.c_double)
    u8p = C.POINTER(C.c_uint8)
    L.oc_rng_seed.argtypes = [C.c_uint64, u64p]
    L.oc_rng_next.argtypes = [u64p]
    L.oc_rng_next.restype = C.c_uint64
    L.oc_rng_gen_range.argtypes = [u64p, C.c_uint64]
    L.oc_rng_gen_range.restype = C.c_uint64
    L.oc_rng_gen_range_incl.argtypes = [u64p, C.c_uint64]
    L.oc_rng_gen_range_incl.restype = C.c_uint64
    L.oc_rng_gen_f64.argtypes = [u64p]
    L.oc_rng_gen_f64.restype = C.c_double
    L.oc_blake3.argtypes = [C.c_char_p, C.c_size_t, u8p]
    L.oc_derive_seed_u64.argtypes = [C.c_uint64, C.c_uint64]
    L.oc_derive_seed_u64.restype = C.c_uint64
    L.oc_sp_slow.argtypes = [C.c_int, C.c_uint32]
    L.oc_sp_fast.argtypes = [C.c_int, C.c_uint32]
    L.oc_sp_table2d.argtypes = [u8p]
    L.oc_sp_table3d.argtypes = [u64p]
    L.oc_full_offsets.argtypes = [C.c_int, C.POINTER(C.c_int32)]
    L.oc_model_new.argtypes = [C.c_int, u64p, C.c_int, C.c_int, C.c_int, C.c_int,
                               f64p, f64p, u32p, u32p, f64p, C.c_uint64]
    L.oc_model_new.restype = C.c_void_p
    L.oc_model_free.argtypes = [C.c_void_p]
    L.oc_set_halo_mode.argtypes = [C.c_void_p, C.c_int]
    L.oc_init_explicit.argtypes = [C.c_void_p, u32p, C.c_uint64, u64p]
    L.oc_init_scatter.argtypes = [C.c_void_p, u64p]
    L.oc_load_state.argtypes = [C.c_void_p, u32p, C.c_uint64, u64p]
    L.oc_run_mcs.argtypes = [C.c_void_p, C.c_uint64]
    L.oc_global_energy.argtypes = [C.c_void_p]
    L.oc_global_energy.restype = C.c_double
    for name in ("oc_n_sites", "oc_n_cells", "oc_mcs"):
        getattr(L, name).argtypes = [C.c_void_p]
        getattr(L, name).restype = C.c_uint64
    L.oc_energy.argtypes = [C.c_void_p]
    L.oc_energy.restype = C.c_double
    L.oc_rng_state.argtypes = [C.c_void_p, u64p]
    for name in ("oc_lattice", "oc_padded_lattice", "oc_volume", "oc_interface",
                 "oc_type_index", "oc_contact"):
        getattr(L, name).argtypes = [C.c_void_p, u32p]
    L.oc_sum_position.argtypes = [C.c_void_p, i64p]
    L.oc_sum_outer.argtypes = [C.c_void_p, i64p]
    L.oc_counters.argtypes = [C.c_void_p, u64p]
    L.oc_recompute.argtypes = [C.c_void_p, u32p, i64p, i64p, u32p]
    L.oc_label.argtypes = [C.c_void_p, u32p, C.c_uint64, u32p, u8p, C.c_uint64, u64p]
    L.oc_label.restype = C.c_uint64
    L.oc_metropolis_log.argtypes = [C.c_void_p, f64p, C.c_uint64]
    L.oc_metropolis_log.restype = C.c_uint64
    L.oc_exp_neg.argtypes = [f64p, C.c_uint64, f64p]
    L.oc_set_modes.argtypes = [C.c_void_p, C.c_int, C.c_int, f64p, u32p]
    L.oc_set_connectivity_check.argtypes = [C.c_void_p, C.c_int]
    L.oc_birth_mcs.argtypes = [C.c_void_p, u32p]
    L.oc_set_birth_mcs.argtypes = [C.c_void_p, u32p]
    L.oc_act_delta.argtypes = [C.c_void_p, u64p, C.c_int, C.c_uint64]
    L.oc_act_delta.restype = C.c_double
    L.oc_root.argtypes = [f64p, u32p, C.c_uint64, f64p]
    L.oc_set_proposal.argtypes = [C.c_void_p, C.c_int]
    L.oc_edge_sites.argtypes = [C.c_void_p, u64p, C.c_uint64]
    L.oc_edge_sites.restype = C.c_uint64
    L.oc_set_edge_sites.argtypes = [C.c_void_p, u64p, C.c_uint64]
    L.oc_binomial.argtypes = [u64p, C.c_uint64, C.c_double]
    L.oc_binomial.restype = C.c_uint64
    _lib = L
    return L


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


# ---------------------------------------------------------------- RNG / hash

class Rng:
    """Xoshiro256++ seeded like `rng_from_seed` (crates/cpm-core/src/rng.rs:12-14)."""

    def __init__(self, seed: int):
        self.s = np.zeros(4, dtype=np.uint64)
        lib().oc_rng_seed(seed, _ptr(self.s, C.c_uint64))

    def next_u64(self) -> int:
        return int(lib().oc_rng_next(_ptr(self.s, C.c_uint64)))

    def gen_range(self, n: int) -> int:
        return int(lib().oc_rng_gen_range(_ptr(self.s, C.c_uint64), n))

    def gen_range_incl(self, hi: int) -> int:
        return int(lib().oc_rng_gen_range_incl(_ptr(self.s, C.c_uint64), hi))

    def gen_f64(self) -> float:
        return float(lib().oc_rng_gen_f64(_ptr(self.s, C.c_uint64)))

    def binomial(self, n: int, p: float) -> int:
        """rand_distr 0.4.3 Binomial::new(n, p).sample on this stream."""
        return int(lib().oc_binomial(_ptr(self.s, C.c_uint64), n, p))


def blake3_hex(data: bytes) -> str:
    out = (C.c_uint8 * 32)()
    lib().oc_blake3(data, len(data), out)
    return bytes(out).hex()


def exp_neg(x: np.ndarray) -> np.ndarray:
    """exp(-x) with the platform libm, element by element (the reference's
    f64::exp is glibc's on its CI)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().oc_exp_neg(_ptr(x, C.c_double), x.size, _ptr(out, C.c_double))
    return out


def root(x: np.ndarray, n: np.ndarray) -> np.ndarray:
    """pow(x, 1 / n) with this platform's libm (the reference's f64::powf)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.uint32)
    out = np.empty_like(x)
    lib().oc_root(_ptr(x, C.c_double), _ptr(n, C.c_uint32), x.size, _ptr(out, C.c_double))
    return out


def derive_seed_u64(master: int, index: int) -> int:
    return int(lib().oc_derive_seed_u64(master, index))


def sp_table2d() -> np.ndarray:
    out = np.zeros(256, dtype=np.uint8)
    lib().oc_sp_table2d(_ptr(out, C.c_uint8))
    return out


_T3 = None


def sp_table3d_words() -> np.ndarray:
    global _T3
    if _T3 is None:
        out = np.zeros(1 << 20, dtype=np.uint64)
        lib().oc_sp_table3d(_ptr(out, C.c_uint64))
        _T3 = out
    return _T3


def full_offsets(dim: int) -> np.ndarray:
    out = np.zeros(26 * 3, dtype=np.int32)
    n = lib().oc_full_offsets(dim, _ptr(out, C.c_int32))
    return out[: n * dim].reshape(n, dim)


# ---------------------------------------------------------------- config

@dataclass
class CellTypeSpec:
    name: str
    target_volume: int
    target_interface: int
    lambda_volume: float
    lambda_interface: float
    lambda_act: float = 0.0
    max_act: int = 0


@dataclass
class Config:
    """The resolved configuration the oracle needs (config.rs:200-257 defaults)."""

    grid: tuple
    boundary: str
    seed: int
    cell_types: list
    cell_counts: list
    adhesion: list
    copy_neighborhood: str = "von_neumann"
    energy_neighborhood: str | None = None
    connectivity_neighborhood: str = "von_neumann"
    init_lattice: list | None = None
    init_cell_type_of: list | None = None
    warn_on_default_init: bool = True
    burn_in_mcs: int = 0
    readout_mcs: int = 1
    sampling_interval_mcs: int = 1
    acceptance: str = "metropolis"  # or "metropolis_hastings" (config.rs:140-200)
    act: bool = False               # active_terms.act (config.rs:84-103)
    proposal: str = "uniform"       # or "edge_list" (config.rs:42-70)

    @property
    def dim(self) -> int:
        return len(self.grid)

    @property
    def energy_nb(self) -> str:
        if self.energy_neighborhood is not None:
            return self.energy_neighborhood
        return "moore" if self.dim == 2 else "eighteen"


def theta_len(n_types: int) -> int:
    """theta.rs:43-51"""
    return 4 * n_types + (n_types + 1) * (n_types + 2) // 2


def theta_extract(cfg: Config) -> list:
    """theta.rs:55-78"""
    n = len(cfg.cell_types)
    th = [t.lambda_volume for t in cfg.cell_types]
    th += [t.lambda_interface for t in cfg.cell_types]
    for i in range(n + 1):
        for j in range(i, n + 1):
            th.append(float(cfg.adhesion[i][j]))
    th += [t.lambda_act for t in cfg.cell_types]
    th += [float(t.max_act) for t in cfg.cell_types]
    return th


def theta_inject(cfg: Config, theta) -> Config:
    """theta.rs:93-147 (returns a modified copy)."""
    import copy

    out = copy.deepcopy(cfg)
    n = len(out.cell_types)
    assert len(theta) == theta_len(n)
    cur = 0
    for t in out.cell_types:
        t.lambda_volume = float(theta[cur]); cur += 1
    for t in out.cell_types:
        t.lambda_interface = float(theta[cur]); cur += 1
    adh = [list(map(float, r)) for r in out.adhesion]
    for i in range(n + 1):
        for j in range(i, n + 1):
            adh[i][j] = adh[j][i] = float(theta[cur]); cur += 1
    out.adhesion = adh
    for t in out.cell_types:
        t.lambda_act = float(theta[cur]); cur += 1
    for t in out.cell_types:
        t.max_act = int(_round_half_away(theta[cur])); cur += 1
    return out


def _round_half_away(x: float) -> float:
    return math.floor(abs(x) + 0.5) * (1 if x >= 0 else -1)


# ---------------------------------------------------------------- hashes

def _serde(obj) -> str:
    return json.dumps(obj, separators=(",", ":"), ensure_ascii=False)


def _nb_json(kind: str) -> dict:
    return {"kind": NB_SERDE[kind]}


def convention_json(cfg: Config) -> bytes:
    """serde_json::to_vec of ConventionView (config.rs:423-497): struct fields
    in declaration order, enums as serde renders them."""
    if cfg.init_lattice is not None:
        init = {"Explicit": {"lattice": [int(v) for v in cfg.init_lattice],
                             "cell_type_of": [int(v) for v in cfg.init_cell_type_of]}}
    else:
        init = {"Default": {"warn_on_default_init": bool(cfg.warn_on_default_init)}}
    view = {
        "dimensionality": cfg.dim,
        "grid": [int(g) for g in cfg.grid],
        "copy_neighborhood": _nb_json(cfg.copy_neighborhood),
        "energy_neighborhood": _nb_json(cfg.energy_nb),
        "connectivity_neighborhood": _nb_json(cfg.connectivity_neighborhood),
        "boundary": "Periodic" if cfg.boundary == "periodic" else "Fixed",
        "acceptance": ("MetropolisHastings" if cfg.acceptance == "metropolis_hastings"
                       else "Metropolis"),
        "proposal": "EdgeList" if cfg.proposal == "edge_list" else "Uniform",
        "cell_types": [{"name": t.name, "target_volume": int(t.target_volume),
                        "target_interface": int(t.target_interface)} for t in cfg.cell_types],
        "active_terms": {"volume": True, "interface": True, "adhesion": True,
                         "act": bool(cfg.act)},
        "initialization": init,
        "burn_in_mcs": int(cfg.burn_in_mcs),
        "readout_mcs": int(cfg.readout_mcs),
        "sampling_interval_mcs": int(cfg.sampling_interval_mcs),
        "min_cell_volume": 1,
        "seed_policy": "Independent",
        "requested_readouts": [],
        "convention_version": CONVENTION_VERSION,
    }
    return _serde(view).encode("utf-8")


def convention_hash(cfg: Config) -> str:
    return blake3_hex(convention_json(cfg))


def model_hash(cfg: Config, lattice, cell_type_of) -> str:
    """config.rs:542-556"""
    view = {"convention_hash": convention_hash(cfg),
            "lattice": [int(v) for v in lattice],
            "cell_type_of": [int(v) for v in cell_type_of]}
    return blake3_hex(_serde(view).encode("utf-8"))


# ---------------------------------------------------------------- model

class Model:
    """`CPM<D>` + `initialize` (model.rs:44-76; initialization.rs:39-55)."""

    def __init__(self, cfg: Config, halo_mirror: bool = False):
        self.cfg = cfg
        L = lib()
        T = len(cfg.cell_types)
        dims = np.array(cfg.grid, dtype=np.uint64)
        lv = np.array([t.lambda_volume for t in cfg.cell_types], dtype=np.float64)
        li = np.array([t.lambda_interface for t in cfg.cell_types], dtype=np.float64)
        tv = np.array([t.target_volume for t in cfg.cell_types], dtype=np.uint32)
        ti = np.array([t.target_interface for t in cfg.cell_types], dtype=np.uint32)
        J = np.ascontiguousarray(np.array(cfg.adhesion, dtype=np.float64).reshape(T + 1, T + 1))
        self._keep = (dims, lv, li, tv, ti, J)
        self.h = L.oc_model_new(cfg.dim, _ptr(dims, C.c_uint64),
                                PERIODIC if cfg.boundary == "periodic" else FIXED,
                                NB[cfg.copy_neighborhood], NB[cfg.energy_nb], T,
                                _ptr(lv, C.c_double), _ptr(li, C.c_double),
                                _ptr(tv, C.c_uint32), _ptr(ti, C.c_uint32),
                                _ptr(J, C.c_double), cfg.seed)
        if halo_mirror:
            L.oc_set_halo_mode(self.h, 1)
        if cfg.acceptance not in ("metropolis", "metropolis_hastings"):
            raise ValueError(f"unknown acceptance {cfg.acceptance!r}")
        la = np.array([t.lambda_act for t in cfg.cell_types], dtype=np.float64)
        ma = np.array([t.max_act for t in cfg.cell_types], dtype=np.uint32)
        L.oc_set_modes(self.h, int(cfg.acceptance == "metropolis_hastings"), int(bool(cfg.act)),
                       _ptr(la, C.c_double), _ptr(ma, C.c_uint32))
        if cfg.proposal not in ("uniform", "edge_list"):
            raise ValueError(f"unknown proposal {cfg.proposal!r}")
        L.oc_set_proposal(self.h, int(cfg.proposal == "edge_list"))
        self.used_default = cfg.init_lattice is None
        if cfg.init_lattice is not None:
            lat = np.ascontiguousarray(np.asarray(cfg.init_lattice, dtype=np.uint32).ravel())
            cto = np.ascontiguousarray(np.asarray(cfg.init_cell_type_of, dtype=np.uint64))
            rc = L.oc_init_explicit(self.h, _ptr(lat, C.c_uint32), len(cto), _ptr(cto, C.c_uint64))
        else:
            cc = np.array(cfg.cell_counts, dtype=np.uint64)
            rc = L.oc_init_scatter(self.h, _ptr(cc, C.c_uint64))
        if rc != 0:
            raise ValueError(f"oracle initialisation failed (code {rc})")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_model_free(self.h)
            self.h = None

    def load_state(self, lattice, cell_type_of) -> None:
        """Replace the state by an arbitrary lattice (no checks) and recompute
        everything derived from it — the brute-force checker's view."""
        lat = np.ascontiguousarray(np.asarray(lattice, dtype=np.uint32).ravel())
        cto = np.ascontiguousarray(np.asarray(cell_type_of, dtype=np.uint64))
        lib().oc_load_state(self.h, _ptr(lat, C.c_uint32), len(cto), _ptr(cto, C.c_uint64))

    # --- dynamics
    def run_mcs(self, n: int) -> None:
        lib().oc_run_mcs(self.h, n)

    # --- state
    @property
    def n_sites(self) -> int:
        return int(lib().oc_n_sites(self.h))

    @property
    def n_cells(self) -> int:
        return int(lib().oc_n_cells(self.h))

    @property
    def mcs(self) -> int:
        return int(lib().oc_mcs(self.h))

    @property
    def energy(self) -> float:
        return float(lib().oc_energy(self.h))

    def global_energy(self) -> float:
        return float(lib().oc_global_energy(self.h))

    def rng_state(self) -> np.ndarray:
        out = np.zeros(4, dtype=np.uint64)
        lib().oc_rng_state(self.h, _ptr(out, C.c_uint64))
        return out

    def _u32(self, fn, n):
        out = np.zeros(n, dtype=np.uint32)
        fn(self.h, _ptr(out, C.c_uint32))
        return out

    def lattice(self) -> np.ndarray:
        return self._u32(lib().oc_lattice, self.n_sites).reshape(self.cfg.grid)

    def padded_lattice(self) -> np.ndarray:
        n = int(np.prod([g + 2 for g in self.cfg.grid]))
        return self._u32(lib().oc_padded_lattice, n)

    def volume(self):
        return self._u32(lib().oc_volume, self.n_cells)

    def interface(self):
        return self._u32(lib().oc_interface, self.n_cells)

    def type_index(self):
        return self._u32(lib().oc_type_index, self.n_cells)

    def contact(self):
        n = self.n_cells + 1
        return self._u32(lib().oc_contact, n * n).reshape(n, n)

    def sum_position(self):
        D = self.cfg.dim
        out = np.zeros(self.n_cells * D, dtype=np.int64)
        lib().oc_sum_position(self.h, _ptr(out, C.c_int64))
        return out.reshape(self.n_cells, D)

    def sum_outer(self):
        D = self.cfg.dim
        out = np.zeros(self.n_cells * D * D, dtype=np.int64)
        lib().oc_sum_outer(self.h, _ptr(out, C.c_int64))
        return out.reshape(self.n_cells, D, D)

    def counters(self):
        out = np.zeros(5, dtype=np.uint64)
        lib().oc_counters(self.h, _ptr(out, C.c_uint64))
        return out

    def recompute(self):
        n, D = self.n_cells, self.cfg.dim
        itf = np.zeros(n, dtype=np.uint32)
        sp = np.zeros(n * D, dtype=np.int64)
        so = np.zeros(n * D * D, dtype=np.int64)
        ct = np.zeros((n + 1) * (n + 1), dtype=np.uint32)
        lib().oc_recompute(self.h, _ptr(itf, C.c_uint32), _ptr(sp, C.c_int64),
                           _ptr(so, C.c_int64), _ptr(ct, C.c_uint32))
        return itf, sp.reshape(n, D), so.reshape(n, D, D), ct.reshape(n + 1, n + 1)

    def label(self):
        n = self.n_cells
        frag = np.zeros(max(n, 1), dtype=np.uint32)
        cap = self.n_sites
        mv = np.zeros(cap, dtype=np.uint32)
        mt = np.zeros(cap, dtype=np.uint8)
        nm = C.c_uint64(0)
        nf = lib().oc_label(self.h, _ptr(frag, C.c_uint32), len(frag), _ptr(mv, C.c_uint32),
                            _ptr(mt, C.c_uint8), cap, C.byref(nm))
        k = int(nm.value)
        return frag[:nf].tolist(), mv[:k].tolist(), [bool(x) for x in mt[:k]]

    def snapshot(self) -> dict:
        return {
            "lattice": self.lattice(), "volume": self.volume(), "interface": self.interface(),
            "sum_position": self.sum_position(), "sum_outer": self.sum_outer(),
            "contact": self.contact(), "energy": self.energy, "rng": self.rng_state(),
            "mcs": self.mcs, "birth_mcs": self.birth_mcs(),
        }

    # --- Act term and acceptance probes (energy.rs:329-471; monte_carlo.rs)
    def birth_mcs(self) -> np.ndarray:
        out = np.zeros(self.n_sites, dtype=np.uint32)
        lib().oc_birth_mcs(self.h, _ptr(out, C.c_uint32))
        return out.reshape(self.cfg.grid)

    def set_birth_mcs(self, birth) -> None:
        b = np.ascontiguousarray(np.asarray(birth, dtype=np.uint32).ravel())
        assert b.size == self.n_sites
        lib().oc_set_birth_mcs(self.h, _ptr(b, C.c_uint32))

    def act_delta(self, target, copy_k: int, mcs: int) -> float:
        tc = np.array(target, dtype=np.uint64)
        return lib().oc_act_delta(self.h, _ptr(tc, C.c_uint64), copy_k, mcs)

    def edge_sites(self) -> np.ndarray:
        """The EdgeList proposal's edge sites in list order (interior
        row-major indices); empty under the uniform proposal."""
        out = np.zeros(self.n_sites, dtype=np.uint64)
        n = lib().oc_edge_sites(self.h, _ptr(out, C.c_uint64), out.size)
        return out[:n]

    def set_edge_sites(self, sites) -> None:
        a = np.ascontiguousarray(np.asarray(sites, dtype=np.uint64))
        lib().oc_set_edge_sites(self.h, _ptr(a, C.c_uint64), a.size)

    def set_connectivity_check(self, on: bool) -> None:
        lib().oc_set_connectivity_check(self.h, int(bool(on)))

    def model_hash(self) -> str:
        return model_hash(self.cfg, self.lattice().ravel(), self.type_index())

    def record_metropolis(self, cap: int) -> None:
        """Start recording (dH, u) of every uphill Metropolis test (up to cap)."""
        self._mlog = np.zeros(2 * cap, dtype=np.float64)
        lib().oc_metropolis_log(self.h, _ptr(self._mlog, C.c_double), cap)

    def metropolis_draws(self) -> np.ndarray:
        """The recorded (dH, u) pairs, shape (k, 2); stops recording."""
        seen = int(lib().oc_metropolis_log(self.h, None, 0))
        k = min(seen, len(self._mlog) // 2)
        return self._mlog[: 2 * k].reshape(k, 2).copy()


# ---------------------------------------------------------------- readouts

def symmetric_eigenvalues(g: np.ndarray) -> list:
    """output.rs:193-243 closed forms."""
    D = g.shape[0]
    if D == 2:
        a, b, d = g[0, 0], g[0, 1], g[1, 1]
        tr = a + d
        disc = math.sqrt(max((a - d) * (a - d) + 4.0 * b * b, 0.0))
        return [(tr + disc) / 2.0, (tr - disc) / 2.0]
    p1 = g[0, 1] * g[0, 1] + g[0, 2] * g[0, 2] + g[1, 2] * g[1, 2]
    if p1 == 0.0:
        return [g[0, 0], g[1, 1], g[2, 2]]
    q = (g[0, 0] + g[1, 1] + g[2, 2]) / 3.0
    p2 = (g[0, 0] - q) ** 2 + (g[1, 1] - q) ** 2 + (g[2, 2] - q) ** 2 + 2.0 * p1
    p = math.sqrt(p2 / 6.0)

    def b(i, j):
        return (g[i, j] - (q if i == j else 0.0)) / p

    det = (b(0, 0) * (b(1, 1) * b(2, 2) - b(1, 2) * b(2, 1))
           - b(0, 1) * (b(1, 0) * b(2, 2) - b(1, 2) * b(2, 0))
           + b(0, 2) * (b(1, 0) * b(2, 1) - b(1, 1) * b(2, 0)))
    r = min(max(det / 2.0, -1.0), 1.0)
    phi = math.acos(r) / 3.0
    e1 = q + 2.0 * p * math.cos(phi)
    e3 = q + 2.0 * p * math.cos(phi + 2.0 * math.pi / 3.0)
    return [e1, 3.0 * q - e1 - e3, e3]


def shape_descriptors(outer: np.ndarray, volume: int, centroid) -> tuple:
    """output.rs:147-187"""
    D = len(centroid)
    v = float(volume)
    g = np.zeros((D, D))
    for a in range(D):
        for b in range(D):
            g[a, b] = float(outer[a, b]) / v - centroid[a] * centroid[b]
    ev = sorted(symmetric_eigenvalues(g), reverse=True)
    s = 0.0
    for x in ev:
        s += x
    ssq = 0.0
    for x in ev:
        ssq += x * x
    if s > 0.0:
        d = float(D)
        kappa = (d / (d - 1.0)) * (ssq / (s * s)) - 1.0 / (d - 1.0)
    else:
        kappa = 0.0
    return ev, s, kappa


@dataclass
class Sample:
    mcs: int
    energy: float
    volume: np.ndarray
    interface: np.ndarray
    centroid: np.ndarray
    eigenvalues: np.ndarray
    rg2: np.ndarray
    kappa: np.ndarray
    contact: np.ndarray
    fragmented: list
    medium_volumes: list
    medium_touch: list
    lattice: np.ndarray | None = None


def take_sample(m: Model, include_lattice: bool = False) -> Sample:
    """output.rs:289-337"""
    vol = m.volume()
    sp = m.sum_position()
    so = m.sum_outer()
    D = m.cfg.dim
    n = m.n_cells
    cen = np.zeros((n, D))
    ev = np.zeros((n, D))
    rg = np.zeros(n)
    ka = np.zeros(n)
    for i in range(n):
        c = [float(sp[i, d]) / float(vol[i]) for d in range(D)]
        cen[i] = c
        e, s, k = shape_descriptors(so[i], int(vol[i]), c)
        ev[i], rg[i], ka[i] = e, s, k
    frag, mv, mt = m.label()
    return Sample(m.mcs, m.energy, vol, m.interface(), cen, ev, rg, ka, m.contact(),
                  frag, mv, mt, m.lattice() if include_lattice else None)


def not_equilibrated(trace: list) -> bool:
    """output.rs:388-420"""
    n = len(trace)
    if n < 3:
        return False
    xs = [float(i) for i in range(n)]
    xm = 0.0
    for x in xs:
        xm += x
    xm /= n
    ym = 0.0
    for y in trace:
        ym += y
    ym /= n
    num = den = 0.0
    for i in range(n):
        num += (xs[i] - xm) * (trace[i] - ym)
        den += (xs[i] - xm) ** 2
    if den == 0.0:
        return False
    slope = num / den
    icpt = ym - slope * xm
    rv = 0.0
    for i in range(n):
        rv += (trace[i] - (icpt + slope * xs[i])) ** 2
    rv /= n
    rs = math.sqrt(rv)
    if rs == 0.0:
        return abs(slope) * (n - 1) > 0.0
    return abs(slope) * (n - 1) > 2.0 * rs


@dataclass
class RunResult:
    samples: list = field(default_factory=list)
    status: tuple = ("Ok", None, None)
    energy_trace: list = field(default_factory=list)
    model_hash: str = ""
    convention_hash: str = ""
    derived_phi: float = 0.0


def run_with(m: Model, include_lattice: bool = False) -> RunResult:
    """output.rs:434-529"""
    cfg = m.cfg
    interval = cfg.sampling_interval_mcs
    burn = cfg.burn_in_mcs
    start = burn - int(float(burn) * 0.25)
    trace = []
    burned = 0
    while burned < burn:
        chunk = min(interval, burn - burned)
        m.run_mcs(chunk)
        burned += chunk
        if burned >= start:
            trace.append(m.energy)
    n_samples = max(cfg.readout_mcs // max(interval, 1), 1)
    samples = []
    for _ in range(n_samples):
        m.run_mcs(interval)
        samples.append(take_sample(m, include_lattice))
    status = status_of(samples, trace, m.n_cells)
    vol = m.volume()
    return RunResult(samples, status, trace, m.model_hash(), convention_hash(cfg),
                     float(int(vol.sum())) / float(m.n_sites))


def status_of(samples: list, trace: list, n_cells: int) -> tuple:
    """The status rules of run_with (output.rs:470-509, 531-559), in their
    order: NotEquilibrated, Fragmented (first sample with a fragmented
    cell), CellLost (a cell at volume <= 1 over the trailing half of the
    samples), Degenerate (the last sample's contact graph joins every
    cell); (kind, cell id, mcs)."""
    status = ("NotEquilibrated", None, None) if not_equilibrated(trace) else ("Ok", None, None)
    if status[0] == "Ok":
        for s in samples:
            if s.fragmented:
                status = ("Fragmented", s.fragmented[0], s.mcs)
                break
    if status[0] == "Ok" and samples:
        window = int(math.ceil(len(samples) * 0.5))
        window = min(max(window, 1), len(samples))
        trailing = samples[len(samples) - window:]
        for i in range(n_cells):
            if all(int(s.volume[i]) <= 1 for s in trailing):
                status = ("CellLost", i + 1, trailing[0].mcs)
                break
    if status[0] == "Ok" and samples and _fully_connected(samples[-1].contact, n_cells):
        status = ("Degenerate", None, None)
    return status


def _fully_connected(contact: np.ndarray, n_cells: int) -> bool:
    """output.rs:531-559"""
    if n_cells < 2:
        return False
    adj = {}
    for a, b in zip(*np.nonzero(contact)):
        if a == 0 or b == 0:
            continue
        adj.setdefault(int(a), []).append(int(b))
        adj.setdefault(int(b), []).append(int(a))
    if not adj:
        return False
    start = next(iter(adj))
    seen = {start}
    stack = [start]
    while stack:
        x = stack.pop()
        for y in adj.get(x, []):
            if y not in seen:
                seen.add(y)
                stack.append(y)
    return len(seen) == n_cells
