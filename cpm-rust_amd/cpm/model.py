"""The user-facing builder API, backed by the MI355X engine.

Mirrors the reference's `cpm.CPM` (python/cpm/model.py:22-182) together with
the builder semantics of its compiled extension
(crates/cpm-py/src/builder.rs:105-525): the same constructor keywords and
defaults, the same `ValueError` messages, the same warning policy, and the
same `RunResult` shape.  `run` and `run_batch` execute on the GPU: the
Monte Carlo steps run in exact-replay mode by default, which reproduces the
reference's serial trajectory bit for bit.
"""

from __future__ import annotations

import copy
import ctypes as C
import warnings
from typing import Any

import numpy as np

from . import _native as N
from .warnings import CPMInitializationWarning

_INIT_WARNING = (
    "default (scatter-and-grow) initial placement used; call "
    "initialize(lattice=...) for an explicit placement, or "
    "initialize(warn_on_default_init=False) to silence this"
)
_BATCH_INIT_WARNING = (
    "default (scatter-and-grow) initial placement used in at "
    "least one batch member; call initialize(lattice=...) for "
    "an explicit placement, or "
    "initialize(warn_on_default_init=False) to silence this"
)


def _parse_boundary(s: str) -> int:
    """builder.rs:35-43"""
    if s in N.BOUNDARY:
        return N.BOUNDARY[s]
    raise ValueError(f'unknown boundary {_rust_debug_str(s)}; expected "fixed" or "periodic"')


def _parse_neighborhood(s: str) -> int:
    """builder.rs:45-57"""
    if s in N.NB:
        return N.NB[s]
    raise ValueError(
        f"unknown neighbourhood {_rust_debug_str(s)}; expected one of \"von_neumann\", "
        "\"moore\", \"eighteen\", \"twenty_six\", \"twenty_six_weighted\"")


def _parse_acceptance(s: str) -> int:
    """builder.rs:59-67"""
    if s in N.ACCEPT:
        return N.ACCEPT[s]
    raise ValueError(f"unknown acceptance mode {_rust_debug_str(s)}; expected \"metropolis\" "
                     "or \"metropolis_hastings\"")


def _parse_proposal(s: str) -> int:
    """builder.rs:69-77"""
    if s in N.PROPOSAL:
        return N.PROPOSAL[s]
    raise ValueError(f"unknown proposal mode {_rust_debug_str(s)}; expected \"uniform\" or "
                     "\"edge_list\"")


def _parse_active_terms(d: dict) -> bool:
    """builder.rs:79-96: only `act` changes behaviour."""
    act = False
    for key, value in d.items():
        if key not in ("volume", "interface", "adhesion", "act"):
            raise ValueError(f"unknown active_terms key {_rust_debug_str(key)}; expected one of "
                             "\"volume\", \"interface\", \"adhesion\", \"act\"")
        if key == "act":
            act = bool(value)
    return act


def _rust_debug_str(s: str) -> str:
    out = ['"']
    for ch in s:
        if ch == '"':
            out.append('\\"')
        elif ch == "\\":
            out.append("\\\\")
        elif ch == "\n":
            out.append("\\n")
        elif ch == "\t":
            out.append("\\t")
        elif ch == "\r":
            out.append("\\r")
        else:
            out.append(ch)
    out.append('"')
    return "".join(out)


def _rust_f64(v: float) -> str:
    if v != v:
        return "NaN"
    if v in (float("inf"), float("-inf")):
        return "inf" if v > 0 else "-inf"
    r = repr(float(v))
    if "e" in r or "E" in r:
        r = np.format_float_positional(v, trim="-")
    if r.endswith(".0"):
        r = r[:-2]
    return r


def _validate_cell_type(t: dict) -> None:
    """cell.rs:40-66"""
    if not t["name"]:
        raise ValueError("cell type name must not be empty")
    for label in ("lambda_volume", "lambda_interface", "lambda_act"):
        v = float(t[label])
        if not np.isfinite(v) or v < 0.0:
            raise ValueError(f"{label} must be finite and non-negative, got {_rust_f64(v)}")
    if t["lambda_act"] > 0.0 and t["max_act"] == 0:
        raise ValueError("max_act must be positive when lambda_act is positive")


class CPM:
    """One CPM trajectory. Configure with the constructor and the
    ``register_cell_type``/``add_cells``/``set_adhesion``/``initialize``
    builder calls, then call :meth:`run`.  Nothing executes until
    :meth:`run` (or :meth:`run_batch`) is called.
    """

    def __init__(
        self,
        grid: tuple[int, ...],
        boundary: str,
        seed: int,
        copy_neighborhood: str = "von_neumann",
        energy_neighborhood: str | None = None,
        connectivity_neighborhood: str = "von_neumann",
        acceptance: str | None = None,
        proposal: str | None = None,
        active_terms: dict[str, bool] | None = None,
    ) -> None:
        b = _parse_boundary(boundary)
        copy_nb = _parse_neighborhood(copy_neighborhood)
        energy_nb = N.NB_DEFAULT if energy_neighborhood is None else _parse_neighborhood(
            energy_neighborhood)
        conn_nb = _parse_neighborhood(connectivity_neighborhood)
        acc = N.ACCEPT["metropolis"] if acceptance is None else _parse_acceptance(acceptance)
        prop = N.PROPOSAL["uniform"] if proposal is None else _parse_proposal(proposal)
        act = _parse_active_terms(active_terms) if active_terms is not None else False
        grid = [int(g) for g in grid]
        if len(grid) not in (2, 3):
            raise ValueError(
                "grid must have length 2 or 3 (dimensionality is inferred from it, "
                f"not from an explicit `dimension` argument); got length {len(grid)}")
        if any(g < 0 for g in grid):
            raise OverflowError("grid dimensions must be non-negative")
        self._spec = {
            "grid": grid, "boundary": b, "seed": int(seed), "copy_neighborhood": copy_nb,
            "energy_neighborhood": energy_nb, "connectivity_neighborhood": conn_nb,
            "acceptance": acc, "proposal": prop, "act": act, "cell_types": [],
            "cell_counts": [], "adhesion": None, "init": None,
        }
        self._warn_on_default_init = True

    # ------------------------------------------------------------ builder
    def register_cell_type(
        self,
        name: str,
        target_volume: int,
        target_interface: int,
        lambda_volume: float,
        lambda_interface: float,
        lambda_act: float = 0.0,
        max_act: int = 0,
    ) -> None:
        """builder.rs:197-238"""
        if any(t["name"] == name for t in self._spec["cell_types"]):
            raise ValueError(f"cell type {_rust_debug_str(name)} has already been added")
        t = {"name": str(name), "target_volume": int(target_volume),
             "target_interface": int(target_interface), "lambda_volume": float(lambda_volume),
             "lambda_interface": float(lambda_interface), "lambda_act": float(lambda_act),
             "max_act": int(max_act)}
        for k in ("target_volume", "target_interface", "max_act"):
            if not 0 <= t[k] <= 0xFFFFFFFF:
                raise OverflowError(f"{k} does not fit in u32")
        _validate_cell_type(t)
        self._spec["cell_types"].append(t)
        self._spec["cell_counts"].append(0)

    def add_cells(self, cell_type: str, n: int) -> None:
        """builder.rs:240-252"""
        for i, t in enumerate(self._spec["cell_types"]):
            if t["name"] == cell_type:
                self._spec["cell_counts"][i] = int(n)
                return
        raise ValueError(
            f"unknown cell type {_rust_debug_str(cell_type)}; call register_cell_type first")

    def set_adhesion(self, matrix: Any) -> None:
        """`matrix` is `(n_types + 1) x (n_types + 1)` (index 0 = medium),
        symmetric. Accepts a nested list or anything iterable row by row."""
        self._spec["adhesion"] = [[float(v) for v in row] for row in matrix]

    def initialize(
        self,
        lattice: Any | None = None,
        cell_type_of: Any | None = None,
        warn_on_default_init: bool = True,
    ) -> None:
        """Configures the initial placement; does not execute it
        (builder.rs:254-289)."""
        self._warn_on_default_init = warn_on_default_init
        if lattice is not None:
            if cell_type_of is None:
                raise ValueError("cell_type_of is required when lattice is supplied")
            self._spec["init"] = {"lattice": [int(v) for v in np.asarray(lattice).ravel()],
                                  "cell_type_of": [int(v) for v in cell_type_of],
                                  "warn": warn_on_default_init}
        else:
            self._spec["init"] = {"lattice": None, "cell_type_of": None,
                                  "warn": warn_on_default_init}

    # ------------------------------------------------------------ running
    def _windowed(self, burn_in_mcs, readout_mcs, sampling_interval_mcs) -> dict:
        spec = copy.deepcopy(self._spec)
        spec["burn_in_mcs"] = int(burn_in_mcs)
        spec["readout_mcs"] = int(readout_mcs)
        spec["sampling_interval_mcs"] = int(sampling_interval_mcs)
        return spec

    def run(
        self,
        burn_in_mcs: int,
        readout_mcs: int,
        sampling_interval_mcs: int,
        include_lattice: bool = False,
        mode: str = "exact_replay",
    ):
        """Runs the entire trajectory in one call and returns a `RunResult`
        (`.metadata`, `.status`, `.samples`).  `mode="sublattice"` selects
        the checkerboard-parallel update order instead of exact replay."""
        spec = self._windowed(burn_in_mcs, readout_mcs, sampling_interval_mcs)
        L = N.lib()
        holder = N.ConfigHolder(spec)
        model = C.c_void_p()
        N.check(L.cpm_model_create(holder.ref(), C.byref(model)))
        try:
            used = C.c_int32(0)
            N.check(L.cpm_model_initialize(model, C.byref(used)))
            res = C.c_void_p()
            N.check(L.cpm_run(N.engine(), model, 1 if include_lattice else 0, N.MODE[mode],
                              C.byref(res)))
        finally:
            L.cpm_model_destroy(model)
        result = N.RunResult(res, spec["grid"], [t["name"] for t in spec["cell_types"]])
        if used.value and self._warn_on_default_init:
            warnings.warn(_INIT_WARNING, CPMInitializationWarning, stacklevel=2)
        return result

    def _theta_holder(self) -> N.ConfigHolder:
        """builder.rs:391-403: placeholder windows, then full validation."""
        spec = copy.deepcopy(self._spec)
        spec.setdefault("burn_in_mcs", 1)
        spec.setdefault("readout_mcs", 1)
        spec.setdefault("sampling_interval_mcs", 1)
        holder = N.ConfigHolder(spec)
        probe = C.c_void_p()
        N.check(N.lib().cpm_model_create(holder.ref(), C.byref(probe)))
        N.lib().cpm_model_destroy(probe)
        return holder

    def theta_names(self) -> list[str]:
        """The canonical `theta` parameter names (theta.rs:151-178)."""
        holder = self._theta_holder()
        L = N.lib()
        n = L.cpm_theta_len(len(self._spec["cell_types"]))
        names = []
        buf = C.create_string_buffer(4096)
        for i in range(n):
            N.check(L.cpm_theta_name(holder.ref(), i, buf, len(buf)))
            names.append(buf.value.decode("utf-8"))
        return names

    def extract_theta(self) -> list[float]:
        """The current configuration's `theta` (theta.rs:55-78)."""
        holder = self._theta_holder()
        L = N.lib()
        n = L.cpm_theta_len(len(self._spec["cell_types"]))
        out = np.zeros(n, dtype=np.float64)
        N.check(L.cpm_theta_extract(holder.ref(), out.ctypes.data_as(N.f64p), n))
        return [float(v) for v in out]

    def run_batch(
        self,
        thetas: Any,
        burn_in_mcs: int,
        readout_mcs: int,
        sampling_interval_mcs: int,
        master_seed: int,
        include_lattice: bool = False,
        mode: str = "exact_replay",
    ) -> list[Any]:
        """Runs one independent trajectory per entry in `thetas`, all resident
        on the GPU together.  Each simulation's seed is
        `derive_seed_u64(master_seed, index)`, so results do not depend on
        how replicas are scheduled.  At most one `CPMInitializationWarning`
        fires for the whole call."""
        thetas_list = [[float(v) for v in theta] for theta in thetas]
        spec = self._windowed(burn_in_mcs, readout_mcs, sampling_interval_mcs)
        L = N.lib()
        expected = L.cpm_theta_len(len(spec["cell_types"]))
        for index, theta in enumerate(thetas_list):  # batch.rs:115-124
            if len(theta) != expected:
                raise ValueError(
                    f"theta at index {index} has length {len(theta)}, expected {expected}")
        holder = N.ConfigHolder(spec)
        n = len(thetas_list)
        flat = np.ascontiguousarray(np.array(thetas_list, dtype=np.float64).reshape(n, expected)
                                    if n else np.zeros(1))
        handles = (C.c_void_p * max(n, 1))()
        any_used = C.c_int32(0)
        N.check(L.cpm_run_batch(N.engine(), holder.ref(), flat.ctypes.data_as(N.f64p), n,
                                expected, int(master_seed), 1 if include_lattice else 0,
                                N.MODE[mode], handles, C.byref(any_used)))
        names = [t["name"] for t in spec["cell_types"]]
        results = [N.RunResult(C.c_void_p(handles[i]), spec["grid"], names) for i in range(n)]
        if any_used.value and self._warn_on_default_init:
            warnings.warn(_BATCH_INIT_WARNING, CPMInitializationWarning, stacklevel=2)
        return results
