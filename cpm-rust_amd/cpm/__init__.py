"""Cellular Potts Model simulator on AMD Instinct MI355X.

The same Python surface as the reference package (python/cpm/__init__.py):
`CPM` is the builder API, `cells_for_density` the density helper.  The
Monte Carlo engine behind it is a HIP library for gfx950
(`libcpm_engine.so`, C ABI in include/cpm_engine.h).  `Replica` exposes the
engine at the granularity of cpm-core's `CPM<D>` / `run_mcs`, and
`run_batch_configs` batches fully independent configurations, and `Server`
is a serving loop (requests admitted by HBM footprint, continuous batching).
"""

from .core import Replica, derive_seed_u64, run_batch_configs
from .model import CPM
from .parameters import cells_for_density
from .serve import Request, Server
from .warnings import CPMInitializationWarning, CPMLowResolutionWarning

__all__ = [
    "CPM",
    "Replica",
    "Request",
    "Server",
    "cells_for_density",
    "derive_seed_u64",
    "run_batch_configs",
    "CPMInitializationWarning",
    "CPMLowResolutionWarning",
]
