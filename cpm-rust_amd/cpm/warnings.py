"""Warnings raised by the Python-facing API (python/cpm/warnings.py:12-36)."""


class CPMInitializationWarning(UserWarning):
    """Raised when a simulation falls back to default (scatter-and-grow)
    initial placement instead of an explicitly supplied one.

    Suppressible via ``warnings.simplefilter("ignore",
    CPMInitializationWarning)`` or ``CPM.initialize(warn_on_default_init=False)``.
    """


class CPMLowResolutionWarning(UserWarning):
    """Raised by :func:`cpm.parameters.cells_for_density` when the implied
    cell radius falls below about 5 lattice sites."""
