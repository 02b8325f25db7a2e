"""MI355X-native explicit time step for PeriDEM models.

The engine (peridem_amd/lib/libpdem_hip.so, C-ABI in include/pdem.h) runs the
peridynamic bond forces, contact search and forces, boundary conditions and
node update on one GPU.  This package holds the host-side setup (model
building, material constants, synthetic decks) and the ctypes binding.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
