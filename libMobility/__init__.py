"""``libMobility`` import surface over this library.

``from libMobility import NBody, SelfMobility, PSE, DPStokes`` works as with
the reference package (solvers/__init__.py:1-6), so code and test files that
import all four classes load.  ``NBody`` and ``SelfMobility`` run on the GPU
through ``libmobility_amd``.  PSE and DPStokes (triply and doubly periodic
spectral solvers) are not part of this library: constructing them raises
``RuntimeError``.
"""
from libmobility_amd import NBody
from libmobility_amd.self_mobility import SelfMobility


class PSE:
    """Not part of this library (the spectral Ewald solver)."""

    def __init__(self, *args, **kwargs):
        raise RuntimeError("[libmobility_amd] PSE is not part of this library; "
                           "NBody and SelfMobility are")


class DPStokes:
    """Not part of this library (the doubly periodic Stokes solver)."""

    def __init__(self, *args, **kwargs):
        raise RuntimeError("[libmobility_amd] DPStokes is not part of this library; "
                           "NBody and SelfMobility are")


__all__ = ["PSE", "SelfMobility", "NBody", "DPStokes"]
