"""libmobility_amd: libMobility's NBody solver (with its Lanczos noise) and
SelfMobility on MI355X (gfx950).

    from libmobility_amd import NBody
    solver = NBody("open", "open", "open")
    solver.setParameters()
    solver.initialize(viscosity=1.0, hydrodynamicRadius=1.0)
    solver.setPositions(positions)
    U, _ = solver.Mdot(forces)
"""
from .nbody import NBody
from .self_mobility import SelfMobility

__all__ = ["NBody", "SelfMobility"]
__version__ = "0.1.0"
