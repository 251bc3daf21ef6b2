"""libmobility_amd: libMobility's NBody solver and Lanczos noise on MI355X (gfx950).

    from libmobility_amd import NBody
    solver = NBody("open", "open", "open")
    solver.setParameters()
    solver.initialize(viscosity=1.0, hydrodynamicRadius=1.0)
    solver.setPositions(positions)
    U, _ = solver.Mdot(forces)
"""
from .nbody import NBody

__all__ = ["NBody"]
__version__ = "0.1.0"
