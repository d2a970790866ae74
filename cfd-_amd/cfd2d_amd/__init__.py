"""cfd2d_amd: the cfd2d projection time step on AMD MI355X (gfx950).

Drop-in for the ``Simulation.step``/``run`` path of the reference package
``cfd2d`` (kaeptnzursee1910/CFD-): same Grid / BoundaryCondition / Simulation
/ case factories / case.json / checkpoint.npz interfaces, with the step
executed by hand-written HIP kernels through the C ABI of
``include/cfd2d_hip.h`` (``libcfd2d_hip.so``, built by ``cfd-_amd/build.py``).
"""

from .boundary import BoundaryCondition
from .grid import Grid
from .solver import Simulation, StopSimulation

__all__ = ["Grid", "BoundaryCondition", "Simulation", "StopSimulation"]
__version__ = "0.1.0"
