"""MI355X-native grand-potential 2-D finite-difference step (gfx950 HIP kernels
behind a C ABI), a drop-in for the solver loop of the MicroSim serial
grand-potential solver."""
from .gpstep import Solver, GPError, lib, read_input, fill_domain, vtk_read, vtk_write, temp_bottom  # noqa: F401
