"""cupymag_amd -- the GSPM LLG time step of CuPyMag on AMD MI355X.

Host side (this package): YAML schema, mesh, native host assembly, and the
reference's Python call signatures (solve_cg, SimulationOperators,
ComputeDerivatives, VolumeAverage, Micromagnetics.main, the ``cupymag`` CLI).
Device side: libcupymag_hip.so (hand-written HIP for gfx950) behind the C
ABI in include/cupymag_hip.h, loaded with ctypes.
"""
__version__ = "0.9.1"
