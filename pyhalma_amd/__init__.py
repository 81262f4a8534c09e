"""pyhalma_amd -- MI355X-native halo-finding hot path (FoF grouping, per-group
reductions, O(N^2) pairwise-potential unbinding) behind pyHALMA's Python interface.

Drop-in modules (same names and signatures as the reference's):
  pyhalma_amd.fof              python_scripts/fof.py
  pyhalma_amd.halo_properties  python_scripts/halo_properties.py (reductions, energy test)
  pyhalma_amd.halo_gas         python_scripts/halo_gas.py (pair potential, RPS, most bound)
  pyhalma_amd.particle         fortran_modules/particle (f2py) pair-potential routines
Batched device-resident API: pyhalma_amd.engine.
"""
__all__ = ["engine", "fof", "halo_properties", "halo_gas", "particle", "synth"]
