"""MI355X-native GIMIC current-density hot path.

Host side in Python over the C-ABI HIP library libgimic_hip.so:
  lib        ctypes binding (Context: batched tensors, plane flux)
  gimic      the pygimic-facing Gimic class on the drop-in gimic_* entry points
  grid       reference grid builders (std/base, bond, file; even/gauss/lobatto)
  flux       plane-flux integration with the reference's stdout block
  synth      seeded synthetic workloads
"""
__all__ = ["lib", "gimic", "grid", "flux", "synth", "build"]
