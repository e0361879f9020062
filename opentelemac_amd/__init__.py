"""MI355X-native implementation of TELEMAC's BIEF linear-system hot path.

Modules:
  bief    -- host layer over the HIP C ABI (include/bief_hip.h): Context,
             MATVEC, SOLVE, DIRICH, LUMP
  mesh    -- synthetic meshes (channel, Poisson-disk estuary, prisms)
  system  -- host assembly of the telemac2d wave-equation system
  build   -- build recipes of the native parts
"""
__all__ = ["bief", "mesh", "system", "build"]
