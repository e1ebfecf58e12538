"""MI355X-native per-timestep SPH pipeline (FluidGrid hot path).

The compute path is libfluidgrid_amd.so (hand-written HIP for gfx950 behind
the reference's C-ABI); `fluidgrid_amd.lib` loads it and fails loudly when it
is missing.  `abi` holds the struct layouts, `scenes` the seeded inputs.
"""
