"""physxcoopa_amd -- MI355X-native rigid-body contact step behind physxcoopa's PhysicsWorld API.

The compute path is libpxg.so (HIP kernels for gfx950 behind the C ABI in include/pxg.h);
`physxcoopa_amd.world.PhysicsWorld` is the Python mirror of the reference's PhysicsWorld over it.
"""
__version__ = "0.1.0"
