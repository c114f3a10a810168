"""dynlattice_amd: MI355X-native explicit-dynamics path of a geometrically
exact rod lattice (Simo-Reissner rods, elastoplastic return map, block-lumped
mass, SO(3) central-difference update, rod-rod penalty contact).

The compute path is the HIP library ``libdlgpu.so`` behind the C ABI declared
in include/dl_gpu.h; :mod:`dynlattice_amd.gpu` binds it.  This package holds
the host-side mirror of the reference's model/module interface and the mesh
generators.
"""
from . import lattice  # noqa: F401
