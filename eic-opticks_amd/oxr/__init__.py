"""oxr: MI355X-native optical-photon propagation engine behind eic-opticks' simulate interface.

Host-side package: ctypes binding of the C ABI (liboxr.so), CSGFoundry-layout scene
builders, genstep builders and the build script. The compute path is the HIP library.
"""
from . import abi, csg, gensteps, scenes  # noqa: F401

__all__ = ['abi', 'csg', 'gensteps', 'scenes']
