"""MI355X-native per-step cloth solve behind AndoSim's ando_barrier_core API.

Native parts (built in-tree by andosim_amd.build):
  libandosim_hip.so        gfx950 kernels and the C ABI of include/andosim_hip.h
  ando_barrier_core*.so    Python module mirroring the reference's pybind11 binding
"""
import os

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
__all__ = ["PACKAGE_DIR"]
