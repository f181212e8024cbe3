"""MI355X-native signal-generation path of FEelMRI.

Reference-shaped drop-ins (numpy in, numpy out):
``feelmri_amd.MRI.Signal``, ``feelmri_amd.BlochSimulator.solve_mri``,
``feelmri_amd.Assemble.MassAssemble``, ``feelmri_amd.POD.tensordot_modes_weights``.
Device-resident API: ``feelmri_amd.device``.  Kernels: libfeelmri_hip.so
(C ABI in include/feelmri_hip.h), built by ``python -m feelmri_amd.build``.
"""
__version__ = "0.1.0"

from . import _lib  # noqa: F401  (binding only; loads nothing until first call)
