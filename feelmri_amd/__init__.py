"""MI355X-native signal-generation path of FEelMRI.

Reference-shaped drop-ins (numpy in, numpy out):
``feelmri_amd.MRI.Signal``, ``feelmri_amd.BlochSimulator.solve_mri``,
``feelmri_amd.Assemble.MassAssemble``, ``feelmri_amd.POD.tensordot_modes_weights``.
Device-resident API: ``feelmri_amd.device``.  Kernels: libfeelmri_hip.so
(C ABI in include/feelmri_hip.h), built by ``python -m feelmri_amd.build``.
"""
__version__ = "0.2.0"


def __getattr__(name):
    # the binding (and torch, whose HIP runtime it binds to) is imported on
    # first use, so host-only modules such as ``partition`` import without it
    if name == "_lib":
        import importlib

        return importlib.import_module(f"{__name__}._lib")
    raise AttributeError(f"module 'feelmri_amd' has no attribute {name!r}")
