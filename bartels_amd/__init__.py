"""bartels_amd: MI355X-native linear-tetmesh hyperelastic hot path (energy / gradient / Hessian,
PSD projection, deterministic CSR assembly, device Newton) behind a C ABI
(``include/bartels_hip.h``), with a host-side mirror of the reference's ``sim::linear_tetmesh_*``
interface.  The HIP library must be built (``python -m bartels_amd.build``); nothing falls back
to the CPU."""
from . import meshgen  # noqa: F401  (pure numpy, no device)


def __getattr__(name):
    # the ctypes-backed modules load the HIP library on first use
    if name in ("tetmesh", "newton", "_lib"):
        import importlib

        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
