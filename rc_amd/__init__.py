"""rc_amd -- MI355X (gfx950) w-projection gridder/degridder with the C ABI of
the SKA-SDP RC MS4 kernels.  The compute runs in rc_amd/librcgrid.so (HIP);
this package holds the host-side data formats, the ctypes binding and the
PyTorch operator."""
from .tables import GCFTable, VisChunks, OVER  # noqa: F401

__all__ = ["GCFTable", "VisChunks", "OVER"]
