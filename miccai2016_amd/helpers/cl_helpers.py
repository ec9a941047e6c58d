"""Device plug-in seam (reference: helpers/cl_helpers.py).

``ClProgramLoader`` keeps the reference's name for the injected device object
so that code constructing one keeps working; it is the HIP DeviceContext.
"""
from ..device import DeviceContext


class ClProgramLoader(DeviceContext):
    """Binds the GPU (device 0 unless given) for the processing classes."""

    def __init__(self, device=0):
        super().__init__(device)
