"""astra.pythonutils (python/astra/pythonutils.py:28-83): geom_size and
GPULink, a raw device pointer for astra.data3d.link."""
from .functions import geom_size  # noqa: F401


class GPULink:
    """A device pointer to float32 data for astra.data3d.link
    (python/astra/pythonutils.py:67-83): x is the fastest-changing
    coordinate, z the slowest; pitch is the row width in bytes
    (4 * x for a contiguous block, or the pitch of a pitched allocation).
    The memory must stay allocated while the data object exists."""

    def __init__(self, ptr, x, y, z, pitch):
        self.ptr = ptr
        self.x = x
        self.y = y
        self.z = z
        self.pitch = pitch
