"""snisr — MI355X-native tiled sparse-voxel ray-marching renderer (HIP kernels for gfx950
behind the C ABI of include/snisr.h)."""
from . import scene  # noqa: F401  (pure numpy, importable without a GPU)

__all__ = ["scene"]
