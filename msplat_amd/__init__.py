"""msplat_amd: a gfx950 (MI355X) differentiable Gaussian-splatting rasterizer.

Same Python API as msplat (msplat/__init__.py:11-19): ``project_point``,
``compute_cov3d``, ``ewa_project``, ``compute_sh``, ``sort_gaussian``,
``alpha_blending`` and ``rasterization``.  Every op runs hand-written HIP
kernels from ``_lib/libmsplat_amd.so`` through a C ABI; there is no CPU or
PyTorch fallback.
"""

from .alpha_blending import alpha_blending
from .compute_cov3d import compute_cov3d
from .compute_sh import compute_sh, compute_sh_points
from .ewa_project import ewa_project
from .loss import l2_loss
from .project_point import project_point
from .rasterization import rasterization
from .rasterization_sh import rasterization_sh, rasterization_sh_l2
from .sort_gaussian import sort_gaussian

__all__ = [
    "project_point",
    "compute_cov3d",
    "ewa_project",
    "sort_gaussian",
    "compute_sh",
    "alpha_blending",
    "rasterization",
    # extensions with no reference counterpart: view direction + SH fused,
    # the training step's squared-error loss in one launch each way
    "compute_sh_points",
    "l2_loss",
    "rasterization_sh",
    "rasterization_sh_l2",
]

__version__ = "0.1.0"
