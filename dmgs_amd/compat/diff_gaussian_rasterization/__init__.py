"""Import shim: put ``dmgs_amd/compat`` on ``sys.path`` (or PYTHONPATH) and the
reference's ``from diff_gaussian_rasterization import
GaussianRasterizationSettings, GaussianRasterizer``
(gaussian_renderer/__init__.py:14) resolves to the gfx950 implementation."""
from dmgs_amd.rasterizer import (GaussianRasterizationSettings, GaussianRasterizer,  # noqa: F401
                                 _RasterizeGaussians, rasterize_gaussians)
