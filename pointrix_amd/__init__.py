"""pointrix_amd: MI355X-native (gfx950 HIP) differentiable Gaussian rasterisation.

Drop-in for the msplat path of pointrix: ``pointrix_amd.ops`` has the six
msplat functions, ``pointrix_amd.renderer.MsplatRender`` the renderer plugin.
"""
from . import ops, scene  # noqa: F401
from .renderer import (RENDERER_REGISTRY, MsplatNormalRender, MsplatRender,  # noqa: F401
                       RenderFeatures, parse_renderer)

__version__ = "0.1.0"
