"""diffvg_amd -- MI355X-native implementation of diffvg's render path.

Import it where pydiffvg was imported (``import diffvg_amd as pydiffvg``): the data
classes, ``RenderFunction`` and the device helpers keep pydiffvg's names and
signatures.  Rendering runs hand-written HIP kernels for gfx950 through the C-ABI
library ``libdiffvg_amd.so``; there is no CPU fallback.
"""
from .device import *  # noqa: F401,F403
from .shape import *  # noqa: F401,F403
from .color import *  # noqa: F401,F403
from .pixel_filter import *  # noqa: F401,F403
from . import diffvg  # noqa: F401
from .render_pytorch import (RenderFunction, OutputType, set_print_timing, set_render_mode,  # noqa: F401
                             set_direct_leaf_gradients,
                             get_render_mode, synchronize)
from .batch import Batch  # noqa: F401
from .parse_svg import svg_to_scene  # noqa: F401
from .save_svg import save_svg  # noqa: F401
from .optimize_svg import SvgOptimizationSettings, OptimizableSvg  # noqa: F401
