"""vszipcu — MI355X BM3D engine, host-side mirror of the vszipcu plugin's BM3D path.

    from vszipcu import core, Clip, GRAYS
    out = core.vszipcu.BM3Dv2(clip, sigma=[5, 3, 3], chroma=True)

See core.py (filter layer) and device.py (kernel layer on device-resident tensors).
"""
from . import synth  # noqa: F401
from ._lib import LIB_PATH, load  # noqa: F401
from .core import (GRAY, GRAY16, GRAYH, GRAYS, RGB, RGBS, YUV, YUV420PS, YUV444PS, Clip, Error, Frame,  # noqa: F401
                   VideoFormat, core)

__all__ = ["core", "Clip", "Frame", "Error", "VideoFormat", "GRAYS", "YUV444PS", "YUV420PS", "RGBS", "synth"]
