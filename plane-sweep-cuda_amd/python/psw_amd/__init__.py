"""psw_amd -- Python binding of libpsw, the MI355X multiview-stereo library.

Importing the package loads plane-sweep-cuda_amd/lib/libpsw.so and fails if it
has not been built.  Scene generation (psw_amd.scene) is numpy only.
"""
from . import scene  # noqa: F401
from ._lib import LIB_PATH, PswError, lib  # noqa: F401
from .device import DeviceArray, Event, Stream, device_count, device_name, synchronize  # noqa: F401
from . import stages  # noqa: F401
