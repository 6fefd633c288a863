"""Device selection, as gputools' R functions do it.

chooseGpu (R/chooseGpu.R:1-5) and getGpuId (R/getGpuId.R:1-6) call the .C
exports rsetDevice / rgetDevice (src/rinterface.cpp:79-85), whose bodies
(setDevice / getDevice, src/correlation.cpp:163-175) check the id against
the device count and make it the process's current device.
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import lib


def _raise_if_failed():
    err = _lib.last_error()
    if err:
        raise _lib.GputoolsError(_lib.GT_ERR_INVALID if "not valid" in err else _lib.GT_ERR_HIP, err)


def chooseGpu(deviceId=0) -> list:
    """Make ``deviceId`` the current GPU.  Like R's ``as.integer`` the id is
    truncated toward zero; an id outside 0 .. count-1 fails with "The gpu id
    number is not valid.".  Returns the argument list, as ``.C`` does."""
    dev = ctypes.c_int(int(deviceId))
    lib().rsetDevice(ctypes.byref(dev))
    _raise_if_failed()
    return [dev.value]


def getGpuId() -> int:
    """The current GPU's id (``.C("rgetDevice", deviceId = integer(1))$deviceId``)."""
    dev = ctypes.c_int(0)
    lib().rgetDevice(ctypes.byref(dev))
    _raise_if_failed()
    return dev.value
