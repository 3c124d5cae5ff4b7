"""galario.double on the GPU: double-precision galario API (python/libcommon.pyx
built with DOUBLE_PRECISION, python/galario_config.pxi.in:26-33)."""
import atexit as _atexit

from ._api import populate as _populate

_populate(globals(), double=True)

_init()
_atexit.register(_cleanup)
