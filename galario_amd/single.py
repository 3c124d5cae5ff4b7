"""galario.single on the GPU: single-precision galario API (python/libcommon.pyx
built without DOUBLE_PRECISION, python/galario_config.pxi.in:35-40)."""
import atexit as _atexit

from ._api import populate as _populate

_populate(globals(), double=False)

_init()
_atexit.register(_cleanup)
