"""galario: the import path of galario's Python API, computed on an AMD MI355X.

Same layout as galario's package (python/__init__.py.in:20-31): ``single``
and ``double`` submodules with the API of python/libcommon.pyx, ``HAVE_CUDA``,
``__version__`` and the constants ``arcsec, au, cgs_to_Jy, pc, deg``, so that
user code written for galario (``from galario.double import chi2Profile``,
docs/quickstart.rst:144-145) runs unchanged.

``galario.double`` and ``galario.single`` are the modules ``galario_amd.double``
and ``galario_amd.single`` themselves (one module object each, registered
under both names), backed by the HIP library; there is no CPU build and no
``*_cuda`` variant (``HAVE_CUDA`` is False).  Importing fails with an
ImportError when the HIP library has not been built.
"""
import sys as _sys

from galario_amd import HAVE_CUDA, HAVE_HIP, __version__  # noqa: F401
from galario_amd import double, single  # noqa: F401

_sys.modules[__name__ + '.double'] = double
_sys.modules[__name__ + '.single'] = single

from galario_amd.double import arcsec, au, cgs_to_Jy, pc, deg  # noqa: E402,F401
