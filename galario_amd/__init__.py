"""galario_amd: galario's model-to-visibility path on AMD Instinct MI355X (gfx950).

Drop-in for the `galario` package (python/__init__.py.in:20-31):
``galario_amd.double`` and ``galario_amd.single`` expose galario's API
(sampleImage, chi2Image, sampleProfile, chi2Profile, get_image_size, sweep,
uv_rotate, interpolate, apply_phase_vis, reduce_chi2, _fft2d, _fftshift,
_fftshift_axis0, ...) computed by hand-written HIP kernels; ``import
galario_amd as galario`` keeps user code unchanged.
"""
__version__ = '1.2.2'

# galario's CUDA flag: there is no CUDA build here.  `double`/`single` run on the GPU.
HAVE_CUDA = False
HAVE_HIP = True

from . import single  # noqa: E402
from . import double  # noqa: E402

from .double import arcsec, au, cgs_to_Jy, pc, deg  # noqa: E402,F401
