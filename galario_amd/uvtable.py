"""uv-table ingestion: the on-disk step before the model-to-visibility path.

galario's quickstart reads an ASCII uv table of five whitespace-separated
columns u [m], v [m], Re [Jy], Im [Jy], w [Jy^-2] ('#' comment lines) and
converts u, v to units of the observing wavelength
(docs/quickstart.rst:56-63):

    u, v, Re, Im, w = np.require(np.loadtxt("uvtable.txt", unpack=True), requirements='C')
    u /= wle; v /= wle

``load_uvtable`` returns the same five C-contiguous float64 arrays (pandas'
C parser when it is importable, numpy otherwise; both parse the decimal text
to the same doubles).  ``dataset_from_uvtable`` uploads them once, in
locality order, as a device-resident ``torch_api.Dataset``.
"""
from __future__ import annotations

import numpy as np


def load_uvtable(path, wavelength: float = 1.0, dtype=np.float64):
    """(u, v, Re, Im, w): u, v divided by ``wavelength`` [m] (1.0 keeps meters)."""
    if not wavelength > 0:
        raise ValueError(f'wavelength must be positive, got {wavelength}')
    data = None
    try:
        import pandas as pd
        frame = pd.read_csv(path, sep=r'\s+', comment='#', header=None, dtype=np.float64, engine='c')
        data = frame.to_numpy(dtype=np.float64).T
    except ImportError:
        data = np.loadtxt(path, unpack=True, dtype=np.float64)
    if data.ndim != 2 or data.shape[0] != 5:
        raise ValueError(f'{path}: expected 5 columns (u, v, Re, Im, w), got shape {data.shape}')
    u, v, re, im, w = (np.ascontiguousarray(c, dtype=dtype) for c in data)
    u /= dtype(wavelength)
    v /= dtype(wavelength)
    return u, v, re, im, w


def dataset_from_uvtable(path, wavelength: float = 1.0, device='cuda', dtype=None, sort: bool = True):
    """Load a uv table and upload it as a ``torch_api.Dataset`` (float64 unless ``dtype`` says otherwise)."""
    import torch

    from .torch_api import Dataset
    u, v, re, im, w = load_uvtable(path, wavelength)
    return Dataset(u, v, re, im, w, device=device, dtype=dtype or torch.float64, sort=sort)
