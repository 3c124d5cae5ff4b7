"""CPU: uv-table ingestion (docs/quickstart.rst:56-63 reads it with np.loadtxt and divides u, v by the
wavelength).  The loader must give the same doubles as that recipe."""
import numpy as np
import pytest

from galario_amd.uvtable import load_uvtable


def _write(tmp_path, rows, header=True):
    p = tmp_path / 'uvtable.txt'
    with open(p, 'w') as f:
        if header:
            f.write('# Example uvtable\n# u [m]  v [m]  Re [Jy]  Im [Jy]  w [Jy^-2]\n')
        for r in rows:
            f.write('\t'.join(f'{x:.5f}' for x in r) + '\n')
    return p


def test_load_matches_quickstart_recipe(tmp_path):
    rng = np.random.default_rng(0)
    rows = np.column_stack([rng.uniform(-300, 300, (200, 2)), rng.normal(0, 0.1, (200, 2)), rng.uniform(100, 250, 200)])
    p = _write(tmp_path, rows)
    u, v, re, im, w = load_uvtable(p, wavelength=1e-3)
    ref = np.require(np.loadtxt(p, unpack=True), requirements='C')
    ref[0] /= 1e-3
    ref[1] /= 1e-3
    for a, b in zip((u, v, re, im, w), ref):
        assert a.flags['C_CONTIGUOUS'] and a.dtype == np.float64
        assert np.array_equal(a, b)


def test_load_rejects_bad_tables(tmp_path):
    p = tmp_path / 'bad.txt'
    p.write_text('1 2 3\n4 5 6\n')
    with pytest.raises(ValueError, match='5 columns'):
        load_uvtable(p)
    with pytest.raises(ValueError, match='wavelength'):
        load_uvtable(_write(tmp_path, [[1, 2, 3, 4, 5]]), wavelength=0.0)


def test_uvtable_head_fixture_round_trip(tmp_path):
    """The committed first rows of galario's docs/uvtable.txt (tests/golden/uvtable_head.npz), written back as
    text with 5 decimals as in the original file, load back to within the text rounding."""
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'uvtable_head.npz'))
    keys = list(d.keys())
    cols = [d[k] for k in ('u', 'v', 're', 'im', 'w')] if 'u' in keys else [d[keys[0]][:, i] for i in range(5)]
    rows = np.column_stack(cols)
    p = _write(tmp_path, rows)
    u, v, re, im, w = load_uvtable(p)
    np.testing.assert_allclose(np.column_stack([u, v, re, im, w]), rows, rtol=0, atol=6e-6)   # 5-decimal text
