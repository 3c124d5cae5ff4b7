"""Host-side checks of the device-resident Dataset's bookkeeping (no GPU):
its locality order and the uv extent its calls pass to the library."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')


def _uv(n=5000, seed=0):
    rng = np.random.default_rng(seed)
    b = np.exp(rng.uniform(np.log(1e3), np.log(1e6), n))
    th = rng.uniform(0, 2 * np.pi, n)
    return b * np.cos(th), b * np.sin(th)


def test_locality_order_is_a_permutation_grouping_neighbours():
    from galario_amd.torch_api import locality_order
    u, v = _uv()
    order = locality_order(u, v)
    assert np.array_equal(np.sort(order), np.arange(len(u)))
    # consecutive points of the order are much closer (Hermitian-folded) than random pairs
    fu, fv = np.abs(u), np.copysign(1.0, u) * v
    step = np.hypot(np.diff(fu[order]), np.diff(fv[order]))
    rand = np.hypot(np.diff(fu), np.diff(fv))
    assert np.median(step) < 0.05 * np.median(rand)


def test_locality_order_is_the_stable_sort_of_the_key():
    """The two 16-bit radix passes give the stable sort of the 32-bit Morton key,
    ties (repeated points) included."""
    from galario_amd.torch_api import locality_key, locality_order
    u, v = _uv(20000, seed=1)
    u[::7], v[::7] = u[3], v[3]
    u[1::11] = -u[1::11]
    key = locality_key(u, v)
    assert key.dtype == np.uint32 and int(key.max()) >= 1 << 16
    assert np.array_equal(locality_order(u, v), np.argsort(key.astype(np.uint64), kind='stable'))
    # the key interleaves the bits, |u| on the even positions and the folded v on the odd ones:
    # (u, v) = (1, 0) quantises to (65535, 32767), whose spread bits are 0x55555555 and 0x15555555
    k0 = int(locality_key(np.array([1.0]), np.array([0.0]))[0])
    assert k0 == 0x55555555 | (0x15555555 << 1)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_dataset_extent_and_order_on_the_host(dtype):
    from galario_amd.torch_api import Dataset
    u, v = _uv()
    re, im, w = np.ones_like(u), np.zeros_like(u), np.ones_like(u)
    ds = Dataset(u, v, re, im, w, device='cpu', dtype=dtype)
    order = ds.order.numpy()
    assert np.array_equal(np.sort(order), np.arange(len(u)))
    # stored values are the caller's, in the locality order
    npdt = np.float64 if dtype == torch.float64 else np.float32
    assert np.array_equal(ds.u.numpy(), u[order].astype(npdt))
    assert np.array_equal(ds.v.numpy(), v[order].astype(npdt))
    # the extent is the largest radius of the stored (rounded) values
    us, vs = ds.u.numpy().astype(np.float64), ds.v.numpy().astype(np.float64)
    assert ds.uvmax == pytest.approx(float(np.sqrt(np.max(us * us + vs * vs))), rel=1e-15)
    assert np.all(np.hypot(us, vs) <= ds.uvmax * (1 + 1e-15))


def test_empty_dataset_has_no_extent():
    from galario_amd.torch_api import Dataset
    ds = Dataset(np.zeros(0), np.zeros(0), device='cpu')
    assert ds.n == 0 and ds.uvmax == 0.0
