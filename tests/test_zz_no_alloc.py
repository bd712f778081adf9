"""pp_gram_schmidt_noalloc (scratch in a fixed device array of the library)
against pp_gram_schmidt (stream-ordered allocations) through the C-ABI: the
same kernels in the same order, so the results must be bitwise equal.  Two
calls in a row, the second with a smaller input, then the size limit."""
import ctypes

import numpy as np
import pytest

from pyprop_amd import _native as nat
from pyprop_amd import device as dev

pytestmark = pytest.mark.gpu


def _basis(rng, k, n):
    V0 = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
    Q, _ = np.linalg.qr(V0.T)
    return Q.T.copy()


def _run(fn, V0, w0, k, passes):
    n = w0.size
    V = dev.to_device(V0) if k > 0 else None
    w = dev.to_device(w0)
    h = np.zeros(max(2 * k, 2))
    nrm = ctypes.c_double()
    rc = fn(dev.ptr(V) if V is not None else None, n, k, dev.ptr(w), passes, nat.dptr(h), ctypes.byref(nrm),
            dev.stream_ptr())
    return rc, h[:2 * k].copy(), w.cpu().numpy(), nrm.value


@pytest.mark.parametrize("passes", [1, 2])
def test_noalloc_gram_schmidt_equals_allocating_entry_point(passes):
    L = nat.lib()
    rng = np.random.default_rng(20261204 + passes)
    # first call: k = 96, n = 150,001; second call: smaller in both k and n
    for k, n in [(96, 150001), (5, 3001), (0, 777)]:
        print("gram_schmidt noalloc vs alloc: k = %d, n = %d, passes = %d" % (k, n, passes), flush=True)
        V0 = _basis(rng, k, n) if k > 0 else None
        w0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        rc_a, h_a, w_a, n_a = _run(L.pp_gram_schmidt, V0, w0, k, passes)
        rc_b, h_b, w_b, n_b = _run(L.pp_gram_schmidt_noalloc, V0, w0, k, passes)
        assert rc_a == 0 and rc_b == 0, (rc_a, rc_b)
        assert np.array_equal(h_a, h_b)
        assert np.array_equal(w_a, w_b)
        assert n_a == n_b
        if k > 0:   # and both against numpy
            href = V0.conj() @ w0
            assert np.abs((h_b[0::2] + 1j * h_b[1::2]) - href).max() < 1e-12 * np.abs(href).max()
        wref = w0 - (V0.T @ (V0.conj() @ w0) if k > 0 else 0)
        assert abs(n_b - np.linalg.norm(wref)) < 1e-12 * np.linalg.norm(wref)


def test_noalloc_gram_schmidt_rejects_basis_beyond_its_scratch():
    print("gram_schmidt noalloc: k = 129 is refused", flush=True)
    n, k = 64, 129
    V = dev.to_device(np.zeros((k, n), dtype=np.complex128))
    w = dev.to_device(np.ones(n, dtype=np.complex128))
    h = np.zeros(2 * k)
    nrm = ctypes.c_double()
    rc = nat.lib().pp_gram_schmidt_noalloc(dev.ptr(V), n, k, dev.ptr(w), 1, nat.dptr(h), ctypes.byref(nrm),
                                           dev.stream_ptr())
    assert rc == nat.PP_EARG
    assert np.array_equal(w.cpu().numpy(), np.ones(n))
