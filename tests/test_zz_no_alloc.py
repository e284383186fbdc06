"""wenox_l_local_noalloc (device arrays, the library's fixed scratch array,
no allocation) against wenox_l_local (host arrays, the drop-in arena) through
the C ABI: bit-for-bit equal for the exact variant and for the fast variant,
on two calls in a row with the second grid smaller than the first (the
scratch array keeps the first call's contents), characteristic and
component-wise, HLLC/HLLE blend and HLLE, on a strided 8-channel view.  A
grid whose face arrays exceed the scratch array is rejected before anything
is launched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
G = 1.4


def _state(nx, ny, seed):
    rng = np.random.default_rng(seed)
    shape = (nx + 6, ny + 6)
    rho = rng.uniform(0.3, 2.0, shape)
    u = rng.uniform(-1.0, 1.0, shape)
    v = rng.uniform(-1.0, 1.0, shape)
    p = rng.uniform(0.2, 2.0, shape)
    q = np.zeros(shape + (8,))
    q[..., 0], q[..., 1], q[..., 2] = rho, rho * u, rho * v
    q[..., 3] = p / (G - 1) + 0.5 * rho * (u * u + v * v)
    q[..., 4], q[..., 5], q[..., 6], q[..., 7] = rho, u, v, p
    return q


class _Hip:
    """The HIP runtime the library itself links (libamdhip64.so.7, already
    loaded with it), for the test's own device buffers and stream."""

    def __init__(self):
        from wenox import _lib
        _lib.load()
        h = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        h.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipStreamSynchronize.argtypes = [vp]
        h.hipStreamDestroy.argtypes = [vp]
        self.h = h
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def download(self, p, shape):
        out = np.empty(shape)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


@pytest.fixture
def hip():
    h = _Hip()
    yield h
    h.close()


def _view(ptr, a):
    from wenox import _lib
    return _lib.View3(ptr.value, (ctypes.c_int64 * 3)(*a.shape),
                      (ctypes.c_int64 * 3)(*[s // 8 for s in a.strides]))


@pytest.mark.parametrize("fast", [False, True])
def test_noalloc_matches_l_local(hip, fast):
    import weno_ext
    from wenox import _lib
    L = _lib.load()
    stream = ctypes.c_void_p()
    assert hip.h.hipStreamCreate(ctypes.byref(stream)) == 0
    weno_ext.set_variant(fast)
    try:
        for (nx, ny), seed in (((61, 47), 3), ((23, 18), 4)):   # the second call is smaller
            q = _state(nx, ny, seed)
            dq = hip.upload(q)
            out = hip.alloc(nx * ny * 4 * 8)
            for cons in (False, True):
                for force_hlle in (False, True):
                    print(f"[noalloc] {nx}x{ny} fast={fast} cons={cons} hlle={force_hlle}", flush=True)
                    ref = weno_ext.l_local_py(q, 1 / nx, G, force_hlle, (0.5, 2.0), cons)
                    assert hip.h.hipMemset(out, 0xFF, nx * ny * 4 * 8) == 0   # NaN everywhere
                    v = _view(dq, q)
                    rc = L.wenox_l_local_noalloc(ctypes.byref(v), 1 / nx, G, int(force_hlle), 0.5, 2.0,
                                                 int(cons), out, stream)
                    assert rc == 0, _lib.last_error()
                    assert hip.h.hipStreamSynchronize(stream) == 0
                    np.testing.assert_array_equal(hip.download(out, (nx, ny, 4)), ref)
    finally:
        weno_ext.set_variant(False)
        hip.h.hipStreamDestroy(stream)


def test_noalloc_rejects_oversized_grid(hip):
    from wenox import _lib
    L = _lib.load()
    m = n = 606
    q = np.zeros((m, n, 4))
    dq = hip.upload(q)
    out = hip.upload(np.zeros((m - 6, n - 6, 4)))
    v = _view(dq, q)
    rc = L.wenox_l_local_noalloc(ctypes.byref(v), 1e-3, G, 0, 0.5, 2.0, 0, out, None)
    assert rc == -2
    assert "MiB of scratch" in _lib.last_error()
    assert float(np.abs(hip.download(out, (m - 6, n - 6, 4))).max()) == 0.0
