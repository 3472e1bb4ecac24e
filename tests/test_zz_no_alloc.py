"""mhx_lu_solve_noalloc (LU solve on caller-owned device buffers, scratch in a static device
array, no allocation) against mhx_lu_solve (the allocating entry point) through the C ABI: two
calls in a row, the second of smaller order, plus its argument checks."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    """The caller's device buffers, allocated with the HIP runtime (the caller owns them)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        self.L.hipDeviceSynchronize.argtypes = []

    def free_bytes(self):
        f, t = C.c_size_t(), C.c_size_t()
        assert self.L.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    def upload(self, X):
        """X (n x m complex) as a column-major device buffer."""
        h = np.asfortranarray(X, dtype=np.complex128)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), h.nbytes) == 0
        assert self.L.hipMemcpy(p, h.ctypes.data, h.nbytes, 1) == 0  # host to device
        return p, h.shape

    def download(self, buf):
        p, shape = buf
        h = np.zeros(shape, dtype=np.complex128, order="F")
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(h.ctypes.data, p, h.nbytes, 2) == 0  # device to host
        return h

    def free(self, *bufs):
        for p, _ in bufs:
            self.L.hipFree(p)


@pytest.fixture(scope="module")
def env():
    import gc

    import mhx
    from mhx import _lib
    gc.collect()
    print("device context", flush=True)
    ctx = mhx.context()
    hip = Hip()
    print(f"free device memory {hip.free_bytes() / 2 ** 30:.1f} GiB", flush=True)
    return hip, mhx, _lib, ctx


def _call(env, Ad, Bd, n, R, trans):
    hip, mhx, _lib, ctx = env
    ctx.check(_lib.lib().mhx_lu_solve_noalloc(ctx.h, n, R, Ad[0], trans, Bd[0]))


@pytest.mark.parametrize("trans", [0, 2])
def test_noalloc_matches_allocating_entry(env, trans):
    hip, mhx, _lib, ctx = env
    for n, R, seed in ((300, 7, 1), (130, 3, 2)):
        print(f"noalloc n={n} R={R} trans={trans}", flush=True)
        rng = np.random.default_rng(seed)
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        B = rng.standard_normal((n, R)) + 1j * rng.standard_normal((n, R))
        ref = mhx.lu_solve(A, B, trans=trans)
        Ad, Bd = hip.upload(A), hip.upload(B)
        _call(env, Ad, Bd, n, R, trans)
        X = hip.download(Bd)
        hip.free(Ad, Bd)
        assert np.linalg.norm(X - ref) <= 1e-12 * np.linalg.norm(ref)
        Aop = A if trans == 0 else A.conj().T
        assert np.linalg.norm(Aop @ X - B) <= 1e-11 * np.linalg.norm(B)


def test_noalloc_checks(env):
    hip, mhx, _lib, ctx = env
    print("noalloc argument checks", flush=True)
    one = hip.upload(np.zeros((1, 1)))
    # a workspace larger than the scratch (order 8000 with 16 right-hand sides: 20 MB) is refused
    # before any launch, as is an order above the factorization's limit
    with pytest.raises(mhx.DimensionMismatch):
        _call(env, one, one, 8000, 16, 0)
    with pytest.raises(mhx.DimensionMismatch):
        _call(env, one, one, 20000, 1, 0)
    hip.free(one)
    # host memory is refused
    h = np.zeros((4, 4), dtype=np.complex128)
    with pytest.raises(ValueError):
        ctx.check(_lib.lib().mhx_lu_solve_noalloc(ctx.h, 4, 1, C.c_void_p(h.ctypes.data), 0, C.c_void_p(h.ctypes.data)))
    # a singular matrix is reported, and the next call is unaffected (the scratch's info is cleared)
    Z, Bz = hip.upload(np.zeros((8, 8))), hip.upload(np.ones((8, 1)))
    with pytest.raises(mhx.SingularException):
        _call(env, Z, Bz, 8, 1, 0)
    hip.free(Z, Bz)
    n = 4096
    rng = np.random.default_rng(5)
    A = np.eye(n) * 0.5 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / n
    B = rng.standard_normal((n, 16)) + 1j * rng.standard_normal((n, 16))
    Ad, Bd = hip.upload(A), hip.upload(B)
    _call(env, Ad, Bd, n, 16, 0)
    X = hip.download(Bd)
    hip.free(Ad, Bd)
    assert np.linalg.norm(A @ X - B) <= 1e-12 * np.linalg.norm(B)
