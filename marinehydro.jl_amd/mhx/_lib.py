"""ctypes binding of libmhx.so (C ABI in include/mhx.h).

The library is built in-tree by `make -C marinehydro.jl_amd` (see
__graft_entry__.build()).  There is no fallback: if the shared library is
missing or cannot be loaded, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libmhx.so")

MHX_OK = 0
MHX_EINVAL = -1
MHX_EDIM = -2
MHX_ESINGULAR = -3
MHX_EDEVICE = -4
MHX_ENOMEM = -5
MHX_EQUAD = -6

GF_WU = 0
GF_DELHOMMEAU = 1
GF_SURROGATE = 2
MODE_PER_OMEGA = 0
MODE_PER_PROBLEM = 1
MODE_BATCHED = 2
MODE_GRAPH = 3  # MHX_MODE_GRAPH: the batched schedule captured into a HIP graph and replayed
SURR_SILU, SURR_RELU = 0, 1  # MHX_SURR_SILU / MHX_SURR_RELU
NOALLOC_SCRATCH_BYTES = 16 << 20  # MHX_NOALLOC_SCRATCH_BYTES
LU_AUTO = 0
LU_FP64 = 1
LU_MIXED = 2
LU_MIXED_BF16 = 3
PIVOT_VERIFIED = 0
PIVOT_ALWAYS = 1

EXPORTED = ["mhx_create", "mhx_destroy", "mhx_last_error", "mhx_set_constants", "mhx_assemble", "mhx_solve",
            "mhx_lu_solve", "mhx_lu_solve_batched", "mhx_lu_solve_noalloc", "mhx_hydro_coeffs", "mhx_body_create", "mhx_body_destroy", "mhx_body_coeffs",
            "mhx_phase_time", "mhx_reset_timing", "mhx_exact_work", "mhx_set_batch", "mhx_surrogate_train",
            "mhx_surrogate_save", "mhx_surrogate_load", "mhx_surrogate_params", "mhx_surrogate_eval",
            "mhx_surrogate_loss_grad", "mhx_surrogate_set_activation", "mhx_surrogate_get_activation", "mhx_hydro_coeffs_fs_jvp", "mhx_set_lu_mode", "mhx_solver_stats", "mhx_set_pivoting", "mhx_set_lu_streams", "mhx_set_autotune", "mhx_lu_config", "mhx_set_timing", "mhx_set_refine_limit",
            "mhx_pivot_stats", "mhx_hydro_coeffs_jvp",
            "mhx_hydro_coeffs_vjp", "mhx_sweep_coeffs", "mhx_body_create_lid", "mhx_body_coeffs_fs"]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)
_fp = C.POINTER(C.c_float)


class Options(C.Structure):
    _fields_ = [("rho", C.c_double), ("g", C.c_double), ("device", C.c_int), ("timing", C.c_int)]


class CMesh(C.Structure):
    _fields_ = [("nv", C.c_int64), ("nf", C.c_int64), ("vertices", _dp), ("faces", _ip), ("centers", _dp),
                ("normals", _dp), ("areas", _dp), ("radii", _dp)]


class CMeshTangent(C.Structure):
    _fields_ = [("vertices", _dp), ("centers", _dp), ("normals", _dp), ("areas", _dp)]


class CMeshGrad(C.Structure):
    _fields_ = [("vertices", _dp), ("centers", _dp), ("normals", _dp), ("areas", _dp), ("dofs", _dp), ("omegas", _dp)]


class CSweepBody(C.Structure):
    _fields_ = [("body", C.c_void_p), ("nw", C.c_int64), ("omegas", _dp), ("A_out", _dp), ("B_out", _dp),
                ("Fd_out", _dp), ("Fk_out", _dp)]


SWEEP_GRAPH = 1


class MHXError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mhx error {code}: {msg}")
        self.code = code


class DimensionMismatch(ValueError):
    """The reference's DimensionMismatch."""


class SingularException(ArithmeticError):
    """LAPACK-style exactly-zero pivot."""


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"libmhx.so not built at {LIB_PATH}; run `make -C {PKG_ROOT}`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.mhx_create.argtypes = [C.POINTER(vp), C.POINTER(Options)]
        L.mhx_destroy.argtypes = [vp]
        L.mhx_destroy.restype = None
        L.mhx_last_error.argtypes = [vp]
        L.mhx_last_error.restype = C.c_char_p
        L.mhx_set_constants.argtypes = [vp, C.c_double, C.c_double]
        L.mhx_assemble.argtypes = [vp, C.POINTER(CMesh), C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.mhx_solve.argtypes = [vp, C.c_int64, C.c_int64, _dp, _dp, _dp, C.c_int, _dp, _dp]
        L.mhx_lu_solve.argtypes = [vp, C.c_int64, C.c_int64, _dp, C.c_int, _dp]
        L.mhx_lu_solve_batched.argtypes = [vp, C.c_int64, C.c_int64, C.c_int64, _dp, C.c_int, _dp]
        L.mhx_lu_solve_noalloc.argtypes = [vp, C.c_int64, C.c_int64, C.c_void_p, C.c_int, C.c_void_p]
        L.mhx_hydro_coeffs.argtypes = [vp, C.POINTER(CMesh), C.c_int64, _dp, C.c_int64, _dp, C.c_int64, _dp,
                                       C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.mhx_hydro_coeffs_jvp.argtypes = [vp, C.POINTER(CMesh), C.POINTER(CMeshTangent), C.c_int64, _dp, _dp,
                                           C.c_int64, _dp, _dp, C.c_int64, _dp, C.c_int, C.c_int] + [_dp] * 8
        L.mhx_hydro_coeffs_vjp.argtypes = [vp, C.POINTER(CMesh), C.c_int64, _dp, C.c_int64, _dp, C.c_int64, _dp,
                                           C.c_int, C.c_int] + [_dp] * 8 + [C.POINTER(CMeshGrad)]
        L.mhx_sweep_coeffs.argtypes = [vp, C.POINTER(CSweepBody), C.c_int64, C.c_int64, _dp, C.c_int, C.c_int,
                                       C.c_int]
        L.mhx_body_create.argtypes = [vp, C.POINTER(CMesh), C.c_int64, _dp, C.POINTER(vp)]
        L.mhx_body_create_lid.argtypes = [vp, C.POINTER(CMesh), C.c_int64, C.c_int64, _dp, C.POINTER(vp)]
        L.mhx_body_destroy.argtypes = [vp]
        L.mhx_body_destroy.restype = None
        L.mhx_body_coeffs.argtypes = [vp, vp, C.c_int64, _dp, C.c_int64, _dp, C.c_int, C.c_int, C.c_int,
                                      _dp, _dp, _dp, _dp]
        L.mhx_body_coeffs_fs.argtypes = [vp, vp, C.c_int64, _dp, C.c_int64, _dp, C.c_int64, _dp, _dp, C.c_int, C.c_int,
                                         _dp, _dp, _dp, _dp]
        L.mhx_phase_time.argtypes = [vp, C.c_char_p, _dp, C.POINTER(C.c_int64)]
        L.mhx_reset_timing.argtypes = [vp]
        L.mhx_reset_timing.restype = None
        L.mhx_exact_work.argtypes = [vp, _ip, _dp]
        L.mhx_set_batch.argtypes = [vp, C.c_int64, C.c_double]
        L.mhx_set_lu_mode.argtypes = [vp, C.c_int]
        L.mhx_set_lu_streams.argtypes = [vp, C.c_int]
        L.mhx_set_autotune.argtypes = [vp, C.c_int]
        L.mhx_lu_config.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.mhx_set_timing.argtypes = [vp, C.c_int]
        L.mhx_set_refine_limit.argtypes = [vp, C.c_int]
        L.mhx_solver_stats.argtypes = [vp, _ip, _dp]
        L.mhx_set_pivoting.argtypes = [vp, C.c_int]
        L.mhx_pivot_stats.argtypes = [vp, _ip]
        L.mhx_surrogate_train.argtypes = [vp, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_uint64, _dp]
        L.mhx_surrogate_save.argtypes = [vp, C.c_char_p]
        L.mhx_surrogate_load.argtypes = [vp, C.c_char_p]
        L.mhx_surrogate_params.argtypes = [vp, _fp, _fp, _ip]
        L.mhx_surrogate_eval.argtypes = [vp, C.c_int64, _dp, _fp]
        L.mhx_surrogate_loss_grad.argtypes = [vp, C.c_int64, _fp, _fp, _fp, _fp, _fp, _fp]
        L.mhx_hydro_coeffs_fs_jvp.argtypes = [vp, C.POINTER(CMesh), C.c_int64, C.c_int64, _dp, C.c_int64, _dp, _dp,
                                               C.c_int64, _dp, C.c_int64, _dp, _dp, C.c_int, C.c_int] + [_dp] * 8
        L.mhx_surrogate_set_activation.argtypes = [vp, C.c_int]
        L.mhx_surrogate_get_activation.argtypes = [vp, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def fptr(a):
    return a.ctypes.data_as(_fp) if a is not None else None


def cmesh(mesh):
    """Borrow a Mesh's arrays as the C struct (the Mesh must outlive the call)."""
    return CMesh(mesh.vertices.shape[0], mesh.faces.shape[0], ptr(mesh.vertices), mesh.faces.ctypes.data_as(_ip),
                 ptr(mesh.centers), ptr(mesh.normals), ptr(mesh.areas), ptr(mesh.radii))


class Context:
    """One HIP stream plus timing state; wraps mhx_create / mhx_destroy."""

    def __init__(self, rho=1000.0, g=9.81, device=0, timing=False):
        self.h = C.c_void_p()
        opt = Options(rho, g, device, 1 if timing else 0)
        rc = lib().mhx_create(C.byref(self.h), C.byref(opt))
        if rc != MHX_OK:
            raise MHXError(rc, "mhx_create failed (no HIP device or bad options)")

    def check(self, rc):
        if rc == MHX_OK:
            return
        msg = lib().mhx_last_error(self.h).decode()
        if rc == MHX_EINVAL:
            raise ValueError(msg)
        if rc == MHX_EDIM:
            raise DimensionMismatch(msg)
        if rc == MHX_ESINGULAR:
            raise SingularException(msg)
        raise MHXError(rc, msg)

    def set_constants(self, rho, g):
        self.check(lib().mhx_set_constants(self.h, rho, g))

    def phase_time(self, phase):
        ms = C.c_double()
        cnt = C.c_int64()
        self.check(lib().mhx_phase_time(self.h, phase.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def exact_work(self):
        counts = np.zeros(3, dtype=np.int64)
        fl = C.c_double()
        self.check(lib().mhx_exact_work(self.h, counts.ctypes.data_as(_ip), C.byref(fl)))
        return counts, fl.value

    def set_batch(self, max_items, max_bytes=64.0 * 2 ** 30):
        self.check(lib().mhx_set_batch(self.h, int(max_items), float(max_bytes)))

    def set_timing(self, on):
        """Per-phase timing events on or off (from the next call on)."""
        self.check(lib().mhx_set_timing(self.h, int(bool(on))))

    def set_refine_limit(self, steps):
        """Refinement steps of the mixed-precision solver before an item is redone in FP64 (-1:
        the solver's default; 0 sends every system down the FP64 path)."""
        self.check(lib().mhx_set_refine_limit(self.h, int(steps)))

    def reset_timing(self):
        lib().mhx_reset_timing(self.h)

    def set_lu_mode(self, mode):
        """LU_AUTO (mixed precision for the surrogate wave term), LU_FP64, LU_MIXED or LU_MIXED_BF16
        (the FP32 factor's trailing updates on split-bf16 MFMA)."""
        self.check(lib().mhx_set_lu_mode(self.h, int(mode)))

    def set_lu_streams(self, streams):
        """Streams (1 .. 4) a batched FP32 factorization is split over; results do not depend on it."""
        self.check(lib().mhx_set_lu_streams(self.h, int(streams)))

    def set_autotune(self, on):
        """Autotuning of the batched FP32 factorization's streams and panels per split update."""
        self.check(lib().mhx_set_autotune(self.h, int(bool(on))))

    def lu_config(self):
        """(streams, panels per split update) of the last batched FP32 factorization."""
        a, b = C.c_int(), C.c_int()
        self.check(lib().mhx_lu_config(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def solver_stats(self):
        """(systems solved in mixed precision, refinement steps, FP64 fallbacks, max residual)."""
        c = np.zeros(3, dtype=np.int64)
        r = C.c_double()
        self.check(lib().mhx_solver_stats(self.h, c.ctypes.data_as(_ip), C.byref(r)))
        return int(c[0]), int(c[1]), int(c[2]), r.value

    def set_pivoting(self, mode):
        """PIVOT_VERIFIED (default: diagonal pivots with LAPACK's rule checked on every candidate, a
        matrix that breaks it factored again with row interchanges) or PIVOT_ALWAYS."""
        self.check(lib().mhx_set_pivoting(self.h, int(mode)))

    def pivot_stats(self):
        """(factorizations that kept the verified diagonal pivots, factorizations with the pivot
        search) since reset_timing."""
        c = np.zeros(2, dtype=np.int64)
        self.check(lib().mhx_pivot_stats(self.h, c.ctypes.data_as(_ip)))
        return int(c[0]), int(c[1])

    # ---- learned surrogate of the exact wave term ----
    def surrogate_train(self, ngrid=2048, steps=8000, batch=4096, lr=2e-3, seed=20261203, activation=None):
        """Generate ngrid^2 exact samples on the device and train the network (hidden-layer
        activation "silu" or "relu"; None keeps the context's); returns the validation MSE (scaled
        units) on the held-out samples."""
        if activation is not None:
            self.set_surrogate_activation(activation)
        vl = C.c_double()
        self.check(lib().mhx_surrogate_train(self.h, int(ngrid), int(steps), int(batch), float(lr), int(seed),
                                             C.byref(vl)))
        return vl.value

    def surrogate_save(self, path):
        self.check(lib().mhx_surrogate_save(self.h, os.fsencode(path)))

    def surrogate_load(self, path):
        self.check(lib().mhx_surrogate_load(self.h, os.fsencode(path)))

    def surrogate_param_count(self):
        n = C.c_int64()
        self.check(lib().mhx_surrogate_params(self.h, None, None, C.byref(n)))
        return n.value

    def surrogate_params(self):
        """(FP32 parameter vector, 32 x 2 Fourier matrix) of the trained network."""
        n = self.surrogate_param_count()
        p = np.zeros(n, dtype=np.float32)
        B = np.zeros((32, 2), dtype=np.float32)
        self.check(lib().mhx_surrogate_params(self.h, fptr(p), fptr(B), None))
        return p, B

    def surrogate_eval(self, r, v):
        """BF16-MFMA network at points (r, v): (m, 6) raw outputs (see surrogate_residual)."""
        rv = np.ascontiguousarray(np.stack([np.asarray(r, float).ravel(), np.asarray(v, float).ravel()], 1))
        out = np.zeros((rv.shape[0], 6), dtype=np.float32)
        self.check(lib().mhx_surrogate_eval(self.h, rv.shape[0], ptr(rv), fptr(out)))
        return out

    def surrogate_residual(self, r, v):
        """The network's correction to GFWu at (r, v) (v = k|z + zeta| >= 0): (m, 6) real array
        Re/Im of dG, dG/dr, dG/dz, i.e. the raw outputs / 100, the imaginary ones also times
        exp(-min(v, 5))."""
        out = self.surrogate_eval(r, v).astype(np.float64) / 100.0
        env = np.exp(-np.minimum(np.asarray(v, float).ravel(), 5.0))
        out[:, 1::2] *= env[:, None]
        return out

    def set_surrogate_activation(self, activation):
        """"silu" or "relu": the activation of the next training and of surrogate_loss_grad."""
        code = {"silu": SURR_SILU, "relu": SURR_RELU}[activation]
        self.check(lib().mhx_surrogate_set_activation(self.h, code))

    def surrogate_activation(self):
        a = C.c_int()
        self.check(lib().mhx_surrogate_get_activation(self.h, C.byref(a)))
        return {SURR_SILU: "silu", SURR_RELU: "relu"}[a.value]

    def surrogate_loss_grad(self, U, Y, params, fourier):
        """One FP32 forward/backward of the training kernels: (loss, gradient)."""
        U = np.ascontiguousarray(U, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        params = np.ascontiguousarray(params, dtype=np.float32)
        fourier = np.ascontiguousarray(fourier, dtype=np.float32)
        g = np.zeros_like(params)
        loss = C.c_float()
        self.check(lib().mhx_surrogate_loss_grad(self.h, U.shape[0], fptr(U), fptr(Y), fptr(params), fptr(fourier),
                                                 fptr(g), C.byref(loss)))
        return loss.value, g

    def close(self):
        if self.h:
            lib().mhx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def complex_out(shape):
    return np.zeros(shape, dtype=np.complex128, order="F")


def as_c128(a):
    return np.asfortranarray(a, dtype=np.complex128)
