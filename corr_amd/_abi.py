"""ctypes binding of libcorr_gpu.so (declared in include/corr_abi.h).

The library is the only compute path: if it is missing or cannot be loaded,
every entry point raises ``CorrLibraryError``.  torch must be imported before
the library is loaded so that both share one HIP runtime.
"""
from __future__ import annotations

import atexit
import ctypes as C
import threading
from pathlib import Path

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

LIB_PATH = Path(__file__).resolve().parent / "libcorr_gpu.so"

CORR_OK, CORR_EINVAL, CORR_ENOMEM, CORR_EHIP, CORR_ENOTCONV, CORR_ENOTPD, CORR_ENODEV = range(7)
PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2
METRIC_EIGEN, METRIC_CHOLESKY = 0, 1
CCSD_REFERENCE, CCSD_DIIS = 0, 1

PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}
# corr_dgemm alone also takes the int8-emulated FP64 GEMM (Ozaki splitting).
PREC_INT8X5 = 3
GEMM_PRECISIONS = {**PRECISIONS, "int8x5": PREC_INT8X5}
METRICS = {"eigen": METRIC_EIGEN, "cholesky": METRIC_CHOLESKY}

_STATUS = {
    CORR_EINVAL: "invalid argument",
    CORR_ENOMEM: "out of device memory",
    CORR_EHIP: "HIP runtime error",
    CORR_ENOTCONV: "not converged",
    CORR_ENOTPD: "metric not positive definite",
    CORR_ENODEV: "no gfx950 device",
}


class CorrLibraryError(RuntimeError):
    """libcorr_gpu.so is missing or unusable."""


class CorrError(RuntimeError):
    def __init__(self, fn: str, status: int, detail: str):
        self.status = status
        super().__init__(f"{fn} failed with status {status} ({_STATUS.get(status, '?')}): {detail}")


class MP2Energy(C.Structure):
    _fields_ = [("e_os", C.c_double), ("e_ss", C.c_double), ("e_mp2", C.c_double), ("e_scs", C.c_double)]

    def as_dict(self) -> dict:
        return {"EMP2_OS": self.e_os, "EMP2_SS": self.e_ss, "EMP2": self.e_mp2, "EMP2_SCS": self.e_scs}


class CCSDOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("maxiter", C.c_int), ("diis_size", C.c_int), ("prec", C.c_int),
                ("e_tol", C.c_double), ("r_tol", C.c_double)]


class CCSDResult(C.Structure):
    _fields_ = [("e_hf_check", C.c_double), ("e_one_check", C.c_double), ("e_fsum_check", C.c_double),
                ("e_mp2", C.c_double), ("e_ccsd", C.c_double),
                ("d_e", C.c_double), ("r_norm", C.c_double), ("iterations", C.c_int),
                ("converged", C.c_int)]


class SCFOpts(C.Structure):
    _fields_ = [("maxiter", C.c_int), ("diis_size", C.c_int), ("core_guess_check", C.c_int), ("reserved", C.c_int),
                ("diis_thresh", C.c_double), ("damp", C.c_double), ("e_tol", C.c_double)]


class SCFResult(C.Structure):
    _fields_ = [("e_scf", C.c_double), ("e_one", C.c_double), ("e_two", C.c_double), ("e_total", C.c_double),
                ("last_de", C.c_double), ("last_error", C.c_double), ("iterations", C.c_int),
                ("converged", C.c_int), ("diis", C.c_int), ("diis_start", C.c_int), ("diis_singular", C.c_int),
                ("core_guess", C.c_int)]


class LocOpts(C.Structure):
    _fields_ = [("max_sweeps", C.c_int), ("tol", C.c_double)]


class LocResult(C.Structure):
    _fields_ = [("spread_initial", C.c_double), ("spread", C.c_double), ("sweeps", C.c_int),
                ("converged", C.c_int)]


class CorrBasis(C.Structure):
    _fields_ = [("nshell", C.c_int64), ("l", C.POINTER(C.c_int32)), ("nprim", C.POINTER(C.c_int32)),
                ("center", C.POINTER(C.c_double)), ("exps", C.POINTER(C.c_double)),
                ("coefs", C.POINTER(C.c_double))]


P = C.c_void_p
I64 = C.c_int64
I = C.c_int
D = C.c_double

# name -> argtypes (all return int status except where noted)
SIGNATURES = {
    "corr_version": [],
    "corr_ctx_create": [I, C.POINTER(P)],
    "corr_ctx_destroy": [P],
    "corr_ctx_set_stream": [P, P],
    "corr_ctx_synchronize": [P],
    "corr_ctx_last_error": [P],
    "corr_ctx_trim": [P],
    "corr_ctx_set_scratch_limit": [P, I64],
    "corr_ctx_enable_timing": [P, I],
    "corr_ctx_kernel_time": [P, C.c_char_p, C.POINTER(D), C.POINTER(I64)],
    "corr_dmalloc": [P, C.POINTER(P), I64],
    "corr_dfree": [P, P],
    "corr_h2d": [P, P, P, I64],
    "corr_d2h": [P, P, P, I64],
    "corr_dgemm": [P, I, I, I64, I64, I64, D, P, I64, I64, P, I64, I64, D, P, I64, I64, I64, I],
    "corr_build_fmo": [P, I64, P, P, P],
    "corr_ao_unfold": [P, I64, I64, C.POINTER(I64), P, P, P],
    "corr_motrans": [P, I64, P, P, P, P, I],
    "corr_canonical_mp2": [P, I64, I64, P, P, C.POINTER(MP2Energy)],
    "corr_ri_metric": [P, I64, P, I, P, C.POINTER(I64)],
    "corr_ri_bpq": [P, I64, I64, P, P, P, I],
    "corr_transform_ri": [P, I64, I64, I64, P, P, P, I],
    "corr_ri_form_bia": [P, I64, I64, I64, P, P, P, I, P, I],
    "corr_ri_bia_from_j": [P, I64, I64, I64, P, P, P, I, P, P, I, P],
    "corr_ri_mp2_energy_from_j": [P, I64, I64, I64, P, P, P, I, P, P, I, C.POINTER(MP2Energy), P],
    "corr_rimp2_energy": [P, I64, I64, I64, P, P, P, I, C.POINTER(MP2Energy)],
    "corr_rimp2_energy_noalloc": [P, I64, I64, I64, P, P, P, I, C.POINTER(MP2Energy)],
    "corr_rimp2_gradient": [P, I64, I64, I64, P, P, P, D, D, I, P, P, P, C.POINTER(MP2Energy)],
    "corr_ccsd": [P, I64, I64, P, P, P, D, C.POINTER(CCSDOpts), P, P, P, C.POINTER(CCSDResult)],
    "corr_ccsd_ints_so": [P, I64, P, P],
    "corr_ccsd_update": [P, I64, I64, P, P, P, P, P, P, I],
    "corr_ccsd_energy": [P, I64, I64, P, P, P, P, C.POINTER(D)],
    "corr_build_fmat": [P, I64, P, P, P, P],
    "corr_build_fmat_list": [P, I64, I64, C.POINTER(I64), P, P, P, P, P],
    "corr_build_pmat": [P, I64, I64, P, D, P],
    "corr_non_canonical_mp2": [P, I64, I64, P, P, I, P, P, P],
    "corr_int1e": [P, C.POINTER(CorrBasis), I64, P, P, P, P, P],
    "corr_int2c": [P, C.POINTER(CorrBasis), P],
    "corr_int3c": [P, C.POINTER(CorrBasis), C.POINTER(CorrBasis), P],
    "corr_int4c": [P, C.POINTER(CorrBasis), P],
    "corr_int1e_dipole": [P, C.POINTER(CorrBasis), P, P],
    "corr_localize": [P, I64, I64, P, I, P, P, I64, C.POINTER(LocOpts), P, C.POINTER(LocResult)],
    "corr_run_scf": [P, I64, I64, P, P, P, P, P, P, P, D, C.POINTER(SCFOpts), P, P, C.POINTER(SCFResult)],
    "corr_build_fmat_ri": [P, I64, I64, P, P, P, P],
    "corr_run_scf_ri": [P, I64, I64, P, P, I64, P, P, P, P, P, D, C.POINTER(SCFOpts), P, P, C.POINTER(SCFResult)],
}

_lib = None
_lock = threading.Lock()
MISSING: list[str] = []   # declared entry points the loaded library lacks


def lib() -> C.CDLL:
    """Loads libcorr_gpu.so once; raises CorrLibraryError if it is absent."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise CorrLibraryError(
                    f"{LIB_PATH} not found: build it with `python -m corr_amd.build` "
                    "(there is no CPU fallback)")
            try:
                h = C.CDLL(str(LIB_PATH))
            except OSError as exc:
                raise CorrLibraryError(f"cannot load {LIB_PATH}: {exc}") from exc
            for name, args in SIGNATURES.items():
                fn = getattr(h, name, None)
                if fn is None:
                    MISSING.append(name)
                    continue
                fn.argtypes = args
                fn.restype = C.c_char_p if name == "corr_ctx_last_error" else C.c_int
            _lib = h
        return _lib


class Context:
    """One corr_ctx per device, bound to torch's current stream per call."""

    _per_device: dict[int, "Context"] = {}

    def __init__(self, device: int):
        self.device = device
        self.handle = P()
        st = lib().corr_ctx_create(device, C.byref(self.handle))
        if st != CORR_OK:
            raise CorrError("corr_ctx_create", st, f"device {device}")

    @classmethod
    def get(cls, device: int | None = None) -> "Context":
        if device is None:
            device = torch.cuda.current_device()
        ctx = cls._per_device.get(device)
        if ctx is None:
            if not cls._per_device:
                atexit.register(cls.destroy_all)
            ctx = cls._per_device[device] = cls(device)
        return ctx

    @classmethod
    def destroy_all(cls) -> None:
        """Destroys every context (scratch, streams, events) while the HIP
        runtime is still up; registered with atexit on first use."""
        for dev, ctx in list(cls._per_device.items()):
            if ctx.handle:
                lib().corr_ctx_destroy(ctx.handle)
                ctx.handle = P()
            del cls._per_device[dev]

    def bind_stream(self) -> None:
        s = torch.cuda.current_stream(self.device).cuda_stream
        self.check("corr_ctx_set_stream", lib().corr_ctx_set_stream(self.handle, P(s)))

    def check(self, fn: str, status: int) -> None:
        if status != CORR_OK:
            msg = lib().corr_ctx_last_error(self.handle)
            raise CorrError(fn, status, msg.decode() if msg else "")

    def call(self, fn: str, *args) -> None:
        self.bind_stream()
        self.check(fn, getattr(lib(), fn)(self.handle, *args))


    def enable_timing(self, on: bool = True) -> None:
        self.check("corr_ctx_enable_timing", lib().corr_ctx_enable_timing(self.handle, int(on)))

    def kernel_time(self, name: str | None) -> tuple[float, int]:
        ms, n = C.c_double(0.0), C.c_int64(0)
        self.check("corr_ctx_kernel_time", lib().corr_ctx_kernel_time(
            self.handle, name.encode() if name else None, C.byref(ms), C.byref(n)))
        return ms.value, n.value


def ptr(t: torch.Tensor | None) -> P:
    if t is None:
        return P()
    return P(t.data_ptr())
