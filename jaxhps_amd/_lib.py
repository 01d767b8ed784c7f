"""ctypes binding of libhps_mi355x.so (the C ABI declared in include/hps_mi355x.h).

There is no fallback: if the library or a GPU is missing, every stage raises
``HPSLibraryError``.  torch is used only for device memory and the stream.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libhps_mi355x.so"

HPS_OK = 0
HPS_F64, HPS_F32, HPS_C128, HPS_C64 = 0, 1, 2, 3
HPS_KIND_DTN, HPS_KIND_ITI = 0, 1

DTYPE_CODE = {
    torch.float64: HPS_F64,
    torch.float32: HPS_F32,
    torch.complex128: HPS_C128,
    torch.complex64: HPS_C64,
}

# name -> (restype, argtypes)
_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_long
_D = ctypes.c_double
_S = ctypes.c_size_t
SIGNATURES = {
    "hps_create": (_I, [_P, ctypes.POINTER(_P)]),
    "hps_destroy": (_I, [_P]),
    "hps_set_stream": (_I, [_P, _P]),
    "hps_last_error": (ctypes.c_char_p, [_P]),
    "hps_version": (_I, []),
    "hps_gemm_batched": (_I, [_P, _I, _I, _I, _I, _D, _P, _L, _L, _P, _L, _L, _D, _P, _L, _L, _P, _L, _L, _I]),
    "hps_getrf_workspace": (_S, [_I, _I]),
    "hps_dinv_size": (_S, [_I]),
    "hps_getrf_batched": (_I, [_P, _I, _I, _P, _L, _L, _P, _P, _P, _I, _P, _S]),
    "hps_getrs_batched": (_I, [_P, _I, _I, _I, _P, _L, _L, _P, _P, _P, _L, _L, _P, _L, _L, _I]),
    "hps_leaf_workspace": (_S, [_I, _I, _I, _I, _I, _I, _I]),
    "hps_leaf_solve": (
        _I,
        [_P, _I, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _S],
    ),
    "hps_set_merge_pivoting": (_I, [_P, _I]),
    "hps_set_leaf_pivoting": (_I, [_P, _I]),
    "hps_merge_workspace": (_S, [_I, _I, _I, _I, _I, _I]),
    "hps_merge_level": (
        _I,
        [_P, _I, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _P, _I, _P, _I, _P, _P, _P, _P, _P,
         _P, _P, _P, _P, _P, _P, _P, _P, _P, _S],
    ),
    "hps_up_workspace": (_S, [_I, _I, _I, _I, _I]),
    "hps_up_level": (
        _I,
        [_P, _I, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
         _S],
    ),
    "hps_block_inverses_elems": (_S, [_I]),
    "hps_block_inverses_workspace": (_S, [_I, _I, _I]),
    "hps_block_inverses": (_I, [_P, _I, _I, _I, _P, _P, _P, _P, _S]),
    "hps_down_workspace": (_S, [_I, _I, _I, _I]),
    "hps_down_level": (_I, [_P, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _S]),
    "hps_static_scratch_bytes": (_S, []),
    "hps_down_level_static": (_I, [_P, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P]),
    "hps_leaf_apply": (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P]),
    "hps_residual_iti": (_I, [_P] + [_I] * 9 + [_P, _P, _P, _I] + [_P] * 7),
    "hps_residual_dtn": (_I, [_P] + [_I] * 9 + [_P, _P, _P, _I] + [_P] * 9),
    "hps_refine_update": (_I, [_P, _I, _I, _L, _P, _P, _I, _P]),
    "hps_gemm_batched_conjt": (_I, [_P, _I, _I, _I, _I, _D, _P, _L, _L, _P, _L, _L, _D, _P, _L, _L, _P, _L, _L, _I]),
    "hps_getrs_conjt_workspace": (_S, [_I, _I, _I, _I]),
    "hps_getrs_batched_conjt": (_I, [_P, _I, _I, _I, _P, _L, _L, _P, _P, _P, _L, _L, _P, _L, _L, _I, _P, _S]),
    "hps_adjoint_down_level": (_I, [_P] + [_I] * 7 + [_P] * 6),
    "hps_adjoint_up_workspace": (_S, [_I, _I, _I, _I, _I]),
    "hps_adjoint_up_level": (_I, [_P] + [_I] * 8 + [_P] * 12 + [_S]),
    "hps_coef_gradient": (_I, [_P] + [_I] * 7 + [_P, _P, _I, _P, _P, _P]),
    "hps_interp_grid": (_I, [_P, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _I, _P, _P, _P, _P, _I, _P, _P, _P]),
    "hps_plane_wave": (_I, [_P, _L, _I, _P, _P, _D, _L, _P, _P]),
    "hps_scale_batched": (_I, [_P, _I, _I, _L, _L, _P, _P]),
    "hps_pack_blocks": (_I, [_P, _I, _I, _L, _P]),
    "hps_iti_to_dtn_workspace": (_S, [_I]),
    "hps_iti_to_dtn": (_I, [_P, _I, _I, _P, _L, _D, _P, _L, _P, _P, _S]),
    "hps_scattering_workspace": (_S, [_I, _I]),
    "hps_scattering_system": (_I, [_P, _I, _I, _I, _P, _P, _P, _P, _P, _D, _P, _P, _P, _S]),
    "hps_scattering_impedance": (_I, [_P, _I, _I, _I, _P, _P, _P, _P, _P, _D, _D, _P, _P, _P, _S]),
    "hps_mlp_gemm_bf16": (_I, [_P, _I, _I, _I, _I, _I, _P, _I, _P, _I, ctypes.c_float, _P, _I, _P, _I]),
    "hps_mlp_tanh_backward": (_I, [_P, _L, _P, _P]),
    "hps_mlp_bias_grad": (_I, [_P, _I, _I, _P, _P]),
    "hps_mlp_adam": (_I, [_P, _L, _P, _P, _P, _P] + [ctypes.c_float] * 4 + [_I]),
    "hps_mlp_lowrank_loss": (_I, [_P, _I, _I, _I, _P, _P, ctypes.c_float, _P, _P]),
    "hps_mlp_lowrank_apply": (_I, [_P, _I, _I, _I, _P, _P, _D, _P]),
    "hps_convert": (_I, [_P, _I, _I, _L, _P, _P]),
    "hps_prof_enable": (_I, [_I]),
    "hps_prof_reset": (_I, []),
    "hps_prof_count": (_I, []),
    "hps_prof_query": (_I, [_I, ctypes.c_char_p, _I, ctypes.POINTER(_L), ctypes.POINTER(_D), ctypes.POINTER(_D), ctypes.POINTER(_D)]),
}


class HPSLibraryError(RuntimeError):
    """The native library is missing, failed to load, or reported an error."""


class HPSSingularError(ValueError):
    """A leaf or merge matrix had an exactly zero pivot column."""


_lib = None


def load_library() -> ctypes.CDLL:
    """Load libhps_mi355x.so and bind its signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise HPSLibraryError(
            f"{LIB_PATH} not found; build it with `python -m jaxhps_amd.csrc.build` "
            "(the HIP path has no CPU fallback)"
        )
    try:
        lib = ctypes.CDLL(str(LIB_PATH))
    except OSError as e:  # pragma: no cover - depends on the environment
        raise HPSLibraryError(f"failed to load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _require_gpu():
    if not torch.cuda.is_available():
        raise HPSLibraryError("jaxhps_amd needs a ROCm GPU; torch.cuda.is_available() is False")


_handles: dict = {}


def handle():
    """Library handle bound to torch's current stream on the current device."""
    _require_gpu()
    lib = load_library()
    stream = torch.cuda.current_stream()
    key = (stream.device_index, stream.cuda_stream)
    h = _handles.get(key)
    if h is None:
        hp = ctypes.c_void_p()
        rc = lib.hps_create(ctypes.c_void_p(stream.cuda_stream), ctypes.byref(hp))
        if rc != HPS_OK:
            raise HPSLibraryError("hps_create failed")
        h = hp
        _handles[key] = h
    return h


def check(rc: int, h, what: str):
    if rc != HPS_OK:
        msg = load_library().hps_last_error(h)
        raise HPSLibraryError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise HPSLibraryError("expected a device tensor")
    if not t.is_contiguous():
        raise HPSLibraryError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


def dtype_code(dt) -> int:
    try:
        return DTYPE_CODE[dt]
    except KeyError as e:
        raise HPSLibraryError(f"unsupported dtype {dt}") from e


# --------------------------------------------------------------------------
# thin typed wrappers
# --------------------------------------------------------------------------
def gemm(A, B, C=None, alpha=1.0, beta=0.0, out=None):
    """Batched D = alpha A @ B + beta C on 3D contiguous tensors (batch, rows, cols).
    A 2D operand broadcasts over the batch."""
    a3 = A if A.dim() == 3 else A.unsqueeze(0)
    b3 = B if B.dim() == 3 else B.unsqueeze(0)
    batch = max(a3.shape[0], b3.shape[0])
    M, K = a3.shape[1], a3.shape[2]
    N = b3.shape[2]
    c3 = None if C is None else (C if C.dim() == 3 else C.unsqueeze(0))
    # the kernel indexes every operand by the common batch: a mismatch would read past the end
    if b3.shape[1] != K:
        raise ValueError(f"gemm: inner dimensions differ ({tuple(A.shape)} @ {tuple(B.shape)})")
    for name, t, shp in (("A", a3, (M, K)), ("B", b3, (K, N)), ("C", c3, (M, N)), ("out", out, (M, N))):
        if t is None:
            continue
        t3 = t if t.dim() == 3 else t.unsqueeze(0)
        if t3.shape[0] not in (1, batch) or tuple(t3.shape[1:]) != shp or not t3.is_contiguous():
            raise ValueError(f"gemm: {name} has shape {tuple(t.shape)}, expected (1 or {batch}, {shp[0]}, {shp[1]}) "
                             "contiguous")
        if t.dtype != A.dtype:
            raise ValueError(f"gemm: {name} has dtype {t.dtype}, A has {A.dtype}")
    if out is None:
        out = torch.empty((batch, M, N), dtype=A.dtype, device=A.device)
    elif (out if out.dim() == 3 else out.unsqueeze(0)).shape[0] != batch:
        raise ValueError(f"gemm: out must hold the whole batch of {batch}")
    sA = 0 if a3.shape[0] == 1 and batch > 1 else M * K
    sB = 0 if b3.shape[0] == 1 and batch > 1 else K * N
    sC = 0 if c3 is None or (c3.shape[0] == 1 and batch > 1) else M * N
    lib = load_library()
    h = handle()
    rc = lib.hps_gemm_batched(
        h, dtype_code(A.dtype), M, N, K, float(alpha), ptr(a3), K, sA, ptr(b3), N, sB, float(beta),
        ptr(c3) if c3 is not None else None, N, sC, ptr(out), N, M * N, batch,
    )
    check(rc, h, "hps_gemm_batched")
    return out


def dinv_size(n: int) -> int:
    return int(load_library().hps_dinv_size(n))


def getrf(A):
    """In-place batched LU of (batch, n, n) A.  Returns (perm, dinv, info)."""
    lib = load_library()
    h = handle()
    b, n, _ = A.shape
    perm = torch.empty((b, n), dtype=torch.int32, device=A.device)
    dinv = torch.empty((b, dinv_size(n)), dtype=A.dtype, device=A.device)
    info = torch.empty((b,), dtype=torch.int32, device=A.device)
    ws = workspace(lib.hps_getrf_workspace(n, b), A.device)
    rc = lib.hps_getrf_batched(
        h, dtype_code(A.dtype), n, ptr(A), n, n * n, ptr(perm), ptr(dinv), ptr(info), b, ptr(ws), ws.numel()
    )
    check(rc, h, "hps_getrf_batched")
    return perm, dinv, info


def getrs(LU, perm, dinv, R):
    """X = A^{-1} R for (batch, n, nrhs) R from getrf's factors."""
    lib = load_library()
    h = handle()
    b, n, _ = LU.shape
    nrhs = R.shape[2]
    X = torch.empty_like(R)
    rc = lib.hps_getrs_batched(
        h, dtype_code(LU.dtype), n, nrhs, ptr(LU), n, n * n, ptr(perm), ptr(dinv), ptr(R), nrhs, n * nrhs, ptr(X),
        nrhs, n * nrhs, b,
    )
    check(rc, h, "hps_getrs_batched")
    return X


def raise_on_info(info: torch.Tensor, what: str):
    bad = torch.nonzero(info).flatten()
    if bad.numel():
        i = int(bad[0])
        raise HPSSingularError(
            f"{what}: batch element {i} has an exactly zero pivot in column {int(info[i]) - 1} "
            f"({bad.numel()} singular element(s))"
        )
