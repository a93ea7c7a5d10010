"""ctypes binding of libadmp_amd.so (the C ABI declared in include/admp_amd.h).

The library is the only compute path of this package: importing works
everywhere (so that shape validation and host logic can be used and tested on
a CPU-only machine), but every call that reaches device code raises when the
library is missing or the tensors are not on a ROCm GPU.  There is no CPU
fallback.
"""

from __future__ import annotations

import ctypes
import os
import weakref
from collections import OrderedDict
from pathlib import Path
from typing import Optional, Tuple

import torch

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libadmp_amd.so"

ADMP_OK = 0
ADMP_OVERFLOW = 1
ADMP_ERR_SCRATCH = -4
ADMP_NOALLOC_SCRATCH_BYTES = 64 << 20
ADMP_BATCH_NO_OVERLAP = 1
ADMP_F64 = 0
ADMP_F32 = 1
KCONV = {"reference": 0, "columns": 1}
PRECISION = {"float64": ADMP_F64, "float32": ADMP_F32}

_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64
_c_dbl = ctypes.c_double
_c_vp = ctypes.c_void_p
_c_sz = ctypes.c_size_t
_P = ctypes.POINTER

# name -> (restype, argtypes); mirrors include/admp_amd.h
_SIGNATURES = {
    "admp_version": (ctypes.c_int, []),
    "admp_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "admp_nblist_workspace_bytes": (ctypes.c_int, [_c_i64, _c_vp, _c_dbl, _P(_c_sz)]),
    "admp_nblist_count": (ctypes.c_int, [_c_vp, _c_i64, _c_vp, _c_dbl, _c_vp, _c_sz,
                                         _P(_c_i64), _c_vp]),
    "admp_nblist_fill": (ctypes.c_int, [_c_vp, _c_i64, _c_vp, _c_dbl, _c_vp, _c_sz, _c_vp,
                                        _c_i64, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "admp_pair_scales": (ctypes.c_int, [_c_vp, _c_i64, _c_vp, _c_dbl, _c_dbl, _c_vp, _c_vp]),
    "admp_max_displacement": (ctypes.c_int, [_c_vp, _c_vp, _c_i64, _c_vp, _c_vp, _c_vp]),
    "admp_pair_ds": (ctypes.c_int, [_c_vp, _c_vp, _c_i64, _c_vp, _c_vp, _c_vp, _c_vp]),
    "admp_pair_ds_backward_workspace_bytes": (ctypes.c_int, [_c_i64, _P(_c_sz)]),
    "admp_pair_ds_backward": (ctypes.c_int, [_c_vp, _c_vp, _c_i64, _c_vp, _c_vp, _c_vp, _c_vp,
                                             _c_vp, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp, _c_sz,
                                             _c_vp]),
    "admp_pair_energy_workspace_bytes": (ctypes.c_int, [_c_i64, _P(_c_sz)]),
    "admp_pair_energy": (ctypes.c_int, [_c_i32, _c_i64, _c_i64, _c_i64, _c_vp, _c_vp, _c_vp,
                                        _c_vp, _c_vp, _c_dbl, _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp,
                                        _c_vp, _c_sz, _c_vp]),
    "admp_pme_workspace_bytes": (ctypes.c_int, [_c_vp, _c_i32, _c_i64, _P(_c_sz)]),
    "admp_pme_spread": (ctypes.c_int, [_c_vp, _c_vp, _c_i64, _c_vp, _c_vp, _c_i32, _c_vp, _c_vp,
                                       _c_sz, _c_vp]),
    "admp_pme_convolve": (ctypes.c_int, [_c_vp, _c_vp, _c_vp, _c_i32, _c_i32, _c_dbl, _c_dbl,
                                         _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp, _c_sz, _c_vp]),
    "admp_pme_box_grad": (ctypes.c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "admp_ewald_self_terms": (ctypes.c_int, [_c_vp, _c_i64, _c_i64, _c_vp, _c_dbl, _c_dbl, _c_vp,
                                             _c_vp, _c_vp]),
    "admp_ewald_self_terms_backward": (ctypes.c_int, [_c_vp, _c_i64, _c_i64, _c_vp, _c_dbl,
                                                      _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp,
                                                      _c_vp]),
    "admp_pme_gather": (ctypes.c_int, [_c_vp, _c_vp, _c_i64, _c_vp, _c_vp, _c_i32, _c_vp, _c_vp,
                                       _c_vp, _c_vp, _c_i32, _c_vp, _c_sz, _c_vp]),
    "admp_pme_fused_workspace_bytes": (ctypes.c_int, [_c_i64, _c_vp, _c_dbl, _c_vp, _c_i32,
                                                      _P(_c_sz)]),
    "admp_pme_fused": (ctypes.c_int, [_c_vp, _c_vp, _c_i64, _c_vp, _c_dbl, _c_dbl, _c_vp, _c_i32,
                                      _c_i32, _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp, _c_sz, _c_vp]),
    "admp_pme_fused_noalloc": (ctypes.c_int, [_c_vp, _c_vp, _c_i64, _c_vp, _c_dbl, _c_dbl, _c_vp,
                                              _c_i32, _c_i32, _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp]),
    "admp_pme_batch_workspace_bytes": (ctypes.c_int, [_c_i32, _c_vp, _c_vp, _c_vp, _c_vp, _c_i32,
                                                      _P(_c_sz)]),
    "admp_pme_batch_system_bytes": (ctypes.c_int, [_c_i32, _c_vp, _c_vp, _c_vp, _c_vp, _c_i32,
                                                   _c_vp]),
    "admp_pme_batch": (ctypes.c_int, [_c_i32, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_i32, _c_i32,
                                      _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_sz,
                                      _c_vp]),
    "admp_pme_batch_ex": (ctypes.c_int, [_c_i32, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_i32,
                                         _c_i32, _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp,
                                         _c_i32, _c_vp, _c_sz, _c_vp]),
    "admp_pme_batch_upload": (ctypes.c_int, [_c_i32, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_i32,
                                             _c_i32, _c_dbl, _c_vp, _c_sz, _c_vp]),
    "admp_pme_batch_run": (ctypes.c_int, [_c_i32, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_i32,
                                          _c_i32, _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp,
                                          _c_vp, _c_sz, _c_vp]),
    "admp_qeq_workspace_bytes": (ctypes.c_int, [_c_vp, _P(_c_sz)]),
    "admp_qeq_matvec": (ctypes.c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_sz, _c_vp]),
    "admp_qeq_solve": (ctypes.c_int, [_c_vp, _c_vp, _c_dbl, _c_i32, _c_i32, _c_vp, _P(_c_i32),
                                      _P(_c_dbl), _c_vp, _c_sz, _c_vp]),
    "admp_qeq_solve_batch_workspace_bytes": (ctypes.c_int, [_c_i32, _c_vp, _c_vp, _c_vp, _c_vp,
                                                            _c_i32, _c_i64, _P(_c_sz)]),
    "admp_qeq_solve_batch": (ctypes.c_int, [_c_i32, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_i32,
                                            _c_i32, _c_dbl, _c_i32, _c_vp, _c_vp, _c_vp, _c_vp,
                                            _c_i64, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp,
                                            _c_vp, _c_i32, _c_i32, _c_vp, _c_vp, _c_vp, _c_vp,
                                            _c_sz, _c_vp]),
    "admp_profile_begin": (ctypes.c_int, [_c_i32]),
    "admp_profile_collect": (ctypes.c_int, [_P(_c_dbl), _P(_c_i32)]),
}

_lib: Optional[ctypes.CDLL] = None


class AdmpError(RuntimeError):
    """A non-OK status returned by the HIP library."""


def library_path() -> Path:
    return Path(os.environ.get("ADMP_AMD_LIB", str(_LIB_PATH)))


def load() -> ctypes.CDLL:
    """Load (once) and return the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not path.exists():
        raise ImportError(
            f"torch_admp_amd: HIP library not found at {path}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C torch-admp_amd/csrc`")
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, what: str) -> None:
    if status != ADMP_OK:
        msg = load().admp_status_string(status).decode()
        raise AdmpError(f"{what} failed with status {status}: {msg}")


def ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_of(t: torch.Tensor):
    return ctypes.c_void_p(torch._C._cuda_getCurrentRawStream(t.device.index))


def require_gpu(*tensors: Optional[torch.Tensor]) -> None:
    """Fail loudly unless every tensor lives on a ROCm GPU."""
    for t in tensors:
        if t is not None and t.device.type != "cuda":
            raise RuntimeError(
                "torch_admp_amd runs only on a ROCm GPU (no CPU fallback); got a tensor on "
                f"{t.device}")


_BOX_ARRAYS: "OrderedDict[int, Tuple[weakref.ref, int, ctypes.Array]]" = OrderedDict()


def box_array(box: Optional[torch.Tensor], box_host: Optional[torch.Tensor] = None):
    """Host copy of a (3, 3) box as a ctypes double[9] (None for open
    boundaries).  ``box_host``, a CPU copy the caller already holds, saves the
    device-to-host copy (a stream synchronisation) of a device box; the array
    made from it is kept while that tensor lives unchanged, for the other
    calls of the same evaluation."""
    if box is None:
        return None
    if box_host is not None:
        ent = _BOX_ARRAYS.get(id(box_host))
        if ent is not None and ent[0]() is box_host and ent[1] == box_host._version:
            return ent[2]
    src = box_host if box_host is not None else box
    vals = src.detach().reshape(9).to("cpu", torch.float64).tolist()
    arr = (ctypes.c_double * 9)(*vals)
    if box_host is not None:
        _BOX_ARRAYS[id(box_host)] = (weakref.ref(box_host), box_host._version, arr)
        while len(_BOX_ARRAYS) > 16:
            _BOX_ARRAYS.popitem(last=False)
    return arr


# Host copies of device boxes made by the neighbour lists, keyed by the device
# address of the box data: a later module call on the same box (or a view of
# it) reuses the copy instead of copying again, which would wait for the
# stream.  An entry holds only while the tensor it was made from is alive and
# its version counter (bumped by every in-place change, also through views)
# is unchanged; changes written through ``.data`` are not seen.
_HOST_BOXES: "OrderedDict[Tuple[str, int], Tuple[weakref.ref, int, torch.Tensor]]" = OrderedDict()


def remember_host_box(box: torch.Tensor, host: torch.Tensor) -> None:
    key = (str(box.device), box.data_ptr())
    _HOST_BOXES[key] = (weakref.ref(box), box._version, host)
    _HOST_BOXES.move_to_end(key)
    while len(_HOST_BOXES) > 8:
        _HOST_BOXES.popitem(last=False)


def host_box(box: torch.Tensor) -> torch.Tensor:
    """float64 CPU copy of a (3, 3) device box: the neighbour list's copy of
    the same data when still valid, else a fresh copy."""
    ent = _HOST_BOXES.get((str(box.device), box.data_ptr()))
    if ent is not None:
        ref, version, host = ent
        base = ref()
        if base is not None and base._version == version and box._version == version:
            return host.reshape(box.shape)
    return box.detach().to(torch.device("cpu"), torch.float64)


def kmesh_array(kmesh) -> ctypes.Array:
    return (ctypes.c_int32 * 3)(*[int(k) for k in kmesh])


def workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


_STREAM_WS: "dict[Tuple[int, int, str], torch.Tensor]" = {}


def stream_workspace(kind: str, nbytes: int, device) -> torch.Tensor:
    """A workspace of at least nbytes kept per (device, current stream,
    kind).  Work on one stream runs in order, so a later call of the same kind
    on that stream may reuse it; calls on other streams get their own."""
    dev = torch.device(device)
    key = (dev.index if dev.index is not None else torch.cuda.current_device(),
           torch._C._cuda_getCurrentRawStream(dev.index), kind)
    ws = _STREAM_WS.get(key)
    if ws is None or ws.numel() < nbytes:
        if ws is None and len(_STREAM_WS) >= 64:
            _STREAM_WS.clear()
        ws = _STREAM_WS[key] = workspace(nbytes, dev)
    return ws
