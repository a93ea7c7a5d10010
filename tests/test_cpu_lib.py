"""CPU-side checks of the C-ABI library: it loads, exports every symbol that
include/*.h declares, and rejects invalid arguments without touching a GPU."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(admp_\w+)\s*\(", txt, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    from torch_admp_amd import _lib

    lib = _lib.load()
    declared = _declared()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_lib._SIGNATURES), "ctypes signature table out of sync with the header"


def test_status_strings_and_version():
    from torch_admp_amd import _lib

    lib = _lib.load()
    assert lib.admp_version() >= 100
    assert lib.admp_status_string(0) == b"ok"
    assert lib.admp_status_string(-1) == b"invalid argument"


def test_invalid_arguments_rejected_on_host():
    from torch_admp_amd import _lib

    lib = _lib.load()
    nbytes = ctypes.c_size_t()
    bad_box = (ctypes.c_double * 9)(*([0.0] * 9))
    assert lib.admp_nblist_workspace_bytes(10, bad_box, 5.0, ctypes.byref(nbytes)) == -1
    assert lib.admp_nblist_workspace_bytes(10, None, -1.0, ctypes.byref(nbytes)) == -1
    km = (ctypes.c_int32 * 3)(0, 16, 16)
    assert lib.admp_pme_workspace_bytes(km, 0, 10, ctypes.byref(nbytes)) == -1
    km = (ctypes.c_int32 * 3)(16, 16, 16)
    assert lib.admp_pme_workspace_bytes(km, 7, 10, ctypes.byref(nbytes)) == -1
    assert lib.admp_pme_workspace_bytes(km, 1, 10, ctypes.byref(nbytes)) == 0
    assert nbytes.value >= 16 * 16 * 9 * 8
    box = (ctypes.c_double * 9)(10.0, 0, 0, 0, 10.0, 0, 0, 0, 10.0)
    assert lib.admp_nblist_workspace_bytes(1000, box, 3.0, ctypes.byref(nbytes)) == 0
    assert lib.admp_pme_fused_workspace_bytes(1000, box, 3.0, km, 0, ctypes.byref(nbytes)) == 0
    # profiling: collect without begin is an invalid call
    stage = (ctypes.c_double * 6)()
    calls = ctypes.c_int32()
    assert lib.admp_profile_collect(stage, ctypes.byref(calls)) == -1


def test_cpu_tensors_fail_loudly():
    import torch

    from torch_admp_amd.nblist import TorchNeighborList

    with pytest.raises(RuntimeError, match="ROCm GPU"):
        TorchNeighborList(3.0)(torch.rand(10, 3, dtype=torch.float64) * 10,
                               torch.eye(3, dtype=torch.float64) * 10)


def test_batch_plan_on_host():
    """Batch planning is host-only: sizes grow with the batch and invalid
    systems are rejected."""
    import numpy as np
    import torch

    from torch_admp_amd import ops
    from torch_admp_amd._lib import AdmpError

    one = ops.PMEBatchPlan([300], [np.eye(3) * 20.0], 6.0, 0.4, [24, 24, 24])
    two = ops.PMEBatchPlan([300, 500], [np.eye(3) * 20.0, np.eye(3) * 25.0], [6.0, 7.0],
                           [0.4, 0.35], [[24, 24, 24], [30, 30, 30]], precision="float32")
    assert two.workspace_bytes > one.workspace_bytes > 0
    assert two.n_total == 800 and list(two.offsets) == [0, 300, 800]
    with pytest.raises(AdmpError, match="status -1"):
        ops.PMEBatchPlan([10], [np.zeros((3, 3))], 6.0, 0.4, [24, 24, 24])
    with pytest.raises(AdmpError, match="status -1"):
        ops.PMEBatchPlan([10], [np.eye(3) * 10.0], 6.0, 0.4, [0, 24, 24])
    # a box far below the cutoff (7^3 periodic images per dimension triple) is planned
    assert ops.PMEBatchPlan([10], [np.eye(3) * 3.0], 9.0, 0.4, [8, 8, 8]).workspace_bytes > 0
    with pytest.raises(ValueError):
        ops.PMEBatchPlan([10, 20], [np.eye(3) * 10.0] * 2, [6.0], 0.4, [8, 8, 8])
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        one.run(torch.zeros(300, 3, dtype=torch.float64), torch.zeros(300, dtype=torch.float64))


def test_host_extension_binds_and_rejects_host_tensors():
    """The native binding of the fused call loads, resolves its C-ABI entry
    points in the HIP library and refuses host tensors before any GPU call."""
    import torch

    from torch_admp_amd import ops

    ext = ops._host_ext()
    if ext is None:
        pytest.skip("host extension not built")
    pos = torch.zeros(4, 3, dtype=torch.float64)
    q = torch.zeros(4, dtype=torch.float64)
    box = torch.eye(3, dtype=torch.float64) * 10.0
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        ext.pme_fused(pos, q, box, 4.0, 0.3, 8, 8, 8, 0, 0, 138.935, 0)
