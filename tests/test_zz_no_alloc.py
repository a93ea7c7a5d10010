"""admp_pme_fused_noalloc (the fused evaluation in the library's static
scratch, no workspace argument) against admp_pme_fused with a caller-owned
workspace, through the C ABI, on two calls in a row with the second one
smaller: the 24,000-atom water box (64^3), then a 1000-ion box (32^3), in
float64 (energy 1e-12 relative, forces 1e-10, dE/dbox 1e-9: both run the
same kernels; float64 atomics leave last-bit differences) and float32 (1e-6
relative energy, 1e-5 eV/A forces, dE/dbox 1e-4: float32 atomics and the
in-cell order change partial sums).  A problem larger
than the scratch is refused with ADMP_ERR_SCRATCH before any launch."""

import ctypes

import numpy as np
import pytest
import torch

from torch_admp_amd import _lib
from torch_admp_amd.workloads import ionic_box, spc_water

pytestmark = pytest.mark.gpu


def _call(fn, pos, q, box, rc, kappa, km, prec, with_ws):
    lib = _lib.load()
    n = pos.shape[0]
    dev = pos.device
    boxh = (ctypes.c_double * 9)(*box.reshape(-1).tolist())
    kmh = (ctypes.c_int32 * 3)(*km)
    out = torch.full((14 + 3 * n,), float("nan"), dtype=torch.float64, device=dev)
    energy, dedbox, forces = out[:5], out[5:14], out[14:].view(n, 3)
    stream = _lib.stream_of(pos)
    if with_ws:
        nb = ctypes.c_size_t()
        _lib.check(lib.admp_pme_fused_workspace_bytes(n, boxh, rc, kmh, prec, ctypes.byref(nb)),
                   "workspace")
        ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
        st = lib.admp_pme_fused(_lib.ptr(pos), _lib.ptr(q), n, boxh, rc, kappa, kmh, prec, 0,
                                14.399645468667817, _lib.ptr(energy), _lib.ptr(forces),
                                _lib.ptr(dedbox), _lib.ptr(ws), nb.value, stream)
    else:
        st = lib.admp_pme_fused_noalloc(_lib.ptr(pos), _lib.ptr(q), n, boxh, rc, kappa, kmh, prec,
                                        0, 14.399645468667817, _lib.ptr(energy),
                                        _lib.ptr(forces), _lib.ptr(dedbox), stream)
    _lib.check(st, fn)
    torch.cuda.synchronize()
    return energy.cpu(), forces.cpu(), dedbox.cpu()


@pytest.mark.parametrize("prec", [_lib.ADMP_F64, _lib.ADMP_F32])
def test_noalloc_matches_workspace_entry_point(dev, prec):
    water = spc_water(8000, seed=0)
    ions = ionic_box(1000, 20.0, seed=0)
    cases = [(water, 8.0, 0.41116789234315304, [64, 64, 64]),
             (ions, 6.0, 0.5482238564575374, [32, 32, 32])]
    etol, ftol, gtol = (1e-12, 1e-10, 1e-9) if prec == _lib.ADMP_F64 else (1e-6, 1e-5, 1e-4)
    for (pos_np, q_np, box_np), rc, kappa, km in cases:     # the second call is smaller
        pos = torch.tensor(pos_np, device=dev)
        q = torch.tensor(q_np, device=dev)
        e0, f0, g0 = _call("admp_pme_fused_noalloc", pos, q, box_np, rc, kappa, km, prec, False)
        e1, f1, g1 = _call("admp_pme_fused", pos, q, box_np, rc, kappa, km, prec, True)
        assert torch.isfinite(e0).all() and torch.isfinite(f0).all() and torch.isfinite(g0).all()
        assert float((e0 - e1).abs().max()) <= etol * float(e1.abs().max()), pos.shape
        assert float((f0 - f1).abs().max()) <= ftol * max(1.0, float(f1.abs().max())), pos.shape
        assert float((g0 - g1).abs().max()) <= gtol * float(g1.abs().max()), pos.shape


def test_noalloc_refuses_problems_beyond_the_scratch(dev):
    """8192 atoms at a 160^3 float64 mesh need about 66 MiB: refused on the
    host before any launch, with the status text naming the limit."""
    lib = _lib.load()
    L, n, km = 43.4, 8192, [160, 160, 160]
    boxh = (ctypes.c_double * 9)(L, 0, 0, 0, L, 0, 0, 0, L)
    kmh = (ctypes.c_int32 * 3)(*km)
    nb = ctypes.c_size_t()
    _lib.check(lib.admp_pme_fused_workspace_bytes(n, boxh, 6.0, kmh, _lib.ADMP_F64,
                                                  ctypes.byref(nb)), "workspace")
    assert nb.value > _lib.ADMP_NOALLOC_SCRATCH_BYTES
    rng = np.random.default_rng(0)
    pos = torch.tensor(rng.random((n, 3)) * L, device=dev)
    q = torch.zeros(n, dtype=torch.float64, device=dev)
    out = torch.zeros(14 + 3 * n, dtype=torch.float64, device=dev)
    st = lib.admp_pme_fused_noalloc(_lib.ptr(pos), _lib.ptr(q), n, boxh, 6.0, 0.5, kmh,
                                    _lib.ADMP_F64, 0, 14.4, _lib.ptr(out), _lib.ptr(out[14:]),
                                    _lib.ptr(out[5:14]), _lib.stream_of(pos))
    assert st == _lib.ADMP_ERR_SCRATCH
    assert b"static scratch" in lib.admp_status_string(st)
    assert float(out.abs().max()) == 0.0
