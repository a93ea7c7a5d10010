#!/usr/bin/env python3
"""PME energy + force evaluations per second on one MI355X.

Workload: SPC-like water box, 8000 rigid 3-site molecules (24,000 atoms) in a
62.09 A cube at 1 g/cm^3, rcut 8 A, ethresh 1e-5 (kappa 0.4112 1/A), mesh 64^3,
B-spline order 6; one step = one energy + forces + dE/dbox evaluation from
positions already resident in HBM (cell-list real space, spread, r2c FFT,
influence function, c2r FFT, gather), float32 bulk arithmetic with float64
accumulation.

Printed as ONE JSON line.  Besides the contract fields it carries the per-stage
GPU times (HIP events around each stage of the fused launch sequence), the
error of the timed result against the float64 CPU oracle, the float64 GPU path
and the drop-in module path (TorchNeighborList + CoulombForceModule + autograd)
timed the same way.
"""

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "torch-admp_amd"))

METRIC = ("PME E+F evals/s at 24k atoms; batched p50/p99 ms vs plain at equal FP64 error")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the float64 and drop-in module timings")
    return ap.parse_args()


def timed(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(systems are independent, see DESIGN.md for how the path shards)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from torch_admp_amd import _lib
    from torch_admp_amd.nblist import TorchNeighborList
    from torch_admp_amd.pme import CoulombForceModule
    from torch_admp_amd.utils import calc_grads
    from torch_admp_amd.workloads import spc_water

    dev = torch.device("cuda:0")
    n_mol, rc, ethresh, kmesh = 8000, 8.0, 1e-5, 64
    pos_np, q_np, box_np = spc_water(n_mol, seed=2024)
    n = pos_np.shape[0]
    pos = torch.tensor(pos_np, dtype=torch.float64, device=dev)
    q = torch.tensor(q_np, dtype=torch.float64, device=dev)
    box = torch.tensor(box_np, dtype=torch.float64, device=dev)
    box_h = torch.tensor(box_np, dtype=torch.float64)  # launch parameter, kept on the host

    mod32 = CoulombForceModule(rc, ethresh, kmesh=kmesh, precision="float32")
    mod64 = CoulombForceModule(rc, ethresh, kmesh=kmesh, precision="float64")
    lib = _lib.load()

    # ---- timed region: fused float32-bulk path --------------------------------
    sec, (e32, f32, gb32) = timed(lambda: mod32.energy_forces(pos, box_h, q), args.steps,
                                  args.warmup, torch)
    value = 1.0 / sec

    # ---- stage timing with HIP events on the launch stream ---------------------
    _lib.check(lib.admp_profile_begin(args.steps), "admp_profile_begin")
    for _ in range(args.steps):
        mod32.energy_forces(pos, box_h, q)
    stage = (ctypes.c_double * 6)()
    ncalls = ctypes.c_int32()
    _lib.check(lib.admp_profile_collect(stage, ctypes.byref(ncalls)), "admp_profile_collect")
    names = ["bin", "real_space", "spread", "fft_convolve", "gather", "finalize"]
    stage_ms = {nm: stage[i] / max(ncalls.value, 1) for i, nm in enumerate(names)}

    # algorithmic bytes per launch (fp32 bulk): see DESIGN.md
    K3 = kmesh ** 3
    alg_bytes = {
        "spread": n * 216 * 4,                                    # stencil-expanded mesh updates
        "gather": n * 216 * 4,                                    # stencil-expanded mesh reads
        "fft_convolve": 2 * (4 * K3 + 8 * kmesh * kmesh * (kmesh // 2 + 1)),
        "real_space": 28 * n,                                     # xyz+q in, F out
        "bin": 28 * n,
    }
    dominant = max(("real_space", "spread", "fft_convolve", "gather"), key=lambda k: stage_ms[k])
    t_dom = stage_ms[dominant] * 1e-3
    achieved = alg_bytes[dominant] / t_dom / 1e9
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": None, "alg_bytes_per_launch": alg_bytes[dominant],
                "avg_launch_us": round(stage_ms[dominant] * 1e3, 2)}
    stage_gbs = {k: round(alg_bytes[k] / (stage_ms[k] * 1e-3) / 1e9, 1)
                 for k in alg_bytes if stage_ms.get(k, 0) > 0}

    extras = {}
    # ---- float64 GPU path and the drop-in module path -------------------------
    e64, f64, gb64 = mod64.energy_forces(pos, box_h, q)
    if not args.no_extras:
        sec64, _ = timed(lambda: mod64.energy_forces(pos, box_h, q), max(args.steps // 5, 3), 2,
                         torch)
        extras["fused_fp64_evals_per_s"] = round(1.0 / sec64, 2)

        def dropin():
            p = pos.clone().requires_grad_(True)
            nbl = TorchNeighborList(rc)
            pairs = nbl(p, box)
            e = mod64(p, box, pairs, nbl.get_ds(), nbl.get_buffer_scales(), {"charge": q})
            return e, -calc_grads(e, p)

        secd, (ed, fd) = timed(dropin, 3, 1, torch)
        extras["dropin_module_fp64_evals_per_s"] = round(1.0 / secd, 3)
        extras["dropin_vs_fused_fp64_energy_rel"] = float(abs(ed.detach()[0] - e64) / abs(e64))

    err32_vs64 = {"energy_rel": float(abs(e32 - e64) / abs(e64)),
                  "force_rmse_eV_per_A": float(((f32 - f64) ** 2).mean().sqrt())}

    # ---- CPU baseline: the float64 oracle (reference algorithm) ---------------
    cpu = None
    oracle_err = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT))
        from oracle import nblist_ref, pme_ref

        cores = min(16, os.cpu_count() or 1)
        torch.set_num_threads(cores)
        p_c = torch.tensor(pos_np, dtype=torch.float64)
        b_c = torch.tensor(box_np, dtype=torch.float64)
        q_c = torch.tensor(q_np, dtype=torch.float64)
        orc = pme_ref.CoulombOracle(rc, ethresh, kmesh=kmesh)
        t0 = time.perf_counter()
        pairs = nblist_ref.nblist_pbc_cells(p_c, b_c, rc)
        bs = nblist_ref.buffer_scales(pairs)
        e_ref, f_ref, gb_ref = pme_ref.energy_forces(
            p_c, b_c, q_c, pairs, lambda p, b: nblist_ref.pairs_ds(p, pairs, b), bs, orc)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": round(1.0 / t_cpu, 5), "unit": "evals/s", "cores": cores,
               "kind": "port",
               "sample": "1 full E+F+dE/dbox evaluation of the same 24k-atom box "
                         f"({t_cpu:.1f} s): reference cell-table neighbour list, "
                         "CoulombForceModule restatement, autograd forces, float64"}
        e_ref = float(e_ref)
        oracle_err = {
            "fp32_energy_rel": float(abs(float(e32) - e_ref) / abs(e_ref)),
            "fp32_force_rmse_eV_per_A": float(((f32.cpu() - f_ref) ** 2).mean().sqrt()),
            "fp64_energy_rel": float(abs(float(e64) - e_ref) / abs(e_ref)),
            "fp64_force_rmse_eV_per_A": float(((f64.cpu() - f_ref) ** 2).mean().sqrt()),
            "oracle_energy_eV": e_ref,
        }

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "accumulation": "f64",
        "data": "synthetic",
        "config": {"workload": "SPC-like water box, 8000 rigid 3-site molecules (24,000 atoms), "
                               "62.09 A cube, rcut 8 A, ethresh 1e-5, mesh 64^3, order 6, "
                               "energy + forces + dE/dbox",
                   "n_atoms": n, "rcut": rc, "kmesh": [kmesh] * 3, "kappa": mod32.kappa,
                   "path": "fused single-system (admp_pme_fused)"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
        "stage_alg_GBps": stage_gbs,
        "error_fp32_vs_fp64_gpu": err32_vs64,
        "error_vs_fp64_oracle": oracle_err,
        "energy_eV": float(e32),
        **extras,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
