"""The float32 erfc of the real-space kernel (torch-admp_amd/csrc/real_device.hpp,
coul_pair for float) against scipy.special.erfc.

The kernel evaluates erfc(kappa r) = exp(-kappa^2 r^2) t S(t), t = 1 / (1 + 0.47
kappa r) (the form of Abramowitz & Stegun 7.1.26) with S a degree-10
polynomial whose coefficients are read here from the kernel source.  This
test repeats the kernel's float32 operation sequence in numpy (fused
multiply-adds as one float64 operation rounded to float32, the exp2 argument
split into a rounded part and its first-order correction) and compares with
erfc of the same float32 squared distance in float64:
  * the fit itself, evaluated in float64: relative error below 5e-8;
  * the float32 evaluation over r in [0.01, 8] A and kappa 0.3-0.7 A^-1:
    relative error below 4e-7, mean |error| below 1e-7, mean error (bias)
    below 1e-8.
"""

import re
from pathlib import Path

import numpy as np
from scipy.special import erfc

SRC = Path(__file__).resolve().parents[1] / "torch-admp_amd" / "csrc" / "real_device.hpp"
f32 = np.float32


def _coefficients():
    text = SRC.read_text()
    body = text[text.index("__device__ __forceinline__ void coul_pair(float r2"):]
    body = body[:body.index("*e = qq * ec_ir;")]
    first = re.search(r"float s = ([-0-9.e+]+)f;", body).group(1)
    rest = re.findall(r"s = fmaf\(s, t, ([-0-9.e+]+)f\);", body)
    consts = dict(re.findall(r"(kLog2eH|kLog2eL|kLn2) = ([-0-9.e+]+)f", body))
    p = float(re.search(r"fmaf\(([0-9.]+)f, C\.kappa \* r, 1\.0f\)", body).group(1))
    return [float(first)] + [float(c) for c in rest], consts, p


def _fma(a, b, c):
    return f32(np.float64(a) * np.float64(b) + np.float64(c))


def _kernel_erfc(r, kappa, coeffs, consts, p):
    """erfc(kappa sqrt(r2)) as the kernel evaluates it, with r2 = fl32(r)^2."""
    kappa = f32(kappa)
    k2 = np.float64(kappa) ** 2
    k2h = f32(k2)
    k2l = f32(k2 - np.float64(k2h))
    r2 = f32(np.float64(f32(r)) ** 2)
    ir = f32(1.0 / np.sqrt(np.float64(r2)))
    rr = f32(r2 * ir)
    y = f32(k2h * r2)
    yl = f32(_fma(k2h, r2, -y) + f32(k2l * r2))
    l2h, l2l, ln2 = (f32(float(consts[k])) for k in ("kLog2eH", "kLog2eL", "kLn2"))
    a = f32(y * l2h)
    al = f32(_fma(y, l2h, -a) + _fma(y, l2l, f32(yl * l2h)))
    ex = f32(np.exp2(np.float64(-a)))
    ex = _fma(-ex, f32(al * ln2), ex)
    t = f32(1.0 / np.float64(_fma(f32(p), f32(kappa * rr), f32(1.0))))
    s = f32(coeffs[0])
    for c in coeffs[1:]:
        s = _fma(s, t, f32(c))
    return float(f32(f32(ex * t) * s)), float(np.sqrt(np.float64(r2)) * np.float64(kappa))


def test_fit_in_exact_arithmetic():
    coeffs, _, p = _coefficients()
    assert len(coeffs) == 11 and p == 0.47
    x = np.linspace(0.0, 6.0, 100001)
    t = 1.0 / (1.0 + p * x)
    s = np.zeros_like(x)
    for c in [float(f32(c)) for c in coeffs]:
        s = s * t + c
    assert np.abs(np.exp(-x * x) * t * s / erfc(x) - 1.0).max() < 5e-8


def test_float32_evaluation_error():
    coeffs, consts, p = _coefficients()
    rel = []
    for kappa in (0.3, 0.411168, 0.548224, 0.7):
        for r in np.linspace(0.01, 8.0, 4001):
            v, x = _kernel_erfc(r, kappa, coeffs, consts, p)
            rel.append((v - erfc(x)) / erfc(x))
    rel = np.array(rel)
    assert np.abs(rel).max() < 4e-7
    assert np.abs(rel).mean() < 1e-7
    assert abs(rel.mean()) < 1e-8
