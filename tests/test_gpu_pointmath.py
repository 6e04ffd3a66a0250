"""The plastic body's point functions on the device against the reference.

* stress_invariant, mohcouf and mohcouq on the 600 records the reference's
  own compiled routines evaluated (tests/golden/constitutive.npz): the
  invariants sigm and dsbar (arithmetic and sqrt only) bit for bit; theta,
  f and dq1..dq3, which pass through asin, sin, cos and tan, to within the
  last bits those functions differ in (f relative to the size of its terms,
  since it cancels near the yield surface).
* The device asin, sin, cos and tan against the host libm (the functions
  the reference calls) on the argument ranges of the point math: how many
  results differ, and by how many ulps at most.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import golden
from specfem3d_geotech_amd import _lib as L

pytestmark = pytest.mark.gpu


def ulps(a, b):
    """Distance in units in the last place (same sign, finite)."""
    ia = np.asarray(a, np.float64).view(np.int64)
    ib = np.asarray(b, np.float64).view(np.int64)
    return np.abs(ia - ib)


def device_math(ctx, fn, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    L.check(L.lib().sgeo_math(ctx.h, fn, x.size, L.ptr(x), L.ptr(y)), "sgeo_math")
    return y


def test_constitutive_records_match_reference(ctx):
    z = golden("constitutive")
    rec = np.ascontiguousarray(z["inputs"], np.float64)
    n = rec.shape[0]
    out = np.zeros((n, 7))
    L.check(L.lib().sgeo_constitutive(ctx.h, n, L.ptr(rec), L.ptr(out)),
            "sgeo_constitutive")
    ref = z["outputs"][:, :7]
    names = ["sigm", "dsbar", "theta", "f", "dq1", "dq2", "dq3"]
    for k, nm in enumerate(names):
        d = ulps(out[:, k], ref[:, k])
        print(f"{nm}: {int((d != 0).sum())} of {n} differ, max {int(d.max())} ulp")
    assert np.array_equal(out[:, 0], ref[:, 0])
    assert np.array_equal(out[:, 1], ref[:, 1])
    # theta = asin(.)/3, dq1 = sin(psi) and what follows from them: the last
    # bits of the device's asin, sin, cos and tan
    for k in (2, 4, 5, 6):
        rel = np.abs(out[:, k] - ref[:, k]) / np.maximum(np.abs(ref[:, k]), 1e-300)
        assert rel.max() <= 1e-12, names[k]
    # f = sin(phi) sigm + dsbar (...) - c cos(phi) cancels near the yield
    # surface: its differences are last bits of the terms it is made of
    phi, c = rec[:, 6], rec[:, 7]
    scale = np.abs(ref[:, 0]) + np.abs(ref[:, 1]) + np.abs(c)
    assert (np.abs(out[:, 3] - ref[:, 3]) <= 1e-14 * scale).all()


@pytest.mark.parametrize("fn,lo,hi", [(0, -1.0, 1.0), (1, -0.6, 0.6), (2, -0.6, 0.6),
                                      (3, -1.5, 1.5), (1, 0.0, 0.8), (2, 0.0, 0.8)])
def test_device_math_against_host_libm(ctx, fn, lo, hi):
    x = np.random.default_rng(fn * 7 + int(hi * 10)).uniform(lo, hi, 200000)
    y = device_math(ctx, fn, x)
    ref = [np.arcsin, np.sin, np.cos, np.tan][fn](x)
    d = ulps(y, ref)
    frac = float((d != 0).mean())
    print(f"{['asin', 'sin', 'cos', 'tan'][fn]} on [{lo}, {hi}]: {frac:.4%} differ, "
          f"max {int(d.max())} ulp")
    assert d.max() <= 4
