"""cg_tissue_pseudo_ecg_noalloc (the allocation-free pseudo-ECG) through the C-ABI: its
results equal cg_tissue_pseudo_ecg's bit for bit on two calls in a row, the second on a
smaller tissue with fewer electrodes (the library's fixed scratch keeps the first call's
contents and must not leak them into the second); it refuses more electrodes than its
scratch holds and a per-cell conductivity that is not on the device yet."""
import ctypes as C
import sys

import numpy as np
import pytest

from tests.gpu_cases import rotating_fibre_sigma

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _tissue(gpu, shape, h, sigma, seed):
    coeffs = {"ionicModel": "BuenoOrovio", "tissue": "epicardialCells", "chi": 1.4e5,
              "cm": 0.01, "solutionAlgorithm": "implicit", "conductivity": sigma}
    dom = gpu.Monodomain(gpu.HexMesh(*shape, h), coeffs, dt=2e-5)
    n = int(np.prod(shape))
    dom.set_field(gpu._capi.FIELD_VM, -0.084 + 0.1 * np.random.default_rng(seed).random(n))
    return dom


def _call(L, fn, dom, E):
    pos = np.ascontiguousarray(E, dtype=np.float64)
    out = np.full(len(pos), np.nan)
    rc = getattr(L, fn)(dom.t, len(pos), pos.ctypes.data_as(C.c_void_p),
                        out.ctypes.data_as(C.c_void_p))
    return rc, out


def test_noalloc_pseudo_ecg_equals_allocating_entry(gpu):
    L = gpu._capi.lib()
    rng = np.random.default_rng(11)
    # call 1: per-cell rotating-fibre conductivity, 16 electrodes (the scratch's maximum)
    shape, h = (40, 24, 18), 2e-4
    _say("noalloc ECG call 1: shape %s, 16 electrodes" % (shape,))
    big = _tissue(gpu, shape, h, rotating_fibre_sigma(*shape), 4)
    ext = np.array(shape) * h
    E1 = rng.uniform(-1, 2, size=(16, 3)) * ext
    rc, ref1 = _call(L, "cg_tissue_pseudo_ecg", big, E1)      # uploads the conductivity
    assert rc == 0
    rc, got1 = _call(L, "cg_tissue_pseudo_ecg_noalloc", big, E1)
    assert rc == 0
    assert np.array_equal(got1, ref1), np.abs(got1 - ref1).max()
    # call 2: a smaller tissue (fewer blocks of partial sums) and 3 electrodes, uniform
    shape, h = (10, 8, 6), 2e-4
    _say("noalloc ECG call 2: shape %s, 3 electrodes" % (shape,))
    small = _tissue(gpu, shape, h, [0.17, 0, 0, 0, 0.019, 0, 0, 0, 0.019], 5)
    E2 = rng.uniform(-0.5, 1.5, size=(3, 3)) * np.array(shape) * h
    rc, got2 = _call(L, "cg_tissue_pseudo_ecg_noalloc", small, E2)
    assert rc == 0
    rc, ref2 = _call(L, "cg_tissue_pseudo_ecg", small, E2)
    assert rc == 0
    assert np.array_equal(got2, ref2), np.abs(got2 - ref2).max()
    _say("  equal bit for bit: %s / %s" % (got1[:2], got2))
    # refusals: 17 electrodes; per-cell conductivity not yet on the device
    rc, _ = _call(L, "cg_tissue_pseudo_ecg_noalloc", small, np.zeros((17, 3)))
    assert rc == 1                                             # CG_ERR_ARG
    fresh = _tissue(gpu, (8, 6, 4), h, rotating_fibre_sigma(8, 6, 4), 6)
    rc, _ = _call(L, "cg_tissue_pseudo_ecg_noalloc", fresh, E2)
    assert rc == 9                                             # CG_ERR_STATE
