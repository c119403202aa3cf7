"""The oracle's Q profiles against the reference's own src/absorb.h, compiled from where it lies
into oracle/_ref/libabsorb_ref.so (oracle/Makefile, target `ref`; built only where the reference
tree is present, skipped elsewhere).  Bit for bit, 3D and 2D, with and without a free surface,
at the propagator tests' parameters and at wide sponges.  The GPU tests compare the libraries'
Q arrays with the oracle's, so this pins them too."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as o3
from oracle import oracle2d as o2

F32 = np.float32
REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libabsorb_ref.so")

CASES = [
    # (nsponge, qMin, qInterior, freqQ, dt)
    (6, 0.5, 80.0, 10.0, 1e-3),
    (5, 1.0, 100.0, 10.0, 1e-3),
    (20, 0.1, 100.0, 10.0, 1e-3),
    (50, 0.1, 100.0, 10.0, 1e-3),
    (40, 0.1, 100.0, 5.0, 5e-4),
    (13, 2.0, 30.0, 12.5, 2e-3),
]


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF):
        pytest.skip("reference absorb.h build not present (needs the reference tree)")
    lib = ctypes.CDLL(REF)
    fp = ctypes.POINTER(ctypes.c_float)
    L, Fl = ctypes.c_long, ctypes.c_float
    lib.ref_setup_q_3d.argtypes = [L, L, L, L, L, Fl, Fl, Fl, Fl, fp]
    lib.ref_setup_q_2d.argtypes = [L, L, L, L, Fl, Fl, Fl, Fl, fp]
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("fs", [False, True])
def test_3d_q_profile_bit_exact_vs_reference(ref, case, fs):
    ns, qmin, qint, fq, dt = case
    nx, ny, nz = 2 * ns + 7, 2 * ns + 4, 2 * ns + 9
    want = np.zeros((nx, ny, nz), F32)
    ref.ref_setup_q_3d(int(fs), nx, ny, nz, ns, F32(dt), F32(fq), F32(qmin), F32(qint), _p(want))
    got = o3.setup_q(fs, nx, ny, nz, ns, dt, fq, qmin, qint)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("fs", [False, True])
def test_2d_q_profile_bit_exact_vs_reference(ref, case, fs):
    ns, qmin, qint, fq, dt = case
    nx, nz = 2 * ns + 11, 2 * ns + 5
    want = np.zeros((nx, nz), F32)
    ref.ref_setup_q_2d(int(fs), nx, nz, ns, F32(dt), F32(fq), F32(qmin), F32(qint), _p(want))
    got = o2.setup_q(fs, nx, nz, ns, dt, fq, qmin, qint)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_2d_ramp_uses_the_float_log(ref):
    """The case that separates the float and double logs of the 2D ramp (1,544 of 30,000 cells
    differ at these parameters when the log is taken in double)."""
    ns, qmin, qint, fq, dt = 6, 0.5, 80.0, 10.0, 1e-3
    nx, nz = 150, 200
    want = np.zeros((nx, nz), F32)
    ref.ref_setup_q_2d(0, nx, nz, ns, F32(dt), F32(fq), F32(qmin), F32(qint), _p(want))
    got = o2.setup_q(False, nx, nz, ns, dt, fq, qmin, qint)
    assert np.array_equal(got, want)
    # the double-log ramp differs from the reference somewhere in the sponge
    lq = np.log(np.float64(qmin)) + np.arange(ns) / (ns - 1) * (np.log(np.float64(qint)) - np.log(np.float64(qmin)))
    qprof_double = np.exp(lq.astype(F32))
    lq_f = np.float64(np.log(F32(qmin))) + np.arange(ns) / (ns - 1) * (np.float64(np.log(F32(qint))) -
                                                                      np.float64(np.log(F32(qmin))))
    qprof_float = np.exp(lq_f.astype(F32))
    assert not np.array_equal(qprof_double, qprof_float)


@pytest.mark.parametrize("fs", [False, True])
def test_unphysical_q_clears_what_the_reference_clears(ref, fs):
    """qMin, qInterior < FLT_EPSILON: the reference prints a warning and clears sizeof(float *)
    bytes, the array's first two values (src/absorb.h:44,107), leaving the rest as it was; a freshly
    allocated (zeroed) array therefore stays all zero, the case test/absorb.jl:85-99 checks."""
    nx, ny, nz = 13, 11, 15
    want = np.ones((nx, ny, nz), F32)
    ref.ref_setup_q_3d(int(fs), nx, ny, nz, 5, F32(1e-3), F32(10.0), F32(0.0), F32(0.0), _p(want))
    got = o3.setup_q(fs, nx, ny, nz, 5, 1e-3, 10.0, 0.0, 0.0, out=np.ones((nx, ny, nz), F32))
    assert np.array_equal(got, want)
    assert not want.ravel()[:2].any() and (want.ravel()[2:] == 1).all()
    want2 = np.ones((nx, nz), F32)
    ref.ref_setup_q_2d(int(fs), nx, nz, 5, F32(1e-3), F32(10.0), F32(0.0), F32(0.0), _p(want2))
    got2 = o2.setup_q(fs, nx, nz, 5, 1e-3, 10.0, 0.0, 0.0, out=np.ones((nx, nz), F32))
    assert np.array_equal(got2, want2)
