"""GPU: the device's sin and cos of a particle's yaw against glibc.

The reference computes sin/cos of a float yaw with an unqualified sin(), which
under its GCC 5 toolchain is C's double sin, rounded once to float where it
is stored (ParticleFilter.cpp:110-111, Grid3d.cpp:139-149).  The device does
the same with its own double sin/cos.  Predict exposes them exactly: with
x = y = 0, delta = (1, 0, 0, 0) and zero deviations, x becomes cos(a) and y
becomes sin(a) (ParticleFilter.cpp:108-118: x += ca*1 - sa*0).  Over 2e6
yaws (and the values where sin or cos is 0 or 1) the device must equal glibc's
double routines rounded to float bit for bit; the rate at which glibc's
sinf/cosf (the reading a GCC 6 or later build would take) differ is printed.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F32 = np.float32


def test_device_sin_cos_equal_glibc_double_rounded(gpu):
    libm = ctypes.CDLL("libm.so.6")
    libm.sin.restype = libm.cos.restype = ctypes.c_double
    libm.sin.argtypes = libm.cos.argtypes = [ctypes.c_double]
    libm.sinf.restype = libm.cosf.restype = ctypes.c_float
    libm.sinf.argtypes = libm.cosf.argtypes = [ctypes.c_float]
    rng = np.random.default_rng(23)
    special = np.array([np.pi, -np.pi, np.pi / 2, -np.pi / 2, np.pi / 4, 1e-30, 1e-3, 3.0, 100.0, 1000.5, -12345.6],
                       np.float64)
    yaw = np.concatenate([rng.uniform(-np.pi, np.pi, 2000000), rng.uniform(-50.0, 50.0, 200000),
                          special]).astype(F32)
    n = yaw.size
    p = np.zeros((n, 7), F32)
    p[:, 3] = yaw
    pf = gpu.ParticleFilter()
    pf.set_particles(p)
    pf.predict(0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, np.zeros((n, 4), F32))
    got = pf.particles()
    sin_ref = np.array([libm.sin(float(a)) for a in yaw], np.float64).astype(F32)
    cos_ref = np.array([libm.cos(float(a)) for a in yaw], np.float64).astype(F32)
    bad_sin = int((got[:, 1].view(np.int32) != sin_ref.view(np.int32)).sum())
    bad_cos = int((got[:, 0].view(np.int32) != cos_ref.view(np.int32)).sum())
    k = 200000
    sinf = np.array([libm.sinf(float(a)) for a in yaw[:k]], F32)
    cosf = np.array([libm.cosf(float(a)) for a in yaw[:k]], F32)
    print("device vs glibc double rounded: sin %d, cos %d of %d yaws differ; glibc sinf/cosf differ from it on "
          "%.2e / %.2e of yaws" % (bad_sin, bad_cos, n, (sinf.view(np.int32) != sin_ref[:k].view(np.int32)).mean(),
                                   (cosf.view(np.int32) != cos_ref[:k].view(np.int32)).mean()), flush=True)
    assert bad_sin == 0 and bad_cos == 0
