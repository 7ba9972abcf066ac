"""rocFFT as an independent cross-check of the hand-written demag FFT.

The field is formed a second way, without any of the library's FFT code:
the oracle's Newell tensor planes (calt, 2d_i4/demag_fft.cu:27-91) and the
magnetisation are transformed by rocFFT's real-to-complex 2D FFT (ROCm's
library, double precision), multiplied by numpy, transformed back by
rocFFT's complex-to-real FFT and scaled by strength / (Px Py); the zero-
padded grid keeps the first nx x ny outputs (displacement 0 at index 0,
the never-used displacement -n cleared), the periodic grid applies the
reference's centring and shift (demag_fft.cu:143-161).  Demag_Apply on the
same input must agree to 1e-12 relative per component -- including the
1536 x 512 padded workload grid (3072 x 1024 transforms).
"""
import ctypes as C

import numpy as np
import pytest

import llg_amd as la
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-12

_RF = None


def _rocfft():
    global _RF
    if _RF is None:
        L = C.CDLL("/opt/rocm/lib/librocfft.so")
        L.rocfft_setup.restype = C.c_int
        L.rocfft_plan_create.restype = C.c_int
        L.rocfft_plan_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int,
                                         C.c_size_t, C.POINTER(C.c_size_t), C.c_size_t,
                                         C.c_void_p]
        L.rocfft_execute.restype = C.c_int
        L.rocfft_execute.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                     C.c_void_p]
        L.rocfft_plan_destroy.argtypes = [C.c_void_p]
        assert L.rocfft_setup() == 0
        _RF = L
    return _RF


def _dev(v):
    return C.cast(la.lib().N_VGetDeviceArrayPointer_Hip(v.v), C.c_void_p)


def _fft2(real, Px, Py, inverse=False, spec=None):
    """batched 2D real FFT on the device through rocFFT: real (B, Py, Px)
    -> Hermitian (B, Py, Px/2 + 1), or the inverse (unnormalised)"""
    RF = _rocfft()
    H = Px // 2 + 1
    B = real.shape[0] if not inverse else spec.shape[0]
    vr = la.HipVector(B * Px * Py, None, None if inverse else real)
    vc = la.HipVector(2 * B * Py * H, None,
                      None if not inverse else np.ascontiguousarray(spec).view(np.float64).ravel())
    la.lib().LLG_DeviceSynchronize()
    plan = C.c_void_p()
    lengths = (C.c_size_t * 2)(Px, Py)          # fastest-varying first
    kind = 3 if inverse else 2                  # real_inverse / real_forward
    assert RF.rocfft_plan_create(C.byref(plan), 1, kind, 1, 2, lengths, B, None) == 0
    src, dst = (vc, vr) if inverse else (vr, vc)
    ib, ob = (C.c_void_p * 1)(_dev(src)), (C.c_void_p * 1)(_dev(dst))
    assert RF.rocfft_execute(plan, ib, ob, None) == 0
    la.lib().LLG_DeviceSynchronize()
    RF.rocfft_plan_destroy(plan)
    if inverse:
        return vr.get().reshape(B, Py, Px)
    return vc.get().view(np.complex128).reshape(B, Py, H)


def _field_rocfft(y, nx, ny, strength, padded):
    Px, Py = (2 * nx, 2 * ny) if padded else (nx, ny)
    T = orc.calt(1.0, Px, Py)                   # N(i - Px/2, j - Py/2)
    K = T[[0, 1, 4, 8]]                         # xx, xy, yy, zz
    if padded:
        K = np.fft.ifftshift(K, axes=(1, 2)).copy()   # displacement 0 at index 0
        K[:, Py // 2, :] = 0.0                   # displacement -ny, -nx never used
        K[:, :, Px // 2] = 0.0
    m = np.zeros((3, Py, Px))
    m[:, :ny, :nx] = y.reshape(3, ny, nx)
    Kh = _fft2(K, Px, Py)
    Mh = _fft2(m, Px, Py)
    Hh = np.stack([Kh[0] * Mh[0] + Kh[1] * Mh[1], Kh[1] * Mh[0] + Kh[2] * Mh[1], Kh[3] * Mh[2]])
    h = _fft2(None, Px, Py, inverse=True, spec=Hh) * (strength / (Px * Py))
    if not padded:
        # reference centring: output index i holds displacement sums shifted by n/2
        h = np.roll(np.roll(h, -(Py // 2), axis=1), -(Px // 2), axis=2)
    return h[:, :ny, :nx].reshape(3, -1)


@pytest.mark.parametrize("nx,ny,padded", [(1536, 512, True), (512, 512, False), (96, 40, True)],
                         ids=["1536x512_padded", "512x512_periodic", "96x40_padded"])
def test_demag_field_matches_rocfft_pipeline(nx, ny, padded):
    y = np.random.default_rng(nx + ny).standard_normal(3 * nx * ny)
    strength = 4.0
    print(f"[gpu] rocFFT cross-check {nx}x{ny} padded={padded}", flush=True)
    h_ref = _field_rocfft(y, nx, ny, strength, padded)
    L = la.lib()
    d = (L.Demag_InitPadded if padded else L.Demag_Init)(nx, ny, 1.0, strength)
    vy = la.HipVector(3 * nx * ny, None, y)
    vh = la.HipVector(3 * nx * ny, None)
    L.Demag_SetStream(d, L.N_VGetStream_Hip(vy.v))
    L.Demag_Apply(d, L.N_VGetDeviceArrayPointer_Hip(vy.v), L.N_VGetDeviceArrayPointer_Hip(vh.v))
    h = vh.get().reshape(3, -1)
    L.Demag_Destroy(d)
    errs = [np.max(np.abs(h[k] - h_ref[k])) / np.max(np.abs(h_ref[k])) for k in range(3)]
    print(f"[gpu] rocFFT cross-check {nx}x{ny}: rel err per component "
          f"{['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < TOL
