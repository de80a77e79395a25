"""Independent numpy formulations used to pin the C oracle (test infrastructure only).

They are written from the reference's definitions with library routines — numpy's fftshift (which
equals the reference's ceil-mid roll, signal_complex.h:24-29), pocketfft (np.fft.fft) for FFTW's
forward DFT (dft.cpp:26,72), np.interp for the clamped linear interpolation on the implicit
linspace axis (backprojection.cpp:52-65,120) and the textbook Taylor window (Carrara, Goodman and
Majewski, "Spotlight Synthetic Aperture Radar", the form RITSAR implements) — so that a dropped
term, a wrong sign, index or axis in the oracle shows up as a mismatch.
"""
from __future__ import annotations

import math

import numpy as np


def taylor_textbook(num: int, s_l: float, nbar: int, int_ratio: bool) -> np.ndarray:
    """Textbook Taylor weights w(xi) = 1 + 2 sum_m F_m cos(2 pi m xi), normalised to max 1, with
    F_m = (-1)^(m+1)/2 prod_n (1 - m^2 / sigma^2 / (A^2 + (n - 1/2)^2)) / prod_{n != m} (1 - m^2/n^2).
    `int_ratio` floors m^2/n^2 (the Int(32) reading of the reference's F_m_den)."""
    A = float(np.float64(np.arccosh(np.float32(10.0) ** np.float32(s_l / 20.0)))) / math.pi
    sigma2 = nbar ** 2 / (A ** 2 + (nbar - 0.5) ** 2)
    F = []
    for m in range(1, nbar):
        num_ = np.prod([1 - m * m / sigma2 / (A * A + (n - 0.5) ** 2) for n in range(1, nbar)])
        den = 1.0
        for n in range(1, nbar):
            if n == m:
                continue
            ratio = (m * m) // (n * n) if int_ratio else np.float32(m * m) / np.float32(n * n)
            den *= 1 - float(ratio)
        F.append((-1) ** (m + 1) / 2 * num_ / den)
    xi = np.linspace(-0.5, 0.5, num)
    w = 1 + 2 * sum(F[m - 1] * np.cos(2 * math.pi * m * xi) for m in range(1, nbar))
    return w / w.max(), F


def range_profiles_twin(phs, k_r, ws, wp, n_fft, pulse_shift_quirk=True):
    """Q for every pulse: window * |k_r| filter, centred zero pad, fftshift, FFT, fftshift."""
    phs = np.asarray(phs, np.complex64)
    npulses, ns = phs.shape
    filt = np.abs(np.asarray(k_r, np.float32))
    prod = (phs * filt[None, :]).astype(np.complex64)           # f32 product, rounded
    filt_win = prod.astype(np.complex128) * np.outer(wp, ws)     # f64 window
    xpad = (n_fft - ns) // 2                                     # floor division
    padded = np.zeros((npulses, n_fft), np.complex128)
    src_lo, src_hi = max(0, -xpad), min(ns, n_fft - xpad)
    padded[:, xpad + src_lo:xpad + src_hi] = filt_win[:, src_lo:src_hi]
    axes = (0, 1) if pulse_shift_quirk else (1,)
    spec = np.fft.fft(np.fft.fftshift(padded, axes=axes), axis=1)
    return np.fft.fftshift(spec, axes=axes)


def backproject_twin(Q, ns, delta_r, k_c, pos, nr0, pixel_locs, nu, nv):
    """Output image complex128 (nv, nu) from profiles Q (np, N)."""
    Q = np.asarray(Q)
    npulses, n_fft = Q.shape
    lsa = math.floor(-ns * delta_r / 2)
    lsb = math.floor(ns * delta_r / 2)
    xp = np.linspace(lsa, lsb, n_fft)
    r = np.asarray(pixel_locs, np.float64)                       # (3, nu*nv)
    p64 = np.asarray(pos, np.float32).astype(np.float64)         # (np, 3)
    dist = np.sqrt(((r[None, :, :] - p64[:, :, None]) ** 2).sum(axis=1))   # (np, npix)
    dr = np.asarray(nr0, np.float64)[:, None] - dist
    k_c = float(np.float32(k_c))
    img = np.zeros(r.shape[1], np.complex128)
    for p in range(npulses):
        qh = np.interp(dr[p], xp, Q[p].real) + 1j * np.interp(dr[p], xp, Q[p].imag)
        img += qh * np.exp(-1j * k_c * dr[p])
    img *= np.exp(1j * k_c * dr[npulses // 2])
    return img.reshape(nv, nu)[::-1, :]
