"""Seismogram outputs: the reference's sheet text file and the surface
transfer function used to compare against published canyon results.

* write_sheet(): the format of write_sheet_points / write_sheet_step /
  write_sheet_close (io.cpp:844-889): "POINTS = [ x z x z ... ];" with
  %.10lf, then one "%lf " line per sample, then " ];" and "SAMP = %lf ;".
* transfer(): |FFT_NFFT(u)| * dt * 2 at the first frequency bin >= f
  (sheets/obtainsingle.m:2-22), NFFT = 2^nextpow2(L), normalised by the same
  quantity for the incident Ricker wavelet (sheets/transmex.m:1-10, where
  the wavelet is sampled on [-5, 5] s at 0.01 s).
"""
from __future__ import annotations

import math

import numpy as np


def write_sheet(path, points, samples, samp):
    """points (n, 2); samples (nsamp, n, 2) in (x, z) per receiver; samp =
    time between samples."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1)
    with open(path, "w") as f:
        f.write("POINTS = [ ")
        for v in pts:
            f.write("%.10f " % v)
        f.write(" ];")
        f.write("\nFIELD = [ ")
        for row in np.asarray(samples, dtype=np.float64).reshape(len(samples), -1):
            for v in row:
                f.write("%f " % v)
            f.write("\n")
        f.write(" ];\n")
        f.write("SAMP = %f ;" % samp)


def _amplitude(trace, dt, f):
    L = trace.shape[0]
    nfft = 1 << int(math.ceil(math.log2(L)))
    Y = np.fft.rfft(trace, n=nfft, axis=0)
    freqs = (1.0 / dt) / 2 * np.linspace(0, 1, nfft // 2 + 1)
    n = int(np.nonzero(freqs >= f)[0][0])
    return dt * 2 * np.abs(Y[n])


def ricker_amplitude(fpeak, f=1.0):
    t = np.linspace(-5, 5, 1001)
    a = (math.pi * fpeak * t) ** 2
    return float(_amplitude((1 - 2 * a) * np.exp(-a), 0.01, f))


def transfer(traces, dt, fpeak, f=1.0):
    """Surface transfer-function amplitude at f for traces (nsamp, nrecv)."""
    return _amplitude(np.asarray(traces, dtype=np.float64), dt, f) / ricker_amplitude(fpeak, f)
