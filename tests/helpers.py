"""Shared numerics for the parity tests."""
import numpy as np

# max |error| bound for FP32 paths, relative to the input RMS
RMS_TOL = 1e-4


def rms(x) -> float:
    x = np.asarray(x, dtype=np.complex128)
    return float(np.sqrt(np.mean(np.abs(x) ** 2))) if x.size else 0.0


def max_err(a, b) -> float:
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    assert a.shape == b.shape, f"shape {a.shape} vs {b.shape}"
    return float(np.max(np.abs(a - b))) if a.size else 0.0


def assert_close_rms(got, want, x_rms, tol=RMS_TOL, what=""):
    e = max_err(got, want)
    assert e <= tol * x_rms, f"{what}: max |err| {e:.3e} > {tol:g} * rms {x_rms:.3e}"
    return e


def chunks(n, sizes):
    """Split range(n) by a repeating list of chunk sizes."""
    out, i, k = [], 0, 0
    while i < n:
        m = min(sizes[k % len(sizes)], n - i)
        out.append((i, i + m))
        i += m
        k += 1
    return out
