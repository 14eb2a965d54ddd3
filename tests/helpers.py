"""Host-side helpers shared by the tests and the benchmark checks."""
import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform_host(n, seed, stream, lo=-1.0, hi=1.0, dtype=np.float32, idx=None):
    """Host restatement of the device generator gat_fill_uniform
    (gatlas_amd/csrc/kernels/fill.hip): element i of stream s is
    lo + (hi-lo) * (mix64((seed*G ^ s<<48 ^ i) + G) >> 40) * 2^-24."""
    with np.errstate(over="ignore"):
        key = np.array([seed], dtype=np.uint64) * _G ^ (np.uint64(stream) << np.uint64(48))
        i = np.arange(n, dtype=np.uint64) if idx is None else np.asarray(idx, dtype=np.uint64)
        z = _mix64((key ^ i) + _G)
    u = (z >> np.uint64(40)).astype(np.float64) * (1.0 / 16777216.0)
    return (lo + (hi - lo) * u).astype(dtype)


def splitmix64_stream(seed, n):
    """Plain SplitMix64 outputs (x += G; out = mix64(x)) for known-answer checks."""
    with np.errstate(over="ignore"):
        x = np.uint64(seed) + _G * np.arange(1, n + 1, dtype=np.uint64)
        return _mix64(x)


def op(M_, rows, cols, trans):
    """op(X) as a (rows x cols) view of the flat row-major storage."""
    return M_.reshape(cols, rows).T if trans else M_.reshape(rows, cols)
