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


def e4m3_table():
    """Value of every OCP FP8 e4m3 byte (bias 7, no infinities, 0x7F/0xFF NaN):
    exponent 0 -> m * 2^-9, else (1 + m/8) * 2^(e-7)."""
    t = np.zeros(256, dtype=np.float64)
    for b in range(256):
        e, m = (b >> 3) & 15, b & 7
        v = m * 2.0 ** -9 if e == 0 else (1 + m / 8.0) * 2.0 ** (e - 7)
        if (b & 0x7F) == 0x7F:
            v = np.nan
        t[b] = -v if b & 0x80 else v
    return t


def bf16_bits_to_f32(u16):
    """BF16 bit patterns (uint16) -> float32 values."""
    return (np.asarray(u16, dtype=np.uint32) << np.uint32(16)).view(np.float32)


def dequant_fp8(q, scales, rows, cols):
    """q (rows x cols e4m3 bytes) times its 128 x 128 block scales, in FP64."""
    t = e4m3_table()
    v = t[q.reshape(rows, cols)]
    cb = (cols + 127) // 128
    s = np.asarray(scales, dtype=np.float64).reshape(-1, cb)
    return v * np.repeat(np.repeat(s, 128, axis=0)[:rows], 128, axis=1)[:, :cols]


def op(M_, rows, cols, trans):
    """op(X) as a (rows x cols) view of the flat row-major storage."""
    return M_.reshape(cols, rows).T if trans else M_.reshape(rows, cols)
