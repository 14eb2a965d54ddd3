"""Pins the CPU oracle (oracle/oracle.c) before anything is judged by it.

Each check is chosen so that a plausible slip in the oracle (wrong index or
transpose, a dropped term, reversed summation, a fused multiply-add where
the reference rounds twice, a wrong LCG constant) fails it.
"""
import ctypes
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as ora


# ---------------------------------------------------------------- drand48
def _glibc_drand48(n, seed=None):
    """drand48() values from glibc in a fresh process (unseeded, or srand48)."""
    code = ("import ctypes,sys\n"
            "L=ctypes.CDLL('libc.so.6'); L.drand48.restype=ctypes.c_double\n"
            "s=%r\n"
            "if s is not None: L.srand48(ctypes.c_long(s))\n"
            "print(' '.join(repr(L.drand48()) for _ in range(%d)))\n" % (seed, n))
    out = subprocess.check_output([sys.executable, "-c", code], text=True)
    return [float(v) for v in out.split()]


def test_lcg_matches_glibc_unseeded():
    # The reference never seeds drand48 (OCLAppUtil.hpp:28-33): glibc's
    # zero-initialised state is the stream its paranoid inputs come from.
    want = _glibc_drand48(64)
    r = ora.Drand48(0)
    got = [r.next() for _ in range(64)]
    assert got == want
    assert got[0] == 11.0 / 2.0 ** 48  # (a*0 + c) / 2^48


def test_lcg_matches_glibc_seeded():
    seed = 12345
    want = _glibc_drand48(32, seed)
    r = ora.Drand48((seed << 16) | 0x330E)  # srand48: high 32 bits = seed, low 16 = 0x330E
    assert [r.next() for _ in range(32)] == want


def test_posrand_skips_exact_zero():
    # Find X with (a*X + c) mod 2^48 == 0; the draw after X is exactly 0.0,
    # which posrand must reject and replace by the following draw.
    a, c, m = 0x5DEECE66D, 0xB, 1 << 48
    x = (-c * pow(a, -1, m)) % m
    r = ora.Drand48(x)
    v = r.posrand(np.float64)
    r2 = ora.Drand48(0)  # state after the zero draw is 0
    assert v == r2.next()
    assert v != 0.0


def test_fillrand_float_converts_each_draw():
    r1, r2 = ora.Drand48(0), ora.Drand48(0)
    f = r1.fillrand(100, np.float32)
    d = np.array([r2.next() for _ in range(100)])
    assert np.array_equal(f, d.astype(np.float32))


# ------------------------------------------------- reference host products
def _int_mats(rng, M, N, K):
    A = rng.integers(-8, 9, size=M * K).astype(np.float32)
    B = rng.integers(-8, 9, size=K * N).astype(np.float32)
    return A, B


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_ref_matmul_integer_exact(ta, tb):
    # Small integers: every float operation is exact, so the result must be
    # the exact integer product for the stated storage layout.
    rng = np.random.default_rng(1)
    M, N, K = 12, 8, 20
    A, B = _int_mats(rng, M, N, K)
    opA = A.reshape(K, M).T if ta else A.reshape(M, K)
    opB = B.reshape(N, K).T if tb else B.reshape(K, N)
    want = opA.astype(np.int64) @ opB.astype(np.int64)
    got = ora.ref_matmul(A, B, M, N, K, transA=ta, transB=tb)
    assert np.array_equal(got.reshape(M, N), want.astype(np.float32))
    # GEMM form alpha*C + beta*Cold with exact small values
    Cold = rng.integers(-4, 5, size=M * N).astype(np.float32)
    got = ora.ref_matmul(A, B, M, N, K, transA=ta, transB=tb, gemm=True, alpha=2.0, beta=-3.0, Cold=Cold)
    assert np.array_equal(got.reshape(M, N), (2 * want - 3 * Cold.reshape(M, N)).astype(np.float32))


def test_ref_matmul_packed_are_independent():
    rng = np.random.default_rng(2)
    M, N, K, P = 4, 8, 4, 3
    A = rng.random(P * M * K, dtype=np.float32)
    B = rng.random(P * K * N, dtype=np.float32)
    got = ora.ref_matmul(A, B, M, N, K, packed=P)
    for p in range(P):
        one = ora.ref_matmul(A[p * M * K:(p + 1) * M * K], B[p * K * N:(p + 1) * K * N], M, N, K)
        assert np.array_equal(got[p * M * N:(p + 1) * M * N], one)


def test_ref_matmul_sums_k_ascending():
    # 2^24 + 1 rounds back to 2^24 in float: k-ascending gives 0, any order
    # that adds the 1 last (or pairs the +-2^24 first) gives 1.
    A = np.ones(3, dtype=np.float32)
    B = np.array([2.0 ** 24, 1.0, -2.0 ** 24], dtype=np.float32)
    assert ora.ref_matmul(A, B, 1, 1, 3)[0] == 0.0
    assert ora.ref_matmul(A, B[::-1].copy(), 1, 1, 3)[0] == 1.0


def test_ref_matmul_rounds_multiply_and_add_separately():
    # a*a = 1 + 2^-11 + 2^-24 rounds (tie to even) to 1 + 2^-11, so the
    # unfused chain gives exactly 0; a fused multiply-add would give 2^-24.
    a = np.float32(1 + 2.0 ** -12)
    A = np.array([a, 1.0], dtype=np.float32)
    B = np.array([a, -(1 + 2.0 ** -11)], dtype=np.float32)
    assert ora.ref_matmul(A, B, 1, 1, 2)[0] == 0.0
    # epilogue: alpha*acc + beta*c, also two roundings
    Cold = np.array([-(1 + 2.0 ** -11)], dtype=np.float32)
    got = ora.ref_matmul(np.array([1.0], np.float32), np.array([1.0], np.float32), 1, 1, 1, gemm=True,
                         alpha=float(a), beta=1.0, Cold=Cold * 0 + np.float32(-(1 + 2.0 ** -11)))
    # alpha*1 = a exactly; a + (-(1+2^-11)) = 2^-12 - 2^-11 exact
    assert got[0] == np.float32(a) - np.float32(1 + 2.0 ** -11)


def test_ref_matmul_double_integer_exact():
    rng = np.random.default_rng(3)
    M, N, K = 6, 10, 7
    A = rng.integers(-100, 100, size=M * K).astype(np.float64)
    B = rng.integers(-100, 100, size=K * N).astype(np.float64)
    got = ora.ref_matmul(A, B, M, N, K, transA=True, transB=False)
    want = A.reshape(K, M).T @ B.reshape(K, N)
    assert np.array_equal(got.reshape(M, N), want)


@pytest.mark.parametrize("ta", [0, 1])
def test_ref_matvec_integer_exact(ta):
    rng = np.random.default_rng(4)
    M, N = 16, 24
    A = rng.integers(-8, 9, size=M * N).astype(np.float32)
    x = rng.integers(-8, 9, size=N).astype(np.float32)
    opA = A.reshape(N, M).T if ta else A.reshape(M, N)
    want = opA.astype(np.int64) @ x.astype(np.int64)
    assert np.array_equal(ora.ref_matvec(A, x, M, N, transA=ta), want.astype(np.float32))
    y0 = rng.integers(-8, 9, size=M).astype(np.float32)
    got = ora.ref_matvec(A, x, M, N, transA=ta, gemv=True, alpha=3.0, beta=2.0, yold=y0)
    assert np.array_equal(got, (3 * want + 2 * y0).astype(np.float32))


def test_ref_matvec_sums_j_ascending():
    A = np.ones(3, dtype=np.float32)
    x = np.array([2.0 ** 24, 1.0, -2.0 ** 24], dtype=np.float32)
    assert ora.ref_matvec(A, x, 1, 3)[0] == 0.0


def test_ref_saxpy_matches_numpy_two_roundings():
    rng = np.random.default_rng(5)
    x = rng.random(1000, dtype=np.float32)
    y = rng.random(1000, dtype=np.float32)
    alpha = np.float32(0.7)
    want = (alpha * x).astype(np.float32) + y  # numpy: each op correctly rounded
    assert np.array_equal(ora.ref_saxpy(alpha, x, y), want)


def test_absdiff_and_paranoid_rule():
    a = np.array([1.0, 2.0], np.float32)
    b = np.array([1.5, 2.0], np.float32)
    assert ora.absdiff(a, b) == 0.5
    assert not ora.paranoid_pass(a, b)          # 0.5 < 1/2 is false
    assert ora.paranoid_pass(a, np.array([1.25, 2.0], np.float32))
    assert ora.check_array(np.full(5, 7.0, np.float32), 7.0)
    assert not ora.check_array(np.array([7.0, 7.0, 6.0], np.float32), 7.0)


# ------------------------------------------------------------- FP64 oracle
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_gemm64_matches_numpy(ta, tb):
    rng = np.random.default_rng(6)
    M, N, K = 37, 29, 41
    A = (rng.random(M * K) * 2 - 1).astype(np.float32)
    B = (rng.random(K * N) * 2 - 1).astype(np.float32)
    C0 = (rng.random(M * N) * 2 - 1).astype(np.float32)
    opA = (A.reshape(K, M).T if ta else A.reshape(M, K)).astype(np.float64)
    opB = (B.reshape(N, K).T if tb else B.reshape(K, N)).astype(np.float64)
    want = 1.5 * opA @ opB - 0.5 * C0.reshape(M, N).astype(np.float64)
    mag_want = 1.5 * np.abs(opA) @ np.abs(opB) + 0.5 * np.abs(C0.reshape(M, N))
    C64, MAG = ora.gemm64(A, B, M, N, K, transA=ta, transB=tb, alpha=1.5, beta=-0.5, C0=C0)
    np.testing.assert_allclose(C64, want, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(MAG, mag_want, rtol=1e-13, atol=1e-13)
    # float64 inputs path
    C64d, _ = ora.gemm64(A.astype(np.float64), B.astype(np.float64), M, N, K, transA=ta, transB=tb, alpha=1.5,
                         beta=-0.5, C0=C0.astype(np.float64))
    np.testing.assert_allclose(C64d, want, rtol=1e-13, atol=1e-13)
    # row sampling returns exactly the selected rows
    rows = [0, 5, 36]
    Cs, _ = ora.gemm64(A, B, M, N, K, transA=ta, transB=tb, alpha=1.5, beta=-0.5, C0=C0, rows=rows)
    np.testing.assert_array_equal(Cs, C64[rows])


@pytest.mark.parametrize("ta", [0, 1])
def test_gemv64_matches_numpy(ta):
    rng = np.random.default_rng(7)
    M, N = 33, 50
    A = (rng.random(M * N) * 2 - 1).astype(np.float32)
    x = (rng.random(N) * 2 - 1).astype(np.float32)
    y0 = (rng.random(M) * 2 - 1).astype(np.float32)
    opA = (A.reshape(N, M).T if ta else A.reshape(M, N)).astype(np.float64)
    want = 2.0 * opA @ x.astype(np.float64) + 1.0 * y0
    y64, mag = ora.gemv64(A, x, M, N, transA=ta, alpha=2.0, beta=1.0, y0=y0)
    np.testing.assert_allclose(y64, want, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(mag, 2.0 * np.abs(opA) @ np.abs(x.astype(np.float64)) + np.abs(y0), rtol=1e-13)


def test_error_metrics_hand_case():
    got = np.array([1.0, 2.0, 4.0], np.float32)
    ref = np.array([1.0, 2.5, 4.0])
    mag = np.array([1.0, 5.0, 4.0])
    cw, nw, amax = ora.error_metrics(got, ref, mag)
    assert cw == pytest.approx(0.1)
    assert nw == pytest.approx(0.5 / np.sqrt(1 + 6.25 + 16))
    assert amax == pytest.approx(0.5)


def test_fp32_reference_chain_within_higham_bound():
    # |fl(sum) - exact| <= K*u*sum|a||b| (u = 2^-24) for the k-ascending chain.
    rng = np.random.default_rng(8)
    M, N, K = 16, 16, 512
    A = (rng.random(M * K) * 2 - 1).astype(np.float32)
    B = (rng.random(K * N) * 2 - 1).astype(np.float32)
    C = ora.ref_matmul(A, B, M, N, K).reshape(M, N)
    C64, MAG = ora.gemm64(A, B, M, N, K)
    assert np.all(np.abs(C - C64) <= K * 2.0 ** -24 * MAG * 1.0001)


def test_port_equals_reference_chain_bitwise():
    # The CPU baseline port reorders loops but not any element's k chain.
    rng = np.random.default_rng(9)
    M, N, K = 24, 40, 64
    for ta, tb in [(0, 0), (1, 0), (0, 1), (1, 1)]:
        A = rng.random(M * K, dtype=np.float32)
        B = rng.random(K * N, dtype=np.float32)
        Cold = rng.random(M * N, dtype=np.float32)
        want = ora.ref_matmul(A, B, M, N, K, transA=ta, transB=tb, gemm=True, alpha=0.3, beta=0.9, Cold=Cold)
        C = Cold.copy()
        ora.port_matmul_rows(A, B, C, M, N, K, 0, M, gemm=True, alpha=0.3, beta=0.9, transA=ta, transB=tb)
        assert np.array_equal(C, want)
    A = rng.random(M * N, dtype=np.float32)
    x = rng.random(N, dtype=np.float32)
    y = np.zeros(M, np.float32)
    ora.port_matvec_rows(A, x, y, N, 0, M)
    assert np.array_equal(y, ora.ref_matvec(A, x, M, N))


def test_vectorised_port_keeps_every_elements_rounding():
    # The port is built -O3 (vectorised across output elements, per-ISA
    # clones): ragged widths, general forms and the SAXPY port still give the
    # reference chain's bits, where a fused multiply-add or a reassociated
    # sum would not.
    rng = np.random.default_rng(11)
    M, N, K = 37, 1029, 211
    A = (rng.random(M * K, dtype=np.float32) * 2 - 1)
    B = (rng.random(K * N, dtype=np.float32) * 2 - 1)
    Cold = (rng.random(M * N, dtype=np.float32) * 2 - 1)
    want = ora.ref_matmul(A, B, M, N, K, gemm=True, alpha=1.5, beta=-0.5, Cold=Cold)
    C = Cold.copy()
    ora.port_matmul_rows(A, B, C, M, N, K, 0, M, gemm=True, alpha=1.5, beta=-0.5)
    assert np.array_equal(C.view(np.uint32), want.view(np.uint32))
    Av = (rng.random(M * N, dtype=np.float32) * 2 - 1)
    x = (rng.random(N, dtype=np.float32) * 2 - 1)
    y0 = (rng.random(M, dtype=np.float32) * 2 - 1)
    y = y0.copy()
    ora.port_matvec_rows(Av, x, y, N, 0, M, gemv=True, alpha=0.7109375, beta=-1.3125)
    ref = ora.ref_matvec(Av, x, M, N, gemv=True, alpha=0.7109375, beta=-1.3125, yold=y0)
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))
    n = 100003
    xs = (rng.random(n, dtype=np.float32) * 2 - 1)
    ys = (rng.random(n, dtype=np.float32) * 2 - 1)
    z = np.empty(n, np.float32)
    ora.port_saxpy(0.7109375, xs, ys, z)
    want = ora.ref_saxpy(np.float32(0.7109375), xs, ys)
    assert np.array_equal(z.view(np.uint32), want.view(np.uint32))
    fused = (np.float64(0.7109375) * xs.astype(np.float64) + ys).astype(np.float32)
    assert not np.array_equal(fused, want)   # the check can tell one rounding from two


def test_planted_mistake_is_caught():
    # A transposed-operand slip planted in a copy of the oracle's result must
    # fail the integer-exact check used above.
    rng = np.random.default_rng(10)
    M, N, K = 8, 8, 8
    A, B = _int_mats(rng, M, N, K)
    good = ora.ref_matmul(A, B, M, N, K).reshape(M, N)
    planted = ora.ref_matmul(A, B, M, N, K, transA=True).reshape(M, N)  # wrong layout
    want = A.reshape(M, K).astype(np.int64) @ B.reshape(K, N).astype(np.int64)
    assert np.array_equal(good, want)
    assert not np.array_equal(planted, want)
