"""Host-side format helpers used as checkers, pinned against independent
implementations (torch's float8_e4m3fn and bfloat16 casts) and published
known answers (SplitMix64)."""
import numpy as np
import torch

from tests.helpers import bf16_bits_to_f32, e4m3_table, splitmix64_stream, uniform_host


def test_e4m3_table_matches_torch_decode():
    t = e4m3_table()
    ref = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).to(torch.float64).numpy()
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(t), nan)
    assert np.array_equal(t[~nan], ref[~nan])
    assert t[0x38] == 1.0 and t[0x7E] == 448.0 and t[0x01] == 2.0 ** -9 and t[0x08] == 2.0 ** -6


def test_torch_e4m3_encode_known_values():
    # round to nearest even on 3 mantissa bits: ties 1.0625 -> 1.0, 1.1875 -> 1.25
    vals = torch.tensor([1.0, 448.0, 0.5, 2.0 ** -9, 2.0 ** -6, 1.0625, 1.1875, -3.5], dtype=torch.float32)
    got = vals.to(torch.float8_e4m3fn).view(torch.uint8).tolist()
    assert got == [0x38, 0x7E, 0x30, 0x01, 0x08, 0x38, 0x3A, 0xC6]


def test_bf16_bits_roundtrip_torch():
    x = torch.tensor([1.0, -2.5, 3.1415926, 1e-30, 65504.0], dtype=torch.float32)
    b = x.to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(bf16_bits_to_f32(b), x.to(torch.bfloat16).to(torch.float32).numpy())


def test_splitmix64_known_answers():
    # SplitMix64 from state 0 (Steele, Lea, Flood 2014; reference outputs)
    z = splitmix64_stream(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform_host_uses_the_splitmix_finaliser():
    # element i is the SplitMix64 finaliser of (key ^ i) + G; for seed 0,
    # stream 0, element 0 is therefore the first SplitMix64 output.
    u = uniform_host(1, 0, 0, lo=0.0, hi=1.0, dtype=np.float64)
    z = splitmix64_stream(0, 1)
    assert u[0] == float(int(z[0]) >> 40) / 2.0 ** 24
    v = uniform_host(1000, 5, 2)
    assert v.dtype == np.float32 and v.min() >= -1.0 and v.max() < 1.0
