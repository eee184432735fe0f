"""CPU check of the exact three-term bf16 split that the ensemble's two-byte
product applies to the FP32 slip rates (split3 in interact_amd/csrc/ensemble.hip):
h = bf16(f), m = bf16(f - h), l = bf16(f - h - m), all rounded to nearest even,
gives h + m + l == f exactly for every finite FP32 value whose residuals stay
in the normal range, which is what makes each tile's product the bf16 entries
times the FP32 slip rates.  The rounding is emulated on the FP32 bit patterns,
as the device code does; two terms (a plausible shortcut) do not suffice."""
import numpy as np


def _bf16_rn(f):
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u << 16).astype(np.uint32).view(np.float32)


def _split(f, rnd):
    h = rnd(f)
    r1 = (f - h).astype(np.float32)
    m = rnd(r1)
    r2 = (r1 - m).astype(np.float32)
    l = rnd(r2)
    return h, m, l, r2


def test_three_bf16_terms_sum_to_the_fp32_value():
    rng = np.random.default_rng(7)
    mant = rng.integers(0, 1 << 23, size=200000, dtype=np.uint32)
    expo = rng.integers(60, 200, size=mant.size, dtype=np.uint32)   # 2^-67 .. 2^72
    sign = rng.integers(0, 2, size=mant.size, dtype=np.uint32) << 31
    f = (sign | (expo << 23) | mant).view(np.float32)
    # slip rates of the cycles: plate rate to seismic, both signs, and ties
    edge = np.array([1e-9, -1e-9, 3.17e-10, 1.0, -2.5, 0.9160873, 1e-30,
                     np.float32(1.0) + np.float32(2.0 ** -8), np.float32(1.0) + np.float32(3 * 2.0 ** -9)],
                    dtype=np.float32)
    f = np.concatenate([f, edge])
    h, m, l, r2 = _split(f, _bf16_rn)
    assert np.array_equal(l, r2)                     # the last residual is a bf16 value
    total = h.astype(np.float64) + m.astype(np.float64) + l.astype(np.float64)
    assert np.array_equal(total, f.astype(np.float64))
    # each term is a bf16 value (low 16 bits zero)
    for t in (h, m, l):
        assert not np.any(t.view(np.uint32) & 0xFFFF)
    # two terms leave a residual for most values (16 of the 24 significant bits)
    two = h.astype(np.float64) + m.astype(np.float64)
    assert np.mean(two != f.astype(np.float64)) > 0.9
