"""The BF16x3 split emulation the GPU tests compare against, pinned by the
properties that fix it: |x - hi| <= 2^-8 |x|, |x - hi - lo| <= 2^-16 |x|,
bf16 values split exactly, and round-to-nearest-even on ties."""
import numpy as np

from bf16_emulation import bf16_rne, bf16x3_split


def test_split_accuracy_and_exact_values():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(1000) * np.exp(rng.uniform(-20, 20, 1000))
    hi, lo = bf16x3_split(x)
    assert np.all(np.abs(x - hi) <= np.abs(x) * 2.0 ** -8)
    assert np.all(np.abs(x - hi - lo) <= np.abs(x) * 2.0 ** -16)
    h2, l2 = bf16x3_split(hi)
    assert np.array_equal(h2, hi) and not l2.any()


def test_round_to_nearest_even_ties():
    # 1 + 2^-8 lies halfway between bf16 neighbours 1 and 1 + 2^-7: ties go to even (1)
    assert bf16_rne(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties go to the even mantissa 1 + 2^-6
    assert bf16_rne(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    assert bf16_rne(np.array([-1.0 - 2.0 ** -8]))[0] == -1.0
