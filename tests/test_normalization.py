"""The -o normalisation factor (RawDataPreparation::FindNormalizationFactor,
raw_data_preparation.cpp:661-686; utils::Median, utils.cpp:1442-1462), on
the host through the C-ABI.  The hand-worked cases fix the reference's
selection: the walk starts at the first maximum, so values above 0.9 * max
that come before it are not counted; the threshold is 0.9 * max computed in
double and stored as float (the comparator's `const RType &`,
raw_data_preparation.cpp:649-658), compared in float; an even count takes the
mean of the two middle values."""
import numpy as np
import pytest

import avionic_amd as A


@pytest.mark.parametrize("values, expected", [
    # one value: itself
    ([3.0], 3.0),
    # 0.95 precedes the maximum and is not counted; kept {1.0, 0.92, 0.99}
    ([0.95, 1.0, 0.5, 0.92, 0.99], 0.99),
    # maximum first: kept {1.0, 0.95, 0.91, 0.97}; even count -> (0.95 + 0.97) / 2
    ([1.0, 0.95, 0.2, 0.91, 0.97], 0.96),
    # a repeated maximum: the walk starts at the first; kept {2, 2, 1.9}
    ([1.0, 2.0, 0.0, 2.0, 1.9], 2.0),
    # below 0.9 * max everything else: the maximum alone
    ([0.1, 0.5, 4.0, 0.2], 4.0),
])
def test_hand_worked_cases(values, expected):
    assert A.normalization_factor(values) == pytest.approx(expected, rel=1e-7)


def test_threshold_is_the_float_rounded_product():
    # 0.9 * 1.0 in double rounds to float32(0.9) = 0.89999998: an element
    # exactly at that float threshold is kept (a double comparison would drop
    # it); the next float below it is not
    thr = np.float32(0.9 * 1.0)
    assert A.normalization_factor([1.0, thr]) == \
        pytest.approx(0.5 * float(np.float32(1.0) + thr), rel=1e-7)
    below = np.nextafter(thr, np.float32(0))
    assert A.normalization_factor([1.0, below]) == 1.0
    assert float(thr) < 0.9  # the case where double and float comparisons differ


def test_matches_sort_based_median_on_random_magnitudes():
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 64, 1000, 4097):
        v = np.abs(rng.standard_normal(n)).astype(np.float32)
        i = int(np.argmax(v))
        thr = np.float32(float(v[i]) * 0.9)
        kept = np.sort(v[i:][v[i:] >= thr])
        k = kept.size
        med = kept[k // 2] if k % 2 else np.float32(0.5 * float(kept[k // 2 - 1] + kept[k // 2]))
        assert A.normalization_factor(v) == np.float32(med)


def test_rejects_empty_and_all_zero():
    with pytest.raises(Exception):
        A.normalization_factor([])
    with pytest.raises(Exception, match="not positive"):
        A.normalization_factor(np.zeros(5))
