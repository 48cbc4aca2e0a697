"""Variant table of the correlation kernel (host-side calls only: no GPU).

The library builds several launch variants per window size; the autotuner
picks among them at run time.  These checks pin which variants exist, that
selection is per size and reversible, and that unbuilt variants are refused.
"""

import pytest

from pivtools_amd import _native
from pivtools_amd import autotune as AT

BUILT = {8: [1, 65], 12: [1, 65], 16: [1, 2, 16, 65, 96, 97, 98], 24: [1, 2, 65], 32: [1, 2, 4, 16, 65],
         48: [2, 4, 16, 68], 64: [2, 4, 16, 36, 68], 96: [4, 36, 68], 128: [4, 8, 36, 68]}
DEFAULT = {8: 1, 12: 1, 16: 65, 24: 1, 32: 65, 48: 68, 64: 68, 96: 4, 128: 68}


@pytest.mark.parametrize("n", sorted(BUILT))
def test_variants_built_per_size(n):
    assert AT.variants(n) == BUILT[n]
    assert AT.default_variant(n) == DEFAULT[n]
    assert DEFAULT[n] in BUILT[n]


def test_unsupported_size_has_no_variants():
    assert AT.variants(20) == []
    assert _native.core().pivamd_xcorr_default_variant(20) < 0
    assert AT.get_variant(20) == -1


def test_set_variant_is_per_size_and_reversible():
    try:
        assert AT.get_variant(32) == 0 and AT.get_variant(64) == 0
        assert AT.set_variant(32, 2) == 0
        assert AT.get_variant(32) == 2 and AT.get_variant(64) == 0
        assert AT.set_variant(32, 16) == 2
        AT.apply({64: {"variant": 2}})
        assert AT.get_variant(64) == 2
        with pytest.raises(ValueError):
            AT.set_variant(128, 1)       # 128 points do not fit one lane
        with pytest.raises(ValueError):
            AT.set_variant(96, 16)       # the MFMA form stops at 64
        assert AT.get_variant(128) == 0 and AT.get_variant(96) == 0
    finally:
        AT.set_variant(32, 0)
        AT.set_variant(64, 0)
    assert AT.get_variant(32) == 0 and AT.get_variant(64) == 0


def test_agree_rule():
    import numpy as np
    x = np.zeros((4, 5), np.float32)
    y = np.zeros_like(x)
    assert AT.agree((x, y), (x + 5e-4, y))
    assert not AT.agree((x, y), (x + 2e-3, y))
    xn = x.copy()
    xn[0, 0] = np.nan
    assert not AT.agree((x, y), (xn, y), max_flip=0.0)
