"""Host checks of float identities the carving kernels rely on (no GPU needed).

The window bounds of a child (visual_hull_cuda.cu:236-247) are
clamp(int64(roundf(v)), 0, n - 1).  The kernels clamp first and then truncate
c + 0.49999997f; these tests check that against roundf for every float the
clamp can produce (roundf is the identity below 0.25 and above 2^23 only the
clamp matters).
"""

import numpy as np

HALF_DOWN = np.float32(0.49999997)


def _roundf(c: np.ndarray) -> np.ndarray:
    # roundf (halves away from zero) of non-negative floats, exact in float64
    return np.floor(c.astype(np.float64) + 0.5)


def test_half_down_is_the_float_below_one_half():
    assert HALF_DOWN == np.nextafter(np.float32(0.5), np.float32(0.0))


def test_truncating_plus_half_down_is_roundf_for_every_clamped_float():
    lo = int(np.float32(0.25).view(np.uint32))
    hi = int(np.float32(2.0**23).view(np.uint32))
    chunk = 1 << 23
    bad = 0
    for start in range(lo, hi, chunk):
        c = np.arange(start, min(start + chunk, hi), dtype=np.uint32).view(np.float32)
        got = np.trunc(c + HALF_DOWN)  # float32 add, rounded to nearest even
        bad += int(np.count_nonzero(got != _roundf(c)))
    assert bad == 0


def test_truncating_plus_one_half_fails_just_below_one_half():
    # the reason for HALF_DOWN: c + 0.5f rounds 0.49999997 up to 1
    c = np.array([HALF_DOWN], dtype=np.float32)
    assert np.trunc(c + np.float32(0.5))[0] == 1.0
    assert _roundf(c)[0] == 0.0
