"""The MA(1) wave scans' lane-exchange plans (CPU, numpy).

The likelihood composes the moving-average recurrence eps_i = a_i eps_{i-1}
+ b_i over a wave of 64 lanes with an inclusive prefix scan of (a, b) maps
(wave_affine_scan in astroemperor_amd/csrc/emp_kernels.h): row_shr 1, 2, 4, 8
within rows of 16 lanes (lanes without a source take the identity), then
row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3.  The
adjoint composes lambda_i = b_i + a_i lambda_{i+1} with a suffix scan
(wave_affine_suffix_scan): row_shl 1, 2, 4, 8 within rows, then rows 0 and
2 take the first lane of the next row and rows 0 and 1 lane 32.  This test
restates both plans lane by lane, with the DPP controls' semantics (row_shr:n
reads lane i - n of the row, row_shl:n lane i + n, row_bcast:15 lane 15 of
the previous row), and checks them against the serial recurrences; the
cross-row order of a whole-wave shuffle scan (lane i + 16, then i + 32)
after row-clipped steps leaves a gap and must fail.
"""
import numpy as np
import pytest

LANES = np.arange(64)


def serial_prefix(a, b):
    A, B = np.empty(64), np.empty(64)
    ca, cb = 1.0, 0.0
    for i in range(64):
        ca, cb = a[i] * ca, a[i] * cb + b[i]
        A[i], B[i] = ca, cb
    return A, B


def serial_suffix(a, b):
    A, B = np.empty(64), np.empty(64)
    ca, cb = 1.0, 0.0
    for i in range(63, -1, -1):
        ca, cb = a[i] * ca, a[i] * cb + b[i]
        A[i], B[i] = ca, cb
    return A, B


def compose(a, b, pa, pb, mask):
    """Lane's map (a, b) composed with the source's (pa, pb) where mask."""
    return np.where(mask, a * pa, a), np.where(mask, a * pb + b, b)


def dpp_row_shift(x, n, fill):
    """row_shr:n (n > 0) or row_shl:-n (n < 0) within rows of 16 lanes."""
    src = LANES - n
    ok = (src >= 0) & (src < 64) & (src // 16 == LANES // 16)
    return np.where(ok, x[np.clip(src, 0, 63)], fill)


def prefix_plan(a, b):
    for n in (1, 2, 4, 8):
        pa, pb = dpp_row_shift(a, n, 1.0), dpp_row_shift(b, n, 0.0)
        a, b = a * pa, a * pb + b
    # row_bcast:15 into rows 1 and 3 (row_mask 0xa): lane 15 of the row before
    src = (LANES // 16 - 1) * 16 + 15
    m = np.isin(LANES // 16, (1, 3))
    a, b = compose(a, b, a[np.clip(src, 0, 63)], b[np.clip(src, 0, 63)], m)
    # row_bcast:31 into rows 2 and 3 (row_mask 0xc)
    m = LANES >= 32
    a, b = compose(a, b, a[31], b[31], m)
    return a, b


def suffix_plan(a, b, shuffle_cross_rows=False):
    for n in (1, 2, 4, 8):
        pa, pb = dpp_row_shift(a, -n, 1.0), dpp_row_shift(b, -n, 0.0)
        a, b = a * pa, a * pb + b
    if shuffle_cross_rows:
        for o in (16, 32):
            src = LANES + o
            m = src < 64
            a, b = compose(a, b, a[np.clip(src, 0, 63)], b[np.clip(src, 0, 63)], m)
        return a, b
    src = (LANES & 48) + 16
    m = (LANES & 16) == 0
    a, b = compose(a, b, a[src % 64], b[src % 64], m)
    m = LANES < 32
    a, b = compose(a, b, a[32], b[32], m)
    return a, b


@pytest.fixture(scope="module")
def maps():
    rng = np.random.default_rng(5)
    out = []
    for _ in range(200):
        a = rng.uniform(-0.5, 0.5, 64)
        a[rng.uniform(size=64) < 0.05] = 0.0        # segment starts
        out.append((a, rng.normal(size=64)))
    return out


def test_dpp_row_shift_semantics():
    x = LANES.astype(float)
    assert list(dpp_row_shift(x, 1, -1)[:17]) == [-1] + list(range(15)) + [-1]
    assert list(dpp_row_shift(x, -1, -1)[:17]) == list(range(1, 16)) + [-1, 17]
    assert list(dpp_row_shift(x, -8, -7)[:16]) == list(range(8, 16)) + [-7] * 8


def test_prefix_plan_matches_serial(maps):
    for a, b in maps:
        A, B = serial_prefix(a, b)
        A2, B2 = prefix_plan(a, b)
        np.testing.assert_allclose(A2, A, rtol=1e-13, atol=1e-300)
        np.testing.assert_allclose(B2, B, rtol=1e-12, atol=1e-14)


def test_suffix_plan_matches_serial(maps):
    for a, b in maps:
        A, B = serial_suffix(a, b)
        A2, B2 = suffix_plan(a, b)
        np.testing.assert_allclose(A2, A, rtol=1e-13, atol=1e-300)
        np.testing.assert_allclose(B2, B, rtol=1e-12, atol=1e-14)


def test_shuffle_cross_rows_after_row_steps_leave_a_gap(maps):
    worst = 0.0
    for a, b in maps:
        _, B = serial_suffix(a, b)
        _, B2 = suffix_plan(a, b, shuffle_cross_rows=True)
        worst = max(worst, np.max(np.abs(B2 - B)))
    assert worst > 1e-3
