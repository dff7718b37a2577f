"""The ZNCC kernels test only the first and last sample of a window row against the right image's
bounds. That equals testing every sample because, within a row, the float sample position strictly
increases with the column: each step adds 1 - gx >= 0.1 (|gx| <= 0.9) and the roundings of the
position's terms are a few ulps of values below 2^17. This evaluates the kernels' float32 expressions,
in their operation order, over random pixels, disparities, slopes and rows in the supported ranges
(images up to 32768 columns, full-resolution pairs up to 2 x 32768) and checks the order directly.
"""
import numpy as np

F = np.float32


def refine_positions(x, d, gx, gy, v, r):
    # refine_kernel: xb = (float)(x + u) - ((d + gx * u) + gy * v), gx * u formed once per column
    u = np.arange(-r, r + 1, dtype=np.int64)
    gxu = (gx[:, None] * u[None, :].astype(F)).astype(F)
    t = (d[:, None] + gxu).astype(F)
    t = (t + (gy * v.astype(F))[:, None]).astype(F)
    return ((x[:, None] + u[None, :]).astype(F) - t).astype(F)


def detail_positions(x, D, gx, gy, Y, n):
    # detail_kernel: x0 = 2x - n/2 + 1, cx = 2x + 0.5, xb = (float)(x0 + i) - ((D + gx * ((float)(x0 + i) - cx)) + dy)
    x0 = 2 * x - n // 2 + 1
    cx = (F(2) * x.astype(F) + F(0.5)).astype(F)
    X = (x0[:, None] + np.arange(n)[None, :]).astype(F)
    gxx = (gx[:, None] * (X - cx[:, None]).astype(F)).astype(F)
    cy = (F(2) * (Y // 2).astype(F) + F(0.5)).astype(F)
    dy = (gy * (Y.astype(F) - cy).astype(F)).astype(F)
    t = ((D[:, None] + gxx).astype(F) + dy[:, None]).astype(F)
    return (X - t).astype(F)


def test_refine_rows_increase_strictly():
    rng = np.random.default_rng(7)
    for r in (1, 2, 3, 4, 5):
        m = 200_000
        x = rng.integers(0, 32768, m)
        d = rng.uniform(-1100.0, 33000.0, m).astype(F)
        gx = rng.uniform(-0.9, 0.9, m).astype(F)
        gx[:1000] = F(0.9)  # the steepest slopes
        gy = rng.uniform(-0.9, 0.9, m).astype(F)
        v = rng.integers(-r, r + 1, m)
        xb = refine_positions(x, d, gx, gy, v, r)
        assert (np.diff(xb, axis=1) > 0).all(), r


def test_detail_rows_increase_strictly():
    rng = np.random.default_rng(8)
    for n in (2, 4, 6, 8, 10, 12, 32):
        m = 200_000
        x = rng.integers(0, 32768, m)
        D = rng.uniform(-2200.0, 66000.0, m).astype(F)
        gx = rng.uniform(-0.9, 0.9, m).astype(F)
        gx[:1000] = F(0.9)
        gy = rng.uniform(-0.9, 0.9, m).astype(F)
        Y = rng.integers(0, 65536, m)
        xb = detail_positions(x, D, gx, gy, Y, n)
        assert (np.diff(xb, axis=1) > 0).all(), n
