"""CPU emulation of the far-tile squared-distance expansions used by the
all-pairs kernel (far paths 3 and 4): the precision budget that lets tiles
past 0.5 Re (Re = the wave's largest distance from the block origin) use
r^2 = |x_i|^2 + |x_j|^2 - 2 x_i . x_j.  The GPU parity tests check the
kernels themselves; these check the arithmetic scheme, in numpy with FP32
rounding, against the exact squared distance of the same FP32 inputs."""
import numpy as np
import pytest

EPS32 = 2.0 ** -24


def bf16(x):
    """Round-to-nearest-even FP32 -> bf16 (kept in an FP32 container)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def split3(x):
    x = np.float32(x)
    h = bf16(x)
    r = np.float32(x - h)
    m = bf16(r)
    return [h, m, bf16(np.float32(r - m))]


# K layout of the weighted bf16 product (A = source row scaled by w = m^(-2/3),
# B = target column): cross terms of split orders a + b <= 2 per axis,
# w |x_j|^2 against 1, and w against |x_i|^2.
RA = [0, 0, 1, 0, 2, 1]
RB = [0, 1, 0, 2, 0, 1]


def source_row(xj, w):
    wx = [split3(np.float32(np.float64(w) * np.float64(v))) for v in xj]
    wq = split3(np.float32(np.float64(w) * np.dot(xj.astype(np.float64), xj)))
    ws = split3(w)
    out = np.zeros(32, np.float32)
    for d in range(3):
        for c in range(6):
            out[6 * d + c] = wx[d][RA[c]]
    out[18:21] = wq
    for c in range(6):
        out[21 + c] = ws[RA[c]]
    return out


def target_col(xi):
    tp = [split3(v) for v in xi]
    cs = split3(np.float32(np.dot(xi.astype(np.float64), xi)))
    out = np.zeros(32, np.float32)
    for d in range(3):
        for c in range(6):
            out[6 * d + c] = bf16(np.float32(-2.0) * tp[d][RB[c]])
    out[18:21] = 1.0
    for c in range(6):
        out[21 + c] = cs[RB[c]]
    return out


def samples(reach, n=4000, seed=0):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        xi = rng.normal(size=3) * 0.01
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = np.linalg.norm(xi) * (reach + rng.exponential(3.0))
        yield xi.astype(np.float32), (xi + d * r).astype(np.float32)


def test_bf16_split_is_exact():
    rng = np.random.default_rng(3)
    xs = (rng.normal(size=2000) * 10.0 ** rng.integers(-8, 8, 2000)).astype(np.float32)
    for x in xs:
        h, m, l = split3(x)
        assert float(h) + float(m) + float(l) == float(x)


@pytest.mark.parametrize("reach,budget", [(0.5, 48), (1.0, 16), (1.5, 12), (4.0, 12)])
def test_bf16_product_r2_within_budget(reach, budget):
    """Far path 4: 27 exact bf16 products accumulated in FP32 give w r^2,
    within about 3 eps (1 + 2 / reach)^2 of the exact value (75 eps at the
    kernel's reach of 0.5 Re; the sampled pairs stay near half of that)."""
    rng = np.random.default_rng(7)
    errs = []
    for xi, xj in samples(reach):
        w = np.float32(1.0 / np.cbrt(rng.uniform(0.1, 10.0) ** 2))
        a, b = source_row(xj, w), target_col(xi)
        acc = np.float32(0.0)
        for k in range(32):
            acc = np.float32(acc + np.float32(a[k] * b[k]))
        exact = float(w) * float(np.sum((xj.astype(np.float64) - xi.astype(np.float64)) ** 2))
        errs.append(abs(float(acc) - exact) / exact)
    errs = np.array(errs)
    assert errs.max() < budget * EPS32 and np.median(errs) < 2 * EPS32


@pytest.mark.parametrize("reach,budget", [(0.5, 24), (2.0, 16)])
def test_weighted_expansion_within_budget(reach, budget):
    """Far path 3: w r^2 from four FP32 FMAs, w = m^(-2/3)."""
    rng = np.random.default_rng(5)
    errs = []
    for xi, xj in samples(reach):
        m = rng.uniform(0.1, 10.0)
        w = np.float32(1.0 / np.cbrt(m * m))
        W = [np.float32(w * xj[0]), np.float32(w * xj[1]), np.float32(w * xj[2]),
             np.float32(w * np.dot(xj.astype(np.float64), xj))]
        ci = np.float32(np.dot(xi.astype(np.float64), xi))
        b = [np.float32(-2.0) * xi[0], np.float32(-2.0) * xi[1], np.float32(-2.0) * xi[2]]
        fma = lambda a, x, c: np.float32(np.float64(a) * np.float64(x) + np.float64(c))
        r2w = fma(W[0], b[0], fma(W[1], b[1], fma(W[2], b[2], fma(w, ci, W[3]))))
        exact = float(w) * float(np.sum((xj.astype(np.float64) - xi.astype(np.float64)) ** 2))
        errs.append(abs(float(r2w) - exact) / exact)
    assert max(errs) < budget * EPS32
