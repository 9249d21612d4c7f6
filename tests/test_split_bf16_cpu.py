"""The arithmetic of the split-bf16 GEMM (gemm_bf16.hip), restated on the CPU: every fp32 value
x = p0 + p1 + p2 with p_t the bf16 rounding (to nearest even) of what the earlier parts leave;
each residual exact in fp32, three parts exact for normal numbers, two parts within 2^-16 |x|;
the part products with s + t < NS summed exactly (bf16 x bf16 products are exact in fp32) leave
a real product within 3 * 2^-24 |a||b| (NS = 3) or 3 * 2^-16 |a||b| (NS = 2). Host-only: no GPU."""
import numpy as np


def _bf16(x):
    """Round float32 to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def _split(x, ns):
    parts, r = [], np.asarray(x, dtype=np.float32)
    for _ in range(ns):
        p = _bf16(r)
        parts.append(p)
        r = (r - p).astype(np.float32)
    return parts, r


def test_split_is_exact_and_bounded():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(200000) * np.exp(rng.uniform(-20, 20, 200000))).astype(np.float32)
    p3, r3 = _split(x, 3)
    # residuals exact: the float64 sum of parts and residual reproduces x
    assert np.array_equal(sum(p.astype(np.float64) for p in p3) + r3, x.astype(np.float64))
    assert np.all(r3 == 0)                                   # three parts hold the 24-bit significand
    p2, r2 = _split(x, 2)
    assert np.all(np.abs(r2) <= 2.0 ** -16 * np.abs(x))       # |r| <= 2^-8NS |x|
    for p in p3:   # every part is a bfloat16 value
        assert np.array_equal(p, _bf16(p))


def test_part_products_error():
    rng = np.random.default_rng(1)
    a = rng.standard_normal(100000).astype(np.float32)
    b = rng.standard_normal(100000).astype(np.float32)
    exact = a.astype(np.float64) * b.astype(np.float64)
    for ns, bound in ((3, 3 * 2.0 ** -24), (2, 3 * 2.0 ** -16)):
        pa, _ = _split(a, ns)
        pb, _ = _split(b, ns)
        got = np.zeros_like(exact)
        for s in range(ns):
            for t in range(ns - s):
                prod = pa[s].astype(np.float64) * pb[t].astype(np.float64)
                assert np.array_equal(prod.astype(np.float32).astype(np.float64), prod)   # exact in fp32
                got += prod
        err = np.abs(got - exact) / (np.abs(a).astype(np.float64) * np.abs(b))
        assert err.max() <= bound, (ns, err.max())
