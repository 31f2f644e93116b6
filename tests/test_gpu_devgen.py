"""Device-side generators against the host ones: the uniform column and the R-MAT graph are the
same values as the host generators (amj_gen_uniform, amj_gen_rmat + amj_graph_orient), bit for
bit; the Zipf column follows P(rank r) proportional to r^-s (chi-square against the law) and its
rank -> key map is a bijection of [0, k)."""
import ctypes as C

import numpy as np
import pytest

from accmultiwayjoins_amd import _native as N
from accmultiwayjoins_amd import generators as G

pytestmark = pytest.mark.gpu


class DevBuf:
    def __init__(self, ctx, n):
        self.ctx, self.n = ctx, n
        self.p = C.c_void_p()
        assert N.lib().amj_device_alloc(ctx.h, 4 * max(n, 1), C.byref(self.p)) == 0

    def host(self, n=None):
        out = np.empty(self.n if n is None else n, np.int32)
        assert N.lib().amj_memcpy(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.p,
                                  4 * len(out)) == 0
        return out

    def free(self):
        N.lib().amj_device_free(self.ctx.h, self.p, 4 * max(self.n, 1))


@pytest.mark.parametrize("lo,hi", [(0, 1000), (-5, 7), (-(1 << 31), 1 << 31), (3, 4)])
def test_device_uniform_equals_host(ctx, lo, hi):
    n = 1_000_003
    b = DevBuf(ctx, n)
    assert N.lib().amj_gen_device_uniform(ctx.h, b.p, n, lo, hi, 42, 7) == 0
    assert np.array_equal(b.host(), G.uniform(n, lo, hi, 42, 7))
    b.free()


@pytest.mark.parametrize("s", [0.0, 0.5, 1.0, 1.2])
def test_device_zipf_follows_the_law(ctx, s):
    n, k = 2_000_000, 50
    b = DevBuf(ctx, n)
    assert N.lib().amj_gen_device_zipf(ctx.h, b.p, n, k, C.c_double(s), 9, 3) == 0
    v = b.host()
    b.free()
    assert v.min() >= 0 and v.max() < k
    counts = np.bincount(v, minlength=k)
    # the rank -> key map is a bijection: sorting key counts recovers the rank law
    obs = np.sort(counts)[::-1].astype(np.float64)
    p = np.arange(1, k + 1, dtype=np.float64) ** -s
    exp = n * p / p.sum()
    chi2 = float(((obs - exp) ** 2 / exp).sum())
    assert chi2 < 150.0, chi2  # 49 degrees of freedom (sorting only lowers it)
    if s == 0.0:
        assert counts.min() > 0.9 * n / k


@pytest.mark.parametrize("scale", [8, 12, 14])
def test_device_rmat_equals_host(ctx, scale):
    m = 16 << scale
    s, d = DevBuf(ctx, m), DevBuf(ctx, m)
    k = C.c_uint64()
    assert N.lib().amj_gen_device_rmat(ctx.h, scale, 16, C.c_double(0.57), C.c_double(0.19),
                                       C.c_double(0.19), 42 + scale, s.p, d.p, C.byref(k)) == 0
    hs, hd = G.rmat_oriented(scale, 16, seed=42 + scale)
    assert k.value == len(hs)
    assert np.array_equal(s.host(k.value), hs)
    assert np.array_equal(d.host(k.value), hd)
    s.free()
    d.free()
