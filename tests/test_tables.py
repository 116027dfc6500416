"""The FP64 sin/cos table and reduction of the far-field kernels, checked on
the CPU: every entry of taichi-bem_amd/csrc/bem_sincos_table.h is the
correctly rounded {sin, cos}(2 pi j / 2048), and the operation sequence of
sincos_f64_t2048 (csrc/bem_device.h), replayed with exact fused
multiply-adds, stays within 6e-16 of the 200-bit values."""
import os
import re
import struct

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "taichi-bem_amd", "csrc", "bem_sincos_table.h")


def _table():
    txt = open(HEADER).read()
    vals = re.findall(r"\{([^{},]+), ([^{},]+)\}", txt.split("kSinCos2048[2048] = {", 1)[1])
    return np.array([[float(a), float(b)] for a, b in vals])


def test_table_entries_are_correctly_rounded():
    tab = _table()
    assert tab.shape == (2048, 2)
    mp.mp.prec = 200
    for j in range(2048):
        a = 2 * mp.pi * j / 2048
        for got, exact in ((tab[j, 0], mp.sin(a)), (tab[j, 1], mp.cos(a))):
            # Within half an ulp of the exact value; the zero crossings hold
            # the 200-bit residue (below 1e-60) in place of zero.
            half_ulp = np.spacing(abs(float(exact))) / 2 if abs(exact) > 1e-300 else 0.0
            assert abs(mp.mpf(got) - exact) <= half_ulp + 1e-60, j


def _fma(a, b, c):
    return float(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))


def _loint(d):
    return struct.unpack("<II", struct.pack("<d", d))[0]


def test_t2048_sequence_error():
    """The kernel's steps: m = fma(x, 2048 / 2pi, 1.5 2^52), n = m - magic,
    t = fma(-n, 2pi / 2048, x), the entry at byte offset (lo(m) << 4) masked
    to the table, sin t = t - t^3/6, cos t = 1 - t^2/2 + t^4/24, then the
    angle sums (each fma rounded once, each product rounded once)."""
    tab = _table()
    mp.mp.prec = 200
    inv_h, h, magic = 2048.0 / 6.283185307179586, 6.283185307179586 / 2048.0, 6755399441055744.0
    rng = np.random.default_rng(4)
    xs = list(rng.uniform(-12.0, 12.0, 250)) + list(rng.uniform(0.0, 0.01, 50)) + [0.0, np.pi, 2 * np.pi, 6.6]
    worst = 0.0
    for x in xs:
        m = _fma(x, inv_h, magic)
        n = m - magic
        t = _fma(-n, h, x)
        t2 = t * t
        st = _fma(t * t2, -1.6666666666666666667e-1, t)
        ct = _fma(t2, _fma(t2, 4.1666666666666666667e-2, -0.5), 1.0)
        off = (_loint(m) << 4) & (16 * 2048 - 1)
        es, ec = tab[off // 16]
        sn = _fma(es, ct, ec * st)
        cs = _fma(ec, ct, -(es * st))
        worst = max(worst, abs(float(mp.sin(mp.mpf(x))) - sn), abs(float(mp.cos(mp.mpf(x))) - cs))
    assert worst <= 6e-16
