"""The fit model's exp (palmtracer-cpp_amd/csrc/pt_exp.h): its 2^(i/128) table recomputed from the
mathematics, and the oracle's C-library exp pinned against correctly rounded values. CPU only; the
device routine's bit-equality with that exp is a GPU test (test_gpu_parity.py)."""
import os
import re
import struct
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np

HEADER = os.path.join(os.path.dirname(__file__), "..", "palmtracer-cpp_amd", "csrc", "pt_exp.h")


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def test_exp_table_is_two_to_the_i_over_128():
    """Entry 2i = bits of t_i, entry 2i+1 = bits of H_i - (i << 45), with H_i the double nearest
    2^(i/128) and t_i the double nearest 2^(i/128)/H_i - 1 (80 significant digits)."""
    text = open(HEADER).read()
    body = text[text.index("kExpTable[256]"):text.index("};", text.index("kExpTable[256]"))]
    vals = [int(v, 16) for v in re.findall(r"0x([0-9a-f]{16})ull", body)]
    assert len(vals) == 256
    getcontext().prec = 80
    for i in range(128):
        v = Decimal(2) ** (Decimal(i) / Decimal(128))
        h = float(Fraction(v))                     # float(Fraction) rounds to nearest
        t = float(Fraction(v / Decimal(h) - 1))
        assert vals[2 * i] == _bits(t), i
        assert vals[2 * i + 1] == (_bits(h) - (i << 45)) % 2**64, i
    # The table's own invariant: H_i * (1 + t_i) approximates 2^(i/128) to far below an ulp of H_i.
    for i in (1, 37, 64, 127):
        h = struct.unpack("<d", struct.pack("<Q", (vals[2 * i + 1] + (i << 45)) % 2**64))[0]
        t = struct.unpack("<d", struct.pack("<Q", vals[2 * i]))[0]
        rel = abs(Fraction(h) * (1 + Fraction(t)) - Fraction(Decimal(2) ** (Decimal(i) / Decimal(128)))) / Fraction(h)
        assert rel < Fraction(1, 2**100)


def test_oracle_exp_within_half_ulp_of_correctly_rounded(oracle_mod):
    """The oracle's exp (the C library's, which the reference's Gaussian model calls) differs from the
    correctly rounded exp by at most one ulp, and equals it in almost every case (its published bound
    is 0.502 ulp): arguments over the fit model's range and the subnormal and overflow edges."""
    getcontext().prec = 60
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-40.0, 1.0, 3000), rng.uniform(-745.0, 709.0, 1000),
                        np.array([0.0, -0.0, 1e-300, -1e-17, 1.0, -1.0, 709.78, -708.4, -745.13])])
    y = oracle_mod.libm_exp(x)
    exact_equal = 0
    for xi, yi in zip(x, y):
        cr = float(Fraction(Decimal(float(xi)).exp()))
        if cr == yi:
            exact_equal += 1
        else:
            assert abs(_bits(cr) - _bits(float(yi))) == 1, (xi, yi, cr)
    assert exact_equal >= 0.99 * len(x)
