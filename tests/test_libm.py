"""mose_pow (mose-gpu_amd/csrc/mose_libm.h) against the host's libm pow, bit for bit.

mose_pow restates glibc's pow so that the device's chemistry controllers (ERR**0.25, FACCON**0.8, QNEWT**p) and
the Troe factor's 10**x take the same bits as the reference's libm-linked Fortran.  The oracle compiles the same
header; oc_pow_batch evaluates it (mode 0) or the host's pow (mode 1) over the same arguments.  The GPU side of the
same check is in test_gpu_chem.py.
"""
import importlib.util
from pathlib import Path

import numpy as np
import pytest

from oracle_bind import _d, oracle_lib

CSRC = Path(__file__).resolve().parent.parent / "mose-gpu_amd" / "csrc"


def both(x, y):
    L = oracle_lib()
    x = np.ascontiguousarray(x, np.float64); y = np.ascontiguousarray(y, np.float64)
    a = np.empty_like(x); b = np.empty_like(x)
    L.oc_pow_batch(0, x.size, _d(x), _d(y), _d(a))
    L.oc_pow_batch(1, x.size, _d(x), _d(y), _d(b))
    return a, b


def same_bits(a, b):
    nan = np.isnan(a) & np.isnan(b)
    return nan | (a.view(np.uint64) == b.view(np.uint64))


def arguments(kind, n, rng):
    if kind == "controllers":       # ERR**0.25, (ERR^2/ERRACC)**0.25, FACCON**0.8, QNEWT**(-1/k), k = 4..10
        x = 10.0 ** rng.uniform(-24, 6, n)
        y = rng.choice(np.array([0.25, 0.8] + [-1.0 / k for k in range(4, 11)]), n)
    elif kind == "troe":            # 10**x, x = log10(Fcent) / (1 + f1^2) in [-3, 0]
        x = np.full(n, 10.0); y = rng.uniform(-3.0, 0.0, n)
    elif kind == "general":
        x = 10.0 ** rng.uniform(-300, 300, n); y = rng.uniform(-4, 4, n)
    elif kind == "bits":            # any bit patterns: negative, subnormal, inf, nan, huge and tiny exponents
        x = rng.integers(0, 2**64, n, dtype=np.uint64).view(np.float64)
        y = rng.integers(0, 2**64, n, dtype=np.uint64).view(np.float64)
    elif kind == "negative_integer_y":
        x = -rng.uniform(0, 3000, n); y = rng.integers(-20, 21, n).astype(np.float64)
    elif kind == "subnormal_x":
        x = rng.integers(1, 2**52, n, dtype=np.uint64).view(np.float64); y = rng.uniform(-2, 2, n)
    elif kind == "near_over_underflow":   # |y log x| in [700, 1100]: the rescaled branches of the exp
        x = 2.0 ** rng.uniform(-1, 1, n)
        lx = np.log(x); lx[lx == 0] = 1e-3
        y = rng.choice([-1.0, 1.0], n) * rng.uniform(700, 1100, n) / lx
    elif kind == "near_one":
        x = 1.0 + rng.uniform(-1e-10, 1e-10, n); y = rng.uniform(-5e11, 5e11, n)
    else:
        raise ValueError(kind)
    return x, y


KINDS = ["controllers", "troe", "general", "bits", "negative_integer_y", "subnormal_x", "near_over_underflow",
         "near_one"]


@pytest.mark.parametrize("kind", KINDS)
def test_mose_pow_is_the_host_pow(kind):
    rng = np.random.default_rng(KINDS.index(kind))
    x, y = arguments(kind, 1_000_000, rng)
    a, b = both(x, y)
    bad = ~same_bits(a, b)
    assert bad.sum() == 0, (kind, x[bad][:4], y[bad][:4], a[bad][:4], b[bad][:4])


def test_special_values():
    sv = np.array([0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 0.5, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 1.7e308, 3.0])
    x, y = np.meshgrid(sv, sv)
    a, b = both(x.ravel(), y.ravel())
    assert same_bits(a, b).all()


def test_sqrt_sqrt_is_not_pow():
    """The former controller form differs from the host pow on a large share of arguments: the check above can
    fail."""
    rng = np.random.default_rng(7)
    x = 10.0 ** rng.uniform(-20, 5, 100_000)
    a, _ = both(x, np.full_like(x, 0.25))
    assert np.mean(np.sqrt(np.sqrt(x)) != a) > 0.05


def test_tables_are_the_generator_output():
    """The committed tables are what gen_libm_tables.py computes from the construction (90-digit arithmetic)."""
    spec = importlib.util.spec_from_file_location("gen_libm_tables", CSRC / "gen_libm_tables.py")
    g = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(g)
    text = (CSRC / "mose_libm_tables.h").read_text()
    for a, b, c in g.log_table():
        assert f"{{{a.hex()}, {b.hex()}, {c.hex()}}}," in text
    for t, s in g.exp_table():
        assert f"0x{t:016x}ULL, 0x{s:016x}ULL," in text
    rows = g.log_table()
    assert rows[74][0] == 1.0 and rows[75][0] == 1.0      # c = 1 on both subintervals beside x = 1: log c = 0
    inv = np.array([r[0] for r in rows])
    assert np.all(inv * 256 == np.round(inv * 256))      # 1/c = j/N or j/2N: z/c - 1 exact
