"""Pins of the oracle's E-value fit (Util/Statistics.hs:37-100) and SP ion
matching (Sequence/IonSeries.hs:99-147, Sequence/Match.hs:33-40) against
closed forms, hand-derived values, numpy's least squares and brute force.

The reference's result files predate the Ions and E-value columns, so there
are no golden vectors for these functions; every check here uses a method
other than the oracle's own formula.
"""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

getcontext().prec = 50

MONO_BW = 1.0005079
H = 1.007276
H2O = 18.010565


# ---------------------------------------------------------------- E-values

def test_evalue_fit_recovers_an_exact_line(orc):
    """Histogram heights 1000, 100, 10, 1 in four consecutive bins of width
    350 lie on logh = const - j: the fit must return beta = -10000/350 and
    alpha = 3 - log10(350) + 13 - 1/2 (bin 13 is the first, centre
    (350*13 - 175)/10000).  497 values below the slice (one at the minimum,
    496 more in bin 13) and 50 above it (in bin 16) make a cut position off
    by one change a height, hence the fit: n = 1658, round(0.3 n) = 497,
    round(0.97 n) = 1608."""
    k13 = 350.0 * 13 - 176.0          # y = x + 1 - 0 = 350*13 - 175, mid-bin
    sc = ([0.0] + [k13] * (496 + 1000) + [350.0 * 14 - 176.0] * 100
          + [350.0 * 15 - 176.0] * 10 + [350.0 * 16 - 176.0] * (1 + 50))
    assert len(sc) == 1658
    rng = np.random.default_rng(3)
    a, b, m, nused = orc.evalue_fit(rng.permutation(np.array(sc, np.float32)))
    assert (m, nused) == (4, 1111)
    assert b == pytest.approx(-10000.0 / 350.0, rel=1e-12)
    assert a == pytest.approx(3 - math.log10(350.0) + 13 - 0.5, rel=1e-12)


@pytest.mark.parametrize("n,perc,expect", [
    (15, 3.0, 15 - 4),            # round(4.5) = 4: halves to even
    (5, 3.0, 5 - 2),              # round(1.5) = 2
    (50, 3.0, 48 - 15),           # Float 48.5 -> 48
    (550, 7.0, 512 - 165),        # Float 511.5 -> 512; in Double 511.4999.. -> 511
    (135317, 3.0, 131258 - 40595),  # Float 131257.5 -> 131258; Double -> 131257
])
def test_evalue_cut_positions(orc, n, perc, expect):
    """cut_end = round((1 - perc/100) * n) in Float and cut_beg =
    round(0.3 * n) in Double (Statistics.hs:72-73), both halves to even:
    with every score but the minimum above 4000, the kept count is
    cut_end - cut_beg (values derived by hand)."""
    sc = np.concatenate([[0.0], 10000.0 + np.arange(n - 1)]).astype(np.float32)
    _, _, _, nused = orc.evalue_fit(sc[::-1].copy(), 350, perc)
    assert nused == expect


def _numpy_fit(sc, w=350, perc=3.0):
    """The fit by numpy: cut by Python's round (halves to even), bins by
    np.ceil, counts by np.unique, slope and intercept by np.polyfit."""
    s = np.sort(np.asarray(sc, np.float64))
    n = len(s)
    ce = round(float((np.float32(1) - np.float32(perc) / np.float32(100)) * np.float32(n)))
    cb = round(0.3 * n)
    y = (s[cb:ce] + 1.0) - s[0]
    y = y[y > 4000]
    k, c = np.unique(np.ceil(y / w), return_counts=True)
    x = (k * w - w / 2) / 10000.0
    lh = np.log10(c / w)
    if len(x) < 2:
        return math.nan, math.nan, len(x), len(y)
    beta, alpha = np.polyfit(x, lh, 1)
    return alpha, beta, len(x), len(y)


@pytest.mark.parametrize("seed", range(6))
def test_evalue_fit_against_polyfit(orc, seed):
    """linearRegression is ordinary least squares (statistics-linreg):
    the oracle's (alpha, beta) agree with numpy's polyfit over an
    independently built histogram, to 1e-9."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(50, 40000))
    sc = (rng.gamma(2.0, 2500.0, n) - rng.uniform(0, 3000)).astype(np.float32)
    a, b, m, nused = orc.evalue_fit(sc)
    ea, eb, em, enu = _numpy_fit(sc)
    assert (m, nused) == (em, enu)
    assert m >= 2
    assert b == pytest.approx(eb, rel=1e-9, abs=1e-9)
    assert a == pytest.approx(ea, rel=1e-9, abs=1e-9)


def test_evalue_fit_degenerate_cases(orc):
    """No scores, one score, and a single histogram point leave 0/0 in the
    slope: NaN, as Haskell's Double division gives.  Two points: the line
    through them (closed form)."""
    for sc in ([], [5.0], [0.0] + [9000.0] * 40):
        a, b, m, _ = orc.evalue_fit(np.array(sc, np.float32))
        assert math.isnan(a) and math.isnan(b) and m <= 1
    # bins 26 (count 20) and 28 (count 2) after cut: 10 at the minimum 0,
    # then 20 at y = 9000 and 2 at y = 9700, 1 high value cut away
    sc = [0.0] * 10 + [8999.0] * 20 + [9699.0] * 2 + [20000.0]
    n = len(sc)
    assert round(0.3 * n) == 10 and round(float(np.float32(0.97) * np.float32(n))) == 32
    a, b, m, nused = orc.evalue_fit(np.array(sc, np.float32))
    x1, x2 = (350 * 26 - 175) / 1e4, (350 * 28 - 175) / 1e4
    y1, y2 = math.log10(20 / 350), math.log10(2 / 350)
    assert (m, nused) == (2, 22)
    assert b == pytest.approx((y2 - y1) / (x2 - x1), rel=1e-12)
    assert a == pytest.approx(y1 - b * x1, rel=1e-12)


def test_evalue_is_a_power_of_ten(orc):
    """Statistics.hs:41: E = 10 ** (alpha + beta * score / 10000), checked
    against 50-digit decimal arithmetic."""
    for a, b, s in [(2.5, -3.0, 12345.0), (-1.0, -28.5, 40000.0), (0.3, 0.0, 1.0)]:
        e = orc.evalue(a, b, s)
        z = Decimal(a) + Decimal(b) * (Decimal(s) / Decimal(10000))
        assert e == pytest.approx(float(Decimal(10) ** z), rel=1e-14)


# ---------------------------------------------------------------- peaks

def _brute_peaks(x):
    top = sorted(range(len(x)), key=lambda i: (-float(x[i]), i))[:200]
    acc = [0.0] * (len(x) // 3 + 1)
    for i in top:
        acc[i // 3] = max(acc[i // 3], float(x[i]))
    return np.array([v > 0 for v in acc], np.uint8)


@pytest.mark.parametrize("seed", range(5))
def test_extract_peaks_brute_force(orc, seed):
    """extractPeaks: the 200 largest values folded by max into bins i div 3;
    brute force by a full sort, with ties (lowest bins first), fewer and
    more than 200 positive values and lengths not a multiple of 3."""
    rng = np.random.default_rng(seed)
    L = int(rng.integers(150, 3000))
    x = rng.normal(0, 1, L).astype(np.float32)
    if seed % 2:
        x = np.round(x * 4) / 4            # many ties, also at the 200th value
    if seed == 4:
        x = -np.abs(x)
        x[rng.choice(L, 50, replace=False)] = 1.0   # 50 positives only
    np.testing.assert_array_equal(orc.extract_peaks(x), _brute_peaks(x))


# ---------------------------------------------------------------- ions

def _pep_db(orc, seq: bytes):
    return orc.prepare_ions(np.frombuffer(seq, np.uint8))


def test_match_ions_hand_example(orc):
    """AGSR (residues 0..3 of AGSRK) against flags {24, 72, 58, 6, 124} and,
    for charge 3+, also {12, 118, 62}; bins derived by hand from the m/z
    ladder (b: 72.044, 129.066, 216.098, 372.199; r = 390.2096, y = r - b).
    init drops the last ion (bin 124 / 62), tail the first (y bin 106)."""
    mass = orc.mass_table()
    ions = _pep_db(orc, b"AGSRK")
    res = np.float32(0)
    for ch in b"AGSR":
        res = np.float32(res + mass[ch - 65])
    on = np.zeros(200, np.uint8)
    on[[24, 72, 58, 6, 124]] = 1
    nb, ny, tot, bf, yf = orc.match_ions(ions, 0, 3, res, mass, 1, on, flags=True)
    assert (nb, ny, tot) == (2, 2, 6)
    assert list(bf) == [1, 0, 1] and list(yf) == [0, 1, 1]
    on[[12, 118, 62]] = 1
    nb, ny, tot, bf, yf = orc.match_ions(ions, 0, 3, res, mass, 2, on, flags=True)
    assert list(bf) == [1, 0, 1, 1, 1, 0, 0]
    assert list(yf) == [0, 1, 1, 1, 0, 0, 0]
    assert (nb, ny, tot) == (4, 3, 6)
    # modified sites lengthen the printed peptide: total grows by 2 each
    assert orc.match_ions(ions, 0, 3, res, mass, 2, on, nmod=2)[2] == 10


def test_match_ions_database_end(orc):
    """lookup (Location.hs:49-55): a peptide whose end index is the
    database's length keeps residues tc .. tn - 1 only."""
    mass = orc.mass_table()
    ions = _pep_db(orc, b"GGAGSR")
    on = np.ones(400, np.uint8)
    res = np.float32(0)
    for ch in b"AGSR":
        res = np.float32(res + mass[ch - 65])
    nb, ny, tot = orc.match_ions(ions, 2, 6, res, mass, 1, on)
    assert tot == 6 and nb == 3 and ny == 3


def _decimal_bins(masses, r, zmax):
    """Ion bins in exact decimal arithmetic from the FP32 masses; None where
    a quotient lies within 1e-4 of a half (FP32 rounding could decide)."""
    bw3 = Decimal(float(np.float32(MONO_BW))) * 3
    hh = Decimal(float(np.float32(H)))
    ions = []
    for c in range(1, zmax + 1):
        p = Decimal(0)
        for m in masses:
            p += Decimal(float(m))
            ions.append((p + hh * c) / c)

    def b(v):
        q = v / bw3
        f = q - q.to_integral_value(rounding="ROUND_FLOOR")
        if abs(f - Decimal("0.5")) < Decimal("1e-4"):
            return None
        return int(q.to_integral_value(rounding="ROUND_HALF_EVEN"))

    return [b(v) for v in ions], [b(Decimal(float(r)) - v) for v in ions]


@pytest.mark.parametrize("seed", range(4))
def test_match_ions_against_decimal_arithmetic(orc, seed):
    """Random peptides and flags: every ion whose bin is not within 1e-4 of
    a rounding half is matched exactly as 50-digit arithmetic says."""
    rng = np.random.default_rng(seed)
    mass = orc.mass_table({"C": 57.021464})
    letters = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", np.uint8)
    checked = 0
    for _ in range(40):
        n = int(rng.integers(4, 30))
        seq = rng.choice(letters, n + 2)
        ions = orc.prepare_ions(seq)
        res = np.float32(0)
        for ch in seq[1:n + 1]:
            res = np.float32(res + mass[ch - 65])
        zmax = int(rng.integers(1, 4))
        on = (rng.random(1400) < 0.3).astype(np.uint8)
        nb, ny, tot, bf, yf = orc.match_ions(ions, 1, n, res, mass, zmax, on, flags=True)
        r = np.float32(np.float32(res + np.float32(H)) + np.float32(H2O))
        bb, yb = _decimal_bins([mass[ch - 65] for ch in seq[1:n + 1]], r, zmax)
        for k in range(len(bb) - 1):
            if bb[k] is not None:
                assert bf[k] == (0 <= bb[k] < len(on) and on[bb[k]] == 1)
                checked += 1
        for k in range(1, len(yb)):
            if yb[k] is not None:
                assert yf[k - 1] == (0 <= yb[k] < len(on) and on[yb[k]] == 1)
                checked += 1
        assert (nb, ny, tot) == (int(bf.sum()), int(yf.sum()), 2 * (n - 1))
    assert checked > 1000
