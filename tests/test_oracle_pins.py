"""Pins of the CPU oracle against closed forms, brute force, independent
high-precision arithmetic and the reference's own known answers.

Each test names the reference function it pins.  A check that merely
re-typed the oracle's formula would pin nothing; these use a different
method (exact rationals, enumeration, linear scans, hand-derived values).
"""
import itertools
import json
import math
import os
from decimal import Decimal, getcontext

import numpy as np
import pytest

getcontext().prec = 50

# combinatorics.h:56, last row of the reference's Pascal table (a fixture)
REF_COMB_ROW_42 = [1, 42, 861, 11480, 111930, 850668, 5245786, 26978328, 118030185,
                   445891810, 1471442973, 4280561376]
REF_COMB_ROW_30 = [1, 30, 435, 4060, 27405, 142506, 593775, 2035800, 5852925, 14307150,
                   30045015, 54627300]


def test_choose_matches_reference_table_and_math(orc):
    """combinatorics.h:13-88: table rows and C(n,k) for n <= 42, k <= 11."""
    assert [orc.choose(42, k) for k in range(12)] == REF_COMB_ROW_42
    assert [orc.choose(30, k) for k in range(12)] == REF_COMB_ROW_30
    for n in range(43):
        for k in range(12):
            assert orc.choose(n, k) == (math.comb(n, k) if k <= n else 0)
    # the uint32 loop outside the table (combinatorics.h:59-77)
    # (cases whose uint32 intermediates do not wrap)
    for n, k in [(43, 2), (50, 3), (60, 1), (44, 44), (10, 11), (100, 4)]:
        assert orc.choose(n, k) == math.comb(n, k)


def test_unrank_combinatorial_number_system(orc):
    """combinatorics.h:102-134 unrankComb: ranks <= total/2 follow the
    combinatorial number system (sum C(a_i, k-i) == rank, a descending); larger
    ranks are the reflection n-1-a of the subset whose rank is total-1-rank."""
    for n in range(1, 13):
        for k in range(1, n + 1):
            total = math.comb(n, k)
            for rank in range(total):
                a = [int(v) for v in orc.unrank_comb(n, k, rank)]
                assert len(set(a)) == k and all(0 <= v < n for v in a)
                if total // 2 >= rank:
                    assert a == sorted(a, reverse=True)
                    assert sum(math.comb(a[i], k - i) for i in range(k)) == rank
                else:
                    d = [n - 1 - v for v in a]
                    assert sum(math.comb(d[i], k - i) for i in range(k)) == total - 1 - rank


def test_unrank_is_not_a_bijection(orc):
    """The reflection trick of unrankComb (combinatorics.h:113-131) is not a
    bijection: for n=5, k=2 (derived by hand) ranks 5 and 7 both give {2,3}
    and {0,4} never appears.  The enumeration reproduces this exactly."""
    subsets = [frozenset(int(v) for v in orc.unrank_comb(5, 2, r)) for r in range(10)]
    assert subsets[5] == subsets[7] == frozenset({2, 3})
    assert frozenset({0, 4}) not in subsets
    bij = {(n, k) for n in range(1, 11) for k in range(1, n + 1)
           if len({frozenset(orc.unrank_comb(n, k, r).tolist()) for r in range(math.comb(n, k))})
           == math.comb(n, k)}
    assert (1, 1) in bij and (4, 2) in bij and (5, 2) not in bij


def test_mod_info_hand_derived(orc):
    """Fragment.hs:260-289 makeModInfo."""
    # hfx.params: mod_Q = -17, max-modable-acids = 2  =>  one vector [1]
    ma, mam, cnt, sm, dl = orc.mod_info({"Q": -17.0}, 2)
    assert ma == b"Q" and cnt.tolist() == [[1]] and sm.tolist() == [1]
    assert dl[0] == np.float32(-17.0)
    # two acids, max 2: [[0,1],[1,0]] (derived by hand from addModDimensions)
    ma, mam, cnt, sm, dl = orc.mod_info({"S": 79.9663, "M": 15.9949}, 2)
    assert ma == b"MS" and cnt.tolist() == [[0, 1], [1, 0]]
    # M,S,T,Y with max 4: every vector with sum 1..3 in lexicographic order (34)
    ma, mam, cnt, sm, dl = orc.mod_info({"M": 15.9949, "S": 79.9663, "T": 79.9663,
                                         "Y": 79.9663}, 4)
    expect = [list(v) for v in itertools.product(range(5), repeat=4) if 0 < sum(v) < 4]
    assert cnt.tolist() == expect and len(expect) == 34
    assert sm.tolist() == [sum(v) for v in expect]
    # delta = foldl' (+) 0 (m_i * c_i) in FP32
    for v, d in zip(expect, dl):
        acc = np.float32(0)
        for c, m in zip(v, mam):
            acc = np.float32(acc + np.float32(m * np.float32(c)))
        assert d == acc


def test_prepare_ions_packing(orc):
    """prepare_ions.cu:29-96 + utils.h:92-101: (c-'A') | (1+index in ma) << 5."""
    seq = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWYMSTY", np.uint8)
    out = orc.prepare_ions(seq, b"MSTY")
    for c, o in zip(seq, out):
        mod = b"MSTY".find(bytes([c])) + 1
        assert o & 0x1F == c - 65 and o >> 5 == mod


def test_windows_against_linear_scan(orc):
    """filter.cu:75-127 lower/upper bound == linear count of r < lo, r <= hi."""
    rng = np.random.default_rng(7)
    r = np.round(rng.uniform(400, 4000, 5000), 1).astype(np.float32)  # many duplicates
    idx, rs = orc.sort_idx(r)
    assert np.all(np.diff(rs) >= 0)
    masses = np.concatenate([rng.uniform(390, 4010, 300), rs[::97] + np.float32(3.0),
                             rs[::89] - np.float32(3.0)]).astype(np.float32)
    b, e, n, tot = orc.find_spec_cands_by_mass(r, idx, masses, 3.0)
    for s, m in enumerate(masses):
        lo, hi = np.float32(m - np.float32(3.0)), np.float32(m + np.float32(3.0))
        assert b[s] == np.count_nonzero(rs < lo)
        assert e[s] == np.count_nonzero(rs <= hi)
    assert tot == int(n.sum())
    # modification windows, filter_by_mod.cu:35-143
    delta = np.array([0.0, 15.9949, 79.9663, -17.0], np.float32)
    b, e, sc, tot = orc.find_begin_end(r, idx, delta, np.float32(2000.5), 3.0)
    lo0, hi0 = np.float32(2000.5) - np.float32(3.0), np.float32(2000.5) + np.float32(3.0)
    acc = 0
    for i, d in enumerate(delta):
        assert b[i] == np.count_nonzero(rs < np.float32(lo0 - d))
        assert e[i] == np.count_nonzero(rs <= np.float32(hi0 - d))
        assert sc[i] == acc
        acc += int(e[i] - b[i])
    assert tot == acc


def test_sort_is_stable(orc):
    """sort.cu:28-35: equal keys keep index order."""
    keys = np.array([3, 1, 2, 1, 3, 0, 1], np.float32)
    idx, _ = orc.sort_idx(keys, init=10)
    assert idx.tolist() == [15, 11, 13, 16, 12, 10, 14]


def _exact_peaks(masses, r, zmax):
    """ion_series.cu:111-156 in exact decimal arithmetic: per ladder position
    and charge the m/z of A, B, B+1, B-1, B-H2O, B-NH3, Y, Y+1, Y-1, Y-NH3."""
    D = lambda v: Decimal(float(np.float32(v)))  # noqa: E731
    H, H2O, NH3, CO = D(1.0078246), D(18.01056), D(17.02655), D(27.9949)
    bw = D(1.0005079)
    out = []
    b = Decimal(0)
    for m in masses:
        b += D(m)
        y = D(r) - b
        for z in range(1, zmax + 1):
            Z = Decimal(z)
            A = (b - CO + Z * H) / Z
            B = (b + Z * H) / Z
            Y = (y + H2O + Z * H) / Z
            q = [A / bw, B / bw, None, None, (B - H2O / Z) / bw, (B - NH3 / Z) / bw,
                 Y / bw, None, None, (Y - NH3 / Z) / bw]
            out.append(q)
    return out


@pytest.mark.parametrize("pep", ["GK", "SAMPLER", "PEPTIDEK", "WYWYWYWR", "ACDEFGHIKLMNPQRSTVWY"])
def test_ladder_bins_match_exact_arithmetic(orc, pep):
    """ion_series.cu:59-156: every bin equals round(exact m/z / bin width)
    unless the exact quotient lies within 1e-3 of a rounding boundary."""
    mass = orc.mass_table()
    masses = np.array([mass[ord(c) - 65] for c in pep], np.float32)
    r = np.float32(0)
    for m in masses:
        r = np.float32(r + m)
    for zmax in (1, 2, 3, 4):
        bins, w = orc.peaks_list(masses, r, zmax)
        assert len(bins) == 10 * len(pep) * zmax
        exact = _exact_peaks(masses, r, zmax)
        bins = bins.reshape(-1, 10)
        w = w.reshape(-1, 10)
        assert (w == np.array([10, 50, 25, 25, 10, 10, 50, 25, 25, 10], np.uint8)).all()
        checked = 0
        for row, q in zip(bins, exact):
            for j, v in enumerate(q):
                if v is None:
                    continue
                frac = v - int(v)
                if abs(frac - Decimal("0.5")) < Decimal("1e-3"):
                    continue
                assert row[j] == int(v.to_integral_value()), (pep, zmax, j, v)
                checked += 1
            # x+1 / x-1 are index neighbours of the B and Y bins (Q18)
            assert row[2] == row[1] + 1 and row[3] == row[1] - 1
            assert row[7] == row[6] + 1 and row[8] == row[6] - 1
        assert checked > 0


def test_dense_row_is_max_of_peak_list(orc):
    """ion_series.cu:104-108 addIon: keep 0 <= x < len, max weight per bin."""
    rng = np.random.default_rng(3)
    mass = orc.mass_table()
    letters = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", np.uint8)
    for trial in range(40):
        pep = letters[rng.integers(0, 20, rng.integers(2, 40))]
        ions = orc.prepare_ions(pep)
        masses = mass[pep - 65]
        r = np.float32(0)
        for m in masses:
            r = np.float32(r + m)
        zmax = int(rng.integers(1, 5))
        L = int(rng.integers(200, 3000))
        dense = orc.peaks_dense(ions, 0, len(pep), r, mass, zmax, L)
        bins, w = orc.peaks_list(masses, r, zmax)
        ref = np.zeros(L, np.uint8)
        for x, wt in zip(bins, w):
            if 0 <= x < L:
                ref[x] = max(ref[x], wt)
        assert (dense == ref).all()


def test_fixed_point_score_is_exact(orc):
    """The exact score equals the float64 dot within 2^-37 per term."""
    rng = np.random.default_rng(11)
    for _ in range(50):
        L = int(rng.integers(100, 2100))
        w = rng.choice(np.array([0, 0, 0, 0, 10, 25, 50], np.uint8), L)
        x = rng.normal(0, 5, L).astype(np.float32)
        fx = orc.score_fx(w, x)
        ref = float(np.dot(w.astype(np.float64), x.astype(np.float64)))
        assert abs(orc.fx_to_double(fx) - ref) <= 1e-10 * max(1.0, abs(ref)) + L * 2.0**-37
        f32 = float(orc.score_f32(w, x))
        assert abs(f32 - ref) <= 1e-5 * np.sum(np.abs(w * x.astype(np.float64))) + 1e-6


def _f32_window_sum(a, st, en):
    s = np.float32(0)
    for j in range(st, en):
        s = np.float32(s + a[j])
    return s


def test_preprocess_closed_forms(orc):
    """Correlation.hs:37-109 on constructed spectra."""
    bw = np.float32(1.0005079)
    # one peak: normalised value 50, x_p = 50 - 50/150, neighbours -50/150
    pcr, z = np.float32(900.0), np.float32(2.0)
    p = 400
    mz = np.array([np.float32(p) * bw, 1999.0], np.float32)
    it = np.array([4.0, 0.0], np.float32)
    rc, x = orc.preprocess(pcr, z, mz, it, mean_mode=0)
    assert rc == 0
    # len = 75 + ceiling(min(50 + pcr*z, trunc(1999/10)*10)) = 75 + 1850
    assert len(x) == 75 + 1850
    assert x[p] == np.float32(np.float32(50) - np.float32(np.float32(50) / np.float32(150)))
    assert x[p + 30] == np.float32(-(np.float32(50) / np.float32(150)))
    assert x[p - 75] == x[p + 30] and x[p - 76] == 0 and x[p + 76] == 0
    # i < 75: the window is a[0..150] for every such i (Q13)
    p = 100
    mz = np.array([np.float32(p) * bw, 1999.0], np.float32)
    rc, x = orc.preprocess(pcr, z, mz, it, mean_mode=0)
    assert np.all(x[:75] == x[0]) and x[0] == np.float32(-(np.float32(50) / np.float32(150)))
    assert x[176] == 0 and x[175] == x[0]
    # a flat run of equal peaks: a = 50 on the run, 151 terms / 150 inside it
    run = np.arange(500, 900)
    mz = np.concatenate([run.astype(np.float32) * bw, [1999.0]]).astype(np.float32)
    it = np.concatenate([np.full(len(run), 9.0), [0.0]]).astype(np.float32)
    for mode in (0, 1):
        rc, x = orc.preprocess(pcr, z, mz, it, mean_mode=mode)
        inner = x[600:800]
        if mode == 0:
            s = _f32_window_sum(np.full(151, 50, np.float32), 0, 151)
            assert np.all(inner == np.float32(np.float32(50) - np.float32(s / np.float32(150))))
        else:
            assert np.all(inner == np.float32(50.0 - 7550.0 / 150.0))
    # regions: a peak in region k is normalised by that region's max only
    sel = (len(x) - 75) // 10
    mz = np.array([10.0 * bw, (sel + 10) * bw, (sel + 20) * bw, 1999.0], np.float32)
    it = np.array([1.0, 4.0, 1.0, 0.0], np.float32)
    rc, x = orc.preprocess(pcr, z, mz, it, mean_mode=1, remove_precursor=False)
    # region 0 holds one peak (-> 50); region 1 holds sqrt 4 = 2 (-> 50) and
    # sqrt 1 = 1 (-> 25); the windows of bins sel+10 and sel+20 hold both
    assert x[10] == np.float32(50.0 - 50.0 / 150.0)
    assert x[sel + 10] == np.float32(50.0 - 75.0 / 150.0)
    assert x[sel + 20] == np.float32(25.0 - 75.0 / 150.0)
    # precursor removal: peaks within 15 of the precursor are dropped (Q12)
    mz = np.array([880.0, 885.5, 905.0, 914.9, 915.0, 1999.0], np.float32)
    it = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 0.0], np.float32)
    rc, x = orc.preprocess(pcr, z, mz, it, mean_mode=1)
    kept = [int(np.rint(np.float32(v) / bw)) for v in (880.0, 915.0)]
    dropped = [int(np.rint(np.float32(v) / bw)) for v in (885.5, 905.0, 914.9)]
    xm = x + np.float32(0)
    assert all(xm[k] > 40 for k in kept) and all(xm[k] < 1 for k in dropped)
    # sel == 0: the reference divides by zero -> error
    # (peaks only below 10 m/z: cutoff 0, len 75, sel = 0 div 10)
    rc, _ = orc.preprocess(pcr, z, np.array([5.0, 9.0], np.float32),
                           np.array([1.0, 1.0], np.float32))
    assert rc == -2
    # last peak at 15 m/z: cutoff 10, len 85, sel = 1 is still valid
    rc, x = orc.preprocess(pcr, z, np.array([5.0, 15.0], np.float32),
                           np.array([1.0, 1.0], np.float32))
    assert rc == 0 and len(x) == 85


def test_ranking_ties_prefer_lower_position(orc):
    """Search.hs:119-123,:276 + stable sort: equal scores rank by position."""
    from hfx_amd import synth

    mass = orc.mass_table()
    # the same tryptic peptide in three proteins -> identical residual and score
    prots = [b"AAKSAGEPLLTRGG", b"CCRSAGEPLLTRWW", b"SAGEPLLTRK"]
    seq = np.frombuffer(b"".join(prots), np.uint8)
    seg = np.cumsum([0] + [len(p) for p in prots]).astype(np.uint32)
    r, tc, tn = orc.digest(seq, seg, mass, rule=1, missed=0, min_mass=200)
    ions = orc.prepare_ions(seq)
    idx, rs = orc.sort_idx(r)
    target = [i for i in range(len(r)) if bytes(seq[tc[i]:tn[i] + 1]) == b"SAGEPLLTR"]
    assert len(target) == 3
    prot = synth.Proteins(seq=seq, seg=seg)
    # spectrum made from the first copy: all three copies tie at the top
    sub = np.array([target[0]])
    sp = synth.make_spectra(prot, r[sub], tc[sub], tn[sub], 1, seed=5)
    X, lens, st = orc.preprocess_batch(sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, 2112)
    masses = np.array([orc.ms2_mass(p, c) for p, c in zip(sp.pcr, sp.charge)], np.float32)
    zmax = np.ones(1, np.uint32)
    res = orc.search_plain(ions, r, tc, tn, mass, idx, X, lens, zmax, masses, 50.0, 5, True)
    fx = res["all_fx"][: int(res["num"][0])]
    order = sorted(range(len(fx)), key=lambda i: (-fx[i], i))
    cnt = int(res["top_cnt"][0])
    assert res["top_pos"][0][:cnt].tolist() == order[:cnt]
    # the three copies tie, rank first, in ascending position (= fragment
    # index, the stable residual sort's tie order)
    top3 = res["top_pos"][0][:3]
    assert len(set(fx[top3].tolist())) == 1
    pep_of_pos = idx[res["begin"][0] + top3]
    assert pep_of_pos.tolist() == sorted(target)


def test_digest_hand_example(orc):
    """Fragment.hs:363-405 on 'AKPGRK' (rule 1, the reference cuts before P)."""
    mass = orc.mass_table()
    seq = np.frombuffer(b"AKPGRK", np.uint8)
    seg = np.array([0, 6], np.uint32)
    r, tc, tn = orc.digest(seq, seg, mass, rule=1, missed=1, min_mass=0, max_mass=1e9)
    m = {c: mass[ord(c) - 65] for c in "AKPGR"}
    ak = np.float32(m["A"] + m["K"])
    pgr = np.float32(np.float32(m["P"] + m["G"]) + m["R"])
    k = m["K"]
    # every piece ends in K/R: tn is the inclusive last index throughout
    assert tc.tolist() == [0, 0, 2, 2, 5]
    assert tn.tolist() == [1, 4, 4, 5, 5]
    assert r.tolist() == [ak, np.float32(ak + pgr), pgr, np.float32(pgr + k), k]
    # a protein ending without a cleavage residue: its last piece has tn = end
    seq2 = np.frombuffer(b"AKPGRKW", np.uint8)
    r3, tc3, tn3 = orc.digest(seq2, np.array([0, 7], np.uint32), mass, rule=1, missed=1,
                              min_mass=0, max_mass=1e9)
    assert tc3.tolist() == [0, 0, 2, 2, 5, 5, 6]
    assert tn3.tolist() == [1, 4, 4, 5, 5, 7, 7]
    # P rule extension: no cut after K before P
    r2, tc2, tn2 = orc.digest(seq, seg, mass, rule=1, missed=0, min_mass=0, max_mass=1e9,
                              p_rule=1)
    assert tc2.tolist() == [0, 5] and tn2.tolist() == [4, 5]


@pytest.mark.parametrize("d", [1.0005079, 1.0, 2.0, 3.0, 4.0])
def test_fast_division_is_exact(orc, d):
    """The dense GPU kernel divides by the bin width and the fragment charge
    as q = x*R, r = fma(-q, d, x), fma(r, R, q) with R = RN(1/d).  Every float
    x in [1e-3, 65536) must give the correctly rounded x / d (exhaustive:
    ~2.2e8 values per divisor); sign symmetry covers negative x, and outside
    that range every bin rounds to 0 or lies beyond the last bin."""
    assert orc.check_fast_div(np.float32(d), np.float32(1e-3), np.float32(65536.0)) == 0


def test_fast_division_check_detects_a_wrong_reciprocal(orc):
    """Planted mistake: without the FMA correction, q = x * RN(1/d) is off by
    an ulp for a large share of x, and the checker must count them."""
    for d in (1.0005079, 3.0):
        assert orc.check_fast_div(np.float32(d), np.float32(1.0), np.float32(2.0), False) > 1000
    # powers of two have exact reciprocals: nothing to correct
    assert orc.check_fast_div(np.float32(2.0), np.float32(1.0), np.float32(2.0), False) == 0


# ---------------------------------------------------------------------------
# mass table and precursor masses
# ---------------------------------------------------------------------------

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
H_PROTON = Decimal("1.007276")
H2O = Decimal("18.010565")
E_MASS = 0.000549  # electron mass: hydrogen atom minus proton


def _reference_matches():
    with open(os.path.join(GOLDEN, "reference_out_peptides.json")) as f:
        files = json.load(f)
    return [(m["peptide"], m["mh"]) for e in files for m in e["matches"]]


def _f32_residue_sum(mass, pep):
    r = np.float32(0)
    for c in pep:
        r = np.float32(r + mass[ord(c) - 65])
    return r


def test_mass_table_against_reference_outputs(orc):
    """Mass.hs:147-215 (the monoisotopic table), Config.hs:266-272 (static
    add) and Location.hs:58 ((M+H)+ = residues + H + H2O) against the 35
    matches printed in the reference's own result files (tests/spectra/*.out,
    written with add_C = 57.0 as in hfx.params; all 20 standard residues
    occur).  The files were written with the hydrogen atom's mass where the
    current constant is the proton's, so every printed value sits one
    electron mass (5.49e-4) above the residue sum + 1.007276 + 18.010565;
    printing to 4 decimals and the FP32 residue sum leave 2e-4 around that.
    A transposed or mistyped entry moves a sum by far more."""
    mass = orc.mass_table({"C": 57.0})
    rows = _reference_matches()
    assert len(rows) == 35
    assert set("".join(p for p, _ in rows)) == set("ACDEFGHIKLMNPQRSTVWY")
    for pep, mh in rows:
        calc = float(_f32_residue_sum(mass, pep)) + float(H_PROTON + H2O)
        assert abs(float(mh) - calc - E_MASS) <= 2e-4, (pep, mh, calc)
    # the static modification adds to its residue only
    m0, m1 = orc.mass_table(), orc.mass_table({"C": 57.0})
    assert m1[2] == np.float32(m0[2] + np.float32(57.0))
    assert (np.delete(m1, 2) == np.delete(m0, 2)).all()


def test_average_masses_exceed_monoisotopic(orc):
    """Mass.hs (average table): every residue's average mass exceeds its
    monoisotopic mass by the heavy-isotope share, 0.03-0.2 Da for the
    standard residues; a transposed entry breaks this."""
    mono, avg = orc.mass_table(mono=True), orc.mass_table(mono=False)
    for c in "ACDEFGHIKLMNPQRSTVWY":
        d = float(avg[ord(c) - 65]) - float(mono[ord(c) - 65])
        assert 0.03 < d < 0.2, (c, d)


@pytest.mark.parametrize("z", [1, 2, 3, 4])
def test_precursor_masses_physical_identities(orc, z):
    """Spectrum/Data.hs:58-59 ms2Mass and DTA.hs:51.  For a peptide of
    residue sum R (neutral mass M = R + H2O) the z+ precursor has
    m/z = (M + z*H+)/z, and ms2Mass of that m/z must give back R; a DTA
    file's (M+H)+ line converts to the same m/z.  Checked in exact decimals
    on the reference's 35 printed peptides: a dropped or sign-flipped term
    is off by at least 1 Da, FP32 rounding by at most 1e-3."""
    mass = orc.mass_table({"C": 57.0})
    for pep, _ in _reference_matches():
        R = Decimal(float(_f32_residue_sum(mass, pep)))
        M = R + H2O
        mz = (M + z * H_PROTON) / z
        got = orc.ms2_mass(np.float32(float(mz)), np.float32(z))
        assert abs(Decimal(float(got)) - R) <= Decimal("1e-3"), (pep, z, got, R)
        MH = M + H_PROTON
        pcr = orc.dta_precursor(np.float32(float(MH)), np.float32(z))
        assert abs(Decimal(float(pcr)) - mz) <= Decimal("3e-4"), (pep, z, pcr, mz)


# ---------------------------------------------------------------------------
# variable-modification enumeration
# ---------------------------------------------------------------------------

VAR = {"M": 15.9949, "S": 79.9663, "T": 79.9663}
MAX_MA = 3  # at most 2 modified residues per peptide


@pytest.fixture(scope="module")
def mdb(orc):
    """150 proteins, tryptic with one missed cleavage, 5-25 residues; 12
    spectrum masses, 8 of them a modified peptide's mass."""
    from hfx_amd import synth

    prot = synth.make_proteins(150, seed=61)
    mass = orc.mass_table()
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=1, p_rule=1, min_len=5,
                           max_len=25)
    ma, mam, cnt, msum, dl = orc.mod_info(VAR, MAX_MA)
    rng = np.random.default_rng(62)
    seqs = [bytes(prot.seq[tc[p]:tn[p]]) for p in range(len(r))]
    modable = [p for p in range(len(r)) if sum(seqs[p].count(a) for a in b"MST") >= 2]
    masses = []
    for k in range(8):
        p = modable[int(rng.integers(len(modable)))]
        m = int(rng.integers(len(dl)))
        masses.append(np.float32(r[p] + dl[m]) + np.float32(rng.uniform(-0.5, 0.5)))
    masses += list(rng.uniform(900, 2500, 4))
    return dict(seq=prot.seq, mass=mass, r=r, tc=tc, tn=tn, seqs=seqs, ma=ma, mam=mam, cnt=cnt,
                msum=msum, dl=dl, masses=np.array(masses, np.float32), eps=np.float32(1.5),
                order=np.argsort(r, kind="stable"))


def _brute_rows(d, mass_s):
    """(peptide, mod, counts) rows of one spectrum by a linear scan of the
    residual-sorted peptides per modification, counts by bytes.count."""
    rs = d["r"][d["order"]]
    rows = []
    for m, delta in enumerate(d["dl"]):
        lo = np.float32(np.float32(mass_s - d["eps"]) - delta)
        hi = np.float32(np.float32(mass_s + d["eps"]) - delta)
        for p in d["order"][(rs >= lo) & (rs <= hi)]:
            c = [d["seqs"][p].count(bytes([a])) for a in d["ma"]]
            if all(c[i] >= d["cnt"][m][i] for i in range(len(c))):
                rows.append((int(p), m, c))
    return rows


def _oracle_rows(orc, d):
    idx, _ = orc.sort_idx(d["r"])
    S, M = len(d["masses"]), len(d["dl"])
    b = np.zeros(S * M, np.uint32)
    e = np.zeros(S * M, np.uint32)
    sc = np.zeros(S * M, np.uint32)
    cand = np.zeros(S, np.uint32)
    for s in range(S):
        bb, ee, ss, tot = orc.find_begin_end(d["r"], idx, d["dl"], d["masses"][s], d["eps"])
        b[s * M:(s + 1) * M], e[s * M:(s + 1) * M], sc[s * M:(s + 1) * M] = bb, ee, ss
        cand[s] = tot
    scan = np.concatenate([[0], np.cumsum(cand)]).astype(np.uint32)
    ions = orc.prepare_ions(d["seq"], d["ma"])
    return orc.find_modable_peptides(scan, scan[-1], ions, d["tc"], d["tn"], idx, b, e, sc,
                                     d["cnt"], d["ma"])


def test_find_modable_peptides_brute_force(orc, mdb):
    """filter_by_mod.cu:150-378 findModablePeptides (with findBeginEnd_f,
    :35-143): rows spectrum-major, modification-minor, residual order within
    a window; a row is kept when every modable acid occurs in [tc, tn) at
    least as often as the modification asks; spec_num_valid_scan counts the
    kept rows of earlier spectra."""
    d = mdb
    nv, pep, pmod, pcnt, vs = _oracle_rows(orc, d)
    want, scan = [], []
    for ms in d["masses"]:
        scan.append(len(want))
        want += _brute_rows(d, ms)
    assert nv == len(want) > 100
    assert pep.tolist() == [w[0] for w in want]
    assert pmod.tolist() == [w[1] for w in want]
    assert pcnt.reshape(nv, -1).tolist() == [w[2] for w in want]
    assert vs.tolist() == scan
    # the windows also held rows that lack a modable acid and were dropped
    rs = d["r"][d["order"]]
    in_windows = sum(int(np.count_nonzero(
        (rs >= np.float32(np.float32(ms - d["eps"]) - dl)) &
        (rs <= np.float32(np.float32(ms + d["eps"]) - dl))))
        for ms in d["masses"] for dl in d["dl"])
    assert in_windows > nv


def test_calc_total_mod_cands_by_enumeration(orc, mdb):
    """calc_mod_cands.cu:27-119: modified peptides per row = the number of
    ways to pick d_i of the D_i occurrences of every modable acid, counted
    here by enumerating the subsets; the row of suffix products."""
    d = mdb
    nv, pep, pmod, pcnt, vs = _oracle_rows(orc, d)
    na = len(d["ma"])
    tot, num, comb = orc.calc_total_mod_cands(d["cnt"], pep, pmod, pcnt, na)
    pcnt = pcnt.reshape(nv, na)
    comb = comb.reshape(nv, na)
    want_total = 0
    for i in range(nv):
        ways = [sum(1 for _ in itertools.combinations(range(int(pcnt[i][a])),
                                                       int(d["cnt"][pmod[i]][a])))
                for a in range(na)]
        assert num[i] == math.prod(ways)
        assert comb[i].tolist() == [math.prod(ways[a:]) for a in range(na)]
        want_total += math.prod(ways)
    assert tot == want_total > nv


def test_prepare_gen_mod_expansion(orc, mdb):
    """prepare_gen_mod.cu:23-129: one record per modified peptide, (peptide,
    modification, rank 0..n-1, row), and the exclusive scan of the
    modification's residue count."""
    d = mdb
    nv, pep, pmod, pcnt, vs = _oracle_rows(orc, d)
    tot, num, comb = orc.calc_total_mod_cands(d["cnt"], pep, pmod, pcnt, len(d["ma"]))
    lu, mp_pep, mp_mod, mp_rank, mp_ith, mp_scan = orc.prepare_gen_mod(d["msum"], pep, pmod, num,
                                                                       tot)
    rec = [(int(pep[i]), int(pmod[i]), k, i) for i in range(nv) for k in range(int(num[i]))]
    assert list(zip(mp_pep.tolist(), mp_mod.tolist(), mp_rank.tolist(), mp_ith.tolist())) == rec
    sums = [int(d["msum"][m]) for _, m, _, _ in rec]
    assert mp_scan.tolist() == [0] + list(itertools.accumulate(sums))[:-1]
    assert lu == sum(sums)


def test_rank_split_is_a_mixed_radix_bijection(orc, mdb):
    """gen_mod_pep.cu:127-157 add_ma_ranks: rank -> per-acid ranks is the
    mixed-radix decomposition over C(D_i, d_i) (digit i = rank // prod_{j>i}
    C_j mod C_i), a bijection of [0, prod C) onto the digit boxes; every
    modified peptide's unrank slice is the per-acid unrankComb of its
    digits (gen_mod_pep.cu:159-312)."""
    d = mdb
    nv, pep, pmod, pcnt, vs = _oracle_rows(orc, d)
    na = len(d["ma"])
    tot, num, comb = orc.calc_total_mod_cands(d["cnt"], pep, pmod, pcnt, na)
    lu, mp_pep, mp_mod, mp_rank, mp_ith, mp_scan = orc.prepare_gen_mod(d["msum"], pep, pmod, num,
                                                                       tot)
    unrank = orc.gen_mod_cands(lu, d["cnt"], mp_ith, mp_rank, mp_scan, pmod, pcnt, comb, na)
    pcnt = pcnt.reshape(nv, na)
    comb = comb.reshape(nv, na)
    row_first = np.concatenate([[0], np.cumsum(num)])
    boxes_seen = 0
    for i in range(nv):
        m = int(pmod[i])
        C_ = [math.comb(int(pcnt[i][a]), int(d["cnt"][m][a])) for a in range(na)]
        digits = set()
        for rank in range(int(num[i])):
            want = [(rank // math.prod(C_[a + 1:])) % C_[a] for a in range(na)]
            got = orc.add_ma_ranks(rank, comb[i]).tolist()
            assert got == want, (i, rank)
            digits.add(tuple(got))
            mp = int(row_first[i]) + rank
            sl = unrank[int(mp_scan[mp]): int(mp_scan[mp]) + int(d["msum"][m])].tolist()
            exp = []
            for a in range(na):
                k = int(d["cnt"][m][a])
                if k:
                    exp += orc.unrank_comb(int(pcnt[i][a]), k, want[a]).tolist()
            assert sl == exp, (i, rank)
        assert digits == set(itertools.product(*[range(c) for c in C_]))
        boxes_seen += sum(c > 1 for c in C_) >= 2
    assert boxes_seen > 0  # rows whose rank splits over two acids


def test_modified_ladder_hand_built(orc):
    """check_series.cu:61-99 genMIons + mion_series.cu:226,277: the i-th
    occurrence of a modable acid (counting that acid only) is modified when i
    is in its unrank slice; a modified residue weighs table + variable mass;
    the residual is r + delta.  Positions below are worked out by hand for
    AMSGMSTPMK (M at 1, 4, 8; S at 2, 5; T at 6); every bin is the rounded
    exact m/z of the hand-modified residue masses, except the bins an ion
    whose exact quotient lies within 1e-3 of a rounding boundary could
    reach, which are left out of the comparison."""
    pep = "AMSGMSTPMK"
    ma, mam = b"MST", np.array([15.9949, 79.9663, 79.9663], np.float32)
    mass = orc.mass_table()
    ions = orc.prepare_ions(np.frombuffer(pep.encode(), np.uint8), ma)
    base = np.array([mass[ord(c) - 65] for c in pep], np.float32)
    r = _f32_residue_sum(mass, pep)
    cases = [  # (row of modified counts M,S,T; unrank slice; modified positions)
        ([1, 0, 0], [2], {8: 0}),
        ([1, 1, 0], [0, 1], {1: 0, 5: 1}),
        ([0, 1, 1], [0, 0], {2: 1, 6: 2}),
        ([2, 0, 0], [2, 0], {8: 0, 1: 0}),
        ([0, 2, 0], [1, 0], {5: 1, 2: 1}),
    ]
    L = 2100
    for row, slice_, pos in cases:
        delta = np.float32(0)
        for c, mv in zip(row, mam):
            delta = np.float32(delta + np.float32(mv * np.float32(c)))
        masses = base.copy()
        for k, a in pos.items():
            masses[k] = np.float32(masses[k] + mam[a])
        rr = np.float32(r + delta)
        for zmax in (1, 2, 3):
            want = np.zeros(L, np.uint8)
            sure = np.ones(L, bool)
            wts = [10, 50, 25, 25, 10, 10, 50, 25, 25, 10]
            nbr = {0: (0,), 1: (-1, 0, 1), 4: (0,), 5: (0,), 6: (-1, 0, 1), 9: (0,)}
            for q in _exact_peaks(masses, rr, zmax):
                xb, xy = int(q[1].to_integral_value()), int(q[6].to_integral_value())
                bins = [int(q[0].to_integral_value()), xb, xb + 1, xb - 1,
                        int(q[4].to_integral_value()), int(q[5].to_integral_value()),
                        xy, xy + 1, xy - 1, int(q[9].to_integral_value())]
                for j, v in enumerate(q):
                    if v is not None and abs((v - int(v)) - Decimal("0.5")) < Decimal("1e-3"):
                        for o in nbr[j]:
                            for k in (int(v) + o, int(v) + 1 + o):
                                if 0 <= k < L:
                                    sure[k] = False
                for x, w in zip(bins, wts):
                    if 0 <= x < L:
                        want[x] = max(want[x], w)
            got = orc.mpeaks_dense(ions, 0, len(pep), r, mass, zmax, L, row, delta, ma, mam,
                                   slice_)
            assert np.count_nonzero(want[sure]) > 20 * zmax and (~sure).sum() <= 4
            assert (got[sure] == want[sure]).all(), (row, slice_, zmax)


def _dense_from_list(masses, r, zmax, L, orc):
    bins, w = orc.peaks_list(masses, r, zmax)
    row = np.zeros(L, np.uint8)
    for x, wt in zip(bins, w):
        if 0 <= x < L:
            row[x] = max(row[x], wt)
    return row


def _fx(row, x):
    nz = np.nonzero(row)[0]
    t = np.ldexp(row[nz].astype(np.float64) * x[nz].astype(np.float64), 36)
    return int(sum(int(v) for v in np.rint(t)))


def test_search_mods_brute_force(orc, mdb):
    """Search.hs:282-622 searchWithMods, composed here from independent
    pieces: windows and modable rows by linear scan (as above), ranks by the
    mixed radix, occurrence subsets by unrankComb (pinned above), modified
    residue masses placed by str.find, theoretical rows from the pinned peak
    list, exact scores by float64 rounding, ranking by Python's sort on
    (score desc, position asc)."""
    d = mdb
    S, topn = len(d["masses"]), 5
    rng = np.random.default_rng(63)
    stride = 2112
    X = np.zeros((S, stride), np.float32)
    lens = rng.integers(1500, stride + 1, S).astype(np.uint32)
    for s in range(S):
        X[s, : lens[s]] = rng.normal(0, 3, lens[s]).astype(np.float32)
    zmax = rng.integers(1, 5, S).astype(np.uint32)
    idx, _ = orc.sort_idx(d["r"])
    ions = orc.prepare_ions(d["seq"], d["ma"])
    ref = orc.search_mods(ions, d["r"], d["tc"], d["tn"], d["mass"], idx, d["ma"], d["mam"],
                          d["cnt"], d["msum"], d["dl"], X, lens, zmax, d["masses"], d["eps"], topn)
    na = len(d["ma"])
    k_mp = 0
    for s in range(S):
        scored = []
        for p, m, c in _brute_rows(d, d["masses"][s]):
            C_ = [math.comb(c[a], int(d["cnt"][m][a])) for a in range(na)]
            for rank in range(math.prod(C_)):
                digits = [(rank // math.prod(C_[a + 1:])) % C_[a] for a in range(na)]
                masses = np.array([d["mass"][ch - 65] for ch in d["seqs"][p]], np.float32)
                sl = []
                for a in range(na):
                    k = int(d["cnt"][m][a])
                    if not k:
                        continue
                    occ = orc.unrank_comb(c[a], k, digits[a]).tolist()
                    sl += occ
                    at = [i for i, ch in enumerate(d["seqs"][p]) if ch == d["ma"][a]]
                    for o in occ:
                        masses[at[o]] = np.float32(masses[at[o]] + d["mam"][a])
                assert ref["mpep_pep"][k_mp] == p and ref["mpep_mod"][k_mp] == m
                off = int(ref["mpep_unrank_off"][k_mp])
                assert ref["unrank"][off: off + len(sl)].tolist() == sl
                k_mp += 1
                rr = np.float32(d["r"][p] + d["dl"][m])
                row = _dense_from_list(masses, rr, int(zmax[s]), int(lens[s]), orc)
                scored.append(_fx(row, X[s]))
        assert ref["spec_num_mpep"][s] == len(scored)
        order = sorted(range(len(scored)), key=lambda i: (-scored[i], i))[:topn]
        c = int(ref["top_cnt"][s])
        assert c == min(topn, len(scored))
        assert ref["top_pos"][s, :c].tolist() == order
        assert ref["top_fx"][s, :c].tolist() == [scored[i] for i in order]
    assert k_mp == len(ref["mpep_pep"]) > 200


def test_search_plain_brute_force(orc, mdb):
    """Search.hs:98-124 searchWithoutMods (filter.cu:75-127 windows,
    ion_series.cu ladders, the dense dot, Search.hs:119-123 ranking) composed
    from the pinned pieces: linear-scan windows, the pinned peak list,
    float64-rounded exact scores, Python's sort; FP32 scores within
    1e-5 * sum w|x| of the exact."""
    d = mdb
    S, topn = len(d["masses"]), 4
    rng = np.random.default_rng(64)
    stride = 2112
    X = np.zeros((S, stride), np.float32)
    lens = rng.integers(1000, stride + 1, S).astype(np.uint32)
    for s in range(S):
        X[s, : lens[s]] = rng.normal(0, 3, lens[s]).astype(np.float32)
    zmax = np.array([1, 2, 3, 4, 5, 1, 2, 3, 4, 2, 3, 1], np.uint32)[:S]
    idx, _ = orc.sort_idx(d["r"])
    ions = orc.prepare_ions(d["seq"])
    ref = orc.search_plain(ions, d["r"], d["tc"], d["tn"], d["mass"], idx, X, lens, zmax,
                           d["masses"], d["eps"], topn, want_all=True)
    rs = d["r"][d["order"]]
    for s in range(S):
        lo = np.float32(d["masses"][s] - d["eps"])
        hi = np.float32(d["masses"][s] + d["eps"])
        win = np.nonzero((rs >= lo) & (rs <= hi))[0]
        assert ref["num"][s] == len(win)
        if len(win):
            assert ref["begin"][s] == win[0]
        fx = []
        for i in win:
            p = d["order"][i]
            masses = np.array([d["mass"][ch - 65] for ch in d["seqs"][p]], np.float32)
            z = int(zmax[s]) if zmax[s] <= 4 else 0  # no ladder above fragment charge 4
            row = _dense_from_list(masses, d["r"][p], z, int(lens[s]), orc)
            fx.append(_fx(row, X[s]))
            a = int(ref["off"][s]) + len(fx) - 1
            assert ref["all_fx"][a] == fx[-1]
            scale = float(np.sum(row * np.abs(X[s, : lens[s]].astype(np.float64))))
            assert abs(float(ref["all_f32"][a]) - fx[-1] * 2.0**-36) <= 1e-5 * scale + 1e-6
        order = sorted(range(len(fx)), key=lambda i: (-fx[i], i))[:topn]
        c = int(ref["top_cnt"][s])
        assert ref["top_pos"][s, :c].tolist() == order
        assert ref["top_fx"][s, :c].tolist() == [fx[i] for i in order]


def test_prepare_scoring_is_window_concatenation(orc):
    """prepare_scoring.cu:20-67: spectrum s's candidates are
    idx_sorted[begin_s : begin_s + num_s], spectra in order."""
    rng = np.random.default_rng(65)
    idx = rng.permutation(1000).astype(np.uint32)
    begin = rng.integers(0, 900, 20).astype(np.uint32)
    num = rng.integers(0, 100, 20).astype(np.uint32)
    out = orc.prepare_scoring(idx, begin, num)
    want = np.concatenate([idx[b:b + n] for b, n in zip(begin, num)])
    assert out.tolist() == want.tolist()
