"""Pins of the CPU oracle against closed forms, brute force, independent
high-precision arithmetic and the reference's own known answers.

Each test names the reference function it pins.  A check that merely
re-typed the oracle's formula would pin nothing; these use a different
method (exact rationals, enumeration, linear scans, hand-derived values).
"""
import itertools
import math
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
