"""CPU tests of the C ABI library: it loads, exports every declared symbol,
and its host helpers (digestion, mass table, modification table, precursor
mass) agree bit for bit with the oracle.  No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import hfx_amd
from hfx_amd import pipeline, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "hfx_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt)
    skip = {"if", "for", "while", "sizeof", "defined", "return"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def test_library_exports_every_declared_symbol():
    L = hfx_amd.lib()
    decl = declared_functions()
    assert len(decl) > 30
    missing = [n for n in decl if not hasattr(L, n) and n not in hfx_amd._OPTIONAL]
    assert not missing, missing
    # the ctypes table covers the header too
    assert set(decl) <= set(hfx_amd.exported_symbols())


def test_mass_table_matches_oracle(orc):
    for static, mono in [({}, True), ({"C": 57.02146}, True), ({"C": 57.0, "M": 1.5}, False)]:
        assert (pipeline.mass_table(static, mono) == orc.mass_table(static, mono)).all()


@pytest.mark.parametrize("rule,missed,p_rule,lmin,lmax", [
    (1, 0, 0, 0, 0), (1, 2, 0, 0, 0), (1, 1, 1, 7, 30), (2, 1, 0, 0, 0), (13, 2, 1, 5, 0),
    (0, 0, 0, 0, 0), (6, 3, 0, 0, 25),
])
def test_digest_matches_oracle(orc, rule, missed, p_rule, lmin, lmax):
    prot = synth.make_proteins(300, seed=rule * 10 + missed)
    mass = orc.mass_table({"C": 57.02146})
    lo, hi = (400.0, 7200.0) if rule else (0.0, 1e9)
    a = pipeline.digest(prot.seq, prot.seg, mass, rule, missed, lo, hi, p_rule, lmin, lmax)
    b = orc.digest(prot.seq, prot.seg, mass, rule, missed, lo, hi, p_rule, lmin, lmax)
    assert len(a[0]) == len(b[0]) and len(a[0]) > 0
    for x, y in zip(a, b):
        assert (x == y).all()


def test_mod_info_matches_oracle(orc):
    L = hfx_amd.lib()
    for var, mx in [({"Q": -17.0}, 2), ({"M": 15.9949, "S": 79.9663, "T": 79.9663, "Y": 79.9663}, 4),
                    ({"M": 15.9949, "N": 0.984}, 20), ({}, 3)]:
        v = np.zeros(26, np.float32)
        for k, x in var.items():
            v[ord(k) - 65] = x
        ma = np.zeros(26, np.uint8)
        mam = np.zeros(26, np.float32)
        na = C.c_uint32()
        nm = C.c_uint32()
        cap = 1 << 14
        cnt = np.zeros(cap * 26, np.uint32)
        sm = np.zeros(cap, np.uint32)
        dl = np.zeros(cap, np.float32)
        hfx_amd.check(L.hfx_mod_info(v.ctypes.data, mx, ma.ctypes.data, mam.ctypes.data, C.byref(na),
                                     cnt.ctypes.data, sm.ctypes.data, dl.ctypes.data, cap,
                                     C.byref(nm)), "hfx_mod_info")
        o = orc.mod_info(var, mx)
        a, m = na.value, nm.value
        assert bytes(ma[:a]) == o[0]
        assert (mam[:a] == o[1]).all()
        assert (cnt[: m * a].reshape(m, a) == o[2]).all()
        assert (sm[:m] == o[3]).all() and (dl[:m] == o[4]).all()


def test_mod_info_against_brute_force():
    """hfx_mod_info walks only the count vectors with sum < max_ma; the table
    equals a filter of every vector in 0..max_ma per acid (lexicographic, the
    last acid fastest, zero dropped), and a table too large to hold is
    refused at once with its count (saturated) instead of enumerated."""
    import itertools
    rng = np.random.default_rng(4)
    for _ in range(40):
        acids = sorted(rng.choice(list("ACDEFGHIKLMNPQRSTVWY"), rng.integers(1, 6), replace=False))
        var = {a: float(np.float32(rng.uniform(-20, 80))) or 1.0 for a in acids}
        mx = int(rng.integers(0, 6))
        want = [v for v in itertools.product(range(mx + 1), repeat=len(acids))
                if 0 < sum(v) < mx]
        mi = pipeline.mod_info(var, mx)
        assert mi.mod_ma_count.tolist() == [list(v) for v in want], (var, mx)
        assert mi.mod_sum.tolist() == [sum(v) for v in want]
    L = hfx_amd.lib()
    v = np.full(26, 1.5, np.float32)
    ma = np.zeros(26, np.uint8)
    mam = np.zeros(26, np.float32)
    na, nm = C.c_uint32(), C.c_uint32()
    rc = L.hfx_mod_info(v.ctypes.data, 60, ma.ctypes.data, mam.ctypes.data, C.byref(na), None,
                        None, None, 1024, C.byref(nm))
    assert rc != 0 and na.value == 26 and nm.value == 2**32 - 1
    # C(28, 2) - 1 vectors of one or two sites over 26 acids
    rc = L.hfx_mod_info(v.ctypes.data, 3, ma.ctypes.data, mam.ctypes.data, C.byref(na), None,
                        None, None, 0, C.byref(nm))
    assert rc != 0 and nm.value == 377
    with pytest.raises(hfx_amd.HfxError):
        pipeline.mod_info({c: 1.0 for c in "ACDEFGHIKLMNPQRSTVWY"}, 40)


def test_precursor_helpers_match_oracle(orc):
    L = hfx_amd.lib()
    rng = np.random.default_rng(0)
    for _ in range(2000):
        p = np.float32(rng.uniform(300, 2000))
        z = np.float32(rng.integers(1, 5))
        assert np.float32(L.hfx_ms2_mass(p, z)) == orc.ms2_mass(p, z)
        assert np.float32(L.hfx_dta_precursor(p, z)) == orc.dta_precursor(p, z)


def test_pipeline_refuses_without_gpu():
    n = C.c_int(0)
    if hfx_amd.lib().hfx_device_count(C.byref(n)) == 0 and n.value > 0:
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError):
        pipeline.empty(16)


@pytest.mark.parametrize("case", ["short", "start", "decreasing", "beyond"])
def test_spectra_offsets_checked_before_upload(case):
    """Malformed peak offsets are refused on the host: the preprocessing
    kernel reads peaks off[s]..off[s+1] of every spectrum."""
    mz = np.linspace(200.0, 900.0, 30, dtype=np.float32)
    it = np.ones(30, np.float32)
    pcr = np.array([500.0, 600.0, 700.0], np.float32)
    ch = np.array([2.0, 2.0, 3.0], np.float32)
    off = {"short": [0, 10, 20], "start": [1, 10, 20, 30], "decreasing": [0, 20, 10, 30],
           "beyond": [0, 10, 20, 31]}[case]
    with pytest.raises(ValueError):
        pipeline.DeviceSpectra(None, pcr, ch, mz, it, np.array(off, np.uint32))


def test_batch_sizing_and_split():
    """hfx_batch_required_bytes grows with the spectra and the scores, the
    dense form needs more than the sparse one; hfx_batch_split
    (groupScorePlan, Execution.hs:77-93) covers the spectra with consecutive
    batches that each fit the budget and could not take their next spectrum,
    and refuses a spectrum that alone exceeds the budget."""
    L = hfx_amd.lib()

    def need(S, total, prec=0):
        b = C.c_uint64()
        assert L.hfx_batch_required_bytes(S, total, 5, 2_700_000, prec, C.byref(b)) == 0
        return b.value

    assert need(100, 10**6) < need(200, 10**6) < need(200, 2 * 10**6)
    assert need(1000, 10**6, 0) < need(1000, 10**6, 1)
    rng = np.random.default_rng(5)
    S = 5000
    num = rng.integers(0, 20000, size=S).astype(np.uint64)
    off = np.zeros(S + 1, np.uint64)
    np.cumsum(num, out=off[1:])
    for prec in (0, 1):
        budget = need(S, int(off[-1]), prec) // 7
        first = np.zeros(S + 2, np.uint32)
        nb = C.c_uint32()
        assert L.hfx_batch_split(off.ctypes.data, S, 5, 2_700_000, prec, budget, first.ctypes.data,
                                 len(first), C.byref(nb)) == 0
        k = nb.value
        assert 7 <= k <= 9
        cuts = first[: k + 1].astype(np.int64)
        assert cuts[0] == 0 and cuts[-1] == S and np.all(np.diff(cuts) > 0)
        for a, e in zip(cuts[:-1], cuts[1:]):
            assert need(int(e - a), int(off[e] - off[a]), prec) <= budget
            if e < S:
                assert need(int(e + 1 - a), int(off[e + 1] - off[a]), prec) > budget
    one = need(1, int(num.max()), 0)
    nb = C.c_uint32()
    rc = L.hfx_batch_split(off.ctypes.data, S, 5, 2_700_000, 0, one // 2, None, 0, C.byref(nb))
    assert rc == 3 and "alone needs" in hfx_amd.last_error()
    bad = off.copy()
    bad[7] = bad[8] + 1
    rc = L.hfx_batch_split(bad.ctypes.data, S, 5, 2_700_000, 0, one * 4, None, 0, C.byref(nb))
    assert rc != 0 and "decrease at spectrum 7" in hfx_amd.last_error()


def _reference_merge(plain, mods, n):
    """Search.hs:118-123 with sequestXC's retrieval (Search.hs:257-276),
    literally: each search returns `reverse` of its scores stably sorted
    ascending (b40c radix sort over window order), the two lists are
    concatenated modified first, stably sorted ascending by score
    (matchScoreOrder) and reversed."""
    def retrieved(items):
        return list(reversed(sorted(items, key=lambda t: t[0])))

    m = [(s, 0, p) for s, p in retrieved([(s, p) for p, s in enumerate(plain)])]
    mm = [(s, 1, p) for s, p in retrieved([(s, p) for p, s in enumerate(mods)])]
    return list(reversed(sorted(mm + m, key=lambda t: t[0])))[:n]


@pytest.mark.parametrize("seed", range(8))
def test_merge_rankings_against_reference_order(seed):
    """merge_rankings (score descending; unmodified first, then lower
    position, on equal scores) against a literal restatement of the
    reference's sorts and reverses, with many equal scores."""
    rng = np.random.default_rng(seed)
    S, N, n = 6, 5, 5
    plains = [list(rng.integers(0, 6, int(rng.integers(0, 12))).astype(float)) for _ in range(S)]
    modss = [list(rng.integers(0, 6, int(rng.integers(0, 12))).astype(float)) for _ in range(S)]

    def top(lst):
        order = sorted(range(len(lst)), key=lambda p: (-lst[p], p))[:N]
        pos = np.zeros(N, np.uint32)
        ex = np.zeros(N)
        pos[:len(order)] = order
        ex[:len(order)] = [lst[p] for p in order]
        return pos, ex, len(order)

    def result(lists, cls):
        tp = np.zeros((S, N), np.uint32)
        te = np.zeros((S, N))
        tc = np.zeros(S, np.uint32)
        for s, lst in enumerate(lists):
            tp[s], te[s], tc[s] = top(lst)
        z = np.zeros(S, np.uint32)
        if cls is pipeline.PlainResult:
            return cls(begin=z, num=z, off=z, top_pos=tp, top_exact=te, top_f32=te.astype(np.float32),
                       top_cnt=tc, flagged=z[:0])
        return cls(spec_cand=z, num_valid=0, num_mpep=0, len_unrank=0, mp_off=z, top_pos=tp,
                   top_exact=te, top_f32=te.astype(np.float32), top_cnt=tc, flagged=z[:0],
                   stages={})

    got = pipeline.merge_rankings(result(plains, pipeline.PlainResult),
                                  result(modss, pipeline.ModResult), n)
    for s in range(S):
        assert got[s] == _reference_merge(plains[s], modss[s], n), s


def test_stats_entry_points_reject_bad_arguments():
    """hfx_evalue_fit / hfx_sp_peaks / hfx_sp_ions validate before any device
    work: a zero bin width, a cut-off outside [0, 100], a lone second score
    set, too few flag words, a stride beyond the flag capacity."""
    L = hfx_amd.lib()
    buf = (C.c_uint64 * 64)()
    p = C.cast(buf, C.c_void_p)
    assert L.hfx_evalue_fit(None, p, None, None, p, p, None, None, 1, 0, 3.0, None) == 1
    assert L.hfx_evalue_fit(None, p, None, None, p, p, None, None, 1, 350, 101.0, None) == 1
    assert L.hfx_evalue_fit(None, p, None, None, p, p, p, None, 1, 350, 3.0, None) == 1
    assert L.hfx_evalue_fit(None, None, None, None, p, p, None, None, 1, 350, 3.0, None) == 1
    assert L.hfx_sp_peak_words(2112) == (2112 // 3 + 1 + 31) // 32
    sp = hfx_amd.HfxSpectra(p.value, 2112, p.value, p.value, p.value, 1)
    assert L.hfx_sp_peaks(None, p, L.hfx_sp_peak_words(2112) - 1, C.byref(sp), None) == 1
    big = hfx_amd.HfxSpectra(p.value, 24576, p.value, p.value, p.value, 1)
    assert L.hfx_sp_peaks(None, p, 4096, C.byref(big), None) == 1
    db = hfx_amd.HfxDb()
    assert L.hfx_sp_ions(None, p, p, p, p, p, None, 1, p, 1, C.byref(db), C.byref(sp), 1,
                         None) == 1
    assert "flag words" in hfx_amd.last_error()


def test_mod_groups_cut_spectra_by_memory():
    """The modified search's batches (Execution.hs:77-93 cuts a search's
    spectra by the memory its buffers need): consecutive ranges covering
    every spectrum once, each within the budget, a new range only where the
    next spectrum would overflow it, and an error for a spectrum that alone
    exceeds the budget."""
    rng = np.random.default_rng(11)
    per = pipeline.MPEP_BYTES
    for trial in range(200):
        n = int(rng.integers(0, 60))
        counts = rng.integers(0, 50, n) * (rng.random(n) < 0.8)
        budget = per * int(rng.integers(max(int(counts.max(initial=0)), 1), 400))
        g = pipeline.mod_groups(counts, budget)
        if n == 0:
            assert g == []
            continue
        assert [a for a, _ in g] == [0] + [b for _, b in g[:-1]]
        assert (g[-1][1] if g else 0) == n
        for k, (a, b) in enumerate(g):
            assert b > a
            assert int(counts[a:b].sum()) * per <= budget
            if k + 1 < len(g):
                assert (int(counts[a:b].sum()) + int(counts[b])) * per > budget
    assert pipeline.mod_groups([], per) == []
    with pytest.raises(hfx_amd.HfxError):
        pipeline.mod_groups([1, 5, 2], 4 * per)
