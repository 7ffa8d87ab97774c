"""GPU parity of the plain (unmodified) search path against the CPU oracle.

Integer and index work (packing, sorts, windows, bins, rankings) must be
bit-exact; FP32 scores must be within 1e-4 of the exact score relative to
max(|exact|, 100) (100 raw units = 0.01 displayed XCorr); exact re-scores
and preprocessing (same mean mode) are bit-exact.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STRIDE = 2112
REL = 1e-4
FLOOR = 100.0


@pytest.fixture(scope="module")
def env(small_workload, orc):
    import hfx_amd
    from hfx_amd import pipeline

    w = small_workload
    ctx = pipeline.Ctx(0)
    hdb = pipeline.HostDB(ions=w["prot"].seq, r=w["r"], tc=w["tc"], tn=w["tn"], mass=w["mass"])
    db = pipeline.DeviceDB(hdb)
    sp = w["sp"]
    dsp = pipeline.DeviceSpectra(ctx, sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE,
                                 mean_mode=1)
    pipeline.sync()
    return dict(ctx=ctx, db=db, dsp=dsp, w=w, L=hfx_amd.lib(), pl=pipeline)


def test_prepare_ions(env, orc):
    pl, L = env["pl"], env["L"]
    rng = np.random.default_rng(1)
    letters = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", np.uint8)
    for n in (1, 100, 511, 513, 100003):
        seq = letters[rng.integers(0, 20, n)]
        for ma in (b"", b"MSTY", b"Q"):
            t = pl.dev(seq)
            tma = pl.dev(np.frombuffer(ma, np.uint8)) if ma else None
            L.prepare_ions(pl.ptr(t), n, pl.ptr(tma) if tma is not None else None, len(ma))
            pl.sync()
            assert (pl.host(t, np.uint8, n) == orc.prepare_ions(seq, ma)).all(), (n, ma)


def test_db_packing_and_sort(env, orc):
    db, w = env["db"], env["w"]
    assert (db.packed_ions() == orc.prepare_ions(w["prot"].seq)).all()
    idx, rs = orc.sort_idx(w["r"])
    assert (db.idx_sorted() == idx).all()
    assert (env["pl"].host(db.t_rsorted, np.float32, db.num_pep) == rs).all()


def test_sort_b40c_f_stable(env):
    pl, L = env["pl"], env["L"]
    rng = np.random.default_rng(2)
    for n in (1, 7, 2048, 2049, 300001):
        keys = rng.integers(-50, 50, n).astype(np.float32) / np.float32(4)
        keys[::11] = -0.0
        vals = np.arange(n, dtype=np.uint32)
        tk, tv = pl.dev(keys), pl.dev(vals)
        L.sort_b40c_f(pl.ptr(tk), pl.ptr(tv), n)
        pl.sync()
        # radix order treats -0.0 < +0.0; ties keep input order
        key2 = keys.view(np.uint32).astype(np.int64)
        order_key = np.where(key2 & 0x80000000, -(key2 & 0x7FFFFFFF) - 1, key2)
        o = np.argsort(order_key, kind="stable")
        assert (pl.host(tv, np.uint32, n) == vals[o]).all()
        assert (pl.host(tk, np.float32, n).view(np.uint32) == keys[o].view(np.uint32)).all()


def test_windows(env, orc):
    pl, L, db, dsp, w = env["pl"], env["L"], env["db"], env["dsp"], env["w"]
    S = dsp.num_spec
    idx, _ = orc.sort_idx(w["r"])
    rng = np.random.default_rng(3)
    masses = np.concatenate([dsp.mass, rng.uniform(300, 8000, 200).astype(np.float32)])
    for eps in (3.0, 0.01, 50.0):
        b, e, n, tot = orc.find_spec_cands_by_mass(w["r"], idx, masses, eps)
        tb, te, tnn = pl.empty(4 * len(masses)), pl.empty(4 * len(masses)), pl.empty(4 * len(masses))
        tm = pl.dev(masses)
        got = L.findSpecCandsByMass(pl.ptr(tb), pl.ptr(te), pl.ptr(tnn), pl.ptr(db.t_r),
                                    pl.ptr(db.t_idx), db.num_pep, pl.ptr(tm), len(masses), eps)
        assert got == tot
        assert (pl.host(tb, np.uint32, len(masses)) == b).all()
        assert (pl.host(te, np.uint32, len(masses)) == e).all()
        assert (pl.host(tnn, np.uint32, len(masses)) == n).all()
        # batched form over the sorted copy
        env["L"].hfx_window(env["ctx"].h, pl.ptr(tb), pl.ptr(tnn), C.byref(db.desc), pl.ptr(tm),
                            len(masses), eps, pl.stream_handle())
        pl.sync()
        assert (pl.host(tb, np.uint32, len(masses)) == b).all()
        assert (pl.host(tnn, np.uint32, len(masses)) == n).all()
        # prepare_scoring, prepare_scoring.cu:47-67
        out = pl.empty(4 * max(tot, 1))
        tb2, tn2 = pl.dev(b), pl.dev(n)
        L.prepare_scoring(pl.ptr(out), pl.ptr(db.t_idx), pl.ptr(tb2), pl.ptr(tn2), len(masses))
        pl.sync()
        assert (pl.host(out, np.uint32, tot) == orc.prepare_scoring(idx, b, n)).all()


def test_find_begin_end(env, orc):
    pl, L, db, w = env["pl"], env["L"], env["db"], env["w"]
    idx, _ = orc.sort_idx(w["r"])
    delta = np.array([15.9949, 79.9663, 95.9612, -17.0, 0.0, 159.9326], np.float32)
    td = pl.dev(delta)
    for mass in (np.float32(1234.5), np.float32(2500.25), np.float32(100.0)):
        b, e, sc, tot = orc.find_begin_end(w["r"], idx, delta, mass, 3.0)
        tb, te, ts = pl.empty(4 * 6), pl.empty(4 * 6), pl.empty(4 * 6)
        got = L.findBeginEnd_f(pl.ptr(tb), pl.ptr(te), pl.ptr(ts), pl.ptr(db.t_r),
                               pl.ptr(db.t_idx), db.num_pep, pl.ptr(td), 6, mass, 3.0)
        assert got == tot
        assert (pl.host(tb, np.uint32, 6) == b).all() and (pl.host(te, np.uint32, 6) == e).all()
        assert (pl.host(ts, np.uint32, 6) == sc).all()


@pytest.mark.parametrize("mode", [0, 1])
def test_preprocess_bit_exact(env, orc, mode):
    pl, w = env["pl"], env["w"]
    sp = w["sp"]
    dsp = pl.DeviceSpectra(env["ctx"], sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE,
                           mean_mode=mode)
    X, lens, st = orc.preprocess_batch(sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE, mode)
    assert (st == 0).all()
    assert (dsp.status() == 0).all()
    assert (dsp.lengths() == lens).all()
    assert (dsp.x().view(np.uint32) == X.view(np.uint32)).all()


def test_preprocess_edge_cases(env, orc):
    pl = env["pl"]
    # no peaks; sel == 0; precursor removal off; a peak exactly at the cutoff
    pcr = np.array([500.0, 500.0, 700.0, 650.0], np.float32)
    ch = np.array([2, 2, 3, 2], np.float32)
    peaks = [[], [(5.0, 1.0), (9.0, 2.0)], [(100.0, 1.0), (690.0, 3.0), (710.0, 2.0), (1500.0, 1.0)],
             [(200.0, 1.0), (1350.0, 2.0), (1360.0, 0.5)]]
    mz = np.array([p[0] for pk in peaks for p in pk], np.float32)
    it = np.array([p[1] for pk in peaks for p in pk], np.float32)
    off = np.cumsum([0] + [len(p) for p in peaks]).astype(np.uint32)
    for remove in (True, False):
        dsp = pl.DeviceSpectra(env["ctx"], pcr, ch, mz, it, off, STRIDE, remove_precursor=remove)
        stat = dsp.status()
        assert stat[0] == 1 and stat[1] == 2
        for s in (2, 3):
            rc, x = orc.preprocess(pcr[s], ch[s], mz[off[s]:off[s + 1]], it[off[s]:off[s + 1]],
                                   remove, True, 1, STRIDE)
            assert rc == 0 and stat[s] == 0
            got = dsp.x()[s]
            assert dsp.lengths()[s] == len(x)
            assert (got[: len(x)].view(np.uint32) == x.view(np.uint32)).all()


def test_add_ions_dense_rows(env, orc):
    pl, L, db, w = env["pl"], env["L"], env["db"], env["w"]
    rng = np.random.default_rng(4)
    ions = db.packed_ions()
    sel = rng.integers(0, db.num_pep, 300).astype(np.uint32)
    tsel = pl.dev(sel)
    for zmax in (1, 2, 3, 4):
        for length in (2075, 1000, 333):
            out = pl.empty(4 * len(sel) * length)
            L.addIons(pl.ptr(out), pl.ptr(db.t_r), pl.ptr(db.t_mass), pl.ptr(db.t_ions),
                      pl.ptr(db.t_tc), pl.ptr(db.t_tn), pl.ptr(tsel), len(sel), zmax, length)
            pl.sync()
            got = pl.host(out, np.float32).reshape(len(sel), length)
            for i, p in enumerate(sel):
                ref = orc.peaks_dense(ions, int(w["tc"][p]), int(w["tn"][p]), w["r"][p], w["mass"],
                                      zmax, length)
                assert (got[i] == ref.astype(np.float32)).all(), (zmax, length, i)


def test_mvm_rows(env):
    pl, L = env["pl"], env["L"]
    rng = np.random.default_rng(5)
    m, n = 777, 2075
    A = rng.choice(np.array([0, 10, 25, 50], np.float32), (m, n))
    x = rng.normal(0, 3, n).astype(np.float32)
    ta, tx, ty = pl.dev(A), pl.dev(x), pl.empty(4 * m)
    hfx = __import__("hfx_amd")
    hfx.check(L.hfx_mvm_rows(env["ctx"].h, pl.ptr(ty), pl.ptr(ta), n, pl.ptr(tx), m, n,
                             pl.stream_handle()), "mvm")
    pl.sync()
    ref = A.astype(np.float64) @ x.astype(np.float64)
    got = pl.host(ty, np.float32, m)
    assert np.all(np.abs(got - ref) <= REL * np.maximum(np.abs(ref), FLOOR))


def _oracle_plain(orc, env, topn, want_all=True, spectra=None):
    w, db, dsp = env["w"], env["db"], env["dsp"]
    idx, _ = orc.sort_idx(w["r"])
    X = dsp.x()
    lens = dsp.lengths()
    return orc.search_plain(db.packed_ions(), w["r"], w["tc"], w["tn"], w["mass"], idx, X, lens,
                            dsp.zmax, dsp.mass, 3.0, topn, want_all)


def test_sparse_scores_and_topn(env, orc):
    pl = env["pl"]
    ref = _oracle_plain(orc, env, 5)
    srch = pl.PlainSearch(env["ctx"], env["db"], env["dsp"], eps=3.0, topn=5, margin=16)
    res = srch.run(keep_scores=True)
    assert (res.begin == ref["begin"]).all() and (res.num == ref["num"]).all()
    assert int(res.off[-1]) == len(res.scores) == int(ref["off"][-1])
    exact = orc.fx_to_double(ref["all_fx"][: len(res.scores)])
    err = np.abs(res.scores.astype(np.float64) - exact)
    assert np.all(err <= REL * np.maximum(np.abs(exact), FLOOR)), err.max()
    # rankings and exact scores identical to the oracle
    assert (res.top_cnt == ref["top_cnt"]).all()
    for s in range(len(res.top_cnt)):
        c = int(res.top_cnt[s])
        assert res.top_pos[s, :c].tolist() == ref["top_pos"][s, :c].tolist(), s
        assert (res.top_exact[s, :c] == orc.fx_to_double(ref["top_fx"][s, :c])).all()
    # the spectrum's source peptide is found at rank 1 for most spectra
    w = env["w"]
    idx = env["db"].idx_sorted()
    src = w["sp"].source
    hits = sum(int(idx[res.begin[s] + res.top_pos[s, 0]]) == int(src[s]) or
               (w["r"][idx[res.begin[s] + res.top_pos[s, 0]]] == w["r"][src[s]])
               for s in range(len(src)) if res.top_cnt[s])
    assert hits >= 0.9 * len(src)


def test_exact_fallback_matches(env, orc):
    pl, L = env["pl"], env["L"]
    import hfx_amd

    ref = _oracle_plain(orc, env, 5, want_all=False)
    srch = pl.PlainSearch(env["ctx"], env["db"], env["dsp"], eps=3.0, topn=5)
    srch.window()
    srch.score()
    S = env["dsp"].num_spec
    # every spectrum through the exact route (exact score of every candidate)
    srch.exact(np.arange(S, dtype=np.uint32))
    pl.sync()
    pos = pl.host(srch.t_pos, np.uint32, S * 5).reshape(S, 5)
    ex = pl.host(srch.t_exact, np.float64, S * 5).reshape(S, 5)
    cnt = pl.host(srch.t_cnt, np.uint32, S)
    assert (cnt == ref["top_cnt"]).all()
    for s in range(S):
        c = int(cnt[s])
        assert pos[s, :c].tolist() == ref["top_pos"][s, :c].tolist()
        assert (ex[s, :c] == orc.fx_to_double(ref["top_fx"][s, :c])).all()


def test_empty_windows_and_high_charge(env, orc):
    pl, w = env["pl"], env["w"]
    sp = w["sp"]
    # precursors far outside the database mass range give empty windows;
    # charge 7 (max fragment charge 6 > 4) scores 0 like the reference's no-launch
    # the 7+ precursor m/z of spectrum 0's peptide keeps the same window
    pcr7 = (2.0 * float(sp.pcr[0]) + 5.0 * 1.007276) / 7.0
    pcr = np.array([90.0, pcr7, sp.pcr[1]], np.float32)
    ch = np.array([2.0, 7.0, 2.0], np.float32)
    s0, s1 = int(sp.off[0]), int(sp.off[2])
    mz = np.concatenate([sp.mz[s0:s1], sp.mz[s0:int(sp.off[1])]])
    it = np.concatenate([sp.inten[s0:s1], sp.inten[s0:int(sp.off[1])]])
    n0, n1 = int(sp.off[1] - sp.off[0]), int(sp.off[2] - sp.off[1])
    off = np.array([0, n0, n0 + n1, n0 + n1 + n0], np.uint32)
    ch = ch * 1
    dsp = pl.DeviceSpectra(env["ctx"], pcr, ch, mz, it, off, STRIDE)
    srch = pl.PlainSearch(env["ctx"], env["db"], dsp, eps=3.0, topn=5)
    res = srch.run(keep_scores=True)
    assert res.num[0] == 0 and res.top_cnt[0] == 0
    a, b = int(res.off[1]), int(res.off[2])
    assert b > a and np.all(res.scores[a:b] == 0)


@pytest.mark.parametrize("copies,topn", [(400, 10), (400, 5), (400, 1), (120, 5)])
def test_tie_heavy_window_takes_exact_route(orc, copies, topn):
    """`copies` proteins carry the same tryptic peptide: a spectrum made from
    it has that many candidates with one identical theoretical spectrum, so
    its N-th place is an exact tie.  With 400 copies the tie holds far more
    candidates than the re-score set (256 past the selection): the spectrum
    goes down the on-device exact route (flagged); with 120 the tied band is
    re-scored inside the top-N kernel.  N of 1 and 5 select through the
    per-thread lists (the band taken from the lists), N of 10 through the
    radix select.  The top N (the tied copies in ascending position order,
    then the rest) equals the oracle's.  Both scoring forms feed the same
    top-N."""
    from hfx_amd import pipeline, synth

    base = synth.make_proteins(600, seed=61)
    motif = np.frombuffer(b"GASGTPAGEGSK", np.uint8)
    rng = np.random.default_rng(62)
    parts, seg = [base.seq], list(base.seg)
    for _ in range(copies):
        head = np.frombuffer(b"MKR", np.uint8)
        tail = synth.make_proteins(1, seed=int(rng.integers(1 << 30)), min_len=20, max_len=60).seq
        p = np.concatenate([head, motif, tail])
        parts.append(p)
        seg.append(seg[-1] + len(p))
    prot = synth.Proteins(seq=np.concatenate(parts).astype(np.uint8),
                          seg=np.asarray(seg, np.uint32))
    mass = orc.mass_table()
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=1)
    sp = synth.make_spectra(prot, r, tc, tn, 24, seed=63, charges=(2,))
    # spectrum 0 from the shared peptide
    hit = [i for i in range(len(r))
           if synth.peptide_residues(prot, int(tc[i]), int(tn[i])) == motif.tobytes()]
    assert len(hit) == copies
    one = synth.make_spectra(prot, r[hit[0]:hit[0] + 1], tc[hit[0]:hit[0] + 1],
                             tn[hit[0]:hit[0] + 1], 1, seed=64, charges=(2,))
    n0 = int(one.off[1])
    mz = np.concatenate([one.mz[:n0], sp.mz[sp.off[1]:]])
    inten = np.concatenate([one.inten[:n0], sp.inten[sp.off[1]:]])
    off = np.concatenate([[0], n0 + (sp.off[1:] - sp.off[1])]).astype(np.uint32)
    pcr = np.concatenate([one.pcr[:1], sp.pcr[1:]]).astype(np.float32)
    ch = np.concatenate([one.charge[:1], sp.charge[1:]]).astype(np.float32)

    ctx = pipeline.Ctx(0)
    db = pipeline.DeviceDB(pipeline.HostDB(prot.seq, r, tc, tn, mass))
    dsp = pipeline.DeviceSpectra(ctx, pcr, ch, mz, inten, off, 2112)
    idx, _ = orc.sort_idx(r)
    ref = orc.search_plain(orc.prepare_ions(prot.seq), r, tc, tn, mass, idx, dsp.x(),
                           dsp.lengths(), dsp.zmax, dsp.mass, 3.0, topn)
    for dense in (False, True):
        res = pipeline.PlainSearch(ctx, db, dsp, eps=3.0, topn=topn).run(dense=dense)
        assert (0 in res.flagged.tolist()) == (copies - topn > 256), dense
        assert (res.top_cnt == ref["top_cnt"]).all()
        for s in range(dsp.num_spec):
            c = int(res.top_cnt[s])
            assert res.top_pos[s, :c].tolist() == ref["top_pos"][s, :c].tolist(), (dense, s)
            assert (res.top_exact[s, :c] == orc.fx_to_double(ref["top_fx"][s, :c])).all()


@pytest.mark.parametrize("topn", [10, 5])
def test_exact_route_past_the_lds_selection(orc, topn):
    """5000 proteins carry the same tryptic peptide: spectrum 0's window holds
    5000 candidates tied exactly, more than the exact route's LDS selection
    takes (4096 live keys), so its top N come from the radix select over
    global memory.  With N = 5 the per-thread lists fill with tied values
    (each thread sees about twenty), so the band is swept instead of taken
    from the lists.  Top N equal the oracle's for both scoring forms."""
    from hfx_amd import pipeline, synth

    copies = 5000
    base = synth.make_proteins(300, seed=71)
    motif = np.frombuffer(b"GASGTPAGEGSK", np.uint8)
    tails = synth.make_proteins(copies, seed=72, min_len=20, max_len=60)
    parts, seg = [base.seq], list(base.seg)
    for k in range(copies):
        # "MKR" + motif + "G" + tail: the motif is cleaved on both sides
        tail = tails.seq[int(tails.seg[k]):int(tails.seg[k + 1])]
        p = np.concatenate([np.frombuffer(b"MKR", np.uint8), motif,
                            np.frombuffer(b"G", np.uint8), tail])
        parts.append(p)
        seg.append(seg[-1] + len(p))
    prot = synth.Proteins(seq=np.concatenate(parts).astype(np.uint8),
                          seg=np.asarray(seg, np.uint32))
    mass = orc.mass_table()
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=1)
    sp = synth.make_spectra(prot, r, tc, tn, 16, seed=73, charges=(2,))
    hit = [i for i in range(len(r))
           if synth.peptide_residues(prot, int(tc[i]), int(tn[i])) == motif.tobytes()]
    assert len(hit) == copies
    one = synth.make_spectra(prot, r[hit[0]:hit[0] + 1], tc[hit[0]:hit[0] + 1],
                             tn[hit[0]:hit[0] + 1], 1, seed=74, charges=(2,))
    n0 = int(one.off[1])
    mz = np.concatenate([one.mz[:n0], sp.mz[sp.off[1]:]])
    inten = np.concatenate([one.inten[:n0], sp.inten[sp.off[1]:]])
    off = np.concatenate([[0], n0 + (sp.off[1:] - sp.off[1])]).astype(np.uint32)
    pcr = np.concatenate([one.pcr[:1], sp.pcr[1:]]).astype(np.float32)
    ch = np.concatenate([one.charge[:1], sp.charge[1:]]).astype(np.float32)

    ctx = pipeline.Ctx(0)
    db = pipeline.DeviceDB(pipeline.HostDB(prot.seq, r, tc, tn, mass))
    dsp = pipeline.DeviceSpectra(ctx, pcr, ch, mz, inten, off, 2112)
    idx, _ = orc.sort_idx(r)
    ref = orc.search_plain(orc.prepare_ions(prot.seq), r, tc, tn, mass, idx, dsp.x(),
                           dsp.lengths(), dsp.zmax, dsp.mass, 3.0, topn)
    assert int(ref["num"][0]) > copies
    for dense in (False, True):
        res = pipeline.PlainSearch(ctx, db, dsp, eps=3.0, topn=topn).run(dense=dense)
        assert 0 in res.flagged.tolist(), dense
        assert (res.top_cnt == ref["top_cnt"]).all()
        for s in range(dsp.num_spec):
            c = int(res.top_cnt[s])
            assert res.top_pos[s, :c].tolist() == ref["top_pos"][s, :c].tolist(), (dense, s)
            assert (res.top_exact[s, :c] == orc.fx_to_double(ref["top_fx"][s, :c])).all()


@pytest.mark.parametrize("n", [0, 1, 1000, 8191, 8192, 8193, 20_000, 32_767, 32_768, 32_769,
                               8192 * 7, 8192 * 7 + 1, 5_450_001])
def test_scan_offsets(env, n):
    """hfx_scan_offsets (one workgroup below 8192 counts, three passes of
    8192-count tiles above) against numpy's cumulative sum, with counts up to
    2^32 - 1 so that the 64-bit offsets carry."""
    pl, L = env["pl"], env["L"]
    rng = np.random.default_rng(n)
    cnt = rng.integers(0, 2**32 - 1, size=n, dtype=np.uint64).astype(np.uint32)
    if n > 2:
        cnt[: n // 3] = rng.integers(0, 50, size=n // 3).astype(np.uint32)
    t_cnt = pl.dev(cnt) if n else pl.empty(4)
    t_off = pl.empty(8 * (n + 1))
    from hfx_amd import check

    check(L.hfx_scan_offsets(env["ctx"].h, t_off.p, t_cnt.p, n, None), "hfx_scan_offsets")
    pl.sync()
    got = pl.host(t_off, np.uint64, n + 1)
    want = np.zeros(n + 1, np.uint64)
    np.cumsum(cnt.astype(np.uint64), out=want[1:])
    assert np.array_equal(got, want)


def test_long_candidates_take_the_general_path(orc):
    """Proteins without K or R digest into peptides of up to 110 residues
    (mass-limited at 7200 Da): candidates of more than 64 ladder positions do
    not fit a wave's lanes and take the general path of the plain kernel,
    mixed with packed short ones; scores within the FP32 bound of the
    oracle's exact scores and identical top-N."""
    from hfx_amd import pipeline, synth

    rng = np.random.default_rng(71)
    light = np.frombuffer(b"GGGAAS", np.uint8)
    plain = np.frombuffer(b"ARNDCQEGHILKMFPSTWYV", np.uint8)
    seqs = []
    for i in range(300):
        if i % 2 == 0:  # long K/R-free stretches
            seqs.append(light[rng.integers(0, len(light), rng.integers(70, 140))])
        else:
            seqs.append(plain[rng.integers(0, len(plain), rng.integers(100, 400))])
    seg = np.zeros(len(seqs) + 1, np.uint32)
    np.cumsum([len(q) for q in seqs], out=seg[1:])
    prot = synth.Proteins(seq=np.ascontiguousarray(np.concatenate(seqs)), seg=seg)
    mass = orc.mass_table({})
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=1)
    nres = (tn.astype(np.int64) - tc.astype(np.int64))
    assert (nres > 64).sum() > 60
    sp = synth.make_spectra(prot, r, tc, tn, 60, seed=72, charges=(2, 3), charge_p=(0.5, 0.5),
                            double_for_3plus=True)
    # half the spectra moved to the precursors of long peptides
    long_idx = np.nonzero(nres > 64)[0]
    ctx = pipeline.Ctx(0)
    db = pipeline.DeviceDB(pipeline.HostDB(prot.seq, r, tc, tn, mass))
    pcr = sp.pcr.copy()
    H = 1.007276
    for s in range(0, 60, 2):
        q = int(long_idx[s % len(long_idx)])
        z = float(sp.charge[s])
        pcr[s] = np.float32((float(r[q]) + 18.010565 + z * H) / z)
    dsp = pipeline.DeviceSpectra(ctx, pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE)
    srch = pipeline.PlainSearch(ctx, db, dsp, eps=3.0, topn=5)
    res = srch.run(keep_scores=True)
    idx, _ = orc.sort_idx(r)
    ref = orc.search_plain(orc.prepare_ions(prot.seq), r, tc, tn, mass, idx, dsp.x(),
                           dsp.lengths(), dsp.zmax, dsp.mass, 3.0, 5, want_all=True)
    # long candidates are in the windows
    isort = idx.astype(np.int64)
    win_long = sum(int((nres[isort[int(res.begin[s]):int(res.begin[s] + res.num[s])]] > 64).sum())
                   for s in range(60))
    assert win_long > 20
    exact = orc.fx_to_double(ref["all_fx"])
    X, lens = dsp.x(), dsp.lengths()
    x1 = np.array([np.abs(X[s, : lens[s]].astype(np.float64)).sum() for s in range(60)])
    bound = np.repeat(80 * 2.0**-24 * 50 * x1 * 1.001 + 1e-6, np.diff(res.off.astype(np.int64)))
    assert np.all(np.abs(res.scores.astype(np.float64) - exact) <= bound)
    assert (res.top_cnt == ref["top_cnt"]).all()
    for s in range(60):
        c = int(res.top_cnt[s])
        assert res.top_pos[s, :c].tolist() == ref["top_pos"][s, :c].tolist(), s
        assert (res.top_exact[s, :c] == orc.fx_to_double(ref["top_fx"][s, :c])).all(), s


@pytest.mark.parametrize("n", [1, 7, 2048, 100_003])
def test_descending_sorts_and_range_indices(env, n):
    """sort_rf / sort_idx_rf (sort.cu:37-76, thrust::greater): stable
    descending sorts, checked against numpy's stable sort of the negated
    keys (the keys hold many ties and both signs); findIndicesInRange_f
    (filter.cu:38-73): the ascending indices of the keys inside [min, max],
    bounds inclusive."""
    pl, L = env["pl"], env["L"]
    rng = np.random.default_rng(n)
    keys = (rng.integers(-40, 40, size=n) * 0.25).astype(np.float32)
    vals = rng.integers(0, 2**32 - 1, size=n, dtype=np.uint64).astype(np.uint32)
    want = np.argsort(-keys.astype(np.float64), kind="stable")
    tk, tv = pl.dev(keys), pl.dev(vals)
    L.sort_rf(tk.p, tv.p, n)
    pl.sync()
    assert np.array_equal(pl.host(tk, np.float32, n), keys[want])
    assert np.array_equal(pl.host(tv, np.uint32, n), vals[want])
    tk = pl.dev(keys)
    ti = pl.empty(4 * n)
    L.sort_idx_rf(tk.p, ti.p, n, 5)
    pl.sync()
    assert np.array_equal(pl.host(ti, np.uint32, n), want.astype(np.uint32) + 5)
    assert np.array_equal(pl.host(tk, np.float32, n), keys[want])
    for lo, hi in ((-1.0, 2.5), (3.0, 3.0), (20.0, 30.0), (-100.0, 100.0)):
        tk = pl.dev(keys)
        to = pl.empty(4 * n)
        cnt = L.findIndicesInRange_f(tk.p, to.p, n, lo, hi)
        exp = np.nonzero((keys >= np.float32(lo)) & (keys <= np.float32(hi)))[0]
        assert cnt == len(exp), (lo, hi)
        assert np.array_equal(pl.host(to, np.uint32, cnt), exp.astype(np.uint32)), (lo, hi)


def test_mvm_ff_on_a_handle_stream(env):
    """mvm_ff (mvm.cu:21-56, cublasSgemv over a row-major matrix): y = A x on
    the stream its handle carries, within the FP32 bound of a float64 dot
    (at most 64 + 6 roundings per row sum: n / 64 terms a lane, then a
    six-level butterfly)."""
    pl, L = env["pl"], env["L"]
    from hfx_amd import check

    rng = np.random.default_rng(3)
    for m, n in ((1, 1), (37, 2075), (1000, 512)):
        A = rng.standard_normal((m, n)).astype(np.float32)
        x = rng.standard_normal(n).astype(np.float32)
        tA, tx, ty = pl.dev(A), pl.dev(x), pl.empty(4 * m)
        h = C.c_void_p()
        check(L.hfx_blas_create(C.byref(h)), "hfx_blas_create")
        st = pl.Stream()
        check(L.hfx_blas_set_stream(h, st.p), "hfx_blas_set_stream")
        L.mvm_ff(h, ty.p, tA.p, tx.p, m, n)
        st.sync()
        y = pl.host(ty, np.float32, m).astype(np.float64)
        exact = A.astype(np.float64) @ x.astype(np.float64)
        bound = (n // 64 + 8) * 2.0**-24 * (np.abs(A).astype(np.float64) @ np.abs(x)) + 1e-30
        assert np.all(np.abs(y - exact) <= bound), (m, n)
        check(L.hfx_blas_destroy(h), "hfx_blas_destroy")


def test_scores_that_break_the_bound_are_reported_short(env, orc):
    """hfx_topn_rel given FP32 scores that are not within its stated bound of
    the exact ones (every score 0, bound factors 0; eps 10 for windows of
    hundreds of candidates, top 1): the re-score band then holds the whole
    window, big windows take the exact route, and where the selected
    candidate's exact score is positive no candidate's FP32 score reaches
    tau.  Those spectra are flagged short with top_cnt 0; every other
    spectrum gets the oracle's top 1.  A following select on real scores
    clears the mark."""
    pl, w, dsp = env["pl"], env["w"], env["dsp"]
    S = dsp.num_spec
    srch = pl.PlainSearch(env["ctx"], env["db"], dsp, eps=10.0, topn=1)
    total = srch.window()
    num = pl.host(srch.t_num, np.uint32, S)
    srch.t_scores.upload(np.zeros(total, np.float32))
    srch.select_rel(0.0, 0.0)
    pl.sync()
    flags = pl.host(srch.t_flag, np.uint32, S)
    cnt = pl.host(srch.t_cnt, np.uint32, S)
    pos = pl.host(srch.t_pos, np.uint32, S)
    ex = pl.host(srch.t_exact, np.float64, S)
    short = (flags & pl.FLAG_SHORT) != 0
    print(f"{int((num > 257).sum())} windows above 257 candidates, {int(short.sum())} short")
    assert short.any() and (flags[short] & pl.FLAG_EXACT).all()
    assert (cnt[short] == 0).all() and (cnt[~short] == np.minimum(1, num[~short])).all()
    assert (srch.short() == np.nonzero(short)[0]).all()
    idx, _ = orc.sort_idx(w["r"])
    ref = orc.search_plain(env["db"].packed_ions(), w["r"], w["tc"], w["tn"], w["mass"], idx,
                           dsp.x(), dsp.lengths(), dsp.zmax, dsp.mass, 10.0, 1)
    for s in np.nonzero(~short)[0]:
        if cnt[s]:
            assert pos[s] == ref["top_pos"][s, 0] and ex[s] == orc.fx_to_double(ref["top_fx"][s, 0])
    srch.score()
    srch.select()
    pl.sync()
    assert len(srch.short()) == 0
