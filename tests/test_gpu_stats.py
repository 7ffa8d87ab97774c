"""GPU parity of the E-value fit (hfx_evalue_fit, hfx_evalues) and the SP ion
matches (hfx_sp_peaks, hfx_sp_ions*) against the CPU oracle.

Integer results -- histogram points, kept scores, peak flags, matched and
total ion counts -- must be identical.  alpha and beta are FP64 sums over the
same histogram in another order with the device's log: within 1e-9
relative (absolute 1e-9 near zero); E-values within 1e-9 relative.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STRIDE = 2112
TOL = 1e-9
VAR = {"M": 15.9949, "S": 79.9663, "T": 79.9663, "Y": 79.9663}


@pytest.fixture(scope="module")
def senv(small_workload, orc):
    import hfx_amd
    from hfx_amd import pipeline

    w = small_workload
    ctx = pipeline.Ctx(0)
    db = pipeline.DeviceDB(pipeline.HostDB(ions=w["prot"].seq, r=w["r"], tc=w["tc"], tn=w["tn"],
                                           mass=w["mass"]))
    sp = w["sp"]
    dsp = pipeline.DeviceSpectra(ctx, sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE,
                                 mean_mode=1)
    ps = pipeline.PlainSearch(ctx, db, dsp, eps=3.0, topn=5)
    res = ps.run(keep_scores=True)
    return dict(w=w, ctx=ctx, db=db, dsp=dsp, ps=ps, res=res, L=hfx_amd.lib(), pl=pipeline)


@pytest.fixture(scope="module")
def menv(orc):
    import hfx_amd
    from hfx_amd import pipeline, synth

    prot = synth.make_proteins(400, seed=41)
    mass = orc.mass_table()
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=1, p_rule=1, min_len=7,
                           max_len=30)
    sp = synth.make_spectra(prot, r, tc, tn, 24, seed=42, charges=(1, 2, 3, 4))
    ctx = pipeline.Ctx(0)
    mi = pipeline.mod_info(VAR, 4)
    db = pipeline.DeviceDB(pipeline.HostDB(prot.seq, r, tc, tn, mass), ma=mi.ma)
    dsp = pipeline.DeviceSpectra(ctx, sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE)
    ms = pipeline.ModSearch(ctx, db, dsp, mi, eps=3.0, topn=5)
    mres = ms.run(keep_scores=True)
    ps = pipeline.PlainSearch(ctx, db, dsp, eps=3.0, topn=5)
    pres = ps.run(keep_scores=True)
    return dict(prot=prot, mass=mass, r=r, tc=tc, tn=tn, ctx=ctx, mi=mi, db=db, dsp=dsp, ms=ms,
                mres=mres, ps=ps, pres=pres, L=hfx_amd.lib(), pl=pipeline)


def _close(a, b):
    if np.isnan(b):
        return np.isnan(a)
    return abs(a - b) <= TOL * max(abs(b), 1.0)


def _fit_direct(pl, L, segs_a, segs_b=None, bin_width=350, perc=3.0):
    """hfx_evalue_fit through the C ABI over per-spectrum score lists."""
    S = len(segs_a)

    def pack(segs):
        off = np.zeros(S + 1, np.uint64)
        off[1:] = np.cumsum([len(x) for x in segs])
        flat = np.concatenate([np.asarray(x, np.float32) for x in segs] + [np.zeros(1, np.float32)])
        return pl.dev(flat), pl.dev(off)

    ta, oa = pack(segs_a)
    tb = ob = None
    if segs_b is not None:
        tb, ob = pack(segs_b)
    fit, npt, nu = pl.empty(16 * S), pl.empty(4 * S), pl.empty(8 * S)
    pl.check(L.hfx_evalue_fit(None, pl.ptr(fit), pl.ptr(npt), pl.ptr(nu), pl.ptr(ta), pl.ptr(oa),
                              pl.ptr(tb) if tb else None, pl.ptr(ob) if ob else None, S,
                              bin_width, perc, None), "hfx_evalue_fit")
    pl.sync()
    return (pl.host(fit, np.float64, 2 * S).reshape(S, 2), pl.host(npt, np.uint32, S),
            pl.host(nu, np.uint64, S))


def _check_fits(orc, got, segs, bin_width=350, perc=3.0):
    fit, npt, nu = got
    for s, sc in enumerate(segs):
        a, b, m, nused = orc.evalue_fit(np.asarray(sc, np.float32), bin_width, perc)
        assert (int(npt[s]), int(nu[s])) == (m, nused), s
        assert _close(fit[s, 0], a) and _close(fit[s, 1], b), (s, fit[s], a, b)


def test_evalue_fit_plain_search(senv, orc):
    """Every spectrum of the small workload, over its FP32 window scores."""
    res, ps = senv["res"], senv["ps"]
    fit, npt, nu = ps.evalue_fit()
    segs = [res.scores[int(res.off[s]):int(res.off[s + 1])] for s in range(len(res.num))]
    _check_fits(orc, (fit, npt, nu), segs)
    assert (npt >= 2).sum() > len(segs) // 2


def _synthetic_segments(rng):
    g = lambda n, k=2500.0: (rng.gamma(2.0, k, n) - 3000.0).astype(np.float32)
    wide = g(5000)
    wide[:1500] = rng.uniform(4e5, 9e5, 1500)       # kept keys past the first 1024 ...
    wide[1500:2000] = rng.uniform(2e6, 2.1e6, 500)  # ... and a gap between ranges
    ties = np.repeat(np.float32([-500.0, 4800.0, 5200.0, 9000.0]), [300, 400, 200, 100])
    return [np.zeros(0, np.float32), np.float32([7.0]), g(5), g(3000), g(20000), wide,
            np.full(777, 123.5, np.float32), ties, g(8192), g(8193),
            np.float32([0.0] * 10 + [8999.0] * 20 + [9699.0] * 2 + [20000.0])]


def test_evalue_fit_edge_segments(senv, orc):
    """Empty and one-score spectra, more scores than the LDS copy holds,
    key spreads wider than one histogram pass (with a gap), ties, a single
    histogram point, the two-point closed form; also with a cut-off of 7%
    and a bin width of 400."""
    pl, L = senv["pl"], senv["L"]
    segs = _synthetic_segments(np.random.default_rng(5))
    _check_fits(orc, _fit_direct(pl, L, segs), segs)
    _check_fits(orc, _fit_direct(pl, L, segs, perc=7.0, bin_width=400), segs, 400, 7.0)


def test_evalue_fit_pools_two_score_sets(senv, orc):
    """Search.hs:118: the unmodified and modified scores of a spectrum form
    one distribution; splitting it between the two sets changes nothing."""
    pl, L = senv["pl"], senv["L"]
    rng = np.random.default_rng(6)
    segs = _synthetic_segments(rng)
    cut = [int(rng.integers(0, len(x) + 1)) for x in segs]
    got = _fit_direct(pl, L, [x[:c] for x, c in zip(segs, cut)],
                      [x[c:] for x, c in zip(segs, cut)])
    _check_fits(orc, got, segs)


def test_evalue_fit_non_finite_gives_nan(senv):
    pl, L = senv["pl"], senv["L"]
    fit, npt, nu = _fit_direct(pl, L, [np.float32([1.0, np.inf, 9000.0] * 10)])
    assert np.isnan(fit).all() and npt[0] == 0 and nu[0] == 0


def test_evalues_apply(senv, orc):
    res, ps = senv["res"], senv["ps"]
    fit, _, _ = ps.evalue_fit()
    ev = ps.evalues(res.top_exact, res.top_cnt)
    for s in range(len(res.top_cnt)):
        for k in range(ev.shape[1]):
            if k >= res.top_cnt[s]:
                assert ev[s, k] == 0.0
                continue
            want = orc.evalue(fit[s, 0], fit[s, 1], float(res.top_exact[s, k]))
            if np.isnan(want):
                assert np.isnan(ev[s, k])
            else:
                assert ev[s, k] == pytest.approx(want, rel=TOL)


def _unpack(words, npk):
    bits = np.unpackbits(np.asarray(words, np.uint32).view(np.uint8), bitorder="little")
    return bits[:npk]


def _peaks_device(pl, L, desc, S, stride):
    W = L.hfx_sp_peak_words(stride)
    t = pl.empty(4 * W * max(S, 1))
    pl.check(L.hfx_sp_peaks(None, pl.ptr(t), W, C.byref(desc), None), "hfx_sp_peaks")
    pl.sync()
    return pl.host(t, np.uint32, W * S).reshape(S, W)


def test_sp_peaks_workload(senv, orc):
    pl, L, dsp = senv["pl"], senv["L"], senv["dsp"]
    got = _peaks_device(pl, L, dsp.desc, dsp.num_spec, dsp.x_stride)
    X, lens = dsp.x(), dsp.lengths()
    for s in range(dsp.num_spec):
        n = int(lens[s])
        np.testing.assert_array_equal(_unpack(got[s], n // 3 + 1),
                                      orc.extract_peaks(X[s, :n]), err_msg=str(s))
        # nothing beyond the spectrum's own bins
        assert _unpack(got[s], got.shape[1] * 32)[n // 3 + 1:].sum() == 0


@pytest.mark.parametrize("stride", [3072, 12000])
def test_sp_peaks_synthetic_rows(senv, orc, stride):
    """Ties at the 200th value, fewer than 200 positive values, empty and
    one-value rows; stride 3072: the wave-per-spectrum kernel up to its
    longest row; stride 12000: the workgroup kernel, also with rows longer
    than its LDS copy."""
    import hfx_amd

    pl, L = senv["pl"], senv["L"]
    rng = np.random.default_rng(7)
    rows = []
    for n in (0, 1, 2, 3, 250, 1000, 2999, 3072, 8192, 8193, 11999):
        if n > stride:
            continue
        x = rng.normal(0, 1, n).astype(np.float32)
        rows.append(x)
        rows.append((np.round(x * 3) / 3).astype(np.float32))   # many ties
    rows.append(np.where(rng.random(min(4000, stride)) < 0.02, 1.0, -1.0)
                .astype(np.float32))  # about 60-80 positive
    S = len(rows)
    X = np.zeros((S, stride), np.float32)
    lens = np.zeros(S, np.uint32)
    for s, x in enumerate(rows):
        X[s, :len(x)] = x
        lens[s] = len(x)
    tx, tl = pl.dev(X), pl.dev(lens)
    tz, tm = pl.dev(np.ones(S, np.uint32)), pl.dev(np.zeros(S, np.float32))
    desc = hfx_amd.HfxSpectra(pl.ptr(tx), stride, pl.ptr(tl), pl.ptr(tz), pl.ptr(tm), S)
    got = _peaks_device(pl, L, desc, S, stride)
    for s, x in enumerate(rows):
        np.testing.assert_array_equal(_unpack(got[s], len(x) // 3 + 1), orc.extract_peaks(x),
                                      err_msg=str(s))


def _oracle_flags(orc, dsp):
    X, lens = dsp.x(), dsp.lengths()
    return [orc.extract_peaks(X[s, :int(lens[s])]) for s in range(dsp.num_spec)]


def test_sp_ions_plain_topn(senv, orc):
    w, res, ps, dsp, db = senv["w"], senv["res"], senv["ps"], senv["dsp"], senv["db"]
    nb, ny, tot = ps.sp_ions()
    flags = _oracle_flags(orc, dsp)
    idx = db.idx_sorted()
    ions = orc.prepare_ions(w["prot"].seq)
    checked = 0
    for s in range(dsp.num_spec):
        for k in range(nb.shape[1]):
            if k >= res.top_cnt[s]:
                assert nb[s, k] == ny[s, k] == tot[s, k] == 0
                continue
            p = int(idx[int(res.begin[s]) + int(res.top_pos[s, k])])
            want = orc.match_ions(ions, int(w["tc"][p]), int(w["tn"][p]), w["r"][p], w["mass"],
                                  int(dsp.zmax[s]), flags[s])
            assert (int(nb[s, k]), int(ny[s, k]), int(tot[s, k])) == want, (s, k)
            checked += 1
    assert checked > 100


def test_sp_ions_list(senv, orc):
    """hfx_sp_ions over arbitrary (peptide, spectrum, modified sites) pairs,
    with peptides at the database's end and 'no peptide' entries."""
    pl, L, w, dsp, db = senv["pl"], senv["L"], senv["w"], senv["dsp"], senv["db"]
    rng = np.random.default_rng(8)
    P = len(w["r"])
    M = 3000
    pep = rng.integers(0, P, M).astype(np.uint32)
    last = np.nonzero(w["tn"] >= len(w["prot"].seq))[0]
    if len(last):
        pep[:len(last)] = last
    pep[-5:] = 0xFFFFFFFF
    spec = rng.integers(0, dsp.num_spec, M).astype(np.uint32)
    nmod = rng.integers(0, 4, M).astype(np.uint32)
    W = L.hfx_sp_peak_words(dsp.x_stride)
    t_on = pl.empty(4 * W * dsp.num_spec)
    pl.check(L.hfx_sp_peaks(None, pl.ptr(t_on), W, C.byref(dsp.desc), None), "hfx_sp_peaks")
    outs = [pl.empty(4 * M) for _ in range(3)]
    tp, ts, tn = pl.dev(pep), pl.dev(spec), pl.dev(nmod)
    pl.check(L.hfx_sp_ions(None, *[pl.ptr(o) for o in outs], pl.ptr(tp), pl.ptr(ts), pl.ptr(tn), M,
                           pl.ptr(t_on), W, C.byref(db.desc), C.byref(dsp.desc), 1, None),
             "hfx_sp_ions")
    pl.sync()
    nb, ny, tot = [pl.host(o, np.uint32, M) for o in outs]
    flags = _oracle_flags(orc, dsp)
    ions = orc.prepare_ions(w["prot"].seq)
    for i in range(M):
        if pep[i] == 0xFFFFFFFF:
            assert nb[i] == ny[i] == tot[i] == 0
            continue
        p, s = int(pep[i]), int(spec[i])
        want = orc.match_ions(ions, int(w["tc"][p]), int(w["tn"][p]), w["r"][p], w["mass"],
                              int(dsp.zmax[s]), flags[s], nmod=int(nmod[i]))
        assert (int(nb[i]), int(ny[i]), int(tot[i])) == want, i


def test_mods_pooled_fit_and_ions(menv, orc):
    """Modified matches: E-values over both searches' scores, SP ions on the
    unmodified peptide with the printed length grown by the modified
    sites."""
    pl, ps, ms, pres, mres, dsp = (menv[k] for k in ("pl", "ps", "ms", "pres", "mres", "dsp"))
    fit, npt, nu = ps.evalue_fit(ms)
    segs = [np.concatenate([pres.scores[int(pres.off[s]):int(pres.off[s + 1])],
                            mres.scores[int(mres.mp_off[s]):int(mres.mp_off[s + 1])]])
            for s in range(dsp.num_spec)]
    _check_fits(orc, (fit, npt, nu), segs)
    nb, ny, tot = ms.sp_ions()
    flags = _oracle_flags(orc, dsp)
    ions = orc.prepare_ions(menv["prot"].seq, menv["mi"].ma)
    st, mi = mres.stages, menv["mi"]
    checked = 0
    for s in range(dsp.num_spec):
        for k in range(int(mres.top_cnt[s])):
            i = int(mres.mp_off[s]) + int(mres.top_pos[s, k])
            p, m = int(st["mp_pep"][i]), int(st["mp_mod"][i])
            want = orc.match_ions(ions, int(menv["tc"][p]), int(menv["tn"][p]), menv["r"][p],
                                  menv["mass"], int(dsp.zmax[s]), flags[s],
                                  nmod=int(mi.mod_sum[m]))
            assert (int(nb[s, k]), int(ny[s, k]), int(tot[s, k])) == want, (s, k)
            checked += 1
    assert checked > 20


def test_finish_matches(menv, orc):
    """Search.hs:115-141: merged order, E-value and Ions of every reported
    match."""
    pl, ps, ms, pres, mres, dsp = (menv[k] for k in ("pl", "ps", "ms", "pres", "mres", "dsp"))
    out = pl.finish_matches(ps, pres, 5, ms, mres)
    fit, _, _ = ps.evalue_fit(ms)
    merged = pl.merge_rankings(pres, mres, 5)
    flags = _oracle_flags(orc, dsp)
    ions = orc.prepare_ions(menv["prot"].seq, menv["mi"].ma)
    idx = menv["db"].idx_sorted()
    st, mi = mres.stages, menv["mi"]
    for s, row in enumerate(out):
        assert [(m.score, m.is_mod, m.index) for m in row] == merged[s]
        for m in row:
            want = orc.evalue(fit[s, 0], fit[s, 1], m.score)
            assert (np.isnan(m.evalue) and np.isnan(want)) or m.evalue == pytest.approx(want,
                                                                                      rel=TOL)
            if m.is_mod:
                i = int(mres.mp_off[s]) + m.index
                p, nm = int(st["mp_pep"][i]), int(mi.mod_sum[int(st["mp_mod"][i])])
            else:
                p, nm = int(idx[int(pres.begin[s]) + m.index]), 0
            b, y, t = orc.match_ions(ions, int(menv["tc"][p]), int(menv["tn"][p]), menv["r"][p],
                                     menv["mass"], int(dsp.zmax[s]), flags[s], nmod=nm)
            assert m.ions == (b + y, t)


def test_files_to_ranked_output(tmp_path, orc):
    """The whole path from files: a FASTA database (readFasta + makeSeqDB) and
    DTA spectra, the search, E-values and SP ions on the device, and the
    ranked tables (printScanResults) -- every printed match the lookup of its
    window position, with the device's ion counts and E-value."""
    from hfx_amd import formats, pipeline, synth

    prot = synth.make_proteins(200, seed=77)
    seq = prot.seq.tobytes().decode()
    fasta = "".join(f">sp|P{i:05d}|PROT{i}_TEST Protein {i} of the test database\n"
                    f"{seq[prot.seg[i]:prot.seg[i + 1]]}\n" for i in range(prot.count))
    mass = orc.mass_table({"C": 57.021464})
    db = formats.read_fasta(mass, text=fasta, rule=1, missed=1, min_mass=400.0, max_mass=7200.0)
    assert db.ions.tobytes() == prot.seq.tobytes()
    sp = synth.make_spectra(prot, db.r, db.tc, db.tn, 6, seed=78, charges=(2, 3),
                            static_mods={"C": 57.021464})
    pcr, z, mzs, its = [], [], [], []
    for s in range(sp.count):
        mh = float(np.float32(sp.pcr[s]) * sp.charge[s] - (sp.charge[s] - 1) * 1.007276)
        lines = [f"{mh:.4f} {int(sp.charge[s])}"]
        a, b = int(sp.off[s]), int(sp.off[s + 1])
        lines += [f"{float(m):.5f} {float(i):.5f}" for m, i in zip(sp.mz[a:b], sp.inten[a:b])]
        p, c, m, i = formats.parse_dta("\n".join(lines) + "\n")
        pcr.append(p), z.append(c), mzs.append(m), its.append(i)
    off = np.concatenate([[0], np.cumsum([len(m) for m in mzs])]).astype(np.uint32)
    ctx = pipeline.Ctx(0)
    ddb = pipeline.DeviceDB(pipeline.HostDB(db.ions, db.r, db.tc, db.tn, mass))
    dsp = pipeline.DeviceSpectra(ctx, np.float32(pcr), np.float32(z), np.concatenate(mzs),
                                 np.concatenate(its), off, STRIDE)
    ps = pipeline.PlainSearch(ctx, ddb, dsp, eps=3.0, topn=5)
    res = ps.run()
    out = pipeline.finish_matches(ps, res, 5)
    idx = ddb.idx_sorted()
    for s, matches in enumerate(out):
        rows = []
        for m in matches:
            k = int(idx[int(res.begin[s]) + m.index])
            label, p, mh = db.lookup(k)
            rows.append(dict(mh=float(mh), xcorr=float(np.float32(m.score / 10000.0)),
                             ions=m.ions, evalue=m.evalue, reference=db.headers[p].split()[0],
                             peptide=label, header=db.headers[p]))
        text = formats.format_scan("", f"spec{s}.dta", "test.fasta", 1, float(pcr[s]), float(z[s]),
                                   3.0, rows, 5, 3)
        lines = text.split("\n")
        assert lines[2].startswith("Database    :: test.fasta")
        body = [l for l in lines[8:8 + len(rows)]]
        for r, l in zip(rows, body):
            cells = l.split()
            assert cells[4] == f"{r['ions'][0]}/{r['ions'][1]}" and cells[6] == r["peptide"]
            assert cells[5] == r["reference"]
            assert cells[7] == formats.format_number(r["evalue"], 2, 4)
