"""Full-size GPU parity on the two benchmark workloads:

  tryptic-20k-10k  20,000 proteins (<= 2 missed cleavages, fixed C+57.02146),
                   10,000 spectra at 2+/3+, top-10;
  tryptic-20k-20k  the same database, 20,000 spectra at 2+/3+, top-5;

through the plain FP32 form, the batched FP32 form, the batched BF16-split
dense form and the batched integer dense form, each against the oracle
(dense-row scoring, exact fixed-point scores, OpenMP):
  - windows identical; top-N positions and exact scores identical for every
    spectrum;
  - every score of every form within 1e-4 * max(|exact|, FLOOR * scale),
    scale = sum_b w_b |x_b| over the pair's theoretical bins (the same
    search over |x|), FLOOR = 1/4: relative 1e-4 wherever the score is at
    least a quarter of its own magnitude scale, 2.5e-5 of that scale where
    heavy cancellation of the mean-subtracted x makes it smaller.  Both forms
    err in proportion to the scale, not to the score: the hi/lo split
    represents each x_b to 2^-17 relative, and each of the <= 2 * 2 * 4 * 49
    nonzero products adds at most 2^-24 of the running sum, a worst case of
    ~5e-5 * scale whose random-sign typical value is ~1e-6 * scale;
  - every BF16 dense score within its derived bound (2^-16 + 210 * 2^-24) *
    50 * sum|x|, every integer-form score within max|x| / 8355711 * 50 len +
    2^-24 * 50 sum|x|;
  - the dense scores of each spectrum's top N within 1e-4 relative of their
    exact values.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STRIDE = 2112
FLOOR = 0.25
REL_DENSE = 2.0**-16 + 210 * 2.0**-24


@pytest.fixture(scope="module", params=["tryptic-20k-10k", "tryptic-20k-20k"])
def fenv(request, orc):
    from hfx_amd import pipeline, synth

    name = request.param
    w = synth.WORKLOADS[name]
    prot = synth.make_proteins(w["proteins"], seed=w["seed"])
    mass = orc.mass_table(w["static"])
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=w["missed"],
                           p_rule=w["p_rule"], min_len=w["min_len"], max_len=w["max_len"])
    sp = synth.make_spectra(prot, r, tc, tn, w["spectra"], seed=w["seed"] + 100,
                            charges=w["charges"], charge_p=w["charge_p"], static_mods=w["static"],
                            double_for_3plus=w["double3"])
    ctx = pipeline.Ctx(0)
    db = pipeline.DeviceDB(pipeline.HostDB(prot.seq, r, tc, tn, mass))
    dsp = pipeline.DeviceSpectra(ctx, sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE)
    assert (dsp.status() == 0).all()
    idx, _ = orc.sort_idx(r)
    ions = orc.prepare_ions(prot.seq)
    X, lens = dsp.x(), dsp.lengths()
    topn = w["topn"]
    ref = orc.search_plain(ions, r, tc, tn, mass, idx, X, lens, dsp.zmax, dsp.mass, 3.0, topn,
                           want_all=True)
    absx = orc.search_plain(ions, r, tc, tn, mass, idx, np.abs(X), lens, dsp.zmax, dsp.mass, 3.0,
                            topn, want_all=True)
    x1 = np.array([np.abs(X[s, : lens[s]].astype(np.float64)).sum() for s in range(dsp.num_spec)])
    reps = np.diff(ref["off"].astype(np.int64))
    xm = np.array([np.abs(X[s, : lens[s]]).max() if lens[s] else 0.0
                   for s in range(dsp.num_spec)], np.float64)
    b8 = (xm / 8355711.0 * 50 * lens.astype(np.float64) + 2.0**-24 * 50 * x1) * 1.001 + 1e-6
    print(f"{name}: {len(r)} peptides, {dsp.num_spec} spectra, {int(ref['off'][-1])} scores")
    return dict(name=name, ctx=ctx, db=db, dsp=dsp, ref=ref, topn=topn, pl=pipeline,
                r=r, tc=tc, tn=tn, mass=mass, ions=ions, idx=idx,
                exact=orc.fx_to_double(ref["all_fx"]), scale=orc.fx_to_double(absx["all_fx"]),
                x1=np.repeat(x1, reps), bound8=np.repeat(b8, reps))


def check_topn(res, ref, orc, S):
    assert (res.begin == ref["begin"]).all() and (res.num == ref["num"]).all()
    assert (res.top_cnt == ref["top_cnt"]).all()
    bad = [s for s in range(S)
           if res.top_pos[s, : res.top_cnt[s]].tolist() != ref["top_pos"][s, : ref["top_cnt"][s]].tolist()
           or not (res.top_exact[s, : res.top_cnt[s]] ==
                   orc.fx_to_double(ref["top_fx"][s, : ref["top_cnt"][s]])).all()]
    assert not bad, bad[:10]


def check_scores(fenv, res, form):
    exact, scale = fenv["exact"], fenv["scale"]
    assert len(res.scores) == len(exact)
    err = np.abs(res.scores.astype(np.float64) - exact)
    floor = np.maximum(np.abs(exact), FLOOR * scale)
    big = np.abs(exact) >= FLOOR * scale
    es = err / np.maximum(scale, 1e-30)
    print(f"{fenv['name']} {form}: {len(err)} scores, max err/scale {float(es.max()):.3g}, "
          f"99.99% {float(np.quantile(es, 0.9999)):.3g}; max rel over the {int(big.sum())} "
          f"scores >= {FLOOR} scale: {float((err[big] / floor[big]).max()) if big.any() else 0:.3g}")
    assert np.all(err <= 1e-4 * floor + 1e-6), float((err / np.maximum(floor, 1e-30)).max())
    return err


def test_plain_full(fenv, orc):
    res = fenv["pl"].PlainSearch(fenv["ctx"], fenv["db"], fenv["dsp"], eps=3.0,
                                 topn=fenv["topn"]).run(keep_scores=True)
    check_topn(res, fenv["ref"], orc, fenv["dsp"].num_spec)
    check_scores(fenv, res, "plain fp32")


def test_batched_fp32_full(fenv, orc):
    bs = fenv["pl"].BatchSearch(fenv["ctx"], fenv["db"], fenv["dsp"], eps=3.0, topn=fenv["topn"])
    bs.run()
    res = bs.results(keep_scores=True)
    check_topn(res, fenv["ref"], orc, fenv["dsp"].num_spec)
    check_scores(fenv, res, "batched fp32")


def test_batched_i8_full(fenv, orc):
    """the integer form: within max|x| / 8355711 * 50 len + 2^-24 * 50
    sum|x| of the exact score, besides the common 1e-4 criterion"""
    bs = fenv["pl"].BatchSearch(fenv["ctx"], fenv["db"], fenv["dsp"], eps=3.0, topn=fenv["topn"])
    bs.set_precision(bs.I8X3_DENSE)
    bs.run()
    res = bs.results(keep_scores=True)
    check_topn(res, fenv["ref"], orc, fenv["dsp"].num_spec)
    err = check_scores(fenv, res, "batched i8x3")
    assert np.all(err <= fenv["bound8"])
    assert len(res.short) == 0


def test_batched_dense_full(fenv, orc):
    bs = fenv["pl"].BatchSearch(fenv["ctx"], fenv["db"], fenv["dsp"], eps=3.0, topn=fenv["topn"])
    bs.set_precision(bs.BF16X2_DENSE)
    bs.run()
    res = bs.results(keep_scores=True)
    check_topn(res, fenv["ref"], orc, fenv["dsp"].num_spec)
    err = check_scores(fenv, res, "batched dense")
    assert np.all(err <= REL_DENSE * 50 * fenv["x1"] * 1.001 + 1e-6)
    for s in range(fenv["dsp"].num_spec):
        c = int(res.top_cnt[s])
        ex = res.top_exact[s, :c]
        assert np.all(np.abs(res.top_f32[s, :c] - ex) <= 1e-4 * np.maximum(np.abs(ex), 1.0)), s


def test_stats_full(fenv, orc):
    """E-value fit of every spectrum over its plain FP32 scores (histogram
    points and kept counts identical, alpha and beta within 1e-9) and the SP
    ion counts of every reported match, against the oracle."""
    pl, dsp = fenv["pl"], fenv["dsp"]
    ps = pl.PlainSearch(fenv["ctx"], fenv["db"], dsp, eps=3.0, topn=fenv["topn"])
    res = ps.run(keep_scores=True)
    fit, npt, nu = ps.evalue_fit()
    bad = []
    for s in range(dsp.num_spec):
        a, b, m, nused = orc.evalue_fit(res.scores[int(res.off[s]):int(res.off[s + 1])])
        ok = (int(npt[s]), int(nu[s])) == (m, nused)
        for got, want in ((fit[s, 0], a), (fit[s, 1], b)):
            ok &= (np.isnan(got) and np.isnan(want)) or abs(got - want) <= 1e-9 * max(abs(want), 1)
        if not ok:
            bad.append(s)
    print(f"{fenv['name']}: {int((npt >= 2).sum())} of {dsp.num_spec} spectra with a fit")
    assert not bad, bad[:10]
    nb, ny, tot = ps.sp_ions()
    X, lens = dsp.x(), dsp.lengths()
    idx, ions = fenv["idx"], fenv["ions"]
    r, tc, tn, mass = fenv["r"], fenv["tc"], fenv["tn"], fenv["mass"]
    for s in range(dsp.num_spec):
        flags = orc.extract_peaks(X[s, :int(lens[s])])
        for k in range(int(res.top_cnt[s])):
            p = int(idx[int(res.begin[s]) + int(res.top_pos[s, k])])
            want = orc.match_ions(ions, int(tc[p]), int(tn[p]), r[p], mass, int(dsp.zmax[s]),
                                  flags)
            assert (int(nb[s, k]), int(ny[s, k]), int(tot[s, k])) == want, (s, k)
