"""GPU parity of the variable-modification path against the CPU oracle.

Every enumeration array (windows per modification, modable rows, counts,
combination suffix products, modified-peptide records, unrank slices) must be
bit-identical; modified theoretical rows identical; FP32 scores within 1e-4
of the exact score relative to max(|exact|, 100); rankings and exact scores
identical.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STRIDE = 2112
VAR = {"M": 15.9949, "S": 79.9663, "T": 79.9663, "Y": 79.9663}
MAX_MA = 4  # at most 3 modified residues per peptide


@pytest.fixture(scope="module")
def menv(orc):
    import hfx_amd
    from hfx_amd import pipeline, synth

    prot = synth.make_proteins(600, seed=31)
    mass = orc.mass_table()
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=1, p_rule=1, min_len=7,
                           max_len=30)
    sp = synth.make_spectra(prot, r, tc, tn, 40, seed=32, charges=(1, 2, 3, 4))
    ctx = pipeline.Ctx(0)
    mi = pipeline.mod_info(VAR, MAX_MA)
    db = pipeline.DeviceDB(pipeline.HostDB(prot.seq, r, tc, tn, mass), ma=mi.ma)
    dsp = pipeline.DeviceSpectra(ctx, sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE)
    ms = pipeline.ModSearch(ctx, db, dsp, mi, eps=3.0, topn=5)
    res = ms.run(keep_scores=True)
    return dict(prot=prot, mass=mass, r=r, tc=tc, tn=tn, sp=sp, ctx=ctx, mi=mi, db=db, dsp=dsp,
                ms=ms, res=res, L=hfx_amd.lib(), pl=pipeline)


def test_mod_table_and_packing(menv, orc):
    mi = menv["mi"]
    o = orc.mod_info(VAR, MAX_MA)
    assert mi.ma == o[0] == b"MSTY" and (mi.mod_ma_count == o[2]).all() and mi.num_mod == 34
    assert (menv["db"].packed_ions() == orc.prepare_ions(menv["prot"].seq, mi.ma)).all()


def test_enumeration_bit_exact(menv, orc):
    res, st, mi, db, dsp = menv["res"], menv["res"].stages, menv["mi"], menv["db"], menv["dsp"]
    r, tc, tn = menv["r"], menv["tc"], menv["tn"]
    S, M, na = dsp.num_spec, mi.num_mod, mi.num_ma
    idx, _ = orc.sort_idx(r)
    ions = orc.prepare_ions(menv["prot"].seq, mi.ma)
    b = np.zeros(S * M, np.uint32)
    e = np.zeros(S * M, np.uint32)
    sc = np.zeros(S * M, np.uint32)
    cand = np.zeros(S, np.uint32)
    for s in range(S):
        bb, ee, ss, tot = orc.find_begin_end(r, idx, mi.mod_delta, dsp.mass[s], 3.0)
        b[s * M:(s + 1) * M], e[s * M:(s + 1) * M], sc[s * M:(s + 1) * M] = bb, ee, ss
        cand[s] = tot
    assert (st["begin"] == b).all() and (st["end"] == e).all() and (st["num_scan"] == sc).all()
    assert (st["spec_cand"] == cand).all()
    scan = np.concatenate([[0], np.cumsum(cand)]).astype(np.uint32)
    nv, pep, pmod, pcnt, vs = orc.find_modable_peptides(scan, scan[-1], ions, tc, tn, idx, b, e, sc,
                                                        mi.mod_ma_count, mi.ma)
    assert res.num_valid == nv > 0
    assert (st["valid_pep"] == pep).all() and (st["valid_mod"] == pmod).all()
    assert (st["valid_cnt"] == pcnt).all() and (st["valid_scan"] == vs).all()
    nmp, num, comb = orc.calc_total_mod_cands(mi.mod_ma_count, pep, pmod, pcnt, na)
    assert res.num_mpep == nmp > 0
    assert (st["pep_num_mpep"] == num).all() and (st["comb"] == comb).all()
    lu, mp_pep, mp_mod, mp_rank, mp_ith, mp_scan = orc.prepare_gen_mod(mi.mod_sum, pep, pmod, num,
                                                                       nmp)
    assert res.len_unrank == lu
    for k, v in (("mp_pep", mp_pep), ("mp_mod", mp_mod), ("mp_rank", mp_rank), ("mp_ith", mp_ith),
                 ("mp_scan", mp_scan)):
        assert (st[k] == v).all(), k
    unrank = orc.gen_mod_cands(lu, mi.mod_ma_count, mp_ith, mp_rank, mp_scan, pmod, pcnt, comb, na)
    assert (st["unrank"] == unrank).all()
    # Search.hs:381-386 per-spectrum blocks
    row_off = np.concatenate([[0], np.cumsum(num.astype(np.uint64))])
    expect = np.array([row_off[v] for v in vs] + [row_off[-1]], np.uint64)
    assert (res.mp_off == expect).all()


def test_add_mod_ions_dense_rows(menv, orc):
    pl, L, res, mi, db = menv["pl"], menv["L"], menv["res"], menv["mi"], menv["db"]
    st = res.stages
    ions = db.packed_ions()
    rng = np.random.default_rng(9)
    sel = rng.choice(res.num_mpep, size=min(400, res.num_mpep), replace=False)
    ms = menv["ms"]
    for zmax in (1, 2, 3, 5):
        length = 2075
        out = pl.empty(4 * res.num_mpep * length)
        t = ms._keep
        L.addModIons(out.p, db.t_r.p, db.t_mass.p, db.t_ions.p, db.t_tc.p, db.t_tn.p, t[0][0].p,
                     t[0][1].p, t[1].p, t[0][4].p, res.len_unrank, res.num_mpep, ms.t_mod_cnt.p,
                     ms.t_mod_delta.p, ms.t_ma.p, ms.t_ma_mass.p, mi.num_ma, length, zmax, None)
        pl.sync()
        got = pl.host(out, np.float32).reshape(res.num_mpep, length)
        for i in sel:
            p, m = int(st["mp_pep"][i]), int(st["mp_mod"][i])
            u0 = int(st["mp_scan"][i])
            k = int(mi.mod_sum[m])
            ref = orc.mpeaks_dense(ions, int(menv["tc"][p]), int(menv["tn"][p]), menv["r"][p],
                                   menv["mass"], zmax, length, mi.mod_ma_count[m],
                                   mi.mod_delta[m], mi.ma, mi.ma_mass, st["unrank"][u0:u0 + k])
            assert (got[i] == ref.astype(np.float32)).all(), (zmax, i)


def test_mod_scores_and_ranking(menv, orc):
    res, mi, dsp = menv["res"], menv["mi"], menv["dsp"]
    idx, _ = orc.sort_idx(menv["r"])
    ions = orc.prepare_ions(menv["prot"].seq, mi.ma)
    ref = orc.search_mods(ions, menv["r"], menv["tc"], menv["tn"], menv["mass"], idx, mi.ma,
                          mi.ma_mass, mi.mod_ma_count, mi.mod_sum, mi.mod_delta, dsp.x(),
                          dsp.lengths(), dsp.zmax, dsp.mass, 3.0, 5)
    assert (np.diff(res.mp_off).astype(np.uint32) == ref["spec_num_mpep"]).all()
    # every FP32 score against the exact score of its modified peptide
    st = res.stages
    X, lens = dsp.x(), dsp.lengths()
    rng = np.random.default_rng(3)
    for s in range(dsp.num_spec):
        a, b = int(res.mp_off[s]), int(res.mp_off[s + 1])
        for i in rng.choice(np.arange(a, b), size=min(40, b - a), replace=False) if b > a else []:
            p, m = int(st["mp_pep"][i]), int(st["mp_mod"][i])
            u0, k = int(st["mp_scan"][i]), int(mi.mod_sum[m])
            w = orc.mpeaks_dense(ions, int(menv["tc"][p]), int(menv["tn"][p]), menv["r"][p],
                                 menv["mass"], int(dsp.zmax[s]), int(lens[s]), mi.mod_ma_count[m],
                                 mi.mod_delta[m], mi.ma, mi.ma_mass, st["unrank"][u0:u0 + k])
            ex = orc.fx_to_double(orc.score_fx(w, X[s, : lens[s]]))
            assert abs(float(res.scores[i]) - ex) <= 1e-4 * max(abs(ex), 100.0)
    assert (res.top_cnt == ref["top_cnt"]).all()
    for s in range(dsp.num_spec):
        c = int(res.top_cnt[s])
        assert res.top_pos[s, :c].tolist() == ref["top_pos"][s, :c].tolist(), s
        assert (res.top_exact[s, :c] == orc.fx_to_double(ref["top_fx"][s, :c])).all()


def test_merged_ranking(menv, orc):
    """Search.hs:118-123: unmodified before modified on equal scores."""
    pl = menv["pl"]
    plain = pl.PlainSearch(menv["ctx"], menv["db"], menv["dsp"], eps=3.0, topn=5).run()
    merged = pl.merge_rankings(plain, menv["res"], 5)
    for s, items in enumerate(merged):
        scores = [t[0] for t in items]
        assert scores == sorted(scores, reverse=True)
        for a, b in zip(items, items[1:]):
            if a[0] == b[0]:
                assert (a[1], a[2]) < (b[1], b[2])


def test_exact_route_matches(menv, orc):
    """hfx_topn_exact_mod for every spectrum: the all-exact ranking equals the
    oracle's (which ranks by exact score)."""
    res, mi, dsp, ms = menv["res"], menv["mi"], menv["dsp"], menv["ms"]
    idx, _ = orc.sort_idx(menv["r"])
    ions = orc.prepare_ions(menv["prot"].seq, mi.ma)
    ref = orc.search_mods(ions, menv["r"], menv["tc"], menv["tn"], menv["mass"], idx, mi.ma,
                          mi.ma_mass, mi.mod_ma_count, mi.mod_sum, mi.mod_delta, dsp.x(),
                          dsp.lengths(), dsp.zmax, dsp.mass, 3.0, 5)
    pos, ex, cnt = ms.exact(np.arange(dsp.num_spec, dtype=np.uint32))
    assert (cnt == ref["top_cnt"]).all()
    for s in range(dsp.num_spec):
        c = int(cnt[s])
        assert pos[s, :c].tolist() == ref["top_pos"][s, :c].tolist(), s
        assert (ex[s, :c] == orc.fx_to_double(ref["top_fx"][s, :c])).all()
        # the gated route agrees with the all-exact route
        assert pos[s, :c].tolist() == res.top_pos[s, :c].tolist(), s


@pytest.fixture(scope="module")
def cfg3(orc):
    """The uneven modified workload's database (2,000 proteins, M ox + S/T/Y
    phospho, at most 3 sites) with 1,000 of its spectra (1+..4+, half from
    modified peptides)."""
    import hfx_amd
    from hfx_amd import pipeline, synth

    w = synth.WORKLOADS["mods-2k-50k"]
    prot = synth.make_proteins(w["proteins"], seed=w["seed"])
    mass = orc.mass_table(w["static"])
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=w["missed"],
                           p_rule=w["p_rule"], min_len=w["min_len"], max_len=w["max_len"])
    sp = synth.make_spectra(prot, r, tc, tn, 1000, seed=w["seed"] + 100, charges=w["charges"],
                            charge_p=w["charge_p"], double_for_3plus=w["double3"],
                            var_mods=w["var"])
    ctx = pipeline.Ctx(0)
    mi = pipeline.mod_info(w["var"], w["max_ma"])
    db = pipeline.DeviceDB(pipeline.HostDB(prot.seq, r, tc, tn, mass), ma=mi.ma)
    dsp = pipeline.DeviceSpectra(ctx, sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE)
    ms = pipeline.ModSearch(ctx, db, dsp, mi, eps=3.0, topn=w["topn"])
    return dict(prot=prot, mass=mass, r=r, tc=tc, tn=tn, sp=sp, ctx=ctx, mi=mi, db=db,
                dsp=dsp, ms=ms, w=w, L=hfx_amd.lib(), pl=pipeline)


def test_mod_windows_all_spectra_match_per_spectrum_calls(cfg3):
    """hfx_mod_windows (one launch) against findBeginEnd_f called once per
    spectrum (Search.hs:317-322): identical windows, scans and counts."""
    a = cfg3["ms"].run(per_spectrum=True)
    b = cfg3["ms"].run()
    for k in ("begin", "end", "num_scan"):
        assert (a.stages[k] == b.stages[k]).all(), k
    assert (a.spec_cand == b.spec_cand).all()
    assert (a.top_cnt == b.top_cnt).all()
    for s in range(len(a.top_cnt)):
        c = int(a.top_cnt[s])
        assert (a.top_pos[s, :c] == b.top_pos[s, :c]).all(), s
        assert (a.top_exact[s, :c] == b.top_exact[s, :c]).all(), s


def test_mods_uneven_workload_against_oracle(cfg3, orc):
    """1,000 spectra of the uneven workload: per-spectrum modified-candidate
    counts (heavy-tailed), top-5 positions and exact scores identical to the
    oracle's search."""
    mi, dsp = cfg3["mi"], cfg3["dsp"]
    res = cfg3["ms"].run()
    idx, _ = orc.sort_idx(cfg3["r"])
    ions = orc.prepare_ions(cfg3["prot"].seq, mi.ma)
    ref = orc.search_mods(ions, cfg3["r"], cfg3["tc"], cfg3["tn"], cfg3["mass"], idx, mi.ma,
                          mi.ma_mass, mi.mod_ma_count, mi.mod_sum, mi.mod_delta, dsp.x(),
                          dsp.lengths(), dsp.zmax, dsp.mass, 3.0, cfg3["w"]["topn"])
    n = np.diff(res.mp_off).astype(np.uint32)
    assert (n == ref["spec_num_mpep"]).all()
    print(f"modified candidates per spectrum: {int(n.sum())} total, max {int(n.max())}, "
          f"median {int(np.median(n))}, zero in {int((n == 0).sum())}")
    assert (res.top_cnt == ref["top_cnt"]).all()
    bad = [s for s in range(dsp.num_spec)
           if res.top_pos[s, :res.top_cnt[s]].tolist() != ref["top_pos"][s, :ref["top_cnt"][s]].tolist()
           or not (res.top_exact[s, :res.top_cnt[s]] ==
                   orc.fx_to_double(ref["top_fx"][s, :ref["top_cnt"][s]])).all()]
    assert not bad, bad[:10]


def test_mods_split_by_memory_matches_one_batch(cfg3):
    """search_mods_split with a budget that cuts the 1,000 spectra into
    several batches: the per-spectrum modified-peptide counts it sizes with
    equal the one-batch run's, every batch fits the budget, and the merged
    result (windows, counts, offsets, top N, exact scores, flags) equals the
    one-batch run's."""
    pl, ms = cfg3["pl"], cfg3["ms"]
    one = ms.run(keep_stages=False)
    per = ms.mpep_per_spectrum()
    assert (per == np.diff(one.mp_off)).all()
    budget = int(per.sum()) * pl.MPEP_BYTES // 5
    res, groups = pl.search_mods_split(cfg3["ctx"], cfg3["db"], cfg3["dsp"], cfg3["mi"], 3.0,
                                       cfg3["w"]["topn"], budget=budget)
    print(f"{len(groups)} batches for {int(per.sum())} modified peptides")
    assert len(groups) >= 5
    for a, b in groups:
        assert int(per[a:b].sum()) * pl.MPEP_BYTES <= budget
    assert (res.spec_cand == one.spec_cand).all()
    assert res.num_valid == one.num_valid and res.num_mpep == one.num_mpep
    assert res.len_unrank == one.len_unrank
    assert (res.mp_off == one.mp_off).all()
    assert (res.top_cnt == one.top_cnt).all()
    for s in range(len(one.top_cnt)):
        c = int(one.top_cnt[s])
        assert (res.top_pos[s, :c] == one.top_pos[s, :c]).all(), s
        assert (res.top_exact[s, :c] == one.top_exact[s, :c]).all(), s
        assert (res.top_f32[s, :c] == one.top_f32[s, :c]).all(), s
    assert res.flagged.tolist() == one.flagged.tolist()
    assert res.short.tolist() == one.short.tolist()


def test_mods_full_uneven_workload_against_oracle(orc):
    """The whole uneven workload searched on the GPU (50,000 spectra, 107 M
    modified candidates); every fifth spectrum (10,000) checked against the
    oracle's search on all host threads: per-spectrum counts, top-5
    positions and exact scores identical."""
    import time

    from hfx_amd import pipeline, synth

    w = synth.WORKLOADS["mods-2k-50k"]
    prot = synth.make_proteins(w["proteins"], seed=w["seed"])
    mass = orc.mass_table(w["static"])
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=w["missed"],
                           p_rule=w["p_rule"], min_len=w["min_len"], max_len=w["max_len"])
    sp = synth.make_spectra(prot, r, tc, tn, w["spectra"], seed=w["seed"] + 100,
                            charges=w["charges"], charge_p=w["charge_p"],
                            double_for_3plus=w["double3"], var_mods=w["var"])
    ctx = pipeline.Ctx(0)
    mi = pipeline.mod_info(w["var"], w["max_ma"])
    db = pipeline.DeviceDB(pipeline.HostDB(prot.seq, r, tc, tn, mass), ma=mi.ma)
    dsp = pipeline.DeviceSpectra(ctx, sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE)
    res = pipeline.ModSearch(ctx, db, dsp, mi, eps=3.0, topn=w["topn"]).run(keep_stages=False)
    n = np.diff(res.mp_off).astype(np.uint32)
    sel = np.arange(0, dsp.num_spec, 5)
    t0 = time.time()
    idx, _ = orc.sort_idx(r)
    ref = orc.search_mods(orc.prepare_ions(prot.seq, mi.ma), r, tc, tn, mass, idx, mi.ma,
                          mi.ma_mass, mi.mod_ma_count, mi.mod_sum, mi.mod_delta, dsp.x()[sel],
                          dsp.lengths()[sel], dsp.zmax[sel], dsp.mass[sel], 3.0, w["topn"],
                          cap_mpep=int(n[sel].sum()) + 1024,
                          cap_unrank=int(res.len_unrank) + 1024)
    print(f"{dsp.num_spec} spectra, {int(n.sum())} modified candidates; {len(sel)} checked, "
          f"oracle {time.time() - t0:.1f}s", flush=True)
    assert (n[sel] == ref["spec_num_mpep"]).all()
    assert (res.top_cnt[sel] == ref["top_cnt"]).all()
    bad = [int(s) for i, s in enumerate(sel)
           if res.top_pos[s, :res.top_cnt[s]].tolist() != ref["top_pos"][i, :ref["top_cnt"][i]].tolist()
           or not (res.top_exact[s, :res.top_cnt[s]] ==
                   orc.fx_to_double(ref["top_fx"][i, :ref["top_cnt"][i]])).all()]
    assert not bad, bad[:10]
