"""GPU parity of the precision form (dense BF16 hi+lo MFMA XCorr) against the
CPU oracle.

Every dense score lies within the stated bound of its exact score,
|dense - exact| <= (2^-16 + 210 * 2^-24) * 50 * sum|x|, and within 1e-4 of
its magnitude scale sum_b w_b |x_b| (the exact score itself can be a small
difference of large terms: x is mean-subtracted); the dense scores of each
spectrum's top N are within 1e-4 of their exact values; top-N positions and
exact scores equal the oracle's; the result does not depend on how spectra
are blocked; the batched dense form (graph-replayed, autotuned) gives the
same rankings.  Spectra cover fragment charges 1-4 (precursor 2+ to 5+), two
spectra whose fragment charge exceeds 4 (scored 0) and an empty window.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STRIDE = 2112
REL = 2.0**-16 + 210 * 2.0**-24


@pytest.fixture(scope="module")
def denv(orc):
    from hfx_amd import pipeline, synth

    prot = synth.make_proteins(3000, seed=51)
    mass = orc.mass_table({"C": 57.02146})
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=2)
    sp = synth.make_spectra(prot, r, tc, tn, 400, seed=52, charges=(2, 3, 4, 5),
                            charge_p=(0.4, 0.3, 0.2, 0.1), static_mods={"C": 57.02146},
                            double_for_3plus=True)
    H = 1.007276
    for s, z in ((7, 7.0), (19, 9.0)):
        z0 = float(sp.charge[s])
        sp.pcr[s] = np.float32((float(sp.pcr[s]) * z0 - z0 * H + z * H) / z)
        sp.charge[s] = z
    sp.pcr[3] = np.float32(90.0)  # far below every peptide: empty window
    ctx = pipeline.Ctx(0)
    db = pipeline.DeviceDB(pipeline.HostDB(prot.seq, r, tc, tn, mass))
    dsp = pipeline.DeviceSpectra(ctx, sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE)
    idx, _ = orc.sort_idx(r)
    ions = orc.prepare_ions(prot.seq)
    ref = orc.search_plain(ions, r, tc, tn, mass, idx, dsp.x(), dsp.lengths(), dsp.zmax, dsp.mass,
                           3.0, 10, want_all=True)
    # sum_b w_b |x_b| of every scored pair: the same search over |x|
    absx = orc.search_plain(ions, r, tc, tn, mass, idx, np.abs(dsp.x()), dsp.lengths(), dsp.zmax,
                            dsp.mass, 3.0, 10, want_all=True)
    ref["all_scale"] = orc.fx_to_double(absx["all_fx"])
    return dict(ctx=ctx, db=db, dsp=dsp, ref=ref, pl=pipeline)


def _bound(dsp, res):
    """per-score error bound of the dense form"""
    X, lens = dsp.x(), dsp.lengths()
    x1 = np.array([np.abs(X[s, : lens[s]].astype(np.float64)).sum() for s in range(dsp.num_spec)])
    per_spec = REL * 50.0 * x1 * 1.001 + 1e-6
    return np.repeat(per_spec, np.diff(res.off.astype(np.int64)))


def check_scores(res, ref, dsp, orc):
    assert (res.begin == ref["begin"]).all() and (res.num == ref["num"]).all()
    exact = orc.fx_to_double(ref["all_fx"][: len(res.scores)])
    scale = ref["all_scale"][: len(res.scores)]
    err = np.abs(res.scores.astype(np.float64) - exact)
    bound = _bound(dsp, res)
    ratio = float((err / bound).max()) if len(err) else 0.0
    rel_scale = float((err / np.maximum(scale, 1e-30)).max())
    over = float(np.mean(err > 1e-4 * np.maximum(np.abs(exact), 100.0)))
    print(f"dense: {len(err)} scores, max |err| = {err.max():.3g}, max err/bound = {ratio:.3g}, "
          f"max err/sum(w|x|) = {rel_scale:.3g}, share above 1e-4*max(|exact|,100) = {over:.3g}")
    assert np.all(err <= bound), ratio
    assert np.all(err <= 1e-4 * scale + 1e-6), rel_scale
    # the dense scores of the reported top N
    for s in range(len(res.top_cnt)):
        c = int(res.top_cnt[s])
        ex = res.top_exact[s, :c]
        assert np.all(np.abs(res.top_f32[s, :c] - ex) <= 1e-4 * np.maximum(np.abs(ex), 1.0)), s


def check_topn(res, ref, orc):
    assert (res.top_cnt == ref["top_cnt"]).all()
    for s in range(len(res.top_cnt)):
        c = int(res.top_cnt[s])
        assert res.top_pos[s, :c].tolist() == ref["top_pos"][s, :c].tolist(), s
        assert (res.top_exact[s, :c] == orc.fx_to_double(ref["top_fx"][s, :c])).all(), s


def test_dense_scores_within_bound(denv, orc):
    pl = denv["pl"]
    srch = pl.PlainSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=10)
    res = srch.run(keep_scores=True, dense=True)
    check_scores(res, denv["ref"], denv["dsp"], orc)
    check_topn(res, denv["ref"], orc)
    assert res.num[3] == 0
    for s in (7, 19):
        a, b = int(res.off[s]), int(res.off[s + 1])
        assert b > a and np.all(res.scores[a:b] == 0)


def test_dense_blocking_does_not_change_scores(denv, orc):
    """Each score is one (row, column) of the contraction: the block a spectrum
    lands in changes neither its row values nor its column, so any spectrum
    order gives identical scores."""
    pl, dsp = denv["pl"], denv["dsp"]
    srch = pl.PlainSearch(denv["ctx"], denv["db"], dsp, eps=3.0, topn=10)
    srch.window()
    srch.score_dense(None)
    a = pl.host(srch.t_scores, np.float32, srch.total).copy()
    order = np.lexsort((dsp.mass, dsp.zmax)).astype(np.uint32)
    srch.score_dense(order)
    b = pl.host(srch.t_scores, np.float32, srch.total).copy()
    rev = order[::-1].copy()
    srch.score_dense(rev)
    c = pl.host(srch.t_scores, np.float32, srch.total).copy()
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_batch_dense_matches_oracle(denv, orc):
    pl = denv["pl"]
    bs = pl.BatchSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=10)
    bs.set_precision(bs.BF16X2_DENSE)
    assert bs.info()["precision"] == 1
    bs.run()
    res = bs.results(keep_scores=True)
    check_scores(res, denv["ref"], denv["dsp"], orc)
    check_topn(res, denv["ref"], orc)


SEG = 32  # order positions per segment of the device's block cut


def test_batch_dense_blocks_cover_the_order(denv):
    """The run orders the spectra by (fragment charge, precursor mass, index)
    on the device, and its dense blocks are consecutive runs of one to eight
    spectra of that order covering every spectrum, cut inside segments of 32
    order positions; their total cost -- charge levels x rows of the hull of
    the windows, summed over blocks -- is the minimum over all such cuts (an
    independent dynamic programme per segment here; ties may cut
    differently), and no worse than fixed runs of eight."""
    pl, dsp = denv["pl"], denv["dsp"]
    bs = pl.BatchSearch(denv["ctx"], denv["db"], dsp, eps=3.0, topn=10)
    bs.set_precision(bs.BF16X2_DENSE)
    bs.run()
    res = bs.results()
    S = dsp.num_spec
    z = dsp.zmax.astype(np.int64)
    z = np.where(z <= 4, z, 0)
    order = np.lexsort((np.arange(S), dsp.mass, z))
    assert np.array_equal(bs.order(), order)
    st = bs.dense_blocks().astype(np.int64)
    assert st[0] == 0 and st[-1] == S
    sizes = np.diff(st)
    assert sizes.min() >= 1 and sizes.max() <= 8
    # no block crosses a segment boundary
    assert all((a // SEG) == ((c - 1) // SEG) for a, c in zip(st[:-1], st[1:]))
    b, n = res.begin.astype(np.int64), res.num.astype(np.int64)

    def cost(lo, hi):
        blk = order[lo:hi]
        blk = blk[(z[blk] > 0) & (n[blk] > 0)]
        if len(blk) == 0:
            return 0
        return len(set(z[blk].tolist())) * int((b[blk] + n[blk]).max() - b[blk].min())

    plan = sum(cost(int(a), int(c)) for a, c in zip(st[:-1], st[1:]))
    total = 0
    for p0 in range(0, S, SEG):
        p1 = min(S, p0 + SEG)
        best = [0] + [None] * (p1 - p0)
        for j in range(1, p1 - p0 + 1):
            best[j] = min(best[j - k] + cost(p0 + j - k, p0 + j) for k in range(1, min(8, j) + 1))
        total += best[p1 - p0]
    assert plan == total
    assert plan <= sum(cost(i, min(S, i + 8)) for i in range(0, S, 8))


@pytest.mark.parametrize("chunk", [16, 48, 512, 4096])
def test_batch_dense_chunks_agree(denv, orc, chunk):
    pl = denv["pl"]
    bs = pl.BatchSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=10)
    bs.set_precision(bs.BF16X2_DENSE)
    bs.set_config(0, chunk, 256)
    assert bs.info()["chunk"] == chunk
    bs.run()
    res = bs.results(keep_scores=True)
    check_topn(res, denv["ref"], orc)


def test_batch_dense_graph_and_autotune(denv, orc):
    pl = denv["pl"]
    bs = pl.BatchSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=10)
    bs.set_precision(bs.BF16X2_DENSE)
    G, Cn, T = bs.autotune()
    assert G == 0 and Cn in (256, 512, 1024, 2048, 4096) and T == 1024
    bs.run()
    first = bs.results(keep_scores=True)
    check_topn(first, denv["ref"], orc)
    bs.capture()
    for _ in range(3):
        bs.replay()
        again = bs.results(keep_scores=True)
        assert np.array_equal(again.scores, first.scores)
        check_topn(again, denv["ref"], orc)


def test_dense_light_first_residue(orc):
    """A static modification that makes glycine 20 Da: peptides that start
    with it have a-ions below bin 0, so their tiles take the checked peak
    writes (the ladder marks them); scores and rankings still follow the
    oracle."""
    from hfx_amd import pipeline, synth

    prot = synth.make_proteins(800, seed=61)
    static = {"C": 57.02146, "G": -37.0}
    mass = orc.mass_table(static)
    assert 19.0 < float(mass[ord("G") - 65]) < 21.0
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=2)
    ions = orc.prepare_ions(prot.seq)
    first_g = sum(1 for i in range(len(tc)) if ions[tc[i]] == ord("G") - 65)
    assert first_g > 100
    sp = synth.make_spectra(prot, r, tc, tn, 160, seed=62, charges=(2, 3, 4),
                            charge_p=(0.5, 0.3, 0.2), static_mods=static)
    ctx = pipeline.Ctx(0)
    db = pipeline.DeviceDB(pipeline.HostDB(prot.seq, r, tc, tn, mass))
    dsp = pipeline.DeviceSpectra(ctx, sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE)
    idx, _ = orc.sort_idx(r)
    ref = orc.search_plain(ions, r, tc, tn, mass, idx, dsp.x(), dsp.lengths(), dsp.zmax, dsp.mass,
                           3.0, 10, want_all=True)
    absx = orc.search_plain(ions, r, tc, tn, mass, idx, np.abs(dsp.x()), dsp.lengths(), dsp.zmax,
                            dsp.mass, 3.0, 10, want_all=True)
    ref["all_scale"] = orc.fx_to_double(absx["all_fx"])
    res = pipeline.PlainSearch(ctx, db, dsp, eps=3.0, topn=10).run(keep_scores=True, dense=True)
    check_scores(res, ref, dsp, orc)
    check_topn(res, ref, orc)
    bs = pipeline.BatchSearch(ctx, db, dsp, eps=3.0, topn=10)
    bs.set_precision(bs.BF16X2_DENSE)
    bs.run()
    res2 = bs.results(keep_scores=True)
    check_scores(res2, ref, dsp, orc)
    check_topn(res2, ref, orc)


def _expected_topn(ref, orc, s, N):
    """the oracle's top N of spectrum s for any N: every exact score of the
    window ordered by (exact desc, position asc)"""
    a, b = int(ref["off"][s]), int(ref["off"][s + 1])
    fx = ref["all_fx"][a:b]
    order = np.lexsort((np.arange(b - a), -fx))[:N]
    return order.tolist(), orc.fx_to_double(fx[order])


@pytest.mark.parametrize("N,margin", [(1, 0), (3, 5), (64, 0), (64, 64)])
def test_batch_dense_topn_depths(denv, orc, N, margin):
    """Selections of T = N + margin from 1 to 128 keys, and the band of
    candidates within 2B of the N-th score re-scored with them: the top N
    equal the oracle's at every depth."""
    pl, ref = denv["pl"], denv["ref"]
    bs = pl.BatchSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=N, margin=margin)
    bs.set_precision(bs.BF16X2_DENSE)
    bs.run()
    res = bs.results()
    for s in range(len(res.top_cnt)):
        pos, ex = _expected_topn(ref, orc, s, N)
        c = int(res.top_cnt[s])
        assert c == len(pos), s
        assert res.top_pos[s, :c].tolist() == pos, (s, N, margin)
        assert (res.top_exact[s, :c] == ex).all(), (s, N, margin)


def test_plain_sparse_topn_deep_margin(denv, orc):
    """The plain form's top-N at the largest selection (N 64, margin 64)."""
    pl, ref = denv["pl"], denv["ref"]
    srch = pl.PlainSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=64, margin=64)
    res = srch.run()
    for s in range(len(res.top_cnt)):
        pos, ex = _expected_topn(ref, orc, s, 64)
        c = int(res.top_cnt[s])
        assert c == len(pos) and res.top_pos[s, :c].tolist() == pos, s
        assert (res.top_exact[s, :c] == ex).all(), s
