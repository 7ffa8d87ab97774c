"""GPU parity of the integer precision form (int8 MFMA over three base-256
digits of each spectrum) against the CPU oracle, on the workload of
test_gpu_dense.py (fragment charges 1-4, two spectra above charge 4, an
empty window):

  - every score within its stated bound,
    |score - exact| <= max|x| / 8355711 * 50 * len + 2^-24 * 50 * sum|x|
    (the rounding of x * 2^k to integers, then one FP32 rounding), and
    within 1e-4 * max(|exact|, sum_b w_b |x_b| / 4);
  - top-N positions and exact scores equal to the oracle's, batched and
    unbatched, at every chunk size and top-N depth, through graph replays;
  - scores independent of the spectrum order (block membership);
  - the block cut: runs of one to sixteen spectra, minimal cost.
"""
import numpy as np
import pytest

from test_gpu_dense import _expected_topn, check_topn, denv  # noqa: F401

pytestmark = pytest.mark.gpu

STRIDE = 2112
XMAX = 8355711.0


def _bound8(dsp, res):
    X, lens = dsp.x(), dsp.lengths()
    xs = [np.abs(X[s, : lens[s]].astype(np.float64)) for s in range(dsp.num_spec)]
    per = np.array([(x.max() / XMAX * 50 * len(x) + 2.0**-24 * 50 * x.sum()) * 1.001 + 1e-6
                    if len(x) else 1e-6 for x in xs])
    return np.repeat(per, np.diff(res.off.astype(np.int64)))


def check_scores8(res, ref, dsp, orc):
    assert (res.begin == ref["begin"]).all() and (res.num == ref["num"]).all()
    exact = orc.fx_to_double(ref["all_fx"][: len(res.scores)])
    scale = ref["all_scale"][: len(res.scores)]
    err = np.abs(res.scores.astype(np.float64) - exact)
    bound = _bound8(dsp, res)
    floor = np.maximum(np.abs(exact), 0.25 * scale)
    print(f"i8x3: {len(err)} scores, max |err| = {err.max():.3g}, "
          f"max err/bound = {float((err / bound).max()):.3g}, "
          f"max err/sum(w|x|) = {float((err / np.maximum(scale, 1e-30)).max()):.3g}, "
          f"max err/max(|exact|, scale/4) = {float((err / np.maximum(floor, 1e-30)).max()):.3g}")
    assert np.all(err <= bound)
    assert np.all(err <= 1e-4 * floor + 1e-6)


def test_i8_scores_within_bound(denv, orc):  # noqa: F811
    pl = denv["pl"]
    srch = pl.PlainSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=10)
    res = srch.run(keep_scores=True, dense="i8x3")
    check_scores8(res, denv["ref"], denv["dsp"], orc)
    check_topn(res, denv["ref"], orc)
    assert len(res.short) == 0
    assert res.num[3] == 0
    for s in (7, 19):
        a, b = int(res.off[s]), int(res.off[s + 1])
        assert b > a and np.all(res.scores[a:b] == 0)


def test_i8_blocking_does_not_change_scores(denv, orc):  # noqa: F811
    pl, dsp = denv["pl"], denv["dsp"]
    srch = pl.PlainSearch(denv["ctx"], denv["db"], dsp, eps=3.0, topn=10)
    srch.window()
    srch.score_dense(None, form="i8x3")
    a = pl.host(srch.t_scores, np.float32, srch.total).copy()
    order = np.lexsort((dsp.mass, dsp.zmax)).astype(np.uint32)
    srch.score_dense(order, form="i8x3")
    b = pl.host(srch.t_scores, np.float32, srch.total).copy()
    srch.score_dense(order[::-1].copy(), form="i8x3")
    c = pl.host(srch.t_scores, np.float32, srch.total).copy()
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_batch_i8_matches_oracle(denv, orc):  # noqa: F811
    pl = denv["pl"]
    bs = pl.BatchSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=10)
    bs.set_precision(bs.I8X3_DENSE)
    assert bs.info()["precision"] == 2
    bs.run()
    res = bs.results(keep_scores=True)
    check_scores8(res, denv["ref"], denv["dsp"], orc)
    check_topn(res, denv["ref"], orc)
    assert len(res.short) == 0


def test_batch_i8_blocks_cover_the_order(denv):  # noqa: F811
    """Blocks of one to sixteen spectra of the run's order, cut inside
    segments of 32 order positions at minimal total cost (charge levels x
    rows of the hull of the windows)."""
    pl, dsp = denv["pl"], denv["dsp"]
    bs = pl.BatchSearch(denv["ctx"], denv["db"], dsp, eps=3.0, topn=10)
    bs.set_precision(bs.I8X3_DENSE)
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
    assert sizes.min() >= 1 and sizes.max() <= 16
    assert all((a // 32) == ((c - 1) // 32) for a, c in zip(st[:-1], st[1:]))
    b, n = res.begin.astype(np.int64), res.num.astype(np.int64)

    def cost(lo, hi):
        blk = order[lo:hi]
        blk = blk[(z[blk] > 0) & (n[blk] > 0)]
        if len(blk) == 0:
            return 0
        return len(set(z[blk].tolist())) * int((b[blk] + n[blk]).max() - b[blk].min())

    plan = sum(cost(int(a), int(c)) for a, c in zip(st[:-1], st[1:]))
    total = 0
    for p0 in range(0, S, 32):
        p1 = min(S, p0 + 32)
        best = [0] + [None] * (p1 - p0)
        for j in range(1, p1 - p0 + 1):
            best[j] = min(best[j - k] + cost(p0 + j - k, p0 + j) for k in range(1, min(16, j) + 1))
        total += best[p1 - p0]
    assert plan == total


@pytest.mark.parametrize("chunk", [64, 192, 512, 4096])
def test_batch_i8_chunks_agree(denv, orc, chunk):  # noqa: F811
    pl = denv["pl"]
    bs = pl.BatchSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=10)
    bs.set_precision(bs.I8X3_DENSE)
    bs.set_config(0, chunk, 256)
    assert bs.info()["chunk"] == chunk
    bs.run()
    check_topn(bs.results(), denv["ref"], orc)


def test_batch_i8_graph_and_autotune(denv, orc):  # noqa: F811
    pl = denv["pl"]
    bs = pl.BatchSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=10)
    bs.set_precision(bs.I8X3_DENSE)
    G, Cn, T = bs.autotune()
    assert G == 0 and Cn in (256, 512, 1024, 2048, 4096) and T == 1024
    bs.run()
    first = bs.results(keep_scores=True)
    check_topn(first, denv["ref"], orc)
    bs.capture()
    for _ in range(3):
        bs.replay()
        pl.sync()
        again = bs.results(keep_scores=True)
        assert np.array_equal(again.scores, first.scores)
        check_topn(again, denv["ref"], orc)


@pytest.mark.parametrize("N,margin", [(1, 0), (3, 5), (64, 0), (64, 64)])
def test_batch_i8_topn_depths(denv, orc, N, margin):  # noqa: F811
    pl, ref = denv["pl"], denv["ref"]
    bs = pl.BatchSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=N, margin=margin)
    bs.set_precision(bs.I8X3_DENSE)
    bs.run()
    res = bs.results()
    for s in range(len(res.top_cnt)):
        pos, ex = _expected_topn(ref, orc, s, N)
        c = int(res.top_cnt[s])
        assert c == len(pos), s
        assert res.top_pos[s, :c].tolist() == pos, (s, N, margin)
        assert (res.top_exact[s, :c] == ex).all(), (s, N, margin)


def test_i8_light_first_residue(orc):
    """Glycine made 20 Da: peptides starting with it put a-ions below bin 0
    (the checked peak writes); scores and rankings still follow the oracle."""
    from hfx_amd import pipeline, synth

    prot = synth.make_proteins(800, seed=61)
    static = {"C": 57.02146, "G": -37.0}
    mass = orc.mass_table(static)
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=2)
    ions = orc.prepare_ions(prot.seq)
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
    res = pipeline.PlainSearch(ctx, db, dsp, eps=3.0, topn=10).run(keep_scores=True, dense="i8x3")
    check_scores8(res, ref, dsp, orc)
    check_topn(res, ref, orc)
    bs = pipeline.BatchSearch(ctx, db, dsp, eps=3.0, topn=10)
    bs.set_precision(bs.I8X3_DENSE)
    bs.run()
    res2 = bs.results(keep_scores=True)
    check_scores8(res2, ref, dsp, orc)
    check_topn(res2, ref, orc)


def test_tuned_shapes_persist_in_a_cache_file(denv, orc, tmp_path, monkeypatch):  # noqa: F811
    """HFX_TUNE_CACHE: a batch takes each dense form's chunk from the file's
    line for its shape (S, pairs and charge levels bucketed as the library
    does), malformed lines and shapes no kernel supports are ignored (the
    later lines would otherwise replace the valid ones), and a tuning
    rewrites the file with the chunk it chose; rankings are unchanged."""
    pl = denv["pl"]
    path = tmp_path / "tune.txt"
    bs = pl.BatchSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=10)
    inf = bs.info()
    lg = lambda v: max(int(v), 1).bit_length() - 1  # noqa: E731
    key = f"{lg(inf['spectra'])} {lg(inf['scores'])} {inf['levels']}"
    path.write_text(f"int8 {key} 512\nbf16 {key} 1024\nint8 not a line\n"
                    f"int8 {key} 100\nint8 {key} 16384\nbf16 {key} 24\n"
                    f"sparse {key} 3 64 300\nsparse {key} 8 0 256\n")
    monkeypatch.setenv("HFX_TUNE_CACHE", str(path))
    bs.set_precision(bs.I8X3_DENSE)
    assert bs.info()["chunk"] == 512
    other = pl.BatchSearch(denv["ctx"], denv["db"], denv["dsp"], eps=3.0, topn=10)
    oinf = other.info()
    assert oinf["G"] in (4, 8) and 1 <= oinf["C"] <= 4096 and oinf["threads"] in (256, 512)
    other.set_precision(other.BF16X2_DENSE)
    assert other.info()["chunk"] == 1024
    G, Cn, T = bs.autotune()
    lines = path.read_text().splitlines()
    assert f"int8 {key} {Cn}" in lines and f"bf16 {key} 1024" in lines, lines
    bs.run()
    check_topn(bs.results(), denv["ref"], orc)
