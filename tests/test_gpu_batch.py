"""GPU parity of the batched path (peptide ion index + persistent
multi-spectrum XCorr + graph replay + autotuner) against the CPU oracle:
identical top-N positions and exact scores, FP32 scores within 1e-4 of the
exact score relative to max(|exact|, 100), every configuration and every
graph replay giving the same results."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STRIDE = 2112


@pytest.fixture(scope="module")
def benv(orc):
    from hfx_amd import pipeline, synth

    prot = synth.make_proteins(3000, seed=41)
    mass = orc.mass_table({"C": 57.02146})
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=2)
    sp = synth.make_spectra(prot, r, tc, tn, 300, seed=42, charges=(2, 3), charge_p=(0.6, 0.4),
                            static_mods={"C": 57.02146}, double_for_3plus=True)
    # two spectra with max fragment charge > 4 (7+ and 8+ precursors of the
    # same peptides, same window): scored 0
    H = 1.007276
    for s, z in ((5, 7.0), (17, 8.0)):
        z0 = float(sp.charge[s])
        sp.pcr[s] = np.float32((float(sp.pcr[s]) * z0 - z0 * H + z * H) / z)
        sp.charge[s] = z
    ctx = pipeline.Ctx(0)
    db = pipeline.DeviceDB(pipeline.HostDB(prot.seq, r, tc, tn, mass))
    dsp = pipeline.DeviceSpectra(ctx, sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE)
    idx, _ = orc.sort_idx(r)
    ref = orc.search_plain(orc.prepare_ions(prot.seq), r, tc, tn, mass, idx, dsp.x(), dsp.lengths(),
                           dsp.zmax, dsp.mass, 3.0, 10, want_all=True)
    return dict(ctx=ctx, db=db, dsp=dsp, ref=ref, pl=pipeline)


def check_against_oracle(res, ref, orc):
    assert (res.begin == ref["begin"]).all() and (res.num == ref["num"]).all()
    exact = orc.fx_to_double(ref["all_fx"][: len(res.scores)])
    err = np.abs(res.scores.astype(np.float64) - exact)
    assert np.all(err <= 1e-4 * np.maximum(np.abs(exact), 100.0)), float(err.max())
    assert (res.top_cnt == ref["top_cnt"]).all()
    for s in range(len(res.top_cnt)):
        c = int(res.top_cnt[s])
        assert res.top_pos[s, :c].tolist() == ref["top_pos"][s, :c].tolist(), s
        assert (res.top_exact[s, :c] == orc.fx_to_double(ref["top_fx"][s, :c])).all(), s


def test_batch_run_matches_oracle(benv, orc):
    bs = benv["pl"].BatchSearch(benv["ctx"], benv["db"], benv["dsp"], eps=3.0, topn=10)
    bs.run()
    info = bs.info()
    assert info["levels"] == 2 and info["peptides"] > 0 and info["entries"] > 0
    res = bs.results(keep_scores=True)
    check_against_oracle(res, benv["ref"], orc)
    for s in (5, 17):
        a, b = int(res.off[s]), int(res.off[s + 1])
        assert b > a and np.all(res.scores[a:b] == 0)


@pytest.mark.parametrize("cfg", [(4, 32, 256), (4, 256, 512), (8, 7, 256), (8, 128, 512)])
def test_batch_configs_agree(benv, orc, cfg):
    bs = benv["pl"].BatchSearch(benv["ctx"], benv["db"], benv["dsp"], eps=3.0, topn=10)
    bs.set_config(*cfg)
    bs.run()
    check_against_oracle(bs.results(keep_scores=True), benv["ref"], orc)


def test_graph_replays_and_autotune(benv, orc):
    pl = benv["pl"]
    bs = pl.BatchSearch(benv["ctx"], benv["db"], benv["dsp"], eps=3.0, topn=10)
    G, Cn, T = bs.autotune()
    assert G in (4, 8) and Cn in (32, 64, 128, 256) and T in (256, 512)
    bs.run()
    first = bs.results(keep_scores=True)
    check_against_oracle(first, benv["ref"], orc)
    bs.capture()
    for _ in range(3):
        bs.replay()
        pl.sync()  # the program's own synchronisation between replays
        again = bs.results(keep_scores=True)
        assert (again.scores.view(np.uint32) == first.scores.view(np.uint32)).all()
        assert (again.top_pos == first.top_pos).all()
        assert (again.top_exact == first.top_exact).all()
