"""hfx_xcorr_sparse_noalloc: the plain sparse XCorr with its tile table in
the library's static device array instead of an allocation.

On the GPU it must give exactly the scores of hfx_xcorr_sparse (same
kernels) -- and so the oracle's exact scores within the FP32 bound -- on two
calls in a row, the second with fewer spectra (the array keeps the first
call's table).  On the CPU: a call beyond HFX_NOALLOC_MAX_SPECTRA spectra is
refused before anything reaches the device."""
import ctypes as C

import numpy as np
import pytest

STRIDE = 2112
MAX_SPECTRA = 1 << 20


def test_noalloc_refuses_more_spectra_than_its_array():
    import hfx_amd

    L = hfx_amd.lib()
    fake = C.c_void_p(16)  # never dereferenced: the size check comes first
    db = hfx_amd.HfxDb(fake, fake, fake, fake, fake, fake, fake, 10, 10)
    sp = hfx_amd.HfxSpectra(fake, STRIDE, fake, fake, fake, MAX_SPECTRA + 1)
    rc = L.hfx_xcorr_sparse_noalloc(fake, fake, fake, fake, 100, C.byref(db), C.byref(sp), None)
    assert rc == 1  # HFX_ERR_ARG
    assert "at most 1048576" in hfx_amd.last_error()
    rc = L.hfx_xcorr_sparse_noalloc(None, fake, fake, fake, 100, C.byref(db), C.byref(sp), None)
    assert rc == 1 and "null argument" in hfx_amd.last_error()


@pytest.mark.gpu
def test_noalloc_matches_the_allocating_entry_point(orc):
    import hfx_amd
    from hfx_amd import check, pipeline, synth

    L = hfx_amd.lib()
    prot = synth.make_proteins(1500, seed=61)
    mass = orc.mass_table({"C": 57.02146})
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=1)
    sp = synth.make_spectra(prot, r, tc, tn, 120, seed=62, charges=(2, 3), charge_p=(0.5, 0.5),
                            static_mods={"C": 57.02146}, double_for_3plus=True)
    ctx = pipeline.Ctx(0)
    db = pipeline.DeviceDB(pipeline.HostDB(prot.seq, r, tc, tn, mass))
    dsp = pipeline.DeviceSpectra(ctx, sp.pcr, sp.charge, sp.mz, sp.inten, sp.off, STRIDE)
    idx, _ = orc.sort_idx(r)
    ref = orc.search_plain(orc.prepare_ions(prot.seq), r, tc, tn, mass, idx, dsp.x(),
                           dsp.lengths(), dsp.zmax, dsp.mass, 3.0, 5, want_all=True)
    srch = pipeline.PlainSearch(ctx, db, dsp, eps=3.0, topn=5)
    total = srch.window()
    off = pipeline.host(srch.t_off, np.uint64, dsp.num_spec + 1)
    X, lens = dsp.x(), dsp.lengths()
    for S in (dsp.num_spec, 37):  # the second call smaller than the first
        n = int(off[S])
        desc = hfx_amd.HfxSpectra(dsp.desc.d_x, dsp.desc.x_stride, dsp.desc.d_len,
                                  dsp.desc.d_zmax, dsp.desc.d_mass, S)
        t_a, t_b = pipeline.empty(4 * n), pipeline.empty(4 * n)
        check(L.hfx_xcorr_sparse(ctx.h, t_a.p, srch.t_off.p, srch.t_begin.p, srch.t_num.p, n,
                                 C.byref(db.desc), C.byref(desc), None), "hfx_xcorr_sparse")
        check(L.hfx_xcorr_sparse_noalloc(t_b.p, srch.t_off.p, srch.t_begin.p, srch.t_num.p, n,
                                         C.byref(db.desc), C.byref(desc), None),
              "hfx_xcorr_sparse_noalloc")
        pipeline.sync()
        a = pipeline.host(t_a, np.float32, n)
        b = pipeline.host(t_b, np.float32, n)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), S
        exact = orc.fx_to_double(ref["all_fx"][:n])
        x1 = np.array([np.abs(X[s, : lens[s]].astype(np.float64)).sum() for s in range(S)])
        bound = np.repeat(80 * 2.0**-24 * 50 * x1 * 1.001 + 1e-6, np.diff(off[: S + 1].astype(np.int64)))
        assert np.all(np.abs(b.astype(np.float64) - exact) <= bound), S
    assert total == int(off[-1])
