"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker; the product never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
FX_SHIFT = 36

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_ms2_mass.restype = C.c_float
        L.orc_ms2_mass.argtypes = [C.c_float, C.c_float]
        L.orc_dta_precursor.restype = C.c_float
        L.orc_dta_precursor.argtypes = [C.c_float, C.c_float]
        L.orc_choose.restype = C.c_uint32
        L.orc_choose.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_unrank_comb.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_score_fx.restype = C.c_int64
        L.orc_score_fx.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
        L.orc_score_f32.restype = C.c_float
        L.orc_score_f32.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
        L.orc_find_spec_cands_by_mass.restype = C.c_uint32
        L.orc_find_begin_end.restype = C.c_uint32
        L.orc_find_begin_end.argtypes = [C.c_void_p] * 5 + [C.c_uint32, C.c_void_p, C.c_uint32,
                                                             C.c_float, C.c_float]
        L.orc_find_modable_peptides.restype = C.c_uint32
        L.orc_calc_total_mod_cands.restype = C.c_uint32
        L.orc_prepare_gen_mod.restype = C.c_uint32
        L.orc_peaks_list.restype = C.c_uint32
        L.orc_check_fast_div.restype = C.c_uint64
        L.orc_check_fast_div.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def u32(a):
    return np.ascontiguousarray(a, np.uint32)


def f32(a):
    return np.ascontiguousarray(a, np.float32)


def mass_table(static: dict | None = None, mono: bool = True) -> np.ndarray:
    add = np.zeros(26, np.float32)
    for k, v in (static or {}).items():
        add[ord(k) - 65] = np.float32(v)
    out = np.zeros(26, np.float32)
    lib().orc_mass_table(_p(add), C.c_int(int(mono)), _p(out))
    return out


def ms2_mass(pcr: float, charge: float) -> np.float32:
    return np.float32(lib().orc_ms2_mass(pcr, charge))


def dta_precursor(mh: float, charge: float) -> np.float32:
    return np.float32(lib().orc_dta_precursor(mh, charge))


RULES = ["", "KR", "FWY", "R", "M", "W", "P", "E", "K", "R", "D", "FWYL", "ALIV", "ALIVKRWFY"]


def rule_mask(rule: int) -> np.ndarray:
    m = np.zeros(26, np.uint8)
    for ch in RULES[rule]:
        m[ord(ch) - 65] = 1
    return m


def digest(seq, seg, mass26, rule=1, missed=2, min_mass=400.0, max_mass=7200.0, p_rule=0,
           min_len=0, max_len=0):
    seq = np.ascontiguousarray(seq, np.uint8)
    seg = u32(seg)
    cap = int(len(seq) * (missed + 1) + 16)
    r = np.zeros(cap, np.float32)
    tc = np.zeros(cap, np.uint32)
    tn = np.zeros(cap, np.uint32)
    n = C.c_uint64()
    rc = lib().orc_digest(_p(seq), _p(seg), C.c_uint32(len(seg) - 1), _p(f32(mass26)),
                          _p(rule_mask(rule)), C.c_uint32(missed), C.c_float(min_mass),
                          C.c_float(max_mass), C.c_int(p_rule), C.c_uint32(min_len),
                          C.c_uint32(max_len), _p(r), _p(tc), _p(tn), C.c_uint64(cap), C.byref(n))
    assert rc == 0
    k = n.value
    return r[:k].copy(), tc[:k].copy(), tn[:k].copy()


def sort_idx(keys, init=0):
    keys = f32(keys)
    idx = np.zeros(len(keys), np.uint32)
    sk = np.zeros(len(keys), np.float32)
    lib().orc_sort_idx(_p(keys), C.c_uint32(len(keys)), C.c_uint32(init), _p(idx), _p(sk))
    return idx, sk


def prepare_ions(ions, ma: bytes = b""):
    a = np.array(ions, np.uint8, copy=True)
    mab = np.frombuffer(ma, np.uint8).copy() if ma else np.zeros(1, np.uint8)
    lib().orc_prepare_ions(_p(a), C.c_uint64(len(a)), _p(mab), C.c_uint32(len(ma)))
    return a


def mod_info(var: dict, max_ma: int):
    v = np.zeros(26, np.float32)
    for k, x in var.items():
        v[ord(k) - 65] = np.float32(x)
    ma = np.zeros(26, np.uint8)
    mam = np.zeros(26, np.float32)
    na = C.c_uint32()
    cap = 1 << 16
    cnt = np.zeros(cap * 26, np.uint32)
    sm = np.zeros(cap, np.uint32)
    dl = np.zeros(cap, np.float32)
    nm = C.c_uint32()
    rc = lib().orc_mod_info(_p(v), C.c_uint32(max_ma), _p(ma), _p(mam), C.byref(na), _p(cnt),
                            _p(sm), _p(dl), C.c_uint32(cap), C.byref(nm))
    assert rc == 0
    a, m = na.value, nm.value
    return (bytes(ma[:a]), mam[:a].copy(), cnt[: m * a].reshape(m, a).copy(), sm[:m].copy(),
            dl[:m].copy())


def find_spec_cands_by_mass(r, idx_sorted, masses, eps):
    S = len(masses)
    b = np.zeros(S, np.uint32)
    e = np.zeros(S, np.uint32)
    n = np.zeros(S, np.uint32)
    tot = lib().orc_find_spec_cands_by_mass(_p(b), _p(e), _p(n), _p(f32(r)), _p(u32(idx_sorted)),
                                            C.c_uint32(len(r)), _p(f32(masses)), C.c_uint32(S),
                                            C.c_float(eps))
    return b, e, n, int(tot)


def prepare_scoring(idx_sorted, begin, num):
    out = np.zeros(int(np.sum(num, dtype=np.uint64)), np.uint32)
    lib().orc_prepare_scoring(_p(out), _p(u32(idx_sorted)), _p(u32(begin)), _p(u32(num)),
                              C.c_uint32(len(begin)))
    return out


def find_begin_end(r, idx_sorted, mod_delta, mass, eps):
    M = len(mod_delta)
    b = np.zeros(M, np.uint32)
    e = np.zeros(M, np.uint32)
    sc = np.zeros(M, np.uint32)
    tot = lib().orc_find_begin_end(_p(b), _p(e), _p(sc), _p(f32(r)), _p(u32(idx_sorted)),
                                   C.c_uint32(len(r)), _p(f32(mod_delta)), C.c_uint32(M),
                                   C.c_float(mass), C.c_float(eps))
    return b, e, sc, int(tot)


def choose(n, k) -> int:
    return int(lib().orc_choose(n, k))


def unrank_comb(n, k, rank):
    out = np.zeros(max(k, 1), np.uint32)
    lib().orc_unrank_comb(_p(out), n, k, rank)
    return out[:k].copy()


def peaks_dense(ions_packed, tc, tn, r, mass26, zmax, length):
    w = np.zeros(length, np.uint8)
    lib().orc_peaks_dense(_p(w), C.c_uint32(length), _p(np.ascontiguousarray(ions_packed, np.uint8)),
                          C.c_uint32(tc), C.c_uint32(tn), C.c_float(r), _p(f32(mass26)),
                          C.c_uint32(zmax))
    return w


def mpeaks_dense(ions_packed, tc, tn, r, mass26, zmax, length, mod_row, mod_delta, ma, ma_mass,
                 unrank):
    w = np.zeros(length, np.uint8)
    mab = np.frombuffer(ma, np.uint8).copy()
    lib().orc_mpeaks_dense(_p(w), C.c_uint32(length),
                           _p(np.ascontiguousarray(ions_packed, np.uint8)), C.c_uint32(tc),
                           C.c_uint32(tn), C.c_float(r), _p(f32(mass26)), C.c_uint32(zmax),
                           _p(u32(mod_row)), C.c_float(mod_delta), _p(mab), _p(f32(ma_mass)),
                           C.c_uint32(len(ma)), _p(u32(unrank) if len(unrank) else np.zeros(1, np.uint32)))
    return w


def peaks_list(masses, r, zmax):
    masses = f32(masses)
    cap = 10 * 4 * max(len(masses), 1)
    bins = np.zeros(cap, np.int32)
    w = np.zeros(cap, np.uint8)
    n = lib().orc_peaks_list(_p(bins), _p(w), C.c_uint32(cap), _p(masses), C.c_uint32(len(masses)),
                             C.c_float(r), C.c_uint32(zmax))
    return bins[:n].copy(), w[:n].copy()


def score_fx(w, x) -> int:
    w = np.ascontiguousarray(w, np.uint8)
    return int(lib().orc_score_fx(_p(w), _p(f32(x)), C.c_uint32(len(w))))


def score_f32(w, x) -> np.float32:
    w = np.ascontiguousarray(w, np.uint8)
    return np.float32(lib().orc_score_f32(_p(w), _p(f32(x)), C.c_uint32(len(w))))


def fx_to_double(fx) -> np.ndarray:
    return np.ldexp(np.asarray(fx, np.float64), -FX_SHIFT)


def preprocess(pcr, charge, mz, inten, remove_precursor=True, mono=True, mean_mode=1, cap=4096):
    out = np.zeros(cap, np.float32)
    n = C.c_uint32()
    mz = f32(mz)
    inten = f32(inten)
    rc = lib().orc_preprocess(_p(out), C.c_uint32(cap), C.byref(n), C.c_float(pcr),
                              C.c_float(charge), _p(mz), _p(inten), C.c_uint32(len(mz)),
                              C.c_int(int(remove_precursor)), C.c_int(int(mono)),
                              C.c_int(mean_mode))
    return rc, out[: n.value].copy() if rc == 0 else None


def preprocess_batch(pcr, charge, mz, inten, off, stride, mean_mode=1, remove_precursor=True):
    S = len(pcr)
    X = np.zeros((S, stride), np.float32)
    lens = np.zeros(S, np.uint32)
    status = np.zeros(S, np.int32)
    for s in range(S):
        a, b = int(off[s]), int(off[s + 1])
        rc, x = preprocess(pcr[s], charge[s], mz[a:b], inten[a:b], remove_precursor, True,
                           mean_mode, cap=stride)
        status[s] = rc
        if rc == 0:
            X[s, : len(x)] = x
            lens[s] = len(x)
    return X, lens, status


def find_modable_peptides(spec_num_cand_scan, num_cand_total, ions_packed, tc, tn, idx_sorted,
                          begin, end, num_pep_scan, mod_ma_count, ma: bytes):
    S = len(spec_num_cand_scan) - 1
    na = len(ma)
    M = len(mod_ma_count)
    n = int(num_cand_total)
    pep = np.zeros(max(n, 1), np.uint32)
    mod = np.zeros(max(n, 1), np.uint32)
    cnt = np.zeros(max(n * na, 1), np.uint32)
    vs = np.zeros(max(S, 1), np.uint32)
    mab = np.frombuffer(ma, np.uint8).copy()
    nv = lib().orc_find_modable_peptides(
        _p(pep), _p(mod), _p(cnt), _p(vs), _p(u32(spec_num_cand_scan)), C.c_uint32(S),
        C.c_uint32(n), _p(np.ascontiguousarray(ions_packed, np.uint8)), _p(u32(tc)), _p(u32(tn)),
        _p(u32(idx_sorted)), _p(u32(begin)), _p(u32(end)), _p(u32(num_pep_scan)),
        _p(u32(np.asarray(mod_ma_count).reshape(-1))), C.c_uint32(M), _p(mab), C.c_uint32(na))
    return int(nv), pep[:nv].copy(), mod[:nv].copy(), cnt[: nv * na].copy(), vs[:S].copy()


def calc_total_mod_cands(mod_ma_count, pep_idx, pep_mod_idx, pep_ma_count, num_ma):
    n = len(pep_idx)
    num = np.zeros(max(n, 1), np.uint32)
    comb = np.zeros(max(n * num_ma, 1), np.uint32)
    tot = lib().orc_calc_total_mod_cands(
        _p(num), _p(comb), _p(u32(np.asarray(mod_ma_count).reshape(-1))), _p(u32(pep_idx)),
        _p(u32(pep_mod_idx)), _p(u32(pep_ma_count)), C.c_uint32(n),
        C.c_uint32(len(mod_ma_count)), C.c_uint32(num_ma))
    return int(tot), num[:n].copy(), comb[: n * num_ma].copy()


def prepare_gen_mod(mod_sum, pep_idx, pep_mod_idx, pep_num_mpep, num_mpep):
    m = int(num_mpep)
    outs = [np.zeros(max(m, 1), np.uint32) for _ in range(5)]
    lu = lib().orc_prepare_gen_mod(*[_p(o) for o in outs], _p(u32(mod_sum)), _p(u32(pep_idx)),
                                   _p(u32(pep_mod_idx)), _p(u32(pep_num_mpep)),
                                   C.c_uint32(len(pep_idx)), C.c_uint32(m))
    return (int(lu),) + tuple(o[:m].copy() for o in outs)


def add_ma_ranks(rank, scan_row):
    row = u32(scan_row)
    out = np.zeros(max(len(row), 1), np.uint32)
    lib().orc_add_ma_ranks(_p(out), C.c_uint32(rank), _p(row), C.c_uint32(len(row)))
    return out[: len(row)].copy()


def gen_mod_cands(len_unrank, mod_ma_count, mpep_ith, mpep_rank, mpep_scan, pep_mod_idx,
                  pep_ma_count, comb_scan, num_ma):
    out = np.zeros(max(int(len_unrank), 1), np.uint32)
    lib().orc_gen_mod_cands(_p(out), _p(u32(np.asarray(mod_ma_count).reshape(-1))),
                            _p(u32(mpep_ith)), _p(u32(mpep_rank)), _p(u32(mpep_scan)),
                            _p(u32(pep_mod_idx)), _p(u32(pep_ma_count)), _p(u32(comb_scan)),
                            C.c_uint32(len(mpep_ith)), C.c_uint32(num_ma))
    return out[: int(len_unrank)].copy()


def search_mods(ions_packed, r, tc, tn, mass26, idx_sorted, ma: bytes, ma_mass, mod_ma_count,
                mod_sum, mod_delta, X, lens, zmax, masses, eps, topn, nthreads=0,
                cap_mpep=1 << 24, cap_unrank=1 << 26):
    S, stride = X.shape
    X = np.ascontiguousarray(X, np.float32)
    na = len(ma)
    snm = np.zeros(S, np.uint32)
    top = np.zeros(S * topn, np.uint32)
    fx = np.zeros(S * topn, np.int64)
    cnt = np.zeros(S, np.uint32)
    mp_pep = np.zeros(cap_mpep, np.uint32)
    mp_mod = np.zeros(cap_mpep, np.uint32)
    mp_uoff = np.zeros(cap_mpep, np.uint32)
    unrank = np.zeros(cap_unrank, np.uint32)
    nmp = C.c_uint64()
    lu = C.c_uint64()
    mab = np.frombuffer(ma, np.uint8).copy()
    rc = lib().orc_search_mods(
        _p(np.ascontiguousarray(ions_packed, np.uint8)), _p(f32(r)), _p(u32(tc)), _p(u32(tn)),
        _p(f32(mass26)), _p(u32(idx_sorted)), C.c_uint32(len(r)), _p(mab), _p(f32(ma_mass)),
        C.c_uint32(na), _p(u32(np.asarray(mod_ma_count).reshape(-1))), _p(u32(mod_sum)),
        _p(f32(mod_delta)), C.c_uint32(len(mod_delta)), _p(X), C.c_uint32(stride), _p(u32(lens)),
        _p(u32(zmax)), _p(f32(masses)), C.c_uint32(S), C.c_float(eps), C.c_uint32(topn),
        _p(snm), _p(top), _p(fx), _p(cnt), _p(mp_pep), _p(mp_mod), _p(mp_uoff), _p(unrank),
        C.c_uint64(cap_mpep), C.c_uint64(cap_unrank), C.byref(nmp), C.byref(lu),
        C.c_int(nthreads))
    assert rc == 0
    m, l = nmp.value, lu.value
    return dict(spec_num_mpep=snm, top_pos=top.reshape(S, topn), top_fx=fx.reshape(S, topn),
                top_cnt=cnt, mpep_pep=mp_pep[:m].copy(), mpep_mod=mp_mod[:m].copy(),
                mpep_unrank_off=mp_uoff[:m].copy(), unrank=unrank[:l].copy())


def search_plain(ions_packed, r, tc, tn, mass26, idx_sorted, X, lens, zmax, masses, eps, topn,
                 want_all=False, nthreads=0):
    S, stride = X.shape
    X = np.ascontiguousarray(X, np.float32)
    b = np.zeros(S, np.uint32)
    n = np.zeros(S, np.uint32)
    pos = np.zeros(S * topn, np.uint32)
    fx = np.zeros(S * topn, np.int64)
    cnt = np.zeros(S, np.uint32)
    all32 = allfx = off = None
    if want_all:
        bb, ee, nn, tot = find_spec_cands_by_mass(r, idx_sorted, masses, eps)
        off = np.zeros(S + 1, np.uint64)
        np.cumsum(nn, out=off[1:])
        all32 = np.zeros(max(int(off[-1]), 1), np.float32)
        allfx = np.zeros(max(int(off[-1]), 1), np.int64)
    rc = lib().orc_search_plain(
        _p(np.ascontiguousarray(ions_packed, np.uint8)), _p(f32(r)), _p(u32(tc)), _p(u32(tn)),
        _p(f32(mass26)), _p(u32(idx_sorted)), C.c_uint32(len(r)), _p(X), C.c_uint32(stride),
        _p(u32(lens)), _p(u32(zmax)), _p(f32(masses)), C.c_uint32(S), C.c_float(eps),
        C.c_uint32(topn), _p(b), _p(n), _p(pos), _p(fx), _p(cnt),
        _p(all32) if want_all else None, _p(allfx) if want_all else None,
        _p(off) if want_all else None, C.c_int(nthreads))
    assert rc == 0
    res = dict(begin=b, num=n, top_pos=pos.reshape(S, topn), top_fx=fx.reshape(S, topn),
               top_cnt=cnt)
    if want_all:
        res.update(off=off, all_f32=all32, all_fx=allfx)
    return res


def check_fast_div(d: float, lo: float, hi: float, corrected: bool = True) -> int:
    """Mismatches of the FMA-corrected reciprocal division over [lo, hi)
    (corrected=False: of the uncorrected x * RN(1/d))."""
    return int(lib().orc_check_fast_div(C.c_float(d), C.c_float(lo), C.c_float(hi),
                                        int(corrected)))


def evalue_fit(scores, bin_width: int = 350, perc: float = 3.0):
    """(alpha, beta, histogram points, scores kept) of one spectrum's scores
    (Statistics.hs:37-100)."""
    s = f32(scores)
    a, b, nu = C.c_double(), C.c_double(), C.c_uint64()
    L = lib()
    L.orc_evalue_fit.restype = C.c_uint32
    m = L.orc_evalue_fit(_p(s), C.c_uint64(len(s)), C.c_uint32(bin_width), C.c_float(perc),
                         C.byref(a), C.byref(b), C.byref(nu))
    return a.value, b.value, int(m), int(nu.value)


def evalue(alpha: float, beta: float, score: float) -> float:
    L = lib()
    L.orc_evalue.restype = C.c_double
    L.orc_evalue.argtypes = [C.c_double] * 3
    return float(L.orc_evalue(alpha, beta, score))


def extract_peaks(x) -> np.ndarray:
    """uint8[len/3 + 1]: 1 where extractPeaks leaves a value above zero."""
    x = f32(x)
    out = np.zeros(len(x) // 3 + 1, np.uint8)
    lib().orc_extract_peaks(_p(x), C.c_uint32(len(x)), _p(out))
    return out


def match_ions(ions_packed, tc, tn, res, mass26, zmax, on, mono=True, nmod=0, flags=False):
    """(b matched, y matched, total) of one peptide against one spectrum's
    peak flags (IonSeries.hs:136-147, Match.hs:33-40); flags=True also
    returns the b and y flag lists."""
    ions = np.ascontiguousarray(ions_packed, np.uint8)
    on = np.ascontiguousarray(on, np.uint8)
    nb, ny, tot = C.c_uint32(), C.c_uint32(), C.c_uint32()
    ntake = tn - tc + 1 if tn < len(ions) else tn - tc
    nI = max(ntake * max(int(zmax), 1), 1)
    bf = np.zeros(nI, np.uint8)
    yf = np.zeros(nI, np.uint8)
    lib().orc_match_ions(_p(ions), C.c_uint64(len(ions)), C.c_uint32(tc), C.c_uint32(tn),
                         C.c_float(res), _p(f32(mass26)), C.c_uint32(zmax), C.c_int(int(mono)),
                         _p(on), C.c_uint32(len(on)), C.c_uint32(nmod), C.byref(nb), C.byref(ny),
                         C.byref(tot), _p(bf), _p(yf))
    if flags:
        return nb.value, ny.value, tot.value, bf[:nI - 1].copy(), yf[:nI - 1].copy()
    return nb.value, ny.value, tot.value
