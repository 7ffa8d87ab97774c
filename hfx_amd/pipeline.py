"""Host-side orchestration over the C ABI (mirrors Search.hs / Fragment.hs).

DeviceDB       Fragment.hs:241-258 withDeviceDB + Main.hs:83-86 prepareIons:
               uploads the digested database, packs residues, sorts residuals.
DeviceSpectra  the preprocessed spectra of one batch (Correlation.hs on GPU).
search_plain   Search.hs:129-277 searchWithoutMods for a batch: window ->
               sparse XCorr -> top-N (+ exact re-score, + fallback).

Device memory, streams and events come from the library's own runtime
helpers (hfx_malloc, hfx_memcpy_*, hfx_event_*), so that one HIP runtime is
loaded in the process; every computation is a call into libhfx_hip.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import HfxDb, HfxError, HfxModset, HfxSpectra, check, lib

# per-spectrum flags of the top-N (include/hfx_hip.h)
FLAG_EXACT = 1
FLAG_SHORT = 2

_gpu_checked = False


def require_gpu():
    global _gpu_checked
    if _gpu_checked:
        return
    n = C.c_int(0)
    rc = lib().hfx_device_count(C.byref(n))
    if rc != 0 or n.value < 1:
        raise RuntimeError("hfx_amd.pipeline needs a GPU (no HIP device visible)")
    _gpu_checked = True


class DevBuf:
    """A device allocation (hfx_malloc) freed with the object."""

    __slots__ = ("p", "nbytes")

    def __init__(self, nbytes: int):
        require_gpu()
        self.nbytes = max(int(nbytes), 1)
        p = C.c_void_p()
        check(lib().hfx_malloc(C.byref(p), self.nbytes), "hfx_malloc")
        self.p = p.value

    def __del__(self):
        try:
            if self.p:
                lib().hfx_free(self.p)
        except Exception:
            pass

    def upload(self, arr: np.ndarray, stream=None):
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise ValueError("upload larger than the buffer")
        check(lib().hfx_memcpy_h2d(self.p, a.ctypes.data, a.nbytes, stream), "hfx_memcpy_h2d")
        check(lib().hfx_stream_sync(stream), "hfx_stream_sync")

    def download(self, dtype, count=None, stream=None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else int(count)
        out = np.empty(n, dt)
        if n:
            check(lib().hfx_memcpy_d2h(out.ctypes.data, self.p, out.nbytes, stream),
                  "hfx_memcpy_d2h")
            check(lib().hfx_stream_sync(stream), "hfx_stream_sync")
        return out

    def zero(self, stream=None):
        check(lib().hfx_memset(self.p, 0, self.nbytes, stream), "hfx_memset")


class DevView(DevBuf):
    """A window [offset, offset + nbytes) of another device buffer, kept
    alive with it; never freed itself."""

    __slots__ = ("owner",)

    def __init__(self, owner: DevBuf, offset: int, nbytes: int):
        if offset < 0 or offset + nbytes > owner.nbytes:
            raise ValueError("view outside its buffer")
        self.owner = owner
        self.p = owner.p + int(offset)
        self.nbytes = max(int(nbytes), 1)

    def __del__(self):
        pass


def dev(arr: np.ndarray) -> DevBuf:
    """Upload a numpy array to a new device buffer (byte-exact)."""
    a = np.ascontiguousarray(arr)
    b = DevBuf(a.nbytes)
    if a.nbytes:
        b.upload(a)
    return b


def empty(nbytes: int) -> DevBuf:
    return DevBuf(nbytes)


def host(t: DevBuf, dtype, count=None) -> np.ndarray:
    return t.download(dtype, count)


def ptr(t: DevBuf) -> int:
    return t.p


def stream_handle():
    """The stream every pipeline call is issued on (the null stream)."""
    return None


def sync():
    check(lib().hfx_device_sync(), "hfx_device_sync")


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib().hfx_event_create(C.byref(p)), "hfx_event_create")
        self.p = p.value

    def record(self, stream=None):
        check(lib().hfx_event_record(self.p, stream), "hfx_event_record")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib().hfx_event_elapsed_ms(C.byref(ms), self.p, end.p), "hfx_event_elapsed_ms")
        return float(ms.value)

    def __del__(self):
        try:
            lib().hfx_event_destroy(self.p)
        except Exception:
            pass


class Ctx:
    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        check(lib().hfx_ctx_create(device, C.byref(self.h)), "hfx_ctx_create")

    def __del__(self):
        try:
            if self.h:
                lib().hfx_ctx_destroy(self.h)
        except Exception:
            pass


@dataclass
class HostDB:
    ions: np.ndarray   # uint8 residue letters (unpacked)
    r: np.ndarray      # float32
    tc: np.ndarray     # uint32
    tn: np.ndarray     # uint32
    mass: np.ndarray   # float32[26]


def mass_table(static: dict | None = None, mono: bool = True) -> np.ndarray:
    add = np.zeros(26, np.float32)
    for k, v in (static or {}).items():
        add[ord(k) - 65] = np.float32(v)
    out = np.zeros(26, np.float32)
    check(lib().hfx_mass_table(add.ctypes.data, int(mono), out.ctypes.data), "hfx_mass_table")
    return out


def digest(seq: np.ndarray, seg: np.ndarray, mass26: np.ndarray, rule: int = 1, missed: int = 2,
           min_mass: float = 400.0, max_mass: float = 7200.0, p_rule: int = 0,
           min_len: int = 0, max_len: int = 0):
    L = lib()
    mask = np.zeros(26, np.uint8)
    check(L.hfx_digestion_rule(rule, mask.ctypes.data), "hfx_digestion_rule")
    seq = np.ascontiguousarray(seq, np.uint8)
    seg = np.ascontiguousarray(seg, np.uint32)
    n = C.c_uint64()
    L.hfx_digest(seq.ctypes.data, seg.ctypes.data, len(seg) - 1, mass26.ctypes.data,
                 mask.ctypes.data, missed, min_mass, max_mass, p_rule, min_len, max_len,
                 None, None, None, 0, C.byref(n))
    cap = n.value
    r = np.zeros(cap, np.float32)
    tc = np.zeros(cap, np.uint32)
    tn = np.zeros(cap, np.uint32)
    check(L.hfx_digest(seq.ctypes.data, seg.ctypes.data, len(seg) - 1, mass26.ctypes.data,
                       mask.ctypes.data, missed, min_mass, max_mass, p_rule, min_len, max_len,
                       r.ctypes.data, tc.ctypes.data, tn.ctypes.data, cap, C.byref(n)),
          "hfx_digest")
    return r, tc, tn


class DeviceDB:
    """The device-resident database (hfx_db)."""

    def __init__(self, hdb: HostDB, ma: bytes = b""):
        L = lib()
        self.host = hdb
        self.num_pep = len(hdb.r)
        self.num_ions = len(hdb.ions)
        self.t_ions = dev(hdb.ions.astype(np.uint8))
        self.t_r = dev(hdb.r.astype(np.float32))
        self.t_tc = dev(hdb.tc.astype(np.uint32))
        self.t_tn = dev(hdb.tn.astype(np.uint32))
        self.t_mass = dev(hdb.mass.astype(np.float32))
        self.t_rsorted = dev(hdb.r.astype(np.float32))
        self.t_idx = empty(4 * self.num_pep)
        self.t_ma = dev(np.frombuffer(ma, np.uint8)) if ma else None
        # Main.hs:83-86 prepareIons, Fragment.hs:257 sort_idx (null stream)
        L.prepare_ions(ptr(self.t_ions), self.num_ions,
                       ptr(self.t_ma) if self.t_ma is not None else None, len(ma))
        L.sort_idx_f(ptr(self.t_rsorted), ptr(self.t_idx), self.num_pep, 0)
        sync()
        self.desc = HfxDb(ptr(self.t_ions), ptr(self.t_r), ptr(self.t_tc), ptr(self.t_tn),
                          ptr(self.t_mass), ptr(self.t_idx), ptr(self.t_rsorted),
                          self.num_pep, self.num_ions)

    def idx_sorted(self) -> np.ndarray:
        return host(self.t_idx, np.uint32, self.num_pep)

    def packed_ions(self) -> np.ndarray:
        return host(self.t_ions, np.uint8, self.num_ions)


class DeviceSpectra:
    """Peaks uploaded and preprocessed on the GPU (hfx_spectra)."""

    def __init__(self, ctx: Ctx, pcr, charge, mz, inten, off, x_stride: int = 2112,
                 remove_precursor: bool = True, mono: bool = True, mean_mode: int = 1):
        L = lib()
        self.num_spec = n = len(pcr)
        self.x_stride = x_stride
        # the kernel reads peaks off[s]..off[s+1] of every spectrum: check the
        # layout here, where the host arrays are at hand
        off_h = np.asarray(off)
        if len(charge) != n or len(off_h) != n + 1:
            raise ValueError(f"{n} spectra need {n} charges and {n + 1} offsets, "
                             f"got {len(charge)} and {len(off_h)}")
        if n and (off_h[0] != 0 or np.any(np.diff(off_h.astype(np.int64)) < 0)
                  or int(off_h[-1]) > min(len(mz), len(inten))):
            raise ValueError("peak offsets must start at 0, not decrease and stay within "
                             "the m/z and intensity arrays")
        self.t_pcr = dev(np.asarray(pcr, np.float32))
        self.t_charge = dev(np.asarray(charge, np.float32))
        self.t_mz = dev(np.asarray(mz, np.float32))
        self.t_int = dev(np.asarray(inten, np.float32))
        self.t_off = dev(np.asarray(off, np.uint32))
        self.t_x = empty(4 * n * x_stride)
        self.t_len = empty(4 * n)
        self.t_status = empty(4 * n)
        check(L.hfx_preprocess(ctx.h, ptr(self.t_x), x_stride, ptr(self.t_len),
                               ptr(self.t_status), ptr(self.t_mz), ptr(self.t_int),
                               ptr(self.t_off), ptr(self.t_pcr), ptr(self.t_charge), n,
                               int(remove_precursor), int(mono), int(mean_mode),
                               stream_handle()), "hfx_preprocess")
        # Search.hs:182 max fragment charge max(1, z-1); Data.hs:58-59 window centre
        ch = np.asarray(charge, np.float32)
        zmax = np.maximum(1, np.rint(np.maximum(np.float32(1), ch - np.float32(1)))).astype(np.uint32)
        masses = np.array([L.hfx_ms2_mass(float(p), float(c)) for p, c in zip(pcr, charge)],
                          np.float32)
        self.zmax = zmax
        self.mass = masses
        self.t_zmax = dev(zmax)
        self.t_mass = dev(masses)
        self.desc = HfxSpectra(ptr(self.t_x), x_stride, ptr(self.t_len), ptr(self.t_zmax),
                               ptr(self.t_mass), n)
        sync()  # later work may run on other (non-blocking) streams

    def lengths(self) -> np.ndarray:
        return host(self.t_len, np.uint32, self.num_spec)

    def status(self) -> np.ndarray:
        return host(self.t_status, np.uint32, self.num_spec)

    def x(self) -> np.ndarray:
        return host(self.t_x, np.float32).reshape(self.num_spec, self.x_stride)

    def view(self, s0: int, s1: int) -> "SpectraView":
        """Spectra [s0, s1) as a search input, sharing this object's device
        arrays (no copy)."""
        return SpectraView(self, s0, s1)


class SpectraView:
    """Spectra [s0, s1) of a DeviceSpectra: the same preprocessed arrays
    from spectrum s0 on, with the fields a search reads."""

    def __init__(self, sp: DeviceSpectra, s0: int, s1: int):
        if not 0 <= s0 <= s1 <= sp.num_spec:
            raise ValueError(f"spectra {s0}..{s1} of {sp.num_spec}")
        n, st = s1 - s0, sp.x_stride
        self.base, self.s0 = sp, s0
        self.num_spec, self.x_stride = n, st
        self.zmax, self.mass = sp.zmax[s0:s1], sp.mass[s0:s1]
        self.t_x = DevView(sp.t_x, 4 * s0 * st, 4 * n * st)
        self.t_len = DevView(sp.t_len, 4 * s0, 4 * n)
        self.t_status = DevView(sp.t_status, 4 * s0, 4 * n)
        self.t_zmax = DevView(sp.t_zmax, 4 * s0, 4 * n)
        self.t_mass = DevView(sp.t_mass, 4 * s0, 4 * n)
        self.desc = HfxSpectra(ptr(self.t_x), st, ptr(self.t_len), ptr(self.t_zmax),
                               ptr(self.t_mass), n)

    lengths = DeviceSpectra.lengths
    status = DeviceSpectra.status
    x = DeviceSpectra.x


@dataclass
class PlainResult:
    begin: np.ndarray
    num: np.ndarray
    off: np.ndarray
    top_pos: np.ndarray     # [S, N] position within the window
    top_exact: np.ndarray   # [S, N] exact score (float64)
    top_f32: np.ndarray     # [S, N]
    top_cnt: np.ndarray
    flagged: np.ndarray     # spectra completed by the exact fallback
    scores: np.ndarray | None = None
    # spectra whose FP32 scores broke the stated error bound: their top N
    # holds only the candidates the exact route could vouch for
    short: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))


class PlainSearch:
    """searchWithoutMods for one batch; buffers reused across calls."""

    def __init__(self, ctx: Ctx, db: DeviceDB, sp: DeviceSpectra, eps: float, topn: int,
                 margin: int = 0):
        self.ctx, self.db, self.sp = ctx, db, sp
        self.eps, self.topn, self.margin = float(eps), int(topn), int(margin)
        S = sp.num_spec
        self.t_begin = empty(4 * S)
        self.t_num = empty(4 * S)
        self.t_off = empty(8 * (S + 1))
        self.t_pos = empty(4 * S * topn)
        self.t_exact = empty(8 * S * topn)
        self.t_f32 = empty(4 * S * topn)
        self.t_cnt = empty(4 * S)
        self.t_flag = empty(4 * S)
        self.t_scores = None
        self.total = None

    def window(self):
        L = lib()
        S = self.sp.num_spec
        st = stream_handle()
        check(L.hfx_window(self.ctx.h, ptr(self.t_begin), ptr(self.t_num), C.byref(self.db.desc),
                           ptr(self.sp.t_mass), S, self.eps, st), "hfx_window")
        check(L.hfx_scan_offsets(self.ctx.h, ptr(self.t_off), ptr(self.t_num), S, st),
              "hfx_scan_offsets")
        off = host(self.t_off, np.uint64, S + 1)
        self.total = int(off[-1])
        if self.t_scores is None or self.t_scores.nbytes < 4 * max(self.total, 1):
            self.t_scores = empty(4 * max(self.total, 1))
        return self.total

    def score(self):
        L = lib()
        st = stream_handle()
        check(L.hfx_xcorr_sparse(self.ctx.h, ptr(self.t_scores), ptr(self.t_off),
                                 ptr(self.t_begin), ptr(self.t_num), self.total,
                                 C.byref(self.db.desc), C.byref(self.sp.desc), st),
              "hfx_xcorr_sparse")

    def select(self):
        L = lib()
        st = stream_handle()
        check(L.hfx_topn(self.ctx.h, ptr(self.t_pos), ptr(self.t_exact), ptr(self.t_f32),
                         ptr(self.t_cnt), ptr(self.t_flag), self.topn, self.margin,
                         ptr(self.t_scores), ptr(self.t_off), self.total, ptr(self.t_begin),
                         ptr(self.t_num), C.byref(self.db.desc), C.byref(self.sp.desc), st),
              "hfx_topn")

    # |score - exact| <= REL * 50 * sum|x| for hfx_xcorr_dense_bf16x2 scores
    REL_DENSE = float(np.float32(2.0**-16 + 210 * 2.0**-24))
    REL_F32 = float(np.float32(80 * 2.0**-24))
    # |score - exact| <= 2^-24 * 50 * sum|x| + max|x| / 8355711 * 50 * len
    # for hfx_xcorr_dense_i8x3 scores
    REL_I8 = float(2.0**-24)
    RELQ_I8 = float(np.float32(1.0 / 8355711.0))

    def score_dense(self, order: np.ndarray | None = None, form: str = "bf16x2"):
        """Scores by hfx_xcorr_dense_bf16x2 (blocks of eight spectra in
        `order`, index order if None) or, form "i8x3", hfx_xcorr_dense_i8x3
        (blocks of sixteen); same layout as score()."""
        t_order = dev(np.asarray(order, np.uint32)) if order is not None else None
        fn = {"bf16x2": "hfx_xcorr_dense_bf16x2", "i8x3": "hfx_xcorr_dense_i8x3"}[form]
        if self.total:
            check(getattr(lib(), fn)(self.ctx.h, ptr(self.t_scores), ptr(self.t_off),
                                     ptr(self.t_begin), ptr(self.t_num),
                                     ptr(t_order) if t_order is not None else None,
                                     C.byref(self.db.desc), C.byref(self.sp.desc),
                                     stream_handle()), fn)
        sync()

    def select_rel(self, rel0: float, rel1: float = 0.0):
        """hfx_topn_rel: top-N for scores within the stated error bound."""
        check(lib().hfx_topn_rel(self.ctx.h, ptr(self.t_pos), ptr(self.t_exact), ptr(self.t_f32),
                                 ptr(self.t_cnt), ptr(self.t_flag), self.topn, self.margin,
                                 float(rel0), float(rel1), ptr(self.t_scores), ptr(self.t_off),
                                 self.total, ptr(self.t_begin), ptr(self.t_num),
                                 C.byref(self.db.desc), C.byref(self.sp.desc), stream_handle()),
              "hfx_topn_rel")

    def select_rel_q(self, rel0: float, rel1: float, relq: float):
        """hfx_topn_rel_q: the bound with the max|x| * 50 * len term."""
        check(lib().hfx_topn_rel_q(self.ctx.h, ptr(self.t_pos), ptr(self.t_exact),
                                   ptr(self.t_f32), ptr(self.t_cnt), ptr(self.t_flag), self.topn,
                                   self.margin, float(rel0), float(rel1), float(relq),
                                   ptr(self.t_scores), ptr(self.t_off), self.total,
                                   ptr(self.t_begin), ptr(self.t_num), C.byref(self.db.desc),
                                   C.byref(self.sp.desc), stream_handle()), "hfx_topn_rel_q")

    def evalue_fit(self, mods: "ModSearch | None" = None, bin_width: int = 350,
                   perc: float = 3.0):
        """hfx_evalue_fit over this search's scores (after run()), pooled with
        the modified-peptide scores of `mods` (after its run()), as
        Search.hs:118-121 pools allMatches.  Keeps the device fit for
        evalues(); returns (fit [S, 2] alpha/beta, points [S], kept [S])."""
        S = self.sp.num_spec
        self.t_fit = empty(16 * S)
        t_np, t_nu = empty(4 * S), empty(8 * S)
        sb = ob = None
        if mods is not None:
            sb, ob = ptr(mods._scored[0]), ptr(mods._scored[1])
        check(lib().hfx_evalue_fit(self.ctx.h, ptr(self.t_fit), ptr(t_np), ptr(t_nu),
                                   ptr(self.t_scores), ptr(self.t_off), sb, ob, S,
                                   int(bin_width), float(perc), stream_handle()),
              "hfx_evalue_fit")
        sync()
        return (host(self.t_fit, np.float64, 2 * S).reshape(S, 2), host(t_np, np.uint32, S),
                host(t_nu, np.uint64, S))

    def evalues(self, scores: np.ndarray, cnt: np.ndarray | None = None) -> np.ndarray:
        """hfx_evalues: E-values of [S, N] scores under the last fit."""
        S = self.sp.num_spec
        sc = np.ascontiguousarray(scores, np.float64).reshape(S, -1)
        N = sc.shape[1]
        t_sc, t_ev = dev(sc), empty(8 * S * N)
        t_cnt = dev(np.asarray(cnt, np.uint32)) if cnt is not None else None
        check(lib().hfx_evalues(self.ctx.h, ptr(t_ev), ptr(self.t_fit), ptr(t_sc),
                                ptr(t_cnt) if t_cnt is not None else None, N, S,
                                stream_handle()), "hfx_evalues")
        sync()
        return host(t_ev, np.float64, S * N).reshape(S, N)

    def sp_ions(self, mono: bool = True):
        """hfx_sp_ions_topn over the last top N: (b matched, y matched,
        total), each [S, N]."""
        S, N = self.sp.num_spec, self.topn
        t = [empty(4 * S * N) for _ in range(3)]
        check(lib().hfx_sp_ions_topn(self.ctx.h, *[ptr(x) for x in t], ptr(self.t_pos),
                                     ptr(self.t_cnt), ptr(self.t_begin), N, C.byref(self.db.desc),
                                     C.byref(self.sp.desc), int(mono), stream_handle()),
              "hfx_sp_ions_topn")
        sync()
        return tuple(host(x, np.uint32, S * N).reshape(S, N) for x in t)

    def exact(self, spectra: np.ndarray):
        """hfx_topn_exact for the listed spectra (every candidate exact)."""
        lst = np.asarray(spectra, np.uint32)
        if len(lst):
            t_list = dev(lst)
            check(lib().hfx_topn_exact(self.ctx.h, ptr(self.t_pos), ptr(self.t_exact),
                                       ptr(self.t_f32), ptr(self.t_cnt), self.topn, ptr(t_list),
                                       len(lst), ptr(self.t_scores), ptr(self.t_off), self.total,
                                       ptr(self.t_begin), ptr(self.t_num), C.byref(self.db.desc),
                                       C.byref(self.sp.desc), stream_handle()), "hfx_topn_exact")

    def flagged(self) -> np.ndarray:
        """Spectra that took the exact route in the last select()."""
        f = host(self.t_flag, np.uint32, self.sp.num_spec)
        return np.nonzero(f & FLAG_EXACT)[0].astype(np.uint32)

    def short(self) -> np.ndarray:
        """Spectra of the last select() whose FP32 scores broke the stated
        error bound (fewer than min(N, window) ranked)."""
        f = host(self.t_flag, np.uint32, self.sp.num_spec)
        return np.nonzero(f & FLAG_SHORT)[0].astype(np.uint32)

    def run(self, keep_scores: bool = False, dense: bool | str = False) -> PlainResult:
        """window -> XCorr (FP32 sparse; dense=True: BF16 hi+lo; dense="i8x3":
        int8 digits) -> top-N."""
        S, N = self.sp.num_spec, self.topn
        self.window()
        if self.total:
            if dense == "i8x3":
                self.score_dense(form="i8x3")
                self.select_rel_q(self.REL_I8, 0.0, self.RELQ_I8)
            elif dense:
                self.score_dense()
                self.select_rel(self.REL_DENSE)
            else:
                self.score()
                self.select()
            flagged, short = self.flagged(), self.short()
        else:
            flagged = short = np.zeros(0, np.uint32)
            self.t_cnt.zero()
        sync()
        return PlainResult(
            begin=host(self.t_begin, np.uint32, S).copy(),
            num=host(self.t_num, np.uint32, S).copy(),
            off=host(self.t_off, np.uint64, S + 1).copy(),
            top_pos=host(self.t_pos, np.uint32, S * N).reshape(S, N).copy(),
            top_exact=host(self.t_exact, np.float64, S * N).reshape(S, N).copy(),
            top_f32=host(self.t_f32, np.float32, S * N).reshape(S, N).copy(),
            top_cnt=host(self.t_cnt, np.uint32, S).copy(),
            flagged=flagged,
            scores=host(self.t_scores, np.float32, self.total).copy() if keep_scores else None,
            short=short,
        )


class Stream:
    """A non-blocking HIP stream (hfx_stream_create)."""

    def __init__(self):
        p = C.c_void_p()
        check(lib().hfx_stream_create(C.byref(p)), "hfx_stream_create")
        self.p = p.value

    def sync(self):
        check(lib().hfx_stream_sync(self.p), "hfx_stream_sync")

    def __del__(self):
        try:
            lib().hfx_stream_destroy(self.p)
        except Exception:
            pass


class PepIndex:
    """The per-database peptide index (hfx_pepix): b ladders of every
    peptide (dense form) and/or their sparse ion lists for fragment charge
    levels 1..levels (FP32 batched form).  Built once per database."""

    LADDER = 1
    IONS = 2
    INFO = ("forms", "levels", "max_len", "ladder_positions", "ion_entries", "bytes")

    def __init__(self, ctx: Ctx, db: DeviceDB, forms: int = LADDER | IONS, levels: int = 4,
                 max_len: int = 4096, stream: "Stream | None" = None):
        self.ctx, self.db = ctx, db
        h = C.c_void_p()
        check(lib().hfx_pepix_build(ctx.h, C.byref(db.desc), int(forms), int(levels),
                                    int(max_len), stream.p if stream is not None else None,
                                    C.byref(h)), "hfx_pepix_build")
        self.h = h.value

    def info(self) -> dict:
        a = np.zeros(6, np.uint64)
        check(lib().hfx_pepix_info(self.h, a.ctypes.data), "hfx_pepix_info")
        return {k: int(v) for k, v in zip(self.INFO, a)}

    def __del__(self):
        try:
            lib().hfx_pepix_destroy(self.h)
        except Exception:
            pass


class BatchSearch:
    """The batched path (hfx_batch_*): plan once, then run or replay a
    captured graph; results in PlainResult form.  With `index` the batch
    reads that per-database index; without, it builds its own on first use."""

    INFO = ("spectra", "scores", "peptides", "levels", "entries", "G", "C", "threads", "lstr",
            "zero", "precision", "chunk")
    FP32_SPARSE = 0
    BF16X2_DENSE = 1
    I8X3_DENSE = 2

    def __init__(self, ctx: Ctx, db: DeviceDB, sp: DeviceSpectra, eps: float, topn: int,
                 margin: int = 0, stream: Stream | None = None, index: PepIndex | None = None):
        L = lib()
        self.ctx, self.db, self.sp, self.topn = ctx, db, sp, int(topn)
        self.index = index
        self.stream = stream or Stream()
        h = C.c_void_p()
        check(L.hfx_batch_create(ctx.h, C.byref(h)), "hfx_batch_create")
        self.h = h.value
        if index is not None:
            check(L.hfx_batch_plan_ix(self.h, index.h, C.byref(sp.desc), float(eps), self.topn,
                                      int(margin), self.stream.p), "hfx_batch_plan_ix")
        else:
            check(L.hfx_batch_plan(self.h, C.byref(db.desc), C.byref(sp.desc), float(eps),
                                   self.topn, int(margin), self.stream.p), "hfx_batch_plan")
        self.captured = False

    def order(self) -> np.ndarray:
        """the spectrum order of the last run: (fragment charge level, precursor
        mass, index)"""
        out = np.zeros(self.sp.num_spec, np.uint32)
        check(lib().hfx_batch_order(self.h, out.ctypes.data), "hfx_batch_order")
        return out

    def info(self) -> dict:
        a = np.zeros(12, np.uint64)
        check(lib().hfx_batch_info(self.h, a.ctypes.data), "hfx_batch_info")
        return {k: int(v) for k, v in zip(self.INFO, a)}

    def dense_blocks(self) -> np.ndarray:
        """block starts of the dense form as positions in the plan's (fragment
        charge, precursor mass) order, the last one the spectrum count"""
        nb = C.c_uint32(0)
        check(lib().hfx_batch_dense_blocks(self.h, None, 0, C.byref(nb)), "hfx_batch_dense_blocks")
        out = np.zeros(nb.value + 1, np.uint32)
        check(lib().hfx_batch_dense_blocks(self.h, out.ctypes.data, len(out), C.byref(nb)),
              "hfx_batch_dense_blocks")
        return out

    def autotune(self):
        cfg = np.zeros(3, np.uint32)
        check(lib().hfx_batch_autotune(self.h, self.stream.p, cfg.ctypes.data), "hfx_batch_autotune")
        self.captured = False
        return tuple(int(v) for v in cfg)

    def time_phases(self) -> dict:
        ms = np.zeros(5, np.float32)
        check(lib().hfx_batch_time(self.h, self.stream.p, ms.ctypes.data), "hfx_batch_time")
        return dict(zip(("window", "prep", "score", "topn", "kernel"), (float(v) for v in ms)))

    def set_config(self, G: int, Cn: int, threads: int):
        check(lib().hfx_batch_set_config(self.h, G, Cn, threads), "hfx_batch_set_config")

    def set_precision(self, precision: int):
        """FP32_SPARSE (ion index + gathers), BF16X2_DENSE (BF16 MFMA) or
        I8X3_DENSE (int8 MFMA over three digits of each spectrum)."""
        check(lib().hfx_batch_set_precision(self.h, int(precision)), "hfx_batch_set_precision")
        self.captured = False

    def run(self):
        check(lib().hfx_batch_run(self.h, self.stream.p), "hfx_batch_run")

    def capture(self):
        check(lib().hfx_batch_capture(self.h, self.stream.p), "hfx_batch_capture")
        self.captured = True

    def replay(self):
        check(lib().hfx_batch_replay(self.h, self.stream.p), "hfx_batch_replay")

    def results(self, keep_scores: bool = False) -> PlainResult:
        self.stream.sync()
        ptrs = (C.c_void_p * 10)()
        check(lib().hfx_batch_outputs(self.h, ptrs), "hfx_batch_outputs")
        S, N = self.sp.num_spec, self.topn
        inf = self.info()

        def get(i, dtype, n):
            out = np.empty(n, dtype)
            if n:
                check(lib().hfx_memcpy_d2h(out.ctypes.data, ptrs[i], out.nbytes, self.stream.p),
                      "d2h")
            return out

        flags = get(4, np.uint32, S)
        r = PlainResult(
            begin=get(7, np.uint32, S), num=get(8, np.uint32, S), off=get(6, np.uint64, S + 1),
            top_pos=get(0, np.uint32, S * N).reshape(S, N),
            top_exact=get(1, np.float64, S * N).reshape(S, N),
            top_f32=get(2, np.float32, S * N).reshape(S, N), top_cnt=get(3, np.uint32, S),
            flagged=np.nonzero(flags & FLAG_EXACT)[0].astype(np.uint32),
            scores=get(5, np.float32, inf["scores"]) if keep_scores else None,
            short=np.nonzero(flags & FLAG_SHORT)[0].astype(np.uint32))
        self.stream.sync()
        return r

    def __del__(self):
        try:
            lib().hfx_batch_destroy(self.h)
        except Exception:
            pass


@dataclass
class ModInfo:
    """Fragment.hs:260-289 HostModInfo (from hfx_mod_info)."""

    ma: bytes
    ma_mass: np.ndarray
    mod_ma_count: np.ndarray  # [num_mod, num_ma]
    mod_sum: np.ndarray
    mod_delta: np.ndarray

    @property
    def num_ma(self) -> int:
        return len(self.ma)

    @property
    def num_mod(self) -> int:
        return len(self.mod_delta)


# the largest modification table mod_info builds (count vectors of up to 26
# acids, 4 bytes a count)
MAX_MODS = 1 << 20


def mod_info(var: dict, max_ma: int) -> ModInfo:
    L = lib()
    v = np.zeros(26, np.float32)
    for k, x in var.items():
        v[ord(k) - 65] = np.float32(x)
    ma = np.zeros(26, np.uint8)
    mam = np.zeros(26, np.float32)
    na, nm = C.c_uint32(), C.c_uint32()
    L.hfx_mod_info(v.ctypes.data, max_ma, ma.ctypes.data, mam.ctypes.data, C.byref(na), None,
                   None, None, 0, C.byref(nm))
    a, m = na.value, nm.value
    if m > MAX_MODS:
        raise HfxError(f"hfx_mod_info: {m} modifications of at most {max_ma - 1} sites over "
                       f"{a} acids exceed {MAX_MODS}")
    cnt = np.zeros(max(m * a, 1), np.uint32)
    sm = np.zeros(max(m, 1), np.uint32)
    dl = np.zeros(max(m, 1), np.float32)
    check(L.hfx_mod_info(v.ctypes.data, max_ma, ma.ctypes.data, mam.ctypes.data, C.byref(na),
                         cnt.ctypes.data, sm.ctypes.data, dl.ctypes.data, m, C.byref(nm)),
          "hfx_mod_info")
    return ModInfo(bytes(ma[:a]), mam[:a].copy(), cnt[: m * a].reshape(m, a).copy(), sm[:m].copy(),
                   dl[:m].copy())


@dataclass
class ModResult:
    spec_cand: np.ndarray     # candidates by mass per spectrum (rows before modability)
    num_valid: int
    num_mpep: int
    len_unrank: int
    mp_off: np.ndarray        # [S+1] first modified peptide of each spectrum
    top_pos: np.ndarray       # [S, N] index within the spectrum's modified peptides
    top_exact: np.ndarray
    top_f32: np.ndarray
    top_cnt: np.ndarray
    flagged: np.ndarray
    stages: dict              # every intermediate array, for parity checks
    scores: np.ndarray | None = None
    short: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))


class ModSearch:
    """Search.hs:282-394 searchWithMods for one batch, through the reference
    ABI (findBeginEnd_f, findModablePeptides, calcTotalModCands,
    prepareGenMod, genModCands) and the sparse modified-peptide scorer."""

    def __init__(self, ctx: Ctx, db: DeviceDB, sp: DeviceSpectra, mi: ModInfo, eps: float,
                 topn: int, margin: int = 0):
        self.ctx, self.db, self.sp, self.mi = ctx, db, sp, mi
        self.eps, self.topn, self.margin = float(eps), int(topn), int(margin)
        self.t_ma = dev(np.frombuffer(mi.ma, np.uint8))
        self.t_ma_mass = dev(mi.ma_mass)
        self.t_mod_cnt = dev(mi.mod_ma_count.reshape(-1))
        self.t_mod_sum = dev(mi.mod_sum)
        self.t_mod_delta = dev(mi.mod_delta)
        self._pool: dict[str, DevBuf] = {}

    def _buf(self, name: str, nbytes: int) -> DevBuf:
        """A device buffer of at least nbytes for one of run()'s arrays: the
        last run's, when it is large enough (a run's arrays stay valid until
        the next run), else a new one."""
        b = self._pool.get(name)
        if b is None or b.nbytes < max(int(nbytes), 1):
            self._pool.pop(name, None)
            b = self._pool[name] = empty(nbytes)
        return b

    def mpep_per_spectrum(self) -> np.ndarray:
        """The modified peptides each spectrum would score, from the stages
        whose arrays grow with the candidates by mass (the windows,
        findModablePeptides, calcTotalModCands), without expanding them."""
        L = lib()
        S, M, na = self.sp.num_spec, self.mi.num_mod, self.mi.num_ma
        db = self.db
        if S == 0:
            return np.zeros(0, np.uint64)
        t_b, t_e, t_sc, t_cand = empty(4 * S * M), empty(4 * S * M), empty(4 * S * M), empty(4 * S)
        check(L.hfx_mod_windows(self.ctx.h, t_b.p, t_e.p, t_sc.p, t_cand.p, C.byref(db.desc),
                                self.sp.t_mass.p, S, self.t_mod_delta.p, M, self.eps, None),
              "hfx_mod_windows")
        spec_cand = host(t_cand, np.uint32, S)
        scan = np.zeros(S + 1, np.uint32)
        np.cumsum(spec_cand, out=scan[1:])
        total = int(scan[-1])
        t_scan = dev(scan)
        t_pep, t_pmod = empty(4 * total), empty(4 * total)
        t_pcnt, t_vscan = empty(4 * total * na), empty(4 * S)
        nv = L.findModablePeptides(t_pep.p, t_pmod.p, t_pcnt.p, t_vscan.p, t_scan.p, S, total,
                                   db.t_ions.p, db.t_tc.p, db.t_tn.p, db.t_idx.p, t_b.p, t_e.p,
                                   t_sc.p, self.t_mod_cnt.p, M, self.t_ma.p, na)
        if nv == 0:
            return np.zeros(S, np.uint64)
        t_nmp, t_comb = empty(4 * nv), empty(4 * nv * na)
        L.calcTotalModCands(t_nmp.p, t_comb.p, self.t_mod_cnt.p, t_pep.p, t_pmod.p, t_pcnt.p,
                            nv, M, na)
        sync()
        # valid rows are spectrum-major; vscan[s] is spectrum s's first
        rows = np.zeros(nv + 1, np.uint64)
        np.cumsum(host(t_nmp, np.uint32, nv), out=rows[1:])
        vb = np.append(host(t_vscan, np.uint32, S).astype(np.int64), nv)
        return rows[vb[1:]] - rows[vb[:-1]]

    def run(self, keep_scores: bool = False, per_spectrum: bool = False,
            keep_stages: bool = True) -> ModResult:
        """per_spectrum: the windows by one findBeginEnd_f call per spectrum
        (Search.hs:317-322, the reference ABI); otherwise hfx_mod_windows
        for all spectra in one launch.  keep_stages: copy every intermediate
        array to the host (for parity checks)."""
        L = lib()
        S, M, na = self.sp.num_spec, self.mi.num_mod, self.mi.num_ma
        db = self.db
        t_b, t_e, t_sc = (self._buf(k, 4 * S * M) for k in ("b", "e", "sc"))
        if per_spectrum:
            spec_cand = np.zeros(S, np.uint32)
            for s in range(S):
                o = 4 * s * M
                spec_cand[s] = L.findBeginEnd_f(t_b.p + o, t_e.p + o, t_sc.p + o, db.t_r.p,
                                                db.t_idx.p, db.num_pep, self.t_mod_delta.p, M,
                                                float(self.sp.mass[s]), self.eps)
        else:
            t_cand = self._buf("cand", 4 * S)
            check(L.hfx_mod_windows(self.ctx.h, t_b.p, t_e.p, t_sc.p, t_cand.p,
                                    C.byref(db.desc), self.sp.t_mass.p, S, self.t_mod_delta.p,
                                    M, self.eps, None), "hfx_mod_windows")
            spec_cand = host(t_cand, np.uint32, S)
        scan = np.zeros(S + 1, np.uint32)
        np.cumsum(spec_cand, out=scan[1:])
        total = int(scan[-1])
        t_scan = self._buf("scan", scan.nbytes)
        t_scan.upload(scan)
        keep = host if keep_stages else (lambda *a: None)
        st = {"spec_cand": spec_cand, "begin": keep(t_b, np.uint32, S * M),
              "end": keep(t_e, np.uint32, S * M), "num_scan": keep(t_sc, np.uint32, S * M)}
        t_pep, t_pmod = self._buf("pep", 4 * total), self._buf("pmod", 4 * total)
        t_pcnt, t_vscan = self._buf("pcnt", 4 * total * na), self._buf("vscan", 4 * S)
        nv = L.findModablePeptides(t_pep.p, t_pmod.p, t_pcnt.p, t_vscan.p, t_scan.p, S, total,
                                   db.t_ions.p, db.t_tc.p, db.t_tn.p, db.t_idx.p, t_b.p, t_e.p,
                                   t_sc.p, self.t_mod_cnt.p, M, self.t_ma.p, na)
        st.update(valid_pep=keep(t_pep, np.uint32, nv), valid_mod=keep(t_pmod, np.uint32, nv),
                  valid_cnt=keep(t_pcnt, np.uint32, nv * na), valid_scan=keep(t_vscan, np.uint32, S))
        t_nmp, t_comb = self._buf("nmp", 4 * nv), self._buf("comb", 4 * nv * na)
        nmp = L.calcTotalModCands(t_nmp.p, t_comb.p, self.t_mod_cnt.p, t_pep.p, t_pmod.p, t_pcnt.p,
                                  nv, M, na) if nv else 0
        st.update(pep_num_mpep=keep(t_nmp, np.uint32, nv), comb=keep(t_comb, np.uint32, nv * na))
        t_mp = [self._buf(f"mp{k}", 4 * nmp) for k in range(5)]
        lu = L.prepareGenMod(*[t.p for t in t_mp], self.t_mod_sum.p, t_pep.p, t_pmod.p, t_nmp.p,
                             nv, nmp) if nmp else 0
        for name, t in zip(("mp_pep", "mp_mod", "mp_rank", "mp_ith", "mp_scan"), t_mp):
            st[name] = keep(t, np.uint32, nmp)
        t_unrank = self._buf("unrank", 4 * max(lu, 1))
        if nmp:
            L.genModCands(t_unrank.p, self.t_mod_cnt.p, t_mp[3].p, t_mp[2].p, t_mp[4].p, t_pmod.p,
                          t_pcnt.p, t_comb.p, nmp, na)
        sync()
        st["unrank"] = keep(t_unrank, np.uint32, lu)
        t_mpoff = self._buf("mpoff", 8 * (S + 1))
        check(L.hfx_mpep_offsets(self.ctx.h, t_mpoff.p, t_nmp.p, nv, t_vscan.p, S, None),
              "hfx_mpep_offsets")
        mp_off = host(t_mpoff, np.uint64, S + 1)
        self.modset = HfxModset(t_mp[0].p, t_mp[1].p, t_mp[4].p, t_unrank.p, self.t_mod_cnt.p,
                                self.t_mod_delta.p, self.t_ma_mass.p, na, nmp)
        N = self.topn
        t_scores = self._buf("scores", 4 * max(nmp, 1))
        t_pos, t_ex = self._buf("pos", 4 * S * N), self._buf("ex", 8 * S * N)
        t_f32, t_cnt, t_flag = (self._buf(k, n) for k, n in (("f32", 4 * S * N), ("cnt", 4 * S),
                                                              ("flag", 4 * S)))
        flagged = short = np.zeros(0, np.uint32)
        if nmp:
            check(L.hfx_xcorr_sparse_mod(self.ctx.h, t_scores.p, t_mpoff.p, nmp,
                                         C.byref(self.modset), C.byref(db.desc),
                                         C.byref(self.sp.desc), None), "hfx_xcorr_sparse_mod")
            check(L.hfx_topn_mod(self.ctx.h, t_pos.p, t_ex.p, t_f32.p, t_cnt.p, t_flag.p, N,
                                 self.margin, t_scores.p, t_mpoff.p, C.byref(self.modset),
                                 C.byref(db.desc), C.byref(self.sp.desc), None), "hfx_topn_mod")
            # flagged spectra were re-ranked exactly inside hfx_topn_mod
            flags = host(t_flag, np.uint32, S)
            flagged = np.nonzero(flags & FLAG_EXACT)[0].astype(np.uint32)
            short = np.nonzero(flags & FLAG_SHORT)[0].astype(np.uint32)
        else:
            t_cnt.zero()
        sync()
        # keep the device arrays alive as long as the result references them
        self._keep = (t_mp, t_unrank, t_pep, t_pmod, t_pcnt, t_nmp, t_comb, t_mpoff)
        self._scored = (t_scores, t_mpoff, nmp)
        self._top = (t_pos, t_cnt)
        return ModResult(
            spec_cand=spec_cand, num_valid=int(nv), num_mpep=int(nmp), len_unrank=int(lu),
            mp_off=mp_off, top_pos=host(t_pos, np.uint32, S * N).reshape(S, N),
            top_exact=host(t_ex, np.float64, S * N).reshape(S, N),
            top_f32=host(t_f32, np.float32, S * N).reshape(S, N),
            top_cnt=host(t_cnt, np.uint32, S), flagged=flagged, stages=st,
            scores=host(t_scores, np.float32, nmp) if keep_scores else None, short=short)


    def exact(self, spectra: np.ndarray):
        """hfx_topn_exact_mod for the listed spectra after run(): every modified
        candidate scored exactly.  Returns (top_pos, top_exact, top_cnt)."""
        t_scores, t_mpoff, nmp = self._scored
        S, N = self.sp.num_spec, self.topn
        t_pos, t_ex = empty(4 * S * N), empty(8 * S * N)
        t_f32, t_cnt = empty(4 * S * N), empty(4 * S)
        t_cnt.zero()
        lst = np.asarray(spectra, np.uint32)
        if nmp and len(lst):
            t_l = dev(lst)
            check(lib().hfx_topn_exact_mod(self.ctx.h, t_pos.p, t_ex.p, t_f32.p, t_cnt.p, N, t_l.p,
                                           len(lst), t_scores.p, t_mpoff.p, C.byref(self.modset),
                                           C.byref(self.db.desc), C.byref(self.sp.desc), None),
                  "hfx_topn_exact_mod")
        sync()
        return (host(t_pos, np.uint32, S * N).reshape(S, N),
                host(t_ex, np.float64, S * N).reshape(S, N), host(t_cnt, np.uint32, S))


    def sp_ions(self, mono: bool = True):
        """hfx_sp_ions_topn_mod over the last top N: (b matched, y matched,
        total), each [S, N]; matched on the unmodified peptide."""
        S, N = self.sp.num_spec, self.topn
        t_pos, t_cnt = self._top
        t = [empty(4 * S * N) for _ in range(3)]
        check(lib().hfx_sp_ions_topn_mod(self.ctx.h, *[ptr(x) for x in t], ptr(t_pos), ptr(t_cnt),
                                         ptr(self._scored[1]), N, C.byref(self.modset),
                                         C.byref(self.db.desc), C.byref(self.sp.desc), int(mono),
                                         stream_handle()), "hfx_sp_ions_topn_mod")
        sync()
        return tuple(host(x, np.uint32, S * N).reshape(S, N) for x in t)


# Device bytes one modified peptide holds while its batch is searched: the
# five arrays of prepareGenMod (20), its score (4), its unrank slice (at most
# four sites, 16), the top-N exact route's per-score scratch (about 8.2) and
# its share of the per-spectrum arrays; an upper bound.
MPEP_BYTES = 64


def mod_groups(spec_mpep, budget: int, per: int = MPEP_BYTES) -> list[tuple[int, int]]:
    """Consecutive spectrum ranges [s0, s1) whose modified peptides take at
    most `budget` bytes at `per` bytes each, as few as possible (a greedy
    cut in spectrum order).  A spectrum that alone exceeds the budget is an
    error."""
    cap = int(budget) // int(per)
    groups, s0, acc = [], 0, 0
    for s, c in enumerate(np.asarray(spec_mpep, np.int64).tolist()):
        if c > cap:
            raise HfxError(f"spectrum {s}: {c} modified peptides need {c * per} bytes, "
                           f"more than the {budget}-byte budget")
        if acc + c > cap:
            groups.append((s0, s))
            s0, acc = s, 0
        acc += c
    if s0 < len(spec_mpep):
        groups.append((s0, len(spec_mpep)))
    return groups


def search_mods_split(ctx: Ctx, db: DeviceDB, sp: DeviceSpectra, mi: ModInfo, eps: float,
                      topn: int, margin: int = 0, budget: int | None = None):
    """searchWithMods over batches of spectra sized to device memory
    (Execution.hs:77-93 cuts a search's spectra by the memory its buffers
    need): each spectrum's modified-peptide count comes from the stages
    that do not expand them, consecutive spectra are grouped so that a
    group's modified peptides fit `budget` bytes (default: half the free
    device memory), and each group is searched on a view of the spectra.
    Returns the merged result (positions index each spectrum's own modified
    peptides, as in one batch) and the groups."""
    if budget is None:
        free, total = C.c_size_t(), C.c_size_t()
        check(lib().hfx_mem_info(C.byref(free), C.byref(total)), "hfx_mem_info")
        budget = free.value // 2
    per = ModSearch(ctx, db, sp, mi, eps, topn, margin).mpep_per_spectrum()
    groups = mod_groups(per, budget)
    S, N = sp.num_spec, int(topn)
    spec_cand = np.zeros(S, np.uint32)
    mp_off = np.zeros(S + 1, np.uint64)
    top_pos = np.zeros((S, N), np.uint32)
    top_exact = np.zeros((S, N), np.float64)
    top_f32 = np.zeros((S, N), np.float32)
    top_cnt = np.zeros(S, np.uint32)
    flagged, short = [], []
    nv = nmp = lu = 0
    for s0, s1 in groups:
        r = ModSearch(ctx, db, sp.view(s0, s1), mi, eps, topn, margin).run(keep_stages=False)
        spec_cand[s0:s1] = r.spec_cand
        mp_off[s0:s1 + 1] = r.mp_off + np.uint64(nmp)
        top_pos[s0:s1], top_exact[s0:s1], top_f32[s0:s1] = r.top_pos, r.top_exact, r.top_f32
        top_cnt[s0:s1] = r.top_cnt
        flagged.append(r.flagged + np.uint32(s0))
        short.append(r.short + np.uint32(s0))
        nv, nmp, lu = nv + r.num_valid, nmp + r.num_mpep, lu + r.len_unrank
    cat = (lambda xs: np.concatenate(xs).astype(np.uint32) if xs else np.zeros(0, np.uint32))
    res = ModResult(spec_cand=spec_cand, num_valid=nv, num_mpep=nmp, len_unrank=lu,
                    mp_off=mp_off, top_pos=top_pos, top_exact=top_exact, top_f32=top_f32,
                    top_cnt=top_cnt, flagged=cat(flagged), stages={}, short=cat(short))
    return res, groups


@dataclass
class Match:
    """One reported match (Match.hs:17-27): exact score (XCorr * 10000),
    unmodified (0) or modified (1), window or modified-peptide position,
    E-value, and Ions = (matched, total)."""
    score: float
    is_mod: int
    index: int
    evalue: float
    ions: tuple


def finish_matches(ps: PlainSearch, plain: PlainResult, n: int, ms: "ModSearch | None" = None,
                   mods: ModResult | None = None, mono: bool = True, bin_width: int = 350,
                   perc: float = 3.0) -> list[list[Match]]:
    """Search.hs:115-141 (and :292-299 with modifications): per spectrum the
    best n matches of both searches in reported order, each with its E-value
    (fit over every score of the spectrum, unmodified and modified, on the
    device) and its SP ion count (on the device)."""
    ps.evalue_fit(ms if mods is not None else None, bin_width, perc)
    ev_p = ps.evalues(plain.top_exact, plain.top_cnt)
    nb, ny, tot = ps.sp_ions(mono)
    if mods is not None:
        ev_m = ps.evalues(mods.top_exact, mods.top_cnt)
        mnb, mny, mtot = ms.sp_ions(mono)
    out = []
    for s, items in enumerate(merge_rankings(plain, mods, n)):
        row = []
        for sc, is_mod, idx in items:
            if is_mod:
                k = int(np.nonzero(mods.top_pos[s, : mods.top_cnt[s]] == idx)[0][0])
                row.append(Match(sc, 1, idx, float(ev_m[s, k]),
                                 (int(mnb[s, k] + mny[s, k]), int(mtot[s, k]))))
            else:
                k = int(np.nonzero(plain.top_pos[s, : plain.top_cnt[s]] == idx)[0][0])
                row.append(Match(sc, 0, idx, float(ev_p[s, k]),
                                 (int(nb[s, k] + ny[s, k]), int(tot[s, k]))))
        out.append(row)
    return out


def merge_rankings(plain: PlainResult, mods: ModResult | None, n: int):
    """Search.hs:118-123: per spectrum the unmodified and modified matches
    merged by score descending; on equal scores unmodified matches first, then
    lower index.  Returns per spectrum a list of (exact score, is_mod, index)."""
    out = []
    for s in range(len(plain.top_cnt)):
        items = [(float(plain.top_exact[s, k]), 0, int(plain.top_pos[s, k]))
                 for k in range(int(plain.top_cnt[s]))]
        if mods is not None:
            items += [(float(mods.top_exact[s, k]), 1, int(mods.top_pos[s, k]))
                      for k in range(int(mods.top_cnt[s]))]
        items.sort(key=lambda t: (-t[0], t[1], t[2]))
        out.append(items[:n])
    return out
