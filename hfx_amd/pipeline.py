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
from dataclasses import dataclass

import numpy as np

from . import HfxDb, HfxSpectra, check, lib

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

    def lengths(self) -> np.ndarray:
        return host(self.t_len, np.uint32, self.num_spec)

    def status(self) -> np.ndarray:
        return host(self.t_status, np.uint32, self.num_spec)

    def x(self) -> np.ndarray:
        return host(self.t_x, np.float32).reshape(self.num_spec, self.x_stride)


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


class PlainSearch:
    """searchWithoutMods for one batch; buffers reused across calls."""

    def __init__(self, ctx: Ctx, db: DeviceDB, sp: DeviceSpectra, eps: float, topn: int,
                 margin: int = 16):
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
                         ptr(self.t_scores), ptr(self.t_off), ptr(self.t_begin), ptr(self.t_num),
                         C.byref(self.db.desc), C.byref(self.sp.desc), st), "hfx_topn")

    def fallback(self) -> np.ndarray:
        S = self.sp.num_spec
        flags = host(self.t_flag, np.uint32, S)
        lst = np.nonzero(flags)[0].astype(np.uint32)
        if len(lst):
            t_list = dev(lst)
            check(lib().hfx_topn_exact(self.ctx.h, ptr(self.t_pos), ptr(self.t_exact),
                                       ptr(self.t_f32), ptr(self.t_cnt), self.topn, ptr(t_list),
                                       len(lst), ptr(self.t_begin), ptr(self.t_num),
                                       C.byref(self.db.desc), C.byref(self.sp.desc),
                                       stream_handle()), "hfx_topn_exact")
        return lst

    def run(self, keep_scores: bool = False) -> PlainResult:
        S, N = self.sp.num_spec, self.topn
        self.window()
        if self.total:
            self.score()
            self.select()
            flagged = self.fallback()
        else:
            flagged = np.zeros(0, np.uint32)
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
        )
