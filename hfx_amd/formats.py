"""Host formats around the search path, over libhfx_hip.so's C ABI
(hfx_io.cpp): FASTA -> sequence database (Fasta.hs, Fragment.hs:197-233
makeSeqDB) and its .index file (Index.hs:43-81), DTA and MGF spectra
(Spectrum/DTA.hs, Spectrum/MGF.hs), the display lookup (Location.hs:42-58)
and the result tables (Util/PrettyPrint.hs:83-180).  Host memory only; no
GPU is needed."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import check, lib


class HfxScanHeader(C.Structure):
    _fields_ = [("title", C.c_char_p), ("spectrum_file", C.c_char_p), ("database", C.c_char_p),
                ("enzyme", C.c_char_p), ("precursor", C.c_float), ("charge", C.c_float),
                ("tolerance", C.c_float)]


class HfxMatchRow(C.Structure):
    _fields_ = [("mh", C.c_float), ("xcorr", C.c_float), ("ions_matched", C.c_uint32),
                ("ions_total", C.c_uint32), ("evalue", C.c_double), ("reference", C.c_char_p),
                ("peptide", C.c_char_p), ("header", C.c_char_p)]


@dataclass
class SeqDB:
    """Fragment.hs:145-166 SequenceDB on the host."""
    headers: list
    ions: np.ndarray       # uint8 residue letters
    ion_seg: np.ndarray    # uint32 [proteins + 1]
    r: np.ndarray          # float32 residual per fragment
    tc: np.ndarray         # uint32
    tn: np.ndarray         # uint32
    frag_seg: np.ndarray   # uint32 [proteins + 1]
    params: str = ""
    h: int | None = None   # the library's handle

    def __del__(self):
        if self.h:
            try:
                lib().hfx_seqdb_free(self.h)
            except Exception:
                pass

    def lookup(self, k: int):
        """(label "X.PEP.Y", protein, (M+H)+) of fragment k."""
        cap = 4096
        n, p, mh = C.c_uint32(), C.c_uint32(), C.c_float()
        while True:
            buf = C.create_string_buffer(cap)
            check(lib().hfx_seqdb_lookup(self.h, k, buf, cap, C.byref(n), C.byref(p),
                                         C.byref(mh)), "hfx_seqdb_lookup")
            if n.value < cap:  # the whole label fitted (n is its full length)
                return buf.value.decode(), int(p.value), np.float32(mh.value)
            cap = n.value + 1

    def write_index(self, path: str, params: str):
        check(lib().hfx_seqdb_write_index(self.h, path.encode(), params.encode()),
              "hfx_seqdb_write_index")


def _seqdb(h) -> SeqDB:
    L = lib()
    info = np.zeros(5, np.uint64)
    check(L.hfx_seqdb_info(h, info.ctypes.data), "hfx_seqdb_info")
    npr, nres, hb, nf, pb = (int(v) for v in info)
    ions = np.zeros(max(nres, 1), np.uint8)
    seg = np.zeros(npr + 1, np.uint32)
    hdr = np.zeros(max(hb, 1), np.uint8)
    hoff = np.zeros(npr + 1, np.uint64)
    r = np.zeros(max(nf, 1), np.float32)
    tc = np.zeros(max(nf, 1), np.uint32)
    tn = np.zeros(max(nf, 1), np.uint32)
    fseg = np.zeros(npr + 1, np.uint32)
    par = np.zeros(max(pb, 1), np.uint8)
    check(L.hfx_seqdb_copy(h, ions.ctypes.data, seg.ctypes.data, hdr.ctypes.data,
                           hoff.ctypes.data, r.ctypes.data, tc.ctypes.data, tn.ctypes.data,
                           fseg.ctypes.data, par.ctypes.data), "hfx_seqdb_copy")
    hb_ = hdr.tobytes()
    headers = [hb_[int(hoff[i]):int(hoff[i + 1])].decode("latin-1") for i in range(npr)]
    return SeqDB(headers, ions[:nres], seg, r[:nf], tc[:nf], tn[:nf], fseg,
                 par[:pb].tobytes().decode("latin-1"), h)


def read_fasta(mass26, path: str | None = None, text: str | bytes | None = None, rule: int = 1,
               missed: int = 2, min_mass: float = 400.0, max_mass: float = 7200.0) -> SeqDB:
    """readFasta + makeSeqDB (".gz" files inflated)."""
    h = C.c_void_p()
    t = text.encode() if isinstance(text, str) else text
    m = np.ascontiguousarray(mass26, np.float32)
    check(lib().hfx_seqdb_from_fasta(path.encode() if path else None, t, len(t) if t else 0,
                                     m.ctypes.data, rule, missed, min_mass, max_mass,
                                     C.byref(h)), "hfx_seqdb_from_fasta")
    return _seqdb(h.value)


def read_index(path: str) -> SeqDB:
    h = C.c_void_p()
    check(lib().hfx_seqdb_read_index(path.encode(), C.byref(h)), "hfx_seqdb_read_index")
    return _seqdb(h.value)


def parse_dta(text: str | bytes):
    """(precursor m/z, charge, m/z, intensity) of a DTA file's text."""
    t = text.encode() if isinstance(text, str) else text
    cap = t.count(b"\n") + 2
    mz, it = np.zeros(cap, np.float32), np.zeros(cap, np.float32)
    pcr, z, n = C.c_float(), C.c_float(), C.c_uint32()
    check(lib().hfx_dta_parse(t, len(t), C.byref(pcr), C.byref(z), mz.ctypes.data,
                              it.ctypes.data, cap, C.byref(n)), "hfx_dta_parse")
    k = int(n.value)
    return np.float32(pcr.value), np.float32(z.value), mz[:k].copy(), it[:k].copy()


def parse_mgf(text: str | bytes):
    """(titles, precursor m/z, charge, offsets, m/z, intensity)."""
    t = text.encode() if isinstance(text, str) else text
    L = lib()
    h = C.c_void_p()
    check(L.hfx_mgf_parse(t, len(t), C.byref(h)), "hfx_mgf_parse")
    try:
        info = np.zeros(3, np.uint64)
        check(L.hfx_mgf_info(h, info.ctypes.data), "hfx_mgf_info")
        ns, npk, tb = (int(v) for v in info)
        pcr, z = np.zeros(max(ns, 1), np.float32), np.zeros(max(ns, 1), np.float32)
        off = np.zeros(ns + 1, np.uint32)
        mz, it = np.zeros(max(npk, 1), np.float32), np.zeros(max(npk, 1), np.float32)
        tit, toff = np.zeros(max(tb, 1), np.uint8), np.zeros(ns + 1, np.uint64)
        check(L.hfx_mgf_copy(h, pcr.ctypes.data, z.ctypes.data, off.ctypes.data, mz.ctypes.data,
                             it.ctypes.data, tit.ctypes.data, toff.ctypes.data), "hfx_mgf_copy")
    finally:
        L.hfx_mgf_free(h)
    tb_ = tit.tobytes()
    titles = [tb_[int(toff[i]):int(toff[i + 1])].decode("latin-1") for i in range(ns)]
    return titles, pcr[:ns], z[:ns], off, mz[:npk], it[:npk]


def format_number(v: float, kind: int, digits: int = 4, is_float: bool = False) -> str:
    cap = 320 + max(int(digits), 0)  # a sign, 309 integer digits, a point, the decimals
    buf = C.create_string_buffer(cap)
    check(lib().hfx_format_number(float(v), int(is_float), kind, digits, buf, cap),
          "hfx_format_number")
    return buf.value.decode()


def _text(fn, *args) -> str:
    n = C.c_uint64()
    check(fn(*args, None, 0, C.byref(n)), fn.__name__)
    buf = C.create_string_buffer(int(n.value) + 1)
    check(fn(*args, buf, int(n.value) + 1, C.byref(n)), fn.__name__)
    return buf.raw[:int(n.value)].decode("latin-1")


def format_table(rows: list, q: int) -> str:
    """ppAsRows q over rows of cell strings."""
    nr, nc = len(rows), max(len(r) for r in rows)
    cells = (C.c_char_p * (nr * nc))(*[(r[j] if j < len(r) else "").encode()
                                       for r in rows for j in range(nc)])
    return _text(lib().hfx_format_table, cells, nr, nc, q)


def format_detail(headers: list) -> str:
    hs = (C.c_char_p * len(headers))(*[h.encode() for h in headers])
    return _text(lib().hfx_format_detail, hs, len(headers))


def format_scan(title: str, spectrum_file: str, database: str, rule: int, precursor: float,
                charge: float, tolerance: float, rows: list, nshow: int = 5,
                ndetail: int = 3) -> str:
    """printScanResults for one scan; rows: dicts with mh, xcorr, ions
    (matched, total), evalue, reference, peptide, header."""
    keep = []

    def b(s):
        keep.append(s.encode())
        return keep[-1]

    hdr = HfxScanHeader(b(title), b(spectrum_file), b(database), lib().hfx_digestion_rule_name(rule),
                        precursor, charge, tolerance)
    arr = (HfxMatchRow * max(len(rows), 1))()
    for i, r in enumerate(rows):
        arr[i] = HfxMatchRow(r["mh"], r["xcorr"], r["ions"][0], r["ions"][1], r["evalue"],
                             b(r["reference"]), b(r["peptide"]), b(r["header"]))
    return _text(lib().hfx_format_scan, C.byref(hdr), arr, len(rows), nshow, ndetail)
