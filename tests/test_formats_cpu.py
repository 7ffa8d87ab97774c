"""CPU tests of the host formats (hfx_io.cpp through hfx_amd.formats): FASTA
-> sequence database, the .index file, DTA and MGF spectra, the display
lookup and the result tables.

The reference's own result files (tests/spectra/*.out, extracted to
tests/golden/) pin the table layout (ppAsRows, also of their detail blocks)
and the number format: all 35 printed (M+H)+ values come out of the
fragment-mass arithmetic (piecewise FP32 residue sums, Q8) through
showFFloat (Just 4) byte for byte.  The files were written by an older
version of the tables (no Title row, no Ions and E-value columns, the
hydrogen atom's mass); the layout functions are the same.
"""
import ctypes as C
import gzip
import json
import math
import os
import struct
from decimal import ROUND_HALF_EVEN, Decimal

import numpy as np
import pytest

import hfx_amd
from hfx_amd import formats

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")

FASTA = """junk before the first header is dropped
>sp|P1|ONE First protein - test
MKWVTFISLLLLFSSAYSRGVFRRDTHKSEIAHRFKDLGEEHFK
GLVLIAFSQYLQQCPFDEHVKLVNELTEFAKTCVADESHAGCEKSLHTLFGDELCK
>sp|P2|TWO Second protein
MAPKRGSTYLLIVAMWERRPGSAFIKTRE
>sp|P3|THREE
GGGGK
"""


def _proteins():
    heads = ["sp|P1|ONE First protein - test", "sp|P2|TWO Second protein", "sp|P3|THREE"]
    seqs = ["MKWVTFISLLLLFSSAYSRGVFRRDTHKSEIAHRFKDLGEEHFK"
            "GLVLIAFSQYLQQCPFDEHVKLVNELTEFAKTCVADESHAGCEKSLHTLFGDELCK",
            "MAPKRGSTYLLIVAMWERRPGSAFIKTRE", "GGGGK"]
    return heads, seqs


def test_fasta_to_sequence_database(orc):
    """readFasta + makeSeqDB: headers, residues joined across lines, segments,
    and the digest of each protein identical to the oracle's."""
    mass = orc.mass_table({"C": 57.021464})
    db = formats.read_fasta(mass, text=FASTA, rule=1, missed=1, min_mass=200.0, max_mass=7200.0)
    heads, seqs = _proteins()
    assert db.headers == heads
    assert db.ions.tobytes() == "".join(seqs).encode()
    assert db.ion_seg.tolist() == [0, len(seqs[0]), len(seqs[0]) + len(seqs[1]),
                                   sum(len(s) for s in seqs)]
    r, tc, tn = orc.digest(db.ions, db.ion_seg, mass, rule=1, missed=1, min_mass=200.0,
                           max_mass=7200.0)
    assert (db.r == r).all() and (db.tc == tc).all() and (db.tn == tn).all()
    # each protein's first fragment
    first = [int(np.searchsorted(tc, s)) for s in db.ion_seg[:-1]] + [len(r)]
    assert db.frag_seg.tolist() == first


def test_fasta_gzip_file(tmp_path, orc):
    mass = orc.mass_table()
    p = tmp_path / "db.fasta.gz"
    with gzip.open(p, "wb") as f:
        f.write(FASTA.encode())
    a = formats.read_fasta(mass, path=str(p), rule=1, missed=2)
    b = formats.read_fasta(mass, text=FASTA, rule=1, missed=2)
    assert a.headers == b.headers and (a.r == b.r).all() and (a.tc == b.tc).all()


def _decode_index(raw: bytes):
    """Index.hs: an independent reading of the Data.Binary stream (Int =
    8-byte big-endian, Word32 big-endian, Integer = tag 0 + Int32, Float =
    decodeFloat's (Integer, Int) pair) after the gzip layer."""
    n = struct.unpack(">q", raw[:8])[0]
    params = raw[8:8 + n].decode()
    body = gzip.decompress(raw[8 + n:])
    pos = 0

    def take(k):
        nonlocal pos
        v = body[pos:pos + k]
        pos += k
        return v

    def i64():
        return struct.unpack(">q", take(8))[0]

    heads = [take(i64()).decode() for _ in range(i64())]
    ions = take(i64())
    seg = [struct.unpack(">I", take(4))[0] for _ in range(i64())]
    rs = []
    for _ in range(i64()):
        assert take(1) == b"\x00"
        m = struct.unpack(">i", take(4))[0]
        e = i64()
        rs.append(math.ldexp(m, e))
    c = [struct.unpack(">I", take(4))[0] for _ in range(i64())]
    nn = [struct.unpack(">I", take(4))[0] for _ in range(i64())]
    fs = [struct.unpack(">I", take(4))[0] for _ in range(i64())]
    assert pos == len(body)
    return params, heads, ions, seg, rs, c, nn, fs


def test_index_file_round_trip_and_encoding(tmp_path, orc):
    """writeIndex / readIndex: the database read back is identical, and the
    bytes follow Data.Binary's encoding (read by an independent decoder)."""
    mass = orc.mass_table({"C": 57.021464})
    db = formats.read_fasta(mass, text=FASTA, rule=1, missed=2, min_mass=200.0)
    params = ("\ndatabase         = test.fasta\nmissed-cleavages = 2\ndigestion-rule   = 1\n"
              "add_C = 57.021464\n")
    p = str(tmp_path / "test.index")
    db.write_index(p, params)
    back = formats.read_index(p)
    assert back.params == params and back.headers == db.headers
    for k in ("ions", "ion_seg", "r", "tc", "tn", "frag_seg"):
        assert (getattr(back, k) == getattr(db, k)).all(), k
    par, heads, ions, seg, rs, c, nn, fs = _decode_index(open(p, "rb").read())
    assert par == params and heads == db.headers and ions == db.ions.tobytes()
    assert seg == db.ion_seg.tolist() and c == db.tc.tolist() and nn == db.tn.tolist()
    assert fs == db.frag_seg.tolist()
    assert np.array(rs, np.float32).tolist() == db.r.tolist()
    assert all(float(np.float32(v)) == v for v in rs)  # exact (mantissa, exponent) pairs


def test_index_rejects_a_truncated_file(tmp_path, orc):
    mass = orc.mass_table()
    db = formats.read_fasta(mass, text=FASTA)
    p = tmp_path / "t.index"
    db.write_index(str(p), "x")
    raw = p.read_bytes()
    p.write_bytes(raw[: len(raw) - 9])
    with pytest.raises(Exception):
        formats.read_index(str(p))


def test_lookup_labels(orc):
    """Location.hs:42-58: "X.PEP.Y" with '-' at a protein's ends; a
    protein's final piece (tn one past its end) shows the next protein's
    first residue (Q28) and a '-' flank; (M+H)+ = (r + H) + H2O."""
    mass = orc.mass_table()
    db = formats.read_fasta(mass, text=FASTA, rule=1, missed=0, min_mass=0.0, max_mass=1e9)
    heads, seqs = _proteins()
    allres = "".join(seqs)
    for k in range(len(db.r)):
        label, prot, mh = db.lookup(k)
        a, b = int(db.ion_seg[prot]), int(db.ion_seg[prot + 1])
        c, n = int(db.tc[k]), int(db.tn[k])
        assert a <= c < b
        pep = allres[c:n + 1] if n < len(allres) else allres[c:n]
        left = allres[c - 1] if c > a else "-"
        right = allres[n + 1] if n < b - 1 else "-"
        assert label == f"{left}.{pep}.{right}", k
        assert mh == np.float32(np.float32(db.r[k] + np.float32(1.007276)) + np.float32(18.010565))
    # protein 1's last piece "GSAFIKTRE" ends at its end: shows P3's G
    last = [db.lookup(k)[0] for k in range(len(db.r)) if db.lookup(k)[1] == 1][-1]
    assert last == "R.EG.-"
    # a label longer than the first buffer comes back whole
    long = "ACDEFGHIKLMNPQRSTVWY" * 250
    db = formats.read_fasta(mass, text=f">big\n{long}\n", rule=0, missed=0, min_mass=0.0,
                            max_mass=1e9)
    assert db.lookup(0)[0] == f"-.{long}.-"


def test_dta_parse(orc):
    """DTA.hs:36-67: MH+ and charge, then pairs; the precursor m/z is
    (MH + (z - 1) H) / z; a fractional charge and an empty file are errors."""
    pcr, z, mz, it = formats.parse_dta("1595.7012 3\n200.1 10\n300.25 20.5\n400 0.5\n")
    assert z == 3 and pcr == orc.dta_precursor(1595.7012, 3.0)
    assert mz.tolist() == np.float32([200.1, 300.25, 400]).tolist()
    assert it.tolist() == np.float32([10, 20.5, 0.5]).tolist()
    for bad in ("1595.7 2.5\n100 1\n", "", "1595.7 2\nabc 1\n"):
        with pytest.raises(Exception):
            formats.parse_dta(bad)


def test_mgf_parse():
    """MGF.hs: samples between BEGIN IONS and END IONS; TITLE, CHARGE (with
    its sign) and PEPMASS (intensity ignored) kept; charge and mass required."""
    text = ("BEGIN IONS\nTITLE=scan 1\nPEPMASS=532.28 12000\nCHARGE=2+\n100.1 5\n200.2 6\n"
            "END IONS\n\nBEGIN IONS\nPEPMASS=800.5\nCHARGE=3+\nRTINSECONDS=12\n150 1\n"
            "END IONS\n")
    titles, pcr, z, off, mz, it = formats.parse_mgf(text)
    assert titles == ["scan 1", ""]
    assert pcr.tolist() == np.float32([532.28, 800.5]).tolist()
    assert z.tolist() == [2.0, 3.0] and off.tolist() == [0, 2, 3]
    assert mz.tolist() == np.float32([100.1, 200.2, 150]).tolist()
    with pytest.raises(Exception):
        formats.parse_mgf("BEGIN IONS\nPEPMASS=500\n100 1\nEND IONS\n")


# ------------------------------------------------------------ tables

def _out_files():
    src = os.path.join(GOLD, "reference_out_tables.json")
    return json.load(open(src))


def test_table_layout_reproduces_reference_files():
    """ppAsRows 1 over the printed cells of each result file gives back its
    table byte for byte; ppAsRows 0 with "::" its configuration block."""
    for f in _out_files():
        rows = [r.split("\t") for r in f["table_cells"]]
        assert formats.format_table(rows, 1) == "\n".join(f["table_lines"]), f["file"]
        cfg = [r.split("\t") for r in f["config_cells"]]
        assert formats.format_table(cfg, 0) == "\n".join(f["config_lines"]), f["file"]


def test_detail_block_of_the_reference_files():
    """The files' detail blocks were laid out by ppAsRows 1 over (" n.",
    header): reproduced byte for byte."""
    for f in _out_files():
        rows = [[f" {i + 1}.", h] for i, h in enumerate(f["detail_headers"])]
        assert formats.format_table(rows, 1) == "\n".join(f["detail_lines"]), f["file"]


def test_detail_block_boxes():
    """The current printResultsDetail (PrettyPrint.hs:137-144): "n."
    right-aligned in 3 columns, two spaces, the header flowed into lines of
    at most 95 characters (Boxes' para), every line padded to the widest;
    hand-derived."""
    long = " ".join(["word%02d" % i for i in range(20)])   # 139 characters
    out = formats.format_detail(["P1|A short header", long])
    lines = out.split("\n")[:-1]
    first = " ".join(["word%02d" % i for i in range(13)])   # 90 characters: 14 words would be 97
    rest = " ".join(["word%02d" % i for i in range(13, 20)])
    w = 5 + 90
    assert lines == [(" 1.  P1|A short header").ljust(w), (" 2.  " + first).ljust(w),
                     ("     " + rest).ljust(w)]


def _f32_pieces(mass, pep):
    """Q8: the residual of a fragment spanning several cleavage pieces is
    the FP32 sum of the pieces' FP32 sums (pieces end after K or R)."""
    pieces, cur = [], ""
    for ch in pep:
        cur += ch
        if ch in "KR":
            pieces.append(cur)
            cur = ""
    if cur:
        pieces.append(cur)
    tot = np.float32(0)
    for pc in pieces:
        s = np.float32(0)
        for ch in pc:
            s = np.float32(s + mass[ord(ch) - 65])
        tot = np.float32(tot + s)
    return tot


def test_printed_masses_reproduced_exactly(orc):
    """All 35 (M+H)+ values of the result files: (residual + H) + H2O in
    FP32 with the hydrogen atom's mass those files were written with, the
    residual as Q8's piecewise sum, printed by showFFloat (Just 4)."""
    mass = orc.mass_table({"C": 57.0})
    rows = json.load(open(os.path.join(GOLD, "reference_out_peptides.json")))
    n = 0
    for e in rows:
        for m in e["matches"]:
            res = _f32_pieces(mass, m["peptide"])
            mh = np.float32(np.float32(res + np.float32(1.0078246)) + np.float32(18.010565))
            assert formats.format_number(float(mh), 1, 4, is_float=True) == m["mh"], m
            n += 1
    assert n == 35


def _py_show_ffloat(x, dec, is_float):
    """showFFloat via Python's shortest repr and Decimal, halves to even."""
    s = np.format_float_positional(np.float32(x) if is_float else np.float64(x), unique=True,
                                   trim="-")
    d = Decimal(s).quantize(Decimal(1).scaleb(-dec), rounding=ROUND_HALF_EVEN)
    out = format(d, "f")
    return out


def _py_show_efloat(x, dec):
    s = repr(float(x))
    d = Decimal(s)
    if d == 0:
        return "0." + "0" * dec + "e0"
    e = d.adjusted()
    m = (d.scaleb(-e)).quantize(Decimal(1).scaleb(-dec), rounding=ROUND_HALF_EVEN)
    if m >= 10:
        m = (m / 10).quantize(Decimal(1).scaleb(-dec), rounding=ROUND_HALF_EVEN)
        e += 1
    return f"{m}e{e}"


def test_number_formats_against_decimal():
    """showFFloat (Just 4) of Floats and Doubles, showEFloat (Just 4) and
    show, against Python's shortest repr with Decimal rounding (halves to
    even, as GHC's roundTo), on random, tiny, huge and exact-half values."""
    rng = np.random.default_rng(9)
    vals = list(rng.uniform(0, 5000, 300)) + list(10.0 ** rng.uniform(-12, 12, 300))
    vals += [0.0, 1.0, 0.5, 0.00005, 0.00015, 1.23445, 2.5e-5, 1595.6831, 123456789.0]
    for v in vals:
        for is_float in (True, False):
            x = float(np.float32(v)) if is_float else float(v)
            assert formats.format_number(x, 1, 4, is_float) == _py_show_ffloat(x, 4, is_float), x
        assert formats.format_number(v, 2, 4) == _py_show_efloat(v, 4), v
    assert formats.format_number(-1.5, 1, 4) == "-1.5000"
    assert formats.format_number(float("nan"), 2, 4) == "NaN"
    # show: fixed between 0.1 and 10^7, else d.ddde<n>
    for v, want in [(3.0, "3.0"), (0.1, "0.1"), (0.01, "1.0e-2"), (1595.7012, "1595.7012"),
                    (1.0e7, "1.0e7"), (9999999.0, "9999999.0"), (123.456, "123.456")]:
        assert formats.format_number(v, 0, 0, is_float=True) == want, v
    # a Float is the value rounded to FP32: Doubles beyond its range are
    # infinite, below its smallest subnormal zero
    for kind in (0, 1, 2):
        assert formats.format_number(1e300, kind, 4, is_float=True) == "Infinity"
        assert formats.format_number(-1e300, kind, 4, is_float=True) == "-Infinity"
    assert formats.format_number(1e-300, 0, 0, is_float=True) == "0.0"
    assert formats.format_number(1e300, 0, 0) == "1.0e300"
    # the widest fixed form comes back whole; a buffer too small is an error
    big = formats.format_number(-1.7976931348623157e308, 1, 4)
    assert len(big) == 1 + 309 + 5 and big.startswith("-17976931348623157") and big.endswith(".0000")
    buf = C.create_string_buffer(8)
    assert hfx_amd.lib().hfx_format_number(12345.678, 0, 1, 4, buf, 8) != 0
    assert buf.value == b"12345.6" and "needs 11 bytes" in hfx_amd.last_error()


def test_format_scan_layout():
    """printScanResults for one scan: the configuration block (Title,
    Spectrum, Database, Enzyme, (M+H)+ Mass with the charge), the match
    table with Ions and E-value, the detail block, blank lines between."""
    rows = [dict(mh=1595.6831, xcorr=2.8196, ions=(18, 26), evalue=1.234e-5,
                 reference="P13726|TF_HUMAN", peptide="R.KSTDSPVECMGQEK.G",
                 header="P13726|TF_HUMAN Tissue factor precursor - Homo sapiens (Human)"),
            dict(mh=1597.7794, xcorr=2.2370, ions=(12, 28), evalue=0.031,
                 reference="Q5FWE3|PRRT3_HUMAN", peptide="R.LACPAPSGKSEVPER.P",
                 header="Q5FWE3|PRRT3_HUMAN Proline-rich transmembrane protein 3")]
    out = formats.format_scan("", "a.dta", "db.fasta", 1, 532.85376, 3.0, 3.0, rows, 5, 3)
    lines = out.split("\n")
    assert lines[0].startswith("Title       ::")
    assert lines[3] == "Enzyme      :: Trypsin 1 KR P".ljust(len(lines[3]))
    # (precursor * z) - ((z * H) - 1) in Float, shown with its shortest digits
    z, h = np.float32(3.0), np.float32(1.007276)
    m = np.float32(np.float32(np.float32(532.85376) * z) - np.float32(np.float32(z * h) - 1))
    assert m == np.float32(1596.5394)
    assert lines[4].startswith("(M+H)+ Mass :: 1596.5394 ~ 3.0 (+3)")
    assert lines[5] == ""
    assert lines[6].split() == ["#", "(M+H)+", "deltCn", "XCorr", "Ions", "Reference",
                                "Peptide", "E-value"]
    assert lines[8].split() == ["1.", "1595.6831", "0.0000", "2.8196", "18/26",
                                "P13726|TF_HUMAN", "R.KSTDSPVECMGQEK.G", "1.2340e-5"]
    assert lines[9].split()[2] == "0.2066" and lines[9].split()[-1] == "3.1000e-2"
    assert lines[10] == ""
    assert lines[11].startswith(" 1.  P13726|TF_HUMAN Tissue factor")
    assert out.endswith("\n\n")
