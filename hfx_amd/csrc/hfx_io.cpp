/*
 * hfx_io.cpp -- the host formats on either side of the search path.
 *
 *   FASTA -> sequence database   Sequence/Fasta.hs:30-35, Fragment.hs:197-233
 *                                (makeSeqDB: headers, ionSeq, ionSeg, digest)
 *   .index files                 Sequence/Index.hs:43-81 (writeIndex,
 *                                readIndex; Data.Binary + gzip)
 *   DTA / MGF spectra            Spectrum/DTA.hs:36-67, Spectrum/MGF.hs:38-100
 *   display lookup               Sequence/Location.hs:42-58 ("X.PEP.Y")
 *   result tables                Util/PrettyPrint.hs:83-180 (printConfig,
 *                                printResults, printResultsDetail) with
 *                                Haskell's show / showFFloat / showEFloat
 *
 * Plain C++ on the host (the reference's side of these is compiled Haskell);
 * nothing here touches the device.
 */
#include <charconv>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <zlib.h>

#include "../../include/hfx_hip.h"

void hfx_set_error(const char *fmt, ...);

namespace {

/* a container's bytes to dst (when dst is given): nothing for an empty one,
 * whose data() may be null, which memcpy does not accept even for 0 bytes */
template <class V>
void copy_out(void *dst, const V &v)
{
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
}

/* ------------------------------------------------------------------ */
/* Files                                                                 */
/* ------------------------------------------------------------------ */

/* the whole file, inflated when its name ends in ".gz" (Fasta.hs:31-33) */
bool read_file(const char *path, std::string &out)
{
    const size_t n = std::strlen(path);
    out.clear();
    if (n > 3 && std::strcmp(path + n - 3, ".gz") == 0) {
        gzFile f = gzopen(path, "rb");
        if (!f) return false;
        char buf[1 << 16];
        int k;
        while ((k = gzread(f, buf, sizeof buf)) > 0) out.append(buf, (size_t)k);
        const bool ok = k == 0;
        gzclose(f);
        return ok;
    }
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, k);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

/* L.lines: split on '\n'; a final line without '\n' still counts */
template <class F>
void for_lines(const char *t, uint64_t len, F &&f)
{
    uint64_t i = 0;
    while (i < len) {
        uint64_t j = i;
        while (j < len && t[j] != '\n') ++j;
        f(t + i, j - i);
        i = j + 1;
    }
}

/* ------------------------------------------------------------------ */
/* Data.Binary encodings (binary's Binary instances, big-endian)         */
/* ------------------------------------------------------------------ */

void put_u8(std::string &o, uint8_t v) { o.push_back((char)v); }
void put_u32(std::string &o, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) o.push_back((char)(v >> s));
}
void put_u64(std::string &o, uint64_t v)
{
    for (int s = 56; s >= 0; s -= 8) o.push_back((char)(v >> s));
}
/* Int: 64-bit big-endian */
void put_int(std::string &o, int64_t v) { put_u64(o, (uint64_t)v); }
/* Integer: tag 0 + Int32 when it fits, else tag 1, sign byte, and the
 * little-endian base-256 digits of |v| as a list (length Int + bytes) */
void put_integer(std::string &o, int64_t v)
{
    if (v >= INT32_MIN && v <= INT32_MAX) {
        put_u8(o, 0);
        put_u32(o, (uint32_t)(int32_t)v);
        return;
    }
    put_u8(o, 1);
    put_u8(o, v < 0 ? 0 : 1);
    uint64_t a = v < 0 ? (uint64_t)(-(v + 1)) + 1 : (uint64_t)v;
    std::string d;
    while (a) {
        d.push_back((char)(a & 0xFF));
        a >>= 8;
    }
    put_int(o, (int64_t)d.size());
    o += d;
}
/* Float: put (decodeFloat x) = (Integer mantissa, Int exponent), the
 * mantissa of 24 significant bits (0 for 0) */
void put_float(std::string &o, float x)
{
    int64_t m = 0;
    int64_t e = 0;
    if (x != 0.0f && std::isfinite(x)) {
        int ex = 0;
        const float f = std::frexp(x, &ex); /* x = f 2^ex, 0.5 <= |f| < 1 */
        m = (int64_t)std::ldexp(f, 24);
        e = ex - 24;
    }
    put_integer(o, m);
    put_int(o, e);
}

struct Reader {
    const unsigned char *p;
    uint64_t n, i = 0;
    bool ok = true;
    bool need(uint64_t k)
    {
        if (i + k > n) ok = false;
        return ok;
    }
    uint8_t u8()
    {
        if (!need(1)) return 0;
        return p[i++];
    }
    uint32_t u32()
    {
        if (!need(4)) return 0;
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) v = (v << 8) | p[i++];
        return v;
    }
    uint64_t u64()
    {
        if (!need(8)) return 0;
        uint64_t v = 0;
        for (int k = 0; k < 8; ++k) v = (v << 8) | p[i++];
        return v;
    }
    int64_t integer()
    {
        const uint8_t tag = u8();
        if (tag == 0) return (int64_t)(int32_t)u32();
        const uint8_t sign = u8();
        const uint64_t k = u64();
        if (k > 8 || !need(k)) {
            ok = false;
            return 0;
        }
        uint64_t a = 0;
        for (uint64_t j = 0; j < k; ++j) a |= (uint64_t)p[i + j] << (8 * j);
        i += k;
        return sign ? (int64_t)a : -(int64_t)a;
    }
    float flt()
    {
        const int64_t m = integer();
        const int64_t e = (int64_t)u64();
        return (float)std::ldexp((double)m, (int)e);
    }
    std::string bytes()
    {
        const uint64_t k = u64();
        if (!need(k)) return std::string();
        std::string s((const char *)p + i, k);
        i += k;
        return s;
    }
};

bool gzip_bytes(const std::string &in, std::string &out)
{
    z_stream z{};
    if (deflateInit2(&z, Z_DEFAULT_COMPRESSION, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK)
        return false;
    out.resize(deflateBound(&z, in.size()) + 64);
    z.next_in = (Bytef *)in.data();
    z.avail_in = (uInt)in.size();
    z.next_out = (Bytef *)&out[0];
    z.avail_out = (uInt)out.size();
    const int rc = deflate(&z, Z_FINISH);
    out.resize(z.total_out);
    deflateEnd(&z);
    return rc == Z_STREAM_END;
}

bool gunzip_bytes(const unsigned char *p, uint64_t n, std::string &out)
{
    z_stream z{};
    if (inflateInit2(&z, 15 + 32) != Z_OK) return false;
    z.next_in = (Bytef *)p;
    z.avail_in = (uInt)n;
    out.clear();
    char buf[1 << 16];
    int rc;
    do {
        z.next_out = (Bytef *)buf;
        z.avail_out = sizeof buf;
        rc = inflate(&z, Z_NO_FLUSH);
        if (rc != Z_OK && rc != Z_STREAM_END) {
            inflateEnd(&z);
            return false;
        }
        out.append(buf, sizeof buf - z.avail_out);
    } while (rc != Z_STREAM_END);
    inflateEnd(&z);
    return true;
}

/* ------------------------------------------------------------------ */
/* Haskell number formatting (GHC.Float formatRealFloat)                 */
/* ------------------------------------------------------------------ */

/* floatToDigits 10: the shortest digits that read back to x, and e with
 * x = 0.d1d2... 10^e (x > 0) */
void float_digits(double x, bool is_float, std::vector<int> &ds, int &e)
{
    char buf[64];
    std::to_chars_result r = is_float
        ? std::to_chars(buf, buf + sizeof buf, (float)x, std::chars_format::scientific)
        : std::to_chars(buf, buf + sizeof buf, x, std::chars_format::scientific);
    *r.ptr = 0;
    ds.clear();
    const char *c = buf;
    for (; *c && *c != 'e'; ++c)
        if (*c >= '0' && *c <= '9') ds.push_back(*c - '0');
    e = std::atoi(c + 1) + 1;
    while (ds.size() > 1 && ds.back() == 0) ds.pop_back();
}

/* roundTo 10 d is: d digits after rounding, halves to the even neighbour
 * (GHC's roundTo since base 4.6) or up (half_even = false) */
int round_to(int d, const std::vector<int> &is, std::vector<int> &out, bool half_even)
{
    /* f n e xs -> (carry, digits) */
    std::vector<int> res((size_t)d, 0);
    int carry = 0;
    if ((size_t)d < is.size()) {
        const int x = is[(size_t)d];
        bool rest0 = true;
        for (size_t k = (size_t)d + 1; k < is.size(); ++k) rest0 = rest0 && is[k] == 0;
        const bool prev_even = d == 0 ? true : (is[(size_t)d - 1] % 2 == 0);
        if (half_even && x == 5 && prev_even && rest0)
            carry = 0;
        else
            carry = x >= 5 ? 1 : 0;
    }
    for (int k = d - 1; k >= 0; --k) {
        const int v = ((size_t)k < is.size() ? is[(size_t)k] : 0) + carry;
        res[(size_t)k] = v == 10 ? 0 : v;
        carry = v == 10 ? 1 : 0;
    }
    out.clear();
    if (carry) out.push_back(1);
    out.insert(out.end(), res.begin(), res.end());
    return carry;
}

std::string digits_str(const std::vector<int> &v, size_t a, size_t b)
{
    std::string s;
    for (size_t k = a; k < b && k < v.size(); ++k) s.push_back((char)('0' + v[k]));
    return s;
}

/* kind 0: show (FFGeneric, shortest digits); 1: showFFloat (Just dec);
 * 2: showEFloat (Just dec) */
std::string hs_format(double x, bool is_float, int kind, int dec, bool half_even = true)
{
    /* a Float is the value rounded to FP32 (beyond its range: Infinity) */
    if (is_float) x = (double)(float)x;
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x < 0 ? "-Infinity" : "Infinity";
    if (x < 0 || (x == 0 && std::signbit(x))) return "-" + hs_format(-x, is_float, kind, dec, half_even);
    std::vector<int> is;
    int e = 0;
    if (x == 0) {
        is = {0};
        e = 0;
    } else {
        float_digits(x, is_float, is, e);
    }
    const bool zero = is.size() == 1 && is[0] == 0;
    auto fixed_nothing = [&]() {
        if (e <= 0) return "0." + std::string((size_t)-e, '0') + digits_str(is, 0, is.size());
        std::string ip, fp;
        for (int k = 0; k < e; ++k) ip.push_back((char)('0' + ((size_t)k < is.size() ? is[(size_t)k] : 0)));
        fp = digits_str(is, (size_t)e, is.size());
        if (fp.empty()) fp = "0";
        return ip + "." + fp;
    };
    auto expo_nothing = [&]() {
        if (zero) return std::string("0.0e0");
        std::string s(1, (char)('0' + is[0]));
        s += ".";
        s += is.size() == 1 ? std::string("0") : digits_str(is, 1, is.size());
        return s + "e" + std::to_string(e - 1);
    };
    if (kind == 0) return (e < 0 || e > 7) ? expo_nothing() : fixed_nothing();
    if (kind == 1) {
        const int d = dec < 0 ? 0 : dec;
        std::vector<int> r;
        if (e >= 0) {
            const int ei = round_to(d + e, is, r, half_even);
            std::string all = digits_str(r, 0, r.size());
            const size_t cut = (size_t)(e + ei);
            std::string ls = all.substr(0, cut), rs = all.substr(cut);
            if (ls.empty()) ls = "0";
            return rs.empty() ? ls : ls + "." + rs;
        }
        std::vector<int> shifted((size_t)-e, 0);
        shifted.insert(shifted.end(), is.begin(), is.end());
        const int ei = round_to(d, shifted, r, half_even);
        std::vector<int> v;
        if (ei > 0) {
            v = r;
        } else {
            v.push_back(0);
            v.insert(v.end(), r.begin(), r.end());
        }
        std::string s(1, (char)('0' + v[0]));
        if (v.size() > 1) s += "." + digits_str(v, 1, v.size());
        return s;
    }
    /* kind 2 */
    const int d = dec < 1 ? 1 : dec;
    if (zero) return "0." + std::string((size_t)d, '0') + "e0";
    std::vector<int> r;
    const int ei = round_to(d + 1, is, r, half_even);
    if (ei > 0) r.pop_back();
    std::string s(1, (char)('0' + r[0]));
    s += "." + digits_str(r, 1, r.size());
    return s + "e" + std::to_string(e - 1 + ei);
}

/* ------------------------------------------------------------------ */
/* Layout (PrettyPrint.hs:183-196 ppAsRows; Text.PrettyPrint.Boxes)      */
/* ------------------------------------------------------------------ */

/* ppAsRows q: every cell padded to its column's width plus q spaces, the
 * cells of a row joined by one space (hsep; empty cells take none) */
std::string pp_rows(const std::vector<std::vector<std::string>> &rows, int q)
{
    size_t ncol = 0;
    for (auto &r : rows) ncol = std::max(ncol, r.size());
    std::vector<size_t> w(ncol, 0);
    for (auto &r : rows)
        for (size_t c = 0; c < r.size(); ++c) w[c] = std::max(w[c], r[c].size());
    std::string out;
    for (size_t i = 0; i < rows.size(); ++i) {
        std::string line;
        bool any = false;
        for (size_t c = 0; c < rows[i].size(); ++c) {
            std::string cell = rows[i][c] + std::string(w[c] - rows[i][c].size() + (size_t)q, ' ');
            if (cell.empty()) continue;
            if (any) line += " ";
            line += cell;
            any = true;
        }
        if (i) out += "\n";
        out += line;
    }
    return out;
}

/* Boxes' para: words flowed greedily into lines of at most n characters */
std::vector<std::string> para_lines(const std::string &t, size_t n)
{
    std::vector<std::string> lines, words;
    std::string cur;
    for (size_t i = 0; i <= t.size(); ++i) {
        if (i == t.size() || std::isspace((unsigned char)t[i])) {
            if (!cur.empty()) words.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(t[i]);
        }
    }
    std::string line;
    for (auto &wd : words) {
        if (!line.empty() && line.size() + 1 + wd.size() > n) {
            lines.push_back(line.substr(0, n));
            line.clear();
        }
        line += (line.empty() ? "" : " ") + wd;
    }
    if (!line.empty() || lines.empty()) lines.push_back(line.substr(0, n));
    return lines;
}

int put_out(const std::string &s, char *out, uint64_t cap, uint64_t *out_len)
{
    if (out_len) *out_len = s.size();
    if (out && cap) {
        const uint64_t k = std::min<uint64_t>(cap - 1, s.size());
        std::memcpy(out, s.data(), k);
        out[k] = 0;
    }
    return HFX_OK;
}

} // namespace

/* ------------------------------------------------------------------ */
/* Sequence database                                                     */
/* ------------------------------------------------------------------ */

struct hfx_seqdb {
    std::vector<std::string> headers;   /* dbHeader */
    std::vector<uint8_t> ions;          /* dbIon: residue letters */
    std::vector<uint32_t> ion_seg;      /* dbIonSeg [nprot + 1] */
    std::vector<float> r;               /* dbFrag residual */
    std::vector<uint32_t> c, n;         /* dbFrag c, n */
    std::vector<uint32_t> frag_seg;     /* dbFragSeg [nprot + 1] */
    std::string params;                 /* an index file's parameter block */
};

/* mkSeqs . L.lines: '>' starts a protein whose header is the rest of the
 * line; its residues are the following lines up to the next header, joined
 * as they are (lines before the first header are dropped) */
static void parse_fasta(const char *t, uint64_t len, hfx_seqdb &db)
{
    db.headers.clear();
    db.ions.clear();
    db.ion_seg.assign(1, 0);
    bool open = false;
    for_lines(t, len, [&](const char *l, uint64_t k) {
        if (k && l[0] == '>') {
            if (open) db.ion_seg.push_back((uint32_t)db.ions.size());
            db.headers.emplace_back(l + 1, k - 1);
            open = true;
        } else if (open) {
            db.ions.insert(db.ions.end(), (const uint8_t *)l, (const uint8_t *)l + k);
        }
    });
    if (open) db.ion_seg.push_back((uint32_t)db.ions.size());
}

/* makeSeqDB (Fragment.hs:197-233): the digest of every protein, and each
 * protein's first fragment (dbFragSeg) */
static int seqdb_digest(hfx_seqdb &db, const float *mass26, int rule, uint32_t missed,
                        float min_mass, float max_mass)
{
    uint8_t mask[26];
    int rc = hfx_digestion_rule(rule, mask);
    if (rc) return rc;
    const uint32_t np = (uint32_t)db.headers.size();
    db.r.clear();
    db.c.clear();
    db.n.clear();
    db.frag_seg.assign(1, 0);
    std::vector<float> r;
    std::vector<uint32_t> tc, tn;
    for (uint32_t p = 0; p < np; ++p) {
        const uint32_t seg[2] = {db.ion_seg[p], db.ion_seg[p + 1]};
        const uint64_t cap = (uint64_t)(seg[1] - seg[0]) * (missed + 1) + 16;
        r.resize(cap);
        tc.resize(cap);
        tn.resize(cap);
        uint64_t k = 0;
        rc = hfx_digest(db.ions.data(), seg, 1, mass26, mask, missed, min_mass, max_mass, 0, 0, 0,
                        r.data(), tc.data(), tn.data(), cap, &k);
        if (rc) return rc;
        db.r.insert(db.r.end(), r.begin(), r.begin() + (int64_t)k);
        db.c.insert(db.c.end(), tc.begin(), tc.begin() + (int64_t)k);
        db.n.insert(db.n.end(), tn.begin(), tn.begin() + (int64_t)k);
        db.frag_seg.push_back((uint32_t)db.r.size());
    }
    return HFX_OK;
}

extern "C" int hfx_seqdb_from_fasta(const char *path, const char *text, uint64_t len,
                                    const float *h_mass26, int rule, uint32_t missed,
                                    float min_mass, float max_mass, hfx_seqdb **out)
{
    if (!out || !h_mass26 || (!path && !text)) {
        hfx_set_error("hfx_seqdb_from_fasta: null argument");
        return HFX_ERR_ARG;
    }
    std::string buf;
    if (path) {
        if (!read_file(path, buf)) {
            hfx_set_error("hfx_seqdb_from_fasta: cannot read %s", path);
            return HFX_ERR_ARG;
        }
        text = buf.data();
        len = buf.size();
    }
    hfx_seqdb *db = new hfx_seqdb;
    parse_fasta(text, len, *db);
    const int rc = seqdb_digest(*db, h_mass26, rule, missed, min_mass, max_mass);
    if (rc) {
        delete db;
        return rc;
    }
    *out = db;
    return HFX_OK;
}

extern "C" int hfx_seqdb_free(hfx_seqdb *db)
{
    delete db;
    return HFX_OK;
}

extern "C" int hfx_seqdb_info(const hfx_seqdb *db, uint64_t *info)
{
    if (!db || !info) {
        hfx_set_error("hfx_seqdb_info: null argument");
        return HFX_ERR_ARG;
    }
    uint64_t hb = 0;
    for (auto &h : db->headers) hb += h.size();
    info[0] = db->headers.size();
    info[1] = db->ions.size();
    info[2] = hb;
    info[3] = db->r.size();
    info[4] = db->params.size();
    return HFX_OK;
}

extern "C" int hfx_seqdb_copy(const hfx_seqdb *db, uint8_t *h_ions, uint32_t *h_ion_seg,
                              char *h_headers, uint64_t *h_header_off, float *h_r, uint32_t *h_tc,
                              uint32_t *h_tn, uint32_t *h_frag_seg, char *h_params)
{
    if (!db) {
        hfx_set_error("hfx_seqdb_copy: null database");
        return HFX_ERR_ARG;
    }
    copy_out(h_ions, db->ions);
    copy_out(h_ion_seg, db->ion_seg);
    if (h_headers || h_header_off) {
        uint64_t o = 0;
        for (size_t i = 0; i < db->headers.size(); ++i) {
            if (h_header_off) h_header_off[i] = o;
            if (h_headers) copy_out(h_headers + o, db->headers[i]);
            o += db->headers[i].size();
        }
        if (h_header_off) h_header_off[db->headers.size()] = o;
    }
    copy_out(h_r, db->r);
    copy_out(h_tc, db->c);
    copy_out(h_tn, db->n);
    copy_out(h_frag_seg, db->frag_seg);
    copy_out(h_params, db->params);
    return HFX_OK;
}

/* writeIndex: encode (the parameter block as a lazy ByteString), then
 * gzip (encode db) with SequenceDB's instance order: headers, ions, ion
 * segments, residuals, c, n, fragment segments (Fragment.hs:157-166);
 * vectors as Int length + elements, ByteStrings as Int length + bytes,
 * Word32 big-endian, Float as decodeFloat's (Integer, Int) */
extern "C" int hfx_seqdb_write_index(const hfx_seqdb *db, const char *path, const char *params)
{
    if (!db || !path || !params) {
        hfx_set_error("hfx_seqdb_write_index: null argument");
        return HFX_ERR_ARG;
    }
    std::string head, body, gz;
    const std::string pb = params;
    put_int(head, (int64_t)pb.size());
    head += pb;
    put_int(body, (int64_t)db->headers.size());
    for (auto &h : db->headers) {
        put_int(body, (int64_t)h.size());
        body += h;
    }
    put_int(body, (int64_t)db->ions.size());
    body.append((const char *)db->ions.data(), db->ions.size());
    put_int(body, (int64_t)db->ion_seg.size());
    for (uint32_t v : db->ion_seg) put_u32(body, v);
    put_int(body, (int64_t)db->r.size());
    for (float v : db->r) put_float(body, v);
    put_int(body, (int64_t)db->c.size());
    for (uint32_t v : db->c) put_u32(body, v);
    put_int(body, (int64_t)db->n.size());
    for (uint32_t v : db->n) put_u32(body, v);
    put_int(body, (int64_t)db->frag_seg.size());
    for (uint32_t v : db->frag_seg) put_u32(body, v);
    if (!gzip_bytes(body, gz)) {
        hfx_set_error("hfx_seqdb_write_index: deflate failed");
        return HFX_ERR_DATA;
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        hfx_set_error("hfx_seqdb_write_index: cannot open %s", path);
        return HFX_ERR_ARG;
    }
    const bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size() &&
                    std::fwrite(gz.data(), 1, gz.size(), f) == gz.size();
    std::fclose(f);
    if (!ok) {
        hfx_set_error("hfx_seqdb_write_index: write failed");
        return HFX_ERR_DATA;
    }
    return HFX_OK;
}

/* readIndex: the parameter block, then the inflated database */
extern "C" int hfx_seqdb_read_index(const char *path, hfx_seqdb **out)
{
    if (!path || !out) {
        hfx_set_error("hfx_seqdb_read_index: null argument");
        return HFX_ERR_ARG;
    }
    std::string raw, body;
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        hfx_set_error("hfx_seqdb_read_index: cannot open %s", path);
        return HFX_ERR_ARG;
    }
    char buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) raw.append(buf, k);
    std::fclose(f);
    Reader h{(const unsigned char *)raw.data(), raw.size()};
    hfx_seqdb *db = new hfx_seqdb;
    db->params = h.bytes();
    if (!h.ok || !gunzip_bytes(h.p + h.i, h.n - h.i, body)) {
        delete db;
        hfx_set_error("hfx_seqdb_read_index: %s is not an index file", path);
        return HFX_ERR_DATA;
    }
    Reader b{(const unsigned char *)body.data(), body.size()};
    const uint64_t nh = b.u64();
    for (uint64_t i = 0; i < nh && b.ok; ++i) db->headers.push_back(b.bytes());
    const uint64_t ni = b.u64();
    if (b.need(ni)) {
        db->ions.assign(b.p + b.i, b.p + b.i + ni);
        b.i += ni;
    }
    auto u32v = [&](std::vector<uint32_t> &v) {
        const uint64_t m = b.u64();
        if (!b.need(m * 4)) return;
        v.resize(m);
        for (auto &x : v) x = b.u32();
    };
    u32v(db->ion_seg);
    const uint64_t nr = b.u64();
    if (b.ok && nr <= body.size()) {
        db->r.resize(nr);
        for (auto &x : db->r) x = b.flt();
    } else {
        b.ok = false;
    }
    u32v(db->c);
    u32v(db->n);
    u32v(db->frag_seg);
    if (!b.ok || db->ion_seg.size() != db->headers.size() + 1 ||
        db->c.size() != db->r.size() || db->n.size() != db->r.size()) {
        delete db;
        hfx_set_error("hfx_seqdb_read_index: %s is truncated or inconsistent", path);
        return HFX_ERR_DATA;
    }
    *out = db;
    return HFX_OK;
}

/* lookup (Location.hs:42-58): fragment k as "X.PEP.Y" -- flanking residues
 * or '-', the residues tc..tn (tn's own residue kept unless it is past the
 * database, Q28), its protein and (M+H)+ = (r + H) + H2O (host constants) */
extern "C" int hfx_seqdb_lookup(const hfx_seqdb *db, uint32_t k, char *out, uint32_t cap,
                                uint32_t *out_len, uint32_t *out_protein, float *out_mh)
{
    if (!db || k >= db->r.size()) {
        hfx_set_error("hfx_seqdb_lookup: fragment %u out of range", k);
        return HFX_ERR_ARG;
    }
    /* searchVector (tail dbFragSeg) k: the protein whose fragments hold k */
    uint32_t lo = 0, hi = (uint32_t)db->headers.size();
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (db->frag_seg[mid] <= k) lo = mid; else hi = mid;
    }
    const uint32_t p = lo;
    const uint32_t a = db->ion_seg[p], b = db->ion_seg[p + 1];
    const uint32_t c = db->c[k], n = db->n[k];
    const uint64_t nions = db->ions.size();
    const uint32_t ntake = (uint64_t)n < nions ? n - c + 1 : n - c;
    std::string s;
    s.push_back(c > a ? (char)db->ions[c - 1] : '-');
    s.push_back('.');
    s.append((const char *)db->ions.data() + c, ntake);
    s.push_back('.');
    s.push_back(n < b - 1 ? (char)db->ions[(size_t)n + 1] : '-');
    if (out_protein) *out_protein = p;
    if (out_mh) *out_mh = (db->r[k] + 1.007276f) + 18.010565f;
    if (out_len) *out_len = (uint32_t)s.size();
    if (out && cap) {
        const uint32_t m = std::min<uint32_t>(cap - 1, (uint32_t)s.size());
        std::memcpy(out, s.data(), m);
        out[m] = 0;
    }
    return HFX_OK;
}

/* ------------------------------------------------------------------ */
/* Spectra                                                               */
/* ------------------------------------------------------------------ */

/* readF2 pairs, one per line (Util/Parsec.hs); blank lines end nothing */
static bool parse_pair(const char *l, uint64_t k, float &a, float &b)
{
    std::string s(l, k);
    const char *c = s.c_str();
    char *e1 = nullptr, *e2 = nullptr;
    a = std::strtof(c, &e1);
    if (e1 == c) return false;
    b = std::strtof(e1, &e2);
    if (e2 == e1) return false;
    for (const char *q = e2; *q; ++q)
        if (!std::isspace((unsigned char)*q)) return false;
    return true;
}

/* DTA (DTA.hs:36-67): first line MH+ and charge, then m/z intensity pairs;
 * the precursor m/z is (MH + (z - 1) H) / z, z must be an integer */
extern "C" int hfx_dta_parse(const char *text, uint64_t len, float *out_pcr, float *out_charge,
                             float *h_mz, float *h_inten, uint32_t cap, uint32_t *out_npk)
{
    if (!text || !out_pcr || !out_charge || !out_npk) {
        hfx_set_error("hfx_dta_parse: null argument");
        return HFX_ERR_ARG;
    }
    std::vector<std::pair<float, float>> v;
    bool bad = false;
    for_lines(text, len, [&](const char *l, uint64_t k) {
        bool blank = true;
        for (uint64_t i = 0; i < k; ++i) blank = blank && std::isspace((unsigned char)l[i]);
        if (blank) return;
        float a, b;
        if (!parse_pair(l, k, a, b)) bad = true;
        else v.emplace_back(a, b);
    });
    if (bad) {
        hfx_set_error("hfx_dta_parse: a line is not a pair of numbers");
        return HFX_ERR_DATA;
    }
    if (v.empty()) {
        hfx_set_error("hfx_dta_parse: empty spectrum");
        return HFX_ERR_DATA;
    }
    const float m = v[0].first, c = v[0].second;
    if (std::trunc(c) != c) {
        hfx_set_error("hfx_dta_parse: invalid peptide charge state, expecting integer");
        return HFX_ERR_DATA;
    }
    *out_pcr = hfx_dta_precursor(m, c);
    *out_charge = c;
    *out_npk = (uint32_t)(v.size() - 1);
    for (size_t i = 1; i < v.size() && i - 1 < cap; ++i) {
        if (h_mz) h_mz[i - 1] = v[i].first;
        if (h_inten) h_inten[i - 1] = v[i].second;
    }
    return HFX_OK;
}

struct hfx_mgf {
    std::vector<std::string> titles;
    std::vector<float> pcr, charge, mz, inten;
    std::vector<uint32_t> off{0};
};

/* leading signed number of a parameter value (readSigned readFloat) */
static bool lead_number(const std::string &s, float &v)
{
    const char *c = s.c_str();
    while (*c == ' ' || *c == '\t') ++c;
    char *e = nullptr;
    v = std::strtof(c, &e);
    return e != c;
}

/* MGF (MGF.hs:38-100): BEGIN IONS, KEY=value lines (TITLE, CHARGE, PEPMASS
 * kept, PEPMASS's intensity ignored), m/z intensity lines, END IONS; a
 * sample without charge or mass is an error */
extern "C" int hfx_mgf_parse(const char *text, uint64_t len, hfx_mgf **out)
{
    if (!text || !out) {
        hfx_set_error("hfx_mgf_parse: null argument");
        return HFX_ERR_ARG;
    }
    hfx_mgf *m = new hfx_mgf;
    bool in = false, bad = false;
    std::string title;
    float charge = 0, pep = 0;
    for_lines(text, len, [&](const char *l, uint64_t k) {
        if (bad) return;
        std::string s(l, k);
        while (!s.empty() && (s.back() == '\r' || s.back() == ' ')) s.pop_back();
        if (!in) {
            if (s == "BEGIN IONS") {
                in = true;
                title.clear();
                charge = pep = 0;
            } else if (!s.empty()) {
                bad = true;
            }
            return;
        }
        if (s == "END IONS") {
            if (charge == 0 || pep == 0) {
                bad = true;
                return;
            }
            m->titles.push_back(title);
            m->pcr.push_back(pep);
            m->charge.push_back(charge);
            m->off.push_back((uint32_t)m->mz.size());
            in = false;
            return;
        }
        const size_t eq = s.find('=');
        if (eq != std::string::npos && !s.empty() && std::isalpha((unsigned char)s[0])) {
            const std::string key = s.substr(0, eq), val = s.substr(eq + 1);
            if (key == "TITLE") title = val;
            else if (key == "CHARGE") lead_number(val, charge);
            else if (key == "PEPMASS") lead_number(val, pep);
            return;
        }
        if (s.empty()) return;
        float a, b;
        if (!parse_pair(s.data(), s.size(), a, b)) {
            bad = true;
            return;
        }
        m->mz.push_back(a);
        m->inten.push_back(b);
    });
    if (bad || in) {
        delete m;
        hfx_set_error("hfx_mgf_parse: malformed sample (or missing charge state or peptide mass)");
        return HFX_ERR_DATA;
    }
    *out = m;
    return HFX_OK;
}

extern "C" int hfx_mgf_info(const hfx_mgf *m, uint64_t *info)
{
    if (!m || !info) {
        hfx_set_error("hfx_mgf_info: null argument");
        return HFX_ERR_ARG;
    }
    uint64_t tb = 0;
    for (auto &t : m->titles) tb += t.size();
    info[0] = m->pcr.size();
    info[1] = m->mz.size();
    info[2] = tb;
    return HFX_OK;
}

extern "C" int hfx_mgf_copy(const hfx_mgf *m, float *h_pcr, float *h_charge, uint32_t *h_off,
                            float *h_mz, float *h_inten, char *h_titles, uint64_t *h_title_off)
{
    if (!m) {
        hfx_set_error("hfx_mgf_copy: null argument");
        return HFX_ERR_ARG;
    }
    copy_out(h_pcr, m->pcr);
    copy_out(h_charge, m->charge);
    copy_out(h_off, m->off);
    copy_out(h_mz, m->mz);
    copy_out(h_inten, m->inten);
    uint64_t o = 0;
    for (size_t i = 0; i < m->titles.size(); ++i) {
        if (h_title_off) h_title_off[i] = o;
        if (h_titles) copy_out(h_titles + o, m->titles[i]);
        o += m->titles[i].size();
    }
    if (h_title_off) h_title_off[m->titles.size()] = o;
    return HFX_OK;
}

extern "C" int hfx_mgf_free(hfx_mgf *m)
{
    delete m;
    return HFX_OK;
}

/* ------------------------------------------------------------------ */
/* Result tables                                                         */
/* ------------------------------------------------------------------ */

/* the digestion rules' description lines (Config.hs:416-430: output text,
 * printed by printConfig without its first word) */
static const char *const RULE_TEXT[14] = {
    "  0:  No enzyme              0      -           -",
    "  1.  Trypsin                1      KR          P",
    "  2.  Chymotrypsin           1      FWY         P",
    "  3.  Clostripain            1      R           -",
    "  4.  Cyanogen_Bromide       1      M           -",
    "  5.  IodosoBenzoate         1      W           -",
    "  6.  Proline_Endopept       1      P           -",
    "  7.  Staph_Protease         1      E           -",
    "  8.  Trypsin_K              1      K           P",
    "  9.  Trypsin_R              1      R           P",
    "  10. AspN                   0      D           -",
    "  11. Cymotryp/Modified      1      FWYL        P",
    "  12. Elastase               1      ALIV        P",
    "  13. Elastase/Tryp/Chymo    1      ALIVKRWFY   P"};

extern "C" const char *hfx_digestion_rule_name(int rule)
{
    return rule >= 0 && rule < 14 ? RULE_TEXT[rule] : nullptr;
}

/* kind 0: show; 1: showFFloat (Just digits); 2: showEFloat (Just digits);
 * is_float: the value is a Float (its own shortest digits) */
extern "C" int hfx_format_number(double v, int is_float, int kind, int digits, char *out,
                                 uint32_t cap)
{
    if (kind < 0 || kind > 2) {
        hfx_set_error("hfx_format_number: kind %d", kind);
        return HFX_ERR_ARG;
    }
    const std::string s = hs_format(v, is_float != 0, kind, digits);
    uint64_t n = 0;
    put_out(s, out, cap, &n);
    if (!out || s.size() >= cap) {
        hfx_set_error("hfx_format_number: the text needs %llu bytes, the buffer holds %u",
                      (unsigned long long)s.size() + 1, cap);
        return HFX_ERR_ARG;
    }
    return HFX_OK;
}

/* ppAsRows q over nrows x ncols cells (row-major, NUL-terminated strings) */
extern "C" int hfx_format_table(const char *const *cells, uint32_t nrows, uint32_t ncols, int q,
                                char *out, uint64_t cap, uint64_t *out_len)
{
    if (!cells && nrows * ncols) {
        hfx_set_error("hfx_format_table: null cells");
        return HFX_ERR_ARG;
    }
    if (q < 0) {
        hfx_set_error("hfx_format_table: negative column spacing %d", q);
        return HFX_ERR_ARG;
    }
    std::vector<std::vector<std::string>> rows(nrows, std::vector<std::string>(ncols));
    for (uint32_t i = 0; i < nrows; ++i)
        for (uint32_t j = 0; j < ncols; ++j) rows[i][j] = cells[(size_t)i * ncols + j];
    return put_out(pp_rows(rows, q), out, cap, out_len);
}

/* printResultsDetail (PrettyPrint.hs:137-144): "n." right-aligned in 3
 * columns, two spaces, the protein's full header flowed into lines of 95;
 * every line padded to the widest */
extern "C" int hfx_format_detail(const char *const *headers, uint32_t n, char *out, uint64_t cap,
                                 uint64_t *out_len)
{
    std::vector<std::string> lines;
    size_t w = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const std::vector<std::string> pl = para_lines(headers[i] ? headers[i] : "", 95);
        size_t pw = 0;
        for (auto &l : pl) pw = std::max(pw, l.size());
        std::string num = std::to_string(i + 1) + ".";
        if (num.size() < 3) num = std::string(3 - num.size(), ' ') + num;
        for (size_t k = 0; k < pl.size(); ++k) {
            std::string l = (k == 0 ? num : std::string(num.size(), ' ')) + "  " + pl[k];
            l += std::string(num.size() + 2 + pw - l.size(), ' ');
            w = std::max(w, l.size());
            lines.push_back(l);
        }
    }
    std::string s;
    for (auto &l : lines) s += l + std::string(w - l.size(), ' ') + "\n";
    return put_out(s, out, cap, out_len);
}

/* printScanResults for one scan (PrettyPrint.hs:83-180): printConfig, the
 * table of the first nshow matches (printResults: #, (M+H)+, deltCn,
 * XCorr, Ions, Reference, Peptide, E-value; deltCn = (s0 - s) / s0 in
 * Float against the first match), and the headers of the first ndetail
 * (printResultsDetail); each block followed by a blank line. */
extern "C" int hfx_format_scan(const hfx_scan_header *hdr, const hfx_match_row *rows, uint32_t n,
                               uint32_t nshow, uint32_t ndetail, char *out, uint64_t cap,
                               uint64_t *out_len)
{
    if (!hdr || (!rows && n)) {
        hfx_set_error("hfx_format_scan: null argument");
        return HFX_ERR_ARG;
    }
    std::string s;
    /* printConfig: ppAsRows 0 with "::" between the two columns */
    {
        const float z = hdr->charge;
        const float mass = (hdr->precursor * z) - ((z * 1.007276f) - 1.0f);
        const long zi = std::lround(std::nearbyint(z));
        std::string signed_int = std::string(1, zi < 0 ? '-' : '+') + std::to_string(zi);
        std::string enz;
        {
            /* words . snd . digestionRule, without its first word */
            std::vector<std::string> ws;
            std::string cur;
            const std::string e = hdr->enzyme ? hdr->enzyme : "";
            for (size_t i = 0; i <= e.size(); ++i) {
                if (i == e.size() || std::isspace((unsigned char)e[i])) {
                    if (!cur.empty()) ws.push_back(cur);
                    cur.clear();
                } else {
                    cur.push_back(e[i]);
                }
            }
            for (size_t i = 1; i < ws.size(); ++i) enz += (i > 1 ? " " : "") + ws[i];
        }
        std::vector<std::vector<std::string>> cfg = {
            {"Title", "::", hdr->title ? hdr->title : ""},
            {"Spectrum", "::", hdr->spectrum_file ? hdr->spectrum_file : ""},
            {"Database", "::", hdr->database ? hdr->database : ""},
            {"Enzyme", "::", enz},
            {"(M+H)+ Mass", "::",
             hs_format(mass, true, 0, 0) + " ~ " + hs_format(hdr->tolerance, true, 0, 0) + " (" +
                 signed_int + ")"}};
        s += pp_rows(cfg, 0) + "\n\n";
    }
    const uint32_t ns = std::min(nshow, n);
    std::vector<std::vector<std::string>> t = {
        {" # ", " (M+H)+  ", "deltCn", "XCorr", "Ions", "Reference", "Peptide", "E-value"},
        {"---", "---------", "------", "-----", "----", "---------", "-------", "-------"}};
    const float s0 = n ? rows[0].xcorr : 0.0f;
    for (uint32_t i = 0; i < ns; ++i) {
        const hfx_match_row &m = rows[i];
        const float dcn = (s0 - m.xcorr) / s0;
        t.push_back({" " + std::to_string(i + 1) + ".", hs_format(m.mh, true, 1, 4),
                     hs_format((double)dcn, false, 1, 4), hs_format((double)m.xcorr, false, 1, 4),
                     std::to_string(m.ions_matched) + "/" + std::to_string(m.ions_total),
                     m.reference ? m.reference : "", m.peptide ? m.peptide : "",
                     hs_format(m.evalue, false, 2, 4)});
    }
    s += pp_rows(t, 1) + "\n\n";
    const uint32_t nd = std::min(ndetail, n);
    std::vector<const char *> hs(nd);
    for (uint32_t i = 0; i < nd; ++i) hs[i] = rows[i].header;
    uint64_t dl = 0;
    std::string det;
    {
        hfx_format_detail(hs.data(), nd, nullptr, 0, &dl);
        det.resize(dl + 1);
        hfx_format_detail(hs.data(), nd, &det[0], dl + 1, &dl);
        det.resize(dl);
    }
    s += det + "\n";
    return put_out(s, out, cap, out_len);
}
