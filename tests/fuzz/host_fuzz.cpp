/*
 * Randomised robustness run of the library's host code (the file parsers,
 * the index reader, the formatters, digestion and the modification table),
 * built with AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_host_fuzz.py: every call must return (an error code is fine)
 * without a memory error, undefined behaviour or an unbounded loop.
 *
 *   host_fuzz <iterations> <seed> <scratch file>
 *
 * Inputs are valid DTA / MGF / FASTA texts with random edits (bytes
 * replaced, removed, inserted, the text cut short), index files written by
 * the library and then damaged, and random cells, rows, capacities and
 * numbers for the formatters.
 */
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/hfx_hip.h"

/* the library's error sink (defined in its runtime, which is not linked) */
static char g_err[512];
void hfx_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

namespace {

std::mt19937_64 rng;

uint64_t pick(uint64_t n) { return n ? rng() % n : 0; }

std::string edit(const std::string &s)
{
    static const char alpha[] = "0123456789.-+eE \n\r\tABCDEFGHIKLMNPQRSTVWYZ=>#";
    std::string t = s;
    const int edits = 1 + (int)pick(pick(2) ? 2 : 8);
    for (int m = 0; m < edits; ++m) {
        if (t.empty()) {
            t.push_back('A');
            continue;
        }
        const size_t p = pick(t.size());
        switch (pick(5)) {
        case 0: t[p] = alpha[pick(sizeof alpha - 1)]; break;
        case 1: t.erase(p, 1 + pick(8)); break;
        case 2: t.insert(p, 1 + pick(4), alpha[pick(sizeof alpha - 1)]); break;
        case 3: t.resize(p); break;
        default: t[p] = (char)(rng() & 255); break;
        }
    }
    return t;
}

std::string word(int maxlen)
{
    std::string s;
    const int n = (int)pick(maxlen + 1);
    for (int i = 0; i < n; ++i) {
        const int c = (int)pick(40);
        s.push_back(c < 26 ? (char)('A' + c) : c < 30 ? ' ' : c < 32 ? '\n' : c < 36 ? '|' : '.');
    }
    return s;
}

void spectra_files()
{
    static const std::string dta = "1500.75 2\n100.1 5.0\n200.2 7.5\n300.3 1e2\n450.5 3\n";
    static const std::string mgf =
        "BEGIN IONS\nTITLE=scan 1\nCHARGE=2+\nPEPMASS=750.5 1000\n100.1 5\n200.2 6\nEND IONS\n"
        "BEGIN IONS\nTITLE=scan 2\nCHARGE=3+\nPEPMASS=512.25\n110.1 5\n220.2 6\n330 1\nEND IONS\n";
    {
        const std::string t = edit(dta);
        std::vector<float> mz(4096), in(4096);
        float pcr, ch;
        uint32_t npk;
        hfx_dta_parse(t.data(), t.size(), &pcr, &ch, mz.data(), in.data(),
                      pick(2) ? 4096 : (uint32_t)pick(8), &npk);
    }
    const std::string t = edit(mgf);
    hfx_mgf *m = nullptr;
    if (hfx_mgf_parse(t.data(), t.size(), &m) == HFX_OK && m) {
        uint64_t info[3];
        hfx_mgf_info(m, info);
        std::vector<float> pcr(info[0] + 1), ch(info[0] + 1), mz(info[1] + 1), in(info[1] + 1);
        std::vector<uint32_t> off(info[0] + 1);
        std::vector<char> titles(info[2] + 1);
        std::vector<uint64_t> toff(info[0] + 1);
        hfx_mgf_copy(m, pcr.data(), ch.data(), off.data(), mz.data(), in.data(), titles.data(),
                     toff.data());
        hfx_mgf_free(m);
    }
}

void databases(const char *scratch)
{
    static const std::string fasta = ">prot1 desc\nMKWVTFISLLLLFSSAYSRGVFRRDTHKSEIAHRFKDLGE\n"
                                     ">prot2\nACDEFGHIKLMNPQRSTVWY\nKRKRPKPR\n>\n\n>p3\nMPEPTIDEK\n";
    float mass26[26];
    for (int i = 0; i < 26; ++i) mass26[i] = 50.0f + 3.7f * (float)i;
    const std::string t = edit(fasta);
    hfx_seqdb *db = nullptr;
    if (hfx_seqdb_from_fasta(nullptr, t.data(), t.size(), mass26, (int)pick(14), (uint32_t)pick(3),
                             100.0f, 5000.0f, &db) != HFX_OK || !db)
        return;
    uint64_t info[5];
    hfx_seqdb_info(db, info);
    char out[256];
    uint32_t len, prot;
    float mh;
    for (uint64_t k = 0; k < info[3] + 2; ++k)
        hfx_seqdb_lookup(db, (uint32_t)k, out, pick(2) ? sizeof out : 8, &len, &prot, &mh);
    if (hfx_seqdb_write_index(db, scratch, "params x=1") == HFX_OK) {
        std::string bytes;
        if (FILE *f = fopen(scratch, "rb")) {
            char buf[4096];
            size_t n;
            while ((n = fread(buf, 1, sizeof buf, f)) > 0) bytes.append(buf, n);
            fclose(f);
        }
        for (int r = 0; r < 4 && !bytes.empty(); ++r) {
            std::string b = bytes;
            const int edits = 1 + (int)pick(4);
            for (int q = 0; q < edits && !b.empty(); ++q) {
                const size_t p = pick(b.size());
                if (pick(3) == 0)
                    b.resize(p);
                else
                    b[p] = (char)(rng() & 255);
            }
            if (FILE *g = fopen(scratch, "wb")) {
                fwrite(b.data(), 1, b.size(), g);
                fclose(g);
            }
            hfx_seqdb *back = nullptr;
            if (hfx_seqdb_read_index(scratch, &back) == HFX_OK && back) {
                uint64_t i2[5];
                hfx_seqdb_info(back, i2);
                for (uint64_t k = 0; k < i2[3] + 1 && k < 64; ++k)
                    hfx_seqdb_lookup(back, (uint32_t)k, out, sizeof out, &len, &prot, &mh);
                hfx_seqdb_free(back);
            }
        }
    }
    hfx_seqdb_free(db);
}

void formatters()
{
    {
        char out[64];
        union {
            uint64_t u;
            double d;
        } x;
        x.u = rng();
        hfx_format_number(x.d, (int)pick(2), (int)pick(3), (int)pick(12) - 1, out,
                          pick(2) ? sizeof out : 4);
    }
    const uint32_t nr = (uint32_t)pick(6), nc = (uint32_t)pick(6);
    std::vector<std::string> cells(nr * nc);
    for (auto &c : cells) c = word(12);
    std::vector<const char *> cp(nr * nc + 1);
    for (size_t i = 0; i < cells.size(); ++i) cp[i] = cells[i].c_str();
    const uint64_t cap = pick(3) == 0 ? pick(16) : 4096;
    std::vector<char> out(cap ? cap : 1);
    uint64_t ol = 0;
    hfx_format_table(cp.data(), nr, nc, (int)pick(4) - 1, cap ? out.data() : nullptr, cap, &ol);
    std::vector<const char *> hp(nr + 1);
    for (uint32_t i = 0; i < nr; ++i) hp[i] = cells.empty() ? "" : cp[i % cells.size()];
    hfx_format_detail(hp.data(), nr, cap ? out.data() : nullptr, cap, &ol);

    const std::string title = word(10), file = word(10), dbn = word(10);
    const char *enzyme = hfx_digestion_rule_name((int)pick(16) - 1);
    hfx_scan_header h{title.c_str(), file.c_str(), dbn.c_str(), enzyme ? enzyme : "",
                      (float)pick(5000) / 3.0f, (float)pick(5), (float)pick(10) / 7.0f};
    const uint32_t n = (uint32_t)pick(6);
    std::vector<std::string> strs(3 * n);
    for (auto &s : strs) s = word(20);
    std::vector<hfx_match_row> rows(n + 1);
    for (uint32_t i = 0; i < n; ++i) {
        union {
            uint64_t u;
            double d;
        } e;
        e.u = rng();
        rows[i] = hfx_match_row{(float)pick(100000) / 7.0f, ((float)pick(2000) - 1000.0f) / 300.0f,
                                (uint32_t)pick(50), (uint32_t)pick(50),
                                pick(2) ? e.d : 1e-3 * (double)pick(1000), strs[3 * i].c_str(),
                                strs[3 * i + 1].c_str(), strs[3 * i + 2].c_str()};
    }
    const uint64_t scap = pick(3) == 0 ? pick(32) : 8192;
    std::vector<char> sout(scap ? scap : 1);
    hfx_format_scan(&h, rows.data(), n, (uint32_t)pick(8), (uint32_t)pick(8),
                    scap ? sout.data() : nullptr, scap, &ol);
}

void digestion_and_mods()
{
    const uint32_t np = (uint32_t)pick(5);
    std::vector<uint32_t> seg(np + 1, 0);
    std::string seq;
    for (uint32_t p = 0; p < np; ++p) {
        seq += word(60);
        seg[p + 1] = (uint32_t)seq.size();
    }
    float mass26[26];
    for (int i = 0; i < 26; ++i) mass26[i] = 50.0f + 4.1f * (float)i;
    uint8_t rule[26];
    hfx_digestion_rule((int)pick(14), rule);
    const uint64_t cap = pick(2) ? pick(4) : 4096;
    std::vector<float> r(cap + 1);
    std::vector<uint32_t> tc(cap + 1), tn(cap + 1);
    uint64_t got = 0;
    hfx_digest((const uint8_t *)seq.data(), seg.data(), np, mass26, rule, (uint32_t)pick(4), 0.0f,
               (float)pick(5000), (int)pick(2), (uint32_t)pick(4), (uint32_t)pick(100), r.data(),
               tc.data(), tn.data(), cap, &got);

    float var26[26];
    for (int i = 0; i < 26; ++i) var26[i] = pick(4) == 0 ? ((float)pick(200) - 100.0f) / 3.0f : 0.0f;
    const uint32_t max_ma = (uint32_t)pick(pick(8) == 0 ? 64 : 6);
    const uint32_t mcap = pick(2) ? (uint32_t)pick(4) : 1024;
    uint8_t ma[26];
    float mam[26];
    uint32_t nma = 0, nmod = 0;
    std::vector<uint32_t> cnt((size_t)(mcap + 1) * 26), sum(mcap + 1);
    std::vector<float> delta(mcap + 1);
    hfx_mod_info(var26, max_ma, ma, mam, &nma, cnt.data(), sum.data(), delta.data(), mcap, &nmod);
}

} // namespace

int main(int argc, char **argv)
{
    if (argc != 4) {
        fprintf(stderr, "usage: host_fuzz <iterations> <seed> <scratch file>\n");
        return 2;
    }
    const long iters = atol(argv[1]);
    rng.seed(strtoull(argv[2], nullptr, 10));
    for (long it = 0; it < iters; ++it) {
        spectra_files();
        if (it % 20 == 0) databases(argv[3]);
        formatters();
        if (it % 4 == 0) digestion_and_mods();
    }
    printf("host_fuzz: %ld iterations\n", iters);
    return 0;
}
