/*
 * hfx_host.cpp -- host-side database and configuration helpers of the C ABI.
 *
 * These restate the Haskell host functions whose outputs the kernels consume,
 * so that a caller without the Haskell driver (the C++ search driver, tests,
 * benchmarks) can build the same device arrays:
 *   hfx_mass_table      Config.hs:266-272 initializeAAMasses, Mass.hs:147-215
 *   hfx_digestion_rule  Config.hs:415-431 getDigestionRule
 *   hfx_digest          Fragment.hs:363-405 digest / fragment / splice
 *   hfx_mod_info        Fragment.hs:260-289 makeModInfo, Config.hs:134-140
 *   hfx_ms2_mass        Spectrum/Data.hs:58-59 ms2Mass
 *   hfx_dta_precursor   Spectrum/DTA.hs:51
 */
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/hfx_hip.h"

extern void hfx_set_error(const char *fmt, ...);

namespace {

/* host constants, Mass.hs:19-27 */
constexpr float kHostH2O = 18.010565f;
constexpr float kHostH = 1.007276f;

/* residue masses A..Z, Mass.hs:147-215 */
constexpr float kMono[26] = {71.037114f,  114.53494f,  103.009185f, 115.026943f, 129.042593f,
                             147.068414f, 57.021464f,  137.058912f, 113.084064f, 113.16472f,
                             128.094963f, 113.084064f, 131.040485f, 114.042927f, 114.07931f,
                             97.052764f,  128.058578f, 156.101111f, 87.032028f,  101.047679f,
                             150.04344f,  99.068414f,  186.079313f, 113.08406f,  163.06332f,
                             128.55059f};
constexpr float kAvg[26] = {71.0788f,  114.5962f, 103.1388f, 115.0886f, 129.1155f, 147.1766f,
                            57.0519f,  137.1411f, 113.1594f, 113.240178f, 128.1741f, 113.1594f,
                            131.1926f, 114.1038f, 114.1472f, 97.1167f,  128.1307f, 156.1875f,
                            87.0782f,  101.1051f, 150.0388f, 99.1326f,  186.2132f, 113.1594f,
                            163.1760f, 128.6231f};

/* cleavage residues per rule, Config.hs:415-431 */
const char *const kRules[14] = {"",    "KR",   "FWY",  "R",    "M",   "W",    "P",
                                "E",   "K",    "R",    "D",    "FWYL", "ALIV", "ALIVKRWFY"};

struct Span {
    uint32_t first, last_excl; /* residues [first, last_excl) */
    uint32_t tn;               /* reference tn: last index, or end for the final piece */
    float mass;
};

} // namespace

extern "C" int hfx_mass_table(const float *static_add26, int mono, float *out26)
{
    if (!out26) {
        hfx_set_error("hfx_mass_table: null output");
        return HFX_ERR_ARG;
    }
    const float *base = mono ? kMono : kAvg;
    for (int i = 0; i < 26; ++i) out26[i] = (static_add26 ? static_add26[i] : 0.0f) + base[i];
    return HFX_OK;
}

extern "C" int hfx_digestion_rule(int rule, uint8_t *mask26)
{
    if (rule < 0 || rule > 13 || !mask26) {
        hfx_set_error("getDigestionRule: unknown enzyme digestion rule %d", rule);
        return HFX_ERR_ARG;
    }
    memset(mask26, 0, 26);
    for (const char *p = kRules[rule]; *p; ++p) mask26[*p - 'A'] = 1;
    return HFX_OK;
}

extern "C" float hfx_ms2_mass(float pcr, float charge)
{
    const float a = pcr * charge;
    const float b = (charge - 1.0f) * kHostH;
    return (a - b) - (kHostH + kHostH2O);
}

extern "C" float hfx_dta_precursor(float mh, float charge)
{
    return (mh + (charge - 1.0f) * kHostH) / charge;
}

/*
 * Digest every protein seq[seg[p] .. seg[p+1]) (residue letters A..Z).
 * Pieces end after each residue in the rule (with p_rule: not when followed
 * by P); a piece's mass is the left-to-right FP32 sum of its residue masses;
 * each piece is emitted followed by its fusions with the next 1..missed pieces
 * (mass a + b in FP32); a peptide is kept when min <= (m + H2O) + H <= max and,
 * as harness extensions, its residue count lies in [min_len, max_len] (0 = no
 * bound).  Outputs (r, tc, tn) in emission order; returns HFX_ERR_ARG when cap
 * is exceeded (*out_n then holds the count needed).
 */
extern "C" int hfx_digest(const uint8_t *seq, const uint32_t *seg, uint32_t nprot,
                          const float *mass26, const uint8_t *rule26, uint32_t missed,
                          float min_mass, float max_mass, int p_rule, uint32_t min_len,
                          uint32_t max_len, float *out_r, uint32_t *out_tc, uint32_t *out_tn,
                          uint64_t cap, uint64_t *out_n)
{
    if (!seq || !seg || !mass26 || !rule26 || !out_n) {
        hfx_set_error("hfx_digest: null argument");
        return HFX_ERR_ARG;
    }
    std::vector<Span> spans;
    uint64_t n = 0;
    bool overflow = false;
    for (uint32_t p = 0; p < nprot; ++p) {
        const uint32_t b = seg[p], e = seg[p + 1];
        spans.clear();
        uint32_t start = b;
        float acc = 0.0f;
        for (uint32_t i = b; i < e; ++i) {
            const uint8_t ch = seq[i];
            if (ch < 'A' || ch > 'Z') {
                hfx_set_error("hfx_digest: invalid residue 0x%02x at %u", ch, i);
                return HFX_ERR_DATA;
            }
            acc = acc + mass26[ch - 'A'];
            bool cut = rule26[ch - 'A'] != 0;
            if (cut && p_rule && i + 1 < e && seq[i + 1] == 'P') cut = false;
            if (cut) {
                spans.push_back(Span{start, i + 1, i, acc});
                start = i + 1;
                acc = 0.0f;
            }
        }
        if (start < e) spans.push_back(Span{start, e, e, acc});
        const size_t np = spans.size();
        for (size_t k = 0; k < np; ++k) {
            float m = spans[k].mass;
            for (size_t j = k; j < np && j <= k + missed; ++j) {
                if (j > k) m = m + spans[j].mass;
                const float full = (m + kHostH2O) + kHostH;
                const uint32_t nres = spans[j].last_excl - spans[k].first;
                bool keep = (min_mass <= full) && (full <= max_mass);
                if (min_len && nres < min_len) keep = false;
                if (max_len && nres > max_len) keep = false;
                if (!keep) continue;
                if (n < cap && out_r && out_tc && out_tn) {
                    out_r[n] = m;
                    out_tc[n] = spans[k].first;
                    out_tn[n] = spans[j].tn;
                } else {
                    overflow = true;
                }
                ++n;
            }
        }
    }
    *out_n = n;
    if (overflow) {
        hfx_set_error("hfx_digest: %llu peptides exceed capacity %llu", (unsigned long long)n,
                      (unsigned long long)cap);
        return HFX_ERR_ARG;
    }
    return HFX_OK;
}

/*
 * Modification table: modable acids = letters with a non-zero variable mass in
 * alphabetical order; modification vectors = count vectors over those acids,
 * each count 0..max_ma, in lexicographic order with the last acid varying
 * fastest (addModDimensions), the zero vector dropped, sum < max_ma kept.
 */
extern "C" int hfx_mod_info(const float *var26, uint32_t max_ma, uint8_t *out_ma,
                            float *out_ma_mass, uint32_t *out_num_ma, uint32_t *out_mod_ma_count,
                            uint32_t *out_mod_sum, float *out_mod_delta, uint32_t cap_mod,
                            uint32_t *out_num_mod)
{
    if (!var26 || !out_ma || !out_ma_mass || !out_num_ma || !out_num_mod) {
        hfx_set_error("hfx_mod_info: null argument");
        return HFX_ERR_ARG;
    }
    uint32_t na = 0;
    for (int c = 0; c < 26; ++c)
        if (var26[c] != 0.0f) {
            out_ma[na] = (uint8_t)('A' + c);
            out_ma_mass[na] = var26[c];
            ++na;
        }
    *out_num_ma = na;
    *out_num_mod = 0;
    if (na == 0) return HFX_OK;
    /* the count first: vectors of na counts with sum <= max_ma - 1 number
     * C(na + max_ma - 1, k) with k = min(na, max_ma - 1), the zero vector
     * among them; C(n, i) rises with i up to k <= n / 2, so it saturates at
     * 2^40, far beyond any table a caller can hold, once a partial product
     * passes that */
    uint64_t total = 0;
    if (max_ma > 0) {
        const uint64_t n = (uint64_t)max_ma - 1 + na, k = std::min<uint64_t>(na, max_ma - 1);
        unsigned __int128 c = 1;
        for (uint64_t i = 0; i < k && c <= ((uint64_t)1 << 40); ++i) c = c * (n - i) / (i + 1);
        total = c > ((uint64_t)1 << 40) ? (uint64_t)1 << 40 : (uint64_t)c - 1;
    }
    if (total > cap_mod) {
        *out_num_mod = total > UINT32_MAX ? UINT32_MAX : (uint32_t)total;
        hfx_set_error("hfx_mod_info: %llu modifications exceed capacity %u",
                      (unsigned long long)total, cap_mod);
        return HFX_ERR_ARG;
    }
    /* lexicographic order, the last acid fastest, over the vectors with sum
     * < max_ma only: the next one raises the rightmost count whose raise
     * (with every count right of it cleared) keeps the sum below max_ma */
    std::vector<uint32_t> v(na, 0);
    uint32_t nm = 0, sum = 0;
    for (;;) {
        int d = (int)na - 1;
        while (d >= 0 && sum + 1 >= max_ma) {
            sum -= v[d];
            v[d--] = 0;
        }
        if (d < 0) break;
        ++v[d];
        ++sum;
        if (nm < cap_mod && out_mod_ma_count && out_mod_sum && out_mod_delta) {
            float delta = 0.0f;
            for (uint32_t a = 0; a < na; ++a) {
                out_mod_ma_count[(size_t)nm * na + a] = v[a];
                delta = delta + out_ma_mass[a] * (float)v[a];
            }
            out_mod_sum[nm] = sum;
            out_mod_delta[nm] = delta;
        }
        ++nm;
    }
    *out_num_mod = nm;
    if (nm > cap_mod) {
        hfx_set_error("hfx_mod_info: %u modifications exceed capacity %u", nm, cap_mod);
        return HFX_ERR_ARG;
    }
    return HFX_OK;
}
