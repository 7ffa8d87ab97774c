/*
 * hfx_oracle.cpp -- CPU restatement of HFX's GPU search path (see hfx_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never by the product.
 *
 * Built with -O2 -ffp-contract=off so that every FP32 expression below is
 * evaluated exactly as written (one IEEE rounding per operation).
 */
#include "hfx_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------
 * Constants
 * ------------------------------------------------------------------- */

/* Device constants used by the ion ladders (ion_series.h:16-26). */
static const float DEV_BIN_WIDTH_MONO = 1.0005079f;
static const float DEV_H2O = 18.01056f;
static const float DEV_NH3 = 17.02655f;
static const float DEV_CO = 27.9949f;
static const float DEV_H = 1.0078246f;

/* Host constants used for precursor, window and display (Mass.hs:19-27). */
static const float HOST_H2O = 18.010565f;
static const float HOST_H = 1.007276f;

/* Residue masses, A..Z (Mass.hs:147-215). */
static const float MONO[26] = {
    71.037114f,  /* A */ 114.53494f,  /* B */ 103.009185f, /* C */ 115.026943f, /* D */
    129.042593f, /* E */ 147.068414f, /* F */ 57.021464f,  /* G */ 137.058912f, /* H */
    113.084064f, /* I */ 113.16472f,  /* J */ 128.094963f, /* K */ 113.084064f, /* L */
    131.040485f, /* M */ 114.042927f, /* N */ 114.07931f,  /* O */ 97.052764f,  /* P */
    128.058578f, /* Q */ 156.101111f, /* R */ 87.032028f,  /* S */ 101.047679f, /* T */
    150.04344f,  /* U */ 99.068414f,  /* V */ 186.079313f, /* W */ 113.08406f,  /* X */
    163.06332f,  /* Y */ 128.55059f   /* Z */
};
static const float AVG[26] = {
    71.0788f,  /* A */ 114.5962f,   /* B */ 103.1388f, /* C */ 115.0886f, /* D */
    129.1155f, /* E */ 147.1766f,   /* F */ 57.0519f,  /* G */ 137.1411f, /* H */
    113.1594f, /* I */ 113.240178f, /* J */ 128.1741f, /* K */ 113.1594f, /* L */
    131.1926f, /* M */ 114.1038f,   /* N */ 114.1472f, /* O */ 97.1167f,  /* P */
    128.1307f, /* Q */ 156.1875f,   /* R */ 87.0782f,  /* S */ 101.1051f, /* T */
    150.0388f, /* U */ 99.1326f,    /* V */ 186.2132f, /* W */ 113.1594f, /* X */
    163.1760f, /* Y */ 128.6231f    /* Z */
};

/* ion byte packing (utils.h:92-101) */
static inline uint8_t acid_idx(uint8_t ion) { return ion & 0x1Fu; }
static inline uint8_t acid_char(uint8_t ion) { return (uint8_t)((ion & 0x1Fu) + 'A'); }

/* ---------------------------------------------------------------------
 * Mass table and configuration
 * ------------------------------------------------------------------- */

/* Config.hs:266-272 initializeAAMasses: table = static_add + base mass
 * (U.accum (+) with the static modification as the accumulator seed). */
void orc_mass_table(const float *static_add26, int mono, float *out26)
{
    const float *base = mono ? MONO : AVG;
    for (int i = 0; i < 26; ++i) {
        float s = static_add26 ? static_add26[i] : 0.0f;
        out26[i] = s + base[i];
    }
}

/* Spectrum/Data.hs:58-59 ms2Mass: (pcr*z) - ((z-1)*H) - (H + H2O), host constants. */
float orc_ms2_mass(float pcr, float charge)
{
    float a = pcr * charge;
    float b = (charge - 1.0f) * HOST_H;
    float c = HOST_H + HOST_H2O;
    return (a - b) - c;
}

/* Spectrum/DTA.hs:51: precursor m/z from the DTA (M+H)+ line. */
float orc_dta_precursor(float mh, float charge)
{
    return (mh + (charge - 1.0f) * HOST_H) / charge;
}

/* Fragment.hs:260-289 makeModInfo and Config.hs:134-140 getMA/getMA_Mass.
 * Modable acids are the letters with a non-zero variable mass, alphabetical.
 * Modification vectors enumerate every count vector with entries 0..max_ma in
 * lexicographic order (last acid fastest, the order addModDimensions
 * produces), drop the all-zero vector (tail), keep those whose sum is
 * strictly below max_ma.  mod_delta = foldl' (+) 0 (m_i * c_i). */
int orc_mod_info(const float *var26, uint32_t max_ma,
                 uint8_t *out_ma, float *out_ma_mass, uint32_t *out_num_ma,
                 uint32_t *out_mod_ma_count, uint32_t *out_mod_sum,
                 float *out_mod_delta, uint32_t cap_mod, uint32_t *out_num_mod)
{
    uint32_t num_ma = 0;
    for (int i = 0; i < 26; ++i) {
        if (var26[i] != 0.0f) {
            out_ma[num_ma] = (uint8_t)('A' + i);
            out_ma_mass[num_ma] = var26[i];
            ++num_ma;
        }
    }
    *out_num_ma = num_ma;
    *out_num_mod = 0;
    if (num_ma < 1) return 0;

    std::vector<uint32_t> c(num_ma, 0);
    uint32_t nmod = 0;
    bool first = true;
    for (;;) {
        if (!first) {
            uint32_t sum = 0;
            for (uint32_t i = 0; i < num_ma; ++i) sum += c[i];
            if (sum < max_ma) {
                if (nmod >= cap_mod) return -1;
                float delta = 0.0f;
                for (uint32_t i = 0; i < num_ma; ++i) {
                    out_mod_ma_count[nmod * num_ma + i] = c[i];
                    delta = delta + out_ma_mass[i] * (float)c[i];
                }
                out_mod_sum[nmod] = sum;
                out_mod_delta[nmod] = delta;
                ++nmod;
            }
        }
        first = false;
        /* odometer, last dimension fastest */
        int d = (int)num_ma - 1;
        while (d >= 0) {
            if (c[d] < max_ma) { ++c[d]; break; }
            c[d] = 0;
            --d;
        }
        if (d < 0) break;
    }
    *out_num_mod = nmod;
    return 0;
}

/* ---------------------------------------------------------------------
 * Database digestion (Fragment.hs:363-405)
 * ------------------------------------------------------------------- */

namespace {
struct Piece {
    float mass;
    uint32_t tc, tn, nres;
};
}

/* digest = filter inrange . splice . fragment.
 *   fragment (Fragment.hs:376-390): cut after the first residue matching the
 *     rule; piece mass = foldl' (+) 0 of residue masses; tn = index of the
 *     piece's last residue, except for the protein's final piece where tn is
 *     one past the end.
 *   splice (:397-405): each piece followed by its fusions with the next
 *     1..missed pieces, masses added left to right in FP32 (fuse).
 *   inrange (:368-369): min <= (m + H2O) + H <= max, host constants.
 * Harness extensions (not expressible in the reference, applied identically on
 * every side): p_rule = no cleavage before P; min_len/max_len = residue count
 * filter (0 disables). */
int orc_digest(const uint8_t *seq, const uint32_t *seg, uint32_t nprot,
               const float *mass26, const uint8_t *rule26, uint32_t missed,
               float min_mass, float max_mass,
               int p_rule, uint32_t min_len, uint32_t max_len,
               float *out_r, uint32_t *out_tc, uint32_t *out_tn,
               uint64_t cap, uint64_t *out_n)
{
    uint64_t n = 0;
    std::vector<Piece> pieces;
    for (uint32_t p = 0; p < nprot; ++p) {
        uint32_t beg = seg[p], end = seg[p + 1];
        pieces.clear();
        uint32_t c = beg;
        while (c < end) {
            uint32_t i = c;
            bool found = false;
            for (; i < end; ++i) {
                uint8_t ch = seq[i];
                if (ch >= 'A' && ch <= 'Z' && rule26[ch - 'A']) {
                    if (p_rule && i + 1 < end && seq[i + 1] == 'P') continue;
                    found = true;
                    break;
                }
            }
            Piece pc;
            float m = 0.0f;
            uint32_t last = found ? i : end - 1;
            for (uint32_t j = c; j <= last; ++j) m = m + mass26[seq[j] - 'A'];
            pc.mass = m;
            pc.tc = c;
            pc.tn = found ? i : end;
            pc.nres = last - c + 1;
            pieces.push_back(pc);
            if (!found) break;
            c = i + 1;
        }
        for (size_t k = 0; k < pieces.size(); ++k) {
            Piece acc = pieces[k];
            for (uint32_t j = 0;; ++j) {
                float mp = (acc.mass + HOST_H2O) + HOST_H;
                bool ok = (min_mass <= mp) && (mp <= max_mass);
                if (min_len && acc.nres < min_len) ok = false;
                if (max_len && acc.nres > max_len) ok = false;
                if (ok) {
                    if (n >= cap) return -1;
                    out_r[n] = acc.mass;
                    out_tc[n] = acc.tc;
                    out_tn[n] = acc.tn;
                    ++n;
                }
                if (j >= missed || k + j + 1 >= pieces.size()) break;
                const Piece &b = pieces[k + j + 1];
                acc.mass = acc.mass + b.mass;
                acc.tn = b.tn;
                acc.nres += b.nres;
            }
        }
    }
    *out_n = n;
    return 0;
}

/* sort.cu:28-35 sort_idx_f: idx = init..init+n-1 stably sorted by key ascending
 * (thrust sort_by_key on float keys is a stable radix sort). */
void orc_sort_idx(const float *keys, uint32_t n, uint32_t init, uint32_t *out_idx,
                  float *out_sorted_keys)
{
    std::vector<uint32_t> o(n);
    for (uint32_t i = 0; i < n; ++i) o[i] = i;
    std::stable_sort(o.begin(), o.end(),
                     [keys](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    for (uint32_t i = 0; i < n; ++i) {
        out_idx[i] = o[i] + init;
        if (out_sorted_keys) out_sorted_keys[i] = keys[o[i]];
    }
}

/* prepare_ions.cu:29-96 + utils.h:92-101: ion = (c - 'A') | mod << 5 with mod
 * the 1-based index of c in the modable-acid list (0 if not modable). */
void orc_prepare_ions(uint8_t *ions, uint64_t n, const uint8_t *ma, uint32_t num_ma)
{
    for (uint64_t i = 0; i < n; ++i) {
        uint8_t c = ions[i];
        uint8_t mod = 0;
        for (uint32_t j = 0; j < num_ma; ++j)
            if (ma[j] == c) { mod = (uint8_t)(j + 1); break; }
        ions[i] = (uint8_t)((c - 'A') | (mod << 5));
    }
}

/* ---------------------------------------------------------------------
 * Precursor windows
 * ------------------------------------------------------------------- */

static uint32_t lower_bound_perm(const float *r, const uint32_t *idx, uint32_t n, float v)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (r[idx[mid]] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}
static uint32_t upper_bound_perm(const float *r, const uint32_t *idx, uint32_t n, float v)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (v < r[idx[mid]]) hi = mid; else lo = mid + 1;
    }
    return lo;
}

/* filter.cu:75-127 findSpecCandsByMass: begin = lower_bound(r_sorted, m-eps),
 * end = upper_bound(r_sorted, m+eps), bounds formed in FP32; returns sum. */
uint32_t orc_find_spec_cands_by_mass(uint32_t *begin, uint32_t *end, uint32_t *num,
                                     const float *r, const uint32_t *idx_sorted,
                                     uint32_t num_pep, const float *masses,
                                     uint32_t num_spec, float eps)
{
    uint32_t total = 0;
    for (uint32_t s = 0; s < num_spec; ++s) {
        float lo = masses[s] - eps;
        float hi = masses[s] + eps;
        begin[s] = lower_bound_perm(r, idx_sorted, num_pep, lo);
        end[s] = upper_bound_perm(r, idx_sorted, num_pep, hi);
        num[s] = end[s] - begin[s];
        total += num[s];
    }
    return total;
}

/* prepare_scoring.cu:20-67: contiguous per-spectrum candidate index lists. */
void orc_prepare_scoring(uint32_t *out, const uint32_t *idx_sorted,
                         const uint32_t *begin, const uint32_t *num, uint32_t num_spec)
{
    uint32_t off = 0;
    for (uint32_t s = 0; s < num_spec; ++s) {
        for (uint32_t i = 0; i < num[s]; ++i) out[off + i] = idx_sorted[begin[s] + i];
        off += num[s];
    }
}

/* filter_by_mod.cu:35-143 findBeginEnd_f: for each modification the residual
 * range [(m-eps)-delta, (m+eps)-delta]; exclusive scan of the counts; returns
 * the total. */
uint32_t orc_find_begin_end(uint32_t *begin, uint32_t *end, uint32_t *num_scan,
                            const float *r, const uint32_t *idx_sorted, uint32_t num_pep,
                            const float *mod_delta, uint32_t num_mod, float mass, float eps)
{
    if (num_mod == 0) return 0;
    float lo0 = mass - eps, hi0 = mass + eps;
    uint32_t acc = 0;
    for (uint32_t i = 0; i < num_mod; ++i) {
        float lo = lo0 - mod_delta[i];
        float hi = hi0 - mod_delta[i];
        begin[i] = lower_bound_perm(r, idx_sorted, num_pep, lo);
        end[i] = upper_bound_perm(r, idx_sorted, num_pep, hi);
        num_scan[i] = acc;
        acc += end[i] - begin[i];
    }
    return acc;
}

/* ---------------------------------------------------------------------
 * Combinatorics (combinatorics.h:13-134)
 * ------------------------------------------------------------------- */

/* combinatorics.h:59-77 _choose: the reference's uint32 fallback loop. */
static uint32_t choose_loop(uint32_t n, uint32_t k)
{
    if (k == 0 || n == k) return 1;
    if (n < k) return 0;
    uint32_t t = 1;
    if (k < n - k) {
        for (uint32_t i = n; i >= n - k + 1; --i) t = t * i / (n - i + 1);
    } else {
        for (uint32_t i = n; i >= k + 1; --i) t = t * i / (n - i + 1);
    }
    return t;
}

/* combinatorics.h:80-88 choose: Pascal table for n <= 42, k <= 11 (the table
 * holds C(n,k), zero for k > n), loop otherwise.  The table is computed here
 * by Pascal's rule; tests compare it with math.comb. */
namespace {
struct PascalTable {
    uint32_t t[43][12];
    PascalTable()
    {
        for (int a = 0; a <= 42; ++a)
            for (int b = 0; b <= 11; ++b)
                t[a][b] = (b == 0) ? 1u : (a == 0 ? 0u : t[a - 1][b - 1] + t[a - 1][b]);
    }
};
const PascalTable kPascal; /* initialised before any call (static storage) */
}

uint32_t orc_choose(uint32_t n, uint32_t k)
{
    if (n > 42 || k > 11) return choose_loop(n, k);
    return kPascal.t[n][k];
}

/* combinatorics.h:90-134 largestV + unrankComb: colex unranking of a k-subset
 * of {0..n-1}; ranks above total/2 are unranked through the dual
 * x' = total-1-rank and reflected (n-1-a). */
void orc_unrank_comb(uint32_t *ans, uint32_t n, uint32_t k, uint32_t rank)
{
    uint32_t a = n, b = k;
    uint32_t total = orc_choose(n, k);
    uint32_t x;
    bool dual = false;
    if ((total / 2) >= rank) x = rank;
    else { x = total - 1 - rank; dual = true; }
    for (uint32_t i = 0; i < k; ++i) {
        uint32_t v = a - 1;
        while (orc_choose(v, b) > x) --v;
        ans[i] = v;
        x = x - orc_choose(ans[i], b);
        a = ans[i];
        b = b - 1;
    }
    if (dual)
        for (uint32_t i = 0; i < k; ++i) ans[i] = (n - 1) - ans[i];
}

/* ---------------------------------------------------------------------
 * Modified-peptide enumeration
 * ------------------------------------------------------------------- */

/* filter_by_mod.cu:150-378 findModablePeptides.
 *  1. fillPepAndModIdx (:151-190): rows (pep, mod) spectrum-major, mod-minor,
 *     peptides in residual-sorted order within each range.
 *  2. fillPepMACount (:193-231): count of each modable acid in [tc, tn).
 *  3. checkModable (:234-262): keep rows whose counts cover the mod vector.
 *  4. stable compaction of rows and count rows (remove_if).
 *  5. spec_num_valid_scan[s] = lower_bound(kept row ids, spec_num_cand_scan[s]). */
uint32_t orc_find_modable_peptides(uint32_t *out_pep_idx, uint32_t *out_pep_mod_idx,
                                   uint32_t *out_pep_ma_count, uint32_t *out_spec_num_valid_scan,
                                   const uint32_t *spec_num_cand_scan, uint32_t num_spec,
                                   uint32_t num_cand_total, const uint8_t *ions,
                                   const uint32_t *tc, const uint32_t *tn,
                                   const uint32_t *idx_sorted, const uint32_t *begin,
                                   const uint32_t *end, const uint32_t *num_pep_scan,
                                   const uint32_t *mod_ma_count, uint32_t num_mod,
                                   const uint8_t *ma, uint32_t num_ma)
{
    std::vector<uint32_t> pep(num_cand_total), mod(num_cand_total);
    for (uint32_t s = 0; s < num_spec; ++s)
        for (uint32_t m = 0; m < num_mod; ++m) {
            size_t off = (size_t)s * num_mod + m;
            uint32_t o = num_pep_scan[off] + spec_num_cand_scan[s];
            for (uint32_t i = begin[off]; i != end[off]; ++i) {
                pep[o] = idx_sorted[i];
                mod[o] = m;
                ++o;
            }
        }
    std::vector<uint32_t> cnt((size_t)num_cand_total * num_ma);
    for (uint32_t row = 0; row < num_cand_total; ++row)
        for (uint32_t a = 0; a < num_ma; ++a) {
            uint32_t c = 0;
            for (uint32_t j = tc[pep[row]]; j < tn[pep[row]]; ++j)
                if (acid_char(ions[j]) == ma[a]) ++c;
            cnt[(size_t)row * num_ma + a] = c;
        }
    std::vector<uint32_t> kept_row;
    uint32_t nv = 0;
    for (uint32_t row = 0; row < num_cand_total; ++row) {
        bool ok = true;
        for (uint32_t a = 0; a < num_ma; ++a)
            if (cnt[(size_t)row * num_ma + a] < mod_ma_count[mod[row] * num_ma + a]) ok = false;
        if (!ok) continue;
        out_pep_idx[nv] = pep[row];
        out_pep_mod_idx[nv] = mod[row];
        for (uint32_t a = 0; a < num_ma; ++a)
            out_pep_ma_count[(size_t)nv * num_ma + a] = cnt[(size_t)row * num_ma + a];
        kept_row.push_back(row);
        ++nv;
    }
    for (uint32_t s = 0; s < num_spec; ++s)
        out_spec_num_valid_scan[s] =
            (uint32_t)(std::lower_bound(kept_row.begin(), kept_row.end(), spec_num_cand_scan[s]) -
                       kept_row.begin());
    return nv;
}

/* calc_mod_cands.cu:27-119: per row num_mpep = prod_i C(D_i, d_i); the row of
 * suffix products pep_ma_num_comb_scan[i] = prod_{j>=i} C(D_j, d_j). */
uint32_t orc_calc_total_mod_cands(uint32_t *out_pep_num_mpep, uint32_t *out_pep_ma_num_comb_scan,
                                  const uint32_t *mod_ma_count, const uint32_t *pep_idx,
                                  const uint32_t *pep_mod_idx, const uint32_t *pep_ma_count,
                                  uint32_t num_cand, uint32_t num_mod, uint32_t num_ma)
{
    (void)pep_idx;
    (void)num_mod;
    uint32_t total = 0;
    for (uint32_t p = 0; p < num_cand; ++p) {
        uint32_t num = 1;
        uint32_t m = pep_mod_idx[p];
        for (int32_t i = (int32_t)num_ma - 1; i >= 0; --i) {
            num *= orc_choose(pep_ma_count[num_ma * p + i], mod_ma_count[num_ma * m + i]);
            out_pep_ma_num_comb_scan[num_ma * p + i] = num;
        }
        out_pep_num_mpep[p] = num;
        total += num;
    }
    return total;
}

/* prepare_gen_mod.cu:23-129: expand each row into its modified peptides
 * (pep_idx, mod_idx, rank 0..k-1, row, sum of mod counts), exclusive scan of
 * the sums; returns the unrank length. */
uint32_t orc_prepare_gen_mod(uint32_t *out_mpep_pep_idx, uint32_t *out_mpep_pep_mod_idx,
                             uint32_t *out_mpep_rank, uint32_t *out_mpep_ith_cand,
                             uint32_t *out_mpep_mod_ma_count_sum_scan,
                             const uint32_t *mod_ma_count_sum, const uint32_t *pep_idx,
                             const uint32_t *pep_mod_idx, const uint32_t *pep_num_mpep,
                             uint32_t num_cand, uint32_t num_mpep)
{
    uint32_t o = 0, acc = 0;
    for (uint32_t c = 0; c < num_cand; ++c) {
        uint32_t mm = mod_ma_count_sum[pep_mod_idx[c]];
        for (uint32_t i = 0; i < pep_num_mpep[c]; ++i) {
            if (o >= num_mpep) return acc;
            out_mpep_pep_idx[o] = pep_idx[c];
            out_mpep_pep_mod_idx[o] = pep_mod_idx[c];
            out_mpep_rank[o] = i;
            out_mpep_ith_cand[o] = c;
            out_mpep_mod_ma_count_sum_scan[o] = acc;
            acc += mm;
            ++o;
        }
    }
    return acc;
}

/* gen_mod_pep.cu:127-157 add_ma_ranks: mixed-radix split of the mpep rank over
 * the suffix products. */
void orc_add_ma_ranks(uint32_t *ranks, uint32_t rank, const uint32_t *scan_row, uint32_t num_ma)
{
    uint32_t x = rank;
    for (uint32_t i = 0; i < num_ma; ++i) {
        if (i != num_ma - 1) {
            uint32_t nxt = scan_row[i + 1];
            if (x >= nxt) { ranks[i] = x / nxt; x = x % nxt; }
            else ranks[i] = 0;
        } else {
            ranks[i] = x;
        }
    }
}

/* gen_mod_pep.cu:159-312 genModCands: per modified peptide, unrank each
 * modified acid's occurrence combination into its slice of out_unrank. */
void orc_gen_mod_cands(uint32_t *out_unrank, const uint32_t *mod_ma_count,
                       const uint32_t *mpep_ith_cand, const uint32_t *mpep_rank,
                       const uint32_t *mpep_mod_ma_count_sum_scan,
                       const uint32_t *pep_mod_idx, const uint32_t *pep_ma_count,
                       const uint32_t *pep_ma_num_comb_scan, uint32_t num_mpep, uint32_t num_ma)
{
    std::vector<uint32_t> ranks(num_ma);
    for (uint32_t mp = 0; mp < num_mpep; ++mp) {
        uint32_t ith = mpep_ith_cand[mp];
        uint32_t pm = pep_mod_idx[ith];
        orc_add_ma_ranks(ranks.data(), mpep_rank[mp], pep_ma_num_comb_scan + (size_t)ith * num_ma, num_ma);
        uint32_t start = mpep_mod_ma_count_sum_scan[mp], pos = 0;
        for (uint32_t i = 0; i < num_ma; ++i) {
            uint32_t k = mod_ma_count[pm * num_ma + i];
            if (k != 0) {
                orc_unrank_comb(out_unrank + start + pos, pep_ma_count[(size_t)ith * num_ma + i], k,
                                ranks[i]);
                pos += k;
            }
        }
    }
}

/* ---------------------------------------------------------------------
 * Theoretical spectra (ion_series.cu:59-232, mion_series.cu:143-328)
 * ------------------------------------------------------------------- */

/* binMZ (ion_series.cu:98-102) with IEEE division: rint(mz / 1.0005079). */
static inline int32_t bin_mz(float mz)
{
    float q = mz / DEV_BIN_WIDTH_MONO;
    return (int32_t)rintf(q);
}

/* Emit every peak of ladder position b (and y = r - b) for charges 1..zmax in
 * the reference's order: addIonsAB (A, B, B+1, B-1, B-H2O, B-NH3) then addIonsY
 * (Y, Y+1, Y-1, Y-NH3), charges ascending (ion_series.cu:111-156, :242-245). */
template <class Emit>
static inline void emit_position(float b, float y, uint32_t zmax, Emit &&emit)
{
    for (uint32_t z = 1; z <= zmax; ++z) {
        float fz = (float)z;
        float zh = DEV_H * fz;
        emit(bin_mz(((b - DEV_CO) + zh) / fz), 10);
        float mb = (b + zh) / fz;
        int32_t xb = bin_mz(mb);
        emit(xb, 50);
        emit(xb + 1, 25);
        emit(xb - 1, 25);
        emit(bin_mz(mb - DEV_H2O / fz), 10);
        emit(bin_mz(mb - DEV_NH3 / fz), 10);
        float my = ((y + DEV_H2O) + zh) / fz;
        int32_t xy = bin_mz(my);
        emit(xy, 50);
        emit(xy + 1, 25);
        emit(xy - 1, 25);
        emit(bin_mz(my - DEV_NH3 / fz), 10);
    }
}

/* Ladder over the n residue masses (the reference's loop j in [tc, tn)):
 * b = sequential FP32 prefix, y = r - b. */
template <class Emit>
static inline void ladder(const float *masses, uint32_t n, float r, uint32_t zmax, Emit &&emit)
{
    float b = 0.0f;
    for (uint32_t k = 0; k < n; ++k) {
        b = b + masses[k];
        float y = r - b;
        emit_position(b, y, zmax, emit);
    }
}

uint32_t orc_peaks_list(int32_t *bins, uint8_t *wts, uint32_t cap, const float *masses,
                        uint32_t n, float r, uint32_t zmax)
{
    uint32_t cnt = 0;
    ladder(masses, n, r, zmax, [&](int32_t x, uint8_t w) {
        if (cnt < cap) { bins[cnt] = x; wts[cnt] = w; }
        ++cnt;
    });
    return cnt;
}

static inline void dense_from_masses(uint8_t *w, uint32_t len, const float *masses, uint32_t n,
                                     float r, uint32_t zmax)
{
    std::memset(w, 0, len);
    ladder(masses, n, r, zmax, [&](int32_t x, uint8_t wt) {
        /* addIon (ion_series.cu:104-108): keep 0 <= x < len, max per bin */
        if (x >= 0 && (uint32_t)x < len && w[x] < wt) w[x] = wt;
    });
}

void orc_peaks_dense(uint8_t *w, uint32_t len, const uint8_t *ions, uint32_t tc, uint32_t tn,
                     float r, const float *mass26, uint32_t zmax)
{
    uint32_t n = tn > tc ? tn - tc : 0;
    std::vector<float> m(n);
    for (uint32_t k = 0; k < n; ++k) m[k] = mass26[acid_idx(ions[tc + k])];
    dense_from_masses(w, len, m.data(), n, r, zmax);
}

/* Residue masses of one modified peptide.  The occurrence index of a modable
 * acid counts only that acid (check_series.cu:61-99 genMIons, consistent with
 * the per-acid counts of filter_by_mod.cu:216-230 and the unrank domain
 * [0, D_i)).  A modified residue weighs table + ma_mass (mion_series.cu:277);
 * the residual is r + mod_delta (mion_series.cu:226). */
static void mod_masses(std::vector<float> &m, const uint8_t *ions, uint32_t tc, uint32_t tn,
                       const float *mass26, const uint32_t *mod_row, const uint8_t *ma,
                       const float *ma_mass, uint32_t num_ma, const uint32_t *unrank)
{
    uint32_t n = tn > tc ? tn - tc : 0;
    m.resize(n);
    uint32_t occ[32] = {0};
    for (uint32_t k = 0; k < n; ++k) {
        uint8_t ch = acid_char(ions[tc + k]);
        float v = mass26[acid_idx(ions[tc + k])];
        for (uint32_t a = 0; a < num_ma; ++a) {
            if (ma[a] != ch) continue;
            uint32_t ith = occ[a]++;
            uint32_t off = 0;
            for (uint32_t q = 0; q < a; ++q) off += mod_row[q];
            for (uint32_t q = 0; q < mod_row[a]; ++q)
                if (unrank[off + q] == ith) { v = v + ma_mass[a]; break; }
            break;
        }
        m[k] = v;
    }
}

void orc_mpeaks_dense(uint8_t *w, uint32_t len, const uint8_t *ions, uint32_t tc, uint32_t tn,
                      float r, const float *mass26, uint32_t zmax,
                      const uint32_t *mod_ma_count_row, float mod_delta,
                      const uint8_t *ma, const float *ma_mass, uint32_t num_ma,
                      const uint32_t *unrank)
{
    std::vector<float> m;
    mod_masses(m, ions, tc, tn, mass26, mod_ma_count_row, ma, ma_mass, num_ma, unrank);
    float rr = r + mod_delta;
    dense_from_masses(w, len, m.data(), (uint32_t)m.size(), rr, zmax);
}

/* ---------------------------------------------------------------------
 * Scores
 * ------------------------------------------------------------------- */

static inline int64_t fx_term(uint8_t w, float x)
{
    double p = (double)w * (double)x; /* exact: <= 6 x 24 significant bits */
    return (int64_t)std::llrint(std::ldexp(p, ORC_FX_SHIFT));
}

/* Exact score: sum of round(w*x*2^36) over the bins, int64 (order-free). */
int64_t orc_score_fx(const uint8_t *w, const float *x, uint32_t len)
{
    int64_t s = 0;
    for (uint32_t b = 0; b < len; ++b)
        if (w[b]) s += fx_term(w[b], x[b]);
    return s;
}

/* Reference-like FP32 dot (Search.hs:242-245 sgemv): products rounded to FP32,
 * summed in ascending bin order (cuBLAS's order is unspecified). */
float orc_score_f32(const uint8_t *w, const float *x, uint32_t len)
{
    float s = 0.0f;
    for (uint32_t b = 0; b < len; ++b)
        if (w[b]) s = s + (float)w[b] * x[b];
    return s;
}

double orc_fx_to_double(int64_t fx) { return std::ldexp((double)fx, -ORC_FX_SHIFT); }

uint64_t orc_check_fast_div(float d, float lo, float hi, int corrected)
{
    volatile float dv = d; /* keep 1/d a run-time IEEE division */
    const float R = 1.0f / dv;
    uint32_t a, b;
    std::memcpy(&a, &lo, 4);
    std::memcpy(&b, &hi, 4);
    uint64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t u = a; u < (int64_t)b; ++u) {
        const uint32_t bits = (uint32_t)u;
        float x;
        std::memcpy(&x, &bits, 4);
        const float q = x * R;
        const float r = std::fma(-q, d, x);
        const float q2 = corrected ? std::fma(r, R, q) : q;
        if (q2 != x / d) ++bad;
    }
    return bad;
}

/* ---------------------------------------------------------------------
 * Spectrum preprocessing (Correlation.hs:29-109)
 * ------------------------------------------------------------------- */

/* sequestXCorr = crossCorrolation . normaliseByRegion . observedIntensity.
 * observedIntensity (:57-71): cutoff = min(50 + pcr*z, trunc(last_mz/10)*10);
 *   len = 75 + ceiling cutoff; keep peaks with mz <= cutoff (and, when
 *   removing the precursor, mz <= pcr-15 || pcr+15 <= mz); bin = round half
 *   even (mz / bin width); value = max sqrt(I) per bin.
 * normaliseByRegion (:84-96): sel = (len-75) div 10; region = min 10 (i div
 *   sel); value 50*(e/max_region) when max_region > 1e-6 and i < 10*sel, else 0.
 *   sel == 0 is a division by zero in the reference: returned as error -2.
 * crossCorrolation (:105-109): x_i = e_i - sum(a[max(0,i-75) .. +150]) / 150.
 *   mean_mode 0 sums in FP32 from the left (foldl'), mode 1 sums in FP64 and
 *   rounds x_i = e_i - S/150 once. */
int orc_preprocess(float *out, uint32_t cap, uint32_t *out_len, float pcr, float charge,
                   const float *mz, const float *inten, uint32_t npk,
                   int remove_precursor, int mono, int mean_mode)
{
    if (npk == 0) return -1;
    float bw = mono ? 1.0005079f : 1.0011413f;
    float mzcut = 50.0f + pcr * charge;
    float q = mz[npk - 1] / 10.0f;
    float mzlim = (float)((int64_t)std::trunc(q) * 10);
    float cutoff = std::min(mzcut, mzlim);
    /* Haskell min x y = if x <= y then x else y */
    cutoff = (mzcut <= mzlim) ? mzcut : mzlim;
    int64_t lenl = 75 + (int64_t)std::ceil(cutoff);
    if (lenl <= 0 || (uint64_t)lenl > cap) return -3;
    uint32_t len = (uint32_t)lenl;
    *out_len = len;
    std::vector<float> s(len, 0.0f);
    float plo = pcr - 15.0f, phi = pcr + 15.0f;
    for (uint32_t i = 0; i < npk; ++i) {
        float x = mz[i];
        bool keep = remove_precursor ? (x <= cutoff && (x <= plo || phi <= x)) : (x <= cutoff);
        if (!keep) continue;
        int32_t bn = (int32_t)rintf(x / bw);
        float v = sqrtf(inten[i]);
        if (bn < 0 || (uint32_t)bn >= len) return -4;
        /* Haskell max x y = if x <= y then y else x */
        s[bn] = (s[bn] <= v) ? v : s[bn];
    }
    if (len < 85) return -2; /* sel = (len-75) div 10 <= 0: division by zero or negative regions */
    int32_t sel = ((int32_t)len - 75) / 10;
    float rmax[11] = {0};
    for (uint32_t i = 0; i < len; ++i) {
        int32_t rg = std::min<int32_t>(10, (int32_t)i / sel);
        rmax[rg] = (rmax[rg] <= s[i]) ? s[i] : rmax[rg];
    }
    std::vector<float> a(len);
    for (uint32_t i = 0; i < len; ++i) {
        int32_t rg = std::min<int32_t>(10, (int32_t)i / sel);
        float m = rmax[rg];
        a[i] = (m > 1e-6f && (int32_t)i < sel * 10) ? 50.0f * (s[i] / m) : 0.0f;
    }
    for (uint32_t i = 0; i < len; ++i) {
        uint32_t st = i >= 75 ? i - 75 : 0;
        uint32_t en = std::min<uint32_t>(len, st + 151);
        if (mean_mode == 0) {
            float sum = 0.0f;
            for (uint32_t j = st; j < en; ++j) sum = sum + a[j];
            out[i] = a[i] - sum / 150.0f;
        } else {
            double sum = 0.0;
            for (uint32_t j = st; j < en; ++j) sum = sum + (double)a[j];
            out[i] = (float)((double)a[i] - sum / 150.0);
        }
    }
    for (uint32_t i = len; i < cap; ++i) out[i] = 0.0f;
    return 0;
}

/* ---------------------------------------------------------------------
 * Whole searches
 * ------------------------------------------------------------------- */

namespace {
/* top-N by (fx desc, position asc): Search.hs:119-123, :276 and the stable
 * ascending radix sort of sort_b40c.cu:34-41 followed by reverse. */
struct TopN {
    uint32_t n = 0, cap = 0;
    std::vector<int64_t> fx;
    std::vector<uint32_t> pos;
    void init(uint32_t c) { cap = c; n = 0; fx.assign(c, 0); pos.assign(c, 0); }
    void push(int64_t f, uint32_t p)
    {
        if (cap == 0) return;
        if (n == cap && !(f > fx[n - 1])) return; /* later position loses ties */
        uint32_t i = (n < cap) ? n++ : n - 1;
        while (i > 0 && f > fx[i - 1]) { fx[i] = fx[i - 1]; pos[i] = pos[i - 1]; --i; }
        fx[i] = f;
        pos[i] = p;
    }
};
}

int orc_search_plain(const uint8_t *ions, const float *r, const uint32_t *tc, const uint32_t *tn,
                     const float *mass26, const uint32_t *idx_sorted, uint32_t num_pep,
                     const float *x, uint32_t x_stride, const uint32_t *len,
                     const uint32_t *zmax, const float *mass, uint32_t num_spec, float eps,
                     uint32_t topn,
                     uint32_t *out_begin, uint32_t *out_num,
                     uint32_t *out_top_pos, int64_t *out_top_fx, uint32_t *out_top_cnt,
                     float *out_all_f32, int64_t *out_all_fx, const uint64_t *all_off,
                     int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    std::vector<float> rs(num_pep);
    for (uint32_t i = 0; i < num_pep; ++i) rs[i] = r[idx_sorted[i]];
#pragma omp parallel
    {
        std::vector<uint8_t> w;
        std::vector<float> m;
        TopN top;
#pragma omp for schedule(dynamic, 1)
        for (int64_t si = 0; si < (int64_t)num_spec; ++si) {
            uint32_t s = (uint32_t)si;
            float lo = mass[s] - eps, hi = mass[s] + eps;
            uint32_t b = (uint32_t)(std::lower_bound(rs.begin(), rs.end(), lo) - rs.begin());
            uint32_t e = (uint32_t)(std::upper_bound(rs.begin(), rs.end(), hi) - rs.begin());
            out_begin[s] = b;
            out_num[s] = e - b;
            uint32_t L = len[s];
            w.resize(L);
            const float *xs = x + (size_t)s * x_stride;
            top.init(topn);
            for (uint32_t p = b; p < e; ++p) {
                uint32_t pep = idx_sorted[p];
                uint32_t n = tn[pep] > tc[pep] ? tn[pep] - tc[pep] : 0;
                m.resize(n);
                for (uint32_t k = 0; k < n; ++k) m[k] = mass26[acid_idx(ions[tc[pep] + k])];
                /* ion_series.cu:334-342: no addIons launch for max charge > 4,
                 * the zeroed rows score 0 */
                dense_from_masses(w.data(), L, m.data(), n, r[pep], zmax[s] <= 4 ? zmax[s] : 0);
                int64_t fx = 0;
                float f32 = 0.0f;
                for (uint32_t q = 0; q < L; ++q) {
                    if (!w[q]) continue;
                    fx += fx_term(w[q], xs[q]);
                    f32 = f32 + (float)w[q] * xs[q];
                }
                if (out_all_f32) out_all_f32[all_off[s] + (p - b)] = f32;
                if (out_all_fx) out_all_fx[all_off[s] + (p - b)] = fx;
                top.push(fx, p - b);
            }
            out_top_cnt[s] = top.n;
            for (uint32_t i = 0; i < top.n; ++i) {
                out_top_pos[(size_t)s * topn + i] = top.pos[i];
                out_top_fx[(size_t)s * topn + i] = top.fx[i];
            }
        }
    }
    return 0;
}

/* Search.hs:282-622 searchWithMods for one batch of spectra: windows per
 * (spectrum, mod), modability filter, enumeration, modified ladders, exact
 * scores, per-spectrum top-N by (score desc, mpep index asc). */
int orc_search_mods(const uint8_t *ions, const float *r, const uint32_t *tc, const uint32_t *tn,
                    const float *mass26, const uint32_t *idx_sorted, uint32_t num_pep,
                    const uint8_t *ma, const float *ma_mass, uint32_t num_ma,
                    const uint32_t *mod_ma_count, const uint32_t *mod_sum,
                    const float *mod_delta, uint32_t num_mod,
                    const float *x, uint32_t x_stride, const uint32_t *len,
                    const uint32_t *zmax, const float *mass, uint32_t num_spec, float eps,
                    uint32_t topn,
                    uint32_t *out_spec_num_mpep, uint32_t *out_top_mpep, int64_t *out_top_fx,
                    uint32_t *out_top_cnt, uint32_t *out_mpep_pep_idx,
                    uint32_t *out_mpep_mod_idx, uint32_t *out_mpep_unrank_off,
                    uint32_t *out_unrank, uint64_t cap_mpep, uint64_t cap_unrank,
                    uint64_t *out_num_mpep, uint64_t *out_len_unrank, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    *out_num_mpep = 0;
    *out_len_unrank = 0;
    if (num_mod == 0 || num_spec == 0) {
        for (uint32_t s = 0; s < num_spec; ++s) { out_spec_num_mpep[s] = 0; out_top_cnt[s] = 0; }
        return 0;
    }
    size_t sm = (size_t)num_spec * num_mod;
    std::vector<uint32_t> beg(sm), end(sm), scan(sm), spec_num_cand(num_spec);
    for (uint32_t s = 0; s < num_spec; ++s)
        spec_num_cand[s] = orc_find_begin_end(&beg[(size_t)s * num_mod], &end[(size_t)s * num_mod],
                                              &scan[(size_t)s * num_mod], r, idx_sorted, num_pep,
                                              mod_delta, num_mod, mass[s], eps);
    std::vector<uint32_t> cand_scan(num_spec + 1, 0);
    for (uint32_t s = 0; s < num_spec; ++s) cand_scan[s + 1] = cand_scan[s] + spec_num_cand[s];
    uint32_t total = cand_scan[num_spec];
    std::vector<uint32_t> pep(total), pmod(total), pcnt((size_t)total * num_ma), valid_scan(num_spec);
    uint32_t nv = orc_find_modable_peptides(pep.data(), pmod.data(), pcnt.data(), valid_scan.data(),
                                            cand_scan.data(), num_spec, total, ions, tc, tn,
                                            idx_sorted, beg.data(), end.data(), scan.data(),
                                            mod_ma_count, num_mod, ma, num_ma);
    std::vector<uint32_t> num_mp(nv), comb((size_t)nv * num_ma);
    uint32_t nmp = orc_calc_total_mod_cands(num_mp.data(), comb.data(), mod_ma_count, pep.data(),
                                            pmod.data(), pcnt.data(), nv, num_mod, num_ma);
    std::vector<uint32_t> mp_pep(nmp), mp_mod(nmp), mp_rank(nmp), mp_ith(nmp), mp_scan(nmp);
    uint32_t lu = orc_prepare_gen_mod(mp_pep.data(), mp_mod.data(), mp_rank.data(), mp_ith.data(),
                                      mp_scan.data(), mod_sum, pep.data(), pmod.data(),
                                      num_mp.data(), nv, nmp);
    std::vector<uint32_t> unrank(lu);
    orc_gen_mod_cands(unrank.data(), mod_ma_count, mp_ith.data(), mp_rank.data(), mp_scan.data(),
                      pmod.data(), pcnt.data(), comb.data(), nmp, num_ma);
    /* Search.hs:381-386: spectrum s owns rows [valid_scan[s], valid_scan[s+1]) */
    std::vector<uint64_t> sp_first(num_spec + 1, 0);
    {
        std::vector<uint64_t> row_scan(nv + 1, 0);
        for (uint32_t i = 0; i < nv; ++i) row_scan[i + 1] = row_scan[i] + num_mp[i];
        for (uint32_t s = 0; s < num_spec; ++s) sp_first[s] = row_scan[valid_scan[s]];
        sp_first[num_spec] = nmp;
        for (uint32_t s = 0; s < num_spec; ++s)
            out_spec_num_mpep[s] = (uint32_t)(sp_first[s + 1] - sp_first[s]);
    }
    *out_num_mpep = nmp;
    *out_len_unrank = lu;
    if (nmp <= cap_mpep && lu <= cap_unrank) {
        for (uint32_t i = 0; i < nmp; ++i) {
            if (out_mpep_pep_idx) out_mpep_pep_idx[i] = mp_pep[i];
            if (out_mpep_mod_idx) out_mpep_mod_idx[i] = mp_mod[i];
            if (out_mpep_unrank_off) out_mpep_unrank_off[i] = mp_scan[i];
        }
        if (out_unrank) for (uint32_t i = 0; i < lu; ++i) out_unrank[i] = unrank[i];
    }
#pragma omp parallel
    {
        std::vector<uint8_t> w;
        std::vector<float> m;
        TopN top;
#pragma omp for schedule(dynamic, 1)
        for (int64_t si = 0; si < (int64_t)num_spec; ++si) {
            uint32_t s = (uint32_t)si;
            uint32_t L = len[s];
            w.resize(L);
            const float *xs = x + (size_t)s * x_stride;
            top.init(topn);
            for (uint64_t i = sp_first[s]; i < sp_first[s + 1]; ++i) {
                uint32_t p = mp_pep[i], md = mp_mod[i];
                mod_masses(m, ions, tc[p], tn[p], mass26, mod_ma_count + (size_t)md * num_ma, ma,
                           ma_mass, num_ma, unrank.data() + mp_scan[i]);
                float rr = r[p] + mod_delta[md];
                /* mion_series.cu:335-556 dispatches max charges 1..10 only */
                dense_from_masses(w.data(), L, m.data(), (uint32_t)m.size(), rr,
                                  zmax[s] <= 10 ? zmax[s] : 0);
                int64_t fx = 0;
                for (uint32_t q = 0; q < L; ++q)
                    if (w[q]) fx += fx_term(w[q], xs[q]);
                top.push(fx, (uint32_t)(i - sp_first[s]));
            }
            out_top_cnt[s] = top.n;
            for (uint32_t i = 0; i < top.n; ++i) {
                out_top_mpep[(size_t)s * topn + i] = top.pos[i];
                out_top_fx[(size_t)s * topn + i] = top.fx[i];
            }
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------
 * E-values (Util/Statistics.hs) and SP ion matches (Sequence/IonSeries.hs,
 * Sequence/Match.hs, Sequence/Location.hs)
 * ------------------------------------------------------------------- */

/* evalues / linearFitLine / histogram (Statistics.hs:37-100) over one
 * spectrum's match scores in any order (Search.hs:118-121 sorts allMatches
 * ascending first):
 *   n = #scores; cut_end = round((1 - perc/100) * n) in Float (the
 *   percentage is a Float, Config.hs:101); cut_beg = round(30/100 * n) in
 *   Double; GHC's round takes halves to the even neighbour;
 *   sc'' = [(x + 1) + (-1) * head sc | x <- drop cut_beg (take cut_end sc)];
 *   sc' = filter (> 4000) sc'';
 *   histogram: bin i = ceiling (x / w), centre i * w - w / 2, uniqCount of
 *   the ascending centres, height count / w;
 *   x = centre / 10000, logh = logBase 10 h = log h / log 10;
 *   (alpha, beta) = linearRegression x logh -- statistics-linreg (a
 *   dependency outside the reference tree): ordinary least squares,
 *   beta = cov(x, logh) / var(x), alpha = mean logh - beta * mean x.
 * Fewer than two histogram points leave 0/0 in beta, i.e. NaN, as in
 * Haskell.  Returns the number of histogram points; *nused = #sc'. */
uint32_t orc_evalue_fit(const float *scores, uint64_t n, uint32_t bin_width, float perc,
                        double *alpha, double *beta, uint64_t *nused)
{
    *alpha = NAN;
    *beta = NAN;
    *nused = 0;
    if (n == 0) return 0;
    std::vector<double> sc(n);
    for (uint64_t i = 0; i < n; ++i) sc[i] = (double)scores[i];
    std::sort(sc.begin(), sc.end());
    const float fe = (1.0f - perc / 100.0f) * (float)n;
    const int64_t cut_end = (int64_t)std::nearbyint((double)fe);
    const int64_t cut_beg = (int64_t)std::nearbyint((30.0 / 100.0) * (double)n);
    const int64_t ce = std::min<int64_t>(std::max<int64_t>(cut_end, 0), (int64_t)n);
    const int64_t cb = std::max<int64_t>(cut_beg, 0);
    const double h0 = sc[0];
    std::vector<double> kept;
    for (int64_t i = cb; i < ce; ++i) {
        const double y = (sc[i] + 1.0) + (-1.0) * h0;
        if (y > 4000.0) kept.push_back(y);
    }
    *nused = kept.size();
    const double w = (double)bin_width, half = w / 2.0;
    std::vector<double> xs, ys;
    for (size_t i = 0; i < kept.size();) {
        const double c = std::ceil(kept[i] / w) * w - half;
        size_t j = i + 1;
        while (j < kept.size() && std::ceil(kept[j] / w) * w - half == c) ++j;
        xs.push_back(c / 10000.0);
        ys.push_back(std::log((double)(j - i) / w) / std::log(10.0));
        i = j;
    }
    const size_t m = xs.size();
    double mx = 0.0, my = 0.0;
    for (size_t i = 0; i < m; ++i) {
        mx += xs[i];
        my += ys[i];
    }
    mx /= (double)m;
    my /= (double)m;
    double sxx = 0.0, sxy = 0.0;
    for (size_t i = 0; i < m; ++i) {
        const double dx = xs[i] - mx;
        sxx += dx * dx;
        sxy += dx * (ys[i] - my);
    }
    *beta = sxy / sxx;
    *alpha = my - *beta * mx;
    return (uint32_t)m;
}

/* Statistics.hs:41: E(x) = 10 ** (alpha + beta * (x / 10000)) */
double orc_evalue(double alpha, double beta, double score)
{
    return std::pow(10.0, alpha + beta * (score / 10000.0));
}

/* extractPeaks (IonSeries.hs:99-107): the 200 largest values of the
 * spectrum, each folded with max into entry i div 3 of len/3 + 1 zeros;
 * out[t] = 1 where that entry ends above zero (the test matchIonSequence
 * applies, IonSeries.hs:146-147).  Ties at the 200th value keep the lowest
 * bins (the reference leaves them to vector-algorithms' introselect). */
void orc_extract_peaks(const float *x, uint32_t len, uint8_t *out)
{
    std::vector<uint32_t> idx(len);
    for (uint32_t i = 0; i < len; ++i) idx[i] = i;
    const size_t k = std::min<size_t>(200, len);
    std::partial_sort(idx.begin(), idx.begin() + k, idx.end(), [&](uint32_t a, uint32_t b) {
        return x[a] > x[b] || (x[a] == x[b] && a < b);
    });
    std::vector<float> peaks(len / 3 + 1, 0.0f);
    for (size_t i = 0; i < k; ++i) {
        const uint32_t t = idx[i] / 3;
        peaks[t] = peaks[t] <= x[idx[i]] ? x[idx[i]] : peaks[t];
    }
    for (uint32_t t = 0; t < len / 3 + 1; ++t) out[t] = peaks[t] > 0.0f;
}

/* lookup (Location.hs:42-58) + matchIonSequence (IonSeries.hs:136-147) +
 * scoreSP (Match.hs:33-40) for one peptide against one spectrum's peak flags
 * (orc_extract_peaks, npk = len/3 + 1 entries):
 *   residues ions[tc .. tn], the last one dropped when tn is the database's
 *   end; fragment mass r = (res + H) + H2O with the host constants;
 *   ions = concat over c = 1 .. zmax of ionMZ(p_k, c) = (p_k + H * c) / c,
 *   p = scanl1 (+) of the residue masses (FP32, left to right);
 *   bin x = round (x / bin width / 3), halves to even; matched when
 *   0 <= bin < npk and the flag is set;
 *   b = init (matched ions), y = tail (matched (map (r -) ions));
 *   total = (#residues - 1 + modified sites) * 2 (modifyFragment adds one
 *   '*' per site, Fragment.hs:84-106).
 * b_flags / y_flags (optional) receive the n_ions - 1 flags of each list. */
void orc_match_ions(const uint8_t *ions, uint64_t num_ions, uint32_t tc, uint32_t tn, float res,
                    const float *mass26, uint32_t zmax, int mono, const uint8_t *on, uint32_t npk,
                    uint32_t nmod, uint32_t *out_nb, uint32_t *out_ny, uint32_t *out_total,
                    uint8_t *b_flags, uint8_t *y_flags)
{
    const uint32_t ntake = (uint64_t)tn < num_ions ? tn - tc + 1 : tn - tc;
    const float r = (res + HOST_H) + HOST_H2O;
    const float bw = mono ? 1.0005079f : 1.0011413f;
    auto matched = [&](float v) {
        const float q = std::nearbyint((v / bw) / 3.0f);
        return q >= 0.0f && q < (float)npk && on[(uint32_t)q] != 0;
    };
    const uint32_t nz = zmax < 1 ? 1 : zmax;
    const uint64_t n_ions = (uint64_t)ntake * nz;
    uint32_t nb = 0, ny = 0;
    uint64_t j = 0;
    for (uint32_t c = 1; c <= nz; ++c) {
        const float cf = (float)c;
        float p = 0.0f;
        for (uint32_t k = 0; k < ntake; ++k, ++j) {
            p = p + mass26[acid_idx(ions[tc + k])];
            const float mz = (p + HOST_H * cf) / cf;
            if (j + 1 < n_ions) {
                const bool m = matched(mz);
                nb += m;
                if (b_flags) b_flags[j] = m;
            }
            if (j >= 1) {
                const bool m = matched(r - mz);
                ny += m;
                if (y_flags) y_flags[j - 1] = m;
            }
        }
    }
    *out_nb = nb;
    *out_ny = ny;
    *out_total = (ntake - 1 + nmod) * 2;
}
