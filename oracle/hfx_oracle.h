/*
 * hfx_oracle.h -- CPU restatement of HFX's GPU search path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product (hfx_amd/, libhfx_hip.so) never
 * links or calls it.
 *
 * Every function restates one reference function; the citation is on the
 * definition in hfx_oracle.cpp.  Canonical arithmetic (fixed here and used
 * identically by the HIP kernels):
 *   - FP32 wherever the reference is FP32, no contraction (-ffp-contract=off);
 *   - IEEE correctly rounded division and sqrt;
 *   - rint (round half to even) for every m/z -> bin conversion;
 *   - b-ion ladder = sequential left-to-right FP32 prefix of residue masses;
 *   - "exact" score = sum over bins of round(w*x*2^36) in int64, i.e. an
 *     order-independent fixed-point accumulation (2^-36 resolution) that both
 *     the oracle and the GPU reproduce bit for bit; reported as a double.
 *
 * Every function is pinned in tests/test_oracle_pins.py by a method other
 * than its own formula:
 *   mass table            the 35 matches printed in the reference's result
 *                         files (tests/golden/reference_out_peptides.json);
 *                         average > monoisotopic for every residue
 *   ms2_mass, dta_precursor   m/z <-> neutral mass identities, exact decimals
 *   mod_info              hand-derived vectors, itertools.product
 *   digest, sort_idx, prepare_ions   hand examples, stability, bit packing
 *   windows (both)        linear scans of the sorted residuals
 *   prepare_scoring       concatenation of the windows
 *   choose, unrank_comb   the reference's Pascal table rows, math.comb, the
 *                         combinatorial number system (and its reflection)
 *   find_modable_peptides linear scans + bytes.count
 *   calc_total_mod_cands  enumeration of occurrence subsets
 *   prepare_gen_mod       explicit expansion, itertools.accumulate
 *   add_ma_ranks, gen_mod_cands   mixed-radix digits and their bijection
 *   peaks_list            exact decimal m/z and bins
 *   peaks_dense           per-bin max of the peak list
 *   mpeaks_dense          hand-modified peptides, exact decimal bins
 *   score_fx, score_f32   float64 dot, error bound
 *   preprocess            closed forms on constructed spectra
 *   search_plain, search_mods   brute-force searches composed of the above
 *   check_fast_div        a planted wrong reciprocal it must detect
 */
#ifndef HFX_ORACLE_H
#define HFX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* fixed-point scale of the exact score: score_fx = sum round(w*x*2^36) */
#define ORC_FX_SHIFT 36

/* ---- mass tables, configuration ------------------------------------- */
void  orc_mass_table(const float *static_add26, int mono, float *out26);
float orc_ms2_mass(float pcr, float charge);
float orc_dta_precursor(float mh, float charge);
int   orc_mod_info(const float *var26, uint32_t max_ma,
                   uint8_t *out_ma, float *out_ma_mass, uint32_t *out_num_ma,
                   uint32_t *out_mod_ma_count, uint32_t *out_mod_sum,
                   float *out_mod_delta, uint32_t cap_mod, uint32_t *out_num_mod);

/* ---- database ------------------------------------------------------- */
int  orc_digest(const uint8_t *seq, const uint32_t *seg, uint32_t nprot,
                const float *mass26, const uint8_t *rule26, uint32_t missed,
                float min_mass, float max_mass,
                int p_rule, uint32_t min_len, uint32_t max_len,
                float *out_r, uint32_t *out_tc, uint32_t *out_tn,
                uint64_t cap, uint64_t *out_n);
void orc_sort_idx(const float *keys, uint32_t n, uint32_t init, uint32_t *out_idx,
                  float *out_sorted_keys);
void orc_prepare_ions(uint8_t *ions, uint64_t n, const uint8_t *ma, uint32_t num_ma);

/* ---- windows -------------------------------------------------------- */
uint32_t orc_find_spec_cands_by_mass(uint32_t *begin, uint32_t *end, uint32_t *num,
                                     const float *r, const uint32_t *idx_sorted,
                                     uint32_t num_pep, const float *masses,
                                     uint32_t num_spec, float eps);
void orc_prepare_scoring(uint32_t *out, const uint32_t *idx_sorted,
                         const uint32_t *begin, const uint32_t *num, uint32_t num_spec);
uint32_t orc_find_begin_end(uint32_t *begin, uint32_t *end, uint32_t *num_scan,
                            const float *r, const uint32_t *idx_sorted, uint32_t num_pep,
                            const float *mod_delta, uint32_t num_mod, float mass, float eps);

/* ---- modified-peptide enumeration ----------------------------------- */
uint32_t orc_choose(uint32_t n, uint32_t k);
void     orc_unrank_comb(uint32_t *ans, uint32_t n, uint32_t k, uint32_t rank);
uint32_t orc_find_modable_peptides(uint32_t *out_pep_idx, uint32_t *out_pep_mod_idx,
                                   uint32_t *out_pep_ma_count, uint32_t *out_spec_num_valid_scan,
                                   const uint32_t *spec_num_cand_scan, uint32_t num_spec,
                                   uint32_t num_cand_total, const uint8_t *ions,
                                   const uint32_t *tc, const uint32_t *tn,
                                   const uint32_t *idx_sorted, const uint32_t *begin,
                                   const uint32_t *end, const uint32_t *num_pep_scan,
                                   const uint32_t *mod_ma_count, uint32_t num_mod,
                                   const uint8_t *ma, uint32_t num_ma);
uint32_t orc_calc_total_mod_cands(uint32_t *out_pep_num_mpep, uint32_t *out_pep_ma_num_comb_scan,
                                  const uint32_t *mod_ma_count, const uint32_t *pep_idx,
                                  const uint32_t *pep_mod_idx, const uint32_t *pep_ma_count,
                                  uint32_t num_cand, uint32_t num_mod, uint32_t num_ma);
uint32_t orc_prepare_gen_mod(uint32_t *out_mpep_pep_idx, uint32_t *out_mpep_pep_mod_idx,
                             uint32_t *out_mpep_rank, uint32_t *out_mpep_ith_cand,
                             uint32_t *out_mpep_mod_ma_count_sum_scan,
                             const uint32_t *mod_ma_count_sum, const uint32_t *pep_idx,
                             const uint32_t *pep_mod_idx, const uint32_t *pep_num_mpep,
                             uint32_t num_cand, uint32_t num_mpep);
/* the per-acid ranks of one modified-peptide rank (mixed radix over the
 * suffix products scan_row[i] = prod_{j>=i} C(D_j, d_j)) */
void orc_add_ma_ranks(uint32_t *ranks, uint32_t rank, const uint32_t *scan_row, uint32_t num_ma);
void orc_gen_mod_cands(uint32_t *out_unrank, const uint32_t *mod_ma_count,
                       const uint32_t *mpep_ith_cand, const uint32_t *mpep_rank,
                       const uint32_t *mpep_mod_ma_count_sum_scan,
                       const uint32_t *pep_mod_idx, const uint32_t *pep_ma_count,
                       const uint32_t *pep_ma_num_comb_scan, uint32_t num_mpep, uint32_t num_ma);

/* ---- theoretical spectra -------------------------------------------- */
/* dense max-weight row (values 0/10/25/50) of an unmodified candidate */
void orc_peaks_dense(uint8_t *w, uint32_t len, const uint8_t *ions, uint32_t tc, uint32_t tn,
                     float r, const float *mass26, uint32_t zmax);
/* same, for one modified peptide (unrank = this mpep's slice) */
void orc_mpeaks_dense(uint8_t *w, uint32_t len, const uint8_t *ions, uint32_t tc, uint32_t tn,
                      float r, const float *mass26, uint32_t zmax,
                      const uint32_t *mod_ma_count_row, float mod_delta,
                      const uint8_t *ma, const float *ma_mass, uint32_t num_ma,
                      const uint32_t *unrank);
/* every peak (bin, weight) in generation order before the per-bin max */
uint32_t orc_peaks_list(int32_t *bins, uint8_t *wts, uint32_t cap, const float *masses,
                        uint32_t n, float r, uint32_t zmax);

/* ---- scores --------------------------------------------------------- */
int64_t orc_score_fx(const uint8_t *w, const float *x, uint32_t len);
float   orc_score_f32(const uint8_t *w, const float *x, uint32_t len);
double  orc_fx_to_double(int64_t fx);

/* ---- arithmetic checks ---------------------------------------------- */
/* Counts the floats x in [lo, hi) (lo, hi >= 0) for which
 * q = x*R, r = fma(-q, d, x), fma(r, R, q) with R = RN(1/d) differs from the
 * correctly rounded x / d (the division the GPU ladders use); exhaustive.
 * corrected = 0 checks q = x*R alone (a deliberately wrong formula, so that
 * tests can see the checker catch it). */
uint64_t orc_check_fast_div(float d, float lo, float hi, int corrected);

/* ---- spectrum preprocessing ----------------------------------------- */
/* mean_mode 0: reference FP32 sequential window sum; 1: FP64 window sum */
int orc_preprocess(float *out, uint32_t cap, uint32_t *out_len, float pcr, float charge,
                   const float *mz, const float *inten, uint32_t npk,
                   int remove_precursor, int mono, int mean_mode);

/* ---- whole searches (parity at scale, CPU baseline) ------------------ */
int orc_search_plain(const uint8_t *ions, const float *r, const uint32_t *tc, const uint32_t *tn,
                     const float *mass26, const uint32_t *idx_sorted, uint32_t num_pep,
                     const float *x, uint32_t x_stride, const uint32_t *len,
                     const uint32_t *zmax, const float *mass, uint32_t num_spec, float eps,
                     uint32_t topn,
                     uint32_t *out_begin, uint32_t *out_num,
                     uint32_t *out_top_pos, int64_t *out_top_fx, uint32_t *out_top_cnt,
                     float *out_all_f32, int64_t *out_all_fx, const uint64_t *all_off,
                     int nthreads);

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
                    uint64_t *out_num_mpep, uint64_t *out_len_unrank, int nthreads);

/* ---- E-values and SP ions ------------------------------------------ */
/* (alpha, beta) of one spectrum's scores; returns the histogram points */
uint32_t orc_evalue_fit(const float *scores, uint64_t n, uint32_t bin_width, float perc,
                        double *alpha, double *beta, uint64_t *nused);
double   orc_evalue(double alpha, double beta, double score);
/* out[len/3 + 1]: 1 where extractPeaks leaves a value above zero */
void     orc_extract_peaks(const float *x, uint32_t len, uint8_t *out);
void     orc_match_ions(const uint8_t *ions, uint64_t num_ions, uint32_t tc, uint32_t tn,
                        float res, const float *mass26, uint32_t zmax, int mono,
                        const uint8_t *on, uint32_t npk, uint32_t nmod, uint32_t *out_nb,
                        uint32_t *out_ny, uint32_t *out_total, uint8_t *b_flags,
                        uint8_t *y_flags);

#ifdef __cplusplus
}
#endif
#endif
