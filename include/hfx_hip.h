/*
 * hfx_hip.h -- C ABI of libhfx_hip.so, the MI355X-native SEQUEST search core.
 *
 * Drop-in for HFX's GPU search path.  Two groups of entry points:
 *
 *  1. The reference's own C ABI (src/common/include/algorithms.h), same names,
 *     same argument meaning, same return convention, so that
 *     src/haskell/Foreign/CUDA/Algorithms.hs binds them unchanged.  These run
 *     on the null stream (except addModIons, which takes a stream like the
 *     reference) and the count-returning ones synchronise, as the reference's
 *     Thrust-based versions do.  Errors: the reference has none; here a failure
 *     returns 0 / does nothing and records a message read by hfx_last_error().
 *
 *  2. New hfx_* entry points with an int status (HFX_OK == 0), an explicit
 *     stream (a hipStream_t passed as void*), and caller-owned buffers; scratch
 *     comes from an hfx_ctx created once.  They replace the dense
 *     addIons + cuBLAS sgemv + b40c sort sequence of Search.hs:177-277 with a
 *     fused sparse XCorr and a top-N select, move fast-XCorr preprocessing
 *     (Correlation.hs:37-109) onto the GPU, and replace the CUDA binding
 *     package (packages/cuda-0.4.1.0) with thin memory/stream helpers.
 *
 * All pointers named d_* are device pointers; h_* are host pointers.
 */
#ifndef HFX_HIP_H
#define HFX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes                                                          */
/* ------------------------------------------------------------------ */
#define HFX_OK 0
#define HFX_ERR_ARG 1     /* invalid argument / size beyond a stated limit */
#define HFX_ERR_HIP 2     /* a HIP runtime call failed */
#define HFX_ERR_NOMEM 3   /* device allocation failed */
#define HFX_ERR_DATA 4    /* input data the reference cannot process */

typedef struct hfx_ctx hfx_ctx;

/* Device-resident peptide database (Fragment.hs:178-188 DeviceSeqDB plus the
 * sorted residual copy that sort_idx_f leaves behind, Fragment.hs:250-257). */
typedef struct hfx_db {
    const uint8_t *d_ions;        /* packed residues, prepare_ions layout        */
    const float *d_r;             /* residual mass per peptide (no water)        */
    const uint32_t *d_tc;         /* first residue index                         */
    const uint32_t *d_tn;         /* ladder end (Fragment.hs:386-390)            */
    const float *d_mass;          /* [26] residue mass table incl. static mods   */
    const uint32_t *d_idx_sorted; /* peptide indices sorted by residual          */
    const float *d_r_sorted;      /* residuals in that order                     */
    uint32_t num_pep;
    uint64_t num_ions;
} hfx_db;

/* Preprocessed spectra resident on the device. */
typedef struct hfx_spectra {
    const float *d_x;        /* [num_spec * x_stride] fast-XCorr arrays           */
    uint32_t x_stride;       /* floats between consecutive spectra (>= max len)   */
    const uint32_t *d_len;   /* [num_spec] array length (Correlation.hs:62-67)    */
    const uint32_t *d_zmax;  /* [num_spec] max fragment charge max(1, z-1)        */
    const float *d_mass;     /* [num_spec] window centre ms2Mass (Data.hs:58-59)  */
    uint32_t num_spec;
} hfx_spectra;

/* Modified peptides of a batch, as genModCands and its predecessors leave
 * them (Search.hs:358-394): per modified peptide its peptide, modification
 * and the offset of its unrank slice. */
typedef struct hfx_modset {
    const uint32_t *d_mpep_pep_idx;    /* [num_mpep]                            */
    const uint32_t *d_mpep_mod_idx;    /* [num_mpep]                            */
    const uint32_t *d_mpep_unrank_off; /* [num_mpep] mpep_mod_ma_count_sum_scan */
    const uint32_t *d_unrank;          /* [len_unrank]                          */
    const uint32_t *d_mod_ma_count;    /* [num_mod * num_ma]                    */
    const float *d_mod_delta;          /* [num_mod]                             */
    const float *d_ma_mass;            /* [num_ma]                              */
    uint32_t num_ma;
    uint64_t num_mpep;
} hfx_modset;

/* ------------------------------------------------------------------ */
/* Runtime: replaces packages/cuda-0.4.1.0 (Runtime/Marshal.chs:126-376, */
/* Runtime/Stream.chs:54-89, Runtime/Device.chs:267-273,                 */
/* Driver/Marshal.chs:548) and the cuBLAS handle of Execution.hs:49-53.  */
/* ------------------------------------------------------------------ */
const char *hfx_last_error(void);
int hfx_ctx_create(int device, hfx_ctx **out);
int hfx_ctx_destroy(hfx_ctx *ctx);
int hfx_malloc(void **d_ptr, size_t bytes);
int hfx_free(void *d_ptr);
int hfx_host_alloc(void **h_ptr, size_t bytes); /* pinned, Marshal.chs mallocHostArray */
int hfx_host_free(void *h_ptr);
int hfx_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes, void *stream);
int hfx_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes, void *stream);
int hfx_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream);
int hfx_memset(void *d_dst, int value, size_t bytes, void *stream);
int hfx_stream_create(void **stream);
int hfx_stream_destroy(void *stream);
int hfx_stream_sync(void *stream);
int hfx_device_sync(void);
int hfx_mem_info(size_t *free_bytes, size_t *total_bytes);
int hfx_device_count(int *count);
/* Runtime/Event.chs: events for timing work on a stream. */
int hfx_event_create(void **ev);
int hfx_event_destroy(void *ev);
int hfx_event_record(void *ev, void *stream);
int hfx_event_elapsed_ms(float *ms, void *start, void *stop);
/* Context.chs:38-62 cublasCreate/Destroy/SetStream: the handle mvm_ff takes
 * (it carries only the stream) */
int hfx_blas_create(void **handle);
int hfx_blas_destroy(void *handle);
int hfx_blas_set_stream(void *handle, void *stream);

/* ------------------------------------------------------------------ */
/* Reference ABI (algorithms.h), same names and meaning.                 */
/* ------------------------------------------------------------------ */

/* algorithms.h:48, sort.cu:28-35: d_idx = init..init+N-1 stably sorted by the
 * keys (ascending); the keys are sorted in place. */
void sort_idx_f(float *d_keys_raw, uint32_t *d_idx_raw, uint32_t N, uint32_t init);

/* algorithms.h:42, sort.cu:20-26: stable ascending key-value sort in place. */
void sort_val_f(float *d_keys, uint32_t *d_vals, uint32_t N);

/* algorithms.h:54, sort_b40c.cu:34-41: stable ascending key-value sort in
 * place (b40c radix sort; the alternate-buffer selector quirk is not kept). */
void sort_b40c_f(float *d_keys, uint32_t *d_vals, uint32_t N);

/* algorithms.h:37, sort.cu:59-76: stable descending key-value sort in place
 * (thrust::greater keys; equal keys keep their order). */
void sort_rf(float *d_keys, uint32_t *d_vals, uint32_t N);

/* algorithms.h:49, sort.cu:37-57: d_idx = init..init+N-1 stably sorted by the
 * keys, descending; the keys are sorted in place. */
void sort_idx_rf(float *d_keys_raw, uint32_t *d_idx_raw, uint32_t N, uint32_t init);

/* algorithms.h:89-97, filter.cu:38-73: the indices i (ascending) with
 * min <= d_in[i] <= max written to d_out; returns their number. */
uint32_t findIndicesInRange_f(const float *d_in, uint32_t *d_out, const uint32_t N,
                              const float min, const float max);

/* algorithms.h:293, mvm.cu:21-56: d_y = A x for row-major A (m rows of n)
 * on the stream of `handle` (an hfx_blas_create handle in place of the
 * cuBLAS one: Algorithms.hs:370 passes CUBLAS.Handle, here a Ptr ()). */
void mvm_ff(void *handle, float *d_y, const float *d_A, const float *d_x, const uint32_t m,
            const uint32_t n);

/* algorithms.h:80-87, prepare_ions.cu:105-132: pack each residue byte as
 * (c - 'A') | mod << 5 (all N bytes, also N < 512). */
void prepare_ions(uint8_t *d_ions, uint32_t N, uint8_t *d_ma, uint32_t num_ma);

/* algorithms.h:99-113, filter.cu:75-127: per spectrum the residual-sorted
 * window [lower_bound(m-eps), upper_bound(m+eps)); returns the total. */
uint32_t findSpecCandsByMass(uint32_t *d_spec_begin_raw, uint32_t *d_spec_end_raw,
                             uint32_t *d_spec_num_pep_raw, const float *d_r_raw,
                             const uint32_t *d_pep_idx_r_sorted_raw, const uint32_t num_pep,
                             const float *d_spec_masses_raw, const uint32_t num_spec,
                             const float eps);

/* algorithms.h:115-131, filter_by_mod.cu:35-143: per modification the range
 * [(m-eps)-delta, (m+eps)-delta], exclusive scan of counts; returns total. */
uint32_t findBeginEnd_f(uint32_t *d_begin_raw, uint32_t *d_end_raw, uint32_t *d_num_pep_scan_raw,
                        const float *d_r_raw, const uint32_t *d_pep_idx_r_sorted_raw,
                        const uint32_t num_pep, const float *d_mod_delta_raw,
                        const uint32_t num_mod, const float mass_, const float eps);

/* algorithms.h:133-159, filter_by_mod.cu:265-378. */
uint32_t findModablePeptides(uint32_t *d_out_pep_idx_raw, uint32_t *d_out_pep_mod_idx_raw,
                             uint32_t *d_out_pep_ma_count_raw, uint32_t *d_out_spec_num_valid_scan,
                             const uint32_t *d_spec_num_cand_scan, const uint32_t num_spec,
                             uint32_t num_cand_total, const uint8_t *d_ions,
                             const uint32_t *d_tc, const uint32_t *d_tn,
                             const uint32_t *d_pep_idx_r_sorted_raw, const uint32_t *d_begin_raw,
                             const uint32_t *d_end_raw, const uint32_t *d_num_pep_scan_raw,
                             const uint32_t *d_mod_ma_count, const uint32_t num_mod,
                             const uint8_t *d_ma, const uint32_t num_ma);

/* algorithms.h:161-173, calc_mod_cands.cu:76-119. */
uint32_t calcTotalModCands(uint32_t *d_out_pep_num_mpep_raw, uint32_t *d_out_pep_ma_num_comb_scan_raw,
                           const uint32_t *d_mod_ma_count_raw, const uint32_t *d_pep_idx_raw,
                           const uint32_t *d_pep_mod_idx_raw, const uint32_t *d_pep_ma_count_raw,
                           const uint32_t nPep, const uint32_t num_mod, const uint32_t num_ma);

/* algorithms.h:175-192, prepare_gen_mod.cu:79-129. */
uint32_t prepareGenMod(uint32_t *d_out_mpep_pep_idx, uint32_t *d_out_mpep_pep_mod_idx,
                       uint32_t *d_out_mpep_rank, uint32_t *d_out_mpep_ith_valid,
                       uint32_t *d_out_mpep_mod_ma_count_sum_scan,
                       const uint32_t *d_mod_ma_count_sum, const uint32_t *d_pep_idx,
                       const uint32_t *d_pep_mod_idx, const uint32_t *d_pep_num_mpep,
                       const uint32_t num_pep, const uint32_t num_mpep);

/* algorithms.h:194-212, gen_mod_pep.cu:229-312. */
void genModCands(uint32_t *d_out_mpep_unrank, const uint32_t *d_mod_ma_count,
                 const uint32_t *d_mpep_ith_valid, const uint32_t *d_mpep_rank,
                 const uint32_t *d_mpep_mod_ma_count_sum_scan, const uint32_t *d_pep_mod_idx,
                 const uint32_t *d_pep_ma_count, const uint32_t *d_pep_ma_num_comb_scan,
                 const uint32_t total, const uint32_t mod_num_ma);

/* algorithms.h:214-241, mion_series.cu:443-556: dense theoretical rows of
 * modified peptides (max per bin, written as floats 0/10/25/50). */
void addModIons(float *d_out_mspec, const float *d_residual, const float *d_mass,
                const uint8_t *d_ions, const uint32_t *d_tc, const uint32_t *d_tn,
                const uint32_t *d_mpep_pep_idx, const uint32_t *d_mpep_pep_mod_idx,
                const uint32_t *d_mpep_unrank, const uint32_t *d_mpep_mod_ma_count_sum_scan,
                const uint32_t len_unrank, const uint32_t num_mpep,
                const uint32_t *d_mod_ma_count, const float *d_mod_delta, const uint8_t *d_ma,
                const float *d_ma_mass, const uint32_t num_ma, const uint32_t len_spec,
                const uint32_t max_charge, void *stream);

/* algorithms.h:247-260, ion_series.cu:319-343: dense theoretical rows. */
void addIons(float *d_spec, const float *d_residual, const float *d_mass, const uint8_t *d_ions,
             const uint32_t *d_tc, const uint32_t *d_tn, const uint32_t *d_idx,
             const uint32_t num_idx, const uint32_t max_charge, const uint32_t len_spec);

/* algorithms.h:279-287, prepare_scoring.cu:47-67. */
void prepare_scoring(uint32_t *d_out_spec_pep_idx, const uint32_t *d_pep_idx_r_sorted_raw,
                     const uint32_t *d_spec_begin, const uint32_t *d_spec_num_pep,
                     const uint32_t num_spec);

/* ------------------------------------------------------------------ */
/* Host helpers: the Haskell host functions whose outputs the kernels    */
/* consume, for callers without the Haskell driver.                      */
/* ------------------------------------------------------------------ */

/* Config.hs:266-272 initializeAAMasses: out26[c] = static_add26[c] + mass(c). */
int hfx_mass_table(const float *static_add26, int mono, float *out26);

/* Config.hs:415-431 getDigestionRule: mask26[c] = 1 for cleavage residues. */
int hfx_digestion_rule(int rule, uint8_t *mask26);

/* Fragment.hs:363-405 digest (fragment + splice + mass filter) of the
 * proteins seq[seg[p] .. seg[p+1]); p_rule / min_len / max_len are harness
 * extensions (0 = off).  Writes (r, tc, tn) per peptide. */
int hfx_digest(const uint8_t *h_seq, const uint32_t *h_seg, uint32_t nprot, const float *h_mass26,
               const uint8_t *h_rule26, uint32_t missed, float min_mass, float max_mass, int p_rule,
               uint32_t min_len, uint32_t max_len, float *h_out_r, uint32_t *h_out_tc,
               uint32_t *h_out_tn, uint64_t cap, uint64_t *out_n);

/* Fragment.hs:260-289 makeModInfo (modable acids, modification vectors,
 * their count sums and mass deltas). */
int hfx_mod_info(const float *h_var26, uint32_t max_ma, uint8_t *h_out_ma, float *h_out_ma_mass,
                 uint32_t *out_num_ma, uint32_t *h_out_mod_ma_count, uint32_t *h_out_mod_sum,
                 float *h_out_mod_delta, uint32_t cap_mod, uint32_t *out_num_mod);

/* Spectrum/Data.hs:58-59 ms2Mass (window centre) and DTA.hs:51 precursor. */
float hfx_ms2_mass(float pcr, float charge);
float hfx_dta_precursor(float mh, float charge);

/* ------------------------------------------------------------------ */
/* New entry points                                                      */
/* ------------------------------------------------------------------ */

/* Replaces CUBLAS.sgemv(T) of Search.hs:242-245 / :527-529
 * (packages/cublas/.../Level2.chs:65-77): d_y[i] = sum_j d_A[i*lda + j] * d_x[j]
 * for i < m, j < n, row-major A (the dense rows addIons writes). */
int hfx_mvm_rows(hfx_ctx *ctx, float *d_y, const float *d_A, uint32_t lda, const float *d_x,
                 uint32_t m, uint32_t n, void *stream);

/* Fast-XCorr preprocessing on the GPU, Correlation.hs:37-109, for num_spec
 * spectra whose peaks (sorted by m/z) are d_mz/d_int[d_peak_off[s] ..
 * d_peak_off[s+1]).  Writes d_x[s*x_stride .. +len) (zero padded to
 * x_stride), d_len[s] and d_status[s] (0 ok, 2 = the reference's sel == 0
 * division by zero, 3 = len > x_stride, 1 = no peaks).  mean_mode 0 sums the
 * 151-bin window in FP32 like the reference; 1 sums it in FP64. */
int hfx_preprocess(hfx_ctx *ctx, float *d_x, uint32_t x_stride, uint32_t *d_len,
                   uint32_t *d_status, const float *d_mz, const float *d_int,
                   const uint32_t *d_peak_off, const float *d_pcr, const float *d_charge,
                   uint32_t num_spec, int remove_precursor, int mono, int mean_mode,
                   void *stream);

/* Window filter for all spectra of a batch in one launch (findSpecCandsByMass
 * without the synchronous return): d_begin/d_num per spectrum. */
int hfx_window(hfx_ctx *ctx, uint32_t *d_begin, uint32_t *d_num, const hfx_db *db,
               const float *d_spec_mass, uint32_t num_spec, float eps, void *stream);

/* Exclusive scan of n uint32 counts into n+1 uint64 offsets. */
int hfx_scan_offsets(hfx_ctx *ctx, uint64_t *d_off, const uint32_t *d_count, uint32_t n,
                     void *stream);

/* Plain FP32 sparse XCorr (replaces addIons + sgemv, Search.hs:177-245):
 * for spectrum s and its i-th window candidate (residual-sorted position
 * d_begin[s] + i), d_scores[d_off[s] + i] = sum over the candidate's
 * theoretical bins b < len_s of maxweight(b) * x_s[b], FP32.  `total` is
 * sum_s d_num[s] (findSpecCandsByMass's return value). */
int hfx_xcorr_sparse(hfx_ctx *ctx, float *d_scores, const uint64_t *d_off,
                     const uint32_t *d_begin, const uint32_t *d_num, uint64_t total,
                     const hfx_db *db, const hfx_spectra *sp, void *stream);

/* hfx_xcorr_sparse for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops): no counterpart in the
 * reference.  Same arguments and results as hfx_xcorr_sparse, without the
 * context: the tile table (a few bytes per spectrum) lives in one static
 * device array of the library, so at most HFX_NOALLOC_MAX_SPECTRA spectra per
 * call (larger calls return HFX_ERR_ARG), and no hipMalloc or other
 * allocating routine is called.  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (different streams, threads) race.
 * The first call on a device looks the array's address up (not under
 * capture). */
#define HFX_NOALLOC_MAX_SPECTRA (1u << 20)
int hfx_xcorr_sparse_noalloc(float *d_scores, const uint64_t *d_off, const uint32_t *d_begin,
                             const uint32_t *d_num, uint64_t total, const hfx_db *db,
                             const hfx_spectra *sp, void *stream);

/* Per-spectrum top-N (replaces sort_b40c_f + full D2H, Search.hs:250-276):
 * selects the top N + margin candidates by FP32 score, re-scores them exactly
 * (int64 fixed point, 2^-36 resolution, reported as double), and writes the
 * best N ordered by (exact score desc, candidate position asc):
 * d_top_pos[s*N + k] (position within the spectrum's window), d_top_exact,
 * d_top_f32, d_top_cnt[s].  Every candidate whose FP32 score lies within twice
 * the FP32 error bound below the N-th best FP32 score is re-scored too (every
 * candidate that could reach the N-th exact score is among them); a spectrum
 * with more such candidates than the re-score set holds (exact ties) takes the
 * exact route on the device: all its candidates scored exactly and the top N
 * selected on the exact scores; d_flag[s] has HFX_FLAG_EXACT set for those
 * spectra.  The exact route keeps every candidate whose FP32 score lies at
 * most the error bound below the N-th exact score of the re-score set; if
 * fewer than min(N, d_num[s]) remain, some FP32 score broke the bound: only
 * those are ranked, d_top_cnt[s] counts them and HFX_FLAG_SHORT is set.
 * `total` = sum_s d_num[s].  N <= 64, N + margin <= 128. */
#define HFX_FLAG_EXACT 1u
#define HFX_FLAG_SHORT 2u
int hfx_topn(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
             uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N, uint32_t margin,
             const float *d_scores, const uint64_t *d_off, uint64_t total,
             const uint32_t *d_begin, const uint32_t *d_num, const hfx_db *db,
             const hfx_spectra *sp, void *stream);

/* hfx_topn for scores with a caller-stated error bound: every score of
 * spectrum s is within (rel0 + rel1 * z_s) * 50 * sum_b |x_s[b]| + 1e-6 of its
 * exact value, z_s = fragment charges generated for s.  hfx_topn is
 * hfx_topn_rel with rel0 = 80 * 2^-24, rel1 = 0 (FP32 sparse sums); scores of
 * hfx_xcorr_dense_bf16x2 need rel0 = 2^-16 + 210 * 2^-24, rel1 = 0. */
int hfx_topn_rel(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                 uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N, uint32_t margin, float rel0,
                 float rel1, const float *d_scores, const uint64_t *d_off, uint64_t total,
                 const uint32_t *d_begin, const uint32_t *d_num, const hfx_db *db,
                 const hfx_spectra *sp, void *stream);

/* Precision form of the XCorr (new; replaces addIons + sgemv like
 * hfx_xcorr_sparse, same output layout): spectra are taken eight at a time in
 * the order d_order (null: index order; sort by fragment charge, then
 * precursor mass, for speed), every candidate of the eight windows gets one
 * theoretical row, and rows x spectra are contracted on the MFMA units with
 * the weights in BF16 (exact) and each spectrum split into hi + lo BF16,
 * FP32 accumulation.  |score - exact| <= (2^-16 + 210 * 2^-24) * 50 *
 * sum_b |x[b]|.  Needs sp->x_stride <= 2176; peptides of at most 768
 * residues. */
int hfx_xcorr_dense_bf16x2(hfx_ctx *ctx, float *d_scores, const uint64_t *d_off,
                           const uint32_t *d_begin, const uint32_t *d_num, const uint32_t *d_order,
                           const hfx_db *db, const hfx_spectra *sp, void *stream);

/* Integer precision form (new; same layout as hfx_xcorr_sparse): blocks of
 * sixteen spectra in the order d_order, one theoretical row per candidate of
 * the block's union window, contracted on the MFMA units in int8: weights
 * 0/10/25/50 exactly, each spectrum as three balanced base-256 digits of
 * X = rint(x * 2^k) (k the largest with max|x| * 2^k <= 8355711), int32
 * sums exact, the score formed in FP64 and rounded once to FP32.
 * |score - exact| <= max|x| / 8355711 * 50 * len + 2^-24 * 50 * sum|x|
 * (top-N: hfx_topn_rel_q with rel0 = 2^-24, rel1 = 0, relq = 1 / 8355711).
 * Needs sp->x_stride <= 2176. */
int hfx_xcorr_dense_i8x3(hfx_ctx *ctx, float *d_scores, const uint64_t *d_off,
                         const uint32_t *d_begin, const uint32_t *d_num, const uint32_t *d_order,
                         const hfx_db *db, const hfx_spectra *sp, void *stream);

/* hfx_topn_rel with one more term in the bound: relq * max_b |x_s[b]| * 50
 * * len_s (scores whose error follows the spectrum's largest value, as the
 * integer form's do). */
int hfx_topn_rel_q(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                   uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N, uint32_t margin, float rel0,
                   float rel1, float relq, const float *d_scores, const uint64_t *d_off,
                   uint64_t total, const uint32_t *d_begin, const uint32_t *d_num,
                   const hfx_db *db, const hfx_spectra *sp, void *stream);

/* Exact top-N of the listed spectra (every candidate scored in fixed point;
 * d_top_f32 taken from d_scores). */
int hfx_topn_exact(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                   uint32_t *d_top_cnt, uint32_t N, const uint32_t *d_spec_list,
                   uint32_t num_list, const float *d_scores, const uint64_t *d_off,
                   uint64_t total, const uint32_t *d_begin, const uint32_t *d_num,
                   const hfx_db *db, const hfx_spectra *sp, void *stream);

/* findBeginEnd_f (filter_by_mod.cu:35-143) for num_spec spectra in one
 * launch, where Search.hs:317-322 calls it once per spectrum: for spectrum s
 * and modification i, [d_begin[s*M+i], d_end[s*M+i]) is the residual-sorted
 * window of peptides with (m_s - eps) - delta_i <= r <= (m_s + eps) - delta_i
 * (FP32), d_num_scan[s*M+i] the exclusive scan of the window sizes over i and
 * d_spec_cand[s] their sum (findBeginEnd_f's return).  M = num_mod.  No host
 * synchronisation. */
int hfx_mod_windows(hfx_ctx *ctx, uint32_t *d_begin, uint32_t *d_end, uint32_t *d_num_scan,
                    uint32_t *d_spec_cand, const hfx_db *db, const float *d_spec_mass,
                    uint32_t num_spec, const float *d_mod_delta, uint32_t num_mod, float eps,
                    void *stream);

/* Search.hs:381-386 on the device: spectrum s owns the modified peptides of
 * rows [d_spec_num_valid_scan[s], d_spec_num_valid_scan[s+1]) (the last bound
 * is num_rows), i.e. d_mp_off[s] = sum of d_pep_num_mpep over the rows before
 * them; d_mp_off has num_spec + 1 entries. */
int hfx_mpep_offsets(hfx_ctx *ctx, uint64_t *d_mp_off, const uint32_t *d_pep_num_mpep,
                     uint32_t num_rows, const uint32_t *d_spec_num_valid_scan, uint32_t num_spec,
                     void *stream);

/* Sparse FP32 XCorr of modified peptides (replaces addModIons + sgemv,
 * Search.hs:477-529): d_scores[d_mp_off[s] + k] for the k-th modified
 * peptide of spectrum s. */
int hfx_xcorr_sparse_mod(hfx_ctx *ctx, float *d_scores, const uint64_t *d_mp_off, uint64_t total,
                         const hfx_modset *ms, const hfx_db *db, const hfx_spectra *sp,
                         void *stream);

/* hfx_topn / hfx_topn_exact over modified peptides; positions are indices
 * within the spectrum's block of modified peptides (Search.hs:534-622). */
int hfx_topn_mod(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                 uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N, uint32_t margin,
                 const float *d_scores, const uint64_t *d_mp_off, const hfx_modset *ms,
                 const hfx_db *db, const hfx_spectra *sp, void *stream);
int hfx_topn_exact_mod(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                       uint32_t *d_top_cnt, uint32_t N, const uint32_t *d_spec_list,
                       uint32_t num_list, const float *d_scores, const uint64_t *d_mp_off,
                       const hfx_modset *ms, const hfx_db *db, const hfx_spectra *sp,
                       void *stream);

/* ------------------------------------------------------------------ */
/* E-values and SP ions of the matches (Search.hs:115-141, :292-299).   */
/* The reference computes both on the host, over every score of a       */
/* spectrum copied back from the device (Search.hs:250-276).            */
/* ------------------------------------------------------------------ */
/* Replaces Util/Statistics.hs:37-100 evalues/linearFitLine/histogram:
 * per spectrum s the (alpha, beta) of the least-squares line through the
 * histogram of its scores -- the scores of set a in [d_off_a[s],
 * d_off_a[s+1]) pooled with those of the optional set b (the modified
 * matches, Search.hs:118) -- after the reference's cut (sorted positions
 * [round(0.3 n), round((1 - cutoff/100) n)), shift by 1 - min, keep > 4000),
 * in bins of bin_width (evalue-bin-width, Config.hs:100-101).  d_fit[2s] =
 * alpha, d_fit[2s+1] = beta; NaN when fewer than two histogram points (the
 * reference's 0/0).  d_npts (optional): histogram points; d_nused (optional):
 * scores kept.  No host synchronisation or allocation. */
int hfx_evalue_fit(hfx_ctx *ctx, double *d_fit, uint32_t *d_npts, uint64_t *d_nused,
                   const float *d_scores_a, const uint64_t *d_off_a, const float *d_scores_b,
                   const uint64_t *d_off_b, uint32_t num_spec, uint32_t bin_width,
                   float cutoff_perc, void *stream);
/* Statistics.hs:41: d_ev[s][k] = 10 ** (alpha_s + beta_s * d_score[s][k] /
 * 10000) for k < d_cnt[s] (all N when d_cnt is null), 0 beyond. */
int hfx_evalues(hfx_ctx *ctx, double *d_ev, const double *d_fit, const double *d_score,
                const uint32_t *d_cnt, uint32_t N, uint32_t num_spec, void *stream);
/* Replaces IonSeries.hs:99-107 extractPeaks: per spectrum `words` uint32
 * flag words, bit t set when bin t (= position div 3) holds one of the 200
 * largest values and that value is above zero.  words >=
 * hfx_sp_peak_words(x_stride); strides up to 24575. */
uint32_t hfx_sp_peak_words(uint32_t x_stride);
int hfx_sp_peaks(hfx_ctx *ctx, uint32_t *d_on, uint32_t words, const hfx_spectra *sp,
                 void *stream);
/* Replaces IonSeries.hs:136-147 matchIonSequence + Match.hs:33-40 scoreSP
 * (with Location.hs:42-58 lookup) for num_match (peptide, spectrum) pairs:
 * d_nb / d_ny = matched b / y ions, d_total = (residues - 1 + d_nmod) * 2
 * (d_nmod optional: modified sites, Fragment.hs:84-106); a peptide index of
 * 0xFFFFFFFF (or any index past the database, or a spectrum past
 * num_spec) gives zeros.  mono selects the bin width (Correlation.hs:29-30). */
int hfx_sp_ions(hfx_ctx *ctx, uint32_t *d_nb, uint32_t *d_ny, uint32_t *d_total,
                const uint32_t *d_pep, const uint32_t *d_spec, const uint32_t *d_nmod,
                uint64_t num_match, const uint32_t *d_on, uint32_t words, const hfx_db *db,
                const hfx_spectra *sp, int mono, void *stream);
/* hfx_sp_peaks + hfx_sp_ions over the top N of an unmodified search
 * (hfx_topn layout: [S][N], entries beyond d_top_cnt[s] get zeros), flags in
 * the context's scratch (allocated on first use). */
int hfx_sp_ions_topn(hfx_ctx *ctx, uint32_t *d_nb, uint32_t *d_ny, uint32_t *d_total,
                     const uint32_t *d_top_pos, const uint32_t *d_top_cnt,
                     const uint32_t *d_begin, uint32_t N, const hfx_db *db,
                     const hfx_spectra *sp, int mono, void *stream);
/* the same over the top N of hfx_topn_mod: matched on the unmodified
 * peptide (Search.hs:296-299), total grown by the modified sites */
int hfx_sp_ions_topn_mod(hfx_ctx *ctx, uint32_t *d_nb, uint32_t *d_ny, uint32_t *d_total,
                         const uint32_t *d_top_pos, const uint32_t *d_top_cnt,
                         const uint64_t *d_mp_off, uint32_t N, const hfx_modset *ms,
                         const hfx_db *db, const hfx_spectra *sp, int mono, void *stream);

/* ------------------------------------------------------------------ */
/* Host formats on either side of the path (hfx_io.cpp): FASTA -> the    */
/* sequence database and its .index file, DTA/MGF spectra, the display   */
/* lookup and the result tables.  Host memory only.                      */
/* ------------------------------------------------------------------ */
/* Fragment.hs:145-166 SequenceDB: headers, residues, per-protein residue
 * segments, fragments (residual, c, n) and per-protein fragment segments. */
typedef struct hfx_seqdb hfx_seqdb;
/* Fasta.hs:30-35 readFasta (".gz" inflated) + Fragment.hs:197-233
 * makeSeqDB: the proteins of a FASTA file (path) or text (path null), each
 * digested by `rule` (Config.hs:416-430) with `missed` missed cleavages and
 * min_mass <= (M+H)+ <= max_mass. */
int hfx_seqdb_from_fasta(const char *path, const char *text, uint64_t len, const float *h_mass26,
                         int rule, uint32_t missed, float min_mass, float max_mass,
                         hfx_seqdb **out);
int hfx_seqdb_free(hfx_seqdb *db);
/* info = {proteins, residues, header bytes, fragments, parameter bytes} */
int hfx_seqdb_info(const hfx_seqdb *db, uint64_t *info);
/* any output may be null; h_header_off and the segments take count + 1 */
int hfx_seqdb_copy(const hfx_seqdb *db, uint8_t *h_ions, uint32_t *h_ion_seg, char *h_headers,
                   uint64_t *h_header_off, float *h_r, uint32_t *h_tc, uint32_t *h_tn,
                   uint32_t *h_frag_seg, char *h_params);
/* Index.hs:43-63 writeIndex: the parameter block (config-file text) as a
 * Data.Binary ByteString, then the database gzip-compressed in
 * Data.Binary's encoding; Index.hs:70-81 readIndex reads both back. */
int hfx_seqdb_write_index(const hfx_seqdb *db, const char *path, const char *params);
int hfx_seqdb_read_index(const char *path, hfx_seqdb **out);
/* Location.hs:42-58 lookup: fragment k as "X.PEP.Y" (flanks or '-'; the
 * residue at n kept unless n is past the database, Q28), its protein and
 * (M+H)+ = (residual + H) + H2O. */
int hfx_seqdb_lookup(const hfx_seqdb *db, uint32_t k, char *out, uint32_t cap, uint32_t *out_len,
                     uint32_t *out_protein, float *out_mh);
/* Config.hs:416-430: the description line of a digestion rule (0..13) */
const char *hfx_digestion_rule_name(int rule);
/* DTA.hs:36-67: first line MH+ and charge (an integer), then m/z intensity
 * pairs; *out_pcr = (MH + (z - 1) H) / z.  Up to cap peaks copied. */
int hfx_dta_parse(const char *text, uint64_t len, float *out_pcr, float *out_charge, float *h_mz,
                  float *h_inten, uint32_t cap, uint32_t *out_npk);
/* MGF.hs:38-100: BEGIN IONS / TITLE, CHARGE, PEPMASS / peaks / END IONS */
typedef struct hfx_mgf hfx_mgf;
int hfx_mgf_parse(const char *text, uint64_t len, hfx_mgf **out);
/* info = {spectra, peaks, title bytes} */
int hfx_mgf_info(const hfx_mgf *m, uint64_t *info);
int hfx_mgf_copy(const hfx_mgf *m, float *h_pcr, float *h_charge, uint32_t *h_off, float *h_mz,
                 float *h_inten, char *h_titles, uint64_t *h_title_off);
int hfx_mgf_free(hfx_mgf *m);
/* PrettyPrint.hs:83-180.  Numbers as Haskell prints them: kind 0 show,
 * 1 showFFloat (Just digits), 2 showEFloat (Just digits); is_float: the
 * value is a Float (its own shortest digits).  A text that does not fit
 * cap - 1 bytes is cut there and HFX_ERR_ARG returned (it needs at most
 * 320 + digits bytes). */
int hfx_format_number(double v, int is_float, int kind, int digits, char *out, uint32_t cap);
/* ppAsRows q over nrows x ncols NUL-terminated cells (row-major) */
int hfx_format_table(const char *const *cells, uint32_t nrows, uint32_t ncols, int q, char *out,
                     uint64_t cap, uint64_t *out_len);
/* printResultsDetail: the full headers of n matches */
int hfx_format_detail(const char *const *headers, uint32_t n, char *out, uint64_t cap,
                      uint64_t *out_len);
typedef struct hfx_scan_header {
    const char *title;         /* ms2info (empty for DTA files) */
    const char *spectrum_file;
    const char *database;
    const char *enzyme;        /* hfx_digestion_rule_name(rule) */
    float precursor;           /* ms2precursor, m/z */
    float charge;              /* ms2charge */
    float tolerance;           /* mass-tolerance */
} hfx_scan_header;
typedef struct hfx_match_row {
    float mh;                  /* fragmass, (M+H)+ */
    float xcorr;               /* scoreXC / 10000 */
    uint32_t ions_matched, ions_total;
    double evalue;
    const char *reference;     /* fraglabel: first word of the header */
    const char *peptide;       /* "X.PEP.Y" ('*' after modified residues) */
    const char *header;        /* the protein's full header */
} hfx_match_row;
/* printScanResults for one scan: printConfig, printResults of the first
 * nshow rows, printResultsDetail of the first ndetail; out gets at most
 * cap - 1 bytes and a NUL, *out_len the full length. */
int hfx_format_scan(const hfx_scan_header *hdr, const hfx_match_row *rows, uint32_t n,
                    uint32_t nshow, uint32_t ndetail, char *out, uint64_t cap, uint64_t *out_len);

/* ------------------------------------------------------------------ */
/* Batched search (new; Main.hs:129-142 scores 16 spectra per group).    */
/* A batch of any number of spectra is planned once (windows, index      */
/* sizes, every buffer allocated; synchronises), then run without host   */
/* synchronisation or allocation -- directly or as a captured hipGraph.  */
/* Results use the layouts of hfx_xcorr_sparse / hfx_topn.               */
/* ------------------------------------------------------------------ */
/* Per-database peptide index (new): what the batched forms precompute per
 * peptide, built once per database and shared by every batch --
 * HFX_PEPIX_LADDER: the b ladder of every peptide in residual-sorted order
 *   (sequential FP32 prefix of its residue masses, 12 bytes a position; the
 *   dense form's input);
 * HFX_PEPIX_IONS: every peptide's theoretical spectrum for fragment charge
 *   levels 1..levels (level l = charges 1..l, max weight per bin), bins below
 *   max_len, as sorted uint16 entries bin | code << 14 (the FP32 batched
 *   form's input).
 * The reference has no counterpart: it rebuilds every theoretical spectrum
 * per (spectrum, candidate) pair (Search.hs:177-245).  Synchronises. */
typedef struct hfx_pepix hfx_pepix;
#define HFX_PEPIX_LADDER 1u
#define HFX_PEPIX_IONS 2u
int hfx_pepix_build(hfx_ctx *ctx, const hfx_db *db, uint32_t forms, uint32_t levels,
                    uint32_t max_len, void *stream, hfx_pepix **out);
int hfx_pepix_destroy(hfx_pepix *ix);
/* info = {forms, levels, max_len, ladder positions, ion entries, device bytes} */
int hfx_pepix_info(const hfx_pepix *ix, uint64_t *info);

typedef struct hfx_batch hfx_batch;
int hfx_batch_create(hfx_ctx *ctx, hfx_batch **out);
int hfx_batch_destroy(hfx_batch *b);
/* Plans a batch: allocates every buffer for these spectra (synchronises).
 * db and sp must stay valid while the batch is used; hfx_batch_run reads the
 * spectra afresh each time.  Without an attached index (hfx_batch_plan_ix,
 * hfx_batch_use_index) the batch builds its own over the whole database the
 * first time a form needs it (outside any capture). */
int hfx_batch_plan(hfx_batch *b, const hfx_db *db, const hfx_spectra *sp, float eps, uint32_t N,
                   uint32_t margin, void *stream);
int hfx_batch_plan_ix(hfx_batch *b, const hfx_pepix *ix, const hfx_spectra *sp, float eps,
                      uint32_t N, uint32_t margin, void *stream);
/* Device bytes a batch plan of S spectra with `total` scored pairs (the sum
 * of their window sizes, e.g. findSpecCandsByMass's return) allocates in
 * the given precision form, the per-database index excepted (host
 * arithmetic; Search.hs:427-433 sizes its work the same way).
 * hfx_batch_plan fails with HFX_ERR_NOMEM, allocating nothing, when its
 * batch exceeds the device's free memory. */
int hfx_batch_required_bytes(uint32_t S, uint64_t total, uint32_t N, uint32_t num_pep,
                             uint32_t precision, uint64_t *bytes);
/* Execution.hs:77-93 groupScorePlan: cuts spectra 0..S-1 (score offsets
 * h_off[0..S]) into consecutive batches, each the longest run whose plan
 * fits `budget` bytes; h_first (optional, cap entries) receives the first
 * spectrum of each batch and then S; *num_batches their number.  A spectrum
 * that alone exceeds the budget returns HFX_ERR_NOMEM (the reference would
 * split its candidates). */
int hfx_batch_split(const uint64_t *h_off, uint32_t S, uint32_t N, uint32_t num_pep,
                    uint32_t precision, uint64_t budget, uint32_t *h_first, uint32_t cap,
                    uint32_t *num_batches);
int hfx_batch_use_index(hfx_batch *b, const hfx_pepix *ix);
/* One batched search, every per-batch step on the device and without host
 * synchronisation or allocation: windows -> spectrum order (fragment charge
 * level, precursor mass) -> FP32 sparse form: blocks of G spectra,
 * persistent multi-spectrum XCorr over the index's ion lists; dense form:
 * block cut, hi/lo BF16 rows of the spectra, MFMA XCorr over the index's
 * ladders -> top-N with exact re-score. */
int hfx_batch_run(hfx_batch *b, void *stream);
#define HFX_BATCH_FP32_SPARSE 0u  /* default after hfx_batch_plan */
#define HFX_BATCH_BF16X2_DENSE 1u /* hfx_xcorr_dense_bf16x2 scores */
#define HFX_BATCH_I8X3_DENSE 2u   /* hfx_xcorr_dense_i8x3 scores */
int hfx_batch_set_precision(hfx_batch *b, uint32_t precision);
/* capture hfx_batch_run on a non-null stream into a graph; replay it */
int hfx_batch_capture(hfx_batch *b, void *stream);
int hfx_batch_replay(hfx_batch *b, void *stream);
/* time the scoring kernel over (spectra per block G, tile length C,
 * workgroup threads) -- dense form: candidates per work item C -- and keep
 * the fastest; out_cfg = {G, C, threads} ({0, C, workgroup threads} for the
 * dense forms: 1024 in both) */
int hfx_batch_autotune(hfx_batch *b, void *stream, uint32_t *out_cfg);
int hfx_batch_set_config(hfx_batch *b, uint32_t G, uint32_t C, uint32_t threads);
/* one run, phase by phase, with event timings in ms (5 entries): {windows,
 * per-batch preparation (order; dense: block cut + hi/lo rows; sparse:
 * tiles), scoring, top-N, the scoring kernel alone} */
int hfx_batch_time(hfx_batch *b, void *stream, float *ms);
/* info = {spectra, scores, database peptides, charge levels, index entries,
 *         G, C, threads, LDS stride, 0, precision, dense chunk} (12 entries) */
int hfx_batch_info(const hfx_batch *b, uint64_t *info);
/* blocks of the dense form's last run (new): the spectra, in fragment charge
 * then precursor mass order, cut into runs of at most eight (BF16 form) or
 * sixteen (integer form) whose precursor windows overlap most.  *num_blocks = their number; h_bstart (host,
 * optional) receives min(cap, num_blocks + 1) block starts as positions in
 * that order, the last one the spectrum count.  Synchronises. */
int hfx_batch_dense_blocks(const hfx_batch *b, uint32_t *h_bstart, uint32_t cap,
                           uint32_t *num_blocks);
/* the spectrum order of the last run (S entries, host).  Synchronises. */
int hfx_batch_order(const hfx_batch *b, uint32_t *h_order);
/* device pointers {top_pos, top_exact, top_f32, top_cnt, flag, scores,
 *                  score offsets, window begin, window size, index entries} */
int hfx_batch_outputs(const hfx_batch *b, void **ptrs);

#ifdef __cplusplus
}
#endif
#endif
