/*
 * hfx_topn.hip -- per-spectrum top-N with an exact re-score.
 *
 * The reference sorts every score of a spectrum with a b40c radix sort and
 * copies them all to the host (Search.hs:250-276, sort_b40c.cu:34-41); the
 * final order is descending score, ties by ascending candidate position
 * (stable ascending sort, then reverse twice: Search.hs:119-123, :276).
 *
 * Here one workgroup per spectrum:
 *   1. radix-selects the T = min(N + margin, n) best FP32 scores (key desc,
 *      position asc);
 *   2. re-scores them exactly: int64 fixed point, sum of round(w*x*2^36),
 *      independent of summation order, so the GPU and the CPU oracle agree
 *      bit for bit;
 *   3. takes the N-th best exact score exN and re-scores every other candidate
 *      whose FP32 score is >= exN - B (B bounds |FP32 - exact|), so that no
 *      candidate left out can reach the top N;
 *   4. orders the re-scored set by (exact desc, position asc), writes the best
 *      N.  If more than TN_EXT candidates fall within B, the spectrum is
 *      flagged and hfx_topn_exact scores all of its candidates exactly.
 * Bound: |f32 - exact| <= 80 * 2^-24 * sum|w*x| (<= 71 roundings per score)
 * and sum|w*x| <= 50 * sum_b |x_b| (each bin counts once, weights <= 50).
 */
#include "hfx_cand.h"

#define TN_THREADS 256
#define TN_WAVES (TN_THREADS / HFX_WAVE)
#define TN_MAXT 128
#define TN_EXT 256
#define TN_MAXC (TN_MAXT + TN_EXT)

__device__ __forceinline__ bool before_fx(int64_t fa, uint32_t pa, int64_t fb, uint32_t pb)
{
    return fa > fb || (fa == fb && pa < pb);
}

template <class Cands>
__global__ void __launch_bounds__(TN_THREADS)
    k_topn(uint32_t *top_pos, double *top_exact, float *top_f32, uint32_t *top_cnt,
           uint32_t *flag, uint32_t N, uint32_t margin, const float *scores, Cands c,
           hfx_spectra sp)
{
    __shared__ float xs[HFX_MAX_LEN];
    __shared__ uint32_t maps[TN_WAVES][HFX_MAP_WORDS];
    __shared__ float mtab[32];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t cp[TN_MAXC];
    __shared__ int64_t cfx[TN_MAXC];
    __shared__ float cf32[TN_MAXC];
    __shared__ uint32_t wcnt[TN_WAVES];
    __shared__ float sh_x1[TN_WAVES];
    __shared__ uint32_t sh_digit, sh_rem, sh_gt, sh_last, sh_ext;
    __shared__ int64_t sh_exN;

    const uint32_t s = blockIdx.x, tid = threadIdx.x, wave = tid / HFX_WAVE, lane = hfx_lane();
    const uint32_t n = c.count(s);
    const float *sc = scores + c.base(s);
    if (n == 0) {
        if (tid == 0) {
            top_cnt[s] = 0;
            flag[s] = 0;
        }
        return;
    }
    const uint32_t T = min(N + margin, n);

    /* 1. radix select of the T largest keys (MSB first, 8-bit digits) */
    uint32_t prefix = 0, pmask = 0;
    if (tid == 0) sh_rem = T;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[tid] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < n; i += TN_THREADS) {
            const uint32_t k = hfx_f2key(sc[i]);
            if ((k & pmask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t rem = sh_rem, cum = 0, d = 0;
            for (int dd = 255; dd >= 0; --dd) {
                if (cum + hist[dd] >= rem) { d = (uint32_t)dd; break; }
                cum += hist[dd];
            }
            sh_digit = d;
            sh_rem = rem - cum;
        }
        __syncthreads();
        prefix |= sh_digit << shift;
        pmask |= 255u << shift;
    }
    const uint32_t K = prefix;
    const uint32_t need_eq = sh_rem; /* >= 1 */
    const uint32_t n_gt = T - need_eq;
    if (tid == 0) { sh_gt = 0; sh_ext = 0; sh_last = 0; }
    __syncthreads();
    for (uint32_t i = tid; i < n; i += TN_THREADS) {
        if (hfx_f2key(sc[i]) > K) {
            const uint32_t slot = atomicAdd(&sh_gt, 1u);
            if (slot < n_gt) cp[slot] = i;
        }
    }
    /* keys equal to the threshold: the need_eq smallest positions */
    uint32_t eq_base = 0;
    for (uint32_t base = 0; base < n && eq_base < need_eq; base += TN_THREADS) {
        const uint32_t i = base + tid;
        const bool f = i < n && hfx_f2key(sc[i]) == K;
        const uint64_t bal = __ballot(f);
        const uint32_t wr = (uint32_t)__builtin_popcountll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wcnt[wave] = (uint32_t)__builtin_popcountll(bal);
        __syncthreads();
        uint32_t wpre = 0, tot = 0;
        for (uint32_t w = 0; w < TN_WAVES; ++w) {
            if (w < wave) wpre += wcnt[w];
            tot += wcnt[w];
        }
        const uint32_t r = eq_base + wpre + wr;
        if (f && r < need_eq) {
            cp[n_gt + r] = i;
            if (r == need_eq - 1) sh_last = i;
        }
        eq_base += tot;
        __syncthreads();
    }

    /* 2. exact re-score of the selection */
    const uint32_t len = hfx_stage_spectrum(xs, mtab, c.db, sp, s);
    uint32_t *map = maps[wave];
    for (uint32_t w = lane; w < HFX_MAP_WORDS; w += HFX_WAVE) map[w] = 0;
    __syncthreads();
    const uint32_t zeff = c.zeff(sp.d_zmax[s]);
    for (uint32_t i = wave; i < T; i += TN_WAVES) {
        c.peaks(map, mtab, s, cp[i], len, zeff);
        hfx_wave_sync();
        const int64_t fx = hfx_wave_dot_fx(map, xs, len);
        hfx_wave_sync();
        if (lane == 0) {
            cfx[i] = fx;
            cf32[i] = sc[cp[i]];
        }
    }
    float x1 = 0.0f;
    for (uint32_t i = tid; i < len; i += TN_THREADS) x1 += fabsf(xs[i]);
    x1 = hfx_wave_sum_f32(x1);
    if (lane == 0) sh_x1[wave] = x1;
    __syncthreads();

    /* 3. every other candidate that could still reach the N-th exact score */
    uint32_t M = T;
    uint32_t fl = 0;
    if (n > T) {
        for (uint32_t i = tid; i < T; i += TN_THREADS) {
            uint32_t rk = 0;
            for (uint32_t j = 0; j < T; ++j) rk += before_fx(cfx[j], cp[j], cfx[i], cp[i]) ? 1u : 0u;
            if (rk == N - 1) sh_exN = cfx[i];
        }
        __syncthreads();
        float xa = 0.0f;
        for (uint32_t w = 0; w < TN_WAVES; ++w) xa += sh_x1[w];
        const double bound = 80.0 * 5.9604644775390625e-08 * 50.0 * (double)xa * 1.001 + 1e-6;
        const double tau = (double)sh_exN * (1.0 / HFX_FX_SCALE) - bound;
        const uint32_t last = sh_last;
        for (uint32_t i = tid; i < n; i += TN_THREADS) {
            const float f = sc[i];
            const uint32_t k = hfx_f2key(f);
            const bool selected = k > K || (k == K && i <= last);
            if (!selected && (double)f >= tau) {
                const uint32_t slot = atomicAdd(&sh_ext, 1u);
                if (slot < TN_EXT) cp[T + slot] = i;
            }
        }
        __syncthreads();
        const uint32_t ne = sh_ext;
        if (ne > TN_EXT) {
            fl = 1;
        } else {
            for (uint32_t i = T + wave; i < T + ne; i += TN_WAVES) {
                c.peaks(map, mtab, s, cp[i], len, zeff);
                hfx_wave_sync();
                const int64_t fx = hfx_wave_dot_fx(map, xs, len);
                hfx_wave_sync();
                if (lane == 0) {
                    cfx[i] = fx;
                    cf32[i] = sc[cp[i]];
                }
            }
            M = T + ne;
        }
        __syncthreads();
    }

    /* 4. order by (exact desc, position asc), emit N */
    const uint32_t outn = min(N, n);
    for (uint32_t i = tid; i < M; i += TN_THREADS) {
        const int64_t fi = cfx[i];
        const uint32_t pi = cp[i];
        uint32_t rk = 0;
        for (uint32_t j = 0; j < M; ++j) rk += before_fx(cfx[j], cp[j], fi, pi) ? 1u : 0u;
        if (rk < outn) {
            top_pos[(size_t)s * N + rk] = pi;
            top_exact[(size_t)s * N + rk] = (double)fi * (1.0 / HFX_FX_SCALE);
            top_f32[(size_t)s * N + rk] = cf32[i];
        }
    }
    if (tid == 0) {
        top_cnt[s] = outn;
        flag[s] = fl;
    }
}

/* ------------------------------------------------------------------ */
/* Exact top-N over every candidate (fallback for flagged spectra)       */
/* ------------------------------------------------------------------ */

#define TX_MAXN 64

template <class Cands>
__global__ void __launch_bounds__(TN_THREADS)
    k_topn_exact(uint32_t *top_pos, double *top_exact, float *top_f32, uint32_t *top_cnt,
                 uint32_t N, const uint32_t *list, const uint32_t *gate, Cands c, hfx_spectra sp)
{
    /* list == nullptr: block b is spectrum b; gate: only spectra flagged there */
    if (gate && !gate[list ? list[blockIdx.x] : blockIdx.x]) return;
    __shared__ float xs[HFX_MAX_LEN];
    __shared__ uint32_t maps[TN_WAVES][HFX_MAP_WORDS];
    __shared__ float mtab[32];
    __shared__ int64_t lfx[TN_WAVES * TX_MAXN];
    __shared__ uint32_t lpos[TN_WAVES * TX_MAXN];
    __shared__ float lf32[TN_WAVES * TX_MAXN];
    __shared__ uint32_t lcnt[TN_WAVES];
    const uint32_t s = list ? list[blockIdx.x] : blockIdx.x;
    const uint32_t tid = threadIdx.x, wave = tid / HFX_WAVE, lane = hfx_lane();
    const uint32_t n = c.count(s);
    const uint32_t len = hfx_stage_spectrum(xs, mtab, c.db, sp, s);
    uint32_t *map = maps[wave];
    for (uint32_t w = lane; w < HFX_MAP_WORDS; w += HFX_WAVE) map[w] = 0;
    if (lane == 0) lcnt[wave] = 0;
    __syncthreads();
    const uint32_t zeff = c.zeff(sp.d_zmax[s]);
    int64_t *wf = lfx + wave * TX_MAXN;
    uint32_t *wp = lpos + wave * TX_MAXN;
    float *w32 = lf32 + wave * TX_MAXN;
    for (uint32_t i = wave; i < n; i += TN_WAVES) {
        float f32;
        int64_t fx;
        c.peaks(map, mtab, s, i, len, zeff);
        hfx_wave_sync();
        hfx_wave_dot_both(map, xs, len, &f32, &fx);
        hfx_wave_sync();
        if (lane == 0) {
            /* positions arrive ascending per wave: equal scores keep order */
            uint32_t cnum = lcnt[wave];
            if (cnum < N || fx > wf[cnum - 1]) {
                uint32_t j = cnum < N ? cnum++ : cnum - 1;
                while (j > 0 && fx > wf[j - 1]) {
                    wf[j] = wf[j - 1];
                    wp[j] = wp[j - 1];
                    w32[j] = w32[j - 1];
                    --j;
                }
                wf[j] = fx;
                wp[j] = i;
                w32[j] = f32;
                lcnt[wave] = cnum;
            }
        }
        hfx_wave_sync();
    }
    __syncthreads();
    const uint32_t outn = min(N, n);
    for (uint32_t e = tid; e < TN_WAVES * TX_MAXN; e += TN_THREADS) {
        const uint32_t w = e / TX_MAXN, k = e % TX_MAXN;
        if (k >= lcnt[w]) continue;
        const int64_t fi = lfx[e];
        const uint32_t pi = lpos[e];
        uint32_t rk = 0;
        for (uint32_t w2 = 0; w2 < TN_WAVES; ++w2)
            for (uint32_t k2 = 0; k2 < lcnt[w2]; ++k2) {
                const uint32_t e2 = w2 * TX_MAXN + k2;
                rk += before_fx(lfx[e2], lpos[e2], fi, pi) ? 1u : 0u;
            }
        if (rk < outn) {
            top_pos[(size_t)s * N + rk] = pi;
            top_exact[(size_t)s * N + rk] = (double)fi * (1.0 / HFX_FX_SCALE);
            top_f32[(size_t)s * N + rk] = lf32[e];
        }
    }
    if (tid == 0) top_cnt[s] = outn;
}

/* ------------------------------------------------------------------ */
/* Entry points                                                          */
/* ------------------------------------------------------------------ */

template <class Cands>
static int launch_topn(uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                       uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N, uint32_t margin,
                       const float *d_scores, const Cands &c, const hfx_spectra *sp, void *stream)
{
    HFX_REQUIRE(d_top_pos && d_top_exact && d_top_f32 && d_top_cnt && d_flag && d_scores && sp,
                HFX_ERR_ARG, "hfx_topn: null argument");
    HFX_REQUIRE(N >= 1 && N + margin <= TN_MAXT, HFX_ERR_ARG,
                "hfx_topn: N + margin must be in 1..%d (N=%u margin=%u)", TN_MAXT, N, margin);
    if (sp->num_spec == 0) return HFX_OK;
    hipLaunchKernelGGL(k_topn<Cands>, dim3(sp->num_spec), dim3(TN_THREADS), 0, hfx_stream(stream),
                       d_top_pos, d_top_exact, d_top_f32, d_top_cnt, d_flag, N, margin, d_scores, c,
                       *sp);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

template <class Cands>
static int launch_topn_exact(uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                             uint32_t *d_top_cnt, uint32_t N, const uint32_t *d_list,
                             uint32_t num_list, const Cands &c, const hfx_spectra *sp,
                             void *stream)
{
    HFX_REQUIRE(d_top_pos && d_top_exact && d_top_f32 && d_top_cnt && d_list && sp, HFX_ERR_ARG,
                "hfx_topn_exact: null argument");
    HFX_REQUIRE(N >= 1 && N <= TX_MAXN, HFX_ERR_ARG, "hfx_topn_exact: N must be in 1..%d",
                TX_MAXN);
    if (num_list == 0) return HFX_OK;
    hipLaunchKernelGGL(k_topn_exact<Cands>, dim3(num_list), dim3(TN_THREADS), 0, hfx_stream(stream),
                       d_top_pos, d_top_exact, d_top_f32, d_top_cnt, N, d_list,
                       (const uint32_t *)nullptr, c, *sp);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

/* hfx_topn followed, in the same stream and without a host round trip, by
 * the exact fallback for the spectra it flags (used by the batched path). */
int hfx_topn_gated(uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32, uint32_t *d_top_cnt,
                   uint32_t *d_flag, uint32_t N, uint32_t margin, const float *d_scores,
                   const uint64_t *d_off, const uint32_t *d_begin, const uint32_t *d_num,
                   const hfx_db *db, const hfx_spectra *sp, void *stream)
{
    HfxPlainCands c{*db, d_begin, d_num, d_off};
    int rc = launch_topn(d_top_pos, d_top_exact, d_top_f32, d_top_cnt, d_flag, N, margin, d_scores,
                         c, sp, stream);
    if (rc != HFX_OK || sp->num_spec == 0) return rc;
    HFX_REQUIRE(N <= TX_MAXN, HFX_ERR_ARG, "N must be <= %d", TX_MAXN);
    hipLaunchKernelGGL(k_topn_exact<HfxPlainCands>, dim3(sp->num_spec), dim3(TN_THREADS), 0,
                       hfx_stream(stream), d_top_pos, d_top_exact, d_top_f32, d_top_cnt, N,
                       (const uint32_t *)nullptr, (const uint32_t *)d_flag, c, *sp);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

extern "C" int hfx_topn(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                        uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N, uint32_t margin,
                        const float *d_scores, const uint64_t *d_off, const uint32_t *d_begin,
                        const uint32_t *d_num, const hfx_db *db, const hfx_spectra *sp, void *stream)
{
    (void)ctx;
    HFX_REQUIRE(d_off && d_begin && d_num && db, HFX_ERR_ARG, "hfx_topn: null argument");
    HfxPlainCands c{*db, d_begin, d_num, d_off};
    return launch_topn(d_top_pos, d_top_exact, d_top_f32, d_top_cnt, d_flag, N, margin, d_scores, c,
                       sp, stream);
}

extern "C" int hfx_topn_exact(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact,
                              float *d_top_f32, uint32_t *d_top_cnt, uint32_t N,
                              const uint32_t *d_spec_list, uint32_t num_list, const uint32_t *d_begin,
                              const uint32_t *d_num, const hfx_db *db, const hfx_spectra *sp,
                              void *stream)
{
    (void)ctx;
    HFX_REQUIRE(d_begin && d_num && db, HFX_ERR_ARG, "hfx_topn_exact: null argument");
    HfxPlainCands c{*db, d_begin, d_num, nullptr};
    return launch_topn_exact(d_top_pos, d_top_exact, d_top_f32, d_top_cnt, N, d_spec_list, num_list,
                             c, sp, stream);
}

extern "C" int hfx_topn_mod(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact,
                            float *d_top_f32, uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N,
                            uint32_t margin, const float *d_scores, const uint64_t *d_mp_off,
                            const hfx_modset *ms, const hfx_db *db, const hfx_spectra *sp,
                            void *stream)
{
    (void)ctx;
    HFX_REQUIRE(d_mp_off && ms && db, HFX_ERR_ARG, "hfx_topn_mod: null argument");
    HFX_REQUIRE(ms->num_ma <= HFX_MAX_MA, HFX_ERR_ARG, "hfx_topn_mod: num_ma %u", ms->num_ma);
    HfxModCands c{*db, *ms, d_mp_off};
    return launch_topn(d_top_pos, d_top_exact, d_top_f32, d_top_cnt, d_flag, N, margin, d_scores, c,
                       sp, stream);
}

extern "C" int hfx_topn_exact_mod(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact,
                                  float *d_top_f32, uint32_t *d_top_cnt, uint32_t N,
                                  const uint32_t *d_spec_list, uint32_t num_list,
                                  const uint64_t *d_mp_off, const hfx_modset *ms, const hfx_db *db,
                                  const hfx_spectra *sp, void *stream)
{
    (void)ctx;
    HFX_REQUIRE(d_mp_off && ms && db, HFX_ERR_ARG, "hfx_topn_exact_mod: null argument");
    HFX_REQUIRE(ms->num_ma <= HFX_MAX_MA, HFX_ERR_ARG, "hfx_topn_exact_mod: num_ma %u", ms->num_ma);
    HfxModCands c{*db, *ms, d_mp_off};
    return launch_topn_exact(d_top_pos, d_top_exact, d_top_f32, d_top_cnt, N, d_spec_list, num_list,
                             c, sp, stream);
}
