/*
 * hfx_topn.hip -- per-spectrum top-N with an exact re-score.
 *
 * The reference sorts every score of a spectrum with a b40c radix sort and
 * copies them all to the host (Search.hs:250-276, sort_b40c.cu:34-41); the
 * final order is descending score, ties by ascending candidate position
 * (stable ascending sort, then reverse twice: Search.hs:119-123, :276).
 *
 * Here one workgroup per spectrum:
 *   1. radix-selects the T+1 best FP32 scores (key desc, position asc), T =
 *      min(N + margin, n);
 *   2. re-scores the best T exactly: int64 fixed point, sum of
 *      round(w * x * 2^36), independent of summation order, so the GPU and
 *      the CPU oracle agree bit for bit;
 *   3. orders them by (exact desc, position asc) and writes the best N;
 *   4. flags the spectrum when the (T+1)-th FP32 score plus the FP32 error
 *      bound could still reach the N-th exact score (then hfx_topn_exact
 *      scores every candidate exactly).
 * FP32 error bound of a score: |f32 - exact| <= 80 * 2^-24 * sum|w*x| and
 * sum|w*x| <= 50 * sum_b |x_b| (each bin counts once, weights <= 50).
 */
#include "hfx_common.h"

#define TN_THREADS 256
#define TN_WAVES (TN_THREADS / HFX_WAVE)
#define TN_MAXT 128

struct TopnArgs {
    uint32_t *top_pos;
    double *top_exact;
    float *top_f32;
    uint32_t *top_cnt;
    uint32_t *flag;
    uint32_t N, margin;
    const float *scores;
    const uint64_t *off;
    const uint32_t *begin;
    const uint32_t *num;
    hfx_db db;
    hfx_spectra sp;
};

__device__ __forceinline__ bool before_key(uint32_t ka, uint32_t pa, uint32_t kb, uint32_t pb)
{
    return ka > kb || (ka == kb && pa < pb);
}
__device__ __forceinline__ bool before_fx(int64_t fa, uint32_t pa, int64_t fb, uint32_t pb)
{
    return fa > fb || (fa == fb && pa < pb);
}

/* Stage the spectrum and mass table; return the effective length. */
__device__ __forceinline__ uint32_t stage_spectrum(float *xs, float *mtab, const hfx_db &db,
                                                   const hfx_spectra &sp, uint32_t s)
{
    const uint32_t len = min(sp.d_len[s], min(sp.x_stride, (uint32_t)HFX_MAX_LEN));
    const float *xg = sp.d_x + (size_t)s * sp.x_stride;
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) xs[i] = xg[i];
    if (threadIdx.x < 32) mtab[threadIdx.x] = threadIdx.x < 26 ? db.d_mass[threadIdx.x] : 0.0f;
    return len;
}

/* Exact and FP32 score of window position `pos` of spectrum s by one wave. */
__device__ __forceinline__ void wave_score(uint32_t *map, const float *xs, const float *mtab,
                                           const hfx_db &db, uint32_t p, uint32_t len,
                                           uint32_t zmax, float *f32, int64_t *fx)
{
    const uint32_t pep = db.d_idx_sorted[p];
    const uint32_t tc = db.d_tc[pep], tn = db.d_tn[pep];
    const uint32_t n = tn > tc ? tn - tc : 0;
    const uint8_t *ions = db.d_ions + tc;
    const uint32_t zeff = zmax <= HFX_MAX_CHARGE ? zmax : 0;
    hfx_wave_peaks(map, n, db.d_r[pep], zeff, len,
                   [&](uint32_t j) { return mtab[HFX_ACID_IDX(ions[j])]; });
    hfx_wave_sync();
    hfx_wave_dot_both(map, xs, len, f32, fx);
    hfx_wave_sync();
}

__global__ void __launch_bounds__(TN_THREADS) k_topn(TopnArgs a)
{
    __shared__ float xs[HFX_MAX_LEN];
    __shared__ uint32_t maps[TN_WAVES][HFX_MAP_WORDS];
    __shared__ float mtab[32];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sel_k[TN_MAXT + 1], sel_p[TN_MAXT + 1];
    __shared__ uint32_t srt_k[TN_MAXT + 1], srt_p[TN_MAXT + 1];
    __shared__ int64_t ex_fx[TN_MAXT];
    __shared__ float ex_f32[TN_MAXT];
    __shared__ uint32_t wcnt[TN_WAVES];
    __shared__ uint32_t sh_digit, sh_rem, sh_gt;
    __shared__ float sh_x1[TN_WAVES];

    const uint32_t s = blockIdx.x, tid = threadIdx.x, wave = tid / HFX_WAVE, lane = hfx_lane();
    const uint32_t n = a.num[s];
    const uint32_t N = a.N;
    const float *sc = a.scores + a.off[s];
    if (n == 0) {
        if (tid == 0) {
            a.top_cnt[s] = 0;
            a.flag[s] = 0;
        }
        return;
    }
    const uint32_t T = min(N + a.margin, n);
    const uint32_t want = min(T + 1, n);

    /* 1. radix select of the `want` largest keys (MSB first, 8-bit digits) */
    uint32_t prefix = 0, pmask = 0;
    if (tid == 0) sh_rem = want;
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
    const uint32_t need_eq = sh_rem;
    const uint32_t n_gt = want - need_eq;
    if (tid == 0) sh_gt = 0;
    __syncthreads();
    for (uint32_t i = tid; i < n; i += TN_THREADS) {
        const uint32_t k = hfx_f2key(sc[i]);
        if (k > K) {
            const uint32_t slot = atomicAdd(&sh_gt, 1u);
            if (slot < n_gt) { sel_k[slot] = k; sel_p[slot] = i; }
        }
    }
    /* keys equal to the threshold: the need_eq smallest positions, in order */
    uint32_t eq_base = 0;
    for (uint32_t base = 0; base < n && eq_base < need_eq; base += TN_THREADS) {
        const uint32_t i = base + tid;
        const bool f = i < n && hfx_f2key(sc[i]) == K;
        const uint64_t bal = __ballot(f);
        const uint32_t wr = __builtin_popcountll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wcnt[wave] = (uint32_t)__builtin_popcountll(bal);
        __syncthreads();
        uint32_t wpre = 0, tot = 0;
        for (uint32_t w = 0; w < TN_WAVES; ++w) {
            if (w < wave) wpre += wcnt[w];
            tot += wcnt[w];
        }
        const uint32_t r = eq_base + wpre + wr;
        if (f && r < need_eq) { sel_k[n_gt + r] = K; sel_p[n_gt + r] = i; }
        eq_base += tot;
        __syncthreads();
    }
    __syncthreads();
    /* order the selection by (key desc, position asc) */
    for (uint32_t i = tid; i < want; i += TN_THREADS) {
        const uint32_t ki = sel_k[i], pi = sel_p[i];
        uint32_t rk = 0;
        for (uint32_t j = 0; j < want; ++j) rk += before_key(sel_k[j], sel_p[j], ki, pi) ? 1u : 0u;
        srt_k[rk] = ki;
        srt_p[rk] = pi;
    }
    /* 2. exact re-score of the best T */
    const uint32_t len = stage_spectrum(xs, mtab, a.db, a.sp, s);
    uint32_t *map = maps[wave];
    for (uint32_t w = lane; w < HFX_MAP_WORDS; w += HFX_WAVE) map[w] = 0;
    __syncthreads();
    const uint32_t zmax = a.sp.d_zmax[s];
    const uint32_t b0 = a.begin[s];
    for (uint32_t i = wave; i < T; i += TN_WAVES) {
        float f32;
        int64_t fx;
        wave_score(map, xs, mtab, a.db, b0 + srt_p[i], len, zmax, &f32, &fx);
        if (lane == 0) {
            ex_fx[i] = fx;
            ex_f32[i] = sc[srt_p[i]];
        }
    }
    /* sum |x| for the FP32 error bound */
    float x1 = 0.0f;
    for (uint32_t i = tid; i < len; i += TN_THREADS) x1 += fabsf(xs[i]);
    x1 = hfx_wave_sum_f32(x1);
    if (lane == 0) sh_x1[wave] = x1;
    __syncthreads();
    /* 3. order the re-scored set by (exact desc, position asc), emit N */
    const uint32_t outn = min(N, T);
    for (uint32_t i = tid; i < T; i += TN_THREADS) {
        const int64_t fi = ex_fx[i];
        const uint32_t pi = srt_p[i];
        uint32_t rk = 0;
        for (uint32_t j = 0; j < T; ++j) rk += before_fx(ex_fx[j], srt_p[j], fi, pi) ? 1u : 0u;
        if (rk < outn) {
            a.top_pos[(size_t)s * N + rk] = pi;
            a.top_exact[(size_t)s * N + rk] = (double)fi * (1.0 / HFX_FX_SCALE);
            a.top_f32[(size_t)s * N + rk] = ex_f32[i];
            if (rk == outn - 1) hist[0] = i; /* index of the N-th, for the flag */
        }
    }
    __syncthreads();
    if (tid == 0) {
        a.top_cnt[s] = outn;
        uint32_t fl = 0;
        if (n > T) {
            float xa = 0.0f;
            for (uint32_t w = 0; w < TN_WAVES; ++w) xa += sh_x1[w];
            const double bound = 80.0 * 5.9604644775390625e-08 * 50.0 * (double)xa * 1.001 + 1e-6;
            const double exN = (double)ex_fx[hist[0]] * (1.0 / HFX_FX_SCALE);
            const double f32ex = (double)hfx_key2f(srt_k[T]);
            fl = (exN <= f32ex + bound) ? 1u : 0u;
        }
        a.flag[s] = fl;
    }
}

extern "C" int hfx_topn(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                        uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N, uint32_t margin,
                        const float *d_scores, const uint64_t *d_off, const uint32_t *d_begin,
                        const uint32_t *d_num, const hfx_db *db, const hfx_spectra *sp, void *stream)
{
    (void)ctx;
    HFX_REQUIRE(d_top_pos && d_top_exact && d_top_f32 && d_top_cnt && d_flag && d_scores && d_off &&
                    d_begin && d_num && db && sp,
                HFX_ERR_ARG, "hfx_topn: null argument");
    HFX_REQUIRE(N >= 1 && N + margin <= TN_MAXT, HFX_ERR_ARG,
                "hfx_topn: N + margin must be in 1..%d (N=%u margin=%u)", TN_MAXT, N, margin);
    if (sp->num_spec == 0) return HFX_OK;
    TopnArgs a{d_top_pos, d_top_exact, d_top_f32, d_top_cnt, d_flag, N, margin,
               d_scores,  d_off,       d_begin,   d_num,     *db,    *sp};
    hipLaunchKernelGGL(k_topn, dim3(sp->num_spec), dim3(TN_THREADS), 0, hfx_stream(stream), a);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

/* ------------------------------------------------------------------ */
/* Exact top-N over every candidate (fallback for flagged spectra)       */
/* ------------------------------------------------------------------ */

#define TX_MAXN 64

struct TopnExactArgs {
    uint32_t *top_pos;
    double *top_exact;
    float *top_f32;
    uint32_t *top_cnt;
    uint32_t N;
    const uint32_t *list;
    const uint32_t *begin;
    const uint32_t *num;
    hfx_db db;
    hfx_spectra sp;
};

__global__ void __launch_bounds__(TN_THREADS) k_topn_exact(TopnExactArgs a)
{
    __shared__ float xs[HFX_MAX_LEN];
    __shared__ uint32_t maps[TN_WAVES][HFX_MAP_WORDS];
    __shared__ float mtab[32];
    __shared__ int64_t lfx[TN_WAVES * TX_MAXN];
    __shared__ uint32_t lpos[TN_WAVES * TX_MAXN];
    __shared__ float lf32[TN_WAVES * TX_MAXN];
    __shared__ uint32_t lcnt[TN_WAVES];
    const uint32_t s = a.list[blockIdx.x];
    const uint32_t tid = threadIdx.x, wave = tid / HFX_WAVE, lane = hfx_lane();
    const uint32_t n = a.num[s], N = a.N;
    const uint32_t len = stage_spectrum(xs, mtab, a.db, a.sp, s);
    uint32_t *map = maps[wave];
    for (uint32_t w = lane; w < HFX_MAP_WORDS; w += HFX_WAVE) map[w] = 0;
    if (lane == 0) lcnt[wave] = 0;
    __syncthreads();
    const uint32_t zmax = a.sp.d_zmax[s], b0 = a.begin[s];
    int64_t *wf = lfx + wave * TX_MAXN;
    uint32_t *wp = lpos + wave * TX_MAXN;
    float *w32 = lf32 + wave * TX_MAXN;
    for (uint32_t i = wave; i < n; i += TN_WAVES) {
        float f32;
        int64_t fx;
        wave_score(map, xs, mtab, a.db, b0 + i, len, zmax, &f32, &fx);
        if (lane == 0) {
            /* positions arrive ascending per wave: equal scores keep order */
            uint32_t c = lcnt[wave];
            if (c < N || fx > wf[c - 1]) {
                uint32_t j = c < N ? c++ : c - 1;
                while (j > 0 && fx > wf[j - 1]) {
                    wf[j] = wf[j - 1];
                    wp[j] = wp[j - 1];
                    w32[j] = w32[j - 1];
                    --j;
                }
                wf[j] = fx;
                wp[j] = i;
                w32[j] = f32;
                lcnt[wave] = c;
            }
        }
        hfx_wave_sync();
    }
    __syncthreads();
    uint32_t tot = 0;
    for (uint32_t w = 0; w < TN_WAVES; ++w) tot += lcnt[w];
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
            a.top_pos[(size_t)s * N + rk] = pi;
            a.top_exact[(size_t)s * N + rk] = (double)fi * (1.0 / HFX_FX_SCALE);
            a.top_f32[(size_t)s * N + rk] = lf32[e];
        }
    }
    (void)tot;
    if (tid == 0) a.top_cnt[s] = outn;
}

extern "C" int hfx_topn_exact(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact,
                              float *d_top_f32, uint32_t *d_top_cnt, uint32_t N,
                              const uint32_t *d_spec_list, uint32_t num_list, const uint32_t *d_begin,
                              const uint32_t *d_num, const hfx_db *db, const hfx_spectra *sp,
                              void *stream)
{
    (void)ctx;
    HFX_REQUIRE(d_top_pos && d_top_exact && d_top_f32 && d_top_cnt && d_spec_list && d_begin &&
                    d_num && db && sp,
                HFX_ERR_ARG, "hfx_topn_exact: null argument");
    HFX_REQUIRE(N >= 1 && N <= TX_MAXN, HFX_ERR_ARG, "hfx_topn_exact: N must be in 1..%d", TX_MAXN);
    if (num_list == 0) return HFX_OK;
    TopnExactArgs a{d_top_pos, d_top_exact, d_top_f32, d_top_cnt, N, d_spec_list, d_begin, d_num,
                    *db,       *sp};
    hipLaunchKernelGGL(k_topn_exact, dim3(num_list), dim3(TN_THREADS), 0, hfx_stream(stream), a);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}
