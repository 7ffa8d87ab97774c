/*
 * hfx_topn.hip -- per-spectrum top-N with an exact re-score.
 *
 * The reference sorts every score of a spectrum with a b40c radix sort and
 * copies them all to the host (Search.hs:250-276, sort_b40c.cu:34-41); the
 * final order is descending score, ties by ascending candidate position
 * (stable ascending sort, then reverse twice: Search.hs:119-123, :276).
 *
 * Here one workgroup per spectrum:
 *   1. selects the T = min(N + margin, n) best FP32 scores (key desc,
 *      position asc): for T <= TN_LT in one sweep, each thread keeping its
 *      TN_LT best in registers and the waves merging them; otherwise by a
 *      radix select;
 *   2. takes the N-th best FP32 score fN and collects every other candidate
 *      whose FP32 score is >= fN - 2B (B bounds |FP32 - exact|): the N-th
 *      best exact score exN is >= fN - B, so no candidate left out can reach
 *      the top N;
 *   3. re-scores the selection and that band exactly, in one round: int64
 *      fixed point, sum of round(w*x*2^36), independent of summation order,
 *      so the GPU and the CPU oracle agree bit for bit;
 *   4. orders the re-scored set by (exact desc, position asc), writes the best
 *      N.  If the band holds more than TN_EXT candidates, the spectrum is
 *      flagged and the exact route scores every candidate at or above
 *      exN - B exactly.
 * Bound: |score - exact| <= (rel0 + rel1 * z) * 50 * sum_b |x_b|
 *                           + relq * max_b |x_b| * 50 * len + 1e-6, with
 * z the spectrum's fragment charge count and sum|w*x| <= 50 * sum_b |x_b|
 * (each bin counts once, weights <= 50).  The sparse FP32 scores use
 * rel0 = 80 * 2^-24 (<= 71 roundings per score), rel1 = 0; the dense BF16
 * scores pass their own (split error + one rounding per non-zero bin).
 */
#define HFX_TOPN_REL_F32 (80.0f * 5.9604644775390625e-08f)
#include "hfx_cand.h"

#define TN_THREADS 256
#define TN_WAVES (TN_THREADS / HFX_WAVE)
#define TN_MAXT 128
#define TN_EXT 256
#define TN_MAXC (TN_MAXT + TN_EXT)

/* T at most this: the selection keeps per-thread lists in registers */
#define TN_LT 6

/* the first select pass histograms the key's 11 high bits */
#define TN_HBITS 11
#define TN_HBINS (1u << TN_HBITS)

/* The staged spectrum and the waves' maps live in dynamic LDS sized by the
 * spectra's row stride: bins of a row, rounded up to whole map words.  The
 * spectrum's region is at least TN_HBINS words: before the spectrum is
 * staged it holds the select's histogram. */
__host__ __device__ inline uint32_t tn_xcap(uint32_t x_stride)
{
    return ((x_stride < HFX_MAX_LEN ? x_stride : HFX_MAX_LEN) + 7u) & ~7u;
}
__host__ __device__ inline size_t tn_xs_bytes(uint32_t x_stride)
{
    return (size_t)(tn_xcap(x_stride) > TN_HBINS ? tn_xcap(x_stride) : TN_HBINS) * 4;
}
__host__ __device__ inline size_t tn_dyn_bytes(uint32_t x_stride)
{
    return tn_xs_bytes(x_stride) + (size_t)TN_WAVES * (tn_xcap(x_stride) / 8) * 4;
}

__device__ __forceinline__ bool before_fx(int64_t fa, uint32_t pa, int64_t fb, uint32_t pb)
{
    return fa > fb || (fa == fb && pa < pb);
}

/* One radix-select step on an nb-bucket histogram: the digit holding the
 * rem-th largest element (counting from bucket nb - 1 down) and the rank
 * left inside it.  Wave 0 does it: lane l sums its run of nb / 64 buckets
 * (one at least), an inclusive scan over the lanes finds the run that holds
 * rank rem, and that lane walks it.  Every thread must call it (it ends with
 * a barrier); all get the same (digit, rem). */
__device__ __forceinline__ void tn_select_digit(const uint32_t *hist, uint32_t nb,
                                                uint32_t *sh_digit, uint32_t *sh_rem)
{
    const uint32_t tid = threadIdx.x, lane = hfx_lane();
    if (tid < HFX_WAVE) {
        const uint32_t rem = *sh_rem;
        const uint32_t per = (nb + HFX_WAVE - 1) / HFX_WAVE;
        const int top = (int)nb - 1 - (int)(lane * per);
        uint32_t sum = 0;
        for (uint32_t k = 0; k < per; ++k)
            if (top - (int)k >= 0) sum += hist[top - (int)k];
        uint32_t incl = sum;
#pragma unroll
        for (uint32_t o = 1; o < HFX_WAVE; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        const uint32_t excl = incl - sum;
        if (excl < rem && rem <= incl) {
            uint32_t cum = excl;
            for (uint32_t k = 0; k < per && top - (int)k >= 0; ++k) {
                const uint32_t h = hist[top - (int)k];
                if (cum + h >= rem) {
                    *sh_digit = (uint32_t)(top - (int)k);
                    *sh_rem = rem - cum;
                    break;
                }
                cum += h;
            }
        }
    }
    __syncthreads();
}

/* the largest of the wave's 64-bit values, in every lane */
__device__ __forceinline__ uint64_t tn_wave_max_u64(uint64_t v)
{
#pragma unroll
    for (uint32_t o = 1; o < HFX_WAVE; o <<= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, o, 64);
        const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), o, 64);
        const uint64_t y = (uint64_t)hi << 32 | lo;
        v = y > v ? y : v;
    }
    return v;
}

/* Histogram increment of digit d for the active lanes of a wave: lanes that
 * share the first pending lane's digit add their count with one atomic (a
 * few rounds: score keys crowd into few high-byte buckets), the rest add one
 * each. */
__device__ __forceinline__ void tn_hist_add(uint32_t *hist, uint32_t d, bool act)
{
    const uint32_t lane = hfx_lane();
    uint64_t m = __ballot(act);
    for (int it = 0; it < 4 && m; ++it) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(m);
        const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)d, (int)leader);
        const uint64_t same = __ballot(((m >> lane) & 1ull) && d == v);
        if (lane == leader) atomicAdd(&hist[v], (uint32_t)__builtin_popcountll(same));
        m &= ~same;
    }
    if ((m >> lane) & 1ull) atomicAdd(&hist[d], 1u);
}

/* Launch order of k_topn: spectra by window size, largest first, in classes
 * of 256 candidates (any order inside a class).  A workgroup's time grows
 * with its window; started first, the long ones no longer finish last. */
#define TN_OCLS 64
#define TN_OTHREADS 1024
template <class Cands>
__global__ void __launch_bounds__(TN_OTHREADS) k_topn_order(uint32_t *order, Cands c, uint32_t S)
{
    __shared__ uint32_t cnt[TN_OCLS];
    const uint32_t tid = threadIdx.x;
    auto cls = [&](uint32_t s) -> uint32_t {
        return TN_OCLS - 1u - min(c.count(s) >> 8, TN_OCLS - 1u);
    };
    if (tid < TN_OCLS) cnt[tid] = 0;
    static_assert(TN_OCLS <= TN_OTHREADS, "one thread clears each class counter");
    __syncthreads();
    for (uint32_t s = tid; s < S; s += TN_OTHREADS) atomicAdd(&cnt[cls(s)], 1u);
    __syncthreads();
    if (tid == 0)
        for (uint32_t k = 0, a = 0; k < TN_OCLS; ++k) {
            const uint32_t v = cnt[k];
            cnt[k] = a;
            a += v;
        }
    __syncthreads();
    for (uint32_t s = tid; s < S; s += TN_OTHREADS) order[atomicAdd(&cnt[cls(s)], 1u)] = s;
}

template <class Cands>
__global__ void __launch_bounds__(TN_THREADS)
    k_topn(uint32_t *top_pos, double *top_exact, float *top_f32, uint32_t *top_cnt,
           uint32_t *flag, uint32_t N, uint32_t margin, const float *scores, Cands c,
           hfx_spectra sp, float rel0, float rel1, float relq, double *tau_out,
           uint32_t *band_out,
           const uint32_t *order)
{
    extern __shared__ __align__(16) unsigned char tn_dyn[];
    const uint32_t xcap = tn_xcap(sp.x_stride), mwords = xcap / 8;
    float *xs = (float *)tn_dyn;
    uint32_t *hist = (uint32_t *)tn_dyn; /* until the spectrum is staged */
    uint32_t *maps = (uint32_t *)(tn_dyn + tn_xs_bytes(sp.x_stride));
    __shared__ float mtab[32];
    __shared__ uint32_t cp[TN_MAXC];
    __shared__ int64_t cfx[TN_MAXC];
    __shared__ float cf32[TN_MAXC];
    __shared__ uint32_t wcnt[TN_WAVES];
    __shared__ float sh_x1[TN_WAVES], sh_xm[TN_WAVES];
    __shared__ uint32_t sh_digit, sh_rem, sh_gt, sh_last, sh_ext, sh_wide;
    __shared__ uint64_t sh_wtop[TN_WAVES * TN_LT];
    __shared__ float sh_fN;

    const uint32_t s = order[blockIdx.x], tid = threadIdx.x, wave = tid / HFX_WAVE;
    const uint32_t lane = hfx_lane();
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

    /* every sweep below visits this thread's keys through for_keys(f), f(i,
     * key), eight loads in flight at a time (clamped into the window, so
     * that none is conditional); i >= n past the window, where all lanes of
     * a wave still take part (for the ballots) */
    auto for_keys = [&](auto &&f) {
        constexpr uint32_t U = 8;
        for (uint32_t base = 0; base < n; base += U * TN_THREADS) {
            uint32_t kk[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                kk[u] = hfx_f2key(sc[min(base + u * TN_THREADS + tid, n - 1u)]);
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) f(base + u * TN_THREADS + tid, kk[u]);
        }
    };

    /* 1. the T largest keys (key desc, position asc).  Radix select, first
     * over the key's 11 high bits (sign, exponent, two mantissa bits: a
     * quarter octave a bucket), then 8, 8 and 5 bits; after each pass, once
     * the keys at or above the prefix found so far -- M of them, known from
     * the histograms -- number at most TN_MAXC, they are collected and ranked
     * exactly.  Otherwise (long runs of near-equal scores) the full threshold
     * key K is found and the keys above it, then the smallest positions
     * equal to it, are collected. */
    uint32_t prefix = 0, pmask = 0;
    if (tid == 0) {
        sh_rem = T;
        sh_gt = 0;
        sh_ext = 0;
        sh_last = 0;
        sh_wide = 0;
    }
    /* T <= TN_LT: one sweep instead.  Each thread keeps its TN_LT best
     * (key, ~position) values in registers, sorted; the waves merge them.
     * Anything a thread left out ranks below its TN_LT-th value, which the
     * band step uses to take the band from the lists. */
    const bool local = T <= TN_LT;
    uint64_t lst[TN_LT];
#pragma unroll
    for (uint32_t j = 0; j < TN_LT; ++j) lst[j] = 0; /* 0: empty (no score has key 0 with position ~0) */
    bool ranked = false;
    if (local) {
        for_keys([&](uint32_t i, uint32_t k) {
            uint64_t x = i < n ? (uint64_t)k << 32 | (0xffffffffu - i) : 0ull;
            if (x > lst[TN_LT - 1]) {
#pragma unroll
                for (uint32_t j = 0; j < TN_LT; ++j) {
                    const uint64_t a = lst[j];
                    lst[j] = x > a ? x : a;
                    x = x > a ? a : x;
                }
            }
        });
        /* the wave's T best: T rounds of a wave max over the lists' heads */
        uint32_t h = 0;
        for (uint32_t r = 0; r < T; ++r) {
            uint64_t hv = 0;
#pragma unroll
            for (uint32_t j = 0; j < TN_LT; ++j) hv = h == j ? lst[j] : hv;
            const uint64_t m = tn_wave_max_u64(hv);
            if (m != 0 && hv == m) ++h;
            if (lane == 0) sh_wtop[wave * TN_LT + r] = m;
        }
        __syncthreads();
        /* the T best of the waves' 4T, ranked exactly */
        if (tid < TN_WAVES * T) {
            const uint64_t v = sh_wtop[(tid / T) * TN_LT + tid % T];
            uint32_t rk = 0;
            for (uint32_t b = 0; b < TN_WAVES * T; ++b)
                rk += sh_wtop[(b / T) * TN_LT + b % T] > v ? 1u : 0u;
            if (v != 0 && rk < T) {
                const uint32_t pa = 0xffffffffu - (uint32_t)v;
                cp[rk] = pa;
                if (rk == N - 1u) sh_fN = hfx_key2f((uint32_t)(v >> 32));
                if (rk == T - 1u) {
                    sh_digit = (uint32_t)(v >> 32);
                    sh_last = pa;
                }
            }
        }
        __syncthreads();
        ranked = true;
    }
    for (int pass = 0; pass < 4 && !local; ++pass) {
        const uint32_t shift = pass == 0 ? 32u - TN_HBITS : pass == 1 ? 13u : pass == 2 ? 5u : 0u;
        const uint32_t nb = pass == 0 ? TN_HBINS : pass == 3 ? 32u : 256u;
        for (uint32_t b = tid; b < nb; b += TN_THREADS) hist[b] = 0;
        __syncthreads();
        for_keys([&](uint32_t i, uint32_t k) {
            tn_hist_add(hist, (k >> shift) & (nb - 1u), i < n && (k & pmask) == prefix);
        });
        __syncthreads();
        tn_select_digit(hist, nb, &sh_digit, &sh_rem);
        prefix |= sh_digit << shift;
        pmask |= (nb - 1u) << shift;
        const uint32_t M = T - sh_rem + hist[sh_digit];
        __syncthreads(); /* every thread has read hist before it is cleared */
        if (M <= TN_MAXC) {
            /* (key, ~position) packed so that a larger value ranks first;
             * cfx holds them until the exact scores replace them */
            uint64_t *kv = (uint64_t *)cfx;
            for_keys([&](uint32_t i, uint32_t k) {
                if (i < n && (k & pmask) >= prefix)
                    kv[atomicAdd(&sh_gt, 1u)] = (uint64_t)k << 32 | (0xffffffffu - i);
            });
            __syncthreads();
            for (uint32_t a = tid; a < M; a += TN_THREADS) {
                const uint64_t va = kv[a];
                uint32_t rk = 0;
                for (uint32_t b = 0; b < M; ++b) rk += kv[b] > va ? 1u : 0u;
                const uint32_t pa = 0xffffffffu - (uint32_t)va;
                if (rk < T) cp[rk] = pa;
                if (rk == T - 1u) {
                    sh_digit = (uint32_t)(va >> 32);
                    sh_last = pa;
                }
            }
            __syncthreads();
            ranked = true;
            break;
        }
    }
    const uint32_t K = ranked ? sh_digit : prefix;
    if (!ranked) {
        const uint32_t need_eq = sh_rem; /* >= 1 */
        const uint32_t n_gt = T - need_eq;
        for_keys([&](uint32_t i, uint32_t k) {
            if (i < n && k > K) {
                const uint32_t slot = atomicAdd(&sh_gt, 1u);
                if (slot < n_gt) cp[slot] = i;
            }
        });
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
    }

    /* 2. the spectrum, the bound B on |FP32 - exact| (from sum |x|), and the
     * N-th best FP32 score fN of the selection */
    const uint32_t len = hfx_stage_spectrum(xs, mtab, c.db, sp, s);
    uint32_t *map = maps + wave * mwords;
    for (uint32_t w = lane; w < mwords; w += HFX_WAVE) map[w] = 0;
    for (uint32_t i = tid; i < T; i += TN_THREADS) cf32[i] = sc[cp[i]];
    __syncthreads();
    float x1 = 0.0f, xm = 0.0f;
    for (uint32_t i = tid; i < len; i += TN_THREADS) {
        x1 += fabsf(xs[i]);
        xm = fmaxf(xm, fabsf(xs[i]));
    }
    x1 = hfx_wave_sum_f32(x1);
    for (uint32_t o = 32; o > 0; o >>= 1) xm = fmaxf(xm, __shfl_xor(xm, o, 64));
    if (lane == 0) {
        sh_x1[wave] = x1;
        sh_xm[wave] = xm;
    }
    if (n > T && !local)
        for (uint32_t i = tid; i < T; i += TN_THREADS) {
            const uint32_t ki = hfx_f2key(cf32[i]), pi = cp[i];
            uint32_t rk = 0;
            for (uint32_t j = 0; j < T; ++j) {
                const uint32_t kj = hfx_f2key(cf32[j]);
                rk += (kj > ki || (kj == ki && cp[j] < pi)) ? 1u : 0u;
            }
            if (rk == N - 1) sh_fN = cf32[i];
        }
    __syncthreads();
    const uint32_t zeff = c.zeff(sp.d_zmax[s]);
    float xa = 0.0f, xmax = 0.0f;
    for (uint32_t w = 0; w < TN_WAVES; ++w) {
        xa += sh_x1[w];
        xmax = fmaxf(xmax, sh_xm[w]);
    }
    const double rel = (double)rel0 + (double)rel1 * (double)zeff;
    const double bound = rel * 50.0 * (double)xa * 1.001 +
                         (double)relq * (double)xmax * 50.0 * (double)len * 1.001 + 1e-6;

    /* 3. the extension: every other candidate that could reach the N-th
     * exact score exN.  The N best FP32 scores are selected and each has an
     * exact score >= fN - B, so exN >= fN - B: a candidate with FP32 score
     * below fN - 2B has an exact score below exN - B + B = exN, under N
     * others.  (A relative 1e-12 margin covers the double roundings.)  All
     * candidates at or above fN - 2B are re-scored with the selection, in one
     * round.  More than TN_EXT of them: the spectrum is flagged, the
     * selection alone is re-scored, and the exact route takes every
     * candidate at or above exN - B. */
    uint32_t ne = 0, fl = 0;
    if (n > T) {
        const double fN = (double)sh_fN;
        const double tau2 = fN - 2.0 * bound - 1e-12 * (fabs(fN) + 2.0 * bound);
        const uint32_t last = sh_last;
        /* from the lists: a thread whose TN_LT-th value is below tau2 left
         * out only values below tau2; if every thread's is, the band is the
         * lists' unselected values at or above tau2 */
        bool wide = false;
        if (local) {
            const uint64_t vT = (uint64_t)K << 32 | (0xffffffffu - last);
            wide = lst[TN_LT - 1] != 0 &&
                   (double)hfx_key2f((uint32_t)(lst[TN_LT - 1] >> 32)) >= tau2;
            if (wide) sh_wide = 1u;
            __syncthreads();
            wide = sh_wide != 0;
            if (!wide)
#pragma unroll
                for (uint32_t j = 0; j < TN_LT; ++j) {
                    const uint64_t v = lst[j];
                    if (v != 0 && v < vT && (double)hfx_key2f((uint32_t)(v >> 32)) >= tau2) {
                        const uint32_t slot = atomicAdd(&sh_ext, 1u);
                        if (slot < TN_EXT) cp[T + slot] = 0xffffffffu - (uint32_t)v;
                    }
                }
        }
        if (!local || wide)
            for_keys([&](uint32_t i, uint32_t k) {
                const bool selected = k > K || (k == K && i <= last);
                if (i < n && !selected && (double)hfx_key2f(k) >= tau2) {
                    const uint32_t slot = atomicAdd(&sh_ext, 1u);
                    if (slot < TN_EXT) cp[T + slot] = i;
                }
            });
        __syncthreads();
        ne = sh_ext;
        if (ne > TN_EXT) {
            fl = 1;
            if (tid == 0) band_out[s] = ne; /* sizes the exact route's tiles */
            ne = 0;
        }
    }

    /* 4. exact re-score of the selection and the extension: the wave's
     * candidates i = wave + 4 j, 64 at a time, lane j fetching candidate i's
     * metadata */
    const uint32_t M = T + ne;
    for (uint32_t c0 = 0; c0 < M; c0 += TN_WAVES * HFX_WAVE) {
        typename Cands::Prep mine{};
        if (c0 + wave + TN_WAVES * lane < M) mine = c.prep(s, cp[c0 + wave + TN_WAVES * lane]);
        /* the next candidate's first residues are loaded while this one is
         * scored */
        uint32_t res = c0 + wave < M ? c.first_res(Cands::take(mine, 0)) : 0u;
        for (uint32_t i = c0 + wave, j = 0; i < M && j < HFX_WAVE; i += TN_WAVES, ++j) {
            const uint32_t cur_res = res;
            if (i + TN_WAVES < M && j + 1 < HFX_WAVE) res = c.first_res(Cands::take(mine, j + 1));
            c.peaks_first(map, mtab, Cands::take(mine, j), cur_res, len, zeff);
            hfx_wave_sync();
            const int64_t fx = hfx_wave_dot_fx(map, xs, len);
            hfx_wave_sync();
            if (lane == 0) {
                cfx[i] = fx;
                if (i >= T) cf32[i] = sc[cp[i]];
            }
        }
    }
    __syncthreads();
    if (fl) {
        /* exN, the N-th best exact score of the selection, for the exact
         * route's threshold */
        for (uint32_t i = tid; i < T; i += TN_THREADS) {
            uint32_t rk = 0;
            for (uint32_t j = 0; j < T; ++j)
                rk += before_fx(cfx[j], cp[j], cfx[i], cp[i]) ? 1u : 0u;
            if (rk == N - 1) tau_out[s] = (double)cfx[i] * (1.0 / HFX_FX_SCALE) - bound;
        }
    }

    /* 5. order by (exact desc, position asc), emit N */
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
/* Exact route: every candidate of a flagged spectrum scored exactly     */
/* (tile-parallel, persistent), then a 64-bit radix select of the top N. */
/* Runs on the device after k_topn without a host round trip.            */
/* ------------------------------------------------------------------ */

#define TX_MAXN 64
#define TX_CHUNK 256
/* a flagged spectrum's tiles also number at least its band count / TX_LIVE,
 * so that the candidates the exact route re-scores spread over the tiles */
#define TX_LIVE 32
/* a wave fetches the metadata of its share of a candidate list one lane
 * per candidate */
static_assert(TN_MAXT / TN_WAVES <= HFX_WAVE && TN_EXT / TN_WAVES <= HFX_WAVE &&
                  TX_CHUNK / TN_WAVES <= HFX_WAVE,
              "one lane per candidate of a wave's share");

/* tiles per flagged spectrum, and the flagged spectra listed (in any
 * order) for the final select; counter[1] counts them */
template <class Cands>
__global__ void k_flag_tiles(uint32_t *tiles, uint32_t *list, uint32_t *counter,
                             const uint32_t *flag, const uint32_t *band, Cands c, uint32_t S)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const bool f = flag[s] != 0u;
    uint32_t nt = 0;
    if (f) {
        const uint32_t n = c.count(s), b = min(band[s], n);
        nt = max((n + TX_CHUNK - 1) / TX_CHUNK, (b + TX_LIVE - 1) / TX_LIVE);
    }
    tiles[s] = nt;
    if (f) list[atomicAdd(&counter[1], 1u)] = s;
}

__global__ void k_tx_reset(uint32_t *counter)
{
    counter[0] = 0;
    counter[1] = 0;
}

/* tile -> spectrum, one thread per spectrum writing its own tiles */
__global__ void k_tile_spec(uint32_t *tile_spec, const uint64_t *tile_off, uint32_t S)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S)
        for (uint64_t t = tile_off[s]; t < tile_off[s + 1]; ++t) tile_spec[t] = s;
}

__global__ void k_zero_flags(uint32_t *flag, uint32_t S)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S) flag[i] = 0;
}
__global__ void k_mark_list(uint32_t *flag, double *tau, uint32_t *band, const uint32_t *list,
                            uint32_t num_list, uint32_t S)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < num_list && list[i] < S) {
        flag[list[i]] = 1;
        tau[list[i]] = -INFINITY; /* every candidate scored exactly */
        band[list[i]] = 0xffffffffu;
    }
}

/* persistent: workgroup g takes tiles g, g + grid, ... of the flagged
 * spectra (about one each: no shared counter to queue on).  A spectrum's nt
 * tiles interleave its candidates (tile k takes k, k + nt, k + 2 nt, ...):
 * candidates near tau sit together in the window's order, and so are
 * spread over the tiles. */
template <class Cands>
__global__ void __launch_bounds__(TN_THREADS)
    k_exact_tiles(int64_t *fxbuf, const uint64_t *tile_off, const uint32_t *tile_spec, Cands c,
                  hfx_spectra sp, const float *scores, const double *tau)
{
    extern __shared__ __align__(16) unsigned char tn_dyn[];
    const uint32_t xcap = tn_xcap(sp.x_stride), mwords = xcap / 8;
    float *xs = (float *)tn_dyn;
    uint32_t *maps = (uint32_t *)(tn_dyn + tn_xs_bytes(sp.x_stride));
    __shared__ float mtab[32];
    const uint32_t wave = threadIdx.x / HFX_WAVE, lane = hfx_lane();
    const uint32_t S = sp.num_spec;
    const uint64_t ntiles = tile_off[S];
    uint32_t *map = maps + wave * mwords;
    for (uint32_t w = lane; w < mwords; w += HFX_WAVE) map[w] = 0;
    uint32_t cur = 0xffffffffu, len = 0;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t s = tile_spec[t];
        if (s != cur) {
            len = hfx_stage_spectrum(xs, mtab, c.db, sp, s);
            cur = s;
            __syncthreads();
        }
        const uint32_t k = (uint32_t)(t - tile_off[s]);
        const uint32_t nt = (uint32_t)(tile_off[s + 1] - tile_off[s]);
        const uint32_t n = c.count(s);
        const uint32_t zeff = c.zeff(sp.d_zmax[s]);
        const double ts = tau[s];
        /* the wave's candidates: tile slots m = wave + 4 j (j < 64), lane j
         * reading candidate i = k + nt m's score.  A candidate below tau
         * cannot reach the top N (its exact score is below the N-th exact
         * score): it keeps the smallest key, stored by its own lane.  The
         * others' metadata is fetched by their lanes, and the wave re-scores
         * them one by one. */
        const uint64_t tb = c.base(s);
        const uint64_t mi64 = k + (uint64_t)nt * (wave + TN_WAVES * lane);
        const uint32_t mi = (uint32_t)mi64;
        bool live = false;
        if (wave + TN_WAVES * lane < TX_CHUNK && mi64 < n) {
            live = !((double)scores[tb + mi] < ts);
            if (!live) fxbuf[tb + mi] = INT64_MIN;
        }
        typename Cands::Prep mine{};
        if (live) mine = c.prep(s, mi);
        const uint64_t lm = __ballot(live);
        /* the next live candidate's first residues are loaded while this one
         * is scored */
        uint32_t res = lm ? c.first_res(Cands::take(mine, (uint32_t)__builtin_ctzll(lm))) : 0u;
        for (uint64_t m = lm; m; m &= m - 1ull) {
            const uint32_t j = (uint32_t)__builtin_ctzll(m);
            const uint32_t cur_res = res;
            const uint64_t rest = m & (m - 1ull);
            if (rest) res = c.first_res(Cands::take(mine, (uint32_t)__builtin_ctzll(rest)));
            c.peaks_first(map, mtab, Cands::take(mine, j), cur_res, len, zeff);
            hfx_wave_sync();
            const int64_t fx = hfx_wave_dot_fx(map, xs, len);
            hfx_wave_sync();
            if (lane == 0) fxbuf[tb + hfx_take(mi, j)] = fx;
        }
        __syncthreads();
    }
}

/* top N of a flagged spectrum by (exact desc, position asc).
 *
 * Only the live candidates count: the exact route leaves INT64_MIN for every
 * candidate below tau, and more than N are live.  They are gathered into LDS
 * in one sweep (up to FX_CAP of them) and selected there by radix over the
 * bits in which their 64-bit keys differ, 11 bits a pass, then, for keys
 * tied in full, over the complemented position.  Windows with more live
 * candidates (or fewer than N) take the 8-bit radix select over global
 * memory.  A persistent grid walks the list of flagged spectra. */
#define FX_CAP 4096
/* a workgroup per flagged spectrum, of few: sixteen waves each, so that a
 * sweep over the window has 8192 loads in flight */
#define FX_THREADS 1024
#define FX_WAVES (FX_THREADS / HFX_WAVE)
#define FX_DBITS 11

template <class Cands>
__global__ void __launch_bounds__(FX_THREADS)
    k_topn_fx(uint32_t *top_pos, double *top_exact, float *top_f32, uint32_t *top_cnt,
              uint32_t *flag, uint32_t N, const uint32_t *list, const uint32_t *list_cnt,
              const int64_t *fxbuf, const float *scores, Cands c)
{
    __shared__ uint64_t lk[FX_CAP];
    __shared__ uint32_t lp[FX_CAP];
    __shared__ uint32_t hist[1u << FX_DBITS];
    __shared__ uint32_t sel_p[TX_MAXN];
    __shared__ int64_t sel_f[TX_MAXN];
    __shared__ uint32_t wcnt[FX_WAVES];
    __shared__ uint32_t sh_digit, sh_rem, sh_gt, sh_cnt;
    __shared__ uint64_t sh_diff;
    const uint32_t tid = threadIdx.x, wave = tid / HFX_WAVE, lane = hfx_lane();
    const uint32_t F = *list_cnt;

    for (uint32_t q = blockIdx.x; q < F; q += gridDim.x) {
        const uint32_t s = list[q];
        const uint32_t n = c.count(s);
        const int64_t *fx = fxbuf + c.base(s);
        const float *sc = scores + c.base(s);
        uint32_t want = min(N, n);
        if (want == 0) {
            if (tid == 0) top_cnt[s] = 0;
            continue;
        }
        auto key = [&](uint32_t i) { return (uint64_t)fx[i] ^ 0x8000000000000000ull; };
        /* f(i, key) over this thread's keys, eight loads in flight
         * (clamped into the window); i >= n past it, where all lanes
         * still take part */
        auto for_keys = [&](auto &&f) {
            constexpr uint32_t U = 8;
            for (uint32_t base = 0; base < n; base += U * FX_THREADS) {
                uint64_t kk[U];
#pragma unroll
                for (uint32_t u = 0; u < U; ++u)
                    kk[u] = key(min(base + u * FX_THREADS + tid, n - 1u));
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) f(base + u * FX_THREADS + tid, kk[u]);
            }
        };

        /* 1. the live keys (key 0 is INT64_MIN) into LDS, one LDS atomic
         * per wave and batch */
        if (tid == 0) {
            sh_cnt = 0;
            sh_gt = 0;
            sh_diff = 0;
            sh_rem = want;
        }
        __syncthreads();
        for_keys([&](uint32_t i, uint64_t k) {
            const bool live = i < n && k != 0ull;
            const uint64_t b = __ballot(live);
            if (b) {
                const uint32_t leader = (uint32_t)__builtin_ctzll(b);
                uint32_t wb = 0;
                if (lane == leader) wb = atomicAdd(&sh_cnt, (uint32_t)__builtin_popcountll(b));
                wb = (uint32_t)__builtin_amdgcn_readlane((int)wb, (int)leader);
                const uint32_t slot =
                    wb + (uint32_t)__builtin_popcountll(b & ((1ull << lane) - 1ull));
                if (live && slot < FX_CAP) {
                    lk[slot] = k;
                    lp[slot] = i;
                }
            }
        });
        __syncthreads();
        const uint32_t M = sh_cnt;
        if (M < want) {
            /* Fewer live candidates than N: the N-th exact score of the
             * selection lies above tau + bound, so at least N candidates
             * have exact scores >= tau + bound, and each of them whose FP32
             * score fell below tau lies further below its exact score than
             * the stated bound allows.  Only the live candidates are ranked
             * and the spectrum is marked for the host. */
            want = M;
            if (tid == 0) {
                flag[s] |= HFX_FLAG_SHORT;
                sh_rem = want;
                if (want == 0) top_cnt[s] = 0;
            }
            __syncthreads();
            if (want == 0) continue;
        }
        if (M <= FX_CAP) {
            /* 2. the bits in which the live keys differ */
            const uint64_t k0 = lk[0];
            uint64_t d = 0;
            for (uint32_t i = tid; i < M; i += FX_THREADS) d |= lk[i] ^ k0;
            if (d) atomicOr((unsigned long long *)&sh_diff, (unsigned long long)d);
            __syncthreads();
            const uint64_t diff = sh_diff;
            /* 3. radix select over the key: prefix/pmask fix the bits
             * found so far; done once the bucket is taken whole */
            const int top = diff ? 63 - __builtin_clzll(diff) : -1;
            /* bits above the top differing one (none when it is bit 63:
             * 2 << 63 wraps to 0) */
            uint64_t pmask = diff ? ~((2ull << top) - 1ull) : ~0ull;
            uint64_t prefix = k0 & pmask;
            int hi = top;
            bool whole = false;
            while (hi >= 0 && !whole) {
                const int lo = hi - (FX_DBITS - 1) > 0 ? hi - (FX_DBITS - 1) : 0;
                const uint32_t width = (uint32_t)(hi - lo + 1), nb = 1u << width;
                for (uint32_t b = tid; b < nb; b += FX_THREADS) hist[b] = 0;
                __syncthreads();
                for (uint32_t i = tid; i < M; i += FX_THREADS) {
                    const uint64_t k = lk[i];
                    if ((k & pmask) == prefix)
                        atomicAdd(&hist[(uint32_t)(k >> lo) & (nb - 1u)], 1u);
                }
                __syncthreads();
                tn_select_digit(hist, nb, &sh_digit, &sh_rem);
                prefix |= (uint64_t)sh_digit << lo;
                pmask |= (uint64_t)(nb - 1u) << lo;
                whole = hist[sh_digit] == sh_rem;
                hi = lo - 1;
                __syncthreads();
            }
            /* 4. keys tied in full (pmask all ones): the rem smallest
             * positions among them, by radix over ~position */
            uint32_t qprefix = 0, qmask = 0;
            if (!whole) {
                for (int qh = 31; qh >= 0 && !whole;) {
                    const int lo = qh - (FX_DBITS - 1) > 0 ? qh - (FX_DBITS - 1) : 0;
                    const uint32_t width = (uint32_t)(qh - lo + 1), nb = 1u << width;
                    for (uint32_t b = tid; b < nb; b += FX_THREADS) hist[b] = 0;
                    __syncthreads();
                    for (uint32_t i = tid; i < M; i += FX_THREADS) {
                        const uint32_t q = ~lp[i];
                        if (lk[i] == prefix && (q & qmask) == qprefix)
                            atomicAdd(&hist[(q >> lo) & (nb - 1u)], 1u);
                    }
                    __syncthreads();
                    tn_select_digit(hist, nb, &sh_digit, &sh_rem);
                    qprefix |= sh_digit << lo;
                    qmask |= (nb - 1u) << lo;
                    whole = hist[sh_digit] == sh_rem;
                    qh = lo - 1;
                    __syncthreads();
                }
            }
            /* 5. the selection: keys above the prefix, and the bucket
             * (or the tied key's chosen positions) -- want of them */
            for (uint32_t i = tid; i < M; i += FX_THREADS) {
                const uint64_t k = lk[i];
                const uint32_t q = ~lp[i];
                const bool take = qmask == 0u
                                      ? (k & pmask) >= prefix
                                      : (k > prefix || (k == prefix && (q & qmask) >= qprefix));
                if (take) {
                    const uint32_t slot = atomicAdd(&sh_gt, 1u);
                    if (slot < want) {
                        sel_p[slot] = lp[i];
                        sel_f[slot] = (int64_t)(k ^ 0x8000000000000000ull);
                    }
                }
            }
            __syncthreads();
        } else {
            /* the 8-bit radix select over global memory */
            uint64_t prefix = 0, pmask = 0;
            for (int shift = 56; shift >= 0; shift -= 8) {
                if (tid < 256u) hist[tid] = 0;
                __syncthreads();
                for_keys([&](uint32_t i, uint64_t k) {
                    tn_hist_add(hist, (uint32_t)(k >> shift) & 255u,
                                i < n && (k & pmask) == prefix);
                });
                __syncthreads();
                tn_select_digit(hist, 256u, &sh_digit, &sh_rem);
                prefix |= (uint64_t)sh_digit << shift;
                pmask |= 255ull << shift;
            }
            const uint64_t K = prefix;
            const uint32_t need_eq = sh_rem, n_gt = want - need_eq;
            for_keys([&](uint32_t i, uint64_t k) {
                if (i < n && k > K) {
                    const uint32_t slot = atomicAdd(&sh_gt, 1u);
                    if (slot < n_gt) {
                        sel_p[slot] = i;
                        sel_f[slot] = (int64_t)(k ^ 0x8000000000000000ull);
                    }
                }
            });
            uint32_t eq_base = 0;
            for (uint32_t base = 0; base < n && eq_base < need_eq; base += FX_THREADS) {
                const uint32_t i = base + tid;
                const bool f = i < n && key(i) == K;
                const uint64_t bal = __ballot(f);
                const uint32_t wr = (uint32_t)__builtin_popcountll(bal & ((1ull << lane) - 1ull));
                if (lane == 0) wcnt[wave] = (uint32_t)__builtin_popcountll(bal);
                __syncthreads();
                uint32_t wpre = 0, tot = 0;
                for (uint32_t w = 0; w < FX_WAVES; ++w) {
                    if (w < wave) wpre += wcnt[w];
                    tot += wcnt[w];
                }
                const uint32_t r = eq_base + wpre + wr;
                if (f && r < need_eq) {
                    sel_p[n_gt + r] = i;
                    sel_f[n_gt + r] = fx[i];
                }
                eq_base += tot;
                __syncthreads();
            }
            __syncthreads();
        }
        for (uint32_t i = tid; i < want; i += FX_THREADS) {
            uint32_t rk = 0;
            for (uint32_t j = 0; j < want; ++j)
                rk += before_fx(sel_f[j], sel_p[j], sel_f[i], sel_p[i]) ? 1u : 0u;
            top_pos[(size_t)s * N + rk] = sel_p[i];
            top_exact[(size_t)s * N + rk] = (double)sel_f[i] * (1.0 / HFX_FX_SCALE);
            top_f32[(size_t)s * N + rk] = sc[sel_p[i]];
        }
        if (tid == 0) top_cnt[s] = want;
        __syncthreads();
    }
}

/* ------------------------------------------------------------------ */
/* Launch helpers                                                        */
/* ------------------------------------------------------------------ */

/* tiles: max(ceil(n / TX_CHUNK), ceil(band / TX_LIVE)) per spectrum, band <= n */
static uint64_t tx_max_tiles(uint32_t S, uint64_t total)
{
    return total / TX_CHUNK + total / TX_LIVE + 2ull * S;
}

size_t hfx_topn_scratch_bytes(uint32_t S, uint64_t total)
{
    return total * 8 + ((size_t)S + 1) * 8 + (size_t)S * 8 + (size_t)S * 4 + 256 +
           tx_max_tiles(S, total) * 4 + (size_t)S * 8;
}

/* exact-route scratch:
 * fxbuf[total] | tile_off[S+1] | tau[S] | tiles[S] | counters (256 B) |
 * tile_spec[max tiles] | list[S] | band[S] */
struct TxScratch {
    int64_t *fxbuf;
    uint64_t *tile_off;
    double *tau;
    uint32_t *tiles, *counter, *tile_spec, *list, *band;
};
static TxScratch tx_carve(void *p, uint32_t S, uint64_t total)
{
    TxScratch x;
    x.fxbuf = (int64_t *)p;
    x.tile_off = (uint64_t *)(x.fxbuf + total);
    x.tau = (double *)(x.tile_off + S + 1);
    x.tiles = (uint32_t *)(x.tau + S);
    x.counter = x.tiles + S;
    x.tile_spec = x.counter + 64;
    x.list = x.tile_spec + tx_max_tiles(S, total);
    x.band = x.list + S;
    return x;
}
static int tx_scratch(hfx_ctx *ctx, void *given, uint32_t S, uint64_t total, void **out)
{
    *out = given;
    if (given) return HFX_OK;
    return hfx_ctx_scratch(ctx, HFX_SLOT_TOPN, hfx_topn_scratch_bytes(S, total), out);
}

template <class Cands>
static int exact_route(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                       uint32_t *d_top_cnt, uint32_t N, uint32_t *d_flag,
                       const float *d_scores, const Cands &c, uint64_t total,
                       const hfx_spectra *sp, void *scratch, hipStream_t st)
{
    const uint32_t S = sp->num_spec;
    if (S == 0) return HFX_OK;
    HFX_REQUIRE(N >= 1 && N <= TX_MAXN, HFX_ERR_ARG, "top-N: N must be in 1..%d", TX_MAXN);
    void *p;
    int rc;
    if ((rc = tx_scratch(ctx, scratch, S, total, &p))) return rc;
    const TxScratch x = tx_carve(p, S, total);
    hipLaunchKernelGGL(k_tx_reset, dim3(1), dim3(1), 0, st, x.counter);
    hipLaunchKernelGGL(k_flag_tiles<Cands>, dim3((S + 255) / 256), dim3(256), 0, st, x.tiles,
                       x.list, x.counter, d_flag, x.band, c, S);
    if ((rc = hfx_scan_offsets(ctx, x.tile_off, x.tiles, S, (void *)st)) != HFX_OK) return rc;
    hipLaunchKernelGGL(k_tile_spec, dim3((S + 255) / 256), dim3(256), 0, st, x.tile_spec,
                       x.tile_off, S);
    hipLaunchKernelGGL(k_exact_tiles<Cands>, dim3((uint32_t)hfx_ctx_num_cus(ctx) * 8),
                       dim3(TN_THREADS), tn_dyn_bytes(sp->x_stride), st, x.fxbuf, x.tile_off,
                       x.tile_spec, c, *sp, d_scores, x.tau);
    hipLaunchKernelGGL(k_topn_fx<Cands>, dim3(min(S, (uint32_t)hfx_ctx_num_cus(ctx))),
                       dim3(FX_THREADS), 0, st, d_top_pos, d_top_exact, d_top_f32, d_top_cnt,
                       d_flag, N, x.list, x.counter + 1, x.fxbuf, d_scores, c);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

template <class Cands>
static int launch_topn(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                       uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N, uint32_t margin,
                       const float *d_scores, const Cands &c, uint64_t total,
                       const hfx_spectra *sp, float rel0, float rel1, float relq, void *scratch,
                       void *stream)
{
    HFX_REQUIRE(ctx && d_top_pos && d_top_exact && d_top_f32 && d_top_cnt && d_flag && d_scores &&
                    sp,
                HFX_ERR_ARG, "hfx_topn: null argument");
    HFX_REQUIRE(N >= 1 && N <= TX_MAXN && N + margin <= TN_MAXT, HFX_ERR_ARG,
                "hfx_topn: need 1 <= N <= %d and N + margin <= %d (N=%u margin=%u)", TX_MAXN,
                TN_MAXT, N, margin);
    if (sp->num_spec == 0) return HFX_OK;
    hipStream_t st = hfx_stream(stream);
    void *p;
    int rc;
    if ((rc = tx_scratch(ctx, scratch, sp->num_spec, total, &p))) return rc;
    const TxScratch x = tx_carve(p, sp->num_spec, total);
    /* the launch order lives in the flagged-spectrum list, which the exact
     * route fills only after k_topn */
    hipLaunchKernelGGL(k_topn_order<Cands>, dim3(1), dim3(TN_OTHREADS), 0, st, x.list, c,
                       sp->num_spec);
    hipLaunchKernelGGL(k_topn<Cands>, dim3(sp->num_spec), dim3(TN_THREADS),
                       tn_dyn_bytes(sp->x_stride), st, d_top_pos, d_top_exact, d_top_f32,
                       d_top_cnt, d_flag, N, margin, d_scores, c, *sp, rel0, rel1, relq, x.tau,
                       x.band, x.list);
    HFX_CHECK_LAUNCH();
    return exact_route(ctx, d_top_pos, d_top_exact, d_top_f32, d_top_cnt, N, d_flag, d_scores, c,
                       total, sp, p, st);
}

template <class Cands>
static int launch_topn_exact(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact,
                             float *d_top_f32, uint32_t *d_top_cnt, uint32_t N,
                             const uint32_t *d_list, uint32_t num_list, const float *d_scores,
                             const Cands &c, uint64_t total, const hfx_spectra *sp, void *stream)
{
    HFX_REQUIRE(ctx && d_top_pos && d_top_exact && d_top_f32 && d_top_cnt && d_list && sp,
                HFX_ERR_ARG, "hfx_topn_exact: null argument");
    const uint32_t S = sp->num_spec;
    if (S == 0 || num_list == 0) return HFX_OK;
    hipStream_t st = hfx_stream(stream);
    void *p, *q;
    int rc;
    if ((rc = hfx_ctx_scratch(ctx, HFX_SLOT_FLAGS, 4ull * S + 64, &p)) != HFX_OK) return rc;
    if ((rc = tx_scratch(ctx, nullptr, S, total, &q))) return rc;
    uint32_t *flag = (uint32_t *)p;
    const TxScratch x = tx_carve(q, S, total);
    hipLaunchKernelGGL(k_zero_flags, dim3((S + 255) / 256), dim3(256), 0, st, flag, S);
    hipLaunchKernelGGL(k_mark_list, dim3((num_list + 255) / 256), dim3(256), 0, st, flag, x.tau,
                       x.band, d_list, num_list, S);
    return exact_route(ctx, d_top_pos, d_top_exact, d_top_f32, d_top_cnt, N, flag, d_scores, c,
                       total, sp, q, st);
}

/* used by the batched path: the whole top-N on the device */
int hfx_topn_device(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                    uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N, uint32_t margin,
                    const float *d_scores, const uint64_t *d_off, uint64_t total,
                    const uint32_t *d_begin, const uint32_t *d_num, const hfx_db *db,
                    const hfx_spectra *sp, float rel0, float rel1, float relq, void *scratch,
                    void *stream)
{
    HFX_REQUIRE(rel0 >= 0.0f && rel1 >= 0.0f && relq >= 0.0f && rel0 < 1.0f && rel1 < 1.0f &&
                    relq < 1.0f,
                HFX_ERR_ARG,
                "top-N: error bound factors must lie in [0, 1) (rel0=%g rel1=%g relq=%g)", rel0,
                rel1, relq);
    HfxPlainCands c{*db, d_begin, d_num, d_off};
    return launch_topn(ctx, d_top_pos, d_top_exact, d_top_f32, d_top_cnt, d_flag, N, margin,
                       d_scores, c, total, sp, rel0, rel1, relq, scratch, stream);
}

extern "C" int hfx_topn(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                        uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N, uint32_t margin,
                        const float *d_scores, const uint64_t *d_off, uint64_t total,
                        const uint32_t *d_begin, const uint32_t *d_num, const hfx_db *db,
                        const hfx_spectra *sp, void *stream)
{
    HFX_REQUIRE(d_off && d_begin && d_num && db, HFX_ERR_ARG, "hfx_topn: null argument");
    return hfx_topn_device(ctx, d_top_pos, d_top_exact, d_top_f32, d_top_cnt, d_flag, N, margin,
                           d_scores, d_off, total, d_begin, d_num, db, sp, HFX_TOPN_REL_F32, 0.0f,
                           0.0f, nullptr, stream);
}

extern "C" int hfx_topn_rel(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact,
                            float *d_top_f32, uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N,
                            uint32_t margin, float rel0, float rel1, const float *d_scores,
                            const uint64_t *d_off, uint64_t total, const uint32_t *d_begin,
                            const uint32_t *d_num, const hfx_db *db, const hfx_spectra *sp,
                            void *stream)
{
    HFX_REQUIRE(d_off && d_begin && d_num && db, HFX_ERR_ARG, "hfx_topn_rel: null argument");
    return hfx_topn_device(ctx, d_top_pos, d_top_exact, d_top_f32, d_top_cnt, d_flag, N, margin,
                           d_scores, d_off, total, d_begin, d_num, db, sp, rel0, rel1, 0.0f,
                           nullptr, stream);
}

extern "C" int hfx_topn_rel_q(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact,
                              float *d_top_f32, uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N,
                              uint32_t margin, float rel0, float rel1, float relq,
                              const float *d_scores, const uint64_t *d_off, uint64_t total,
                              const uint32_t *d_begin, const uint32_t *d_num, const hfx_db *db,
                              const hfx_spectra *sp, void *stream)
{
    HFX_REQUIRE(d_off && d_begin && d_num && db, HFX_ERR_ARG, "hfx_topn_rel_q: null argument");
    return hfx_topn_device(ctx, d_top_pos, d_top_exact, d_top_f32, d_top_cnt, d_flag, N, margin,
                           d_scores, d_off, total, d_begin, d_num, db, sp, rel0, rel1, relq,
                           nullptr, stream);
}

extern "C" int hfx_topn_exact(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact,
                              float *d_top_f32, uint32_t *d_top_cnt, uint32_t N,
                              const uint32_t *d_spec_list, uint32_t num_list,
                              const float *d_scores, const uint64_t *d_off, uint64_t total,
                              const uint32_t *d_begin, const uint32_t *d_num, const hfx_db *db,
                              const hfx_spectra *sp, void *stream)
{
    HFX_REQUIRE(d_off && d_begin && d_num && db && d_scores, HFX_ERR_ARG,
                "hfx_topn_exact: null argument");
    HfxPlainCands c{*db, d_begin, d_num, d_off};
    return launch_topn_exact(ctx, d_top_pos, d_top_exact, d_top_f32, d_top_cnt, N, d_spec_list,
                             num_list, d_scores, c, total, sp, stream);
}

extern "C" int hfx_topn_mod(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact,
                            float *d_top_f32, uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N,
                            uint32_t margin, const float *d_scores, const uint64_t *d_mp_off,
                            const hfx_modset *ms, const hfx_db *db, const hfx_spectra *sp,
                            void *stream)
{
    HFX_REQUIRE(d_mp_off && ms && db, HFX_ERR_ARG, "hfx_topn_mod: null argument");
    HFX_REQUIRE(ms->num_ma <= HFX_MAX_MA, HFX_ERR_ARG, "hfx_topn_mod: num_ma %u", ms->num_ma);
    HfxModCands c{*db, *ms, d_mp_off};
    return launch_topn(ctx, d_top_pos, d_top_exact, d_top_f32, d_top_cnt, d_flag, N, margin,
                       d_scores, c, ms->num_mpep, sp, HFX_TOPN_REL_F32, 0.0f, 0.0f, nullptr,
                       stream);
}

extern "C" int hfx_topn_exact_mod(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact,
                                  float *d_top_f32, uint32_t *d_top_cnt, uint32_t N,
                                  const uint32_t *d_spec_list, uint32_t num_list,
                                  const float *d_scores, const uint64_t *d_mp_off,
                                  const hfx_modset *ms, const hfx_db *db, const hfx_spectra *sp,
                                  void *stream)
{
    HFX_REQUIRE(d_mp_off && ms && db && d_scores, HFX_ERR_ARG, "hfx_topn_exact_mod: null argument");
    HFX_REQUIRE(ms->num_ma <= HFX_MAX_MA, HFX_ERR_ARG, "hfx_topn_exact_mod: num_ma %u", ms->num_ma);
    HfxModCands c{*db, *ms, d_mp_off};
    return launch_topn_exact(ctx, d_top_pos, d_top_exact, d_top_f32, d_top_cnt, N, d_spec_list,
                             num_list, d_scores, c, ms->num_mpep, sp, stream);
}
