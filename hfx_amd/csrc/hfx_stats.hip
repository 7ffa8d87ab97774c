/*
 * hfx_stats.hip -- E-values and SP ion matches of a search, on the device.
 *
 * E-values (Util/Statistics.hs:37-100, applied per spectrum by
 * Search.hs:118-123 to every match of the spectrum, unmodified and modified):
 * the reference sorts all scores ascending, keeps sorted positions
 * [round(0.3 n), round((1 - p/100) n)), shifts them by 1 - min, drops those at
 * or below 4000, histograms the rest in bins of width w and fits log10 of the
 * heights against the bin centres by least squares (statistics-linreg).
 *
 * No sort is needed here: x -> bin is monotone, so the sorted positions of the
 * scores of one bin form the range [C(k-1), C(k)) of the cumulative bin counts
 * C, and the kept count of bin k is that range's overlap with the cut.  One
 * workgroup per spectrum:
 *   1. min and max of the spectrum's scores (kept in LDS when they fit);
 *   2. key(x) = 0 for (x + 1) - min <= 4000, else ceil(((x + 1) - min) / w)
 *      - kbase (>= 1), counted per key in LDS (EV_KEYS keys per pass; a wider
 *      spread takes one pass per range of keys);
 *   3. a block scan of the counts gives every key's sorted range, hence the
 *      kept counts and the histogram points (x_k, log10(c_k / w));
 *   4. least squares in FP64, two passes: means, then centred sums.
 * The score array is read once from HBM (twice when a spectrum holds more than
 * EV_CACHE scores); the rest is LDS work.
 *
 * SP ions (IonSeries.hs:99-147, Match.hs:33-40): extractPeaks keeps the 200
 * largest values of the spectrum and folds them by max into bins i div 3; a
 * b or y ion matches when its bin round(m/z / bin width / 3) holds a value
 * above zero.  k_sp_peaks_w turns each spectrum into those flags (one bit
 * per bin; one wave per spectrum, the row in registers, the 200th value
 * selected bit by bit over the float bits by ballot counts; k_sp_peaks does
 * the same with a workgroup for rows longer than 3072); k_sp_* match one
 * peptide per thread along its FP32 b ladder.
 */
#include "hfx_common.h"

#include <math.h>

#define ST_THREADS 256
#define ST_WAVES (ST_THREADS / HFX_WAVE)

#define EV_CACHE 8192               /* scores of a spectrum kept in LDS       */
#define EV_KEYS 1024                /* histogram keys per pass                */
#define EV_KPT (EV_KEYS / ST_THREADS)

#define SP_TOP 200                  /* IonSeries.hs:104                       */
#define SP_CACHE 8192               /* spectrum values kept in LDS            */
#define SP_WORDS 256                /* flag words: len / 3 + 1 <= 8192 bins   */

/* Host constants of the ion ladder (Mass.hs:19-27) */
#define SP_H 1.007276f
#define SP_H2O 18.010565f

/* ------------------------------------------------------------------ */
/* Block primitives (ST_THREADS threads)                                */
/* ------------------------------------------------------------------ */

template <typename T>
__device__ __forceinline__ T st_wave_sum(T v)
{
#pragma unroll
    for (int m = HFX_WAVE / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, HFX_WAVE);
    return v;
}

/* sum over the block, partial sums added in wave order (deterministic) */
template <typename T>
__device__ __forceinline__ T st_block_sum(T v, T *red)
{
    v = st_wave_sum(v);
    __syncthreads();
    if (threadIdx.x % HFX_WAVE == 0) red[threadIdx.x / HFX_WAVE] = v;
    __syncthreads();
    T t = red[0];
#pragma unroll
    for (int i = 1; i < ST_WAVES; ++i) t += red[i];
    return t;
}

/* exclusive prefix sum over the block in thread order */
template <typename T>
__device__ __forceinline__ T st_block_excl_scan(T v, T *red)
{
    const int lane = threadIdx.x % HFX_WAVE, w = threadIdx.x / HFX_WAVE;
    T x = v;
#pragma unroll
    for (int d = 1; d < HFX_WAVE; d <<= 1) {
        const T y = __shfl_up(x, d, HFX_WAVE);
        if (lane >= d) x += y;
    }
    __syncthreads();
    if (lane == HFX_WAVE - 1) red[w] = x;
    __syncthreads();
    T off = 0;
    for (int i = 0; i < w; ++i) off += red[i];
    return off + x - v;
}

/* ------------------------------------------------------------------ */
/* E-value fit                                                           */
/* ------------------------------------------------------------------ */

/* ------------------------------------------------------------------ */
/* Wave reductions over DPP lane moves (no LDS round trips)              */
/* ------------------------------------------------------------------ */

/* quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror:
 * every lane of a row of 16 ends with the row's total; row_bcast15 (rows 1
 * and 3) and row_bcast31 (rows 2 and 3) carry the rows into lane 63. */
#define ST_DPP32(old, v, ctrl, rm) \
    ((uint32_t)__builtin_amdgcn_update_dpp((int)(old), (int)(v), (ctrl), (rm), 0xF, false))

template <int CTRL, int RM>
__device__ __forceinline__ uint64_t st_dpp64(uint64_t v, uint64_t old)
{
    const uint32_t lo = ST_DPP32((uint32_t)old, (uint32_t)v, CTRL, RM);
    const uint32_t hi = ST_DPP32((uint32_t)(old >> 32), (uint32_t)(v >> 32), CTRL, RM);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t st_readlane63(uint64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
    return ((uint64_t)hi << 32) | lo;
}

/* op-reduction of a 64-bit value over the wave; ident is op's identity */
template <class Op>
__device__ __forceinline__ uint64_t st_wave_reduce64(uint64_t v, uint64_t ident, Op op)
{
    v = op(v, st_dpp64<0xB1, 0xF>(v, ident));
    v = op(v, st_dpp64<0x4E, 0xF>(v, ident));
    v = op(v, st_dpp64<0x141, 0xF>(v, ident));
    v = op(v, st_dpp64<0x140, 0xF>(v, ident));
    v = op(v, st_dpp64<0x142, 0xA>(v, ident));
    v = op(v, st_dpp64<0x143, 0xC>(v, ident));
    return st_readlane63(v);
}

__device__ __forceinline__ double st_wsum_f64(double v)
{
    const uint64_t r = st_wave_reduce64(__double_as_longlong(v), 0ull, [](uint64_t a, uint64_t b) {
        return (uint64_t)__double_as_longlong(__longlong_as_double(a) + __longlong_as_double(b));
    });
    return __longlong_as_double(r);
}

/* inclusive prefix sum of uint32 over the wave: row_shr 1, 2, 4, 8 within
 * rows of 16, then row_bcast15 / row_bcast31 across rows */
__device__ __forceinline__ uint32_t st_wave_incl_scan32(uint32_t x)
{
    x += ST_DPP32(0u, x, 0x111, 0xF);
    x += ST_DPP32(0u, x, 0x112, 0xF);
    x += ST_DPP32(0u, x, 0x114, 0xF);
    x += ST_DPP32(0u, x, 0x118, 0xF);
    x += ST_DPP32(0u, x, 0x142, 0xA);
    x += ST_DPP32(0u, x, 0x143, 0xC);
    return x;
}

/* ------------------------------------------------------------------ */
/* E-value fit                                                           */
/* ------------------------------------------------------------------ */

struct EvArgs {
    double *fit;        /* [S][2] alpha, beta          */
    uint32_t *npts;     /* [S] histogram points, opt.  */
    uint64_t *nused;    /* [S] scores kept, opt.       */
    const float *sa;
    const uint64_t *oa;
    const float *sb;    /* second score set, opt.      */
    const uint64_t *ob;
    double w;
    float perc;
};

#define EV_NSUM 6   /* sum dx, sum y, sum dx^2, sum dx y, points, kept */

struct EvShared {
    float cache[EV_CACHE];
    uint32_t hist[EV_KEYS];       /* counts per key, then kept counts of the points */
    uint16_t pkey[EV_KEYS];       /* key (less the range's base) of each point      */
    double redd[EV_NSUM * ST_WAVES];
    uint64_t redu[2 * ST_WAVES];
    uint32_t red32[ST_WAVES];
};

/* one spectrum's view: n scores, the cut, the key map */
struct EvSpec {
    const float *sa, *sb;
    uint64_t na, n;
    bool cached;
    int64_t cb, ce;
    double h0, w, kb, x0;

    double rw;   /* RN(1 / w) */

    /* 0 for (x + 1) - min <= 4000 (filtered out, Statistics.hs:76), else
     * the bin ceiling(y / w) less kbase: monotone in x.  q = y * RN(1/w) is
     * within 2 ulp of y / w (below 2^20: 5e-10), so its ceiling is that of
     * the correctly rounded quotient unless q lies within 1e-9 of an
     * integer: only then (or for q >= 2^20) divide. */
    __device__ __forceinline__ uint64_t key(float v) const
    {
        const double y = ((double)v + 1.0) - h0;
        if (!(y > 4000.0)) return 0;
        const double q = y * rw;
        const double f = q - floor(q);
        const double c = (f > 1e-9 && f < 1.0 - 1e-9 && q < 1048576.0) ? ceil(q) : ceil(y / w);
        const double kd = c - kb;
        return kd < 4.6e18 ? (uint64_t)kd : (uint64_t)4.6e18;
    }

    /* valBin = i * w - w / 2 with i = ceiling(y / w); x = valBin / 10000
     * (here times RN(1e-4): within an ulp) */
    __device__ __forceinline__ double xof(uint64_t k) const
    {
        return (((double)k + kb) * w - w / 2.0) * 1e-4;
    }
};

/* Histogram of keys [base, base + EV_KEYS); the kept part of each key's
 * scores becomes a point (x, logh), added to the thread's sums (x shifted by
 * x0).  Key 0 (filtered) is counted with the keys below the range, in
 * registers.  next = the smallest key above the range (~0 if none). */
__device__ __forceinline__ void ev_window(const EvSpec &e, EvShared &sh, uint64_t base,
                                          double *acc, uint64_t &next)
{
    const uint32_t tid = threadIdx.x, lane = tid % HFX_WAVE, wv = tid / HFX_WAVE;
    for (uint32_t k = tid; k < EV_KEYS; k += ST_THREADS) sh.hist[k] = 0;
    __syncthreads();
    uint32_t below = 0;
    uint64_t nxt = ~(uint64_t)0;
    auto count = [&](float v) {
        const uint64_t k = e.key(v);
        if (k < base || k == 0)
            ++below;
        else if (k - base < EV_KEYS)
            atomicAdd(&sh.hist[k - base], 1u);
        else
            nxt = k < nxt ? k : nxt;
    };
    if (e.cached) {
        /* four scores per thread at a time: their keys' latencies overlap */
        uint64_t i = tid;
        for (; i + 3 * ST_THREADS < e.n; i += 4 * ST_THREADS) {
            float v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = sh.cache[i + u * ST_THREADS];
#pragma unroll
            for (int u = 0; u < 4; ++u) count(v[u]);
        }
        for (; i < e.n; i += ST_THREADS) count(sh.cache[i]);
    } else {
        for (uint64_t i = tid; i < e.na; i += ST_THREADS) count(e.sa[i]);
        for (uint64_t i = tid; i < e.n - e.na; i += ST_THREADS) count(e.sb[i]);
    }
    /* block: sum of below, min of nxt (one barrier pair; the first barrier
     * also ends the counting) */
    const uint64_t bsum = st_wave_reduce64(below, 0ull, [](uint64_t a, uint64_t b) { return a + b; });
    nxt = st_wave_reduce64(nxt, ~0ull, [](uint64_t a, uint64_t b) { return a < b ? a : b; });
    __syncthreads();
    if (lane == 0) {
        sh.redu[2 * wv] = bsum;
        sh.redu[2 * wv + 1] = nxt;
    }
    __syncthreads();
    uint64_t tb = sh.redu[0];
    next = sh.redu[1];
#pragma unroll
    for (int i = 1; i < ST_WAVES; ++i) {
        tb += sh.redu[2 * i];
        next = sh.redu[2 * i + 1] < next ? sh.redu[2 * i + 1] : next;
    }
    /* sorted position of each key: exclusive block scan of the counts */
    uint32_t c[EV_KPT];
    uint32_t tot = 0;
#pragma unroll
    for (int j = 0; j < EV_KPT; ++j) {
        c[j] = sh.hist[tid * EV_KPT + j];
        tot += c[j];
    }
    const uint32_t incl = st_wave_incl_scan32(tot);
    if (lane == HFX_WAVE - 1) sh.red32[wv] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (uint32_t i = 0; i < wv; ++i) woff += sh.red32[i];
    int64_t pre = (int64_t)tb + (int64_t)(woff + incl - tot);
    uint32_t kept[EV_KPT], nv = 0;
#pragma unroll
    for (int j = 0; j < EV_KPT; ++j) {
        const uint64_t kk = base + tid * EV_KPT + j;
        const int64_t lo = pre, hi = pre + c[j];
        pre = hi;
        const int64_t kp = min(hi, e.ce) - max(lo, e.cb);
        kept[j] = (kk >= 1 && kp > 0) ? (uint32_t)kp : 0u;
        nv += kept[j] != 0;
    }
    /* the points, compacted in key order, one per thread at a time */
    const uint32_t incl2 = st_wave_incl_scan32(nv);
    __syncthreads();   /* every thread has read its counts and the first scan */
    if (lane == HFX_WAVE - 1) sh.red32[wv] = incl2;
    __syncthreads();
    uint32_t o = incl2 - nv, npt = 0;
#pragma unroll
    for (int i = 0; i < ST_WAVES; ++i) {
        o += i < (int)wv ? sh.red32[i] : 0u;
        npt += sh.red32[i];
    }
#pragma unroll
    for (int j = 0; j < EV_KPT; ++j)
        if (kept[j]) {
            sh.hist[o] = kept[j];
            sh.pkey[o] = (uint16_t)(tid * EV_KPT + j);
            ++o;
        }
    __syncthreads();
    for (uint32_t i = tid; i < npt; i += ST_THREADS) {
        /* logh = logBase 10 (count / w) = log h / log 10 (here log10: within
         * an ulp) */
        const double dx = e.xof(base + sh.pkey[i]) - e.x0;
        const double y = log10((double)sh.hist[i] / e.w);
        acc[0] += dx;
        acc[1] += y;
        acc[2] += dx * dx;
        acc[3] += dx * y;
        acc[4] += 1.0;
        acc[5] += (double)sh.hist[i];
    }
    __syncthreads();   /* the next range reuses hist */
}

/* min and max of p[0..cnt), copied to dst[0..cnt) when COPY; 16-byte loads
 * over the aligned middle, up to eight per thread in flight */
template <bool COPY>
__device__ __forceinline__ void ev_scan(const float *p, uint64_t cnt, float *dst, float &lo,
                                        float &hi)
{
    const uint32_t tid = threadIdx.x;
    const uint64_t h = min(cnt, (uint64_t)(((16u - ((uint32_t)(uintptr_t)p & 15u)) & 15u) / 4u));
    auto one = [&](uint64_t i, float v) {
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
        if (COPY) dst[i] = v;
    };
    if (tid < h) one(tid, p[tid]);
    const float4 *p4 = (const float4 *)(p + h);
    const uint64_t n4 = (cnt - h) / 4;
    for (uint64_t i = tid; i < n4; i += 8 * ST_THREADS) {
        float4 q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            q[u] = i + u * ST_THREADS < n4 ? p4[i + u * ST_THREADS]
                                           : make_float4(p[0], p[0], p[0], p[0]);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint64_t o = h + 4 * (i + u * ST_THREADS);
            lo = fminf(fminf(lo, q[u].x), fminf(q[u].y, fminf(q[u].z, q[u].w)));
            hi = fmaxf(fmaxf(hi, q[u].x), fmaxf(q[u].y, fmaxf(q[u].z, q[u].w)));
            if (COPY && i + u * ST_THREADS < n4) {
                dst[o] = q[u].x;
                dst[o + 1] = q[u].y;
                dst[o + 2] = q[u].z;
                dst[o + 3] = q[u].w;
            }
        }
    }
    for (uint64_t t = h + 4 * n4 + tid; t < cnt; t += ST_THREADS) one(t, p[t]);
}

__global__ __launch_bounds__(ST_THREADS) void k_evfit(EvArgs a)
{
    __shared__ EvShared sh;
    const uint32_t s = blockIdx.x, tid = threadIdx.x, lane = tid % HFX_WAVE;
    const uint32_t wv = tid / HFX_WAVE;
    EvSpec e;
    const uint64_t a0 = a.oa[s];
    e.na = a.oa[s + 1] - a0;
    e.sa = a.sa + a0;
    uint64_t nb = 0;
    e.sb = nullptr;
    if (a.sb) {
        const uint64_t b0 = a.ob[s];
        nb = a.ob[s + 1] - b0;
        e.sb = a.sb + b0;
    }
    e.n = e.na + nb;
    auto no_fit = [&]() {
        if (tid == 0) {
            a.fit[2 * s] = NAN;
            a.fit[2 * s + 1] = NAN;
            if (a.npts) a.npts[s] = 0;
            if (a.nused) a.nused[s] = 0;
        }
    };
    /* positions are counted in 32 bits */
    if (e.n == 0 || e.n > 0xFFFFFFFFull) {
        no_fit();
        return;
    }
    e.cached = e.n <= EV_CACHE;

    /* 1. min, max (and the LDS copy) */
    float lo = INFINITY, hi = -INFINITY;
    if (e.cached) {
        ev_scan<true>(e.sa, e.na, sh.cache, lo, hi);
        if (nb) ev_scan<true>(e.sb, nb, sh.cache + e.na, lo, hi);
    } else {
        ev_scan<false>(e.sa, e.na, nullptr, lo, hi);
        if (nb) ev_scan<false>(e.sb, nb, nullptr, lo, hi);
    }
    {
        const uint64_t mm = ((uint64_t)__float_as_uint(hi) << 32) | __float_as_uint(lo);
        const uint64_t r = st_wave_reduce64(mm, ((uint64_t)__float_as_uint(-INFINITY) << 32) |
                                                    __float_as_uint(INFINITY),
                                            [](uint64_t x, uint64_t y) {
            const float l = fminf(__uint_as_float((uint32_t)x), __uint_as_float((uint32_t)y));
            const float h = fmaxf(__uint_as_float((uint32_t)(x >> 32)),
                                  __uint_as_float((uint32_t)(y >> 32)));
            return ((uint64_t)__float_as_uint(h) << 32) | __float_as_uint(l);
        });
        __syncthreads();
        if (lane == 0) sh.redu[wv] = r;
        __syncthreads();
        lo = INFINITY;
        hi = -INFINITY;
#pragma unroll
        for (int i = 0; i < ST_WAVES; ++i) {
            lo = fminf(lo, __uint_as_float((uint32_t)sh.redu[i]));
            hi = fmaxf(hi, __uint_as_float((uint32_t)(sh.redu[i] >> 32)));
        }
    }
    if (!(isfinite(lo) && isfinite(hi))) {
        /* no histogram over non-finite scores */
        no_fit();
        return;
    }

    /* the cut (Statistics.hs:72-73): Float for cut_end, Double for cut_beg,
     * halves to even */
    const float fe = (1.0f - a.perc / 100.0f) * (float)e.n;
    e.ce = min((int64_t)e.n, max((int64_t)0, (int64_t)__builtin_rintf(fe)));
    e.cb = max((int64_t)0, (int64_t)rint((30.0 / 100.0) * (double)e.n));
    e.h0 = (double)lo;
    e.w = a.w;
    e.rw = 1.0 / a.w;
    e.kb = fmax(floor(4000.0 / a.w) - 1.0, 0.0);
    /* sums are taken about x0, the middle of the key range of the scores,
     * and centred at the end (no cancellation beyond the data's own spread) */
    const uint64_t khi = e.key(hi);
    e.x0 = khi ? 0.5 * (e.xof(1) + e.xof(khi)) : 0.0;

    /* 2-3. histogram points, one pass per occupied range of EV_KEYS keys */
    double acc[EV_NSUM] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    uint64_t next = 0;
    for (uint64_t base = 0; base != ~(uint64_t)0; base = next) ev_window(e, sh, base, acc, next);

    /* 4. least squares: sums over the block (wave order, deterministic) */
#pragma unroll
    for (int q = 0; q < EV_NSUM; ++q) acc[q] = st_wsum_f64(acc[q]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < EV_NSUM; ++q) sh.redd[EV_NSUM * wv + q] = acc[q];
    __syncthreads();
    if (tid == 0) {
        double t[EV_NSUM];
#pragma unroll
        for (int q = 0; q < EV_NSUM; ++q) {
            t[q] = sh.redd[q];
            for (int i = 1; i < ST_WAVES; ++i) t[q] += sh.redd[EV_NSUM * i + q];
        }
        const double m = t[4];
        const double sxx = t[2] - t[0] * t[0] / m;
        const double sxy = t[3] - t[0] * t[1] / m;
        const double beta = sxy / sxx;
        const double mx = e.x0 + t[0] / m, my = t[1] / m;
        a.fit[2 * s] = my - beta * mx;
        a.fit[2 * s + 1] = beta;
        if (a.npts) a.npts[s] = (uint32_t)m;
        if (a.nused) a.nused[s] = (uint64_t)t[5];
    }
}

/* Statistics.hs:41: E = 10 ** (alpha + beta * (score / 10000)) */
__global__ void k_evapply(double *ev, const double *fit, const double *score,
                          const uint32_t *cnt, uint32_t N, uint32_t S)
{
    const uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (uint64_t)N * S) return;
    const uint32_t s = (uint32_t)(m / N), k = (uint32_t)(m % N);
    if (cnt && k >= cnt[s]) {
        ev[m] = 0.0;
        return;
    }
    ev[m] = pow(10.0, fit[2 * s] + fit[2 * s + 1] * (score[m] / 10000.0));
}

/* ------------------------------------------------------------------ */
/* SP ions                                                               */
/* ------------------------------------------------------------------ */

struct SpShared {
    float xs[SP_CACHE];
    uint32_t flags[SP_WORDS];
    uint32_t redu[ST_WAVES];
};

/* positive values by their bits (order preserving), the rest as 0 */
__device__ __forceinline__ uint32_t sp_key(float v) { return v > 0.0f ? __float_as_uint(v) : 0u; }

/* extractPeaks (IonSeries.hs:99-107) of one spectrum per workgroup: bit t of
 * the W flag words set where bin t = i div 3 holds one of the 200 largest
 * values and that value is above zero.  Ties at the 200th value keep the
 * lowest positions. */
__global__ __launch_bounds__(ST_THREADS) void k_sp_peaks(uint32_t *on, uint32_t W, hfx_spectra sp)
{
    __shared__ SpShared sh;
    const uint32_t s = blockIdx.x, tid = threadIdx.x;
    const uint32_t L = min(sp.d_len[s], sp.x_stride);
    const float *row = sp.d_x + (size_t)s * sp.x_stride;
    const bool cached = L <= SP_CACHE;
    if (cached)
        for (uint32_t i = tid; i < L; i += ST_THREADS) sh.xs[i] = row[i];
    for (uint32_t k = tid; k < W; k += ST_THREADS) sh.flags[k] = 0;
    __syncthreads();
    const float *src = cached ? sh.xs : row;

    uint32_t P = 0;
    for (uint32_t i = tid; i < L; i += ST_THREADS) P += src[i] > 0.0f;
    P = st_block_sum(P, sh.redu);

    if (P > SP_TOP) {
        /* T = key of the 200th largest value: the largest T with
         * #{key >= T} >= 200, bit by bit from the top */
        uint32_t T = 0;
        for (int bit = 30; bit >= 0; --bit) {
            const uint32_t cand = T | (1u << bit);
            uint32_t c = 0;
            for (uint32_t i = tid; i < L; i += ST_THREADS) c += sp_key(src[i]) >= cand;
            c = st_block_sum(c, sh.redu);
            if (c >= SP_TOP) T = cand;
        }
        uint32_t gt = 0;
        for (uint32_t i = tid; i < L; i += ST_THREADS) gt += sp_key(src[i]) > T;
        gt = st_block_sum(gt, sh.redu);
        const uint32_t need = SP_TOP - gt;
        /* ties in position order: contiguous runs per thread */
        const uint32_t per = (L + ST_THREADS - 1) / ST_THREADS;
        const uint32_t i0 = min(L, tid * per), i1 = min(L, i0 + per);
        uint32_t ties = 0;
        for (uint32_t i = i0; i < i1; ++i) ties += sp_key(src[i]) == T;
        uint32_t rank = st_block_excl_scan(ties, sh.redu);
        for (uint32_t i = i0; i < i1; ++i) {
            const uint32_t k = sp_key(src[i]);
            const bool sel = k > T || (k == T && rank++ < need);
            if (sel) atomicOr(&sh.flags[(i / 3) >> 5], 1u << ((i / 3) & 31));
        }
    } else {
        for (uint32_t i = tid; i < L; i += ST_THREADS)
            if (src[i] > 0.0f) atomicOr(&sh.flags[(i / 3) >> 5], 1u << ((i / 3) & 31));
    }
    __syncthreads();
    for (uint32_t k = tid; k < W; k += ST_THREADS) on[(size_t)s * W + k] = sh.flags[k];
}

#define SPW_MAXE 48                                          /* values per lane */
#define SPW_WORDS ((SPW_MAXE * HFX_WAVE / 3 + 1 + 31) / 32)     /* 33 flag words  */

/* k_sp_peaks for rows of at most E * 64 values (E <= SPW_MAXE): one wave per
 * spectrum, the row in registers (value j of a lane is position lane + 64 j,
 * 0 past the row, which no count sees), every count a ballot; the flag words
 * go through LDS. */
template <int E>
__global__ __launch_bounds__(ST_THREADS) void k_sp_peaks_w(uint32_t *on, uint32_t W,
                                                            hfx_spectra sp)
{
    __shared__ uint32_t words[ST_WAVES * SPW_WORDS];
    const uint32_t wv = threadIdx.x / HFX_WAVE, lane = threadIdx.x % HFX_WAVE;
    const uint32_t s = blockIdx.x * ST_WAVES + wv;
    uint32_t *fw = words + wv * SPW_WORDS;
    for (uint32_t k = lane; k < SPW_WORDS; k += HFX_WAVE) fw[k] = 0;
    __syncthreads();
    const bool live = s < sp.num_spec;
    const uint32_t L = live ? min(min(sp.d_len[s], sp.x_stride), (uint32_t)(E * HFX_WAVE)) : 0u;
    const float *row = sp.d_x + (size_t)(live ? s : 0) * sp.x_stride;
    float v[E];
#pragma unroll
    for (int j = 0; j < E; ++j) {
        /* unconditional loads (past the row: its first value, discarded) */
        const uint32_t i = lane + HFX_WAVE * j;
        const float t = row[i < L ? i : 0u];
        v[j] = i < L ? t : 0.0f;
    }
    uint32_t kv[E];
#pragma unroll
    for (int j = 0; j < E; ++j) kv[j] = sp_key(v[j]);
    uint32_t P = 0;
#pragma unroll
    for (int j = 0; j < E; ++j) P += __popcll(__builtin_amdgcn_ballot_w64(kv[j] != 0));
    uint32_t T = 0, need = 0;
    const bool all_pos = P <= SP_TOP;
    if (!all_pos) {
        /* T = key of the 200th largest value, bit by bit from the top */
        for (int bit = 30; bit >= 0; --bit) {
            const uint32_t cand = T | (1u << bit);
            uint32_t c = 0;
#pragma unroll
            for (int j = 0; j < E; ++j) c += __popcll(__builtin_amdgcn_ballot_w64(kv[j] >= cand));
            if (c >= SP_TOP) T = cand;
        }
        uint32_t gt = 0;
#pragma unroll
        for (int j = 0; j < E; ++j) gt += __popcll(__builtin_amdgcn_ballot_w64(kv[j] > T));
        need = SP_TOP - gt;
    }
    /* ties of T in position order: rows in order, lanes in order within */
    const uint64_t below = __lanemask_lt();
    uint32_t base = 0;
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const uint32_t k = kv[j];
        bool sel;
        if (all_pos) {
            sel = k != 0;
        } else {
            const uint64_t m = __builtin_amdgcn_ballot_w64(k == T);
            sel = k > T || (k == T && base + (uint32_t)__popcll(m & below) < need);
            base += (uint32_t)__popcll(m);
        }
        if (sel) {
            const uint32_t t = (lane + HFX_WAVE * j) / 3;
            atomicOr(&fw[t >> 5], 1u << (t & 31));
        }
    }
    __syncthreads();
    if (live)
        for (uint32_t k = lane; k < W; k += HFX_WAVE)
            on[(size_t)s * W + k] = k < SPW_WORDS ? fw[k] : 0u;
}

struct SpArgs {
    uint32_t *nb, *ny, *tot;
    const uint32_t *on;
    uint32_t W;
    float bw;
    hfx_db db;
    const uint32_t *len, *zmax;
    uint32_t stride, num_spec;
};

/* bin = round(v / bin width / 3), halves to even; matched when in range and
 * flagged (IonSeries.hs:142-147) */
__device__ __forceinline__ uint32_t sp_on(const uint32_t *f, uint32_t npk, float v, float bw)
{
    const float q = __builtin_rintf((v / bw) / 3.0f);
    if (!(q >= 0.0f && q < (float)npk)) return 0;
    const uint32_t i = (uint32_t)q;
    return (f[i >> 5] >> (i & 31)) & 1u;
}

/* lookup (Location.hs:42-58) + matchIonSequence (IonSeries.hs:136-147) +
 * scoreSP (Match.hs:33-40) of peptide p against spectrum s */
__device__ __forceinline__ void sp_match(const SpArgs &a, uint64_t m, uint32_t p, uint32_t s,
                                         uint32_t nmod)
{
    const uint32_t tc = a.db.d_tc[p], tn = a.db.d_tn[p];
    const uint32_t ntake = (uint64_t)tn < a.db.num_ions ? tn - tc + 1 : tn - tc;
    const float r = (a.db.d_r[p] + SP_H) + SP_H2O;
    const uint32_t npk = min(a.len[s], a.stride) / 3 + 1;
    const uint32_t *f = a.on + (size_t)s * a.W;
    const uint32_t nz = max(a.zmax[s], 1u);
    const uint64_t nI = (uint64_t)ntake * nz;
    uint32_t nb = 0, ny = 0;
    uint64_t j = 0;
    for (uint32_t c = 1; c <= nz; ++c) {
        const float cf = (float)c;
        const float hc = SP_H * cf;
        float pm = 0.0f;
        for (uint32_t k = 0; k < ntake; ++k, ++j) {
            pm = pm + a.db.d_mass[HFX_ACID_IDX(a.db.d_ions[tc + k])];
            const float mz = (pm + hc) / cf;
            if (j + 1 < nI) nb += sp_on(f, npk, mz, a.bw);     /* init */
            if (j >= 1) ny += sp_on(f, npk, r - mz, a.bw);     /* tail */
        }
    }
    a.nb[m] = nb;
    a.ny[m] = ny;
    a.tot[m] = (ntake - 1 + nmod) * 2;
}

__device__ __forceinline__ void sp_none(const SpArgs &a, uint64_t m)
{
    a.nb[m] = 0;
    a.ny[m] = 0;
    a.tot[m] = 0;
}

__global__ void k_sp_list(SpArgs a, const uint32_t *pep, const uint32_t *spec,
                          const uint32_t *nmod, uint64_t M)
{
    const uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const uint32_t p = pep[m], s = spec[m];
    if (p >= a.db.num_pep || s >= a.num_spec) {   /* 0xFFFFFFFF: no peptide */
        sp_none(a, m);
        return;
    }
    sp_match(a, m, p, s, nmod ? nmod[m] : 0u);
}

/* top-N of an unmodified search: entry k of spectrum s is the window
 * position top_pos[s][k] (sorted peptide begin[s] + position) */
__global__ void k_sp_topn(SpArgs a, const uint32_t *top_pos, const uint32_t *top_cnt,
                          const uint32_t *begin, uint32_t N, uint32_t S)
{
    const uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (uint64_t)N * S) return;
    const uint32_t s = (uint32_t)(m / N), k = (uint32_t)(m % N);
    if (k >= top_cnt[s]) {
        sp_none(a, m);
        return;
    }
    sp_match(a, m, a.db.d_idx_sorted[begin[s] + top_pos[m]], s, 0u);
}

/* top-N of a modified search: entry k of spectrum s is modified peptide
 * mp_off[s] + top_pos[s][k]; matched on its unmodified peptide
 * (Search.hs:296-299), the printed length grows by its modified sites */
__global__ void k_sp_topn_mod(SpArgs a, const uint32_t *top_pos, const uint32_t *top_cnt,
                              const uint64_t *mp_off, uint32_t N, uint32_t S, hfx_modset ms)
{
    const uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (uint64_t)N * S) return;
    const uint32_t s = (uint32_t)(m / N), k = (uint32_t)(m % N);
    if (k >= top_cnt[s]) {
        sp_none(a, m);
        return;
    }
    const uint64_t i = mp_off[s] + top_pos[m];
    const uint32_t md = ms.d_mpep_mod_idx[i];
    uint32_t nmod = 0;
    for (uint32_t q = 0; q < ms.num_ma; ++q) nmod += ms.d_mod_ma_count[(size_t)md * ms.num_ma + q];
    sp_match(a, m, ms.d_mpep_pep_idx[i], s, nmod);
}

/* ------------------------------------------------------------------ */
/* Host                                                                  */
/* ------------------------------------------------------------------ */

extern "C" int hfx_evalue_fit(hfx_ctx *ctx, double *d_fit, uint32_t *d_npts, uint64_t *d_nused,
                              const float *d_scores_a, const uint64_t *d_off_a,
                              const float *d_scores_b, const uint64_t *d_off_b,
                              uint32_t num_spec, uint32_t bin_width, float cutoff_perc,
                              void *stream)
{
    (void)ctx;
    HFX_REQUIRE(d_fit && d_off_a && d_scores_a, HFX_ERR_ARG, "hfx_evalue_fit: null argument");
    HFX_REQUIRE(!d_scores_b == !d_off_b, HFX_ERR_ARG,
                "hfx_evalue_fit: second score set needs both scores and offsets");
    HFX_REQUIRE(bin_width >= 1, HFX_ERR_ARG, "hfx_evalue_fit: bin width 0");
    HFX_REQUIRE(cutoff_perc >= 0.0f && cutoff_perc <= 100.0f, HFX_ERR_ARG,
                "hfx_evalue_fit: cut-off percentage %g outside [0, 100]", (double)cutoff_perc);
    if (num_spec == 0) return HFX_OK;
    EvArgs a{d_fit, d_npts, d_nused, d_scores_a, d_off_a, d_scores_b, d_off_b,
             (double)bin_width, cutoff_perc};
    hipLaunchKernelGGL(k_evfit, dim3(num_spec), dim3(ST_THREADS), 0, hfx_stream(stream), a);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

extern "C" int hfx_evalues(hfx_ctx *ctx, double *d_ev, const double *d_fit,
                           const double *d_score, const uint32_t *d_cnt, uint32_t N,
                           uint32_t num_spec, void *stream)
{
    (void)ctx;
    HFX_REQUIRE(d_ev && d_fit && d_score, HFX_ERR_ARG, "hfx_evalues: null argument");
    const uint64_t M = (uint64_t)N * num_spec;
    if (M == 0) return HFX_OK;
    hipLaunchKernelGGL(k_evapply, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                       hfx_stream(stream), d_ev, d_fit, d_score, d_cnt, N, num_spec);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

extern "C" uint32_t hfx_sp_peak_words(uint32_t x_stride) { return (x_stride / 3 + 1 + 31) / 32; }

extern "C" int hfx_sp_peaks(hfx_ctx *ctx, uint32_t *d_on, uint32_t words, const hfx_spectra *sp,
                            void *stream)
{
    (void)ctx;
    HFX_REQUIRE(d_on && sp, HFX_ERR_ARG, "hfx_sp_peaks: null argument");
    HFX_REQUIRE(words >= hfx_sp_peak_words(sp->x_stride), HFX_ERR_ARG,
                "hfx_sp_peaks: %u flag words < %u for stride %u", words,
                hfx_sp_peak_words(sp->x_stride), sp->x_stride);
    HFX_REQUIRE(hfx_sp_peak_words(sp->x_stride) <= SP_WORDS, HFX_ERR_ARG,
                "hfx_sp_peaks: spectrum stride %u above %u", sp->x_stride,
                3u * (32u * SP_WORDS - 1u) + 2u);
    if (sp->num_spec == 0) return HFX_OK;
    const dim3 gw((sp->num_spec + ST_WAVES - 1) / ST_WAVES);
    const uint32_t rows = (sp->x_stride + HFX_WAVE - 1) / HFX_WAVE;
    if (rows <= 12)
        hipLaunchKernelGGL(k_sp_peaks_w<12>, gw, dim3(ST_THREADS), 0, hfx_stream(stream), d_on,
                           words, *sp);
    else if (rows <= 24)
        hipLaunchKernelGGL(k_sp_peaks_w<24>, gw, dim3(ST_THREADS), 0, hfx_stream(stream), d_on,
                           words, *sp);
    else if (rows <= 36)
        hipLaunchKernelGGL(k_sp_peaks_w<36>, gw, dim3(ST_THREADS), 0, hfx_stream(stream), d_on,
                           words, *sp);
    else if (rows <= SPW_MAXE)
        hipLaunchKernelGGL(k_sp_peaks_w<SPW_MAXE>, gw, dim3(ST_THREADS), 0, hfx_stream(stream),
                           d_on, words, *sp);
    else
        hipLaunchKernelGGL(k_sp_peaks, dim3(sp->num_spec), dim3(ST_THREADS), 0,
                           hfx_stream(stream), d_on, words, *sp);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

static int sp_args(SpArgs &a, uint32_t *d_nb, uint32_t *d_ny, uint32_t *d_total,
                   const uint32_t *d_on, uint32_t words, const hfx_db *db, const hfx_spectra *sp,
                   int mono)
{
    HFX_REQUIRE(d_nb && d_ny && d_total && d_on && db && sp, HFX_ERR_ARG,
                "hfx_sp_ions: null argument");
    HFX_REQUIRE(words >= hfx_sp_peak_words(sp->x_stride), HFX_ERR_ARG,
                "hfx_sp_ions: %u flag words < %u", words, hfx_sp_peak_words(sp->x_stride));
    a.nb = d_nb;
    a.ny = d_ny;
    a.tot = d_total;
    a.on = d_on;
    a.W = words;
    a.bw = mono ? HFX_BIN_WIDTH_MONO : HFX_BIN_WIDTH_AVG; /* Correlation.hs:29-30 */
    a.db = *db;
    a.len = sp->d_len;
    a.zmax = sp->d_zmax;
    a.stride = sp->x_stride;
    a.num_spec = sp->num_spec;
    return HFX_OK;
}

extern "C" int hfx_sp_ions(hfx_ctx *ctx, uint32_t *d_nb, uint32_t *d_ny, uint32_t *d_total,
                           const uint32_t *d_pep, const uint32_t *d_spec, const uint32_t *d_nmod,
                           uint64_t num_match, const uint32_t *d_on, uint32_t words,
                           const hfx_db *db, const hfx_spectra *sp, int mono, void *stream)
{
    (void)ctx;
    SpArgs a;
    int rc = sp_args(a, d_nb, d_ny, d_total, d_on, words, db, sp, mono);
    if (rc) return rc;
    HFX_REQUIRE(d_pep && d_spec, HFX_ERR_ARG, "hfx_sp_ions: null match list");
    if (num_match == 0) return HFX_OK;
    hipLaunchKernelGGL(k_sp_list, dim3((unsigned)((num_match + 255) / 256)), dim3(256), 0,
                       hfx_stream(stream), a, d_pep, d_spec, d_nmod, num_match);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

static int sp_topn_flags(hfx_ctx *ctx, const hfx_spectra *sp, uint32_t **on, uint32_t *W,
                         void *stream)
{
    HFX_REQUIRE(ctx && sp, HFX_ERR_ARG, "hfx_sp_ions_topn: null argument");
    *W = hfx_sp_peak_words(sp->x_stride);
    void *p = nullptr;
    int rc = hfx_ctx_scratch(ctx, HFX_SLOT_STATS, (size_t)4 * *W * (sp->num_spec ? sp->num_spec : 1), &p);
    if (rc) return rc;
    *on = (uint32_t *)p;
    return hfx_sp_peaks(ctx, *on, *W, sp, stream);
}

extern "C" int hfx_sp_ions_topn(hfx_ctx *ctx, uint32_t *d_nb, uint32_t *d_ny, uint32_t *d_total,
                                const uint32_t *d_top_pos, const uint32_t *d_top_cnt,
                                const uint32_t *d_begin, uint32_t N, const hfx_db *db,
                                const hfx_spectra *sp, int mono, void *stream)
{
    HFX_REQUIRE(d_top_pos && d_top_cnt && d_begin && db && sp, HFX_ERR_ARG,
                "hfx_sp_ions_topn: null argument");
    uint32_t *on = nullptr, W = 0;
    int rc = sp_topn_flags(ctx, sp, &on, &W, stream);
    if (rc) return rc;
    SpArgs a;
    rc = sp_args(a, d_nb, d_ny, d_total, on, W, db, sp, mono);
    if (rc) return rc;
    const uint64_t M = (uint64_t)N * sp->num_spec;
    if (M == 0) return HFX_OK;
    hipLaunchKernelGGL(k_sp_topn, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                       hfx_stream(stream), a, d_top_pos, d_top_cnt, d_begin, N, sp->num_spec);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

extern "C" int hfx_sp_ions_topn_mod(hfx_ctx *ctx, uint32_t *d_nb, uint32_t *d_ny,
                                    uint32_t *d_total, const uint32_t *d_top_pos,
                                    const uint32_t *d_top_cnt, const uint64_t *d_mp_off,
                                    uint32_t N, const hfx_modset *ms, const hfx_db *db,
                                    const hfx_spectra *sp, int mono, void *stream)
{
    HFX_REQUIRE(d_top_pos && d_top_cnt && d_mp_off && ms && db && sp, HFX_ERR_ARG,
                "hfx_sp_ions_topn_mod: null argument");
    uint32_t *on = nullptr, W = 0;
    int rc = sp_topn_flags(ctx, sp, &on, &W, stream);
    if (rc) return rc;
    SpArgs a;
    rc = sp_args(a, d_nb, d_ny, d_total, on, W, db, sp, mono);
    if (rc) return rc;
    const uint64_t M = (uint64_t)N * sp->num_spec;
    if (M == 0) return HFX_OK;
    hipLaunchKernelGGL(k_sp_topn_mod, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                       hfx_stream(stream), a, d_top_pos, d_top_cnt, d_mp_off, N, sp->num_spec,
                       *ms);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}
