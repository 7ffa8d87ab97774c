/*
 * hfx_batch.hip -- batched search: peptide ion index + persistent
 * work-stealing multi-spectrum XCorr, hipGraph capture, autotuner.
 *
 * The reference scores each spectrum on its own: every (spectrum, candidate)
 * pair rebuilds the candidate's theoretical spectrum (Search.hs:177-245).  A
 * candidate's peaks depend only on the peptide and the fragment charges, and
 * neighbouring precursors share most of their window, so a batch of thousands
 * of spectra here
 *   1. builds, once per batch, the sparse theoretical spectrum of every
 *      peptide any window touches, for fragment charge levels 1..L (level l =
 *      charges 1..l, max per bin): sorted uint16 entries bin | code << 14
 *      (k_ix_build, one wave per peptide, from the same ladder/bin code as the
 *      plain path);
 *   2. sorts the spectra by precursor mass and groups G consecutive spectra
 *      into blocks; a tile is (block, C consecutive candidates of the block's
 *      union window); a persistent kernel pulls tiles from an atomic counter
 *      (work stealing: heavy-tailed windows are cut into equal tiles), keeps
 *      the block's G spectra in LDS, streams each candidate's entries once
 *      (coalesced) and gathers every spectrum of the block that holds the
 *      candidate in its window, reducing the G partial sums with one
 *      reduce-scatter butterfly;
 *   3. selects the top N with the exact re-score of hfx_topn (flagged spectra
 *      completed in the same launch sequence, no host round trip).
 * Scores land in the same layout as hfx_xcorr_sparse, so every result is
 * checked against the same oracle.  Once planned, a batch runs without host
 * synchronisation or allocation and can be captured in a hipGraph.
 */
#include "hfx_cand.h"

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <vector>

#define IX_THREADS 256
#define IX_WAVES (IX_THREADS / HFX_WAVE)
#define BT_GMAX 8

/* ------------------------------------------------------------------ */
/* Ion index                                                             */
/* ------------------------------------------------------------------ */

/* capacity of position q at level l: every emitted peak fills at most one
 * bin, 10 peaks per ladder position and charge, and no more than lcap bins */
__global__ void k_ix_caps(uint32_t *caps, hfx_db db, uint32_t p0, uint32_t P, uint32_t L,
                          uint32_t lcap)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P) return;
    const uint32_t pep = db.d_idx_sorted[p0 + q];
    const uint32_t tc = db.d_tc[pep], tn = db.d_tn[pep];
    const uint32_t n = tn > tc ? tn - tc : 0;
    for (uint32_t l = 1; l <= L; ++l) caps[(size_t)q * L + l - 1] = min(10u * n * l, lcap);
}

/* compact the map (bins ascending) into entries; returns the count */
__device__ __forceinline__ uint32_t wave_compact(const uint32_t *map, uint32_t nw, uint16_t *out)
{
    const uint32_t lane = hfx_lane();
    uint32_t base = 0;
    for (uint32_t w0 = 0; w0 < nw; w0 += HFX_WAVE) {
        const uint32_t w = w0 + lane;
        const uint32_t word = w < nw ? map[w] : 0u;
        const uint32_t nz = (word | (word >> 1) | (word >> 2) | (word >> 3)) & 0x11111111u;
        const uint32_t c = (uint32_t)__builtin_popcount(nz);
        uint32_t incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= (uint32_t)o) incl += y;
        }
        uint32_t pos = base + incl - c;
        uint32_t wd = word;
        while (wd) {
            const uint32_t sh = __builtin_ctz(wd) & ~3u;
            const uint32_t nib = (wd >> sh) & 0xFu;
            wd &= ~(0xFu << sh);
            out[pos++] = (uint16_t)(((w << 3) + (sh >> 2)) | ((uint32_t)__builtin_popcount(nib) << 14));
        }
        base += (uint32_t)__shfl(incl, 63, 64);
    }
    return base;
}

__global__ void __launch_bounds__(IX_THREADS)
    k_ix_build(uint16_t *ent, uint32_t *cnt, const uint64_t *off, hfx_db db, uint32_t p0, uint32_t P,
               uint32_t L, uint32_t lcap)
{
    __shared__ uint32_t maps[IX_WAVES][HFX_MAP_WORDS];
    __shared__ float mtab[32];
    const uint32_t wave = threadIdx.x / HFX_WAVE, lane = hfx_lane();
    uint32_t *map = maps[wave];
    if (threadIdx.x < 32) mtab[threadIdx.x] = threadIdx.x < 26 ? db.d_mass[threadIdx.x] : 0.0f;
    for (uint32_t w = lane; w < HFX_MAP_WORDS; w += HFX_WAVE) map[w] = 0;
    __syncthreads();
    const uint32_t nw = (lcap + 7) >> 3;
    for (uint32_t q = blockIdx.x * IX_WAVES + wave; q < P; q += gridDim.x * IX_WAVES) {
        const uint32_t pep = db.d_idx_sorted[p0 + q];
        const uint32_t tc = db.d_tc[pep], tn = db.d_tn[pep];
        const uint32_t n = tn > tc ? tn - tc : 0;
        const float r = db.d_r[pep];
        for (uint32_t l = 1; l <= L; ++l) {
            /* level l = charges 1..l: add charge l to the map of level l-1 */
            hfx_wave_peaks_z(map, n, r, l, l, lcap, HfxPlainMass{db.d_ions + tc, mtab});
            hfx_wave_sync();
            const size_t k = (size_t)q * L + l - 1;
            const uint32_t c = wave_compact(map, nw, ent + off[k]);
            if (lane == 0) cnt[k] = c;
            hfx_wave_sync();
        }
        for (uint32_t w = lane; w < nw; w += HFX_WAVE) map[w] = 0;
        hfx_wave_sync();
    }
}

/* ------------------------------------------------------------------ */
/* Blocks, tiles                                                          */
/* ------------------------------------------------------------------ */

__global__ void k_bt_blocks(uint32_t *ubeg, uint32_t *uend, uint32_t *ntiles, const uint32_t *order,
                            const uint32_t *begin, const uint32_t *num, uint32_t S, uint32_t G,
                            uint32_t B, uint32_t C)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    uint32_t lo = 0xffffffffu, hi = 0;
    for (uint32_t g = 0; g < G; ++g) {
        const uint32_t i = b * G + g;
        if (i >= S) break;
        const uint32_t s = order[i];
        if (num[s] == 0) continue;
        lo = min(lo, begin[s]);
        hi = max(hi, begin[s] + num[s]);
    }
    if (hi <= lo) { lo = 0; hi = 0; }
    ubeg[b] = lo;
    uend[b] = hi;
    ntiles[b] = (hi - lo + C - 1) / C;
}

__global__ void k_bt_reset(uint32_t *counter) { *counter = 0; }

/* zmax > 4: the reference launches nothing and the zero rows score 0 */
__global__ void k_bt_zero_unscored(float *scores, const uint64_t *off, const uint32_t *zmax,
                                   uint32_t S)
{
    const uint32_t s = blockIdx.x;
    if (s >= S || hfx_zeff(zmax[s]) != 0) return;
    for (uint64_t i = off[s] + threadIdx.x; i < off[s + 1]; i += blockDim.x) scores[i] = 0.0f;
}

struct BtArgs {
    float *scores;
    const uint64_t *score_off;
    const uint32_t *begin, *num, *order;
    uint32_t S, G, B, C;
    const uint32_t *ubeg, *uend;
    const uint64_t *tile_off;
    const uint16_t *ent;
    const uint32_t *cnt;
    const uint64_t *ent_off;
    uint32_t p0, P, L, lstr;
    uint32_t *counter;
    hfx_spectra sp;
};

__device__ __forceinline__ float code_w(uint32_t c)
{
    return c == 3 ? 50.0f : (c == 2 ? 25.0f : (c == 1 ? 10.0f : 0.0f));
}

/* Sum each of the G accumulators over the wave: a reduce-scatter butterfly
 * (halving the accumulators per lane at each of the first log2 G steps) then
 * a plain butterfly.  Lane l ends with the sum of accumulator
 * g = bit-reverse-free index ((l >> 5) & 1) * 4 + ((l >> 4) & 1) * 2 + ((l >> 3) & 1)
 * for G = 8 (first bits only for smaller G). */
template <int G>
__device__ __forceinline__ float reduce_scatter(float (&acc)[G], uint32_t lane)
{
    static_assert(G == 4 || G == 8, "G must be 4 or 8");
    float v[G];
#pragma unroll
    for (int k = 0; k < G; ++k) v[k] = acc[k];
    /* each step pairs lanes across `mask`: the lower lane keeps the lower half
     * of its live accumulators, the upper lane the upper half, each adding
     * the partner's copy of the accumulators it keeps */
    int mask = 32;
#pragma unroll
    for (int n = G; n > 1; n >>= 1, mask >>= 1) {
        const bool up = lane & (uint32_t)mask;
#pragma unroll
        for (int k = 0; k < n / 2; ++k) {
            const float send = up ? v[k] : v[k + n / 2];
            const float keep = up ? v[k + n / 2] : v[k];
            v[k] = keep + __shfl_xor(send, mask, 64);
        }
    }
    float s = v[0];
#pragma unroll
    for (; mask > 0; mask >>= 1) s += __shfl_xor(s, mask, 64);
    return s;
}

template <int G>
__device__ __forceinline__ uint32_t owner_of(uint32_t lane)
{
    if constexpr (G == 8) return ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
    else return ((lane >> 5) & 1) * 2 + ((lane >> 4) & 1);
}

template <int G>
__global__ void __launch_bounds__(512) k_batch(BtArgs a)
{
    extern __shared__ float bx[]; /* G spectra, stride lstr */
    __shared__ uint32_t m_s[BT_GMAX], m_beg[BT_GMAX], m_end[BT_GMAX], m_z[BT_GMAX], m_len[BT_GMAX];
    __shared__ uint64_t m_off[BT_GMAX];
    __shared__ uint32_t sh_t, sh_blk, sh_lblk;
    const uint32_t tid = threadIdx.x, lane = hfx_lane();
    const uint32_t nwaves = blockDim.x / HFX_WAVE, wave = tid / HFX_WAVE;
    const uint64_t ntiles = a.tile_off[a.B];
    uint32_t cur = 0xffffffffu;
    for (;;) {
        if (tid == 0) sh_t = atomicAdd(a.counter, 1u);
        __syncthreads();
        const uint64_t t0 = sh_t;
        if (t0 >= ntiles) break;
        /* tiles from the last to the first: blocks follow the plan's order by
         * charge level, then precursor mass, so the heaviest come first and
         * the lightest finish the launch */
        const uint64_t t = ntiles - 1u - t0;
        /* block of the tile: largest b with tile_off[b] <= t */
        uint32_t lo = 0, hi = a.B;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.tile_off[mid] <= t) lo = mid; else hi = mid;
        }
        const uint32_t b = lo;
        if (b != cur) {
            /* stage the block's spectra (the previous tile's readers are done:
             * the barrier above follows every wave's last read) */
            if (tid < BT_GMAX) {
                const uint32_t i = b * a.G + tid;
                uint32_t s = 0xffffffffu, be = 0, en = 0, z = 0, ln = 0;
                uint64_t of = 0;
                if (tid < a.G && i < a.S) {
                    s = a.order[i];
                    be = a.begin[s];
                    en = be + a.num[s];
                    z = hfx_zeff(a.sp.d_zmax[s]);
                    ln = min(a.sp.d_len[s], min(a.lstr, a.sp.x_stride));
                    of = a.score_off[s];
                }
                m_s[tid] = s; m_beg[tid] = be; m_end[tid] = en; m_z[tid] = z; m_len[tid] = ln;
                m_off[tid] = of;
            }
            __syncthreads();
            if (tid == 0) {
                uint32_t lb = 0;
                for (int g = 0; g < G; ++g) lb = max(lb, m_z[g]);
                sh_lblk = min(lb, a.L);
            }
            for (int g = 0; g < G; ++g) {
                const uint32_t s = m_s[g];
                if (s == 0xffffffffu) continue;
                const float *xg = a.sp.d_x + (size_t)s * a.sp.x_stride;
                for (uint32_t i = tid; i < m_len[g]; i += blockDim.x) bx[g * a.lstr + i] = xg[i];
            }
            cur = b;
            __syncthreads();
        }
        const uint32_t lblk = sh_lblk;
        const uint32_t chunk = (uint32_t)(t - a.tile_off[b]);
        const uint32_t pa = a.ubeg[b] + chunk * a.C;
        const uint32_t pb = min(pa + a.C, a.uend[b]);
        for (uint32_t p = pa + wave; p < pb; p += nwaves) {
            const uint32_t q = p - a.p0;
            for (uint32_t l = 1; l <= lblk; ++l) {
                bool act[G];
                bool any = false;
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    act[g] = m_z[g] == l && m_beg[g] <= p && p < m_end[g];
                    any |= act[g];
                }
                if (!any) continue;
                const size_t k = (size_t)q * a.L + l - 1;
                const uint16_t *e = a.ent + a.ent_off[k];
                const uint32_t ne = a.cnt[k];
                float acc[G];
#pragma unroll
                for (int g = 0; g < G; ++g) acc[g] = 0.0f;
                for (uint32_t j = lane; j < ne; j += HFX_WAVE) {
                    const uint32_t v = e[j];
                    const uint32_t bin = v & 0x3fffu;
                    const float w = code_w(v >> 14);
#pragma unroll
                    for (int g = 0; g < G; ++g)
                        if (act[g] && bin < m_len[g]) acc[g] = acc[g] + w * bx[g * a.lstr + bin];
                }
                const float sum = reduce_scatter<G>(acc, lane);
                const uint32_t g = owner_of<G>(lane);
                if ((lane & 7) == 0 && act[g])
                    a.scores[m_off[g] + (p - m_beg[g])] = sum;
            }
        }
        __syncthreads();
    }
}

/* ------------------------------------------------------------------ */
/* The batch object                                                       */
/* ------------------------------------------------------------------ */

struct hfx_batch {
    hfx_ctx *ctx;
    hfx_db db;
    hfx_spectra sp;
    float eps;
    uint32_t N, margin;
    uint32_t S, G, C, threads, B, L, lstr, lcap, p0, P;
    uint64_t total, n_ent, tiles;
    bool planned;
    /* device buffers (owned) */
    uint32_t *begin, *num, *order, *ubeg, *uend, *ntiles, *caps, *cnt, *counter;
    uint64_t *score_off, *tile_off, *ent_off;
    uint16_t *ent;
    float *scores;
    uint32_t *top_pos, *top_cnt, *flag;
    double *top_exact;
    float *top_f32;
    uint8_t *tx_scratch; /* top-N exact route */
    uint32_t precision;  /* HFX_BATCH_FP32_SPARSE or HFX_BATCH_BF16X2_DENSE */
    uint32_t CH;         /* dense: candidates per work item */
    uint8_t *dn_scratch; /* dense: hi/lo spectra, blocks (null if unsupported) */
    void *dn_mem[2];     /* dense: the ladders of the plan's windows */
    hfx_dn_ladder dn_lad;
    uint32_t *dn_bstart; /* dense: [DB + 1] block starts in order */
    uint32_t DB;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

static void bt_free(hfx_batch *b)
{
    void *ps[] = {b->begin,    b->num,     b->order,  b->ubeg,      b->uend,    b->ntiles,
                  b->caps,     b->cnt,     b->counter, b->score_off, b->tile_off, b->ent_off,
                  b->ent,      b->scores,  b->top_pos, b->top_cnt,   b->flag,     b->top_exact,
                  b->top_f32,  b->tx_scratch, b->dn_scratch, b->dn_mem[0], b->dn_mem[1],
                  b->dn_bstart};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    hfx_ctx *ctx = b->ctx;
    memset(b, 0, sizeof *b);
    b->ctx = ctx;
}

extern "C" int hfx_batch_create(hfx_ctx *ctx, hfx_batch **out)
{
    HFX_REQUIRE(ctx && out, HFX_ERR_ARG, "hfx_batch_create: null argument");
    hfx_batch *b = (hfx_batch *)calloc(1, sizeof(hfx_batch));
    HFX_REQUIRE(b, HFX_ERR_NOMEM, "hfx_batch_create: out of host memory");
    b->ctx = ctx;
    *out = b;
    return HFX_OK;
}

extern "C" int hfx_batch_destroy(hfx_batch *b)
{
    if (!b) return HFX_OK;
    (void)hipDeviceSynchronize();
    bt_free(b);
    free(b);
    return HFX_OK;
}

template <class T>
static int dalloc(T **p, size_t n)
{
    hipError_t e = hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        hfx_set_error("batch allocation of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
        return HFX_ERR_NOMEM;
    }
    return HFX_OK;
}

/* Shape-bucketed autotuner results of this process: (S bucket, P bucket, L)
 * -> (G, C, threads). */
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, uint32_t>, std::tuple<uint32_t, uint32_t, uint32_t>> g_tune;
static std::map<std::tuple<int, int, uint32_t>, uint32_t> g_tune_dense;

static std::tuple<int, int, uint32_t> tune_key(const hfx_batch *b)
{
    auto lg = [](uint64_t v) { int k = 0; while (v > 1) { v >>= 1; ++k; } return k; };
    return std::make_tuple(lg(b->S), lg(b->P), b->L);
}

extern "C" int hfx_batch_plan(hfx_batch *b, const hfx_db *db, const hfx_spectra *sp, float eps,
                              uint32_t N, uint32_t margin, void *stream)
{
    HFX_REQUIRE(b && db && sp, HFX_ERR_ARG, "hfx_batch_plan: null argument");
    HFX_REQUIRE(N >= 1 && N + margin <= 128 && N <= 64, HFX_ERR_ARG,
                "hfx_batch_plan: N + margin must be <= 128 and N <= 64");
    hipStream_t st = hfx_stream(stream);
    bt_free(b);
    b->db = *db;
    b->sp = *sp;
    b->eps = eps;
    b->N = N;
    b->margin = margin;
    const uint32_t S = b->S = sp->num_spec;
    int rc;
    if ((rc = dalloc(&b->begin, S)) || (rc = dalloc(&b->num, S)) ||
        (rc = dalloc(&b->score_off, (size_t)S + 1)) || (rc = dalloc(&b->order, S)))
        return rc;
    if ((rc = hfx_window(b->ctx, b->begin, b->num, db, sp->d_mass, S, eps, stream))) return rc;
    if ((rc = hfx_scan_offsets(b->ctx, b->score_off, b->num, S, stream))) return rc;
    /* host view of the windows, masses and charges (plan only) */
    std::vector<uint32_t> hb(S), hn(S), hz(S), hl(S);
    std::vector<float> hm(S);
    HFX_CHECK_HIP(hipMemcpyAsync(hb.data(), b->begin, 4ull * S, hipMemcpyDeviceToHost, st));
    HFX_CHECK_HIP(hipMemcpyAsync(hn.data(), b->num, 4ull * S, hipMemcpyDeviceToHost, st));
    HFX_CHECK_HIP(hipMemcpyAsync(hz.data(), sp->d_zmax, 4ull * S, hipMemcpyDeviceToHost, st));
    HFX_CHECK_HIP(hipMemcpyAsync(hl.data(), sp->d_len, 4ull * S, hipMemcpyDeviceToHost, st));
    HFX_CHECK_HIP(hipMemcpyAsync(hm.data(), sp->d_mass, 4ull * S, hipMemcpyDeviceToHost, st));
    HFX_CHECK_HIP(hipMemcpyAsync(&b->total, b->score_off + S, 8, hipMemcpyDeviceToHost, st));
    HFX_CHECK_HIP(hipStreamSynchronize(st));
    std::vector<uint32_t> ord(S);
    std::iota(ord.begin(), ord.end(), 0u);
    /* by fragment charge level, then precursor mass: blocks of neighbouring
     * spectra share one ladder level and most of their windows */
    auto zlev = [&](uint32_t s) { return hz[s] <= HFX_MAX_CHARGE ? hz[s] : 0u; };
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) {
        const uint32_t zx = zlev(x), zy = zlev(y);
        return zx != zy ? zx < zy : hm[x] < hm[y];
    });
    HFX_CHECK_HIP(hipMemcpyAsync(b->order, ord.data(), 4ull * S, hipMemcpyHostToDevice, st));
    /* dense blocks: the order cut into runs of at most eight spectra that
     * minimise the sum over blocks of (charge levels x rows of the hull of
     * the windows) -- a block spanning a gap in precursor mass would score
     * every candidate in the gap.  A per-block constant (B staging, tile
     * rounding) measured no better than none at tryptic-20k-10k. */
    std::vector<uint32_t> bcut;
    {
        constexpr uint64_t FIXED = 0;
        std::vector<uint64_t> best(S + 1, UINT64_MAX);
        std::vector<uint32_t> from(S + 1, 0);
        best[0] = 0;
        for (uint32_t j = 1; j <= S; ++j) {
            uint32_t lo = 0xffffffffu, hi = 0, lv = 0;
            for (uint32_t k = 1; k <= 8 && k <= j; ++k) {
                const uint32_t x = ord[j - k], z = zlev(x);
                if (z && hn[x]) {
                    lo = std::min(lo, hb[x]);
                    hi = std::max(hi, hb[x] + hn[x]);
                    lv |= 1u << z;
                }
                const uint64_t rows = hi > lo ? hi - lo : 0;
                const uint64_t c = best[j - k] + rows * (uint64_t)__builtin_popcount(lv) + FIXED;
                if (c < best[j]) {
                    best[j] = c;
                    from[j] = j - k;
                }
            }
        }
        for (uint32_t j = S; j > 0; j = from[j]) bcut.push_back(from[j]);
        std::reverse(bcut.begin(), bcut.end());
        bcut.push_back(S);
    }
    b->DB = (uint32_t)bcut.size() - 1;
    if ((rc = dalloc(&b->dn_bstart, bcut.size()))) return rc;
    HFX_CHECK_HIP(hipMemcpyAsync(b->dn_bstart, bcut.data(), 4ull * bcut.size(),
                                 hipMemcpyHostToDevice, st));
    HFX_CHECK_HIP(hipStreamSynchronize(st));
    uint32_t p0 = 0xffffffffu, p1 = 0, L = 0, lmax = 0;
    for (uint32_t s = 0; s < S; ++s) {
        lmax = std::max(lmax, std::min(hl[s], std::min(sp->x_stride, (uint32_t)HFX_MAX_LEN)));
        if (hn[s] == 0) continue;
        p0 = std::min(p0, hb[s]);
        p1 = std::max(p1, hb[s] + hn[s]);
        uint32_t z = hz[s] <= HFX_MAX_CHARGE ? hz[s] : 0u;
        L = std::max(L, z);
    }
    if (p1 <= p0) { p0 = 0; p1 = 0; }
    b->p0 = p0;
    b->P = p1 - p0;
    b->L = std::max(L, 1u);
    b->lcap = std::max(lmax, 8u);
    b->lstr = (b->lcap + 15) & ~15u;
    /* defaults, replaced by the autotuner's choice for this shape */
    b->G = 8;
    b->C = 64;
    b->threads = 512;
    b->CH = 256;
    b->precision = HFX_BATCH_FP32_SPARSE;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(tune_key(b));
        if (it != g_tune.end()) std::tie(b->G, b->C, b->threads) = it->second;
        auto jt = g_tune_dense.find(tune_key(b));
        if (jt != g_tune_dense.end()) b->CH = jt->second;
    }
    const uint32_t P = b->P, Lv = b->L;
    /* index capacities and offsets */
    if ((rc = dalloc(&b->caps, (size_t)P * Lv)) || (rc = dalloc(&b->ent_off, (size_t)P * Lv + 1)) ||
        (rc = dalloc(&b->cnt, (size_t)P * Lv)))
        return rc;
    if (P) {
        hipLaunchKernelGGL(k_ix_caps, dim3((P + 255) / 256), dim3(256), 0, st, b->caps, *db, p0, P,
                           Lv, b->lcap);
        HFX_REQUIRE((uint64_t)P * Lv < 0xffffffffull, HFX_ERR_ARG, "index too large");
        if ((rc = hfx_scan_offsets(b->ctx, b->ent_off, b->caps, P * Lv, stream))) return rc;
        HFX_CHECK_HIP(hipMemcpyAsync(&b->n_ent, b->ent_off + (size_t)P * Lv, 8,
                                     hipMemcpyDeviceToHost, st));
        HFX_CHECK_HIP(hipStreamSynchronize(st));
    }
    if ((rc = dalloc(&b->ent, b->n_ent))) return rc;
    if ((rc = dalloc(&b->scores, b->total)) || (rc = dalloc(&b->counter, 1)) ||
        (rc = dalloc(&b->top_pos, (size_t)S * N)) || (rc = dalloc(&b->top_exact, (size_t)S * N)) ||
        (rc = dalloc(&b->top_f32, (size_t)S * N)) || (rc = dalloc(&b->top_cnt, S)) ||
        (rc = dalloc(&b->flag, S)))
        return rc;
    /* the top-N exact route's scratch is owned by the batch, so that a
     * captured graph never allocates and never sees it move */
    if ((rc = dalloc(&b->tx_scratch, hfx_topn_scratch_bytes(S, b->total)))) return rc;
    if (sp->x_stride <= hfx_dense_max_len()) {
        if ((rc = dalloc(&b->dn_scratch, hfx_dense_scratch_bytes(S)))) return rc;
        /* the spectra's hi/lo BF16 rows and the b ladders of every candidate
         * of the plan, built once */
        if ((rc = hfx_dense_split(sp, b->dn_scratch, stream))) return rc;
        if ((rc = hfx_dense_ladder_build(b->ctx, db, b->p0, b->P, 1, &b->dn_lad, b->dn_mem,
                                         stream)))
            return rc;
    }
    /* blocks are sized for the smallest G the tuner may pick (4) */
    const uint32_t bmax = (S + 3) / 4;
    if ((rc = dalloc(&b->ubeg, bmax)) || (rc = dalloc(&b->uend, bmax)) ||
        (rc = dalloc(&b->ntiles, bmax)) || (rc = dalloc(&b->tile_off, (size_t)bmax + 1)))
        return rc;
    b->planned = true;
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

/* blocks + tiles for the current (G, C) */
static int bt_tiles(hfx_batch *b, hipStream_t st)
{
    b->B = (b->S + b->G - 1) / b->G;
    if (b->B == 0) return HFX_OK;
    hipLaunchKernelGGL(k_bt_blocks, dim3((b->B + 255) / 256), dim3(256), 0, st, b->ubeg, b->uend,
                       b->ntiles, b->order, b->begin, b->num, b->S, b->G, b->B, b->C);
    return hfx_scan_offsets(b->ctx, b->tile_off, b->ntiles, b->B, (void *)st);
}

static int bt_score(hfx_batch *b, hipStream_t st)
{
    if (b->B == 0 || b->total == 0) return HFX_OK;
    hipLaunchKernelGGL(k_bt_reset, dim3(1), dim3(1), 0, st, b->counter);
    BtArgs a{b->scores, b->score_off, b->begin, b->num,  b->order,   b->S,     b->G,
             b->B,      b->C,         b->ubeg,  b->uend, b->tile_off, b->ent,  b->cnt,
             b->ent_off, b->p0,       b->P,     b->L,    b->lstr,    b->counter, b->sp};
    const size_t smem = (size_t)b->G * b->lstr * 4;
    const int waves_per_wg = (int)(b->threads / HFX_WAVE);
    (void)waves_per_wg;
    const uint32_t grid = (uint32_t)hfx_ctx_num_cus(b->ctx) * (b->G == 8 ? 2u : 4u) *
                          (b->threads == 256 ? 2u : 1u);
    /* gfx950 gives a workgroup up to 160 KiB of LDS; ask for what this
     * configuration stages (G spectra of lstr floats) */
    if (b->G == 8) {
        HFX_CHECK_HIP(hipFuncSetAttribute((const void *)k_batch<8>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        hipLaunchKernelGGL(k_batch<8>, dim3(grid), dim3(b->threads), smem, st, a);
    } else {
        HFX_CHECK_HIP(hipFuncSetAttribute((const void *)k_batch<4>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        hipLaunchKernelGGL(k_batch<4>, dim3(grid), dim3(b->threads), smem, st, a);
    }
    hipLaunchKernelGGL(k_bt_zero_unscored, dim3(b->S), dim3(256), 0, st, b->scores, b->score_off,
                       b->sp.d_zmax, b->S);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

/* the phases of one batched search */
static int bt_windows(hfx_batch *b, hipStream_t st)
{
    int rc;
    if ((rc = hfx_window(b->ctx, b->begin, b->num, &b->db, b->sp.d_mass, b->S, b->eps, (void *)st)))
        return rc;
    return hfx_scan_offsets(b->ctx, b->score_off, b->num, b->S, (void *)st);
}

static int bt_index(hfx_batch *b, hipStream_t st)
{
    if (b->precision != HFX_BATCH_FP32_SPARSE || b->P == 0) return HFX_OK;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((b->P + IX_WAVES - 1) / IX_WAVES,
                                                       (uint64_t)hfx_ctx_num_cus(b->ctx) * 16);
    hipLaunchKernelGGL(k_ix_build, dim3(grid), dim3(IX_THREADS), 0, st, b->ent, b->cnt, b->ent_off,
                       b->db, b->p0, b->P, b->L, b->lcap);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

static int bt_scores(hfx_batch *b, hipStream_t st)
{
    int rc;
    if (b->precision == HFX_BATCH_BF16X2_DENSE) {
        if (b->total == 0) return HFX_OK;
        return hfx_dense_scores(b->ctx, b->scores, b->score_off, b->begin, b->num, b->order,
                                b->dn_bstart, b->DB, &b->dn_lad, &b->sp, b->CH, b->dn_scratch, 1,
                                (void *)st);
    }
    if ((rc = bt_tiles(b, st))) return rc;
    return bt_score(b, st);
}

/* |dense - exact| <= (2^-16 + 210 * 2^-24) * 50 * sum|x|: the hi/lo split
 * (2^-16 per bin) plus at most 100 FP32 additions on any term's path (32 in
 * one MFMA, one per k-step after it) and the hi + lo sum, each allowed 2 ulp */
#define BT_REL_DENSE (1.52587890625e-05f + 210.0f * 5.9604644775390625e-08f)
#define BT_REL_F32 (80.0f * 5.9604644775390625e-08f)

static int bt_topn(hfx_batch *b, hipStream_t st)
{
    const float rel0 = b->precision == HFX_BATCH_BF16X2_DENSE ? BT_REL_DENSE : BT_REL_F32;
    return hfx_topn_device(b->ctx, b->top_pos, b->top_exact, b->top_f32, b->top_cnt, b->flag,
                           b->N, b->margin, b->scores, b->score_off, b->total, b->begin, b->num,
                           &b->db, &b->sp, rel0, 0.0f, b->tx_scratch, (void *)st);
}

extern "C" int hfx_batch_run(hfx_batch *b, void *stream)
{
    HFX_REQUIRE(b && b->planned, HFX_ERR_ARG, "hfx_batch_run: batch not planned");
    hipStream_t st = hfx_stream(stream);
    int rc;
    if (b->S == 0) return HFX_OK;
    if ((rc = bt_windows(b, st)) || (rc = bt_index(b, st)) || (rc = bt_scores(b, st))) return rc;
    return bt_topn(b, st);
}

extern "C" int hfx_batch_set_precision(hfx_batch *b, uint32_t precision)
{
    HFX_REQUIRE(b && b->planned, HFX_ERR_ARG, "hfx_batch_set_precision: batch not planned");
    HFX_REQUIRE(precision == HFX_BATCH_FP32_SPARSE || precision == HFX_BATCH_BF16X2_DENSE,
                HFX_ERR_ARG, "hfx_batch_set_precision: unknown precision %u", precision);
    HFX_REQUIRE(precision != HFX_BATCH_BF16X2_DENSE || b->dn_scratch, HFX_ERR_ARG,
                "hfx_batch_set_precision: the dense form takes spectra of up to %u bins "
                "(x_stride %u)", hfx_dense_max_len(), b->sp.x_stride);
    if (precision != b->precision && b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    b->precision = precision;
    return HFX_OK;
}

extern "C" int hfx_batch_capture(hfx_batch *b, void *stream)
{
    HFX_REQUIRE(b && b->planned, HFX_ERR_ARG, "hfx_batch_capture: batch not planned");
    HFX_REQUIRE(stream != nullptr, HFX_ERR_ARG, "hfx_batch_capture: needs a non-null stream");
    hipStream_t st = hfx_stream(stream);
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    HFX_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = hfx_batch_run(b, stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    HFX_REQUIRE(rc == HFX_OK, rc, "hfx_batch_capture: run failed during capture: %s",
                hfx_last_error());
    HFX_CHECK_HIP(e);
    b->graph = g;
    HFX_CHECK_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    return HFX_OK;
}

extern "C" int hfx_batch_replay(hfx_batch *b, void *stream)
{
    HFX_REQUIRE(b && b->exec, HFX_ERR_ARG, "hfx_batch_replay: no captured graph");
    HFX_CHECK_HIP(hipGraphLaunch(b->exec, hfx_stream(stream)));
    return HFX_OK;
}

/* Times the scoring kernel of the planned batch for each (G, C, threads) in
 * a small grid and keeps the fastest; the choice is remembered for batches of
 * the same shape bucket in this process. */
extern "C" int hfx_batch_autotune(hfx_batch *b, void *stream, uint32_t *out_cfg)
{
    HFX_REQUIRE(b && b->planned, HFX_ERR_ARG, "hfx_batch_autotune: batch not planned");
    hipStream_t st = hfx_stream(stream);
    int rc;
    /* index and windows once */
    if ((rc = hfx_batch_run(b, stream))) return rc;
    hipEvent_t e0, e1;
    HFX_CHECK_HIP(hipEventCreate(&e0));
    HFX_CHECK_HIP(hipEventCreate(&e1));
    /* a warm run, then the fastest of three timed runs (the minimum is the
     * least disturbed by clocks and other work) */
    auto time2 = [&](float *ms) -> int {
        int r;
        if ((r = bt_scores(b, st))) return r; /* warm */
        *ms = 1e30f;
        for (int k = 0; k < 3; ++k) {
            float t = 0.0f;
            HFX_CHECK_HIP(hipEventRecord(e0, st));
            if ((r = bt_scores(b, st))) return r;
            HFX_CHECK_HIP(hipEventRecord(e1, st));
            HFX_CHECK_HIP(hipEventSynchronize(e1));
            HFX_CHECK_HIP(hipEventElapsedTime(&t, e0, e1));
            *ms = t < *ms ? t : *ms;
        }
        return HFX_OK;
    };
    float best = 1e30f;
    uint32_t bG = b->G, bC = b->C, bT = b->threads, bCH = b->CH;
    if (b->precision == HFX_BATCH_BF16X2_DENSE) {
        const uint32_t CHs[5] = {256, 512, 1024, 2048, 4096};
        for (uint32_t CH : CHs) {
            b->CH = CH;
            float ms = 0;
            if ((rc = time2(&ms))) return rc;
            if (ms < best) { best = ms; bCH = CH; }
        }
    } else {
        const uint32_t Gs[2] = {4, 8}, Cs[4] = {32, 64, 128, 256}, Ts[2] = {256, 512};
        for (uint32_t G : Gs)
            for (uint32_t C : Cs)
                for (uint32_t T : Ts) {
                    b->G = G;
                    b->C = C;
                    b->threads = T;
                    float ms = 0;
                    if ((rc = time2(&ms))) return rc;
                    if (ms < best) { best = ms; bG = G; bC = C; bT = T; }
                }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->G = bG;
    b->C = bC;
    b->threads = bT;
    b->CH = bCH;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (b->precision == HFX_BATCH_BF16X2_DENSE) g_tune_dense[tune_key(b)] = bCH;
        else g_tune[tune_key(b)] = std::make_tuple(bG, bC, bT);
    }
    if (b->exec) {
        /* the captured graph holds the old launch shapes */
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    if (out_cfg) {
        out_cfg[0] = bG; out_cfg[1] = bC; out_cfg[2] = bT;
        if (b->precision == HFX_BATCH_BF16X2_DENSE) { out_cfg[0] = 0; out_cfg[1] = bCH; out_cfg[2] = 256; }
    }
    return HFX_OK;
}

/* Runs the planned batch phase by phase with events between the phases:
 * ms[0] windows + offsets, ms[1] ion index (sparse form only), ms[2] blocks,
 * tiles + scoring kernel, ms[3] top-N (+ exact route). */
extern "C" int hfx_batch_time(hfx_batch *b, void *stream, float *ms)
{
    HFX_REQUIRE(b && b->planned && ms, HFX_ERR_ARG, "hfx_batch_time: batch not planned");
    hipStream_t st = hfx_stream(stream);
    hipEvent_t ev[5];
    for (auto &e : ev) HFX_CHECK_HIP(hipEventCreate(&e));
    int rc = HFX_OK;
    HFX_CHECK_HIP(hipEventRecord(ev[0], st));
    if (!rc) rc = bt_windows(b, st);
    HFX_CHECK_HIP(hipEventRecord(ev[1], st));
    if (!rc) rc = bt_index(b, st);
    HFX_CHECK_HIP(hipEventRecord(ev[2], st));
    if (!rc) rc = bt_scores(b, st);
    HFX_CHECK_HIP(hipEventRecord(ev[3], st));
    if (!rc) rc = bt_topn(b, st);
    HFX_CHECK_HIP(hipEventRecord(ev[4], st));
    HFX_CHECK_HIP(hipEventSynchronize(ev[4]));
    for (int i = 0; i < 4; ++i) HFX_CHECK_HIP(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

extern "C" int hfx_batch_set_config(hfx_batch *b, uint32_t G, uint32_t C, uint32_t threads)
{
    HFX_REQUIRE(b && b->planned, HFX_ERR_ARG, "hfx_batch_set_config: batch not planned");
    if (b->precision == HFX_BATCH_BF16X2_DENSE) {
        /* dense form: C = candidates per work item (G, threads fixed) */
        HFX_REQUIRE(C >= 16 && C % 16 == 0 && C <= 8192, HFX_ERR_ARG,
                    "hfx_batch_set_config: dense chunk must be a multiple of 16 in 16..8192");
        b->CH = C;
        return HFX_OK;
    }
    HFX_REQUIRE((G == 4 || G == 8) && C >= 1 && C <= 4096 && (threads == 256 || threads == 512),
                HFX_ERR_ARG, "hfx_batch_set_config: G in {4,8}, C in 1..4096, threads in {256,512}");
    b->G = G;
    b->C = C;
    b->threads = threads;
    return HFX_OK;
}

extern "C" int hfx_batch_dense_blocks(const hfx_batch *b, uint32_t *h_bstart, uint32_t cap,
                                      uint32_t *num_blocks)
{
    HFX_REQUIRE(b && num_blocks, HFX_ERR_ARG, "hfx_batch_dense_blocks: null argument");
    HFX_REQUIRE(b->planned, HFX_ERR_ARG, "hfx_batch_dense_blocks: batch not planned");
    *num_blocks = b->DB;
    if (h_bstart && cap) {
        const size_t n = std::min<size_t>(cap, (size_t)b->DB + 1);
        HFX_CHECK_HIP(hipMemcpy(h_bstart, b->dn_bstart, 4 * n, hipMemcpyDeviceToHost));
    }
    return HFX_OK;
}

extern "C" int hfx_batch_info(const hfx_batch *b, uint64_t *info /*[12]*/)
{
    HFX_REQUIRE(b && info, HFX_ERR_ARG, "hfx_batch_info: null argument");
    info[0] = b->S; info[1] = b->total; info[2] = b->P; info[3] = b->L; info[4] = b->n_ent;
    info[5] = b->G; info[6] = b->C; info[7] = b->threads; info[8] = b->lstr; info[9] = b->p0;
    info[10] = b->precision; info[11] = b->CH;
    return HFX_OK;
}

extern "C" int hfx_batch_outputs(const hfx_batch *b, void **ptrs /*[10]*/)
{
    HFX_REQUIRE(b && ptrs, HFX_ERR_ARG, "hfx_batch_outputs: null argument");
    ptrs[0] = b->top_pos; ptrs[1] = b->top_exact; ptrs[2] = b->top_f32; ptrs[3] = b->top_cnt;
    ptrs[4] = b->flag; ptrs[5] = b->scores; ptrs[6] = b->score_off; ptrs[7] = b->begin;
    ptrs[8] = b->num; ptrs[9] = b->ent;
    return HFX_OK;
}
