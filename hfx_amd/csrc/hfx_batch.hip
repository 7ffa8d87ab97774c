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

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
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
    __shared__ uint32_t sh_t, sh_lblk;
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
/* Per-database peptide index                                            */
/* ------------------------------------------------------------------ */

/* Everything the batched forms precompute per peptide: it depends on the
 * database only, so it is built once and shared by every batch. */
struct hfx_pepix {
    hfx_ctx *ctx;
    hfx_db db;
    uint32_t forms, L, lcap;
    /* HFX_PEPIX_LADDER: b ladders of every peptide, residual-sorted order */
    void *lad_mem[2];
    hfx_dn_ladder lad;
    /* HFX_PEPIX_IONS: peptide q (residual-sorted position), level l <= L:
     * cnt[q L + l - 1] entries at ent + ent_off[q L + l - 1] */
    uint32_t *caps, *cnt;
    uint64_t *ent_off;
    uint16_t *ent;
    uint64_t n_ent, n_pos, bytes;
};

template <class T>
static int dalloc(T **p, size_t n)
{
    hipError_t e = hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        hfx_set_error("device allocation of %zu bytes failed: %s", n * sizeof(T),
                      hipGetErrorString(e));
        return HFX_ERR_NOMEM;
    }
    return HFX_OK;
}

extern "C" int hfx_pepix_destroy(hfx_pepix *ix)
{
    if (!ix) return HFX_OK;
    (void)hipDeviceSynchronize();
    void *ps[] = {ix->lad_mem[0], ix->lad_mem[1], ix->caps, ix->cnt, ix->ent_off, ix->ent};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    free(ix);
    return HFX_OK;
}

extern "C" int hfx_pepix_build(hfx_ctx *ctx, const hfx_db *db, uint32_t forms, uint32_t levels,
                               uint32_t max_len, void *stream, hfx_pepix **out)
{
    HFX_REQUIRE(ctx && db && out, HFX_ERR_ARG, "hfx_pepix_build: null argument");
    HFX_REQUIRE(forms && !(forms & ~(HFX_PEPIX_LADDER | HFX_PEPIX_IONS)), HFX_ERR_ARG,
                "hfx_pepix_build: forms 0x%x", forms);
    HFX_REQUIRE(!(forms & HFX_PEPIX_IONS) || (levels >= 1 && levels <= HFX_MAX_CHARGE &&
                                                max_len >= 8 && max_len <= HFX_MAX_LEN),
                HFX_ERR_ARG, "hfx_pepix_build: levels 1..%d and max_len 8..%d", HFX_MAX_CHARGE,
                HFX_MAX_LEN);
    *out = nullptr;
    hfx_pepix *ix = (hfx_pepix *)calloc(1, sizeof(hfx_pepix));
    HFX_REQUIRE(ix, HFX_ERR_NOMEM, "hfx_pepix_build: out of host memory");
    ix->ctx = ctx;
    ix->db = *db;
    ix->forms = forms;
    hipStream_t st = hfx_stream(stream);
    const uint32_t P = db->num_pep;
    int rc = HFX_OK;
    if (forms & HFX_PEPIX_LADDER) {
        rc = hfx_dense_ladder_build(ctx, db, 0, P, 1, &ix->lad, ix->lad_mem, stream);
        if (!rc) {
            ix->n_pos = 0;
            rc = hipMemcpyAsync(&ix->n_pos, ix->lad.off + P, 8, hipMemcpyDeviceToHost, st) ||
                         hipStreamSynchronize(st)
                     ? HFX_ERR_HIP
                     : HFX_OK;
            ix->bytes += ((uint64_t)P + 1) * 8 + (uint64_t)P * 8 +
                         std::max<uint64_t>(ix->n_pos, 1) * sizeof(hfx_dn_pos);
        }
    }
    if (!rc && (forms & HFX_PEPIX_IONS)) {
        ix->L = levels;
        ix->lcap = max_len;
        const uint64_t PL = (uint64_t)P * levels;
        if (PL >= 0xffffffffull) {
            hfx_set_error("hfx_pepix_build: %u peptides x %u levels", P, levels);
            rc = HFX_ERR_ARG;
        }
        if (!rc) rc = dalloc(&ix->caps, PL);
        if (!rc) rc = dalloc(&ix->ent_off, PL + 1);
        if (!rc) rc = dalloc(&ix->cnt, PL);
        if (!rc && P) {
            hipLaunchKernelGGL(k_ix_caps, dim3((P + 255) / 256), dim3(256), 0, st, ix->caps, *db,
                               0u, P, levels, max_len);
            rc = hfx_scan_offsets(ctx, ix->ent_off, ix->caps, (uint32_t)PL, stream);
            if (!rc && (hipMemcpyAsync(&ix->n_ent, ix->ent_off + PL, 8, hipMemcpyDeviceToHost,
                                       st) != hipSuccess ||
                        hipStreamSynchronize(st) != hipSuccess)) {
                hfx_set_error("hfx_pepix_build: reading the index size failed");
                rc = HFX_ERR_HIP;
            }
        }
        if (!rc) rc = dalloc(&ix->ent, ix->n_ent);
        if (!rc && P) {
            const uint32_t grid = (uint32_t)std::min<uint64_t>(
                (P + IX_WAVES - 1) / IX_WAVES, (uint64_t)hfx_ctx_num_cus(ctx) * 16);
            hipLaunchKernelGGL(k_ix_build, dim3(grid), dim3(IX_THREADS), 0, st, ix->ent, ix->cnt,
                               ix->ent_off, *db, 0u, P, levels, max_len);
        }
        ix->bytes += PL * 4 * 2 + (PL + 1) * 8 + ix->n_ent * 2;
    }
    if (!rc && hipGetLastError() != hipSuccess) {
        hfx_set_error("hfx_pepix_build: kernel launch failed");
        rc = HFX_ERR_HIP;
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess) {
        hfx_set_error("hfx_pepix_build: synchronize failed");
        rc = HFX_ERR_HIP;
    }
    if (rc) {
        hfx_pepix_destroy(ix);
        return rc;
    }
    *out = ix;
    return HFX_OK;
}

extern "C" int hfx_pepix_info(const hfx_pepix *ix, uint64_t *info /*[6]*/)
{
    HFX_REQUIRE(ix && info, HFX_ERR_ARG, "hfx_pepix_info: null argument");
    info[0] = ix->forms; info[1] = ix->L; info[2] = ix->lcap; info[3] = ix->n_pos;
    info[4] = ix->n_ent; info[5] = ix->bytes;
    return HFX_OK;
}

/* ------------------------------------------------------------------ */
/* Spectrum order and dense blocks, on the device                        */
/* ------------------------------------------------------------------ */

/* Spectra are ordered by (fragment charge level, precursor mass, index):
 * bucket = level * nbz + (window begin >> sh) -- the window's first
 * residual-sorted position is a monotone function of the precursor mass --
 * counted, scattered, then ranked inside each bucket by (mass, index). */
__global__ void k_ord_fill(uint32_t *a, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0u;
}

__device__ __forceinline__ uint32_t ord_bucket(uint32_t s, const uint32_t *begin,
                                               const uint32_t *zmax, uint32_t sh, uint32_t nbz)
{
    return hfx_zeff(zmax[s]) * nbz + (begin[s] >> sh);
}

__global__ void k_ord_count(uint32_t *cnt, const uint32_t *begin, const uint32_t *zmax, uint32_t S,
                            uint32_t sh, uint32_t nbz)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) atomicAdd(&cnt[ord_bucket(s, begin, zmax, sh, nbz)], 1u);
}

/* counts down to zero, leaving cnt cleared */
__global__ void k_ord_scatter(uint32_t *tmp, uint32_t *cnt, const uint64_t *boff,
                              const uint32_t *begin, const uint32_t *zmax, uint32_t S, uint32_t sh,
                              uint32_t nbz)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const uint32_t k = ord_bucket(s, begin, zmax, sh, nbz);
    tmp[boff[k] + atomicSub(&cnt[k], 1u) - 1u] = s;
}

__global__ void k_ord_rank(uint32_t *order, const uint32_t *tmp, const uint64_t *boff,
                           const uint32_t *begin, const uint32_t *zmax, const float *mass,
                           uint32_t S, uint32_t sh, uint32_t nbz)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    const uint32_t s = tmp[i];
    const uint32_t k = ord_bucket(s, begin, zmax, sh, nbz);
    const uint32_t a = (uint32_t)boff[k], e = (uint32_t)boff[k + 1];
    const float ms = mass[s];
    uint32_t rank = 0;
    for (uint32_t j = a; j < e; ++j) {
        const uint32_t t = tmp[j];
        const float mt = mass[t];
        rank += (mt < ms || (mt == ms && t < s)) ? 1u : 0u;
    }
    order[a + rank] = s;
}

/* Dense blocks: the order cut into runs of at most DN_NS spectra that
 * minimise the sum over blocks of (charge levels x rows of the hull of the
 * windows) -- a block spanning a gap in precursor mass would score every
 * candidate in the gap.  A dynamic programme per segment of CUT_SEG order
 * positions (segments are cut independently, which costs about 1% more
 * rows than one programme over the whole order), one row of 16 lanes a
 * segment: lane i holds the window of the (i+1)-th newest position and the
 * best cost of the positions before it.  Each step shifts both by one lane
 * (DPP row_shr:1, the new position entering lane 0), takes the hull of the
 * newest i+1 windows by inclusive row scans (min of starts, max of ends, OR
 * of charge levels) and the best block length by a row minimum of
 * (cost << 4 | i): the shortest block among equal costs, the one a serial
 * scan of lengths 1, 2, ... with strict improvement keeps. */
#define CUT_SEG 32
#define CUT_ROW 16
#define CUT_WG 256
#define CUT_SPW (CUT_WG / CUT_ROW) /* segments a workgroup */

template <int CTRL>
__device__ __forceinline__ uint32_t cut_dpp(uint32_t old, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ uint64_t cut_dpp64(uint64_t old, uint64_t v)
{
    return (uint64_t)cut_dpp<CTRL>((uint32_t)(old >> 32), (uint32_t)(v >> 32)) << 32 |
           cut_dpp<CTRL>((uint32_t)old, (uint32_t)v);
}

template <int CUT_NS>
__global__ void __launch_bounds__(CUT_WG) k_cut_seg(uint32_t *seg_cuts, uint32_t *seg_n,
                                                   const uint32_t *order, const uint32_t *begin,
                                                   const uint32_t *num, const uint32_t *zmax,
                                                   uint32_t S)
{
    static_assert(CUT_NS <= CUT_ROW, "every block length has its lane");
    __shared__ uint32_t lo_s[CUT_SPW * CUT_SEG], hi_s[CUT_SPW * CUT_SEG], lv_s[CUT_SPW * CUT_SEG];
    __shared__ uint8_t from_s[CUT_SPW][CUT_SEG + 1];
    const uint32_t base = blockIdx.x * CUT_SPW * CUT_SEG;
    for (uint32_t i = threadIdx.x; i < CUT_SPW * CUT_SEG; i += CUT_WG) {
        const uint32_t p = base + i;
        uint32_t lo = 0xffffffffu, hi = 0, lv = 0;
        if (p < S) {
            const uint32_t s = order[p], z = hfx_zeff(zmax[s]), n = num[s];
            if (z && n) {
                lo = begin[s];
                hi = lo + n;
                lv = 1u << z;
            }
        }
        lo_s[i] = lo;
        hi_s[i] = hi;
        lv_s[i] = lv;
    }
    __syncthreads();
    const uint32_t r = threadIdx.x / CUT_ROW, i = threadIdx.x % CUT_ROW;
    const uint32_t seg = blockIdx.x * CUT_SPW + r;
    const uint32_t p0 = seg * CUT_SEG;
    if (p0 >= S) return; /* the whole row leaves together */
    const uint32_t len = min((uint32_t)CUT_SEG, S - p0);
    const uint32_t *lo = lo_s + r * CUT_SEG, *hi = hi_s + r * CUT_SEG, *lv = lv_s + r * CUT_SEG;
    uint8_t *from = from_s[r];
    uint32_t wlo = 0xffffffffu, whi = 0, wlv = 0;
    uint64_t bst = i == 0 ? 0ull : ~0ull;
    for (uint32_t j = 1; j <= len; ++j) {
        wlo = cut_dpp<0x111>(lo[j - 1], wlo);
        whi = cut_dpp<0x111>(hi[j - 1], whi);
        wlv = cut_dpp<0x111>(lv[j - 1], wlv);
        uint32_t l = wlo, h = whi, v = wlv;
        l = min(l, cut_dpp<0x111>(0xffffffffu, l));
        h = max(h, cut_dpp<0x111>(0u, h));
        v |= cut_dpp<0x111>(0u, v);
        l = min(l, cut_dpp<0x112>(0xffffffffu, l));
        h = max(h, cut_dpp<0x112>(0u, h));
        v |= cut_dpp<0x112>(0u, v);
        l = min(l, cut_dpp<0x114>(0xffffffffu, l));
        h = max(h, cut_dpp<0x114>(0u, h));
        v |= cut_dpp<0x114>(0u, v);
        l = min(l, cut_dpp<0x118>(0xffffffffu, l));
        h = max(h, cut_dpp<0x118>(0u, h));
        v |= cut_dpp<0x118>(0u, v);
        /* lane i: the block of the newest i + 1 positions after the best cut
         * of the ones before (bst, finite for i < j) */
        const uint64_t rows = h > l ? (uint64_t)(h - l) : 0ull;
        uint64_t key = ~0ull;
        if (i < (uint32_t)CUT_NS && i < j)
            key = (bst + rows * (uint64_t)__builtin_popcount(v)) << 4 | i;
        uint64_t t;
        t = cut_dpp64<0xB1>(0ull, key);  /* quad_perm [1,0,3,2] */
        key = t < key ? t : key;
        t = cut_dpp64<0x4E>(0ull, key);  /* quad_perm [2,3,0,1] */
        key = t < key ? t : key;
        t = cut_dpp64<0x141>(0ull, key); /* row_half_mirror */
        key = t < key ? t : key;
        t = cut_dpp64<0x140>(0ull, key); /* row_mirror */
        key = t < key ? t : key;
        bst = cut_dpp64<0x111>(key >> 4, bst);
        if (i == 0) from[j] = (uint8_t)(j - 1u - (uint32_t)(key & 15u));
    }
    if (i != 0) return;
    /* block starts, back from the segment's end, then written in order */
    uint32_t cnt = 0;
    for (uint32_t j = len; j > 0; j = from[j]) ++cnt;
    uint32_t w = cnt;
    for (uint32_t j = len; j > 0; j = from[j]) seg_cuts[(size_t)seg * CUT_SEG + --w] = p0 + from[j];
    seg_n[seg] = cnt;
}

/* bstart[0 .. nb) = the segments' block starts in order, bstart[nb .. cap]
 * = S (empty blocks), nb = seg_off[nseg] */
__global__ void k_cut_place(uint32_t *bstart, const uint32_t *seg_cuts, const uint32_t *seg_n,
                            const uint64_t *seg_off, uint32_t nseg, uint32_t cap, uint32_t S)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t seg = t / CUT_SEG, i = t % CUT_SEG;
    if (seg < nseg && i < seg_n[seg]) bstart[seg_off[seg] + i] = seg_cuts[t];
    const uint32_t nb = (uint32_t)seg_off[nseg];
    if (t <= cap && t >= nb) bstart[t] = S;
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
    uint32_t S, G, C, threads, B, L, lstr;
    uint64_t total;
    bool planned;
    const hfx_pepix *ix; /* the caller's index, or own */
    hfx_pepix *own;      /* built on first need when the caller gave none */
    /* device buffers (owned) */
    uint32_t *begin, *num, *order, *ubeg, *uend, *ntiles, *counter;
    uint64_t *score_off, *tile_off;
    float *scores;
    uint32_t *top_pos, *top_cnt, *flag;
    double *top_exact;
    float *top_f32;
    uint8_t *tx_scratch; /* top-N exact route */
    /* spectrum order: buckets */
    uint32_t ord_sh, ord_nbz, ord_nb;
    uint32_t *ord_cnt, *ord_tmp;
    uint64_t *ord_off;
    uint32_t precision; /* HFX_BATCH_FP32_SPARSE, _BF16X2_DENSE or _I8X3_DENSE */
    uint32_t CH;        /* dense: candidates per work item */
    uint8_t *dn_scratch; /* dense: hi/lo spectra, blocks (allocated with the dense form) */
    uint32_t *dn_bstart; /* dense: [S + 1] block starts in order (S past the last block) */
    uint32_t *cut_seg, *cut_n;
    uint64_t *cut_off;
    uint32_t nseg;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

static void bt_free(hfx_batch *b)
{
    void *ps[] = {b->begin,    b->num,       b->order,     b->ubeg,      b->uend,
                  b->ntiles,   b->counter,   b->score_off, b->tile_off,  b->scores,
                  b->top_pos,  b->top_cnt,   b->flag,      b->top_exact, b->top_f32,
                  b->tx_scratch, b->ord_cnt, b->ord_tmp,   b->ord_off,   b->dn_scratch,
                  b->dn_bstart, b->cut_seg,  b->cut_n,     b->cut_off};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    hfx_ctx *ctx = b->ctx;
    const hfx_pepix *ix = b->ix;
    hfx_pepix *own = b->own;
    memset(b, 0, sizeof *b);
    b->ctx = ctx;
    b->own = own;
    b->ix = ix == own ? nullptr : ix; /* the caller's index stays attached */
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
    if (b->own) hfx_pepix_destroy(b->own);
    free(b);
    return HFX_OK;
}

extern "C" int hfx_batch_use_index(hfx_batch *b, const hfx_pepix *ix)
{
    HFX_REQUIRE(b, HFX_ERR_ARG, "hfx_batch_use_index: null batch");
    HFX_REQUIRE(!b->planned || !ix ||
                    (ix->db.d_idx_sorted == b->db.d_idx_sorted && ix->db.num_pep == b->db.num_pep),
                HFX_ERR_ARG, "hfx_batch_use_index: the index belongs to another database");
    b->ix = ix;
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    return HFX_OK;
}

/* Shape-bucketed autotuner results of this process: (S bucket, P bucket, L)
 * -> (G, C, threads). */
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, uint32_t>, std::tuple<uint32_t, uint32_t, uint32_t>> g_tune;
static std::map<std::tuple<int, int, uint32_t>, uint32_t> g_tune_dense, g_tune_dense8;

/* the two dense forms share the ladders, block cut and work items */
static inline bool bt_is_dense(uint32_t precision)
{
    return precision == HFX_BATCH_BF16X2_DENSE || precision == HFX_BATCH_I8X3_DENSE;
}
static std::map<std::tuple<int, int, uint32_t>, uint32_t> &bt_tune_dense(uint32_t precision)
{
    return precision == HFX_BATCH_I8X3_DENSE ? g_tune_dense8 : g_tune_dense;
}
/* the dense forms' per-batch scratch: the larger of the two */
static size_t bt_dense_scratch(uint32_t S)
{
    return std::max(hfx_dense_scratch_bytes(S), hfx_dense8_scratch_bytes(S));
}

/* the launch shapes the kernels support: sparse G in {4, 8}, C in 1..4096,
 * threads in {256, 512}; a dense chunk a multiple of its tile (16 candidates
 * for BF16, 64 for int8) up to 8192 */
static bool tune_sparse_ok(uint32_t G, uint32_t C, uint32_t threads)
{
    return (G == 4 || G == 8) && C >= 1 && C <= 4096 && (threads == 256 || threads == 512);
}
static uint32_t tune_dense_step(uint32_t precision)
{
    return precision == HFX_BATCH_I8X3_DENSE ? 64u : 16u;
}
static bool tune_dense_ok(uint32_t precision, uint32_t CH)
{
    const uint32_t step = tune_dense_step(precision);
    return CH >= step && CH % step == 0 && CH <= 8192;
}

static std::tuple<int, int, uint32_t> tune_key(const hfx_batch *b)
{
    auto lg = [](uint64_t v) { int k = 0; while (v > 1) { v >>= 1; ++k; } return k; };
    return std::make_tuple(lg(b->S), lg(b->total), b->L);
}

/* The tuned shapes persist across processes when HFX_TUNE_CACHE names a
 * file: one line a shape, "sparse <S bucket> <pairs bucket> <levels> <G> <C>
 * <threads>", "bf16 ... <chunk>" or "int8 ... <chunk>" (other lines are
 * ignored).  It is read when a lookup first sees the path (its entries
 * replace this process's for the same shapes) and rewritten whole, through a
 * temporary file and a rename, after each tuning.  Callers hold g_tune_mu. */
static std::string g_tune_path;
static void tune_load_locked()
{
    const char *p = getenv("HFX_TUNE_CACHE");
    const std::string path = p ? p : "";
    if (path == g_tune_path) return;
    g_tune_path = path;
    if (path.empty()) return;
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return;
    char kind[16];
    int sb, pb;
    unsigned L, a, c, t;
    char line[256];
    while (fgets(line, sizeof line, f)) {
        const int n = sscanf(line, "%15s %d %d %u %u %u %u", kind, &sb, &pb, &L, &a, &c, &t);
        const auto key = std::make_tuple(sb, pb, (uint32_t)L);
        /* entries outside the supported shapes (a damaged or hand-edited
         * file) are skipped, never launched */
        if (n == 7 && !strcmp(kind, "sparse") && tune_sparse_ok(a, c, t))
            g_tune[key] = std::make_tuple((uint32_t)a, (uint32_t)c, (uint32_t)t);
        else if (n == 5 && !strcmp(kind, "bf16") && tune_dense_ok(HFX_BATCH_BF16X2_DENSE, a))
            g_tune_dense[key] = a;
        else if (n == 5 && !strcmp(kind, "int8") && tune_dense_ok(HFX_BATCH_I8X3_DENSE, a))
            g_tune_dense8[key] = a;
    }
    fclose(f);
}
static void tune_save_locked()
{
    if (g_tune_path.empty()) return;
    /* a temporary name of this process's own, so that processes tuning at
     * once never write one file together (the last rename wins) */
    const std::string tmp = g_tune_path + ".tmp." + std::to_string((long)getpid());
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) return;
    for (const auto &e : g_tune)
        fprintf(f, "sparse %d %d %u %u %u %u\n", std::get<0>(e.first), std::get<1>(e.first),
                std::get<2>(e.first), std::get<0>(e.second), std::get<1>(e.second),
                std::get<2>(e.second));
    for (const auto &e : g_tune_dense)
        fprintf(f, "bf16 %d %d %u %u\n", std::get<0>(e.first), std::get<1>(e.first),
                std::get<2>(e.first), e.second);
    for (const auto &e : g_tune_dense8)
        fprintf(f, "int8 %d %d %u %u\n", std::get<0>(e.first), std::get<1>(e.first),
                std::get<2>(e.first), e.second);
    if (fclose(f) == 0) rename(tmp.c_str(), g_tune_path.c_str());
}

/* the form of the peptide index the current precision reads */
static uint32_t bt_form(const hfx_batch *b)
{
    return bt_is_dense(b->precision) ? HFX_PEPIX_LADDER : HFX_PEPIX_IONS;
}

/* makes sure the index holds the current precision's form: the caller's,
 * or one the batch builds (never while a stream is being captured) */
static int bt_need_index(hfx_batch *b, hipStream_t st)
{
    const uint32_t f = bt_form(b);
    if (b->ix && (b->ix->forms & f) && (f != HFX_PEPIX_IONS || b->ix->L >= b->L)) return HFX_OK;
    HFX_REQUIRE(!b->ix || b->ix == b->own, HFX_ERR_ARG,
                "batch: the attached peptide index lacks the %s form this precision reads",
                f == HFX_PEPIX_LADDER ? "ladder" : "ion list");
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HFX_CHECK_HIP(hipStreamIsCapturing(st, &cs));
    HFX_REQUIRE(cs == hipStreamCaptureStatusNone, HFX_ERR_ARG,
                "batch: the peptide index is built on first use, not under capture "
                "(run or set the precision once before capturing)");
    uint32_t forms = f;
    if (b->own) forms |= b->own->forms;
    hfx_pepix *nx = nullptr;
    int rc = hfx_pepix_build(b->ctx, &b->db, forms, HFX_MAX_CHARGE, HFX_MAX_LEN, (void *)st, &nx);
    if (rc) return rc;
    if (b->own) hfx_pepix_destroy(b->own);
    b->own = nx;
    b->ix = nx;
    return HFX_OK;
}

/* Device bytes hfx_batch_plan allocates for S spectra with `total` scored
 * pairs (the precision form's own buffers included, the per-database index
 * not): the same sums as the allocations below. */
static uint64_t bt_bytes(uint32_t S, uint64_t total, uint32_t N, uint32_t num_pep,
                         uint32_t precision)
{
    auto al = [](uint64_t n, uint64_t sz) { return (std::max<uint64_t>(n, 1) * sz + 255) & ~255ull; };
    uint32_t sh = 0;
    while (((uint64_t)num_pep >> sh) > std::max<uint64_t>(S / 2, 64)) ++sh;
    const uint64_t nb = (uint64_t)(HFX_MAX_CHARGE + 1) * ((num_pep >> sh) + 1);
    const uint64_t bmax = (S + 3) / 4;
    uint64_t b = al(S, 4) * 3 + al((uint64_t)S + 1, 8)                  /* begin, num, order, off */
                 + al(nb, 4) + al(nb + 1, 8) + al(S, 4)                   /* order buckets */
                 + al(total, 4) + al(1, 4)                               /* scores, counter */
                 + al((uint64_t)S * N, 4) + al((uint64_t)S * N, 8) + al((uint64_t)S * N, 4) +
                 al(S, 4) * 2                                            /* top-N */
                 + al(hfx_topn_scratch_bytes(S, total), 1)               /* exact route */
                 + al(bmax, 4) * 3 + al(bmax + 1, 8);                    /* sparse blocks */
    if (bt_is_dense(precision)) {
        const uint64_t nseg = (S + CUT_SEG - 1) / CUT_SEG;
        b += al(bt_dense_scratch(S), 1) + al((uint64_t)S + 1, 4) + al(nseg * CUT_SEG, 4) +
             al(nseg, 4) + al(nseg + 1, 8);
    }
    return b;
}

extern "C" int hfx_batch_required_bytes(uint32_t S, uint64_t total, uint32_t N, uint32_t num_pep,
                                        uint32_t precision, uint64_t *bytes)
{
    HFX_REQUIRE(bytes, HFX_ERR_ARG, "hfx_batch_required_bytes: null out");
    HFX_REQUIRE(precision <= HFX_BATCH_I8X3_DENSE, HFX_ERR_ARG,
                "hfx_batch_required_bytes: unknown precision %u", precision);
    *bytes = bt_bytes(S, total, N, num_pep, precision);
    return HFX_OK;
}

/* Search.hs:427-433 / Execution.hs:77-93 (groupScorePlan): consecutive runs
 * of spectra whose batch fits `budget` device bytes, each as long as it can
 * be (h_off: the S + 1 score offsets, window sizes summed).  A spectrum that
 * alone needs more than the budget is an error (the reference splits such a
 * spectrum's candidates instead; a batch here scores whole windows). */
extern "C" int hfx_batch_split(const uint64_t *h_off, uint32_t S, uint32_t N, uint32_t num_pep,
                               uint32_t precision, uint64_t budget, uint32_t *h_first,
                               uint32_t cap, uint32_t *num_batches)
{
    HFX_REQUIRE(h_off && num_batches, HFX_ERR_ARG, "hfx_batch_split: null argument");
    HFX_REQUIRE(precision <= HFX_BATCH_I8X3_DENSE, HFX_ERR_ARG,
                "hfx_batch_split: unknown precision %u", precision);
    for (uint32_t s = 0; s < S; ++s)
        HFX_REQUIRE(h_off[s + 1] >= h_off[s], HFX_ERR_ARG,
                    "hfx_batch_split: score offsets decrease at spectrum %u", s);
    uint32_t nb = 0, a = 0;
    while (a < S) {
        /* the longest run [a, e) that fits: the requirement grows with the
         * run, so a binary search over its end */
        uint32_t lo = a, hi = S;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo + 1) / 2;
            if (bt_bytes(mid - a, h_off[mid] - h_off[a], N, num_pep, precision) <= budget) lo = mid;
            else hi = mid - 1;
        }
        HFX_REQUIRE(lo > a, HFX_ERR_NOMEM,
                    "hfx_batch_split: spectrum %u alone needs %llu bytes, budget %llu", a,
                    (unsigned long long)bt_bytes(1, h_off[a + 1] - h_off[a], N, num_pep, precision),
                    (unsigned long long)budget);
        if (h_first && nb < cap) h_first[nb] = a;
        ++nb;
        a = lo;
    }
    if (h_first && nb < cap) h_first[nb] = S;
    *num_batches = nb;
    return HFX_OK;
}

extern "C" int hfx_batch_plan(hfx_batch *b, const hfx_db *db, const hfx_spectra *sp, float eps,
                              uint32_t N, uint32_t margin, void *stream)
{
    HFX_REQUIRE(b && db && sp, HFX_ERR_ARG, "hfx_batch_plan: null argument");
    HFX_REQUIRE(N >= 1 && N + margin <= 128 && N <= 64, HFX_ERR_ARG,
                "hfx_batch_plan: N + margin must be <= 128 and N <= 64");
    hipStream_t st = hfx_stream(stream);
    bt_free(b);
    if (b->own && (b->own->db.d_idx_sorted != db->d_idx_sorted ||
                   b->own->db.num_pep != db->num_pep || b->own->db.d_ions != db->d_ions)) {
        hfx_pepix_destroy(b->own);
        b->own = nullptr;
    }
    if (b->ix && (b->ix->db.d_idx_sorted != db->d_idx_sorted || b->ix->db.num_pep != db->num_pep))
        HFX_REQUIRE(false, HFX_ERR_ARG, "hfx_batch_plan: the attached index belongs to another "
                                        "database");
    if (!b->ix) b->ix = b->own;
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
    /* the windows once, to size the score arrays (the run computes them again) */
    if ((rc = hfx_window(b->ctx, b->begin, b->num, db, sp->d_mass, S, eps, stream))) return rc;
    if ((rc = hfx_scan_offsets(b->ctx, b->score_off, b->num, S, stream))) return rc;
    std::vector<uint32_t> hz(S), hl(S);
    if (S) {
        HFX_CHECK_HIP(hipMemcpyAsync(hz.data(), sp->d_zmax, 4ull * S, hipMemcpyDeviceToHost, st));
        HFX_CHECK_HIP(hipMemcpyAsync(hl.data(), sp->d_len, 4ull * S, hipMemcpyDeviceToHost, st));
    }
    HFX_CHECK_HIP(hipMemcpyAsync(&b->total, b->score_off + S, 8, hipMemcpyDeviceToHost, st));
    HFX_CHECK_HIP(hipStreamSynchronize(st));
    {
        /* fail with a clear message, before allocating, when the batch
         * cannot fit (hfx_batch_split cuts a larger set into batches that do) */
        size_t fr = 0, tot = 0;
        HFX_CHECK_HIP(hipMemGetInfo(&fr, &tot));
        const uint64_t need = bt_bytes(S, b->total, N, db->num_pep, HFX_BATCH_FP32_SPARSE);
        HFX_REQUIRE(need <= fr, HFX_ERR_NOMEM,
                    "hfx_batch_plan: %u spectra with %llu scores need %llu device bytes, %zu free "
                    "(split them with hfx_batch_split)",
                    S, (unsigned long long)b->total, (unsigned long long)need, fr);
    }
    uint32_t L = 0, lmax = 0;
    for (uint32_t s = 0; s < S; ++s) {
        lmax = std::max(lmax, std::min(hl[s], std::min(sp->x_stride, (uint32_t)HFX_MAX_LEN)));
        L = std::max(L, hz[s] <= HFX_MAX_CHARGE ? hz[s] : 0u);
    }
    b->L = std::max(L, 1u);
    b->lstr = (std::max(lmax, 8u) + 15) & ~15u;
    /* order buckets: about two spectra a bucket on average */
    b->ord_sh = 0;
    while (((uint64_t)db->num_pep >> b->ord_sh) > std::max<uint64_t>(S / 2, 64)) ++b->ord_sh;
    b->ord_nbz = (db->num_pep >> b->ord_sh) + 1;
    b->ord_nb = (HFX_MAX_CHARGE + 1) * b->ord_nbz;
    if ((rc = dalloc(&b->ord_cnt, b->ord_nb)) || (rc = dalloc(&b->ord_off, (size_t)b->ord_nb + 1)) ||
        (rc = dalloc(&b->ord_tmp, S)))
        return rc;
    /* defaults, replaced by the autotuner's choice for this shape */
    b->G = 8;
    b->C = 64;
    b->threads = 512;
    b->CH = 256;
    b->precision = HFX_BATCH_FP32_SPARSE;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load_locked();
        auto it = g_tune.find(tune_key(b));
        if (it != g_tune.end()) std::tie(b->G, b->C, b->threads) = it->second;
        auto jt = g_tune_dense.find(tune_key(b));
        if (jt != g_tune_dense.end()) b->CH = jt->second;
    }
    if ((rc = dalloc(&b->scores, b->total)) || (rc = dalloc(&b->counter, 1)) ||
        (rc = dalloc(&b->top_pos, (size_t)S * N)) || (rc = dalloc(&b->top_exact, (size_t)S * N)) ||
        (rc = dalloc(&b->top_f32, (size_t)S * N)) || (rc = dalloc(&b->top_cnt, S)) ||
        (rc = dalloc(&b->flag, S)))
        return rc;
    /* the top-N exact route's scratch is owned by the batch, so that a
     * captured graph never allocates and never sees it move */
    if ((rc = dalloc(&b->tx_scratch, hfx_topn_scratch_bytes(S, b->total)))) return rc;
    /* blocks are sized for the smallest G the tuner may pick (4) */
    const uint32_t bmax = (S + 3) / 4;
    if ((rc = dalloc(&b->ubeg, bmax)) || (rc = dalloc(&b->uend, bmax)) ||
        (rc = dalloc(&b->ntiles, bmax)) || (rc = dalloc(&b->tile_off, (size_t)bmax + 1)))
        return rc;
    b->planned = true;
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

extern "C" int hfx_batch_plan_ix(hfx_batch *b, const hfx_pepix *ix, const hfx_spectra *sp,
                                 float eps, uint32_t N, uint32_t margin, void *stream)
{
    HFX_REQUIRE(b && ix, HFX_ERR_ARG, "hfx_batch_plan_ix: null argument");
    b->ix = ix;
    return hfx_batch_plan(b, &ix->db, sp, eps, N, margin, stream);
}

/* the dense form's per-batch buffers (hi/lo rows of every spectrum, block
 * bookkeeping), allocated when the form is first chosen */
static int bt_dense_alloc(hfx_batch *b)
{
    if (b->dn_scratch) return HFX_OK;
    int rc;
    b->nseg = (b->S + CUT_SEG - 1) / CUT_SEG;
    if ((rc = dalloc(&b->dn_scratch, bt_dense_scratch(b->S))) ||
        (rc = dalloc(&b->dn_bstart, (size_t)b->S + 1)) ||
        (rc = dalloc(&b->cut_seg, (size_t)b->nseg * CUT_SEG)) || (rc = dalloc(&b->cut_n, b->nseg)) ||
        (rc = dalloc(&b->cut_off, (size_t)b->nseg + 1)))
        return rc;
    return HFX_OK;
}

/* ---- the phases of one batched search ---- */

/* windows and score offsets */
static int bt_windows(hfx_batch *b, hipStream_t st)
{
    int rc;
    if ((rc = hfx_window(b->ctx, b->begin, b->num, &b->db, b->sp.d_mass, b->S, b->eps, (void *)st)))
        return rc;
    return hfx_scan_offsets(b->ctx, b->score_off, b->num, b->S, (void *)st);
}

/* spectrum order (fragment charge level, precursor mass, index) */
static int bt_order(hfx_batch *b, hipStream_t st)
{
    const uint32_t S = b->S, g = (S + 255) / 256;
    hipLaunchKernelGGL(k_ord_fill, dim3((b->ord_nb + 255) / 256), dim3(256), 0, st, b->ord_cnt,
                       b->ord_nb);
    hipLaunchKernelGGL(k_ord_count, dim3(g), dim3(256), 0, st, b->ord_cnt, b->begin, b->sp.d_zmax,
                       S, b->ord_sh, b->ord_nbz);
    int rc;
    if ((rc = hfx_scan_offsets(b->ctx, b->ord_off, b->ord_cnt, b->ord_nb, (void *)st))) return rc;
    hipLaunchKernelGGL(k_ord_scatter, dim3(g), dim3(256), 0, st, b->ord_tmp, b->ord_cnt,
                       b->ord_off, b->begin, b->sp.d_zmax, S, b->ord_sh, b->ord_nbz);
    hipLaunchKernelGGL(k_ord_rank, dim3(g), dim3(256), 0, st, b->order, b->ord_tmp, b->ord_off,
                       b->begin, b->sp.d_zmax, b->sp.d_mass, S, b->ord_sh, b->ord_nbz);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

/* dense blocks (segment DP; eight spectra a block in the BF16 form, sixteen
 * in the integer form) and the spectra's hi/lo rows or digits */
static int bt_dense_prep(hfx_batch *b, hipStream_t st)
{
    const uint32_t S = b->S;
    const bool i8 = b->precision == HFX_BATCH_I8X3_DENSE;
    if (i8)
        hipLaunchKernelGGL(k_cut_seg<16>, dim3((b->nseg + CUT_SPW - 1) / CUT_SPW), dim3(CUT_WG), 0,
                           st, b->cut_seg, b->cut_n, b->order, b->begin, b->num, b->sp.d_zmax, S);
    else
        hipLaunchKernelGGL(k_cut_seg<8>, dim3((b->nseg + CUT_SPW - 1) / CUT_SPW), dim3(CUT_WG), 0,
                           st, b->cut_seg, b->cut_n, b->order, b->begin, b->num, b->sp.d_zmax, S);
    int rc;
    if ((rc = hfx_scan_offsets(b->ctx, b->cut_off, b->cut_n, b->nseg, (void *)st))) return rc;
    const uint32_t nt = std::max(b->nseg * CUT_SEG, S + 1);
    hipLaunchKernelGGL(k_cut_place, dim3((nt + 255) / 256), dim3(256), 0, st, b->dn_bstart,
                       b->cut_seg, b->cut_n, b->cut_off, b->nseg, S, S);
    HFX_CHECK_LAUNCH();
    return i8 ? hfx_dense8_split(&b->sp, b->dn_scratch, (void *)st)
              : hfx_dense_split(&b->sp, b->dn_scratch, (void *)st);
}

/* blocks of G spectra in order + tiles for the current (G, C) */
static int bt_tiles(hfx_batch *b, hipStream_t st)
{
    b->B = (b->S + b->G - 1) / b->G;
    if (b->B == 0) return HFX_OK;
    hipLaunchKernelGGL(k_bt_blocks, dim3((b->B + 255) / 256), dim3(256), 0, st, b->ubeg, b->uend,
                       b->ntiles, b->order, b->begin, b->num, b->S, b->G, b->B, b->C);
    return hfx_scan_offsets(b->ctx, b->tile_off, b->ntiles, b->B, (void *)st);
}

static int bt_prep(hfx_batch *b, hipStream_t st)
{
    int rc;
    if ((rc = bt_order(b, st))) return rc;
    if (bt_is_dense(b->precision)) return bt_dense_prep(b, st);
    return bt_tiles(b, st);
}

static int bt_score(hfx_batch *b, hipStream_t st, void **ev)
{
    if (b->B == 0 || b->total == 0) return HFX_OK;
    const hfx_pepix *ix = b->ix;
    hipLaunchKernelGGL(k_bt_reset, dim3(1), dim3(1), 0, st, b->counter);
    BtArgs a{b->scores, b->score_off, b->begin,   b->num,     b->order,   b->S,
             b->G,      b->B,         b->C,       b->ubeg,    b->uend,    b->tile_off,
             ix->ent,   ix->cnt,      ix->ent_off, 0u,        b->db.num_pep, ix->L,
             b->lstr,   b->counter,   b->sp};
    const size_t smem = (size_t)b->G * b->lstr * 4;
    const uint32_t grid = (uint32_t)hfx_ctx_num_cus(b->ctx) * (b->G == 8 ? 2u : 4u) *
                          (b->threads == 256 ? 2u : 1u);
    /* gfx950 gives a workgroup up to 160 KiB of LDS; ask for what this
     * configuration stages (G spectra of lstr floats) */
    if (b->G == 8) {
        HFX_CHECK_HIP(hipFuncSetAttribute((const void *)k_batch<8>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        if (ev) HFX_CHECK_HIP(hipEventRecord((hipEvent_t)ev[0], st));
        hipLaunchKernelGGL(k_batch<8>, dim3(grid), dim3(b->threads), smem, st, a);
    } else {
        HFX_CHECK_HIP(hipFuncSetAttribute((const void *)k_batch<4>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        if (ev) HFX_CHECK_HIP(hipEventRecord((hipEvent_t)ev[0], st));
        hipLaunchKernelGGL(k_batch<4>, dim3(grid), dim3(b->threads), smem, st, a);
    }
    if (ev) HFX_CHECK_HIP(hipEventRecord((hipEvent_t)ev[1], st));
    hipLaunchKernelGGL(k_bt_zero_unscored, dim3(b->S), dim3(256), 0, st, b->scores, b->score_off,
                       b->sp.d_zmax, b->S);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

/* ev (optional): events recorded around the scoring kernel itself */
static int bt_scores(hfx_batch *b, hipStream_t st, void **ev = nullptr)
{
    if (b->precision == HFX_BATCH_I8X3_DENSE) {
        if (b->total == 0) return HFX_OK;
        return hfx_dense8_scores(b->ctx, b->scores, b->score_off, b->begin, b->num, b->order,
                                 b->dn_bstart, b->S, &b->ix->lad, &b->sp, b->CH, b->dn_scratch, 1,
                                 (void *)st, ev);
    }
    if (b->precision == HFX_BATCH_BF16X2_DENSE) {
        if (b->total == 0) return HFX_OK;
        return hfx_dense_scores(b->ctx, b->scores, b->score_off, b->begin, b->num, b->order,
                                b->dn_bstart, b->S, &b->ix->lad, &b->sp, b->CH, b->dn_scratch, 1,
                                (void *)st, ev);
    }
    return bt_score(b, st, ev);
}

/* |dense - exact| <= (2^-16 + 210 * 2^-24) * 50 * sum|x|: the hi/lo split
 * (2^-16 per bin) plus at most 100 FP32 additions on any term's path (32 in
 * one MFMA, one per k-step after it) and the hi + lo sum, each allowed 2 ulp */
#define BT_REL_DENSE (1.52587890625e-05f + 210.0f * 5.9604644775390625e-08f)
#define BT_REL_F32 (80.0f * 5.9604644775390625e-08f)
/* integer form: |score - exact| <= max|x| / 8355711 * 50 len (the rounding of
 * x 2^k to integers) + 2^-24 * 50 sum|x| (the one FP32 rounding) */
#define BT_REL_I8 5.9604644775390625e-08f
#define BT_RELQ_I8 (1.0f / 8355711.0f)

static int bt_topn(hfx_batch *b, hipStream_t st)
{
    const bool i8 = b->precision == HFX_BATCH_I8X3_DENSE;
    const float rel0 = i8 ? BT_REL_I8 : b->precision == HFX_BATCH_BF16X2_DENSE ? BT_REL_DENSE
                                                                             : BT_REL_F32;
    return hfx_topn_device(b->ctx, b->top_pos, b->top_exact, b->top_f32, b->top_cnt, b->flag,
                           b->N, b->margin, b->scores, b->score_off, b->total, b->begin, b->num,
                           &b->db, &b->sp, rel0, 0.0f, i8 ? BT_RELQ_I8 : 0.0f, b->tx_scratch,
                           (void *)st);
}

extern "C" int hfx_batch_run(hfx_batch *b, void *stream)
{
    HFX_REQUIRE(b && b->planned, HFX_ERR_ARG, "hfx_batch_run: batch not planned");
    hipStream_t st = hfx_stream(stream);
    int rc;
    if (b->S == 0) return HFX_OK;
    if ((rc = bt_need_index(b, st))) return rc;
    if ((rc = bt_windows(b, st)) || (rc = bt_prep(b, st)) || (rc = bt_scores(b, st))) return rc;
    return bt_topn(b, st);
}

extern "C" int hfx_batch_set_precision(hfx_batch *b, uint32_t precision)
{
    HFX_REQUIRE(b && b->planned, HFX_ERR_ARG, "hfx_batch_set_precision: batch not planned");
    HFX_REQUIRE(precision == HFX_BATCH_FP32_SPARSE || bt_is_dense(precision), HFX_ERR_ARG,
                "hfx_batch_set_precision: unknown precision %u", precision);
    HFX_REQUIRE(!bt_is_dense(precision) || b->sp.x_stride <= hfx_dense_max_len(),
                HFX_ERR_ARG,
                "hfx_batch_set_precision: the dense form takes spectra of up to %u bins "
                "(x_stride %u)", hfx_dense_max_len(), b->sp.x_stride);
    int rc;
    if (bt_is_dense(precision) && (rc = bt_dense_alloc(b))) return rc;
    if (bt_is_dense(precision) && precision != b->precision) {
        /* the tuned chunk of this form for batches of this shape, if any */
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load_locked();
        auto &m = bt_tune_dense(precision);
        auto jt = m.find(tune_key(b));
        b->CH = jt != m.end() ? jt->second : 256;
    }
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
    int rc;
    if ((rc = bt_need_index(b, st))) return rc;
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    /* kernel attributes are set outside the capture */
    HFX_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    rc = hfx_batch_run(b, stream);
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

/* Times the scoring kernel of the planned batch for each configuration in a
 * small grid and keeps the fastest; the choice is remembered for batches of
 * the same shape bucket in this process.  Sparse form: (G, C, threads);
 * dense form: the chunk of candidates per work item. */
extern "C" int hfx_batch_autotune(hfx_batch *b, void *stream, uint32_t *out_cfg)
{
    HFX_REQUIRE(b && b->planned, HFX_ERR_ARG, "hfx_batch_autotune: batch not planned");
    hipStream_t st = hfx_stream(stream);
    int rc;
    /* index, windows and order once */
    if ((rc = hfx_batch_run(b, stream))) return rc;
    hipEvent_t e0, e1;
    HFX_CHECK_HIP(hipEventCreate(&e0));
    HFX_CHECK_HIP(hipEventCreate(&e1));
    /* a warm run, then the fastest of three timed runs (the minimum is the
     * least disturbed by clocks and other work) */
    auto time2 = [&](float *ms) -> int {
        int r;
        if ((r = bt_prep(b, st)) || (r = bt_scores(b, st))) return r; /* warm */
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
    if (bt_is_dense(b->precision)) {
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
        tune_load_locked();
        if (bt_is_dense(b->precision)) bt_tune_dense(b->precision)[tune_key(b)] = bCH;
        else g_tune[tune_key(b)] = std::make_tuple(bG, bC, bT);
        tune_save_locked();
    }
    if (b->exec) {
        /* the captured graph holds the old launch shapes */
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    /* leave the batch's outputs those of the chosen configuration */
    if ((rc = hfx_batch_run(b, stream))) return rc;
    if (out_cfg) {
        out_cfg[0] = bG; out_cfg[1] = bC; out_cfg[2] = bT;
        if (bt_is_dense(b->precision)) {
            out_cfg[0] = 0;
            out_cfg[1] = bCH;
            out_cfg[2] = 1024u; /* workgroup threads of both dense kernels */
        }
    }
    return HFX_OK;
}

/* Runs the planned batch phase by phase with events between the phases:
 * ms[0] windows + offsets, ms[1] per-batch preparation (spectrum order;
 * dense form: block cut and hi/lo rows; sparse form: tiles), ms[2] scoring,
 * ms[3] top-N (+ exact route). */
extern "C" int hfx_batch_time(hfx_batch *b, void *stream, float *ms)
{
    HFX_REQUIRE(b && b->planned && ms, HFX_ERR_ARG, "hfx_batch_time: batch not planned");
    hipStream_t st = hfx_stream(stream);
    int rc = HFX_OK;
    for (int i = 0; i < 5; ++i) ms[i] = 0.0f;
    if (b->S == 0) return HFX_OK;
    if ((rc = bt_need_index(b, st))) return rc;
    hipEvent_t ev[7];
    for (auto &e : ev) HFX_CHECK_HIP(hipEventCreate(&e));
    void *kev[2] = {ev[5], ev[6]};
    HFX_CHECK_HIP(hipEventRecord(ev[0], st));
    if (!rc) rc = bt_windows(b, st);
    HFX_CHECK_HIP(hipEventRecord(ev[1], st));
    if (!rc) rc = bt_prep(b, st);
    HFX_CHECK_HIP(hipEventRecord(ev[2], st));
    const bool scored = b->total != 0 && (bt_is_dense(b->precision) || b->B != 0);
    if (!rc) rc = bt_scores(b, st, kev);
    HFX_CHECK_HIP(hipEventRecord(ev[3], st));
    if (!rc) rc = bt_topn(b, st);
    HFX_CHECK_HIP(hipEventRecord(ev[4], st));
    HFX_CHECK_HIP(hipEventSynchronize(ev[4]));
    for (int i = 0; i < 4; ++i) HFX_CHECK_HIP(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
    if (scored && !rc) HFX_CHECK_HIP(hipEventElapsedTime(&ms[4], ev[5], ev[6]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

extern "C" int hfx_batch_set_config(hfx_batch *b, uint32_t G, uint32_t C, uint32_t threads)
{
    HFX_REQUIRE(b && b->planned, HFX_ERR_ARG, "hfx_batch_set_config: batch not planned");
    if (bt_is_dense(b->precision)) {
        /* dense forms: C = candidates per work item (G, threads fixed) */
        const uint32_t step = tune_dense_step(b->precision);
        HFX_REQUIRE(tune_dense_ok(b->precision, C), HFX_ERR_ARG,
                    "hfx_batch_set_config: dense chunk must be a multiple of %u in %u..8192",
                    step, step);
        b->CH = C;
        return HFX_OK;
    }
    HFX_REQUIRE(tune_sparse_ok(G, C, threads), HFX_ERR_ARG,
                "hfx_batch_set_config: G in {4,8}, C in 1..4096, threads in {256,512}");
    b->G = G;
    b->C = C;
    b->threads = threads;
    return HFX_OK;
}

extern "C" int hfx_batch_dense_blocks(const hfx_batch *b, uint32_t *h_bstart, uint32_t cap,
                                      uint32_t *num_blocks)
{
    HFX_REQUIRE(b && num_blocks, HFX_ERR_ARG, "hfx_batch_dense_blocks: null argument");
    HFX_REQUIRE(b->planned && b->dn_scratch, HFX_ERR_ARG,
                "hfx_batch_dense_blocks: no dense run of this batch");
    uint64_t nb = 0;
    HFX_CHECK_HIP(hipDeviceSynchronize());
    HFX_CHECK_HIP(hipMemcpy(&nb, b->cut_off + b->nseg, 8, hipMemcpyDeviceToHost));
    *num_blocks = (uint32_t)nb;
    if (h_bstart && cap) {
        const size_t n = std::min<size_t>(cap, (size_t)nb + 1);
        HFX_CHECK_HIP(hipMemcpy(h_bstart, b->dn_bstart, 4 * n, hipMemcpyDeviceToHost));
    }
    return HFX_OK;
}

extern "C" int hfx_batch_order(const hfx_batch *b, uint32_t *h_order)
{
    HFX_REQUIRE(b && b->planned && h_order, HFX_ERR_ARG, "hfx_batch_order: batch not planned");
    HFX_CHECK_HIP(hipDeviceSynchronize());
    if (b->S) HFX_CHECK_HIP(hipMemcpy(h_order, b->order, 4ull * b->S, hipMemcpyDeviceToHost));
    return HFX_OK;
}

extern "C" int hfx_batch_info(const hfx_batch *b, uint64_t *info /*[12]*/)
{
    HFX_REQUIRE(b && info, HFX_ERR_ARG, "hfx_batch_info: null argument");
    const hfx_pepix *ix = b->ix;
    info[0] = b->S; info[1] = b->total; info[2] = b->db.num_pep; info[3] = b->L;
    info[4] = ix ? ix->n_ent : 0; info[5] = b->G; info[6] = b->C; info[7] = b->threads;
    info[8] = b->lstr; info[9] = 0; info[10] = b->precision; info[11] = b->CH;
    return HFX_OK;
}

extern "C" int hfx_batch_outputs(const hfx_batch *b, void **ptrs /*[10]*/)
{
    HFX_REQUIRE(b && ptrs, HFX_ERR_ARG, "hfx_batch_outputs: null argument");
    ptrs[0] = b->top_pos; ptrs[1] = b->top_exact; ptrs[2] = b->top_f32; ptrs[3] = b->top_cnt;
    ptrs[4] = b->flag; ptrs[5] = b->scores; ptrs[6] = b->score_off; ptrs[7] = b->begin;
    ptrs[8] = b->num; ptrs[9] = b->ix ? (void *)b->ix->ent : nullptr;
    return HFX_OK;
}
