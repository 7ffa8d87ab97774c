/*
 * hfx_xcorr.hip -- plain FP32 sparse XCorr, dense drop-in rows, row MVM.
 *
 * The reference scores a spectrum by materialising a dense theoretical row
 * per candidate (addIons, ion_series.cu:150-232: memset + global atomicMax +
 * int->float rewrite) and multiplying with cuBLAS sgemv (Search.hs:242-245).
 * Here the theoretical peaks of one candidate live in a 1 KiB-per-wave LDS
 * nibble map (one nibble per bin, max per bin through a unary weight code)
 * and are dotted straight against the spectrum held in LDS: no dense row ever
 * reaches HBM.
 *
 * Work decomposition: one workgroup (4 waves) per (spectrum, chunk of
 * XC_CHUNK candidates) tile; the spectrum is staged in LDS once per tile; each
 * wave scores one candidate at a time (ladder -> map -> dot).
 */
#include "hfx_cand.h"

#include <stdlib.h>

#include <mutex>
#include <type_traits>

#define XC_THREADS 256
#define XC_WAVES (XC_THREADS / HFX_WAVE)
#define XC_CHUNK 64

/* largest s with off[s] <= t, off has n+1 entries, off[0] = 0 */
__device__ __forceinline__ uint32_t find_segment(const uint64_t *off, uint32_t n, uint64_t t)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

/* ------------------------------------------------------------------ */
/* Plain candidates, packed: several candidates per wave                 */
/* ------------------------------------------------------------------ */

/* One workgroup (four waves) per tile of XP_CHUNK candidates of one
 * spectrum; each wave takes a quarter of the tile and packs up to XP_SLOTS
 * consecutive candidates of at most 64 ladder positions in all into its 64
 * lanes, one lane per (candidate, ladder position):
 *   - the b ladder is the sequential FP32 prefix of each candidate's residue
 *     masses: at step j every lane at position j adds its residue mass to its
 *     left neighbour's finished sum (a DPP shift by one lane), so lane k
 *     holds ((m_0 + m_1) + ...) + m_k, the association of a sequential loop;
 *   - each lane ORs the unary weight codes of its peaks into its candidate's
 *     nibble map (OR of codes = code of the maximum weight), then claims
 *     them back with an atomic AND that clears them: the first claimant of a
 *     bin receives its final code and every bin counts once, so no lane
 *     scans the whole map, and the map is clean for the next pack;
 *   - the lanes' partial sums are reduced per candidate.
 * A candidate of more than 64 positions takes a wave of its own through the
 * general path (hfx_wave_peaks, hfx_wave_dot_f32). */
#define XP_THREADS 256
#define XP_WAVES (XP_THREADS / HFX_WAVE)
#define XP_CHUNK 128
#define XP_RANGE (XP_CHUNK / XP_WAVES) /* candidates a wave takes from a tile */
#define XP_SLOTS 4
/* b-ladder positions a wave's buffer holds (a chunk of its candidates at a
 * time; any one packable candidate has at most 64) */
#define XP_LB 256
/* Peaks outside [0, len) are clamped onto margin bins whose spectrum values
 * are zero (so they add 0 when claimed): singles onto bin -1 - 8 (lane & 7)
 * or len + 8 (lane & 7), triple centres onto -2 - 8 (lane & 7) or
 * len + 1 + 8 (lane & 7) (none of a clamped triple's nibbles is inside
 * [0, len)), spread over eight words a side so that they do not queue on
 * one address.  Clamped nibbles lie in [-59, len + 58]; the staged spectrum
 * is zero from bin -XP_PAD to len + XP_PAD and the map words cover bins
 * -XP_PAD .. xstr + XP_PAD. */
#define XP_SPREAD 8
#define XP_PAD (8 * XP_SPREAD)
static_assert(XP_RANGE <= HFX_WAVE, "a wave's candidates fit its lanes");
static_assert(XP_PAD % 8 == 0 && XP_PAD >= 8 * (XP_SPREAD - 1) + 3, "clamped peaks inside the margins");

#define XP_MAGIC 12582912.0f /* 1.5 * 2^23: x + XP_MAGIC rounds x to an integer, ties to even */
#define XP_MAGIC_BITS 0x4B400000u

/* map words per candidate for spectra of at most xstr bins: bin x in word
 * (x + XP_PAD) / 8, x in [-XP_PAD, xstr + XP_PAD) */
static uint32_t xp_map_words(uint32_t xstr) { return (xstr + 2 * XP_PAD + 7) / 8 + 1; }
static size_t xp_lds_bytes(uint32_t xstr)
{
    const size_t xf = ((size_t)xstr + 2 * XP_PAD + 3) & ~(size_t)3;
    return xf * 4 + (size_t)XP_WAVES * XP_SLOTS * xp_map_words(xstr) * 4 +
           (size_t)XP_WAVES * XP_LB * 4;
}

/* rint(mz / 1.0005079) as the bits of the integer-valued float q + 1.5 2^23
 * (q the correctly rounded quotient): u - XP_MAGIC_BITS is the bin for
 * |bin| < 2^22 */
__device__ __forceinline__ uint32_t xp_binu(float mz)
{
    const float R = 1.0f / HFX_BIN_WIDTH_MONO;
    return __float_as_uint(hfx_div(mz, HFX_BIN_WIDTH_MONO, R) + XP_MAGIC);
}

/* division by the fragment charge, correctly rounded (exact products for
 * 1, 2, 4) */
template <int Z>
__device__ __forceinline__ float xp_zdiv(float x)
{
    if constexpr (Z == 3) return hfx_div(x, 3.0f, 1.0f / 3.0f);
    else return x * (1.0f / (float)Z);
}

/* weight of a nibble holding the OR of unary codes: 0, 1, 3, 7 -> 0, 10,
 * 25, 50, one byte lookup (v_perm: selector c picks byte c of the pool
 * 0x32000000:19000A00) */
__device__ __forceinline__ float xp_weight(uint32_t nib)
{
    return (float)(__builtin_amdgcn_perm(0x32000000u, 0x19000A00u, nib) & 0xffu);
}

/* median of three unsigned (one v_med3_u32; the compiler does not fold a
 * min of a max with non-constant bounds) */
__device__ __forceinline__ uint32_t xp_med3(uint32_t v, uint32_t lo, uint32_t hi)
{
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(v), "v"(lo), "v"(hi));
    return r;
}

/* the six bins of one ladder position at charge Z, in the FP32 operation
 * order of ion_series.cu:111-156 (addIonsAB, addIonsY), as xp_binu bits
 * (XP_MAGIC_BITS + bin) clamped into [lo, hi] (singles: a, b - H2O,
 * b - NH3, y - NH3) and [lo3, hi3] (triple centres b, y) */
struct XpClamp {
    uint32_t lo, hi, lo3, hi3;
};
template <int Z>
__device__ __forceinline__ void xp_bins(float b, float y, const XpClamp &c, uint32_t (&u)[6])
{
    constexpr float zh = HFX_DEV_H * (float)Z;
    constexpr float h2o = HFX_DEV_H2O / (float)Z, nh3 = HFX_DEV_NH3 / (float)Z;
    const float ma = xp_zdiv<Z>((b - HFX_DEV_CO) + zh);
    const float mb = xp_zdiv<Z>(b + zh);
    const float my = xp_zdiv<Z>((y + HFX_DEV_H2O) + zh);
    u[0] = xp_med3(xp_binu(ma), c.lo, c.hi);
    u[1] = xp_med3(xp_binu(mb), c.lo3, c.hi3);
    u[2] = xp_med3(xp_binu(mb - h2o), c.lo, c.hi);
    u[3] = xp_med3(xp_binu(mb - nh3), c.lo, c.hi);
    u[4] = xp_med3(xp_binu(my), c.lo3, c.hi3);
    u[5] = xp_med3(xp_binu(my - nh3), c.lo, c.hi);
}

/* the words and nibble shifts of one position's peaks: singles at word
 * w[i], triples over words w[i], w[i] + 1 from nibble shift sh[i]; u[i] the
 * bits of the single's bin or of the bin before the triple's centre.  rowm
 * and xsm are the map row and the staged spectrum shifted back by
 * XP_MAGIC_BITS bins (never dereferenced themselves): the magic constant's
 * low three bits are zero, so u >> 3 and u & 7 give the word and nibble */
struct XpPeaks {
    uint32_t *w[6];
    uint32_t sh[6];
    uint32_t u[6];
};

__device__ __forceinline__ void xp_locate(uint32_t *rowm, const uint32_t (&u)[6], XpPeaks &pk)
{
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const uint32_t ui = (i == 1 || i == 4) ? u[i] - 1u : u[i];
        pk.u[i] = ui;
        pk.w[i] = rowm + (ui >> 3);
        pk.sh[i] = (ui & 7u) * 4u;
    }
}

/* ORs the peaks of one position into the candidate's map: the four
 * weight-10 peaks alone, each 25/50/25 triple as one 12-bit field over two
 * words */
__device__ __forceinline__ void xp_put(const XpPeaks &pk)
{
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        if (i == 1 || i == 4) {
            const uint64_t t = (uint64_t)(HFX_CODE25 | HFX_CODE50 << 4 | HFX_CODE25 << 8)
                               << pk.sh[i];
            atomicOr(pk.w[i], (uint32_t)t);
            atomicOr(pk.w[i] + 1, (uint32_t)(t >> 32));
        } else {
            atomicOr(pk.w[i], HFX_CODE10 << pk.sh[i]);
        }
    }
}

/* claims the peaks of one position (clearing them) and adds weight * x of
 * every nibble it receives (clamped peaks sit on zero spectrum values) */
__device__ __forceinline__ float xp_claim(const XpPeaks &pk, const float *xsm, float acc)
{
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const uint32_t sh = pk.sh[i];
        if (i == 1 || i == 4) {
            const uint64_t t = (uint64_t)0xfffu << sh;
            const uint32_t o0 = atomicAnd(pk.w[i], ~(uint32_t)t);
            const uint32_t o1 = atomicAnd(pk.w[i] + 1, ~(uint32_t)(t >> 32));
            const uint32_t f = (uint32_t)(((uint64_t)o1 << 32 | o0) >> sh);
            const float *xv = xsm + pk.u[i];
            acc = __builtin_fmaf(xp_weight(f & 15u), xv[0], acc);
            acc = __builtin_fmaf(xp_weight((f >> 4) & 15u), xv[1], acc);
            acc = __builtin_fmaf(xp_weight((f >> 8) & 15u), xv[2], acc);
        } else {
            const uint32_t old = atomicAnd(pk.w[i], ~(0xfu << sh));
            acc = __builtin_fmaf(xp_weight((old >> sh) & 15u), xsm[pk.u[i]], acc);
        }
    }
    return acc;
}

/* v + (the value of the lane DPP control CTRL selects; 0 where none, and in
 * rows outside ROWS) */
template <int CTRL, int ROWS>
__device__ __forceinline__ float xp_dpp_add(float v)
{
    return v + __uint_as_float((uint32_t)__builtin_amdgcn_update_dpp(
                   0, (int)__float_as_uint(v), CTRL, ROWS, 0xf, true));
}

/* Inclusive segmented sum over the wave: lane l ends with the sum of the
 * lanes of its segment from the segment's first lane (kpos = 0) to l, added
 * in a fixed tree (at most six additions on any term's path).  Within rows
 * of 16 lanes by row shifts of 1, 2, 4, 8, then the last lanes of rows 0
 * and 2 into rows 1 and 3 and the last lane of row 1 into rows 2 and 3 (the
 * wave scan of gfx9 DPP), each step taken only where the source lane lies in
 * the same segment. */
__device__ __forceinline__ float xp_segsum(float v, uint32_t kpos, uint32_t lane)
{
    const uint32_t r = lane & 15u;
    float t;
    t = xp_dpp_add<0x111 /* row_shr:1 */, 0xf>(v);
    v = kpos >= 1 && r >= 1 ? t : v;
    t = xp_dpp_add<0x112 /* row_shr:2 */, 0xf>(v);
    v = kpos >= 2 && r >= 2 ? t : v;
    t = xp_dpp_add<0x114 /* row_shr:4 */, 0xf>(v);
    v = kpos >= 4 && r >= 4 ? t : v;
    t = xp_dpp_add<0x118 /* row_shr:8 */, 0xf>(v);
    v = kpos >= 8 && r >= 8 ? t : v;
    /* lane 15 of rows 0 and 2 into rows 1 and 3 */
    t = xp_dpp_add<0x142 /* row_bcast:15 */, 0xa>(v);
    v = (lane & 16u) && kpos > r ? t : v;
    /* lane 31 into rows 2 and 3 */
    t = xp_dpp_add<0x143 /* row_bcast:31 */, 0xc>(v);
    v = lane >= 32u && kpos > lane - 32u ? t : v;
    return v;
}

/* a pack: K candidates, candidate g on lanes [st_g, st_{g+1}), first
 * residue tc_g, residual r_g; wave-uniform values (no arrays: a lane picks
 * its candidate's by comparisons) */
struct XpPack {
    uint32_t K;
    uint32_t st1, st2, st3, tot;
    uint32_t lb0, lb1, lb2, lb3; /* first ladder position of each in the wave's buffer */
    float r0, r1, r2, r3;
};

/* one pack at fragment charges 1..Z; the candidates' b ladders are in lb */
template <int Z>
__device__ __forceinline__ void xp_pack(uint32_t *maps, uint32_t mstr, const float *xs,
                                        const float *lb, uint32_t len, const XpPack pk,
                                        float *out)
{
    const uint32_t lane = hfx_lane();
    const bool g1 = lane >= pk.st1, g2 = lane >= pk.st2, g3 = lane >= pk.st3;
    const uint32_t slot = (uint32_t)g1 + (uint32_t)g2 + (uint32_t)g3;
    const bool act = lane < pk.tot;
    const uint32_t k0 = g3 ? pk.st3 : g2 ? pk.st2 : g1 ? pk.st1 : 0u;
    const uint32_t k1 = g3 ? pk.tot : g2 ? min(pk.st3, pk.tot) : g1 ? min(pk.st2, pk.tot)
                                                                   : min(pk.st1, pk.tot);
    const uint32_t l0 = g3 ? pk.lb3 : g2 ? pk.lb2 : g1 ? pk.lb1 : pk.lb0;
    const float rl = g3 ? pk.r3 : g2 ? pk.r2 : g1 ? pk.r1 : pk.r0;
    const uint32_t kpos = lane - k0;
    const float b = act ? lb[l0 + kpos] : 0.0f;
    const float y = rl - b;
    /* bin x of this lane's candidate map is in word rowp[x >> 3]; rowm and
     * xsm index by the xp_binu bits instead */
    uint32_t *rowm = maps + slot * mstr + XP_PAD / 8 - (XP_MAGIC_BITS >> 3);
    const float *xsm = xs - XP_MAGIC_BITS;
    const uint32_t sp = (lane & (XP_SPREAD - 1)) * 8u;
    const XpClamp cl{XP_MAGIC_BITS - 1u - sp, XP_MAGIC_BITS + len + sp, XP_MAGIC_BITS - 2u - sp,
                     XP_MAGIC_BITS + len + 1u + sp};
    /* charges 1..Z: the bins of each (and, for Z <= 2, their words and
     * shifts) kept in registers from the OR pass to the claim pass */
    uint32_t xz[Z][6];
    float acc = 0.0f;
    if constexpr (Z <= 2) {
        XpPeaks pks[Z];
        if (act) {
            xp_bins<1>(b, y, cl, xz[0]);
            xp_locate(rowm, xz[0], pks[0]);
            if constexpr (Z >= 2) { xp_bins<2>(b, y, cl, xz[1]); xp_locate(rowm, xz[1], pks[1]); }
#pragma unroll
            for (int z = 0; z < Z; ++z) xp_put(pks[z]);
        }
        hfx_wave_sync();
        if (act) {
#pragma unroll
            for (int z = 0; z < Z; ++z) acc = xp_claim(pks[z], xsm, acc);
        }
    } else {
        if (act) {
            xp_bins<1>(b, y, cl, xz[0]);
            xp_bins<2>(b, y, cl, xz[1]);
            xp_bins<3>(b, y, cl, xz[2]);
            if constexpr (Z >= 4) xp_bins<4>(b, y, cl, xz[3]);
#pragma unroll
            for (int z = 0; z < Z; ++z) {
                XpPeaks pk;
                xp_locate(rowm, xz[z], pk);
                xp_put(pk);
            }
        }
        hfx_wave_sync();
        if (act) {
#pragma unroll
            for (int z = 0; z < Z; ++z) {
                XpPeaks pk;
                xp_locate(rowm, xz[z], pk);
                acc = xp_claim(pk, xsm, acc);
            }
        }
    }
    const float sc = xp_segsum(acc, kpos, lane);
    hfx_wave_sync();
    /* the last lane of each candidate holds its score; a candidate without
     * positions scores 0 */
    if (act && lane + 1 == k1) out[slot] = sc;
    if (lane < pk.K) {
        const uint32_t a = lane == 0 ? 0u : lane == 1 ? pk.st1 : lane == 2 ? pk.st2 : pk.st3;
        const uint32_t e = lane == 0 ? min(pk.st1, pk.tot) : lane == 1 ? min(pk.st2, pk.tot)
                         : lane == 2 ? min(pk.st3, pk.tot) : pk.tot;
        if (a == e) out[lane] = 0.0f;
    }
}

template <class Cands>
__global__ void __launch_bounds__(XP_THREADS)
    k_xcorr_pack(float *scores, const uint64_t *tile_off, Cands c, hfx_spectra sp, uint32_t mstr)
{
    extern __shared__ __align__(16) float xp_lds[];
    __shared__ float mtab[32];
    const uint32_t S = sp.num_spec;
    const uint64_t t = blockIdx.x;
    if (t >= tile_off[S]) return;
    uint32_t lo = 0, hi = S;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tile_off[mid] <= t) lo = mid; else hi = mid;
    }
    const uint32_t s = lo;
    const uint32_t first = (uint32_t)(t - tile_off[s]) * XP_CHUNK;
    const uint32_t cnt = min((uint32_t)XP_CHUNK, c.count(s) - first);
    const uint32_t xstr = sp.x_stride;
    const uint32_t xf = (xstr + 2 * XP_PAD + 3) & ~3u;
    float *xs = xp_lds + XP_PAD;
    uint32_t *maps_all = (uint32_t *)(xp_lds + xf);
    const uint32_t lane = hfx_lane(), wave = threadIdx.x / HFX_WAVE;
    uint32_t *maps = maps_all + wave * XP_SLOTS * mstr;
    float *lb = (float *)(maps_all + XP_WAVES * XP_SLOTS * mstr) + wave * XP_LB;
    const uint32_t len = hfx_stage_spectrum(xs, mtab, c.db, sp, s);
    /* zeros around the spectrum: clamped peaks read them */
    for (uint32_t i = len + threadIdx.x; i < xf - XP_PAD; i += XP_THREADS) xs[i] = 0.0f;
    for (uint32_t i = threadIdx.x; i < XP_PAD; i += XP_THREADS) xp_lds[i] = 0.0f;
    for (uint32_t i = lane; i < XP_SLOTS * mstr; i += HFX_WAVE) maps[i] = 0u;
    __syncthreads();
    const uint32_t zeff = c.zeff(sp.d_zmax[s]);
    /* this wave's candidates [c0, c0 + rn) of the tile */
    const uint32_t c0 = wave * XP_RANGE;
    const uint32_t rn = __builtin_amdgcn_readfirstlane(
        c0 < cnt ? min((uint32_t)XP_RANGE, cnt - c0) : 0u);
    float *out = scores + c.base(s) + first + c0;
    if (rn == 0) return;
    if (zeff == 0) {
        /* no fragment charge generated (ion_series.cu:334-342): score 0 */
        if (lane < rn) out[lane] = 0.0f;
        return;
    }
    /* lane i < rn: candidate c0 + i's first residue, positions, residual */
    typename Cands::Prep pr{};
    if (lane < rn) pr = c.prep(s, first + c0 + lane);
    /* the packed peaks serve fragment charges 1..4; above (modified
     * candidates only: addModIons goes to 10) every candidate takes the
     * general path */
    const bool general = zeff > 4;
    /* ladder positions of the packable candidates (long ones take none):
     * exclusive prefix over the wave's candidates */
    const uint32_t nl = lane < rn && pr.n <= HFX_WAVE && !general ? pr.n : 0u;
    uint32_t incl = nl;
#pragma unroll
    for (int o = 1; o < HFX_WAVE; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o, 64);
        if (lane >= (uint32_t)o) incl += v;
    }
    const uint32_t pofs = incl - nl;
    for (uint32_t p = 0; p < rn;) {
        /* a chunk [p, q) of candidates whose ladders fit the wave's buffer,
         * computed one candidate a lane: the sequential FP32 prefix of its
         * residue masses */
        const uint32_t base = hfx_take(pofs, p);
        const uint64_t fit = __ballot(lane >= p && lane < rn && incl - base <= XP_LB);
        const uint32_t q = p + (uint32_t)__builtin_popcountll(fit);
        if (lane >= p && lane < q && nl) c.ladder(pr, mtab, lb + (pofs - base), nl);
        hfx_wave_sync();
        while (p < q) {
            const uint32_t n0 = hfx_take(pr.n, p);
            if (n0 > HFX_WAVE || general) {
                /* a long candidate: the general path over chunks of 64 */
                const typename Cands::Prep qq = Cands::take(pr, p);
                c.peaks_prep(maps + XP_PAD / 8, mtab, qq, len, zeff);
                hfx_wave_sync();
                const float sc = hfx_wave_dot_f32(maps + XP_PAD / 8, xs, len);
                hfx_wave_sync();
                if (lane == 0) out[p] = sc;
                ++p;
                continue;
            }
            /* up to XP_SLOTS consecutive candidates of the chunk with at most
             * 64 positions in all (wave-uniform); lanes past q are not taken */
            const uint32_t n1 = hfx_take(pr.n, min(p + 1u, (uint32_t)HFX_WAVE - 1u));
            const uint32_t n2 = hfx_take(pr.n, min(p + 2u, (uint32_t)HFX_WAVE - 1u));
            const uint32_t n3 = hfx_take(pr.n, min(p + 3u, (uint32_t)HFX_WAVE - 1u));
            const bool t1 = p + 1 < q && n0 + n1 <= HFX_WAVE;
            const bool t2 = t1 && p + 2 < q && n0 + n1 + n2 <= HFX_WAVE;
            const bool t3 = t2 && p + 3 < q && n0 + n1 + n2 + n3 <= HFX_WAVE;
            XpPack pk;
            pk.K = 1u + (uint32_t)t1 + (uint32_t)t2 + (uint32_t)t3;
            pk.st1 = t1 ? n0 : 0xffffffffu;
            pk.st2 = t2 ? n0 + n1 : 0xffffffffu;
            pk.st3 = t3 ? n0 + n1 + n2 : 0xffffffffu;
            pk.tot = n0 + (t1 ? n1 : 0u) + (t2 ? n2 : 0u) + (t3 ? n3 : 0u);
            pk.lb0 = hfx_take(pofs, p) - base;
            pk.lb1 = hfx_take(pofs, min(p + 1u, (uint32_t)HFX_WAVE - 1u)) - base;
            pk.lb2 = hfx_take(pofs, min(p + 2u, (uint32_t)HFX_WAVE - 1u)) - base;
            pk.lb3 = hfx_take(pofs, min(p + 3u, (uint32_t)HFX_WAVE - 1u)) - base;
            pk.r0 = hfx_take(pr.r, p);
            pk.r1 = hfx_take(pr.r, min(p + 1u, (uint32_t)HFX_WAVE - 1u));
            pk.r2 = hfx_take(pr.r, min(p + 2u, (uint32_t)HFX_WAVE - 1u));
            pk.r3 = hfx_take(pr.r, min(p + 3u, (uint32_t)HFX_WAVE - 1u));
            switch (zeff) {
            case 1: xp_pack<1>(maps, mstr, xs, lb, len, pk, out + p); break;
            case 2: xp_pack<2>(maps, mstr, xs, lb, len, pk, out + p); break;
            case 3: xp_pack<3>(maps, mstr, xs, lb, len, pk, out + p); break;
            default: xp_pack<4>(maps, mstr, xs, lb, len, pk, out + p); break;
            }
            p += pk.K;
        }
        hfx_wave_sync(); /* the buffer is rewritten by the next chunk */
    }
}

template <class Cands>
__global__ void k_tile_count(uint32_t *tiles, Cands c, uint32_t num_spec, uint32_t chunk)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < num_spec) tiles[s] = (c.count(s) + chunk - 1) / chunk;
}

template <class Cands>
__global__ void __launch_bounds__(XC_THREADS)
    k_xcorr(float *scores, const uint64_t *tile_off, Cands c, hfx_spectra sp)
{
    __shared__ float xs[HFX_MAX_LEN];
    __shared__ uint32_t maps[XC_WAVES][HFX_MAP_WORDS];
    __shared__ float mtab[32];
    const uint32_t S = sp.num_spec;
    const uint64_t t = blockIdx.x;
    if (t >= tile_off[S]) return;
    const uint32_t s = find_segment(tile_off, S, t);
    const uint32_t first = (uint32_t)(t - tile_off[s]) * XC_CHUNK;
    const uint32_t cnt = min((uint32_t)XC_CHUNK, c.count(s) - first);
    const uint32_t wave = threadIdx.x / HFX_WAVE;
    uint32_t *map = maps[wave];
    const uint32_t len = hfx_stage_spectrum(xs, mtab, c.db, sp, s);
    for (uint32_t w = hfx_lane(); w < HFX_MAP_WORDS; w += HFX_WAVE) map[w] = 0;
    __syncthreads();
    const uint32_t zeff = c.zeff(sp.d_zmax[s]);
    float *out = scores + c.base(s) + first;
    for (uint32_t i = wave; i < cnt; i += XC_WAVES) {
        c.peaks(map, mtab, s, first + i, len, zeff);
        hfx_wave_sync();
        const float sc = hfx_wave_dot_f32(map, xs, len);
        hfx_wave_sync();
        if (hfx_lane() == 0) out[i] = sc;
    }
}

/* device scratch of the sparse XCorr for S spectra: per-spectrum tile counts
 * and their offsets */
static size_t xc_scratch_bytes(uint32_t S) { return (size_t)S * 4 + ((size_t)S + 1) * 8 + 64; }

template <class Cands>
static int launch_xcorr(void *scratch, float *d_scores, const Cands &c, uint64_t total,
                        const hfx_spectra *sp, void *stream)
{
    const uint32_t S = sp->num_spec;
    if (S == 0 || total == 0) return HFX_OK;
    HFX_REQUIRE(sp->x_stride <= HFX_MAX_LEN, HFX_ERR_ARG,
                "sparse XCorr: spectra of up to %d bins (x_stride %u)", HFX_MAX_LEN, sp->x_stride);
    /* unmodified and modified candidates take the packed kernel */
    constexpr bool PACK = std::is_same<Cands, HfxPlainCands>::value ||
                          std::is_same<Cands, HfxModCands>::value;
    const uint32_t chunk = PACK ? XP_CHUNK : XC_CHUNK;
    hipStream_t st = hfx_stream(stream);
    int rc;
    uint32_t *tiles = (uint32_t *)scratch;
    uint64_t *tile_off = (uint64_t *)(((uintptr_t)(tiles + S) + 15) & ~(uintptr_t)15);
    hipLaunchKernelGGL(k_tile_count<Cands>, dim3((S + 255) / 256), dim3(256), 0, st, tiles, c, S,
                       chunk);
    if ((rc = hfx_scan_offsets(nullptr, tile_off, tiles, S, stream)) != HFX_OK) return rc;
    /* sum_s ceil(n_s / chunk) <= total / chunk + S */
    const uint64_t grid = (total + chunk - 1) / chunk + S;
    HFX_REQUIRE(grid < (1ull << 31), HFX_ERR_ARG, "xcorr: %llu tiles", (unsigned long long)grid);
    if constexpr (PACK) {
        hipLaunchKernelGGL(k_xcorr_pack<Cands>, dim3((uint32_t)grid), dim3(XP_THREADS),
                           xp_lds_bytes(sp->x_stride), st, d_scores, tile_off, c, *sp,
                           xp_map_words(sp->x_stride));
    } else {
        hipLaunchKernelGGL(k_xcorr<Cands>, dim3((uint32_t)grid), dim3(XC_THREADS), 0, st,
                           d_scores, tile_off, c, *sp);
    }
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

extern "C" int hfx_xcorr_sparse(hfx_ctx *ctx, float *d_scores, const uint64_t *d_off,
                                const uint32_t *d_begin, const uint32_t *d_num, uint64_t total,
                                const hfx_db *db, const hfx_spectra *sp, void *stream)
{
    HFX_REQUIRE(ctx && d_scores && d_off && d_begin && d_num && db && sp, HFX_ERR_ARG,
                "hfx_xcorr_sparse: null argument");
    HfxPlainCands c{*db, d_begin, d_num, d_off};
    void *p = nullptr;
    int rc;
    if (sp->num_spec &&
        (rc = hfx_ctx_scratch(ctx, HFX_SLOT_TILES, xc_scratch_bytes(sp->num_spec), &p)) != HFX_OK)
        return rc;
    return launch_xcorr(p, d_scores, c, total, sp, stream);
}

extern "C" int hfx_xcorr_sparse_mod(hfx_ctx *ctx, float *d_scores, const uint64_t *d_mp_off,
                                    uint64_t total, const hfx_modset *ms, const hfx_db *db,
                                    const hfx_spectra *sp, void *stream)
{
    HFX_REQUIRE(ctx && d_scores && d_mp_off && ms && db && sp, HFX_ERR_ARG,
                "hfx_xcorr_sparse_mod: null argument");
    HFX_REQUIRE(ms->num_ma <= HFX_MAX_MA, HFX_ERR_ARG, "hfx_xcorr_sparse_mod: num_ma %u > %d",
                ms->num_ma, HFX_MAX_MA);
    HfxModCands c{*db, *ms, d_mp_off};
    void *p = nullptr;
    int rc;
    if (sp->num_spec &&
        (rc = hfx_ctx_scratch(ctx, HFX_SLOT_TILES, xc_scratch_bytes(sp->num_spec), &p)) != HFX_OK)
        return rc;
    return launch_xcorr(p, d_scores, c, total, sp, stream);
}

/* ------------------------------------------------------------------ */
/* hfx_xcorr_sparse for callers that cannot allocate                     */
/* ------------------------------------------------------------------ */

/* The tile table of up to HFX_NOALLOC_MAX_SPECTRA spectra lives in this
 * file-scope device array (12 MiB, held by every GPU the library runs on
 * until the process ends).  In host code `scratch` names a host-side stand-in,
 * never the device memory: the device address comes from
 * hipGetSymbolAddress, once per device, and only that address is passed on. */
#define NA_BYTES (((size_t)HFX_NOALLOC_MAX_SPECTRA * 4 + ((size_t)HFX_NOALLOC_MAX_SPECTRA + 1) * 8 + \
                   64 + 255) & ~(size_t)255)
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

#define NA_MAX_DEVICES 64
static std::mutex g_na_mu;
static void *g_na_ptr[NA_MAX_DEVICES];

/* the current device's address of `scratch`, looked up on the first call on
 * that device */
static int na_scratch(void **out)
{
    int dev = 0;
    HFX_CHECK_HIP(hipGetDevice(&dev));
    HFX_REQUIRE(dev >= 0 && dev < NA_MAX_DEVICES, HFX_ERR_ARG, "no-alloc scratch: device %d", dev);
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        HFX_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        HFX_REQUIRE(p != nullptr, HFX_ERR_HIP, "no-alloc scratch: no device address");
        g_na_ptr[dev] = p;
    }
    *out = g_na_ptr[dev];
    return HFX_OK;
}

extern "C" int hfx_xcorr_sparse_noalloc(float *d_scores, const uint64_t *d_off,
                                        const uint32_t *d_begin, const uint32_t *d_num,
                                        uint64_t total, const hfx_db *db, const hfx_spectra *sp,
                                        void *stream)
{
    HFX_REQUIRE(d_scores && d_off && d_begin && d_num && db && sp, HFX_ERR_ARG,
                "hfx_xcorr_sparse_noalloc: null argument");
    HFX_REQUIRE(sp->num_spec <= HFX_NOALLOC_MAX_SPECTRA, HFX_ERR_ARG,
                "hfx_xcorr_sparse_noalloc: %u spectra, at most %u in one call", sp->num_spec,
                (unsigned)HFX_NOALLOC_MAX_SPECTRA);
    static_assert(NA_BYTES >= ((size_t)HFX_NOALLOC_MAX_SPECTRA * 4 +
                               ((size_t)HFX_NOALLOC_MAX_SPECTRA + 1) * 8 + 64),
                  "no-alloc scratch holds the tile table of the largest call");
    if (sp->num_spec == 0 || total == 0) return HFX_OK;
    void *p = nullptr;
    int rc;
    if ((rc = na_scratch(&p))) return rc;
    HfxPlainCands c{*db, d_begin, d_num, d_off};
    return launch_xcorr(p, d_scores, c, total, sp, stream);
}

/* ------------------------------------------------------------------ */
/* Dense rows for the reference ABI (addIons, ion_series.cu:319-343)     */
/* ------------------------------------------------------------------ */

/* One wave per row: map -> dense row of floats 0/10/25/50, every element
 * written (no memset needed, no cross-block race). */
__device__ __forceinline__ void wave_map_to_row(uint32_t *map, float *row, uint32_t len)
{
    const uint32_t nw = (len + 7) >> 3;
    for (uint32_t w = hfx_lane(); w < nw; w += HFX_WAVE) {
        uint32_t word = map[w];
        map[w] = 0;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
            uint32_t b = (w << 3) + k;
            if (b < len) row[b] = hfx_code_weight((word >> (4 * k)) & 0xFu);
        }
    }
}

__global__ void __launch_bounds__(XC_THREADS)
    k_add_ions(float *spec, const float *residual, const float *mass, const uint8_t *ions_all,
               const uint32_t *tc_a, const uint32_t *tn_a, const uint32_t *idx, uint32_t num_idx,
               uint32_t zmax, uint32_t len)
{
    __shared__ uint32_t maps[XC_WAVES][HFX_MAP_WORDS];
    __shared__ float mtab[32];
    const uint32_t wave = threadIdx.x / HFX_WAVE;
    uint32_t *map = maps[wave];
    if (threadIdx.x < 32) mtab[threadIdx.x] = threadIdx.x < 26 ? mass[threadIdx.x] : 0.0f;
    for (uint32_t w = hfx_lane(); w < HFX_MAP_WORDS; w += HFX_WAVE) map[w] = 0;
    __syncthreads();
    for (uint32_t row = blockIdx.x * XC_WAVES + wave; row < num_idx; row += gridDim.x * XC_WAVES) {
        const uint32_t pep = idx[row];
        const uint32_t tc = tc_a[pep], tn = tn_a[pep];
        const uint32_t n = tn > tc ? tn - tc : 0;
        hfx_wave_peaks(map, n, residual[pep], zmax, len, HfxPlainMass{ions_all + tc, mtab});
        hfx_wave_sync();
        wave_map_to_row(map, spec + (size_t)row * len, len);
        hfx_wave_sync();
    }
}

extern "C" void addIons(float *d_spec, const float *d_residual, const float *d_mass,
                        const uint8_t *d_ions, const uint32_t *d_tc, const uint32_t *d_tn,
                        const uint32_t *d_idx, const uint32_t num_idx, const uint32_t max_charge,
                        const uint32_t len_spec)
{
    if (num_idx == 0) return;
    if (max_charge < 1 || max_charge > HFX_MAX_CHARGE) {
        hfx_set_error("addIons: max_charge %u outside 1..4 (no kernel launched)", max_charge);
        return;
    }
    if (len_spec > HFX_MAX_LEN) {
        hfx_set_error("addIons: len_spec %u > %d", len_spec, HFX_MAX_LEN);
        return;
    }
    uint32_t blocks = min((num_idx + XC_WAVES - 1) / XC_WAVES, 65535u);
    hipLaunchKernelGGL(k_add_ions, dim3(blocks), dim3(XC_THREADS), 0, 0, d_spec, d_residual, d_mass,
                       d_ions, d_tc, d_tn, d_idx, num_idx, max_charge, len_spec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) hfx_set_error("addIons: %s", hipGetErrorString(e));
}

/* ------------------------------------------------------------------ */
/* Row MVM replacing cuBLAS sgemv(T)                                    */
/* ------------------------------------------------------------------ */

__global__ void __launch_bounds__(256) k_mvm_rows(float *y, const float *A, uint32_t lda,
                                                  const float *x, uint32_t m, uint32_t n)
{
    const uint32_t row = blockIdx.x * 4 + threadIdx.x / HFX_WAVE;
    if (row >= m) return;
    const float *a = A + (size_t)row * lda;
    float acc = 0.0f;
    for (uint32_t j = hfx_lane(); j < n; j += HFX_WAVE) acc = acc + a[j] * x[j];
    acc = hfx_wave_sum_f32(acc);
    if (hfx_lane() == 0) y[row] = acc;
}

/* mvm_ff (mvm.cu:21-56): y = A x for row-major A of m rows and n columns,
 * on the handle's stream (cublasSgemv(CUBLAS_OP_T) over the transposed
 * view, Level2.chs:65-77).  The handle replaces the cuBLAS one of
 * Execution.hs:49-53: it only carries the stream set by hfx_blas_set_stream
 * (cublasSetStream_v2). */
struct hfx_blas {
    hipStream_t stream;
};

extern "C" int hfx_blas_create(void **handle)
{
    HFX_REQUIRE(handle, HFX_ERR_ARG, "hfx_blas_create: null out");
    hfx_blas *b = (hfx_blas *)calloc(1, sizeof(hfx_blas));
    HFX_REQUIRE(b, HFX_ERR_NOMEM, "hfx_blas_create: out of host memory");
    *handle = b;
    return HFX_OK;
}

extern "C" int hfx_blas_destroy(void *handle)
{
    free(handle);
    return HFX_OK;
}

extern "C" int hfx_blas_set_stream(void *handle, void *stream)
{
    HFX_REQUIRE(handle, HFX_ERR_ARG, "hfx_blas_set_stream: null handle");
    ((hfx_blas *)handle)->stream = hfx_stream(stream);
    return HFX_OK;
}

extern "C" void mvm_ff(void *handle, float *d_y, const float *d_A, const float *d_x,
                       const uint32_t m, const uint32_t n)
{
    if (m == 0) return;
    const hipStream_t st = handle ? ((hfx_blas *)handle)->stream : nullptr;
    hipLaunchKernelGGL(k_mvm_rows, dim3((m + 3) / 4), dim3(256), 0, st, d_y, d_A, n, d_x, m, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) hfx_set_error("mvm_ff: %s", hipGetErrorString(e));
}

extern "C" int hfx_mvm_rows(hfx_ctx *ctx, float *d_y, const float *d_A, uint32_t lda,
                            const float *d_x, uint32_t m, uint32_t n, void *stream)
{
    (void)ctx;
    HFX_REQUIRE(d_y && d_A && d_x, HFX_ERR_ARG, "hfx_mvm_rows: null argument");
    HFX_REQUIRE(n <= lda, HFX_ERR_ARG, "hfx_mvm_rows: n %u > lda %u", n, lda);
    if (m == 0) return HFX_OK;
    hipLaunchKernelGGL(k_mvm_rows, dim3((m + 3) / 4), dim3(256), 0, hfx_stream(stream), d_y, d_A,
                       lda, d_x, m, n);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}
