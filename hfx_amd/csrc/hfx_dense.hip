/*
 * hfx_dense.hip -- XCorr of spectra that share a precursor window as a dense
 * BF16 MFMA contraction (precision form), FP32 accumulation.
 *
 * The reference builds, per (spectrum, candidate), a dense theoretical row of
 * len bins (weights 50/25/10, max per bin; ion_series.cu:111-232) and takes
 * its dot product with the spectrum through cuBLAS sgemv (Search.hs:242-245).
 * Here eight spectra that are neighbours in precursor mass form a block; every
 * candidate of the block's union window gets ONE theoretical row, and the
 * rows meet all eight spectra in one matrix product on the MFMA units:
 *
 *   A [16 candidates x K bins]   theoretical weights, exact in BF16
 *   B [K bins x 16 columns]      column 2g = hi(x_g), 2g+1 = lo(x_g), where
 *                                hi = bf16_rne(x), lo = bf16_rne(x - hi)
 *   C [16 x 16] FP32             score(c, g) = C[c][2g] + C[c][2g+1]
 *
 * with v_mfma_f32_16x16x32_bf16 (gfx950).  Error: |x - hi - lo| <= 2^-16 |x|
 * per bin, products are exact in FP32, and a zero term adds exactly, so
 * |score - exact| <= (2^-16 + (m + 2) 2^-24) * sum_b w_b |x_b| with m the
 * number of non-zero bins of the row (<= 10 * residues * fragment charges);
 * the top-N stage re-scores exactly with that bound.
 *
 * Layout and work split (one persistent workgroup of sixteen waves per CU):
 *   - k_dn_split writes every spectrum once as hi/lo BF16 rows of DN_MAXK bins
 *     (zero beyond its length), bins permuted inside each group of eight so
 *     that the A fragment built below needs no shuffles;
 *   - the b ladders of the candidates (sequential FP32 sums of residue masses,
 *     bit-exact with the sparse path) are built once, in residual-sorted
 *     order, so that the ladder positions of a tile of consecutive
 *     candidates are one contiguous range;
 *   - a work item is (block, fragment charge level, chunk of CH candidates);
 *     the workgroup stages the block's 16 B columns in LDS (once per block);
 *   - four waves share a map and take 16 candidates: each lane takes one
 *     ladder position (of its wave's four candidates) and ORs the unary
 *     weight codes of its peaks into the map in LDS (OR of codes = max
 *     weight), bin x in word x/8 of its row, so a lane's A word at k-step ks
 *     is word 4 ks + k-group; ladder positions are loaded a tile ahead;
 *   - A words (8 nibbles) become eight BF16 weights with four v_perm lookups
 *     (codes 0/1/3/7 select the high and low bytes of 0, 10, 25, 50) and four
 *     byte interleaves; the map is cleared as it is read; each wave of the
 *     map contracts every fourth group of 128 bins and the four partial sums
 *     are added in wave order.
 * Bins use the same canonical arithmetic as the sparse path: the divisions
 * by the bin width and by the fragment charge are computed as
 * q = x*R, r = fma(-q, d, x), q + r*R, which equals the correctly rounded
 * x / d for every x the ladders produce (checked exhaustively on the CPU).
 */
#include "hfx_cand.h"

#include <algorithm>
#include <type_traits>

#define DN_NS 8                     /* spectra per block */
#define DN_COLS 16                  /* MFMA N: hi and lo of each spectrum */
#define DN_PAIRS 4                  /* four maps, one 16-candidate tile each */
#define DN_SPLIT 4                  /* waves sharing a map */
#define DN_WAVES (DN_PAIRS * DN_SPLIT)
#define DN_CPW (16 / DN_SPLIT)      /* candidates whose ladder positions a wave takes */
#define DN_THREADS (DN_WAVES * HFX_WAVE)
#define DN_ROWS 16                  /* candidates per wave tile (MFMA M) */
#define DN_MAXK 2176                /* bins of a row (multiple of 128) */
#define DN_NKS (DN_MAXK / 32)       /* k-steps of 32 bins */
/* map words per row: bin x in word 1 + x/8; 290 = 2 mod 32, so the 16 rows x
 * 2 k-groups that a half-wave reads or clears fall in 32 distinct banks */
#define DN_MSTR 290
static_assert(DN_MSTR >= DN_MAXK / 8 + 2 && DN_MSTR % 32 == 2, "map row stride");
/* the clamped peaks' words past kq / 8 (kq <= DN_MAXK) stay inside the row */
static_assert(1 + DN_MAXK / 8 + 16 < DN_MSTR, "clamped peak words within a map row");
#define DN_CSTR (DN_MAXK + 8)       /* BF16 per B column in LDS (padded) */
#define DN_PRE 2                    /* ladder positions per lane loaded a tile ahead */
#define DN_NONE 0xffffffffu

#define DN_B_BYTES (DN_COLS * DN_CSTR * 2)
#define DN_MAP_BYTES (DN_PAIRS * DN_ROWS * DN_MSTR * 4)
#define DN_COMB_BYTES (DN_PAIRS * DN_SPLIT * HFX_WAVE * 16)
#define DN_LDS_BYTES (DN_B_BYTES + DN_MAP_BYTES + DN_COMB_BYTES)
/* the kernel's fixed __shared__ arrays (m_*, sh_t) share the CU's 160 KiB
 * with the dynamic part */
#define DN_FIXED_LDS (DN_NS * 4 * 4 + DN_NS * 8 + 2 * 4 + 16)
static_assert(DN_LDS_BYTES + DN_FIXED_LDS <= 160 * 1024, "dense kernel LDS exceeds a CU");
static_assert(DN_SPLIT == 2 || DN_SPLIT == 4, "a wave takes 16 / DN_SPLIT candidates and 4 / DN_SPLIT rows of C");

typedef __bf16 dn_bf16x8 __attribute__((ext_vector_type(8)));
typedef float dn_f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t dn_u32x4 __attribute__((ext_vector_type(4)));

/* ------------------------------------------------------------------ */
/* Arithmetic                                                            */
/* ------------------------------------------------------------------ */

/* round-to-nearest-even FP32 -> BF16 (finite inputs) */
__device__ __forceinline__ uint32_t dn_bf16(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

/* Eight unary weight codes (one nibble per bin, bins 0..7 of a group) to the
 * eight BF16 of an A fragment, in slot order bins 0,2,4,6,1,3,5,7.  Byte pools
 * for v_perm: code c selects byte c; lo bytes of (0, 10, 25, 50) = 00 20 C8 48,
 * hi bytes = 00 41 41 42. */
__device__ __forceinline__ dn_u32x4 dn_weights(uint32_t w)
{
    const uint32_t e = w & 0x0f0f0f0fu;
    const uint32_t o = (w >> 4) & 0x0f0f0f0fu;
    const uint32_t lo_e = __builtin_amdgcn_perm(0x48000000u, 0xc8002000u, e);
    const uint32_t hi_e = __builtin_amdgcn_perm(0x42000000u, 0x41004100u, e);
    const uint32_t lo_o = __builtin_amdgcn_perm(0x48000000u, 0xc8002000u, o);
    const uint32_t hi_o = __builtin_amdgcn_perm(0x42000000u, 0x41004100u, o);
    dn_u32x4 a;
    a.x = __builtin_amdgcn_perm(hi_e, lo_e, 0x05010400u);
    a.y = __builtin_amdgcn_perm(hi_e, lo_e, 0x07030602u);
    a.z = __builtin_amdgcn_perm(hi_o, lo_o, 0x05010400u);
    a.w = __builtin_amdgcn_perm(hi_o, lo_o, 0x07030602u);
    return a;
}

/* Lane moves inside rows of 16 lanes through DPP (no LDS round trip):
 * row_ror:n -- lane i takes lane (i + n) mod 16 of its row; quad_perm
 * [1,0,3,2] swaps neighbours. */
#define DN_ROR(n) (0x120 + (n))
#define DN_SWAP1 0xb1
template <int CTRL>
__device__ __forceinline__ uint32_t dn_dpp(uint32_t old, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dn_dppf(float v)
{
    return __uint_as_float(dn_dpp<CTRL>(__float_as_uint(v), __float_as_uint(v)));
}

/* maximum over each row of 16 lanes of non-negative floats (ordered as their
 * bit patterns), left in every lane of the row (the rotations by 8, 4, 2, 1
 * reach all 16 lanes whichever way they turn) */
__device__ __forceinline__ float dn_row_max(float f)
{
    uint32_t v = __float_as_uint(f);
    v = max(v, dn_dpp<DN_ROR(8)>(v, v));
    v = max(v, dn_dpp<DN_ROR(4)>(v, v));
    v = max(v, dn_dpp<DN_ROR(2)>(v, v));
    v = max(v, dn_dpp<DN_ROR(1)>(v, v));
    return __uint_as_float(v);
}

/* ------------------------------------------------------------------ */
/* Spectra as hi/lo BF16 rows                                            */
/* ------------------------------------------------------------------ */

/* xb[s][0][:] = hi, xb[s][1][:] = lo, DN_MAXK bins each; within each group of
 * eight bins slot j holds bin {0,2,4,6,1,3,5,7}[j]. */
__global__ void k_dn_split(uint16_t *xb, hfx_spectra sp)
{
    const uint32_t s = blockIdx.x;
    const uint32_t len = min(sp.d_len[s], min(sp.x_stride, (uint32_t)DN_MAXK));
    const float *x = sp.d_x + (size_t)s * sp.x_stride;
    uint16_t *hi = xb + (size_t)s * 2 * DN_MAXK, *lo = hi + DN_MAXK;
    for (uint32_t q = threadIdx.x; q < DN_MAXK / 8; q += blockDim.x) {
        uint32_t h[8], l[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t bin = q * 8 + j;
            const float v = bin < len ? x[bin] : 0.0f;
            h[j] = dn_bf16(v);
            l[j] = dn_bf16(v - __uint_as_float(h[j] << 16));
        }
        uint4 vh, vl;
        vh.x = h[0] | h[2] << 16; vh.y = h[4] | h[6] << 16;
        vh.z = h[1] | h[3] << 16; vh.w = h[5] | h[7] << 16;
        vl.x = l[0] | l[2] << 16; vl.y = l[4] | l[6] << 16;
        vl.z = l[1] | l[3] << 16; vl.w = l[5] | l[7] << 16;
        *(uint4 *)(hi + q * 8) = vh;
        *(uint4 *)(lo + q * 8) = vl;
    }
}

/* ------------------------------------------------------------------ */
/* Blocks                                                                 */
/* ------------------------------------------------------------------ */

/* read-only, wave-uniform tables: scalar loads through the constant address
 * space (nothing in the kernel writes them) */
template <class T>
__device__ __forceinline__ T dn_sload(const T *p, uint32_t i)
{
    return ((const __attribute__((address_space(4))) T *)p)[i];
}

__device__ __forceinline__ uint32_t dn_order(const uint32_t *order, uint32_t i)
{
    return order ? order[i] : i;
}

/* spectra (positions in order) of block b: [bstart[b], bstart[b+1]), at
 * most NS of them, or NS at a time without a table */
__device__ __forceinline__ uint32_t dn_bfirst(const uint32_t *bstart, uint32_t b, uint32_t NS)
{
    return bstart ? bstart[b] : b * NS;
}
__device__ __forceinline__ uint32_t dn_bend(const uint32_t *bstart, uint32_t b, uint32_t S,
                                            uint32_t NS)
{
    return bstart ? bstart[b + 1] : min(S, b * NS + NS);
}

/* block b: union window, fragment charge levels present (bit z), longest
 * row kb (multiple of 128), work items */
template <uint32_t NS>
__global__ void k_dn_blocks(uint32_t *ubeg, uint32_t *uend, uint32_t *binfo, uint32_t *ntiles,
                            const uint32_t *order, const uint32_t *bstart, const uint32_t *begin,
                            const uint32_t *num, hfx_spectra sp, uint32_t B, uint32_t CH)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    uint32_t lo = 0xffffffffu, hi = 0, mask = 0, kmax = 0;
    const uint32_t i0 = dn_bfirst(bstart, b, NS), i1 = dn_bend(bstart, b, sp.num_spec, NS);
    /* the block's spectra first, then their windows, charges and lengths:
     * two rounds of independent loads */
    uint32_t sv[NS];
#pragma unroll
    for (uint32_t g = 0; g < NS; ++g) sv[g] = i0 + g < i1 ? dn_order(order, i0 + g) : DN_NONE;
#pragma unroll
    for (uint32_t g = 0; g < NS; ++g) {
        const uint32_t s = sv[g];
        if (s == DN_NONE) continue;
        const uint32_t z = hfx_zeff(sp.d_zmax[s]), n = num[s], bg = begin[s], ln = sp.d_len[s];
        if (z == 0 || n == 0) continue;
        lo = min(lo, bg);
        hi = max(hi, bg + n);
        mask |= 1u << z;
        kmax = max(kmax, min(ln, min(sp.x_stride, (uint32_t)DN_MAXK)));
    }
    if (hi <= lo) { lo = 0; hi = 0; mask = 0; }
    const uint32_t kb = min((kmax + 127u) & ~127u, (uint32_t)DN_MAXK);
    ubeg[b] = lo;
    uend[b] = hi;
    binfo[b] = kb | mask << 16;
    ntiles[b] = ((hi - lo + CH - 1) / CH) * (uint32_t)__builtin_popcount(mask);
}

__global__ void k_dn_reset(uint32_t *counter) { *counter = 0; }

/* spectra whose fragment charge is not generated score 0 (no launch in the
 * reference, ion_series.cu:334-342) */
__global__ void k_dn_zero_unscored(float *scores, const uint64_t *off, const uint32_t *zmax,
                                   uint32_t S)
{
    const uint32_t s = blockIdx.x;
    if (s >= S || hfx_zeff(zmax[s]) != 0) return;
    for (uint64_t i = off[s] + threadIdx.x; i < off[s + 1]; i += blockDim.x) scores[i] = 0.0f;
}

/* ------------------------------------------------------------------ */
/* The contraction                                                        */
/* ------------------------------------------------------------------ */

/* the pipeline registers of one tile */
struct DnPipe {
    hfx_dn_pos pos[DN_PRE]; /* the half's ladder positions, lane + 64 k */
    uint64_t rng;           /* lanes 0, 1: the half's range in the tile after */
    uint64_t L;             /* first position */
    uint32_t np;            /* number of positions */
    float top;              /* lane < 16: top of candidate lane */
};

struct DnArgs {
    float *scores;
    const uint64_t *score_off;
    const uint32_t *begin, *num, *order;
    const uint32_t *bstart; /* [B + 1] block starts in order, or null: eight a block */
    uint32_t B, CH;
    const uint32_t *ubeg, *uend, *binfo;
    const uint64_t *tile_off;
    const uint16_t *xb;
    uint32_t *counter;
    float *dummy; /* [64]: target of the score stores that have no score */
    hfx_dn_ladder lad;
    hfx_spectra sp;
};

/* Map rows: bin x lives in word 1 + x/8 of its row (word 0 is a margin that
 * the contraction never reads, so a bin in [-8, 0) needs no check).  `row`
 * below points at word 1. */

/* Unchecked forms, for tiles whose every bin lies in [1, kq): a bin comes as
 * the bits u of x + 1.5 * 2^23 -- the float addition rounds x to the nearest
 * integer, ties to even (as rint), and leaves it in the low mantissa bits, so
 * u = 0x4B400000 + rint(x) for |x| < 2^22.  The constant's low three bits are
 * zero: u & 7 = x & 7 and u >> 3 = x >> 3 + DN_MAGIC_WORDS. */
#define DN_MAGIC 12582912.0f
#define DN_MAGIC_WORDS (0x4B400000u >> 3)

__device__ __forceinline__ uint32_t dn_binu(float mz)
{
    const float R = 1.0f / HFX_BIN_WIDTH_MONO;
    return __float_as_uint(hfx_div(mz, HFX_BIN_WIDTH_MONO, R) + DN_MAGIC);
}
/* rowm: the row's word-0 address less DN_MAGIC_WORDS words (a 32-bit LDS
 * address that wraps; only rowm + (u >> 3) is ever dereferenced) */
__device__ __forceinline__ void dn_put_u(uint32_t *rowm, uint32_t u, uint32_t code)
{
    atomicOr(rowm + (u >> 3), code << ((u & 7u) << 2));
}
/* 25 at x-1, 50 at x, 25 at x+1: the nibbles 3, 7, 3 shifted into place as
 * one 64-bit value over the word of x-1 and the next (zero when nothing
 * spills) */
__device__ __forceinline__ void dn_put3_u(uint32_t *rowm, uint32_t u)
{
    const uint32_t um = u - 1u;
    const uint64_t t = (uint64_t)(HFX_CODE25 | HFX_CODE50 << 4 | HFX_CODE25 << 8)
                       << ((um & 7u) << 2);
    uint32_t *w = rowm + (um >> 3);
    atomicOr(w, (uint32_t)t);
    atomicOr(w + 1, (uint32_t)(t >> 32));
}

/* Clamped form, for tiles whose peaks may fall at or past kq (or, with a
 * light first residue, below 0): the bits u of a bin are clamped into
 * [bin -8, bin kq + 8 (lane % 16)] (one v_med3).  Bins below 0 land in the
 * row's margin word 0, which the contraction never reads; bins at or past kq
 * in words kq / 8 .. kq / 8 + 16 (kq is a multiple of 128; a triple that
 * starts in the last of a lane's words spills into the next), which this
 * tile's contraction does not read and which are cleared before the next
 * tile's peaks.  The lane-dependent ceiling spreads the clamped writes over
 * sixteen words, so that they do not all queue on one address. */
#define DN_ULO (0x4B400000u - 8u)
#define DN_TRASH 17u /* words past kq / 8 that clamped peaks may touch */
__device__ __forceinline__ uint32_t dn_clamp(uint32_t u, uint32_t uhi)
{
    return min(max(u, DN_ULO), uhi);
}

/* division by the fragment charge Z: exact as a product for Z = 1, 2, 4 */
template <int Z>
__device__ __forceinline__ float dn_zdiv(float x)
{
    if constexpr (Z == 3) return hfx_div(x, 3.0f, 1.0f / 3.0f);
    else return x * (1.0f / (float)Z);
}

/* peaks of one ladder position (b, y) at charge Z into a map row: the ion
 * series of ion_series.cu:111-156 with the FP32 operation order of the
 * sparse path (zh = H * Z, H2O / Z and NH3 / Z fold to the same FP32 values);
 * CHECK keeps only bins in [0, kq) (the others go to words this tile's
 * contraction does not read) */
template <int Z, bool CHECK, uint32_t SPREAD = 15u>
__device__ __forceinline__ void dn_peaks(uint32_t *row, float b, float y, uint32_t kq,
                                         uint32_t lane)
{
    constexpr float zh = HFX_DEV_H * (float)Z;
    constexpr float h2o = HFX_DEV_H2O / (float)Z, nh3 = HFX_DEV_NH3 / (float)Z;
    const float ma = dn_zdiv<Z>((b - HFX_DEV_CO) + zh);
    const float mb = dn_zdiv<Z>(b + zh);
    const float my = dn_zdiv<Z>((y + HFX_DEV_H2O) + zh);
    if constexpr (CHECK) {
        uint32_t *rowm = row - DN_MAGIC_WORDS;
        const uint32_t uhi = 0x4B400000u + kq + 8u * (lane & SPREAD);
        dn_put_u(rowm, dn_clamp(dn_binu(ma), uhi), HFX_CODE10);
        dn_put3_u(rowm, dn_clamp(dn_binu(mb) - 1u, uhi) + 1u);
        dn_put_u(rowm, dn_clamp(dn_binu(mb - h2o), uhi), HFX_CODE10);
        dn_put_u(rowm, dn_clamp(dn_binu(mb - nh3), uhi), HFX_CODE10);
        dn_put3_u(rowm, dn_clamp(dn_binu(my) - 1u, uhi) + 1u);
        dn_put_u(rowm, dn_clamp(dn_binu(my - nh3), uhi), HFX_CODE10);
    } else {
        uint32_t *rowm = row - DN_MAGIC_WORDS;
        dn_put_u(rowm, dn_binu(ma), HFX_CODE10);
        dn_put3_u(rowm, dn_binu(mb));
        dn_put_u(rowm, dn_binu(mb - h2o), HFX_CODE10);
        dn_put_u(rowm, dn_binu(mb - nh3), HFX_CODE10);
        dn_put3_u(rowm, dn_binu(my));
        dn_put_u(rowm, dn_binu(my - nh3), HFX_CODE10);
    }
}

template <bool CHECK, uint32_t SPREAD = 15u>
__device__ __forceinline__ void dn_peaks_all(uint32_t *row, float b, float y, uint32_t kq,
                                             uint32_t lane, uint32_t lvl)
{
    dn_peaks<1, CHECK, SPREAD>(row, b, y, kq, lane);
    if (lvl >= 2) dn_peaks<2, CHECK, SPREAD>(row, b, y, kq, lane);
    if (lvl >= 3) dn_peaks<3, CHECK, SPREAD>(row, b, y, kq, lane);
    if (lvl >= 4) dn_peaks<4, CHECK, SPREAD>(row, b, y, kq, lane);
}

__global__ void __launch_bounds__(DN_THREADS, 1) k_dense(DnArgs a)
{
    extern __shared__ __align__(16) unsigned char dn_lds[];
    uint16_t *Bl = (uint16_t *)dn_lds;
    uint32_t *maps = (uint32_t *)(dn_lds + DN_B_BYTES);
    dn_f32x4 *comb_all = (dn_f32x4 *)(dn_lds + DN_B_BYTES + DN_MAP_BYTES);
    __shared__ uint32_t m_s[DN_NS], m_beg[DN_NS], m_end[DN_NS], m_z[DN_NS];
    __shared__ uint64_t m_off[DN_NS];
    __shared__ uint32_t sh_t[2];

    /* the DN_SPLIT waves q, q + 4, ... (h = wave / 4) share map q and pass
     * their partial sums through comb */
    const uint32_t tid = threadIdx.x, lane = hfx_lane(), wave = tid / HFX_WAVE;
    const uint32_t q = wave % DN_PAIRS, h = wave / DN_PAIRS;
    uint32_t *map = maps + q * DN_ROWS * DN_MSTR;
    dn_f32x4 *comb = comb_all + q * DN_SPLIT * HFX_WAVE;
    for (uint32_t i = tid; i < DN_PAIRS * DN_ROWS * DN_MSTR; i += DN_THREADS) maps[i] = 0;
    const uint64_t ntiles = dn_sload(a.tile_off, a.B);
    const uint32_t row = lane & 15u, kg = lane >> 4;
    uint32_t cur = DN_NONE;
    /* work items are claimed one ahead: the claim of item k+1 is in flight
     * while item k is processed */
    if (tid == 0) sh_t[0] = atomicAdd(a.counter, 1u);
    uint32_t next = 0; /* thread 0: the claim of the following item */
    for (uint32_t slot = 0;; slot ^= 1u) {
        __syncthreads(); /* every wave is done with the previous item */
        /* wave-uniform from here: block metadata comes through scalar loads */
        const uint32_t t0 = __builtin_amdgcn_readfirstlane(sh_t[slot]);
        if (t0 >= ntiles) break;
        /* items from the last to the first: the plan orders blocks by charge
         * level, then precursor mass, so the heaviest come first and the
         * lightest finish the launch */
        const uint32_t t = (uint32_t)ntiles - 1u - t0;
        if (tid == 0) next = atomicAdd(a.counter, 1u);
        uint32_t lo = 0, hi = a.B;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (dn_sload(a.tile_off, mid) <= t) lo = mid; else hi = mid;
        }
        const uint32_t b = lo;
        const uint32_t info = dn_sload(a.binfo, b);
        const uint32_t kb = info & 0xffffu, mask = info >> 16;
        if (b != cur) {
            if (tid < DN_NS) {
                const uint32_t i = dn_bfirst(a.bstart, b, DN_NS) + tid;
                uint32_t s = DN_NONE, be = 0, en = 0, z = 0;
                uint64_t of = 0;
                if (i < dn_bend(a.bstart, b, a.sp.num_spec, DN_NS)) {
                    s = dn_order(a.order, i);
                    be = a.begin[s];
                    en = be + a.num[s];
                    z = hfx_zeff(a.sp.d_zmax[s]);
                    of = a.score_off[s];
                }
                m_s[tid] = s; m_beg[tid] = be; m_end[tid] = en; m_z[tid] = z; m_off[tid] = of;
            }
            __syncthreads();
            /* all loads of the block's B columns in flight before any store */
            const uint32_t nvec = kb / 8;
            constexpr uint32_t PER = (DN_COLS * (DN_MAXK / 8) + DN_THREADS - 1) / DN_THREADS;
            uint4 val[PER];
#pragma unroll
            for (uint32_t k = 0; k < PER; ++k) {
                const uint32_t i = tid + k * DN_THREADS;
                val[k] = make_uint4(0u, 0u, 0u, 0u);
                if (i < DN_COLS * nvec) {
                    const uint32_t col = i / nvec, v = i - col * nvec, s = m_s[col >> 1];
                    if (s != DN_NONE)
                        val[k] = *(const uint4 *)(a.xb + ((size_t)s * 2 + (col & 1u)) * DN_MAXK +
                                                  v * 8);
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < PER; ++k) {
                const uint32_t i = tid + k * DN_THREADS;
                if (i < DN_COLS * nvec) {
                    const uint32_t col = i / nvec, v = i - col * nvec;
                    *(uint4 *)(Bl + col * DN_CSTR + v * 8) = val[k];
                }
            }
            cur = b;
            __syncthreads();
        }
        /* the claim made above is published for the next round here, long
         * after it was issued (sh_t[slot ^ 1] is not read before the next
         * barrier) */
        if (tid == 0) sh_t[slot ^ 1u] = next;
        const uint32_t beg = dn_sload(a.ubeg, b), end = dn_sload(a.uend, b);
        /* windows outside the ladder have no positions to read: no scores */
        if (beg < a.lad.q0 || end > a.lad.q0 + a.lad.nq) continue;
        const uint32_t nch = (end - beg + a.CH - 1) / a.CH;
        const uint32_t local = (uint32_t)(t - dn_sload(a.tile_off, b));
        const uint32_t li = local / nch, chunk = local - li * nch;
        uint32_t m = mask;
        for (uint32_t qq = 0; qq < li; ++qq) m &= m - 1u;
        const uint32_t lvl = (uint32_t)__builtin_ctz(m);
        const uint32_t cbeg = beg + chunk * a.CH, cend = min(cbeg + a.CH, end);
        const uint32_t step = DN_PAIRS * DN_ROWS;
        /* every pair runs the same number of tiles (barriers below are
         * workgroup-wide), rounded up to even for the two-tile loop; tiles
         * past the chunk are empty */
        const uint32_t ntile = ((cend - cbeg + step - 1) / step + 1u) & ~1u;
        const uint32_t pstart = cbeg + q * DN_ROWS;
        /* the spectrum of this lane's C column (lane % 16 = hi or lo of
         * spectrum g): charge level, window and score offset */
        const uint32_t g = row >> 1;
        const uint32_t zg = m_z[g], bg = m_beg[g], eg = m_end[g];
        const uint64_t og = m_off[g];
        float *const dummy = a.dummy + lane;

        /* software pipeline over the map's tiles: at the start of tile j the
         * ladder range of tile j+2 and the ladder positions and tops of tile
         * j+1 are issued, each consumed a tile later; loads and stores are
         * unconditional (clamped indices, a dummy target), so the number in
         * flight does not depend on the data and the waits stay exact.  Wave
         * h of a map takes the positions of its candidates DN_CPW h ..
         * DN_CPW (h + 1) - 1. */
        auto rows_of = [&](uint32_t t0) -> uint32_t {
            return t0 < cend ? min((uint32_t)DN_ROWS, cend - t0) : 0u;
        };
        /* lane 0: the half's first position in the tile at t0, lane 1: one
         * past its last (both offset 0 for an empty tile) */
        auto load_range = [&](uint32_t t0) -> uint64_t {
            const uint32_t nr = rows_of(t0);
            const uint32_t c = min(DN_CPW * h + (lane ? DN_CPW : 0u), nr);
            return a.lad.off[nr ? t0 + c - a.lad.q0 : 0u];
        };
        /* lane < 16: the top of candidate lane (meaningful below the tile's
         * row count) */
        auto load_top = [&](uint32_t t0) -> float {
            return a.lad.top[lane < rows_of(t0) ? t0 + lane - a.lad.q0 : 0u];
        };
        auto lane64 = [&](uint64_t v, uint32_t l) -> uint64_t {
            return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l) |
                   (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l) << 32;
        };
        /* lane l takes positions l, l + 64, ...; past the half's count it
         * reloads the last one (or position 0); a wave-uniform base and a
         * 32-bit index per lane */
        auto load_pos = [&](hfx_dn_pos(&pv)[DN_PRE], uint64_t L0, uint32_t np) {
            const hfx_dn_pos *base = np ? a.lad.pos + L0 : a.lad.pos;
            const uint32_t last = np ? np - 1u : 0u;
#pragma unroll
            for (uint32_t k = 0; k < DN_PRE; ++k) pv[k] = base[min(lane + k * HFX_WAVE, last)];
        };
        DnPipe X, Y;
        {
            /* X.rng before X's positions and top, as in the loop, so that the
             * wait for it at the loop head is the same on both entries */
            const uint64_t r0 = load_range(pstart);
            X.rng = load_range(pstart + step);
            __builtin_amdgcn_sched_barrier(0);
            X.L = lane64(r0, 0);
            X.np = (uint32_t)(lane64(r0, 1) - X.L);
            load_pos(X.pos, X.L, X.np);
            X.top = load_top(pstart);
        }
        auto tile = [&](DnPipe &cur, DnPipe &nxt, uint32_t p0) {
            const uint32_t nrows = rows_of(p0);
            /* issue: range of tile j+2, positions and tops of tile j+1 */
            nxt.rng = load_range(p0 + 2 * step);
            nxt.L = lane64(cur.rng, 0);
            nxt.np = (uint32_t)(lane64(cur.rng, 1) - nxt.L);
            load_pos(nxt.pos, nxt.L, nxt.np);
            nxt.top = load_top(p0 + step);

            /* every peak bin of a position lies below top / 1.0005 + 2 for
             * top = max(residual, last b) of its candidate: kq covers the
             * tile, and where it is not cut at the spectrum's length the
             * unchecked writes apply */
            const float mm = __uint_as_float(__builtin_amdgcn_readfirstlane(
                __float_as_uint(dn_row_max(lane < nrows ? cur.top : 0.0f))));
            const uint32_t kneed = (uint32_t)(mm + 24.0f) + 4u;
            const uint32_t kq = nrows ? min(kb, (kneed + 127u) & ~127u) : 0u;
            auto peaks = [&](auto chk) {
                constexpr bool CHECK = decltype(chk)::value;
#pragma unroll
                for (uint32_t k = 0; k < DN_PRE; ++k) {
                    const hfx_dn_pos &v = cur.pos[k];
                    if (lane + k * HFX_WAVE < cur.np)
                        dn_peaks_all<CHECK>(map + (v.q - p0) * DN_MSTR + 1, v.b, v.y, kq, lane,
                                            lvl);
                }
                for (uint32_t i = DN_PRE * HFX_WAVE + lane; i < cur.np; i += HFX_WAVE) {
                    const hfx_dn_pos v = a.lad.pos[cur.L + i];
                    dn_peaks_all<CHECK>(map + (v.q - p0) * DN_MSTR + 1, v.b, v.y, kq, lane, lvl);
                }
            };
            if (kneed <= kq) peaks(std::false_type{});
            else peaks(std::true_type{});
            __syncthreads(); /* (b) the tile's map rows are complete */
            /* the clamped peaks' words past the tile's bins, cleared before
             * the next tile's peaks (which follow barrier (c)) */
            if (kneed > kq && h == DN_SPLIT - 1)
                for (uint32_t k = lane; k < DN_ROWS * DN_TRASH; k += HFX_WAVE) {
                    const uint32_t r = k / DN_TRASH;
                    map[r * DN_MSTR + 1 + (kq >> 3) + (k - r * DN_TRASH)] = 0u;
                }

            /* 16 x 16 x kq contraction over groups of four k-steps (128
             * bins): wave h of the map takes groups h, h + DN_SPLIT, ... and
             * clears their map words.  The split depends only on the group
             * index, so each wave's sum over a row's non-zero k-steps -- and
             * the score -- is the same whichever tile the row falls in.  Two
             * accumulator chains; the next group's operands are read ahead (a
             * read past the last group stays inside the LDS allocation,
             * unused).  Lane (row, kg) takes word 4 ks + kg of its row at
             * k-step ks. */
            dn_f32x4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
            const uint32_t ngrp = kq >> 7;
            uint32_t *arow = map + row * DN_MSTR + 1 + kg;
            const uint16_t *bcol = Bl + row * DN_CSTR + kg * 8;
            auto mma4 = [&](const dn_u32x4 &cw, const dn_u32x4 &b0, const dn_u32x4 &b1,
                            const dn_u32x4 &b2, const dn_u32x4 &b3) {
                acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    __builtin_bit_cast(dn_bf16x8, dn_weights(cw.x)),
                    __builtin_bit_cast(dn_bf16x8, b0), acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    __builtin_bit_cast(dn_bf16x8, dn_weights(cw.y)),
                    __builtin_bit_cast(dn_bf16x8, b1), acc1, 0, 0, 0);
                acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    __builtin_bit_cast(dn_bf16x8, dn_weights(cw.z)),
                    __builtin_bit_cast(dn_bf16x8, b2), acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    __builtin_bit_cast(dn_bf16x8, dn_weights(cw.w)),
                    __builtin_bit_cast(dn_bf16x8, b3), acc1, 0, 0, 0);
            };
            /* group offsets: G = DN_SPLIT groups on, 16 words / 128 BF16 a group */
            constexpr uint32_t GW = 16 * DN_SPLIT, GB = 128 * DN_SPLIT;
            if (h < ngrp) {
                uint32_t *aw = arow + 16 * h;
                const uint16_t *bw = bcol + 128 * h;
                dn_u32x4 ca = {aw[0], aw[4], aw[8], aw[12]};
                dn_u32x4 a0 = *(const dn_u32x4 *)(bw + 0), a1 = *(const dn_u32x4 *)(bw + 32);
                dn_u32x4 a2 = *(const dn_u32x4 *)(bw + 64), a3 = *(const dn_u32x4 *)(bw + 96);
                for (uint32_t g = h; g < ngrp; g += 2 * DN_SPLIT) {
                    /* group g (in ca, a0..a3); read group g + G */
                    aw[0] = 0u; aw[4] = 0u; aw[8] = 0u; aw[12] = 0u;
                    const dn_u32x4 cb = {aw[GW], aw[GW + 4], aw[GW + 8], aw[GW + 12]};
                    const dn_u32x4 e0 = *(const dn_u32x4 *)(bw + GB);
                    const dn_u32x4 e1 = *(const dn_u32x4 *)(bw + GB + 32);
                    const dn_u32x4 e2 = *(const dn_u32x4 *)(bw + GB + 64);
                    const dn_u32x4 e3 = *(const dn_u32x4 *)(bw + GB + 96);
                    mma4(ca, a0, a1, a2, a3);
                    if (g + DN_SPLIT >= ngrp) break;
                    /* group g + G (in cb, e0..e3); read group g + 2G */
                    aw[GW] = 0u; aw[GW + 4] = 0u; aw[GW + 8] = 0u; aw[GW + 12] = 0u;
                    ca = dn_u32x4{aw[2 * GW], aw[2 * GW + 4], aw[2 * GW + 8], aw[2 * GW + 12]};
                    a0 = *(const dn_u32x4 *)(bw + 2 * GB);
                    a1 = *(const dn_u32x4 *)(bw + 2 * GB + 32);
                    a2 = *(const dn_u32x4 *)(bw + 2 * GB + 64);
                    a3 = *(const dn_u32x4 *)(bw + 2 * GB + 96);
                    mma4(cb, e0, e1, e2, e3);
                    aw += 2 * GW;
                    bw += 2 * GB;
                }
            }
            const dn_f32x4 part = acc0 + acc1;
            comb[h * HFX_WAVE + lane] = part;
            __syncthreads(); /* (c) every wave's partial sums are in */
            /* C[row 4 kg + i][column lane % 16]: wave h stores rows
             * i = (4 / DN_SPLIT) h + i2, summing the partials of waves
             * 0, 1, ... in that order; an even column's lane adds its odd
             * neighbour (hi + lo) */
            constexpr uint32_t RPW = 4 / DN_SPLIT;
#pragma unroll
            for (uint32_t i2 = 0; i2 < RPW; ++i2) {
                const uint32_t i = RPW * h + i2;
                const float *cf = (const float *)comb;
                float c4 = cf[(0 * HFX_WAVE + lane) * 4 + i];
#pragma unroll
                for (uint32_t w2 = 1; w2 < DN_SPLIT; ++w2) c4 = c4 + cf[(w2 * HFX_WAVE + lane) * 4 + i];
                const float v = c4 + dn_dppf<DN_SWAP1>(c4);
                const uint32_t rr_i = 4u * kg + i, p = p0 + rr_i;
                const bool ok =
                    (lane & 1u) == 0 && rr_i < nrows && zg == lvl && bg <= p && p < eg;
                *(ok ? a.scores + og + (p - bg) : dummy) = v;
            }
        };
        for (uint32_t j = 0; j < ntile; j += 2) {
            tile(X, Y, pstart + j * step);
            tile(Y, X, pstart + (j + 1) * step);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Integer form: int8 MFMA over three limbs of each spectrum              */
/* ------------------------------------------------------------------ */
/*
 * Same candidates, maps and peaks as above; the contraction is exact:
 *
 *   A [16 candidates x 64 bins]   weights 0/10/25/50 as int8
 *   B [64 bins x 16 spectra]      one base-256 digit of each X_s
 *   C [16 x 16] int32             sum of w * digit, exact
 *
 * with v_mfma_i32_16x16x64_i8, once per digit for each A fragment.  Spectrum
 * s is scaled by the power of two 2^k_s (the largest with
 * max|x_s| 2^k_s <= 8355711) and rounded to integers X = rint(x 2^k_s),
 * written as balanced base-256 digits X = d2 2^16 + d1 2^8 + d0, each in
 * [-128, 127] (8355711 = 0x7F7F7F is the largest |X| they reach).  Then
 * score = (a2 2^16 + a1 2^8 + a0) 2^-k_s is formed exactly in FP64 and rounded
 * once to FP32:
 *   |score - exact| <= 2^-(k_s+1) sum_b w_b + 2^-24 |score|
 *                   <= max|x_s| / 8355711 * 50 len_s + 2^-24 * 50 sum|x_s|
 * (2^-k_s < 2 max|x_s| / 8355711).
 *
 * Sixteen spectra a block (the MFMA's N), 64 candidates a tile (four 16-row
 * blocks of A); one persistent workgroup of sixteen waves per CU (four a
 * SIMD, so that each holds up to 128 registers) in two groups of eight that
 * take turns, with two maps and two sum buffers:
 *
 *   phase k:  group k % 2      forms the scores of tile k - 2 and writes the
 *                              peaks of tile k into map k % 2;
 *             group (k+1) % 2  contracts tile k - 1 from map (k+1) % 2;
 *             one workgroup barrier,
 *
 * so the peaks (VALU, LDS atomics) of one tile overlap the contraction (VALU,
 * MFMA, LDS reads) of the previous one.
 *   - peaks: the tile's ladder positions go to the group's waves 64 at a time
 *     (group wave g takes positions 64 g + lane, 64 (g + 8) + lane, ...);
 *   - contraction: group wave g takes k-steps g, g + 8, g + 16, ... (64 bins
 *     each) of the four row blocks; the B fragments of its k-steps (three
 *     digits each) stay in its registers for the whole block, loaded once
 *     from HBM, so B never passes through LDS; lane (row, kg) reads map words
 *     8 ks + kg and 8 ks + 4 + kg of its row (clearing them), conflict-free,
 *     and one v_perm per four bins turns codes into weight bytes; each wave
 *     folds its three digit sums into one int64 per output (a2 2^16 +
 *     a1 2^8 + a0, exact) and adds it into LDS (ds_add_u64, exact in any
 *     order).
 * B's K order (the split's layout): byte 16 kg + j of a k-step's 64 is bin
 * (j < 8 ? 0 : 32) + 8 kg + {0,2,4,6,1,3,5,7}[j % 8].
 */
#define D8_NS 16
#define D8_RB 4                   /* 16-row blocks of A a tile */
#define D8_ROWS (16 * D8_RB)
#define D8_WAVES 16
#define D8_GW 8                   /* waves a group */
#define D8_THREADS (D8_WAVES * HFX_WAVE)
#define D8_GT (D8_GW * HFX_WAVE)  /* threads a group */
#define D8_MAXK DN_MAXK           /* bins of a row: 34 k-steps of 64 */
#define D8_NKS (D8_MAXK / 64)
#define D8_KPW ((D8_NKS + D8_GW - 1) / D8_GW) /* k-steps a wave holds B for */
/* map rows: bin x in word 1 + x / 8 (word 0 a margin); clamped peaks spread
 * over four words past kq / 8 (D8_TRASH words with a triple's spill); a row
 * stride of 2 mod 4 words keeps the contraction's reads conflict-free */
#define D8_SPREAD 3u
#define D8_TRASH 5u
#define D8_MSTR 278
static_assert(D8_MSTR >= 1 + D8_MAXK / 8 + D8_TRASH && D8_MSTR % 4 == 2, "map row stride");
#define D8_MAP_BYTES (D8_ROWS * D8_MSTR * 4)
/* int64 sums [row block][16 rows][16 columns] */
#define D8_COMB_BYTES (D8_RB * 16 * 16 * 8)
#define D8_LDS_BYTES (2 * D8_MAP_BYTES + 2 * D8_COMB_BYTES)
#define D8_FIXED_LDS (D8_NS * 4 * 5 + D8_NS * 8 + 2 * 4 + 2 * 4 + 16)
#define D8_XMAX 8355711.0f
static_assert(D8_LDS_BYTES + D8_FIXED_LDS <= 160 * 1024, "integer dense kernel LDS exceeds a CU");
static_assert(D8_MAXK % 128 == 0, "rows of whole k-steps");
/* a wave's digit sums: at most D8_KPW k-steps of 64 bins, weights <= 50,
 * digits in [-128, 127]; 257 times that stays below 2^31 */
static_assert((int64_t)D8_KPW * 64 * 50 * 128 * 257 < ((int64_t)1 << 31), "int32 digit fold");

typedef int32_t d8_i32x4 __attribute__((ext_vector_type(4)));

/* bin (0..63 within a k-step) of MFMA K index kk (0..63) */
__host__ __device__ inline uint32_t d8_bin_of(uint32_t kk)
{
    const uint32_t kg = kk >> 4, j = kk & 15u, t = j & 7u;
    return (j < 8u ? 0u : 32u) + 8u * kg + (t < 4u ? 2u * t : 2u * (t - 4u) + 1u);
}

/* xq[s][d][K order] = digit d of X, qs[s] = 2^-k_s */
__global__ void __launch_bounds__(256) k_d8_split(int8_t *xq, float *qs, hfx_spectra sp)
{
    __shared__ float red[256 / HFX_WAVE];
    const uint32_t s = blockIdx.x;
    const uint32_t len = min(sp.d_len[s], min(sp.x_stride, (uint32_t)D8_MAXK));
    const float *x = sp.d_x + (size_t)s * sp.x_stride;
    /* the row once, in registers: thread t takes the bins of output words
     * t + 256 u (the K order permutes bins within each 64), which cover
     * every bin once, so max |x| comes from them too */
    constexpr uint32_t NW = D8_MAXK / 4, NU = (NW + 255) / 256;
    float v[NU][4];
    float m = 0.0f;
#pragma unroll
    for (uint32_t u = 0; u < NU; ++u) {
        const uint32_t t = threadIdx.x + 256 * u;
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            const uint32_t pos = 4 * t + i;
            const uint32_t bin = (pos & ~63u) + d8_bin_of(pos & 63u);
            v[u][i] = t < NW && bin < len ? x[bin] : 0.0f;
            m = fmaxf(m, fabsf(v[u][i]));
        }
    }
    for (uint32_t o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (hfx_lane() == 0) red[threadIdx.x / HFX_WAVE] = m;
    __syncthreads();
    m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    /* k: m 2^k in [2^22, 2^23) from m's exponent, one less if that passes
     * 8355711; at most 126 (a spectrum of denormals keeps fewer bits) */
    int k = 0;
    if (m > 0.0f) {
        const int e = (int)((__float_as_uint(m) >> 23) & 255u) - 127;
        k = 22 - e;
        if (k > 126) k = 126;
        else if (ldexpf(m, k) > D8_XMAX) --k;
    }
    const float sc = ldexpf(1.0f, k);
    if (threadIdx.x == 0) qs[s] = ldexpf(1.0f, -k);
    int8_t *o = xq + (size_t)s * 3 * D8_MAXK;
#pragma unroll
    for (uint32_t u = 0; u < NU; ++u) {
        const uint32_t t = threadIdx.x + 256 * u;
        if (t >= NW) break;
        uint32_t w[3] = {0u, 0u, 0u};
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            const int32_t X = (int32_t)rintf(v[u][i] * sc);
            const int32_t d0 = ((X + 128) & 255) - 128;
            const int32_t X1 = (X - d0) >> 8;
            const int32_t d1 = ((X1 + 128) & 255) - 128;
            const int32_t d2 = (X1 - d1) >> 8;
            w[0] |= (uint32_t)(d0 & 255) << (8 * i);
            w[1] |= (uint32_t)(d1 & 255) << (8 * i);
            w[2] |= (uint32_t)(d2 & 255) << (8 * i);
        }
#pragma unroll
        for (uint32_t d = 0; d < 3; ++d) ((uint32_t *)(o + d * D8_MAXK))[t] = w[d];
    }
}

/* four unary codes (one a byte, low nibble) to their weight bytes 0/10/25/50:
 * code c selects byte c of the pool 00 0A -- 19 | -- -- -- 32 */
__device__ __forceinline__ uint32_t d8_wbytes(uint32_t codes)
{
    return __builtin_amdgcn_perm(0x32000000u, 0x19000A00u, codes);
}

struct D8Pipe {
    hfx_dn_pos pos; /* position 64 wave + lane of the tile */
    uint64_t rng;   /* lanes 0, 1: the range of the tile after */
    uint64_t L;     /* first position */
    uint32_t np;    /* number of positions */
    float top;      /* lane < 32: top of candidate lane */
};

struct D8Args {
    float *scores;
    const uint64_t *score_off;
    const uint32_t *begin, *num, *order;
    const uint32_t *bstart; /* [B + 1] block starts in order, or null: sixteen a block */
    uint32_t B, CH;
    const uint32_t *ubeg, *uend, *binfo;
    const uint64_t *tile_off;
    const int8_t *xq;
    const float *qs;
    uint32_t *counter;
    float *dummy;
    hfx_dn_ladder lad;
    hfx_spectra sp;
};

__global__ void __launch_bounds__(D8_THREADS, 1) k_dense8(D8Args a)
{
    extern __shared__ __align__(16) unsigned char d8_lds[];
    uint32_t *maps = (uint32_t *)d8_lds;                          /* [2][64 rows][D8_MSTR] */
    uint64_t *combs = (uint64_t *)(d8_lds + 2 * D8_MAP_BYTES);    /* [2][4][16][16] */
    __shared__ uint32_t m_s[D8_NS], m_beg[D8_NS], m_end[D8_NS], m_z[D8_NS];
    __shared__ float m_q[D8_NS];
    __shared__ uint64_t m_off[D8_NS];
    __shared__ uint32_t sh_t[2];
    __shared__ uint32_t sh_kq[2], sh_kneed[2]; /* of the tile last written into map i */

    const uint32_t tid = threadIdx.x, lane = hfx_lane();
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid / HFX_WAVE);
    const uint32_t grp = wave / D8_GW, gw = wave % D8_GW, gtid = tid % D8_GT;
    for (uint32_t i = tid; i < 2 * D8_ROWS * D8_MSTR; i += D8_THREADS) maps[i] = 0;
    for (uint32_t i = tid; i < 2 * D8_COMB_BYTES / 8; i += D8_THREADS) combs[i] = 0;
    const uint64_t ntiles = dn_sload(a.tile_off, a.B);
    const uint32_t col = lane & 15u, kg = lane >> 4;
    /* B fragments of k-steps gw + 6 j, three digits each */
    d8_i32x4 bq[D8_KPW][3];
    uint32_t cur = DN_NONE;
    if (tid == 0) sh_t[0] = atomicAdd(a.counter, 1u);
    uint32_t next = 0;
    for (uint32_t slot = 0;; slot ^= 1u) {
        __syncthreads();
        const uint32_t t0 = __builtin_amdgcn_readfirstlane(sh_t[slot]);
        if (t0 >= ntiles) break;
        const uint32_t t = (uint32_t)ntiles - 1u - t0;
        if (tid == 0) next = atomicAdd(a.counter, 1u);
        uint32_t lo = 0, hi = a.B;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (dn_sload(a.tile_off, mid) <= t) lo = mid; else hi = mid;
        }
        const uint32_t b = lo;
        const uint32_t info = dn_sload(a.binfo, b);
        const uint32_t kb = info & 0xffffu, mask = info >> 16;
        if (b != cur) {
            if (tid < D8_NS) {
                const uint32_t i = dn_bfirst(a.bstart, b, D8_NS) + tid;
                uint32_t s = DN_NONE, be = 0, en = 0, z = 0;
                uint64_t of = 0;
                float q = 0.0f;
                if (i < dn_bend(a.bstart, b, a.sp.num_spec, D8_NS)) {
                    s = dn_order(a.order, i);
                    be = a.begin[s];
                    en = be + a.num[s];
                    z = hfx_zeff(a.sp.d_zmax[s]);
                    of = a.score_off[s];
                    q = a.qs[s];
                }
                m_s[tid] = s; m_beg[tid] = be; m_end[tid] = en; m_z[tid] = z; m_off[tid] = of;
                m_q[tid] = q;
            }
            __syncthreads();
            /* this wave's B fragments of the block, straight from HBM */
            const uint32_t sc = m_s[col];
#pragma unroll
            for (uint32_t j = 0; j < D8_KPW; ++j) {
                const uint32_t ks = gw + D8_GW * j;
#pragma unroll
                for (uint32_t d = 0; d < 3; ++d) {
                    bq[j][d] = d8_i32x4{0, 0, 0, 0};
                    if (sc != DN_NONE && ks * 64 < kb)
                        bq[j][d] = *(const d8_i32x4 *)(a.xq + ((size_t)sc * 3 + d) * D8_MAXK +
                                                       ks * 64 + kg * 16);
                }
            }
            cur = b;
        }
        if (tid == 0) sh_t[slot ^ 1u] = next;
        const uint32_t beg = dn_sload(a.ubeg, b), end = dn_sload(a.uend, b);
        if (beg < a.lad.q0 || end > a.lad.q0 + a.lad.nq) continue;
        const uint32_t nch = (end - beg + a.CH - 1) / a.CH;
        const uint32_t local = (uint32_t)(t - dn_sload(a.tile_off, b));
        const uint32_t li = local / nch, chunk = local - li * nch;
        uint32_t m = mask;
        for (uint32_t qq = 0; qq < li; ++qq) m &= m - 1u;
        const uint32_t lvl = (uint32_t)__builtin_ctz(m);
        const uint32_t cbeg = beg + chunk * a.CH, cend = min(cbeg + a.CH, end);
        const uint32_t ntile = (cend - cbeg + D8_ROWS - 1) / D8_ROWS;
        float *const dummy = a.dummy + lane;

        auto p0_of = [&](uint32_t k) -> uint32_t { return cbeg + k * D8_ROWS; };
        auto rows_of = [&](uint32_t p) -> uint32_t {
            return p < cend ? min((uint32_t)D8_ROWS, cend - p) : 0u;
        };
        /* lane 0: the tile's first position, lane 1: one past its last */
        auto load_range = [&](uint32_t p) -> uint64_t {
            const uint32_t nr = rows_of(p);
            return a.lad.off[nr ? p + (lane ? nr : 0u) - a.lad.q0 : 0u];
        };
        auto load_top = [&](uint32_t p) -> float {
            return a.lad.top[lane < rows_of(p) ? p + lane - a.lad.q0 : 0u];
        };
        auto lane64 = [&](uint64_t v, uint32_t l) -> uint64_t {
            return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l) |
                   (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l) << 32;
        };
        auto load_pos = [&](uint64_t L0, uint32_t np) -> hfx_dn_pos {
            const hfx_dn_pos *base = np ? a.lad.pos + L0 : a.lad.pos;
            return base[min(HFX_WAVE * gw + lane, np ? np - 1u : 0u)];
        };
        /* this group's pipeline: the tile it writes peaks for next (k0 = grp,
         * then every other tile) is in pc, the range of the one after in rn */
        D8Pipe pc;
        uint64_t rn;
        {
            const uint64_t r0 = load_range(p0_of(grp));
            rn = load_range(p0_of(grp + 2));
            __builtin_amdgcn_sched_barrier(0);
            pc.L = lane64(r0, 0);
            pc.np = (uint32_t)(lane64(r0, 1) - pc.L);
            pc.pos = load_pos(pc.L, pc.np);
            pc.top = load_top(p0_of(grp));
        }
        for (uint32_t k = 0; k < ntile + 2; ++k) {
            const uint32_t mi = k & 1u;
            if (grp == mi) {
                /* 1. the scores of tile k - 2 (its sums, in buffer mi, are
                 * complete since the last barrier) */
                if (k >= 2) {
                    const uint32_t pk = p0_of(k - 2), nr = rows_of(pk);
                    uint64_t *cb = combs + mi * (D8_COMB_BYTES / 8);
                    for (uint32_t o = gtid; o < D8_ROWS * D8_NS; o += D8_GT) {
                        const uint32_t g = o & 15u, r = o >> 4;
                        const int64_t v = (int64_t)cb[o];
                        cb[o] = 0ull;
                        const float sc = (float)((double)v * (double)m_q[g]);
                        const uint32_t p = pk + r;
                        const bool ok = r < nr && m_z[g] == lvl && m_beg[g] <= p && p < m_end[g];
                        *(ok ? a.scores + m_off[g] + (p - m_beg[g]) : dummy) = sc;
                    }
                }
                /* 2. the peaks of tile k into map mi */
                if (k < ntile) {
                    const uint32_t p0 = p0_of(k), nrows = rows_of(p0);
                    /* issue: the range of tile k + 4, the positions and tops of
                     * tile k + 2 (this group's next) */
                    D8Pipe nx;
                    const uint64_t rn2 = load_range(p0_of(k + 4));
                    nx.L = lane64(rn, 0);
                    nx.np = (uint32_t)(lane64(rn, 1) - nx.L);
                    nx.pos = load_pos(nx.L, nx.np);
                    nx.top = load_top(p0_of(k + 2));
                    /* bins of the tile: below top / 1.0005 + 2 for every candidate */
                    const float rm = dn_row_max(lane < nrows ? pc.top : 0.0f);
                    uint32_t mb = (uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(rm), 0);
                    mb = max(mb, (uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(rm), 16));
                    mb = max(mb, (uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(rm), 32));
                    mb = max(mb, (uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(rm), 48));
                    const float mm = __uint_as_float(mb);
                    const uint32_t kneed = (uint32_t)(mm + 24.0f) + 4u;
                    const uint32_t kq = nrows ? min(kb, (kneed + 127u) & ~127u) : 0u;
                    if (gtid == 0) {
                        sh_kq[mi] = kq;
                        sh_kneed[mi] = kneed;
                    }
                    uint32_t *map = maps + mi * (D8_ROWS * D8_MSTR);
                    /* a lane's positions i, i + 384, ...: the first came a
                     * tile ahead, each later one is loaded while the peaks
                     * of the one before are written */
                    auto peaks = [&](auto chk) {
                        constexpr bool CHECK = decltype(chk)::value;
                        constexpr uint32_t STEP = HFX_WAVE * D8_GW;
                        uint32_t i = HFX_WAVE * gw + lane;
                        hfx_dn_pos v = pc.pos;
                        while (i < pc.np) {
                            const uint32_t in = i + STEP;
                            hfx_dn_pos vn = v;
                            if (in < pc.np) vn = a.lad.pos[pc.L + in];
                            dn_peaks_all<CHECK, D8_SPREAD>(map + (v.q - p0) * D8_MSTR + 1, v.b,
                                                           v.y, kq, lane, lvl);
                            v = vn;
                            i = in;
                        }
                    };
                    if (kneed <= kq) peaks(std::false_type{});
                    else peaks(std::true_type{});
                    pc = nx;
                    rn = rn2;
                }
            } else if (k >= 1 && k <= ntile) {
                /* 3. the contraction of tile k - 1 from map mo into sums mo */
                const uint32_t mo = mi ^ 1u;
                const uint32_t nrows = rows_of(p0_of(k - 1));
                const uint32_t kq = __builtin_amdgcn_readfirstlane(sh_kq[mo]);
                const uint32_t kneed = __builtin_amdgcn_readfirstlane(sh_kneed[mo]);
                uint32_t *map = maps + mo * (D8_ROWS * D8_MSTR);
                uint64_t *cb0 = combs + mo * (D8_COMB_BYTES / 8);
                /* the clamped peaks' words past the tile's bins, cleared before
                 * the map's next peaks (after the barrier) */
                if (kneed > kq && gw == D8_GW - 1)
                    for (uint32_t i = lane; i < D8_ROWS * D8_TRASH; i += HFX_WAVE) {
                        const uint32_t r = i / D8_TRASH;
                        map[r * D8_MSTR + 1 + (kq >> 3) + (i - r * D8_TRASH)] = 0u;
                    }
                const uint32_t nks = kq >> 6, nrb = (nrows + 15u) >> 4;
                if (gw < nks) {
                    /* the words of the wave's next k-step (the next row
                     * block's first after a block's last) are loaded before
                     * the current ones are turned into weights */
                    uint32_t n0, n1;
                    {
                        const uint32_t *a0 = map + col * D8_MSTR + 1 + kg + 8 * gw;
                        n0 = a0[0];
                        n1 = a0[4];
                    }
#pragma unroll
                    for (uint32_t r = 0; r < D8_RB; ++r) {
                        if (r < nrb) {
                            d8_i32x4 acc[3] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
                            uint32_t *arow = map + (r * 16u + col) * D8_MSTR + 1 + kg;
#pragma unroll
                            for (uint32_t j = 0; j < D8_KPW; ++j) {
                                const uint32_t ks = gw + D8_GW * j;
                                if (ks < nks) {
                                    uint32_t *aw = arow + 8 * ks;
                                    const uint32_t w0 = n0, w1 = n1;
                                    if (j + 1 < D8_KPW && ks + D8_GW < nks) {
                                        n0 = aw[8 * D8_GW];
                                        n1 = aw[8 * D8_GW + 4];
                                    } else if (r + 1 < nrb) {
                                        n0 = arow[16 * D8_MSTR + 8 * gw];
                                        n1 = arow[16 * D8_MSTR + 8 * gw + 4];
                                    }
                                    aw[0] = 0u;
                                    aw[4] = 0u;
                                    d8_i32x4 av;
                                    av.x = (int32_t)d8_wbytes(w0 & 0x0f0f0f0fu);
                                    av.y = (int32_t)d8_wbytes((w0 >> 4) & 0x0f0f0f0fu);
                                    av.z = (int32_t)d8_wbytes(w1 & 0x0f0f0f0fu);
                                    av.w = (int32_t)d8_wbytes((w1 >> 4) & 0x0f0f0f0fu);
#pragma unroll
                                    for (uint32_t d = 0; d < 3; ++d)
                                        acc[d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                            av, bq[j][d], acc[d], 0, 0, 0);
                                }
                            }
                            /* C[16 r + 4 kg + i][col] of the three digits as
                             * one int64 into the tile's sums: a wave's digit
                             * sums are below 2^31 / 257 in magnitude, so
                             * a1 2^8 + a0 is exact in int32 and one 64-bit
                             * shift-add takes a2 2^16 on */
                            uint64_t *cb = cb0 + r * 256 + (4 * kg) * 16 + col;
#pragma unroll
                            for (uint32_t i = 0; i < 4; ++i) {
                                const int32_t v10 = (int32_t)(((uint32_t)acc[1][i] << 8) +
                                                              (uint32_t)acc[0][i]);
                                const int64_t v = (int64_t)(((uint64_t)(int64_t)acc[2][i] << 16) +
                                                            (uint64_t)(int64_t)v10);
                                atomicAdd((unsigned long long *)(cb + 16 * i),
                                          (unsigned long long)v);
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
}

/* ------------------------------------------------------------------ */
/* Ladders                                                                */
/* ------------------------------------------------------------------ */

__global__ void k_dn_lad_count(uint32_t *cnt, hfx_db db, uint32_t q0, uint32_t nq)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const uint32_t pep = db.d_idx_sorted[q0 + i];
    const uint32_t tc = db.d_tc[pep], tn = db.d_tn[pep];
    cnt[i] = tn > tc ? tn - tc : 0u;
}

/* one thread per peptide: the sequential FP32 b ladder of the sparse path
 * (residue masses from the [26] table by index byte & 31, zero past 25) and
 * y = r - b */
__global__ void k_dn_lad_fill(hfx_dn_pos *pos, float *top, const uint64_t *off, hfx_db db,
                              uint32_t q0, uint32_t nq)
{
    __shared__ float mt[32];
    if (threadIdx.x < 32) mt[threadIdx.x] = threadIdx.x < 26 ? db.d_mass[threadIdx.x] : 0.0f;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const uint32_t pep = db.d_idx_sorted[q0 + i];
    const uint32_t tc = db.d_tc[pep];
    const float r = db.d_r[pep];
    const uint64_t o = off[i];
    const uint32_t n = (uint32_t)(off[i + 1] - o);
    float b = 0.0f, b0 = 0.0f;
    for (uint32_t k = 0; k < n; ++k) {
        b = b + mt[HFX_ACID_IDX(db.d_ions[(uint64_t)tc + k])];
        if (k == 0) b0 = b;
        pos[o + k] = hfx_dn_pos{b, r - b, q0 + i};
    }
    /* with a first residue of 32 Da or more every peak bin is >= 1 (the
     * lowest, y - NH3 at charge 4, is above 1.2); a lighter one (a zero-mass
     * residue code) marks the peptide's tiles for the clamped writes */
    top[i] = n && b0 < 32.0f ? 1e9f : fmaxf(r, b);
}

int hfx_dense_ladder_build(hfx_ctx *ctx, const hfx_db *db, uint32_t q0, uint32_t nq, int own,
                           hfx_dn_ladder *lad, void **mem, void *stream)
{
    HFX_REQUIRE(ctx && db && lad && (!own || mem), HFX_ERR_ARG, "dense ladder: null argument");
    HFX_REQUIRE((uint64_t)q0 + nq <= db->num_pep, HFX_ERR_ARG,
                "dense ladder: peptides %u..%u beyond the database's %u", q0, q0 + nq,
                db->num_pep);
    if (own) mem[0] = mem[1] = nullptr;
    hipStream_t st = hfx_stream(stream);
    const size_t meta = ((size_t)nq + 1) * 8 + (size_t)nq * 8 + 64;
    void *m = nullptr;
    int rc;
    if (own) {
        HFX_CHECK_HIP(hipMalloc(&m, meta));
        mem[0] = m;
    } else if ((rc = hfx_ctx_scratch(ctx, HFX_SLOT_LADDER, meta, &m))) {
        return rc;
    }
    uint64_t *off = (uint64_t *)m;
    float *top = (float *)(off + nq + 1);
    uint32_t *cnt = (uint32_t *)(top + nq);
    const uint32_t g = (nq + 255) / 256;
    if (nq) hipLaunchKernelGGL(k_dn_lad_count, dim3(g), dim3(256), 0, st, cnt, *db, q0, nq);
    if ((rc = hfx_scan_offsets(ctx, off, cnt, nq, stream))) return rc;
    uint64_t total = 0;
    HFX_CHECK_HIP(hipMemcpyAsync(&total, off + nq, 8, hipMemcpyDeviceToHost, st));
    HFX_CHECK_HIP(hipStreamSynchronize(st));
    /* at least one position: the kernel's surplus loads read position 0 */
    const size_t pb = (size_t)std::max<uint64_t>(total, 1) * sizeof(hfx_dn_pos);
    void *pp = nullptr;
    if (own) {
        HFX_CHECK_HIP(hipMalloc(&pp, pb));
        mem[1] = pp;
    } else if ((rc = hfx_ctx_scratch(ctx, HFX_SLOT_LADDER2, pb, &pp))) {
        return rc;
    }
    if (nq)
        hipLaunchKernelGGL(k_dn_lad_fill, dim3(g), dim3(256), 0, st, (hfx_dn_pos *)pp, top, off,
                           *db, q0, nq);
    HFX_CHECK_LAUNCH();
    lad->pos = (const hfx_dn_pos *)pp;
    lad->off = off;
    lad->top = top;
    lad->q0 = q0;
    lad->nq = nq;
    return HFX_OK;
}

size_t hfx_dense_scratch_bytes(uint32_t S)
{
    const size_t B = S; /* blocks of at least one spectrum */
    return (size_t)S * 2 * DN_MAXK * 2 + B * 4 * 4 + (B + 1) * 8 + 1024;
}

uint32_t hfx_dense_max_len(void) { return DN_MAXK; }

int hfx_dense_split(const hfx_spectra *sp, void *scratch, void *stream)
{
    HFX_REQUIRE(sp && scratch, HFX_ERR_ARG, "dense split: null argument");
    HFX_REQUIRE(sp->x_stride <= DN_MAXK, HFX_ERR_ARG,
                "dense XCorr: spectra of up to %d bins (x_stride %u)", DN_MAXK, sp->x_stride);
    if (sp->num_spec)
        hipLaunchKernelGGL(k_dn_split, dim3(sp->num_spec), dim3(256), 0, hfx_stream(stream),
                           (uint16_t *)scratch, *sp);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

int hfx_dense_scores(hfx_ctx *ctx, float *d_scores, const uint64_t *d_off, const uint32_t *d_begin,
                     const uint32_t *d_num, const uint32_t *d_order, const uint32_t *d_bstart,
                     uint32_t nblocks, const hfx_dn_ladder *lad, const hfx_spectra *sp,
                     uint32_t CH, void *scratch, int presplit, void *stream, void **ev)
{
    HFX_REQUIRE(ctx && d_scores && d_off && d_begin && d_num && lad && sp, HFX_ERR_ARG,
                "dense XCorr: null argument");
    HFX_REQUIRE(sp->x_stride <= DN_MAXK, HFX_ERR_ARG,
                "dense XCorr: spectra of up to %d bins (x_stride %u)", DN_MAXK, sp->x_stride);
    HFX_REQUIRE(CH >= DN_ROWS && CH % DN_ROWS == 0 && CH <= 8192, HFX_ERR_ARG,
                "dense XCorr: chunk %u must be a multiple of %d in %d..8192", CH, DN_ROWS, DN_ROWS);
    const uint32_t S = sp->num_spec;
    if (S == 0) return HFX_OK;
    hipStream_t st = hfx_stream(stream);
    void *p = scratch;
    int rc;
    if (!p && (rc = hfx_ctx_scratch(ctx, HFX_SLOT_DENSE, hfx_dense_scratch_bytes(S), &p)))
        return rc;
    HFX_REQUIRE(!d_bstart || (nblocks >= 1 && nblocks <= S), HFX_ERR_ARG,
                "dense XCorr: %u blocks for %u spectra", nblocks, S);
    const uint32_t B = d_bstart ? nblocks : (S + DN_NS - 1) / DN_NS;
    uint16_t *xb = (uint16_t *)p;
    uint32_t *ubeg = (uint32_t *)(xb + (size_t)S * 2 * DN_MAXK);
    uint32_t *uend = ubeg + B, *binfo = uend + B, *ntiles = binfo + B;
    uint64_t *tile_off = (uint64_t *)(((uintptr_t)(ntiles + B) + 7) & ~(uintptr_t)7);
    uint32_t *counter = (uint32_t *)(tile_off + B + 1);
    float *dummy = (float *)(((uintptr_t)(counter + 1) + 255) & ~(uintptr_t)255);
    if (!presplit) hipLaunchKernelGGL(k_dn_split, dim3(S), dim3(256), 0, st, xb, *sp);
    hipLaunchKernelGGL(k_dn_blocks<DN_NS>, dim3((B + 255) / 256), dim3(256), 0, st, ubeg, uend, binfo,
                       ntiles, d_order, d_bstart, d_begin, d_num, *sp, B, CH);
    if ((rc = hfx_scan_offsets(ctx, tile_off, ntiles, B, (void *)st))) return rc;
    hipLaunchKernelGGL(k_dn_reset, dim3(1), dim3(1), 0, st, counter);
    DnArgs args{d_scores, d_off,    d_begin, d_num,   d_order, d_bstart, B,    CH,
                ubeg,     uend,     binfo,   tile_off, xb,      counter,  dummy, *lad,
                *sp};
    HFX_CHECK_HIP(hipFuncSetAttribute((const void *)k_dense,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, DN_LDS_BYTES));
    if (ev) HFX_CHECK_HIP(hipEventRecord((hipEvent_t)ev[0], st));
    hipLaunchKernelGGL(k_dense, dim3((uint32_t)hfx_ctx_num_cus(ctx)), dim3(DN_THREADS),
                       DN_LDS_BYTES, st, args);
    if (ev) HFX_CHECK_HIP(hipEventRecord((hipEvent_t)ev[1], st));
    hipLaunchKernelGGL(k_dn_zero_unscored, dim3(S), dim3(256), 0, st, d_scores, d_off, sp->d_zmax,
                       S);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

extern "C" int hfx_xcorr_dense_bf16x2(hfx_ctx *ctx, float *d_scores, const uint64_t *d_off,
                                      const uint32_t *d_begin, const uint32_t *d_num,
                                      const uint32_t *d_order, const hfx_db *db,
                                      const hfx_spectra *sp, void *stream)
{
    HFX_REQUIRE(db, HFX_ERR_ARG, "hfx_xcorr_dense_bf16x2: null database");
    if (sp && sp->num_spec == 0) return HFX_OK;
    /* ladders of the whole database, rebuilt on every call (a batch plan
     * builds those of its windows once) */
    hfx_dn_ladder lad;
    int rc;
    if ((rc = hfx_dense_ladder_build(ctx, db, 0, db->num_pep, 0, &lad, nullptr, stream)))
        return rc;
    return hfx_dense_scores(ctx, d_scores, d_off, d_begin, d_num, d_order, nullptr, 0, &lad, sp,
                            256, nullptr, 0, stream, nullptr);
}

/* ------------------------------------------------------------------ */
/* Integer form: host side                                               */
/* ------------------------------------------------------------------ */

size_t hfx_dense8_scratch_bytes(uint32_t S)
{
    const size_t B = S; /* blocks of at least one spectrum */
    return (size_t)S * 3 * D8_MAXK + (size_t)S * 4 + 256 + B * 4 * 4 + (B + 1) * 8 + 1024;
}

/* scratch layout: xq [S][3][D8_MAXK] | qs [S] | block arrays */
static void d8_carve(void *p, uint32_t S, int8_t **xq, float **qs, uint8_t **rest)
{
    *xq = (int8_t *)p;
    *qs = (float *)(((uintptr_t)(*xq + (size_t)S * 3 * D8_MAXK) + 255) & ~(uintptr_t)255);
    *rest = (uint8_t *)(*qs + S);
}

int hfx_dense8_split(const hfx_spectra *sp, void *scratch, void *stream)
{
    HFX_REQUIRE(sp && scratch, HFX_ERR_ARG, "dense split: null argument");
    HFX_REQUIRE(sp->x_stride <= D8_MAXK, HFX_ERR_ARG,
                "dense XCorr: spectra of up to %d bins (x_stride %u)", D8_MAXK, sp->x_stride);
    int8_t *xq;
    float *qs;
    uint8_t *rest;
    d8_carve(scratch, sp->num_spec, &xq, &qs, &rest);
    if (sp->num_spec)
        hipLaunchKernelGGL(k_d8_split, dim3(sp->num_spec), dim3(256), 0, hfx_stream(stream), xq,
                           qs, *sp);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

int hfx_dense8_scores(hfx_ctx *ctx, float *d_scores, const uint64_t *d_off,
                      const uint32_t *d_begin, const uint32_t *d_num, const uint32_t *d_order,
                      const uint32_t *d_bstart, uint32_t nblocks, const hfx_dn_ladder *lad,
                      const hfx_spectra *sp, uint32_t CH, void *scratch, int presplit,
                      void *stream, void **ev)
{
    HFX_REQUIRE(ctx && d_scores && d_off && d_begin && d_num && lad && sp, HFX_ERR_ARG,
                "dense XCorr: null argument");
    HFX_REQUIRE(sp->x_stride <= D8_MAXK, HFX_ERR_ARG,
                "dense XCorr: spectra of up to %d bins (x_stride %u)", D8_MAXK, sp->x_stride);
    HFX_REQUIRE(CH >= D8_ROWS && CH % D8_ROWS == 0 && CH <= 8192, HFX_ERR_ARG,
                "dense XCorr: chunk %u must be a multiple of %d in %d..8192", CH, D8_ROWS,
                D8_ROWS);
    const uint32_t S = sp->num_spec;
    if (S == 0) return HFX_OK;
    hipStream_t st = hfx_stream(stream);
    void *p = scratch;
    int rc;
    if (!p && (rc = hfx_ctx_scratch(ctx, HFX_SLOT_DENSE, hfx_dense8_scratch_bytes(S), &p)))
        return rc;
    HFX_REQUIRE(!d_bstart || (nblocks >= 1 && nblocks <= S), HFX_ERR_ARG,
                "dense XCorr: %u blocks for %u spectra", nblocks, S);
    const uint32_t B = d_bstart ? nblocks : (S + D8_NS - 1) / D8_NS;
    int8_t *xq;
    float *qs;
    uint8_t *rest;
    d8_carve(p, S, &xq, &qs, &rest);
    uint32_t *ubeg = (uint32_t *)rest;
    uint32_t *uend = ubeg + B, *binfo = uend + B, *ntiles = binfo + B;
    uint64_t *tile_off = (uint64_t *)(((uintptr_t)(ntiles + B) + 7) & ~(uintptr_t)7);
    uint32_t *counter = (uint32_t *)(tile_off + B + 1);
    float *dummy = (float *)(((uintptr_t)(counter + 1) + 255) & ~(uintptr_t)255);
    if (!presplit) hipLaunchKernelGGL(k_d8_split, dim3(S), dim3(256), 0, st, xq, qs, *sp);
    hipLaunchKernelGGL(k_dn_blocks<D8_NS>, dim3((B + 255) / 256), dim3(256), 0, st, ubeg, uend,
                       binfo, ntiles, d_order, d_bstart, d_begin, d_num, *sp, B, CH);
    if ((rc = hfx_scan_offsets(ctx, tile_off, ntiles, B, (void *)st))) return rc;
    hipLaunchKernelGGL(k_dn_reset, dim3(1), dim3(1), 0, st, counter);
    D8Args args{d_scores, d_off, d_begin,  d_num, d_order, d_bstart, B,    CH,  ubeg,
                uend,     binfo, tile_off, xq,    qs,      counter,  dummy, *lad, *sp};
    HFX_CHECK_HIP(hipFuncSetAttribute((const void *)k_dense8,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, D8_LDS_BYTES));
    if (ev) HFX_CHECK_HIP(hipEventRecord((hipEvent_t)ev[0], st));
    hipLaunchKernelGGL(k_dense8, dim3((uint32_t)hfx_ctx_num_cus(ctx)), dim3(D8_THREADS),
                       D8_LDS_BYTES, st, args);
    if (ev) HFX_CHECK_HIP(hipEventRecord((hipEvent_t)ev[1], st));
    hipLaunchKernelGGL(k_dn_zero_unscored, dim3(S), dim3(256), 0, st, d_scores, d_off, sp->d_zmax,
                       S);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

extern "C" int hfx_xcorr_dense_i8x3(hfx_ctx *ctx, float *d_scores, const uint64_t *d_off,
                                    const uint32_t *d_begin, const uint32_t *d_num,
                                    const uint32_t *d_order, const hfx_db *db,
                                    const hfx_spectra *sp, void *stream)
{
    HFX_REQUIRE(db, HFX_ERR_ARG, "hfx_xcorr_dense_i8x3: null database");
    if (sp && sp->num_spec == 0) return HFX_OK;
    hfx_dn_ladder lad;
    int rc;
    if ((rc = hfx_dense_ladder_build(ctx, db, 0, db->num_pep, 0, &lad, nullptr, stream)))
        return rc;
    return hfx_dense8_scores(ctx, d_scores, d_off, d_begin, d_num, d_order, nullptr, 0, &lad, sp,
                             256, nullptr, 0, stream, nullptr);
}
