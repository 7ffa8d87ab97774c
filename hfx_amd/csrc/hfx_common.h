/*
 * hfx_common.h -- shared device/host definitions of the MI355X SEQUEST core.
 *
 * Canonical arithmetic (identical on the CPU oracle and on gfx950):
 *   FP32 with one IEEE rounding per operation (built with -ffp-contract=off),
 *   correctly rounded division (HIP's default for FP32 '/'), rint (half to even)
 *   for every m/z -> bin conversion, b ladder = sequential FP32 prefix.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/hfx_hip.h"

/* ------------------------------------------------------------------ */
/* Constants                                                            */
/* ------------------------------------------------------------------ */

/* Device ion constants, ion_series.h:16-26. */
#define HFX_BIN_WIDTH_MONO 1.0005079f
#define HFX_BIN_WIDTH_AVG 1.0011413f
#define HFX_DEV_H2O 18.01056f
#define HFX_DEV_NH3 17.02655f
#define HFX_DEV_CO 27.9949f
#define HFX_DEV_H 1.0078246f

/* Largest spectrum length (bins) the LDS-resident paths accept. */
#define HFX_MAX_LEN 4096
/* Words of a per-wave peak map: one nibble per bin. */
#define HFX_MAP_WORDS (HFX_MAX_LEN / 8)
/* Highest fragment charge handled (ion_series.cu:334-342 supports 1..4). */
#define HFX_MAX_CHARGE 4
/* Fixed-point scale of the exact score: round(w*x*2^36) summed in int64. */
#define HFX_FX_SCALE 68719476736.0

#define HFX_WAVE 64

/* ion byte packing, utils.h:92-101 */
#define HFX_ACID_IDX(i) ((i)&0x1Fu)
#define HFX_ACID_MOD(i) (((i)&0xE0u) >> 5)

/* ------------------------------------------------------------------ */
/* Error handling (host)                                                */
/* ------------------------------------------------------------------ */

void hfx_set_error(const char *fmt, ...);

#define HFX_CHECK_HIP(expr)                                                          \
    do {                                                                             \
        hipError_t e__ = (expr);                                                     \
        if (e__ != hipSuccess) {                                                     \
            hfx_set_error("%s:%d: %s: %s", __FILE__, __LINE__, #expr,                \
                          hipGetErrorString(e__));                                   \
            return HFX_ERR_HIP;                                                      \
        }                                                                            \
    } while (0)

#define HFX_CHECK_LAUNCH() HFX_CHECK_HIP(hipGetLastError())

#define HFX_REQUIRE(cond, code, ...)                                                 \
    do {                                                                             \
        if (!(cond)) {                                                               \
            hfx_set_error(__VA_ARGS__);                                              \
            return (code);                                                           \
        }                                                                            \
    } while (0)

static inline hipStream_t hfx_stream(void *s) { return (hipStream_t)s; }

/* Scratch slots of a context (one buffer per slot, grown on demand, never
 * shrunk); entry points that may run back to back use distinct slots. */
enum {
    HFX_SLOT_SORT = 0,    /* radix sort ping-pong buffers        */
    HFX_SLOT_SUM = 1,     /* device-side reductions to the host  */
    HFX_SLOT_SCAN = 2,    /* prepare_scoring offsets             */
    HFX_SLOT_TILES = 3,   /* xcorr tile offsets                  */
    HFX_SLOT_TOPN = 4,    /* top-N bookkeeping                   */
    HFX_SLOT_MVM = 5,     /* addIons / dense helpers             */
    HFX_SLOT_MODS = 6,    /* modified-peptide enumeration        */
    HFX_SLOT_MODS2 = 7,
    HFX_SLOT_BATCH = 8,   /* batched path                        */
    HFX_SLOT_BATCH2 = 9,
    HFX_SLOT_DENSE = 10,  /* BF16 dense path                     */
    HFX_SLOT_FLAGS = 11,  /* spectrum lists turned into flags    */
    HFX_SLOT_LADDER = 12, /* dense ladders: offsets, tops, counts */
    HFX_SLOT_LADDER2 = 13, /* dense ladders: positions           */
    HFX_SLOT_STATS = 14,  /* SP ion peak flags                   */
    HFX_SLOT_SCAN2 = 15,  /* tile sums of the in-place scans     */
};
/* Scratch owned by a context: grown on demand, never shrunk. */
int hfx_ctx_scratch(hfx_ctx *ctx, int slot, size_t bytes, void **out);
int hfx_ctx_pinned(hfx_ctx *ctx, size_t bytes, void **out);
int hfx_ctx_num_cus(hfx_ctx *ctx);
hfx_ctx *hfx_default_ctx(void);
/* sum of n uint32 (wrapping) read back synchronously (null stream) */
uint32_t hfx_sum_u32_to_host(hfx_ctx *ctx, const uint32_t *d_a, uint32_t n);
/* exclusive scan in place, one workgroup */
int hfx_scan_u32_inplace(uint32_t *d_a, uint32_t n, void *stream);
/* device scratch the top-N exact route needs for S spectra, total scores */
size_t hfx_topn_scratch_bytes(uint32_t S, uint64_t total);
/* hfx_topn over the unmodified candidates (batched path; no host sync), with
 * the error bound (rel0 + rel1 * fragment charges) * 50 * sum|x|
 * + relq * max|x| * 50 * len + 1e-6.
 * scratch: hfx_topn_scratch_bytes(S, total) bytes of device memory, or null
 * for the context's own slot */
int hfx_topn_device(hfx_ctx *ctx, uint32_t *d_top_pos, double *d_top_exact, float *d_top_f32,
                    uint32_t *d_top_cnt, uint32_t *d_flag, uint32_t N, uint32_t margin,
                    const float *d_scores, const uint64_t *d_off, uint64_t total,
                    const uint32_t *d_begin, const uint32_t *d_num, const hfx_db *db,
                    const hfx_spectra *sp, float rel0, float rel1, float relq, void *scratch,
                    void *stream);
/* b ladders of the peptides in residual-sorted positions [q0, q0 + nq), for
 * the dense path.  Position k < tn - tc of sorted peptide q is
 * pos[off[q - q0] + k] = {b_k, r - b_k, q}, with b_k the sequential FP32 sum
 * of its residue masses 0..k; top[q - q0] = max(r, b_last) bounds every
 * fragment mass of the peptide (1e9 when a first residue under 32 Da could
 * put a peak below bin 1). */
struct hfx_dn_pos {
    float b, y;
    uint32_t q;
};
struct hfx_dn_ladder {
    const hfx_dn_pos *pos;
    const uint64_t *off; /* [nq + 1] */
    const float *top;    /* [nq] */
    uint32_t q0, nq;
};
/* builds the ladder of sorted peptides [q0, q0 + nq) on the stream and waits
 * for it.  own != 0: two buffers from hipMalloc returned in mem[0..1], which
 * the caller frees; own == 0: the context's ladder slots (valid until the
 * next build on the context). */
int hfx_dense_ladder_build(hfx_ctx *ctx, const hfx_db *db, uint32_t q0, uint32_t nq, int own,
                           hfx_dn_ladder *lad, void **mem, void *stream);
/* dense BF16 hi+lo scores (same layout as hfx_xcorr_sparse); spectra are
 * blocked in d_order (null: index order): block b holds order positions
 * [d_bstart[b], d_bstart[b+1]) of nblocks blocks of at most eight, or eight
 * at a time when d_bstart is null; CH candidates per work item; every
 * window lies inside the ladder's range; scratch:
 * hfx_dense_scratch_bytes(S) bytes or null for the context's slot, whose
 * hi/lo rows hfx_dense_split writes (presplit != 0: already written for
 * these spectra, as a batch plan does once) */
size_t hfx_dense_scratch_bytes(uint32_t S);
uint32_t hfx_dense_max_len(void);
int hfx_dense_split(const hfx_spectra *sp, void *scratch, void *stream);
/* the integer form (int8 MFMA over three base-256 digits of each spectrum,
 * blocks of at most sixteen, chunks a multiple of 64): same arguments;
 * scratch of hfx_dense8_scratch_bytes(S), digits written by
 * hfx_dense8_split */
size_t hfx_dense8_scratch_bytes(uint32_t S);
int hfx_dense8_split(const hfx_spectra *sp, void *scratch, void *stream);
int hfx_dense8_scores(hfx_ctx *ctx, float *d_scores, const uint64_t *d_off,
                      const uint32_t *d_begin, const uint32_t *d_num, const uint32_t *d_order,
                      const uint32_t *d_bstart, uint32_t nblocks, const hfx_dn_ladder *lad,
                      const hfx_spectra *sp, uint32_t CH, void *scratch, int presplit,
                      void *stream, void **ev);
/* ev (optional): two hipEvent_t recorded right before and after the
 * contraction kernel itself */
int hfx_dense_scores(hfx_ctx *ctx, float *d_scores, const uint64_t *d_off, const uint32_t *d_begin,
                     const uint32_t *d_num, const uint32_t *d_order, const uint32_t *d_bstart,
                     uint32_t nblocks, const hfx_dn_ladder *lad, const hfx_spectra *sp,
                     uint32_t CH, void *scratch, int presplit, void *stream, void **ev);

/* ------------------------------------------------------------------ */
/* Device helpers                                                        */
/* ------------------------------------------------------------------ */

#ifdef __HIP_DEVICE_COMPILE__
#define HFX_DEV __device__ __forceinline__
#else
#define HFX_DEV __device__ __host__ inline
#endif

/* binMZ with IEEE division (ion_series.cu:98-102): rint(mz / 1.0005079). */
__device__ __forceinline__ int32_t hfx_bin(float mz)
{
    return (int32_t)__builtin_rintf(mz / HFX_BIN_WIDTH_MONO);
}

/* x / d correctly rounded, given R = RN(1/d): q = RN(x R) is within an ulp,
 * the residual x - q d is exact in one FMA, and one Newton step rounds to
 * the nearest quotient.  Holds for every |x| in [1e-3, 65536) for the bin
 * width and d = 1..4 (exhaustive CPU check); below 1e-3 both sides round to
 * bin 0, above 65536 every bin lies beyond the row. */
__device__ __forceinline__ float hfx_div(float x, float d, float R)
{
    const float q = x * R;
    const float r = __builtin_fmaf(-q, d, x);
    return __builtin_fmaf(r, R, q);
}

/* hfx_bin through hfx_div: the same bin for every theoretical peak */
__device__ __forceinline__ int32_t hfx_bin_q(float mz)
{
    return (int32_t)__builtin_rintf(
        hfx_div(mz, HFX_BIN_WIDTH_MONO, 1.0f / HFX_BIN_WIDTH_MONO));
}

/* Unary weight codes: OR of codes == code of the maximum weight
 * (10 -> 0b001, 25 -> 0b011, 50 -> 0b111); weight = {0,10,25,50}[popc]. */
#define HFX_CODE10 1u
#define HFX_CODE25 3u
#define HFX_CODE50 7u

__device__ __forceinline__ float hfx_code_weight(uint32_t nib)
{
    uint32_t c = __builtin_popcount(nib);
    return c == 3 ? 50.0f : (c == 2 ? 25.0f : (c == 1 ? 10.0f : 0.0f));
}

/* Order-preserving map of an FP32 key to uint32 (larger float -> larger). */
__device__ __forceinline__ uint32_t hfx_f2key(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float hfx_key2f(uint32_t k)
{
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

__device__ __forceinline__ int64_t hfx_fx_term(float w, float x)
{
    double p = (double)w * (double)x * HFX_FX_SCALE; /* exact: 6 x 24 bits, 2^36 */
    return (int64_t)__builtin_rint(p);
}

__device__ __forceinline__ void hfx_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float hfx_wave_sum_f32(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int64_t hfx_wave_sum_i64(int64_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t hfx_lane() { return __lane_id(); }

/*
 * Theoretical peaks of one candidate into a per-wave LDS nibble map.
 *
 * Residue masses come in chunks of 64 from `mass_of(c0, cn)`, called by every
 * lane of the wave (so it may use wave-wide ballots); lane j returns the mass
 * of residue c0 + j for j < cn and 0 otherwise.  The b ladder is the
 * sequential FP32 prefix, reproduced exactly by broadcasting each residue mass
 * to every lane and letting lane k add masses 0..k in order.  For each ladder
 * position and each charge zlo..zmax the peaks of addIonsAB/addIonsY
 * (ion_series.cu:111-156) are OR-ed into `map` (bins >= len dropped, max per
 * bin through the unary weight code).
 */
template <class MassFn>
__device__ __forceinline__ void hfx_wave_peaks_z(uint32_t *map, uint32_t n, float r, uint32_t zlo,
                                                 uint32_t zmax, uint32_t len, MassFn &&mass_of);
template <class MassFn>
__device__ __forceinline__ void hfx_wave_peaks(uint32_t *map, uint32_t n, float r,
                                               uint32_t zmax, uint32_t len, MassFn &&mass_of)
{
    hfx_wave_peaks_z(map, n, r, 1u, zmax, len, mass_of);
}
template <class MassFn>
__device__ __forceinline__ void hfx_wave_peaks_z(uint32_t *map, uint32_t n, float r, uint32_t zlo,
                                                 uint32_t zmax, uint32_t len, MassFn &&mass_of)
{
    const uint32_t lane = hfx_lane();
    float carry = 0.0f;
    for (uint32_t c0 = 0; c0 < n; c0 += HFX_WAVE) {
        const uint32_t cn = min((uint32_t)HFX_WAVE, n - c0);
        const float m = mass_of(c0, cn);
        float b = carry;
        for (uint32_t j = 0; j < cn; ++j) {
            const float mj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m), j));
            if (j <= lane) b = b + mj;
        }
        carry = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b), cn - 1));
        if (lane < cn) {
            const float y = r - b;
            for (uint32_t z = zlo; z <= zmax; ++z) {
                const float fz = (float)z;
                const float zh = HFX_DEV_H * fz;
                /* divisions by z (1..4) through hfx_div: correctly rounded */
                const float rz = 1.0f / fz;
                int32_t x;
#define HFX_PUT(xx, code)                                                            \
    do {                                                                             \
        x = (xx);                                                                    \
        if (x >= 0 && (uint32_t)x < len)                                             \
            atomicOr(&map[(uint32_t)x >> 3], (code) << (((uint32_t)x & 7u) * 4u));     \
    } while (0)
                const float h2o = hfx_div(HFX_DEV_H2O, fz, rz);
                const float nh3 = hfx_div(HFX_DEV_NH3, fz, rz);
                HFX_PUT(hfx_bin_q(hfx_div((b - HFX_DEV_CO) + zh, fz, rz)), HFX_CODE10);
                const float mb = hfx_div(b + zh, fz, rz);
                const int32_t xb = hfx_bin_q(mb);
                HFX_PUT(xb, HFX_CODE50);
                HFX_PUT(xb + 1, HFX_CODE25);
                HFX_PUT(xb - 1, HFX_CODE25);
                HFX_PUT(hfx_bin_q(mb - h2o), HFX_CODE10);
                HFX_PUT(hfx_bin_q(mb - nh3), HFX_CODE10);
                const float my = hfx_div((y + HFX_DEV_H2O) + zh, fz, rz);
                const int32_t xy = hfx_bin_q(my);
                HFX_PUT(xy, HFX_CODE50);
                HFX_PUT(xy + 1, HFX_CODE25);
                HFX_PUT(xy - 1, HFX_CODE25);
                HFX_PUT(hfx_bin_q(my - nh3), HFX_CODE10);
#undef HFX_PUT
            }
        }
    }
}

/* Dot of the map with the LDS spectrum, FP32, clearing the map as it goes.
 * Lane l scans words l, l+64, ... (conflict-free), bins ascending. */
__device__ __forceinline__ float hfx_wave_dot_f32(uint32_t *map, const float *xs, uint32_t len)
{
    const uint32_t lane = hfx_lane();
    const uint32_t nw = (len + 7) >> 3;
    float acc = 0.0f;
    for (uint32_t w = lane; w < nw; w += HFX_WAVE) {
        uint32_t word = map[w];
        if (word) {
            map[w] = 0;
            while (word) {
                uint32_t sh = __builtin_ctz(word) & ~3u;
                uint32_t nib = (word >> sh) & 0xFu;
                word &= ~(0xFu << sh);
                acc = acc + hfx_code_weight(nib) * xs[(w << 3) + (sh >> 2)];
            }
        }
    }
    return hfx_wave_sum_f32(acc);
}

/* FP32 and exact dots of the map in one pass, clearing it. */
__device__ __forceinline__ void hfx_wave_dot_both(uint32_t *map, const float *xs, uint32_t len,
                                                  float *f32_out, int64_t *fx_out)
{
    const uint32_t lane = hfx_lane();
    const uint32_t nw = (len + 7) >> 3;
    float acc = 0.0f;
    int64_t fx = 0;
    for (uint32_t w = lane; w < nw; w += HFX_WAVE) {
        uint32_t word = map[w];
        if (word) {
            map[w] = 0;
            while (word) {
                uint32_t sh = __builtin_ctz(word) & ~3u;
                uint32_t nib = (word >> sh) & 0xFu;
                word &= ~(0xFu << sh);
                const float wt = hfx_code_weight(nib);
                const float xv = xs[(w << 3) + (sh >> 2)];
                acc = acc + wt * xv;
                fx += hfx_fx_term(wt, xv);
            }
        }
    }
    *f32_out = hfx_wave_sum_f32(acc);
    *fx_out = hfx_wave_sum_i64(fx);
}

/* Exact (fixed-point) dot of the map with the LDS spectrum, clearing it. */
__device__ __forceinline__ int64_t hfx_wave_dot_fx(uint32_t *map, const float *xs, uint32_t len)
{
    const uint32_t lane = hfx_lane();
    const uint32_t nw = (len + 7) >> 3;
    int64_t acc = 0;
    for (uint32_t w = lane; w < nw; w += HFX_WAVE) {
        uint32_t word = map[w];
        if (word) {
            map[w] = 0;
            while (word) {
                uint32_t sh = __builtin_ctz(word) & ~3u;
                uint32_t nib = (word >> sh) & 0xFu;
                word &= ~(0xFu << sh);
                acc += hfx_fx_term(hfx_code_weight(nib), xs[(w << 3) + (sh >> 2)]);
            }
        }
    }
    return hfx_wave_sum_i64(acc);
}

/* Residue masses of an unmodified candidate (ion_series.cu:163-168). */
struct HfxPlainMass {
    const uint8_t *ions; /* first residue of the candidate */
    const float *mtab;   /* LDS copy of the 26-entry table */
    __device__ __forceinline__ float operator()(uint32_t c0, uint32_t cn) const
    {
        const uint32_t lane = hfx_lane();
        return lane < cn ? mtab[HFX_ACID_IDX(ions[c0 + lane])] : 0.0f;
    }
};

/* HfxPlainMass with the candidate's first 64 residue codes already in hand
 * (lane j's in first), so that their load can be issued a candidate ahead */
struct HfxPlainMassFirst {
    const uint8_t *ions;
    const float *mtab;
    uint32_t first;
    __device__ __forceinline__ float operator()(uint32_t c0, uint32_t cn) const
    {
        const uint32_t lane = hfx_lane();
        if (lane >= cn) return 0.0f;
        return mtab[HFX_ACID_IDX(c0 == 0 ? first : (uint32_t)ions[c0 + lane])];
    }
};

/* Residue masses of a modified candidate: a residue packed with modable-acid
 * index a+1 (prepare_ions) that is the i-th occurrence of acid a in the
 * ladder, with i listed in this peptide's unrank slice for a, weighs
 * table + ma_mass[a] (mion_series.cu:243-286).  The occurrence index counts
 * acid a only (check_series.cu:61-99), carried across 64-residue chunks. */
#define HFX_MAX_MA 7
struct HfxModMass {
    const uint8_t *ions;
    const float *mtab;
    const uint32_t *mod_row; /* counts of each modable acid in this modification */
    const uint32_t *unrank;  /* this modified peptide's unrank slice */
    const float *ma_mass;
    uint32_t num_ma;
    uint32_t occ[HFX_MAX_MA];
    __device__ __forceinline__ float operator()(uint32_t c0, uint32_t cn)
    {
        const uint32_t lane = hfx_lane();
        const bool act = lane < cn;
        const uint32_t ion = act ? ions[c0 + lane] : 0u;
        const uint32_t mod = HFX_ACID_MOD(ion);
        float m = act ? mtab[HFX_ACID_IDX(ion)] : 0.0f;
        const uint64_t lt = (1ull << lane) - 1ull;
        uint32_t off = 0;
        /* unrolled to the largest acid count, so that occ stays in registers */
#pragma unroll
        for (uint32_t a = 0; a < HFX_MAX_MA; ++a) {
            if (a >= num_ma) break;
            const bool is_a = act && mod == a + 1;
            const uint64_t bal = __ballot(is_a);
            const uint32_t k = mod_row[a];
            if (is_a) {
                const uint32_t ith = occ[a] + (uint32_t)__builtin_popcountll(bal & lt);
                bool hit = false;
                for (uint32_t q = 0; q < k; ++q) hit |= unrank[off + q] == ith;
                if (hit) m = m + ma_mass[a];
            }
            occ[a] += (uint32_t)__builtin_popcountll(bal);
            off += k;
        }
        return m;
    }
};
