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

#include <mutex>

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
    hipStream_t st = hfx_stream(stream);
    int rc;
    uint32_t *tiles = (uint32_t *)scratch;
    uint64_t *tile_off = (uint64_t *)(((uintptr_t)(tiles + S) + 15) & ~(uintptr_t)15);
    hipLaunchKernelGGL(k_tile_count<Cands>, dim3((S + 255) / 256), dim3(256), 0, st, tiles, c, S,
                       (uint32_t)XC_CHUNK);
    if ((rc = hfx_scan_offsets(nullptr, tile_off, tiles, S, stream)) != HFX_OK) return rc;
    /* sum_s ceil(n_s / chunk) <= total / chunk + S */
    const uint64_t grid = (total + XC_CHUNK - 1) / XC_CHUNK + S;
    HFX_REQUIRE(grid < (1ull << 31), HFX_ERR_ARG, "xcorr: %llu tiles", (unsigned long long)grid);
    hipLaunchKernelGGL(k_xcorr<Cands>, dim3((uint32_t)grid), dim3(XC_THREADS), 0, st, d_scores,
                       tile_off, c, *sp);
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
