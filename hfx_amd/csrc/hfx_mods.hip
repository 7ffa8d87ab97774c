/*
 * hfx_mods.hip -- modified-peptide enumeration and modified theoretical rows.
 *
 *   findModablePeptides  filter_by_mod.cu:150-378
 *   calcTotalModCands    calc_mod_cands.cu:27-119
 *   prepareGenMod        prepare_gen_mod.cu:23-129
 *   genModCands          gen_mod_pep.cu:127-312, combinatorics.h:13-134
 *   addModIons           mion_series.cu:174-556 (dense rows, drop-in)
 *   hfx_mpep_offsets     Search.hs:381-386 on the device
 *
 * Integer work throughout; outputs are bit-identical to the reference's
 * (including unrankComb's reflection trick, which repeats some subsets and
 * misses others for e.g. n=5, k=2 -- reproduced, not fixed).  The one
 * deliberate difference: a modified residue's occurrence index counts its own
 * acid (check_series.cu:61-99, filter_by_mod.cu:216-230) where
 * mion_series.cu:252-269 counts every modable acid.
 */
#include "hfx_cand.h"

/* ------------------------------------------------------------------ */
/* Combinatorics                                                         */
/* ------------------------------------------------------------------ */

struct ChooseTab {
    uint32_t t[43][12];
};
static constexpr ChooseTab make_choose()
{
    ChooseTab c{};
    for (int a = 0; a <= 42; ++a)
        for (int b = 0; b <= 11; ++b)
            c.t[a][b] = (b == 0) ? 1u : (a == 0 ? 0u : c.t[a - 1][b - 1] + c.t[a - 1][b]);
    return c;
}
/* C(n, k) for n <= 42, k <= 11: the table of combinatorics.h:13-57 */
__constant__ ChooseTab k_choose = make_choose();

/* combinatorics.h:59-88 */
__device__ __forceinline__ uint32_t d_choose(uint32_t n, uint32_t k)
{
    if (n > 42 || k > 11) {
        if (k == 0 || n == k) return 1;
        if (n < k) return 0;
        uint32_t t = 1;
        if (k < n - k) {
            for (uint32_t i = n; i >= n - k + 1; --i) t = t * i / (n - i + 1);
        } else {
            for (uint32_t i = n; i >= k + 1; --i) t = t * i / (n - i + 1);
        }
        return t;
    }
    return k_choose.t[n][k];
}

/* combinatorics.h:90-134 largestV + unrankComb */
__device__ __forceinline__ void d_unrank_comb(uint32_t *ans, uint32_t n, uint32_t k, uint32_t rank)
{
    uint32_t a = n, b = k;
    const uint32_t total = d_choose(n, k);
    uint32_t x;
    bool dual = false;
    if ((total / 2) >= rank) x = rank;
    else { x = total - 1 - rank; dual = true; }
    for (uint32_t i = 0; i < k; ++i) {
        uint32_t v = a - 1;
        while (d_choose(v, b) > x) --v;
        ans[i] = v;
        x = x - d_choose(v, b);
        a = v;
        b = b - 1;
    }
    if (dual)
        for (uint32_t i = 0; i < k; ++i) ans[i] = (n - 1) - ans[i];
}

/* ------------------------------------------------------------------ */
/* findModablePeptides                                                   */
/* ------------------------------------------------------------------ */

/* counts of each modable acid over [tc, tn) (fillPepMACount :193-231) of
 * the peptide at every residual-sorted position below npos, once: a
 * peptide's rows repeat once per modification */
__global__ void k_pos_ma_count(uint32_t *pcnt, const uint32_t *idx_sorted, uint32_t npos,
                               const uint8_t *ions, const uint32_t *tc, const uint32_t *tn,
                               const uint8_t *ma, uint32_t num_ma)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npos) return;
    const uint32_t pep = idx_sorted[k];
    uint32_t c[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint8_t mac[10];
    for (uint32_t a = 0; a < num_ma; ++a) mac[a] = ma[a];
    for (uint32_t j = tc[pep]; j < tn[pep]; ++j) {
        const uint8_t ch = (uint8_t)(HFX_ACID_IDX(ions[j]) + 'A');
        for (uint32_t a = 0; a < num_ma; ++a) c[a] += (ch == mac[a]) ? 1u : 0u;
    }
    for (uint32_t a = 0; a < num_ma; ++a) pcnt[(size_t)k * num_ma + a] = c[a];
}

/* rows (pep, mod): spectrum-major, mod-minor, residual-sorted within a range
 * (fillPepAndModIdx, filter_by_mod.cu:151-190), with each row's acid counts
 * and the modability test (checkModable :234-262); one wave per
 * (spectrum, mod) */
__global__ void __launch_bounds__(256)
    k_mod_fill(uint32_t *rows_pep, uint32_t *rows_mod, uint32_t *cnt, uint32_t *flag,
               const uint32_t *spec_num_cand_scan, const uint32_t *idx_sorted,
               const uint32_t *begin, const uint32_t *end, const uint32_t *num_pep_scan,
               uint32_t num_mod, uint64_t pairs, const uint32_t *pcnt,
               const uint32_t *mod_ma_count, uint32_t num_ma)
{
    const uint64_t pr = (uint64_t)blockIdx.x * 4 + threadIdx.x / HFX_WAVE;
    if (pr >= pairs) return;
    const uint32_t m = (uint32_t)(pr % num_mod), s = (uint32_t)(pr / num_mod);
    const uint32_t o = num_pep_scan[pr] + spec_num_cand_scan[s];
    const uint32_t b = begin[pr], e = end[pr];
    const uint32_t *need = mod_ma_count + (size_t)m * num_ma;
    for (uint32_t i = b + hfx_lane(); i < e; i += HFX_WAVE) {
        const uint32_t row = o + (i - b);
        rows_pep[row] = idx_sorted[i];
        rows_mod[row] = m;
        bool ok = true;
        for (uint32_t a = 0; a < num_ma; ++a) {
            const uint32_t v = pcnt[(size_t)i * num_ma + a];
            cnt[(size_t)row * num_ma + a] = v;
            ok = ok && v >= need[a];
        }
        flag[row] = ok ? 1u : 0u;
    }
}

/* largest window end (every window position lies below it) */
__global__ void __launch_bounds__(1024) k_max_end(const uint32_t *end, uint64_t n, uint32_t *out)
{
    __shared__ uint32_t w[16];
    uint32_t v = 0;
    for (uint64_t i = threadIdx.x; i < n; i += 1024) v = max(v, end[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < 16; ++i) t = max(t, w[i]);
        *out = t;
    }
}

/* stable compaction (remove_if, :345-358) */
__global__ void k_mod_compact(uint32_t *out_pep, uint32_t *out_mod, uint32_t *out_cnt,
                              const uint32_t *rows_pep, const uint32_t *rows_mod,
                              const uint32_t *cnt, const uint32_t *flag, const uint64_t *pos,
                              uint32_t n, uint32_t num_ma)
{
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n || !flag[row]) return;
    const uint64_t p = pos[row];
    out_pep[p] = rows_pep[row];
    out_mod[p] = rows_mod[row];
    for (uint32_t a = 0; a < num_ma; ++a) out_cnt[p * num_ma + a] = cnt[(size_t)row * num_ma + a];
}

/* lower_bound(kept row ids, spec_num_cand_scan[s]) == number of kept rows
 * before that row (:364-366) */
__global__ void k_mod_valid_scan(uint32_t *out, const uint32_t *spec_num_cand_scan,
                                 const uint64_t *pos, uint32_t num_spec)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < num_spec) out[s] = (uint32_t)pos[spec_num_cand_scan[s]];
}

static int read_u64(const uint64_t *d, uint64_t *out)
{
    HFX_CHECK_HIP(hipMemcpy(out, d, 8, hipMemcpyDeviceToHost));
    return HFX_OK;
}

extern "C" uint32_t findModablePeptides(
    uint32_t *d_out_pep_idx_raw, uint32_t *d_out_pep_mod_idx_raw, uint32_t *d_out_pep_ma_count_raw,
    uint32_t *d_out_spec_num_valid_scan, const uint32_t *d_spec_num_cand_scan,
    const uint32_t num_spec, uint32_t num_cand_total, const uint8_t *d_ions, const uint32_t *d_tc,
    const uint32_t *d_tn, const uint32_t *d_pep_idx_r_sorted_raw, const uint32_t *d_begin_raw,
    const uint32_t *d_end_raw, const uint32_t *d_num_pep_scan_raw, const uint32_t *d_mod_ma_count,
    const uint32_t num_mod, const uint8_t *d_ma, const uint32_t num_ma)
{
    hfx_ctx *ctx = hfx_default_ctx();
    if (!ctx) return 0;
    if (num_ma > 10) {
        hfx_set_error("findModablePeptides: num_ma %u > 10", num_ma);
        return 0;
    }
    const uint32_t n = num_cand_total;
    if (num_spec == 0 || num_mod == 0) return 0;
    const uint64_t pairs = (uint64_t)num_spec * num_mod;
    /* the positions the windows reach (one small read back) */
    uint32_t npos = 0;
    if (n) {
        void *q, *h;
        if (hfx_ctx_scratch(ctx, HFX_SLOT_MODS2, 16, &q) != HFX_OK) return 0;
        if (hfx_ctx_pinned(ctx, 16, &h) != HFX_OK) return 0;
        hipLaunchKernelGGL(k_max_end, dim3(1), dim3(1024), 0, 0, d_end_raw, pairs, (uint32_t *)q);
        if (hipMemcpy(h, q, 4, hipMemcpyDeviceToHost) != hipSuccess) {
            hfx_set_error("findModablePeptides: read back failed");
            return 0;
        }
        npos = *(uint32_t *)h;
    }
    void *p;
    size_t bytes = (size_t)n * (12 + 4 * (size_t)num_ma) + ((size_t)n + 1) * 8 +
                   (size_t)npos * 4 * num_ma + 64;
    if (hfx_ctx_scratch(ctx, HFX_SLOT_MODS, bytes, &p) != HFX_OK) return 0;
    uint64_t *pos = (uint64_t *)p;
    uint32_t *rows_pep = (uint32_t *)(pos + n + 1);
    uint32_t *rows_mod = rows_pep + n;
    uint32_t *flag = rows_mod + n;
    uint32_t *cnt = flag + n;
    uint32_t *pcnt = cnt + (size_t)n * num_ma;
    if (n) {
        hipLaunchKernelGGL(k_pos_ma_count, dim3((npos + 255) / 256), dim3(256), 0, 0, pcnt,
                           d_pep_idx_r_sorted_raw, npos, d_ions, d_tc, d_tn, d_ma, num_ma);
        hipLaunchKernelGGL(k_mod_fill, dim3((uint32_t)((pairs + 3) / 4)), dim3(256), 0, 0, rows_pep,
                           rows_mod, cnt, flag, d_spec_num_cand_scan, d_pep_idx_r_sorted_raw,
                           d_begin_raw, d_end_raw, d_num_pep_scan_raw, num_mod, pairs, pcnt,
                           d_mod_ma_count, num_ma);
    }
    if (hfx_scan_offsets(ctx, pos, flag, n, nullptr) != HFX_OK) return 0;
    if (n)
        hipLaunchKernelGGL(k_mod_compact, dim3((n + 255) / 256), dim3(256), 0, 0,
                           d_out_pep_idx_raw, d_out_pep_mod_idx_raw, d_out_pep_ma_count_raw,
                           rows_pep, rows_mod, cnt, flag, pos, n, num_ma);
    hipLaunchKernelGGL(k_mod_valid_scan, dim3((num_spec + 255) / 256), dim3(256), 0, 0,
                       d_out_spec_num_valid_scan, d_spec_num_cand_scan, pos, num_spec);
    uint64_t total = 0;
    if (read_u64(pos + n, &total) != HFX_OK) return 0;
    return (uint32_t)total;
}

/* ------------------------------------------------------------------ */
/* calcTotalModCands                                                      */
/* ------------------------------------------------------------------ */

__global__ void k_calc_mod_cands(uint32_t *num_mpep, uint32_t *comb_scan, const uint32_t *mod_ma_count,
                                 const uint32_t *pep_mod_idx, const uint32_t *pep_ma_count,
                                 uint32_t n, uint32_t num_ma)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t m = pep_mod_idx[p];
    uint32_t num = 1;
    for (int32_t i = (int32_t)num_ma - 1; i >= 0; --i) {
        num *= d_choose(pep_ma_count[(size_t)num_ma * p + i], mod_ma_count[(size_t)num_ma * m + i]);
        comb_scan[(size_t)num_ma * p + i] = num;
    }
    num_mpep[p] = num;
}

extern "C" uint32_t calcTotalModCands(uint32_t *d_out_pep_num_mpep_raw,
                                      uint32_t *d_out_pep_ma_num_comb_scan_raw,
                                      const uint32_t *d_mod_ma_count_raw, const uint32_t *d_pep_idx_raw,
                                      const uint32_t *d_pep_mod_idx_raw,
                                      const uint32_t *d_pep_ma_count_raw, const uint32_t nPep,
                                      const uint32_t num_mod, const uint32_t num_ma)
{
    (void)d_pep_idx_raw;
    (void)num_mod;
    hfx_ctx *ctx = hfx_default_ctx();
    if (!ctx || nPep == 0) return 0;
    hipLaunchKernelGGL(k_calc_mod_cands, dim3((nPep + 255) / 256), dim3(256), 0, 0,
                       d_out_pep_num_mpep_raw, d_out_pep_ma_num_comb_scan_raw, d_mod_ma_count_raw,
                       d_pep_mod_idx_raw, d_pep_ma_count_raw, nPep, num_ma);
    return hfx_sum_u32_to_host(ctx, d_out_pep_num_mpep_raw, nPep);
}

/* ------------------------------------------------------------------ */
/* prepareGenMod                                                          */
/* ------------------------------------------------------------------ */

__device__ __forceinline__ uint32_t seg_of(const uint64_t *off, uint32_t n, uint64_t t)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void k_gen_expand(uint32_t *o_pep, uint32_t *o_mod, uint32_t *o_rank, uint32_t *o_ith,
                             uint32_t *o_sum, const uint32_t *mod_ma_count_sum,
                             const uint32_t *pep_idx, const uint32_t *pep_mod_idx,
                             const uint64_t *row_off, uint32_t num_rows, uint32_t num_mpep)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_mpep) return;
    const uint32_t row = seg_of(row_off, num_rows, i);
    const uint32_t m = pep_mod_idx[row];
    o_pep[i] = pep_idx[row];
    o_mod[i] = m;
    o_rank[i] = (uint32_t)(i - row_off[row]);
    o_ith[i] = row;
    o_sum[i] = mod_ma_count_sum[m];
}

extern "C" uint32_t prepareGenMod(uint32_t *d_out_mpep_pep_idx, uint32_t *d_out_mpep_pep_mod_idx,
                                  uint32_t *d_out_mpep_rank, uint32_t *d_out_mpep_ith_valid,
                                  uint32_t *d_out_mpep_mod_ma_count_sum_scan,
                                  const uint32_t *d_mod_ma_count_sum, const uint32_t *d_pep_idx,
                                  const uint32_t *d_pep_mod_idx, const uint32_t *d_pep_num_mpep,
                                  const uint32_t num_pep, const uint32_t num_mpep)
{
    hfx_ctx *ctx = hfx_default_ctx();
    if (!ctx || num_pep == 0 || num_mpep == 0) return 0;
    void *p;
    if (hfx_ctx_scratch(ctx, HFX_SLOT_MODS2, ((size_t)num_pep + 1) * 8, &p) != HFX_OK) return 0;
    uint64_t *row_off = (uint64_t *)p;
    if (hfx_scan_offsets(ctx, row_off, d_pep_num_mpep, num_pep, nullptr) != HFX_OK) return 0;
    uint64_t avail = 0;
    if (read_u64(row_off + num_pep, &avail) != HFX_OK) return 0;
    const uint32_t m = (uint32_t)min((uint64_t)num_mpep, avail);
    hipLaunchKernelGGL(k_gen_expand, dim3((m + 255) / 256), dim3(256), 0, 0, d_out_mpep_pep_idx,
                       d_out_mpep_pep_mod_idx, d_out_mpep_rank, d_out_mpep_ith_valid,
                       d_out_mpep_mod_ma_count_sum_scan, d_mod_ma_count_sum, d_pep_idx,
                       d_pep_mod_idx, row_off, num_pep, m);
    const uint32_t total = hfx_sum_u32_to_host(ctx, d_out_mpep_mod_ma_count_sum_scan, m);
    hfx_scan_u32_inplace(d_out_mpep_mod_ma_count_sum_scan, m, nullptr);
    if (hipStreamSynchronize(0) != hipSuccess) hfx_set_error("prepareGenMod: sync failed");
    return total;
}

/* ------------------------------------------------------------------ */
/* genModCands                                                            */
/* ------------------------------------------------------------------ */

__global__ void k_gen_mod_cands(uint32_t *out_unrank, const uint32_t *mod_ma_count,
                                const uint32_t *mpep_ith, const uint32_t *mpep_rank,
                                const uint32_t *mpep_scan, const uint32_t *pep_mod_idx,
                                const uint32_t *pep_ma_count, const uint32_t *comb_scan,
                                uint32_t num_mpep, uint32_t num_ma)
{
    const uint32_t mp = blockIdx.x * blockDim.x + threadIdx.x;
    if (mp >= num_mpep) return;
    const uint32_t ith = mpep_ith[mp];
    const uint32_t pm = pep_mod_idx[ith];
    const uint32_t *row = comb_scan + (size_t)ith * num_ma;
    /* add_ma_ranks, gen_mod_pep.cu:127-157 */
    uint32_t ranks[10];
    uint32_t x = mpep_rank[mp];
    for (uint32_t i = 0; i < num_ma; ++i) {
        if (i != num_ma - 1) {
            const uint32_t nxt = row[i + 1];
            if (x >= nxt) {
                ranks[i] = x / nxt;
                x = x % nxt;
            } else {
                ranks[i] = 0;
            }
        } else {
            ranks[i] = x;
        }
    }
    uint32_t pos = mpep_scan[mp];
    for (uint32_t i = 0; i < num_ma; ++i) {
        const uint32_t k = mod_ma_count[(size_t)pm * num_ma + i];
        if (k != 0) {
            d_unrank_comb(out_unrank + pos, pep_ma_count[(size_t)ith * num_ma + i], k, ranks[i]);
            pos += k;
        }
    }
}

extern "C" void genModCands(uint32_t *d_out_mpep_unrank, const uint32_t *d_mod_ma_count,
                            const uint32_t *d_mpep_ith_valid, const uint32_t *d_mpep_rank,
                            const uint32_t *d_mpep_mod_ma_count_sum_scan,
                            const uint32_t *d_pep_mod_idx, const uint32_t *d_pep_ma_count,
                            const uint32_t *d_pep_ma_num_comb_scan, const uint32_t total,
                            const uint32_t mod_num_ma)
{
    if (total == 0) return;
    if (mod_num_ma < 1 || mod_num_ma > 10) {
        hfx_set_error("genModCands: num_ma %u outside 1..10", mod_num_ma);
        return;
    }
    hipLaunchKernelGGL(k_gen_mod_cands, dim3((total + 255) / 256), dim3(256), 0, 0,
                       d_out_mpep_unrank, d_mod_ma_count, d_mpep_ith_valid, d_mpep_rank,
                       d_mpep_mod_ma_count_sum_scan, d_pep_mod_idx, d_pep_ma_count,
                       d_pep_ma_num_comb_scan, total, mod_num_ma);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) hfx_set_error("genModCands: %s", hipGetErrorString(e));
}

/* ------------------------------------------------------------------ */
/* addModIons: dense rows of modified peptides                           */
/* ------------------------------------------------------------------ */

#define AM_THREADS 256
#define AM_WAVES (AM_THREADS / HFX_WAVE)

__global__ void __launch_bounds__(AM_THREADS)
    k_add_mod_ions(float *spec, const float *residual, const float *mass, const uint8_t *ions_all,
                   const uint32_t *tc_a, const uint32_t *tn_a, const uint32_t *mp_pep,
                   const uint32_t *mp_mod, const uint32_t *unrank, const uint32_t *mp_scan,
                   uint32_t num_mpep, const uint32_t *mod_ma_count, const float *mod_delta,
                   const float *ma_mass, uint32_t num_ma, uint32_t len, uint32_t zmax)
{
    __shared__ uint32_t maps[AM_WAVES][HFX_MAP_WORDS];
    __shared__ float mtab[32];
    const uint32_t wave = threadIdx.x / HFX_WAVE;
    uint32_t *map = maps[wave];
    if (threadIdx.x < 32) mtab[threadIdx.x] = threadIdx.x < 26 ? mass[threadIdx.x] : 0.0f;
    for (uint32_t w = hfx_lane(); w < HFX_MAP_WORDS; w += HFX_WAVE) map[w] = 0;
    __syncthreads();
    for (uint32_t row = blockIdx.x * AM_WAVES + wave; row < num_mpep; row += gridDim.x * AM_WAVES) {
        const uint32_t pep = mp_pep[row], mod = mp_mod[row];
        const uint32_t tc = tc_a[pep], tn = tn_a[pep];
        const uint32_t n = tn > tc ? tn - tc : 0;
        HfxModMass mm{ions_all + tc, mtab, mod_ma_count + (size_t)mod * num_ma,
                      unrank + mp_scan[row], ma_mass, num_ma, {0, 0, 0, 0, 0, 0, 0}};
        hfx_wave_peaks(map, n, residual[pep] + mod_delta[mod], zmax, len, mm);
        hfx_wave_sync();
        const uint32_t nw = (len + 7) >> 3;
        float *out = spec + (size_t)row * len;
        for (uint32_t w = hfx_lane(); w < nw; w += HFX_WAVE) {
            const uint32_t word = map[w];
            map[w] = 0;
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) {
                const uint32_t b = (w << 3) + k;
                if (b < len) out[b] = hfx_code_weight((word >> (4 * k)) & 0xFu);
            }
        }
        hfx_wave_sync();
    }
}

extern "C" void addModIons(float *d_out_mspec, const float *d_residual, const float *d_mass,
                           const uint8_t *d_ions, const uint32_t *d_tc, const uint32_t *d_tn,
                           const uint32_t *d_mpep_pep_idx, const uint32_t *d_mpep_pep_mod_idx,
                           const uint32_t *d_mpep_unrank,
                           const uint32_t *d_mpep_mod_ma_count_sum_scan, const uint32_t len_unrank,
                           const uint32_t num_mpep, const uint32_t *d_mod_ma_count,
                           const float *d_mod_delta, const uint8_t *d_ma, const float *d_ma_mass,
                           const uint32_t num_ma, const uint32_t len_spec,
                           const uint32_t max_charge, void *stream)
{
    (void)len_unrank;
    (void)d_ma;
    if (num_mpep == 0) return;
    /* mion_series.cu:443-556 dispatches MaxCharge 1..10 and NumMA 1..10; the
     * 3-bit packing of prepare_ions limits modable acids to 7 */
    if (max_charge < 1 || max_charge > 10 || num_ma < 1 || num_ma > HFX_MAX_MA ||
        len_spec > HFX_MAX_LEN) {
        hfx_set_error("addModIons: unsupported max_charge %u / num_ma %u / len %u", max_charge,
                      num_ma, len_spec);
        return;
    }
    const uint32_t blocks = min((num_mpep + AM_WAVES - 1) / AM_WAVES, 65535u);
    hipLaunchKernelGGL(k_add_mod_ions, dim3(blocks), dim3(AM_THREADS), 0, hfx_stream(stream),
                       d_out_mspec, d_residual, d_mass, d_ions, d_tc, d_tn, d_mpep_pep_idx,
                       d_mpep_pep_mod_idx, d_mpep_unrank, d_mpep_mod_ma_count_sum_scan, num_mpep,
                       d_mod_ma_count, d_mod_delta, d_ma_mass, num_ma, len_spec, max_charge);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) hfx_set_error("addModIons: %s", hipGetErrorString(e));
}

/* ------------------------------------------------------------------ */
/* Per-spectrum offsets of the modified peptides                          */
/* ------------------------------------------------------------------ */

__global__ void k_mpep_offsets(uint64_t *mp_off, const uint64_t *row_off, uint32_t num_rows,
                               const uint32_t *valid_scan, uint32_t num_spec)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < num_spec) mp_off[s] = row_off[min(valid_scan[s], num_rows)];
    else if (s == num_spec) mp_off[s] = row_off[num_rows];
}

extern "C" int hfx_mpep_offsets(hfx_ctx *ctx, uint64_t *d_mp_off, const uint32_t *d_pep_num_mpep,
                                uint32_t num_rows, const uint32_t *d_spec_num_valid_scan,
                                uint32_t num_spec, void *stream)
{
    HFX_REQUIRE(ctx && d_mp_off && d_spec_num_valid_scan, HFX_ERR_ARG,
                "hfx_mpep_offsets: null argument");
    void *p;
    int rc;
    if ((rc = hfx_ctx_scratch(ctx, HFX_SLOT_MODS2, ((size_t)num_rows + 1) * 8, &p)) != HFX_OK)
        return rc;
    uint64_t *row_off = (uint64_t *)p;
    if ((rc = hfx_scan_offsets(ctx, row_off, d_pep_num_mpep, num_rows, stream)) != HFX_OK) return rc;
    hipLaunchKernelGGL(k_mpep_offsets, dim3((num_spec + 256) / 256), dim3(256), 0,
                       hfx_stream(stream), d_mp_off, row_off, num_rows, d_spec_num_valid_scan,
                       num_spec);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}
