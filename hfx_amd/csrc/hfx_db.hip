/*
 * hfx_db.hip -- database preparation and precursor windows.
 *
 *   prepare_ions       residue byte packing          (prepare_ions.cu:54-132)
 *   sort_idx_f & co    stable LSD radix sort          (sort.cu:20-76, sort_b40c.cu:34-41)
 *   findSpecCandsByMass, hfx_window                   (filter.cu:75-127)
 *   findBeginEnd_f                                    (filter_by_mod.cu:35-143)
 *   prepare_scoring                                   (prepare_scoring.cu:47-67)
 *   hfx_scan_offsets   exclusive scans used by every batched entry point
 */
#include "hfx_common.h"

/* ------------------------------------------------------------------ */
/* prepare_ions                                                          */
/* ------------------------------------------------------------------ */

__global__ void __launch_bounds__(256) k_prepare_ions(uint8_t *ions, uint64_t n, const uint8_t *ma,
                                                      uint32_t num_ma)
{
    __shared__ uint8_t lut[256];
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) {
        uint8_t mod = 0;
        for (uint32_t j = 0; j < num_ma; ++j)
            if (ma[j] == i) { mod = (uint8_t)(j + 1); break; }
        lut[i] = mod;
    }
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint8_t c = ions[i];
        ions[i] = (uint8_t)((uint8_t)(c - 'A') | (uint8_t)(lut[c] << 5));
    }
}

extern "C" void prepare_ions(uint8_t *d_ions, uint32_t N, uint8_t *d_ma, uint32_t num_ma)
{
    if (N == 0) return;
    if (num_ma > 7) { hfx_set_error("prepare_ions: num_ma %u > 7 (3-bit field)", num_ma); return; }
    uint32_t blocks = (uint32_t)min((uint64_t)4096, ((uint64_t)N + 255) / 256);
    const uint8_t *ma = num_ma ? d_ma : nullptr;
    hipLaunchKernelGGL(k_prepare_ions, dim3(blocks), dim3(256), 0, 0, d_ions, (uint64_t)N, ma,
                       num_ma);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) hfx_set_error("prepare_ions: %s", hipGetErrorString(e));
}

/* ------------------------------------------------------------------ */
/* Stable LSD radix sort of (uint32 key, uint32 value), 4-bit digits.    */
/* ------------------------------------------------------------------ */

#define RS_THREADS 256
#define RS_ITEMS 8
#define RS_TILE (RS_THREADS * RS_ITEMS)
#define RS_BUCKETS 16

__global__ void __launch_bounds__(RS_THREADS) k_rs_upsweep(const uint32_t *keys, uint32_t n,
                                                          uint32_t shift, uint32_t *hist,
                                                          uint32_t nb)
{
    __shared__ uint32_t cnt[RS_BUCKETS];
    if (threadIdx.x < RS_BUCKETS) cnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * RS_TILE;
    for (uint32_t i = threadIdx.x; i < RS_TILE; i += RS_THREADS) {
        uint32_t g = base + i;
        if (g < n) atomicAdd(&cnt[(keys[g] >> shift) & 15u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < RS_BUCKETS) hist[threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
}

/* Exclusive scan of n uint32 in place by one workgroup of 1024 threads. */
__global__ void __launch_bounds__(1024) k_scan_u32_inplace(uint32_t *a, uint32_t n)
{
    __shared__ uint32_t warp_tot[16];
    __shared__ uint32_t carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (uint32_t base = 0; base < n; base += 1024) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < n ? a[i] : 0;
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o, 64);
            if (lane >= (uint32_t)o) x += y;
        }
        if (lane == 63) warp_tot[wid] = x;
        __syncthreads();
        uint32_t wpre = 0;
        for (uint32_t w = 0; w < wid; ++w) wpre += warp_tot[w];
        uint32_t carry = carry_s;
        if (i < n) a[i] = carry + wpre + x - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry_s = carry + wpre + x;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(RS_THREADS) k_rs_downsweep(const uint32_t *kin, const uint32_t *vin,
                                                            uint32_t *kout, uint32_t *vout,
                                                            uint32_t n, uint32_t shift,
                                                            const uint32_t *hist, uint32_t nb)
{
    __shared__ uint32_t cnt[RS_BUCKETS * RS_THREADS];
    __shared__ uint32_t tsum[RS_THREADS];
    __shared__ uint32_t dstart[RS_BUCKETS];
    const uint32_t t = threadIdx.x;
    const uint32_t base = blockIdx.x * RS_TILE + t * RS_ITEMS;
    uint32_t k[RS_ITEMS], v[RS_ITEMS];
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        uint32_t g = base + i;
        k[i] = g < n ? kin[g] : 0;
        v[i] = g < n ? vin[g] : 0;
    }
#pragma unroll
    for (int d = 0; d < RS_BUCKETS; ++d) cnt[d * RS_THREADS + t] = 0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i)
        if (base + i < n) cnt[((k[i] >> shift) & 15u) * RS_THREADS + t] += 1;
    __syncthreads();
    /* block exclusive scan over cnt in (digit, thread) order: thread t owns
     * flat entries [16t, 16t+16) */
    uint32_t loc[RS_BUCKETS];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < RS_BUCKETS; ++j) {
        loc[j] = s;
        s += cnt[t * RS_BUCKETS + j];
    }
    tsum[t] = s;
    __syncthreads();
    if (t < 64) {
        /* scan of 256 thread totals by one wave, 4 per lane */
        uint32_t a0 = tsum[4 * t], a1 = tsum[4 * t + 1], a2 = tsum[4 * t + 2], a3 = tsum[4 * t + 3];
        uint32_t tot = a0 + a1 + a2 + a3, x = tot;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o, 64);
            if (t >= (uint32_t)o) x += y;
        }
        uint32_t ex = x - tot;
        tsum[4 * t] = ex;
        tsum[4 * t + 1] = ex + a0;
        tsum[4 * t + 2] = ex + a0 + a1;
        tsum[4 * t + 3] = ex + a0 + a1 + a2;
    }
    __syncthreads();
    uint32_t pre = tsum[t];
#pragma unroll
    for (int j = 0; j < RS_BUCKETS; ++j) cnt[t * RS_BUCKETS + j] = pre + loc[j];
    __syncthreads();
    if (t < RS_BUCKETS) dstart[t] = cnt[t * RS_THREADS];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        if (base + i < n) {
            uint32_t d = (k[i] >> shift) & 15u;
            uint32_t lp = cnt[d * RS_THREADS + t]++;
            uint32_t g = hist[d * nb + blockIdx.x] + (lp - dstart[d]);
            kout[g] = k[i];
            vout[g] = v[i];
        }
    }
}


/* keys in radix order: ascending float order, or (desc) its reverse --
 * the complement of an order-preserving key reverses the order, and a
 * stable sort of the complements is a stable descending sort */
__global__ void k_float_to_key(const float *f, uint32_t *k, uint32_t n, uint32_t desc)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) k[i] = desc ? ~hfx_f2key(f[i]) : hfx_f2key(f[i]);
}
__global__ void k_key_to_float(const uint32_t *k, float *f, uint32_t n, uint32_t desc)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = hfx_key2f(desc ? ~k[i] : k[i]);
}
__global__ void k_iota(uint32_t *v, uint32_t n, uint32_t init)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = init + i;
}

/* Sorts float keys ascending with values, stably, in place. */
static int radix_sort_f32(hfx_ctx *ctx, float *d_keys, uint32_t *d_vals, uint32_t n, bool iota,
                          uint32_t init, hipStream_t st, uint32_t desc = 0)
{
    if (n == 0) return HFX_OK;
    uint32_t nb = (n + RS_TILE - 1) / RS_TILE;
    void *p;
    int rc;
    if ((rc = hfx_ctx_scratch(ctx, HFX_SLOT_SORT, (size_t)n * 16 + (size_t)nb * RS_BUCKETS * 4, &p)) != HFX_OK)
        return rc;
    uint32_t *ka = (uint32_t *)p, *va = ka + n, *kb = va + n, *vb = kb + n, *hist = vb + n;
    uint32_t g = (n + 255) / 256;
    hipLaunchKernelGGL(k_float_to_key, dim3(g), dim3(256), 0, st, d_keys, ka, n, desc);
    if (iota) hipLaunchKernelGGL(k_iota, dim3(g), dim3(256), 0, st, va, n, init);
    else HFX_CHECK_HIP(hipMemcpyAsync(va, d_vals, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    for (uint32_t shift = 0; shift < 32; shift += 4) {
        hipLaunchKernelGGL(k_rs_upsweep, dim3(nb), dim3(RS_THREADS), 0, st, ka, n, shift, hist, nb);
        hipLaunchKernelGGL(k_scan_u32_inplace, dim3(1), dim3(1024), 0, st, hist, nb * RS_BUCKETS);
        hipLaunchKernelGGL(k_rs_downsweep, dim3(nb), dim3(RS_THREADS), 0, st, ka, va, kb, vb, n,
                           shift, hist, nb);
        uint32_t *t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    hipLaunchKernelGGL(k_key_to_float, dim3(g), dim3(256), 0, st, ka, d_keys, n, desc);
    HFX_CHECK_HIP(hipMemcpyAsync(d_vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

extern "C" void sort_idx_f(float *d_keys_raw, uint32_t *d_idx_raw, uint32_t N, uint32_t init)
{
    hfx_ctx *ctx = hfx_default_ctx();
    if (!ctx) return;
    if (radix_sort_f32(ctx, d_keys_raw, d_idx_raw, N, true, init, 0) == HFX_OK)
        (void)hipStreamSynchronize(0);
}

extern "C" void sort_val_f(float *d_keys, uint32_t *d_vals, uint32_t N)
{
    hfx_ctx *ctx = hfx_default_ctx();
    if (!ctx) return;
    if (radix_sort_f32(ctx, d_keys, d_vals, N, false, 0, 0) == HFX_OK) (void)hipStreamSynchronize(0);
}

extern "C" void sort_rf(float *d_keys, uint32_t *d_vals, uint32_t N)
{
    hfx_ctx *ctx = hfx_default_ctx();
    if (!ctx) return;
    if (radix_sort_f32(ctx, d_keys, d_vals, N, false, 0, 0, 1) == HFX_OK)
        (void)hipStreamSynchronize(0);
}

extern "C" void sort_idx_rf(float *d_keys_raw, uint32_t *d_idx_raw, uint32_t N, uint32_t init)
{
    hfx_ctx *ctx = hfx_default_ctx();
    if (!ctx) return;
    if (radix_sort_f32(ctx, d_keys_raw, d_idx_raw, N, true, init, 0, 1) == HFX_OK)
        (void)hipStreamSynchronize(0);
}

/* findIndicesInRange_f (filter.cu:38-73): the indices i, ascending, with
 * min <= d_in[i] <= max; stable compaction (flags, offsets, scatter) */
__global__ void k_range_flags(uint32_t *flag, const float *in, uint32_t n, float lo, float hi)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (lo <= in[i] && in[i] <= hi) ? 1u : 0u;
}
__global__ void k_range_scatter(uint32_t *out, const uint32_t *flag, const uint64_t *pos, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) out[pos[i]] = i;
}

extern "C" uint32_t findIndicesInRange_f(const float *d_in, uint32_t *d_out, const uint32_t N,
                                         const float min, const float max)
{
    if (N == 0) return 0;
    hfx_ctx *ctx = hfx_default_ctx();
    if (!ctx) return 0;
    void *p = nullptr, *h = nullptr;
    const size_t fb = ((size_t)N * 4 + 255) & ~(size_t)255;
    if (hfx_ctx_scratch(ctx, HFX_SLOT_SCAN, fb + ((size_t)N + 1) * 8, &p) != HFX_OK) return 0;
    if (hfx_ctx_pinned(ctx, 16, &h) != HFX_OK) return 0;
    uint32_t *flag = (uint32_t *)p;
    uint64_t *pos = (uint64_t *)((uint8_t *)p + fb);
    const uint32_t g = (N + 255) / 256;
    hipLaunchKernelGGL(k_range_flags, dim3(g), dim3(256), 0, 0, flag, d_in, N, min, max);
    if (hfx_scan_offsets(ctx, pos, flag, N, nullptr) != HFX_OK) return 0;
    hipLaunchKernelGGL(k_range_scatter, dim3(g), dim3(256), 0, 0, d_out, flag, pos, N);
    if (hipMemcpyAsync(h, pos + N, 8, hipMemcpyDeviceToHost, 0) != hipSuccess ||
        hipStreamSynchronize(0) != hipSuccess) {
        hfx_set_error("findIndicesInRange_f: %s", hipGetErrorString(hipGetLastError()));
        return 0;
    }
    return (uint32_t)*(uint64_t *)h;
}

extern "C" void sort_b40c_f(float *d_keys, uint32_t *d_vals, uint32_t N)
{
    hfx_ctx *ctx = hfx_default_ctx();
    if (!ctx) return;
    if (radix_sort_f32(ctx, d_keys, d_vals, N, false, 0, 0) == HFX_OK) (void)hipStreamSynchronize(0);
}

/* ------------------------------------------------------------------ */
/* Precursor windows                                                     */
/* ------------------------------------------------------------------ */

/* lower_bound / upper_bound over r[idx[i]] (thrust's permutation_iterator
 * form, filter.cu:101-115); idx == nullptr reads r directly (sorted copy). */
__device__ __forceinline__ uint32_t lb_perm(const float *r, const uint32_t *idx, uint32_t n, float v)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        float e = idx ? r[idx[mid]] : r[mid];
        if (e < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ uint32_t ub_perm(const float *r, const uint32_t *idx, uint32_t n, float v)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        float e = idx ? r[idx[mid]] : r[mid];
        if (v < e) hi = mid; else lo = mid + 1;
    }
    return lo;
}

__global__ void k_window(uint32_t *begin, uint32_t *end, uint32_t *num, const float *r,
                         const uint32_t *idx, uint32_t num_pep, const float *mass, uint32_t num_spec,
                         float eps)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= num_spec) return;
    float lo = mass[s] - eps, hi = mass[s] + eps;
    uint32_t b = lb_perm(r, idx, num_pep, lo);
    uint32_t e = ub_perm(r, idx, num_pep, hi);
    if (begin) begin[s] = b;
    if (end) end[s] = e;
    num[s] = e - b;
}

/* sum of n uint32 (wrapping, like thrust::reduce on uint32) into *out */
__global__ void __launch_bounds__(1024) k_sum_u32(const uint32_t *a, uint32_t n, uint32_t *out)
{
    __shared__ uint32_t part[16];
    uint32_t s = 0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) s += a[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < 16; ++w) t += part[w];
        *out = t;
    }
}

/* partial sums of n uint32: block b adds elements b, b + G*1024, ... (G
 * blocks of 1024 threads) into part[b] */
__global__ void __launch_bounds__(1024) k_sum_part(const uint32_t *a, uint32_t n, uint32_t *part)
{
    __shared__ uint32_t w[16];
    uint32_t s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 1024 + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * 1024)
        s += a[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < 16; ++i) t += w[i];
        part[blockIdx.x] = t;
    }
}

#define SUM_BLOCKS 512u

uint32_t hfx_sum_u32_to_host(hfx_ctx *ctx, const uint32_t *d_a, uint32_t n)
{
    void *p, *h;
    if (hfx_ctx_scratch(ctx, HFX_SLOT_SUM, 16 + 4 * SUM_BLOCKS, &p) != HFX_OK) return 0;
    if (hfx_ctx_pinned(ctx, 16, &h) != HFX_OK) return 0;
    if (n > 65536u) {
        /* many workgroups, then one over their sums (integer: any order) */
        const uint32_t g = min(SUM_BLOCKS, (n + 8191u) / 8192u);
        uint32_t *part = (uint32_t *)p + 4;
        hipLaunchKernelGGL(k_sum_part, dim3(g), dim3(1024), 0, 0, d_a, n, part);
        hipLaunchKernelGGL(k_sum_u32, dim3(1), dim3(1024), 0, 0, part, g, (uint32_t *)p);
    } else {
        hipLaunchKernelGGL(k_sum_u32, dim3(1), dim3(1024), 0, 0, d_a, n, (uint32_t *)p);
    }
    if (hipMemcpyAsync(h, p, 4, hipMemcpyDeviceToHost, 0) != hipSuccess) return 0;
    if (hipStreamSynchronize(0) != hipSuccess) {
        hfx_set_error("sum_to_host: synchronize failed");
        return 0;
    }
    return *(uint32_t *)h;
}

extern "C" uint32_t findSpecCandsByMass(uint32_t *d_spec_begin_raw, uint32_t *d_spec_end_raw,
                                        uint32_t *d_spec_num_pep_raw, const float *d_r_raw,
                                        const uint32_t *d_pep_idx_r_sorted_raw,
                                        const uint32_t num_pep, const float *d_spec_masses_raw,
                                        const uint32_t num_spec, const float eps)
{
    if (num_spec == 0) return 0;
    hfx_ctx *ctx = hfx_default_ctx();
    if (!ctx) return 0;
    hipLaunchKernelGGL(k_window, dim3((num_spec + 255) / 256), dim3(256), 0, 0, d_spec_begin_raw,
                       d_spec_end_raw, d_spec_num_pep_raw, d_r_raw, d_pep_idx_r_sorted_raw, num_pep,
                       d_spec_masses_raw, num_spec, eps);
    return hfx_sum_u32_to_host(ctx, d_spec_num_pep_raw, num_spec);
}

extern "C" int hfx_window(hfx_ctx *ctx, uint32_t *d_begin, uint32_t *d_num, const hfx_db *db,
                          const float *d_spec_mass, uint32_t num_spec, float eps, void *stream)
{
    HFX_REQUIRE(ctx && db && d_begin && d_num, HFX_ERR_ARG, "hfx_window: null argument");
    if (num_spec == 0) return HFX_OK;
    const float *r = db->d_r_sorted ? db->d_r_sorted : db->d_r;
    const uint32_t *idx = db->d_r_sorted ? nullptr : db->d_idx_sorted;
    hipLaunchKernelGGL(k_window, dim3((num_spec + 255) / 256), dim3(256), 0, hfx_stream(stream),
                       d_begin, (uint32_t *)nullptr, d_num, r, idx, db->num_pep, d_spec_mass,
                       num_spec, eps);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

__global__ void k_begin_end_mod(uint32_t *begin, uint32_t *end, uint32_t *cnt, const float *r,
                                const uint32_t *idx, uint32_t num_pep, const float *delta,
                                uint32_t num_mod, float lo0, float hi0)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_mod) return;
    float lo = lo0 - delta[i], hi = hi0 - delta[i];
    uint32_t b = lb_perm(r, idx, num_pep, lo), e = ub_perm(r, idx, num_pep, hi);
    begin[i] = b;
    end[i] = e;
    cnt[i] = e - b;
}

extern "C" uint32_t findBeginEnd_f(uint32_t *d_begin_raw, uint32_t *d_end_raw,
                                   uint32_t *d_num_pep_scan_raw, const float *d_r_raw,
                                   const uint32_t *d_pep_idx_r_sorted_raw, const uint32_t num_pep,
                                   const float *d_mod_delta_raw, const uint32_t num_mod,
                                   const float mass_, const float eps)
{
    if (num_mod == 0) return 0;
    hfx_ctx *ctx = hfx_default_ctx();
    if (!ctx) return 0;
    float lo0 = mass_ - eps, hi0 = mass_ + eps;
    hipLaunchKernelGGL(k_begin_end_mod, dim3((num_mod + 255) / 256), dim3(256), 0, 0, d_begin_raw,
                       d_end_raw, d_num_pep_scan_raw, d_r_raw, d_pep_idx_r_sorted_raw, num_pep,
                       d_mod_delta_raw, num_mod, lo0, hi0);
    uint32_t total = hfx_sum_u32_to_host(ctx, d_num_pep_scan_raw, num_mod);
    hipLaunchKernelGGL(k_scan_u32_inplace, dim3(1), dim3(1024), 0, 0, d_num_pep_scan_raw, num_mod);
    if (hipStreamSynchronize(0) != hipSuccess) hfx_set_error("findBeginEnd_f: sync failed");
    return total;
}

/* findBeginEnd_f for many spectra in one launch: one thread per spectrum
 * walks the modifications in order, so its exclusive scan of the window sizes
 * is the reference's per-spectrum scan. */
__global__ void k_mod_windows(uint32_t *begin, uint32_t *end, uint32_t *scan, uint32_t *spec_cand,
                              const float *r, const uint32_t *idx, uint32_t num_pep,
                              const float *delta, uint32_t num_mod, const float *mass,
                              uint32_t num_spec, float eps)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= num_spec) return;
    const float lo0 = mass[s] - eps, hi0 = mass[s] + eps;
    uint32_t acc = 0;
    for (uint32_t i = 0; i < num_mod; ++i) {
        const float lo = lo0 - delta[i], hi = hi0 - delta[i];
        const uint32_t b = lb_perm(r, idx, num_pep, lo), e = ub_perm(r, idx, num_pep, hi);
        const size_t o = (size_t)s * num_mod + i;
        begin[o] = b;
        end[o] = e;
        scan[o] = acc;
        acc += e - b;
    }
    spec_cand[s] = acc;
}

extern "C" int hfx_mod_windows(hfx_ctx *ctx, uint32_t *d_begin, uint32_t *d_end,
                               uint32_t *d_num_scan, uint32_t *d_spec_cand, const hfx_db *db,
                               const float *d_spec_mass, uint32_t num_spec,
                               const float *d_mod_delta, uint32_t num_mod, float eps,
                               void *stream)
{
    HFX_REQUIRE(ctx && db && d_begin && d_end && d_num_scan && d_spec_cand && d_spec_mass,
                HFX_ERR_ARG, "hfx_mod_windows: null argument");
    HFX_REQUIRE(num_mod == 0 || d_mod_delta, HFX_ERR_ARG, "hfx_mod_windows: null deltas");
    if (num_spec == 0) return HFX_OK;
    const float *r = db->d_r_sorted ? db->d_r_sorted : db->d_r;
    const uint32_t *idx = db->d_r_sorted ? nullptr : db->d_idx_sorted;
    hipLaunchKernelGGL(k_mod_windows, dim3((num_spec + 255) / 256), dim3(256), 0,
                       hfx_stream(stream), d_begin, d_end, d_num_scan, d_spec_cand, r, idx,
                       db->num_pep, d_mod_delta, num_mod, d_spec_mass, num_spec, eps);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

/* ------------------------------------------------------------------ */
/* Offsets and candidate lists                                           */
/* ------------------------------------------------------------------ */

/* Exclusive scan of n uint32 counts into n+1 uint64 offsets, one workgroup. */
__global__ void __launch_bounds__(1024) k_scan_offsets(uint64_t *off, const uint32_t *cnt, uint32_t n)
{
    __shared__ uint64_t warp_tot[16];
    __shared__ uint64_t carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (uint32_t base = 0; base < n; base += 1024) {
        uint32_t i = base + threadIdx.x;
        uint64_t v = i < n ? cnt[i] : 0;
        uint64_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint64_t y = __shfl_up(x, o, 64);
            if (lane >= (uint32_t)o) x += y;
        }
        if (lane == 63) warp_tot[wid] = x;
        __syncthreads();
        uint64_t wpre = 0;
        for (uint32_t w = 0; w < wid; ++w) wpre += warp_tot[w];
        uint64_t carry = carry_s;
        if (i < n) off[i] = carry + wpre + x - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry_s = carry + wpre + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) off[n] = carry_s;
}

/* Multi-workgroup form (reduce, then scan), with no scratch: the offsets
 * array itself carries the tile totals between the passes.
 *   k_scan_tile_sums   tile b (SC_TILE counts) writes its sum to
 *                      off[min((b + 1) SC_TILE, n)];
 *   k_scan_tile_bases  one workgroup turns those entries, and off[0] = 0,
 *                      into the exclusive prefix of each tile (inclusive
 *                      prefix of the sums: the entry at (b + 1) SC_TILE is
 *                      tile b + 1's base, the one at n the total);
 *   k_scan_tiles       tile b reads its base from off[b SC_TILE] and writes
 *                      off[b SC_TILE + i] for its counts (the first one
 *                      being the base again); no other tile touches them. */
#define SC_THREADS 1024
#define SC_ITEMS 8
#define SC_TILE (SC_THREADS * SC_ITEMS)
/* below this count one workgroup's pass is quicker than three launches */
#define SC_SMALL 8192u

__device__ __forceinline__ uint64_t sc_block_sum(uint64_t v, uint64_t *part)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t t = 0;
#pragma unroll
    for (int w = 0; w < SC_THREADS / 64; ++w) t += part[w];
    return t;
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_tile_sums(uint64_t *off, const uint32_t *cnt,
                                                               uint32_t n)
{
    __shared__ uint64_t part[SC_THREADS / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SC_TILE;
    uint64_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < SC_ITEMS; ++k) {
        const uint64_t i = base + k * SC_THREADS + threadIdx.x;
        if (i < n) s += cnt[i];
    }
    const uint64_t t = sc_block_sum(s, part);
    if (threadIdx.x == 0) off[min(base + SC_TILE, (uint64_t)n)] = t;
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_tile_bases(uint64_t *off, uint32_t n,
                                                                uint32_t ntiles)
{
    __shared__ uint64_t warp_tot[SC_THREADS / 64];
    __shared__ uint64_t carry_s;
    if (threadIdx.x == 0) {
        carry_s = 0;
        off[0] = 0;
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (uint32_t b0 = 0; b0 < ntiles; b0 += SC_THREADS) {
        const uint32_t b = b0 + threadIdx.x;
        const uint64_t at = min((uint64_t)(b + 1) * SC_TILE, (uint64_t)n);
        const uint64_t v = b < ntiles ? off[at] : 0;
        uint64_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint64_t y = __shfl_up(x, o, 64);
            if (lane >= (uint32_t)o) x += y;
        }
        if (lane == 63) warp_tot[wid] = x;
        __syncthreads();
        uint64_t wpre = 0;
        for (uint32_t w = 0; w < wid; ++w) wpre += warp_tot[w];
        const uint64_t carry = carry_s;
        if (b < ntiles) off[at] = carry + wpre + x;
        __syncthreads();
        if (threadIdx.x == SC_THREADS - 1) carry_s = carry + wpre + x;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_tiles(uint64_t *off, const uint32_t *cnt,
                                                           uint32_t n)
{
    __shared__ uint64_t warp_tot[SC_THREADS / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SC_TILE;
    const uint64_t tile_base = off[base];
    /* thread t takes counts t * SC_ITEMS .. + SC_ITEMS - 1 of the tile */
    const uint64_t first = base + (uint64_t)threadIdx.x * SC_ITEMS;
    uint32_t v[SC_ITEMS];
    uint64_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < SC_ITEMS; ++k) {
        v[k] = first + k < n ? cnt[first + k] : 0u;
        s += v[k];
    }
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint64_t x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) warp_tot[wid] = x;
    __syncthreads();
    uint64_t pre = tile_base + x - s;
    for (uint32_t w = 0; w < wid; ++w) pre += warp_tot[w];
#pragma unroll
    for (uint32_t k = 0; k < SC_ITEMS; ++k) {
        if (first + k < n) off[first + k] = pre;
        pre += v[k];
    }
}

/* In-place exclusive scan of n uint32 (wrapping) over many workgroups:
 * tile sums (SC_TILE elements a tile) into scratch, one workgroup scans
 * them, each tile then scans its own elements from its base. */
__global__ void __launch_bounds__(SC_THREADS) k_scan32_tile_sums(const uint32_t *a, uint32_t n,
                                                                 uint32_t *sums)
{
    __shared__ uint64_t part[SC_THREADS / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SC_TILE;
    uint64_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < SC_ITEMS; ++k) {
        const uint64_t i = base + k * SC_THREADS + threadIdx.x;
        if (i < n) s += a[i];
    }
    const uint64_t t = sc_block_sum(s, part);
    if (threadIdx.x == 0) sums[blockIdx.x] = (uint32_t)t;
}

__global__ void __launch_bounds__(SC_THREADS) k_scan32_tiles(uint32_t *a, uint32_t n,
                                                             const uint32_t *sums)
{
    __shared__ uint32_t warp_tot[SC_THREADS / 64];
    const uint64_t first = (uint64_t)blockIdx.x * SC_TILE + (uint64_t)threadIdx.x * SC_ITEMS;
    uint32_t v[SC_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < SC_ITEMS; ++k) {
        v[k] = first + k < n ? a[first + k] : 0u;
        s += v[k];
    }
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) warp_tot[wid] = x;
    __syncthreads();
    uint32_t pre = sums[blockIdx.x] + x - s;
    for (uint32_t w = 0; w < wid; ++w) pre += warp_tot[w];
#pragma unroll
    for (uint32_t k = 0; k < SC_ITEMS; ++k) {
        if (first + k < n) a[first + k] = pre;
        pre += v[k];
    }
}

int hfx_scan_u32_inplace(uint32_t *d_a, uint32_t n, void *stream)
{
    if (n == 0) return HFX_OK;
    hipStream_t st = hfx_stream(stream);
    if (n < SC_SMALL) {
        hipLaunchKernelGGL(k_scan_u32_inplace, dim3(1), dim3(1024), 0, st, d_a, n);
    } else {
        const uint32_t nt = (uint32_t)(((uint64_t)n + SC_TILE - 1) / SC_TILE);
        hfx_ctx *ctx = hfx_default_ctx();
        HFX_REQUIRE(ctx, HFX_ERR_HIP, "hfx_scan_u32_inplace: no context");
        void *p = nullptr;
        int rc = hfx_ctx_scratch(ctx, HFX_SLOT_SCAN2, (size_t)4 * nt, &p);
        if (rc) return rc;
        uint32_t *sums = (uint32_t *)p;
        hipLaunchKernelGGL(k_scan32_tile_sums, dim3(nt), dim3(SC_THREADS), 0, st, d_a, n, sums);
        hipLaunchKernelGGL(k_scan_u32_inplace, dim3(1), dim3(1024), 0, st, sums, nt);
        hipLaunchKernelGGL(k_scan32_tiles, dim3(nt), dim3(SC_THREADS), 0, st, d_a, n, sums);
    }
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

extern "C" int hfx_scan_offsets(hfx_ctx *ctx, uint64_t *d_off, const uint32_t *d_count, uint32_t n,
                                void *stream)
{
    (void)ctx;
    HFX_REQUIRE(d_off != nullptr, HFX_ERR_ARG, "hfx_scan_offsets: null output");
    hipStream_t st = hfx_stream(stream);
    if (n < SC_SMALL) {
        hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, st, d_off, d_count, n);
    } else {
        const uint32_t nt = (uint32_t)(((uint64_t)n + SC_TILE - 1) / SC_TILE);
        hipLaunchKernelGGL(k_scan_tile_sums, dim3(nt), dim3(SC_THREADS), 0, st, d_off, d_count, n);
        hipLaunchKernelGGL(k_scan_tile_bases, dim3(1), dim3(SC_THREADS), 0, st, d_off, n, nt);
        hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(SC_THREADS), 0, st, d_off, d_count, n);
    }
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}

__global__ void k_prepare_scoring(uint32_t *out, const uint32_t *idx, const uint32_t *begin,
                                  const uint32_t *num, const uint64_t *off)
{
    uint32_t s = blockIdx.x;
    uint32_t b = begin[s], n = num[s];
    uint64_t o = off[s];
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) out[o + i] = idx[b + i];
}

extern "C" void prepare_scoring(uint32_t *d_out_spec_pep_idx, const uint32_t *d_pep_idx_r_sorted_raw,
                                const uint32_t *d_spec_begin, const uint32_t *d_spec_num_pep,
                                const uint32_t num_spec)
{
    if (num_spec == 0) return;
    hfx_ctx *ctx = hfx_default_ctx();
    if (!ctx) return;
    void *p;
    if (hfx_ctx_scratch(ctx, HFX_SLOT_SCAN, ((size_t)num_spec + 1) * 8, &p) != HFX_OK) return;
    uint64_t *off = (uint64_t *)p;
    hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, 0, off, d_spec_num_pep, num_spec);
    hipLaunchKernelGGL(k_prepare_scoring, dim3(num_spec), dim3(256), 0, 0, d_out_spec_pep_idx,
                       d_pep_idx_r_sorted_raw, d_spec_begin, d_spec_num_pep, off);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) hfx_set_error("prepare_scoring: %s", hipGetErrorString(e));
}
