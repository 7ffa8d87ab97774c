/*
 * hfx_preprocess.hip -- fast-XCorr spectrum preprocessing on the GPU.
 *
 * Restates Correlation.hs:37-109 (sequestXCorr = crossCorrolation .
 * normaliseByRegion . observedIntensity), one workgroup per spectrum, the
 * whole spectrum resident in LDS:
 *   1. observedIntensity (:57-71): len = 75 + ceiling(min(50 + pcr*z,
 *      trunc(last_mz/10)*10)); peaks with mz <= cutoff (outside pcr +- 15 when
 *      removing the precursor) binned with rint(mz / bin width), max sqrt(I)
 *      per bin (LDS atomicMax on the bit pattern: values are >= 0).
 *   2. normaliseByRegion (:84-96): 10 regions of sel = (len-75) div 10 bins,
 *      50 * e / region max (max > 1e-6 and i < 10*sel), else 0.
 *   3. crossCorrolation (:105-109): x_i = e_i - sum(a[max(0,i-75) .. +150]) / 150,
 *      the 151-bin window summed left to right in FP32 (mean_mode 0, the
 *      reference's foldl') or in FP64 (mean_mode 1).
 */
#include "hfx_common.h"

#define PP_THREADS 256

__global__ void __launch_bounds__(PP_THREADS)
    k_preprocess(float *x, uint32_t stride, uint32_t *len_out, uint32_t *status, const float *mz,
                 const float *inten, const uint32_t *off, const float *pcr_a, const float *chg_a,
                 int remove_precursor, float bw, int mean_mode)
{
    __shared__ uint32_t sbin[HFX_MAX_LEN];
    __shared__ float a[HFX_MAX_LEN];
    __shared__ uint32_t rmax[11];
    const uint32_t s = blockIdx.x, t = threadIdx.x;
    const uint32_t p0 = off[s], p1 = off[s + 1];
    float *xs = x + (size_t)s * stride;
    const float pcr = pcr_a[s], chg = chg_a[s];
    uint32_t st = 0, len = 0;
    if (p1 <= p0) {
        st = 1;
    } else {
        const float mzcut = 50.0f + pcr * chg;
        const float q = mz[p1 - 1] / 10.0f;
        const float mzlim = (float)((int64_t)__builtin_truncf(q) * 10);
        const float cutoff = (mzcut <= mzlim) ? mzcut : mzlim; /* Haskell min */
        const float lenf = 75.0f + __builtin_ceilf(cutoff);
        if (!(lenf >= 1.0f) || lenf > (float)min(stride, (uint32_t)HFX_MAX_LEN)) {
            st = 3;
        } else {
            len = (uint32_t)lenf;
            if (len < 85) st = 2; /* sel = (len-75) div 10 <= 0 */
        }
        if (st == 0) {
            for (uint32_t i = t; i < len; i += PP_THREADS) sbin[i] = 0;
            if (t < 11) rmax[t] = 0;
            __syncthreads();
            const float plo = pcr - 15.0f, phi = pcr + 15.0f;
            for (uint32_t p = p0 + t; p < p1; p += PP_THREADS) {
                const float m = mz[p];
                const bool keep = remove_precursor ? (m <= cutoff && (m <= plo || phi <= m))
                                                   : (m <= cutoff);
                if (keep) {
                    const int32_t b = (int32_t)__builtin_rintf(m / bw);
                    const float v = __builtin_sqrtf(inten[p]);
                    if (b >= 0 && (uint32_t)b < len)
                        atomicMax(&sbin[b], __float_as_uint(v) & 0x7fffffffu);
                }
            }
            __syncthreads();
            const int32_t sel = (int32_t)(len - 75) / 10;
            for (uint32_t i = t; i < len; i += PP_THREADS) {
                const int32_t rg = min(10, (int32_t)i / sel);
                atomicMax(&rmax[rg], sbin[i]);
            }
            __syncthreads();
            for (uint32_t i = t; i < len; i += PP_THREADS) {
                const int32_t rg = min(10, (int32_t)i / sel);
                const float m = __uint_as_float(rmax[rg]);
                const float e = __uint_as_float(sbin[i]);
                a[i] = (m > 1e-6f && (int32_t)i < sel * 10) ? 50.0f * (e / m) : 0.0f;
            }
            __syncthreads();
            for (uint32_t i = t; i < len; i += PP_THREADS) {
                const uint32_t w0 = i >= 75 ? i - 75 : 0;
                const uint32_t w1 = min(len, w0 + 151);
                float v;
                if (mean_mode == 0) {
                    float sum = 0.0f;
                    for (uint32_t j = w0; j < w1; ++j) sum = sum + a[j];
                    v = a[i] - sum / 150.0f;
                } else {
                    double sum = 0.0;
                    for (uint32_t j = w0; j < w1; ++j) sum = sum + (double)a[j];
                    v = (float)((double)a[i] - sum / 150.0);
                }
                xs[i] = v;
            }
        }
    }
    for (uint32_t i = len + t; i < stride; i += PP_THREADS) xs[i] = 0.0f;
    if (t == 0) {
        len_out[s] = len;
        status[s] = st;
    }
}

extern "C" int hfx_preprocess(hfx_ctx *ctx, float *d_x, uint32_t x_stride, uint32_t *d_len,
                              uint32_t *d_status, const float *d_mz, const float *d_int,
                              const uint32_t *d_peak_off, const float *d_pcr, const float *d_charge,
                              uint32_t num_spec, int remove_precursor, int mono, int mean_mode,
                              void *stream)
{
    (void)ctx;
    HFX_REQUIRE(d_x && d_len && d_status && d_peak_off && d_pcr && d_charge, HFX_ERR_ARG,
                "hfx_preprocess: null argument");
    HFX_REQUIRE(mean_mode == 0 || mean_mode == 1, HFX_ERR_ARG, "hfx_preprocess: mean_mode %d",
                mean_mode);
    if (num_spec == 0) return HFX_OK;
    const float bw = mono ? HFX_BIN_WIDTH_MONO : HFX_BIN_WIDTH_AVG;
    hipLaunchKernelGGL(k_preprocess, dim3(num_spec), dim3(PP_THREADS), 0, hfx_stream(stream), d_x,
                       x_stride, d_len, d_status, d_mz, d_int, d_peak_off, d_pcr, d_charge,
                       remove_precursor, bw, mean_mode);
    HFX_CHECK_LAUNCH();
    return HFX_OK;
}
