/*
 * hfx_cand.h -- candidate sets scored by the sparse XCorr and top-N kernels.
 *
 * A candidate set maps (spectrum s, k-th candidate of s) to the residues,
 * residual and residue masses of one theoretical spectrum:
 *   HfxPlainCands  unmodified peptides of the precursor window, in residual-
 *                  sorted order (Search.hs:151-165, prepare_scoring.cu:47-67);
 *   HfxModCands    modified peptides enumerated for the spectrum
 *                  (Search.hs:358-394: rows of the spectrum expanded to their
 *                  modified peptides, in enumeration order).
 */
#pragma once

#include "hfx_common.h"

#include <type_traits>

/* Stage spectrum s of sp in LDS (xs) and the mass table (mtab); returns the
 * effective length, never above HFX_MAX_LEN or the stride. */
__device__ __forceinline__ uint32_t hfx_stage_spectrum(float *xs, float *mtab, const hfx_db &db,
                                                       const hfx_spectra &sp, uint32_t s)
{
    const uint32_t len = min(sp.d_len[s], min(sp.x_stride, (uint32_t)HFX_MAX_LEN));
    const float *xg = sp.d_x + (size_t)s * sp.x_stride;
    /* eight loads in flight per thread (clamped into the row) before the stores */
    constexpr uint32_t U = 8;
    for (uint32_t base = 0; base < len; base += U * blockDim.x) {
        float v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) v[u] = xg[min(base + u * blockDim.x + threadIdx.x, len - 1u)];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t i = base + u * blockDim.x + threadIdx.x;
            if (i < len) xs[i] = v[u];
        }
    }
    if (threadIdx.x < 32) mtab[threadIdx.x] = threadIdx.x < 26 ? db.d_mass[threadIdx.x] : 0.0f;
    return len;
}

__device__ __forceinline__ uint32_t hfx_take(uint32_t v, uint32_t l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ float hfx_take(float v, uint32_t l)
{
    return __uint_as_float(hfx_take(__float_as_uint(v), l));
}

/* Max fragment charge actually generated: the reference launches nothing for
 * max_charge > 4 (ion_series.cu:334-342), leaving zero rows. */
__device__ __forceinline__ uint32_t hfx_zeff(uint32_t zmax)
{
    return zmax <= HFX_MAX_CHARGE ? zmax : 0u;
}

struct HfxPlainCands {
    hfx_db db;
    const uint32_t *begin; /* [S] first residual-sorted position of the window */
    const uint32_t *num;   /* [S] window size */
    const uint64_t *off;   /* [S+1] score offsets */

    __device__ __forceinline__ uint32_t count(uint32_t s) const { return num[s]; }
    __device__ __forceinline__ uint64_t base(uint32_t s) const { return off[s]; }
    __device__ __forceinline__ uint32_t zeff(uint32_t zmax) const { return hfx_zeff(zmax); }

    /* what the peaks of candidate k of spectrum s need from global memory
     * (prep: one lane's fetch; take: lane l's, to the whole wave) */
    struct Prep {
        uint32_t tc, n;
        float r;
    };
    __device__ __forceinline__ Prep prep(uint32_t s, uint32_t k) const
    {
        const uint32_t pep = db.d_idx_sorted[begin[s] + k];
        const uint32_t tc = db.d_tc[pep], tn = db.d_tn[pep];
        return Prep{tc, tn > tc ? tn - tc : 0u, db.d_r[pep]};
    }
    __device__ __forceinline__ static Prep take(const Prep &p, uint32_t l)
    {
        return Prep{hfx_take(p.tc, l), hfx_take(p.n, l), hfx_take(p.r, l)};
    }
    __device__ __forceinline__ void peaks_prep(uint32_t *map, const float *mtab, const Prep &p,
                                               uint32_t len, uint32_t zeff) const
    {
        hfx_wave_peaks(map, p.n, p.r, zeff, len, HfxPlainMass{db.d_ions + p.tc, mtab});
    }
    /* lane l's residue code of the candidate's first 64 (loaded a candidate
     * ahead), and the peaks from it */
    __device__ __forceinline__ uint32_t first_res(const Prep &p) const
    {
        return hfx_lane() < p.n ? (uint32_t)db.d_ions[p.tc + hfx_lane()] : 0u;
    }
    __device__ __forceinline__ void peaks_first(uint32_t *map, const float *mtab, const Prep &p,
                                                uint32_t first, uint32_t len, uint32_t zeff) const
    {
        hfx_wave_peaks(map, p.n, p.r, zeff, len, HfxPlainMassFirst{db.d_ions + p.tc, mtab, first});
    }
    /* peaks of candidate k of spectrum s into the wave's map */
    __device__ __forceinline__ void peaks(uint32_t *map, const float *mtab, uint32_t s, uint32_t k,
                                          uint32_t len, uint32_t zeff) const
    {
        peaks_prep(map, mtab, prep(s, k), len, zeff);
    }
    /* one lane: the candidate's b ladder (sequential FP32 prefix of its
     * first nl residue masses) into dst, four residues loaded ahead (the
     * last one re-read past the end) */
    __device__ __forceinline__ void ladder(const Prep &p, const float *mtab, float *dst,
                                           uint32_t nl) const
    {
        const uint8_t *res = db.d_ions + p.tc;
        float bl = 0.0f;
        for (uint32_t k = 0; k < nl; k += 4) {
            uint32_t a[4];
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) a[j] = res[min(k + j, nl - 1u)];
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                if (k + j < nl) {
                    bl = bl + mtab[HFX_ACID_IDX(a[j])];
                    dst[k + j] = bl;
                }
            }
        }
    }
};

struct HfxModCands {
    hfx_db db;
    hfx_modset ms;
    const uint64_t *off; /* [S+1] first modified peptide of each spectrum */

    __device__ __forceinline__ uint32_t count(uint32_t s) const
    {
        return (uint32_t)(off[s + 1] - off[s]);
    }
    __device__ __forceinline__ uint64_t base(uint32_t s) const { return off[s]; }
    /* addModIons dispatches max charges 1..10 (mion_series.cu:335-556) */
    __device__ __forceinline__ uint32_t zeff(uint32_t zmax) const { return zmax <= 10 ? zmax : 0u; }

    struct Prep {
        uint32_t tc, n, mod, uoff;
        float r;
    };
    __device__ __forceinline__ Prep prep(uint32_t s, uint32_t k) const
    {
        const uint64_t i = off[s] + k;
        const uint32_t pep = ms.d_mpep_pep_idx[i], mod = ms.d_mpep_mod_idx[i];
        const uint32_t tc = db.d_tc[pep], tn = db.d_tn[pep];
        return Prep{tc, tn > tc ? tn - tc : 0u, mod, ms.d_mpep_unrank_off[i],
                    db.d_r[pep] + ms.d_mod_delta[mod]};
    }
    __device__ __forceinline__ static Prep take(const Prep &p, uint32_t l)
    {
        return Prep{hfx_take(p.tc, l), hfx_take(p.n, l), hfx_take(p.mod, l), hfx_take(p.uoff, l),
                    hfx_take(p.r, l)};
    }
    __device__ __forceinline__ void peaks_prep(uint32_t *map, const float *mtab, const Prep &p,
                                               uint32_t len, uint32_t zeff) const
    {
        const uint32_t na = min(ms.num_ma, (uint32_t)HFX_MAX_MA);
        HfxModMass mm{db.d_ions + p.tc, mtab, ms.d_mod_ma_count + (size_t)p.mod * ms.num_ma,
                      ms.d_unrank + p.uoff, ms.d_ma_mass, na, {0, 0, 0, 0, 0, 0, 0}};
        hfx_wave_peaks(map, p.n, p.r, zeff, len, mm);
    }
    /* no read ahead for modified candidates (their masses need the unrank
     * slice as well) */
    __device__ __forceinline__ uint32_t first_res(const Prep &) const { return 0u; }
    __device__ __forceinline__ void peaks_first(uint32_t *map, const float *mtab, const Prep &p,
                                                uint32_t, uint32_t len, uint32_t zeff) const
    {
        peaks_prep(map, mtab, p, len, zeff);
    }
    __device__ __forceinline__ void peaks(uint32_t *map, const float *mtab, uint32_t s, uint32_t k,
                                          uint32_t len, uint32_t zeff) const
    {
        peaks_prep(map, mtab, prep(s, k), len, zeff);
    }
    /* one lane: the modified candidate's b ladder over its first nl (<= 64)
     * positions.  A residue of modable acid a (code field a + 1) that is the
     * i-th of its acid in the peptide weighs table + ma_mass[a] when i is in
     * the acid's part of the unrank slice (HfxModMass's rule, per lane):
     * those parts become one 64-bit occurrence mask per acid. */
    __device__ __forceinline__ void ladder(const Prep &p, const float *mtab, float *dst,
                                           uint32_t nl) const
    {
        const uint32_t na = min(ms.num_ma, (uint32_t)HFX_MAX_MA);
        const uint32_t *row = ms.d_mod_ma_count + (size_t)p.mod * ms.num_ma;
        const uint32_t *u = ms.d_unrank + p.uoff;
        uint64_t mask[HFX_MAX_MA];
        uint32_t occ[HFX_MAX_MA];
        float mam[HFX_MAX_MA];
        uint32_t off = 0;
#pragma unroll
        for (uint32_t a = 0; a < HFX_MAX_MA; ++a) {
            mask[a] = 0;
            occ[a] = 0;
            mam[a] = 0.0f;
            if (a < na) {
                const uint32_t k = row[a];
                for (uint32_t q = 0; q < k; ++q) {
                    const uint32_t i = u[off + q];
                    if (i < 64) mask[a] |= 1ull << i;
                }
                off += k;
                mam[a] = ms.d_ma_mass[a];
            }
        }
        /* the residue loop tests only as many acids as the batch has: one
         * instance for up to four (the usual M, S, T, Y), one for more */
        const uint8_t *res = db.d_ions + p.tc;
        auto walk = [&](auto nac) {
            constexpr uint32_t NA = decltype(nac)::value;
            float bl = 0.0f;
            for (uint32_t k = 0; k < nl; ++k) {
                const uint32_t ion = res[k];
                const uint32_t md = HFX_ACID_MOD(ion);
                float m = mtab[HFX_ACID_IDX(ion)];
#pragma unroll
                for (uint32_t a = 0; a < NA; ++a)
                    if (md == a + 1) {
                        if ((mask[a] >> occ[a]) & 1ull) m = m + mam[a];
                        ++occ[a];
                    }
                bl = bl + m;
                dst[k] = bl;
            }
        };
        if (na <= 4) walk(std::integral_constant<uint32_t, 4>{});
        else walk(std::integral_constant<uint32_t, HFX_MAX_MA>{});
    }
};
