// sim_core.h -- deterministic synthetic SMLM data: shared by the host and device simulators.
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11), keyed by the 64-bit seed, counter = (draw, a, b, tag) where (a, b, tag) name
// the quantity drawn (ROI or frame, pixel or emitter, stream).  Poisson variates: inversion for
// lambda < 10, PTRS transformed rejection (W. Hoermann, "The transformed rejection method for
// generating Poisson random variables", Insurance: Mathematics and Economics 12, 1993) otherwise.
// PSF: Gaussian sampled at pixel centres, pixel (r, c) centred at (c + 0.5, r + 0.5):
//   mu = N / (2 pi sx sy) exp(-(dx^2 / 2sx^2 + dy^2 / 2sy^2)) + b   (photons)
// electrons e = Poisson(QE mu); ADU = min(65535, floor(e / KAdc + 0.5) + Offset).
#pragma once
#include <stdint.h>
#include <math.h>
#include "../../include/qcstorm.h"

#if defined(__HIPCC__)
#define QCS_HD __host__ __device__
#else
#define QCS_HD
#endif

namespace qcs_sim {

enum { TAG_ROI_PARAM = 1, TAG_ROI_PIX = 2, TAG_FRAME_EMIT = 3, TAG_FRAME_PIX = 4, TAG_FRAME_BG = 5, TAG_FRAME_SITE = 6,
       TAG_FRAME_MEMBER = 7, TAG_FRAME_DENSITY = 8 };

QCS_HD inline void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Stream of uniforms for one (a, b, tag) triple.
struct Gen {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int used;
    QCS_HD Gen(uint64_t seed, uint32_t a, uint32_t b, uint32_t tag)
    {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
        ctr[0] = 0; ctr[1] = a; ctr[2] = b; ctr[3] = tag;
        used = 4;
    }
    QCS_HD uint32_t next_u32()
    {
        if (used == 4) {
            philox4x32_10(ctr, key, buf);
            ctr[0]++;
            used = 0;
        }
        return buf[used++];
    }
    // uniform double in [0, 1) with 53 random bits
    QCS_HD double u01()
    {
        uint32_t a = next_u32() >> 5, b = next_u32() >> 6;
        return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
    }
};

// log Gamma(x) for x >= 1 by the Stirling series after shifting x to >= 7.
QCS_HD inline double loggam(double x)
{
    double shift = 0.0;
    while (x < 7.0) { shift += log(x); x += 1.0; }
    double x2 = 1.0 / (x * x);
    double s = (((-1.0 / 1680.0 * x2 + 1.0 / 1260.0) * x2 - 1.0 / 360.0) * x2 + 1.0 / 12.0) / x;
    return (x - 0.5) * log(x) - x + 0.91893853320467274178 + s - shift;
}

QCS_HD inline int poisson(double lam, Gen &g)
{
    if (!(lam > 0.0)) return 0;
    if (lam < 10.0) {
        double p = exp(-lam), F = p, u = g.u01();
        int k = 0;
        while (u > F && k < 1000) {
            k++;
            p *= lam / k;
            F += p;
        }
        return k;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < 1000; it++) {
        double U = g.u01() - 0.5;
        double V = g.u01();
        double us = 0.5 - fabs(U);
        double kf = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return (int)kf;
        if (kf < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + kf * loglam - loggam(kf + 1.0))
            return (int)kf;
    }
    return (int)floor(lam + 0.5);
}

QCS_HD inline uint16_t to_adu(int e, const qcs_sim_params &sp)
{
    double v = floor((double)e / (double)sp.kadc + 0.5) + (double)sp.offset;
    if (v > 65535.0) v = 65535.0;
    if (v < 0.0) v = 0.0;
    return (uint16_t)v;
}

struct Emitter {
    double x, y, z, n, sx, sy;
};

QCS_HD inline void astig_sigma(const qcs_sim_params &sp, double z, double *sx, double *sy)
{
    double tx = (z - sp.c_nm) / sp.d_nm, ty = (z + sp.c_nm) / sp.d_nm;
    *sx = sp.s0 * sqrt(1.0 + tx * tx);
    *sy = sp.s0 * sqrt(1.0 + ty * ty);
}

QCS_HD inline double log_uniform(Gen &g, double lo, double hi)
{
    return exp(log(lo) + g.u01() * (log(hi) - log(lo)));
}

// Emitter k of frame `frame` (uniform position over the frame).
QCS_HD inline Emitter frame_emitter(const qcs_sim_params &sp, uint64_t seed, int frame, int k)
{
    Gen g(seed, (uint32_t)frame, (uint32_t)k, TAG_FRAME_EMIT);
    Emitter e;
    e.x = g.u01() * sp.width;
    e.y = g.u01() * sp.height;
    e.n = log_uniform(g, sp.photons_min, sp.photons_max);
    double s = sp.sigma_min + g.u01() * (sp.sigma_max - sp.sigma_min);
    e.z = 0.0;
    e.sx = e.sy = s;
    if (sp.astig) {
        e.z = sp.z_min + g.u01() * (sp.z_max - sp.z_min);
        astig_sigma(sp, e.z, &e.sx, &e.sy);
    }
    return e;
}

QCS_HD inline double frame_background(const qcs_sim_params &sp, uint64_t seed, int frame)
{
    Gen g(seed, (uint32_t)frame, 0u, TAG_FRAME_BG);
    return sp.bg_min + g.u01() * (sp.bg_max - sp.bg_min);
}

QCS_HD inline int frame_emitter_count(const qcs_sim_params &sp)
{
    return (int)floor((double)sp.density_per_px * sp.width * sp.height + 0.5);
}

// Frames with clusters (n_per_unit_max = M > 1) or a density range (density_max_per_px >
// density_per_px): the frame's density is uniform in [density_per_px, density_max_per_px];
// emitters come in sites of 1..M members (uniform count), the first uniform over the frame, the
// others at a separation in [min_sep, max_sep] from it.  Slot k = site * M + member; a slot
// beyond its site's count is empty.  With M = 1 and no range this is frame_emitter.
QCS_HD inline int frame_members(const qcs_sim_params &sp) { return sp.n_per_unit_max > 1 ? sp.n_per_unit_max : 1; }

QCS_HD inline bool frame_plain(const qcs_sim_params &sp)
{
    return frame_members(sp) == 1 && !(sp.density_max_per_px > sp.density_per_px);
}

QCS_HD inline int frame_slots(const qcs_sim_params &sp, uint64_t seed, int frame)
{
    if (frame_plain(sp)) return frame_emitter_count(sp);
    double dens = sp.density_per_px;
    if (sp.density_max_per_px > sp.density_per_px) {
        Gen g(seed, (uint32_t)frame, 0u, TAG_FRAME_DENSITY);
        dens += g.u01() * ((double)sp.density_max_per_px - sp.density_per_px);
    }
    const int M = frame_members(sp);
    const int sites = (int)floor(dens * sp.width * sp.height / (0.5 * (M + 1)) + 0.5);
    return sites * M;
}

QCS_HD inline bool frame_slot_emitter(const qcs_sim_params &sp, uint64_t seed, int frame, int k, Emitter &e)
{
    if (frame_plain(sp)) {
        e = frame_emitter(sp, seed, frame, k);
        return true;
    }
    const int M = frame_members(sp), site = k / M, m = k - site * M;
    Gen g(seed, (uint32_t)frame, (uint32_t)site, TAG_FRAME_SITE);
    int K = 1 + (int)(g.u01() * M);
    K = K > M ? M : K;
    const double x0 = g.u01() * sp.width, y0 = g.u01() * sp.height;
    if (m >= K) return false;
    Gen gm(seed, (uint32_t)frame, (uint32_t)k, TAG_FRAME_MEMBER);
    if (m == 0) {
        e.x = x0;
        e.y = y0;
    } else {
        const double ang = 6.283185307179586476925 * gm.u01();
        const double sep = sp.min_sep + gm.u01() * (sp.max_sep - sp.min_sep);
        e.x = x0 + sep * cos(ang);
        e.y = y0 + sep * sin(ang);
    }
    e.n = log_uniform(gm, sp.photons_min, sp.photons_max);
    const double s = sp.sigma_min + gm.u01() * (sp.sigma_max - sp.sigma_min);
    e.z = 0.0;
    e.sx = e.sy = s;
    if (sp.astig) {
        e.z = sp.z_min + gm.u01() * (sp.z_max - sp.z_min);
        astig_sigma(sp, e.z, &e.sx, &e.sy);
    }
    return true;
}

// Half-width (pixels) of the window an emitter contributes to.
QCS_HD inline double emitter_reach(const Emitter &e) { return 6.0 * (e.sx > e.sy ? e.sx : e.sy); }

QCS_HD inline double psf_value(const Emitter &e, double px, double py)
{
    double dx = px - e.x, dy = py - e.y;
    return e.n / (2.0 * 3.14159265358979323846 * e.sx * e.sy) *
           exp(-(dx * dx / (2.0 * e.sx * e.sx) + dy * dy / (2.0 * e.sy * e.sy)));
}

// ROI unit i: 1..n_per_unit_max emitters, the first within +-center_jitter of the ROI centre,
// the others at a separation in [min_sep, max_sep] from it; all share sigma (or z model).
QCS_HD inline int roi_emitters(const qcs_sim_params &sp, uint64_t seed, int i, Emitter *em, double *bg)
{
    Gen g(seed, (uint32_t)i, 0u, TAG_ROI_PARAM);
    const double R = sp.width;
    int K = 1;
    if (sp.n_per_unit_max > 1) {
        K = 1 + (int)(g.u01() * sp.n_per_unit_max);
        if (K > sp.n_per_unit_max) K = sp.n_per_unit_max;
    }
    double s = sp.sigma_min + g.u01() * (sp.sigma_max - sp.sigma_min);
    *bg = sp.bg_min + g.u01() * (sp.bg_max - sp.bg_min);
    for (int k = 0; k < K; k++) {
        Emitter &e = em[k];
        if (k == 0) {
            e.x = R / 2.0 + (2.0 * g.u01() - 1.0) * sp.center_jitter;
            e.y = R / 2.0 + (2.0 * g.u01() - 1.0) * sp.center_jitter;
        } else {
            double ang = 6.283185307179586476925 * g.u01();
            double sep = sp.min_sep + g.u01() * (sp.max_sep - sp.min_sep);
            e.x = em[0].x + sep * cos(ang);
            e.y = em[0].y + sep * sin(ang);
        }
        e.n = log_uniform(g, sp.photons_min, sp.photons_max);
        e.z = 0.0;
        e.sx = e.sy = s;
        if (sp.astig) {
            e.z = sp.z_min + g.u01() * (sp.z_max - sp.z_min);
            astig_sigma(sp, e.z, &e.sx, &e.sy);
        }
    }
    return K;
}

// Fill one ROI record (R*(R+1) words) and its truth rows (n_per_unit_max x 8 floats).
QCS_HD inline void roi_unit(const qcs_sim_params &sp, uint64_t seed, int i, int first_frame,
                            uint16_t *rec, float *truth)
{
    const int R = sp.width;
    Emitter em[4];
    double bg;
    int K = roi_emitters(sp, seed, i, em, &bg);
    for (int r = 0; r < R; r++)
        for (int c = 0; c < R; c++) {
            double mu = bg;
            for (int k = 0; k < K; k++) mu += psf_value(em[k], c + 0.5, r + 0.5);
            Gen g(seed, (uint32_t)i, (uint32_t)(r * R + c + 1), TAG_ROI_PIX);
            int e = poisson((double)sp.qe * mu, g);
            rec[r * R + c] = to_adu(e, sp);
        }
    const int frame = first_frame + i;
    uint16_t *h = rec + R * R;
    h[0] = 100;
    h[1] = 100;
    h[2] = (uint16_t)(frame % 65536);
    h[3] = (uint16_t)(frame / 65536);
    h[4] = QCS_ROITYPE_SINGLE;
    for (int k = 5; k < R + 0; k++) h[k] = 0;
    if (truth) {
        const int M = sp.n_per_unit_max > 1 ? sp.n_per_unit_max : 1;
        for (int k = 0; k < M; k++) {
            float *t = truth + (long)k * QCS_TRUTH_FLOATS;
            if (k < K) {
                t[0] = (float)i; t[1] = (float)em[k].x; t[2] = (float)em[k].y; t[3] = (float)em[k].z;
                t[4] = (float)em[k].n; t[5] = (float)bg; t[6] = (float)em[k].sx; t[7] = (float)em[k].sy;
            } else {
                for (int j = 0; j < QCS_TRUTH_FLOATS; j++) t[j] = 0.0f;
                t[0] = -1.0f;
            }
        }
    }
}

}  // namespace qcs_sim
