// sanitize_main.cpp -- drives every oracle routine on small synthetic inputs and edge cases, built
// with AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle sanitize).  TEST
// INFRASTRUCTURE: run by tests/test_sanitize.py; a sanitizer report or a non-zero exit fails it.
// Synthetic frames and ROIs come from the library's header-only simulator (sim_core.h, host build).
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <vector>
extern "C" {
#include "orc.h"
}
#include "../qc-storm_amd/csrc/sim_core.h"

static qcs_params params(int W, int H, int R, int loctype, int multi)
{
    qcs_params p;
    memset(&p, 0, sizeof(p));
    p.KAdc = 0.45f; p.Offset = 100.0f; p.QE = 1.0f; p.ReadNoise_e = 1.3f;
    p.ROISize = R; p.LocType = loctype; p.MultiEmitterFitEn = multi; p.WLEEn = 1;
    p.ConsecFit_DistanceTh_nm = 80.0f; p.PixelSize = 100.0f; p.PixelZoom = 4.0f; p.SNR_th = 5.0f;
    p.ImageWidth = W; p.ImageHigh = H; p.TotalFrameNum = 1;
    p.SRImageWidth = (W * 4 + 3) / 4 * 4; p.SRImageHigh = (H * 4 + 3) / 4 * 4;
    p.MinZDepth = -500.0f; p.MaxZDepth = 500.0f; p.ZDepthCorrFactor = 1.0f;
    p.p1_XGY = -111.11f; p.p1_XLY = -111.11f;
    p.RenderingMode = 0; p.FixedLocPrec = 20.0f;
    p.MaxROIPerBatch = QCS_MAX_POINT_NUM; p.BatchedImageNum = 1;
    return p;
}

static qcs_sim_params sim(int W, int H, int clusters)
{
    qcs_sim_params sp;
    memset(&sp, 0, sizeof(sp));
    sp.width = W; sp.height = H; sp.offset = 100.0f; sp.kadc = 0.45f; sp.qe = 1.0f;
    sp.photons_min = 500.0f; sp.photons_max = 5000.0f; sp.sigma_min = 1.0f; sp.sigma_max = 1.6f;
    sp.bg_min = 10.0f; sp.bg_max = 100.0f; sp.density_per_px = clusters ? 0.02f : 0.003f;
    sp.density_max_per_px = clusters ? 0.05f : 0.0f; sp.center_jitter = 1.0f;
    sp.s0 = 1.2f; sp.c_nm = 250.0f; sp.d_nm = 400.0f; sp.z_min = -500.0f; sp.z_max = 500.0f;
    sp.n_per_unit_max = clusters ? 3 : 1; sp.min_sep = 1.0f; sp.max_sep = 2.5f;
    return sp;
}

static std::vector<uint16_t> frames(const qcs_sim_params &sp, uint64_t seed, int nf)
{
    const size_t npx = (size_t)sp.width * sp.height;
    std::vector<uint16_t> out(npx * nf);
    std::vector<double> img(npx);
    for (int f = 0; f < nf; f++) {
        std::fill(img.begin(), img.end(), 0.0);
        const int slots = qcs_sim::frame_slots(sp, seed, f + 1);
        for (int k = 0; k < slots; k++) {
            qcs_sim::Emitter e;
            if (!qcs_sim::frame_slot_emitter(sp, seed, f + 1, k, e)) continue;
            for (int r = 0; r < sp.height; r++)
                for (int c = 0; c < sp.width; c++) img[(size_t)r * sp.width + c] += qcs_sim::psf_value(e, c + 0.5, r + 0.5);
        }
        const double bg = qcs_sim::frame_background(sp, seed, f + 1);
        for (size_t i = 0; i < npx; i++) {
            qcs_sim::Gen g(seed, (uint32_t)(f + 1), (uint32_t)i, qcs_sim::TAG_FRAME_PIX);
            out[npx * f + i] = qcs_sim::to_adu(qcs_sim::poisson(bg + img[i], g), sp);
        }
    }
    return out;
}

static int run_pipeline(const qcs_params &p, const std::vector<uint16_t> &fr, int nf)
{
    const int R = p.ROISize, cap = 20000;
    std::vector<uint16_t> rois((size_t)cap * QCS_ROI_WORDS(R));
    std::vector<float> wle((size_t)cap * QCS_WLE_FLOATS), var(nf + 1);
    int found = 0;
    const int n = orc_extract(&p, fr.data(), nf, 1, rois.data(), wle.data(), cap, &found, nullptr, var.data());
    std::vector<float> loc((size_t)(2 * n + 2) * QCS_REC_FLOATS), out(loc.size());
    std::vector<double> loc64(loc.size());
    orc_fit_stats st;
    const int rows = orc_fit(&p, rois.data(), wle.data(), n, loc.data(), 2 * n + 2, &st);
    orc_fit_f64(&p, rois.data(), wle.data(), n < 64 ? n : 64, loc64.data(), 2 * 64 + 2, nullptr);
    const int m = orc_compact_sorted(loc.data(), rows, 1, nf, out.data());
    std::vector<float> img((size_t)p.SRImageWidth * p.SRImageHigh), zimg(img.size());
    orc_render(&p, out.data(), m, img.data(), zimg.data());
    qcs_loc_stats ls;
    orc_loc_stats(&p, out.data(), m, &ls);
    orc_consec *cs = orc_consec_create();
    std::vector<float> cout((size_t)(m + 1) * QCS_REC_FLOATS), ratio(QCS_ONTIME_BINS);
    orc_consec_fit(&p, cs, out.data(), m / 2, 0, cout.data(), m + 1, ratio.data());
    orc_consec_fit(&p, cs, out.data() + (size_t)(m / 2) * QCS_REC_FLOATS, m - m / 2, 1, cout.data(), m + 1, ratio.data());
    orc_consec_destroy(cs);
    int hist[2048];
    const float mx = orc_display_max_intensity(img.data(), (long long)img.size(), nullptr, hist);
    std::vector<uint8_t> rgb(img.size() * 4);
    orc_display_rgb(&p, img.data(), zimg.data(), (long long)img.size(), mx, 2, rgb.data());
    printf("R %d loctype %d multi %d: %d ROIs, %d rows, %d kept\n", R, p.LocType, p.MultiEmitterFitEn, n, rows, m);
    return n;
}

int main()
{
    int checks = 0;
    // whole pipelines on simulated frames
    {
        const qcs_sim_params sp = sim(200, 160, 0);
        const std::vector<uint16_t> fr = frames(sp, 11, 3);
        for (int R : {5, 7, 9, 13, 17}) checks += run_pipeline(params(200, 160, R, 0, 0), fr, 3) >= 0;
        checks += run_pipeline(params(200, 160, 7, 0, 1), fr, 3) >= 0;
        checks += run_pipeline(params(200, 160, 9, 1, 1), fr, 3) >= 0;
        checks += run_pipeline(params(200, 160, 11, 1, 0), fr, 3) >= 0;
    }
    // high-density frames: mixed-size extraction
    {
        const qcs_sim_params sp = sim(128, 96, 1);
        const std::vector<uint16_t> fr = frames(sp, 12, 2);
        qcs_params p = params(128, 96, 7, 0, 1);
        p.BatchedImageNum = 2;
        std::vector<uint16_t> r7(4000 * QCS_ROI_WORDS(7)), r9(4000 * QCS_ROI_WORDS(9)), r11(4000 * QCS_ROI_WORDS(11));
        uint16_t *const sets[3] = {r7.data(), r9.data(), r11.data()};
        const int caps[3] = {4000, 50, 10};
        int counts[3], nc = 0, nfound = 0;
        std::vector<int> sizes(4000);
        orc_extract_mixed(&p, fr.data(), 2, 1, 4000, sets, caps, counts, sizes.data(), &nc, &nfound);
        printf("mixed: %d candidates, sets %d %d %d\n", nc, counts[0], counts[1], counts[2]);
        checks++;
    }
    // edge inputs: constant, saturated and tiny frames, empty calls
    for (uint16_t v : {(uint16_t)0, (uint16_t)100, (uint16_t)65535}) {
        std::vector<uint16_t> fr((size_t)64 * 48 * 2, v);
        checks += run_pipeline(params(64, 48, 7, 0, 1), fr, 2) == 0;
    }
    {
        std::vector<uint16_t> fr((size_t)16 * 16, 300);
        fr[8 * 16 + 8] = 5000;
        checks += run_pipeline(params(16, 16, 7, 0, 0), fr, 1) >= 0;
        qcs_params p = params(16, 16, 7, 0, 0);
        float loc[12] = {0};
        orc_fit_stats st;
        checks += orc_fit(&p, nullptr, nullptr, 0, loc, 0, &st) == 0;
        checks += orc_compact_sorted(loc, 0, 1, 1, loc) == 0;
    }
    // FP64 maximum likelihood and full-Fisher CRLB, elliptical three-emitter model included
    {
        std::vector<double> img(11 * 11, 20.0);
        for (int i = 0; i < 11 * 11; i++) img[i] += 400.0 * exp(-((i % 11 - 5) * (i % 11 - 5) + (i / 11 - 5) * (i / 11 - 5)) / 3.0);
        double th[16] = {3000, 5.5, 5.5, 1.2, 1.3, 2000, 4.0, 6.5, 1.2, 1.3, 1500, 7.0, 4.5, 1.2, 1.2, 20};
        double ll = 0.0, cr[16];
        orc_mle_f64(img.data(), 11, 3, 1, th, 200, &ll);
        orc_fisher_crlb_f64(11, 3, 1, th, cr);
        checks++;
    }
    // drift correction of a drifting structure (rows sorted by frame)
    {
        qcs_params p = params(64, 64, 7, 0, 0);
        std::vector<float> rows;
        for (int f = 1; f <= 40; f++)
            for (int k = 0; k < 60; k++) {
                float r[12] = {800.0f, 8.0f + (float)((k * 37) % 48) + 0.01f * f, 8.0f + (float)((k * 53) % 48) - 0.02f * f, 0.0f,
                               1.3f, 1.3f, 5000.0f, 20.0f, 30.0f, 10.0f, 10.0f, (float)f};
                rows.insert(rows.end(), r, r + 12);
            }
        std::vector<float> fx(64), fy(64), gx(64), gy(64);
        int total = 0;
        const int g = orc_drift_correct(&p, rows.data(), (long long)(rows.size() / 12), 10, fx.data(), fy.data(), 64, &total,
                                        gx.data(), gy.data());
        printf("drift: %d groups over %d frames\n", g, total);
        checks++;
    }
    printf("sanitized oracle checks: %d\n", checks);
    return 0;
}
