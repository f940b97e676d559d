/*
 * orc_fit_impl.h -- BFGS Poisson-MLE fitters of the reference, restated for the CPU oracle
 * (TEST INFRASTRUCTURE).  Included twice: once with REAL = float (FP32 oracle) and once with
 * REAL = double (the same schedule in FP64).  The including file defines
 *   REAL, SUF(name), FMA(a,b,c), EXPR(x), LOGR(x), LOGU(x,c) and LOGUB(x,c) (log x + c, fused
 *   where the arithmetic mode fuses it: the multi- and one-emitter losses' table logs),
 *   SQRTR(x), SEPARABLE(), C(x) (literal cast)
 *
 * Reference files (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll):
 *   shared BFGS machinery  DLL/bfgsMLE_Loc_LD/bfgsMLE_core.h:51-238
 *   GS2D 1-emitter         DLL/bfgsMLE_Loc_LD/bfgsMLE_GS2D_core.h:56-575
 *   AS3D 1-emitter         DLL/bfgsMLE_Loc_LD/bfgsMLE_AS3D_core.h:57-607
 *   GS2D 2-emitter         DLL/bfgsMLE_Loc_HD/bfgsMLE_GS2D_2Emitter_core.h:74-698
 *   GS2D 3-emitter         DLL/bfgsMLE_Loc_HD/bfgsMLE_GS2D_3Emitter_core.h:76-710
 *   AS3D 2-emitter         DLL/bfgsMLE_Loc_HD/bfgsMLE_AS3D_2Emitter_core.h:77-750
 *   cascade                DLL/bfgsMLE_Loc_top/bfgs_top.cu:26-204,
 *                          DLL/bfgsMLE_Loc_top/bfgs_top_MoleculeClasify.cu:26-92
 *   CRLB                   DLL/bfgsMLE_Loc_top/LocPrecCalc_GaussianCRLB.h:64-243
 */

#define MAXP   11
#define MAXR   17
#define MAXPIX (MAXR * MAXR)

/* model ids */
#define M_GS2D   0
#define M_AS3D   1
#define M_2E     2
#define M_3E     3
#define M_AS3D2E 4

typedef struct SUF(mdesc) {
    int P, iters, K, ellip;       /* params, BFGS iterations, emitters, elliptical sigma */
    int ia[3], ix[3], iy[3], isx[3], isy[3], ib;
    REAL clampv;                  /* loss clamps params below clampv to clampv (0 or 0.01) */
} SUF(mdesc);

static void SUF(get_desc)(int model, SUF(mdesc) *m)
{
    int k;
    switch (model) {
    case M_GS2D:   /* bfgsMLE_GS2D_core.h:25-37 */
        m->P = 5; m->iters = 8; m->K = 1; m->ellip = 0;
        m->ia[0] = 0; m->ix[0] = 1; m->iy[0] = 2; m->isx[0] = m->isy[0] = 3; m->ib = 4;
        m->clampv = C(0.0); break;
    case M_AS3D:   /* bfgsMLE_AS3D_core.h:25-38, loss clamp :501-505 */
        m->P = 6; m->iters = 11; m->K = 1; m->ellip = 1;
        m->ia[0] = 0; m->ix[0] = 1; m->iy[0] = 2; m->isx[0] = 3; m->isy[0] = 4; m->ib = 5;
        m->clampv = C(0.01); break;
    case M_2E:     /* bfgsMLE_GS2D_2Emitter_core.h:24-45 */
        m->P = 8; m->iters = 13; m->K = 2; m->ellip = 0;
        for (k = 0; k < 2; k++) { m->ia[k] = 3 * k; m->ix[k] = 3 * k + 1; m->iy[k] = 3 * k + 2; m->isx[k] = m->isy[k] = 6; }
        m->ib = 7; m->clampv = C(0.0); break;
    case M_3E:     /* bfgsMLE_GS2D_3Emitter_core.h:24-49 */
        m->P = 11; m->iters = 16; m->K = 3; m->ellip = 0;
        for (k = 0; k < 3; k++) { m->ia[k] = 3 * k; m->ix[k] = 3 * k + 1; m->iy[k] = 3 * k + 2; m->isx[k] = m->isy[k] = 9; }
        m->ib = 10; m->clampv = C(0.0); break;
    default:       /* M_AS3D2E: bfgsMLE_AS3D_2Emitter_core.h:24-49, clamp :590-594 */
        m->P = 11; m->iters = 16; m->K = 2; m->ellip = 1;
        for (k = 0; k < 2; k++) { m->ia[k] = 5 * k; m->ix[k] = 5 * k + 1; m->iy[k] = 5 * k + 2; m->isx[k] = 5 * k + 3; m->isy[k] = 5 * k + 4; }
        m->ib = 10; m->clampv = C(0.01); break;
    }
}

/* Per-ROI fit context: photo-electron pixels q, their logs, c1 = 1 - log q, and the WLE weights. */
typedef struct SUF(roictx) {
    int R;
    REAL q[MAXPIX], lq[MAXPIX], c1[MAXPIX];
    REAL w[MAXPIX];              /* literal weights: exp(-(sx1 dx^2 + sy1 dy^2)) */
    REAL wx[MAXR], wy[MAXR];     /* separable weights: w = wy[r] * wx[c] */
    int weighted;
} SUF(roictx);

/*
 * Poisson deviance with WLE weights, sum over pixels of
 *   w * (I - q - q (log I - log q)),  I = sum_k A_k exp(-(sx dx^2 + sy dy^2)) + B,
 * with the in-place clamping of the argument vector the reference performs
 * (bfgsMLE_GS2D_core.h:458-518; AS3D :484-544; 2E :518-592; 3E :510-589; AS3D-2E :569-644).
 * Literal mode sums in row-major order per pixel as the reference writes it.  Separable mode
 * (the HIP library's arithmetic) forms the pixel term as fma(-q, LOGU(I, c1), I) with
 * c1 = 1 - log q, i.e. I - q (1 + log I - log q), and sums sum_r wy[r] sum_c wx[c] t(r, c);
 * one-emitter models take the binade table log (LOGUB), multi-emitter ones the 32-pair one.
 */
static REAL SUF(loss)(const SUF(mdesc) *m, const SUF(roictx) *rc, REAL *p)
{
    const int R = rc->R;
    int i, r, c, k;
    orc_count_loss_eval();
    for (i = 0; i < m->P; i++)
        if (p[i] < m->clampv) p[i] = m->clampv;

    const REAL b = p[m->ib] * C(64.0);
    REAL total = C(0.0);

    if (SEPARABLE()) {
        REAL ex[3][MAXR], ay[3][MAXR];
        for (k = 0; k < m->K; k++) {
            REAL a = p[m->ia[k]] * C(128.0);
            REAL sx = p[m->isx[k]] * C(0.1);
            REAL sy = p[m->isy[k]] * C(0.1);
            for (c = 0; c < R; c++) {
                REAL dx = ((REAL)c + C(0.5)) - p[m->ix[k]];
                ex[k][c] = EXPR((-(dx * dx)) * sx);
            }
            for (r = 0; r < R; r++) {
                REAL dy = ((REAL)r + C(0.5)) - p[m->iy[k]];
                ay[k][r] = a * EXPR((-(dy * dy)) * sy);
            }
        }
        for (r = 0; r < R; r++) {
            REAL acc = C(0.0);
            for (c = 0; c < R; c++) {
                REAL I;
                if (m->K == 1) {
                    I = FMA(ay[0][r], ex[0][c], b);
                } else {
                    I = ay[0][r] * ex[0][c];
                    for (k = 1; k < m->K; k++) I = FMA(ay[k][r], ex[k][c], I);
                    I = I + b;
                }
                const int pi = r * R + c;
                const REAL u = (m->K == 1) ? LOGUB(I, rc->c1[pi]) : LOGU(I, rc->c1[pi]);
                const REAL t = FMA(-rc->q[pi], u, I);
                acc = rc->weighted ? FMA(rc->wx[c], t, acc) : acc + t;
            }
            total = rc->weighted ? FMA(rc->wy[r], acc, total) : total + acc;
        }
    } else {
        for (r = 0; r < R; r++) {
            const REAL rowpos = (REAL)r + C(0.5);
            for (c = 0; c < R; c++) {
                const REAL colpos = (REAL)c + C(0.5);
                REAL I = C(0.0);
                for (k = 0; k < m->K; k++) {
                    REAL dx = colpos - p[m->ix[k]];
                    REAL dy = rowpos - p[m->iy[k]];
                    REAL e;
                    if (m->ellip)
                        e = EXPR(-FMA((dx * dx) * p[m->isx[k]], C(0.1), ((dy * dy) * p[m->isy[k]]) * C(0.1)));
                    else
                        e = EXPR(((-FMA(dx, dx, dy * dy)) * p[m->isx[k]]) * C(0.1));
                    REAL a = p[m->ia[k]] * C(128.0);
                    if (m->K == 1) I = FMA(a, e, b);
                    else if (k == 0) I = a * e;
                    else I = FMA(a, e, I);
                }
                if (m->K > 1) I = I + b;
                const int pi = r * R + c;
                const REAL q = rc->q[pi];
                REAL t = FMA(-q, LOGR(I) - rc->lq[pi], I - q);
                total = rc->weighted ? FMA(rc->w[pi], t, total) : total + t;
            }
        }
    }
    return total;
}

/* Central-difference gradient, +-0.002 in scaled units; the loss clamps the shared probe
 * vector in place (NumericalGradient_GS2D, bfgsMLE_GS2D_core.h:429-454). */
static void SUF(numgrad)(const SUF(mdesc) *m, const SUF(roictx) *rc, const REAL *x, REAL *g)
{
    REAL t[MAXP];
    int i;
    for (i = 0; i < m->P; i++) t[i] = x[i];
    for (i = 0; i < m->P; i++) {
        t[i] = t[i] - C(0.002);
        REAL ln = SUF(loss)(m, rc, t);
        t[i] = t[i] + C(0.004);
        REAL lp = SUF(loss)(m, rc, t);
        g[i] = (lp - ln) * C(250.0);
        t[i] = t[i] - C(0.002);
    }
}

/* D <- (I - s y'/y's) D (I - y s'/y's) + s s'/y's  (ConstructD0, bfgsMLE_core.h:87-182) */
static void SUF(bfgs_update)(int P, REAL *D, const REAL *s, const REAL *y)
{
    REAL A[MAXP * MAXP], T[MAXP * MAXP], ty[MAXP], ts[MAXP];
    int r, c, k;
    REAL sy = C(0.0);
    for (k = 0; k < P; k++) sy = FMA(s[k], y[k], sy);
    sy = C(1.0) / sy;
    for (k = 0; k < P; k++) { ty[k] = y[k] * sy; ts[k] = s[k] * sy; }
    for (r = 0; r < P; r++)
        for (c = 0; c < P; c++)
            A[r * P + c] = FMA(-s[r], ty[c], (r == c) ? C(1.0) : C(0.0));
    for (r = 0; r < P; r++)
        for (c = 0; c < P; c++) {
            REAL acc = C(0.0);
            for (k = 0; k < P; k++) acc = FMA(A[r * P + k], D[k * P + c], acc);
            T[r * P + c] = acc;
        }
    for (r = 0; r < P; r++)
        for (c = 0; c < P; c++) {
            REAL acc = C(0.0);
            for (k = 0; k < P; k++) acc = FMA(T[r * P + k], A[c * P + k], acc);
            D[r * P + c] = acc;
        }
    for (r = 0; r < P; r++)
        for (c = 0; c < P; c++)
            D[r * P + c] = D[r * P + c] + s[r] * ts[c];
}

/*
 * Fixed-schedule BFGS with bisection line search (MLELocalization_GS2D,
 * bfgsMLE_GS2D_core.h:313-426; the other models use the same loop with their own P and
 * iteration count): direction normalised to mean |d| = 1; bisection over [1e-5, 1] with probe
 * points 1e-4 and 1, 2 + 10 loss evaluations; step clamped at >= 0.01; gradient and inverse
 * Hessian update skipped on the last iteration.
 */
static void SUF(bfgs_run)(const SUF(mdesc) *m, const SUF(roictx) *rc, REAL *x, int iters)
{
    const int P = m->P;
    REAL D[MAXP * MAXP], g[MAXP], d[MAXP], s[MAXP], y[MAXP], go[MAXP];
    REAL xd[2][MAXP], L[2], pos[2];
    int i, j, it, bs;
    for (i = 0; i < P; i++)
        for (j = 0; j < P; j++) D[i * P + j] = (i == j) ? C(1.0) : C(0.0);

    SUF(numgrad)(m, rc, x, g);
    for (i = 0; i < P; i++) d[i] = -g[i];

    for (it = 0; it < iters; it++) {
        REAL tot = C(0.0);
        for (i = 0; i < P; i++) tot = tot + FABSR(d[i]);
        tot = (REAL)P / tot;
        for (i = 0; i < P; i++) d[i] = d[i] * tot;

        pos[0] = C(0.00001);
        pos[1] = C(1.0);
        for (i = 0; i < P; i++) { xd[0][i] = FMA(d[i], C(0.0001), x[i]); xd[1][i] = FMA(d[i], C(1.0), x[i]); }
        L[0] = SUF(loss)(m, rc, xd[0]);
        L[1] = SUF(loss)(m, rc, xd[1]);
        for (bs = 0; bs < 11; bs++) {
            int sel = (L[0] < L[1]);
            pos[sel] = (pos[0] + pos[1]) * C(0.5);
            if (bs < 10) {
                for (i = 0; i < P; i++) xd[sel][i] = FMA(d[i], pos[sel], x[i]);
                L[sel] = SUF(loss)(m, rc, xd[sel]);
            }
        }
        REAL scale = (pos[0] + pos[1]) * C(0.5);
        for (i = 0; i < P; i++) { s[i] = d[i] * scale; x[i] = x[i] + s[i]; }
        for (i = 0; i < P; i++) if (x[i] < C(0.01)) x[i] = C(0.01);   /* IninfConstrain */

        if (it < iters - 1) {
            for (i = 0; i < P; i++) go[i] = g[i];
            SUF(numgrad)(m, rc, x, g);
            for (i = 0; i < P; i++) y[i] = g[i] - go[i];
            SUF(bfgs_update)(P, D, s, y);
            for (i = 0; i < P; i++) {                                 /* MatMulVector */
                REAL acc = C(0.0);
                for (j = 0; j < P; j++) acc = FMA(-D[i * P + j], g[j], acc);
                d[i] = acc;
            }
        }
    }
}

/* pix = max((adu - Offset) * KAdc, 1)  (bfgsMLE_GS2D_core.h:144) */
static void SUF(load_roi)(const qcs_params *p, const uint16_t *roi, int R, SUF(roictx) *rc)
{
    int i;
    rc->R = R;
    const REAL off = (REAL)p->Offset, kadc = (REAL)p->KAdc;
    for (i = 0; i < R * R; i++) {
        REAL v = ((REAL)roi[i] - off) * kadc;
        rc->q[i] = (v > C(1.0)) ? v : C(1.0);
        rc->lq[i] = LOGR(rc->q[i]);
        rc->c1[i] = C(1.0) - rc->lq[i];
    }
}

static REAL SUF(fminr)(REAL a, REAL b)
{   /* fminf semantics: a NaN operand yields the other operand */
    if (a != a) return b;
    if (b != b) return a;
    return a < b ? a : b;
}

/* WLEWeightsCalc (bfgsMLE_core.h:206-238): 1 for MLEFit, else a Gaussian mask centred on the
 * ROI.  The separable evaluation factors the mask into exp(-sx1 dx^2) * exp(-sy1 dy^2). */
static void SUF(wle_weights)(SUF(roictx) *rc, REAL sgx, REAL sgy, int moltype)
{
    const int R = rc->R;
    const REAL ctr = (REAL)R / C(2.0);
    const REAL sx1 = C(1.0) / ((C(2.0) * sgx) * sgx);
    const REAL sy1 = C(1.0) / ((C(2.0) * sgy) * sgy);
    int r, c;
    rc->weighted = 1;
    for (c = 0; c < R; c++) {
        REAL d = ((REAL)c + C(0.5)) - ctr;
        rc->wx[c] = (moltype == QCS_MOLTYPE_MLEFIT) ? C(1.0) : EXPR(-((d * d) * sx1));
        rc->wy[c] = (moltype == QCS_MOLTYPE_MLEFIT) ? C(1.0) : EXPR(-((d * d) * sy1));
    }
    for (r = 0; r < R; r++)
        for (c = 0; c < R; c++) {
            if (moltype == QCS_MOLTYPE_MLEFIT) { rc->w[r * R + c] = C(1.0); continue; }
            REAL dx = ((REAL)c + C(0.5)) - ctr;
            REAL dy = ((REAL)r + C(0.5)) - ctr;
            rc->w[r * R + c] = EXPR(-FMA(dx * dx, sx1, (dy * dy) * sy1));
        }
}

/* border sums: Mean1 = left column, Mean2 = right column, Mean3 = top row, Mean4 = bottom row */
static void SUF(border_sums)(const SUF(roictx) *rc, REAL *m1, REAL *m2, REAL *m3, REAL *m4)
{
    const int R = rc->R;
    int k;
    REAL a = C(0.0), b = C(0.0), c = C(0.0), d = C(0.0);
    for (k = 0; k < R; k++) {
        a = a + rc->q[k * R + 0];
        b = b + rc->q[k * R + R - 1];
        c = c + rc->q[0 * R + k];
        d = d + rc->q[(R - 1) * R + k];
    }
    *m1 = a; *m2 = b; *m3 = c; *m4 = d;
}

/* 3x3 block sums used to place extra emitters (PreEstimation_GS2D_2E, bfgsMLE_GS2D_2Emitter_core.h:634-681) */
static void SUF(block_peaks)(const SUF(roictx) *rc, int npk, REAL *px, REAL *py)
{
    const int R = rc->R, half = R / 2;
    const REAL ctr = (REAL)R / C(2.0);
    const int L = R / 3 + 1, Lh = L / 2;
    int start[3] = {0, half - Lh, R - L};
    REAL rpos[3];
    REAL sum[9];
    int cnt, r, c, k;
    rpos[0] = ctr - (REAL)L / C(2.0);
    rpos[1] = ctr;
    rpos[2] = ctr + (REAL)L / C(2.0);
    for (cnt = 0; cnt < 9; cnt++) {
        REAL s = C(0.0);
        int xb = start[cnt % 3], yb = start[cnt / 3];
        for (r = 0; r < L; r++)
            for (c = 0; c < L; c++) s = s + rc->q[(r + yb) * R + (c + xb)];
        sum[cnt] = s;
    }
    sum[4] = C(0.0);
    for (k = 0; k < npk; k++) {
        int mx = 0;
        for (cnt = 0; cnt < 9; cnt++) if (sum[cnt] > sum[mx]) mx = cnt;
        px[k] = rpos[mx % 3];
        py[k] = rpos[mx / 3];
        sum[mx] = C(0.0);
    }
}

/* initial parameter vector of each model */
static void SUF(preestimate)(int model, const SUF(roictx) *rc, REAL wsx, REAL wsy, REAL *x)
{
    const int R = rc->R, half = R / 2;
    const REAL ctr = (REAL)R / C(2.0);
    REAL m1, m2, m3, m4, bkg, amp;
    SUF(border_sums)(rc, &m1, &m2, &m3, &m4);
    m1 = m1 < m2 ? m1 : m2;   /* min() of two non-NaN sums */
    m3 = m3 < m4 ? m3 : m4;
    if (model == M_2E) {
        /* PreEstimation_GS2D_2E :611-626 uses the smallest of the four border sums */
        REAL mm = m1 < m3 ? m1 : m3;
        bkg = mm / (REAL)R;
    } else {
        /* PreEstimation_GS2D :535-546 */
        bkg = ((m1 + m3) / C(2.0)) / (REAL)R;
    }
    amp = rc->q[half * R + half] - bkg;

    if (model == M_GS2D) {   /* PreEstimation_GS2D :521-575 */
        REAL ps = C(1.0) / ((C(2.0) * wsx) * wsx);
        x[0] = amp * C(0.0078125);
        x[1] = (REAL)half + C(0.5);
        x[2] = (REAL)half + C(0.5);
        x[3] = ps * C(10.0);
        x[4] = bkg * C(0.015625);
    } else if (model == M_AS3D) {  /* PreEstimation_AS3D :547-607 */
        x[0] = amp * C(0.0078125);
        x[1] = (REAL)half + C(0.5);
        x[2] = (REAL)half + C(0.5);
        x[3] = C(10.0) * (C(0.5) / (wsx * wsx));
        x[4] = C(10.0) * (C(0.5) / (wsy * wsy));
        x[5] = bkg * C(0.015625);
    } else {
        const REAL s0 = ((REAL)R / C(2.2)) / C(2.35);
        REAL px[2], py[2];
        SUF(block_peaks)(rc, (model == M_3E) ? 2 : 1, px, py);
        if (model == M_2E) {   /* :684-697 */
            REAL s1 = C(1.0) / ((C(2.0) * s0) * s0);
            x[0] = (C(0.8) * amp) * C(0.0078125); x[1] = ctr; x[2] = ctr;
            x[3] = (C(0.8) * amp) * C(0.0078125); x[4] = px[0]; x[5] = py[0];
            x[6] = s1 * C(10.0);
            x[7] = (C(0.8) * bkg) * C(0.015625);
        } else if (model == M_3E) {   /* bfgsMLE_GS2D_3Emitter_core.h:693-708 */
            REAL s1 = C(1.0) / ((C(2.0) * s0) * s0);
            x[0] = (C(0.7) * amp) * C(0.0078125); x[1] = ctr; x[2] = ctr;
            x[3] = (C(0.7) * amp) * C(0.0078125); x[4] = px[0]; x[5] = py[0];
            x[6] = (C(0.3) * amp) * C(0.0078125); x[7] = px[1]; x[8] = py[1];
            x[9] = s1 * C(10.0);
            x[10] = (C(0.8) * bkg) * C(0.015625);
        } else {   /* M_AS3D2E: bfgsMLE_AS3D_2Emitter_core.h:679-749 */
            REAL s1 = C(10.0) * (C(0.5) / (s0 * s0));
            x[0] = (C(0.7) * amp) * C(0.0078125); x[1] = ctr; x[2] = ctr; x[3] = s1; x[4] = s1;
            x[5] = (C(0.7) * amp) * C(0.0078125); x[6] = px[0]; x[7] = py[0]; x[8] = s1; x[9] = s1;
            x[10] = (C(0.8) * bkg) * C(0.015625);
        }
    }
}

/* Weighted-likelihood setup per model; returns the initial vector in x. */
static void SUF(setup)(const qcs_params *p, int model, const uint16_t *roi, const float *wle,
                       SUF(roictx) *rc, REAL *x, REAL *nn_out)
{
    const int R = p->ROISize;
    SUF(load_roi)(p, roi, R, rc);
    rc->weighted = 0;
    REAL nn = C(1000.0);
    if (model == M_GS2D) {   /* bfgsMLELoc_Gauss2D :182-228 */
        REAL sl = wle ? (REAL)wle[1] : C(0.0), sr = wle ? (REAL)wle[2] : C(0.0);
        REAL su = wle ? (REAL)wle[3] : C(0.0), sd = wle ? (REAL)wle[4] : C(0.0);
        int mt = wle ? (int)wle[5] : 0;
        REAL ws = SUF(fminr)(SUF(fminr)(sl, sr), SUF(fminr)(su, sd));
        ws = SUF(fminr)(ws, ((REAL)R / C(2.0)) / C(2.35));
        SUF(preestimate)(model, rc, ws, ws, x);
        if (mt != QCS_MOLTYPE_MLEFIT) {
            ws = ws / C(1.2);
            ws = SUF(fminr)(ws, ((REAL)R / C(2.0)) / C(2.35));
        }
        if (p->WLEEn == 0 && mt != QCS_MOLTYPE_MULTIFIT) ws = (REAL)R;
        SUF(wle_weights)(rc, ws, ws, mt);
    } else if (model == M_AS3D) {   /* bfgsMLELoc_AS3D :181-228 */
        REAL sl = wle ? (REAL)wle[1] : C(0.0), sr = wle ? (REAL)wle[2] : C(0.0);
        REAL su = wle ? (REAL)wle[3] : C(0.0), sd = wle ? (REAL)wle[4] : C(0.0);
        int mt = wle ? (int)wle[5] : 0;
        REAL wx = SUF(fminr)(sl, sr), wy = SUF(fminr)(su, sd);
        SUF(preestimate)(model, rc, wx, wy, x);
        if (mt != QCS_MOLTYPE_MLEFIT) {
            REAL cap = ((REAL)R / C(1.6)) / C(2.35);
            wx = SUF(fminr)(wx, cap);
            wy = SUF(fminr)(wy, cap);
        }
        if (p->WLEEn == 0 && mt != QCS_MOLTYPE_MULTIFIT) { wx = (REAL)R; wy = (REAL)R; }
        SUF(wle_weights)(rc, wx, wy, mt);
    } else if (model == M_2E) {   /* bfgsMLELoc_Gauss2D_2E :190-227 */
        REAL ws = ((REAL)R / C(1.8)) / C(2.35);
        nn = wle ? (REAL)wle[0] : C(0.0);
        SUF(preestimate)(model, rc, ws, ws, x);
        if (p->WLEEn == 0) ws = (REAL)R;
        SUF(wle_weights)(rc, ws, ws, QCS_MOLTYPE_MULTIFIT);
    } else {   /* 3E and AS3D-2E run unweighted (WLE_ENABLE_MultiEmitterFit_* = 0) */
        SUF(preestimate)(model, rc, C(0.0), C(0.0), x);
    }
    if (nn_out) *nn_out = nn;
}

static void SUF(zero_row)(REAL *row) { int i; for (i = 0; i < QCS_REC_FLOATS; i++) row[i] = C(0.0); }

/* z = sum p_k d^k (branch on the sign of d) * ZDepthCorrFactor  (bfgsMLE_AS3D_core.h:266-279) */
static REAL SUF(zpoly)(const qcs_params *p, REAL d)
{
    REAL d2 = d * d, d3 = d2 * d, d4 = d2 * d2, z;
    if (d >= C(0.0)) {
        z = (REAL)p->p4_XGY * d4;
        z = FMA((REAL)p->p3_XGY, d3, z);
        z = FMA((REAL)p->p2_XGY, d2, z);
        z = FMA((REAL)p->p1_XGY, d, z);
        z = z + (REAL)p->p0_XGY;
    } else {
        z = (REAL)p->p4_XLY * d4;
        z = FMA((REAL)p->p3_XLY, d3, z);
        z = FMA((REAL)p->p2_XLY, d2, z);
        z = FMA((REAL)p->p1_XLY, d, z);
        z = z + (REAL)p->p0_XLY;
    }
    return z * (REAL)p->ZDepthCorrFactor;
}

static REAL SUF(hypot2)(REAL dx, REAL dy) { return SQRTR(FMA(dx, dx, dy * dy)); }

/* ROI header: Xc, Yc, frame */
static void SUF(roi_header)(const uint16_t *roi, int R, int *xo, int *yo, int *fr)
{
    const uint16_t *h = roi + R * R;
    *xo = h[0]; *yo = h[1]; *fr = (int)h[2] + 65536 * (int)h[3];
}

/*
 * One-emitter fit + validity (bfgsMLELoc_Gauss2D :170-309, bfgsMLELoc_AS3D :169-336).
 * Writes the row; returns 1 if accepted, 0 if rejected; *snr_out receives the SNR used
 * for the multi-emitter re-route, *valid_out the NaN-background test.
 */
static int SUF(select_1e)(const qcs_params *p, int model, const REAL *x, const uint16_t *roi, REAL *row,
                          REAL *snr_out, int *valid_out);

static int SUF(fit_1e)(const qcs_params *p, int model, const uint16_t *roi, const float *wle,
                       REAL *row, REAL *snr_out, int *valid_out)
{
    SUF(roictx) rc;
    REAL x[MAXP];
    SUF(mdesc) m;
    SUF(get_desc)(model, &m);
    SUF(setup)(p, model, roi, wle, &rc, x, NULL);
    SUF(bfgs_run)(&m, &rc, x, m.iters);
    return SUF(select_1e)(p, model, x, roi, row, snr_out, valid_out);
}

/* Validity test and row of a one-emitter fit vector x (bfgsMLELoc_Gauss2D :230-309,
 * bfgsMLELoc_AS3D :230-336); the ROI supplies the header only. */
static int SUF(select_1e)(const qcs_params *p, int model, const REAL *x, const uint16_t *roi, REAL *row,
                          REAL *snr_out, int *valid_out)
{
    const int R = p->ROISize, half = R / 2;
    int xo, yo, fr;
    SUF(roi_header)(roi, R, &xo, &yo, &fr);
    const REAL QE = (REAL)p->QE;
    const int multi = p->MultiEmitterFitEn;

    if (model == M_GS2D) {
        const REAL lb = (REAL)R / C(2.0) - C(1.5), ub = (REAL)R / C(2.0) + C(1.5);
        const REAL smax = multi ? ((REAL)R / C(2.0)) / C(2.35) : ((REAL)R / C(1.5)) / C(2.35);
        REAL pp = (x[0] * C(128.0)) / QE;
        REAL bg = (x[4] * C(64.0)) / QE;
        REAL fx = x[1], fy = x[2];
        REAL sg = x[3] * C(0.1);
        sg = SQRTR(C(0.5) / sg);
        if (fx != fx) fx = C(-1.0);
        if (fy != fy) fy = C(-1.0);
        if (sg != sg) sg = C(-1.0);
        REAL snr = (SQRTR(QE) * pp) / SQRTR(pp + bg);
        REAL tp = (((pp * C(2.0)) * C(3.141593)) * sg) * sg;
        int valid = (bg != bg) ? 0 : 1;
        *snr_out = snr; *valid_out = valid;
        if (valid && snr > C(4.0) && sg <= smax && sg >= C(0.6) && fx > lb && fx < ub && fy > lb && fy < ub) {
            row[0] = pp;
            row[1] = (fx - (REAL)half) + (REAL)xo;
            row[2] = (fy - (REAL)half) + (REAL)yo;
            row[3] = C(0.0);
            row[4] = sg; row[5] = sg;
            row[6] = tp; row[7] = bg; row[8] = snr;
            row[9] = C(0.0); row[10] = C(0.0);
            row[11] = (REAL)fr;
            return 1;
        }
        SUF(zero_row)(row);
        return 0;
    } else {
        const REAL lb = (REAL)R / C(2.0) - C(2.5), ub = (REAL)R / C(2.0) + C(2.5);
        const REAL smax = multi ? ((REAL)R / C(2.0)) / C(2.35) : ((REAL)R / C(1.5)) / C(2.35);
        REAL pe = x[0] * C(128.0);
        REAL be = x[5] * C(64.0);
        REAL fx = x[1], fy = x[2];
        REAL sx = SQRTR(C(0.5) / (x[3] * C(0.1)));
        REAL sy = SQRTR(C(0.5) / (x[4] * C(0.1)));
        if (fx != fx) fx = C(-1.0);
        if (fy != fy) fy = C(-1.0);
        if (sx != sx) sx = C(-1.0);
        if (sy != sy) sy = C(-1.0);
        REAL snr = pe / SQRTR(pe + be);   /* in e-, bfgsMLE_AS3D_core.h:257 */
        REAL pp = pe / QE, bg = be / QE;
        REAL tp = (((pp * C(2.0)) * C(3.141593)) * sx) * sy;
        REAL dif = (sx * sx) - (sy * sy);
        REAL z = SUF(zpoly)(p, dif);
        REAL ms = (sx + sy) / C(2.0);
        int valid = (bg != bg) ? 0 : 1;
        *snr_out = snr; *valid_out = valid;
        if (valid && snr > C(3.5) && ms > C(0.8) && ms < smax && fx > lb && fx < ub && fy > lb && fy < ub) {
            row[0] = pp;
            row[1] = (fx - (REAL)half) + (REAL)xo;
            row[2] = (fy - (REAL)half) + (REAL)yo;
            row[3] = z;
            row[4] = sx; row[5] = sy;
            row[6] = tp; row[7] = bg; row[8] = snr;
            row[9] = C(0.0); row[10] = C(0.0);
            row[11] = (REAL)fr;
            return 1;
        }
        SUF(zero_row)(row);
        return 0;
    }
}

/*
 * Multi-emitter fit + emitter selection (bfgsMLELoc_Gauss2D_2E :180-370, _3E :181-361,
 * AS3D_2E :186-419).  Writes the first emitter into `row` and, when accepted, the second into
 * `row2` (returns 2); returns 1 if only the first is written, 0 if rejected (row zeroed).
 */
static int SUF(select_me)(const qcs_params *p, int model, const REAL *x, REAL nn, const uint16_t *roi, REAL *row,
                          REAL *row2);

static int SUF(fit_me)(const qcs_params *p, int model, const uint16_t *roi, const float *wle,
                       REAL *row, REAL *row2)
{
    SUF(roictx) rc;
    REAL x[MAXP], nn;
    SUF(mdesc) m;
    SUF(get_desc)(model, &m);
    SUF(setup)(p, model, roi, wle, &rc, x, &nn);
    SUF(bfgs_run)(&m, &rc, x, m.iters);
    return SUF(select_me)(p, model, x, nn, roi, row, row2);
}

/* Emitter selection of a multi-emitter fit vector x with the nearest-neighbour distance nn the
 * model uses (WLE NearDist for the 2-emitter 2D fit, 1000 otherwise). */
static int SUF(select_me)(const qcs_params *p, int model, const REAL *x, REAL nn, const uint16_t *roi, REAL *row,
                          REAL *row2)
{
    const int R = p->ROISize, half = R / 2;
    const REAL ctr = (REAL)R / C(2.0);
    SUF(mdesc) m;
    int xo, yo, fr, k;
    SUF(get_desc)(model, &m);
    SUF(roi_header)(roi, R, &xo, &yo, &fr);
    const REAL QE = (REAL)p->QE;
    const REAL sqe = SQRTR(QE);
    const int K = m.K;
    REAL pk[3], ex[3], ey[3], sxv[3], syv[3], score[3];

    for (k = 0; k < K; k++) pk[k] = (x[m.ia[k]] * C(128.0)) / QE;
    REAL bg = (x[m.ib] * C(64.0)) / QE;
    for (k = 0; k < K; k++) {
        ex[k] = x[m.ix[k]]; ey[k] = x[m.iy[k]];
        if (m.ellip) {
            sxv[k] = SQRTR(C(0.5) / (x[m.isx[k]] * C(0.1)));
            syv[k] = SQRTR(C(0.5) / (x[m.isy[k]] * C(0.1)));
        } else {
            sxv[k] = syv[k] = SQRTR(C(0.5) / (x[m.isx[k]] * C(0.1)));
        }
    }
    int valid = (bg != bg) ? 0 : 1;
    for (k = 0; k < K; k++) {
        REAL cx = ex[k], cy = ey[k];
        if (FABSR(cx - ctr) >= (REAL)half) cx = C(-100.0);
        if (FABSR(cy - ctr) >= (REAL)half) cy = C(-100.0);
        REAL dist = SUF(hypot2)(cx - ctr, cy - ctr);
        score[k] = pk[k] * ((REAL)R - dist);
    }
    int s0 = 0, s1 = 0;
    for (k = 0; k < K; k++) if (score[k] > score[s0]) s0 = k;
    if (K == 2) s1 = 1 - s0;
    else {
        score[s0] = C(-100.0);
        for (k = 0; k < K; k++) if (score[k] > score[s1]) s1 = k;
    }
    REAL pp0 = pk[s0], pp1 = pk[s1];
    REAL x0 = ex[s0], y0 = ey[s0], x1 = ex[s1], y1 = ey[s1];
    REAL sx0 = sxv[s0], sy0 = syv[s0], sx1 = sxv[s1], sy1 = syv[s1];
    REAL tp0 = (((pp0 * C(2.0)) * C(3.141593)) * sx0) * sy0;
    REAL tp1 = (((pp1 * C(2.0)) * C(3.141593)) * sx1) * sy1;
    REAL snr0 = (sqe * pp0) / SQRTR(pp0 + bg);
    REAL snr1 = (sqe * pp1) / SQRTR(pp1 + bg);
    REAL sep = SUF(hypot2)(x0 - x1, y0 - y1);
    REAL z0 = C(0.0), z1 = C(0.0);
    int ok0, ok1;
    if (m.ellip) {   /* AS3D_2E acceptance :360, :379 */
        const REAL smax = ((REAL)R / C(1.9)) / C(2.35);
        REAL ms0 = (sx0 + sy0) / C(2.0), ms1 = (sx1 + sy1) / C(2.0);
        z0 = SUF(zpoly)(p, (sx0 * sx0) - (sy0 * sy0));
        z1 = SUF(zpoly)(p, (sx1 * sx1) - (sy1 * sy1));
        ok0 = valid && snr0 > C(5.5) && ms0 > C(0.6) && ms0 > C(0.8) && ms0 <= smax;
        ok1 = sep > C(1.5) && snr1 > C(5.2) && nn > ctr && ms1 > C(0.6) && ms1 > C(0.8) && ms1 <= smax &&
              pp1 > pp0 / C(3.0) && x1 > C(1.0) && y1 > C(1.0) && x1 < (REAL)(R - 1) && y1 < (REAL)(R - 1);
    } else {
        const REAL smax = ((REAL)R / C(2.0)) / C(2.35);
        ok0 = valid && snr0 > C(5.5) && sx0 > C(0.6) && sx0 <= smax;
        /* 2E requires NN >= centre (:313), 3E NN > centre (:322) */
        int nn_ok = (model == M_2E) ? (nn >= ctr) : (nn > ctr);
        ok1 = sep > C(1.5) && snr1 > C(5.2) && nn_ok && pp1 > pp0 / C(2.0) &&
              x1 > C(1.0) && y1 > C(1.0) && x1 < (REAL)(R - 1) && y1 < (REAL)(R - 1);
    }
    if (!ok0) { SUF(zero_row)(row); return 0; }
    row[0] = pp0; row[1] = (x0 - (REAL)half) + (REAL)xo; row[2] = (y0 - (REAL)half) + (REAL)yo;
    row[3] = z0; row[4] = sx0; row[5] = sy0; row[6] = tp0; row[7] = bg; row[8] = snr0;
    row[9] = C(0.0); row[10] = C(0.0); row[11] = (REAL)fr;
    if (!ok1) return 1;
    row2[0] = pp1; row2[1] = (x1 - (REAL)half) + (REAL)xo; row2[2] = (y1 - (REAL)half) + (REAL)yo;
    row2[3] = z1; row2[4] = sx1; row2[5] = sy1; row2[6] = tp1; row2[7] = bg; row2[8] = snr1;
    row2[9] = C(0.0); row2[10] = C(0.0); row2[11] = (REAL)fr;
    return 2;
}

/* ---------------- CRLB (LocPrecCalc_GaussianCRLB.h:64-243) ---------------- */
static void SUF(crlb_psf)(int G, REAL N, REAL B, REAL x0, REAL y0, REAL sx, REAL sy, REAL *out)
{
    const int H = G / 2;
    const REAL pi = C(3.141592654);
    const REAL peak = (C(1.0) / (((C(2.0) * pi) * sx) * sy)) * N;
    const REAL ix = C(1.0) / ((C(2.0) * sx) * sx);
    const REAL iy = C(1.0) / ((C(2.0) * sy) * sy);
    int r, c;
    for (r = 0; r < G; r++) {
        REAL cy = (REAL)(r - H);
        for (c = 0; c < G; c++) {
            REAL cx = (REAL)(c - H);
            REAL dx = cx - x0, dy = cy - y0;
            out[r * G + c] = FMA(peak, EXPR(-FMA(dx * dx, ix, (dy * dy) * iy)), B);
        }
    }
}

static REAL SUF(crlb_axis)(int G, const REAL *m0, const REAL *lm0, REAL N, REAL B, REAL x0, REAL y0,
                           REAL sx, REAL sy, int axis)
{
    REAL mp[15 * 15], mn[15 * 15];
    const REAL dh = C(0.02);
    const int n = G * G;
    int i;
    if (axis == 0) {
        SUF(crlb_psf)(G, N, B, x0 + dh, y0, sx, sy, mp);
        SUF(crlb_psf)(G, N, B, (x0 + dh) - C(2.0) * dh, y0, sx, sy, mn);
    } else {
        SUF(crlb_psf)(G, N, B, x0, y0 + dh, sx, sy, mp);
        SUF(crlb_psf)(G, N, B, x0, (y0 + dh) - C(2.0) * dh, sx, sy, mn);
    }
    REAL fisher = C(0.0);
    for (i = 0; i < n; i++) {
        REAL d2 = ((FMA(C(-2.0), m0[i], mp[i]) + mn[i]) / dh) / dh;
        REAL d2l = ((FMA(C(-2.0), lm0[i], LOGR(mp[i])) + LOGR(mn[i])) / dh) / dh;
        fisher = fisher + FMA(-m0[i], d2l, d2);
    }
    if (fisher == C(0.0)) fisher = C(1.0);
    return SQRTR(C(1.0) / fisher);
}

static void SUF(crlb_rows)(const qcs_params *p, REAL *loc, int n)
{
    const int G = (p->LocType == QCS_LOCTYPE_AS3D) ? 15 : 9;
    const REAL QE = (REAL)p->QE, RN = (REAL)p->ReadNoise_e, PS = (REAL)p->PixelSize;
    REAL m0[15 * 15], lm0[15 * 15];
    int i, j;
    for (i = 0; i < n; i++) {
        REAL *row = loc + (long)i * QCS_REC_FLOATS;
        REAL te = row[6] * QE, be = row[7] * QE, sx = row[4], sy = row[5];
        REAL stdx = C(100.0), stdy = C(100.0);
        /* the reference compares the sigmas with the double literal 0.1 */
        if (te > C(1.0) && be > C(1.0) && (double)sx > 0.1 && (double)sy > 0.1 && sx < C(10.0) && sy < C(10.0)) {
            be = FMA(be, QE, RN * RN) / QE;
            SUF(crlb_psf)(G, te, be, C(0.0), C(0.0), sx, sy, m0);
            for (j = 0; j < G * G; j++) lm0[j] = LOGR(m0[j]);
            stdx = SUF(crlb_axis)(G, m0, lm0, te, be, C(0.0), C(0.0), sx, sy, 0) * PS;
            stdy = SUF(crlb_axis)(G, m0, lm0, te, be, C(0.0), C(0.0), sx, sy, 1) * PS;
        }
        row[9] = stdx;
        row[10] = stdy;
    }
}

/* ---------------- cascade (bfgs_top.cu:26-204) ---------------- */
static int SUF(fit_all)(const qcs_params *p, const uint16_t *rois, const float *wle, int n,
                        REAL *loc, int loc_cap, orc_fit_stats *st)
{
    const int R = p->ROISize, W = QCS_ROI_WORDS(R);
    const int as3d = (p->LocType == QCS_LOCTYPE_AS3D);
    int i, rows = n;
    orc_fit_stats s = {0, 0, 0, 0, 0, 0};
    s.n_in = n;
    if (n <= 0) { if (st) *st = s; return 0; }
    if (!p->MultiEmitterFitEn) {
        REAL snr; int valid;
        for (i = 0; i < n; i++)
            SUF(fit_1e)(p, as3d ? M_AS3D : M_GS2D, rois + (long)i * W, wle ? wle + (long)i * QCS_WLE_FLOATS : NULL,
                        loc + (long)i * QCS_REC_FLOATS, &snr, &valid);
        s.n_single = n;
    } else {
        int *l2 = (int *)malloc(sizeof(int) * (size_t)(2 * n + 1));
        int *l3 = (int *)malloc(sizeof(int) * (size_t)(n + 1));
        int n2 = 0, n3 = 0;
        /* gpuMoleculeFitClasify (bfgs_top_MoleculeClasify.cu:57-71): WLE type <= WLEFit -> 1E */
        for (i = 0; i < n; i++) {
            int mt = (int)wle[(long)i * QCS_WLE_FLOATS + QCS_WLE_FLUOTYPE];
            if (mt > QCS_MOLTYPE_WLEFIT) l2[n2++] = i;
        }
        int n2_classified = n2;
        for (i = 0; i < n; i++) {
            int mt = (int)wle[(long)i * QCS_WLE_FLOATS + QCS_WLE_FLUOTYPE];
            if (mt <= QCS_MOLTYPE_WLEFIT) {
                REAL snr; int valid;
                s.n_single++;
                int ok = SUF(fit_1e)(p, as3d ? M_AS3D : M_GS2D, rois + (long)i * W, wle + (long)i * QCS_WLE_FLOATS,
                                     loc + (long)i * QCS_REC_FLOATS, &snr, &valid);
                /* rejected 1E fits with SNR > 6 are re-fitted with two emitters (GS2D_core.h:279-297) */
                if (!ok && valid && snr > C(6.0)) l2[n2++] = i;
            }
        }
        (void)n2_classified;
        s.n_2e = n2;
        /* 2-emitter stage; accepted second emitters append at rows >= n in list order */
        for (int j = 0; j < n2; j++) {
            i = l2[j];
            REAL row2[QCS_REC_FLOATS];
            int r = SUF(fit_me)(p, as3d ? M_AS3D2E : M_2E, rois + (long)i * W, wle + (long)i * QCS_WLE_FLOATS,
                                loc + (long)i * QCS_REC_FLOATS, row2);
            if (r == 0) {
                if (as3d) s.n_rejected++;
                else l3[n3++] = i;
            } else if (r == 2 && rows < loc_cap) {
                memcpy(loc + (long)rows * QCS_REC_FLOATS, row2, sizeof(row2));
                rows++;
            }
        }
        s.n_3e = n3;
        if (!as3d) {
            for (int j = 0; j < n3; j++) {
                i = l3[j];
                REAL row2[QCS_REC_FLOATS];
                int r = SUF(fit_me)(p, M_3E, rois + (long)i * W, wle + (long)i * QCS_WLE_FLOATS,
                                    loc + (long)i * QCS_REC_FLOATS, row2);
                if (r == 0) s.n_rejected++;
                else if (r == 2 && rows < loc_cap) {
                    memcpy(loc + (long)rows * QCS_REC_FLOATS, row2, sizeof(row2));
                    rows++;
                }
            }
        }
        free(l2);
        free(l3);
    }
    SUF(crlb_rows)(p, loc, rows);
    s.n_rows = rows;
    if (st) *st = s;
    return rows;
}

static int SUF(fit_params)(const qcs_params *p, int model, const uint16_t *rois, const float *wle, int n,
                           REAL *theta, int iters)
{
    const int R = p->ROISize, W = QCS_ROI_WORDS(R);
    SUF(mdesc) m;
    SUF(get_desc)(model, &m);
    for (int i = 0; i < n; i++) {
        SUF(roictx) rc;
        REAL x[MAXP];
        SUF(setup)(p, model, rois + (long)i * W, wle ? wle + (long)i * QCS_WLE_FLOATS : NULL, &rc, x, NULL);
        SUF(bfgs_run)(&m, &rc, x, iters > 0 ? iters : m.iters);
        for (int k = 0; k < m.P; k++) theta[(long)i * m.P + k] = x[k];
    }
    return m.P;
}
