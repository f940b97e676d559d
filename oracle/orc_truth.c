/*
 * orc_truth.c -- FP64 maximum-likelihood tools (TEST INFRASTRUCTURE).
 *
 * Not a restatement of the reference: an independent, converged Poisson MLE of the sampled
 * Gaussian PSF model with analytic derivatives, and the full-Fisher Cramer-Rao bound.  The tests
 * use them as the ground-truth yardstick (RMSE vs CRLB) for the reference-schedule fitters.
 *
 * Model, pixel (r, c) centre at (c + 0.5, r + 0.5):
 *   mu = sum_k N_k / (2 pi sx_k sy_k) exp(-(dx^2 / (2 sx_k^2) + dy^2 / (2 sy_k^2))) + b
 * theta: for each emitter [N, x, y] (+ [sx, sy] if astig), then the shared sigma (2D), then b.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define TMAXP 16

static int npar(int K, int astig) { return astig ? 5 * K + 1 : 3 * K + 2; }

/* mu and d mu / d theta at every pixel */
static void model_eval(int R, int K, int astig, const double *th, double *mu, double *jac)
{
    const int P = npar(K, astig);
    for (int r = 0; r < R; r++)
        for (int c = 0; c < R; c++) {
            const int i = r * R + c;
            double *J = jac + (long)i * P;
            memset(J, 0, sizeof(double) * P);
            double m = th[P - 1];
            J[P - 1] = 1.0;
            for (int k = 0; k < K; k++) {
                const int b = astig ? 5 * k : 3 * k;
                double N = th[b], x = th[b + 1], y = th[b + 2];
                double sx = astig ? th[b + 3] : th[3 * K];
                double sy = astig ? th[b + 4] : th[3 * K];
                double dx = (c + 0.5) - x, dy = (r + 0.5) - y;
                double E = exp(-(dx * dx / (2 * sx * sx) + dy * dy / (2 * sy * sy)));
                double g = N / (2 * M_PI * sx * sy) * E;
                m += g;
                J[b] += g / N;
                J[b + 1] += g * dx / (sx * sx);
                J[b + 2] += g * dy / (sy * sy);
                if (astig) {
                    J[b + 3] += g * (dx * dx / (sx * sx * sx) - 1.0 / sx);
                    J[b + 4] += g * (dy * dy / (sy * sy * sy) - 1.0 / sy);
                } else {
                    J[3 * K] += g * ((dx * dx + dy * dy) / (sx * sx * sx) - 2.0 / sx);
                }
            }
            mu[i] = m;
        }
}

static double loglik(const double *img, const double *mu, int n)
{
    double s = 0;
    for (int i = 0; i < n; i++) s += (img[i] > 0 ? img[i] * log(mu[i]) : 0.0) - mu[i];
    return s;
}

/* solve A x = b (P x P, symmetric positive definite) by Gaussian elimination with pivoting */
static int solve(int P, double *A, double *b)
{
    for (int c = 0; c < P; c++) {
        int piv = c;
        for (int r = c + 1; r < P; r++) if (fabs(A[r * P + c]) > fabs(A[piv * P + c])) piv = r;
        if (fabs(A[piv * P + c]) < 1e-300) return -1;
        if (piv != c) {
            for (int k = 0; k < P; k++) { double t = A[c * P + k]; A[c * P + k] = A[piv * P + k]; A[piv * P + k] = t; }
            double t = b[c]; b[c] = b[piv]; b[piv] = t;
        }
        for (int r = c + 1; r < P; r++) {
            double f = A[r * P + c] / A[c * P + c];
            for (int k = c; k < P; k++) A[r * P + k] -= f * A[c * P + k];
            b[r] -= f * b[c];
        }
    }
    for (int r = P - 1; r >= 0; r--) {
        double s = b[r];
        for (int k = r + 1; k < P; k++) s -= A[r * P + k] * b[k];
        b[r] = s / A[r * P + r];
    }
    return 0;
}

static int feasible(int K, int astig, const double *th)
{
    const int P = npar(K, astig);
    for (int k = 0; k < K; k++) {
        const int b = astig ? 5 * k : 3 * k;
        if (!(th[b] > 0)) return 0;
        if (astig && (!(th[b + 3] > 0.2) || !(th[b + 4] > 0.2))) return 0;
    }
    if (!astig && !(th[3 * K] > 0.2)) return 0;
    if (!(th[P - 1] > 0)) return 0;
    return 1;
}

int orc_mle_f64(const double *img, int R, int K, int astig, double *theta, int max_iter, double *ll_out)
{
    const int P = npar(K, astig), n = R * R;
    if (P > TMAXP || R <= 0 || R > 32) return -1;
    double *mu = (double *)malloc(sizeof(double) * n), *jac = (double *)malloc(sizeof(double) * n * P);
    double *mu2 = (double *)malloc(sizeof(double) * n), *jac2 = (double *)malloc(sizeof(double) * n * P);
    double th2[TMAXP], I[TMAXP * TMAXP], A[TMAXP * TMAXP], g[TMAXP], d[TMAXP];
    double lambda = 1e-3;
    model_eval(R, K, astig, theta, mu, jac);
    double ll = loglik(img, mu, n);
    int it, converged = 0;
    for (it = 0; it < max_iter && !converged; it++) {
        memset(I, 0, sizeof(I));
        memset(g, 0, sizeof(g));
        for (int i = 0; i < n; i++) {
            const double *J = jac + (long)i * P;
            double w = img[i] / mu[i] - 1.0;
            for (int a = 0; a < P; a++) {
                g[a] += w * J[a];
                for (int b = 0; b < P; b++) I[a * P + b] += J[a] * J[b] / mu[i];
            }
        }
        int accepted = 0;
        for (int tries = 0; tries < 30 && !accepted; tries++) {
            memcpy(A, I, sizeof(double) * P * P);
            for (int a = 0; a < P; a++) A[a * P + a] *= (1.0 + lambda);
            memcpy(d, g, sizeof(double) * P);
            if (solve(P, A, d) != 0) { lambda *= 10; continue; }
            for (int a = 0; a < P; a++) th2[a] = theta[a] + d[a];
            if (!feasible(K, astig, th2)) { lambda *= 10; continue; }
            model_eval(R, K, astig, th2, mu2, jac2);
            double ll2 = loglik(img, mu2, n);
            if (ll2 >= ll) {
                double step = 0;
                for (int a = 0; a < P; a++) step = fmax(step, fabs(d[a]) / fmax(1e-3, fabs(theta[a])));
                memcpy(theta, th2, sizeof(double) * P);
                double *t = mu; mu = mu2; mu2 = t;
                t = jac; jac = jac2; jac2 = t;
                double dll = ll2 - ll;
                ll = ll2;
                lambda = fmax(lambda * 0.3, 1e-9);
                accepted = 1;
                if (step < 1e-10 || dll < 1e-12) converged = 1;
            } else {
                lambda *= 10;
            }
        }
        if (!accepted) converged = 1;   /* no ascent direction left: at the optimum to precision */
    }
    if (ll_out) *ll_out = ll;
    free(mu); free(jac); free(mu2); free(jac2);
    return it;
}

int orc_fisher_crlb_f64(int R, int K, int astig, const double *theta, double *crlb)
{
    const int P = npar(K, astig), n = R * R;
    if (P > TMAXP) return -1;
    double *mu = (double *)malloc(sizeof(double) * n), *jac = (double *)malloc(sizeof(double) * n * P);
    double I[TMAXP * TMAXP];
    model_eval(R, K, astig, theta, mu, jac);
    memset(I, 0, sizeof(I));
    for (int i = 0; i < n; i++) {
        const double *J = jac + (long)i * P;
        for (int a = 0; a < P; a++)
            for (int b = 0; b < P; b++) I[a * P + b] += J[a] * J[b] / mu[i];
    }
    int rc = 0;
    for (int a = 0; a < P; a++) {
        double A[TMAXP * TMAXP], e[TMAXP];
        memcpy(A, I, sizeof(double) * P * P);
        for (int b = 0; b < P; b++) e[b] = (a == b) ? 1.0 : 0.0;
        if (solve(P, A, e) != 0) { rc = -1; crlb[a] = NAN; continue; }
        crlb[a] = sqrt(e[a]);
    }
    free(mu); free(jac);
    return rc;
}
