// cfit_dev.h -- device building blocks of the converged multi-emitter fitter (gfx950), shared by
// the one-pass kernels (cfit.hip) and the persistent work-queue kernels (cfq.hip).
//
// Models, pixel (r, c) centred at (c + 0.5, r + 0.5), K = 1..3 emitters:
//
//   2D (KK = K):          mu = sum_k N_k / (2 pi s^2) exp(-((c + .5 - x_k)^2 + (r + .5 - y_k)^2) / (2 s^2)) + b
//                         theta = [N_k, x_k, y_k]_k, s, b              (3K + 2 parameters)
//   elliptical (KK = K + 10, astigmatic 3D):
//                         mu = sum_k N_k / (2 pi sx_k sy_k) exp(-(dx^2 / (2 sx_k^2) + dy^2 / (2 sy_k^2))) + b
//                         theta = [N_k, x_k, y_k, sx_k, sy_k]_k, b     (5K + 1 parameters, 16 at K = 3)
//
// The template parameter KK selects model and emitter count.  Per-lane arrays in LDS are
// lane-interleaved (element i of lane l at [i * 64 + l]).
#pragma once
#include <math.h>
#include <type_traits>
#include "qcs_internal.h"

namespace qcs {

struct CFitConst {
    float offset, kadc, inv_qe;
    int max_iter;
    float tol;          // convergence: max |step| in scaled units
    int init_mode;      // CF_INIT_*
    float sigma0;       // heuristic initial PSF width (px)
    float ftol;         // convergence: an accepted step lowered the loss by less than this (0: off)
    int active_set;     // leave bound-pinned parameters out of the step
    int pin_stop;       // elliptical three-emitter fits: end after CF_PIN_RUN iterations at a bound
};
enum { CF_INIT_HEUR = 0, CF_INIT_GIVEN = 1, CF_INIT_ADD = 2 };
enum { CF_CONVERGED = 1, CF_MAXITER = 2, CF_STALLED = 4, CF_FLAT = 8, CF_PINNED = 16 };
constexpr int CF_PIN_RUN = 10;   // accepted iterations in a row at a box bound that end a fit (pin_stop)
constexpr int CF_ELLIPTIC = 10;   // KK = K + CF_ELLIPTIC selects the elliptical model

// model traits: parameter count and indices
template <int KK> struct CM {
    static constexpr bool AS = KK > CF_ELLIPTIC;
    static constexpr int K = AS ? KK - CF_ELLIPTIC : KK;
    static constexpr int E = AS ? 5 : 3;                  // parameters per emitter
    static constexpr int P = AS ? 5 * K + 1 : 3 * K + 2;
    static constexpr int NS = P * (P + 1) / 2;
    static constexpr int IB = P - 1;                      // background
    __host__ __device__ static constexpr int iN(int k) { return E * k; }
    __host__ __device__ static constexpr int ix(int k) { return E * k + 1; }
    __host__ __device__ static constexpr int iy(int k) { return E * k + 2; }
    __host__ __device__ static constexpr int isx(int k) { return AS ? E * k + 3 : 3 * K; }
    __host__ __device__ static constexpr int isy(int k) { return AS ? E * k + 4 : 3 * K; }
};
template <int KK> struct CP { static constexpr int P = CM<KK>::P, NS = CM<KK>::NS; };

// N -> kilo-photons, b -> hundreds of photons, positions and widths in pixels
template <int KK>
__device__ __forceinline__ float cf_scale(int i)
{
    constexpr int P = CM<KK>::P;
    if (i == P - 1) return 100.0f;
    if constexpr (CM<KK>::AS) return (i % 5 == 0) ? 1000.0f : 1.0f;
    else return (i < P - 2 && i % 3 == 0) ? 1000.0f : 1.0f;
}

// packed upper-triangular index of (i, j), i <= j
__device__ __forceinline__ constexpr int tri(int i, int j, int P) { return i * P - (i * (i - 1)) / 2 + (j - i); }

// f(integral_constant<int, i>) for i = B .. E - 1: a loop whose index is a compile-time constant
// in every iteration, whatever the unroller decides (triangular nests included)
template <int B, int E, class F>
__device__ __forceinline__ void cf_sfor(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        cf_sfor<B + 1, E>(f);
    }
}

// Where a lane keeps its observed photons q and its packed inverse Hessian D.  Every D access
// uses a compile-time index (unrolled loops), so the entries a DSplit keeps in registers stay
// there.
struct QLds {   // q of pixel i at p[i * 64] (lane-interleaved floats)
    const float *p;
    __device__ __forceinline__ float operator()(int i) const { return p[i * 64]; }
};
struct QLds16 {   // raw counts at p[i * 64], converted on use exactly as cf_load_q converts them
    const uint16_t *p;
    float offset, kadc, inv_qe;
    __device__ __forceinline__ float operator()(int i) const { return fmaxf(((float)p[i * 64] - offset) * kadc, 1.0f) * inv_qe; }
};
struct DLds {   // D[i] at p[i * 64]
    float *p;
    __device__ __forceinline__ float &operator[](int i) const { return p[i * 64]; }
};
template <int NREG>
struct DSplit {   // D[i] in registers for i < NREG, else at p[(i - NREG) * 64]
    float r[NREG > 0 ? NREG : 1];
    float *p;
    __device__ __forceinline__ float &operator[](int i) { return (NREG > 0 && i < NREG) ? r[i] : p[(i - NREG) * 64]; }
};

// loss sum((mu - q) - q ln(mu / q)) (the Poisson negative log-likelihood up to a constant) and,
// if G, its gradient with respect to theta (physical units)
template <int KK, int R, bool G, class QA>
__device__ __forceinline__ float cf_lossgrad_q(const QA &qa, const float (&th)[CP<KK>::P], float (&g)[CP<KK>::P])
{
    using M = CM<KK>;
    constexpr int K = M::K;
    if constexpr (!M::AS) {
        const float s = th[3 * K], b = th[3 * K + 1];
        const float is2 = 1.0f / (s * s);
        const float h = -0.5f * is2;
        const float nrm = 0.15915494309189535f * is2;   // 1 / (2 pi s^2)
        float ex[K][R], dx[K][R];
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int c = 0; c < R; c++) {
                dx[k][c] = ((float)c + 0.5f) - th[3 * k + 1];
                ex[k][c] = __expf(h * dx[k][c] * dx[k][c]);
            }
        float L = 0.0f;
        float gN[K], gx[K], gy[K], gd2 = 0.0f, gw = 0.0f, gb = 0.0f;
#pragma unroll
        for (int k = 0; k < K; k++) gN[k] = gx[k] = gy[k] = 0.0f;
        // rows stay a loop: a full unroll lets the scheduler hoist every row's LDS loads and
        // exponentials, which costs registers (occupancy) for no gain.  The gradient terms that
        // depend on the row only (dy, dy^2) multiply the row's sums of w g_k once per row.
#pragma unroll 1
        for (int r = 0; r < R; r++) {
            float ay[K], dy[K], sw[K], sx2[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                dy[k] = ((float)r + 0.5f) - th[3 * k + 2];
                ay[k] = th[3 * k] * nrm * __expf(h * dy[k] * dy[k]);
                sw[k] = sx2[k] = 0.0f;
            }
#pragma unroll
            for (int c = 0; c < R; c++) {
                float gk[K];
                float mu = b;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    gk[k] = ay[k] * ex[k][c];
                    mu += gk[k];
                }
                mu = fmaxf(mu, 1e-6f);
                const float q = qa(r * R + c);
                // deviance form: terms are O(1), so the FP32 sum resolves loss changes of ~1e-4
                L += (mu - q) - q * __logf(mu * __builtin_amdgcn_rcpf(q));
                if (G) {
                    const float w = 1.0f - q * __builtin_amdgcn_rcpf(mu);
                    gb += w;
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const float wg = w * gk[k];
                        sw[k] += wg;
                        gx[k] = fmaf(wg, dx[k][c], gx[k]);
                        sx2[k] = fmaf(wg, dx[k][c] * dx[k][c], sx2[k]);
                    }
                }
            }
            if (G) {
#pragma unroll
                for (int k = 0; k < K; k++) {
                    gN[k] += sw[k];
                    gw += sw[k];
                    gy[k] = fmaf(sw[k], dy[k], gy[k]);
                    gd2 = fmaf(sw[k], dy[k] * dy[k], gd2 + sx2[k]);
                }
            }
        }
        if (G) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                g[3 * k] = gN[k] / fmaxf(th[3 * k], 1e-3f);
                g[3 * k + 1] = gx[k] * is2;
                g[3 * k + 2] = gy[k] * is2;
            }
            g[3 * K] = (gd2 * is2 - 2.0f * gw) / s;
            g[3 * K + 1] = gb;
        }
        return L;
    } else {
        const float b = th[M::IB];
        float ex[K][R], dx[K][R], isx2[K], isy2[K], hy[K], nk[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const float sx = th[M::isx(k)], sy = th[M::isy(k)];
            isx2[k] = 1.0f / (sx * sx);
            isy2[k] = 1.0f / (sy * sy);
            hy[k] = -0.5f * isy2[k];
            nk[k] = th[M::iN(k)] * (0.15915494309189535f / (sx * sy));
            const float hx = -0.5f * isx2[k];
#pragma unroll
            for (int c = 0; c < R; c++) {
                dx[k][c] = ((float)c + 0.5f) - th[M::ix(k)];
                ex[k][c] = __expf(hx * dx[k][c] * dx[k][c]);
            }
        }
        float L = 0.0f;
        float gN[K], gx[K], gy[K], gxx[K], gyy[K], gb = 0.0f;
#pragma unroll
        for (int k = 0; k < K; k++) gN[k] = gx[k] = gy[k] = gxx[k] = gyy[k] = 0.0f;
#pragma unroll 1
        for (int r = 0; r < R; r++) {
            float ay[K], dy[K], sw[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                dy[k] = ((float)r + 0.5f) - th[M::iy(k)];
                ay[k] = nk[k] * __expf(hy[k] * dy[k] * dy[k]);
                sw[k] = 0.0f;
            }
#pragma unroll
            for (int c = 0; c < R; c++) {
                float gk[K];
                float mu = b;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    gk[k] = ay[k] * ex[k][c];
                    mu += gk[k];
                }
                mu = fmaxf(mu, 1e-6f);
                const float q = qa(r * R + c);
                L += (mu - q) - q * __logf(mu * __builtin_amdgcn_rcpf(q));
                if (G) {
                    const float w = 1.0f - q * __builtin_amdgcn_rcpf(mu);
                    gb += w;
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const float wg = w * gk[k];
                        sw[k] += wg;
                        gx[k] = fmaf(wg, dx[k][c], gx[k]);
                        gxx[k] = fmaf(wg, dx[k][c] * dx[k][c], gxx[k]);
                    }
                }
            }
            if (G) {   // the row's terms in dy, dy^2 from the row sums of w g_k
#pragma unroll
                for (int k = 0; k < K; k++) {
                    gN[k] += sw[k];
                    gy[k] = fmaf(sw[k], dy[k], gy[k]);
                    gyy[k] = fmaf(sw[k], dy[k] * dy[k], gyy[k]);
                }
            }
        }
        if (G) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                g[M::iN(k)] = gN[k] / fmaxf(th[M::iN(k)], 1e-3f);
                g[M::ix(k)] = gx[k] * isx2[k];
                g[M::iy(k)] = gy[k] * isy2[k];
                g[M::isx(k)] = fmaf(gxx[k], isx2[k], -gN[k]) / th[M::isx(k)];
                g[M::isy(k)] = fmaf(gyy[k], isy2[k], -gN[k]) / th[M::isy(k)];
            }
            g[M::IB] = gb;
        }
        return L;
    }
}

template <int KK, int R, bool G>
__device__ __forceinline__ float cf_lossgrad(const float *qs, const float (&th)[CP<KK>::P], float (&g)[CP<KK>::P])
{
    return cf_lossgrad_q<KK, R, G>(QLds{qs}, th, g);
}

template <int KK, int R>
__device__ __forceinline__ void cf_project(float (&th)[CP<KK>::P])
{
    using M = CM<KK>;
#pragma unroll
    for (int k = 0; k < M::K; k++) {
        th[M::iN(k)] = fmaxf(th[M::iN(k)], 1.0f);
        th[M::ix(k)] = fminf(fmaxf(th[M::ix(k)], -0.5f), (float)R + 0.5f);
        th[M::iy(k)] = fminf(fmaxf(th[M::iy(k)], -0.5f), (float)R + 0.5f);
        if constexpr (M::AS) {
            th[M::isx(k)] = fminf(fmaxf(th[M::isx(k)], 0.4f), 0.5f * (float)R);
            th[M::isy(k)] = fminf(fmaxf(th[M::isy(k)], 0.4f), 0.5f * (float)R);
        }
    }
    if constexpr (!M::AS) th[3 * M::K] = fminf(fmaxf(th[3 * M::K], 0.4f), 0.5f * (float)R);
    th[M::IB] = fmaxf(th[M::IB], 0.01f);
}

// d mu / d theta (scaled units) of one pixel; returns mu
template <int KK>
__device__ __forceinline__ float cf_jacobian(const float (&th)[CP<KK>::P], float cpx, float rpx, float (&J)[CP<KK>::P])
{
    using M = CM<KK>;
    constexpr int K = M::K, P = M::P;
    float mu = th[M::IB];
    if constexpr (!M::AS) {
        const float s = th[3 * K], is2 = 1.0f / (s * s), h = -0.5f * is2, nrm = 0.15915494309189535f * is2;
        float jw = 0.0f;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const float dx = cpx - th[3 * k + 1], dy = rpx - th[3 * k + 2];
            const float d2 = fmaf(dx, dx, dy * dy);
            const float mk = th[3 * k] * nrm * __expf(h * d2);
            mu += mk;
            J[3 * k] = mk / fmaxf(th[3 * k], 1e-3f) * cf_scale<KK>(3 * k);
            J[3 * k + 1] = mk * dx * is2;
            J[3 * k + 2] = mk * dy * is2;
            jw = fmaf(mk, fmaf(d2, is2, -2.0f), jw);
        }
        J[3 * K] = jw / s;
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const float sx = th[M::isx(k)], sy = th[M::isy(k)];
            const float isx2 = 1.0f / (sx * sx), isy2 = 1.0f / (sy * sy);
            const float dx = cpx - th[M::ix(k)], dy = rpx - th[M::iy(k)];
            const float mk = th[M::iN(k)] * (0.15915494309189535f / (sx * sy)) *
                             __expf(fmaf(-0.5f * isx2, dx * dx, -0.5f * isy2 * (dy * dy)));
            mu += mk;
            J[M::iN(k)] = mk / fmaxf(th[M::iN(k)], 1e-3f) * cf_scale<KK>(M::iN(k));
            J[M::ix(k)] = mk * dx * isx2;
            J[M::iy(k)] = mk * dy * isy2;
            J[M::isx(k)] = mk * fmaf(dx * dx, isx2, -1.0f) / sx;
            J[M::isy(k)] = mk * fmaf(dy * dy, isy2, -1.0f) / sy;
        }
    }
    J[P - 1] = cf_scale<KK>(P - 1);
    return mu;
}

// Inverse of the Fisher information sum grad(mu) grad(mu)^T / mu at theta, in scaled units,
// into the packed symmetric D (the BFGS starting inverse Hessian).  Returns false when the
// Cholesky factorisation fails (D then holds no usable matrix).  Up to 66 packed entries (the
// 2D models) the factorisation runs in registers, larger ones (the elliptical models with two or
// three emitters) in the lane's LDS column D.
template <int KK, int R>
__device__ __forceinline__ bool cf_fisher_inverse(const float (&th)[CP<KK>::P], float *D)
{
    constexpr int P = CP<KK>::P, NS = CP<KK>::NS;
    constexpr bool IN_LDS = NS > 66;
    float Freg[IN_LDS ? 1 : NS];
    auto F = [&](int i) -> float & {
        if constexpr (IN_LDS) return D[i * 64];
        else return Freg[i];
    };
#pragma unroll
    for (int i = 0; i < NS; i++) F(i) = 0.0f;
#pragma unroll 1
    for (int r = 0; r < R; r++) {
#pragma unroll
        for (int c = 0; c < R; c++) {
            float J[P];
            const float mu = cf_jacobian<KK>(th, (float)c + 0.5f, (float)r + 0.5f, J);
            const float im = __builtin_amdgcn_rcpf(fmaxf(mu, 1e-6f));
#pragma unroll
            for (int a = 0; a < P; a++) {
                const float ja = J[a] * im;
#pragma unroll
                for (int b2 = a; b2 < P; b2++) F(tri(a, b2, P)) = fmaf(ja, J[b2], F(tri(a, b2, P)));
            }
        }
    }
    // F = U^T U (U upper, in place)
#pragma unroll
    for (int j = 0; j < P; j++) {
        float d = F(tri(j, j, P));
#pragma unroll
        for (int k = 0; k < j; k++) d = fmaf(-F(tri(k, j, P)), F(tri(k, j, P)), d);
        if (!(d > 0.0f)) return false;
        const float u = sqrtf(d), iu = 1.0f / u;
        F(tri(j, j, P)) = u;
#pragma unroll
        for (int i = j + 1; i < P; i++) {
            float v = F(tri(j, i, P));
#pragma unroll
            for (int k = 0; k < j; k++) v = fmaf(-F(tri(k, j, P)), F(tri(k, i, P)), v);
            F(tri(j, i, P)) = v * iu;
        }
    }
    // U^-1 in place, rows bottom-up, each row right to left
#pragma unroll
    for (int i = P - 1; i >= 0; i--) {
        const float iu = 1.0f / F(tri(i, i, P));
#pragma unroll
        for (int j = P - 1; j > i; j--) {
            float v = 0.0f;
#pragma unroll
            for (int k = i + 1; k <= j; k++) v = fmaf(F(tri(i, k, P)), F(tri(k, j, P)), v);
            F(tri(i, j, P)) = -v * iu;
        }
        F(tri(i, i, P)) = iu;
    }
    // F^-1 = U^-1 U^-T, in place row by row (entries (i, j), j >= i, use row i and row j from k >= j)
#pragma unroll
    for (int i = 0; i < P; i++)
#pragma unroll
        for (int j = i; j < P; j++) {
            float v = 0.0f;
#pragma unroll
            for (int k = j; k < P; k++) v = fmaf(F(tri(i, k, P)), F(tri(j, k, P)), v);
            F(tri(i, j, P)) = v;
        }
    if constexpr (!IN_LDS) {
#pragma unroll
        for (int i = 0; i < NS; i++) D[i * 64] = Freg[i];
    }
    return true;
}

// The same inverse with the packed matrix in registers for any model size (the persistent
// queue's prep pass, whose output goes to HBM): every loop is a compile-time loop, and the
// arithmetic and its order are those of cf_fisher_inverse, so the result is bit-identical.
template <int KK, int R>
__device__ __forceinline__ bool cf_fisher_inverse_regs(const float (&th)[CP<KK>::P], float (&F)[CP<KK>::NS])
{
    constexpr int P = CP<KK>::P, NS = CP<KK>::NS;
    cf_sfor<0, NS>([&](auto IC) { F[decltype(IC)::value] = 0.0f; });
#pragma unroll 1
    for (int r = 0; r < R; r++) {
#pragma unroll
        for (int c = 0; c < R; c++) {
            float J[P];
            const float mu = cf_jacobian<KK>(th, (float)c + 0.5f, (float)r + 0.5f, J);
            const float im = __builtin_amdgcn_rcpf(fmaxf(mu, 1e-6f));
            cf_sfor<0, P>([&](auto AC) {
                constexpr int a = decltype(AC)::value;
                const float ja = J[a] * im;
                cf_sfor<a, P>([&](auto BC) {
                    constexpr int b2 = decltype(BC)::value;
                    F[tri(a, b2, P)] = fmaf(ja, J[b2], F[tri(a, b2, P)]);
                });
            });
        }
    }
    // F = U^T U (U upper, in place); a non-positive pivot ends it
    bool ok = true;
    cf_sfor<0, P>([&](auto JC) {
        constexpr int j = decltype(JC)::value;
        if (!ok) return;
        float d = F[tri(j, j, P)];
        cf_sfor<0, j>([&](auto KC) {
            constexpr int k = decltype(KC)::value;
            d = fmaf(-F[tri(k, j, P)], F[tri(k, j, P)], d);
        });
        if (!(d > 0.0f)) {
            ok = false;
            return;
        }
        const float u = sqrtf(d), iu = 1.0f / u;
        F[tri(j, j, P)] = u;
        cf_sfor<j + 1, P>([&](auto IC) {
            constexpr int i = decltype(IC)::value;
            float v = F[tri(j, i, P)];
            cf_sfor<0, j>([&](auto KC) {
                constexpr int k = decltype(KC)::value;
                v = fmaf(-F[tri(k, j, P)], F[tri(k, i, P)], v);
            });
            F[tri(j, i, P)] = v * iu;
        });
    });
    if (!ok) return false;
    // U^-1 in place, rows bottom-up, each row right to left
    cf_sfor<0, P>([&](auto IC) {
        constexpr int i = P - 1 - decltype(IC)::value;
        const float iu = 1.0f / F[tri(i, i, P)];
        cf_sfor<0, P - 1 - i>([&](auto JC) {
            constexpr int j = P - 1 - decltype(JC)::value;
            float v = 0.0f;
            cf_sfor<i + 1, j + 1>([&](auto KC) {
                constexpr int k = decltype(KC)::value;
                v = fmaf(F[tri(i, k, P)], F[tri(k, j, P)], v);
            });
            F[tri(i, j, P)] = -v * iu;
        });
        F[tri(i, i, P)] = iu;
    });
    // F^-1 = U^-1 U^-T
    cf_sfor<0, P>([&](auto IC) {
        constexpr int i = decltype(IC)::value;
        cf_sfor<i, P>([&](auto JC) {
            constexpr int j = decltype(JC)::value;
            float v = 0.0f;
            cf_sfor<j, P>([&](auto KC) {
                constexpr int k = decltype(KC)::value;
                v = fmaf(F[tri(i, k, P)], F[tri(j, k, P)], v);
            });
            F[tri(i, j, P)] = v;
        });
    });
    return true;
}

// mu of one pixel in double (the FP64 likelihood and CRLB)
template <int KK>
__device__ __forceinline__ double cf_mu64(const float (&th)[CP<KK>::P], double cpx, double rpx, double *J)
{
    using M = CM<KK>;
    constexpr int K = M::K, P = M::P;
    double mu = th[M::IB];
    if constexpr (!M::AS) {
        const double s = th[3 * K], is2 = 1.0 / (s * s), nrm = 1.0 / (6.283185307179586 * s * s);
        double jw = 0.0;
        for (int k = 0; k < K; k++) {
            const double dx = cpx - (double)th[3 * k + 1], dy = rpx - (double)th[3 * k + 2];
            const double d2 = dx * dx + dy * dy;
            const double mk = (double)th[3 * k] * nrm * exp(-0.5 * d2 * is2);
            mu += mk;
            if (J) {
                J[3 * k] = mk / (double)th[3 * k];
                J[3 * k + 1] = mk * dx * is2;
                J[3 * k + 2] = mk * dy * is2;
                jw += mk * (d2 * is2 - 2.0) / s;
            }
        }
        if (J) J[3 * K] = jw;
    } else {
        for (int k = 0; k < K; k++) {
            const double sx = th[M::isx(k)], sy = th[M::isy(k)];
            const double isx2 = 1.0 / (sx * sx), isy2 = 1.0 / (sy * sy);
            const double dx = cpx - (double)th[M::ix(k)], dy = rpx - (double)th[M::iy(k)];
            const double mk = (double)th[M::iN(k)] / (6.283185307179586 * sx * sy) * exp(-0.5 * (dx * dx * isx2 + dy * dy * isy2));
            mu += mk;
            if (J) {
                J[M::iN(k)] = mk / (double)th[M::iN(k)];
                J[M::ix(k)] = mk * dx * isx2;
                J[M::iy(k)] = mk * dy * isy2;
                J[M::isx(k)] = mk * (dx * dx * isx2 - 1.0) / sx;
                J[M::isy(k)] = mk * (dy * dy * isy2 - 1.0) / sy;
            }
        }
    }
    if (J) J[P - 1] = 1.0;
    return mu;
}

// FP64 log-likelihood sum(q ln mu - mu) (without the ln q! term, common to every model).  Each
// emitter's Gaussian is separable, A_k gx_k(c) gy_k(r), so an emitter takes 2 R exponentials per
// ROI instead of R^2 (the column factors in registers, the row factor once per row).
template <int KK, int R>
__device__ double cf_loglik64(const float *qs, const float (&th)[CP<KK>::P])
{
    using M = CM<KK>;
    constexpr int K = M::K;
    double ay[K], amp[K], yk[K], gx[K][R];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double sx = th[M::isx(k)], sy = th[M::isy(k)], ax = -0.5 / (sx * sx);
        ay[k] = -0.5 / (sy * sy);
        amp[k] = (double)th[M::iN(k)] / (6.283185307179586 * sx * sy);
        yk[k] = th[M::iy(k)];
        const double xk = th[M::ix(k)];
#pragma unroll
        for (int c = 0; c < R; c++) {
            const double dx = (c + 0.5) - xk;
            gx[k][c] = exp(ax * (dx * dx));
        }
    }
    const double b = th[M::IB];
    double ll = 0.0;
#pragma unroll 1
    for (int r = 0; r < R; r++) {
        double gy[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const double dy = (r + 0.5) - yk[k];
            gy[k] = amp[k] * exp(ay[k] * (dy * dy));
        }
#pragma unroll
        for (int c = 0; c < R; c++) {
            double mu = b;
#pragma unroll
            for (int k = 0; k < K; k++) mu += gy[k] * gx[k][c];
            const double q = qs[(r * R + c) * 64];
            ll += q * log(mu) - mu;
        }
    }
    return ll;
}

// initial vector: heuristic (K = 1), given, or the previous model plus an emitter at the largest
// positive residual (K >= 2; init holds the K-1 model's theta)
template <int KK, int R>
__device__ __forceinline__ void cf_init(const float *qs, const float *init, int mode, float sigma0, float (&th)[CP<KK>::P])
{
    using M = CM<KK>;
    constexpr int P = M::P, K = M::K;
    if (mode == CF_INIT_GIVEN) {
#pragma unroll
        for (int i = 0; i < P; i++) th[i] = init[i];
        return;
    }
    if constexpr (K > 1) {
      if (mode == CF_INIT_ADD) {
        using M0 = CM<KK - 1>;
        constexpr int P0 = M0::P;
        float t0[P0];
#pragma unroll
        for (int i = 0; i < P0; i++) t0[i] = init[i];
        if constexpr (M::AS) {
            // the previous emitters made round at their narrower width: a fit of fewer emitters
            // than the ROI holds widens along the line to the missing one, and the residual of
            // that wide model hides it
#pragma unroll
            for (int k = 0; k < K - 1; k++) {
                const float w = fminf(t0[M0::isx(k)], t0[M0::isy(k)]);
                t0[M0::isx(k)] = w;
                t0[M0::isy(k)] = w;
            }
        }
        const float b = t0[M0::IB];
        float best = -1e30f;
        int bi = (R / 2) * R + R / 2;
        for (int r = 0; r < R; r++)
            for (int c = 0; c < R; c++) {
                float J0[P0];
                float mu;
                if constexpr (!M::AS) {
                    const float s = t0[3 * (K - 1)];
                    const float h = -0.5f / (s * s), nrm = 0.15915494309189535f / (s * s);
                    mu = b;
#pragma unroll
                    for (int k = 0; k < K - 1; k++) {
                        const float dx = ((float)c + 0.5f) - t0[3 * k + 1], dy = ((float)r + 0.5f) - t0[3 * k + 2];
                        mu += t0[3 * k] * nrm * __expf(h * (dx * dx + dy * dy));
                    }
                    (void)J0;
                } else {
                    mu = cf_jacobian<KK - 1>(t0, (float)c + 0.5f, (float)r + 0.5f, J0);
                }
                const float res = qs[(r * R + c) * 64] - mu;
                if (res > best) { best = res; bi = r * R + c; }
            }
#pragma unroll
        for (int k = 0; k < K - 1; k++) {
#pragma unroll
            for (int e = 0; e < M::E; e++) th[M::E * k + e] = t0[M0::E * k + e];
        }
        if constexpr (!M::AS) {
            const float s = t0[3 * (K - 1)];
            th[3 * (K - 1)] = fmaxf(best, 5.0f) * 6.283185307179586f * s * s;
            th[3 * (K - 1) + 1] = (float)(bi % R) + 0.5f;
            th[3 * (K - 1) + 2] = (float)(bi / R) + 0.5f;
            th[3 * K] = s;
        } else {
            // the new emitter takes the mean widths of the others
            float msx = 0.0f, msy = 0.0f;
#pragma unroll
            for (int k = 0; k < K - 1; k++) { msx += t0[M0::isx(k)]; msy += t0[M0::isy(k)]; }
            msx /= (float)(K - 1);
            msy /= (float)(K - 1);
            th[M::iN(K - 1)] = fmaxf(best, 5.0f) * 6.283185307179586f * msx * msy;
            th[M::ix(K - 1)] = (float)(bi % R) + 0.5f;
            th[M::iy(K - 1)] = (float)(bi / R) + 0.5f;
            th[M::isx(K - 1)] = msx;
            th[M::isy(K - 1)] = msy;
        }
        th[M::IB] = b;
        return;
      }
    }
    // heuristic: background from the smaller of the opposite edge means, photons from the excess,
    // position at the excess-weighted centroid (single emitter; extra emitters share it)
    float e[4] = {0, 0, 0, 0};
    for (int i = 0; i < R; i++) {
        e[0] += qs[(0 * R + i) * 64];
        e[1] += qs[((R - 1) * R + i) * 64];
        e[2] += qs[(i * R + 0) * 64];
        e[3] += qs[(i * R + R - 1) * 64];
    }
    const float b = 0.5f * (fminf(e[0], e[1]) + fminf(e[2], e[3])) / (float)R;
    float tot = 0.0f, sx = 0.0f, sy = 0.0f;
    for (int r = 0; r < R; r++)
        for (int c = 0; c < R; c++) {
            const float v = fmaxf(qs[(r * R + c) * 64] - b, 0.0f);
            tot += v;
            sx = fmaf(v, (float)c + 0.5f, sx);
            sy = fmaf(v, (float)r + 0.5f, sy);
        }
    const float cx = tot > 0.0f ? sx / tot : 0.5f * R, cy = tot > 0.0f ? sy / tot : 0.5f * R;
#pragma unroll
    for (int k = 0; k < K; k++) {
        th[M::iN(k)] = fmaxf(tot, 50.0f) / (float)K;
        th[M::ix(k)] = cx + 0.3f * (float)k;
        th[M::iy(k)] = cy;
        if constexpr (M::AS) {
            th[M::isx(k)] = sigma0;
            th[M::isy(k)] = sigma0;
        }
    }
    if constexpr (!M::AS) th[3 * K] = sigma0;
    th[M::IB] = fmaxf(b, 0.5f);
}

// ---- FP64 full-Fisher CRLB of one model (standard deviations of x_k and y_k, px) ----------
// I = sum_pixels grad(mu) grad(mu)^T / mu in double, packed (upper triangle) in registers: every
// loop over parameters is a compile-time loop, so each entry has a fixed register.  The 16-parameter
// elliptical model's 136 doubles spill into the accumulation registers (one wave per SIMD).
// Cholesky I = U^T U, then (I^-1)_ee = |U^-T e_e|^2 for the position indices e.  The divisions of
// the model's derivatives move out of the pixel loop (1/s, 1/sx, 1/sy once per ROI; dmu/dN_k is
// the emitter's normalised Gaussian), and the forward substitution for e_e starts at row e,
// where z is first non-zero.
template <int KK, int R>
__device__ __forceinline__ void cf_crlb64_reg(const float (&th)[CP<KK>::P], float (&cx)[CM<KK>::K], float (&cy)[CM<KK>::K])
{
    using M = CM<KK>;
    constexpr int P = M::P, K = M::K, NS = CP<KK>::NS;
    double I[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) I[i] = 0.0;
    // per emitter: photons, centre, 1/sx^2, 1/sy^2, 1/(2 pi sx sy), 1/sx, 1/sy (2D: one width)
    double nk[K], xk[K], yk[K], ax[K], ay[K], nrm[K], jx[K], jy[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double sx = th[M::isx(k)], sy = th[M::isy(k)];
        nk[k] = th[M::iN(k)];
        xk[k] = th[M::ix(k)];
        yk[k] = th[M::iy(k)];
        ax[k] = 1.0 / (sx * sx);
        ay[k] = 1.0 / (sy * sy);
        nrm[k] = 1.0 / (6.283185307179586 * sx * sy);
        jx[k] = 1.0 / sx;
        jy[k] = 1.0 / sy;
    }
    const double b = th[M::IB];
    for (int r = 0; r < R; r++) {
        const double rpx = r + 0.5;
        for (int c = 0; c < R; c++) {
            const double cpx = c + 0.5;
            double J[P];
            double mu = b, jw = 0.0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const double dx = cpx - xk[k], dy = rpx - yk[k];
                const double u = dx * dx * ax[k], v = dy * dy * ay[k];
                const double g = nrm[k] * exp(-0.5 * (u + v));
                const double mk = nk[k] * g;
                mu += mk;
                J[M::iN(k)] = g;
                J[M::ix(k)] = mk * dx * ax[k];
                J[M::iy(k)] = mk * dy * ay[k];
                if constexpr (M::AS) {
                    J[M::isx(k)] = mk * (u - 1.0) * jx[k];
                    J[M::isy(k)] = mk * (v - 1.0) * jy[k];
                } else {
                    jw += mk * (u + v - 2.0);
                }
            }
            if constexpr (!M::AS) J[3 * K] = jw * jx[0];
            J[P - 1] = 1.0;
            const double im = 1.0 / mu;
#pragma unroll
            for (int a = 0; a < P; a++) {
                const double ja = J[a] * im;
#pragma unroll
                for (int b2 = a; b2 < P; b2++) I[tri(a, b2, P)] += ja * J[b2];
            }
        }
    }
    // Cholesky I = U^T U in place; a non-positive pivot marks the matrix singular
    bool ok = true;
    cf_sfor<0, P>([&](auto JC) {
        constexpr int j = decltype(JC)::value;
        double d = I[tri(j, j, P)];
        cf_sfor<0, j>([&](auto KC) {
            constexpr int k = decltype(KC)::value;
            d -= I[tri(k, j, P)] * I[tri(k, j, P)];
        });
        ok = ok && d > 0.0;
        const double u = sqrt(d), iu = 1.0 / u;
        I[tri(j, j, P)] = u;
        cf_sfor<j + 1, P>([&](auto IC) {
            constexpr int i = decltype(IC)::value;
            double v = I[tri(j, i, P)];
            cf_sfor<0, j>([&](auto KC) {
                constexpr int k = decltype(KC)::value;
                v -= I[tri(k, j, P)] * I[tri(k, i, P)];
            });
            I[tri(j, i, P)] = v * iu;
        });
    });
    // (I^-1)_ee = |U^-T e_e|^2 for e = x_k, y_k; z_i = 0 for i < e
    cf_sfor<0, 2 * K>([&](auto WC) {
        constexpr int w = decltype(WC)::value;
        constexpr int e = (w & 1) ? M::iy(w >> 1) : M::ix(w >> 1);
        double z[P], var = 0.0;
        cf_sfor<e, P>([&](auto IC) {
            constexpr int i = decltype(IC)::value;
            double v = (i == e) ? 1.0 : 0.0;
            cf_sfor<e, i>([&](auto KC) {
                constexpr int k2 = decltype(KC)::value;
                v -= I[tri(k2, i, P)] * z[k2];
            });
            z[i] = v / I[tri(i, i, P)];
            var += z[i] * z[i];
        });
        const float sd = ok && var > 0.0 ? (float)sqrt(var) : -1.0f;
        if (w & 1) cy[w >> 1] = sd; else cx[w >> 1] = sd;
    });
}

struct CSelConst {
    float llr_half;     // accept K+1 over K when ll(K+1) - ll(K) exceeds this (half the chi^2 bound)
    float n_min;        // minimum photons of every emitter of an accepted model
    float pixel_size, qe;
    int kmax;
    float pG[5], pL[5], zcorr;   // astigmatic z calibration (elliptical model)
};

// some photon count, position or width of theta at its box bound (cf_project); every box bound
// lies outside the range cf_valid accepts
template <int KK, int R>
__device__ __forceinline__ bool cf_at_bound(const float (&th)[CP<KK>::P])
{
    using M = CM<KK>;
    bool b = false;
#pragma unroll
    for (int k = 0; k < M::K; k++) {
        b = b || th[M::iN(k)] <= 1.0f;
        b = b || th[M::ix(k)] <= -0.5f || th[M::ix(k)] >= (float)R + 0.5f;
        b = b || th[M::iy(k)] <= -0.5f || th[M::iy(k)] >= (float)R + 0.5f;
        b = b || th[M::isx(k)] <= 0.4f || th[M::isx(k)] >= 0.5f * (float)R;
        b = b || th[M::isy(k)] <= 0.4f || th[M::isy(k)] >= 0.5f * (float)R;
    }
    return b;
}

template <int KK, int R>
__device__ __forceinline__ bool cf_valid(const float *t, float n_min)
{
    using M = CM<KK>;
    for (int k = 0; k < M::K; k++) {
        if (!(t[M::iN(k)] >= n_min)) return false;
        if (!(t[M::ix(k)] > 0.0f && t[M::ix(k)] < (float)R && t[M::iy(k)] > 0.0f && t[M::iy(k)] < (float)R)) return false;
        if constexpr (M::AS)
            if (!(t[M::isx(k)] > 0.45f && t[M::isx(k)] < 0.45f * (float)R && t[M::isy(k)] > 0.45f &&
                  t[M::isy(k)] < 0.45f * (float)R))
                return false;
    }
    if constexpr (M::AS) return true;
    else return t[3 * M::K] > 0.45f && t[3 * M::K] < 0.45f * (float)R;
}

// z of an elliptical emitter (bfgsMLE_AS3D_core.h:266-279): d = sx^2 - sy^2, polynomial by sign
__device__ __forceinline__ float cf_zpoly(const float (&pc)[5], float d, float corr)
{
    const float d2 = d * d, d3 = d2 * d, d4 = d2 * d2;
    float z = pc[0] * d4;
    z = fmaf(pc[1], d3, z);
    z = fmaf(pc[2], d2, z);
    z = fmaf(pc[3], d, z);
    z = z + pc[4];
    return z * corr;
}

template <int KK, int R>
__device__ void cf_write_rows(const float *t, const uint16_t *hdr, const CSelConst &c, float *rows)
{
    using M = CM<KK>;
    constexpr int K = M::K;
    float th[CP<KK>::P];
    for (int i = 0; i < CP<KK>::P; i++) th[i] = t[i];
    float cx[K], cy[K];
    cf_crlb64_reg<KK, R>(th, cx, cy);
    const float b = th[M::IB];
    const int half = R / 2;
    const float xc = (float)hdr[0], yc = (float)hdr[1];
    const int fr = (int)hdr[2] + (int)hdr[3] * 65536;
    // a record is 48 bytes: three 16-byte stores when the row array is 16-byte aligned
    const bool vec = (reinterpret_cast<uintptr_t>(rows) & 15) == 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        float o[QCS_REC_FLOATS];
        const float sx = th[M::isx(k)], sy = th[M::isy(k)];
        const float peak = th[M::iN(k)] * 0.15915494309189535f / (sx * sy);
        o[QCS_POS_PPHO] = peak;
        o[QCS_POS_XPOS] = (th[M::ix(k)] - (float)half) + xc;
        o[QCS_POS_YPOS] = (th[M::iy(k)] - (float)half) + yc;
        if constexpr (M::AS) {
            const float d = sx * sx - sy * sy;
            o[QCS_POS_ZPOS] = d >= 0.0f ? cf_zpoly(c.pG, d, c.zcorr) : cf_zpoly(c.pL, d, c.zcorr);
        } else {
            o[QCS_POS_ZPOS] = 0.0f;
        }
        o[QCS_POS_SIGX] = sx;
        o[QCS_POS_SIGY] = sy;
        o[QCS_POS_TPHO] = th[M::iN(k)];
        o[QCS_POS_BAKG] = b;
        o[QCS_POS_PSNR] = sqrtf(c.qe) * peak / sqrtf(peak + b);
        o[QCS_POS_CRBX] = cx[k] > 0.0f ? cx[k] * c.pixel_size : -1.0f;
        o[QCS_POS_CRBY] = cy[k] > 0.0f ? cy[k] * c.pixel_size : -1.0f;
        o[QCS_POS_FRME] = (float)fr;
        float *dst = rows + (size_t)k * QCS_REC_FLOATS;
        if (vec) {
#pragma unroll
            for (int v = 0; v < QCS_REC_FLOATS / 4; v++)
                reinterpret_cast<float4 *>(dst)[v] = make_float4(o[4 * v], o[4 * v + 1], o[4 * v + 2], o[4 * v + 3]);
        } else {
#pragma unroll
            for (int i = 0; i < QCS_REC_FLOATS; i++) dst[i] = o[i];
        }
    }
}


// observed photons of ROI roi into the lane's LDS column: 4-byte loads (every ROI record starts
// 4-byte aligned since R (R + 1) is even)
template <int R>
__device__ __forceinline__ void cf_load_q(const uint16_t *rois, int roi, const CFitConst &c, float *qs)
{
    constexpr int RR = R * R, WR = R * (R + 1);
    const uint32_t *src = reinterpret_cast<const uint32_t *>(rois + (size_t)roi * WR);
#pragma unroll
    for (int w = 0; w < (RR + 1) / 2; w++) {
        const uint32_t u = src[w];
        qs[(2 * w) * 64] = fmaxf(((float)(u & 0xffffu) - c.offset) * c.kadc, 1.0f) * c.inv_qe;
        if (2 * w + 1 < RR) qs[(2 * w + 1) * 64] = fmaxf(((float)(u >> 16) - c.offset) * c.kadc, 1.0f) * c.inv_qe;
    }
}

// raw counts of ROI roi into the lane's LDS column (converted on use, QLds16)
template <int R>
__device__ __forceinline__ void cf_load_q16(const uint16_t *rois, int roi, uint16_t *qs)
{
    constexpr int RR = R * R, WR = R * (R + 1);
    const uint32_t *src = reinterpret_cast<const uint32_t *>(rois + (size_t)roi * WR);
#pragma unroll
    for (int w = 0; w < (RR + 1) / 2; w++) {
        const uint32_t u = src[w];
        qs[(2 * w) * 64] = (uint16_t)(u & 0xffffu);
        if (2 * w + 1 < RR) qs[(2 * w + 1) * 64] = (uint16_t)(u >> 16);
    }
}

// A parameter held at its box bound (cf_project) whose descent direction -gu points further out:
// it is left out of the quasi-Newton step (active set), so the line search is not defeated by a
// projection that cancels part of the step.
template <int KK, int R>
__device__ __forceinline__ bool cf_pinned(const float (&th)[CP<KK>::P], int i, float gu)
{
    constexpr int P = CP<KK>::P;
    const float v = th[i];
    if (i == P - 1) return v <= 0.01f && gu > 0.0f;                                          // b
    const int e = CM<KK>::AS ? i % 5 : (i == P - 2 ? 3 : i % 3);
    if (e >= 3) return (v <= 0.4f && gu > 0.0f) || (v >= 0.5f * (float)R && gu < 0.0f);     // widths
    if (e == 0) return v <= 1.0f && gu > 0.0f;                                               // N_k
    return (v <= -0.5f && gu > 0.0f) || (v >= (float)R + 0.5f && gu < 0.0f);                // x_k, y_k
}

// Quasi-Newton direction du = -D gu over the free parameters (scaled units, gu = g * scale;
// every parameter is free unless active_set).
// Returns true when the fit has converged: the free step is below tol in every coordinate with a
// non-positive slope (this includes a zero free gradient).  Otherwise du, slope = du . gu < 0 and
// dmax = max |du| are set; a non-descent direction restarts from the gradient with D = I (first
// = true), and with D = I the step is limited to 0.1 scaled units.
template <int KK, int R, class DA>
__device__ __forceinline__ bool cf_direction(const float (&th)[CP<KK>::P], const float (&g)[CP<KK>::P], DA &D, bool &first,
                                             float tol, bool active_set, float (&du)[CP<KK>::P], float &slope, float &dmax)
{
    constexpr int P = CP<KK>::P;
    float gu[P];
    bool pin[P];
#pragma unroll
    for (int i = 0; i < P; i++) {
        gu[i] = g[i] * cf_scale<KK>(i);
        pin[i] = active_set && cf_pinned<KK, R>(th, i, gu[i]);
    }
    slope = 0.0f;
#pragma unroll
    for (int i = 0; i < P; i++) {
        float a = 0.0f;
#pragma unroll
        for (int jj = 0; jj < P; jj++)
            if (!pin[jj]) a = fmaf(D[i <= jj ? tri(i, jj, P) : tri(jj, i, P)], gu[jj], a);
        du[i] = pin[i] ? 0.0f : -a;
        slope = fmaf(du[i], gu[i], slope);
    }
    {
        float dm = 0.0f;
#pragma unroll
        for (int i = 0; i < P; i++) dm = fmaxf(dm, fabsf(du[i]));
        if (dm < tol && slope <= 0.0f) return true;
    }
    if (!(slope < 0.0f)) {   // not a descent direction: restart from the gradient
#pragma unroll
        for (int i = 0; i < P; i++) du[i] = pin[i] ? 0.0f : -gu[i];
        slope = 0.0f;
#pragma unroll
        for (int i = 0; i < P; i++) slope = fmaf(du[i], gu[i], slope);
#pragma unroll
        for (int i = 0; i < P; i++)
#pragma unroll
            for (int jj = i; jj < P; jj++) D[tri(i, jj, P)] = (i == jj) ? 1.0f : 0.0f;
        first = true;
    }
    if (first) {   // identity inverse Hessian: limit the first step to 0.1 scaled units
        float m = 0.0f;
#pragma unroll
        for (int i = 0; i < P; i++) m = fmaxf(m, fabsf(du[i]));
        if (m > 0.1f) {
            const float f = 0.1f / m;
            slope = 0.0f;
#pragma unroll
            for (int i = 0; i < P; i++) {
                du[i] *= f;
                slope = fmaf(du[i], gu[i], slope);
            }
        }
    }
    dmax = 0.0f;
#pragma unroll
    for (int i = 0; i < P; i++) dmax = fmaxf(dmax, fabsf(du[i]));
    return false;
}

// BFGS update of the packed inverse Hessian after an accepted step from th to tn with gradients
// g -> gn (physical units); with D = I (first) D is first rescaled to (s.y / y.y) I.
template <int KK, class DA>
__device__ __forceinline__ void cf_bfgs_update(DA &D, bool &first, const float (&th)[CP<KK>::P], const float (&tn)[CP<KK>::P],
                                               const float (&g)[CP<KK>::P], const float (&gn)[CP<KK>::P])
{
    constexpr int P = CP<KK>::P;
    float s[P], y[P], sy = 0.0f, yy = 0.0f;
#pragma unroll
    for (int i = 0; i < P; i++) s[i] = (tn[i] - th[i]) / cf_scale<KK>(i);
#pragma unroll
    for (int i = 0; i < P; i++) {
        y[i] = (gn[i] - g[i]) * cf_scale<KK>(i);
        sy = fmaf(s[i], y[i], sy);
        yy = fmaf(y[i], y[i], yy);
    }
    if (!(sy > 1e-12f)) return;   // curvature condition fails: keep D
    if (first) {
        const float sc = sy / yy;
#pragma unroll
        for (int i = 0; i < P; i++)
#pragma unroll
            for (int jj = i; jj < P; jj++) D[tri(i, jj, P)] = (i == jj) ? sc : 0.0f;
        first = false;
    }
    // D <- D - rho (s u^T + u s^T) + (rho^2 y.u + rho) s s^T, u = D y
    const float rho = 1.0f / sy;
    float u[P], yu = 0.0f;
#pragma unroll
    for (int i = 0; i < P; i++) {
        float a = 0.0f;
#pragma unroll
        for (int jj = 0; jj < P; jj++) a = fmaf(D[i <= jj ? tri(i, jj, P) : tri(jj, i, P)], y[jj], a);
        u[i] = a;
        yu = fmaf(y[i], a, yu);
    }
    const float cs = fmaf(rho * rho, yu, rho);
#pragma unroll
    for (int i = 0; i < P; i++)
#pragma unroll
        for (int jj = i; jj < P; jj++) {
            float &e = D[tri(i, jj, P)];
            e = fmaf(cs * s[i], s[jj], fmaf(-rho, fmaf(s[i], u[jj], u[i] * s[jj]), e));
        }
}

// Position of this lane's item in a list that the whole wave appends to (one atomic per wave),
// -1 if the lane takes no slot.  Every lane of the wave must make the call.
__device__ __forceinline__ int wave_append(int *count, bool take)
{
    const unsigned long long m = __ballot(take);
    if (m == 0ull) return -1;
    const int lane = threadIdx.x & 63, leader = __ffsll(m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, leader, 64);
    return take ? base + __popcll(m & ((1ull << lane) - 1ull)) : -1;
}

// Appends from a whole work-group of W waves to NL lists with counters count[0 .. NL - 1]: lane
// flags take[k]; pos[k] receives the lane's slot in list k, or -1.  One global atomic per
// work-group and list instead of one per wave (wave_append): atomics on one address serialise
// at the L2, at about 10 ns each, which outweighs a light kernel's own work over 10^6 items.
// Every thread of the work-group calls it (two barriers).
template <int W, int NL>
struct BlockAppend {
    int woff[W][NL];
    int base[NL];
};

template <int W, int NL>
__device__ __forceinline__ void block_append(BlockAppend<W, NL> &sh, int *count, const bool (&take)[NL], int (&pos)[NL])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long m[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        m[k] = __ballot(take[k]);
        if (lane == 0) sh.woff[w][k] = __popcll(m[k]);
    }
    __syncthreads();
    if (threadIdx.x < NL) {
        const int k = threadIdx.x;
        int s = 0;
        for (int i = 0; i < W; i++) {
            const int c = sh.woff[i][k];
            sh.woff[i][k] = s;
            s += c;
        }
        sh.base[k] = s > 0 ? atomicAdd(count + k, s) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NL; k++)
        pos[k] = take[k] ? sh.base[k] + sh.woff[w][k] + __popcll(m[k] & ((1ull << lane) - 1ull)) : -1;
}

// ---- host side ----------------------------------------------------------------------------

// one set of ROIs of one size in a converged-fitting launch.  Work item j fits ROI
// roi = idx ? idx[j] : j (j < n, or < *n_dev when given); init (stride istride), theta (physical
// units, stride ostride), info (4 ints: iterations, flags, loss evaluations, 0) and ll (FP64
// log-likelihood, may be null) are indexed by roi.
struct CfqSet {
    const uint16_t *rois;
    const int *idx;
    const int *n_dev;
    const float *init;
    float *theta;
    int *info;
    double *ll;
    float *rec;      // work-item records of the persistent path (set by the launcher)
    int n, R, istride, ostride;
    int init_mode;   // start of this set's fits (CF_INIT_*; a launch may mix them)
};

// arguments of one persistent launch (up to three ROI sizes, each with its own queue head)
struct CfqArgs {
    CfqSet set[3];
    int nsets;
    int *heads;
    CFitConst c;
};

CFitConst cfit_const(const qcs_params &p, const qcs_tuning &t, int K, int init_mode);
// persistent work-queue fit of K emitters over the sets (cfq.hip), each set started by its own
// init_mode; the model (2D or elliptical) follows the context's LocType
int cfq_launch(qcs_ctx *ctx, int K, const CfqSet *sets, int nsets, hipStream_t s);

// model key of K emitters under the context's LocType, and its parameter count
inline int cf_kk(const qcs_params &p, int K) { return p.LocType == QCS_LOCTYPE_AS3D ? K + CF_ELLIPTIC : K; }
inline int cf_np(int KK) { return KK > CF_ELLIPTIC ? 5 * (KK - CF_ELLIPTIC) + 1 : 3 * KK + 2; }

#define CF_DISPATCH_KK(KK_, ...)                                                          \
    switch (KK_) {                                                                         \
    case 1: { constexpr int KT = 1; __VA_ARGS__; } break;                                  \
    case 2: { constexpr int KT = 2; __VA_ARGS__; } break;                                  \
    case 3: { constexpr int KT = 3; __VA_ARGS__; } break;                                  \
    case 11: { constexpr int KT = 11; __VA_ARGS__; } break;                                \
    case 12: { constexpr int KT = 12; __VA_ARGS__; } break;                                \
    case 13: { constexpr int KT = 13; __VA_ARGS__; } break;                                \
    default: return fail(QCS_ERR_ARG, "model %d: emitter count not in 1..3", KK_);          \
    }

#define CF_DISPATCH_R(R_, ...)                                                            \
    switch (R_) {                                                                          \
    case 7: { constexpr int RT = 7; __VA_ARGS__; } break;                                  \
    case 9: { constexpr int RT = 9; __VA_ARGS__; } break;                                  \
    case 11: { constexpr int RT = 11; __VA_ARGS__; } break;                                \
    default: return fail(QCS_ERR_ARG, "converged fitting supports ROISize 7, 9, 11 (got %d)", R_); \
    }

}  // namespace qcs
