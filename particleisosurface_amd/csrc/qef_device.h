// Leaf vertex QEF (TNode::NodeCalcNode, src/iso_method_ours.cpp:101-174) on the device.
#pragma once
#include "device_common.h"

namespace pisurf {

// Finite-difference (negated) gradient of the 27 samples (Evaluator::GridEval,
// src/evaluator.cpp:185-251): central inside, one-sided on the faces, step = cellsize / 2.
template <class R> __device__ __forceinline__ void fd_grads27(const R f[27], float cellsize, R g[27][3]) {
    const float step = cellsize / 2.0f;
    const R s1 = R(step), s2 = R(step * 2.0f);
#pragma unroll
    for (int i = 0; i < 27; i++) {
        const int t[3] = {i % 3, (i / 3) % 3, i / 9};
        const int st[3] = {1, 3, 9};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (t[a] == 0) g[i][a] = (f[i] - f[i + st[a]]) / s1;
            else if (t[a] == 2) g[i][a] = (f[i - st[a]] - f[i]) / s1;
            else g[i][a] = (f[i - st[a]] - f[i + st[a]]) / s2;
        }
    }
}

// Sample point s (0..26, index z*9 + y*3 + x) of a node box (TNode::GenerateSampling,
// src/iso_method_ours.cpp:52-75) with correctly rounded, uncontracted float ops.
__device__ __forceinline__ void sample_point(float4 c, int s, float p[3]) {
    // every operation rounded on its own, as the reference computes it (and alike in every kernel)
#pragma clang fp contract(off)
    const int t[3] = {s % 3, (s / 3) % 3, s / 9};
    const float cc[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float mn = cc[a] - c.w, mx = cc[a] + c.w;
        const float u = float(t[a]) / 2.0f;
        p[a] = (1.0f - u) * mn + u * mx;
    }
}

// 4x4 inverse by adjugate / determinant, float (singular input gives inf/NaN entries)
__device__ __forceinline__ void inverse4f(const float A[4][4], float inv[4][4]) {
#pragma clang fp contract(off)
    float cof[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            float m[3][3];
            int ri = 0;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (r == i) continue;
                int ci = 0;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    if (c == j) continue;
                    m[ri][ci++] = A[r][c];
                }
                ri++;
            }
            const float mdet = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                               m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
            cof[i][j] = ((i + j) & 1) ? -mdet : mdet;
        }
    const float det = A[0][0] * cof[0][0] + A[0][1] * cof[0][1] + A[0][2] * cof[0][2] + A[0][3] * cof[0][3];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) inv[i][j] = cof[j][i] / det;
}

// symmetric 4x4 Jacobi eigendecomposition in double (Golub & Van Loan, Alg. 8.5.3)
__device__ __forceinline__ void jacobi4d(double a[4][4], double lam[4], double Q[4][4]) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) Q[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 30; sweep++) {
        double off = 0, dia = 0;
        for (int i = 0; i < 4; i++) {
            dia += a[i][i] * a[i][i];
            for (int j = i + 1; j < 4; j++) off += a[i][j] * a[i][j];
        }
        if (!(off > 5.0e-34 * dia)) break;
        for (int p = 0; p < 3; p++)
            for (int q = p + 1; q < 4; q++) {
                const double apq = a[p][q];
                if (apq == 0.0) continue;
                const double tau = (a[q][q] - a[p][p]) / (2.0 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                for (int r = 0; r < 4; r++) {
                    const double arp = a[r][p], arq = a[r][q];
                    a[r][p] = c * arp - s * arq;
                    a[r][q] = s * arp + c * arq;
                }
                for (int r = 0; r < 4; r++) {
                    const double apr = a[p][r], aqr = a[q][r];
                    a[p][r] = c * apr - s * aqr;
                    a[q][r] = s * apr + c * aqr;
                }
                a[p][q] = a[q][p] = 0.0;
                for (int r = 0; r < 4; r++) {
                    const double qrp = Q[r][p], qrq = Q[r][q];
                    Q[r][p] = c * qrp - s * qrq;
                    Q[r][q] = s * qrp + c * qrq;
                }
            }
    }
    for (int i = 0; i < 4; i++) lam[i] = a[i][i];
}

// Solve the leaf QEF from 27 samples. Returns the candidate vertex pc (4 values) and whether it
// lies inside the node box; err is the DMC error (calcErrorDMC, src/iso_method_ours.cpp:40-50)
// when inside. The samples come through accessors (unrolled, so every index is a constant):
// f_at(i) the value, p_at(i, p) the position (float), g_at(i, g) the gradient. A caller that
// computes positions and gradients in the accessors keeps only the 27 values live.
template <class R, class FA, class PA, class GA>
__device__ __forceinline__ bool qef_core(float4 c, FA f_at, PA p_at, GA g_at, int mode, float svd_tol, R pc[4],
                                         R& err) {
    if (mode == 1) {
#pragma clang fp contract(off)
        float q[15];
#pragma unroll
        for (int k = 0; k < 15; k++) q[k] = 0.0f;
#pragma unroll
        for (int i = 0; i < 27; i++) {
            float pt[3];
            R g[3];
            p_at(i, pt);
            g_at(i, g);
            float pl[5];
            pl[0] = float(g[0]);
            pl[1] = float(g[1]);
            pl[2] = float(g[2]);
            pl[3] = -1.0f;
            // plain operators: contraction is off in this scope (the *_rn helpers would be
            // inlined as contractable operations)
            pl[4] = -((pt[0] * pl[0] + pt[1] * pl[1]) + pt[2] * pl[2]) + float(f_at(i));
            int k = 0;
#pragma unroll
            for (int a = 0; a < 5; a++)
#pragma unroll
                for (int b = a; b < 5; b++) q[k++] += pl[a] * pl[b];
        }
        float A[4][4], B[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const int idx = ((2 * 4 + 3 - a) * a) / 2;
#pragma unroll
            for (int b = a; b < 4; b++) {
                A[a][b] = q[idx + b - a];
                A[b][a] = A[a][b];
            }
            B[a] = -q[idx + 4 - a];
        }
        float inv[4][4];
        inverse4f(A, inv);
#pragma unroll
        for (int a = 0; a < 4; a++) {
            float r = 0.0f;
#pragma unroll
            for (int b = 0; b < 4; b++) r += inv[b][a] * B[b];
            pc[a] = R(r);
        }
    } else {
        double q[5][5];
        for (int a = 0; a < 5; a++)
            for (int b = 0; b < 5; b++) q[a][b] = 0.0;
#pragma unroll
        for (int i = 0; i < 27; i++) {
            float pt[3];
            R g[3];
            p_at(i, pt);
            g_at(i, g);
            double pl[5] = {double(g[0]), double(g[1]), double(g[2]), -1.0, 0.0};
            pl[4] = -(double(pt[0]) * pl[0] + double(pt[1]) * pl[1] + double(pt[2]) * pl[2]) + double(f_at(i));
            for (int a = 0; a < 5; a++)
                for (int b = 0; b < 5; b++) q[a][b] += pl[a] * pl[b];
        }
        double A[4][4], rhs[4];
        const double m[4] = {double(c.x), double(c.y), double(c.z), double(f_at(13))};
        for (int a = 0; a < 4; a++) {
            rhs[a] = -q[a][4];
            for (int b = 0; b < 4; b++) A[a][b] = q[a][b];
            for (int b = 0; b < 4; b++) rhs[a] -= A[a][b] * m[b];
        }
        double lam[4], Q[4][4];
        jacobi4d(A, lam, Q);
        double lmax = 0.0;
        for (int a = 0; a < 4; a++) lmax = fmax(lmax, fabs(lam[a]));
        double x[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; k++) {
            if (fabs(lam[k]) <= double(svd_tol) * lmax || lam[k] == 0.0) continue;
            double dot = 0;
            for (int a = 0; a < 4; a++) dot += Q[a][k] * rhs[a];
            for (int a = 0; a < 4; a++) x[a] += Q[a][k] * dot / lam[k];
        }
        for (int a = 0; a < 4; a++) pc[a] = R(m[a] + x[a]);
    }
    const float mn[3] = {__fsub_rn(c.x, c.w), __fsub_rn(c.y, c.w), __fsub_rn(c.z, c.w)};
    const float mx[3] = {__fadd_rn(c.x, c.w), __fadd_rn(c.y, c.w), __fadd_rn(c.z, c.w)};
    const bool inside = pc[0] >= R(mn[0]) && pc[0] <= R(mx[0]) && pc[1] >= R(mn[1]) && pc[1] <= R(mx[1]) &&
                        pc[2] >= R(mn[2]) && pc[2] <= R(mx[2]);
    err = R(1e30);
    if (inside) {
        err = R(0);
#pragma unroll
        for (int i = 0; i < 27; i++) {
            float pt[3];
            R g[3];
            p_at(i, pt);
            g_at(i, g);
            const R dp = (g[0] * (pc[0] - R(pt[0])) + g[1] * (pc[1] - R(pt[1]))) + g[2] * (pc[2] - R(pt[2]));
            const R e = dp - pc[3];
            err += (e * e) / (R(1) + (g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
        }
    }
    return inside;
}

// the same from given sample positions, values and gradients
template <class R>
__device__ __forceinline__ bool qef_solve(float4 c, const float pts[27][3], const R f[27], const R g[27][3], int mode,
                                          float svd_tol, R pc[4], R& err) {
    return qef_core<R>(
        c, [&](int i) { return f[i]; },
        [&](int i, float p[3]) {
            for (int a = 0; a < 3; a++) p[a] = pts[i][a];
        },
        [&](int i, R gg[3]) {
            for (int a = 0; a < 3; a++) gg[a] = g[i][a];
        },
        mode, svd_tol, pc, err);
}

// Finite-difference (negated) gradient of sample i of a node's 27 (as fd_grads27)
template <class R> __device__ __forceinline__ void fd_grad_at(const R f[27], int i, float cellsize, R g[3]) {
    const float step = cellsize / 2.0f;
    const R s1 = R(step), s2 = R(step * 2.0f);
    const int t[3] = {i % 3, (i / 3) % 3, i / 9};
    const int st[3] = {1, 3, 9};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (t[a] == 0) g[a] = (f[i] - f[i + st[a]]) / s1;
        else if (t[a] == 2) g[a] = (f[i - st[a]] - f[i]) / s1;
        else g[a] = (f[i - st[a]] - f[i + st[a]]) / s2;
    }
}

// the leaf QEF of a node from its 27 sample values: positions (sample_point) and gradients
// (fd_grads27) are computed where they are used, so only the values stay in registers
template <class R>
__device__ __forceinline__ bool qef_solve_samples(float4 c, const R f[27], int mode, float svd_tol, R pc[4], R& err) {
    const float cellsize = __fmul_rn(2.0f, c.w);
    return qef_core<R>(
        c, [&](int i) { return f[i]; }, [&](int i, float p[3]) { sample_point(c, i, p); },
        [&](int i, R gg[3]) { fd_grad_at<R>(f, i, cellsize, gg); }, mode, svd_tol, pc, err);
}

}  // namespace pisurf
