// Per-particle anisotropy (Evaluator::compute_Gs_xMeans, src/evaluator.cpp:656-749).
// One thread per particle in cell-sorted order: lanes of a wave sit in the same or adjacent
// cells, so their 9 candidate row ranges coincide and neighbour reads are shared in L1/L2.
#include "pisurf_internal.h"

namespace pisurf {

template <class R> struct R4 { typedef typename Vec4<R>::T T; };

__device__ __forceinline__ float4 mk4(float x, float y, float z, float w) { return make_float4(x, y, z, w); }
__device__ __forceinline__ double4 mk4(double x, double y, double z, double w) { return make_double4(x, y, z, w); }

// Pass 1: xMean (compute_xMeans, src/evaluator.cpp:465-512).
template <class R, bool K>
__global__ void __launch_bounds__(256) k_xmean(GridDev g, const float4* __restrict__ spos, int64_t n, AnisoParams ap,
                                               typename R4<R>::T* __restrict__ xm) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float4 pp = spos[p];
    int ix, iy, iz;
    cell_of(g, pp.x, pp.y, pp.z, ix, iy, iz);
    int lo[9], hi[9], cnt = 0;
#pragma unroll
    for (int r = 0; r < 9; r++) {
        if (!row_range(g, ix, iy, iz, r, lo[r], hi[r])) lo[r] = hi[r] = 0;
        cnt += hi[r] - lo[r];
    }
    const R px = pp.x, py = pp.y, pz = pp.z;
    if (cnt <= 2) {  // fewer than 3 candidates: xMean = position (src/evaluator.cpp:672-683)
        xm[p] = mk4(px, py, pz, R(0));
        if (K && sizeof(R) == 4) ((float4*)xm)[n + p] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const R nI = ap.nI, nI2 = ap.nI2, inv_h3 = R(1) / (nI * nI * nI);
    // compensated FP32 sums the offsets x_j - x_i (exact for nearby floats) instead of x_j
    constexpr bool REL = K && sizeof(R) == 4;
    constexpr int kLoadAhead = sizeof(R) == 4 ? 8 : 4;
    Acc<R, K> sx, sy, sz, sw;
    auto cand = [&](const float4 q) {
        const R dx = R(q.x) - px, dy = R(q.y) - py, dz = R(q.z) - pz;
        const R d2 = dx * dx + dy * dy + dz * dz;
        if (d2 >= nI2) return;
        const R w = wij<R>(dsqrt<R>(d2), nI, inv_h3);
        sw.add(w);
        sx.add((REL ? dx : R(q.x)) * w);
        sy.add((REL ? dy : R(q.y)) * w);
        sz.add((REL ? dz : R(q.z)) * w);
    };
    for (int r = 0; r < 9; r++) {
        int j = lo[r];
        // kLoadAhead candidates' positions are loaded before any is used, so the loads overlap
        for (; j + kLoadAhead <= hi[r]; j += kLoadAhead) {
            float4 q[kLoadAhead];
#pragma unroll
            for (int k = 0; k < kLoadAhead; k++) q[k] = spos[j + k];
#pragma unroll
            for (int k = 0; k < kLoadAhead; k++)
                if (j + k != (int)p) cand(q[k]);
        }
        for (; j < hi[r]; j++)
            if (j != (int)p) cand(spos[j]);
    }
    R ox = px, oy = py, oz = pz;
    const R ws = sw.get();
    if (REL) {
        float4 lo4 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!ap.use_poly6 && ws > R(0)) {
            // xMean = x + b (mean - x) with b = 0.9 as a float; kept as hi + lo
            const float b = ap.xmean_delta;
            const float d[3] = {float(sx.get() / ws) * b, float(sy.get() / ws) * b, float(sz.get() / ws) * b};
            const float x[3] = {pp.x, pp.y, pp.z};
            float h[3], l[3];
            for (int a = 0; a < 3; a++) {
                h[a] = __fadd_rn(x[a], d[a]);
                l[a] = __fadd_rn(__fsub_rn(x[a], h[a]), d[a]);
            }
            ox = h[0];
            oy = h[1];
            oz = h[2];
            lo4 = make_float4(l[0], l[1], l[2], 0.f);
        }
        xm[p] = mk4(ox, oy, oz, R(0));
        ((float4*)xm)[n + p] = lo4;
        return;
    }
    if (!ap.use_poly6 && ws > R(0)) {
        const R a = R(1.0f - ap.xmean_delta), b = R(ap.xmean_delta);
        ox = px * a + (sx.get() / ws) * b;
        oy = py * a + (sy.get() / ws) * b;
        oz = pz * a + (sz.get() / ws) * b;
    }
    xm[p] = mk4(ox, oy, oz, R(0));
}

// Pass 2: covariance, eigendecomposition, G (compute_G_Yus, src/evaluator.cpp:585-654;
// compute_G_ours, :514-583, when use_poly6). Writes the field records of device_common.h.
template <class R, bool K>
__global__ void __launch_bounds__(256) k_gmat(GridDev g, const float4* __restrict__ spos, int64_t n, AnisoParams ap,
                                              const typename R4<R>::T* __restrict__ xm,
                                              typename R4<R>::T* __restrict__ rec_a,
                                              typename R4<R>::T* __restrict__ rec_g,
                                              float4* __restrict__ rec_lo, unsigned long long* __restrict__ n_fp64) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const auto xi = xm[p];
    const R h2 = R(ap.h2);
    if (!ap.use_ani) {  // isotropic poly6 support |d|^2 <= h^2
        if (K && sizeof(R) == 4) rec_lo[p] = ((const float4*)xm)[n + p];
        rec_a[p] = mk4(xi.x, xi.y, xi.z, h2 * R(1.001));
        rec_g[2 * p] = mk4(R(1), R(0), R(0), R(1));
        rec_g[2 * p + 1] = mk4(R(0), R(1), R(1), R(0));
        return;
    }
    const float4 pp = spos[p];
    int ix, iy, iz;
    cell_of(g, pp.x, pp.y, pp.z, ix, iy, iz);
    const R nI = ap.nI, nI2 = ap.nI2, inv_h3 = R(1) / (nI * nI * nI);
    constexpr bool REL = K && sizeof(R) == 4;
    const float4 li = REL ? ((const float4*)xm)[n + p] : make_float4(0.f, 0.f, 0.f, 0.f);
    Acc<R, K> c[6], sw;
    const R base = ap.use_poly6 ? R(1.5) : h2;
    auto cov_add = [&](R wx, R wy, R wz, R w) {
        sw.add(w);
        c[0].add(wx * wx * w);
        c[1].add(wx * wy * w);
        c[2].add(wx * wz * w);
        c[3].add(wy * wy * w);
        c[4].add(wy * wz * w);
        c[5].add(wz * wz * w);
    };
    // default weight: wij of |xMean_j - xMean_i|; (hi_j - hi_i) is exact for nearby floats, and
    // compensated FP32 adds the residuals
    auto cand_xm = [&](const typename R4<R>::T xj, const float4 lj) {
        R wx = xj.x - xi.x, wy = xj.y - xi.y, wz = xj.z - xi.z;
        if (REL) {
            wx += R(lj.x - li.x);
            wy += R(lj.y - li.y);
            wz += R(lj.z - li.z);
        }
        const R d2 = wx * wx + wy * wy + wz * wz;
        if (d2 >= nI2) return;
        cov_add(wx, wy, wz, wij<R>(dsqrt<R>(d2), nI, inv_h3));
    };
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    constexpr int kLoadAhead = sizeof(R) == 4 ? 8 : 4;
    for (int r = 0; r < 9; r++) {
        int lo, hi;
        if (!row_range(g, ix, iy, iz, r, lo, hi)) continue;
        int j = lo;
        if (!ap.use_poly6) {
            // kLoadAhead candidates' means are loaded before any is used, so the loads overlap
            for (; j + kLoadAhead <= hi; j += kLoadAhead) {
                typename R4<R>::T x4[kLoadAhead];
                float4 l4[kLoadAhead];
#pragma unroll
                for (int k = 0; k < kLoadAhead; k++) {
                    x4[k] = xm[j + k];
                    l4[k] = REL ? ((const float4*)xm)[n + j + k] : zero4;
                }
#pragma unroll
                for (int k = 0; k < kLoadAhead; k++) cand_xm(x4[k], l4[k]);
            }
            for (; j < hi; j++) cand_xm(xm[j], REL ? ((const float4*)xm)[n + j] : zero4);
            continue;
        }
        for (; j < hi; j++) {
            const auto xj = xm[j];
            R wx = xj.x - xi.x, wy = xj.y - xi.y, wz = xj.z - xi.z;
            if (REL) {
                const float4 lj = ((const float4*)xm)[n + j];
                wx += R(lj.x - li.x);
                wy += R(lj.y - li.y);
                wz += R(lj.z - li.z);
            }
            const float4 q = spos[j];
            const R ex = R(q.x) - R(pp.x), ey = R(q.y) - R(pp.y), ez = R(q.z) - R(pp.z);
            if (ex * ex + ey * ey + ez * ez >= nI2) continue;
            const R fx = R(pp.x) - xj.x, fy = R(pp.y) - xj.y, fz = R(pp.z) - xj.z;
            cov_add(wx, wy, wz, wij<R>(dsqrt<R>(fx * fx + fy * fy + fz * fz), nI, inv_h3));
        }
    }
    const R ws = sw.get();
    R cov[6] = {(c[0].get() + base) / ws, c[1].get() / ws, c[2].get() / ws,
                (c[3].get() + base) / ws, c[4].get() / ws, (c[5].get() + base) / ws};
    const R extra = ap.use_poly6 ? R(0.5) : R(1);
    R G[6], det, gmin;
    bool clamped;
    g_from_cov<R>(cov, extra, G, det, gmin, clamped);
    if (K && sizeof(R) == 4 && clamped) {
        // sheet / filament neighbourhood: redo the eigensolve in FP64
        double cd[6], Gd[6], dd, gm;
        bool cl;
        for (int k = 0; k < 6; k++) cd[k] = double(cov[k]);
        g_from_cov<double>(cd, double(extra), Gd, dd, gm, cl);
        for (int k = 0; k < 6; k++) G[k] = R(Gd[k]);
        det = R(dd);
        gmin = R(gm);
        atomicAdd(n_fp64, 1ull);
    }
    // squared world radius outside which |G d|^2 >= (support)^2: 4 h^2 (B-spline) or h^2 (poly6)
    const R supp = ap.use_poly6 ? h2 : R(4) * h2;
    R R2 = supp / (gmin * gmin) * R(1.001);
    if (R2 != R2) R2 = R(INFINITY);  // 0 / 0 covariance: never culled, its NaN reaches every sum
    rec_a[p] = mk4(xi.x, xi.y, xi.z, R2);
    rec_g[2 * p] = mk4(G[0], G[1], G[2], G[3]);
    rec_g[2 * p + 1] = mk4(G[4], G[5], det, R(0));
    if (REL) rec_lo[p] = li;
}

// ---- PISURF_FP32: the reference's float arithmetic, operation by operation ----------------
// Every operation is rounded on its own (contraction off), divisions and square roots are
// correctly rounded, and the double promotions of pow() inside wij and the SVD's cube root are
// reproduced, so the records equal those of a float restatement of compute_Gs_xMeans that keeps
// the reference's operation order.

// wij (src/evaluator.cpp:751-766): pow(d, 2), pow(h, 3) and pow(h - d, 3) in double from floats;
// h3 = pow(h, 3) comes from the host's libm, like the reference's
__device__ __forceinline__ float ref_wij(float d, float h, double h3) {
#pragma clang fp contract(off)
    if (d < h) {
        const double dd = d;
        if (d < h / 2) return float(1.0 - (6.0 * (dd * dd) / h3 * double(h - d)));
        const double t = double(h - d);
        return float(2.0 / h3 * (t * t * t));
    }
    return 0.0f;
}

// symmetric 3x3 eigendecomposition by cyclic Jacobi rotations (Golub & Van Loan Alg. 8.5.3) in
// float, rotation pairs (0,1), (0,2), (1,2), stopping when the off-diagonal mass is below
// (eps^2 / 100) of the diagonal's; eigenvalues returned in descending order with their vectors
__device__ __forceinline__ void ref_jacobi3(const float a_in[3][3], float lam[3], float Q[3][3]) {
#pragma clang fp contract(off)
    float a[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            a[i][j] = a_in[i][j];
            Q[i][j] = (i == j) ? 1.0f : 0.0f;
        }
    const float eps = 1.1920928955078125e-07f;  // float epsilon
    for (int sweep = 0; sweep < 32; sweep++) {
        const float off = (a[0][1] * a[0][1] + a[0][2] * a[0][2]) + a[1][2] * a[1][2];
        const float dia = (a[0][0] * a[0][0] + a[1][1] * a[1][1]) + a[2][2] * a[2][2];
        if (off <= eps * eps * dia * 0.01f || off == 0.0f) break;
        for (int k = 0; k < 3; k++) {
            const int p = (k == 2) ? 1 : 0, q = (k == 0) ? 1 : 2;
            const float apq = a[p][q];
            if (apq == 0.0f) continue;
            const float tau = (a[q][q] - a[p][p]) / (2.0f * apq);
            const float t = (tau >= 0.0f ? 1.0f : -1.0f) / (fabsf(tau) + __builtin_sqrtf(1.0f + tau * tau));
            const float c = 1.0f / __builtin_sqrtf(1.0f + t * t), sn = t * c;
            for (int r = 0; r < 3; r++) {
                const float arp = a[r][p], arq = a[r][q];
                a[r][p] = c * arp - sn * arq;
                a[r][q] = sn * arp + c * arq;
            }
            for (int r = 0; r < 3; r++) {
                const float apr = a[p][r], aqr = a[q][r];
                a[p][r] = c * apr - sn * aqr;
                a[q][r] = sn * apr + c * aqr;
            }
            a[p][q] = a[q][p] = 0.0f;
            for (int r = 0; r < 3; r++) {
                const float qrp = Q[r][p], qrq = Q[r][q];
                Q[r][p] = c * qrp - sn * qrq;
                Q[r][q] = sn * qrp + c * qrq;
            }
        }
    }
    for (int i = 0; i < 3; i++) lam[i] = a[i][i];
    for (int i = 0; i < 2; i++)
        for (int j = i + 1; j < 3; j++)
            if (lam[j] > lam[i]) {
                const float tl = lam[i];
                lam[i] = lam[j];
                lam[j] = tl;
                for (int r = 0; r < 3; r++) {
                    const float tq = Q[r][i];
                    Q[r][i] = Q[r][j];
                    Q[r][j] = tq;
                }
            }
}

// SVD tail of compute_G_Yus / compute_G_ours (src/evaluator.cpp:627-651, :556-580) on the
// symmetric positive-definite covariance (U = V = Q): |lambda| clamped below at max / 4,
// G = Q diag(1 / w) Q^T * pow(w0 w1 w2, 1/3) (times 0.5 for compute_G_ours), the upper triangle
// computed and mirrored; det by cofactors along the first row
__device__ __forceinline__ void ref_g_from_cov(const float cov[3][3], bool poly6, float G[6], float& det,
                                               float& gmin) {
#pragma clang fp contract(off)
    float lam[3], Q[3][3];
    ref_jacobi3(cov, lam, Q);
    float w[3] = {fabsf(lam[0]), fabsf(lam[1]), fabsf(lam[2])};
    const float mx = (w[0] < ((w[1] < w[2]) ? w[2] : w[1])) ? ((w[1] < w[2]) ? w[2] : w[1]) : w[0];
    for (int k = 0; k < 3; k++) w[k] = (w[k] < mx / 4.0f) ? mx / 4.0f : w[k];
    const float inv[3] = {1.0f / w[0], 1.0f / w[1], 1.0f / w[2]};
    const float scale = float(pow(double((w[0] * w[1]) * w[2]), 1.0 / 3.0));
    float M[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++) {
            const float v = ((Q[a][0] * inv[0]) * Q[b][0] + (Q[a][1] * inv[1]) * Q[b][1]) + (Q[a][2] * inv[2]) * Q[b][2];
            M[a][b] = v * scale;
            if (poly6) M[a][b] = M[a][b] * 0.5f;
            M[b][a] = M[a][b];
        }
    G[0] = M[0][0];
    G[1] = M[0][1];
    G[2] = M[0][2];
    G[3] = M[1][1];
    G[4] = M[1][2];
    G[5] = M[2][2];
    det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
          M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    gmin = float(double(scale) / double(mx) * (poly6 ? 0.5 : 1.0));
}

// Pass 1 in the reference's float order: x_j - x_i, |.|^2, sqrt, wij, sum w, sum x_j w, divide,
// blend x (1 - 0.9f) + mean 0.9f
__global__ void __launch_bounds__(256) k_xmean_ref(GridDev g, const float4* __restrict__ spos, int64_t n,
                                                   AnisoParams ap, float4* __restrict__ xm) {
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float4 pp = spos[p];
    int ix, iy, iz;
    cell_of(g, pp.x, pp.y, pp.z, ix, iy, iz);
    int lo[9], hi[9], cnt = 0;
#pragma unroll
    for (int r = 0; r < 9; r++) {
        if (!row_range(g, ix, iy, iz, r, lo[r], hi[r])) lo[r] = hi[r] = 0;
        cnt += hi[r] - lo[r];
    }
    if (cnt <= 2 || ap.use_poly6) {  // fewer than 3 candidates, or no smoothing in poly6 mode
        xm[p] = make_float4(pp.x, pp.y, pp.z, 0.0f);
        return;
    }
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, ws = 0.0f;
    for (int r = 0; r < 9; r++)
        for (int j = lo[r]; j < hi[r]; j++) {
            if (j == (int)p) continue;
            const float4 q = spos[j];
            const float dx = q.x - pp.x, dy = q.y - pp.y, dz = q.z - pp.z;
            const float d2 = (dx * dx + dy * dy) + dz * dz;
            if (d2 >= ap.nI2) continue;
            const float w = ref_wij(__builtin_sqrtf(d2), ap.nI, ap.nI3);
            ws = ws + w;
            sx = sx + q.x * w;
            sy = sy + q.y * w;
            sz = sz + q.z * w;
        }
    float ox = pp.x, oy = pp.y, oz = pp.z;
    if (ws > 0.0f) {
        const float a = 1.0f - ap.xmean_delta, b = ap.xmean_delta;
        sx = sx / ws;
        sy = sy / ws;
        sz = sz / ws;
        ox = pp.x * a + sx * b;
        oy = pp.y * a + sy * b;
        oz = pp.z * a + sz * b;
    }
    xm[p] = make_float4(ox, oy, oz, 0.0f);
}

// Pass 2 in the reference's float order: the covariance starts at diag(h^2) (compute_G_Yus) or
// diag(1 + 0.5) (compute_G_ours), each candidate adds (wd_a wd_b) w to all nine entries, then
// every entry is divided by the weight sum
__global__ void __launch_bounds__(256) k_gmat_ref(GridDev g, const float4* __restrict__ spos, int64_t n,
                                                  AnisoParams ap, const float4* __restrict__ xm,
                                                  float4* __restrict__ rec_a, float4* __restrict__ rec_g) {
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float4 xi = xm[p];
    const float h2 = float(ap.h2);
    if (!ap.use_ani) {  // isotropic poly6 support |d|^2 <= h^2
        rec_a[p] = make_float4(xi.x, xi.y, xi.z, h2 * 1.001f);
        rec_g[2 * p] = make_float4(1.0f, 0.0f, 0.0f, 1.0f);
        rec_g[2 * p + 1] = make_float4(0.0f, 1.0f, 1.0f, 0.0f);
        return;
    }
    const float4 pp = spos[p];
    int ix, iy, iz;
    cell_of(g, pp.x, pp.y, pp.z, ix, iy, iz);
    const bool ours = ap.use_poly6 != 0;
    const float diag = ours ? 1.0f + 0.5f : h2;
    float cov[3][3] = {{diag, 0.0f, 0.0f}, {0.0f, diag, 0.0f}, {0.0f, 0.0f, diag}};
    float ws = 0.0f;
    for (int r = 0; r < 9; r++) {
        int lo, hi;
        if (!row_range(g, ix, iy, iz, r, lo, hi)) continue;
        for (int j = lo; j < hi; j++) {
            const float4 xj = xm[j];
            float d;
            if (ours) {
                const float4 q = spos[j];
                const float ex = q.x - pp.x, ey = q.y - pp.y, ez = q.z - pp.z;
                if ((ex * ex + ey * ey) + ez * ez >= ap.nI2) continue;
                const float fx = pp.x - xj.x, fy = pp.y - xj.y, fz = pp.z - xj.z;
                d = __builtin_sqrtf((fx * fx + fy * fy) + fz * fz);
            } else {
                const float ex = xj.x - xi.x, ey = xj.y - xi.y, ez = xj.z - xi.z;
                const float d2 = (ex * ex + ey * ey) + ez * ez;
                if (d2 >= ap.nI2) continue;
                d = __builtin_sqrtf(d2);
            }
            const float w = ref_wij(d, ap.nI, ap.nI3);
            ws = ws + w;
            const float wd[3] = {xj.x - xi.x, xj.y - xi.y, xj.z - xi.z};
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) cov[a][b] = cov[a][b] + (wd[a] * wd[b]) * w;
        }
    }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) cov[a][b] = cov[a][b] / ws;
    float G[6], det, gmin;
    ref_g_from_cov(cov, ours, G, det, gmin);
    // squared world radius outside which |G d|^2 >= (support)^2: 4 h^2 (B-spline) or h^2 (poly6)
    const double supp = ours ? double(h2) : 4.0 * double(h2);
    float R2 = float(supp / (double(gmin) * double(gmin)) * 1.001);
    if (R2 != R2) R2 = INFINITY;  // 0 / 0 covariance: never culled, its NaN reaches every sum
    rec_a[p] = make_float4(xi.x, xi.y, xi.z, R2);
    rec_g[2 * p] = make_float4(G[0], G[1], G[2], G[3]);
    rec_g[2 * p + 1] = make_float4(G[4], G[5], det, 0.0f);
}

void launch_xmean(int prec, GridDev g, const float4* spos, int64_t n, int64_t, AnisoParams ap, void* xm,
                  hipStream_t s) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (prec == 2) k_xmean<double, false><<<nb, 256, 0, s>>>(g, spos, n, ap, (double4*)xm);
    else if (prec == 1) k_xmean<float, true><<<nb, 256, 0, s>>>(g, spos, n, ap, (float4*)xm);
    else k_xmean_ref<<<nb, 256, 0, s>>>(g, spos, n, ap, (float4*)xm);
}

void launch_gmat(int prec, GridDev g, const float4* spos, int64_t n, int64_t, AnisoParams ap, const void* xm,
                 void* rec_a, void* rec_g, void* rec_lo, unsigned long long* n_fp64, hipStream_t s) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (prec == 2)
        k_gmat<double, false><<<nb, 256, 0, s>>>(g, spos, n, ap, (const double4*)xm, (double4*)rec_a, (double4*)rec_g,
                                                 nullptr, n_fp64);
    else if (prec == 1)
        k_gmat<float, true><<<nb, 256, 0, s>>>(g, spos, n, ap, (const float4*)xm, (float4*)rec_a, (float4*)rec_g,
                                              (float4*)rec_lo, n_fp64);
    else
        k_gmat_ref<<<nb, 256, 0, s>>>(g, spos, n, ap, (const float4*)xm, (float4*)rec_a, (float4*)rec_g);
}

}  // namespace pisurf
