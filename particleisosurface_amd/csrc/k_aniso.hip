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

void launch_xmean(int prec, GridDev g, const float4* spos, int64_t n, int64_t, AnisoParams ap, void* xm,
                  hipStream_t s) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (prec == 2) k_xmean<double, false><<<nb, 256, 0, s>>>(g, spos, n, ap, (double4*)xm);
    else if (prec == 1) k_xmean<float, true><<<nb, 256, 0, s>>>(g, spos, n, ap, (float4*)xm);
    else k_xmean<float, false><<<nb, 256, 0, s>>>(g, spos, n, ap, (float4*)xm);
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
        k_gmat<float, false><<<nb, 256, 0, s>>>(g, spos, n, ap, (const float4*)xm, (float4*)rec_a, (float4*)rec_g,
                                              (float4*)rec_lo, n_fp64);
}

}  // namespace pisurf
