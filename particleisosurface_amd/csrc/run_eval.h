// Field sums for runs of query points that share a hash-cell row, with the candidates staged
// through LDS.
//
// A run is a set of consecutive points (cell-major order) whose cells lie in one (y, z) row of
// the hash grid, with x in [xa, xb]. For each of the nine neighbour rows (z-outer, then y, as in
// GetPIdxList) the union of the run's windows, cells [xa - 1, xb + 1], is one contiguous range of
// sorted particles. It is staged through LDS in chunks. Each thread then walks its own point's
// sub-range: cells [x - 1, x + 1], particles in sorted order. Every point therefore sees exactly
// the candidate sequence of field_sum, with the same per-candidate arithmetic (dist2,
// field_contrib), and its sum is bit-identical to the per-point kernels.
#pragma once
#include "pisurf_internal.h"

namespace pisurf {

constexpr int kRunThreads = 128;
constexpr int kRunChunk = 256;

template <class R, bool LO> struct RunLds {
    typename Vec4<R>::T a[kRunChunk], g0[kRunChunk], g1[kRunChunk];
    float4 lo[LO ? kRunChunk : 1];
};

// Src: point(h, float p[3]) gives point h; store(h, sum, grad) receives its sums (without sigma).
// *visited receives the candidates visited by this thread's points.
template <class R, bool K, bool GRAD, class Src>
__device__ __forceinline__ void run_eval(const GridDev& g, const Recs<R>& rec, const FieldConst& fc, const Src& src,
                                         int h0, int h1, RunLds<R, K && sizeof(R) == 4>& lds,
                                         unsigned long long* visited) {
    typedef typename Vec4<R>::T V4;
    constexpr bool LO = K && sizeof(R) == 4;
    float pa[3], pb[3];
    src.point(h0, pa);
    src.point(h1 - 1, pb);
    int xa, ya, za, xb, yb, zb;
    cell_of(g, pa[0], pa[1], pa[2], xa, ya, za);
    cell_of(g, pb[0], pb[1], pb[2], xb, yb, zb);
    if (xb < xa) xb = xa;  // only in the group of points with empty windows
    // the run's nine union row windows, concatenated in row order: row r is [w0[r], w0[r] + len)
    // at concatenated offset roff[r]
    int w0[9], roff[10];
    roff[0] = 0;
#pragma unroll
    for (int r = 0; r < 9; r++) {
        const int yy = ya + (r % 3) - 1, zz = za + (r / 3) - 1;
        const int ux0 = max(xa - 1, 0), ux1 = min(xb + 1, g.n[0] - 1);
        int a = 0, b = 0;
        if (yy >= 0 && yy < g.n[1] && zz >= 0 && zz < g.n[2] && ux0 <= ux1) {
            const int base = (zz * g.n[1] + yy) * g.n[0];
            a = g.off[base + ux0];
            b = g.off[base + ux1 + 1];
        }
        w0[r] = a;
        roff[r + 1] = roff[r] + (b - a);
    }
    const int total = roff[9];
    // a run of one cell (not the empty-window group): every point's windows are the run's rows
    const bool one_cell = xa == xb && xa >= -1 && xa <= g.n[0] && ya >= -1 && ya <= g.n[1] && za >= -1 &&
                          za <= g.n[2];
    const R inv_h2 = R(fc.inv_h2), h2 = R(fc.h2);
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    unsigned long long vis = 0;
    for (int hb = h0; hb < h1; hb += kRunThreads) {
        const int h = hb + (int)threadIdx.x;
        const bool act = h < h1;
        float p[3] = {0.f, 0.f, 0.f};
        // this point's windows as concatenated ranges (sub-ranges of the run's rows)
        int clo[9], chi[9];
#pragma unroll
        for (int r = 0; r < 9; r++) clo[r] = chi[r] = 0;
        if (act) {
            src.point(h, p);
            if (one_cell) {
                vis += (unsigned long long)total;
            } else {
                int ix, iy, iz;
                cell_of(g, p[0], p[1], p[2], ix, iy, iz);
#pragma unroll
                for (int r = 0; r < 9; r++) {
                    int lo, hi;
                    if (row_range(g, ix, iy, iz, r, lo, hi)) {
                        clo[r] = roff[r] + (lo - w0[r]);
                        chi[r] = roff[r] + (hi - w0[r]);
                        vis += (unsigned long long)(hi - lo);
                    }
                }
            }
        }
        const R px = R(p[0]), py = R(p[1]), pz = R(p[2]);
        Acc<R, K> acc, ga[3];
        for (int cb = 0; cb < total; cb += kRunChunk) {
            const int n = min(kRunChunk, total - cb);
            if (hb > h0 || cb > 0) __syncthreads();  // the previous chunk has been consumed
            if (!(hb > h0 && total <= kRunChunk)) {    // a single chunk stays staged for every point
                for (int t = threadIdx.x; t < n; t += kRunThreads) {
                    const int c = cb + t;
                    int j = 0;
#pragma unroll
                    for (int r = 0; r < 9; r++)  // the row holding concatenated position c
                        if (c >= roff[r] && c < roff[r + 1]) j = w0[r] + (c - roff[r]);
                    lds.a[t] = rec.a[j];
                    lds.g0[t] = rec.g[2 * j];
                    lds.g1[t] = rec.g[2 * j + 1];
                    if (LO) lds.lo[t] = fc.lo[j];
                }
            }
            __syncthreads();
            auto visit = [&](int t0, int t1) {
                for (int t = t0; t < t1; t++) {
                    const V4 a = lds.a[t];
                    const R dx = px - a.x, dy = py - a.y, dz = pz - a.z;
                    const R d2 = dist2(dx, dy, dz);
                    if (d2 < a.w)
                        field_contrib<R, K, GRAD>(fc, inv_h2, h2, dx, dy, dz, d2, lds.g0[t], lds.g1[t],
                                                  LO ? lds.lo[LO ? t : 0] : zero4, acc, ga);
                }
            };
            if (one_cell) {
                if (act) visit(0, n);  // the windows are the whole concatenated sequence
            } else {
#pragma unroll
                for (int r = 0; r < 9; r++) visit(max(clo[r], cb) - cb, min(chi[r], cb + n) - cb);
            }
        }
        if (act) {
            R gr[3] = {ga[0].get(), ga[1].get(), ga[2].get()};
            src.store(h, acc.get(), gr);
        }
    }
    *visited = vis;
}

// run id of a hash cell (extended by one cell on every side): its (y, z) row and x / kx
__device__ __forceinline__ uint64_t run_of_cell(const GridDev& g, int ix, int iy, int iz, int kx) {
    const uint64_t ex = g.n[0] + 2, ey = g.n[1] + 2, ez = g.n[2] + 2;
    const uint64_t per_row = (ex + kx - 1) / kx;
    if (ix < -1 || ix > g.n[0] || iy < -1 || iy > g.n[1] || iz < -1 || iz > g.n[2]) return ey * ez * per_row;
    return ((uint64_t)(iz + 1) * ey + (uint64_t)(iy + 1)) * per_row + (uint64_t)(ix + 1) / kx;
}

}  // namespace pisurf
