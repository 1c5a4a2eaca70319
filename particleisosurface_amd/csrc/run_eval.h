// Neighbour sums for runs of query points that share a hash-cell row, with the candidates staged
// through LDS.
//
// A run is a set of consecutive points (cell-major order) whose cells lie in one (y, z) row of
// the hash grid, with x in [xa, xb]. For each of the nine neighbour rows (z-outer, then y, as in
// GetPIdxList) the union of the run's windows, cells [xa - 1, xb + 1], is one contiguous range of
// sorted particles; the nine ranges are concatenated and staged through LDS in chunks. Each thread
// then walks its own point's sub-ranges: cells [x - 1, x + 1] of every row, particles in sorted
// order. Every point therefore sees exactly the candidate sequence of the per-point kernels, and a
// policy that does the per-point arithmetic gives bit-identical sums.
#pragma once
#include "pisurf_internal.h"

namespace pisurf {

constexpr int kRunThreads = 128;  // block size of the particle runs (normals)
constexpr int kRunChunk = 256;
constexpr int kCullStep = 4;     // candidates tested per cull call (the staging arrays are padded by it)
constexpr int kCullGroup = 32;   // candidates per pass mask

// Policy interface of run_visit:
//   typename Pol::Lds     staging buffers for kRunChunk (+ kCullStep padding) candidates
//   typename Pol::State   per-point accumulators
//   void point(int h, float p[3]) const                 query point h
//   void stage(Lds&, int t, int j) const                candidate j -> slot t
//   void begin(State&, int h, const float p[3]) const   (inactive threads too, with h = -1)
//   unsigned cull<N>(const State&, const Lds&, int t) const
//                                                       bit k set: slot t + k can contribute (slots
//                                                       past the row are dropped by the caller)
//   void take(State&, const Lds&, int t) const          contribution of slot t (one that passed)
//   void end(State&, int h, int ncand) const            active points: ncand = candidates in its
//                                                       27-cell window
// T is the block size. Returns the candidates visited by this thread's points.
//
// Candidates are culled in groups of up to kCullGroup slots of one row, and only the ones that
// pass are then summed, in slot order. The lanes of a wave hold different points, so a per-
// candidate branch would run the contribution whenever any lane needs it, which is nearly always
// even though most candidates lie outside the kernel support; the group's pass mask lets each lane
// step through its own passing candidates, and the wave runs as many contribution steps as its
// busiest lane needs. Every point still adds exactly its passing candidates in the reference's
// order, so the sums are the same as with one test per candidate.
template <int T, class Pol>
__device__ __forceinline__ unsigned long long run_visit(const GridDev& g, const Pol& pol, int h0, int h1,
                                                        typename Pol::Lds& lds) {
    float pa[3], pb[3];
    pol.point(h0, pa);
    pol.point(h1 - 1, pb);
    int xa, ya, za, xb, yb, zb;
    cell_of(g, pa[0], pa[1], pa[2], xa, ya, za);
    cell_of(g, pb[0], pb[1], pb[2], xb, yb, zb);
    if (xb < xa) xb = xa;  // only in the group of points with empty windows
    // the run's nine union row windows, concatenated in row order: row r starts at sorted position
    // w0[r] and at concatenated offset roff[r]
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
    unsigned long long vis = 0;
    for (int hb = h0; hb < h1; hb += T) {
        const int h = hb + (int)threadIdx.x;
        const bool act = h < h1;
        float p[3] = {0.f, 0.f, 0.f};
        // this point's windows as concatenated ranges (sub-ranges of the run's rows)
        int clo[9], chi[9], ncand = 0;
#pragma unroll
        for (int r = 0; r < 9; r++) clo[r] = chi[r] = 0;
        if (act) {
            pol.point(h, p);
            if (one_cell) {
#pragma unroll
                for (int r = 0; r < 9; r++) {
                    clo[r] = roff[r];
                    chi[r] = roff[r + 1];
                }
                ncand = total;
            } else {
                int ix, iy, iz;
                cell_of(g, p[0], p[1], p[2], ix, iy, iz);
#pragma unroll
                for (int r = 0; r < 9; r++) {
                    int lo, hi;
                    if (row_range(g, ix, iy, iz, r, lo, hi)) {
                        clo[r] = roff[r] + (lo - w0[r]);
                        chi[r] = roff[r] + (hi - w0[r]);
                        ncand += hi - lo;
                    }
                }
            }
        }
        vis += (unsigned long long)ncand;
        typename Pol::State st;
        pol.begin(st, act ? h : -1, p);
        for (int cb = 0; cb < total; cb += kRunChunk) {
            const int n = min(kRunChunk, total - cb);
            if (hb > h0 || cb > 0) __syncthreads();  // the previous chunk has been consumed
            if (!(hb > h0 && total <= kRunChunk)) {    // a single chunk stays staged for every point
                for (int t = threadIdx.x; t < n; t += T) {
                    const int c = cb + t;
                    int j = 0;
#pragma unroll
                    for (int r = 0; r < 9; r++)  // the row holding concatenated position c
                        if (c >= roff[r] && c < roff[r + 1]) j = w0[r] + (c - roff[r]);
                    pol.stage(lds, t, j);
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 9; r++) {
                const int t0 = max(clo[r], cb) - cb, t1 = min(chi[r], cb + n) - cb;
                for (int t = t0; t < t1; t += kCullGroup) {
                    const int tn = min(t1, t + kCullGroup);
                    unsigned m = 0;
                    for (int u = t; u < tn; u += kCullStep) m |= pol.template cull<kCullStep>(st, lds, u) << (u - t);
                    // slots from tn on (other rows, or not staged) are tested too and dropped here
                    if (tn - t < 32) m &= (1u << (tn - t)) - 1u;
                    while (m) {
                        const int k = __builtin_ctz(m);
                        m &= m - 1;
                        pol.take(st, lds, t + k);
                    }
                }
            }
        }
        if (act) pol.end(st, h, ncand);
    }
    return vis;
}

// ---- field sums (SingleEval / SingleEvalWithGrad): records of device_common.h ----
template <class R, bool LO> struct RunLds {
    typename Vec4<R>::T a[kRunChunk + kCullStep];  // the padding is read by a cull step at a row's end
    typename Vec4<R>::T g0[kRunChunk], g1[kRunChunk];
    float4 lo[LO ? kRunChunk : 1];
};

// Src: point(h, float p[3]) gives point h; store(h, sum, grad) receives its sums (without sigma).
// MODE: kernel choice of field_contrib.
template <class R, bool K, bool GRAD, int MODE, class Src> struct FieldPol {
    static constexpr bool LO = K && sizeof(R) == 4;
    typedef RunLds<R, LO> Lds;
    struct State {
        R px, py, pz;
        Acc<R, K> acc, ga[3];
    };
    Recs<R> rec;
    FieldConst fc;
    R inv_h2, h2;
    Src src;
    __device__ __forceinline__ void point(int h, float p[3]) const { src.point(h, p); }
    __device__ __forceinline__ void stage(Lds& l, int t, int j) const {
        l.a[t] = rec.a[j];
        l.g0[t] = rec.g[2 * j];
        l.g1[t] = rec.g[2 * j + 1];
        if (LO) l.lo[t] = fc.lo[j];
    }
    __device__ __forceinline__ void begin(State& s, int, const float p[3]) const {
        s.px = R(p[0]);
        s.py = R(p[1]);
        s.pz = R(p[2]);
    }
    // The cull is written !(d2 >= R^2) so that a NaN record passes and poisons the sum, as the
    // reference's kernels do for it. The N positions are read before any is used, so their LDS
    // reads overlap.
    template <int N> __device__ __forceinline__ unsigned cull(const State& s, const Lds& l, int t) const {
        typename Vec4<R>::T a[N];
#pragma unroll
        for (int k = 0; k < N; k++) a[k] = l.a[t + k];
        unsigned m = 0;
#pragma unroll
        for (int k = 0; k < N; k++) {
            const R dx = s.px - a[k].x, dy = s.py - a[k].y, dz = s.pz - a[k].z;
            m |= unsigned(!(dist2(dx, dy, dz) >= a[k].w)) << k;
        }
        return m;
    }
    __device__ __forceinline__ void take(State& s, const Lds& l, int t) const {
        const typename Vec4<R>::T a = l.a[t];
        const R dx = s.px - a.x, dy = s.py - a.y, dz = s.pz - a.z;
        const R d2 = dist2(dx, dy, dz);
        const float4 lo = LO ? l.lo[LO ? t : 0] : make_float4(0.f, 0.f, 0.f, 0.f);
        field_contrib<R, K, GRAD, MODE>(fc, inv_h2, h2, dx, dy, dz, d2, l.g0[t], l.g1[t], lo, s.acc, s.ga);
    }
    __device__ __forceinline__ void end(State& s, int h, int) const {
        R gr[3] = {s.ga[0].get(), s.ga[1].get(), s.ga[2].get()};
        src.store(h, s.acc.get(), gr);
    }
};

template <int T, class R, bool K, bool GRAD, int MODE, class Src>
__device__ __forceinline__ unsigned long long run_eval(const GridDev& g, const Recs<R>& rec, const FieldConst& fc,
                                                       const Src& src, int h0, int h1,
                                                       RunLds<R, K && sizeof(R) == 4>& lds) {
    const FieldPol<R, K, GRAD, MODE, Src> pol{rec, fc, R(fc.inv_h2), R(fc.h2), src};
    return run_visit<T>(g, pol, h0, h1, lds);
}

// run id of a hash cell (extended by one cell on every side): its (y, z) row and x / kx
__device__ __forceinline__ uint64_t run_of_cell(const GridDev& g, int ix, int iy, int iz, int kx) {
    const uint64_t ex = g.n[0] + 2, ey = g.n[1] + 2, ez = g.n[2] + 2;
    const uint64_t per_row = (ex + kx - 1) / kx;
    if (ix < -1 || ix > g.n[0] || iy < -1 || iy > g.n[1] || iz < -1 || iz > g.n[2]) return ey * ez * per_row;
    return ((uint64_t)(iz + 1) * ey + (uint64_t)(iy + 1)) * per_row + (uint64_t)(ix + 1) / kx;
}

// particle runs: block b covers the particles of cells [x0, x0 + kx) of one grid row (contiguous
// in sorted order)
__device__ __forceinline__ void particle_run(const GridDev& g, int kx, int& h0, int& h1) {
    const int per_row = (g.n[0] + kx - 1) / kx;
    const int row = blockIdx.x / per_row, x0 = (blockIdx.x % per_row) * kx;
    const int base = row * g.n[0];
    h0 = g.off[base + x0];
    h1 = g.off[base + min(x0 + kx, g.n[0])];
}

}  // namespace pisurf
