// Device-side building blocks shared by the pisurf kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace pisurf {

constexpr int kWave = 64;

// Uniform grid of HashGrid (src/hash_grid.cpp:4-34): cell (x,y,z) has key z*nx*ny + y*nx + x.
// off[] is the CSR offset table over the key-sorted particles of valid keys (ncell + 1 entries):
// the reference's window for a cell is sorted positions [off[c], off[c+1]) of the full sorted
// list, including the n_invalid particles with key -1 that sort first (src/hash_grid.cpp:103-122).
struct GridDev {
    float b0[3];
    float cell;
    int n[3];
    int ncell;
    const int* off;
};

// HashGrid::CalcXYZIdx (src/hash_grid.cpp:124-129): int((p - B0) / CellSize), truncation toward
// zero, correctly rounded float ops. Points far outside the box get a sentinel that yields no cell.
__device__ __forceinline__ int cell_coord(float p, float b0, float cell, int n) {
    float v = __fdiv_rn(__fsub_rn(p, b0), cell);
    if (!(v > -8.0f && v < float(n) + 8.0f)) return (v > 0.0f) ? n + 8 : -8;
    return (int)v;
}

__device__ __forceinline__ void cell_of(const GridDev& g, float x, float y, float z, int& ix, int& iy, int& iz) {
    ix = cell_coord(x, g.b0[0], g.cell, g.n[0]);
    iy = cell_coord(y, g.b0[1], g.cell, g.n[1]);
    iz = cell_coord(z, g.b0[2], g.cell, g.n[2]);
}

// Row r (0..8, z-outer then y) of the 27-cell neighbourhood of cell (ix,iy,iz) as a contiguous
// range of sorted positions: the three x-neighbours of one (y,z) row have consecutive keys, so
// their windows are adjacent (GetPIdxList order, src/hash_grid.cpp:196-214).
__device__ __forceinline__ bool row_range(const GridDev& g, int ix, int iy, int iz, int r, int& lo, int& hi) {
    const int y = iy + (r % 3) - 1, z = iz + (r / 3) - 1;
    if (y < 0 || y >= g.n[1] || z < 0 || z >= g.n[2]) return false;
    const int x0 = max(ix - 1, 0), x1 = min(ix + 1, g.n[0] - 1);
    if (x0 > x1) return false;
    const int base = (z * g.n[1] + y) * g.n[0];
    lo = g.off[base + x0];
    hi = g.off[base + x1 + 1];
    return lo < hi;
}

// ---- scalar kernels ----------------------------------------------------------------------
// wij (src/evaluator.cpp:751-766): cubic spline with w(0) = 1 and support h.
template <class R> __device__ __forceinline__ R wij(R d, R h, R inv_h3) {
    if (!(d < h)) return R(0);
    if (d < h * R(0.5)) return R(1) - R(6) * d * d * (h - d) * inv_h3;
    const R t = h - d;
    return R(2) * t * t * t * inv_h3;
}

// Cubic B-spline shape without sigma (src/evaluator.cpp:383-396), argument q^2.
template <class R> __device__ __forceinline__ R bspline_q2(R q2);
template <> __device__ __forceinline__ float bspline_q2<float>(float q2) {
    // both pieces, then a select: the same values as branching on q, without divergence
    const float q = __builtin_amdgcn_sqrtf(q2);
    const float inner = 1.0f + q2 * (-1.5f + 0.75f * q);
    const float t = 2.0f - q;
    const float outer = 0.25f * t * t * t;
    return (q2 < 4.0f) ? ((q <= 1.0f) ? inner : outer) : 0.0f;
}
template <> __device__ __forceinline__ double bspline_q2<double>(double q2) {
    if (!(q2 < 4.0)) return 0.0;
    const double q = sqrt(q2);
    if (q <= 1.0) return 1.0 + q2 * (-1.5 + 0.75 * q);
    const double t = 2.0 - q;
    return 0.25 * t * t * t;
}

template <class R> __device__ __forceinline__ R dsqrt(R x);
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <> __device__ __forceinline__ double dsqrt<double>(double x) { return sqrt(x); }

// inclusive prefix sum over the 64 lanes of a wave
__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

// Sum accumulator. The compensated FP32 mode (K) adds its float terms into a double: on gfx950 an
// unpacked FP32 add and an FP64 add issue at the same rate, so a conversion and one FP64 add
// replace the four float operations of a Kahan step, and the sum is at least as accurate. K =
// false: a plain sum in R.
template <class R, bool K> struct Acc {
    typedef typename std::conditional<K && sizeof(R) == 4, double, R>::type S;
    S s = S(0);
    __device__ __forceinline__ void add(R v) { s += S(v); }
    // the same, never fused with the product that made v: the field sums of every kernel then
    // round each contribution the same way, whatever code surrounds them
    __device__ __forceinline__ void add_unfused(R v) {
#pragma clang fp contract(off)
        s += S(v);
    }
    __device__ __forceinline__ R get() const { return R(s); }
};

// ---- per-particle field records (sorted order) ---------------------------------------------
// a[j] = (xMean.x, xMean.y, xMean.z, R2): R2 is a squared world-space radius beyond which the
//        particle's kernel and gradient are exactly zero (|G d| >= sigma_min(G)|d|);
// g[2j], g[2j+1] = (G00, G01, G02, G11), (G12, G22, det, 0)  (G is symmetric).
template <class R> struct Vec4;
template <> struct Vec4<float> { typedef float4 T; };
template <> struct Vec4<double> { typedef double4 T; };

template <class R> struct Recs {
    const typename Vec4<R>::T* a;
    const typename Vec4<R>::T* g;
};

struct FieldConst {
    double inv_h2;      // 1 / Influnce2
    double h2;          // Influnce2
    double sigma;       // Sigma
    double iso;         // iso value
    int use_ani;
    int use_poly6;
    // compensated FP32 only: xMean = a.xyz + lo[j].xyz, the float rounding residual of xMean, so
    // that q - xMean keeps its accuracy when coordinates are large compared with r
    const float4* lo;
};

// Square roots that parity depends on are written __builtin_sqrtf: on gfx950 it lowers to the
// hardware square root followed by a one-ulp correction (correctly rounded), while __fsqrt_rn
// lowers to the bare, scaled v_sqrt_f32.
//
// FP32 without compensation (PISURF_FP32) is the reference's float arithmetic, term by term:
// Bspline_kernel / poly6_kernel return sigma times the shape, so each term carries sigma and the
// sums are finished without it (out_sigma = 1); the other modes sum shapes and scale once.
template <class R, bool K> __host__ __device__ constexpr bool ref_float() { return sizeof(R) == 4 && !K; }
template <class R, bool K> __device__ __forceinline__ R out_sigma(const FieldConst& fc) {
    return ref_float<R, K>() ? R(1) : R(fc.sigma);
}

// Bspline_kernel (src/evaluator.cpp:383-396) as the reference evaluates it: the float ratio
// promoted to double by pow (pow(q, 2) and pow(q, 3) are q*q and (q*q)*q rounded once, q*q being
// exact for a float q), 2 - ratio taken in float, the result narrowed to float
__device__ __forceinline__ float ref_bspline(float ratio, float sigma) {
#pragma clang fp contract(off)
    if (ratio < 2.0f) {
        if (ratio <= 1.0f) {
            const double q = ratio;
            return float(double(sigma) * (1.0 - 1.5 * (q * q) + 0.75 * (q * q * q)));
        }
        const double t = double(2.0f - ratio);
        return float(double(sigma) * (0.25 * (t * t * t)));
    }
    return 0.0f;
}
// poly6_kernel (src/evaluator.cpp:377-381): (h2 - d2)^3 in double, narrowed, times sigma in float
__device__ __forceinline__ float ref_poly6(float d2, float h2, float sigma) {
#pragma clang fp contract(off)
    const double t = double(h2 - d2);
    const float p = (d2 > h2) ? 0.0f : float(t * t * t);
    return p * sigma;
}

// squared distance of a candidate offset (one expression shared by every field kernel, so the
// compiler forms the same operations everywhere)
template <class R> __device__ __forceinline__ R dist2(R dx, R dy, R dz) { return dx * dx + dy * dy + dz * dz; }

// kernel choice of the field sums: read from FieldConst at run time, or fixed at compile time to
// the default anisotropic B-spline (use_ani = 1, use_poly6 = 0)
constexpr int kModeAny = 0, kModeAniBspline = 1;
__host__ __device__ inline int field_mode(const FieldConst& fc) {
    return (fc.use_ani && !fc.use_poly6) ? kModeAniBspline : kModeAny;
}

// contribution of one candidate that passed the cheap reject test: offset (dx, dy, dz) = q - a.xyz,
// records g0 / g1 and, in compensated FP32, the residual lo of the particle mean
// One term of the reference's float sums (AnisotropicInterpolate / IsotropicInterpolate and their
// gradients, src/evaluator.cpp:408-463), every operation rounded on its own in the reference's
// order: G d row by row, |G d|^2, sqrt(|G d|^2 / h^2) for the B-spline, sigma inside the kernel,
// det times the kernel; the poly6-form gradient sigma (-6 G d) (h^2 - |G d|^2)^2 times det.
template <bool GRAD>
__device__ __forceinline__ void ref_float_contrib(bool ani, bool poly6, float h2, float sigma, float dx, float dy,
                                                  float dz, const float4& g0, const float4& g1, Acc<float, false>& acc,
                                                  Acc<float, false>* ga) {
#pragma clang fp contract(off)
    float gx = dx, gy = dy, gz = dz;
    if (ani) {
        gx = (g0.x * dx + g0.y * dy) + g0.z * dz;
        gy = (g0.y * dx + g0.w * dy) + g1.x * dz;
        gz = (g0.z * dx + g1.x * dy) + g1.y * dz;
    }
    const float n2 = (gx * gx + gy * gy) + gz * gz;
    const float k = (poly6 || !ani) ? ref_poly6(n2, h2, sigma) : ref_bspline(__builtin_sqrtf(__fdiv_rn(n2, h2)), sigma);
    acc.s = acc.s + (ani ? g1.z * k : k);
    if (GRAD) {
        const float w = (n2 > h2) ? 0.0f : (h2 - n2) * (h2 - n2);
        const float t[3] = {(sigma * (-6.0f * gx)) * w, (sigma * (-6.0f * gy)) * w, (sigma * (-6.0f * gz)) * w};
#pragma unroll
        for (int a = 0; a < 3; a++) ga[a].s = ga[a].s + (ani ? g1.z * t[a] : t[a]);
    }
}

template <class R, bool KAHAN, bool GRAD, int MODE = kModeAny>
__device__ __forceinline__ void field_contrib(const FieldConst& fc, R inv_h2, R h2, R dx, R dy, R dz, R d2,
                                              const typename Vec4<R>::T& g0, const typename Vec4<R>::T& g1,
                                              const float4& lo, Acc<R, KAHAN>& acc, Acc<R, KAHAN>* ga) {
    const bool ani = MODE == kModeAniBspline || fc.use_ani;
    const bool poly6 = MODE != kModeAniBspline && fc.use_poly6;
    if constexpr (ref_float<R, KAHAN>()) {
        ref_float_contrib<GRAD>(ani, poly6, h2, float(fc.sigma), dx, dy, dz, g0, g1, acc, ga);
        return;
    }
    if (KAHAN && sizeof(R) == 4) {
        dx -= R(lo.x);
        dy -= R(lo.y);
        dz -= R(lo.z);
    }
    if (ani) {
        const R gx = g0.x * dx + g0.y * dy + g0.z * dz;
        const R gy = g0.y * dx + g0.w * dy + g1.x * dz;
        const R gz = g0.z * dx + g1.x * dy + g1.y * dz;
        const R n2 = gx * gx + gy * gy + gz * gz;
        if (poly6) {
            const R t = h2 - n2;
            if (!(n2 > h2)) acc.add_unfused(g1.z * (t * t * t));
        } else {
            acc.add_unfused(g1.z * bspline_q2<R>(n2 * inv_h2));
        }
        if (GRAD && !(n2 > h2)) {
            const R t = h2 - n2;
            const R w = R(-6) * g1.z * t * t;
            ga[0].add_unfused(w * gx);
            ga[1].add_unfused(w * gy);
            ga[2].add_unfused(w * gz);
        }
    } else {
        if (!(d2 > h2)) {
            const R t = h2 - d2;
            acc.add_unfused(t * t * t);
            if (GRAD) {
                const R w = R(-6) * t * t;
                ga[0].add_unfused(w * dx);
                ga[1].add_unfused(w * dy);
                ga[2].add_unfused(w * dz);
            }
        }
    }
}

// the nine (y, z) row windows of the 27-cell neighbourhood of q's cell; returns the candidate count
__device__ __forceinline__ int field_rows(const GridDev& g, float fx, float fy, float fz, int rlo[9], int rhi[9]) {
    int ix, iy, iz;
    cell_of(g, fx, fy, fz, ix, iy, iz);
    int visited = 0;
#pragma unroll
    for (int r = 0; r < 9; r++) {
        if (!row_range(g, ix, iy, iz, r, rlo[r], rhi[r])) rlo[r] = rhi[r] = 0;
        visited += rhi[r] - rlo[r];
    }
    return visited;
}

// Field sum S(q) = sum_j det_j K(|G_j (q - xMean_j)|) over the 27-cell candidates of q (without
// sigma), optionally with the poly6-form gradient sum (src/evaluator.cpp:408-415,437-463).
// Candidates are taken in the reference's order (GetPIdxList: rows z-outer, then y, each row's
// cells x-ascending, sorted particles within a cell). Returns the number of candidates visited.
// LEAN: the row windows are looked up one row at a time in a loop that is not unrolled, which
// keeps far fewer registers live (for kernels whose occupancy the 9 windows and 9 unrolled row
// loops would cut); the candidates and their order are the same.
template <class R, bool KAHAN, bool GRAD, int MODE = kModeAny, bool LEAN = false>
__device__ __forceinline__ int field_sum(const GridDev& g, const Recs<R>& rec, const FieldConst& fc, float fx,
                                         float fy, float fz, R px, R py, R pz, R& sum, R grad[3]) {
    Acc<R, KAHAN> acc, ga[3];
    const R inv_h2 = R(fc.inv_h2);
    const R h2 = R(fc.h2);
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    auto contrib = [&](int j, R dx, R dy, R dz, R d2) {
        const float4 lo = (KAHAN && sizeof(R) == 4) ? fc.lo[j] : zero4;
        field_contrib<R, KAHAN, GRAD, MODE>(fc, inv_h2, h2, dx, dy, dz, d2, rec.g[2 * j], rec.g[2 * j + 1], lo, acc,
                                            ga);
    };
    // Candidates are culled in groups of up to 32 of one row (four records requested before any
    // is used), then only the ones that pass are summed, in order: the lanes of a wave hold
    // different points, and a per-candidate branch would run the contribution whenever any lane
    // needs it. The cull is written !(d2 >= R^2) so that a NaN record passes and poisons the sum,
    // as the reference's kernels do for it.
    auto row = [&](int j0, int e) {
        for (int j = j0; j < e; j += 32) {
            const int jn = min(e, j + 32);
            unsigned m = 0;
            // (the per-row form, LEAN, requests eight records at a time: it is bound by their
            // latency, not by registers)
            constexpr int CS = LEAN ? 8 : 4;
            for (int u = j; u < jn; u += CS) {
                typename Vec4<R>::T a[CS];
#pragma unroll
                for (int k = 0; k < CS; k++) a[k] = rec.a[min(u + k, jn - 1)];
#pragma unroll
                for (int k = 0; k < CS; k++)
                    m |= unsigned(!(dist2(px - a[k].x, py - a[k].y, pz - a[k].z) >= a[k].w)) << (u - j + k);
            }
            if (jn - j < 32) m &= (1u << (jn - j)) - 1u;
            while (m) {
                const int jj = j + __builtin_ctz(m);
                m &= m - 1;
                const auto a = rec.a[jj];
                const R dx = px - a.x, dy = py - a.y, dz = pz - a.z;
                contrib(jj, dx, dy, dz, dist2(dx, dy, dz));
            }
        }
    };
    int visited = 0;
    if constexpr (LEAN) {
        int ix, iy, iz;
        cell_of(g, fx, fy, fz, ix, iy, iz);
#pragma unroll 1
        for (int r = 0; r < 9; r++) {
            int lo, hi;
            if (!row_range(g, ix, iy, iz, r, lo, hi)) continue;
            visited += hi - lo;
            row(lo, hi);
        }
    } else {
        // all 9 row windows first (independent loads in flight together)
        int rlo[9], rhi[9];
        visited = field_rows(g, fx, fy, fz, rlo, rhi);
        for (int r = 0; r < 9; r++) row(rlo[r], rhi[r]);
    }
    sum = acc.get();
    if (GRAD) {
        grad[0] = ga[0].get();
        grad[1] = ga[1].get();
        grad[2] = ga[2].get();
    }
    return visited;
}

// ---- symmetric 3x3 eigendecomposition: cyclic Jacobi (Golub & Van Loan, Alg. 8.5.3) -------
// a = (a00, a01, a02, a11, a12, a22); eigenvalues in lam, eigenvectors in columns of Q.
template <class R> __device__ __forceinline__ void jacobi3(const R a_in[6], R lam[3], R Q[3][3]) {
    R a[3][3] = {{a_in[0], a_in[1], a_in[2]}, {a_in[1], a_in[3], a_in[4]}, {a_in[2], a_in[4], a_in[5]}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Q[i][j] = (i == j) ? R(1) : R(0);
    const R eps = sizeof(R) == 4 ? R(1.2e-7) : R(2.3e-16);
    for (int sweep = 0; sweep < 24; sweep++) {
        const R off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        const R dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (!(off > eps * eps * dia * R(1e-2))) break;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int p = (k == 2) ? 1 : 0, q = (k == 0) ? 1 : 2;
            const R apq = a[p][q];
            if (apq == R(0)) continue;
            const R tau = (a[q][q] - a[p][p]) / (R(2) * apq);
            const R t = (tau >= R(0) ? R(1) : R(-1)) / (fabs(tau) + sqrt(R(1) + tau * tau));
            const R c = R(1) / sqrt(R(1) + t * t), s = t * c;
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const R arp = a[r][p], arq = a[r][q];
                a[r][p] = c * arp - s * arq;
                a[r][q] = s * arp + c * arq;
            }
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const R apr = a[p][r], aqr = a[q][r];
                a[p][r] = c * apr - s * aqr;
                a[q][r] = s * apr + c * aqr;
            }
            a[p][q] = a[q][p] = R(0);
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const R qrp = Q[r][p], qrq = Q[r][q];
                Q[r][p] = c * qrp - s * qrq;
                Q[r][q] = s * qrp + c * qrq;
            }
        }
    }
    lam[0] = a[0][0];
    lam[1] = a[1][1];
    lam[2] = a[2][2];
}

// Anisotropy matrix of compute_G_Yus (src/evaluator.cpp:627-651) from the normalised
// covariance: |eigenvalues| clamped below at max/kr (kr = 4), G = Q diag(s / w) Q^T with the
// volume-preserving s = cbrt(w0 w1 w2). Returns the six entries of symmetric G, det(G) by
// cofactor expansion, and sigma_min(G) = s / max(w).
template <class R>
__device__ __forceinline__ void g_from_cov(const R cov[6], R scale_extra, R G[6], R& det, R& gmin, bool& clamped) {
    R lam[3], Q[3][3];
    jacobi3<R>(cov, lam, Q);
    R w[3] = {fabs(lam[0]), fabs(lam[1]), fabs(lam[2])};
    const R mx = fmax(w[0], fmax(w[1], w[2]));
    const R lo = mx / R(4);
    clamped = (w[0] < lo) || (w[1] < lo) || (w[2] < lo);
    w[0] = fmax(w[0], lo);
    w[1] = fmax(w[1], lo);
    w[2] = fmax(w[2], lo);
    const R s = R(cbrt(double(w[0] * w[1] * w[2])));
    const R d0 = s / w[0] * scale_extra, d1 = s / w[1] * scale_extra, d2 = s / w[2] * scale_extra;
    G[0] = Q[0][0] * d0 * Q[0][0] + Q[0][1] * d1 * Q[0][1] + Q[0][2] * d2 * Q[0][2];
    G[1] = Q[0][0] * d0 * Q[1][0] + Q[0][1] * d1 * Q[1][1] + Q[0][2] * d2 * Q[1][2];
    G[2] = Q[0][0] * d0 * Q[2][0] + Q[0][1] * d1 * Q[2][1] + Q[0][2] * d2 * Q[2][2];
    G[3] = Q[1][0] * d0 * Q[1][0] + Q[1][1] * d1 * Q[1][1] + Q[1][2] * d2 * Q[1][2];
    G[4] = Q[1][0] * d0 * Q[2][0] + Q[1][1] * d1 * Q[2][1] + Q[1][2] * d2 * Q[2][2];
    G[5] = Q[2][0] * d0 * Q[2][0] + Q[2][1] * d1 * Q[2][1] + Q[2][2] * d2 * Q[2][2];
    det = G[0] * (G[3] * G[5] - G[4] * G[4]) - G[1] * (G[1] * G[5] - G[4] * G[2]) + G[2] * (G[1] * G[4] - G[3] * G[2]);
    gmin = s / mx * scale_extra;
}

}  // namespace pisurf
