// Vertex placement: edge bisection (VisitorExtract::calc_vertices, src/visitorextract.cpp:47-84)
// and the per-leaf QEF solve (TNode::NodeCalcNode, src/iso_method_ours.cpp:101-174).
#include "pisurf_internal.h"
#include "qef_device.h"

#include <type_traits>

namespace pisurf {

// One thread per edge. Endpoints are (x, y, z, f) with f of opposite signs (f > 0 is +1).
// The midpoint replaces the endpoint of equal sign and the length is halved (not recomputed)
// until it is at most r/2; then the zero crossing is interpolated, snapping to an endpoint
// when the ratio is below 0.1 or above 0.9.
template <class R, bool K, int MODE>
__global__ void __launch_bounds__(256) k_bisect(GridDev g, Recs<R> rec, FieldConst fc, float half_r,
                                                const double* __restrict__ ends, int64_t n, double* __restrict__ out,
                                                int* __restrict__ nev_out, Counters* ctr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int vis = 0, nev = 0;
    if (i < n) {
        // the edge length and the interpolation round every operation as the reference does
#pragma clang fp contract(off)
        R v1[4], v2[4];
        for (int a = 0; a < 4; a++) {
            v1[a] = R(ends[8 * i + a]);
            v2[a] = R(ends[8 * i + 4 + a]);
        }
        const R ex = v1[0] - v2[0], ey = v1[1] - v2[1], ez = v1[2] - v2[2];
        const R e2 = ex * ex + ey * ey + ez * ez;
        R d = (sizeof(R) == 4) ? R(__builtin_sqrtf(float(e2))) : R(sqrt(double(e2)));
        const R rr = R(half_r);
        // the reference loop has no cap; 256 halvings bound it for any finite input
        for (int it = 0; it < 256 && d > rr; it++) {
            R m[4];
            for (int a = 0; a < 3; a++) m[a] = (v1[a] + v2[a]) / R(2);
            R sum, g3[3];
            vis += field_sum<R, K, false, MODE, true>(g, rec, fc, float(m[0]), float(m[1]), float(m[2]), m[0], m[1],
                                                      m[2], sum, g3);
            nev++;
            m[3] = R(fc.iso) - out_sigma<R, K>(fc) * sum;
            const int sm = m[3] > R(0) ? 1 : -1, s1 = v1[3] > R(0) ? 1 : -1;
            if (sm == s1)
                for (int a = 0; a < 4; a++) v1[a] = m[a];
            else
                for (int a = 0; a < 4; a++) v2[a] = m[a];
            d /= R(2);
        }
        const R ratio = (R(0) - v1[3]) / (v2[3] - v1[3]);
        R o[3];
        if (ratio < R(0.1))
            for (int a = 0; a < 3; a++) o[a] = v1[a];
        else if (ratio > R(0.9))
            for (int a = 0; a < 3; a++) o[a] = v2[a];
        else
            for (int a = 0; a < 3; a++) o[a] = v1[a] + (v2[a] - v1[a]) * ratio;
        for (int a = 0; a < 3; a++) out[3 * i + a] = double(o[a]);
        if (nev_out) nev_out[i] = nev;
    }
    wave_count(&ctr_stripe(ctr)->field_evals, (unsigned long long)nev);
    wave_count(&ctr_stripe(ctr)->candidates, (unsigned long long)vis);
}

void launch_bisect(int prec, GridDev g, const void* ra, const void* rg, FieldConst fc, float half_r, const double* ends,
                   int64_t n, double* out, int* nev, Counters* ctr, hipStream_t s) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    auto go = [&](auto mode) {
        constexpr int M = decltype(mode)::value;
        if (prec == 2) {
            Recs<double> r{(const double4*)ra, (const double4*)rg};
            k_bisect<double, false, M><<<nb, 256, 0, s>>>(g, r, fc, half_r, ends, n, out, nev, ctr);
        } else {
            Recs<float> r{(const float4*)ra, (const float4*)rg};
            if (prec == 1) k_bisect<float, true, M><<<nb, 256, 0, s>>>(g, r, fc, half_r, ends, n, out, nev, ctr);
            else k_bisect<float, false, M><<<nb, 256, 0, s>>>(g, r, fc, half_r, ends, n, out, nev, ctr);
        }
    };
    if (field_mode(fc) == kModeAniBspline) go(std::integral_constant<int, kModeAniBspline>());
    else go(std::integral_constant<int, kModeAny>());
}

// Standalone QEF over given samples (n x 27 x 4) and gradients (n x 27 x 3).
template <class R, bool K>
__global__ void __launch_bounds__(64) k_qef(GridDev g, Recs<R> rec, FieldConst fc, TreeParams tp,
                                            const float4* __restrict__ cen, const double* __restrict__ samples,
                                            const double* __restrict__ grads, int64_t n, int mode,
                                            double4* __restrict__ out, Counters* ctr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int vis = 0;
    if (i < n) {
        const float4 c = cen[i];
        float pts[27][3];
        R f[27], gr[27][3];
        for (int s = 0; s < 27; s++) {
            for (int a = 0; a < 3; a++) {
                pts[s][a] = float(samples[(i * 27 + s) * 4 + a]);
                gr[s][a] = R(grads[(i * 27 + s) * 3 + a]);
            }
            f[s] = R(samples[(i * 27 + s) * 4 + 3]);
        }
        R pc[4], err;
        R vx = R(c.x), vy = R(c.y), vz = R(c.z);
        if (qef_solve<R>(c, pts, f, gr, mode, tp.qef_svd_tol, pc, err)) {
            vx = pc[0];
            vy = pc[1];
            vz = pc[2];
        }
        R sum, g3[3];
        vis = field_sum<R, K, false>(g, rec, fc, float(vx), float(vy), float(vz), vx, vy, vz, sum, g3);
        R fv = R(fc.iso) - out_sigma<R, K>(fc) * sum;
        if (err / R(__fmul_rn(2.0f, c.w)) < R(0.0001)) fv = R(0);
        out[i] = make_double4(double(vx), double(vy), double(vz), double(fv));
    }
    wave_count(&ctr_stripe(ctr)->qef_solves, i < n ? 1ull : 0ull);
    wave_count(&ctr_stripe(ctr)->field_evals, i < n ? 1ull : 0ull);
    wave_count(&ctr_stripe(ctr)->candidates, (unsigned long long)vis);
}

void launch_qef(int prec, GridDev g, const void* ra, const void* rg, FieldConst fc, TreeParams tp, const float4* cen,
                const double* samples, const double* grads, int64_t n, int mode, double4* out, Counters* ctr,
                hipStream_t s) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)((n + 63) / 64);
    if (prec == 2) {
        Recs<double> r{(const double4*)ra, (const double4*)rg};
        k_qef<double, false><<<nb, 64, 0, s>>>(g, r, fc, tp, cen, samples, grads, n, mode, out, ctr);
    } else {
        Recs<float> r{(const float4*)ra, (const float4*)rg};
        if (prec == 1) k_qef<float, true><<<nb, 64, 0, s>>>(g, r, fc, tp, cen, samples, grads, n, mode, out, ctr);
        else k_qef<float, false><<<nb, 64, 0, s>>>(g, r, fc, tp, cen, samples, grads, n, mode, out, ctr);
    }
}

}  // namespace pisurf
