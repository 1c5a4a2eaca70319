// Field evaluation at arbitrary points (Evaluator::SingleEval / SingleEvalWithGrad,
// src/evaluator.cpp:60-125) and particle normals (CalcParticlesNormal, :363-375).
#include "pisurf_internal.h"
#include "run_eval.h"
#include "run_tuner.h"

#include <type_traits>

namespace pisurf {

template <class R, bool K, bool GRAD>
__global__ void __launch_bounds__(256) k_eval_points(GridDev g, Recs<R> rec, FieldConst fc, const float* __restrict__ q,
                                                     int64_t n, double* __restrict__ f, double* __restrict__ grad,
                                                     Counters* ctr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int vis = 0;
    if (i < n) {
        const float x = q[3 * i], y = q[3 * i + 1], z = q[3 * i + 2];
        R s, gr[3];
        vis = field_sum<R, K, GRAD>(g, rec, fc, x, y, z, R(x), R(y), R(z), s, gr);
        const R sigma = out_sigma<R, K>(fc);
        f[i] = double(R(fc.iso) - sigma * s);
        if (GRAD) {
            grad[3 * i] = double(sigma * gr[0]);
            grad[3 * i + 1] = double(sigma * gr[1]);
            grad[3 * i + 2] = double(sigma * gr[2]);
        }
    }
    wave_count(GRAD ? &ctr_stripe(ctr)->grad_evals : &ctr_stripe(ctr)->field_evals, i < n ? 1ull : 0ull);
    wave_count(&ctr_stripe(ctr)->candidates, (unsigned long long)vis);
}

// per-particle form for grids that leave particles outside (shifted cell windows, no runs)
template <class R, bool K>
__global__ void __launch_bounds__(256) k_normals_pt(GridDev g, Recs<R> rec, FieldConst fc, const float4* __restrict__ spos,
                                                 int64_t n, float4* __restrict__ nrm, Counters* ctr) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int vis = 0;
    if (p < n) {
        const float4 pp = spos[p];
        R s, gr[3];
        vis = field_sum<R, K, true>(g, rec, fc, pp.x, pp.y, pp.z, R(pp.x), R(pp.y), R(pp.z), s, gr);
        const R sigma = out_sigma<R, K>(fc);
        nrm[p] = make_float4(float(sigma * gr[0]), float(sigma * gr[1]), float(sigma * gr[2]), 0.0f);
    }
    wave_count(&ctr_stripe(ctr)->grad_evals, p < n ? 1ull : 0ull);
    wave_count(&ctr_stripe(ctr)->candidates, (unsigned long long)vis);
}

// CalcParticlesNormal over runs: the particles of kx consecutive cells of one (y, z) row of the
// grid (their sorted range is contiguous), candidates staged through LDS
template <class R> struct ParticleSrc {
    const float4* spos;
    float4* nrm;
    R sigma;
    __device__ __forceinline__ void point(int h, float p[3]) const {
        const float4 q = spos[h];
        p[0] = q.x;
        p[1] = q.y;
        p[2] = q.z;
    }
    __device__ __forceinline__ void store(int h, R, const R* gr) const {
        nrm[h] = make_float4(float(sigma * gr[0]), float(sigma * gr[1]), float(sigma * gr[2]), 0.0f);
    }
};

template <class R, bool K, int T, int MODE>
__global__ void __launch_bounds__(T) k_run_normals(GridDev g, Recs<R> rec, FieldConst fc,
                                                   const float4* __restrict__ spos, int kx,
                                                   float4* __restrict__ nrm, Counters* ctr) {
    __shared__ RunLds<R, K && sizeof(R) == 4> lds;
    int h0, h1;
    particle_run(g, kx, h0, h1);
    if (h0 >= h1) return;  // uniform over the block
    ParticleSrc<R> src{spos, nrm, out_sigma<R, K>(fc)};
    const unsigned long long vis = run_eval<T, R, K, true, MODE>(g, rec, fc, src, h0, h1, lds);
    Counters* c = ctr_stripe(ctr);
    wave_count(&c->candidates, vis);
    if (threadIdx.x == 0) atomicAdd(&c->grad_evals, (unsigned long long)(h1 - h0));
}

void launch_eval_points(int prec, GridDev g, const void* ra, const void* rg, FieldConst fc, const float* q, int64_t n,
                        double* f, double* grad, Counters* ctr, hipStream_t s) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (prec == 2) {
        Recs<double> r{(const double4*)ra, (const double4*)rg};
        if (grad) k_eval_points<double, false, true><<<nb, 256, 0, s>>>(g, r, fc, q, n, f, grad, ctr);
        else k_eval_points<double, false, false><<<nb, 256, 0, s>>>(g, r, fc, q, n, f, grad, ctr);
    } else {
        Recs<float> r{(const float4*)ra, (const float4*)rg};
        if (prec == 1) {
            if (grad) k_eval_points<float, true, true><<<nb, 256, 0, s>>>(g, r, fc, q, n, f, grad, ctr);
            else k_eval_points<float, true, false><<<nb, 256, 0, s>>>(g, r, fc, q, n, f, grad, ctr);
        } else {
            if (grad) k_eval_points<float, false, true><<<nb, 256, 0, s>>>(g, r, fc, q, n, f, grad, ctr);
            else k_eval_points<float, false, false><<<nb, 256, 0, s>>>(g, r, fc, q, n, f, grad, ctr);
        }
    }
}

void launch_normals(int prec, GridDev g, const void* ra, const void* rg, FieldConst fc, const float4* spos, int64_t n,
                    int64_t n_invalid, float4* nrm, Counters* ctr, Counters* trial_ctr, hipStream_t s) {
    if (n <= 0) return;
    if (n_invalid > 0) {
        const unsigned nb = (unsigned)((n + 255) / 256);
        if (prec == 2) {
            Recs<double> r{(const double4*)ra, (const double4*)rg};
            k_normals_pt<double, false><<<nb, 256, 0, s>>>(g, r, fc, spos, n, nrm, ctr);
        } else {
            Recs<float> r{(const float4*)ra, (const float4*)rg};
            if (prec == 1) k_normals_pt<float, true><<<nb, 256, 0, s>>>(g, r, fc, spos, n, nrm, ctr);
            else k_normals_pt<float, false><<<nb, 256, 0, s>>>(g, r, fc, spos, n, nrm, ctr);
        }
        return;
    }
    // runs of kx cells of a grid row, one workgroup each (empty runs end at once); the shape is
    // tuned per precision and average particles per cell
    const double p = (double)n / std::max<int64_t>((int64_t)g.n[0] * g.n[1] * g.n[2], 1);
    tuned_run_launch(
        prec, kRunNormals, p, trial_ctr, s,
        [&](RunShape c) { return (int64_t)((g.n[0] + c.kx - 1) / c.kx) * g.n[1] * g.n[2] <= INT32_MAX; },
        [&](RunShape c, bool trial) {
            Counters* cc = trial ? trial_ctr : ctr;
            const unsigned nb = (unsigned)((int64_t)((g.n[0] + c.kx - 1) / c.kx) * g.n[1] * g.n[2]);
            auto go = [&](auto tag, auto mode) {
                constexpr int TT = decltype(tag)::value, M = decltype(mode)::value;
                if (prec == 2) {
                    Recs<double> r{(const double4*)ra, (const double4*)rg};
                    k_run_normals<double, false, TT, M><<<nb, TT, 0, s>>>(g, r, fc, spos, c.kx, nrm, cc);
                } else {
                    Recs<float> r{(const float4*)ra, (const float4*)rg};
                    if (prec == 1) k_run_normals<float, true, TT, M><<<nb, TT, 0, s>>>(g, r, fc, spos, c.kx, nrm, cc);
                    else k_run_normals<float, false, TT, M><<<nb, TT, 0, s>>>(g, r, fc, spos, c.kx, nrm, cc);
                }
            };
            using M0 = std::integral_constant<int, kModeAny>;
            using M1 = std::integral_constant<int, kModeAniBspline>;
            const bool m1 = field_mode(fc) == kModeAniBspline;
            if (c.T == 256) m1 ? go(std::integral_constant<int, 256>(), M1()) : go(std::integral_constant<int, 256>(), M0());
            else m1 ? go(std::integral_constant<int, 128>(), M1()) : go(std::integral_constant<int, 128>(), M0());
        });
}

// UniformGrid::gridSampling (src/marching.cpp:72-91): point i = x + dims0 * (y + dims1 * z) at
// float(steps - x) / float(steps) * min + float(x) / float(steps) * max per axis, each operation
// rounded on its own (the reference is built without FMA contraction)
template <class R, bool K>
__global__ void __launch_bounds__(256) k_uniform_grid(GridDev g, Recs<R> rec, FieldConst fc, UniformBox ub,
                                                      int64_t total, float* __restrict__ out, Counters* ctr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int vis = 0;
    if (i < total) {
        const int64_t plane = (int64_t)ub.dims[0] * ub.dims[1];
        const int z = (int)(i / plane);
        const int64_t rem = i - (int64_t)z * plane;
        const int y = (int)(rem / ub.dims[0]);
        const int x = (int)(rem - (int64_t)y * ub.dims[0]);
        const int ix[3] = {x, y, z};
        float p[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            // plain operators with contraction off in this scope: the *_rn helpers are inlined as
            // contractable operations, and a fused product would round differently
#pragma clang fp contract(off)
            const float st = (float)ub.steps[a];
            const float u = (float)(ub.steps[a] - ix[a]) / st;
            const float v = (float)ix[a] / st;
            p[a] = u * ub.lo[a] + v * ub.hi[a];
        }
        R s, gr[3];
        vis = field_sum<R, K, false>(g, rec, fc, p[0], p[1], p[2], R(p[0]), R(p[1]), R(p[2]), s, gr);
        out[i] = float(R(fc.iso) - out_sigma<R, K>(fc) * s);
    }
    wave_count(&ctr_stripe(ctr)->field_evals, i < total ? 1ull : 0ull);
    wave_count(&ctr_stripe(ctr)->candidates, (unsigned long long)vis);
}

void launch_uniform_grid(int prec, GridDev g, const void* ra, const void* rg, FieldConst fc, UniformBox ub,
                         int64_t total, float* out, Counters* ctr, hipStream_t s) {
    if (total <= 0) return;
    const unsigned nb = (unsigned)((total + 255) / 256);
    if (prec == 2) {
        Recs<double> r{(const double4*)ra, (const double4*)rg};
        k_uniform_grid<double, false><<<nb, 256, 0, s>>>(g, r, fc, ub, total, out, ctr);
    } else {
        Recs<float> r{(const float4*)ra, (const float4*)rg};
        if (prec == 1) k_uniform_grid<float, true><<<nb, 256, 0, s>>>(g, r, fc, ub, total, out, ctr);
        else k_uniform_grid<float, false><<<nb, 256, 0, s>>>(g, r, fc, ub, total, out, ctr);
    }
}

}  // namespace pisurf
