// Octree level sweep: node emptiness and curvature, 27-point sampling, classification and
// child generation (SurfReconstructor::genIsoOurs state 0, src/surface_reconstructor.cpp:152-407;
// TNode, src/iso_method_ours.cpp:16-99). One launch of each kernel covers all nodes of a level.
#include <climits>

#include "pisurf_internal.h"
#include "qef_device.h"
#include "run_eval.h"
#include "run_tuner.h"

#include <cmath>
#include <type_traits>

namespace pisurf {

enum : uint8_t { kEmpty = 0, kInternal = 1, kLeaf = 2 };  // include/iso_method_ours.h:25-27

template <class R> struct R4t { typedef typename Vec4<R>::T T; };

// leaf test of afterSampleEval: (cellsize - min_radius) < 1e-8 in double
__device__ __forceinline__ bool leaf_depth(float half, float r) {
    const float cellsize = __fmul_rn(2.0f, half);
    return double(__fsub_rn(cellsize, r)) < 1e-8;
}

// checkEmptyAndCalcCurv (src/surface_reconstructor.cpp:152-222): a group of G lanes per node
// (G = 64 for large boxes, 16 or 8 for the small boxes of the deep levels). The box query enumerates
// cells [lo-1, hi+1] of the clamped box (GetInBoxParticles); each lane walks whole (y,z) rows of
// that range (one contiguous window of sorted positions per row). A particle counts when it lies
// strictly inside the unclamped box grown by r * smooth. Nodes above D_min only need emptiness and
// stop at the first hit.
// nodes whose 27 samples are needed: non-empty, depth >= D_min, and not at leaf depth unless the
// QEF consumes them (the shipped build samples leaf-depth nodes but never uses the result)
__device__ __forceinline__ bool wants_samples(const NodeLevel& lv, const TreeParams& tp, int64_t node) {
    if (lv.empty[node]) return false;
    if (lv.depth[node] < tp.depth_min) return false;
    if (tp.qef_mode == 0 && leaf_depth(lv.c[node].w, tp.radius)) return false;
    return true;
}

// leaf-depth nodes that need their centre value and have no samples (QEF off)
__device__ __forceinline__ bool wants_center(const NodeLevel& lv, const TreeParams& tp, int64_t node) {
    return !lv.empty[node] && tp.qef_mode == 0 && leaf_depth(lv.c[node].w, tp.radius) &&
           lv.depth[node] >= tp.depth_min;
}

template <int G, int PASS>
__global__ void __launch_bounds__(256) k_empty_curv(GridDev g, const float4* __restrict__ spos,
                                                    const float4* __restrict__ nrm, TreeParams tp, NodeLevel lv,
                                                    const void* __restrict__ samples, int rsz) {
    const int64_t node = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    const int gbase = (threadIdx.x & 63) - lane;  // first lane of this group within the wave
    if (node >= lv_count(lv)) return;
    if (PASS == kCurvOnly) {
        // classification reads curv only for sampled nodes whose samples change sign (the exits
        // below are uniform over the node's lanes)
        if (!tp.curv || !wants_samples(lv, tp, node)) return;
        bool sc = false;
        if (rsz == 8) {
            const double* f = (const double*)samples + node * 27;
            const bool s0 = f[0] >= 0.0;
            for (int k = 1; k < 27; k++) sc = sc || ((f[k] >= 0.0) != s0);
        } else {
            const float* f = (const float*)samples + node * 27;
            const bool s0 = f[0] >= 0.0f;
            for (int k = 1; k < 27; k++) sc = sc || ((f[k] >= 0.0f) != s0);
        }
        if (!sc) return;
        if (!tp.calc_normals) {  // no normals: curv = 1 (no normal length)
            if (lane == 0) lv.curv[node] = 1.0f;
            return;
        }
    }
    const float4 c = lv.c[node];
    const int depth = lv.depth[node];
    const float cc[3] = {c.x, c.y, c.z};
    float b1[3], b2[3], bc1[3], bc2[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        b1[a] = __fsub_rn(cc[a], c.w);
        b2[a] = __fadd_rn(cc[a], c.w);
        bc1[a] = fmaxf(b1[a], tp.b_lo[a]);
        bc2[a] = fminf(b2[a], tp.b_hi[a]);
    }
    const bool need_curv = PASS == kCurvOnly || (PASS == kEmptyAndCurv && tp.calc_normals && tp.curv &&
                                                 depth >= tp.depth_min);
    float lx[3], hx[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lx[a] = __fsub_rn(b1[a], tp.margin);
        hx[a] = __fadd_rn(b2[a], tp.margin);
    }
    // cells [lo-1, hi+1] of the clamped box, narrowed to the cells that can hold a particle of the
    // grown box: the cell coordinate is monotone in the position, so a particle with lx < p < hx
    // lies in a cell between those of lx and hx (the particles visited are the same)
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int l = cell_coord(bc1[a], g.b0[a], g.cell, g.n[a]);
        const int h = cell_coord(bc2[a], g.b0[a], g.cell, g.n[a]);
        lo[a] = max(max(l - 1, 0), cell_coord(lx[a], g.b0[a], g.cell, g.n[a]));
        hi[a] = min(min(h + 1, g.n[a] - 1), cell_coord(hx[a], g.b0[a], g.cell, g.n[a]));
    }
    bool found = false;
    float sx = 0.f, sy = 0.f, sz = 0.f, area = 0.f;
    const int ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    const int nrows = (hi[0] >= lo[0] && ny > 0 && nz > 0) ? ny * nz : 0;
    const unsigned long long gmask = (G == 64) ? ~0ull : (((1ull << (G % 64)) - 1) << gbase);
    for (int r0 = 0; r0 < nrows; r0 += G) {
        const int row = r0 + lane;
        if (row < nrows) {
            const int y = lo[1] + row % ny, z = lo[2] + row / ny;
            const int base = (z * g.n[1] + y) * g.n[0];
            const int j0 = g.off[base + lo[0]], j1 = g.off[base + hi[0] + 1];
            auto inside = [&](const float4 p) {
                return p.x > lx[0] && p.x < hx[0] && p.y > lx[1] && p.y < hx[1] && p.z > lx[2] && p.z < hx[2];
            };
            auto add_normal = [&](int j) {
                const float4 n = nrm[j];
                sx += n.x;
                sy += n.y;
                sz += n.z;
                area += sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
            };
            int j = j0;
            // four positions are loaded before any is tested, so the loads overlap
            for (; j + 4 <= j1; j += 4) {
                float4 q[4];
#pragma unroll
                for (int k = 0; k < 4; k++) q[k] = spos[j + k];
                bool in[4];
#pragma unroll
                for (int k = 0; k < 4; k++) in[k] = inside(q[k]);
                if (!need_curv) {
                    if (in[0] || in[1] || in[2] || in[3]) {
                        found = true;
                        break;
                    }
                    continue;
                }
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (in[k]) {
                        found = true;
                        add_normal(j + k);
                    }
            }
            for (; j < j1 && !(found && !need_curv); j++) {
                if (inside(spos[j])) {
                    found = true;
                    if (need_curv) add_normal(j);
                }
            }
        }
        if (!need_curv && (__ballot(found) & gmask)) break;
    }
    const bool any = (__ballot(found) & gmask) != 0;
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o, kWave);
        sy += __shfl_xor(sy, o, kWave);
        sz += __shfl_xor(sz, o, kWave);
        area += __shfl_xor(area, o, kWave);
    }
    if (lane == 0) {
        if (PASS != kCurvOnly) lv.empty[node] = any ? 0 : 1;
        if (PASS != kEmptyOnly) lv.curv[node] = (area == 0.f) ? 1.0f : sqrtf(sx * sx + sy * sy + sz * sz) / area;
    }
}


// compaction of the nodes to evaluate: flag, then (after a scan) the list. A listed node gets its
// 27 samples, or only its centre value (stored in sample slot 13) when it sits at leaf depth.
__global__ void k_sample_flag(TreeParams tp, NodeLevel lv, int* __restrict__ flag, int* __restrict__ cflag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < lv.n) {
        flag[i] = wants_samples(lv, tp, i) ? 1 : 0;
        cflag[i] = wants_center(lv, tp, i) ? 1 : 0;
    }
    if (i == lv.n) flag[i] = cflag[i] = 0;
}

// ---- per-level sample deduplication --------------------------------------------------------
// Neighbouring nodes share sample points on their faces, edges and corners. Entries (listed node,
// sample) are keyed by the point's hash cell group and lattice position, so a shared point's
// copies sort next to each other and the distinct points come out grouped by cell for the run
// evaluation; only copies with bitwise-identical float coordinates are merged, so every value is
// exactly what the node's own evaluation would produce.

// Hash cell group of a point for the run evaluation: its cell (CalcXYZIdx) in the grid
// extended by one cell on every side; points farther out have an empty 27-cell window and share
// the last group. Points of one group have the same candidate list in the same order.
__device__ __forceinline__ uint64_t cell_group(const GridDev& g, float x, float y, float z) {
    int ix, iy, iz;
    cell_of(g, x, y, z, ix, iy, iz);
    const uint64_t ex = g.n[0] + 2, ey = g.n[1] + 2, ez = g.n[2] + 2;
    if (ix < -1 || ix > g.n[0] || iy < -1 || iy > g.n[1] || iz < -1 || iz > g.n[2]) return ex * ey * ez;
    return ((uint64_t)(iz + 1) * ey + (uint64_t)(iy + 1)) * ex + (uint64_t)(ix + 1);
}

// key of sample entry t: the point's cell group, then its lattice position modulo 2^lb per axis
// (lb bits: enough for the lattice steps a cell spans). Equal points have equal keys; distinct
// points that share one are told apart by the heads pass, which compares coordinates.
__global__ void k_sample_keys(GridDev g, NodeLevel lv, const int* __restrict__ list, int64_t nent, float ox,
                              float oy, float oz, int lb, int vbits, uint64_t* __restrict__ key,
                              int* __restrict__ val) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nent) return;
    const int node = list[t / 27];
    const float4 c = lv.c[node];
    float p[3];
    sample_point(c, (int)(t % 27), p);
    const float o[3] = {ox, oy, oz};
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float u = rintf((p[a] - o[a]) / c.w);
        q[a] = (uint32_t)fminf(fmaxf(u, 0.0f), 2097151.0f);
    }
    const uint32_t m = (1u << lb) - 1;
    const uint64_t k = (cell_group(g, p[0], p[1], p[2]) << (3 * lb)) | (q[0] & m) | ((q[1] & m) << lb) |
                       ((uint64_t)(q[2] & m) << (2 * lb));
    if (vbits) {
        key[t] = (k << vbits) | (uint64_t)t;
    } else {
        key[t] = k;
        val[t] = (int)t;
    }
}

__device__ __forceinline__ void entry_point(const NodeLevel& lv, const int* list, int e, float p[3]) {
    sample_point(lv.c[list[e / 27]], e % 27, p);
}
// the point of sorted item value v: entry v of the list, or with ent node * 27 + sample ent[v]
__device__ __forceinline__ void item_point(const NodeLevel& lv, const int* list, const int* ent, int v, float p[3]) {
    if (ent) {
        const int ns = ent[v];
        sample_point(lv.c[ns / 27], ns % 27, p);
    } else {
        entry_point(lv, list, v, p);
    }
}

// head[j] = 1 when no earlier entry of the same key run has the same coordinates; rep[j] = sorted
// position of the first entry with j's coordinates. A run (a lattice point shared by up to 8
// nodes) is short: the 7 entries before j come from the lanes before j by shuffle, and only
// entries of the run before those (or before the wave) are recomputed.
__device__ __forceinline__ bool same_point(const float p[3], const float q[3]) {
    return q[0] == p[0] && q[1] == p[1] && q[2] == p[2] && __float_as_uint(q[0]) == __float_as_uint(p[0]) &&
           __float_as_uint(q[1]) == __float_as_uint(p[1]) && __float_as_uint(q[2]) == __float_as_uint(p[2]);
}
__global__ void __launch_bounds__(256) k_sample_heads(NodeLevel lv, const int* __restrict__ list, SortedKV kv,
                                                      const int* __restrict__ ent, int64_t nent,
                                                      int* __restrict__ block_heads, int* __restrict__ rep) {
    __shared__ int wsum[4];
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    const bool act = j < nent;
    float p[3] = {0.f, 0.f, 0.f};
    int64_t s0 = j;
    if (act) {
        item_point(lv, list, ent, kv.val(j), p);
        const uint64_t kj = kv.key(j);
        while (s0 > 0 && j - s0 < 64 && kv.key(s0 - 1) == kj) s0--;
    }
    // entries [lo, j) are read from the lanes before this one
    const int64_t lo = max(s0, max(j - 7, j - lane));
    int64_t r = j;
    for (int64_t k = s0; k < lo && r == j; k++) {
        float q[3];
        item_point(lv, list, ent, kv.val(k), q);
        if (same_point(p, q)) r = k;
    }
    float q[8][3];
#pragma unroll
    for (int o = 1; o <= 7; o++) {
        q[o][0] = __shfl_up(p[0], o, kWave);
        q[o][1] = __shfl_up(p[1], o, kWave);
        q[o][2] = __shfl_up(p[2], o, kWave);
    }
#pragma unroll
    for (int o = 7; o >= 1; o--)
        if (r == j && j - o >= lo && same_point(p, q[o])) r = j - o;
    if (act) rep[j] = (int)r;
    // heads of this block of 256 entries (a head is its own representative)
    int c = (act && r == j) ? 1 : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, kWave);
    if (lane == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) block_heads[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (blockIdx.x == 0 && threadIdx.x == 0) block_heads[gridDim.x] = 0;  // closes the counts for their scan
}

// hlist[h] = sorted position of the h-th head, hpos[j] = h for each head j (block_off: scanned
// per-block head counts of k_sample_heads' blocks of 256)
__global__ void __launch_bounds__(256) k_head_list(const int* __restrict__ rep, int64_t nent,
                                                   const int* __restrict__ block_off, int* __restrict__ hlist,
                                                   int* __restrict__ hpos) {
    __shared__ int wsum[4];
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int f = (j < nent && rep[j] == (int)j) ? 1 : 0;
    const int incl = wave_incl_scan(f, lane);
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int pos = block_off[blockIdx.x] + incl - f;
    for (int k = 0; k < w; k++) pos += wsum[k];
    if (f) {
        hlist[pos] = (int)j;
        hpos[j] = pos;
    }
}

// ---- groups of points by hash cell, and runs of groups (run_eval.h) ----
// cflag[h] = 1 where the cell group changes along the distinct points (cell-major order)
__global__ void k_cell_flags(SortedKV kv, int gshift, const int* __restrict__ hlist, int64_t nheads,
                             int* __restrict__ cflag) {
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h > nheads) return;
    if (h == nheads) {
        cflag[h] = 0;
        return;
    }
    // hlist NULL: the points are the sorted entries themselves
    const int64_t a = hlist ? hlist[h] : h, b = h > 0 ? (hlist ? hlist[h - 1] : h - 1) : 0;
    cflag[h] = (h == 0 || (kv.key(a) >> gshift) != (kv.key(b) >> gshift)) ? 1 : 0;
}
// cstart[k] = first distinct point of cell group k; cstart[ngroups] = nheads
__global__ void k_cell_starts(const int* __restrict__ cflag, const int* __restrict__ cpos, int64_t nheads,
                              int* __restrict__ cstart) {
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h > nheads) return;
    if (h == nheads) cstart[cpos[h]] = (int)nheads;
    else if (cflag[h]) cstart[cpos[h]] = (int)h;
}

// runs of cell groups: rflag[k] = 1 where the run id (cell row, x / kx) changes along the groups
__global__ void k_run_flags(GridDev g, SortedKV kv, int gshift, const int* __restrict__ hlist,
                            const int* __restrict__ cstart, int64_t ngroups, int kx, int* __restrict__ rflag) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > ngroups) return;
    if (k == ngroups) {
        rflag[k] = 0;
        return;
    }
    auto run_id = [&](int64_t grp) {
        const uint64_t gid = kv.key(hlist ? hlist[cstart[grp]] : cstart[grp]) >> gshift;
        const uint64_t ex = g.n[0] + 2, ey = g.n[1] + 2;
        const int ix = (int)(gid % ex) - 1, iy = (int)((gid / ex) % ey) - 1, iz = (int)(gid / (ex * ey)) - 1;
        return run_of_cell(g, ix, iy, iz, kx);  // the empty-window group maps past every row
    };
    rflag[k] = (k == 0 || run_id(k) != run_id(k - 1)) ? 1 : 0;
}
// rstart[j] = first distinct point of run j; rstart[nruns] = nheads (= cstart[ngroups])
__global__ void k_run_starts(const int* __restrict__ rflag, const int* __restrict__ rpos,
                             const int* __restrict__ cstart, int64_t ngroups, int* __restrict__ rstart) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > ngroups) return;
    if (k == ngroups || rflag[k]) rstart[rpos[k]] = cstart[k];
}

// leaf-depth centres keyed by cell group (the centres are distinct, one per node)
__global__ void k_center_keys(GridDev g, NodeLevel lv, const int* __restrict__ list, int64_t k, int vbits,
                              uint64_t* __restrict__ key, int* __restrict__ val) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= k) return;
    const float4 c = lv.c[list[e]];
    const uint64_t kk = cell_group(g, c.x, c.y, c.z);
    if (vbits) {
        key[e] = (kk << vbits) | (uint64_t)e;
    } else {
        key[e] = kk;
        val[e] = (int)e;
    }
}

template <class R> struct CenterSrc {
    NodeLevel lv;
    const int* list;
    SortedKV kv;
    R* samples;
    R iso, sigma;
    __device__ __forceinline__ void point(int h, float p[3]) const {
        const float4 c = lv.c[list[kv.val(h)]];
        p[0] = c.x;
        p[1] = c.y;
        p[2] = c.z;
    }
    __device__ __forceinline__ void store(int h, R sum, const R*) const {
        samples[(int64_t)list[kv.val(h)] * 27 + 13] = iso - sigma * sum;
    }
};

template <class R, bool K, int T, int MODE>
__global__ void __launch_bounds__(T) k_run_centers(GridDev g, Recs<R> rec, FieldConst fc, NodeLevel lv,
                                                             const int* __restrict__ list, SortedKV kv,
                                                             const int* __restrict__ rstart, R* __restrict__ samples,
                                                             Counters* ctr) {
    __shared__ RunLds<R, K && sizeof(R) == 4> lds;
    const int h0 = rstart[blockIdx.x], h1 = rstart[blockIdx.x + 1];
    CenterSrc<R> src{lv, list, kv, samples, R(fc.iso), out_sigma<R, K>(fc)};
    const unsigned long long vis = run_eval<T, R, K, false, MODE>(g, rec, fc, src, h0, h1, lds);
    Counters* c = ctr_stripe(ctr);
    wave_count(&c->candidates, vis);
    if (threadIdx.x == 0) atomicAdd(&c->field_evals, (unsigned long long)(h1 - h0));
}

// Cell groups of nh points in cell-major order (key[hlist[h]] >> gshift is the cell group of point
// h; hlist NULL: point h is entry h). cwork: flags | positions | group starts, run starts
// ((nh + 2) + (nh + 2) + 2 (nh + 2) ints). Returns the number of groups (one read-back; 0 with *err
// set when the read fails).
static int build_groups(SortedKV kv, int gshift, const int* hlist, int nh, int* cwork, int* tmp, HostRead* rd,
                        hipStream_t s, hipError_t* err) {
    if (nh <= 0) return 0;
    int* flag = cwork;
    int* pos = cwork + (nh + 2);
    int* cstart = cwork + 2 * (int64_t)(nh + 2);
    const unsigned nbh = (unsigned)((nh + 1 + 255) / 256);
    k_cell_flags<<<nbh, 256, 0, s>>>(kv, gshift, hlist, nh, flag);
    const ScanPub pub = scan_read_begin(rd);
    exclusive_scan_i32(flag, pos, nh + 1, tmp, s, pub);
    k_cell_starts<<<nbh, 256, 0, s>>>(flag, pos, nh, cstart);
    int ng = 0;
    const hipError_t e = scan_read_end(&ng, pos + nh, pub, s, rd);
    if (e != hipSuccess) {
        *err = e;
        return 0;
    }
    return ng;
}

// Runs of at most kx cells of one grid row over the ng groups built by build_groups: returns the
// number of runs (one read-back; 0 with *err set when the read fails) and in *rstart_out their
// starts (nruns + 1 entries in cwork).
static int build_run_starts(GridDev g, SortedKV kv, int gshift, const int* hlist, int nh, int ng, int kx, int* cwork,
                            int* tmp, int** rstart_out, HostRead* rd, hipStream_t s, hipError_t* err) {
    int* flag = cwork;
    int* pos = cwork + (nh + 2);
    int* cstart = cwork + 2 * (int64_t)(nh + 2);
    int* rstart = cstart + (ng + 1);  // (ng + 1) + (nruns + 1) <= 2 (nh + 1)
    *rstart_out = rstart;
    if (ng <= 0) return 0;
    const unsigned nbg = (unsigned)((ng + 1 + 255) / 256);
    k_run_flags<<<nbg, 256, 0, s>>>(g, kv, gshift, hlist, cstart, ng, kx, flag);
    const ScanPub pub = scan_read_begin(rd);
    exclusive_scan_i32(flag, pos, ng + 1, tmp, s, pub);
    k_run_starts<<<nbg, 256, 0, s>>>(flag, pos, cstart, ng, rstart);
    int nr = 0;
    const hipError_t e = scan_read_end(&nr, pos + ng, pub, s, rd);
    if (e != hipSuccess) {
        *err = e;
        return 0;
    }
    return nr;
}

// ---- sample points merged per tile of nodes, before the global sort -------------------------
// The listed nodes of a level (one node size) are taken in tiles of 256 consecutive list entries
// (children follow their parents, so a tile is mostly a compact piece of space). Sample s of node
// ip lies on lattice point 2 ip + t(s), t(s) in {0, 1, 2}^3; the nodes that can share it are
// ip + d with d_a in {-1, 0} when t_a = 0, {0} when t_a = 1 and {0, +1} when t_a = 2. Of those,
// the ones with a -1 component (the others come later in Morton order) are looked up in a dense
// LDS map over the tile's bounding box, and the entry refers to the one first in the list, if it
// is earlier than the entry's own node and that sample's coordinates are bitwise identical to its
// own. References point to earlier nodes, so following them ends at an entry with none, which owns
// a record; the global heads pass, which compares coordinates, still merges identical points of
// different records. A tile whose box (grown by one) exceeds the map gives every entry its own
// record. A tile takes the slots of its records from a level-wide counter (*nrec, one atomic per
// tile) and writes per record its sort key as k_sample_keys makes it (packed with the slot when
// vbits > 0, else the slot in rval) and node * 27 + sample of the owning entry (rent), per entry
// the slot of its record (ent_rec). The order of the tiles' slot ranges varies from run to run;
// what is evaluated does not: equal points get equal values whichever record is their head.
constexpr int kStT = 256;           // nodes per tile
constexpr int kStRecs = 27 * kStT;  // records per tile at most
constexpr int kStMap = 4096;        // dense map cells

__global__ void __launch_bounds__(kStT) k_sample_tiles(GridDev g, NodeLevel lv, const int* __restrict__ list,
                                                       int64_t k, float ox, float oy, float oz, int lb, int vbits,
                                                       int* __restrict__ nrec, int* __restrict__ ent_rec,
                                                       uint64_t* __restrict__ rkey, int* __restrict__ rval,
                                                       int* __restrict__ rent) {
    __shared__ short map[kStMap];
    __shared__ float4 cen[kStT];
    __shared__ int nid[kStT];
    __shared__ short ref[kStRecs];    // per tile entry (node * 27 + sample): the entry it refers to, or -1
    __shared__ short recid[kStRecs];  // record of an entry that owns one
    __shared__ int red[6][kStT / kWave];
    __shared__ int wsum[kStT / kWave];
    __shared__ int tbase;
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    const int64_t l0 = (int64_t)blockIdx.x * kStT, li = l0 + t;
    const bool act = li < k;
    int4 ip = make_int4(0, 0, 0, 0);
    float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
    if (act) {
        const int node = list[li];
        c = lv.c[node];
        ip = lv.ip[node];
        cen[t] = c;
        nid[t] = node;
    }
    // the tile's bounding box of node coordinates
    int mn[3] = {act ? ip.x : INT_MAX, act ? ip.y : INT_MAX, act ? ip.z : INT_MAX};
    int mx[3] = {act ? ip.x : INT_MIN, act ? ip.y : INT_MIN, act ? ip.z : INT_MIN};
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mn[a] = min(mn[a], __shfl_xor(mn[a], o, kWave));
            mx[a] = max(mx[a], __shfl_xor(mx[a], o, kWave));
        }
    if (lane == 0)
        for (int a = 0; a < 3; a++) {
            red[a][w] = mn[a];
            red[3 + a][w] = mx[a];
        }
    for (int q = t; q < kStMap; q += kStT) map[q] = -1;
    __syncthreads();
    int b0[3], dim[3];
    bool dense = true;  // uniform over the tile
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int lo = red[a][0], hi = red[3 + a][0];
        for (int q = 1; q < kStT / kWave; q++) {
            lo = min(lo, red[a][q]);
            hi = max(hi, red[3 + a][q]);
        }
        b0[a] = lo - 1;  // one node of margin: the neighbours looked up
        dim[a] = hi - lo + 3;
        if (dim[a] > kStMap) dense = false;
    }
    if (dense && dim[0] * dim[1] * dim[2] > kStMap) dense = false;
    if (dense && act) map[(ip.x - b0[0]) + dim[0] * ((ip.y - b0[1]) + dim[1] * (ip.z - b0[2]))] = (short)t;
    __syncthreads();
    int own = 0;
    if (act) {
        const int m0 = (ip.x - b0[0]) + dim[0] * ((ip.y - b0[1]) + dim[1] * (ip.z - b0[2]));
#pragma unroll
        for (int sm = 0; sm < 27; sm++) {
            const int ta[3] = {sm % 3, (sm / 3) % 3, sm / 9};
            int best = t, bs = sm;
            if (dense) {
#pragma unroll
                for (int d = 1; d < 8; d++) {
                    int dd[3];
                    bool ok = true, earlier = false;
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        const int bit = (d >> a) & 1;
                        dd[a] = bit ? (ta[a] == 0 ? -1 : (ta[a] == 2 ? 1 : 0)) : 0;
                        if (bit && dd[a] == 0) ok = false;
                        if (dd[a] < 0) earlier = true;
                    }
                    if (!ok || !earlier) continue;
                    const int j = map[m0 + dd[0] + dim[0] * (dd[1] + dim[1] * dd[2])];
                    if (j >= 0 && j < best) {
                        best = j;
                        bs = (ta[0] - 2 * dd[0]) + 3 * (ta[1] - 2 * dd[1]) + 9 * (ta[2] - 2 * dd[2]);
                    }
                }
            }
            short r = -1;
            if (best != t) {
                float p[3], q[3];
                sample_point(c, sm, p);
                sample_point(cen[best], bs, q);
                if (same_point(p, q)) r = (short)(best * 27 + bs);
            }
            ref[t * 27 + sm] = r;
            own += r < 0 ? 1 : 0;
        }
    }
    // records numbered in entry order (thread, then sample)
    const int incl = wave_incl_scan(own, lane);
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int u = incl - own, total = 0;
#pragma unroll
    for (int q = 0; q < kStT / kWave; q++) {
        if (q < w) u += wsum[q];
        total += wsum[q];
    }
    if (t == 0) tbase = atomicAdd(nrec, total);
    if (act)
        for (int sm = 0; sm < 27; sm++)
            if (ref[t * 27 + sm] < 0) recid[t * 27 + sm] = (short)u++;
    __syncthreads();
    // entry by entry (coalesced): the record slot of every entry, the key and entry of every record
    const int ne = (int)min<int64_t>(k - l0, kStT) * 27;
    const float o[3] = {ox, oy, oz};
    const uint32_t m = (1u << lb) - 1;
    for (int e = t; e < ne; e += kStT) {
        int r = e;
        while (ref[r] >= 0) r = ref[r];
        const int slot = tbase + recid[r];
        ent_rec[l0 * 27 + e] = slot;
        if (r != e) continue;
        const float4 cc = cen[e / 27];
        float p[3];
        sample_point(cc, e % 27, p);
        uint32_t qq[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float v = rintf((p[a] - o[a]) / cc.w);
            qq[a] = (uint32_t)fminf(fmaxf(v, 0.0f), 2097151.0f);
        }
        const uint64_t kk = (cell_group(g, p[0], p[1], p[2]) << (3 * lb)) | (qq[0] & m) | ((qq[1] & m) << lb) |
                            ((uint64_t)(qq[2] & m) << (2 * lb));
        if (vbits) {
            rkey[slot] = (kk << vbits) | (uint64_t)slot;
        } else {
            rkey[slot] = kk;
            rval[slot] = slot;
        }
        rent[slot] = nid[e / 27] * 27 + e % 27;
    }
}

// the head (index into the evaluated points) of every record slot: rep[j] is the first equal
// record within j's 64-record window, and a longer run can chain further back (rep[x] <= x)
__global__ void k_record_heads(SortedKV kv, const int* __restrict__ rep, const int* __restrict__ hpos, int64_t nrec,
                               int* __restrict__ rec_head) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nrec) return;
    int r = rep[j];
    while (rep[r] != r) r = rep[r];
    rec_head[kv.val(j)] = hpos[r];
}

// every sample entry receives its point's value
__global__ void k_entry_scatter(const int* __restrict__ list, const int* __restrict__ ent_rec,
                                const int* __restrict__ rec_head, int64_t nent, const void* __restrict__ vals, int rsz,
                                void* __restrict__ samples) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    const int64_t dst = (int64_t)list[e / 27] * 27 + e % 27;
    const int64_t src = rec_head[ent_rec[e]];
    if (rsz == 8) ((double*)samples)[dst] = ((const double*)vals)[src];
    else ((float*)samples)[dst] = ((const float*)vals)[src];
}

void launch_sample_tiles(GridDev g, NodeLevel lv, const int* list, int64_t k, const float origin[3], int lb,
                         int vbits, int* nrec, int* ent_rec, uint64_t* rkey, int* rval, int* rent, hipStream_t s) {
    launch_fill_i32(nrec, 1, 0, s);
    if (k > 0)
        k_sample_tiles<<<(unsigned)((k + kStT - 1) / kStT), kStT, 0, s>>>(g, lv, list, k, origin[0], origin[1],
                                                                        origin[2], lb, vbits, nrec, ent_rec, rkey,
                                                                        rval, rent);
}

// distinct sample points of a level (cell-major), evaluated run by run
template <class R> struct SampleSrc {
    NodeLevel lv;
    const int* list;
    SortedKV kv;
    const int* hlist;
    R* out;
    R iso, sigma;
    const int* ent = nullptr;  // sorted item -> node * 27 + sample (nullptr: the items are the list's entries)
    __device__ __forceinline__ void point(int h, float p[3]) const { item_point(lv, list, ent, kv.val(hlist[h]), p); }
    __device__ __forceinline__ void store(int h, R sum, const R*) const { out[h] = iso - sigma * sum; }
};

template <class R, bool K, int T, int MODE>
__global__ void __launch_bounds__(T) k_run_samples(GridDev g, Recs<R> rec, FieldConst fc, NodeLevel lv,
                                                             const int* __restrict__ list, SortedKV kv,
                                                             const int* __restrict__ hlist,
                                                             const int* __restrict__ rstart, R* __restrict__ out,
                                                             const int* __restrict__ ent, Counters* ctr) {
    __shared__ RunLds<R, K && sizeof(R) == 4> lds;
    const int h0 = rstart[blockIdx.x], h1 = rstart[blockIdx.x + 1];
    SampleSrc<R> src{lv, list, kv, hlist, out, R(fc.iso), out_sigma<R, K>(fc), ent};
    const unsigned long long vis = run_eval<T, R, K, false, MODE>(g, rec, fc, src, h0, h1, lds);
    Counters* c = ctr_stripe(ctr);
    wave_count(&c->sample_candidates, vis);
    wave_count(&c->candidates, vis);
    wave_count(&c->run_candidates, vis);
    if (threadIdx.x == 0) {
        atomicAdd(&c->sample_unique, (unsigned long long)(h1 - h0));
        atomicAdd(&c->run_evals, (unsigned long long)(h1 - h0));
    }
}

// every entry receives its point's value (the host counts the entries: 27 per sampled node, the
// reference's evaluations)
__global__ void k_sample_scatter(const int* __restrict__ list, SortedKV kv, const int* __restrict__ rep,
                                 const int* __restrict__ hpos, int64_t nent,
                                 const void* __restrict__ vals, int rsz, void* __restrict__ samples) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nent) return;
    const int e = kv.val(j);
    const int64_t dst = (int64_t)list[e / 27] * 27 + e % 27;
    // rep[j] is the first equal entry within j's 64-entry window; in a longer run of equal keys
    // that entry can itself have an earlier copy (its window reaches further back), so follow the
    // chain to the head (rep[x] <= x, equal only at a head; hpos is written for heads only)
    int r = rep[j];
    while (rep[r] != r) r = rep[r];
    const int64_t src = hpos[r];
    if (rsz == 8) ((double*)samples)[dst] = ((const double*)vals)[src];
    else ((float*)samples)[dst] = ((const float*)vals)[src];
}

// centre values of leaf-depth nodes (sample slot 13), one thread per listed node
template <class R, bool K>
__global__ void __launch_bounds__(256) k_centers(GridDev g, Recs<R> rec, FieldConst fc, NodeLevel lv,
                                                 const int* __restrict__ list, const int* __restrict__ count,
                                                 R* __restrict__ samples, Counters* ctr) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int vis = 0;
    bool did = false;
    if (k < *count) {
        const int64_t node = list[k];
        const float4 c = lv.c[node];
        R sum, gr[3];
        vis = field_sum<R, K, false>(g, rec, fc, c.x, c.y, c.z, R(c.x), R(c.y), R(c.z), sum, gr);
        samples[node * 27 + 13] = R(fc.iso) - out_sigma<R, K>(fc) * sum;
        did = true;
    }
    wave_count(&ctr_stripe(ctr)->field_evals, did ? 1ull : 0ull);
    wave_count(&ctr_stripe(ctr)->candidates, (unsigned long long)vis);
}
__global__ void k_sample_list(NodeLevel lv, const int* __restrict__ flag, const int* __restrict__ pos,
                              int* __restrict__ list) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < lv.n && flag[i]) list[pos[i]] = (int)i;
}

// 27 samples of each listed node; thread t handles sample t % 27 of list entry t / 27, so a wave
// holds samples of 2-3 spatially adjacent nodes and no idle lanes for skipped nodes
template <class R, bool K>
__global__ void __launch_bounds__(256) k_samples(GridDev g, Recs<R> rec, FieldConst fc, TreeParams tp, NodeLevel lv,
                                                 const int* __restrict__ list, const int* __restrict__ count,
                                                 R* __restrict__ samples, Counters* ctr) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t k = t / 27;
    const int s = (int)(t - k * 27);
    int vis = 0;
    bool did = false;
    if (k < *count) {
        const int64_t node = list[k];
        float p[3];
        sample_point(lv.c[node], s, p);
        R sum, gr[3];
        vis = field_sum<R, K, false>(g, rec, fc, p[0], p[1], p[2], R(p[0]), R(p[1]), R(p[2]), sum, gr);
        samples[node * 27 + s] = R(fc.iso) - out_sigma<R, K>(fc) * sum;
        did = true;
    }
    wave_count(&ctr_stripe(ctr)->field_evals, did ? 1ull : 0ull);
    wave_count(&ctr_stripe(ctr)->candidates, (unsigned long long)vis);
    wave_count(&ctr_stripe(ctr)->sample_evals, did ? 1ull : 0ull);
    wave_count(&ctr_stripe(ctr)->sample_candidates, (unsigned long long)vis);
}

// afterSampleEval classification (src/surface_reconstructor.cpp:237-273) + leaf values
template <class R, bool K, bool QEF>
__global__ void __launch_bounds__(256) k_classify(GridDev g, Recs<R> rec, FieldConst fc, TreeParams tp, NodeLevel lv,
                                                  const R* __restrict__ samples, int32_t* __restrict__ sc_out,
                                                  Counters* ctr) {
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int vis = 0;
    unsigned long long evals = 0, qefs = 0;
    const int64_t count = lv_count(lv);
    if (node < count) {
        const float4 c = lv.c[node];
        const int depth = lv.depth[node];
        uint8_t type;
        R fval = R(fc.iso);
        R vx = R(c.x), vy = R(c.y), vz = R(c.z);
        bool sc = false;
        if (lv.empty[node]) {
            type = kEmpty;
        } else {
            const bool isbig = depth < tp.depth_min;
            const bool has_samples = wants_samples(lv, tp, node);
            if (has_samples) {
                const R* f = samples + node * 27;
                const bool s0 = f[0] >= R(0);
                for (int s = 1; s < 27; s++) sc = sc || ((f[s] >= R(0)) != s0);
            }
            const bool leafd = leaf_depth(c.w, tp.radius);
            if (!leafd && (isbig || (sc && double(lv.curv[node]) < 0.99))) {
                type = kInternal;
            } else {
                type = kLeaf;
                if (QEF && tp.qef_mode != 0 && has_samples) {
                    R f[27];
#pragma unroll
                    for (int s = 0; s < 27; s++) f[s] = samples[node * 27 + s];
                    R pc[4], err;
                    if (qef_solve_samples<R>(c, f, tp.qef_mode, tp.qef_svd_tol, pc, err)) {
                        vx = pc[0];
                        vy = pc[1];
                        vz = pc[2];
                    }
                    qefs++;
                    // the field at the vertex is summed by k_qef_values (fewer registers than the
                    // solve needs); until then the value slot says whether the DMC error zeroes it
                    fval = (err / R(__fmul_rn(2.0f, c.w)) < R(0.0001)) ? R(0) : R(1);
                } else {
                    float p13[3];
                    sample_point(c, 13, p13);
                    if (wants_center(lv, tp, node)) {
                        fval = samples[node * 27 + 13];  // evaluated by k_centers
                    } else if (has_samples && p13[0] == c.x && p13[1] == c.y && p13[2] == c.z) {
                        fval = samples[node * 27 + 13];  // the centre sample is the node centre
                    } else {
                        R sum, g3[3];
                        vis += field_sum<R, K, false, kModeAny, true>(g, rec, fc, c.x, c.y, c.z, R(c.x), R(c.y),
                                                                     R(c.z), sum, g3);
                        evals++;
                        fval = R(fc.iso) - out_sigma<R, K>(fc) * sum;
                    }
                }
            }
        }
        lv.type[node] = type;
        lv.flag[node] = (type == kInternal) ? 1 : 0;
        lv.node[node] = make_double4(double(vx), double(vy), double(vz), double(fval));
        lv.child[node] = -1;
        if (sc_out) sc_out[node] = sc ? 1 : 0;
    } else if (node <= lv.n) {
        lv.flag[node] = 0;  // past a device count: nothing to scan; flag[n] closes the flags for their scan
    }
    wave_count(&ctr_stripe(ctr)->field_evals, evals);
    wave_count(&ctr_stripe(ctr)->qef_solves, qefs);
    wave_count(&ctr_stripe(ctr)->candidates, (unsigned long long)vis);
}

// Field values at the QEF vertices of a level's sampled leaves (k_classify<QEF> placed them and
// left 0 in the value slot where the DMC error zeroes the value, 1 elsewhere)
template <class R, bool K>
__global__ void __launch_bounds__(256) k_qef_values(GridDev g, Recs<R> rec, FieldConst fc, TreeParams tp, NodeLevel lv,
                                                    Counters* ctr) {
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int vis = 0;
    unsigned long long evals = 0;
    if (node < lv_count(lv) && lv.type[node] == kLeaf && wants_samples(lv, tp, node)) {
        const double4 nd = lv.node[node];
        const R vx = R(nd.x), vy = R(nd.y), vz = R(nd.z);
        R sum, g3[3];
        vis = field_sum<R, K, false, kModeAny, true>(g, rec, fc, float(vx), float(vy), float(vz), vx, vy, vz, sum, g3);
        evals = 1;
        const R fval = nd.w == 0.0 ? R(0) : R(fc.iso) - out_sigma<R, K>(fc) * sum;
        lv.node[node].w = double(fval);
    }
    wave_count(&ctr_stripe(ctr)->field_evals, evals);
    wave_count(&ctr_stripe(ctr)->candidates, (unsigned long long)vis);
}

// Children of INTERNAL nodes (TNode ctor, src/iso_method_ours.cpp:25-38): child t has
// centre parent +/- half_child per axis (bit0 = x, bit1 = y, bit2 = z) and id = parent * 8 + t.
__global__ void k_children(NodeLevel par, const int* __restrict__ scanned, NodeLevel nxt, int* __restrict__ next_count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (next_count && i == par.n) *next_count = 8 * scanned[par.n];
    if (i >= lv_count(par) || !par.flag[i]) return;
    const int64_t base = (int64_t)scanned[i] * 8;
    if (base + 8 > nxt.n) return;  // past the next level's arrays (its count is then out of bound)
    par.child[i] = (int)base;
    const float4 c = par.c[i];
    const float hc = c.w / 2.0f;
    const uint64_t id = par.id[i];
    const int8_t d = par.depth[i] + 1;
    const int4 pi = par.ip[i];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const float sx = (t & 1) ? hc : -hc, sy = (t & 2) ? hc : -hc, sz = (t & 4) ? hc : -hc;
        nxt.c[base + t] = make_float4(__fadd_rn(c.x, sx), __fadd_rn(c.y, sy), __fadd_rn(c.z, sz), hc);
        nxt.id[base + t] = id * 8ull + (uint64_t)t;
        nxt.depth[base + t] = d;
        nxt.ip[base + t] = make_int4(2 * pi.x + (t & 1), 2 * pi.y + ((t >> 1) & 1), 2 * pi.z + ((t >> 2) & 1), d);
    }
}

// samples and their finite-difference gradients for the sweep API
template <class R>
__global__ void k_sample_grads(TreeParams tp, NodeLevel lv, const R* __restrict__ samples, double* __restrict__ os,
                               double* __restrict__ og) {
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= lv.n) return;
    const float4 c = lv.c[node];
    const bool has = wants_samples(lv, tp, node);
    R f[27], gr[27][3];
    for (int s = 0; s < 27; s++) f[s] = has ? samples[node * 27 + s] : R(0);
    fd_grads27<R>(f, __fmul_rn(2.0f, c.w), gr);
    for (int s = 0; s < 27; s++) {
        float p[3];
        sample_point(c, s, p);
        if (os) {
            os[(node * 27 + s) * 4 + 0] = p[0];
            os[(node * 27 + s) * 4 + 1] = p[1];
            os[(node * 27 + s) * 4 + 2] = p[2];
            os[(node * 27 + s) * 4 + 3] = double(f[s]);
        }
        if (og)
            for (int a = 0; a < 3; a++) og[(node * 27 + s) * 3 + a] = has ? double(gr[s][a]) : 0.0;
    }
}

// ---- phase-split deduplicated sampling and the multi-frame run queue --------------------------
// The same steps as launch_sample_dedupe_eval / launch_centers_runs, cut at every count the host
// must read, so that a sweep over many frames can issue one phase for every frame and then read
// every frame's count at once.

int* heads_count(int* head, int64_t nent) {
    const int64_t nb = (nent + 255) / 256;
    return head + (nb + 1) + nb;
}

void launch_heads(NodeLevel lv, const int* list, SortedKV kv, int64_t nent, int* head, int* rep, int* hlist, int* hpos,
                  int* tmp, hipStream_t s) {
    if (nent <= 0) return;
    const unsigned nb = (unsigned)((nent + 255) / 256);
    int* block_off = head + (nb + 1);
    k_sample_heads<<<nb, 256, 0, s>>>(lv, list, kv, nullptr, nent, head, rep);
    exclusive_scan_i32(head, block_off, nb + 1, tmp, s);
    k_head_list<<<nb, 256, 0, s>>>(rep, nent, block_off, hlist, hpos);
}

int* groups_count(int* cwork, int nh) { return cwork + (nh + 2) + nh; }

void launch_cell_groups(SortedKV kv, int gshift, const int* hlist, int nh, int* cwork, int* tmp, hipStream_t s) {
    if (nh <= 0) return;
    int* flag = cwork;
    int* pos = cwork + (nh + 2);
    int* cstart = cwork + 2 * (int64_t)(nh + 2);
    const unsigned nbh = (unsigned)((nh + 1 + 255) / 256);
    k_cell_flags<<<nbh, 256, 0, s>>>(kv, gshift, hlist, nh, flag);
    exclusive_scan_i32(flag, pos, nh + 1, tmp, s);
    k_cell_starts<<<nbh, 256, 0, s>>>(flag, pos, nh, cstart);
}

int* runs_count(int* cwork, int nh, int ng) { return cwork + (nh + 2) + ng; }
int* runs_start(int* cwork, int nh, int ng) { return cwork + 2 * (int64_t)(nh + 2) + (ng + 1); }

void launch_run_starts(GridDev g, SortedKV kv, int gshift, const int* hlist, int nh, int ng, int kx, int* cwork,
                       int* tmp, hipStream_t s) {
    if (nh <= 0 || ng <= 0) return;
    int* flag = cwork;
    int* pos = cwork + (nh + 2);
    int* cstart = cwork + 2 * (int64_t)(nh + 2);
    const unsigned nbg = (unsigned)((ng + 1 + 255) / 256);
    k_run_flags<<<nbg, 256, 0, s>>>(g, kv, gshift, hlist, cstart, ng, kx, flag);
    exclusive_scan_i32(flag, pos, ng + 1, tmp, s);
    k_run_starts<<<nbg, 256, 0, s>>>(flag, pos, cstart, ng, cstart + (ng + 1));
}

void run_shape_for(double points_per_cell, int* kx, int* t) {
    const RunShape sh = run_shape_guess(points_per_cell);
    *kx = sh.kx;
    *t = sh.T;
}

void launch_sample_scatter(const int* list, SortedKV kv, const int* rep, const int* hpos, int64_t nent,
                           const void* vals, int rsz, void* samples, hipStream_t s) {
    if (nent <= 0) return;
    k_sample_scatter<<<(unsigned)((nent + 255) / 256), 256, 0, s>>>(list, kv, rep, hpos, nent, vals, rsz, samples);
}

SortedKV launch_center_keys_sorted(GridDev g, NodeLevel lv, const int* list, int64_t k, uint64_t* key, int* val,
                                   uint64_t* key2, int* val2, int* tmp, hipStream_t s) {
    const uint64_t groups = (uint64_t)(g.n[0] + 2) * (uint64_t)(g.n[1] + 2) * (uint64_t)(g.n[2] + 2);
    int bits = 1;
    while (bits < 55 && ((uint64_t)1 << bits) <= groups) bits++;
    const int vbits = pack_vbits(bits, k);
    if (k <= 0) return SortedKV{key, val, vbits};
    k_center_keys<<<(unsigned)((k + 255) / 256), 256, 0, s>>>(g, lv, list, k, vbits, key, val);
    return radix_sort_entries(key, val, key2, val2, k, bits, vbits, tmp, s);
}

// items[i] = (frame, run): frame f owns items [off[f], off[f + 1])
__global__ void k_queue_items(const int* __restrict__ off, int nframes, int nitems, int2* __restrict__ items) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitems) return;
    int lo = 0, hi = nframes - 1;
    while (lo < hi) {  // the last frame whose offset is <= i
        const int mid = (lo + hi + 1) / 2;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    items[i] = make_int2(lo, i - off[lo]);
}

// One launch for the runs of every frame of a level: each workgroup takes the next (frame, run)
// item from a device counter until none is left (every workgroup ends when the counter passes
// the item count), and evaluates that run with the frame's own grid, records and constants.
template <class R, bool K, int T, int MODE>
__global__ void __launch_bounds__(T) k_run_queue(const RunFrame* __restrict__ frames, const int2* __restrict__ items,
                                                 int nitems, int* __restrict__ qhead) {
    __shared__ RunLds<R, K && sizeof(R) == 4> lds;
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(qhead, 1);
        __syncthreads();
        const int it = next;
        __syncthreads();
        if (it >= nitems) return;
        const int2 w = items[it];
        const RunFrame& F = frames[w.x];
        const Recs<R> rec{(const typename Vec4<R>::T*)F.ra, (const typename Vec4<R>::T*)F.rg};
        const int h0 = F.rstart[w.y], h1 = F.rstart[w.y + 1];
        Counters* c = ctr_stripe(F.ctr);
        if (F.kind == 0) {
            SampleSrc<R> src{F.lv, F.list, F.kv, F.hlist, (R*)F.out, R(F.fc.iso), out_sigma<R, K>(F.fc)};
            const unsigned long long vis = run_eval<T, R, K, false, MODE>(F.g, rec, F.fc, src, h0, h1, lds);
            wave_count(&c->sample_candidates, vis);
            wave_count(&c->candidates, vis);
            wave_count(&c->run_candidates, vis);
            if (threadIdx.x == 0) {
                atomicAdd(&c->sample_unique, (unsigned long long)(h1 - h0));
                atomicAdd(&c->run_evals, (unsigned long long)(h1 - h0));
            }
        } else {
            CenterSrc<R> src{F.lv, F.list, F.kv, (R*)F.out, R(F.fc.iso), out_sigma<R, K>(F.fc)};
            const unsigned long long vis = run_eval<T, R, K, false, MODE>(F.g, rec, F.fc, src, h0, h1, lds);
            wave_count(&c->candidates, vis);
            if (threadIdx.x == 0) atomicAdd(&c->field_evals, (unsigned long long)(h1 - h0));
        }
    }
}

void launch_run_queue(int prec, int mode, int T, const RunFrame* frames, const int* item_off, int nframes, int nitems,
                      int2* items, int* qhead, int max_blocks, hipStream_t s) {
    if (nitems <= 0) return;
    k_queue_items<<<(unsigned)((nitems + 255) / 256), 256, 0, s>>>(item_off, nframes, nitems, items);
    launch_fill_i32(qhead, 1, 0, s);
    const unsigned nb = (unsigned)std::min(nitems, std::max(max_blocks, 1));
    auto go = [&](auto tag, auto md) {
        constexpr int TT = decltype(tag)::value, M = decltype(md)::value;
        if (prec == 2) k_run_queue<double, false, TT, M><<<nb, TT, 0, s>>>(frames, items, nitems, qhead);
        else if (prec == 1) k_run_queue<float, true, TT, M><<<nb, TT, 0, s>>>(frames, items, nitems, qhead);
        else k_run_queue<float, false, TT, M><<<nb, TT, 0, s>>>(frames, items, nitems, qhead);
    };
    using M0 = std::integral_constant<int, kModeAny>;
    using M1 = std::integral_constant<int, kModeAniBspline>;
    if (T == 256) mode == kModeAniBspline ? go(std::integral_constant<int, 256>(), M1()) : go(std::integral_constant<int, 256>(), M0());
    else mode == kModeAniBspline ? go(std::integral_constant<int, 128>(), M1()) : go(std::integral_constant<int, 128>(), M0());
}

void launch_empty_curv(GridDev g, const float4* spos, const float4* nrm, TreeParams tp, NodeLevel lv, int group,
                       int pass, const void* samples, int rsz, hipStream_t s) {
    if (lv.n <= 0) return;
    const unsigned nb4 = (unsigned)((lv.n * 4 + 255) / 256), nb8 = (unsigned)((lv.n * 8 + 255) / 256);
    const unsigned nb16 = (unsigned)((lv.n * 16 + 255) / 256);
    const unsigned nb64 = (unsigned)((lv.n * kWave + 255) / 256);
    auto go = [&](auto tag) {
        constexpr int P = decltype(tag)::value;
        if (group <= 4) k_empty_curv<4, P><<<nb4, 256, 0, s>>>(g, spos, nrm, tp, lv, samples, rsz);
        else if (group <= 8) k_empty_curv<8, P><<<nb8, 256, 0, s>>>(g, spos, nrm, tp, lv, samples, rsz);
        else if (group <= 16) k_empty_curv<16, P><<<nb16, 256, 0, s>>>(g, spos, nrm, tp, lv, samples, rsz);
        else k_empty_curv<64, P><<<nb64, 256, 0, s>>>(g, spos, nrm, tp, lv, samples, rsz);
    };
    if (pass == kEmptyOnly) go(std::integral_constant<int, kEmptyOnly>());
    else if (pass == kCurvOnly) go(std::integral_constant<int, kCurvOnly>());
    else go(std::integral_constant<int, kEmptyAndCurv>());
}

void launch_sample_list(TreeParams tp, NodeLevel lv, int* flag, int* pos, int* list, int* cflag, int* cpos,
                        int* clist, int* tmp, hipStream_t s, ScanPub pub, ScanPub cpub) {
    if (lv.n <= 0) return;
    k_sample_flag<<<(unsigned)((lv.n + 1 + 255) / 256), 256, 0, s>>>(tp, lv, flag, cflag);
    exclusive_scan_i32(flag, pos, lv.n + 1, tmp, s, pub);
    k_sample_list<<<(unsigned)((lv.n + 255) / 256), 256, 0, s>>>(lv, flag, pos, list);
    exclusive_scan_i32(cflag, cpos, lv.n + 1, tmp, s, cpub);
    k_sample_list<<<(unsigned)((lv.n + 255) / 256), 256, 0, s>>>(lv, cflag, cpos, clist);
}

void launch_sample_dedupe_keys(GridDev g, NodeLevel lv, const int* list, int64_t nent, const float origin[3],
                               int lb, int vbits, uint64_t* key, int* val, hipStream_t s) {
    if (nent <= 0) return;
    k_sample_keys<<<(unsigned)((nent + 255) / 256), 256, 0, s>>>(g, lv, list, nent, origin[0], origin[1], origin[2],
                                                                 lb, vbits, key, val);
}

// one launch of the sample runs with a given block size (every run one workgroup)
static void launch_run_samples(int prec, int T, GridDev g, const void* ra, const void* rg, FieldConst fc, NodeLevel lv,
                               const int* list, SortedKV kv, const int* hlist, const int* rstart, void* vals, int nr,
                               const int* ent, Counters* ctr, hipStream_t s) {
    auto go = [&](auto tag, auto mode) {
        constexpr int TT = decltype(tag)::value, M = decltype(mode)::value;
        if (prec == 2) {
            Recs<double> r{(const double4*)ra, (const double4*)rg};
            k_run_samples<double, false, TT, M><<<nr, TT, 0, s>>>(g, r, fc, lv, list, kv, hlist, rstart, (double*)vals,
                                                                  ent, ctr);
        } else {
            Recs<float> r{(const float4*)ra, (const float4*)rg};
            if (prec == 1)
                k_run_samples<float, true, TT, M><<<nr, TT, 0, s>>>(g, r, fc, lv, list, kv, hlist, rstart, (float*)vals,
                                                                    ent, ctr);
            else
                k_run_samples<float, false, TT, M><<<nr, TT, 0, s>>>(g, r, fc, lv, list, kv, hlist, rstart,
                                                                     (float*)vals, ent, ctr);
        }
    };
    using M0 = std::integral_constant<int, kModeAny>;
    using M1 = std::integral_constant<int, kModeAniBspline>;
    const bool m1 = field_mode(fc) == kModeAniBspline;
    if (T == 256) m1 ? go(std::integral_constant<int, 256>(), M1()) : go(std::integral_constant<int, 256>(), M0());
    else m1 ? go(std::integral_constant<int, 128>(), M1()) : go(std::integral_constant<int, 128>(), M0());
}

void launch_sample_dedupe_eval(int prec, GridDev g, const void* ra, const void* rg, FieldConst fc, NodeLevel lv,
                               const int* list, SortedKV kv, const int* ent, int gshift, int64_t nrec,
                               const int* ent_rec, int64_t nent, int* rec_head, int* head, int* hpos, int* rep,
                               int* hlist, int* cwork, int64_t* nheads_out, int64_t* nruns_out, int* kx_out,
                               void* vals, void* samples, int* tmp, Counters* ctr, Counters* trial_ctr, hipEvent_t e0,
                               hipEvent_t e1, HostRead* rd, hipStream_t s, hipError_t* err) {
    *nheads_out = 0;
    *nruns_out = 0;
    *kx_out = 0;
    if (nrec <= 0) return;
    // head holds the per-block head counts (nb + 1) and their scan (nb + 1)
    const unsigned nb = (unsigned)((nrec + 255) / 256);
    int* block_off = head + (nb + 1);
    k_sample_heads<<<nb, 256, 0, s>>>(lv, list, kv, ent, nrec, head, rep);
    const ScanPub pub = scan_read_begin(rd);
    exclusive_scan_i32(head, block_off, nb + 1, tmp, s, pub);
    k_head_list<<<nb, 256, 0, s>>>(rep, nrec, block_off, hlist, hpos);
    int nh = 0;
    const hipError_t e = scan_read_end(&nh, block_off + nb, pub, s, rd);
    if (e != hipSuccess) {
        *err = e;
        return;
    }
    *nheads_out = nh;
    const int ng = build_groups(kv, gshift, hlist, nh, cwork, tmp, rd, s, err);
    int* rstart = nullptr;
    int nr = 0;
    (void)hipEventRecord(e0, s);
    if (ng > 0) {
        const RunShape sh = tuned_run_launch(
            prec, kRunSamples, (double)nh / ng, trial_ctr, s,
            [&](RunShape c) {
                nr = build_run_starts(g, kv, gshift, hlist, nh, ng, c.kx, cwork, tmp, &rstart, rd, s, err);
                return nr > 0;
            },
            [&](RunShape c, bool trial) {
                launch_run_samples(prec, c.T, g, ra, rg, fc, lv, list, kv, hlist, rstart, vals, nr, ent,
                                   trial ? trial_ctr : ctr, s);
            });
        *kx_out = sh.kx;
    }
    *nruns_out = nr;
    (void)hipEventRecord(e1, s);
    if (!ent_rec) {
        k_sample_scatter<<<nb, 256, 0, s>>>(list, kv, rep, hpos, nrec, vals, prec == 2 ? 8 : 4, samples);
        return;
    }
    k_record_heads<<<nb, 256, 0, s>>>(kv, rep, hpos, nrec, rec_head);
    k_entry_scatter<<<(unsigned)((nent + 255) / 256), 256, 0, s>>>(list, ent_rec, rec_head, nent, vals,
                                                                   prec == 2 ? 8 : 4, samples);
}

void launch_centers(int prec, GridDev g, const void* ra, const void* rg, FieldConst fc, NodeLevel lv, const int* list,
                    const int* count, void* samples, Counters* ctr, hipStream_t s) {
    if (lv.n <= 0) return;
    const unsigned nb = (unsigned)((lv.n + 255) / 256);
    if (prec == 2) {
        Recs<double> r{(const double4*)ra, (const double4*)rg};
        k_centers<double, false><<<nb, 256, 0, s>>>(g, r, fc, lv, list, count, (double*)samples, ctr);
    } else {
        Recs<float> r{(const float4*)ra, (const float4*)rg};
        if (prec == 1) k_centers<float, true><<<nb, 256, 0, s>>>(g, r, fc, lv, list, count, (float*)samples, ctr);
        else k_centers<float, false><<<nb, 256, 0, s>>>(g, r, fc, lv, list, count, (float*)samples, ctr);
    }
}

void launch_centers_runs(int prec, GridDev g, const void* ra, const void* rg, FieldConst fc, NodeLevel lv,
                         const int* list, int64_t k, uint64_t* key, int* val, uint64_t* key2, int* val2, int* cwork,
                         int* tmp, void* samples, Counters* ctr, Counters* trial_ctr, HostRead* rd,
                         hipStream_t s, hipError_t* err) {
    if (k <= 0) return;
    const uint64_t groups = (uint64_t)(g.n[0] + 2) * (uint64_t)(g.n[1] + 2) * (uint64_t)(g.n[2] + 2);
    int bits = 1;
    while (bits < 55 && ((uint64_t)1 << bits) <= groups) bits++;
    const int vbits = pack_vbits(bits, k);
    k_center_keys<<<(unsigned)((k + 255) / 256), 256, 0, s>>>(g, lv, list, k, vbits, key, val);
    const SortedKV kv = radix_sort_entries(key, val, key2, val2, k, bits, vbits, tmp, s);
    const int ng = build_groups(kv, 0, nullptr, (int)k, cwork, tmp, rd, s, err);
    if (ng <= 0) return;
    int* rstart = nullptr;
    int nr = 0;
    auto launch = [&](RunShape c, bool trial) {
        Counters* cc = trial ? trial_ctr : ctr;
        auto go = [&](auto tag, auto mode) {
            constexpr int TT = decltype(tag)::value, M = decltype(mode)::value;
            if (prec == 2) {
                Recs<double> r{(const double4*)ra, (const double4*)rg};
                k_run_centers<double, false, TT, M><<<nr, TT, 0, s>>>(g, r, fc, lv, list, kv, rstart,
                                                                      (double*)samples, cc);
            } else {
                Recs<float> r{(const float4*)ra, (const float4*)rg};
                if (prec == 1)
                    k_run_centers<float, true, TT, M><<<nr, TT, 0, s>>>(g, r, fc, lv, list, kv, rstart,
                                                                        (float*)samples, cc);
                else
                    k_run_centers<float, false, TT, M><<<nr, TT, 0, s>>>(g, r, fc, lv, list, kv, rstart,
                                                                         (float*)samples, cc);
            }
        };
        using M0 = std::integral_constant<int, kModeAny>;
        using M1 = std::integral_constant<int, kModeAniBspline>;
        const bool m1 = field_mode(fc) == kModeAniBspline;
        if (c.T == 256) m1 ? go(std::integral_constant<int, 256>(), M1()) : go(std::integral_constant<int, 256>(), M0());
        else m1 ? go(std::integral_constant<int, 128>(), M1()) : go(std::integral_constant<int, 128>(), M0());
    };
    tuned_run_launch(
        prec, kRunCenters, (double)k / ng, trial_ctr, s,
        [&](RunShape c) {
            nr = build_run_starts(g, kv, 0, nullptr, (int)k, ng, c.kx, cwork, tmp, &rstart, rd, s, err);
            return nr > 0;
        },
        launch);
}

void launch_samples(int prec, GridDev g, const void* ra, const void* rg, FieldConst fc, TreeParams tp, NodeLevel lv,
                    const int* list, const int* count, void* samples, Counters* ctr, hipStream_t s) {
    if (lv.n <= 0) return;
    const unsigned nb = (unsigned)((lv.n * 27 + 255) / 256);
    if (prec == 2) {
        Recs<double> r{(const double4*)ra, (const double4*)rg};
        k_samples<double, false><<<nb, 256, 0, s>>>(g, r, fc, tp, lv, list, count, (double*)samples, ctr);
    } else {
        Recs<float> r{(const float4*)ra, (const float4*)rg};
        if (prec == 1) k_samples<float, true><<<nb, 256, 0, s>>>(g, r, fc, tp, lv, list, count, (float*)samples, ctr);
        else k_samples<float, false><<<nb, 256, 0, s>>>(g, r, fc, tp, lv, list, count, (float*)samples, ctr);
    }
}

void launch_classify(int prec, GridDev g, const void* ra, const void* rg, FieldConst fc, TreeParams tp, NodeLevel lv,
                     const void* samples, int32_t* sc, Counters* ctr, hipStream_t s) {
    if (lv.n <= 0) {
        if (lv.n == 0) launch_fill_i32(lv.flag, 1, 0, s);  // the closing flag of an empty level
        return;
    }
    const unsigned nb = (unsigned)((lv.n + 1 + 255) / 256);
    // the QEF code is compiled in only when used: it takes ~200 registers, which would cut the
    // occupancy of the common case
    auto go = [&](auto qef) {
        constexpr bool Q = decltype(qef)::value;
        if (prec == 2) {
            Recs<double> r{(const double4*)ra, (const double4*)rg};
            k_classify<double, false, Q><<<nb, 256, 0, s>>>(g, r, fc, tp, lv, (const double*)samples, sc, ctr);
        } else {
            Recs<float> r{(const float4*)ra, (const float4*)rg};
            if (prec == 1)
                k_classify<float, true, Q><<<nb, 256, 0, s>>>(g, r, fc, tp, lv, (const float*)samples, sc, ctr);
            else
                k_classify<float, false, Q><<<nb, 256, 0, s>>>(g, r, fc, tp, lv, (const float*)samples, sc, ctr);
        }
    };
    if (tp.qef_mode != 0) {
        go(std::true_type());
        if (prec == 2) {
            Recs<double> r{(const double4*)ra, (const double4*)rg};
            k_qef_values<double, false><<<nb, 256, 0, s>>>(g, r, fc, tp, lv, ctr);
        } else {
            Recs<float> r{(const float4*)ra, (const float4*)rg};
            if (prec == 1) k_qef_values<float, true><<<nb, 256, 0, s>>>(g, r, fc, tp, lv, ctr);
            else k_qef_values<float, false><<<nb, 256, 0, s>>>(g, r, fc, tp, lv, ctr);
        }
    } else {
        go(std::false_type());
    }
}

void launch_children(NodeLevel parent, const int* scanned, NodeLevel next, hipStream_t s) {
    if (parent.n <= 0) return;
    k_children<<<(unsigned)((parent.n + 255) / 256), 256, 0, s>>>(parent, scanned, next, nullptr);
}

void launch_children_counted(NodeLevel parent, const int* scanned, NodeLevel next, int* next_count, hipStream_t s) {
    if (parent.n <= 0) return;
    k_children<<<(unsigned)((parent.n + 1 + 255) / 256), 256, 0, s>>>(parent, scanned, next, next_count);
}

void launch_sample_grads(int prec, TreeParams tp, NodeLevel lv, const void* samples, double* os, double* og,
                         hipStream_t s) {
    if (lv.n <= 0) return;
    const unsigned nb = (unsigned)((lv.n + 63) / 64);
    if (prec == 2) k_sample_grads<double><<<nb, 64, 0, s>>>(tp, lv, (const double*)samples, os, og);
    else k_sample_grads<float><<<nb, 64, 0, s>>>(tp, lv, (const float*)samples, os, og);
}

}  // namespace pisurf
