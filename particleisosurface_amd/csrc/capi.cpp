// pisurf C ABI: context, device buffers, the host level loop of the octree build, and the
// extraction driver. All field arithmetic runs in the HIP kernels (k_*.hip); the host only sets
// up grid geometry, drives launches level by level, and assembles the mesh topology.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pisurf.h"
#include "dmc_tables.h"
#include "pisurf_internal.h"
#include "run_tuner.h"

using namespace pisurf;

namespace {

struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t ensure(size_t bytes) {
        if (bytes == 0) bytes = 16;
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
    template <class T> T* as() const { return (T*)p; }
};

// page-locked host staging (hipHostMalloc; never registration of existing memory)
struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
};

// growable array in page-locked host memory: the mesh is copied off the device at full link speed
template <class T> struct PinnedVec {
    T* p = nullptr;
    size_t n = 0, cap = 0;
    PinnedVec() = default;
    PinnedVec(const PinnedVec&) = delete;
    PinnedVec& operator=(const PinnedVec&) = delete;
    ~PinnedVec() {
        if (p) (void)hipHostFree(p);
    }
    void reserve(size_t m) {
        if (m <= cap) return;
        size_t c2 = std::max(m, cap + cap / 2);
        void* q = nullptr;
        if (hipHostMalloc(&q, c2 * sizeof(T), hipHostMallocDefault) != hipSuccess) throw std::bad_alloc();
        if (n) std::memcpy(q, p, n * sizeof(T));
        if (p) (void)hipHostFree(p);
        p = (T*)q;
        cap = c2;
    }
    void resize(size_t m) {  // new elements are left uninitialised
        reserve(m);
        n = m;
    }
    void assign(size_t m, T v) {
        resize(m);
        for (size_t i = 0; i < m; i++) p[i] = v;
    }
    void push_back(T v) {
        if (n == cap) reserve(n + 1);
        p[n++] = v;
    }
    void clear() { n = 0; }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
    T* data() { return p; }
    const T* data() const { return p; }
};

struct LevelBuf {
    int64_t n = 0;
    DBuf c, depth, id, type, empty, curv, node, flag, child, scanned, ip;
    hipError_t ensure(int64_t m) {
        hipError_t e;
        if ((e = c.ensure(m * sizeof(float4))) != hipSuccess) return e;
        if ((e = ip.ensure(m * sizeof(int4))) != hipSuccess) return e;
        if ((e = depth.ensure(m)) != hipSuccess) return e;
        if ((e = id.ensure(m * 8)) != hipSuccess) return e;
        if ((e = type.ensure(m)) != hipSuccess) return e;
        if ((e = empty.ensure(m)) != hipSuccess) return e;
        if ((e = curv.ensure(m * 4)) != hipSuccess) return e;
        if ((e = node.ensure(m * sizeof(double4))) != hipSuccess) return e;
        if ((e = flag.ensure((m + 1) * 4)) != hipSuccess) return e;
        if ((e = child.ensure(m * 4)) != hipSuccess) return e;
        if ((e = scanned.ensure((m + 1) * 4)) != hipSuccess) return e;
        return hipSuccess;
    }
    NodeLevel view() const {
        NodeLevel v;
        v.c = c.as<float4>();
        v.depth = depth.as<int8_t>();
        v.id = id.as<uint64_t>();
        v.type = type.as<uint8_t>();
        v.empty = empty.as<uint8_t>();
        v.curv = curv.as<float>();
        v.node = node.as<double4>();
        v.flag = flag.as<int>();
        v.child = child.as<int>();
        v.ip = ip.as<int4>();
        v.n = n;
        return v;
    }
};

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// open-addressing map from a 128-bit key to an int (mesh vertex / triangle dedupe)
struct KeyMap {
    std::vector<uint64_t> ka, kb;
    std::vector<int32_t> val;
    size_t mask = 0, count = 0;
    void reset(size_t expect) {
        size_t cap = 1024;
        while (cap < expect * 2) cap <<= 1;
        ka.assign(cap, 0);
        kb.assign(cap, 0);
        val.assign(cap, -1);
        mask = cap - 1;
        count = 0;
    }
    static uint64_t mix(uint64_t x) {
        x ^= x >> 33;
        x *= 0xff51afd7ed558ccdULL;
        x ^= x >> 33;
        x *= 0xc4ceb9fe1a85ec53ULL;
        x ^= x >> 33;
        return x;
    }
    void grow() {
        std::vector<uint64_t> oa, ob;
        std::vector<int32_t> ov;
        oa.swap(ka);
        ob.swap(kb);
        ov.swap(val);
        const size_t cap = oa.size() * 2;
        ka.assign(cap, 0);
        kb.assign(cap, 0);
        val.assign(cap, -1);
        mask = cap - 1;
        for (size_t i = 0; i < oa.size(); i++)
            if (ov[i] >= 0) {
                size_t h = mix(oa[i] * 0x9E3779B97F4A7C15ULL ^ ob[i]) & mask;
                while (val[h] >= 0) h = (h + 1) & mask;
                ka[h] = oa[i];
                kb[h] = ob[i];
                val[h] = ov[i];
            }
    }
    // returns (value, inserted)
    std::pair<int32_t, bool> try_emplace(uint64_t a, uint64_t b, int32_t v) {
        if ((count + 1) * 2 > ka.size()) grow();
        size_t h = mix(a * 0x9E3779B97F4A7C15ULL ^ b) & mask;
        while (val[h] >= 0) {
            if (ka[h] == a && kb[h] == b) return {val[h], false};
            h = (h + 1) & mask;
        }
        ka[h] = a;
        kb[h] = b;
        val[h] = v;
        count++;
        return {v, true};
    }
};

}  // namespace

// ---- working memory of pisurf_neighbors_noalloc -------------------------------------------
// A fixed device array (each GPU the library runs on holds its own copy until the process ends):
// per-query counts and 32-bit offsets for at most PISURF_NOALLOC_MAX_QUERIES queries, the scan's
// workspace and a 64-bit total. Host code never uses the array's name as an address: its device
// address is looked up once per device (hipGetSymbolAddress, in pisurf_create) and kept here.
namespace {
constexpr int64_t kNoAllocQ = PISURF_NOALLOC_MAX_QUERIES;
constexpr int64_t kNoAllocScanInts = 4096;  // scan workspace for kNoAllocQ + 1 counts: 1032 ints
constexpr int64_t kNoAllocInts = 2 * (kNoAllocQ + 1) + kNoAllocScanInts + 4;
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_m;
void* g_scratch[kNoAllocMaxDevices] = {};
}  // namespace
static __device__ __align__(256) int scratch[kNoAllocInts];

struct pisurf_ctx {
    int device = 0;
    bool own_stream = true;       // false for the frame contexts of a batch (the batch's stream)
    void* scratch_ptr = nullptr;  // this device's copy of the no-allocation working array
    std::vector<pisurf_ctx*> sub;  // pisurf_run_batch: one context per frame, kept for reuse
    std::vector<pisurf_ctx*> lanes;  // pisurf_run_frames: contexts of lanes 1 .. K - 1, kept for reuse
    // pisurf_run_batch work space: frame descriptors of the run queue, item offsets, items, counter
    DBuf bq_frames, bq_off, bq_items, bq_head;
    PinnedBuf bq_host;
    size_t rb_used = 0;  // bytes of rb claimed by the pending batched read-backs
    hipStream_t stream = nullptr;
    pisurf_params prm;
    std::string err;
    // particles
    int64_t n = 0;
    DBuf xyz;
    float bbox[6] = {0, 0, 0, 0, 0, 0};
    bool single_layer = false;
    bool bbox_done = false;
    // grid
    bool grid_done = false;
    pisurf_grid_info ginfo;
    GridDev g;
    DBuf keys, vals, keys2, vals2, off, tmp, spos, part, small;
    // anisotropy (records are for prec_rec)
    bool aniso_done = false;
    int prec_rec = 0;
    DBuf xm, rec_a, rec_g, rec_lo;
    // constants
    bool consts_done = false;
    double influnce2 = 0, sigma = 0, iso = 0, max_scalar = 0;
    // normals
    bool normals_done = false;
    DBuf nrm;
    // tree
    bool tree_done = false;
    int dmax = 0, dmin = 0;
    float root_c[3] = {0, 0, 0}, root_half = 0;
    std::vector<LevelBuf*> levels;
    DBuf samples, cl_flag, cl_pos, cl_list;
    DBuf lvcnt;  // node counts of the levels built from device-resident counts
    hipGraphExec_t top_exec = nullptr;  // those levels' launches as an executable graph
    // a frame issues those levels on the mesh-download stream (dstream, idle until the bisection;
    // a stream of their own would take a hardware queue from the lanes of a multi-frame call),
    // with their own scan workspace, right after the grid, so they run beside the anisotropy
    hipEvent_t tev_grid = nullptr, tev_top = nullptr;
    DBuf tmp_top;
    bool top_pending = false;  // issued on dstream, counts not read yet
    bool lane = false;         // meshing as a lane of pisurf_run_frames (whole frames overlap instead)
    int top_t = 0;             // levels 0 .. top_t - 1 built there
    std::vector<int64_t> top_ub;
    std::vector<int> top_n;    // their counts once read (empty: not read yet)
    // sample deduplication: keys / entries (ping-pong), heads, scanned heads, representatives,
    // head list, per-point values
    DBuf sd_key, sd_val, sd_key2, sd_val2, sd_head, sd_hpos, sd_rep, sd_hlist, sd_vals, sd_cwork;
    DBuf sd_entrec, sd_rent, sd_rechead, sd_nrec;  // per-tile sample records
    // scratch for API calls
    DBuf q, out1, out2, out3;
    // multi-frame path: next frame's particles, their staging and copy stream
    DBuf xyz_next;
    PinnedBuf pin;
    hipStream_t cstream = nullptr;
    hipEvent_t cev = nullptr;  // end of a staged copy on cstream (waited for by the staging thread)
    // mesh download stream and the events around the bisection
    hipStream_t dstream = nullptr;
    hipEvent_t eb0 = nullptr, eb1 = nullptr;
    std::vector<hipEvent_t> ev_chunk;  // ends of the bisection chunks
    // device extraction workspace
    DBuf eb[38];
    // mesh
    PinnedVec<double> mesh_v;
    PinnedVec<int32_t> mesh_t;
    // stats
    DBuf counters;
    pisurf_stats st;
    unsigned long long n_fp64_eig = 0;
    unsigned long long host_sample_entries = 0;  // deduplicated sampling: 27 per sampled node
    std::vector<hipEvent_t> ev;  // start/stop pairs around node-sampling launches
    std::vector<int> ev_kind;
    size_t ev_used = 0;

    // read-backs: a polled event and a page-locked staging buffer owned by the context, so calls
    // made from different threads (one at a time) create no per-thread HIP objects
    hipEvent_t wev = nullptr;
    PinnedBuf rb;
    // counts and waits on the context's stream go through its read-back slot (host-mapped memory
    // written by a small kernel, polled by the host)
    HostRead* hr = nullptr;
    hipError_t wait() { return hr ? stream_fence(stream, hr) : stream_wait_on(stream, wev); }
    HostRead* hread() { return hr; }
    hipError_t read(void* host, const void* dptr, size_t bytes) { return read_device(host, dptr, bytes, stream, hr); }
    // a scan total published by the scan itself (scan_read_begin / scan_read_end)
    ScanPub scan_pub(int off = 0) { return scan_read_begin(hr, off); }
    hipError_t read_scan(int* host, const int* dlast, ScanPub p) { return scan_read_end(host, dlast, p, stream, hr); }

    ~pisurf_ctx() {
        for (auto* b : sub) delete b;
        for (auto* l : lanes) {
            (void)hipSetDevice(l->device);
            (void)l->wait();
            delete l;
        }
        host_read_release(hr);
        if (top_exec) (void)hipGraphExecDestroy(top_exec);
        if (tev_grid) (void)hipEventDestroy(tev_grid);
        if (tev_top) (void)hipEventDestroy(tev_top);
        if (wev) (void)hipEventDestroy(wev);
        for (auto* l : levels) delete l;
        for (auto e : ev) (void)hipEventDestroy(e);
        if (cstream) (void)hipStreamDestroy(cstream);
        if (cev) (void)hipEventDestroy(cev);
        if (dstream) (void)hipStreamDestroy(dstream);
        if (eb0) (void)hipEventDestroy(eb0);
        if (eb1) (void)hipEventDestroy(eb1);
        for (auto e : ev_chunk)
            if (e) (void)hipEventDestroy(e);
        if (stream && own_stream) (void)hipStreamDestroy(stream);
    }
    // start/stop events around one evaluation launch; kind 0: per-point sampling kernel, 1: run kernel
    hipError_t event_pair(hipEvent_t& a, hipEvent_t& b, size_t* slot = nullptr) {
        while (ev.size() < ev_used + 2) {
            hipEvent_t e;
            hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            ev.push_back(e);
        }
        a = ev[ev_used];
        b = ev[ev_used + 1];
        if (ev_kind.size() < ev_used / 2 + 1) ev_kind.resize(ev_used / 2 + 1);
        ev_kind[ev_used / 2] = 0;
        if (slot) *slot = ev_used / 2;
        ev_used += 2;
        return hipSuccess;
    }
    // sum the recorded intervals per kind (stream already synchronized)
    void harvest_events() {
        for (size_t i = 0; i + 1 < ev_used; i += 2) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess) ms = 0.f;
            if (ev_kind[i / 2]) {
                st.ms_run_kernel += ms;
                st.run_launches++;
            } else {
                st.ms_sample_kernel += ms;
                st.sample_launches++;
            }
        }
        ev_used = 0;
    }
    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
    int hipfail(hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return PISURF_E_HIP;
    }
    Counters* ctr() { return counters.as<Counters>(); }
    // counters of the timed trial launches of run-shape tuning (never read)
    DBuf trial_counters;
    Counters* trial_ctr() {
        if (!trial_counters.p && trial_counters.ensure(kCtrStripes * sizeof(Counters)) != hipSuccess) return nullptr;
        return trial_counters.as<Counters>();
    }
};

#define HIPCHK(ctx, call)                                   \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return (ctx)->hipfail(e_, #call); \
    } while (0)

#define API_BEGIN try {
#define API_END(ctx)                                                          \
    }                                                                         \
    catch (const std::bad_alloc&) { return (ctx)->fail(PISURF_E_OOM, "host allocation failed"); } \
    catch (...) { return (ctx)->fail(PISURF_E_INVALID, "unexpected host exception"); }

namespace {

int prec_of(const pisurf_ctx* c) { return c->prm.precision; }

int64_t pow2_at_least(int64_t n) {
    int64_t p = 1024;
    while (p < n) p <<= 1;
    return p;
}

int bits_for(int64_t v) {
    int b = 1;
    while (b < 63 && ((int64_t)1 << b) <= v) b++;
    return b;
}

FieldConst field_const(const pisurf_ctx* c) {
    FieldConst fc;
    fc.h2 = c->influnce2;
    fc.inv_h2 = 1.0 / c->influnce2;
    fc.sigma = c->sigma;
    fc.iso = c->iso;
    fc.use_ani = c->prm.use_ani;
    fc.use_poly6 = c->prm.use_poly6;
    fc.lo = c->rec_lo.as<float4>();
    if (prec_of(c) != PISURF_FP64) fc.inv_h2 = double(1.0f / float(c->influnce2));
    return fc;
}

// Cubic B-spline and poly6 of the reference's constants (src/evaluator.cpp:377-396), evaluated
// in double; FP32 modes narrow the result like the reference does.
double bspline_d(double q, double sigma) {
    if (q < 2) {
        if (q <= 1) return sigma * (1 - 1.5 * std::pow(q, 2) + 0.75 * std::pow(q, 3));
        return sigma * (0.25 * std::pow(2 - q, 3));
    }
    return 0.0;
}

// Evaluator ctor constants (src/evaluator.cpp:31-37) and CalculateMaxScalarConstR /
// RecommendIsoValueConstR (src/evaluator.cpp:272-291,323-335).
void compute_constants(pisurf_ctx* c) {
    const float r = c->prm.radius;
    const double h = double(r * 2.0f);
    const bool f32 = prec_of(c) != PISURF_FP64;
    auto nar = [&](double v) { return f32 ? double(float(v)) : v; };
    c->influnce2 = nar(std::pow(h, 2));
    c->sigma = nar(1 / std::pow(h, 6));
    const float r2 = r * r;
    if (c->prm.use_poly6) {
        auto p6 = [&](double d2) {
            double pd = d2 > c->influnce2 ? 0.0 : nar(std::pow(c->influnce2 - d2, 3));
            return nar(pd * c->sigma);
        };
        double k = 0;
        k = nar(k + p6(0.0));
        k = nar(k + p6(nar(4 * r2)));
        k = nar(k + p6(nar(8 * r2)));
        k = nar(k + p6(nar(12 * r2)));
        c->max_scalar = k;
        c->iso = p6(nar(double(c->prm.iso_factor * r2)));
    } else {
        const float sqrt2 = 1.4142135623730950488016887242097f, sqrt3 = 1.7320508075688772935274463415059f;
        double k = 0;
        k = nar(k + nar(bspline_d(0.0, c->sigma)));
        k = nar(k + nar(nar(bspline_d(1.0, c->sigma)) * 6));
        k = nar(k + nar(nar(bspline_d(double(sqrt2), c->sigma)) * 12));
        k = nar(k + nar(nar(bspline_d(double(sqrt3), c->sigma)) * 8));
        c->max_scalar = k;
        c->iso = nar(bspline_d(double(0.01f), c->sigma));
    }
    c->consts_done = true;
}

// HashGrid geometry (src/hash_grid.cpp:4-34), float arithmetic as in the reference
void grid_geometry(const float* bounding, float radius, float inf, float out_b[6], uint32_t n3[3], int64_t& ncell,
                   float& cell) {
    cell = radius * inf;
    for (int i = 0; i < 3; i++) {
        const float length = std::ceil((bounding[2 * i + 1] - bounding[2 * i]) / cell) * cell;
        const float center = (bounding[2 * i] + bounding[2 * i + 1]) / 2;
        out_b[2 * i] = center - length / 2;
        out_b[2 * i + 1] = center + length / 2;
    }
    for (int i = 0; i < 3; i++)
        n3[i] = (uint32_t)std::max(int(std::ceil((out_b[2 * i + 1] - out_b[2 * i]) / cell)), 1);
    ncell = (int64_t)n3[0] * n3[1] * n3[2];
}

int host_cell_coord(float p, float b0, float cell, int n) {
    const float v = (p - b0) / cell;
    if (!(v > -8.0f && v < float(n) + 8.0f)) return (v > 0.0f) ? n + 8 : -8;
    return (int)v;
}

// the tight bounding box on the device: min / max (6 floats) at small + 64, NaN count at small
int bbox_launch(pisurf_ctx* c) {
    if (c->n <= 0) return c->fail(PISURF_E_STATE, "no particles uploaded");
    HIPCHK(c, c->part.ensure(6 * 1024 * sizeof(float) + 64));
    HIPCHK(c, c->small.ensure(256));
    HIPCHK(c, hipMemsetAsync(c->small.as<char>(), 0, 256, c->stream));
    launch_bbox(c->xyz.as<float>(), c->n, c->part.as<float>(), (float*)(c->small.as<char>() + 64), c->small.as<int>(),
                c->stream);
    HIPCHK(c, hipGetLastError());
    return PISURF_OK;
}
const void* bbox_dev(pisurf_ctx* c) { return c->small.as<char>() + 64; }
const int* bbox_nan_dev(pisurf_ctx* c) { return c->small.as<int>(); }
// with c->bbox read back
int bbox_finish(pisurf_ctx* c, int nans) {
    if (nans) return c->fail(PISURF_E_INVALID, "particle positions contain NaN");
    for (int a = 0; a < 6; a++)
        if (!std::isfinite(c->bbox[a])) return c->fail(PISURF_E_INVALID, "particle positions are not finite");
    c->single_layer = (c->bbox[0] == c->bbox[1] || c->bbox[2] == c->bbox[3] || c->bbox[4] == c->bbox[5]);
    c->bbox_done = true;
    return PISURF_OK;
}

int ensure_bbox(pisurf_ctx* c) {
    if (c->bbox_done) return PISURF_OK;
    int rc = bbox_launch(c);
    if (rc) return rc;
    // the NaN count and the box in one read: the box sits 64 bytes after the count
    char buf[64 + 6 * sizeof(float)];
    HIPCHK(c, c->read(buf, bbox_nan_dev(c), sizeof(buf)));
    int nans = 0;
    std::memcpy(&nans, buf, sizeof(int));
    std::memcpy(c->bbox, buf + 64, 6 * sizeof(float));
    return bbox_finish(c, nans);
}

// HashGrid build launches on the bbox already known on the host; the count of particles outside
// the grid is left at grid_ninv_dev for grid_finish
int grid_launch(pisurf_ctx* c, const float* bbox, float inf);
const int* grid_ninv_dev(pisurf_ctx* c) { return c->small.as<int>(); }
int grid_finish(pisurf_ctx* c, int ninv) {
    c->ginfo.n_invalid = ninv;
    c->grid_done = true;
    c->aniso_done = c->normals_done = c->tree_done = false;
    return PISURF_OK;
}

int do_grid(pisurf_ctx* c, const float* bbox, float inf) {
    int rc = ensure_bbox(c);
    if (rc) return rc;
    const double t0 = now_ms();
    rc = grid_launch(c, bbox, inf);
    if (rc) return rc;
    int ninv = 0;
    HIPCHK(c, c->read(&ninv, grid_ninv_dev(c), sizeof(int)));
    grid_finish(c, ninv);
    c->st.ms_grid += now_ms() - t0;
    return PISURF_OK;
}

int grid_launch(pisurf_ctx* c, const float* bbox, float inf) {
    if (!(c->prm.radius > 0.0f) || !std::isfinite(c->prm.radius)) return c->fail(PISURF_E_INVALID, "radius must be > 0");
    if (!(inf > 0.0f)) return c->fail(PISURF_E_INVALID, "inf_factor must be > 0");
    float b[6];
    uint32_t n3[3];
    int64_t ncell;
    float cell;
    grid_geometry(bbox ? bbox : c->bbox, c->prm.radius, inf, b, n3, ncell, cell);
    for (int a = 0; a < 6; a++)
        if (!std::isfinite(b[a])) return c->fail(PISURF_E_INVALID, "grid bounding box is not finite");
    if (ncell >= (int64_t)INT32_MAX - 1) return c->fail(PISURF_E_UNSUPPORTED, "cell count exceeds 2^31 - 2");
    const int64_t n = c->n;
    HIPCHK(c, c->keys.ensure(n * 4));
    HIPCHK(c, c->vals.ensure(n * 4));
    HIPCHK(c, c->keys2.ensure(n * 4));
    HIPCHK(c, c->vals2.ensure(n * 4));
    HIPCHK(c, c->off.ensure((ncell + 1) * 4));
    HIPCHK(c, c->spos.ensure(n * sizeof(float4)));
    const int64_t tmp_ints = std::max(radix_tmp_ints(n), scan_tmp_ints(ncell + 1)) + 64;
    HIPCHK(c, c->tmp.ensure(tmp_ints * 4));
    HIPCHK(c, c->small.ensure(256));
    HIPCHK(c, hipMemsetAsync(c->small.p, 0, 64, c->stream));
    GridDev g;
    for (int a = 0; a < 3; a++) {
        g.b0[a] = b[2 * a];
        g.n[a] = (int)n3[a];
    }
    g.cell = cell;
    g.ncell = (int)ncell;
    g.off = c->off.as<int>();
    int* ninv_d = c->small.as<int>();
    launch_keys(c->xyz.as<float>(), n, g, c->keys.as<uint32_t>(), c->vals.as<int>(), ninv_d, c->stream);
    int bits = 1;
    while (bits < 32 && ((uint64_t)1 << bits) <= (uint64_t)ncell) bits++;
    radix_sort_pairs(c->keys.as<uint32_t>(), c->vals.as<int>(), c->keys2.as<uint32_t>(), c->vals2.as<int>(), n, bits,
                     c->tmp.as<int>(), c->stream);
    launch_cell_offsets(c->keys.as<uint32_t>(), n, ninv_d, ncell, c->off.as<int>(), c->stream);
    launch_gather_pos(c->xyz.as<float>(), c->vals.as<int>(), n, c->spos.as<float4>(), c->stream);
    HIPCHK(c, hipGetLastError());
    c->g = g;
    for (int a = 0; a < 6; a++) c->ginfo.bounding[a] = b[a];
    for (int a = 0; a < 3; a++) c->ginfo.xyz_cell_num[a] = n3[a];
    c->ginfo.cell_num = ncell;
    c->ginfo.cell_size = cell;
    return PISURF_OK;
}

// compute_Gs_xMeans launches; the count of FP64 eigensolves is left at aniso_nfp_dev
int aniso_launch(pisurf_ctx* c);
const void* aniso_nfp_dev(pisurf_ctx* c) { return c->small.as<char>() + 128; }
void aniso_finish(pisurf_ctx* c, unsigned long long nfp) {
    c->n_fp64_eig = nfp;
    c->prec_rec = prec_of(c);
    c->aniso_done = true;
    c->normals_done = c->tree_done = false;
}

int do_aniso(pisurf_ctx* c) {
    if (!c->grid_done) {
        int rc = do_grid(c, nullptr, 4.0f);
        if (rc) return rc;
    }
    const double t0 = now_ms();
    int rc = aniso_launch(c);
    if (rc) return rc;
    unsigned long long nfp = 0;
    HIPCHK(c, c->read(&nfp, aniso_nfp_dev(c), 8));
    aniso_finish(c, nfp);
    c->st.ms_aniso += now_ms() - t0;
    return PISURF_OK;
}

int aniso_launch(pisurf_ctx* c) {
    const int prec = prec_of(c);
    const size_t rb = (prec == PISURF_FP64) ? sizeof(double4) : sizeof(float4);
    // compensated FP32 keeps xMean as hi + lo (two float4 per particle) and a residual record
    HIPCHK(c, c->xm.ensure(c->n * rb * (prec == PISURF_FP32_KAHAN ? 2 : 1)));
    HIPCHK(c, c->rec_a.ensure(c->n * rb));
    HIPCHK(c, c->rec_g.ensure(2 * c->n * rb));
    HIPCHK(c, c->rec_lo.ensure(prec == PISURF_FP32_KAHAN ? c->n * sizeof(float4) : 16));
    HIPCHK(c, c->small.ensure(256));
    HIPCHK(c, hipMemsetAsync(c->small.as<char>() + 128, 0, 8, c->stream));
    AnisoParams ap;
    ap.radius = c->prm.radius;
    ap.nI = c->prm.radius * 4.0f;
    ap.nI2 = ap.nI * ap.nI;
    const double h = double(c->prm.radius * 2.0f);
    ap.h2 = (prec == PISURF_FP64) ? std::pow(h, 2) : double(float(std::pow(h, 2)));
    ap.xmean_delta = 0.9f;
    ap.use_poly6 = c->prm.use_poly6;
    ap.use_ani = c->prm.use_ani;
    ap.nI3 = std::pow(double(ap.nI), 3);
    launch_xmean(prec, c->g, c->spos.as<float4>(), c->n, c->ginfo.n_invalid, ap, c->xm.p, c->stream);
    unsigned long long* nfp = (unsigned long long*)(c->small.as<char>() + 128);
    launch_gmat(prec, c->g, c->spos.as<float4>(), c->n, c->ginfo.n_invalid, ap, c->xm.p, c->rec_a.p, c->rec_g.p,
                c->rec_lo.p, nfp, c->stream);
    HIPCHK(c, hipGetLastError());
    return PISURF_OK;
}

int need_aniso(pisurf_ctx* c) {
    if (!c->aniso_done || c->prec_rec != prec_of(c)) return do_aniso(c);
    return PISURF_OK;
}

// CalcParticlesNormal launches (records and constants ready)
int normals_launch(pisurf_ctx* c) {
    HIPCHK(c, c->nrm.ensure(c->n * sizeof(float4)));
    launch_normals(prec_of(c), c->g, c->rec_a.p, c->rec_g.p, field_const(c), c->spos.as<float4>(), c->n,
                   c->ginfo.n_invalid, c->nrm.as<float4>(), c->ctr(), c->trial_ctr(), c->stream);
    HIPCHK(c, hipGetLastError());
    c->normals_done = true;
    return PISURF_OK;
}

int top_levels_read(pisurf_ctx* c);

int do_normals(pisurf_ctx* c) {
    int rc = need_aniso(c);
    if (rc) return rc;
    if (!c->consts_done) compute_constants(c);
    const double t0 = now_ms();
    rc = normals_launch(c);
    if (rc) return rc;
    if (c->top_pending) {
        // the counts of the levels above D_min (issued on the second stream, long done by now)
        // are read as the fence that ends the normals: one round trip instead of two
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->tev_top, 0));
        if ((rc = top_levels_read(c))) return rc;
    } else {
        HIPCHK(c, c->wait());
    }
    c->st.ms_normals += now_ms() - t0;
    return PISURF_OK;
}

// resizeRootBoxConstR (src/surface_reconstructor.cpp:97-122)
void resize_root(pisurf_ctx* c) {
    const float r = c->prm.radius;
    const float* B = c->bbox;
    const float maxLen = std::max({B[1] - B[0], B[3] - B[2], B[5] - B[4]});
    int D = int(std::ceil(std::log2(std::ceil(double(maxLen / r)))));
    float resizeLen = float(std::pow(2.0, D) * double(r));
    while (resizeLen - maxLen < (4.0f * r * 2)) {
        D++;
        resizeLen = float(std::pow(2.0, D) * double(r));
    }
    if (c->prm.depth_max_override > 0) {
        D = c->prm.depth_max_override;
        resizeLen = float(std::pow(2.0, D) * double(r));
    }
    c->dmax = D;
    c->root_half = resizeLen / 2;
    for (int i = 0; i < 3; i++) c->root_c[i] = (B[2 * i] + B[2 * i + 1]) / 2;
    c->dmin = c->dmax - (c->single_layer ? 1 : 3);
}

TreeParams tree_params(const pisurf_ctx* c) {
    TreeParams tp;
    tp.radius = c->prm.radius;
    tp.margin = c->prm.radius * c->prm.smooth_factor;
    for (int a = 0; a < 3; a++) {
        tp.b_lo[a] = c->ginfo.bounding[2 * a];
        tp.b_hi[a] = c->ginfo.bounding[2 * a + 1];
    }
    tp.depth_min = c->dmin;
    tp.depth_max = c->dmax;
    tp.calc_normals = c->prm.calc_normals;
    tp.curv = 1;
    tp.qef_mode = c->prm.qef_mode;
    tp.qef_svd_tol = c->prm.qef_svd_tol;
    return tp;
}

int frame_ready(pisurf_ctx* c) {
    int rc = need_aniso(c);
    if (rc) return rc;
    if (!c->consts_done) compute_constants(c);
    if (c->prm.calc_normals && !c->normals_done) {
        rc = do_normals(c);
        if (rc) return rc;
    }
    if (!c->normals_done) {  // curvature reads normals; without CALC_P_NORMAL they stay zero
        HIPCHK(c, c->nrm.ensure(c->n * sizeof(float4)));
        HIPCHK(c, hipMemsetAsync(c->nrm.p, 0, c->n * sizeof(float4), c->stream));
    }
    return PISURF_OK;
}

// run the three level kernels on `lv` (nodes already in lv.c / depth / id)
int sweep(pisurf_ctx* c, LevelBuf* L, int32_t* sc_dev, int level_depth) {
    const int prec = prec_of(c);
    const size_t rsz = (prec == PISURF_FP64) ? 8 : 4;
    HIPCHK(c, c->samples.ensure(L->n * 27 * rsz));
    // leaf-depth level (cellsize - r < 1e-8, afterSampleEval): its nodes are leaves whatever their
    // curvature, which the reference computes there and discards
    bool leaf_level = false;
    if (level_depth >= 0) {
        const float half = std::ldexp(c->root_half, -level_depth);
        leaf_level = double((2.0f * half) - c->prm.radius) < 1e-8;
    }
    TreeParams tp = tree_params(c);
    if (leaf_level) tp.curv = 0;
    const FieldConst fc = field_const(c);
    const NodeLevel v = L->view();
    const int64_t need_tmp = scan_tmp_ints(L->n + 1) + 64;
    if ((size_t)need_tmp * 4 > c->tmp.cap) {
        HIPCHK(c, c->wait());
        HIPCHK(c, c->tmp.ensure(need_tmp * 4));
    }
    // lanes per node ~ (y, z) cell rows of the grown node box: (width / cell + 1)^2 on average
    int group = 64;
    if (level_depth >= 0) {
        const double half = std::ldexp((double)c->root_half, -level_depth);
        const double w = (2 * half + 2 * (double)tp.margin) / (double)c->g.cell + 1;
        group = (w * w <= 6) ? 4 : (w * w <= 8) ? 8 : (w * w <= 16) ? 16 : 64;
    }
    // At D_min, where most nodes lie inside the fluid and their samples do not change sign, the
    // frame tests emptiness first and takes the curvature after sampling, for the nodes whose
    // samples change sign; deeper levels hold the children of such nodes, and mixed API sweeps,
    // do both at once. (Above D_min classification reads no curvature.)
    const bool split_curv = level_depth >= 0 && level_depth == c->dmin;
    launch_empty_curv(c->g, c->spos.as<float4>(), c->nrm.as<float4>(), tp, v, group,
                      split_curv ? kEmptyOnly : kEmptyAndCurv, nullptr, 0, c->stream);
    // the level's flag / scanned / child arrays serve as the sample list until classification
    HIPCHK(c, c->cl_flag.ensure((L->n + 1) * 4));
    HIPCHK(c, c->cl_pos.ensure((L->n + 1) * 4));
    HIPCHK(c, c->cl_list.ensure((L->n + 1) * 4));
    // Deduplicated sampling needs one node size per level (lattice keys) and 21-bit lattice
    // coordinates: 2^(depth + 1) half-widths span the root box.
    // levels above D_min sample nothing (wants_samples), so they skip the count read-back
    const bool dedupe = c->prm.sample_dedupe && level_depth >= 0 && level_depth <= 19 && level_depth >= c->dmin;
    // centre values of the leaf-depth level by runs (QEF off); that count is read after the
    // sampling's reads, so it is published far from the bytes they copy
    const bool centre_runs = level_depth >= 0 && leaf_level && c->prm.qef_mode == 0 && level_depth >= c->dmin;
    const ScanPub kpub = dedupe ? c->scan_pub() : ScanPub();
    const ScanPub cpub = centre_runs ? c->scan_pub(kScanPubFar) : ScanPub();
    launch_sample_list(tp, v, L->flag.as<int>(), L->scanned.as<int>(), L->child.as<int>(), c->cl_flag.as<int>(),
                       c->cl_pos.as<int>(), c->cl_list.as<int>(), c->tmp.as<int>(), c->stream, kpub, cpub);
    if (dedupe) {
        int k = 0;
        HIPCHK(c, c->read_scan(&k, L->scanned.as<int>() + L->n, kpub));
        const int64_t nent = (int64_t)k * 27;
        if (nent >= INT32_MAX) return c->fail(PISURF_E_UNSUPPORTED, "octree level has too many sample points");
        if (nent > 0) {
            // entries merged per tile of nodes first (fewer keys to sort)
            const bool tiles = L->n * 27 < INT32_MAX;
            HIPCHK(c, c->sd_key.ensure(nent * 8));
            HIPCHK(c, c->sd_key2.ensure(nent * 8));
            HIPCHK(c, c->sd_val.ensure(nent * 4));
            HIPCHK(c, c->sd_val2.ensure(nent * 4));
            if (tiles) {
                HIPCHK(c, c->sd_entrec.ensure(nent * 4));
                HIPCHK(c, c->sd_rent.ensure(nent * 4));
                HIPCHK(c, c->sd_rechead.ensure(nent * 4));
                HIPCHK(c, c->sd_nrec.ensure(4));
            }
            HIPCHK(c, c->sd_head.ensure((2 * ((nent + 255) / 256) + 2) * 4));
            HIPCHK(c, c->sd_hpos.ensure((nent + 1) * 4));
            HIPCHK(c, c->sd_rep.ensure(nent * 4));
            HIPCHK(c, c->sd_hlist.ensure(nent * 4));
            HIPCHK(c, c->sd_vals.ensure(nent * rsz));
            HIPCHK(c, c->sd_cwork.ensure(4 * (nent + 2) * 4));
            const int64_t need = std::max(radix_tmp_ints(nent), scan_tmp_ints(nent + 1)) + 64;
            if ((size_t)need * 4 > c->tmp.cap) {
                // kernels queued before the count read may still use the workspace
                HIPCHK(c, c->wait());
                HIPCHK(c, c->tmp.ensure(need * 4));
            }
            const float origin[3] = {c->root_c[0] - c->root_half, c->root_c[1] - c->root_half,
                                     c->root_c[2] - c->root_half};
            // distinct points grouped by hash cell (cell-major keys), evaluated in runs of cells
            const uint64_t groups = (uint64_t)(c->g.n[0] + 2) * (uint64_t)(c->g.n[1] + 2) * (uint64_t)(c->g.n[2] + 2);
            // three lattice bits per axis: fewer would save a sort pass at the coarser levels but
            // lengthen the runs of equal keys (shared cells, and the points outside the grid, which
            // form one group), which costs the heads pass more than the sort gains
            const int lb = 3;
            const int kbits = std::min(64, bits_for((int64_t)groups + 1) + 3 * lb);
            int64_t nrec = nent;
            const int vbits = pack_vbits(kbits, nent);
            if (tiles) {
                launch_sample_tiles(c->g, v, L->child.as<int>(), k, origin, lb, vbits, c->sd_nrec.as<int>(),
                                    c->sd_entrec.as<int>(), c->sd_key.as<uint64_t>(), c->sd_val.as<int>(),
                                    c->sd_rent.as<int>(), c->stream);
                int nr = 0;
                HIPCHK(c, c->read(&nr, c->sd_nrec.as<int>(), 4));
                nrec = nr;
            } else {
                launch_sample_dedupe_keys(c->g, v, L->child.as<int>(), nent, origin, lb, vbits,
                                          c->sd_key.as<uint64_t>(), c->sd_val.as<int>(), c->stream);
            }
            const SortedKV skv = radix_sort_entries(c->sd_key.as<uint64_t>(), c->sd_val.as<int>(),
                                                    c->sd_key2.as<uint64_t>(), c->sd_val2.as<int>(), nrec, kbits,
                                                    vbits, c->tmp.as<int>(), c->stream);
            hipEvent_t e0, e1;
            size_t eslot = 0;
            HIPCHK(c, c->event_pair(e0, e1, &eslot));
            int64_t nheads = 0, nruns = 0;
            int kx = 0;
            hipError_t herr = hipSuccess;
            launch_sample_dedupe_eval(prec, c->g, c->rec_a.p, c->rec_g.p, fc, v, L->child.as<int>(), skv,
                                      tiles ? c->sd_rent.as<int>() : nullptr, 3 * lb, nrec,
                                      tiles ? c->sd_entrec.as<int>() : nullptr, nent,
                                      tiles ? c->sd_rechead.as<int>() : nullptr, c->sd_head.as<int>(),
                                      c->sd_hpos.as<int>(), c->sd_rep.as<int>(), c->sd_hlist.as<int>(),
                                      c->sd_cwork.as<int>(), &nheads, &nruns, &kx, c->sd_vals.p, c->samples.p,
                                      c->tmp.as<int>(), c->ctr(), c->trial_ctr(), e0, e1, c->hread(), c->stream,
                                      &herr);
            HIPCHK(c, herr);
            c->host_sample_entries += (unsigned long long)nent;
            c->ev_kind[eslot] = nruns > 0 ? 1 : 0;
        }
    } else {
        hipEvent_t e0, e1;
        HIPCHK(c, c->event_pair(e0, e1));
        HIPCHK(c, hipEventRecord(e0, c->stream));
        launch_samples(prec, c->g, c->rec_a.p, c->rec_g.p, fc, tp, v, L->child.as<int>(),
                       L->scanned.as<int>() + L->n, c->samples.p, c->ctr(), c->stream);
        HIPCHK(c, hipEventRecord(e1, c->stream));
    }
    // centre values of leaf-depth nodes (QEF off): only the leaf-depth level has them
    if (level_depth < 0) {
        launch_centers(prec, c->g, c->rec_a.p, c->rec_g.p, fc, v, c->cl_list.as<int>(), c->cl_pos.as<int>() + L->n,
                       c->samples.p, c->ctr(), c->stream);
    } else if (centre_runs) {
        int k = 0;
        HIPCHK(c, c->read_scan(&k, c->cl_pos.as<int>() + L->n, cpub));
        if (k > 0) {
            HIPCHK(c, c->sd_key.ensure((size_t)k * 8));
            HIPCHK(c, c->sd_key2.ensure((size_t)k * 8));
            HIPCHK(c, c->sd_val.ensure((size_t)k * 4));
            HIPCHK(c, c->sd_val2.ensure((size_t)k * 4));
            HIPCHK(c, c->sd_cwork.ensure(4 * ((size_t)k + 2) * 4));
            const int64_t need = std::max(radix_tmp_ints(k), scan_tmp_ints((int64_t)k + 1)) + 64;
            if ((size_t)need * 4 > c->tmp.cap) {
                // kernels queued before the count read may still use the workspace
                HIPCHK(c, c->wait());
                HIPCHK(c, c->tmp.ensure(need * 4));
            }
            hipError_t herr = hipSuccess;
            launch_centers_runs(prec, c->g, c->rec_a.p, c->rec_g.p, fc, v, c->cl_list.as<int>(), k,
                                c->sd_key.as<uint64_t>(), c->sd_val.as<int>(), c->sd_key2.as<uint64_t>(),
                                c->sd_val2.as<int>(), c->sd_cwork.as<int>(), c->tmp.as<int>(), c->samples.p, c->ctr(),
                                c->trial_ctr(), c->hread(), c->stream, &herr);
            HIPCHK(c, herr);
        }
    }
    if (split_curv)
        launch_empty_curv(c->g, c->spos.as<float4>(), c->nrm.as<float4>(), tp, v, group, kCurvOnly, c->samples.p,
                          (int)rsz, c->stream);
    launch_classify(prec, c->g, c->rec_a.p, c->rec_g.p, fc, tp, v, c->samples.p, sc_dev, c->ctr(), c->stream);
    HIPCHK(c, hipGetLastError());
    return PISURF_OK;
}

// bound on the nodes of a level built from device-resident counts
constexpr int64_t kTopMaxNodes = int64_t(1) << 22;

// the root node (resizeRootBoxConstR) as level 0, every level's count reset
int tree_root(pisurf_ctx* c, hipStream_t s) {
    resize_root(c);
    for (auto* l : c->levels) l->n = 0;
    while (c->levels.empty()) c->levels.push_back(new LevelBuf());
    LevelBuf* L = c->levels[0];
    L->n = 1;
    HIPCHK(c, L->ensure(1));
    const float4 rc4 = make_float4(c->root_c[0], c->root_c[1], c->root_c[2], c->root_half);
    const int8_t d0 = 0;
    const uint64_t id0 = 0;
    const int4 ip0 = make_int4(0, 0, 0, 0);
    HIPCHK(c, hipMemcpyAsync(L->ip.p, &ip0, sizeof(int4), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(L->c.p, &rc4, sizeof(float4), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(L->depth.p, &d0, 1, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(L->id.p, &id0, 8, hipMemcpyHostToDevice, s));
    return PISURF_OK;
}

// The launches of the levels above D_min on stream s with scan workspace tmp (counts stay on the
// device until top_levels_finish reads them).
int top_levels_issue(pisurf_ctx* c, hipStream_t s, DBuf& tmp) {
    c->top_t = 0;
    c->top_n.clear();
    c->top_ub.assign(1, 1);
    const int top = c->dmin;
    if (top <= 0) return PISURF_OK;
    // bounds: a level-l node that is not empty meets the particles' box grown by the margin, so
    // level l holds at most geo(l) non-empty nodes and level l + 1 at most 8 geo(l) nodes
    std::vector<int64_t>& ub = c->top_ub;
    const double m = double(c->prm.radius) * double(c->prm.smooth_factor);
    for (int l = 1; l <= top; l++) {
        const double size = std::ldexp(2.0 * (double)c->root_half, -(l - 1));
        double geo = 1;
        for (int a = 0; a < 3; a++)
            geo *= std::floor(((double)c->ginfo.bounding[2 * a + 1] - (double)c->ginfo.bounding[2 * a] + 2 * m) / size) + 3;
        const double b = 8 * std::min((double)ub.back(), geo);
        if (b > (double)kTopMaxNodes) break;
        ub.push_back((int64_t)b);
    }
    const int t = (int)ub.size() - 1;  // levels 0 .. t - 1 are built here, level t's nodes written
    if (t <= 0) return PISURF_OK;
    auto level = [&](size_t i) {
        while (c->levels.size() <= i) c->levels.push_back(new LevelBuf());
        return c->levels[i];
    };
    // every buffer first: the launches below run while the host goes on
    int64_t tmp_need = 0;
    for (int l = 0; l <= t; l++) {
        LevelBuf* L = level(l);
        HIPCHK(c, L->ensure(ub[l]));
        L->n = ub[l];
        tmp_need = std::max(tmp_need, scan_tmp_ints(ub[l] + 1) + 64);
    }
    if ((size_t)tmp_need * 4 > tmp.cap) {
        HIPCHK(c, hipStreamSynchronize(s));
        HIPCHK(c, tmp.ensure(tmp_need * 4));
    }
    HIPCHK(c, c->lvcnt.ensure(64 * sizeof(int)));
    int* cnt = c->lvcnt.as<int>();
    const int prec = prec_of(c);
    const FieldConst fc = field_const(c);
    const TreeParams tp = tree_params(c);
    auto issue = [&]() {
        launch_fill_i32(cnt, 1, 1, s);
        for (int l = 0; l < t; l++) {
            LevelBuf* L = c->levels[l];
            NodeLevel v = L->view();
            v.ndev = cnt + l;
            const double half = std::ldexp((double)c->root_half, -l);
            const double w = (2 * half + 2 * (double)tp.margin) / (double)c->g.cell + 1;
            const int group = (w * w <= 6) ? 4 : (w * w <= 8) ? 8 : (w * w <= 16) ? 16 : 64;
            launch_empty_curv(c->g, c->spos.as<float4>(), c->nrm.as<float4>(), tp, v, group, kEmptyAndCurv, nullptr, 0,
                              s);
            launch_classify(prec, c->g, c->rec_a.p, c->rec_g.p, fc, tp, v, c->samples.p, nullptr, c->ctr(), s);
            exclusive_scan_i32(L->flag.as<int>(), L->scanned.as<int>(), ub[l] + 1, tmp.as<int>(), s);
            launch_children_counted(v, L->scanned.as<int>(), c->levels[l + 1]->view(), cnt + l + 1, s);
        }
    };
    // With level_graph the launches (kernels only: fills are kernels too) are captured as a HIP
    // graph: the first frame instantiates it, later frames update its parameters in place (the same
    // sequence of kernels whenever the level count and the scans' tile counts repeat) or
    // instantiate anew. The update and the graph launch cost about as much host time as the 40-odd
    // launches they replace, so the launches go to the stream directly by default.
    bool graphed = false;
    if (c->prm.level_graph && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        issue();
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &gr);
        if (ec == hipSuccess && gr) {
            hipError_t e = hipErrorNotInitialized;
            if (c->top_exec) {
                hipGraphNode_t bad = nullptr;
                hipGraphExecUpdateResult res;
                e = hipGraphExecUpdate(c->top_exec, gr, &bad, &res);
                if (e != hipSuccess) {
                    (void)hipGetLastError();
                    (void)hipGraphExecDestroy(c->top_exec);
                    c->top_exec = nullptr;
                }
            }
            if (!c->top_exec) e = hipGraphInstantiate(&c->top_exec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) return c->hipfail(e, "octree level graph");
            HIPCHK(c, hipGraphLaunch(c->top_exec, s));
            graphed = true;
        } else {
            if (gr) (void)hipGraphDestroy(gr);
            (void)hipGetLastError();  // capture refused: the launches go to the stream directly
        }
    }
    if (!graphed) issue();
    HIPCHK(c, hipGetLastError());
    c->top_t = t;
    c->top_n.clear();
    return PISURF_OK;
}

// The counts of the issued levels (one read on the context's stream, which must follow their
// launches); *next = the index of the first level built the ordinary way.
int top_levels_read(pisurf_ctx* c) {
    const int t = c->top_t;
    c->top_n.assign(t > 0 ? t + 1 : 0, 0);
    if (t > 0) HIPCHK(c, c->read(c->top_n.data(), c->lvcnt.as<int>(), (t + 1) * sizeof(int)));
    return PISURF_OK;
}
int top_levels_finish(pisurf_ctx* c, size_t* next) {
    *next = 0;
    const int t = c->top_t;
    if (t <= 0) return PISURF_OK;
    int rc;
    if ((int)c->top_n.size() != t + 1 && (rc = top_levels_read(c))) return rc;
    const std::vector<int>& n = c->top_n;
    for (int l = 0; l <= t; l++) {
        if (n[l] < 0 || n[l] > c->top_ub[l]) return c->fail(PISURF_E_STATE, "octree level count out of its bound");
        c->levels[l]->n = n[l];
    }
    *next = (size_t)t;
    return PISURF_OK;
}

// The levels above D_min take no samples: a node there is INTERNAL exactly when it is not empty
// (afterSampleEval, src/surface_reconstructor.cpp:237-273, with depth < D_min). They are built
// with their node counts kept in device memory and their arrays and grids sized by an upper bound,
// so the host issues every launch without waiting; the counts are read once, together with the
// count of the first level built the ordinary way. Level 0 (the root) is in place. Returns, in
// *next, the index of that first ordinary level.
int top_levels(pisurf_ctx* c, size_t* next) {
    int rc = top_levels_issue(c, c->stream, c->tmp);
    if (rc) return rc;
    return top_levels_finish(c, next);
}

// A frame's root and levels above D_min, issued on a second stream as soon as the grid is built
// (they read the grid, the bounding box and the constants, and classify by emptiness alone: no
// samples, normals or particle records), so they run while the anisotropy and the normals run on
// the context's stream; do_tree joins them. The second stream is the one the mesh download uses
// at the end of the frame.
int tree_start(pisurf_ctx* c) {
    c->top_pending = false;
    if (!c->dstream) HIPCHK(c, hipStreamCreateWithFlags(&c->dstream, hipStreamNonBlocking));
    if (!c->tev_grid) HIPCHK(c, hipEventCreateWithFlags(&c->tev_grid, hipEventDisableTiming));
    if (!c->tev_top) HIPCHK(c, hipEventCreateWithFlags(&c->tev_top, hipEventDisableTiming));
    HIPCHK(c, hipEventRecord(c->tev_grid, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->dstream, c->tev_grid, 0));
    int rc = tree_root(c, c->dstream);
    if (!rc) rc = top_levels_issue(c, c->dstream, c->tmp_top);
    if (rc) return rc;
    HIPCHK(c, hipEventRecord(c->tev_top, c->dstream));
    c->top_pending = true;
    return PISURF_OK;
}

int do_tree(pisurf_ctx* c) {
    int rc = frame_ready(c);
    if (rc) return rc;
    const double t0 = now_ms();
    size_t li = 0;
    auto level = [&](size_t i) {
        while (c->levels.size() <= i) c->levels.push_back(new LevelBuf());
        return c->levels[i];
    };
    if (c->top_pending) {
        // root and levels above D_min already issued on dstream (tree_start)
        c->top_pending = false;
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->tev_top, 0));
        if ((rc = top_levels_finish(c, &li))) return rc;
    } else {
        if ((rc = tree_root(c, c->stream))) return rc;
        if ((rc = top_levels(c, &li))) return rc;
    }
    LevelBuf* L = nullptr;
    while (true) {
        L = level(li);
        if (L->n == 0) break;  // every node above D_min empty
        rc = sweep(c, L, nullptr, (int)li);
        if (rc) return rc;
        const ScanPub pub = c->scan_pub();
        exclusive_scan_i32(L->flag.as<int>(), L->scanned.as<int>(), L->n + 1, c->tmp.as<int>(), c->stream, pub);
        int n_int = 0;
        HIPCHK(c, c->read_scan(&n_int, L->scanned.as<int>() + L->n, pub));
        c->harvest_events();
        if (n_int == 0) break;
        if ((int64_t)n_int * 8 > (int64_t)INT32_MAX / 2) return c->fail(PISURF_E_UNSUPPORTED, "octree level too large");
        LevelBuf* N = level(li + 1);
        N->n = (int64_t)n_int * 8;
        HIPCHK(c, N->ensure(N->n));
        const int64_t need_tmp = scan_tmp_ints(N->n + 1) + 64;
        if ((size_t)need_tmp * 4 > c->tmp.cap) {
            // grow the scan workspace without losing the grid (tmp holds no persistent data)
            HIPCHK(c, c->wait());
            HIPCHK(c, c->tmp.ensure(need_tmp * 4));
        }
        launch_children(L->view(), L->scanned.as<int>(), N->view(), c->stream);
        HIPCHK(c, hipGetLastError());
        li++;
    }
    // the last count read ended the loop: the stream holds only the tail of that scan, which the
    // extraction's launches follow in stream order
    c->tree_done = true;
    c->st.n_nodes = 0;
    for (auto* l : c->levels) c->st.n_nodes += l->n;
    c->st.ms_tree += now_ms() - t0;
    return PISURF_OK;
}

struct HostTree {
    std::vector<uint8_t> type;
    std::vector<double4> node;
    std::vector<uint64_t> id;
    std::vector<int32_t> child;  // global index of the first child, -1
    std::vector<float4> cen;
    std::vector<int8_t> depth;
    std::vector<int64_t> level_off;
};

int download_tree(pisurf_ctx* c, HostTree& t, bool geometry) {
    int64_t tot = 0;
    t.level_off.clear();
    for (auto* l : c->levels) {
        if (l->n == 0) break;
        t.level_off.push_back(tot);
        tot += l->n;
    }
    t.level_off.push_back(tot);
    t.type.resize(tot);
    t.node.resize(tot);
    t.id.resize(tot);
    t.child.resize(tot);
    if (geometry) {
        t.cen.resize(tot);
        t.depth.resize(tot);
    }
    for (size_t li = 0; li + 1 < t.level_off.size(); li++) {
        LevelBuf* l = c->levels[li];
        const int64_t o = t.level_off[li];
        HIPCHK(c, hipMemcpyAsync(&t.type[o], l->type.p, l->n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(&t.node[o], l->node.p, l->n * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(&t.id[o], l->id.p, l->n * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(&t.child[o], l->child.p, l->n * 4, hipMemcpyDeviceToHost, c->stream));
        if (geometry) {
            HIPCHK(c, hipMemcpyAsync(&t.cen[o], l->c.p, l->n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipMemcpyAsync(&t.depth[o], l->depth.p, l->n, hipMemcpyDeviceToHost, c->stream));
        }
    }
    HIPCHK(c, c->wait());
    for (size_t li = 0; li + 1 < t.level_off.size(); li++) {
        const int64_t next = t.level_off[li + 1];
        for (int64_t i = t.level_off[li]; i < next; i++)
            if (t.child[i] >= 0) t.child[i] = (int32_t)(next + t.child[i]);
    }
    return PISURF_OK;
}

// Dual-cell enumeration of the octree (include/traverse.h:12-312 with the VisitorExtract
// callbacks, src/visitorextract.cpp:229-452). The recursion order fixes the order of the dual
// cells and therefore the mesh numbering.
struct Extractor {
    const HostTree& t;
    struct Dual {
        int32_t n[8];
        uint8_t sidx;
    };
    std::vector<Dual> duals;
    explicit Extractor(const HostTree& tr) : t(tr) {}
    bool leaf(int n) const { return t.type[n] == 2 || t.type[n] == 0; }
    int gen(int n, int i) const { return leaf(n) ? n : t.child[n] + i; }
    static int IX(int x, int y, int z) { return x | (y << 1) | (z << 2); }
    void vert(const int* n) {
        bool all = true;
        for (int i = 0; i < 8; i++) all = all && leaf(n[i]);
        if (all) {
            int index = 0;
            for (int i = 0; i < 8; i++)
                if (t.node[n[i]].w > 0) index |= 1 << i;
            const uint8_t code = kDmcCase[index];
            if (code != 0 && code != 255) {
                Dual d;
                const uint32_t perm = kDmcPerm[index];
                for (int i = 0; i < 8; i++) d.n[i] = n[(perm >> (4 * i)) & 15];
                d.sidx = (uint8_t)index;
                duals.push_back(d);
            }
            return;
        }
        int c[8];
        for (int k = 0; k < 8; k++) c[k] = gen(n[k], k ^ 7);
        vert(c);
    }
    void edge(int axis, int n00, int n10, int n01, int n11) {
        if (leaf(n00) && leaf(n10) && leaf(n01) && leaf(n11)) return;
        int c[8];
        // slot (i, a, b) along `axis`, the other two axes (a, b) in x, y, z order
        auto slot = [&](int i, int a, int b) {
            if (axis == 0) return IX(i, a, b);
            if (axis == 1) return IX(a, i, b);
            return IX(a, b, i);
        };
        for (int i = 0; i < 2; i++) {
            c[slot(i, 0, 0)] = gen(n00, slot(i, 1, 1));
            c[slot(i, 1, 0)] = gen(n10, slot(i, 0, 1));
            c[slot(i, 0, 1)] = gen(n01, slot(i, 1, 0));
            c[slot(i, 1, 1)] = gen(n11, slot(i, 0, 0));
        }
        for (int i = 0; i < 2; i++) edge(axis, c[slot(i, 0, 0)], c[slot(i, 1, 0)], c[slot(i, 0, 1)], c[slot(i, 1, 1)]);
        vert(c);
    }
    void face(int axis, int n0, int n1) {
        if (leaf(n0) && leaf(n1)) return;
        int c[8];
        auto slot = [&](int s, int a, int b) {
            if (axis == 0) return IX(s, a, b);
            if (axis == 1) return IX(a, s, b);
            return IX(a, b, s);
        };
        for (int i = 0; i < 4; i++) {
            const int a = i & 1, b = (i >> 1) & 1;
            c[slot(0, a, b)] = gen(n0, slot(1, a, b));
            c[slot(1, a, b)] = gen(n1, slot(0, a, b));
        }
        for (int i = 0; i < 4; i++) {
            const int a = i & 1, b = (i >> 1) & 1;
            face(axis, c[slot(0, a, b)], c[slot(1, a, b)]);
        }
        // the two edge directions perpendicular to `axis`, in x, y, z order
        for (int i = 0; i < 2; i++) {
            for (int e = 0; e < 3; e++) {
                if (e == axis) continue;
                if (e == 0) edge(0, c[IX(i, 0, 0)], c[IX(i, 1, 0)], c[IX(i, 0, 1)], c[IX(i, 1, 1)]);
                else if (e == 1) edge(1, c[IX(0, i, 0)], c[IX(1, i, 0)], c[IX(0, i, 1)], c[IX(1, i, 1)]);
                else edge(2, c[IX(0, 0, i)], c[IX(1, 0, i)], c[IX(0, 1, i)], c[IX(1, 1, i)]);
            }
        }
        vert(c);
    }
    void node(int n) {
        if (leaf(n)) return;
        int c[8];
        for (int i = 0; i < 8; i++) {
            c[i] = gen(n, i);
            node(c[i]);
        }
        for (int i = 0; i < 4; i++) {
            const int a = i & 1, b = (i >> 1) & 1;
            face(0, c[IX(0, a, b)], c[IX(1, a, b)]);
            face(1, c[IX(a, 0, b)], c[IX(a, 1, b)]);
            face(2, c[IX(a, b, 0)], c[IX(a, b, 1)]);
        }
        for (int i = 0; i < 2; i++) {
            edge(0, c[IX(i, 0, 0)], c[IX(i, 1, 0)], c[IX(i, 0, 1)], c[IX(i, 1, 1)]);
            edge(1, c[IX(0, i, 0)], c[IX(1, i, 0)], c[IX(0, i, 1)], c[IX(1, i, 1)]);
            edge(2, c[IX(0, 0, i)], c[IX(1, 0, i)], c[IX(0, 1, i)], c[IX(1, 1, i)]);
        }
        vert(c);
    }
};


int read_int(pisurf_ctx* c, const int* dptr, int64_t& out) {
    int v = 0;
    HIPCHK(c, c->read(&v, dptr, 4));
    out = v;
    return PISURF_OK;
}
// a scan's last element (dlast), published by the scan itself (c->scan_pub() before the scan)
int read_scan_int(pisurf_ctx* c, const int* dlast, ScanPub p, int64_t& out) {
    int v = 0;
    HIPCHK(c, c->read_scan(&v, dlast, p));
    out = v;
    return PISURF_OK;
}

int ensure_tmp(pisurf_ctx* c, int64_t n) {
    const int64_t need = std::max(radix_tmp_ints(n), scan_tmp_ints(n + 1)) + 64;
    if ((size_t)need * 4 > c->tmp.cap) {
        HIPCHK(c, c->wait());
        HIPCHK(c, c->tmp.ensure(need * 4));
    }
    return PISURF_OK;
}

// Extraction on the device (k_extract.hip): leaf corners -> sorted points -> dual cells ->
// triangle corners -> vertices keyed by leaf pairs -> bisection -> triangle dedupe.
int do_extract_device(pisurf_ctx* c) {
    const double t0 = now_ms();
    hipStream_t s = c->stream;
    TreeView tv;
    std::memset(&tv, 0, sizeof(tv));
    int64_t T = 0;
    for (auto* l : c->levels) {
        if (l->n == 0) break;
        if (tv.nlev >= kMaxLevels) return c->fail(PISURF_E_UNSUPPORTED, "too many octree levels");
        tv.off[tv.nlev] = T;
        tv.type[tv.nlev] = l->type.as<uint8_t>();
        tv.child[tv.nlev] = l->child.as<int>();
        T += l->n;
        tv.nlev++;
    }
    tv.dp = tv.nlev - 1;
    if (tv.dp > 21) return c->fail(PISURF_E_UNSUPPORTED, "octree deeper than 21 levels");
    DBuf* E = c->eb;
    // the tree index from which the point pass's descents start
    tv.glev = tree_index_level(tv.dp);
    HIPCHK(c, E[36].ensure(((size_t)1 << (3 * tv.glev)) * 4));
    int* coarse = nullptr;
    if (tv.glev >= 1) {
        HIPCHK(c, E[37].ensure(((size_t)1 << (3 * (tv.glev - 1))) * 4));
        coarse = E[37].as<int>();
    }
    tv.gidx = E[36].as<int>();
    launch_tree_index(tv, E[36].as<int>(), coarse, s);
    int rc;
    // leaves
    HIPCHK(c, E[0].ensure((T + 1) * 4));
    HIPCHK(c, E[1].ensure((T + 1) * 4));
    LeafLevels ll;
    std::memset(&ll, 0, sizeof(ll));
    ll.nlev = tv.nlev;
    ll.dp = tv.dp;
    for (int L = 0; L < tv.nlev; L++) {
        LevelBuf* lb = c->levels[L];
        ll.off[L] = tv.off[L];
        ll.type[L] = lb->type.as<uint8_t>();
        ll.node[L] = lb->node.as<double4>();
        ll.ip[L] = lb->ip.as<int4>();
        ll.tile0[L + 1] = ll.tile0[L] + corner_tiles(lb->n);
    }
    ll.off[tv.nlev] = T;
    launch_leaf_flags(ll, E[0].as<int>(), s);
    if ((rc = ensure_tmp(c, T + 1))) return rc;
    ScanPub pub = c->scan_pub();
    exclusive_scan_i32(E[0].as<int>(), E[1].as<int>(), T + 1, c->tmp.as<int>(), s, pub);
    int64_t nleaf;
    if ((rc = read_scan_int(c, E[1].as<int>() + T, pub, nleaf))) return rc;
    const int* lidx = E[1].as<int>();
    // distinct corner points: per tile of 256 nodes of a level, one record per point (its key and
    // the tile's leaves around it) in the tile's own region; then the keys of all records in one
    // array, sorted
    const int64_t ntiles = ll.tile0[tv.nlev];
    const int64_t nslots = ntiles * corner_tile_records();
    if (nslots >= INT32_MAX) return c->fail(PISURF_E_UNSUPPORTED, "octree too large for the corner records");
    HIPCHK(c, E[2].ensure(std::max<int64_t>(nleaf, 1) * sizeof(double4)));
    HIPCHK(c, E[7].ensure(std::max<int64_t>(nslots, 1) * 8));
    HIPCHK(c, E[29].ensure(std::max<int64_t>(nslots, 1) * 8 * 4));
    HIPCHK(c, E[30].ensure((ntiles + 1) * 4));
    HIPCHK(c, E[31].ensure((ntiles + 1) * 4));
    double4* leaf_node = E[2].as<double4>();
    HIPCHK(c, E[28].ensure(std::max<int64_t>(nleaf, 1)));
    launch_corner_tiles(ll, lidx, leaf_node, E[28].as<uint8_t>(), E[7].as<uint64_t>(), E[29].as<int>(), E[30].as<int>(),
                        s);
    if ((rc = ensure_tmp(c, ntiles + 1))) return rc;
    pub = c->scan_pub();
    exclusive_scan_i32(E[30].as<int>(), E[31].as<int>(), ntiles + 1, c->tmp.as<int>(), s, pub);
    HIPCHK(c, hipGetLastError());
    int64_t nent = 0;
    if ((rc = read_scan_int(c, E[31].as<int>() + ntiles, pub, nent))) return rc;
    const int ckbits = std::max(1, 3 * tv.dp);
    const int cvbits = pack_vbits(ckbits, nslots);
    HIPCHK(c, E[3].ensure(std::max<int64_t>(nent, 1) * 8));
    HIPCHK(c, E[4].ensure(std::max<int64_t>(nent, 1) * 4));
    HIPCHK(c, E[5].ensure(std::max<int64_t>(nent, 1) * 8));
    HIPCHK(c, E[6].ensure(std::max<int64_t>(nent, 1) * 4));
    launch_records_compact(E[30].as<int>(), E[31].as<int>(), ntiles, E[7].as<uint64_t>(), cvbits, E[3].as<uint64_t>(),
                           E[4].as<int>(), s);
    if ((rc = ensure_tmp(c, nent + 1))) return rc;
    const SortedKV ckv = radix_sort_entries(E[3].as<uint64_t>(), E[4].as<int>(), E[5].as<uint64_t>(), E[6].as<int>(),
                                            nent, ckbits, cvbits, c->tmp.as<int>(), s);
    HIPCHK(c, hipGetLastError());
    // points -> dual cells
    // distinct points: run heads of the sorted records, one thread each
    HIPCHK(c, E[25].ensure(std::max<int64_t>(nent, 1) * 4));
    HIPCHK(c, E[9].ensure(head_list_tmp_ints(nent) * 4));
    pub = nent > 0 ? c->scan_pub() : ScanPub();
    launch_head_list(ckv, nent, E[25].as<int>(), E[9].as<int>(), c->tmp.as<int>(), s, pub);
    int64_t npts;
    if ((rc = read_scan_int(c, head_list_count(E[9].as<int>(), nent), pub, npts))) return rc;
    HIPCHK(c, E[26].ensure((npts + 1) * 4));
    HIPCHK(c, E[27].ensure((npts + 1) * 4));
    HIPCHK(c, E[24].ensure(std::max<int64_t>(npts, 1) * 9 * 4));
    launch_points(tv, lidx, E[28].as<uint8_t>(), ckv, nent, E[29].as<int>(), E[25].as<int>(), npts, E[26].as<int>(),
                  E[24].as<int>(), s);
    pub = c->scan_pub();
    exclusive_scan_i32(E[26].as<int>(), E[27].as<int>(), npts + 1, c->tmp.as<int>(), s, pub);
    int64_t nd;
    if ((rc = read_scan_int(c, E[27].as<int>() + npts, pub, nd))) return rc;
    HIPCHK(c, E[10].ensure(std::max<int64_t>(nd, 1) * 9 * 4));
    launch_points_compact(E[26].as<int>(), E[27].as<int>(), E[24].as<int>(), npts, E[10].as<int>(), s);
    // triangles
    HIPCHK(c, E[11].ensure((nd + 1) * 4));
    HIPCHK(c, E[12].ensure((nd + 1) * 4));
    launch_tri_count(E[10].as<int>(), nd, E[11].as<int>(), s);
    pub = c->scan_pub();
    exclusive_scan_i32(E[11].as<int>(), E[12].as<int>(), nd + 1, c->tmp.as<int>(), s, pub);
    int64_t nt;
    if ((rc = read_scan_int(c, E[12].as<int>() + nd, pub, nt))) return rc;
    const int64_t nc = 3 * nt;
    HIPCHK(c, E[13].ensure(std::max<int64_t>(nc, 1) * 8));
    HIPCHK(c, E[14].ensure(std::max<int64_t>(nc, 1) * 4));
    HIPCHK(c, E[15].ensure(std::max<int64_t>(nc, 1) * 8));
    HIPCHK(c, E[16].ensure(std::max<int64_t>(nc, 1) * 4));
    HIPCHK(c, E[17].ensure(std::max<int64_t>(nc, 1) * 8));
    launch_tri_corners(E[10].as<int>(), nd, E[12].as<int>(), E[13].as<uint64_t>(), E[14].as<int>(), E[17].as<int>(), s);
    // vertices: corners keyed by their unordered leaf pair, numbered by first occurrence
    // about six corners per distinct vertex on a surface: a table of 1.25 nc slots or more stays
    // sparse, and holds every key even if all were distinct
    const int64_t vtab = pow2_at_least(nc + nc / 4 + 1);
    HIPCHK(c, E[15].ensure(vtab * sizeof(unsigned)));
    HIPCHK(c, E[18].ensure((nc + 1) * 4));
    HIPCHK(c, E[19].ensure((nc + 1) * 4));
    if ((rc = ensure_tmp(c, nc + 1))) return rc;
    int* vslot = E[14].as<int>();  // the corner indices written there are implicit
    launch_hash_dedupe(E[13].as<uint64_t>(), nc, E[15].as<unsigned>(), vtab, vslot, E[18].as<int>(), s);
    pub = c->scan_pub();
    exclusive_scan_i32(E[18].as<int>(), E[19].as<int>(), nc + 1, c->tmp.as<int>(), s, pub);
    int64_t nv;
    if ((rc = read_scan_int(c, E[19].as<int>() + nc, pub, nv))) return rc;
    HIPCHK(c, E[20].ensure(std::max<int64_t>(nc, 1) * 4));
    HIPCHK(c, E[21].ensure(std::max<int64_t>(nv, 1) * 8 * 8));
    int* cvid = E[20].as<int>();
    launch_vertex_ids_hashed(vslot, E[15].as<unsigned>(), E[19].as<int>(), nc, E[17].as<int>(), leaf_node, cvid,
                             E[21].as<double>(), s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, E[22].ensure(std::max<int64_t>(nv, 1) * 3 * 8));
    if (!c->eb0) HIPCHK(c, hipEventCreate(&c->eb0));
    if (!c->eb1) HIPCHK(c, hipEventCreate(&c->eb1));
    if (!c->dstream) HIPCHK(c, hipStreamCreateWithFlags(&c->dstream, hipStreamNonBlocking));
    // triangle dedupe by sorted vertex triple; the first triangle of a repeated triple stays
    int64_t ntk = nt;
    HIPCHK(c, E[8].ensure((ntk + 1) * 4));
    HIPCHK(c, E[9].ensure((ntk + 1) * 4));
    if ((rc = ensure_tmp(c, ntk + 1))) return rc;
    if (nv < ((int64_t)1 << 21)) {
        // one 63-bit key per triangle, hashed
        const int64_t ttab = pow2_at_least(2 * ntk);
        HIPCHK(c, E[3].ensure(std::max<int64_t>(ntk, 1) * 8));
        HIPCHK(c, E[4].ensure(std::max<int64_t>(ntk, 1) * 4));
        HIPCHK(c, E[5].ensure(ttab * sizeof(unsigned)));
        launch_tri_key63(cvid, nt, E[3].as<uint64_t>(), s);
        launch_hash_dedupe(E[3].as<uint64_t>(), nt, E[5].as<unsigned>(), ttab, E[4].as<int>(), E[8].as<int>(), s);
    } else {
        // larger meshes: two stable sorts (largest id, then the lower pair) group equal triples
        HIPCHK(c, E[3].ensure(std::max<int64_t>(ntk, 1) * 8));
        HIPCHK(c, E[4].ensure(std::max<int64_t>(ntk, 1) * 4));
        HIPCHK(c, E[5].ensure(std::max<int64_t>(ntk, 1) * 8));
        HIPCHK(c, E[6].ensure(std::max<int64_t>(ntk, 1) * 4));
        HIPCHK(c, E[15].ensure(std::max<int64_t>(ntk, 1) * 8));
        HIPCHK(c, E[16].ensure(std::max<int64_t>(ntk, 1) * 8));
        uint32_t* tk = (uint32_t*)E[3].p;
        launch_tri_key_hi(cvid, nt, tk, E[4].as<int>(), s);
        radix_sort_pairs(tk, E[4].as<int>(), (uint32_t*)E[5].p, E[6].as<int>(), nt, bits_for(nv), c->tmp.as<int>(),
                         s);
        launch_tri_key_lo(cvid, E[4].as<int>(), nt, E[5].as<uint64_t>(), E[6].as<int>(), s);
        radix_sort_pairs64(E[5].as<uint64_t>(), E[6].as<int>(), E[15].as<uint64_t>(), E[16].as<int>(), nt,
                           32 + bits_for(nv), c->tmp.as<int>(), s);
        launch_tri_keep(cvid, E[6].as<int>(), nt, E[8].as<int>(), s);
    }
    pub = c->scan_pub();
    exclusive_scan_i32(E[8].as<int>(), E[9].as<int>(), nt + 1, c->tmp.as<int>(), s, pub);
    int64_t ntf;
    if ((rc = read_scan_int(c, E[9].as<int>() + nt, pub, ntf))) return rc;
    HIPCHK(c, E[23].ensure(std::max<int64_t>(ntf, 1) * 3 * 4));
    launch_tri_out(cvid, E[8].as<int>(), E[9].as<int>(), nt, E[23].as<int>(), s);
    HIPCHK(c, hipGetLastError());
    c->st.ms_extract_host += now_ms() - t0;  // device extraction time (topology)
    const double t2 = now_ms();
    // The triangles go to the host on a second stream while the vertices are placed; the
    // bisection runs in chunks, and each chunk's vertices follow the triangles down as soon as
    // the chunk is done.
    c->mesh_t.resize(ntf * 3);
    c->mesh_v.resize(nv * 3);
    HIPCHK(c, hipEventRecord(c->eb0, s));
    HIPCHK(c, hipStreamWaitEvent(c->dstream, c->eb0, 0));
    if (ntf) HIPCHK(c, hipMemcpyAsync(c->mesh_t.data(), E[23].p, ntf * 3 * 4, hipMemcpyDeviceToHost, c->dstream));
    if (c->ev_chunk.empty()) {
        c->ev_chunk.resize(3, nullptr);
        for (auto& e : c->ev_chunk) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    // From 2^20 vertices on, three parts of 40, 35 and 25 %: the download of the last part is
    // the only one not hidden; smaller launches would lose more to the tail of the bisection
    // (threads take different numbers of steps) than the overlap gains.
    int64_t part_end[3] = {nv, nv, nv};
    if (nv >= ((int64_t)1 << 20)) {
        part_end[0] = nv * 2 / 5;
        part_end[1] = nv * 3 / 4;
    }
    int64_t v0 = 0;
    for (int ci = 0; ci < 3 && v0 < nv; ci++) {
        const int64_t m = part_end[ci] - v0;
        if (m <= 0) continue;
        launch_bisect(prec_of(c), c->g, c->rec_a.p, c->rec_g.p, field_const(c), c->prm.radius / 2,
                      E[21].as<double>() + 8 * v0, m, E[22].as<double>() + 3 * v0, nullptr, c->ctr(), s);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev_chunk[ci], s));
        HIPCHK(c, hipStreamWaitEvent(c->dstream, c->ev_chunk[ci], 0));
        HIPCHK(c, hipMemcpyAsync(c->mesh_v.data() + 3 * v0, E[22].as<double>() + 3 * v0, m * 3 * 8,
                                 hipMemcpyDeviceToHost, c->dstream));
        v0 = part_end[ci];
    }
    HIPCHK(c, hipEventRecord(c->eb1, s));
    HIPCHK(c, c->wait());
    HIPCHK(c, stream_wait_on(c->dstream, c->wev));
    {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->eb0, c->eb1) == hipSuccess) c->st.ms_bisect += ms;
    }
    c->st.ms_extract_host += now_ms() - t2;
    c->st.n_dual_cells = nd;
    c->st.n_verts = nv;
    c->st.n_tris = ntf;
    return PISURF_OK;
}

int do_extract(pisurf_ctx* c);

int do_extract_host(pisurf_ctx* c);

int do_extract(pisurf_ctx* c) {
    if (!c->tree_done) return c->fail(PISURF_E_STATE, "octree not built");
    if (c->prm.mesh_order == 0) return do_extract_device(c);
    return do_extract_host(c);
}

int do_extract_host(pisurf_ctx* c) {
    double t0 = now_ms();
    HostTree t;
    int rc = download_tree(c, t, false);
    if (rc) return rc;
    Extractor ex(t);
    if (!t.type.empty()) ex.node(0);
    // mesh topology (VisitorExtract::generate_mesh, src/visitorextract.cpp:203-227; Mesh::insert_vert /
    // insert_tri, src/global.cpp:13-53): a vertex is keyed by the unordered pair of leaf ids of its
    // edge, the first dual cell to use it fixes the edge orientation; triangles are keyed by their
    // sorted vertex triple.
    std::vector<int32_t> vend;  // ordered node pair per vertex
    PinnedVec<int32_t>& tris = c->mesh_t;
    tris.clear();
    KeyMap vmap, tmap;
    vmap.reset(ex.duals.size() * 2 + 16);
    tmap.reset(ex.duals.size() * 3 + 16);
    int canon_slot[256];
    for (int i = 0; i < 256; i++) canon_slot[i] = -1;
    for (int k = 0; k < PISURF_DMC_NCASES; k++) canon_slot[kDmcCanon[k]] = k;
    for (const auto& d : ex.duals) {
        const int slot = canon_slot[kDmcCase[d.sidx]];
        if (slot < 0) continue;
        for (int tt = 0; tt < kDmcNTri[slot]; tt++) {
            int vid[3];
            for (int k = 0; k < 3; k++) {
                const int e = kDmcTri[slot][tt][k];
                const int na = d.n[kDmcEdge[e][0]], nb = d.n[kDmcEdge[e][1]];
                uint64_t ia = t.id[na], ib = t.id[nb];
                if (ia > ib) std::swap(ia, ib);
                auto r = vmap.try_emplace(ia, ib, (int32_t)(vend.size() / 2));
                if (r.second) {
                    vend.push_back(na);
                    vend.push_back(nb);
                }
                vid[k] = r.first + 1;
            }
            int t0v, t1v, t2v;
            if (kDmcMirror[d.sidx]) { t0v = vid[0]; t1v = vid[1]; t2v = vid[2]; }
            else { t0v = vid[2]; t1v = vid[1]; t2v = vid[0]; }
            int s[3] = {t0v, t1v, t2v};
            std::sort(s, s + 3);
            const uint64_t ka = ((uint64_t)(uint32_t)s[0] << 32) | (uint32_t)s[1];
            auto r = tmap.try_emplace(ka, (uint64_t)(uint32_t)s[2], (int32_t)(tris.size() / 3));
            if (r.second) {
                tris.push_back(t0v);
                tris.push_back(t1v);
                tris.push_back(t2v);
            }
        }
    }
    const int64_t nv = (int64_t)vend.size() / 2;
    std::vector<double> ends(nv * 8);
    for (int64_t i = 0; i < nv; i++) {
        const double4 a = t.node[vend[2 * i]], b = t.node[vend[2 * i + 1]];
        double* e = &ends[8 * i];
        e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = a.w;
        e[4] = b.x; e[5] = b.y; e[6] = b.z; e[7] = b.w;
    }
    c->st.n_dual_cells = (int64_t)ex.duals.size();
    c->st.ms_extract_host += now_ms() - t0;
    t0 = now_ms();
    c->mesh_v.assign(nv * 3, 0.0);
    if (nv > 0) {
        HIPCHK(c, c->q.ensure(nv * 8 * sizeof(double)));
        HIPCHK(c, c->out1.ensure(nv * 3 * sizeof(double)));
        HIPCHK(c, hipMemcpyAsync(c->q.p, ends.data(), nv * 8 * sizeof(double), hipMemcpyHostToDevice, c->stream));
        launch_bisect(prec_of(c), c->g, c->rec_a.p, c->rec_g.p, field_const(c), c->prm.radius / 2, c->q.as<double>(), nv,
                      c->out1.as<double>(), nullptr, c->ctr(), c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(c->mesh_v.data(), c->out1.p, nv * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, c->wait());
    }
    c->st.ms_bisect += now_ms() - t0;
    c->st.n_verts = nv;
    c->st.n_tris = (int64_t)tris.size() / 3;
    return PISURF_OK;
}

int sync_counters(pisurf_ctx* c) {
    Counters h;
    Counters hs[kCtrStripes];
    HIPCHK(c, hipMemcpyAsync(hs, c->counters.p, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, c->wait());
    std::memset(&h, 0, sizeof(h));
    for (const Counters& k : hs) {
        h.field_evals += k.field_evals;
        h.grad_evals += k.grad_evals;
        h.candidates += k.candidates;
        h.qef_solves += k.qef_solves;
        h.sample_evals += k.sample_evals;
        h.sample_candidates += k.sample_candidates;
        h.sample_unique += k.sample_unique;
        h.run_evals += k.run_evals;
        h.run_candidates += k.run_candidates;
    }
    c->st.field_evals = (int64_t)(h.field_evals + c->host_sample_entries);
    c->st.grad_evals = (int64_t)h.grad_evals;
    c->st.candidates = (int64_t)h.candidates;
    c->st.qef_solves = (int64_t)h.qef_solves;
    c->st.sample_evals = (int64_t)(h.sample_evals + c->host_sample_entries);
    c->st.sample_candidates = (int64_t)h.sample_candidates;
    c->st.sample_unique = (int64_t)h.sample_unique;
    c->st.run_evals = (int64_t)h.run_evals;
    c->st.run_candidates = (int64_t)h.run_candidates;
    c->st.fp64_eigensolves = (int64_t)c->n_fp64_eig;
    return PISURF_OK;
}

// ---- pisurf_run_batch: the frames of a batch meshed together ------------------------------
// Every frame has its own context (sharing the batch's stream). The frames go through each step
// together: a step's launches are issued for every frame, then the counts the host needs are read
// for all frames at once (one wait), so a step costs one synchronisation for the whole batch
// rather than one per frame. The octree levels advance in lockstep, and at each level one
// launch of the run queue evaluates the distinct sample points (and leaf-depth centres) of every
// frame, its workgroups taking (frame, run) items from a device counter.

constexpr bool kBatchEnabled = false;
constexpr int kMaxFramesInFlight = 16;
constexpr int kDefaultFramesInFlight = 2;

pisurf_ctx* make_frame_ctx(pisurf_ctx* c) {
    pisurf_ctx* f = new (std::nothrow) pisurf_ctx();
    if (!f) return nullptr;
    f->device = c->device;
    f->stream = c->stream;
    f->own_stream = false;
    if (hipEventCreateWithFlags(&f->wev, hipEventDisableTiming) != hipSuccess || f->rb.ensure(kReadbackBytes) != hipSuccess ||
        f->counters.ensure(kCtrStripes * sizeof(Counters)) != hipSuccess || !(f->hr = host_read_acquire())) {
        delete f;
        return nullptr;
    }
    std::memset(&f->st, 0, sizeof(f->st));
    return f;
}

// device-to-host reads of many frames' counts, completed by one wait on the shared stream; each
// frame context stages its reads in its own page-locked buffer
struct BatchReads {
    struct Req {
        pisurf_ctx* c;
        const void* d;
        void* host;
        size_t bytes, off;
    };
    std::vector<Req> req;
    pisurf_ctx* owner;
    explicit BatchReads(pisurf_ctx* o) : owner(o) {}
    int add(pisurf_ctx* c, const void* d, void* host, size_t bytes) {
        const size_t off = (c->rb_used + 15) & ~size_t(15);
        if (off + bytes > kReadbackBytes) return owner->fail(PISURF_E_STATE, "batched read-back buffer full");
        c->rb_used = off + bytes;
        req.push_back({c, d, host, bytes, off});
        HIPCHK(owner, hipMemcpyAsync((char*)c->rb.p + off, d, bytes, hipMemcpyDeviceToHost, owner->stream));
        return PISURF_OK;
    }
    int flush() {
        HIPCHK(owner, owner->wait());
        for (const Req& r : req) {
            std::memcpy(r.host, (const char*)r.c->rb.p + r.off, r.bytes);
            r.c->rb_used = 0;
        }
        req.clear();
        return PISURF_OK;
    }
};

struct BatchFrame {
    pisurf_ctx* c = nullptr;
    int k = 0;             // index of the frame in the caller's arrays
    bool active = false;   // still building its octree
    size_t li = 0;         // current level
    int nans = 0, ninv = 0;
    unsigned long long nfp = 0;
    // the current level
    bool leaf_level = false, dedupe = false, centers = false, split_curv = false;
    TreeParams tp;
    int group = 64;
    int ksamp = 0, kcent = 0, nh = 0, ng = 0, nr = 0, ngc = 0, nrc = 0, n_int = 0;
    int64_t nent = 0;
    SortedKV skv{nullptr, nullptr, 0}, ckv{nullptr, nullptr, 0};
    int T = 128;
};

LevelBuf* level_of(pisurf_ctx* c, size_t i) {
    while (c->levels.size() <= i) c->levels.push_back(new LevelBuf());
    return c->levels[i];
}

// the launches of sweep() up to the first count read (emptiness, sample list, the per-node sampling
// of levels that are not deduplicated); the counts to read are queued
int batch_level_begin(BatchFrame& f, BatchReads& br) {
    pisurf_ctx* c = f.c;
    LevelBuf* L = c->levels[f.li];
    const int ld = (int)f.li;
    const size_t rsz = prec_of(c) == PISURF_FP64 ? 8 : 4;
    HIPCHK(c, c->samples.ensure(L->n * 27 * rsz));
    const float half = std::ldexp(c->root_half, -ld);
    f.leaf_level = double((2.0f * half) - c->prm.radius) < 1e-8;
    f.tp = tree_params(c);
    if (f.leaf_level) f.tp.curv = 0;
    const int64_t need_tmp = scan_tmp_ints(L->n + 1) + 64;
    if ((size_t)need_tmp * 4 > c->tmp.cap) {
        HIPCHK(c, c->wait());
        HIPCHK(c, c->tmp.ensure(need_tmp * 4));
    }
    const double w = (2 * std::ldexp((double)c->root_half, -ld) + 2 * (double)f.tp.margin) / (double)c->g.cell + 1;
    f.group = (w * w <= 6) ? 4 : (w * w <= 8) ? 8 : (w * w <= 16) ? 16 : 64;
    f.split_curv = ld == c->dmin;
    const NodeLevel v = L->view();
    launch_empty_curv(c->g, c->spos.as<float4>(), c->nrm.as<float4>(), f.tp, v, f.group,
                      f.split_curv ? kEmptyOnly : kEmptyAndCurv, nullptr, 0, c->stream);
    HIPCHK(c, c->cl_flag.ensure((L->n + 1) * 4));
    HIPCHK(c, c->cl_pos.ensure((L->n + 1) * 4));
    HIPCHK(c, c->cl_list.ensure((L->n + 1) * 4));
    launch_sample_list(f.tp, v, L->flag.as<int>(), L->scanned.as<int>(), L->child.as<int>(), c->cl_flag.as<int>(),
                       c->cl_pos.as<int>(), c->cl_list.as<int>(), c->tmp.as<int>(), c->stream);
    f.dedupe = c->prm.sample_dedupe && ld <= 19 && ld >= c->dmin;
    f.centers = f.leaf_level && c->prm.qef_mode == 0 && ld >= c->dmin;
    f.ksamp = f.kcent = f.nh = f.ng = f.nr = f.ngc = f.nrc = 0;
    f.nent = 0;
    int rc = PISURF_OK;
    if (f.dedupe) rc = br.add(c, L->scanned.as<int>() + L->n, &f.ksamp, 4);
    else
        launch_samples(prec_of(c), c->g, c->rec_a.p, c->rec_g.p, field_const(c), f.tp, v, L->child.as<int>(),
                       L->scanned.as<int>() + L->n, c->samples.p, c->ctr(), c->stream);
    if (!rc && f.centers) rc = br.add(c, c->cl_pos.as<int>() + L->n, &f.kcent, 4);
    HIPCHK(c, hipGetLastError());
    return rc;
}

int batch_ensure_entries(pisurf_ctx* c, int64_t m, bool heads) {
    HIPCHK(c, c->sd_key.ensure(m * 8));
    HIPCHK(c, c->sd_key2.ensure(m * 8));
    HIPCHK(c, c->sd_val.ensure(m * 4));
    HIPCHK(c, c->sd_val2.ensure(m * 4));
    HIPCHK(c, c->sd_cwork.ensure(4 * (m + 2) * 4));
    if (heads) {
        const size_t rsz = prec_of(c) == PISURF_FP64 ? 8 : 4;
        HIPCHK(c, c->sd_head.ensure((2 * ((m + 255) / 256) + 2) * 4));
        HIPCHK(c, c->sd_hpos.ensure((m + 1) * 4));
        HIPCHK(c, c->sd_rep.ensure(m * 4));
        HIPCHK(c, c->sd_hlist.ensure(m * 4));
        HIPCHK(c, c->sd_vals.ensure(m * rsz));
    }
    const int64_t need = std::max(radix_tmp_ints(m), scan_tmp_ints(m + 1)) + 64;
    HIPCHK(c, c->tmp.ensure(std::max<size_t>(c->tmp.cap, need * 4)));
    return PISURF_OK;
}

// sample keys, their sort and the heads (deduplicated levels); centre keys, their sort and cell
// groups (leaf-depth level)
int batch_level_keys(BatchFrame& f, BatchReads& br) {
    pisurf_ctx* c = f.c;
    LevelBuf* L = c->levels[f.li];
    const NodeLevel v = L->view();
    if (f.ksamp > 0 && f.kcent > 0) return c->fail(PISURF_E_STATE, "samples and centres on one level");
    if (f.ksamp > 0) {
        f.nent = (int64_t)f.ksamp * 27;
        if (f.nent >= INT32_MAX) return c->fail(PISURF_E_UNSUPPORTED, "octree level has too many sample points");
        int rc = batch_ensure_entries(c, f.nent, true);
        if (rc) return rc;
        const float origin[3] = {c->root_c[0] - c->root_half, c->root_c[1] - c->root_half, c->root_c[2] - c->root_half};
        const uint64_t groups = (uint64_t)(c->g.n[0] + 2) * (uint64_t)(c->g.n[1] + 2) * (uint64_t)(c->g.n[2] + 2);
        const int lb = 3;
        const int kbits = std::min(64, bits_for((int64_t)groups + 1) + 3 * lb);
        const int vbits = pack_vbits(kbits, f.nent);
        launch_sample_dedupe_keys(c->g, v, L->child.as<int>(), f.nent, origin, lb, vbits, c->sd_key.as<uint64_t>(),
                                  c->sd_val.as<int>(), c->stream);
        f.skv = radix_sort_entries(c->sd_key.as<uint64_t>(), c->sd_val.as<int>(), c->sd_key2.as<uint64_t>(),
                                   c->sd_val2.as<int>(), f.nent, kbits, vbits, c->tmp.as<int>(), c->stream);
        launch_heads(v, L->child.as<int>(), f.skv, f.nent, c->sd_head.as<int>(), c->sd_rep.as<int>(),
                     c->sd_hlist.as<int>(), c->sd_hpos.as<int>(), c->tmp.as<int>(), c->stream);
        c->host_sample_entries += (unsigned long long)f.nent;
        rc = br.add(c, heads_count(c->sd_head.as<int>(), f.nent), &f.nh, 4);
        if (rc) return rc;
    }
    if (f.kcent > 0) {
        int rc = batch_ensure_entries(c, f.kcent, false);
        if (rc) return rc;
        f.ckv = launch_center_keys_sorted(c->g, v, c->cl_list.as<int>(), f.kcent, c->sd_key.as<uint64_t>(),
                                          c->sd_val.as<int>(), c->sd_key2.as<uint64_t>(), c->sd_val2.as<int>(),
                                          c->tmp.as<int>(), c->stream);
        launch_cell_groups(f.ckv, 0, nullptr, f.kcent, c->sd_cwork.as<int>(), c->tmp.as<int>(), c->stream);
        rc = br.add(c, groups_count(c->sd_cwork.as<int>(), f.kcent), &f.ngc, 4);
        if (rc) return rc;
    }
    HIPCHK(c, hipGetLastError());
    return PISURF_OK;
}

// the distinct sample points' cell groups; the centres' runs
int batch_level_groups(BatchFrame& f, BatchReads& br) {
    pisurf_ctx* c = f.c;
    if (f.nh > 0) {
        launch_cell_groups(f.skv, 9, c->sd_hlist.as<int>(), f.nh, c->sd_cwork.as<int>(), c->tmp.as<int>(), c->stream);
        int rc = br.add(c, groups_count(c->sd_cwork.as<int>(), f.nh), &f.ng, 4);
        if (rc) return rc;
    }
    if (f.kcent > 0 && f.ngc > 0) {
        int kx = 1, T = 128;
        run_shape_for((double)f.kcent / f.ngc, &kx, &T);
        launch_run_starts(c->g, f.ckv, 0, nullptr, f.kcent, f.ngc, kx, c->sd_cwork.as<int>(), c->tmp.as<int>(),
                          c->stream);
        int rc = br.add(c, runs_count(c->sd_cwork.as<int>(), f.kcent, f.ngc), &f.nrc, 4);
        if (rc) return rc;
    }
    HIPCHK(c, hipGetLastError());
    return PISURF_OK;
}

// the distinct sample points' runs
int batch_level_runs(BatchFrame& f, BatchReads& br) {
    pisurf_ctx* c = f.c;
    if (f.nh > 0 && f.ng > 0) {
        int kx = 1;
        run_shape_for((double)f.nh / f.ng, &kx, &f.T);
        launch_run_starts(c->g, f.skv, 9, c->sd_hlist.as<int>(), f.nh, f.ng, kx, c->sd_cwork.as<int>(),
                          c->tmp.as<int>(), c->stream);
        int rc = br.add(c, runs_count(c->sd_cwork.as<int>(), f.nh, f.ng), &f.nr, 4);
        if (rc) return rc;
    }
    HIPCHK(c, hipGetLastError());
    return PISURF_OK;
}

// one launch of the run queue for every frame's runs of the level
int batch_level_eval(pisurf_ctx* c, std::vector<BatchFrame*>& fr) {
    std::vector<RunFrame> rf;
    std::vector<int> off(1, 0);
    int T = 128;
    int64_t most = -1;
    for (BatchFrame* f : fr) {
        pisurf_ctx* s = f->c;
        LevelBuf* L = s->levels[f->li];
        const FieldConst fc = field_const(s);
        if (f->nr > 0) {
            RunFrame r;
            r.g = s->g;
            r.ra = s->rec_a.p;
            r.rg = s->rec_g.p;
            r.fc = fc;
            r.lv = L->view();
            r.list = L->child.as<int>();
            r.kv = f->skv;
            r.hlist = s->sd_hlist.as<int>();
            r.rstart = runs_start(s->sd_cwork.as<int>(), f->nh, f->ng);
            r.out = s->sd_vals.p;
            r.ctr = s->ctr();
            r.kind = 0;
            rf.push_back(r);
            off.push_back(off.back() + f->nr);
            if (f->nh > most) {
                most = f->nh;
                T = f->T;
            }
        }
        if (f->nrc > 0) {
            RunFrame r;
            r.g = s->g;
            r.ra = s->rec_a.p;
            r.rg = s->rec_g.p;
            r.fc = fc;
            r.lv = L->view();
            r.list = s->cl_list.as<int>();
            r.kv = f->ckv;
            r.hlist = nullptr;
            r.rstart = runs_start(s->sd_cwork.as<int>(), f->kcent, f->ngc);
            r.out = s->samples.p;
            r.ctr = s->ctr();
            r.kind = 1;
            rf.push_back(r);
            off.push_back(off.back() + f->nrc);
        }
    }
    const int nf = (int)rf.size();
    if (nf == 0) return PISURF_OK;
    const int nitems = off.back();
    const size_t fbytes = nf * sizeof(RunFrame), obytes = (nf + 1) * sizeof(int);
    HIPCHK(c, c->bq_host.ensure(fbytes + obytes + 64));
    std::memcpy(c->bq_host.p, rf.data(), fbytes);
    std::memcpy((char*)c->bq_host.p + fbytes, off.data(), obytes);
    HIPCHK(c, c->bq_frames.ensure(fbytes));
    HIPCHK(c, c->bq_off.ensure(obytes));
    HIPCHK(c, c->bq_items.ensure((size_t)nitems * sizeof(int2)));
    HIPCHK(c, c->bq_head.ensure(64));
    HIPCHK(c, hipMemcpyAsync(c->bq_frames.p, c->bq_host.p, fbytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->bq_off.p, (char*)c->bq_host.p + fbytes, obytes, hipMemcpyHostToDevice, c->stream));
    launch_run_queue(prec_of(c), field_mode(rf[0].fc), T, c->bq_frames.as<RunFrame>(), c->bq_off.as<int>(), nf, nitems,
                     c->bq_items.as<int2>(), c->bq_head.as<int>(), 2048, c->stream);
    HIPCHK(c, hipGetLastError());
    return PISURF_OK;
}

// scatter, curvature, classification and the children count
int batch_level_classify(BatchFrame& f, BatchReads& br) {
    pisurf_ctx* c = f.c;
    LevelBuf* L = c->levels[f.li];
    const NodeLevel v = L->view();
    const int prec = prec_of(c);
    const size_t rsz = prec == PISURF_FP64 ? 8 : 4;
    if (f.nent > 0)
        launch_sample_scatter(L->child.as<int>(), f.skv, c->sd_rep.as<int>(), c->sd_hpos.as<int>(), f.nent,
                              c->sd_vals.p, (int)rsz, c->samples.p, c->stream);
    if (f.split_curv)
        launch_empty_curv(c->g, c->spos.as<float4>(), c->nrm.as<float4>(), f.tp, v, f.group, kCurvOnly, c->samples.p,
                          (int)rsz, c->stream);
    launch_classify(prec, c->g, c->rec_a.p, c->rec_g.p, field_const(c), f.tp, v, c->samples.p, nullptr, c->ctr(),
                    c->stream);
    launch_fill_i32(L->flag.as<int>() + L->n, 1, 0, c->stream);
    exclusive_scan_i32(L->flag.as<int>(), L->scanned.as<int>(), L->n + 1, c->tmp.as<int>(), c->stream);
    HIPCHK(c, hipGetLastError());
    return br.add(c, L->scanned.as<int>() + L->n, &f.n_int, 4);
}

// the next level, or the end of the frame's octree
int batch_level_next(BatchFrame& f) {
    pisurf_ctx* c = f.c;
    LevelBuf* L = c->levels[f.li];
    if (f.n_int == 0) {
        f.active = false;
        c->tree_done = true;
        c->st.n_nodes = 0;
        for (auto* l : c->levels) c->st.n_nodes += l->n;
        return PISURF_OK;
    }
    if ((int64_t)f.n_int * 8 > (int64_t)INT32_MAX / 2) return c->fail(PISURF_E_UNSUPPORTED, "octree level too large");
    LevelBuf* N = level_of(c, f.li + 1);
    N->n = (int64_t)f.n_int * 8;
    HIPCHK(c, N->ensure(N->n));
    const int64_t need_tmp = scan_tmp_ints(N->n + 1) + 64;
    if ((size_t)need_tmp * 4 > c->tmp.cap) {
        HIPCHK(c, c->wait());
        HIPCHK(c, c->tmp.ensure(need_tmp * 4));
    }
    launch_children(L->view(), L->scanned.as<int>(), N->view(), c->stream);
    HIPCHK(c, hipGetLastError());
    f.li++;
    return PISURF_OK;
}

}  // namespace

extern "C" {

int pisurf_default_params(pisurf_params* p) {
    if (!p) return PISURF_E_INVALID;
    std::memset(p, 0, sizeof(*p));
    p->radius = 0.0f;
    p->smooth_factor = 2.0f;
    p->iso_factor = 1.9f;
    p->use_ani = 1;
    p->use_poly6 = 0;
    p->calc_normals = 1;
    p->oversample = 2;
    p->precision = PISURF_FP32;
    p->qef_mode = PISURF_QEF_OFF;
    p->qef_svd_tol = 1e-3f;
    p->depth_max_override = 0;
    p->sample_dedupe = 1;
    return PISURF_OK;
}

int pisurf_create(pisurf_ctx** out, int device) {
    if (!out) return PISURF_E_INVALID;
    *out = nullptr;
    pisurf_ctx* c = new (std::nothrow) pisurf_ctx();
    if (!c) return PISURF_E_OOM;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        delete c;
        return PISURF_E_HIP;
    }
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return PISURF_E_HIP;
    }
    if (hipEventCreateWithFlags(&c->wev, hipEventDisableTiming) != hipSuccess || c->rb.ensure(kReadbackBytes) != hipSuccess ||
        !(c->hr = host_read_acquire())) {
        delete c;
        return PISURF_E_HIP;
    }
    if (device < kNoAllocMaxDevices) {  // the no-allocation array's address on this device, once
        std::lock_guard<std::mutex> lk(g_scratch_m);
        if (!g_scratch[device]) {
            void* p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
                delete c;
                return PISURF_E_HIP;
            }
            g_scratch[device] = p;
        }
        c->scratch_ptr = g_scratch[device];
    }
    pisurf_default_params(&c->prm);
    std::memset(&c->st, 0, sizeof(c->st));
    if (c->counters.ensure(kCtrStripes * sizeof(Counters)) != hipSuccess ||
        hipMemset(c->counters.p, 0, kCtrStripes * sizeof(Counters)) != hipSuccess) {
        delete c;
        return PISURF_E_HIP;
    }
    *out = c;
    return PISURF_OK;
}

int pisurf_destroy(pisurf_ctx* c) {
    if (!c) return PISURF_E_INVALID;
    (void)hipSetDevice(c->device);
    (void)c->wait();
    delete c;
    return PISURF_OK;
}

const char* pisurf_last_error(const pisurf_ctx* c) { return c ? c->err.c_str() : "null context"; }

void* pisurf_stream(pisurf_ctx* c) { return c ? (void*)c->stream : nullptr; }

int pisurf_run_tuning(int32_t enabled, int32_t clear) {
    RunTuner& t = RunTuner::get();
    if (enabled >= 0) t.set_enabled(enabled != 0);
    if (clear) t.clear();
    return (int)t.size();
}

int pisurf_set_params(pisurf_ctx* c, const pisurf_params* p) {
    if (!c || !p) return PISURF_E_INVALID;
    if (!(p->radius > 0.0f) || !std::isfinite(p->radius)) return c->fail(PISURF_E_INVALID, "radius must be > 0");
    if (p->oversample != 2) return c->fail(PISURF_E_UNSUPPORTED, "only oversample = 2 (27 samples) is supported");
    if (p->precision < 0 || p->precision > 2) return c->fail(PISURF_E_INVALID, "bad precision");
    if (p->qef_mode < 0 || p->qef_mode > 2) return c->fail(PISURF_E_INVALID, "bad qef_mode");
    if (p->depth_max_override < 0 || p->depth_max_override > 21) return c->fail(PISURF_E_INVALID, "bad depth override");
    if (p->mesh_order < 0 || p->mesh_order > 1) return c->fail(PISURF_E_INVALID, "bad mesh_order");
    if (p->sample_dedupe < 0 || p->sample_dedupe > 1) return c->fail(PISURF_E_INVALID, "bad sample_dedupe");
    if (p->frames_in_flight < 0 || p->frames_in_flight > kMaxFramesInFlight)
        return c->fail(PISURF_E_INVALID, "frames_in_flight must be in 0 .. 16");
    if (p->level_graph < 0 || p->level_graph > 1) return c->fail(PISURF_E_INVALID, "bad level_graph");
    const bool geo = p->radius != c->prm.radius || p->use_ani != c->prm.use_ani || p->use_poly6 != c->prm.use_poly6 ||
                     p->precision != c->prm.precision;
    c->prm = *p;
    if (geo) {
        c->grid_done = false;
        c->aniso_done = c->normals_done = c->tree_done = c->consts_done = false;
    }
    c->tree_done = false;
    c->consts_done = false;
    return PISURF_OK;
}

int pisurf_upload_particles(pisurf_ctx* c, const float* xyz, int64_t n, int32_t on_device) {
    if (!c) return PISURF_E_INVALID;
    if (n < 0 || (n > 0 && !xyz)) return c->fail(PISURF_E_INVALID, "bad particle buffer");
    if (n >= (int64_t)INT32_MAX / 2) return c->fail(PISURF_E_UNSUPPORTED, "particle count exceeds 2^30");
    API_BEGIN
    (void)hipSetDevice(c->device);
    HIPCHK(c, c->xyz.ensure(std::max<int64_t>(n, 1) * 3 * sizeof(float)));
    if (n > 0)
        HIPCHK(c, hipMemcpyAsync(c->xyz.p, xyz, n * 3 * sizeof(float),
                                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, c->wait());
    c->n = n;
    c->bbox_done = c->grid_done = c->aniso_done = c->normals_done = c->tree_done = c->consts_done = false;
    return PISURF_OK;
    API_END(c)
}

int pisurf_hashgrid_build(pisurf_ctx* c, const float* bbox, float inf, pisurf_grid_info* out) {
    if (!c) return PISURF_E_INVALID;
    API_BEGIN
    (void)hipSetDevice(c->device);
    if (c->n <= 0) return c->fail(PISURF_E_STATE, "no particles uploaded");
    int rc = do_grid(c, bbox, inf);
    if (rc) return rc;
    if (out) *out = c->ginfo;
    return PISURF_OK;
    API_END(c)
}

int pisurf_hashgrid_download(pisurf_ctx* c, int32_t* index_list, int32_t* cs, int32_t* ce) {
    if (!c) return PISURF_E_INVALID;
    if (!c->grid_done) return c->fail(PISURF_E_STATE, "hash grid not built");
    API_BEGIN
    if (index_list)
        HIPCHK(c, hipMemcpyAsync(index_list, c->vals.p, c->n * 4, hipMemcpyDeviceToHost, c->stream));
    if (cs || ce) {
        const int64_t nc = c->ginfo.cell_num;
        HIPCHK(c, c->out1.ensure(nc * 4));
        HIPCHK(c, c->out2.ensure(nc * 4));
        launch_cell_window(c->g, c->out1.as<int>(), c->out2.as<int>(), c->stream);
        HIPCHK(c, hipGetLastError());
        if (cs) HIPCHK(c, hipMemcpyAsync(cs, c->out1.p, nc * 4, hipMemcpyDeviceToHost, c->stream));
        if (ce) HIPCHK(c, hipMemcpyAsync(ce, c->out2.p, nc * 4, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, c->wait());
    return PISURF_OK;
    API_END(c)
}

int pisurf_neighbors(pisurf_ctx* c, const float* q, int64_t nq, int64_t* offsets, int32_t* ids, int64_t cap,
                     int64_t* total) {
    if (!c) return PISURF_E_INVALID;
    if (!c->grid_done) return c->fail(PISURF_E_STATE, "hash grid not built");
    if (nq < 0 || (nq > 0 && !q)) return c->fail(PISURF_E_INVALID, "bad query buffer");
    API_BEGIN
    HIPCHK(c, c->q.ensure(std::max<int64_t>(nq, 1) * 3 * 4));
    HIPCHK(c, c->out1.ensure((nq + 1) * 4));
    HIPCHK(c, c->out2.ensure((nq + 1) * 4));
    HIPCHK(c, c->tmp.ensure(std::max<size_t>(c->tmp.cap, (scan_tmp_ints(nq + 1) + 64) * 4)));
    if (nq > 0) HIPCHK(c, hipMemcpyAsync(c->q.p, q, nq * 12, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->out1.as<int>() + nq, 0, 4, c->stream));
    launch_neighbor_count(c->g, c->q.as<float>(), nq, c->out1.as<int>(), c->stream);
    // the offsets are scanned in 32 bits: refuse totals they cannot hold
    HIPCHK(c, c->small.ensure(256));
    unsigned long long* tot_d = (unsigned long long*)(c->small.as<char>() + 192);
    HIPCHK(c, hipMemsetAsync(tot_d, 0, 8, c->stream));
    launch_sum_counts(c->out1.as<int>(), nq, tot_d, c->stream);
    HIPCHK(c, hipGetLastError());
    unsigned long long tot64 = 0;
    HIPCHK(c, c->read(&tot64, tot_d, 8));
    if (tot64 >= (unsigned long long)INT32_MAX) {
        if (total) *total = (int64_t)tot64;
        return c->fail(PISURF_E_UNSUPPORTED, "neighbour lists total 2^31 ids or more: query fewer points per call");
    }
    exclusive_scan_i32(c->out1.as<int>(), c->out2.as<int>(), nq + 1, c->tmp.as<int>(), c->stream);
    std::vector<int> off32(nq + 1);
    HIPCHK(c, hipMemcpyAsync(off32.data(), c->out2.p, (nq + 1) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, c->wait());
    const int64_t tot = off32[nq];
    if (total) *total = tot;
    if (offsets)
        for (int64_t i = 0; i <= nq; i++) offsets[i] = off32[i];
    if (ids && tot > 0) {
        HIPCHK(c, c->out3.ensure(tot * 4));
        launch_neighbor_write(c->g, c->q.as<float>(), nq, c->out2.as<int>(), c->vals.as<int>(), c->out3.as<int>(),
                              c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(ids, c->out3.p, std::min(tot, cap) * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, c->wait());
    }
    return PISURF_OK;
    API_END(c)
}

int pisurf_neighbors_noalloc(pisurf_ctx* c, const float* q, int64_t nq, int64_t* offsets, int32_t* ids, int64_t cap,
                             int64_t* total) {
    if (!c) return PISURF_E_INVALID;
    if (!c->grid_done) return c->fail(PISURF_E_STATE, "hash grid not built");
    if (!c->scratch_ptr) return c->fail(PISURF_E_UNSUPPORTED, "no working array on this device");
    if (nq < 0 || (nq > 0 && !q) || !offsets || cap < 0 || (cap > 0 && !ids))
        return c->fail(PISURF_E_INVALID, "bad query or output buffer");
    if (nq > kNoAllocQ)
        return c->fail(PISURF_E_UNSUPPORTED, "at most " + std::to_string(kNoAllocQ) + " queries per call");
    API_BEGIN
    (void)hipSetDevice(c->device);
    int* counts = (int*)c->scratch_ptr;  // nq + 1
    int* off32 = counts + (kNoAllocQ + 1);  // nq + 1
    int* scan_tmp = off32 + (kNoAllocQ + 1);
    unsigned long long* tot_d = (unsigned long long*)(scan_tmp + kNoAllocScanInts);
    if (scan_tmp_ints(nq + 1) > kNoAllocScanInts) return c->fail(PISURF_E_UNSUPPORTED, "scan workspace too small");
    // the array keeps the previous call's contents: the sentinel count and the total start at zero
    launch_fill_i32(counts + nq, 1, 0, c->stream);
    launch_fill_i32((int*)tot_d, 2, 0, c->stream);
    launch_neighbor_count(c->g, q, nq, counts, c->stream);
    launch_sum_counts(counts, nq, tot_d, c->stream);
    HIPCHK(c, hipGetLastError());
    unsigned long long tot64 = 0;
    HIPCHK(c, c->read(&tot64, tot_d, 8));
    if (total) *total = (int64_t)tot64;
    if (tot64 >= (unsigned long long)INT32_MAX)
        return c->fail(PISURF_E_UNSUPPORTED, "neighbour lists total 2^31 ids or more: query fewer points per call");
    if ((int64_t)tot64 > cap) return c->fail(PISURF_E_INVALID, "ids capacity below the total (see *total)");
    exclusive_scan_i32(counts, off32, nq + 1, scan_tmp, c->stream);
    launch_widen_i64(off32, nq + 1, offsets, c->stream);
    if (tot64 > 0) launch_neighbor_write(c->g, q, nq, off32, c->vals.as<int>(), ids, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, c->wait());
    return PISURF_OK;
    API_END(c)
}

int pisurf_box_particles(pisurf_ctx* c, const float* box1, const float* box2, int32_t* ids, int64_t cap,
                         int64_t* total) {
    if (!c || !box1 || !box2) return PISURF_E_INVALID;
    if (!c->grid_done) return c->fail(PISURF_E_STATE, "hash grid not built");
    API_BEGIN
    int lo[3], n3[3];
    int64_t ncells = 1;
    for (int a = 0; a < 3; a++) {
        const float b1 = std::max(box1[a], c->ginfo.bounding[2 * a]);
        const float b2 = std::min(box2[a], c->ginfo.bounding[2 * a + 1]);
        const int l = host_cell_coord(b1, c->g.b0[a], c->g.cell, c->g.n[a]);
        const int h = host_cell_coord(b2, c->g.b0[a], c->g.cell, c->g.n[a]);
        lo[a] = l - 1;
        n3[a] = std::max(0, (h + 1) - (l - 1) + 1);
        ncells *= n3[a];
    }
    int geo[6] = {lo[0], lo[1], lo[2], n3[0], n3[1], n3[2]};
    HIPCHK(c, c->q.ensure(64));
    HIPCHK(c, hipMemcpyAsync(c->q.p, geo, sizeof(geo), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, c->out1.ensure((ncells + 1) * 4));
    HIPCHK(c, c->out2.ensure((ncells + 1) * 4));
    HIPCHK(c, c->tmp.ensure(std::max<size_t>(c->tmp.cap, (scan_tmp_ints(ncells + 1) + 64) * 4)));
    HIPCHK(c, hipMemsetAsync(c->out1.as<int>() + ncells, 0, 4, c->stream));
    launch_box_count(c->g, c->q.as<int>(), c->q.as<int>() + 3, ncells, c->out1.as<int>(), c->stream);
    exclusive_scan_i32(c->out1.as<int>(), c->out2.as<int>(), ncells + 1, c->tmp.as<int>(), c->stream);
    int tot = 0;
    HIPCHK(c, hipMemcpyAsync(&tot, c->out2.as<int>() + ncells, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, c->wait());
    if (total) *total = tot;
    if (ids && tot > 0) {
        HIPCHK(c, c->out3.ensure((int64_t)tot * 4));
        launch_box_write(c->g, c->q.as<int>(), c->q.as<int>() + 3, ncells, c->out2.as<int>(), c->vals.as<int>(),
                         c->out3.as<int>(), c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(ids, c->out3.p, std::min<int64_t>(tot, cap) * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, c->wait());
    }
    return PISURF_OK;
    API_END(c)
}

int pisurf_anisotropy(pisurf_ctx* c) {
    if (!c) return PISURF_E_INVALID;
    API_BEGIN
    (void)hipSetDevice(c->device);
    if (c->n <= 0) return c->fail(PISURF_E_STATE, "no particles uploaded");
    return do_aniso(c);
    API_END(c)
}

int pisurf_download_aniso(pisurf_ctx* c, double* xmean, double* G9, double* det) {
    if (!c) return PISURF_E_INVALID;
    if (!c->aniso_done) return c->fail(PISURF_E_STATE, "anisotropy not computed");
    API_BEGIN
    const int64_t n = c->n;
    const bool f64 = c->prec_rec == PISURF_FP64;
    const size_t rb = f64 ? sizeof(double4) : sizeof(float4);
    std::vector<char> a(n * rb), g(2 * n * rb);
    std::vector<float4> lo(c->prec_rec == PISURF_FP32_KAHAN ? n : 0);
    std::vector<int> sidx(n);
    HIPCHK(c, hipMemcpyAsync(a.data(), c->rec_a.p, n * rb, hipMemcpyDeviceToHost, c->stream));
    if (!lo.empty())
        HIPCHK(c, hipMemcpyAsync(lo.data(), c->rec_lo.p, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(g.data(), c->rec_g.p, 2 * n * rb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(sidx.data(), c->vals.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, c->wait());
    auto rd = [&](const std::vector<char>& b, int64_t k) {
        return f64 ? ((const double*)b.data())[k] : double(((const float*)b.data())[k]);
    };
    for (int64_t p = 0; p < n; p++) {
        const int64_t i = sidx[p];
        if (xmean) {
            for (int k = 0; k < 3; k++) xmean[3 * i + k] = rd(a, 4 * p + k);
            if (!lo.empty()) {
                xmean[3 * i] += lo[p].x;
                xmean[3 * i + 1] += lo[p].y;
                xmean[3 * i + 2] += lo[p].z;
            }
        }
        const double G6[6] = {rd(g, 8 * p), rd(g, 8 * p + 1), rd(g, 8 * p + 2), rd(g, 8 * p + 3), rd(g, 8 * p + 4),
                              rd(g, 8 * p + 5)};
        if (G9) {
            // column-major 3x3 from the symmetric (00, 01, 02, 11, 12, 22)
            const double m[9] = {G6[0], G6[1], G6[2], G6[1], G6[3], G6[4], G6[2], G6[4], G6[5]};
            for (int k = 0; k < 9; k++) G9[9 * i + k] = m[k];
        }
        if (det) det[i] = rd(g, 8 * p + 6);
    }
    return PISURF_OK;
    API_END(c)
}

int pisurf_iso_constants(pisurf_ctx* c, double* ms, double* iso) {
    if (!c) return PISURF_E_INVALID;
    if (!(c->prm.radius > 0.0f)) return c->fail(PISURF_E_STATE, "parameters not set");
    compute_constants(c);
    if (ms) *ms = c->max_scalar;
    if (iso) *iso = c->iso;
    return PISURF_OK;
}

int pisurf_particle_normals(pisurf_ctx* c, double* normals) {
    if (!c) return PISURF_E_INVALID;
    API_BEGIN
    (void)hipSetDevice(c->device);
    if (c->n <= 0) return c->fail(PISURF_E_STATE, "no particles uploaded");
    int rc = do_normals(c);
    if (rc) return rc;
    if (normals) {
        std::vector<float4> h(c->n);
        std::vector<int> sidx(c->n);
        HIPCHK(c, hipMemcpyAsync(h.data(), c->nrm.p, c->n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(sidx.data(), c->vals.p, c->n * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, c->wait());
        for (int64_t p = 0; p < c->n; p++) {
            normals[3 * sidx[p]] = h[p].x;
            normals[3 * sidx[p] + 1] = h[p].y;
            normals[3 * sidx[p] + 2] = h[p].z;
        }
    }
    return PISURF_OK;
    API_END(c)
}

int pisurf_eval_points(pisurf_ctx* c, const float* q, int64_t n, double* f, double* grad) {
    if (!c) return PISURF_E_INVALID;
    if (n < 0 || (n > 0 && (!q || !f))) return c->fail(PISURF_E_INVALID, "bad buffers");
    API_BEGIN
    (void)hipSetDevice(c->device);
    if (c->n <= 0) return c->fail(PISURF_E_STATE, "no particles uploaded");
    int rc = need_aniso(c);
    if (rc) return rc;
    if (!c->consts_done) compute_constants(c);
    if (n == 0) return PISURF_OK;
    HIPCHK(c, c->q.ensure(n * 12));
    HIPCHK(c, c->out1.ensure(n * 8));
    if (grad) HIPCHK(c, c->out2.ensure(n * 24));
    HIPCHK(c, hipMemcpyAsync(c->q.p, q, n * 12, hipMemcpyHostToDevice, c->stream));
    launch_eval_points(prec_of(c), c->g, c->rec_a.p, c->rec_g.p, field_const(c), c->q.as<float>(), n,
                       c->out1.as<double>(), grad ? c->out2.as<double>() : nullptr, c->ctr(), c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(f, c->out1.p, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (grad) HIPCHK(c, hipMemcpyAsync(grad, c->out2.p, n * 24, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, c->wait());
    return PISURF_OK;
    API_END(c)
}

int pisurf_uniform_grid(pisurf_ctx* c, int32_t dims[3], float box[6], float* spacing, float* scalars) {
    if (!c) return PISURF_E_INVALID;
    API_BEGIN
    (void)hipSetDevice(c->device);
    if (c->n <= 0) return c->fail(PISURF_E_STATE, "no particles uploaded");
    int rc = ensure_bbox(c);
    if (rc) return rc;
    // resizeRootBoxConstR of UniformGrid (src/marching.cpp:41-56) on the tight particle box
    const float r = c->prm.radius;
    UniformBox ub;
    int64_t total = 1;
    for (int a = 0; a < 3; a++) {
        const float center = (c->bbox[2 * a] + c->bbox[2 * a + 1]) / 2;
        const float length = c->bbox[2 * a + 1] - c->bbox[2 * a];
        const float len2 = std::ceil(length / r) * r + r * 4;
        ub.lo[a] = center - len2 / 2;
        ub.hi[a] = center + len2 / 2;
        ub.steps[a] = int(len2 / r);
        ub.dims[a] = ub.steps[a] + 1;
        if (ub.steps[a] < 1) return c->fail(PISURF_E_INVALID, "degenerate sampling box");
        total *= ub.dims[a];
        if (total >= ((int64_t)1 << 31)) return c->fail(PISURF_E_UNSUPPORTED, "uniform grid exceeds 2^31 points");
    }
    if (dims)
        for (int a = 0; a < 3; a++) dims[a] = ub.dims[a];
    if (box)
        for (int a = 0; a < 3; a++) {
            box[2 * a] = ub.lo[a];
            box[2 * a + 1] = ub.hi[a];
        }
    if (spacing) *spacing = r;
    if (!scalars) return PISURF_OK;
    rc = need_aniso(c);
    if (rc) return rc;
    if (!c->consts_done) compute_constants(c);
    HIPCHK(c, c->out1.ensure(total * sizeof(float)));
    launch_uniform_grid(prec_of(c), c->g, c->rec_a.p, c->rec_g.p, field_const(c), ub, total, c->out1.as<float>(),
                        c->ctr(), c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(scalars, c->out1.p, total * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, c->wait());
    return PISURF_OK;
    API_END(c)
}

int pisurf_sweep_level(pisurf_ctx* c, const float* centers, const float* halfs, const int32_t* depths, int64_t n,
                       uint8_t* type, double* f, float* curv, int32_t* signchange, double* samples, double* grads) {
    if (!c) return PISURF_E_INVALID;
    if (n < 0 || (n > 0 && (!centers || !halfs || !depths))) return c->fail(PISURF_E_INVALID, "bad node buffers");
    API_BEGIN
    (void)hipSetDevice(c->device);
    if (c->n <= 0) return c->fail(PISURF_E_STATE, "no particles uploaded");
    int rc = frame_ready(c);
    if (rc) return rc;
    resize_root(c);
    if (n == 0) return PISURF_OK;
    LevelBuf L;
    L.n = n;
    HIPCHK(c, L.ensure(n));
    std::vector<float4> c4(n);
    std::vector<int8_t> d8(n);
    std::vector<uint64_t> ids(n, 0);
    for (int64_t i = 0; i < n; i++) {
        c4[i] = make_float4(centers[3 * i], centers[3 * i + 1], centers[3 * i + 2], halfs[i]);
        d8[i] = (int8_t)depths[i];
    }
    HIPCHK(c, hipMemcpyAsync(L.c.p, c4.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(L.depth.p, d8.data(), n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(L.id.p, ids.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, c->out3.ensure(n * 4));
    rc = sweep(c, &L, c->out3.as<int32_t>(), -1);
    if (rc) return rc;
    if (samples || grads) {
        HIPCHK(c, c->out1.ensure(n * 27 * 4 * 8));
        HIPCHK(c, c->out2.ensure(n * 27 * 3 * 8));
        launch_sample_grads(prec_of(c), tree_params(c), L.view(), c->samples.p, c->out1.as<double>(),
                            c->out2.as<double>(), c->stream);
        HIPCHK(c, hipGetLastError());
        if (samples) HIPCHK(c, hipMemcpyAsync(samples, c->out1.p, n * 27 * 4 * 8, hipMemcpyDeviceToHost, c->stream));
        if (grads) HIPCHK(c, hipMemcpyAsync(grads, c->out2.p, n * 27 * 3 * 8, hipMemcpyDeviceToHost, c->stream));
    }
    std::vector<double4> nd(n);
    if (type) HIPCHK(c, hipMemcpyAsync(type, L.type.p, n, hipMemcpyDeviceToHost, c->stream));
    if (curv) HIPCHK(c, hipMemcpyAsync(curv, L.curv.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    if (signchange) HIPCHK(c, hipMemcpyAsync(signchange, c->out3.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(nd.data(), L.node.p, n * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, c->wait());
    c->harvest_events();
    if (f)
        for (int64_t i = 0; i < n; i++) f[i] = nd[i].w;
    return PISURF_OK;
    API_END(c)
}

int pisurf_build_tree(pisurf_ctx* c, pisurf_tree_info* out) {
    if (!c) return PISURF_E_INVALID;
    API_BEGIN
    (void)hipSetDevice(c->device);
    if (c->n <= 0) return c->fail(PISURF_E_STATE, "no particles uploaded");
    int rc = do_tree(c);
    if (rc) return rc;
    if (out) {
        out->depth_max = c->dmax;
        out->depth_min = c->dmin;
        for (int a = 0; a < 3; a++) out->root_center[a] = c->root_c[a];
        out->root_half = c->root_half;
        int64_t tot = 0;
        int nl = 0;
        for (auto* l : c->levels) {
            if (l->n == 0) break;
            tot += l->n;
            nl++;
        }
        out->n_nodes = tot;
        out->n_levels = nl;
        out->n_leaves = 0;
    }
    return PISURF_OK;
    API_END(c)
}

int pisurf_tree_download(pisurf_ctx* c, float* center3, float* half, int32_t* depth, uint8_t* type, uint64_t* id,
                         double* node4, int32_t* child) {
    if (!c) return PISURF_E_INVALID;
    if (!c->tree_done) return c->fail(PISURF_E_STATE, "octree not built");
    API_BEGIN
    HostTree t;
    int rc = download_tree(c, t, true);
    if (rc) return rc;
    const size_t n = t.type.size();
    for (size_t i = 0; i < n; i++) {
        if (center3) {
            center3[3 * i] = t.cen[i].x;
            center3[3 * i + 1] = t.cen[i].y;
            center3[3 * i + 2] = t.cen[i].z;
        }
        if (half) half[i] = t.cen[i].w;
        if (depth) depth[i] = t.depth[i];
        if (type) type[i] = t.type[i];
        if (id) id[i] = t.id[i];
        if (node4) {
            node4[4 * i] = t.node[i].x;
            node4[4 * i + 1] = t.node[i].y;
            node4[4 * i + 2] = t.node[i].z;
            node4[4 * i + 3] = t.node[i].w;
        }
        if (child) child[i] = t.child[i];
    }
    return PISURF_OK;
    API_END(c)
}

int pisurf_qef_solve(pisurf_ctx* c, const float* centers, const float* halfs, const double* samples,
                     const double* grads, int64_t n, int32_t mode, double* node4) {
    if (!c) return PISURF_E_INVALID;
    if (mode != 1 && mode != 2) return c->fail(PISURF_E_INVALID, "mode must be 1 or 2");
    if (n < 0 || (n > 0 && (!centers || !halfs || !samples || !grads || !node4)))
        return c->fail(PISURF_E_INVALID, "bad buffers");
    API_BEGIN
    (void)hipSetDevice(c->device);
    if (c->n <= 0) return c->fail(PISURF_E_STATE, "no particles uploaded");
    int rc = need_aniso(c);
    if (rc) return rc;
    if (!c->consts_done) compute_constants(c);
    if (n == 0) return PISURF_OK;
    std::vector<float4> c4(n);
    for (int64_t i = 0; i < n; i++) c4[i] = make_float4(centers[3 * i], centers[3 * i + 1], centers[3 * i + 2], halfs[i]);
    HIPCHK(c, c->q.ensure(n * sizeof(float4)));
    HIPCHK(c, c->out1.ensure(n * 27 * 4 * 8));
    HIPCHK(c, c->out2.ensure(n * 27 * 3 * 8));
    HIPCHK(c, c->out3.ensure(n * sizeof(double4)));
    HIPCHK(c, hipMemcpyAsync(c->q.p, c4.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->out1.p, samples, n * 27 * 4 * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->out2.p, grads, n * 27 * 3 * 8, hipMemcpyHostToDevice, c->stream));
    TreeParams tp = tree_params(c);
    launch_qef(prec_of(c), c->g, c->rec_a.p, c->rec_g.p, field_const(c), tp, c->q.as<float4>(), c->out1.as<double>(),
               c->out2.as<double>(), n, mode, c->out3.as<double4>(), c->ctr(), c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(node4, c->out3.p, n * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, c->wait());
    return PISURF_OK;
    API_END(c)
}

int pisurf_bisect_edges(pisurf_ctx* c, const double* ends, int64_t n, double* out_xyz, int32_t* n_evals) {
    if (!c) return PISURF_E_INVALID;
    if (n < 0 || (n > 0 && (!ends || !out_xyz))) return c->fail(PISURF_E_INVALID, "bad buffers");
    for (int64_t i = 0; i < 8 * n; i++)
        if (!std::isfinite(ends[i])) return c->fail(PISURF_E_INVALID, "edge endpoints must be finite");
    API_BEGIN
    (void)hipSetDevice(c->device);
    if (c->n <= 0) return c->fail(PISURF_E_STATE, "no particles uploaded");
    int rc = need_aniso(c);
    if (rc) return rc;
    if (!c->consts_done) compute_constants(c);
    if (n == 0) return PISURF_OK;
    HIPCHK(c, c->q.ensure(n * 64));
    HIPCHK(c, c->out1.ensure(n * 24));
    HIPCHK(c, c->out2.ensure(n * 4));
    HIPCHK(c, hipMemcpyAsync(c->q.p, ends, n * 64, hipMemcpyHostToDevice, c->stream));
    launch_bisect(prec_of(c), c->g, c->rec_a.p, c->rec_g.p, field_const(c), c->prm.radius / 2, c->q.as<double>(), n,
                  c->out1.as<double>(), c->out2.as<int>(), c->ctr(), c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(out_xyz, c->out1.p, n * 24, hipMemcpyDeviceToHost, c->stream));
    if (n_evals) HIPCHK(c, hipMemcpyAsync(n_evals, c->out2.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, c->wait());
    return PISURF_OK;
    API_END(c)
}

int pisurf_extract(pisurf_ctx* c, int64_t* nv, int64_t* nt) {
    if (!c) return PISURF_E_INVALID;
    API_BEGIN
    (void)hipSetDevice(c->device);
    int rc = do_extract(c);
    if (rc) return rc;
    if (nv) *nv = (int64_t)c->mesh_v.size() / 3;
    if (nt) *nt = (int64_t)c->mesh_t.size() / 3;
    return PISURF_OK;
    API_END(c)
}

int pisurf_dual_cells(pisurf_ctx* c, int32_t* nodes8, uint8_t* sign_index, int64_t cap, int64_t* total) {
    if (!c) return PISURF_E_INVALID;
    if (!c->tree_done) return c->fail(PISURF_E_STATE, "octree not built");
    if (cap < 0 || (cap > 0 && (!nodes8 || !sign_index))) return c->fail(PISURF_E_INVALID, "bad output buffers");
    API_BEGIN
    HostTree t;
    int rc = download_tree(c, t, false);
    if (rc) return rc;
    Extractor ex(t);
    if (!t.type.empty()) ex.node(0);
    const int64_t nd = (int64_t)ex.duals.size();
    if (total) *total = nd;
    for (int64_t i = 0; i < nd && i < cap; i++) {
        for (int k = 0; k < 8; k++) nodes8[8 * i + k] = ex.duals[i].n[k];
        sign_index[i] = ex.duals[i].sidx;
    }
    return PISURF_OK;
    API_END(c)
}

int pisurf_mesh_download(pisurf_ctx* c, double* verts, int32_t* tris) {
    if (!c) return PISURF_E_INVALID;
    if (verts) std::memcpy(verts, c->mesh_v.data(), c->mesh_v.size() * sizeof(double));
    if (tris) std::memcpy(tris, c->mesh_t.data(), c->mesh_t.size() * sizeof(int32_t));
    return PISURF_OK;
}

int pisurf_run_frame(pisurf_ctx* c, pisurf_tree_info* tree) {
    if (!c) return PISURF_E_INVALID;
    API_BEGIN
    (void)hipSetDevice(c->device);
    if (c->n <= 0) return c->fail(PISURF_E_STATE, "no particles uploaded");
    const double t0 = now_ms();
    c->bbox_done = c->grid_done = c->aniso_done = c->normals_done = c->tree_done = c->consts_done = false;
    c->top_pending = false;
    int rc = do_grid(c, nullptr, 4.0f);
    if (!rc) {
        // The anisotropy passes go to the GPU first; the levels above D_min are issued on their own
        // stream while those run (issuing their 40-odd launches takes the host about 0.4 ms), and
        // only then does the host wait for the anisotropy's count. The constants depend on the
        // parameters only; those levels write the iso value into their empty nodes.
        compute_constants(c);
        const double ta = now_ms();
        rc = aniso_launch(c);
        // (lanes of a multi-frame call overlap whole frames instead)
        if (!rc && !c->lane) rc = tree_start(c);
        if (!rc) {
            unsigned long long nfp = 0;
            HIPCHK(c, c->read(&nfp, aniso_nfp_dev(c), 8));
            aniso_finish(c, nfp);
            c->st.ms_aniso += now_ms() - ta;
        }
    }
    if (!rc) rc = frame_ready(c);
    if (!rc) rc = pisurf_build_tree(c, tree);
    if (!rc) rc = do_extract(c);
    if (rc) return rc;
    c->st.ms_total += now_ms() - t0;
    return PISURF_OK;
    API_END(c)
}

// pisurf_run_frames with K lanes: frame k is meshed on lane k mod K (lane 0 is c, on the calling
// thread; the others are contexts of their own, each on a thread of its own for this call), and the
// meshes are handed to the callback one at a time in frame order.
static int run_frames_lanes(pisurf_ctx* c, int K, int32_t nframes, const float* const* xyz, const int64_t* n,
                            const float* radius, int32_t on_device, pisurf_mesh_cb on_mesh, void* user,
                            pisurf_stats* per_frame) {
    API_BEGIN
    (void)hipSetDevice(c->device);
    while ((int)c->lanes.size() < K - 1) {
        pisurf_ctx* l = nullptr;
        const int rc = pisurf_create(&l, c->device);
        if (rc) return c->fail(rc, "lane context creation failed");
        c->lanes.push_back(l);
    }
    std::vector<pisurf_ctx*> lane(K);
    lane[0] = c;
    for (int i = 1; i < K; i++) lane[i] = c->lanes[i - 1];
    struct LaneMode {  // the contexts mesh as lanes for this call only
        std::vector<pisurf_ctx*>& l;
        explicit LaneMode(std::vector<pisurf_ctx*>& v) : l(v) {
            for (auto* x : l) x->lane = true;
        }
        ~LaneMode() {
            for (auto* x : l) x->lane = false;
        }
    } lane_mode(lane);
    const pisurf_params base = c->prm;
    // one frame on lane L: its particles copied on the lane's stream (the frame's kernels follow
    // in stream order, and the frame's own read-backs have completed the copy before the staging
    // buffer is reused), then the whole frame
    auto mesh_frame = [&](pisurf_ctx* L, int32_t k, pisurf_tree_info* ti) -> int {
        pisurf_params p = base;
        p.radius = radius[k];
        int rc = pisurf_set_params(L, &p);
        if (!rc) rc = pisurf_reset_stats(L);
        if (rc) return rc;
        L->n = n[k];
        L->bbox_done = L->grid_done = L->aniso_done = L->normals_done = L->tree_done = L->consts_done = false;
        if (n[k] > 0) {
            const size_t b = (size_t)n[k] * 3 * sizeof(float);
            HIPCHK(L, L->xyz.ensure(b));
            if (on_device) {
                HIPCHK(L, hipMemcpyAsync(L->xyz.p, xyz[k], b, hipMemcpyDeviceToDevice, L->stream));
            } else {
                HIPCHK(L, L->pin.ensure(b));
                std::memcpy(L->pin.p, xyz[k], b);
                HIPCHK(L, hipMemcpyAsync(L->xyz.p, L->pin.p, b, hipMemcpyHostToDevice, L->stream));
            }
            rc = pisurf_run_frame(L, ti);
        } else {
            L->mesh_v.clear();
            L->mesh_t.clear();
        }
        if (!rc && per_frame) rc = sync_counters(L);
        return rc;
    };
    std::mutex m;
    std::condition_variable cv;
    int32_t delivered = 0;  // frames handed to the callback
    bool stop = false;      // a frame failed: no later frame is delivered
    int fail_rc = PISURF_OK;
    std::string fail_msg;
    auto lane_frames = [&](int li) {
        pisurf_ctx* L = lane[li];
        if (li > 0) (void)hipSetDevice(c->device);
        for (int32_t k = li; k < nframes; k += K) {
            {
                std::lock_guard<std::mutex> lk(m);
                if (stop) return;
            }
            pisurf_tree_info ti;
            std::memset(&ti, 0, sizeof(ti));
            int rc;
            try {
                rc = mesh_frame(L, k, &ti);
            } catch (const std::bad_alloc&) {
                rc = L->fail(PISURF_E_OOM, "out of host memory");
            } catch (...) {
                rc = L->fail(PISURF_E_STATE, "unexpected exception");
            }
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return stop || delivered == k; });
            if (stop) return;
            if (rc) {
                fail_rc = rc;
                fail_msg = L->err + " (frame " + std::to_string(k) + ")";
                stop = true;
                cv.notify_all();
                return;
            }
            if (per_frame) per_frame[k] = L->st;
            if (on_mesh)
                on_mesh(user, k, &ti, L->mesh_v.data(), (int64_t)L->mesh_v.size() / 3, L->mesh_t.data(),
                        (int64_t)L->mesh_t.size() / 3);
            delivered = k + 1;
            cv.notify_all();
        }
    };
    // a lane that ends by an exception stops the others, which may be waiting for its frame
    auto lane_main = [&](int li) {
        try {
            lane_frames(li);
        } catch (...) {
            std::lock_guard<std::mutex> lk(m);
            if (!fail_rc) {
                fail_rc = PISURF_E_STATE;
                fail_msg = "unexpected exception in lane " + std::to_string(li);
            }
            stop = true;
            cv.notify_all();
        }
    };
    struct Joiner {  // the lane threads are joined on every exit path
        std::vector<std::thread> t;
        ~Joiner() {
            for (auto& x : t)
                if (x.joinable()) x.join();
        }
    } threads;
    for (int i = 1; i < K; i++) threads.t.emplace_back(lane_main, i);
    lane_main(0);
    for (auto& x : threads.t) x.join();
    if (fail_rc) {
        c->err = fail_msg;
        return fail_rc;
    }
    return PISURF_OK;
    API_END(c)
}

int pisurf_run_frames(pisurf_ctx* c, int32_t nframes, const float* const* xyz, const int64_t* n,
                      const float* radius, int32_t on_device, pisurf_mesh_cb on_mesh, void* user,
                      pisurf_stats* per_frame) {
    if (!c) return PISURF_E_INVALID;
    if (nframes < 0 || (nframes > 0 && (!xyz || !n || !radius))) return c->fail(PISURF_E_INVALID, "bad frame arrays");
    int64_t nmax = 1;
    for (int32_t k = 0; k < nframes; k++) {
        const std::string fk = " (frame " + std::to_string(k) + ")";
        if (n[k] < 0 || (n[k] > 0 && !xyz[k])) return c->fail(PISURF_E_INVALID, "bad particle buffer" + fk);
        if (n[k] >= (int64_t)INT32_MAX / 2) return c->fail(PISURF_E_UNSUPPORTED, "particle count exceeds 2^30" + fk);
        if (!(radius[k] > 0.0f) || !std::isfinite(radius[k])) return c->fail(PISURF_E_INVALID, "radius must be > 0" + fk);
        nmax = std::max(nmax, n[k]);
    }
    const int lanes_req = c->prm.frames_in_flight == 0 ? kDefaultFramesInFlight : c->prm.frames_in_flight;
    if (std::min<int>(lanes_req, nframes) > 1)
        return run_frames_lanes(c, std::min<int>(lanes_req, nframes), nframes, xyz, n, radius, on_device, on_mesh, user,
                                per_frame);
    API_BEGIN
    (void)hipSetDevice(c->device);
    if (!c->cstream) HIPCHK(c, hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking));
    if (!c->cev) HIPCHK(c, hipEventCreateWithFlags(&c->cev, hipEventDisableTiming));
    const size_t bytes_max = (size_t)nmax * 3 * sizeof(float);
    HIPCHK(c, c->wait());
    HIPCHK(c, c->xyz.ensure(bytes_max));
    HIPCHK(c, c->xyz_next.ensure(bytes_max));
    if (!on_device) HIPCHK(c, c->pin.ensure(bytes_max));
    // frame k -> xyz_next on the copy stream, waited for (runs on a helper thread during the previous frame)
    auto stage = [c, xyz, n, on_device](int32_t k) -> hipError_t {
        const size_t b = (size_t)n[k] * 3 * sizeof(float);
        if (b == 0) return hipSuccess;
        hipError_t e = hipSetDevice(c->device);
        if (e != hipSuccess) return e;
        if (on_device) {
            e = hipMemcpyAsync(c->xyz_next.p, xyz[k], b, hipMemcpyDeviceToDevice, c->cstream);
        } else {
            std::memcpy(c->pin.p, xyz[k], b);
            e = hipMemcpyAsync(c->xyz_next.p, c->pin.p, b, hipMemcpyHostToDevice, c->cstream);
        }
        // the context's event, not the thread's: the staging threads create no HIP objects
        return e != hipSuccess ? e : stream_wait_on(c->cstream, c->cev);
    };
    struct Joiner {  // the helper is joined on every exit path, exceptions included
        std::thread t;
        ~Joiner() {
            if (t.joinable()) t.join();
        }
    };
    const pisurf_params base = c->prm;
    if (nframes > 0) {
        const hipError_t e = stage(0);
        if (e != hipSuccess) return c->hipfail(e, "frame upload (frame 0)");
    }
    for (int32_t k = 0; k < nframes; k++) {
        std::swap(c->xyz.p, c->xyz_next.p);
        std::swap(c->xyz.cap, c->xyz_next.cap);
        hipError_t staged = hipSuccess;
        int rc = PISURF_OK;
        pisurf_tree_info ti;
        std::memset(&ti, 0, sizeof(ti));
        {
            Joiner helper;
            if (k + 1 < nframes) helper.t = std::thread([&stage, &staged, k] { staged = stage(k + 1); });
            pisurf_params p = base;
            p.radius = radius[k];
            rc = pisurf_set_params(c, &p);
            if (!rc) rc = pisurf_reset_stats(c);
            c->n = n[k];
            c->bbox_done = c->grid_done = c->aniso_done = c->normals_done = c->tree_done = c->consts_done = false;
            if (!rc && n[k] > 0) {
                rc = pisurf_run_frame(c, &ti);
            } else if (!rc) {
                c->mesh_v.clear();
                c->mesh_t.clear();
            }
            if (!rc && per_frame) rc = sync_counters(c);
        }
        if (rc) {
            c->err += " (frame " + std::to_string(k) + ")";
            return rc;
        }
        if (staged != hipSuccess) return c->hipfail(staged, "frame upload");
        if (per_frame) per_frame[k] = c->st;
        if (on_mesh)
            on_mesh(user, k, &ti, c->mesh_v.data(), (int64_t)c->mesh_v.size() / 3, c->mesh_t.data(),
                    (int64_t)c->mesh_t.size() / 3);
    }
    return PISURF_OK;
    API_END(c)
}

int pisurf_run_batch(pisurf_ctx* c, int32_t nframes, const float* const* xyz, const int64_t* n, const float* radius,
                     int32_t on_device, pisurf_mesh_cb on_mesh, void* user, pisurf_stats* per_frame) {
    if (!c) return PISURF_E_INVALID;
    if (nframes < 0 || (nframes > 0 && (!xyz || !n || !radius))) return c->fail(PISURF_E_INVALID, "bad frame arrays");
    for (int32_t k = 0; k < nframes; k++) {
        const std::string fk = " (frame " + std::to_string(k) + ")";
        if (n[k] < 0 || (n[k] > 0 && !xyz[k])) return c->fail(PISURF_E_INVALID, "bad particle buffer" + fk);
        if (n[k] >= (int64_t)INT32_MAX / 2) return c->fail(PISURF_E_UNSUPPORTED, "particle count exceeds 2^30" + fk);
        if (!(radius[k] > 0.0f) || !std::isfinite(radius[k])) return c->fail(PISURF_E_INVALID, "radius must be > 0" + fk);
    }
    // Switched off: the first device run of the lockstep batch stopped making progress and the
    // cause has not been found in the code; pisurf_run_frames meshes the same batch.
    if (!kBatchEnabled)
        return c->fail(PISURF_E_UNSUPPORTED, "pisurf_run_batch is disabled (use pisurf_run_frames)");
    API_BEGIN
    (void)hipSetDevice(c->device);
    HIPCHK(c, c->wait());
    while ((int32_t)c->sub.size() < nframes) {
        pisurf_ctx* f = make_frame_ctx(c);
        if (!f) return c->fail(PISURF_E_HIP, "frame context creation failed");
        c->sub.push_back(f);
    }
    auto tag = [&](pisurf_ctx* f, int rc) {
        for (int32_t k = 0; k < nframes; k++)
            if (c->sub[k] == f) {
                c->err = f->err + " (frame " + std::to_string(k) + ")";
                break;
            }
        return rc;
    };
    // every frame's particles on the device, its parameters and statistics reset
    std::vector<BatchFrame> frames;
    for (int32_t k = 0; k < nframes; k++) {
        pisurf_ctx* f = c->sub[k];
        pisurf_params p = c->prm;
        p.radius = radius[k];
        int rc = pisurf_set_params(f, &p);
        if (rc) return tag(f, rc);
        std::memset(&f->st, 0, sizeof(f->st));
        f->host_sample_entries = 0;
        f->n_fp64_eig = 0;
        HIPCHK(c, hipMemsetAsync(f->counters.p, 0, kCtrStripes * sizeof(Counters), c->stream));
        f->n = n[k];
        f->bbox_done = f->grid_done = f->aniso_done = f->normals_done = f->tree_done = f->consts_done = false;
        for (auto* l : f->levels) l->n = 0;
        f->mesh_v.clear();
        f->mesh_t.clear();
        if (n[k] == 0) continue;
        HIPCHK(c, f->xyz.ensure((size_t)n[k] * 12));
        HIPCHK(c, hipMemcpyAsync(f->xyz.p, xyz[k], (size_t)n[k] * 12,
                                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
        BatchFrame b;
        b.c = f;
        b.k = k;
        b.active = true;
        frames.push_back(b);
    }
    // larger frames first: their runs enter the queue first
    std::stable_sort(frames.begin(), frames.end(), [&](const BatchFrame& a, const BatchFrame& b) { return n[a.k] > n[b.k]; });
    BatchReads br(c);
    int rc = PISURF_OK;
    // bounding boxes, hash grids, anisotropy, constants, normals
    for (auto& f : frames) {
        if ((rc = bbox_launch(f.c))) return tag(f.c, rc);
        if ((rc = br.add(f.c, bbox_dev(f.c), f.c->bbox, 6 * sizeof(float)))) return rc;
        if ((rc = br.add(f.c, bbox_nan_dev(f.c), &f.nans, sizeof(int)))) return rc;
    }
    if ((rc = br.flush())) return rc;
    for (auto& f : frames) {
        if ((rc = bbox_finish(f.c, f.nans))) return tag(f.c, rc);
        if ((rc = grid_launch(f.c, nullptr, 4.0f))) return tag(f.c, rc);
        if ((rc = br.add(f.c, grid_ninv_dev(f.c), &f.ninv, sizeof(int)))) return rc;
    }
    if ((rc = br.flush())) return rc;
    for (auto& f : frames) {
        grid_finish(f.c, f.ninv);
        if ((rc = aniso_launch(f.c))) return tag(f.c, rc);
        if ((rc = br.add(f.c, aniso_nfp_dev(f.c), &f.nfp, 8))) return rc;
        compute_constants(f.c);
        if (f.c->prm.calc_normals) {
            if ((rc = normals_launch(f.c))) return tag(f.c, rc);
        } else {
            HIPCHK(c, f.c->nrm.ensure(f.c->n * sizeof(float4)));
            HIPCHK(c, hipMemsetAsync(f.c->nrm.p, 0, f.c->n * sizeof(float4), c->stream));
        }
        // the root of the octree (resizeRootBoxConstR)
        resize_root(f.c);
        LevelBuf* L = level_of(f.c, 0);
        L->n = 1;
        HIPCHK(c, L->ensure(1));
        const float4 rc4 = make_float4(f.c->root_c[0], f.c->root_c[1], f.c->root_c[2], f.c->root_half);
        const int8_t d0 = 0;
        const uint64_t id0 = 0;
        const int4 ip0 = make_int4(0, 0, 0, 0);
        HIPCHK(c, hipMemcpyAsync(L->ip.p, &ip0, sizeof(int4), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(L->c.p, &rc4, sizeof(float4), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(L->depth.p, &d0, 1, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(L->id.p, &id0, 8, hipMemcpyHostToDevice, c->stream));
    }
    if ((rc = br.flush())) return rc;
    for (auto& f : frames) {
        const bool nrm = f.c->normals_done;
        aniso_finish(f.c, f.nfp);
        f.c->normals_done = nrm;
    }
    // the octree levels of all frames in lockstep
    std::vector<BatchFrame*> act;
    for (int guard = 0;; guard++) {
        act.clear();
        for (auto& f : frames)
            if (f.active) act.push_back(&f);
        if (act.empty()) break;
        if (guard > 64) return c->fail(PISURF_E_STATE, "octree deeper than 64 levels");
        for (BatchFrame* f : act)
            if ((rc = batch_level_begin(*f, br))) return tag(f->c, rc);
        if ((rc = br.flush())) return rc;
        for (BatchFrame* f : act)
            if ((rc = batch_level_keys(*f, br))) return tag(f->c, rc);
        if ((rc = br.flush())) return rc;
        for (BatchFrame* f : act)
            if ((rc = batch_level_groups(*f, br))) return tag(f->c, rc);
        if ((rc = br.flush())) return rc;
        for (BatchFrame* f : act)
            if ((rc = batch_level_runs(*f, br))) return tag(f->c, rc);
        if ((rc = br.flush())) return rc;
        if ((rc = batch_level_eval(c, act))) return rc;
        for (BatchFrame* f : act)
            if ((rc = batch_level_classify(*f, br))) return tag(f->c, rc);
        if ((rc = br.flush())) return rc;
        for (BatchFrame* f : act)
            if ((rc = batch_level_next(*f))) return tag(f->c, rc);
    }
    HIPCHK(c, c->wait());
    // extraction and meshes, in the caller's order
    for (int32_t k = 0; k < nframes; k++) {
        pisurf_ctx* f = c->sub[k];
        pisurf_tree_info ti;
        std::memset(&ti, 0, sizeof(ti));
        if (n[k] > 0) {
            if ((rc = do_extract(f))) return tag(f, rc);
            ti.depth_max = f->dmax;
            ti.depth_min = f->dmin;
            for (int a = 0; a < 3; a++) ti.root_center[a] = f->root_c[a];
            ti.root_half = f->root_half;
            int nl = 0;
            for (auto* l : f->levels) {
                if (l->n == 0) break;
                ti.n_nodes += l->n;
                nl++;
            }
            ti.n_levels = nl;
        }
        if (per_frame) {
            if ((rc = sync_counters(f))) return tag(f, rc);
            per_frame[k] = f->st;
        }
        if (on_mesh)
            on_mesh(user, k, &ti, f->mesh_v.data(), (int64_t)f->mesh_v.size() / 3, f->mesh_t.data(),
                    (int64_t)f->mesh_t.size() / 3);
    }
    return PISURF_OK;
    API_END(c)
}

int pisurf_get_stats(pisurf_ctx* c, pisurf_stats* out) {
    if (!c || !out) return PISURF_E_INVALID;
    int rc = sync_counters(c);
    if (rc) return rc;
    *out = c->st;
    return PISURF_OK;
}

int pisurf_reset_stats(pisurf_ctx* c) {
    if (!c) return PISURF_E_INVALID;
    std::memset(&c->st, 0, sizeof(c->st));
    c->host_sample_entries = 0;
    c->n_fp64_eig = 0;
    HIPCHK(c, hipMemsetAsync(c->counters.p, 0, kCtrStripes * sizeof(Counters), c->stream));
    HIPCHK(c, c->wait());
    return PISURF_OK;
}

}  // extern "C"
