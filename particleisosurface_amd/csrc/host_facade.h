// C++ mirror of the reference's HashGrid / Evaluator / SurfReconstructor / TNode / Mesh classes
// over the pisurf C ABI, so code written against include/hash_grid.h, include/evaluator.h,
// include/surface_reconstructor.h, include/iso_method_ours.h and include/global.h of the
// reference compiles against this header with its type names (Eigen types are replaced by
// layout-identical PODs: Vector3f = 3 packed floats, Vector4f = 4, Matrix3f = 9 floats
// column-major). Header-only; link with libpisurf.so.
//
// Threading: the reference calls SingleEval, SingleEvalWithGrad, GetPIdxList and
// GetInBoxParticles from OpenMP loops (src/visitorextract.cpp:44, src/evaluator.cpp:365,658,720,
// src/surface_reconstructor.cpp:348-388). Here every call into the device context holds the
// context's mutex, and point evaluations go through a micro-batch: the points that concurrent
// callers submit while a batch is on the device are evaluated together in the next launch, each
// caller receiving exactly the value a single-point call would give. The batched paths are
// SurfReconstructor::Run (the device octree build and extraction, pisurf_run_frame) and
// Evaluator::BatchEval. Errors that the reference would turn into undefined behaviour throw
// pisurf::Error.
#pragma once
#include <array>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pisurf.h"

namespace pisurf {

struct Vector3f {
    float v[3] = {0, 0, 0};
    Vector3f() = default;
    Vector3f(float x, float y, float z) : v{x, y, z} {}
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float z() const { return v[2]; }
};
static_assert(sizeof(Vector3f) == 12, "Vector3f must be 3 packed floats");
// Eigen::Vector4f as TNode::node uses it: (x, y, z, f)
struct Vector4f {
    float v[4] = {0, 0, 0, 0};
    Vector4f() = default;
    Vector4f(float x, float y, float z, float w) : v{x, y, z, w} {}
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    Vector3f head3() const { return Vector3f(v[0], v[1], v[2]); }
    static Vector4f Zero() { return Vector4f(); }
};
struct Vector3i {
    int v[3] = {0, 0, 0};
    int& operator[](int i) { return v[i]; }
    int operator[](int i) const { return v[i]; }
};
struct Matrix3f {
    float m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // column-major
    float operator()(int r, int c) const { return m[c * 3 + r]; }
    float& operator()(int r, int c) { return m[c * 3 + r]; }
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// one shared device context per reconstructor (the reference shares HashGrid / Evaluator by shared_ptr)
class Device {
  public:
    explicit Device(int device = 0) {
        int rc = pisurf_create(&ctx_, device);
        if (rc != PISURF_OK) throw Error(rc, "pisurf_create failed: no usable GPU");
        pisurf_default_params(&prm_);
    }
    ~Device() { pisurf_destroy(ctx_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    pisurf_ctx* get() const { return ctx_; }
    pisurf_params& params() { return prm_; }
    // every call into the context holds this mutex (the ABI admits one thread at a time)
    std::mutex& mutex() { return call_; }
    void apply() {
        std::lock_guard<std::mutex> lk(call_);
        check(pisurf_set_params(ctx_, &prm_), "set_params");
    }
    void check(int rc, const char* what) const {
        if (rc != PISURF_OK) throw Error(rc, std::string(what) + ": " + pisurf_last_error(ctx_));
    }
    // SingleEval / SingleEvalWithGrad of one point (grad may be NULL), safe from any number of
    // threads: the first caller to find no batch on the device takes every point submitted so far
    // and evaluates them in one pisurf_eval_points call per kind (with and without gradient); the
    // others wait for their results. A point's value does not depend on the batch it rides in.
    void eval_point(const float p[3], double& f, double* grad) {
        PointReq r;
        for (int a = 0; a < 3; a++) r.p[a] = p[a];
        r.grad = grad != nullptr;
        std::unique_lock<std::mutex> lk(queue_m_);
        queue_.push_back(&r);
        while (!r.done) {
            if (!leader_) {
                leader_ = true;
                std::vector<PointReq*> batch;
                batch.swap(queue_);
                lk.unlock();
                std::exception_ptr err;
                try {
                    run_batch(batch);
                } catch (...) {
                    err = std::current_exception();
                }
                lk.lock();
                for (PointReq* b : batch) {
                    b->err = err;
                    b->done = true;
                }
                leader_ = false;
                queue_cv_.notify_all();
            } else {
                queue_cv_.wait(lk);
            }
        }
        lk.unlock();
        if (r.err) std::rethrow_exception(r.err);
        f = r.f;
        if (grad)
            for (int a = 0; a < 3; a++) grad[a] = r.g[a];
    }

  private:
    struct PointReq {
        float p[3];
        double f = 0, g[3] = {0, 0, 0};
        bool grad = false, done = false;
        std::exception_ptr err;
    };
    void run_batch(const std::vector<PointReq*>& batch) {
        for (int kind = 0; kind < 2; kind++) {
            std::vector<PointReq*> sel;
            for (PointReq* b : batch)
                if (b->grad == (kind == 1)) sel.push_back(b);
            if (sel.empty()) continue;
            std::vector<float> xyz(3 * sel.size());
            std::vector<double> f(sel.size()), g(kind ? 3 * sel.size() : 0);
            for (size_t i = 0; i < sel.size(); i++)
                for (int a = 0; a < 3; a++) xyz[3 * i + a] = sel[i]->p[a];
            {
                std::lock_guard<std::mutex> lk(call_);
                check(pisurf_eval_points(ctx_, xyz.data(), (int64_t)sel.size(), f.data(), kind ? g.data() : nullptr),
                      "eval_points");
            }
            for (size_t i = 0; i < sel.size(); i++) {
                sel[i]->f = f[i];
                if (kind)
                    for (int a = 0; a < 3; a++) sel[i]->g[a] = g[3 * i + a];
            }
        }
    }
    pisurf_ctx* ctx_ = nullptr;
    pisurf_params prm_;
    std::mutex call_;
    std::mutex queue_m_;
    std::condition_variable queue_cv_;
    std::vector<PointReq*> queue_;
    bool leader_ = false;
};

// include/hash_grid.h:9-45 (const-radius constructor)
class HashGrid {
  public:
    HashGrid(std::shared_ptr<Device> dev, std::vector<Vector3f>* particles, float* bounding, float radius,
             float inf_factor)
        : dev_(std::move(dev)), Particles(particles) {
        ParticlesNum = (unsigned)Particles->size();
        Radius = radius;
        dev_->params().radius = radius;
        dev_->apply();
        std::lock_guard<std::mutex> lk(dev_->mutex());
        dev_->check(pisurf_upload_particles(dev_->get(), &(*Particles)[0][0], ParticlesNum, 0), "upload");
        pisurf_grid_info gi;
        dev_->check(pisurf_hashgrid_build(dev_->get(), bounding, inf_factor, &gi), "hashgrid_build");
        CellSize = gi.cell_size;
        for (int i = 0; i < 6; i++) Bounding[i] = gi.bounding[i];
        for (int i = 0; i < 3; i++) XYZCellNum[i] = gi.xyz_cell_num[i];
        CellNum = (unsigned long long)gi.cell_num;
        IndexList.resize(ParticlesNum);
        std::vector<int32_t> s(CellNum), e(CellNum);
        dev_->check(pisurf_hashgrid_download(dev_->get(), IndexList.data(), s.data(), e.data()), "hashgrid_download");
        for (unsigned long long c = 0; c < CellNum; c++)
            if (e[c] > s[c]) {
                StartList[(long long)c] = s[c];
                EndList[(long long)c] = e[c];
            }
    }
    // the grid of a context that builds it itself (SurfReconstructor::Run / BuildTree): queries go
    // to the device grid; the table fields (IndexList, StartList, EndList) are not mirrored
    HashGrid(std::shared_ptr<Device> dev, std::vector<Vector3f>* particles, float radius)
        : dev_(std::move(dev)), Particles(particles) {
        ParticlesNum = (unsigned)Particles->size();
        Radius = radius;
        CellSize = radius * 4.0f;
    }
    std::vector<Vector3f>* Particles;
    unsigned int ParticlesNum = 0;
    float Radius = 0, CellSize = 0;
    float Bounding[6] = {0, 0, 0, 0, 0, 0};
    unsigned int XYZCellNum[3] = {0, 0, 0};
    unsigned long long CellNum = 0;
    std::vector<int> IndexList;
    std::map<long long, int> StartList, EndList;

    void GetPIdxList(const Vector3f& pos, std::vector<int>& pIdxList) {
        std::lock_guard<std::mutex> lk(dev_->mutex());
        int64_t off[2], tot = 0;
        dev_->check(pisurf_neighbors(dev_->get(), pos.v, 1, off, nullptr, 0, &tot), "neighbors");
        const size_t base = pIdxList.size();
        pIdxList.resize(base + tot);
        dev_->check(pisurf_neighbors(dev_->get(), pos.v, 1, off, pIdxList.data() + base, tot, &tot), "neighbors");
    }
    void CalcXYZIdx(const Vector3f& pos, Vector3i& xyzIdx) const {
        for (int i = 0; i < 3; i++) xyzIdx[i] = int((pos[i] - Bounding[i * 2]) / CellSize);
    }
    long long CalcCellHash(const Vector3i& xyzIdx) const {
        if (xyzIdx[0] < 0 || xyzIdx[0] >= (int)XYZCellNum[0] || xyzIdx[1] < 0 || xyzIdx[1] >= (int)XYZCellNum[1] ||
            xyzIdx[2] < 0 || xyzIdx[2] >= (int)XYZCellNum[2])
            return -1;
        return (long long)xyzIdx[2] * XYZCellNum[0] * XYZCellNum[1] + (long long)xyzIdx[1] * XYZCellNum[0] + xyzIdx[0];
    }
    void GetInCellList(const long long hash, std::vector<int>& pIdxList) const {
        auto s = StartList.find(hash), e = EndList.find(hash);
        if (s == StartList.end() || e == EndList.end()) return;
        for (int k = s->second; k < e->second; k++) pIdxList.push_back(IndexList[k]);
    }
    void GetInBoxParticles(Vector3f box1, Vector3f box2, std::vector<int>& insides) {
        std::lock_guard<std::mutex> lk(dev_->mutex());
        int64_t tot = 0;
        dev_->check(pisurf_box_particles(dev_->get(), box1.v, box2.v, nullptr, 0, &tot), "box_particles");
        const size_t base = insides.size();
        insides.resize(base + tot);
        dev_->check(pisurf_box_particles(dev_->get(), box1.v, box2.v, insides.data() + base, tot, &tot), "box_particles");
    }
    std::shared_ptr<Device> device() const { return dev_; }

  private:
    std::shared_ptr<Device> dev_;
};

// include/multi_level_researcher.h:8-38. The reference's Evaluator never queries it (every
// GetNeighbors call in src/evaluator.cpp is commented out), so only the constant-radius case
// exists: the radius statistics the callers read (calc_vertices reads getMinRadius when the
// radius varies, src/visitorextract.cpp:52) and no per-radius grids.
class MultiLevelSearcher {
  public:
    MultiLevelSearcher() = default;
    MultiLevelSearcher(std::vector<Vector3f>* particles, float* bounding, std::vector<float>* radiuses,
                       float inf_factor)
        : infFactor(inf_factor), particlesNum(particles ? (unsigned)particles->size() : 0u) {
        (void)bounding;
        if (radiuses && !radiuses->empty()) {
            double sum = 0;
            minRadius = maxRadius = (*radiuses)[0];
            for (float r : *radiuses) {
                minRadius = std::min(minRadius, r);
                maxRadius = std::max(maxRadius, r);
                sum += r;
            }
            avgRadius = float(sum / radiuses->size());
        }
    }
    float getMaxRadius() const { return maxRadius; }
    float getMinRadius() const { return minRadius; }
    float getAvgRadius() const { return avgRadius; }
    float getParticlesNum() const { return float(particlesNum); }

  private:
    float maxRadius = 0, minRadius = 0, avgRadius = 0;
    float infFactor = 0;
    unsigned int particlesNum = 0;
};

// include/evaluator.h:13-108 (const-radius path)
class Evaluator {
  public:
    // Evaluator(hashgrid, searcher, particles, radiuses, radius) (include/evaluator.h:76-80,
    // constructed at src/surface_reconstructor.cpp:432); the device state is the hash grid's
    Evaluator(std::shared_ptr<HashGrid>& hashgrid, std::shared_ptr<MultiLevelSearcher>& searcher,
              std::vector<Vector3f>* global_particles, std::vector<float>* radiuses, float radius)
        : _hashgrid(hashgrid), _searcher(searcher), GlobalPoses(global_particles), GlobalRadius(radiuses),
          Radius(radius), dev_(hashgrid->device()) {
        _GlobalParticlesNum = (int)global_particles->size();
    }
    std::shared_ptr<HashGrid> _hashgrid;
    std::shared_ptr<MultiLevelSearcher> _searcher;
    std::vector<Vector3f>* GlobalPoses;
    std::vector<float>* GlobalRadius;
    int _GlobalParticlesNum = 0;
    float Radius = 0;
    std::vector<Vector3f> PariclesNormals, GlobalxMeans;
    std::vector<Matrix3f> GlobalGs;
    std::vector<float> GlobalDeterminant;

    // safe from concurrent threads (micro-batched, see Device::eval_point)
    void SingleEval(const Vector3f& pos, float& scalar) {
        double f;
        dev_->eval_point(pos.v, f, nullptr);
        scalar = (float)f;
    }
    void SingleEvalWithGrad(const Vector3f& pos, float& scalar, Vector3f& gradient) {
        double f, g[3];
        dev_->eval_point(pos.v, f, g);
        scalar = (float)f;
        for (int a = 0; a < 3; a++) gradient[a] = (float)g[a];
    }
    // batched form of SingleEval (n points, xyz packed)
    void BatchEval(const float* xyz, int64_t n, double* f, double* grad = nullptr) {
        std::lock_guard<std::mutex> lk(dev_->mutex());
        dev_->check(pisurf_eval_points(dev_->get(), xyz, n, f, grad), "eval_points");
    }
    // Evaluator::GridEval (src/evaluator.cpp:127-252): values of the (os+1)^3 sample points and
    // their one-sided / central differences
    void GridEval(float* sample_points, float* field_gradients, float cellsize, bool& signchange, int oversample,
                  bool grad_normalize = false) {
        const int n1 = oversample + 1, ns = n1 * n1 * n1;
        std::vector<float> xyz(3 * ns);
        std::vector<double> f(ns);
        for (int i = 0; i < ns; i++)
            for (int a = 0; a < 3; a++) xyz[3 * i + a] = sample_points[4 * i + a];
        BatchEval(xyz.data(), ns, f.data());
        const float step = cellsize / oversample;
        for (int i = 0; i < ns; i++) {
            sample_points[4 * i + 3] = (float)f[i];
            const bool origin = sample_points[3] >= 0;
            if (!signchange) signchange = origin ^ (sample_points[4 * i + 3] >= 0);
        }
        const int stride[3] = {1, n1, n1 * n1};
        for (int i = 0; i < ns; i++) {
            const int t[3] = {i % n1, (i / n1) % n1, i / (n1 * n1)};
            float g[3];
            for (int a = 0; a < 3; a++) {
                const float* s = sample_points;
                if (t[a] == 0) g[a] = (s[4 * i + 3] - s[4 * (i + stride[a]) + 3]) / step;
                else if (t[a] == oversample) g[a] = (s[4 * (i - stride[a]) + 3] - s[4 * i + 3]) / step;
                else g[a] = (s[4 * (i - stride[a]) + 3] - s[4 * (i + stride[a]) + 3]) / (step * 2);
            }
            if (grad_normalize) {
                const float l = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
                for (int a = 0; a < 3; a++) g[a] = (l > 0) ? g[a] / l : 0.0f;
            }
            for (int a = 0; a < 3; a++) field_gradients[3 * i + a] = g[a];
        }
    }
    void compute_Gs_xMeans() {
        std::lock_guard<std::mutex> lk(dev_->mutex());
        dev_->check(pisurf_anisotropy(dev_->get()), "anisotropy");
        const size_t n = _GlobalParticlesNum;
        std::vector<double> xm(3 * n), G(9 * n), det(n);
        dev_->check(pisurf_download_aniso(dev_->get(), xm.data(), G.data(), det.data()), "download_aniso");
        GlobalxMeans.resize(n);
        GlobalGs.resize(n);
        GlobalDeterminant.resize(n);
        for (size_t i = 0; i < n; i++) {
            for (int a = 0; a < 3; a++) GlobalxMeans[i][a] = (float)xm[3 * i + a];
            for (int k = 0; k < 9; k++) GlobalGs[i].m[k] = (float)G[9 * i + k];
            GlobalDeterminant[i] = (float)det[i];
        }
    }
    void CalcParticlesNormal() {
        std::lock_guard<std::mutex> lk(dev_->mutex());
        std::vector<double> nrm(3 * (size_t)_GlobalParticlesNum);
        dev_->check(pisurf_particle_normals(dev_->get(), nrm.data()), "particle_normals");
        PariclesNormals.resize(_GlobalParticlesNum);
        for (size_t i = 0; i < PariclesNormals.size(); i++)
            for (int a = 0; a < 3; a++) PariclesNormals[i][a] = (float)nrm[3 * i + a];
    }
    void CalculateMaxScalarConstR() { constants(); }
    void RecommendIsoValueConstR() { constants(); }
    bool CheckSplash(const int&) { return false; }  // splash flags are never set (src/evaluator.cpp:254-261)
    float getNeighborFactor() const { return 4.0f; }
    float getSmoothFactor() const { return dev_->params().smooth_factor; }
    void setSmoothFactor(float s) { dev_->params().smooth_factor = s; dev_->apply(); }
    float getIsoFactor() const { return dev_->params().iso_factor; }
    void setIsoFactor(float f) { dev_->params().iso_factor = f; dev_->apply(); }
    float getIsoValue() { constants(); return (float)iso_; }
    float getMaxScalar() { constants(); return (float)max_scalar_; }

  private:
    void constants() {
        std::lock_guard<std::mutex> lk(dev_->mutex());
        dev_->check(pisurf_iso_constants(dev_->get(), &max_scalar_, &iso_), "iso_constants");
    }
    std::shared_ptr<Device> dev_;
    double max_scalar_ = -1, iso_ = 0;
};

// include/global.h:47-115 and src/global.cpp:13-60: vertices and 1-based triangles
struct Triangle {
    std::array<int, 3> v{{0, 0, 0}};
    Triangle() = default;
    Triangle(int a, int b, int c) : v{{a, b, c}} {}
};
class Mesh {
  public:
    explicit Mesh(int mesh_precision = 10000) : MESH_PRECISION(mesh_precision) {}
    int MESH_PRECISION;
    std::vector<Vector3f> vertices;
    std::vector<Triangle> tris;
    unsigned int verticesNum = 0, trianglesNum = 0;
    // Mesh::insert_vert (src/global.cpp:13-32): one vertex per unordered pair of node ids (the
    // reference keys its map by the sorted ids), the first insertion's position kept; returns the
    // vertex's 1-based index
    int insert_vert(unsigned long long id1, unsigned long long id2, const Vector3f& p) {
        if (id2 < id1) std::swap(id1, id2);
        auto r = vertices_map.emplace(std::array<unsigned long long, 2>{{id1, id2}}, (int)verticesNum);
        if (r.second) {
            verticesNum++;
            vertices.push_back(p);
        }
        return r.first->second + 1;
    }
    // Mesh::insert_tri (src/global.cpp:34-53): one triangle per vertex set (keyed by the sorted
    // indices), the first insertion's orientation kept
    void insert_tri(int t0, int t1, int t2) {
        std::array<int, 3> k{{t0, t1, t2}};
        if (k[1] < k[0]) std::swap(k[0], k[1]);
        if (k[2] < k[1]) std::swap(k[1], k[2]);
        if (k[1] < k[0]) std::swap(k[0], k[1]);
        if (tris_map.emplace(k, (int)trianglesNum).second) {
            trianglesNum++;
            tris.push_back(Triangle(t0, t1, t2));
        }
    }
    void reset() {
        vertices_map.clear();
        tris_map.clear();
        vertices.clear();
        tris.clear();
        verticesNum = trianglesNum = 0;
    }
    std::map<std::array<unsigned long long, 2>, int> vertices_map;
    std::map<std::array<int, 3>, int> tris_map;
};

// writeObjFile (src/main.cpp:54-67)
inline void writeObjFile(const Mesh& m, const std::string& fn) {
    FILE* f = std::fopen(fn.c_str(), "w");
    if (!f) throw Error(PISURF_E_INVALID, "cannot open " + fn);
    for (const auto& p : m.vertices) std::fprintf(f, "v %f %f %f\n", p[0], p[1], p[2]);
    for (const auto& t : m.tris) std::fprintf(f, "f %d %d %d\n", t.v[0], t.v[1], t.v[2]);
    std::fclose(f);
}

// writePlyFile (src/main.cpp:69-114): binary little-endian PLY as the reference's rply writer lays
// it out: float32 x, y, z per vertex, faces as a uchar count and 0-based int indices
inline void writePlyFile(const Mesh& m, const std::string& fn) {
    FILE* f = std::fopen(fn.c_str(), "wb");
    if (!f) throw Error(PISURF_E_INVALID, "cannot open " + fn);
    std::fprintf(f,
                 "ply\nformat binary_little_endian 1.0\nelement vertex %ld\nproperty float32 x\nproperty float32 y\n"
                 "property float32 z\nelement face %ld\nproperty list uchar int vertex_indices\nend_header\n",
                 (long)m.vertices.size(), (long)m.tris.size());
    for (const auto& p : m.vertices) std::fwrite(p.v, sizeof(float), 3, f);
    for (const auto& t : m.tris) {
        const unsigned char three = 3;
        const int32_t idx[3] = {t.v[0] - 1, t.v[1] - 1, t.v[2] - 1};
        std::fwrite(&three, 1, 1, f);
        std::fwrite(idx, sizeof(int32_t), 3, f);
    }
    std::fclose(f);
}

class SurfReconstructor;

// Node types of include/iso_method_ours.h:20-23
constexpr char EMPTY = 0, INTERNAL = 1, LEAF = 2, UNCERTAIN = 3;

// TNode (include/iso_method_ours.h:25-78): the octree as the reference's traversal reads it.
// node = (vertex x, y, z, f): the leaf centre and its field value (the QEF vertex with
// PISURF_QEF_*); id = parent id * 8 + octant, root 0.
struct TNode {
    SurfReconstructor* constructor = nullptr;
    Vector3f center;
    float half_length = 0;
    Vector4f node;
    char depth = 0;
    char type = EMPTY;
    unsigned long long id = 0;
    std::array<std::shared_ptr<TNode>, 8> children;
    bool is_leaf() const { return type == LEAF || type == EMPTY; }
};

// one dual grid of traverse_node<trav_vert> + VisitorExtract::on_vert
// (src/visitorextract.cpp:229-257): the 8 leaves permuted by the procedure table's `trans`, and the
// sign index selecting the procedure (bit i set when traversal corner i has f > 0)
struct DualGrid {
    uint8_t sign_index = 0;
    std::array<std::shared_ptr<TNode>, 8> grid;
};

// include/surface_reconstructor.h:21-93
class SurfReconstructor {
  public:
    SurfReconstructor(std::vector<Vector3f>& particles, std::vector<float>& radiuses, Mesh* mesh, float radius,
                      int device = 0)
        : _GlobalParticles(particles), _GlobalRadiuses(radiuses), _RADIUS(radius), _OurMesh(mesh),
          dev_(std::make_shared<Device>(device)) {
        _GlobalParticlesNum = (int)_GlobalParticles.size();
    }
    // Run(iso_factor, smooth_factor) (src/surface_reconstructor.cpp:409-479): hash grid, evaluator,
    // octree and extraction on the device, the mesh copied into the caller's Mesh. precision and
    // qef_mode are pisurf extensions (default: the reference's FP32, QEF off).
    void Run(float iso_factor, float smooth_factor, int precision = PISURF_FP32, int qef_mode = PISURF_QEF_OFF) {
        _OurMesh->reset();
        setup(iso_factor, smooth_factor, precision, qef_mode);
        pisurf_tree_info ti;
        std::vector<double> v;
        std::vector<int32_t> t;
        {
            std::lock_guard<std::mutex> lk(dev_->mutex());
            dev_->check(pisurf_run_frame(dev_->get(), &ti), "run_frame");
            pisurf_stats st;
            dev_->check(pisurf_get_stats(dev_->get(), &st), "get_stats");
            v.resize(3 * st.n_verts);
            t.resize(3 * st.n_tris);
            dev_->check(pisurf_mesh_download(dev_->get(), v.data(), t.data()), "mesh_download");
        }
        tree_info(ti);
        // the mesh arrives whole: the vertex / triangle lists are filled directly (insert_vert /
        // insert_tri's maps are left empty)
        const size_t nv = v.size() / 3, nt = t.size() / 3;
        _OurMesh->vertices.resize(nv);
        _OurMesh->tris.resize(nt);
        for (size_t i = 0; i < nv; i++) _OurMesh->vertices[i] = Vector3f((float)v[3 * i], (float)v[3 * i + 1], (float)v[3 * i + 2]);
        for (size_t i = 0; i < nt; i++) _OurMesh->tris[i] = Triangle(t[3 * i], t[3 * i + 1], t[3 * i + 2]);
        _OurMesh->verticesNum = (unsigned)nv;
        _OurMesh->trianglesNum = (unsigned)nt;
        _STATE = 1;
    }
    // The first half of Run, genIsoOurs state 0 (src/surface_reconstructor.cpp:275-407): the octree
    // is built on the device and getRoot() returns it, for a caller that runs the reference's own
    // extraction (VisitorExtract over traverse_node, then calc_vertices and generate_mesh, which
    // call getEvaluator()->SingleEval from OpenMP threads).
    void BuildTree(float iso_factor, float smooth_factor, int precision = PISURF_FP32, int qef_mode = PISURF_QEF_OFF) {
        _OurMesh->reset();
        setup(iso_factor, smooth_factor, precision, qef_mode);
        pisurf_tree_info ti;
        {
            std::lock_guard<std::mutex> lk(dev_->mutex());
            dev_->check(pisurf_build_tree(dev_->get(), &ti), "build_tree");
        }
        tree_info(ti);
        _STATE = 1;
    }
    int getOverSampleQEF() const { return 2; }
    float getBorder() const { return 0.0f; }
    int getDepthMax() const { return _DEPTH_MAX; }
    int getDepthMin() const { return _DEPTH_MIN; }
    std::shared_ptr<HashGrid> getHashGrid() { return _hashgrid; }
    std::shared_ptr<MultiLevelSearcher> getSearcher() { return _searcher; }
    std::shared_ptr<Evaluator> getEvaluator() { return _evaluator; }
    std::vector<Vector3f>* getGlobalParticles() { return &_GlobalParticles; }
    int getGlobalParticlesNum() const { return _GlobalParticlesNum; }
    float getConstRadius() const { return _RADIUS; }
    int getSTATE() const { return _STATE; }
    // the device octree as TNodes (downloaded on first use after Run / BuildTree)
    std::shared_ptr<TNode> getRoot() {
        if (!_OurRoot && _STATE == 1) download_tree();
        return _OurRoot;
    }
    // the dual grids of the octree in the reference's traversal order (what traverse_node with
    // VisitorExtract::on_vert collects), as TNodes of getRoot()'s tree
    std::vector<DualGrid> DualGrids() {
        getRoot();
        std::vector<int32_t> n8;
        std::vector<uint8_t> si;
        {
            std::lock_guard<std::mutex> lk(dev_->mutex());
            int64_t tot = 0;
            dev_->check(pisurf_dual_cells(dev_->get(), nullptr, nullptr, 0, &tot), "dual_cells");
            n8.resize(8 * tot);
            si.resize(tot);
            dev_->check(pisurf_dual_cells(dev_->get(), n8.data(), si.data(), tot, &tot), "dual_cells");
        }
        std::vector<DualGrid> out(si.size());
        for (size_t i = 0; i < si.size(); i++) {
            out[i].sign_index = si[i];
            for (int k = 0; k < 8; k++) out[i].grid[k] = nodes_.at(n8[8 * i + k]);
        }
        return out;
    }
    std::shared_ptr<Device> device() const { return dev_; }

  private:
    // loadRootBox + HashGrid + Evaluator of Run (src/surface_reconstructor.cpp:409-440); the hash
    // grid's tables and the evaluator's per-particle arrays stay on the device (the mirror fills
    // them on request: HashGrid built directly, Evaluator::compute_Gs_xMeans)
    void setup(float iso_factor, float smooth_factor, int precision, int qef_mode) {
        _OurRoot = nullptr;
        nodes_.clear();
        _STATE = 0;
        pisurf_params& p = dev_->params();
        p.radius = _RADIUS;
        p.iso_factor = iso_factor;
        p.smooth_factor = smooth_factor;
        p.precision = precision;
        p.qef_mode = qef_mode;
        dev_->apply();
        {
            std::lock_guard<std::mutex> lk(dev_->mutex());
            dev_->check(pisurf_upload_particles(dev_->get(), &_GlobalParticles[0][0], _GlobalParticlesNum, 0), "upload");
        }
        _hashgrid = std::make_shared<HashGrid>(dev_, &_GlobalParticles, _RADIUS);
        _evaluator = std::make_shared<Evaluator>(_hashgrid, _searcher, &_GlobalParticles, &_GlobalRadiuses, _RADIUS);
    }
    void tree_info(const pisurf_tree_info& ti) {
        _DEPTH_MAX = ti.depth_max;
        _DEPTH_MIN = ti.depth_min;
        _RootHalfLength = ti.root_half;
        for (int a = 0; a < 3; a++) _RootCenter[a] = ti.root_center[a];
        n_nodes_ = ti.n_nodes;
    }
    void download_tree() {
        const size_t n = (size_t)n_nodes_;
        std::vector<float> c3(3 * n), half(n);
        std::vector<int32_t> depth(n), child(n);
        std::vector<uint8_t> type(n);
        std::vector<uint64_t> id(n);
        std::vector<double> node(4 * n);
        {
            std::lock_guard<std::mutex> lk(dev_->mutex());
            dev_->check(pisurf_tree_download(dev_->get(), c3.data(), half.data(), depth.data(), type.data(), id.data(),
                                             node.data(), child.data()),
                        "tree_download");
        }
        nodes_.resize(n);
        for (size_t i = 0; i < n; i++) {
            auto t = std::make_shared<TNode>();
            t->constructor = this;
            t->center = Vector3f(c3[3 * i], c3[3 * i + 1], c3[3 * i + 2]);
            t->half_length = half[i];
            t->node = Vector4f((float)node[4 * i], (float)node[4 * i + 1], (float)node[4 * i + 2], (float)node[4 * i + 3]);
            t->depth = (char)depth[i];
            t->type = (char)type[i];
            t->id = id[i];
            nodes_[i] = t;
        }
        for (size_t i = 0; i < n; i++)
            if (child[i] >= 0)
                for (int k = 0; k < 8; k++) nodes_[i]->children[k] = nodes_.at(child[i] + k);
        _OurRoot = n ? nodes_[0] : nullptr;
    }
    std::vector<Vector3f> _GlobalParticles;
    std::vector<float> _GlobalRadiuses;
    float _RADIUS = 0;
    int _GlobalParticlesNum = 0;
    Mesh* _OurMesh;
    std::shared_ptr<Device> dev_;
    std::shared_ptr<HashGrid> _hashgrid;
    std::shared_ptr<MultiLevelSearcher> _searcher;  // constant radius: none, as in the reference
    std::shared_ptr<Evaluator> _evaluator;
    int _STATE = 0;
    int _DEPTH_MAX = 0, _DEPTH_MIN = 0;
    float _RootHalfLength = 0, _RootCenter[3] = {0, 0, 0};
    int64_t n_nodes_ = 0;
    std::shared_ptr<TNode> _OurRoot;
    std::vector<std::shared_ptr<TNode>> nodes_;
};

// Image data in VTK's XML format (what vtkXMLImageDataWriter produces for UniformGrid::Run,
// src/marching.cpp:155-185), written with raw appended data: one Float32 point array "scalars",
// x fastest, UInt64 block header.
inline void writeVtiFile(const std::string& fn, const int dims[3], const float origin[3], float spacing,
                         const std::vector<float>& scalars) {
    FILE* f = std::fopen(fn.c_str(), "wb");
    if (!f) throw Error(PISURF_E_INVALID, "cannot open " + fn);
    std::fprintf(f,
                 "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" "
                 "header_type=\"UInt64\">\n  <ImageData WholeExtent=\"0 %d 0 %d 0 %d\" Origin=\"%.9g %.9g %.9g\" "
                 "Spacing=\"%.9g %.9g %.9g\" Direction=\"1 0 0 0 1 0 0 0 1\">\n    <Piece Extent=\"0 %d 0 %d 0 %d\">\n"
                 "      <PointData Scalars=\"scalars\">\n        <DataArray type=\"Float32\" Name=\"scalars\" "
                 "format=\"appended\" offset=\"0\"/>\n      </PointData>\n      <CellData>\n      </CellData>\n"
                 "    </Piece>\n  </ImageData>\n  <AppendedData encoding=\"raw\">\n   _",
                 dims[0] - 1, dims[1] - 1, dims[2] - 1, origin[0], origin[1], origin[2], spacing, spacing, spacing,
                 dims[0] - 1, dims[1] - 1, dims[2] - 1);
    const uint64_t nb = (uint64_t)scalars.size() * sizeof(float);
    std::fwrite(&nb, sizeof(nb), 1, f);
    if (nb) std::fwrite(scalars.data(), 1, nb, f);
    std::fprintf(f, "\n  </AppendedData>\n</VTKFile>\n");
    std::fclose(f);
}

// include/marching.h:18-72 (constant-radius constructor; the variable-radius one is disabled in
// the reference's field path)
class UniformGrid {
  public:
    float _RADIUS = 0;
    int _GlobalParticlesNum = 0;
    float _BoundingBox[6] = {0, 0, 0, 0, 0, 0};
    int dims[3] = {0, 0, 0};
    int steps[3] = {0, 0, 0};
    std::vector<float> _Scalars;
    std::vector<Vector3f> _GlobalParticles;

    UniformGrid(const std::vector<Vector3f>& particles, float radius, int device = 0)
        : _RADIUS(radius), _GlobalParticles(particles), dev_(std::make_shared<Device>(device)) {
        _GlobalParticlesNum = (int)_GlobalParticles.size();
    }
    // Run (src/marching.cpp:93-186): samples the field on the resized box and writes
    // <filepath>/<filename>.vti; iso_value is unused, as in the reference
    void Run(float iso_value, const std::string& filename, const std::string& filepath,
             int precision = PISURF_FP32) {
        (void)iso_value;
        pisurf_params& p = dev_->params();
        p.radius = _RADIUS;
        p.precision = precision;
        dev_->apply();
        dev_->check(pisurf_upload_particles(dev_->get(), &_GlobalParticles[0][0], _GlobalParticlesNum, 0), "upload");
        float sp = 0;
        dev_->check(pisurf_uniform_grid(dev_->get(), dims, _BoundingBox, &sp, nullptr), "uniform_grid");
        for (int a = 0; a < 3; a++) steps[a] = dims[a] - 1;
        _Scalars.resize((size_t)dims[0] * dims[1] * dims[2]);
        dev_->check(pisurf_uniform_grid(dev_->get(), nullptr, nullptr, nullptr, _Scalars.data()), "uniform_grid");
        const float origin[3] = {_BoundingBox[0], _BoundingBox[2], _BoundingBox[4]};
        writeVtiFile(filepath + "/" + filename + ".vti", dims, origin, sp, _Scalars);
    }
    std::shared_ptr<Device> device() const { return dev_; }

  private:
    std::shared_ptr<Device> dev_;
};

}  // namespace pisurf
