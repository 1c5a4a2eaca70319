// C++ mirror of the reference's HashGrid / Evaluator / SurfReconstructor / Mesh classes over the
// pisurf C ABI, so code written against include/hash_grid.h, include/evaluator.h and
// include/surface_reconstructor.h of the reference compiles against this header with its type
// names (Eigen types are replaced by layout-identical PODs: Vector3f = 3 packed floats,
// Matrix3f = 9 floats column-major). Header-only; link with libpisurf.so.
//
// Per-point methods (SingleEval, GetPIdxList, ...) exist for source compatibility and make one
// device call each; the batched path is SurfReconstructor::Run, which drives the device octree
// build (pisurf_run_frame). Errors that the reference would turn into undefined behaviour throw
// pisurf::Error here.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
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
    void apply() { check(pisurf_set_params(ctx_, &prm_), "set_params"); }
    void check(int rc, const char* what) const {
        if (rc != PISURF_OK) throw Error(rc, std::string(what) + ": " + pisurf_last_error(ctx_));
    }

  private:
    pisurf_ctx* ctx_ = nullptr;
    pisurf_params prm_;
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
    std::vector<Vector3f>* Particles;
    unsigned int ParticlesNum = 0;
    float Radius = 0, CellSize = 0;
    float Bounding[6];
    unsigned int XYZCellNum[3];
    unsigned long long CellNum = 0;
    std::vector<int> IndexList;
    std::map<long long, int> StartList, EndList;

    void GetPIdxList(const Vector3f& pos, std::vector<int>& pIdxList) {
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

// include/evaluator.h:13-108 (const-radius path; the variable-radius searcher is not used)
class Evaluator {
  public:
    Evaluator(std::shared_ptr<HashGrid>& hashgrid, std::vector<Vector3f>* global_particles, float radius)
        : _hashgrid(hashgrid), GlobalPoses(global_particles), Radius(radius), dev_(hashgrid->device()) {
        _GlobalParticlesNum = (int)global_particles->size();
    }
    std::shared_ptr<HashGrid> _hashgrid;
    std::vector<Vector3f>* GlobalPoses;
    int _GlobalParticlesNum = 0;
    float Radius = 0;
    std::vector<Vector3f> PariclesNormals, GlobalxMeans;
    std::vector<Matrix3f> GlobalGs;
    std::vector<float> GlobalDeterminant;

    void SingleEval(const Vector3f& pos, float& scalar) {
        double f;
        dev_->check(pisurf_eval_points(dev_->get(), pos.v, 1, &f, nullptr), "eval_points");
        scalar = (float)f;
    }
    void SingleEvalWithGrad(const Vector3f& pos, float& scalar, Vector3f& gradient) {
        double f, g[3];
        dev_->check(pisurf_eval_points(dev_->get(), pos.v, 1, &f, g), "eval_points");
        scalar = (float)f;
        for (int a = 0; a < 3; a++) gradient[a] = (float)g[a];
    }
    // batched form of SingleEval (n points, xyz packed)
    void BatchEval(const float* xyz, int64_t n, double* f, double* grad = nullptr) {
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
    void constants() { dev_->check(pisurf_iso_constants(dev_->get(), &max_scalar_, &iso_), "iso_constants"); }
    std::shared_ptr<Device> dev_;
    double max_scalar_ = -1, iso_ = 0;
};

// include/global.h:47-115 (vertices and 1-based triangles)
struct Triangle {
    int v[3];
};
class Mesh {
  public:
    std::vector<Vector3f> vertices;
    std::vector<Triangle> tris;
    unsigned int verticesNum = 0, trianglesNum = 0;
    void reset() {
        vertices.clear();
        tris.clear();
        verticesNum = trianglesNum = 0;
    }
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

// include/surface_reconstructor.h:21-93
class SurfReconstructor {
  public:
    SurfReconstructor(std::vector<Vector3f>& particles, std::vector<float>& radiuses, Mesh* mesh, float radius,
                      int device = 0)
        : _GlobalParticles(particles), _GlobalRadiuses(radiuses), _RADIUS(radius), _OurMesh(mesh),
          dev_(std::make_shared<Device>(device)) {
        _GlobalParticlesNum = (int)_GlobalParticles.size();
    }
    // Run(iso_factor, smooth_factor) (src/surface_reconstructor.cpp:409-479) on the device;
    // precision / qef are pisurf extensions (default: the reference's FP32, QEF off)
    void Run(float iso_factor, float smooth_factor, int precision = PISURF_FP32, int qef_mode = PISURF_QEF_OFF) {
        _OurMesh->reset();
        pisurf_params& p = dev_->params();
        p.radius = _RADIUS;
        p.iso_factor = iso_factor;
        p.smooth_factor = smooth_factor;
        p.precision = precision;
        p.qef_mode = qef_mode;
        dev_->apply();
        dev_->check(pisurf_upload_particles(dev_->get(), &_GlobalParticles[0][0], _GlobalParticlesNum, 0), "upload");
        pisurf_tree_info ti;
        dev_->check(pisurf_run_frame(dev_->get(), &ti), "run_frame");
        _DEPTH_MAX = ti.depth_max;
        _DEPTH_MIN = ti.depth_min;
        _RootHalfLength = ti.root_half;
        for (int a = 0; a < 3; a++) _RootCenter[a] = ti.root_center[a];
        pisurf_stats st;
        dev_->check(pisurf_get_stats(dev_->get(), &st), "get_stats");
        std::vector<double> v(3 * st.n_verts);
        std::vector<int32_t> t(3 * st.n_tris);
        dev_->check(pisurf_mesh_download(dev_->get(), v.data(), t.data()), "mesh_download");
        _OurMesh->vertices.resize(st.n_verts);
        _OurMesh->tris.resize(st.n_tris);
        for (int64_t i = 0; i < st.n_verts; i++)
            _OurMesh->vertices[i] = Vector3f((float)v[3 * i], (float)v[3 * i + 1], (float)v[3 * i + 2]);
        for (int64_t i = 0; i < st.n_tris; i++)
            _OurMesh->tris[i] = Triangle{{t[3 * i], t[3 * i + 1], t[3 * i + 2]}};
        _OurMesh->verticesNum = (unsigned)st.n_verts;
        _OurMesh->trianglesNum = (unsigned)st.n_tris;
    }
    int getOverSampleQEF() const { return 2; }
    float getBorder() const { return 0.0f; }
    int getDepthMax() const { return _DEPTH_MAX; }
    int getDepthMin() const { return _DEPTH_MIN; }
    std::vector<Vector3f>* getGlobalParticles() { return &_GlobalParticles; }
    int getGlobalParticlesNum() const { return _GlobalParticlesNum; }
    float getConstRadius() const { return _RADIUS; }
    std::shared_ptr<Device> device() const { return dev_; }

  private:
    std::vector<Vector3f> _GlobalParticles;
    std::vector<float> _GlobalRadiuses;
    float _RADIUS = 0;
    int _GlobalParticlesNum = 0;
    Mesh* _OurMesh;
    std::shared_ptr<Device> dev_;
    int _DEPTH_MAX = 0, _DEPTH_MIN = 0;
    float _RootHalfLength = 0, _RootCenter[3] = {0, 0, 0};
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
