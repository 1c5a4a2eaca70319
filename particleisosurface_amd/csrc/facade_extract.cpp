// The reference's extraction stage run by its own kind of caller on the C++ mirror: genIsoOurs
// state 1 (src/surface_reconstructor.cpp:290-324) with VisitorExtract::calc_vertices and
// generate_mesh (src/visitorextract.cpp:41-227) written as the reference writes them, over the
// TNode octree that SurfReconstructor::getRoot() returns after BuildTree, with the bisection's
// SingleEval calls made from OpenMP threads through getEvaluator(). The mesh is then compared
// with the library's own reference-order extraction of the same frame, and the per-point calls
// made from the threads with one batched call on the same points.
//
//   pisurf_extract_demo <particles.f32> <radius> <threads>
// prints one JSON line. particles.f32: raw little-endian float32 x, y, z triples.
#include <omp.h>

#include <array>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <vector>

#include "dmc_tables.h"
#include "host_facade.h"

using pisurf::Vector3f;
using pisurf::Vector4f;

namespace {

// sign (include/utils.h:7-11): f > 0 is outside
int sign(const Vector4f& x) { return x[3] > 0 ? 1 : -1; }

struct DualCell {
    std::array<Vector3f, 12> vertices;
    std::vector<std::array<int, 3>> faces;
};

// VisitorExtract::calc_vertices (src/visitorextract.cpp:41-201): every dual grid's 12 edges in a
// parallel loop; an edge whose leaves differ in sign is bisected with SingleEval until it is no
// longer than r/2, then the zero crossing is interpolated (snapped below 0.1 and above 0.9)
void calc_vertices(pisurf::SurfReconstructor& sr, const std::vector<pisurf::DualGrid>& grids,
                   std::vector<DualCell>& cells) {
    cells.resize(grids.size());
    const auto ev = sr.getEvaluator();
    const float r = sr.getConstRadius() / 2;
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)grids.size(); i++) {
        auto point = [&](int e, int a, int b) {
            const pisurf::TNode& v1 = *grids[i].grid[a];
            const pisurf::TNode& v2 = *grids[i].grid[b];
            if (sign(v1.node) == sign(v2.node)) return;
            Vector4f t1 = v1.node, t2 = v2.node, t;
            const float dx = t1[0] - t2[0], dy = t1[1] - t2[1], dz = t1[2] - t2[2];
            float d = std::sqrt((dx * dx + dy * dy) + dz * dz);
            while (d > r) {
                for (int k = 0; k < 3; k++) t[k] = (t1[k] + t2[k]) / 2;
                ev->SingleEval(t.head3(), t[3]);
                if (sign(t) == sign(t1)) {
                    t1 = t;
                    t = Vector4f::Zero();
                } else if (sign(t) == sign(t2)) {
                    t2 = t;
                    t = Vector4f::Zero();
                } else {
                    break;
                }
                d /= 2;
            }
            const float ratio = (0.0f - t1[3]) / (t2[3] - t1[3]);
            if (ratio < 0.1f) t = t1;
            else if (ratio > 0.9f) t = t2;
            else
                for (int k = 0; k < 4; k++) t[k] = t1[k] + (t2[k] - t1[k]) * ratio;
            cells[i].vertices[e] = t.head3();
        };
        for (int e = 0; e < 12; e++) point(e, kDmcEdge[e][0], kDmcEdge[e][1]);
        // the procedure's triangles (src/visitorextract.cpp:95-199)
        const uint8_t code = kDmcCase[grids[i].sign_index];
        for (int k = 0; k < PISURF_DMC_NCASES; k++)
            if (kDmcCanon[k] == code)
                for (int tt = 0; tt < kDmcNTri[k]; tt++)
                    cells[i].faces.push_back(
                        {{kDmcTri[k][tt][0], kDmcTri[k][tt][1], kDmcTri[k][tt][2]}});
    }
}

// VisitorExtract::generate_mesh (src/visitorextract.cpp:203-227)
void generate_mesh(const std::vector<pisurf::DualGrid>& grids, const std::vector<DualCell>& cells, pisurf::Mesh& m) {
    int ids[3];
    for (size_t i = 0; i < grids.size(); i++)
        for (const auto& f : cells[i].faces) {
            for (int k = 0; k < 3; k++)
                ids[k] = m.insert_vert(grids[i].grid[kDmcEdge[f[k]][0]]->id,
                                       grids[i].grid[kDmcEdge[f[k]][1]]->id, cells[i].vertices[f[k]]);
            if (kDmcMirror[grids[i].sign_index]) m.insert_tri(ids[0], ids[1], ids[2]);
            else m.insert_tri(ids[2], ids[1], ids[0]);
        }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s particles.f32 radius threads\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        return 2;
    }
    std::vector<Vector3f> particles;
    float xyz[3];
    while (std::fread(xyz, sizeof(float), 3, f) == 3) particles.emplace_back(xyz[0], xyz[1], xyz[2]);
    std::fclose(f);
    const float r = std::strtof(argv[2], nullptr);
    const int threads = std::atoi(argv[3]);
    omp_set_num_threads(threads);
    try {
        std::vector<float> radiuses;
        pisurf::Mesh mesh;
        pisurf::SurfReconstructor sr(particles, radiuses, &mesh, r);
        sr.BuildTree(1.9f, 2.0f, PISURF_FP32);
        // the octree as the reference's traversal sees it
        long nodes = 0, leaves = 0;
        std::function<void(const std::shared_ptr<pisurf::TNode>&)> walk = [&](const std::shared_ptr<pisurf::TNode>& n) {
            nodes++;
            if (n->is_leaf()) {
                leaves++;
                return;
            }
            for (const auto& c : n->children) walk(c);
        };
        walk(sr.getRoot());
        const std::vector<pisurf::DualGrid> grids = sr.DualGrids();
        std::vector<DualCell> cells;
        calc_vertices(sr, grids, cells);
        generate_mesh(grids, cells, mesh);

        // the library's own extraction of the same frame, in the reference's order
        pisurf::Mesh lib_mesh;
        pisurf::SurfReconstructor lib(particles, radiuses, &lib_mesh, r);
        lib.device()->params().mesh_order = 1;
        lib.Run(1.9f, 2.0f, PISURF_FP32);
        long vdiff = 0, tdiff = 0;
        const bool same_size = lib_mesh.vertices.size() == mesh.vertices.size() && lib_mesh.tris.size() == mesh.tris.size();
        if (same_size) {
            for (size_t i = 0; i < mesh.vertices.size(); i++)
                for (int a = 0; a < 3; a++)
                    if (mesh.vertices[i][a] != lib_mesh.vertices[i][a]) {
                        vdiff++;
                        break;
                    }
            for (size_t i = 0; i < mesh.tris.size(); i++)
                if (mesh.tris[i].v != lib_mesh.tris[i].v) tdiff++;
        }

        // per-point calls from the threads against one batched call on the same points
        const auto ev = sr.getEvaluator();
        const size_t np = particles.size();
        std::vector<float> q(3 * np);
        for (size_t i = 0; i < np; i++)
            for (int a = 0; a < 3; a++) q[3 * i + a] = particles[i][a] + (a == 0 ? 0.3f : -0.2f) * r;
        std::vector<float> fv(np), fg(np);
        std::vector<Vector3f> gv(np);
        std::vector<size_t> nbn(np);
        const auto grid = sr.getHashGrid();
#pragma omp parallel for schedule(dynamic, 16)
        for (long i = 0; i < (long)np; i++) {
            const Vector3f p(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
            ev->SingleEval(p, fv[i]);
            ev->SingleEvalWithGrad(p, fg[i], gv[i]);
            if (i % 16 == 0) {
                std::vector<int> nb;
                grid->GetPIdxList(p, nb);
                nbn[i] = nb.size();
            }
        }
        std::vector<double> bf(np), bg(3 * np);
        ev->BatchEval(q.data(), (int64_t)np, bf.data(), bg.data());
        long pdiff = 0;
        for (size_t i = 0; i < np; i++) {
            bool ok = fv[i] == (float)bf[i] && fg[i] == (float)bf[i];
            for (int a = 0; a < 3; a++) ok = ok && gv[i][a] == (float)bg[3 * i + a];
            if (!ok) pdiff++;
        }
        std::vector<int64_t> off(np + 1);
        int64_t tot = 0;
        {
            std::lock_guard<std::mutex> lk(sr.device()->mutex());
            sr.device()->check(pisurf_neighbors(sr.device()->get(), q.data(), (int64_t)np, off.data(), nullptr, 0, &tot),
                               "neighbors");
        }
        long ndiff = 0;
        for (size_t i = 0; i < np; i += 16)
            if ((int64_t)nbn[i] != off[i + 1] - off[i]) ndiff++;
        std::printf("{\"threads\": %d, \"state\": %d, \"nodes\": %ld, \"leaves\": %ld, \"dual_grids\": %zu, "
                    "\"verts\": %zu, \"tris\": %zu, \"lib_verts\": %zu, \"lib_tris\": %zu, \"vert_diff\": %ld, "
                    "\"tri_diff\": %ld, \"points\": %zu, \"point_diff\": %ld, \"neighbor_diff\": %ld}\n",
                    threads, sr.getSTATE(), nodes, leaves, grids.size(), mesh.vertices.size(), mesh.tris.size(),
                    lib_mesh.vertices.size(), lib_mesh.tris.size(), same_size ? vdiff : -1L, same_size ? tdiff : -1L,
                    np, pdiff, ndiff);
    } catch (const pisurf::Error& e) {
        std::fprintf(stderr, "pisurf error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
