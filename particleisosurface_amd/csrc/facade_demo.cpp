// Drives one frame through the C++ mirror of the reference classes (host_facade.h), the way the
// reference's runOurs (src/main.cpp:330-405) drives SurfReconstructor, and writes an OBJ.
//   pisurf_facade_demo <particles.f32> <radius> <out.obj> [precision 0|1|2] [mesh_order 0|1] [vti_dir]
// With vti_dir, UniformGrid (the reference's runUniform path) also writes <vti_dir>/grid.vti.
// particles.f32: raw little-endian float32 x,y,z triples.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "host_facade.h"

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s particles.f32 radius out.obj [precision] [mesh_order]\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        return 2;
    }
    std::vector<pisurf::Vector3f> particles;
    float xyz[3];
    while (std::fread(xyz, sizeof(float), 3, f) == 3) particles.emplace_back(xyz[0], xyz[1], xyz[2]);
    std::fclose(f);
    const float r = std::strtof(argv[2], nullptr);
    const int precision = argc > 4 ? std::atoi(argv[4]) : PISURF_FP32;
    const int order = argc > 5 ? std::atoi(argv[5]) : 0;
    try {
        std::vector<float> radiuses;
        pisurf::Mesh mesh;
        pisurf::SurfReconstructor sr(particles, radiuses, &mesh, r);
        sr.device()->params().mesh_order = order;
        sr.Run(1.9f, 2.0f, precision);
        pisurf::writeObjFile(mesh, argv[3]);
        // the per-point interfaces on the same device state
        float bb[6] = {1e30f, -1e30f, 1e30f, -1e30f, 1e30f, -1e30f};
        for (const auto& p : particles)
            for (int a = 0; a < 3; a++) {
                bb[2 * a] = std::min(bb[2 * a], p[a]);
                bb[2 * a + 1] = std::max(bb[2 * a + 1], p[a]);
            }
        auto grid = std::make_shared<pisurf::HashGrid>(sr.device(), sr.getGlobalParticles(), bb, r, 4.0f);
        std::shared_ptr<pisurf::MultiLevelSearcher> searcher;  // constant radius: none, as in the reference
        pisurf::Evaluator ev(grid, searcher, sr.getGlobalParticles(), &radiuses, r);
        ev.compute_Gs_xMeans();
        std::vector<int> nb;
        grid->GetPIdxList(particles[0], nb);
        float f0 = 0;
        ev.SingleEval(particles[0], f0);
        int gdims[3] = {0, 0, 0};
        if (argc > 6) {
            pisurf::UniformGrid ug(particles, r);
            ug.Run(0.0f, "grid", argv[6], precision);
            for (int a = 0; a < 3; a++) gdims[a] = ug.dims[a];
        }
        std::printf("{\"grid_dims\": [%d, %d, %d], ", gdims[0], gdims[1], gdims[2]);
        std::printf("\"verts\": %u, \"tris\": %u, \"depth_max\": %d, \"depth_min\": %d, \"cells\": %llu, "
                    "\"neighbors0\": %zu, \"f0\": %.9g, \"iso\": %.9g}\n",
                    mesh.verticesNum, mesh.trianglesNum, sr.getDepthMax(), sr.getDepthMin(), grid->CellNum,
                    nb.size(), (double)f0, (double)ev.getIsoValue());
    } catch (const pisurf::Error& e) {
        std::fprintf(stderr, "pisurf error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
