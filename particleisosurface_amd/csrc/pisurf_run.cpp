// Frame driver over a data directory, as the reference's runOurs (src/main.cpp:330-405) with
// loadConfigJson (src/main.cpp:116-213): reads <data_dir>/controlData.json, loads every frame
// (.csv, .h5, or raw float32), meshes all frames in one pisurf_run_frames call and
// writes <out_dir>/<frame name>.obj or .ply (OUTPUT_TYPE).
//   pisurf_run <data_dir> [out_dir] [--precision 0|1|2] [--load-only]
// out_dir defaults to <data_dir>/out (src/main.cpp:590). --load-only parses the inputs and prints
// a JSON summary without touching a GPU. Variable radii that really vary are refused: the
// reference's variable-radius field is disabled (it evaluates to the iso value everywhere).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "frame_io.h"
#include "host_facade.h"

namespace {

struct Ctx {
    std::string out_dir, type;
    std::vector<std::string> names;
    std::string err;
    long long verts = 0, tris = 0;
};

void on_mesh(void* user, int32_t k, const pisurf_tree_info*, const double* v, int64_t nv, const int32_t* t,
             int64_t nt) {
    Ctx& c = *static_cast<Ctx*>(user);
    if (!c.err.empty()) return;
    pisurf::Mesh m;
    m.vertices.resize(nv);
    m.tris.resize(nt);
    for (int64_t i = 0; i < nv; i++) m.vertices[i] = pisurf::Vector3f((float)v[3 * i], (float)v[3 * i + 1], (float)v[3 * i + 2]);
    for (int64_t i = 0; i < nt; i++) m.tris[i] = pisurf::Triangle(t[3 * i], t[3 * i + 1], t[3 * i + 2]);
    m.verticesNum = (unsigned)nv;
    m.trianglesNum = (unsigned)nt;
    const std::string& f = c.names[k];
    const std::string base = c.out_dir + "/" + f.substr(0, f.find_last_of('.'));
    try {
        if (c.type == "obj" || c.type == "OBJ") pisurf::writeObjFile(m, base + ".obj");
        else pisurf::writePlyFile(m, base + ".ply");
    } catch (const std::exception& e) {
        c.err = e.what();
    }
    c.verts += nv;
    c.tris += nt;
    std::printf("frame %d %s: %lld vertices, %lld triangles\n", k, f.c_str(), (long long)nv, (long long)nt);
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> pos;
    int precision = PISURF_FP32;
    bool load_only = false;
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "--load-only")) load_only = true;
        else if (!std::strcmp(argv[i], "--precision") && i + 1 < argc) precision = std::atoi(argv[++i]);
        else pos.push_back(argv[i]);
    }
    if (pos.empty() || pos.size() > 2) {
        std::fprintf(stderr, "usage: %s <data_dir> [out_dir] [--precision 0|1|2] [--load-only]\n", argv[0]);
        return 2;
    }
    const std::string dir = pos[0];
    const std::string out_dir = pos.size() > 1 ? pos[1] : dir + "/out";
    try {
        const pisurf::io::Config cfg = pisurf::io::load_config_json(dir);
        if (cfg.data_pathes.empty()) throw pisurf::io::IoError("no frames listed");
        std::vector<std::vector<float>> frames(cfg.data_pathes.size());
        std::vector<float> radius(frames.size());
        for (size_t k = 0; k < frames.size(); k++) {
            radius[k] = pisurf::io::load_frame(dir + "/" + cfg.data_pathes[k], cfg, frames[k]);
            if (!(radius[k] > 0))
                throw pisurf::io::IoError(cfg.data_pathes[k] + ": particle radii vary; only constant radius is supported");
        }
        if (load_only) {
            std::printf("{\"frames\": [");
            for (size_t k = 0; k < frames.size(); k++) {
                const auto& f = frames[k];
                double s[3] = {0, 0, 0};
                for (size_t i = 0; i < f.size(); i++) s[i % 3] += f[i];
                std::printf("%s{\"name\": \"%s\", \"n\": %zu, \"radius\": %.9g, \"sum\": [%.9g, %.9g, %.9g]}",
                            k ? ", " : "", cfg.data_pathes[k].c_str(), f.size() / 3, (double)radius[k], s[0], s[1], s[2]);
            }
            std::printf("], \"output_type\": \"%s\", \"use_ani\": %d, \"use_poly6\": %d, \"calc_p_normal\": %d, "
                        "\"smooth_factor\": %.9g, \"iso_factor\": %.9g}\n",
                        cfg.output_type.c_str(), (int)cfg.use_ani, (int)cfg.use_poly6, (int)cfg.calc_p_normal,
                        (double)cfg.smooth_factor, (double)cfg.iso_factor);
            return 0;
        }
        std::filesystem::create_directories(out_dir);
        pisurf::Device dev(0);
        pisurf_params& p = dev.params();
        p.radius = radius[0];
        p.use_ani = cfg.use_ani;
        p.use_poly6 = cfg.use_poly6;
        p.calc_normals = cfg.calc_p_normal;
        p.smooth_factor = cfg.smooth_factor;
        p.iso_factor = cfg.iso_factor;
        p.precision = precision;
        dev.apply();
        std::vector<const float*> ptr(frames.size());
        std::vector<int64_t> n(frames.size());
        for (size_t k = 0; k < frames.size(); k++) {
            ptr[k] = frames[k].data();
            n[k] = (int64_t)frames[k].size() / 3;
        }
        Ctx c{out_dir, cfg.output_type, cfg.data_pathes, "", 0, 0};
        dev.check(pisurf_run_frames(dev.get(), (int32_t)frames.size(), ptr.data(), n.data(), radius.data(), 0, on_mesh,
                                    &c, nullptr),
                  "run_frames");
        if (!c.err.empty()) throw pisurf::io::IoError(c.err);
        std::printf("{\"frames\": %zu, \"vertices\": %lld, \"triangles\": %lld}\n", frames.size(), c.verts, c.tris);
    } catch (const pisurf::Error& e) {
        std::fprintf(stderr, "pisurf error %d: %s\n", e.code, e.what());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
