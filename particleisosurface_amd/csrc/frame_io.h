// Input front-end of the reference's frame loop program: controlData.json (src/main.cpp:116-213), CSV
// particle tables (src/main.cpp:215-279) and ShonDy HDF5 frames (src/main.cpp:281-328,
// src/hdf5Utils.cpp:50-77), plus raw float32 triples. Header-only C++17 without dependencies: the
// JSON subset is parsed here, and libhdf5 is loaded at run time (dlopen) only when an .h5 frame is
// read, so everything builds and runs where HDF5 is absent.
#pragma once
#include <dlfcn.h>

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <regex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace pisurf {
namespace io {

struct IoError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// The frame-loop globals controlData.json sets, with the reference's defaults (src/main.cpp:33-55).
struct Config {
    bool use_ani = true, use_ours = true, use_poly6 = false, calc_p_normal = true, gen_splash = true;
    bool need_record = false;
    std::string output_type = "ply", suffix;
    std::vector<std::string> data_pathes;  // frame file names, relative to the data directory
    bool is_const_radius = false;          // RADIUS present
    float radius = 0, smooth_factor = 2.0f, iso_factor = 1.9f, iso_value = 0.0f;
};

// ---- a flat JSON object of scalars and strings (all controlData.json holds) ----
struct JsonValue {
    enum Kind { NUM, STR, BOOL, NUL } kind = NUL;
    double num = 0;
    bool b = false;
    std::string str;
};

inline std::map<std::string, JsonValue> parse_flat_json(const std::string& text) {
    std::map<std::string, JsonValue> out;
    size_t i = 0;
    auto ws = [&] {
        while (i < text.size() && std::isspace((unsigned char)text[i])) i++;
    };
    auto fail = [](const std::string& what) { throw IoError("controlData.json: " + what); };
    auto string_at = [&]() {
        if (i >= text.size() || text[i] != '"') fail("expected a string");
        std::string s;
        for (i++; i < text.size() && text[i] != '"'; i++) {
            if (text[i] == '\\' && i + 1 < text.size()) {
                const char e = text[++i];
                s += e == 'n' ? '\n' : e == 't' ? '\t' : e;
            } else {
                s += text[i];
            }
        }
        if (i >= text.size()) fail("unterminated string");
        i++;
        return s;
    };
    ws();
    if (i >= text.size() || text[i] != '{') fail("expected an object");
    i++;
    ws();
    if (i < text.size() && text[i] == '}') return out;
    while (true) {
        ws();
        const std::string key = string_at();
        ws();
        if (i >= text.size() || text[i] != ':') fail("expected ':'");
        i++;
        ws();
        JsonValue v;
        if (i < text.size() && text[i] == '"') {
            v.kind = JsonValue::STR;
            v.str = string_at();
        } else if (text.compare(i, 4, "true") == 0) {
            v.kind = JsonValue::BOOL;
            v.b = true;
            i += 4;
        } else if (text.compare(i, 5, "false") == 0) {
            v.kind = JsonValue::BOOL;
            i += 5;
        } else if (text.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            const char* p = text.c_str() + i;
            char* end = nullptr;
            v.num = std::strtod(p, &end);
            if (end == p) fail("unsupported value for " + key + " (the frame loop reads scalars and strings only)");
            v.kind = JsonValue::NUM;
            i += (size_t)(end - p);
        }
        out[key] = v;
        ws();
        if (i < text.size() && text[i] == ',') {
            i++;
            continue;
        }
        if (i < text.size() && text[i] == '}') break;
        fail("expected ',' or '}'");
    }
    return out;
}

// parseString (src/utils.cpp:14-43): split at any of the separator characters; a trailing empty
// field is dropped
inline std::vector<std::string> parse_string(const std::string& input, const std::string& sep) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : input) {
        if (sep.find(ch) != std::string::npos) {
            out.push_back(cur);
            cur.clear();
        } else {
            cur += ch;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

// regular files of `dir` whose name ends in `suffix`, sorted by name
inline std::vector<std::string> list_dir_suffix(const std::string& dir, const std::string& suffix) {
    std::vector<std::string> names;
    for (const auto& e : std::filesystem::directory_iterator(dir)) {
        if (!e.is_regular_file()) continue;
        const std::string n = e.path().filename().string();
        if (n.size() >= suffix.size() && n.compare(n.size() - suffix.size(), suffix.size(), suffix) == 0)
            names.push_back(n);
    }
    std::sort(names.begin(), names.end());
    return names;
}

// loadConfigJson (src/main.cpp:116-213). Flags accept true/false or numbers, as nlohmann's
// conversions do. With SUFFIX set the reference lists the directory on Windows only; here it is
// listed on every platform, in name order.
inline Config load_config_json(const std::string& dir) {
    std::ifstream in(dir + "/controlData.json");
    if (!in.good()) throw IoError("cannot find ConfigJson file: " + dir + "/controlData.json");
    std::stringstream ss;
    ss << in.rdbuf();
    const auto j = parse_flat_json(ss.str());
    Config c;
    auto flag = [&](const char* k, bool& dst) {
        auto it = j.find(k);
        if (it == j.end()) return;
        if (it->second.kind == JsonValue::BOOL) dst = it->second.b;
        else if (it->second.kind == JsonValue::NUM) dst = it->second.num != 0;
        else throw IoError(std::string("controlData.json: ") + k + " must be a boolean");
    };
    auto num = [&](const char* k, float& dst) -> bool {
        auto it = j.find(k);
        if (it == j.end()) return false;
        if (it->second.kind != JsonValue::NUM) throw IoError(std::string("controlData.json: ") + k + " must be a number");
        dst = (float)it->second.num;
        return true;
    };
    auto str = [&](const char* k, std::string& dst) -> bool {
        auto it = j.find(k);
        if (it == j.end()) return false;
        if (it->second.kind != JsonValue::STR) throw IoError(std::string("controlData.json: ") + k + " must be a string");
        dst = it->second.str;
        return true;
    };
    flag("USE_ANI", c.use_ani);
    flag("USE_OURS", c.use_ours);
    flag("USE_POLY6", c.use_poly6);
    flag("CALC_P_NORMAL", c.calc_p_normal);
    flag("GEN_SPLASH", c.gen_splash);
    flag("NEED_RECORD", c.need_record);
    str("OUTPUT_TYPE", c.output_type);
    str("SUFFIX", c.suffix);
    if (c.suffix.empty()) {
        std::string files;
        if (!str("DATA_FILE", files)) throw IoError("controlData.json: DATA_FILE is required when SUFFIX is empty");
        c.data_pathes = parse_string(files, ",");
    } else {
        c.data_pathes = list_dir_suffix(dir, c.suffix);
    }
    c.is_const_radius = num("RADIUS", c.radius);
    num("SMOOTH_FACTOR", c.smooth_factor);
    num("ISO_FACTOR", c.iso_factor);
    num("ISO_VALUE", c.iso_value);
    return c;
}

// ---- CSV (loadParticlesFromCSV, src/main.cpp:215-279) ----
// The header's quotes are removed and the columns found by the reference's patterns (case
// insensitive): radius "^((.*)radius|r)$", axes "^(.*(:|_|-)( *)(x|0)(\"*)|x)$" and the same with
// y|1 and z|2; the first matching column wins. Rows are read until the first empty line.
// want_radius: read the radius column (the reference does when the radius is not constant).
inline void load_particles_csv(const std::string& path, std::vector<float>& xyz, std::vector<float>& radiuses,
                               bool want_radius) {
    std::ifstream in(path);
    if (!in.good()) throw IoError("cannot open CSV file: " + path);
    std::string line;
    std::getline(in, line);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    line.erase(std::remove(line.begin(), line.end(), '"'), line.end());
    const std::vector<std::string> titles = parse_string(line, ",");
    auto find = [&](const char* pattern) -> size_t {
        const std::regex re(pattern, std::regex::ECMAScript | std::regex::icase);
        for (size_t i = 0; i < titles.size(); i++)
            if (std::regex_match(titles[i], re)) return i;
        return titles.size();
    };
    const size_t xi = find("^(.*(:|_|-)( *)(x|0)(\"*)|x)$");
    const size_t yi = find("^(.*(:|_|-)( *)(y|1)(\"*)|y)$");
    const size_t zi = find("^(.*(:|_|-)( *)(z|2)(\"*)|z)$");
    const size_t ri = want_radius ? find("^((.*)radius|r)$") : titles.size();
    if (xi == titles.size() || yi == titles.size() || zi == titles.size())
        throw IoError("failed to locate axis column in " + path);
    if (want_radius && ri == titles.size()) throw IoError("failed to locate radius column in " + path);
    xyz.clear();
    radiuses.clear();
    size_t row = 1;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) break;
        row++;
        const std::vector<std::string> el = parse_string(line, ",");
        const size_t need = std::max({xi, yi, zi, want_radius ? ri : 0}) + 1;
        if (el.size() < need) throw IoError(path + ": row " + std::to_string(row) + " has too few fields");
        try {
            xyz.push_back(std::stof(el[xi]));
            xyz.push_back(std::stof(el[yi]));
            xyz.push_back(std::stof(el[zi]));
            if (want_radius) radiuses.push_back(std::stof(el[ri]));
        } catch (const std::exception&) {
            throw IoError(path + ": row " + std::to_string(row) + " has a non-numeric field");
        }
    }
}

// ---- ShonDy HDF5 (readShonDyParticleData, src/main.cpp:281-328) ----
// Datasets "position" (flat float64[3N]; N = dims[0] / 3) and "particleRadius" (float64[N]),
// narrowed to float. libhdf5 (1.10 ABI: 64-bit hid_t) is opened by dlopen from
// $PISURF_HDF5_LIB, else "libhdf5.so". The file is opened read-only (the reference asks for
// read-write access and so refuses read-only files; nothing is written).
class Hdf5 {
  public:
    Hdf5() {
        const char* env = std::getenv("PISURF_HDF5_LIB");
        h_ = dlopen(env && *env ? env : "libhdf5.so", RTLD_NOW | RTLD_LOCAL);
        if (!h_) throw IoError(std::string("HDF5 library not available: ") + dlerror());
        sym(open_, "H5open");
        sym(fopen_, "H5Fopen");
        sym(fclose_, "H5Fclose");
        sym(dopen_, "H5Dopen2");
        sym(dclose_, "H5Dclose");
        sym(dspace_, "H5Dget_space");
        sym(sndims_, "H5Sget_simple_extent_ndims");
        sym(sdims_, "H5Sget_simple_extent_dims");
        sym(sclose_, "H5Sclose");
        sym(dread_, "H5Dread");
        if (open_() < 0) throw IoError("H5open failed");
        native_double_ = (const int64_t*)dlsym(h_, "H5T_NATIVE_DOUBLE_g");
        if (!native_double_) throw IoError("HDF5 library lacks H5T_NATIVE_DOUBLE_g");
    }
    ~Hdf5() {
        if (h_) dlclose(h_);
    }
    Hdf5(const Hdf5&) = delete;
    Hdf5& operator=(const Hdf5&) = delete;

    // one dataset as float64 (size(), readDouble: src/hdf5Utils.cpp:59-77)
    std::vector<double> read_doubles(int64_t file, const std::string& name, bool flat_len_only = false) {
        const int64_t ds = dopen_(file, name.c_str(), 0);
        if (ds < 0) throw IoError("missing dataset \"" + name + "\"");
        const int64_t sp = dspace_(ds);
        const int nd = sp >= 0 ? sndims_(sp) : -1;
        std::vector<unsigned long long> dims(nd > 0 ? nd : 1, 0);
        if (nd < 1 || sdims_(sp, dims.data(), nullptr) < 0) {
            if (sp >= 0) sclose_(sp);
            dclose_(ds);
            throw IoError("bad dataspace of \"" + name + "\"");
        }
        sclose_(sp);
        unsigned long long total = 1;
        for (auto d : dims) total *= d;
        (void)flat_len_only;
        std::vector<double> v(total);
        const int rc = total ? dread_(ds, *native_double_, 0, 0, 0, v.data()) : 0;
        dclose_(ds);
        if (rc < 0) throw IoError("cannot read \"" + name + "\" as float64");
        return v;
    }
    int64_t open_file(const std::string& path) {
        const int64_t f = fopen_(path.c_str(), 0u /*H5F_ACC_RDONLY*/, 0 /*H5P_DEFAULT*/);
        if (f < 0) throw IoError("cannot read hdf5 file: " + path);
        return f;
    }
    void close_file(int64_t f) { fclose_(f); }

  private:
    template <class F> void sym(F& f, const char* name) {
        f = (F)dlsym(h_, name);
        if (!f) throw IoError(std::string("HDF5 library lacks ") + name);
    }
    void* h_ = nullptr;
    int (*open_)() = nullptr;
    int64_t (*fopen_)(const char*, unsigned, int64_t) = nullptr;
    int (*fclose_)(int64_t) = nullptr;
    int64_t (*dopen_)(int64_t, const char*, int64_t) = nullptr;
    int (*dclose_)(int64_t) = nullptr;
    int64_t (*dspace_)(int64_t) = nullptr;
    int (*sndims_)(int64_t) = nullptr;
    int (*sdims_)(int64_t, unsigned long long*, unsigned long long*) = nullptr;
    int (*sclose_)(int64_t) = nullptr;
    int (*dread_)(int64_t, int64_t, int64_t, int64_t, int64_t, void*) = nullptr;
    const int64_t* native_double_ = nullptr;
};

inline void load_particles_h5(const std::string& path, std::vector<float>& xyz, std::vector<float>& radiuses,
                              bool want_radius) {
    if (!std::filesystem::exists(path)) throw IoError("cannot find hdf5 file: " + path);
    Hdf5 h5;
    const int64_t f = h5.open_file(path);
    try {
        const std::vector<double> pos = h5.read_doubles(f, "position");
        const size_t n = pos.size() / 3;  // N = dims[0] / 3 on the flat array
        xyz.resize(3 * n);
        for (size_t i = 0; i < 3 * n; i++) xyz[i] = (float)pos[i];
        radiuses.clear();
        if (want_radius) {
            const std::vector<double> rd = h5.read_doubles(f, "particleRadius");
            if (rd.size() < n) throw IoError("particleRadius shorter than the particle count");
            radiuses.resize(n);
            for (size_t i = 0; i < n; i++) radiuses[i] = (float)rd[i];
        }
    } catch (...) {
        h5.close_file(f);
        throw;
    }
    h5.close_file(f);
}

// raw little-endian float32 x, y, z triples (no header)
inline void load_particles_raw(const std::string& path, std::vector<float>& xyz) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw IoError("cannot open " + path);
    xyz.clear();
    float t[3];
    while (std::fread(t, sizeof(float), 3, f) == 3) xyz.insert(xyz.end(), t, t + 3);
    std::fclose(f);
}

// One frame by its suffix (runOurs, src/main.cpp:346-356): ".csv", ".h5", else raw float32.
// A variable-radius frame whose radii agree within 1e-7 becomes constant-radius
// (src/main.cpp:358-366); returns the radius to use, or 0 when the radii truly vary.
inline float load_frame(const std::string& path, const Config& cfg, std::vector<float>& xyz) {
    std::vector<float> rad;
    const bool want_r = !cfg.is_const_radius;
    const std::string ext = std::filesystem::path(path).extension().string();
    if (ext == ".csv") load_particles_csv(path, xyz, rad, want_r);
    else if (ext == ".h5") load_particles_h5(path, xyz, rad, want_r);
    else load_particles_raw(path, xyz);
    if (cfg.is_const_radius) return cfg.radius;
    if (rad.empty()) throw IoError(path + ": no radius column and no RADIUS in controlData.json");
    const auto mm = std::minmax_element(rad.begin(), rad.end());
    if (std::abs(*mm.second - *mm.first) < 1e-7f) return rad[0];
    return 0.0f;
}

}  // namespace io
}  // namespace pisurf
