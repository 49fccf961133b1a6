// project.hpp -- project.json scenes for facet_engine: the reference's
// project.json -> facet scene path (facetInputView and loadProjectScene,
// models/histream/src/facetrt/facetrtcore.cpp:1888-2043, 2553-3025, "core",
// and ProjectJson, models/histream/src/base/projectjson.cpp).
//
// Built: objects read from OBJ files (or a box of the object's dimensions),
// placed from their position files (x, y, z, scale, rotation) on the
// terrain, clipped to the scene box and appended object by object; a ground
// grid of voxel-size cells, flat or on a DEM (ESRI ASCII grid or baseline
// uncompressed GeoTIFF) stretched over the scene; the visibility raster's
// size and layers (large-scene rule, adaptive layers); one material per
// object mesh from the project's spectra (custom spectra over the sensor
// bands), thermal, surface-energy and soil entries; the FacetEB settings the
// project's control, light and meteo blocks give (photovoltaic
// surface-energy entries included).  Not read: reflectance textures;
// PROSPECT, BSM and file spectra are replaced by their entry's own listed
// values taken as a custom spectrum (stated on stderr).  Fire, Fog and
// medium objects are skipped as the reference does.
#pragma once

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <limits>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "image.hpp"
#include "json.hpp"
#include "scene.hpp"

namespace facetcli {

constexpr uint32_t kEnergySamples = 2001;  // 1-nm samples from 400 nm (core N1)

// Reflectance-variance texture (ReflectanceVarianceTexture and
// loadReflectanceVarianceTexture, core:88-111, 623-702): the image's first
// three bands (one band: grey) read as sRGB, converted to linear luminance
// 0.2126 R + 0.7152 G + 0.0722 B, the image mean removed and the values
// divided by the largest |deviation|; sampled bilinearly, wrapping in both
// directions.
struct VarianceTexture {
    int width = 0, height = 0;
    std::vector<float> values;

    float sample(float u, float v) const {
        if (width <= 0 || height <= 0 || values.empty()) return 0.0f;
        const float x = (u - std::floor(u)) * float(width), y = (v - std::floor(v)) * float(height);
        const int x0 = int(std::floor(x)) % width, y0 = int(std::floor(y)) % height;
        const int x1 = (x0 + 1) % width, y1 = (y0 + 1) % height;
        const float fx = x - std::floor(x), fy = y - std::floor(y);
        const auto at = [&](int c, int r) { return values[size_t(r) * size_t(width) + size_t(c)]; };
        const float top = at(x0, y0) * (1.0f - fx) + at(x1, y0) * fx;
        const float bottom = at(x0, y1) * (1.0f - fx) + at(x1, y1) * fx;
        return top * (1.0f - fy) + bottom * fy;
    }
};

inline float srgbToLinear(float v) {
    v = std::clamp(v, 0.0f, 1.0f);
    return v <= 0.04045f ? v / 12.92f : std::pow((v + 0.055f) / 1.055f, 2.4f);
}

inline std::shared_ptr<const VarianceTexture> loadVarianceTexture(const std::filesystem::path& path) {
    static std::map<std::string, std::weak_ptr<const VarianceTexture>> cache;
    const std::string key = std::filesystem::absolute(path).lexically_normal().string();
    if (auto it = cache.find(key); it != cache.end())
        if (auto t = it->second.lock()) return t;
    facetimg::Image img;
    try {
        img = facetimg::readImage(key);
    } catch (const std::exception& e) {
        throw std::runtime_error("Cannot open reflectance variance texture: " + key + " (" + e.what() + ")");
    }
    if (img.width < 2 || img.height < 2)
        throw std::runtime_error("Reflectance variance texture must contain at least 2 x 2 pixels: " + key);
    auto t = std::make_shared<VarianceTexture>();
    t->width = img.width;
    t->height = img.height;
    const size_t n = size_t(img.width) * size_t(img.height);
    const int ch = std::min(3, img.channels);
    t->values.resize(n);
    double mean = 0.0;
    for (size_t p = 0; p < n; ++p) {
        const uint8_t* px = &img.data[p * size_t(img.channels)];
        const float red = srgbToLinear(px[0] / 255.0f);
        const float green = ch > 1 ? srgbToLinear(px[1] / 255.0f) : red;
        const float blue = ch > 2 ? srgbToLinear(px[2] / 255.0f) : red;
        t->values[p] = 0.2126f * red + 0.7152f * green + 0.0722f * blue;
        mean += t->values[p];
    }
    mean /= double(n);
    float largest = 0.0f;
    for (float& v : t->values) {
        v -= float(mean);
        largest = std::max(largest, std::abs(v));
    }
    if (!(largest > 1.0e-6f)) throw std::runtime_error("Reflectance variance texture is effectively constant: " + key);
    for (float& v : t->values) v /= largest;
    cache[key] = t;
    return t;
}

// One material of a project scene: the fields of the reference's
// SurfaceParameters (core:113-143) this engine uses, with its defaults.
struct ProjectMaterial {
    std::string label;                    // spectral / thermal / energy names
    std::vector<float> energyR, energyT;  // per 1-nm sample from 400 nm; empty: r, t
    std::vector<float> bandR, bandT;      // per sensor band (optical bands); empty: r, t
    float r = 0.10f, t = 0.05f;
    float tirR = 0.03f, tirT = 0.0f;
    float sunlitT = 300.0f, shadedT = 296.0f;
    float rss = 120.0f, cs = 1180.0f, rhos = 1800.0f, lambdas = 1.55f, smc = 0.25f;
    float heatCapacity = 60000.0f, convectiveScale = 1.0f;
    int soilMethod = 1;
    bool vegetation = true;
    bool pv = false;  // energyModel "photovoltaic"
    float pvEta25 = 0.22f, pvGamma = -0.0035f, pvBifaciality = 0.0f;
    uint8_t component = 0;  // 0 unknown, 1 soil, 2 vegetation, 3 building, 4 photovoltaic, 5 water
    // reflectance-variance texture of the spectral entry (physicalTexture)
    std::shared_ptr<const VarianceTexture> texture;
    float textureStrength = 0.0f, textureRepeat = 1.0f;

    // shortwave reflectance / transmittance of an energy window whose centre
    // is the 1-nm sample k (makeEnergyShortwaveOptics, core:952-1003, with a
    // Lambertian background)
    float swR(size_t k, bool leaf) const {
        const float fallback = leaf ? 0.10f : 0.20f;
        const float v = energyR.empty() ? r : (k < energyR.size() ? energyR[k] : energyR.back());
        return std::clamp(std::isfinite(v) ? v : fallback, 0.0f, 1.0f);
    }
    float swT(size_t k, bool leaf, float refl) const {
        const float fallback = leaf ? 0.05f : 0.0f;
        const float v = energyT.empty() ? t : (k < energyT.size() ? energyT[k] : energyT.back());
        return std::clamp(std::isfinite(v) ? v : fallback, 0.0f, 1.0f - refl);
    }
};

// The FacetEB settings of a project (the tags facetInputView writes,
// core:1903-2040, read back by the run with the clamps of core:3193-3240).
struct ProjectRun {
    int startNode = 0, endNode = 1;
    float dTime = 1800.0f, latitude = 40.0f, longitude = 116.0f;
    std::string meteoFile;  // empty: the built-in defaults per node
    float skyTemperature = 250.0f, directRatio = 0.8f;
    float tolerance = 0.05f, relaxation = 0.5f;
    uint32_t couplingIterations = 20, spectralWidth = 100;
    int vegetationMethod = 0;
    bool saveEnergy = false, saveRadiation = false;
    float sunZenith = 30.0f, sunAzimuth = 135.0f;
    bool acceleratedSolver = false;  // control.radiationSolver == "accelerated" (fileio_json.cpp:208)
    uint32_t periodicNeighbors = 0;
    std::vector<float> bands;  // sensor bands, nm
    std::string outDir;
};

struct Project {
    Scene scene;
    std::vector<ProjectMaterial> materials;  // objects' meshes, then the ground
    std::vector<uint8_t> facetComponent;     // output component per facet
    ProjectRun run;
    std::string mode;
};

namespace project_detail {

inline std::string lower(std::string s) {
    for (char& c : s) c = char(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

inline std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return {};
    return s.substr(a, s.find_last_not_of(" \t\r\n") - a + 1);
}

// comma-separated names (nameList in the reference's XML reader)
inline std::vector<std::string> names(const std::string& joined) {
    std::vector<std::string> out;
    std::stringstream ss(joined);
    std::string tok;
    while (std::getline(ss, tok, ',')) out.push_back(trim(tok));
    return out;
}

// joinedNames (core:1872-1884): the meshes' values of key, or the fallback
inline std::string joined(const json::Value& object, const char* key, const std::string& fallback) {
    std::string out;
    const json::Value& meshes = json::member(object, "meshes", json::Value::Array);
    for (size_t i = 0; i < meshes.items.size(); ++i) {
        if (i) out += ',';
        out += json::str(meshes.items[i], key, fallback);
    }
    return out.empty() ? fallback : out;
}

// ProjectJson::resolve (projectjson.cpp:89-102): absolute paths as given,
// others relative to the project folder.  A path that does not exist there
// (one written on another system, e.g. "C:\\work\\...\\tree.obj") is looked
// up by its file name in the project folder and its asset folders.
inline std::filesystem::path asset(const std::filesystem::path& dir, const json::Value& root,
                                   const std::string& value) {
    namespace fs = std::filesystem;
    if (value.empty()) return {};
    const fs::path p(value);
    const fs::path direct = p.is_absolute() ? p.lexically_normal() : (dir / p).lexically_normal();
    std::error_code ec;
    if (fs::exists(direct, ec)) return direct;
    const size_t cut = value.find_last_of("/\\");
    const std::string name = cut == std::string::npos ? value : value.substr(cut + 1);
    std::vector<fs::path> dirs{dir};
    const json::Value& folders = json::member(root, "folders", json::Value::Object);
    for (const auto& m : folders.members)
        if (m.second.kind == json::Value::String) dirs.push_back(dir / m.second.s);
    for (const char* sub : {"models", "positions", "spectral", "meteorology", "terrain"}) dirs.push_back(dir / sub);
    for (const fs::path& d : dirs)
        if (fs::exists(d / name, ec)) return (d / name).lexically_normal();
    return direct;
}

// sensorBands (projectjson.cpp:189-213): custom bands plus the continuous
// range when enabled, sorted, unique, at most 800; 10500 nm when empty
inline std::vector<float> sensorBands(const json::Value& sensor) {
    std::vector<float> values;
    std::set<long long> keys;
    const auto add = [&](float v) {
        if (!(v > 0.0f) || !std::isfinite(v)) return;
        if (keys.insert(std::llround(double(v) * 1000000.0)).second) values.push_back(v);
    };
    if (json::boolean(sensor, "continuousBands", false)) {
        const float a = json::num(sensor, "bandStart", 400.0f), b = json::num(sensor, "bandEnd", 2500.0f);
        const float step = json::num(sensor, "bandStep", 10.0f);
        if (a > 0.0f && b >= a && step > 0.0f)
            for (int i = 0; i < 800 && a + float(i) * step <= b + 1e-5f; ++i) add(a + float(i) * step);
    }
    for (float v : json::numbers(sensor.get("bands"))) add(v);
    if (values.empty()) values.push_back(10500.0f);
    std::sort(values.begin(), values.end());
    if (values.size() > 800) values.resize(800);
    return values;
}

inline float nanometres(float w) { return w > 0.0f && w <= 50.0f ? w * 1000.0f : w; }

// interpolateValue (compo.cpp:74-88): one value (or a count that does not
// match the wavelengths) is constant; otherwise linear, held at the ends
inline float interpolate(const std::vector<float>& wl, const std::vector<float>& v, float lambda, float fallback) {
    if (v.empty()) return fallback;
    if (v.size() == 1 || wl.size() != v.size()) return v.front();
    const float x = nanometres(lambda);
    if (x <= wl.front()) return v.front();
    if (x >= wl.back()) return v.back();
    const size_t right = size_t(std::upper_bound(wl.begin(), wl.end(), x) - wl.begin());
    const size_t left = right - 1;
    const float span = wl[right] - wl[left];
    const float w = span > 0.0f ? (x - wl[left]) / span : 0.0f;
    return v[left] + (v[right] - v[left]) * w;
}

inline const json::Value* named(const json::Value& list, const std::string& name) {
    for (const json::Value& item : list.items)
        if (item.isObject() && json::str(item, "name") == name) return &item;
    return nullptr;
}

// loadSurfaceParameters (core:2254-2383) for the spectral entry and the
// thermal entry of the given names; a name without an entry leaves the
// material as it is
inline void surfaceParameters(const json::Value& cfg, const std::string& spectralName, const std::string& thermalName,
                              const std::vector<float>& bands, ProjectMaterial& m,
                              const std::function<std::filesystem::path(const std::string&)>& resolve = {}) {
    const json::Value* sp = named(json::member(cfg, "spectra", json::Value::Array), spectralName);
    if (sp) {
        const std::string model = json::str(*sp, "model", "custom");
        const std::string ml = lower(model);
        if (ml == "prospect" || ml == "bsm" || ml == "file")
            std::cerr << "facet_engine: spectrum '" << spectralName << "' uses the " << model
                      << " model, which this engine does not evaluate; its listed values are used\n";
        const json::Value* rv = sp->get("reflectance");
        const json::Value* tv = sp->get("transmittance");
        json::Value rdef, tdef;
        rdef.kind = tdef.kind = json::Value::String;
        rdef.s = "0.2";
        tdef.s = "0";
        const std::vector<float> R = json::numbers(rv ? rv : &rdef), T = json::numbers(tv ? tv : &tdef);
        std::vector<float> wl;
        for (float b : bands) wl.push_back(nanometres(b));
        m.energyR.assign(kEnergySamples, 0.0f);
        m.energyT.assign(kEnergySamples, 0.0f);
        for (uint32_t k = 0; k < kEnergySamples; ++k) {
            const float lambda = 400.0f + float(k);
            const float ri = interpolate(wl, R, lambda, 0.2f);
            const float r = std::clamp(std::isfinite(ri) ? ri : 0.2f, 0.0f, 1.0f);
            const float ti = interpolate(wl, T, lambda, 0.0f);
            m.energyR[k] = r;
            m.energyT[k] = std::min(std::clamp(std::isfinite(ti) ? ti : 0.0f, 0.0f, 1.0f), 1.0f - r);
        }
        // per sensor band (resolveSpectrals over the sensor wavelengths)
        m.bandR.assign(bands.size(), 0.0f);
        m.bandT.assign(bands.size(), 0.0f);
        for (size_t b = 0; b < bands.size(); ++b) {
            const float ri = interpolate(wl, R, bands[b], 0.2f);
            const float r = std::clamp(std::isfinite(ri) ? ri : 0.2f, 0.0f, 1.0f);
            const float ti = interpolate(wl, T, bands[b], 0.0f);
            m.bandR[b] = r;
            m.bandT[b] = std::min(std::clamp(std::isfinite(ti) ? ti : 0.0f, 0.0f, 1.0f), 1.0f - r);
        }
        m.r = m.energyR.front();
        m.t = m.energyT.front();
        // ref_TIR / tau_TIR, which facetInputView always writes (0.05 / 0)
        m.tirR = json::num(*sp, "refTir", 0.05f);
        m.tirT = json::num(*sp, "tauTir", 0.0f);
        m.label = spectralName;
        // physicalTexture (facetInputView core:1977-1994, read at core:2320-2344)
        m.texture.reset();
        m.textureStrength = 0.0f;
        m.textureRepeat = 1.0f;
        const json::Value& tex = json::member(*sp, "physicalTexture", json::Value::Object);
        if (json::boolean(tex, "enabled", false)) {
            const std::string file = json::str(tex, "fileName");
            if (file.empty()) throw std::runtime_error("Physical texture is enabled but has no file: " + spectralName);
            m.textureStrength = std::clamp(json::num(tex, "strength", 0.2f), 0.0f, 1.0f);
            m.textureRepeat = std::max(0.01f, json::num(tex, "repeatSize", 2.0f));
            m.texture = loadVarianceTexture(resolve ? resolve(file) : std::filesystem::path(file));
        }
    }
    const json::Value* th = named(json::member(cfg, "thermals", json::Value::Array), thermalName);
    if (th) {
        m.sunlitT = json::num(*th, "sunlitTemperature", 305.0f);
        m.shadedT = json::num(*th, "shadedTemperature", 295.0f);
    }
}

// OBJ positions, triangles and their groups (readObjGeometry, core:2559-2608):
// "g" / "o" name the group, "usemtl" when no group is active; polygons are
// fanned; the lowest vertex is moved to y = 0
struct ObjGeometry {
    std::vector<Vec3> positions;
    std::vector<std::array<int, 3>> triangles;
    std::vector<std::string> groups;
};

inline int objIndex(const std::string& tok, size_t count) {
    const int raw = std::stoi(tok.substr(0, tok.find('/')));
    const int idx = raw > 0 ? raw - 1 : int(count) + raw;
    if (idx < 0 || size_t(idx) >= count) throw std::runtime_error("OBJ position index is out of range");
    return idx;
}

inline ObjGeometry readObj(const std::filesystem::path& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Cannot open scene OBJ: " + path.string());
    ObjGeometry g;
    std::string active, line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string rec;
        ss >> rec;
        if (rec == "g" || rec == "o") {
            std::string name;
            if (ss >> name) active = name;
        } else if (rec == "usemtl" && active.empty()) {
            ss >> active;
        } else if (rec == "v") {
            Vec3 p{};
            if (!(ss >> p[0] >> p[1] >> p[2])) throw std::runtime_error("Invalid OBJ vertex in " + path.string());
            g.positions.push_back(p);
        } else if (rec == "f") {
            std::vector<int> face;
            std::string tok;
            while (ss >> tok) face.push_back(objIndex(tok, g.positions.size()));
            for (size_t i = 1; i + 1 < face.size(); ++i) {
                g.triangles.push_back({face[0], face[i], face[i + 1]});
                g.groups.push_back(active);
            }
        }
    }
    if (g.positions.empty() || g.triangles.empty())
        throw std::runtime_error("Scene OBJ contains no usable geometry: " + path.string());
    float ymin = std::numeric_limits<float>::max();
    for (const Vec3& p : g.positions) ymin = std::min(ymin, p[1]);
    for (Vec3& p : g.positions) p[1] -= ymin;
    return g;
}

// a box of the object's dimensions (cubeGeometry, core:2801-2828)
inline ObjGeometry boxGeometry(const std::vector<float>& dims) {
    float lx = dims.size() > 0 ? dims[0] : 1.0f, lz = dims.size() > 1 ? dims[1] : 1.0f;
    float h = dims.size() > 2 ? dims[2] : 1.0f;
    if (!(lx > 0.0f) || !(lz > 0.0f) || !(h > 0.0f)) lx = lz = h = 1.0f;
    const float hx = lx * 0.5f, hz = lz * 0.5f;
    ObjGeometry g;
    g.positions = {{-hx, 0.0f, -hz}, {hx, 0.0f, -hz}, {hx, 0.0f, hz}, {-hx, 0.0f, hz},
                   {-hx, h, -hz},    {hx, h, -hz},    {hx, h, hz},    {-hx, h, hz}};
    g.triangles = {{0, 2, 1}, {0, 3, 2}, {4, 5, 6}, {4, 6, 7}, {0, 1, 5}, {0, 5, 4},
                   {1, 2, 6}, {1, 6, 5}, {2, 3, 7}, {2, 7, 6}, {3, 0, 4}, {3, 4, 7}};
    g.groups.assign(g.triangles.size(), std::string{});
    return g;
}

// A DEM height field (DemHeightField and readDemHeightField,
// core:2617-2697): the first band of a grid of width x height cells, row 0
// the northernmost, no-data and non-finite cells replaced by the minimum of
// the valid ones.  Stretched over the scene: scene x (north) runs over the
// rows from the last to the first, scene z (east) over the columns, sampled
// bilinearly; heights are given above the minimum.
struct DemGrid {
    int width = 0, height = 0;
    float minimum = 0.0f;
    std::vector<float> values;

    bool valid() const { return width >= 2 && height >= 2 && values.size() == size_t(width) * size_t(height); }
    float heightAbove(float northX, float eastZ, float W, float D) const {
        if (!valid()) return 0.0f;
        const float u = std::clamp(eastZ / D, 0.0f, 1.0f) * float(width - 1);
        const float v = (1.0f - std::clamp(northX / W, 0.0f, 1.0f)) * float(height - 1);
        const int c0 = int(std::floor(u)), r0 = int(std::floor(v));
        const int c1 = std::min(width - 1, c0 + 1), r1 = std::min(height - 1, r0 + 1);
        const float fu = u - float(c0), fv = v - float(r0);
        const auto at = [&](int c, int r) { return values[size_t(r) * size_t(width) + size_t(c)]; };
        const float north = at(c0, r0) * (1.0f - fu) + at(c1, r0) * fu;
        const float south = at(c0, r1) * (1.0f - fu) + at(c1, r1) * fu;
        return std::max(0.0f, north * (1.0f - fv) + south * fv - minimum);
    }
};

// ESRI ASCII grid (.asc): "key value" header lines (ncols, nrows, the
// corner or centre coordinates, cellsize, an optional nodata_value), then
// nrows rows of ncols values from north to south
inline DemGrid readAsciiGrid(const std::filesystem::path& path, bool& hasNoData, float& noData) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Cannot open DEM: " + path.string());
    DemGrid g;
    hasNoData = false;
    std::string key;
    while (in >> key) {
        const char c0 = key.empty() ? ' ' : key[0];
        if ((c0 >= '0' && c0 <= '9') || c0 == '-' || c0 == '+' || c0 == '.') break;  // the first value
        double v = 0.0;
        if (!(in >> v)) throw std::runtime_error("Malformed DEM header: " + path.string());
        const std::string k = lower(key);
        if (k == "ncols") g.width = int(v);
        else if (k == "nrows") g.height = int(v);
        else if (k == "nodata_value") {
            hasNoData = true;
            noData = float(v);
        }
        key.clear();
    }
    if (g.width < 1 || g.height < 1 || size_t(g.width) * size_t(g.height) > (size_t(1) << 28))
        throw std::runtime_error("Cannot open DEM: " + path.string());
    const size_t n = size_t(g.width) * size_t(g.height);
    g.values.reserve(n);
    if (!key.empty()) g.values.push_back(std::strtof(key.c_str(), nullptr));
    float x = 0.0f;
    while (g.values.size() < n && in >> x) g.values.push_back(x);
    if (g.values.size() != n) throw std::runtime_error("Cannot read DEM elevations: " + path.string());
    return g;
}

// Baseline GeoTIFF (.tif): the first image, uncompressed, in strips or
// tiles, one or more samples per pixel interleaved (the first is read), 8 to
// 64 bits of unsigned, signed or IEEE samples, little- or big-endian; no-data
// from the GDAL_NODATA tag (42113)
inline DemGrid readGeoTiff(const std::filesystem::path& path, bool& hasNoData, float& noData) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open DEM: " + path.string());
    const std::vector<uint8_t> b((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    const auto fail = [&](const char* why) -> void {
        throw std::runtime_error(std::string("Cannot read DEM (") + why + "): " + path.string());
    };
    if (b.size() < 8) fail("not a TIFF file");
    const bool le = b[0] == 'I' && b[1] == 'I';
    if (!le && !(b[0] == 'M' && b[1] == 'M')) fail("not a TIFF file");
    const auto u16 = [&](size_t o) -> uint32_t {
        if (o + 2 > b.size()) fail("truncated");
        return le ? uint32_t(b[o]) | uint32_t(b[o + 1]) << 8 : uint32_t(b[o]) << 8 | uint32_t(b[o + 1]);
    };
    const auto u32 = [&](size_t o) -> uint32_t {
        if (o + 4 > b.size()) fail("truncated");
        return le ? u16(o) | u16(o + 2) << 16 : u16(o) << 16 | u16(o + 2);
    };
    if (u16(2) != 42) fail("BigTIFF and other variants are not read");
    const size_t ifd = u32(4);
    const uint32_t count = u16(ifd);
    uint32_t width = 0, height = 0, bits = 8, compression = 1, samples = 1, planar = 1, format = 1;
    uint32_t rowsPerStrip = 0, tileW = 0, tileH = 0;
    std::vector<uint32_t> offsets, byteCounts;
    hasNoData = false;
    for (uint32_t e = 0; e < count; ++e) {
        const size_t o = ifd + 2 + 12 * size_t(e);
        const uint32_t tag = u16(o), type = u16(o + 2), n = u32(o + 4);
        const uint32_t unit = type == 3 ? 2 : type == 4 ? 4 : 1;
        const size_t data = size_t(unit) * n <= 4 ? o + 8 : size_t(u32(o + 8));
        const auto value = [&](uint32_t i) -> uint32_t {
            return type == 3 ? u16(data + 2 * size_t(i)) : type == 4 ? u32(data + 4 * size_t(i)) : b.at(data + i);
        };
        switch (tag) {
        case 256: width = value(0); break;
        case 257: height = value(0); break;
        case 258: bits = value(0); break;
        case 259: compression = value(0); break;
        case 277: samples = value(0); break;
        case 278: rowsPerStrip = value(0); break;
        case 284: planar = value(0); break;
        case 322: tileW = value(0); break;
        case 323: tileH = value(0); break;
        case 339: format = value(0); break;
        case 273:
        case 324:
            offsets.resize(n);
            for (uint32_t i = 0; i < n; ++i) offsets[i] = value(i);
            break;
        case 279:
        case 325:
            byteCounts.resize(n);
            for (uint32_t i = 0; i < n; ++i) byteCounts[i] = value(i);
            break;
        case 42113:
            if (type == 2 && data + n <= b.size()) {
                const std::string text(reinterpret_cast<const char*>(b.data() + data), n);
                char* end = nullptr;
                const float v = std::strtof(text.c_str(), &end);
                if (end != text.c_str()) {
                    hasNoData = true;
                    noData = v;
                }
            }
            break;
        default: break;
        }
    }
    if (compression != 1) fail("compressed images are not read");
    if (width < 1 || height < 1 || size_t(width) * height > (size_t(1) << 28)) fail("bad size");
    if (samples < 1 || (samples > 1 && planar != 1)) fail("planar multi-band images are not read");
    if (!((bits == 8 || bits == 16 || bits == 32 || bits == 64) && format >= 1 && format <= 3) ||
        (format == 3 && bits < 32))
        fail("unsupported sample type");
    if (offsets.empty()) fail("no image data");
    const uint32_t bytes = bits / 8, pixelBytes = bytes * samples;
    const auto sample = [&](size_t o) -> float {
        if (o + bytes > b.size()) fail("truncated");
        uint64_t raw = 0;
        for (uint32_t k = 0; k < bytes; ++k) raw |= uint64_t(b[o + (le ? k : bytes - 1 - k)]) << (8 * k);
        if (format == 3) {
            if (bytes == 4) {
                const uint32_t r32 = uint32_t(raw);
                float f;
                std::memcpy(&f, &r32, 4);
                return f;
            }
            double d;
            std::memcpy(&d, &raw, 8);
            return float(d);
        }
        if (format == 2) {
            const int64_t sgn = int64_t(raw << (64 - 8 * bytes)) >> (64 - 8 * bytes);
            return float(sgn);
        }
        return float(raw);
    };
    DemGrid g;
    g.width = int(width);
    g.height = int(height);
    g.values.assign(size_t(width) * height, 0.0f);
    if (tileW && tileH) {
        const uint32_t across = (width + tileW - 1) / tileW, down = (height + tileH - 1) / tileH;
        if (offsets.size() < size_t(across) * down) fail("missing tiles");
        for (uint32_t ty = 0; ty < down; ++ty)
            for (uint32_t tx = 0; tx < across; ++tx) {
                const size_t base = offsets[size_t(ty) * across + tx];
                for (uint32_t y = 0; y < tileH && ty * tileH + y < height; ++y)
                    for (uint32_t x = 0; x < tileW && tx * tileW + x < width; ++x)
                        g.values[size_t(ty * tileH + y) * width + tx * tileW + x] =
                            sample(base + (size_t(y) * tileW + x) * pixelBytes);
            }
    } else {
        const uint32_t rows = rowsPerStrip ? std::min(rowsPerStrip, height) : height;
        if (offsets.size() < size_t((height + rows - 1) / rows)) fail("missing strips");
        for (uint32_t y = 0; y < height; ++y) {
            const size_t base = offsets[y / rows] + size_t(y % rows) * width * pixelBytes;
            for (uint32_t x = 0; x < width; ++x) g.values[size_t(y) * width + x] = sample(base + size_t(x) * pixelBytes);
        }
    }
    return g;
}

// readDemHeightField: a grid of at least 2 x 2 cells with a valid elevation
inline DemGrid readDem(const std::filesystem::path& path) {
    bool hasNoData = false;
    float noData = 0.0f;
    const std::string ext = lower(path.extension().string());
    DemGrid g = ext == ".asc" ? readAsciiGrid(path, hasNoData, noData) : readGeoTiff(path, hasNoData, noData);
    if (g.width < 2 || g.height < 2) throw std::runtime_error("DEM must contain at least 2 x 2 cells: " + path.string());
    g.minimum = std::numeric_limits<float>::infinity();
    for (float v : g.values)
        if (std::isfinite(v) && !(hasNoData && v == noData)) g.minimum = std::min(g.minimum, v);
    if (!std::isfinite(g.minimum)) throw std::runtime_error("DEM contains no valid elevations: " + path.string());
    for (float& v : g.values)
        if (!std::isfinite(v) || (hasNoData && v == noData)) v = g.minimum;
    return g;
}

// Texture factors per facet (finalizeReflectanceTextureFactors,
// core:718-846): for the facets of a textured material of positive
// strength, the texture sampled at the centroid by triplanar projection
// (weights |n_x|, |n_y|, |n_z| on the (z, y), (x, z), (x, y) planes, in
// units of the repeat size), less its area-weighted mean over the
// material; the deviations scaled by the strength, limited so that no
// factor falls to 0.001 or lifts any band of the material above r + t = 1
// (r max(r) factor <= 1 - t, over the band and energy spectra), plus one.
// The area-weighted mean factor of a material is then exactly one.
inline void textureFactors(Scene& scene, const std::vector<ProjectMaterial>& mats) {
    const uint32_t F = scene.facetCount();
    scene.reflectanceFactor.assign(F, 1.0f);
    if (!F || mats.empty()) return;
    struct Stats {
        double area = 0.0, weighted = 0.0, factorSum = 0.0;
        float minDev = 0.0f, maxDev = 0.0f, scale = 0.0f;
        float minFactor = std::numeric_limits<float>::infinity(), maxFactor = std::numeric_limits<float>::lowest();
    };
    std::vector<Stats> st(mats.size());
    std::vector<float> sample(F, 0.0f), area(F, 0.0f);
    for (uint32_t f = 0; f < F; ++f) {
        if (f >= scene.material.size() || scene.material[f] >= mats.size()) continue;
        const ProjectMaterial& m = mats[scene.material[f]];
        if (!m.texture || !(m.textureStrength > 0.0f)) continue;
        const facethip::Vertex& a = scene.vertices[3 * size_t(f)];
        const facethip::Vertex& b = scene.vertices[3 * size_t(f) + 1];
        const facethip::Vertex& c = scene.vertices[3 * size_t(f) + 2];
        const float u[3] = {b.position[0] - a.position[0], b.position[1] - a.position[1], b.position[2] - a.position[2]};
        const float w[3] = {c.position[0] - a.position[0], c.position[1] - a.position[1], c.position[2] - a.position[2]};
        const float cx = u[1] * w[2] - u[2] * w[1], cy = u[2] * w[0] - u[0] * w[2], cz = u[0] * w[1] - u[1] * w[0];
        area[f] = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
        const float px = (a.position[0] + b.position[0] + c.position[0]) / 3.0f;
        const float py = (a.position[1] + b.position[1] + c.position[1]) / 3.0f;
        const float pz = (a.position[2] + b.position[2] + c.position[2]) / 3.0f;
        const float nx = std::abs(a.normal[0]), ny = std::abs(a.normal[1]), nz = std::abs(a.normal[2]);
        const float wsum = std::max(1.0e-6f, nx + ny + nz);
        const float k = 1.0f / std::max(0.01f, m.textureRepeat);
        const VarianceTexture& t = *m.texture;
        sample[f] = (nx * t.sample(pz * k, py * k) + ny * t.sample(px * k, pz * k) + nz * t.sample(px * k, py * k)) /
                    wsum;
        Stats& s = st[scene.material[f]];
        s.area += area[f];
        s.weighted += double(area[f]) * sample[f];
    }
    for (uint32_t f = 0; f < F; ++f) {
        if (!(area[f] > 0.0f)) continue;
        Stats& s = st[scene.material[f]];
        const float mean = s.area > 0.0 ? float(s.weighted / s.area) : 0.0f;
        sample[f] -= mean;
        s.minDev = std::min(s.minDev, sample[f]);
        s.maxDev = std::max(s.maxDev, sample[f]);
    }
    for (size_t i = 0; i < mats.size(); ++i) {
        Stats& s = st[i];
        if (!(s.area > 0.0)) continue;
        const ProjectMaterial& m = mats[i];
        float scale = std::clamp(m.textureStrength, 0.0f, 1.0f);
        if (s.minDev < 0.0f) scale = std::min(scale, 0.999f / -s.minDev);
        // maximumMeanPreservingFactor (core:743-762)
        float maxFactor = std::numeric_limits<float>::infinity();
        const auto bands = [&](const std::vector<float>& r, const std::vector<float>& t) {
            for (size_t b = 0; b < r.size(); ++b) {
                const float base = std::clamp(r[b], 0.0f, 1.0f);
                if (!(base > 1.0e-6f)) continue;
                const float tau = t.empty() ? 0.0f : std::clamp(t[std::min(b, t.size() - 1)], 0.0f, 1.0f - base);
                maxFactor = std::min(maxFactor, (1.0f - tau) / base);
            }
        };
        bands(m.bandR.empty() ? std::vector<float>{m.r} : m.bandR, m.bandT.empty() ? std::vector<float>{m.t} : m.bandT);
        bands(m.energyR, m.energyT);
        if (s.maxDev > 0.0f && std::isfinite(maxFactor))
            scale = std::min(scale, std::max(0.0f, maxFactor - 1.0f) / s.maxDev);
        s.scale = scale;
    }
    for (uint32_t f = 0; f < F; ++f) {
        if (!(area[f] > 0.0f)) continue;
        Stats& s = st[scene.material[f]];
        const float factor = 1.0f + s.scale * sample[f];
        scene.reflectanceFactor[f] = factor;
        s.factorSum += double(area[f]) * factor;
        s.minFactor = std::min(s.minFactor, factor);
        s.maxFactor = std::max(s.maxFactor, factor);
    }
    bool any = false;
    for (size_t i = 0; i < st.size(); ++i)
        if (st[i].area > 0.0) {
            any = true;
            std::cout << "facetrt: physical texture material " << i << ", area-weighted factor mean="
                      << st[i].factorSum / st[i].area << ", range=[" << st[i].minFactor << ',' << st[i].maxFactor
                      << "]" << std::endl;
        }
    if (!any) scene.reflectanceFactor.clear();  // no textured facet: 1 everywhere
}

struct Placement {
    float x = 0.0f, y = 0.0f, z = 0.0f, scale = 1.0f, rotation = 0.0f;
};

// readPlacements (core:2699-2737): "x y z [scale [rotation]]" per line, the
// scene centre when the object has no (usable) positions
inline std::vector<Placement> placements(const std::filesystem::path& file, float W, float D) {
    std::vector<Placement> out;
    if (!file.empty()) {
        std::ifstream in(file);
        if (!in) throw std::runtime_error("Cannot open object positions: " + file.string());
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ss(line);
            Placement p;
            if (!(ss >> p.x >> p.y >> p.z)) continue;
            if (!(ss >> p.scale)) {
                p.scale = 1.0f;
                ss.clear();
            }
            if (!(ss >> p.rotation)) p.rotation = 0.0f;
            if (std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z) && std::isfinite(p.scale) &&
                std::isfinite(p.rotation) && p.scale > 0.0f)
                out.push_back(p);
        }
    }
    if (out.empty()) out.push_back({W * 0.5f, D * 0.5f, 0.0f, 1.0f, 0.0f});
    return out;
}

// transformPosition (core:2739-2755): scale, rotate about +y by -rotation
// degrees, move to the placement (scene x, y -> x, z) centred on the scene
inline Vec3 place(const Vec3& v, const Placement& p, float W, float D) {
    const float rad = -p.rotation * 3.14159265358979323846f / 180.0f;
    const float c = std::cos(rad), s = std::sin(rad);
    const float x = v[0] * p.scale, y = v[1] * p.scale, z = v[2] * p.scale;
    return {c * x + s * z + p.x - W * 0.5f, y + p.z, -s * x + c * z + p.y - D * 0.5f};
}

// clipPolygonToPlane (core:308-341)
inline std::vector<Vec3> clip(const std::vector<Vec3>& src, int axis, float plane, bool keepGreater) {
    std::vector<Vec3> out;
    if (src.empty()) return out;
    const auto inside = [&](const Vec3& p) {
        return keepGreater ? p[axis] >= plane - 1.0e-6f : p[axis] <= plane + 1.0e-6f;
    };
    Vec3 prev = src.back();
    bool prevIn = inside(prev);
    for (const Vec3& cur : src) {
        const bool curIn = inside(cur);
        if (curIn != prevIn) {
            const float den = cur[axis] - prev[axis];
            if (std::abs(den) > 1.0e-8f) {
                const float a = std::clamp((plane - prev[axis]) / den, 0.0f, 1.0f);
                Vec3 q{};
                for (int k = 0; k < 3; ++k) q[k] = prev[k] + a * (cur[k] - prev[k]);
                out.push_back(q);
            }
        }
        if (curIn) out.push_back(cur);
        prev = cur;
        prevIn = curIn;
    }
    return out;
}

// appendTriangleInsideDomain (core:343-360): the triangle clipped to the
// scene box, fanned
inline void appendInside(Scene& s, const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& lo, const Vec3& hi,
                         uint32_t mat) {
    std::vector<Vec3> poly{a, b, c};
    for (int axis = 0; axis < 3 && poly.size() >= 3; ++axis) {
        poly = clip(poly, axis, lo[axis], true);
        poly = clip(poly, axis, hi[axis], false);
    }
    for (size_t i = 1; i + 1 < poly.size(); ++i) appendTriangle(s, poly[0], poly[i], poly[i + 1], mat);
}

}  // namespace project_detail

// loadProjectScene (core:2880-3024) over the project's JSON (the tags of
// facetInputView, core:1888-2043)
inline Project loadProject(const std::string& file) {
    namespace fs = std::filesystem;
    namespace pd = project_detail;
    const fs::path path = fs::absolute(file).lexically_normal();
    if (pd::lower(path.filename().string()) != "project.json")
        throw std::invalid_argument("HiStream accepts project.json only");
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open project.json: " + path.string());
    std::stringstream buf;
    buf << in.rdbuf();
    const json::Value root = json::parse(buf.str());
    if (!root.isObject()) throw std::runtime_error("project.json root must be an object");
    const std::string format = json::str(root, "format", "streamsim-project");
    if (format != "streamsim-project") throw std::runtime_error("Unsupported project format: " + format);
    const int version = json::integer(root, "schemaVersion", 1);
    if (version < 1 || version > 6) throw std::runtime_error("Unsupported project schemaVersion; expected 1-6");
    const json::Value& cfg = json::member(root, "configuration", json::Value::Object);
    if (cfg.members.empty()) throw std::runtime_error("project.json has no configuration object");
    const fs::path dir = path.parent_path();
    const auto resolve = [&](const std::string& v) { return pd::asset(dir, root, v); };
    const json::Value& sc = json::member(cfg, "scene", json::Value::Object);
    const json::Value& light = json::member(cfg, "light", json::Value::Object);
    const json::Value& sensor = json::member(cfg, "sensor", json::Value::Object);
    const json::Value& control = json::member(cfg, "control", json::Value::Object);
    const json::Value& meteo = json::member(cfg, "meteo", json::Value::Object);
    const json::Value& background = json::member(sc, "background", json::Value::Object);
    const json::Value& mats = json::member(cfg, "materials", json::Value::Array);

    Project prj;
    prj.mode = json::str(root, "mode", "");
    ProjectRun& run = prj.run;
    run.bands = pd::sensorBands(sensor);
    // settings as facetInputView writes them and the run reads them back
    run.periodicNeighbors = uint32_t(std::clamp(
        json::integer(control, "periodicTraversalCount",
                      json::boolean(control, "periodicBoundary", false)
                          ? json::integer(control, "periodicNeighborCount", 8)
                          : 0),
        0, 20));
    run.couplingIterations = uint32_t(std::clamp(json::integer(control, "couplingIterations", 20), 1, 100));
    run.tolerance = std::clamp(json::num(control, "temperatureTolerance", 0.05f), 0.001f, 10.0f);
    run.relaxation = std::clamp(json::num(control, "temperatureRelaxation", 0.5f), 0.05f, 1.0f);
    run.spectralWidth =
        uint32_t(std::clamp(int(std::lround(json::num(control, "spectralAccelerationWidth", 100.0f))), 1, 1000));
    run.vegetationMethod = json::integer(control, "vegetationTemperatureMethod", 0);
    run.acceleratedSolver = json::str(control, "radiationSolver", "traditional") == "accelerated";
    run.saveEnergy = json::boolean(sensor, "energyProcess", false);
    run.saveRadiation = json::boolean(sensor, "radiationProcess", false);
    run.sunZenith = json::num(light, "zenith", 30.0f);
    run.sunAzimuth = json::num(light, "azimuth", 135.0f);
    run.skyTemperature = std::clamp(json::num(light, "skyTemperature", 250.0f), 150.0f, 350.0f);
    run.directRatio = std::clamp(json::num(light, "direct", 0.8f), 0.0f, 1.0f);
    run.startNode = std::max(0, json::integer(meteo, "start", 0));
    run.endNode = std::max(run.startNode + 1, json::integer(meteo, "end", run.startNode + 1));
    run.dTime = std::max(1.0f, json::num(meteo, "dTime", 1800.0f));
    run.latitude = json::num(meteo, "latitude", 40.0f);
    run.longitude = json::num(meteo, "longitude", 116.0f);
    {
        // ProjectJson::meteorologyPath: the built-in data unless a path is set
        const std::string v = json::str(meteo, "path", "defined/meteo.txt");
        if (!v.empty() && v != "defined/meteo.txt" && v != "HiStream 内置气象数据")
            run.meteoFile = pd::asset(dir, root, v).string();
    }
    run.outDir = pd::asset(dir, root, json::str(cfg, "outDir", "output")).string();

    // scene box, ground, materials (loadProjectScene)
    const float W = std::max(0.01f, json::num(sc, "x", 60.0f));
    const float D = std::max(0.01f, json::num(sc, "y", 60.0f));
    const float H = std::max(0.01f, json::num(sc, "height", 15.0f));
    const float voxel = std::max(0.01f, json::num(sc, "voxel", 1.0f));
    // DEM ground (facetInputView writes <isDEM> and <DEM> for scene.terrain
    // with a demFile, core:1921, 1944-1945; readDemHeightField, core:2652-2697)
    pd::DemGrid dem;
    if (json::boolean(sc, "terrain", false) && !json::str(sc, "demFile").empty())
        dem = pd::readDem(pd::asset(dir, root, json::str(sc, "demFile")));
    for (const json::Value& m : mats.items)
        if (json::str(m, "energyModel") == "photovoltaic" &&
            json::str(m, "name") == json::str(background, "materialName", "soilset"))
            throw std::runtime_error("Photovoltaic panels must be scene objects, not background ground");
    // the Hapke ground (core:1942-1943 writes the project's background as
    // backgroundHeterogeneity Hapke / Lambert and backgroundAngularStrength;
    // core:2904-2910 keeps the strength, clamped to [0, 1], for Hapke only)
    prj.scene.backgroundAngularStrength =
        json::boolean(background, "heterogeneityEnabled", false)
            ? std::clamp(json::num(background, "angularEffectStrength", 0.5f), 0.0f, 1.0f)
            : 0.0f;
    prj.scene.extentX = W;
    prj.scene.extentZ = D;
    // large scenes rasterise at 2048 pixels and 16 layers (core:2895-2901)
    if (std::max(W, D) >= 1000.0f) {
        prj.scene.rasterSize = 2048u;
        prj.scene.rasterLayers = 16u;
    }

    // materials of the FacetRT parameters (facetRTParametersFromInput,
    // core:2385-2420): the object default from the first object's first
    // spectral / thermal names, the ground from the background entries
    std::vector<const json::Value*> objects;
    for (const json::Value& o : json::member(json::member(cfg, "objects", json::Value::Object), "items",
                                             json::Value::Array)
                                    .items) {
        if (!o.isObject()) continue;
        const std::string type = json::str(o, "type", "Vegetation");
        if (type == "Fire" || type == "Fog" || o.get("medium")) continue;
        objects.push_back(&o);
    }
    ProjectMaterial objectDefault;
    if (!objects.empty()) {
        const json::Value& o = *objects.front();
        const std::string sp = pd::names(pd::joined(o, "spectralName", json::str(o, "spectralName", "soil"))).front();
        const std::string th =
            pd::names(pd::joined(o, "thermalName", json::str(o, "thermalName", "soil_temperature"))).front();
        pd::surfaceParameters(cfg, sp, th, run.bands, objectDefault, resolve);
    }
    ProjectMaterial ground;
    ground.r = 0.20f;
    ground.t = 0.0f;
    ground.tirR = 0.05f;
    ground.sunlitT = 305.0f;
    ground.shadedT = 295.0f;
    ground.component = 1;
    pd::surfaceParameters(cfg, json::str(background, "spectralName", "soil"),
                          json::str(background, "thermalName", "soil_temperature"), run.bands, ground, resolve);
    {  // loadSoilParameters (core:2368-2383) from the background's soil set
        const std::string soilName = json::str(background, "materialName", "soilset");
        for (const json::Value& m : mats.items) {
            if (!m.isObject() || json::str(m, "type", "Soil") != "Soil" || json::str(m, "name", "soilset") != soilName)
                continue;
            const json::Value& pr = json::member(m, "params", json::Value::Object);
            ground.soilMethod =
                std::clamp(json::integer(control, "soilTemperatureMethod", json::integer(pr, "method", 1)), 0, 2);
            ground.rss = std::max(0.0f, json::num(pr, "rss", 2000.0f));
            ground.cs = std::max(1.0f, json::num(pr, "cs", 1180.0f));
            ground.rhos = std::max(1.0f, json::num(pr, "rhos", 1800.0f));
            ground.lambdas = std::max(0.001f, json::num(pr, "lambdas", 1.55f));
            const float smc = json::num(pr, "SMC", 25.0f);
            ground.smc = std::clamp(smc > 1.0f ? smc * 0.01f : smc, 0.0f, 1.0f);
            break;
        }
    }

    for (const json::Value* op : objects) {
        const json::Value& o = *op;
        const std::string type = json::str(o, "type", "Vegetation");
        const std::string name = json::str(o, "name", "object");
        const std::vector<std::string> meshNames = pd::names(pd::joined(o, "name", name));
        const std::vector<std::string> spectral =
            pd::names(pd::joined(o, "spectralName", json::str(o, "spectralName", "soil")));
        const std::vector<std::string> thermal =
            pd::names(pd::joined(o, "thermalName", json::str(o, "thermalName", "soil_temperature")));
        const std::vector<std::string> energy = pd::names(pd::joined(
            o, "materialName",
            json::str(o, "materialName", type == "Vegetation" ? "leaf_c3" : type == "Water" ? "water_set" : "soilset")));
        const size_t count = std::max<size_t>(1, std::max(meshNames.size(), std::max(spectral.size(), thermal.size())));
        std::vector<std::pair<std::string, uint32_t>> bindings;
        const uint8_t component = type == "Soil" ? 1 : type == "Vegetation" ? 2 : type == "Building" ? 3
                                  : type == "Water" ? 5 : 0;
        for (size_t i = 0; i < count; ++i) {
            ProjectMaterial m = objectDefault;
            m.component = component;
            const std::string sp = spectral.empty() ? std::string{} : spectral[std::min(i, spectral.size() - 1)];
            const std::string th = thermal.empty() ? std::string{} : thermal[std::min(i, thermal.size() - 1)];
            pd::surfaceParameters(cfg, sp, th, run.bands, m, resolve);
            const std::string en = energy.empty() ? std::string{} : energy[std::min(i, energy.size() - 1)];
            if (const json::Value* e = pd::named(mats, en)) {
                // surfaceEnergy (core:2015-2026, 2953-2973)
                const json::Value& pr = json::member(*e, "params", json::Value::Object);
                const bool pv = json::str(*e, "energyModel") == "photovoltaic";
                m.pv = pv;
                m.vegetation = json::str(*e, "type") == "Vegetation" && json::str(*e, "energyModel") != "wood" && !pv;
                m.heatCapacity = json::num(pr, "heatCapacityPerArea", pv ? 12000.0f : 60000.0f);
                m.convectiveScale = json::num(pr, "convectiveScale", 1.0f);
                m.pvEta25 = json::num(pr, "eta25", 0.22f);
                m.pvGamma = json::num(pr, "gamma", -0.0035f);
                m.pvBifaciality = json::num(pr, "bifaciality", 0.0f);
                if (pv && (!(m.pvEta25 >= 0.0f && m.pvEta25 <= 0.5f) || !(m.pvGamma >= -0.02f && m.pvGamma <= 0.0f) ||
                           !(m.pvBifaciality >= 0.0f && m.pvBifaciality <= 1.0f) || !(m.heatCapacity >= 0.0f) ||
                           !(m.convectiveScale > 0.0f)))
                    throw std::runtime_error("Invalid photovoltaic material parameters: " + en);
                if (pv && (m.tirT > 1e-6f ||
                           std::any_of(m.energyT.begin(), m.energyT.end(), [](float v) { return v > 1e-6f; })))
                    throw std::runtime_error("Photovoltaic surface must be opaque: " + en);
                if (pv) m.component = 4;
            }
            m.label = sp + "/" + th + "/" + en;
            bindings.push_back({meshNames.empty() ? std::string{} : meshNames[std::min(i, meshNames.size() - 1)],
                                uint32_t(prj.materials.size())});
            prj.materials.push_back(std::move(m));
        }
        const std::string model = json::str(o, "fileName");
        const pd::ObjGeometry geo = model.empty()
                                        ? pd::boxGeometry(json::numbers(o.get("dimensions")))
                                        : pd::readObj(pd::asset(dir, root, model));
        const std::string posFile = json::str(o, "positionFile");
        const Vec3 lo{-W * 0.5f, 0.0f, -D * 0.5f}, hi{W * 0.5f, H, D * 0.5f};
        for (pd::Placement p : pd::placements(posFile.empty() ? fs::path{} : pd::asset(dir, root, posFile), W, D)) {
            p.z += dem.heightAbove(p.x, p.y, W, D);  // objects stand on the terrain (core:2991-2992)
            // appendGeometry (core:2762-2799): a triangle takes the material of
            // the binding whose mesh name equals its group (case-insensitive),
            // else the first binding's
            for (size_t ti = 0; ti < geo.triangles.size(); ++ti) {
                uint32_t mat = bindings.front().second;
                const std::string g = pd::lower(geo.groups[ti]);
                for (const auto& b : bindings)
                    if (!b.first.empty() && pd::lower(b.first) == g) {
                        mat = b.second;
                        break;
                    }
                const auto& t = geo.triangles[ti];
                pd::appendInside(prj.scene, pd::place(geo.positions[t[0]], p, W, D),
                                 pd::place(geo.positions[t[1]], p, W, D), pd::place(geo.positions[t[2]], p, W, D), lo,
                                 hi, mat);
            }
        }
    }
    prj.scene.leafFacetCount = prj.scene.facetCount();

    // appendProjectGround (core:2830-2878): cells of about the voxel size, at
    // most a quarter of the visibility raster per side, corners 1 cm below
    // the terrain (the DEM's height above its minimum, 0 without one)
    const uint32_t groundMat = uint32_t(prj.materials.size());
    prj.materials.push_back(ground);
    const uint32_t maxGrid = prj.scene.rasterSize / 4u;
    const uint32_t cols = std::clamp(std::max(1u, uint32_t(std::ceil(W / voxel))), 1u, maxGrid);
    const uint32_t rows = std::clamp(std::max(1u, uint32_t(std::ceil(D / voxel))), 1u, maxGrid);
    const float cw = W / float(cols), cd = D / float(rows), x0 = -W * 0.5f, z0 = -D * 0.5f;
    const auto h = [&](float x, float z) { return dem.heightAbove(x + W * 0.5f, z + D * 0.5f, W, D) - 0.01f; };
    for (uint32_t r = 0; r < rows; ++r) {
        const float za = z0 + cd * float(r);
        const float zb = r + 1 == rows ? z0 + D : z0 + cd * float(r + 1);
        for (uint32_t c = 0; c < cols; ++c) {
            const float xa = x0 + cw * float(c);
            const float xb = c + 1 == cols ? x0 + W : x0 + cw * float(c + 1);
            const float haa = h(xa, za), hab = h(xa, zb), hba = h(xb, za), hbb = h(xb, zb);
            appendTriangle(prj.scene, {xa, haa, za}, {xb, hbb, zb}, {xb, hba, za}, groundMat);
            appendTriangle(prj.scene, {xa, haa, za}, {xa, hab, zb}, {xb, hbb, zb}, groundMat);
        }
    }
    // adaptiveVisibilityRasterLayers (core:168-188): 16 times the mean
    // fragment depth over the raster (periodic copies included), the layers
    // doubled up to 64 until they hold it
    {
        const double pixels = double(prj.scene.rasterSize) * double(prj.scene.rasterSize);
        const double depth = double(prj.scene.facetCount()) * double(run.periodicNeighbors + 1u) / pixels;
        const uint32_t want = uint32_t(std::ceil(std::max(1.0, depth * 16.0)));
        const uint32_t initial = prj.scene.rasterLayers;
        while (prj.scene.rasterLayers < want && prj.scene.rasterLayers < 64u)
            prj.scene.rasterLayers = std::min(64u, prj.scene.rasterLayers * 2u);
        if (prj.scene.rasterLayers != initial)
            std::cout << "facetrt: automatically increased A-buffer layers from " << initial << " to "
                      << prj.scene.rasterLayers << " for " << prj.scene.facetCount() << " facets" << std::endl;
    }
    pd::textureFactors(prj.scene, prj.materials);
    prj.facetComponent.resize(prj.scene.facetCount());
    for (uint32_t f = 0; f < prj.scene.facetCount(); ++f)
        prj.facetComponent[f] = f < prj.scene.leafFacetCount ? prj.materials[prj.scene.material[f]].component : 1;
    return prj;
}

// The assembled scene as OBJ: "o object" for the object facets, then
// "o STREAMSIM_TERRAIN" for the ground, one "usemtl m<index>" per material run
inline void writeSceneObj(const Scene& s, const std::string& file) {
    std::ofstream out(file, std::ios::trunc);
    if (!out) throw std::runtime_error("Cannot write scene OBJ: " + file);
    out << std::setprecision(9);
    for (const auto& v : s.vertices) out << "v " << v.position[0] << ' ' << v.position[1] << ' ' << v.position[2] << '\n';
    uint32_t prev = std::numeric_limits<uint32_t>::max();
    for (uint32_t f = 0; f < s.facetCount(); ++f) {
        if (f == 0) out << "o object\n";
        if (f == s.leafFacetCount) {
            out << "o STREAMSIM_TERRAIN\n";
            prev = std::numeric_limits<uint32_t>::max();
        }
        if (s.material[f] != prev) {
            out << "usemtl m" << s.material[f] << '\n';
            prev = s.material[f];
        }
        out << "f " << 3 * f + 1 << ' ' << 3 * f + 2 << ' ' << 3 * f + 3 << '\n';
    }
    if (!out) throw std::runtime_error("Cannot write scene OBJ: " + file);
}

}  // namespace facetcli
