// facet_engine.cpp -- command-line engine for the facet modes.
//
//   facet_engine eFacetRT|eFacetEB scene.obj [output.json]
//
// Mirrors the reference executable's facet path for OBJ input
// (models/histream/src/main.cpp:54-79, runFacetRTCore / runFacetEBCore in
// models/histream/src/facetrt/facetrtcore.cpp, "core"):
//   * the scene is read as loadTreeWithSoil does (core:362-473);
//   * eFacetRT: makeOptics, a 64-sweep solve with direct light and one
//     without, light enhancement = max(0, difference), written in the
//     writeWebResult layout (core:1782-1840);
//   * eFacetEB: one meteorological node with the reference's defaults for an
//     input without XML settings (25 C air, 15 hPa, 839 hPa, 1 m/s, no
//     shortwave, 320 W m-2 longwave, 0.05 K tolerance, 20 coupling
//     iterations, relaxation 0.5, dTime 1800 s), written in the faceteb.json
//     layout (core:4088-4140); exit code 2 when the node did not converge.
// stdout carries the reference's PROGRESS / RESULT lines; exit codes follow
// main.cpp (2 usage or non-converged, 1 failure).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "facet_backend.hpp"
#include "facet_hip.h"

namespace {

using Vec3 = std::array<float, 3>;
using Clock = std::chrono::steady_clock;

constexpr float kPi = 3.14159265358979323846f;
constexpr float kBeam = 700.0f;
constexpr float kSky = 100.0f;
constexpr uint32_t kSoilGrid = 48;

struct Scene {
    std::vector<facethip::Vertex> vertices;
    std::vector<uint32_t> material;
    uint32_t leafFacetCount = 0;
    uint32_t facetCount() const { return uint32_t(vertices.size() / 3); }
};

// appendTriangle (core:288-306): normal of cross(b - a, c - a) in float
bool appendTriangle(Scene& s, const Vec3& a, const Vec3& b, const Vec3& c, uint32_t mat) {
    const Vec3 u{b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const Vec3 v{c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const Vec3 n{u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    const float dd = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (dd <= 1.0e-16f) return false;
    const float inv = 1.0f / std::sqrt(dd);
    for (const Vec3* p : {&a, &b, &c}) {
        facethip::Vertex vx;
        for (int k = 0; k < 3; ++k) {
            vx.position[k] = (*p)[k];
            vx.normal[k] = n[k] * inv;
        }
        s.vertices.push_back(vx);
    }
    s.material.push_back(mat);
    return true;
}

// loadTreeWithSoil (core:362-473) for an OBJ without a terrain object
Scene loadObj(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Cannot open tree OBJ: " + path);
    std::vector<Vec3> pos;
    std::vector<std::array<int, 3>> tri;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string rec;
        ss >> rec;
        if (rec == "v") {
            Vec3 p{};
            if (!(ss >> p[0] >> p[1] >> p[2])) throw std::runtime_error("Invalid OBJ vertex");
            pos.push_back(p);
        } else if (rec == "f") {
            std::vector<int> face;
            std::string tok;
            while (ss >> tok) {
                const int raw = std::stoi(tok.substr(0, tok.find('/')));
                const int idx = raw > 0 ? raw - 1 : int(pos.size()) + raw;
                if (idx < 0 || size_t(idx) >= pos.size()) throw std::runtime_error("OBJ position index is out of range");
                face.push_back(idx);
            }
            for (size_t i = 1; i + 1 < face.size(); ++i) tri.push_back({face[0], face[i], face[i + 1]});
        }
    }
    if (pos.empty() || tri.empty()) throw std::runtime_error("Tree OBJ contains no usable geometry");
    Vec3 lo{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
    Vec3 hi{std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(),
            std::numeric_limits<float>::lowest()};
    for (const Vec3& p : pos)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    for (Vec3& p : pos) p[1] -= lo[1];
    Scene s;
    for (const auto& t : tri)
        if (appendTriangle(s, pos[t[0]], pos[t[1]], pos[t[2]], 0u)) ++s.leafFacetCount;
    const float margin = 0.5f * std::max(hi[0] - lo[0], hi[2] - lo[2]);
    const float x0 = lo[0] - margin, x1 = hi[0] + margin, z0 = lo[2] - margin, z1 = hi[2] + margin;
    for (uint32_t row = 0; row < kSoilGrid; ++row) {
        const float za = z0 + (z1 - z0) * float(row) / float(kSoilGrid);
        const float zb = z0 + (z1 - z0) * float(row + 1) / float(kSoilGrid);
        for (uint32_t col = 0; col < kSoilGrid; ++col) {
            const float xa = x0 + (x1 - x0) * float(col) / float(kSoilGrid);
            const float xb = x0 + (x1 - x0) * float(col + 1) / float(kSoilGrid);
            appendTriangle(s, {xa, -0.01f, za}, {xb, -0.01f, zb}, {xb, -0.01f, za}, 1u);
            appendTriangle(s, {xa, -0.01f, za}, {xa, -0.01f, zb}, {xb, -0.01f, zb}, 1u);
        }
    }
    return s;
}

// sunDirection (core:491-498) at the scene defaults zenith 30, azimuth 135
facethip::Direction sunDirection(float zenithDeg, float azimuthDeg) {
    const float z = zenithDeg * kPi / 180.0f, a = azimuthDeg * kPi / 180.0f;
    return {std::sin(z) * std::cos(a), std::cos(z), std::sin(z) * std::sin(a)};
}

// makeOptics (core:500-522)
std::vector<facethip::SurfaceOptics> makeOptics(const Scene& s, const std::vector<float>& sunlit,
                                                 const facethip::Direction& sun) {
    std::vector<facethip::SurfaceOptics> o(size_t(s.facetCount()) * 2);
    for (uint32_t f = 0; f < s.facetCount(); ++f) {
        const bool leaf = f < s.leafFacetCount;
        const float* n = s.vertices[3 * size_t(f)].normal;
        for (uint32_t k = 0; k < 2; ++k) {
            const uint32_t side = 2 * f + k;
            const float sg = k == 0 ? 1.0f : -1.0f;
            const float d = (sg * n[0]) * sun.x + (sg * n[1]) * sun.y + (sg * n[2]) * sun.z;
            o[side].reflectance = leaf ? 0.10f : 0.20f;
            o[side].transmittance = leaf ? 0.05f : 0.0f;
            o[side].directIrradiance = sunlit[side] * std::max(0.0f, d) * kBeam;
        }
    }
    return o;
}

double ms(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

void writeArray(std::ostream& out, const char* name, const std::vector<float>& v, bool comma) {
    out << "  \"" << name << "\": [";
    for (size_t i = 0; i < v.size(); ++i) {
        if (i) out << ',';
        out << v[i];
    }
    out << ']' << (comma ? "," : "") << '\n';
}

void writeHeader(std::ostream& out, const std::string& backend, const Scene& s) {
    out << std::setprecision(9) << "{\n  \"backend\": \"" << backend << "\",\n";
}

void writeVertices(std::ostream& out, const Scene& s) {
    out << "  \"vertexPositions\": [";
    for (size_t i = 0; i < s.vertices.size(); ++i) {
        if (i) out << ',';
        const auto& v = s.vertices[i];
        out << v.position[0] << ',' << v.position[1] << ',' << v.position[2];
    }
    out << "],\n";
}

std::string outputPath(const std::string& input, const std::string& requested, const char* name) {
    if (!requested.empty()) return requested;
    return (std::filesystem::path(input).parent_path() / name).lexically_normal().string();
}

facethip::Config engineConfig() {
    facethip::Config c;
    if (const char* e = std::getenv("FACET_RAYS_PER_SIDE")) c.raysPerSide = uint32_t(std::atoi(e));
    if (const char* e = std::getenv("FACET_SUN_RAYS_PER_SIDE")) c.sunRaysPerSide = uint32_t(std::atoi(e));
    return c;
}

int runFacetRT(const std::string& input, const std::string& outFile) try {
    const std::string result = outputPath(input, outFile, "facetrt.json");
    const Scene scene = loadObj(input);
    const auto t0 = Clock::now();
    std::cout << "PROGRESS\t3\tCreating HIP device context" << std::endl;
    facethip::FacetrtHip model;
    model.initialize(engineConfig());
    model.setGeometry(scene.vertices);
    std::cout << "PROGRESS\t10\tRay-cast form factors" << std::endl;
    const auto tv = Clock::now();
    const facethip::GraphDiagnostics graph = model.buildVisibilityGraph({facethip::Direction{}});
    const auto tv1 = Clock::now();
    const facethip::Direction sun = sunDirection(30.0f, 135.0f);
    const std::vector<float> sunlit = model.computeSunlitFraction(sun);
    const auto ts1 = Clock::now();
    const auto optics = makeOptics(scene, sunlit, sun);
    std::cout << "PROGRESS\t82\tHIP Jacobi radiosity: 64 iterations" << std::endl;
    const facethip::SolveResult sol = model.solve(optics, kSky, 64, 1.0f);
    auto diffuse = optics;
    for (auto& o : diffuse) o.directIrradiance = 0.0f;
    const facethip::SolveResult dsol = model.solve(diffuse, kSky, 64, 1.0f);
    std::vector<float> enh(sol.radiosity.size());
    for (size_t i = 0; i < enh.size(); ++i) enh[i] = std::max(0.0f, sol.radiosity[i] - dsol.radiosity[i]);
    const auto t1 = Clock::now();
    std::filesystem::create_directories(std::filesystem::path(result).parent_path().empty()
                                            ? std::filesystem::path(".")
                                            : std::filesystem::path(result).parent_path());
    std::ofstream out(result, std::ios::binary);
    if (!out) throw std::runtime_error("Cannot write web result: " + result);
    writeHeader(out, "HIP GPU", scene);
    out << "  \"quantity\": \"hemispherical radiosity\",\n  \"units\": \"W m-2\",\n"
        << "  \"facetCount\": " << scene.facetCount() << ",\n  \"leafFacetCount\": " << scene.leafFacetCount
        << ",\n";
    writeVertices(out, scene);
    out << "  \"graph\": {\"directedEdges\":" << graph.directedEdgeCount << ",\"closureError\":"
        << graph.maxClosureError << ",\"fragmentOverflow\":" << graph.fragmentOverflow
        << ",\"hashOverflow\":" << graph.hashOverflow << "},\n"
        << "  \"timing\": {\"totalMs\":" << ms(t0, t1) << ",\"visibilityMs\":" << ms(tv, tv1)
        << ",\"sunlightMs\":" << ms(tv1, ts1) << ",\"solveMs\":" << ms(ts1, t1) << "},\n"
        << "  \"iterations\": " << sol.iterations << ",\n  \"maxDelta\": " << sol.maxDelta << ",\n";
    writeArray(out, "wavelengths", {}, true);
    writeArray(out, "sunlit", sunlit, true);
    writeArray(out, "radiosity", sol.radiosity, true);
    writeArray(out, "lightEnhancement", enh, false);
    out << "}\n";
    out.close();
    std::cout << "RESULT\t" << result << std::endl;
    return 0;
} catch (const std::exception& e) {
    std::cerr << "facetrt: " << e.what() << '\n';
    return 1;
}

int runFacetEB(const std::string& input, const std::string& outFile) try {
    const std::string result = outputPath(input, outFile, "faceteb.json");
    const Scene scene = loadObj(input);
    facethip::FacetrtHip model;
    model.initialize(engineConfig());
    model.setGeometry(scene.vertices);
    const facethip::GraphDiagnostics graph = model.buildVisibilityGraph({facethip::Direction{}});
    fh_context* ctx = model.handle();
    // object material = SurfaceParameters{}, ground = {0.20},{0.0},0.05 (core:113-143, 437-440)
    fh_eb_material mats[2] = {};
    for (int i = 0; i < 2; ++i) {
        mats[i].tirReflectance = i == 0 ? 0.03f : 0.05f;
        mats[i].tirTransmittance = 0.0f;
        mats[i].surfaceResistance = 120.0f;
        mats[i].specificHeat = 1180.0f;
        mats[i].density = 1800.0f;
        mats[i].thermalConductivity = 1.55f;
        mats[i].soilMoisture = 0.25f;
        mats[i].heatCapacityPerArea = 60000.0f;
        mats[i].convectiveScale = 1.0f;
        mats[i].soilTemperatureMethod = 1;
        mats[i].vegetation = 1;
    }
    const float swR[2] = {0.10f, 0.20f}, swT[2] = {0.05f, 0.0f};
    facethip::check(fh_eb_set_scene(ctx, scene.leafFacetCount, scene.material.data(), 2, mats, 1, swR, swT), ctx);
    fh_eb_node_params p{};
    p.sun = {0.0f, -1.0f, 0.0f};  // DOY 1 00:00: the sun is below the horizon
    p.airTemperatureK = std::clamp(25.0f + 273.15f, 220.0f, 340.0f);
    p.vaporPressure = 15.0f;
    p.pressure = 839.0f;
    p.windSpeed = 1.0f;
    p.skyLongwave = 320.0f;
    p.dTime = 1800.0f;
    p.temperatureTolerance = 0.05f;
    p.temperatureRelaxation = 0.5f;
    p.maxCoupling = 20;
    p.radiosityTolerance = 1.0e-5f;
    p.maxSweeps = 2000;
    p.precision = FH_PRECISION_FP32;
    p.warmStart = FH_WARM_PREVIOUS;
    p.firstNode = 1;
    const float zero = 0.0f;
    fh_eb_node_stats st{};
    std::cout << "PROGRESS\t10\tCoupled facet radiation and energy balance" << std::endl;
    facethip::check(fh_eb_run_node(ctx, &p, &zero, &zero, &st), ctx);
    const size_t S = scene.facetCount() * size_t(2);
    std::vector<float> T(S), Blw(S), sunlit(S), Rn(S), H(S), LE(S), G(S);
    facethip::check(fh_eb_get_state(ctx, T.data(), nullptr, Blw.data(), sunlit.data(), Rn.data(), H.data(),
                                    LE.data(), G.data(), nullptr),
                    ctx);
    std::ofstream out(result, std::ios::binary);
    if (!out) throw std::runtime_error("FacetEB cannot write result: " + result);
    writeHeader(out, "HIP GPU Facet RT-EB coupled", scene);
    out << "  \"componentFile\": \".facet_steps/components.bin\",\n"
        << "  \"facetCount\": " << scene.facetCount() << ",\n  \"leafFacetCount\": " << scene.leafFacetCount
        << ",\n  \"time\": \"DOY1_00-00\",\n";
    writeVertices(out, scene);
    out << "  \"graph\": {\"directedEdges\":" << graph.directedEdgeCount << ",\"closureError\":"
        << graph.maxClosureError << ",\"fragmentOverflow\":0,\"hashOverflow\":0},\n"
        << "  \"iterations\": " << st.lwSweeps << ",\n  \"maxDelta\": 0,\n"
        << "  \"couplingIterations\": " << st.couplingIterations << ",\n"
        << "  \"temperatureDelta\": " << st.maxTemperatureDelta << ",\n"
        << "  \"converged\": " << (st.converged ? "true" : "false") << ",\n"
        << "  \"energyResidualUnits\": \"W m-2\",\n"
        << "  \"energyResidualTemperatureEquivalent\": " << st.maxResidualEquivalent << ",\n"
        << "  \"temperatureTolerance\": " << p.temperatureTolerance << ",\n"
        << "  \"iterationBudget\": " << p.maxCoupling << ",\n";
    writeArray(out, "sunlit", sunlit, true);
    writeArray(out, "radiosity", Blw, true);
    writeArray(out, "lightEnhancement", std::vector<float>(S, 0.0f), true);
    writeArray(out, "temperature", T, true);
    writeArray(out, "netRadiation", Rn, true);
    writeArray(out, "sensibleHeat", H, true);
    writeArray(out, "latentHeat", LE, true);
    writeArray(out, "storageHeat", G, true);
    writeArray(out, "photovoltaicPower", std::vector<float>(S, 0.0f), false);
    out << "}\n";
    out.close();
    std::cout << "FACET_CONVERGENCE\t0\t" << (st.converged ? 1 : 0) << '\t' << st.couplingIterations << '\t'
              << st.maxTemperatureDelta << '\t' << st.maxResidualEquivalent << std::endl;
    std::cout << "RESULT\t" << result << std::endl;
    return st.converged ? 0 : 2;
} catch (const std::exception& e) {
    std::cerr << "faceteb: " << e.what() << '\n';
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "";
    const std::string input = argc > 2 ? argv[2] : "";
    const std::string output = argc > 3 ? argv[3] : "";
    if (mode != "eFacetRT" && mode != "eFacetEB") {
        std::cerr << "Usage: facet_engine [eFacetRT|eFacetEB] [scene.obj] [output.json]\n";
        return 2;
    }
    if (input.empty()) {
        std::cerr << (mode == "eFacetRT" ? "facetrt" : "faceteb") << ": an input scene is required\n";
        return 1;
    }
    const std::string ext = std::filesystem::path(input).extension().string();
    if (ext != ".obj" && ext != ".OBJ") {
        std::cerr << (mode == "eFacetRT" ? "facetrt" : "faceteb")
                  << ": only OBJ scenes are read by this engine (project.json is not supported)\n";
        return 1;
    }
    return mode == "eFacetRT" ? runFacetRT(input, output) : runFacetEB(input, output);
}
