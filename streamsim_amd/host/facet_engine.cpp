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
#include <ctime>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "facet_backend.hpp"
#include "../csrc/hapke.hpp"
#include "facet_hip.h"
#include "project.hpp"
#include "scene.hpp"

namespace {

using facetcli::appendTriangle;
using facetcli::Scene;
using facetcli::Vec3;
using Clock = std::chrono::steady_clock;

constexpr float kPi = 3.14159265358979323846f;
constexpr float kBeam = 700.0f;
constexpr float kSky = 100.0f;
constexpr uint32_t kSoilGrid = 48;

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
    s.extentX = x1 - x0;
    s.extentZ = z1 - z0;
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

// FACET_RAYS_PER_SIDE, FACET_SUN_RAYS_PER_SIDE: ray counts; FACET_PERIODIC_NEIGHBORS:
// periodic copies of the scene (0..20, the reference's periodicNeighborCount)
// with the soil grid's extent as the period (periodicSizeX/Z = scene width /
// depth, core:2890-2894)
// (a project's periodicNeighborCount is the default of the variable)
// FACET_VISIBILITY=raster: the reference's raster graph and sunlit fraction
// over the 40 hemisphere directions, at the scene's visibility raster
// (1024 pixels square and 64 fragments per pixel, kRasterSize /
// kRasterLayers, core:75-76; a large project's 2048 and adapted layers,
// core:168-188, 2895-2901), or FACET_RASTER_SIZE pixels square and
// FACET_RASTER_LAYERS fragments per pixel
facethip::Config engineConfig(const Scene& s, uint32_t periodic = 0) {
    facethip::Config c;
    c.visibilityMethod = facethip::VisibilityMethod::RayCast;
    if (const char* e = std::getenv("FACET_VISIBILITY"); e && std::string(e) == "raster") {
        c.visibilityMethod = facethip::VisibilityMethod::Raster;
        c.rasterWidth = c.rasterHeight = s.rasterSize;
        c.maxFragmentsPerPixel = s.rasterLayers;
        if (const char* r = std::getenv("FACET_RASTER_SIZE")) c.rasterWidth = c.rasterHeight = uint32_t(std::atoi(r));
        if (const char* l = std::getenv("FACET_RASTER_LAYERS")) c.maxFragmentsPerPixel = uint32_t(std::atoi(l));
    }
    if (const char* e = std::getenv("FACET_RAYS_PER_SIDE")) c.raysPerSide = uint32_t(std::atoi(e));
    if (const char* e = std::getenv("FACET_SUN_RAYS_PER_SIDE")) c.sunRaysPerSide = uint32_t(std::atoi(e));
    if (const char* e = std::getenv("FACET_PERIODIC_NEIGHBORS")) periodic = uint32_t(std::clamp(std::atoi(e), 0, 20));
    if (periodic) {
        c.periodicNeighborCount = periodic;
        c.periodicSizeX = s.extentX;
        c.periodicSizeZ = s.extentZ;
    }
    return c;
}

// makeHemisphereDirections (core:475-489): 40 golden-angle directions over
// the upper hemisphere
std::vector<facethip::Direction> hemisphereDirections() {
    const uint32_t n = 40u;
    const float goldenAngle = 2.39996322972865332f;
    std::vector<facethip::Direction> d;
    d.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float y = (float(i) + 0.5f) / float(n);
        const float radius = std::sqrt(std::max(0.0f, 1.0f - y * y));
        const float azimuth = goldenAngle * float(i);
        d.push_back({radius * std::cos(azimuth), y, radius * std::sin(azimuth)});
    }
    return d;
}

bool isProject(const std::string& input);
int runFacetRTProject(const std::string& input, const std::string& outFile);

int runFacetRT(const std::string& input, const std::string& outFile) try {
    if (isProject(input)) return runFacetRTProject(input, outFile);
    const std::string result = outputPath(input, outFile, "facetrt.json");
    const Scene scene = loadObj(input);
    const auto t0 = Clock::now();
    std::cout << "PROGRESS\t3\tCreating HIP device context" << std::endl;
    facethip::FacetrtHip model;
    model.initialize(engineConfig(scene));
    model.setGeometry(scene.vertices);
    std::cout << "PROGRESS\t10\tVisibility graph" << std::endl;
    const auto tv = Clock::now();
    const facethip::GraphDiagnostics graph = model.buildVisibilityGraph(hemisphereDirections());
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

// planckRadiance (core:524-536): W m-2 sr-1 um-1
float planckRadiance(float wavelengthNm, float temperature) {
    constexpr double c1 = 1.1910439340652e8, c2 = 14388.291040407;
    const double um = wavelengthNm > 50.0f ? double(wavelengthNm) / 1000.0 : double(wavelengthNm);
    const double x = c2 / (std::max(1.0, double(temperature)) * std::max(1.0e-6, um));
    return float(c1 / (std::pow(um, 5.0) * std::expm1(x)));
}

// eFacetRT on a project.json (runGpu's band loop, core:1672-1774): per
// sensor band the optics of makeSpectralOptics (core:1057-1126, built on the
// device): optical bands r, t of the band (r times the texture factor, the
// Hapke ground), D = sunlit max(0, n.s) 700, sky 50; thermal bands (> 2500 nm) r, t = ref_TIR,
// tau_TIR, emission e = eps (f Planck(T_sunlit) + (1 - f) Planck(T_shaded)),
// sky Planck(T_sky).  All bands in one 64-sweep multi-band solve (the
// reference's kIterations sweeps per band); optical results as apparent
// reflectance against 700 cos(zenith) + 50, clamped to [0, 1]
int runFacetRTProject(const std::string& input, const std::string& outFile) try {
    const facetcli::Project prj = facetcli::loadProject(input);
    const Scene& scene = prj.scene;
    const std::vector<float>& bands = prj.run.bands;
    const uint32_t nb = uint32_t(bands.size());
    const std::string result = outputPath(input, outFile, "facetrt.json");
    const auto t0 = Clock::now();
    std::cout << "PROGRESS\t3\tCreating HIP device context" << std::endl;
    facethip::FacetrtHip model;
    model.initialize(engineConfig(scene, prj.run.periodicNeighbors));
    model.setGeometry(scene.vertices);
    std::cout << "PROGRESS\t10\tVisibility graph" << std::endl;
    const auto tv = Clock::now();
    const facethip::GraphDiagnostics graph = model.buildVisibilityGraph(hemisphereDirections());
    const auto tv1 = Clock::now();
    const facethip::Direction sun = sunDirection(prj.run.sunZenith, prj.run.sunAzimuth);
    const std::vector<float> sunlit = model.computeSunlitFraction(sun);
    const auto ts1 = Clock::now();
    const size_t S = size_t(scene.facetCount()) * 2u;
    std::vector<float> r(S * nb), t(S * nb), e(S * nb, 0.0f), D(S * nb, 0.0f), sky(nb);
    for (uint32_t b = 0; b < nb; ++b)
        sky[b] = bands[b] > 2500.0f ? planckRadiance(bands[b], prj.run.skyTemperature) : kSky * 0.5f;
    // makeSpectralOptics on the device: per material and band the band's r,
    // t (ref_TIR, tau_TIR above 2500 nm), the materials' sunlit / shaded
    // temperatures, the texture factors, 700 W m-2 normal direct
    const size_t nm = prj.materials.size();
    std::vector<float> matR(nm * nb), matT(nm * nb), tSun(nm), tShade(nm), dni(nb);
    for (size_t m = 0; m < nm; ++m) {
        const facetcli::ProjectMaterial& pm = prj.materials[m];
        tSun[m] = pm.sunlitT;
        tShade[m] = pm.shadedT;
        for (uint32_t b = 0; b < nb; ++b) {
            const bool thermal = bands[b] > 2500.0f;
            matR[m * nb + b] = thermal ? pm.tirR : pm.bandR.empty() ? pm.r : pm.bandR[std::min<size_t>(b, pm.bandR.size() - 1)];
            matT[m * nb + b] = thermal ? pm.tirT : pm.bandT.empty() ? pm.t : pm.bandT[std::min<size_t>(b, pm.bandT.size() - 1)];
        }
    }
    for (uint32_t b = 0; b < nb; ++b) dni[b] = bands[b] > 2500.0f ? 0.0f : kBeam;
    fh_spectral_optics so{};
    so.mode = FH_OPTICS_SCENE;
    so.bands = nb;
    so.materialCount = uint32_t(nm);
    so.leafFacetCount = scene.leafFacetCount;
    so.facetMaterial = scene.material.data();
    so.reflectance = matR.data();
    so.transmittance = matT.data();
    so.wavelength = bands.data();
    so.reflectanceFactor = scene.reflectanceFactor.empty() ? nullptr : scene.reflectanceFactor.data();
    so.backgroundAngularStrength = scene.backgroundAngularStrength;
    so.sun = {sun.x, sun.y, sun.z};
    so.sunlit = sunlit.data();
    so.directNormal = dni.data();
    so.sunlitTemperature = tSun.data();
    so.shadedTemperature = tShade.data();
    facethip::check(fh_spectral_optics_build(model.handle(), &so, r.data(), t.data(), e.data(), D.data()),
                    model.handle());
    fh_context* ctx = model.handle();
    std::vector<float> out(S * nb);
    fh_solve_stats st{};
    // one band's optics as the reference's AoS vector (accelerated solves)
    const auto bandOptics = [&](uint32_t b, bool direct) {
        std::vector<fh_surface_optics> o(S);
        for (size_t s = 0; s < S; ++s)
            o[s] = {r[s * nb + b], t[s * nb + b], e[s * nb + b], direct ? D[s * nb + b] : 0.0f};
        return o;
    };
    if (prj.run.acceleratedSolver) {
        // radiationSolver "accelerated" (core:1706-1724): per band
        // solveAccelerated(optics, sky, previous band's radiosity when it is
        // of the same kind, 64, batches of 8, |dB| <= 1e-4)
        std::cout << "PROGRESS\t82\tHIP accelerated spectral solve: " << nb << " bands" << std::endl;
        std::vector<float> prev, cur(S);
        bool prevThermal = false;
        for (uint32_t b = 0; b < nb; ++b) {
            const bool thermal = bands[b] > 2500.0f;
            const std::vector<fh_surface_optics> o = bandOptics(b, true);
            const float* init = !prev.empty() && prevThermal == thermal ? prev.data() : nullptr;
            fh_solve_stats sb{};
            facethip::check(fh_solve_accelerated(ctx, o.data(), sky[b], init, 64u, 8u, 1.0e-4f, 1.0f, cur.data(), &sb),
                            ctx);
            for (size_t s = 0; s < S; ++s) out[s * nb + b] = cur[s];
            if (b == 0) st = sb;
            prev = cur;
            prevThermal = thermal;
        }
    } else {
        std::cout << "PROGRESS\t82\tHIP Jacobi radiosity: " << nb << " bands, 64 iterations" << std::endl;
        facethip::check(fh_solve_spectral(ctx, &so, sky.data(), nullptr, 64u, -1.0f, 1.0f, FH_PRECISION_FP32,
                                          out.data(), &st),
                        ctx);
    }
    const float ref = kBeam * std::max(0.0f, sun.y) + kSky * 0.5f;
    const auto toReflectance = [&](float v) { return std::clamp(v / ref, 0.0f, 1.0f); };
    bool anyThermal = false, allThermal = true;
    std::vector<float> band0(S), enh(S, 0.0f), bandRad;
    for (uint32_t b = 0; b < nb; ++b) {
        const bool thermal = bands[b] > 2500.0f;
        anyThermal = anyThermal || thermal;
        allThermal = allThermal && thermal;
    }
    if (nb > 1) bandRad.resize(S * nb);
    for (uint32_t b = 0; b < nb; ++b) {
        const bool conv = bands[b] <= 2500.0f && ref > 1.0e-8f;
        for (size_t s = 0; s < S; ++s) {
            float v = out[s * nb + b];
            if (conv) v = toReflectance(v);
            if (b == 0) band0[s] = v;
            if (nb > 1) bandRad[size_t(b) * S + s] = v;
        }
    }
    if (bands[0] <= 2500.0f) {  // light enhancement of the first band (a diffuse-only solve)
        std::vector<float> r0(S), t0(S), e0(S), D0(S, 0.0f), d(S);
        for (size_t s = 0; s < S; ++s) {
            r0[s] = r[s * nb];
            t0[s] = t[s * nb];
            e0[s] = e[s * nb];
        }
        fh_solve_stats sd{};
        if (prj.run.acceleratedSolver) {
            const std::vector<fh_surface_optics> o = bandOptics(0, false);
            facethip::check(fh_solve_accelerated(ctx, o.data(), sky[0], nullptr, 64u, 8u, 1.0e-4f, 1.0f, d.data(), &sd),
                            ctx);
        } else {
            facethip::check(fh_solve_bands(ctx, 1u, r0.data(), t0.data(), e0.data(), D0.data(), sky.data(), nullptr,
                                           64u, -1.0f, 1.0f, FH_PRECISION_FP32, d.data(), &sd),
                            ctx);
        }
        for (size_t s = 0; s < S; ++s) enh[s] = std::max(0.0f, band0[s] - (ref > 1.0e-8f ? toReflectance(d[s]) : d[s]));
    }
    const auto t1 = Clock::now();
    std::filesystem::create_directories(std::filesystem::path(result).parent_path().empty()
                                            ? std::filesystem::path(".")
                                            : std::filesystem::path(result).parent_path());
    std::ofstream o(result, std::ios::binary);
    if (!o) throw std::runtime_error("Cannot write web result: " + result);
    writeHeader(o, "HIP GPU", scene);
    o << "  \"quantity\": \""
      << (allThermal ? "spectral radiance" : !anyThermal ? "reflectance" : "reflectance and spectral radiance")
      << "\",\n  \"units\": \"" << (allThermal ? "W m-2 sr-1 um-1" : !anyThermal ? "1" : "band-dependent")
      << "\",\n  \"facetCount\": " << scene.facetCount() << ",\n  \"leafFacetCount\": " << scene.leafFacetCount
      << ",\n";
    writeVertices(o, scene);
    o << "  \"graph\": {\"directedEdges\":" << graph.directedEdgeCount << ",\"closureError\":" << graph.maxClosureError
      << ",\"fragmentOverflow\":" << graph.fragmentOverflow << ",\"hashOverflow\":" << graph.hashOverflow << "},\n"
      << "  \"timing\": {\"totalMs\":" << ms(t0, t1) << ",\"visibilityMs\":" << ms(tv, tv1)
      << ",\"sunlightMs\":" << ms(tv1, ts1) << ",\"solveMs\":" << ms(ts1, t1) << "},\n"
      << "  \"iterations\": " << st.iterations << ",\n  \"maxDelta\": " << st.maxDelta << ",\n";
    writeArray(o, "wavelengths", bands, true);
    writeArray(o, "sunlit", sunlit, true);
    writeArray(o, "radiosity", band0, true);
    writeArray(o, "lightEnhancement", enh, !bandRad.empty());
    if (!bandRad.empty()) writeArray(o, "bandRadiosity", bandRad, false);
    o << "}\n";
    o.close();
    std::cout << "RESULT\t" << result << std::endl;
    return 0;
} catch (const std::exception& ex) {
    std::cerr << "facetrt: " << ex.what() << '\n';
    return 1;
}

// ---- FacetEB run settings ---------------------------------------------------
// The reference reads these from the input's XML tags (core:3193-3240); an
// OBJ input carries none, so they come from the environment here, with the
// reference's defaults and clamps:
//   FACET_EB_START / FACET_EB_END      startTimeNode / endTimeNode (0 / 1)
//   FACET_EB_DTIME                     dTime, s (1800)
//   FACET_EB_METEO                     meteorology file (filePath): a header
//                                      line, then rows "julianTime Tair[C]
//                                      vapourPressure[hPa] pressure[hPa]
//                                      wind[m/s] shortwave[W/m2] longwave[W/m2]"
//   FACET_EB_LATITUDE / _LONGITUDE     Latitude / Longitude (40 / 116)
//   FACET_EB_PROCESS                   energy | radiation | both
//                                      (isEnergyProcess / isRadiationProcess)
//   FACET_EB_TOLERANCE, FACET_EB_COUPLING, FACET_EB_RELAXATION
//                                      temperatureTolerance (0.05), couplingIterations (20),
//                                      temperatureRelaxation (0.5)
//   FACET_EB_SKY_TEMPERATURE           skyTemperature, K (250)
//   FACET_EB_DIRECT_RATIO              directScatteringRatio (0.8)
//   FACET_EB_SPECTRAL_WIDTH            spectralAccelerationWidth, 1-nm samples (100)
//   STREAMSIM_FACET_STEP_SYNC=1        FACET_STEP / FACET_ACK back-pressure (core:3438-3441)
// and two switches of this engine (both off: the reference's numerics):
//   FACET_EB_ADAPTIVE_RELAXATION=1     full Newton steps, the relaxation only where a step reverses
//   FACET_EB_BF16=1                    shortwave solves gather BF16 increments (FP32 fixed point)
double envNumber(const char* name, double fallback) {
    const char* v = std::getenv(name);
    if (!v || !*v) return fallback;
    char* end = nullptr;
    const double x = std::strtod(v, &end);
    return (end != v && std::isfinite(x)) ? x : fallback;
}

struct EbSettings {
    int startNode = 0, endNode = 1;
    float dTime = 1800.0f;
    float latitude = 40.0f, longitude = 116.0f;
    float skyTemperature = 250.0f, directRatio = 0.8f;
    float tolerance = 0.05f, relaxation = 0.5f;
    uint32_t couplingIterations = 20;
    uint32_t spectralWidth = 100;
    bool saveEnergy = false, saveRadiation = false, stepSync = false;
    bool adaptiveRelaxation = false, bf16 = false;
    int vegetationMethod = 0;
    std::string meteoFile;
};

int envSwitches(EbSettings& s);

// the settings a project.json gives (facetInputView's tags); step sync and
// this engine's switches from the environment as for OBJ input
EbSettings projectSettings(const facetcli::ProjectRun& r) {
    EbSettings s;
    s.startNode = r.startNode;
    s.endNode = r.endNode;
    s.dTime = r.dTime;
    s.latitude = r.latitude;
    s.longitude = r.longitude;
    s.skyTemperature = r.skyTemperature;
    s.directRatio = r.directRatio;
    s.tolerance = r.tolerance;
    s.relaxation = r.relaxation;
    s.couplingIterations = r.couplingIterations;
    s.spectralWidth = r.spectralWidth;
    s.saveEnergy = r.saveEnergy;
    s.saveRadiation = r.saveRadiation;
    s.meteoFile = r.meteoFile;
    s.vegetationMethod = r.vegetationMethod;
    envSwitches(s);
    return s;
}

int envSwitches(EbSettings& s) {
    const char* sync = std::getenv("STREAMSIM_FACET_STEP_SYNC");
    s.stepSync = sync && std::string(sync) == "1";
    s.adaptiveRelaxation = envNumber("FACET_EB_ADAPTIVE_RELAXATION", 0.0) != 0.0;
    s.bf16 = envNumber("FACET_EB_BF16", 0.0) != 0.0;
    return 0;
}

EbSettings ebSettings() {
    EbSettings s;
    s.startNode = std::max(0, int(std::lround(envNumber("FACET_EB_START", 0.0))));
    s.endNode = std::max(s.startNode + 1, int(std::lround(envNumber("FACET_EB_END", s.startNode + 1))));
    s.dTime = std::max(1.0f, float(envNumber("FACET_EB_DTIME", 1800.0)));
    s.latitude = float(envNumber("FACET_EB_LATITUDE", 40.0));
    s.longitude = float(envNumber("FACET_EB_LONGITUDE", 116.0));
    s.skyTemperature = std::clamp(float(envNumber("FACET_EB_SKY_TEMPERATURE", 250.0)), 150.0f, 350.0f);
    s.directRatio = std::clamp(float(envNumber("FACET_EB_DIRECT_RATIO", 0.8)), 0.0f, 1.0f);
    s.tolerance = std::clamp(float(envNumber("FACET_EB_TOLERANCE", 0.05)), 0.001f, 10.0f);
    s.relaxation = std::clamp(float(envNumber("FACET_EB_RELAXATION", 0.5)), 0.05f, 1.0f);
    s.couplingIterations = uint32_t(std::clamp(int(std::lround(envNumber("FACET_EB_COUPLING", 20.0))), 1, 100));
    s.spectralWidth = uint32_t(std::clamp(int(std::lround(envNumber("FACET_EB_SPECTRAL_WIDTH", 100.0))), 1, 1000));
    if (const char* v = std::getenv("FACET_EB_PROCESS")) {
        const std::string p = v;
        s.saveEnergy = p == "energy" || p == "both";
        s.saveRadiation = p == "radiation" || p == "both";
    }
    if (const char* v = std::getenv("FACET_EB_METEO")) s.meteoFile = v;
    envSwitches(s);
    return s;
}

// Direct (Esun) and diffuse (Esky) shortwave spectra, 1-nm samples from
// 400 nm, for the energy windows and the sensor bands (ShortwaveSpectrum and
// loadShortwaveSpectrum, core:2458-2551).  A project run reads the two files
// its input names, Esun_.dat and Esky_.dat in the executable's "defined"
// folder (facetInputView, core:1930-1931, base/projectjson.cpp:108); this
// engine looks there too, or in $STREAMSIM_DEFINED_DIR.  Each file is a list
// of numbers (negative or non-finite ones read as 0, at most 2001 kept);
// both are cut to the shorter and scaled together so that the trapezoid
// integral of direct + diffuse is one per micrometre.  Without two readable
// spectra of at least two samples, or with a zero integral, the flat
// spectrum 0.5 directRatio / 0.5 (1 - directRatio) over 2001 samples is
// used (its integral is one too), as for an OBJ input, which names no files.
struct ShortwaveSpectrum {
    static constexpr size_t kMaxSamples = 2001;
    std::vector<float> direct, diffuse;
    std::string source;  // the folder the spectra came from; empty: flat

    size_t count() const { return std::min(direct.size(), diffuse.size()); }
    // the sample nearest a wavelength in nm (micrometres when <= 2.5)
    static float at(const std::vector<float>& v, float wavelength) {
        if (v.empty()) return 0.0f;
        const float nm = wavelength <= 2.5f ? wavelength * 1000.0f : wavelength;
        const long k = std::clamp(std::lround(nm - 400.0f), 0L, long(v.size()) - 1L);
        return v[size_t(k)];
    }
    float directAt(float wavelength) const { return at(direct, wavelength); }
    float diffuseAt(float wavelength) const { return at(diffuse, wavelength); }
    // trapezoid integral over samples [a, b], per micrometre
    static float integrate(const std::vector<float>& v, size_t a, size_t b) {
        if (v.size() < 2 || a >= v.size() - 1) return 0.0f;
        b = std::min(b, v.size() - 1);
        double sum = 0.0;
        for (size_t k = a; k < b; ++k) sum += 0.5 * double(v[k] + v[k + 1]);
        return float(sum * 0.001);
    }

    static std::vector<float> readColumn(const std::filesystem::path& file) {
        std::ifstream in(file);
        std::vector<float> v;
        float x = 0.0f;
        while (v.size() < kMaxSamples && in >> x) v.push_back(std::isfinite(x) ? std::max(0.0f, x) : 0.0f);
        return v;
    }
    static ShortwaveSpectrum flat(float directRatio) {
        ShortwaveSpectrum s;
        s.direct.assign(kMaxSamples, 0.5f * directRatio);
        s.diffuse.assign(kMaxSamples, 0.5f * (1.0f - directRatio));
        return s;
    }
    static ShortwaveSpectrum load(const std::filesystem::path& folder, float directRatio) {
        ShortwaveSpectrum s;
        s.direct = readColumn(folder / "Esun_.dat");
        s.diffuse = readColumn(folder / "Esky_.dat");
        const size_t n = s.count();
        if (n < 2) return flat(directRatio);
        s.direct.resize(n);
        s.diffuse.resize(n);
        double total = 0.0;
        for (size_t k = 0; k + 1 < n; ++k)
            total += 0.5 * double(s.direct[k] + s.direct[k + 1] + s.diffuse[k] + s.diffuse[k + 1]);
        total *= 0.001;
        if (!(total > 0.0) || !std::isfinite(total)) return flat(directRatio);
        const float inv = float(1.0 / total);
        for (float& x : s.direct) x *= inv;
        for (float& x : s.diffuse) x *= inv;
        s.source = folder.string();
        return s;
    }
};

std::filesystem::path definedDirectory() {
    if (const char* v = std::getenv("STREAMSIM_DEFINED_DIR"); v && *v) return v;
    std::error_code ec;
    const std::filesystem::path exe = std::filesystem::read_symlink("/proc/self/exe", ec);
    return ec ? std::filesystem::path("defined") : exe.parent_path() / "defined";
}

// Energy windows of spectralAccelerationWidth samples over the spectrum
// (core:3506-3524): samples [first, last] with last = min(first + width,
// count - 1), the window's optics at its centre sample, its direct and
// diffuse integrals; windows whose integrals sum to at most 1e-10 carry no
// energy and are not solved.
struct EnergyWindow {
    uint32_t first, last;
    float directFraction, diffuseFraction;
    uint32_t centre() const { return (first + last) / 2u; }
};

std::vector<EnergyWindow> energyWindows(const ShortwaveSpectrum& sp, uint32_t width) {
    std::vector<EnergyWindow> w;
    const size_t n = sp.count();
    for (size_t f = 0; f + 1 < n; f += width) {
        const size_t l = std::min(f + width, n - 1);
        const float d = ShortwaveSpectrum::integrate(sp.direct, f, l);
        const float q = ShortwaveSpectrum::integrate(sp.diffuse, f, l);
        if (d + q <= 1.0e-10f) continue;
        w.push_back({uint32_t(f), uint32_t(l), d, q});
    }
    return w;
}

struct MeteoStep {  // core:3172-3181
    int node = 0;
    float julianTime = 0.0f;
    float airTemperature = 25.0f;
    float vaporPressure = 15.0f;
    float pressure = 839.0f;
    float windSpeed = 1.0f;
    float shortwave = 0.0f;
    float longwave = 320.0f;
};

// meteorology rows (core:3240-3268): a file, else defaults at node times
std::vector<MeteoStep> loadMeteorology(const EbSettings& s) {
    std::vector<MeteoStep> rows;
    if (!s.meteoFile.empty()) {
        std::ifstream in(s.meteoFile);
        if (!in) throw std::runtime_error("FacetEB cannot open meteorology: " + s.meteoFile);
        std::string line;
        std::getline(in, line);
        int row = 0;
        while (std::getline(in, line)) {
            std::istringstream v(line);
            MeteoStep m;
            m.node = row++;
            if (!(v >> m.julianTime >> m.airTemperature >> m.vaporPressure >> m.pressure >> m.windSpeed >>
                  m.shortwave >> m.longwave))
                continue;
            rows.push_back(m);
        }
    }
    if (rows.empty()) {
        rows.resize(size_t(s.endNode));
        for (int n = 0; n < s.endNode; ++n) {
            rows[size_t(n)].node = n;
            rows[size_t(n)].julianTime = 1.0f + float(n) * s.dTime / 86400.0f;
        }
    }
    if (s.endNode > int(rows.size())) throw std::runtime_error("FacetEB time range exceeds meteorology rows");
    return rows;
}

// "DOY<day>_HH-MM" (core:3276-3288)
std::string timeToken(float julianTime) {
    int day = int(std::floor(julianTime));
    int minutes = int(std::lround((double(julianTime) - day) * 1440.0));
    if (minutes >= 1440) {
        day += minutes / 1440;
        minutes %= 1440;
    }
    std::ostringstream t;
    t << "DOY" << day << '_' << std::setw(2) << std::setfill('0') << minutes / 60 << '-' << std::setw(2)
      << std::setfill('0') << minutes % 60;
    return t.str();
}

// Solar zenith and azimuth (degrees; azimuth eastward from north) for a
// local clock time in a time zone hours east of UTC.  Low-precision solar
// coordinates after Meeus, Astronomical Algorithms (2nd ed.), ch. 7, 12,
// 25 (mean longitude and anomaly, equation of the centre, apparent longitude,
// obliquity, Greenwich mean sidereal time), with the atmospheric refraction
// correction of Reda & Andreas, Solar Position Algorithm (NREL/TP-560-34302,
// eq. 42).  The reference runs the full SPA (thirdparty/spa.cpp) with
// timezone 8 and elevation 100 m; this form agrees with SPA's published
// example to about 0.01 degree (test_host.py).
struct SunAngles {
    double zenith, azimuth;
};

SunAngles solarPosition(int year, int month, int day, int hour, int minute, double second, double timezone,
                        double longitude, double latitude, double pressure, double temperature) {
    constexpr double d2r = 3.14159265358979323846 / 180.0;
    int y = year, m = month;
    if (m <= 2) {
        y -= 1;
        m += 12;
    }
    const double a = std::floor(y / 100.0), b = 2.0 - a + std::floor(a / 4.0);
    const double dayFrac = day + (hour - timezone + minute / 60.0 + second / 3600.0) / 24.0;
    const double jd = std::floor(365.25 * (y + 4716)) + std::floor(30.6001 * (m + 1)) + dayFrac + b - 1524.5;
    const double t = (jd - 2451545.0) / 36525.0;
    const double l0 = std::fmod(280.46646 + t * (36000.76983 + t * 0.0003032), 360.0);
    const double ma = 357.52911 + t * (35999.05029 - 0.0001537 * t);
    const double c = (1.914602 - t * (0.004817 + 0.000014 * t)) * std::sin(ma * d2r) +
                     (0.019993 - 0.000101 * t) * std::sin(2.0 * ma * d2r) + 0.000289 * std::sin(3.0 * ma * d2r);
    const double omega = 125.04 - 1934.136 * t;
    const double lambda = l0 + c - 0.00569 - 0.00478 * std::sin(omega * d2r);
    const double eps0 = 23.0 + (26.0 + (21.448 - t * (46.815 + t * (0.00059 - t * 0.001813))) / 60.0) / 60.0;
    const double eps = eps0 + 0.00256 * std::cos(omega * d2r);
    const double ra = std::atan2(std::cos(eps * d2r) * std::sin(lambda * d2r), std::cos(lambda * d2r));
    const double dec = std::asin(std::sin(eps * d2r) * std::sin(lambda * d2r));
    const double gmst = 280.46061837 + 360.98564736629 * (jd - 2451545.0) + t * t * (0.000387933 - t / 38710000.0);
    const double h = (gmst + longitude) * d2r - ra;
    const double lat = latitude * d2r;
    const double sinE = std::sin(lat) * std::sin(dec) + std::cos(lat) * std::cos(dec) * std::cos(h);
    const double e0 = std::asin(std::clamp(sinE, -1.0, 1.0)) / d2r;
    double de = 0.0;
    if (e0 >= -(0.26667 + 0.5667))
        de = (pressure / 1010.0) * (283.0 / (273.0 + temperature)) * 1.02 /
             (60.0 * std::tan((e0 + 10.3 / (e0 + 5.11)) * d2r));
    const double az = std::atan2(std::sin(h), std::cos(h) * std::sin(lat) - std::tan(dec) * std::cos(lat)) / d2r;
    return {90.0 - (e0 + de), std::fmod(az + 180.0 + 360.0, 360.0)};
}

// calendar day of a 2019 day of year (the reference's Utils::calculateMonthAndDay(2019, ...))
void monthAndDay(int dayOfYear, int& month, int& day) {
    static const int len[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    int d = std::max(1, dayOfYear);
    month = 1;
    while (month <= 12 && d > len[month - 1]) {
        d -= len[month - 1];
        ++month;
    }
    if (month > 12) {
        month = 12;
        d = 31;
    }
    day = d;
}

// updateSolarPosition (core:3290-3330): whole-minute local time, timezone 8
SunAngles nodeSun(const MeteoStep& m, const EbSettings& s) {
    const int doy = int(std::floor(m.julianTime));
    int month = 1, day = 1;
    monthAndDay(doy, month, day);
    const double hours = (double(m.julianTime) - doy) * 24.0;
    const int hour = int(std::floor(hours));
    const int minute = int(std::floor((hours - hour) * 60.0));
    return solarPosition(2019, month, day, std::clamp(hour, 0, 23), std::clamp(minute, 0, 59), 0.0, 8.0,
                         s.longitude, s.latitude, std::max(100.0f, m.pressure), m.airTemperature);
}

std::string utcTimestamp() {
    const std::time_t now = std::time(nullptr);
    std::tm utc{};
    gmtime_r(&now, &utc);
    std::ostringstream t;
    t << std::put_time(&utc, "%Y-%m-%dT%H:%M:%SZ");
    return t.str();
}

std::string jsonString(const std::string& v) {
    std::string o = "\"";
    for (char ch : v) {
        if (ch == '"' || ch == '\\') o += '\\';
        o += ch;
    }
    return o + "\"";
}

// faceteb_run.json (core:3213-3222, 4030-4042, 4143-4146): the index of this
// invocation's products and completed nodes, rewritten after every node
struct RunIndex {
    std::filesystem::path path;
    std::string startedAt, input;
    int startNode = 0, endNode = 1;
    float tolerance = 0.05f;
    uint32_t budget = 20;
    std::string status = "running", finishedAt;
    std::vector<std::string> products, completed;
    RunIndex() = default;
    RunIndex(const RunIndex&) = delete;
    RunIndex& operator=(const RunIndex&) = delete;
    ~RunIndex() {  // a run that ends by an exception is marked failed (core:66-72)
        if (status != "running" || path.empty()) return;
        status = "failed";
        finishedAt = utcTimestamp();
        try {
            write();
        } catch (...) {
        }
    }
    void write() const {
        std::ofstream o(path, std::ios::trunc);
        o << std::setprecision(9) << "{\n  \"kind\": \"faceteb-run\",\n  \"schemaVersion\": 1,\n  \"status\": "
          << jsonString(status) << ",\n  \"startedAt\": " << jsonString(startedAt) << ",\n";
        if (!finishedAt.empty()) o << "  \"finishedAt\": " << jsonString(finishedAt) << ",\n";
        o << "  \"input\": " << jsonString(input) << ",\n  \"requested\": {\"startNode\": " << startNode
          << ", \"endNodeExclusive\": " << endNode << "},\n  \"convergence\": {\"temperatureTolerance_K\": "
          << tolerance << ", \"iterationBudget\": " << budget
          << ", \"criterion\": \"temperatureDelta and residual/thermal-response <= temperatureTolerance\"},\n"
          << "  \"completed\": [";
        for (size_t i = 0; i < completed.size(); ++i) o << (i ? ",\n    " : "\n    ") << completed[i];
        o << (completed.empty() ? "" : "\n  ") << "],\n  \"products\": [";
        for (size_t i = 0; i < products.size(); ++i) o << (i ? ", " : "") << jsonString(products[i]);
        o << "]\n}\n";
        if (!o) throw std::runtime_error("FacetEB cannot write run index");
    }
};

// process/<model>_T=<token>.{bin,json} (core:3937-4019): 5 floats per side for
// energy [LE, H, G, T, Rn], 3 for radiation [B_SW, B_LW, Rn]
void writeProcess(const std::filesystem::path& dir, const std::string& type, int node, float julianTime,
                  const std::string& token, const SunAngles& sun, const Scene& scene, uint32_t iterations,
                  float dT, bool converged, float residual, float residualEq, const EbSettings& set,
                  const std::vector<float>* cols[5], RunIndex& index) {
    const std::string model = type == "radiation" ? "facetrt" : "faceteb";
    const std::string stem = model + (type == "photovoltaic" ? "_pv_T=" : "_T=") + token;
    const std::filesystem::path pdir = dir / "process";
    std::filesystem::create_directories(pdir);
    const int nf = type == "energy" ? 5 : 3;
    std::ofstream bin(pdir / (stem + ".bin"), std::ios::binary | std::ios::trunc);
    const size_t S = cols[0]->size();
    std::vector<float> rec(size_t(nf) * S);
    for (size_t s = 0; s < S; ++s)
        for (int k = 0; k < nf; ++k) rec[s * nf + k] = (*cols[k])[s];
    bin.write(reinterpret_cast<const char*>(rec.data()), std::streamsize(rec.size() * sizeof(float)));
    bin.close();
    if (!bin) throw std::runtime_error("FacetEB cannot finish process binary");
    std::ofstream js(pdir / (stem + ".json"), std::ios::trunc);
    js << std::setprecision(9) << "{\n  \"kind\": \"facet-" << type << "-process\",\n  \"model\": \"" << model
       << "\",\n  \"processType\": \"" << type << "\",\n  \"geometry\": \"facet\",\n  \"node\": " << node
       << ",\n  \"julianTime\": " << julianTime << ",\n  \"time\": \"" << token << "\",\n  \"solarZenith\": "
       << float(sun.zenith) << ",\n  \"solarAzimuth\": " << float(sun.azimuth) << ",\n  \"facetCount\": "
       << scene.facetCount() << ",\n  \"surfaceCount\": " << S
       << ",\n  \"geometryFile\": \"../faceteb.json\",\n  \"componentFile\": \"../.facet_steps/components.bin\",\n"
       << "  \"dataFile\": \"" << stem << ".bin\",\n  \"dataType\": \"float32-little-endian\",\n"
       << "  \"layout\": \"surface-interleaved\",\n  \"recordFloats\": " << nf << ",\n  \"couplingIterations\": "
       << iterations << ",\n  \"temperatureDelta\": " << dT << ",\n  \"converged\": "
       << (converged ? "true" : "false") << ",\n  \"energyResidual\": " << residual
       << ",\n  \"energyResidualUnits\": \"W m-2\",\n  \"energyResidualTemperatureEquivalent\": " << residualEq
       << ",\n  \"temperatureTolerance\": " << set.tolerance << ",\n  \"iterationBudget\": "
       << set.couplingIterations << ",\n  \"fields\": ";
    if (type == "photovoltaic")
        js << "[{\"id\":\"photovoltaicPower\",\"label\":\"photovoltaic power [W m-2]\",\"offset\":0},"
              "{\"id\":\"temperature\",\"label\":\"temperature [K]\",\"offset\":1},"
              "{\"id\":\"absorbedShortwave\",\"label\":\"absorbed shortwave of the side [W m-2]\",\"offset\":2}]\n}\n";
    else if (type == "radiation")
        js << "[{\"id\":\"shortwaveRadiation\",\"label\":\"shortwave radiosity [W m-2]\",\"offset\":0},"
              "{\"id\":\"longwaveRadiation\",\"label\":\"longwave radiosity [W m-2]\",\"offset\":1},"
              "{\"id\":\"netRadiation\",\"label\":\"net radiation [W m-2]\",\"offset\":2}]\n}\n";
    else
        js << "[{\"id\":\"latentHeat\",\"label\":\"latent heat [W m-2]\",\"offset\":0},"
              "{\"id\":\"sensibleHeat\",\"label\":\"sensible heat [W m-2]\",\"offset\":1},"
              "{\"id\":\"surfaceHeatFlux\",\"label\":\"surface heat flux [W m-2]\",\"offset\":2},"
              "{\"id\":\"temperature\",\"label\":\"temperature [K]\",\"offset\":3},"
              "{\"id\":\"netRadiation\",\"label\":\"net radiation [W m-2]\",\"offset\":4}]\n}\n";
    index.products.push_back("process/" + stem + ".bin");
    index.products.push_back("process/" + stem + ".json");
}

bool isProject(const std::string& input) {
    std::string e = std::filesystem::path(input).extension().string();
    for (char& c : e) c = char(std::tolower(static_cast<unsigned char>(c)));
    return e == ".json";
}

int runFacetEB(const std::string& input, const std::string& outFile) try {
    // project.json: scene, materials and settings from the project; OBJ:
    // loadTreeWithSoil's scene with the default materials, settings from
    // the environment
    const bool project = isProject(input);
    facetcli::Project prj;
    if (project) prj = facetcli::loadProject(input);
    const EbSettings set = project ? projectSettings(prj.run) : ebSettings();
    const Scene scene = project ? prj.scene : loadObj(input);
    const std::vector<MeteoStep> meteo = loadMeteorology(set);
    const std::string result = outputPath(input, outFile, "faceteb.json");
    const std::filesystem::path outDir = std::filesystem::path(result).parent_path().empty()
                                             ? std::filesystem::path(".")
                                             : std::filesystem::path(result).parent_path();
    std::filesystem::create_directories(outDir);
    RunIndex index;
    index.path = outDir / "faceteb_run.json";
    index.startedAt = utcTimestamp();
    index.input = input;
    index.startNode = set.startNode;
    index.endNode = set.endNode;
    index.tolerance = set.tolerance;
    index.budget = set.couplingIterations;
    index.write();
    std::cout << "PROGRESS\t3\tLoading the facet scene and meteorology" << std::endl;

    facethip::FacetrtHip model;
    model.initialize(engineConfig(scene, project ? prj.run.periodicNeighbors : 0u));
    model.setGeometry(scene.vertices);
    const facethip::GraphDiagnostics graph = model.buildVisibilityGraph(hemisphereDirections());
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
    // shortwave spectrum and energy windows (core:3238-3239, 3495-3565): a
    // project's Esun / Esky files, else the flat spectrum
    const ShortwaveSpectrum spectrum = project ? ShortwaveSpectrum::load(definedDirectory(), set.directRatio)
                                               : ShortwaveSpectrum::flat(set.directRatio);
    if (project)
        std::cout << "PROGRESS\t9\tShortwave spectrum: "
                  << (spectrum.source.empty() ? std::string("flat (no Esun_.dat / Esky_.dat)") : spectrum.source)
                  << std::endl;
    const std::vector<EnergyWindow> windows = energyWindows(spectrum, set.spectralWidth);
    const uint32_t nw = uint32_t(windows.size());
    const size_t nSamples = spectrum.count();
    const float directIntegral =
        nSamples >= 2 ? ShortwaveSpectrum::integrate(spectrum.direct, 0, nSamples - 1) : 0.0f;
    std::vector<fh_eb_material> matList(mats, mats + 2);
    std::vector<float> swR(2 * size_t(nw)), swT(2 * size_t(nw));
    for (uint32_t w = 0; w < nw; ++w) {
        swR[w] = 0.10f;
        swT[w] = 0.05f;
        swR[nw + w] = 0.20f;
        swT[nw + w] = 0.0f;
    }
    if (project) {
        // a project's materials: objects' meshes, then the ground; window r,
        // t at the window's centre sample (core:3510-3529, 952-1003)
        const size_t nm = prj.materials.size();
        matList.assign(nm, fh_eb_material{});
        swR.assign(nm * nw, 0.0f);
        swT.assign(nm * nw, 0.0f);
        for (size_t m = 0; m < nm; ++m) {
            const facetcli::ProjectMaterial& pm = prj.materials[m];
            fh_eb_material& e = matList[m];
            e.tirReflectance = pm.tirR;
            e.tirTransmittance = pm.tirT;
            e.surfaceResistance = pm.rss;
            e.specificHeat = pm.cs;
            e.density = pm.rhos;
            e.thermalConductivity = pm.lambdas;
            e.soilMoisture = pm.smc;
            e.heatCapacityPerArea = pm.heatCapacity;
            e.convectiveScale = pm.convectiveScale;
            e.soilTemperatureMethod = pm.soilMethod;
            e.vegetation = pm.vegetation ? 1 : 0;
            e.photovoltaic = pm.pv ? 1 : 0;
            e.pvEta25 = pm.pvEta25;
            e.pvGamma = pm.pvGamma;
            e.pvBifaciality = pm.pvBifaciality;
            const bool leaf = m + 1 < nm;  // the last material is the ground's
            for (uint32_t w = 0; w < nw; ++w) {
                const size_t centre = windows[w].centre();
                const float r = pm.swR(centre, leaf);
                swR[m * nw + w] = r;
                swT[m * nw + w] = pm.swT(centre, leaf, r);
            }
        }
    }
    facethip::check(fh_eb_set_scene(ctx, scene.leafFacetCount, scene.material.data(), uint32_t(matList.size()),
                                    matList.data(), nw, swR.data(), swT.data()),
                    ctx);
    facethip::check(fh_eb_set_background_angular_strength(ctx, scene.backgroundAngularStrength), ctx);
    if (!scene.reflectanceFactor.empty())  // texture factors on the windows' reflectance (core:965-972)
        facethip::check(fh_eb_set_reflectance_factors(ctx, scene.reflectanceFactor.data()), ctx);

    // sensor bands solved after every node (core:3862-3908): the project's
    // sensor bands, else the default 550 nm (core:141, 2389-2400); per
    // material the band r, t (custom spectra; defaults 0.10 / 0.05 for
    // objects, 0.20 / 0.0 for the ground) and the thermal r, t
    struct SensorMaterial {
        std::vector<float> r, t;
        float tirR = 0.03f, tirT = 0.0f;
    };
    std::vector<SensorMaterial> sensorMats(2);
    sensorMats[0].r = {0.10f};
    sensorMats[0].t = {0.05f};
    sensorMats[1].r = {0.20f};
    sensorMats[1].t = {0.0f};
    sensorMats[1].tirR = 0.05f;
    std::vector<float> sensorBands{550.0f};
    if (project) {
        sensorBands = prj.run.bands;
        sensorMats.assign(prj.materials.size(), SensorMaterial{});
        for (size_t m = 0; m < prj.materials.size(); ++m) {
            const facetcli::ProjectMaterial& pm = prj.materials[m];
            sensorMats[m].r = pm.bandR.empty() ? std::vector<float>{pm.r} : pm.bandR;
            sensorMats[m].t = pm.bandT.empty() ? std::vector<float>{pm.t} : pm.bandT;
            sensorMats[m].tirR = pm.tirR;
            sensorMats[m].tirT = pm.tirT;
        }
    }
    const uint32_t nbs = uint32_t(sensorBands.size());

    const std::filesystem::path stepDir = outDir / ".facet_steps";
    std::filesystem::create_directories(stepDir);
    {  // components.bin (core:3418-3429): 1 for ground, the object's output component (0) otherwise
        std::ofstream comp(stepDir / "components.bin", std::ios::binary | std::ios::trunc);
        for (uint32_t f = 0; f < scene.facetCount(); ++f) {
            const uint8_t c = project ? prj.facetComponent[f] : f >= scene.leafFacetCount ? 1 : 0;
            comp.write(reinterpret_cast<const char*>(&c), 1);
        }
        if (!comp) throw std::runtime_error("FacetEB cannot write component identities");
    }
    // photovoltaic summary, one row per node with panels (core:3433-3436)
    std::ofstream pvSummary(outDir / "photovoltaic_summary.csv", std::ios::trunc);
    pvSummary << "node,time,area_m2,temperature_C,incident_W,absorbed_W,electric_W,net_longwave_W,sensible_W,"
                 "storage_W,max_residual_W_m2,energy_kWh\n";
    double pvEnergyKwh = 0.0;
    std::vector<uint8_t> pvFacet(scene.facetCount(), 0);
    std::vector<float> pvArea(scene.facetCount(), 0.0f), pvBif(scene.facetCount(), 0.0f);
    bool anyPv = false;
    for (uint32_t f = 0; f < scene.leafFacetCount && project; ++f) {
        const facetcli::ProjectMaterial& pm = prj.materials[scene.material[f]];
        if (!pm.pv) continue;
        anyPv = true;
        pvFacet[f] = 1;
        pvBif[f] = pm.pvBifaciality;
        const auto& a = scene.vertices[3 * size_t(f)].position;
        const auto& b = scene.vertices[3 * size_t(f) + 1].position;
        const auto& c = scene.vertices[3 * size_t(f) + 2].position;
        const float u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, w[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const float x = u[1] * w[2] - u[2] * w[1], y = u[2] * w[0] - u[0] * w[2], z = u[0] * w[1] - u[1] * w[0];
        pvArea[f] = 0.5f * std::sqrt(x * x + y * y + z * z);  // facetArea (core:333-341)
    }
    index.products.push_back("photovoltaic_summary.csv");
    index.products.push_back(".facet_steps/components.bin");
    if (set.stepSync) {  // shared geometry before the first node (core:3442-3462)
        std::ofstream g(result, std::ios::binary | std::ios::trunc);
        writeHeader(g, "HIP GPU Facet RT-EB coupled", scene);
        g << "  \"componentFile\": \".facet_steps/components.bin\",\n  \"facetCount\": " << scene.facetCount()
          << ",\n  \"leafFacetCount\": " << scene.leafFacetCount << ",\n  \"vertexPositions\": [";
        for (size_t i = 0; i < scene.vertices.size(); ++i) {
            if (i) g << ',';
            const auto& v = scene.vertices[i];
            g << v.position[0] << ',' << v.position[1] << ',' << v.position[2];
        }
        g << "]\n}\n";
        if (!g) throw std::runtime_error("FacetEB cannot write shared geometry");
    }

    const size_t S = scene.facetCount() * size_t(2);
    std::vector<float> T(S), Blw(S), sunlit(S), Rn(S), H(S), LE(S), G(S), sw(S * nw), Bsw(S), Btot(S);
    std::vector<float> absorbedSW(S), pvInc(S), pvPow(S, 0.0f);
    bool allConverged = true;
    fh_eb_node_stats last{};
    std::string lastToken;
    for (int node = set.startNode; node < set.endNode; ++node) {
        const MeteoStep& m = meteo[size_t(node)];
        const SunAngles sun = nodeSun(m, set);
        const std::string token = timeToken(m.julianTime);
        const facethip::Direction sd = sunDirection(float(sun.zenith), float(sun.azimuth));
        const float cz = std::cos(float(sun.zenith) * kPi / 180.0f);
        const float total = std::max(0.0f, m.shortwave);
        const float scale = (cz > 1.0e-4f && directIntegral > 1.0e-8f)
                                ? std::min(1400.0f, total * directIntegral / cz) / directIntegral
                                : 0.0f;
        std::vector<float> dni(nw, 0.0f), dif(nw, 0.0f);
        for (uint32_t w = 0; w < nw && total > 0.0f; ++w) {
            dni[w] = scale * windows[w].directFraction;
            dif[w] = total * windows[w].diffuseFraction;
        }
        fh_eb_node_params p{};
        p.sun = {sd.x, sd.y, sd.z};
        p.airTemperatureK = std::clamp(m.airTemperature + 273.15f, 220.0f, 340.0f);
        p.vaporPressure = m.vaporPressure;
        p.pressure = m.pressure;
        p.windSpeed = m.windSpeed;
        p.skyLongwave = m.longwave > 0.0f ? m.longwave
                                           : float(5.670374419e-8 * std::pow(double(set.skyTemperature), 4.0));
        p.dTime = set.dTime;
        p.temperatureTolerance = set.tolerance;
        p.temperatureRelaxation = set.relaxation;
        p.maxCoupling = set.couplingIterations;
        p.radiosityTolerance = 1.0e-5f;
        p.maxSweeps = 2000;
        p.precision = set.bf16 ? FH_PRECISION_BF16_DC : FH_PRECISION_FP32;
        p.adaptiveRelaxation = set.adaptiveRelaxation ? 1u : 0u;
        p.vegetationMethod = uint32_t(std::clamp(set.vegetationMethod, 0, 1));
        p.warmStart = FH_WARM_PREVIOUS;
        p.firstNode = node == set.startNode ? 1u : 0u;
        fh_eb_node_stats st{};
        facethip::check(fh_eb_run_node(ctx, &p, dni.data(), dif.data(), &st), ctx);
        facethip::check(fh_eb_get_state(ctx, T.data(), absorbedSW.data(), Blw.data(), sunlit.data(), Rn.data(),
                                        H.data(), LE.data(), G.data(), sw.data()),
                        ctx);
        for (size_t s = 0; s < S; ++s) {
            float b = 0.0f;
            for (uint32_t w = 0; w < nw; ++w) b += sw[s * nw + w];
            Bsw[s] = b;
            Btot[s] = b + Blw[s];
        }
        const bool converged = st.converged != 0;
        allConverged = allConverged && converged;
        // sensor products (core:3862-3908, makeStateSpectralOptics
        // core:1135-1199, built on the device): per band the optics at
        // the node's temperatures, direct total * Esun(band) * 0.001 and
        // diffuse total * Esky(band) * 0.001 of the spectrum, sky radiance
        // diffuse / 2 (thermal: Planck at the sky temperature), 64 sweeps;
        // optical bands as apparent reflectance against direct cos(zenith) +
        // diffuse / 2, thermal bands as radiance
        std::vector<float> bandOut(S * nbs, 0.0f);
        if (nbs) {
            std::vector<float> bsky(nbs), dS(nbs, 0.0f), qS(nbs, 0.0f), bB(S * nbs);
            std::vector<uint8_t> opt(nbs);
            for (uint32_t b = 0; b < nbs; ++b) {  // isOpticalWavelength (core:1127-1133)
                opt[b] = (sensorBands[b] <= 2.5f ? sensorBands[b] * 1000.0f : sensorBands[b]) <= 2500.0f;
                dS[b] = opt[b] ? total * spectrum.directAt(sensorBands[b]) * 0.001f : 0.0f;
                qS[b] = opt[b] ? total * spectrum.diffuseAt(sensorBands[b]) * 0.001f : 0.0f;
                bsky[b] = opt[b] ? qS[b] * 0.5f : planckRadiance(sensorBands[b], set.skyTemperature);
            }
            // makeStateSpectralOptics on the device: per material and band the
            // band's r, t (ref_TIR, tau_TIR on thermal bands), the node's side
            // temperatures and sunlit fractions, the texture factors
            const size_t nsm = sensorMats.size();
            std::vector<float> smR(nsm * nbs), smT(nsm * nbs);
            for (size_t q = 0; q < nsm; ++q)
                for (uint32_t b = 0; b < nbs; ++b) {
                    const SensorMaterial& sm = sensorMats[q];
                    smR[q * nbs + b] = opt[b] ? sm.r[std::min<size_t>(b, sm.r.size() - 1)] : sm.tirR;
                    smT[q * nbs + b] = opt[b] ? sm.t[std::min<size_t>(b, sm.t.size() - 1)] : sm.tirT;
                }
            std::vector<uint32_t> smIndex(scene.facetCount());
            for (uint32_t f = 0; f < scene.facetCount(); ++f)
                smIndex[f] = uint32_t(std::min<size_t>(scene.material[f], nsm - 1));
            fh_spectral_optics so{};
            so.mode = FH_OPTICS_STATE;
            so.bands = nbs;
            so.materialCount = uint32_t(nsm);
            so.leafFacetCount = scene.leafFacetCount;
            so.facetMaterial = smIndex.data();
            so.reflectance = smR.data();
            so.transmittance = smT.data();
            so.wavelength = sensorBands.data();
            so.reflectanceFactor = scene.reflectanceFactor.empty() ? nullptr : scene.reflectanceFactor.data();
            so.backgroundAngularStrength = scene.backgroundAngularStrength;
            so.sun = {sd.x, sd.y, sd.z};
            so.sunlit = sunlit.data();
            so.directNormal = dS.data();
            so.temperature = T.data();
            fh_solve_stats bs{};
            facethip::check(fh_solve_spectral(ctx, &so, bsky.data(), nullptr, 64u, -1.0f, 1.0f, FH_PRECISION_FP32,
                                              bB.data(), &bs),
                            ctx);
            const float czs = std::max(0.0f, sd.y);
            for (uint32_t b = 0; b < nbs; ++b) {
                const bool optical = opt[b] != 0;
                const float refIrr = dS[b] * czs + qS[b] * 0.5f;
                for (size_t s = 0; s < S; ++s) {
                    const float v = bB[s * nbs + b];
                    if (!optical) bandOut[s * nbs + b] = v;
                    else if (refIrr > 1.0e-8f) bandOut[s * nbs + b] = std::clamp(v / refIrr, 0.0f, 1.0f);
                }
            }
        }
        // .facet_steps/energy_T=<token>.bin (core:3910-3935): per side
        // [T, B_total, B_SW, B_LW, Rn, H, LE, G, sunlit] + the sensor bands
        {
            const size_t rs = 9 + size_t(nbs);
            std::vector<float> rec(S * rs);
            for (size_t s = 0; s < S; ++s) {
                const float v[9] = {T[s], Btot[s], Bsw[s], Blw[s], Rn[s], H[s], LE[s], G[s], sunlit[s]};
                std::copy(v, v + 9, rec.begin() + std::ptrdiff_t(s * rs));
                std::copy(bandOut.begin() + std::ptrdiff_t(s * nbs), bandOut.begin() + std::ptrdiff_t((s + 1) * nbs),
                          rec.begin() + std::ptrdiff_t(s * rs + 9));
            }
            std::ofstream bin(stepDir / ("energy_T=" + token + ".bin"), std::ios::binary | std::ios::trunc);
            bin.write(reinterpret_cast<const char*>(rec.data()), std::streamsize(rec.size() * sizeof(float)));
            if (!bin) throw std::runtime_error("FacetEB cannot write time-step energy file");
        }
        if (set.saveRadiation) {
            const std::vector<float>* cols[5] = {&Bsw, &Blw, &Rn, nullptr, nullptr};
            writeProcess(outDir, "radiation", node, m.julianTime, token, sun, scene, st.couplingIterations,
                         st.maxTemperatureDelta, converged, st.maxEnergyResidual, st.maxResidualEquivalent, set,
                         cols, index);
        }
        if (set.saveEnergy) {
            const std::vector<float>* cols[5] = {&LE, &H, &G, &T, &Rn};
            writeProcess(outDir, "energy", node, m.julianTime, token, sun, scene, st.couplingIterations,
                         st.maxTemperatureDelta, converged, st.maxEnergyResidual, st.maxResidualEquivalent, set,
                         cols, index);
        }
        {  // photovoltaic facets of the node and their summary row (core:3848-3859)
            std::ofstream pv(outDir / ("photovoltaic_node_" + std::to_string(node) + ".csv"), std::ios::trunc);
            pv << "area_m2,temperature_C,incident_W_m2,absorbed_W_m2,electric_W_m2,residual_W_m2,latent_W_m2,"
                  "effective_W_m2,net_longwave_W_m2,sensible_W_m2,storage_W_m2,facet_index\n";
            double pa = 0, pt = 0, pg = 0, pabs = 0, pp = 0, plw = 0, ph = 0, ps = 0, pres = 0;
            if (anyPv) {
                facethip::check(fh_eb_get_photovoltaic(ctx, pvInc.data(), nullptr, pvPow.data()), ctx);
                for (uint32_t f = 0; f < scene.leafFacetCount; ++f) {
                    if (!pvFacet[f]) continue;
                    const size_t k = 2 * size_t(f);
                    const float a = pvArea[f], ab = absorbedSW[k] + absorbedSW[k + 1];
                    const float inc = pvInc[k] + pvInc[k + 1], power = pvPow[k];
                    const float lw = Rn[k] - ab, res = Rn[k] - H[k] - LE[k] - G[k] - power;
                    pv << a << ',' << T[k] - 273.15f << ',' << inc << ',' << ab << ',' << power << ',' << res << ','
                       << LE[k] << ',' << pvInc[k] + pvBif[f] * pvInc[k + 1] << ',' << lw << ',' << H[k] << ','
                       << G[k] << ',' << f << '\n';
                    pa += a;
                    pt += a * (T[k] - 273.15f);
                    pg += a * inc;
                    pabs += a * ab;
                    pp += a * power;
                    plw += a * lw;
                    ph += a * H[k];
                    ps += a * G[k];
                    pres = std::max(pres, double(std::abs(res)));
                }
            }
            if (pa > 0) {
                pvEnergyKwh += pp * set.dTime / 3600000.0;
                pvSummary << node << ',' << token << ',' << pa << ',' << pt / pa << ',' << pg << ',' << pabs << ','
                          << pp << ',' << plw << ',' << ph << ',' << ps << ',' << pres << ',' << pvEnergyKwh << '\n';
                pvSummary.flush();
                const std::vector<float>* cols[5] = {&pvPow, &T, &absorbedSW, nullptr, nullptr};
                writeProcess(outDir, "photovoltaic", node, m.julianTime, token, sun, scene, st.couplingIterations,
                             st.maxTemperatureDelta, converged, st.maxEnergyResidual, st.maxResidualEquivalent, set,
                             cols, index);
            }
            if (!pv || !pvSummary) throw std::runtime_error("FacetEB cannot finish time-step output");
        }
        index.products.push_back(".facet_steps/energy_T=" + token + ".bin");
        index.products.push_back("photovoltaic_node_" + std::to_string(node) + ".csv");
        {
            std::ostringstream c;
            c << std::setprecision(9) << "{\"node\": " << node << ", \"token\": " << jsonString(token)
              << ", \"julianTime\": " << m.julianTime << ", \"converged\": " << (converged ? "true" : "false")
              << ", \"iterations\": " << st.couplingIterations << ", \"iterationBudget\": " << set.couplingIterations
              << ", \"temperatureDelta\": " << st.maxTemperatureDelta
              << ", \"temperatureTolerance\": " << set.tolerance << ", \"energyResidual\": " << st.maxEnergyResidual
              << ", \"energyResidualUnits\": \"W m-2\", \"energyResidualTemperatureEquivalent\": "
              << st.maxResidualEquivalent << "}";
            index.completed.push_back(c.str());
        }
        index.write();
        const int progress = 10 + int(86.0 * (node - set.startNode + 1) / (set.endNode - set.startNode));
        std::cout << "PROGRESS\t" << progress << "\tFacet RT-EB node " << (node - set.startNode + 1) << "/"
                  << (set.endNode - set.startNode) << " " << token << ", iterations=" << st.couplingIterations
                  << ", max dT=" << st.maxTemperatureDelta << " K" << std::endl;
        std::cout << "FACET_CONVERGENCE\t" << node << '\t' << (converged ? 1 : 0) << '\t' << st.couplingIterations
                  << '\t' << st.maxTemperatureDelta << '\t' << st.maxEnergyResidual << '\t'
                  << st.maxResidualEquivalent << std::endl;
        if (set.stepSync) {  // back-pressure (core:4072-4082)
            std::cout << "FACET_STEP\t" << node << '\t' << token << '\t' << std::setprecision(9) << m.julianTime
                      << std::endl;
            std::string ack;
            if (!std::getline(std::cin, ack) || ack != "FACET_ACK\t" + std::to_string(node))
                throw std::runtime_error("FacetEB time-step output was interrupted or failed");
        }
        last = st;
        lastToken = token;
    }

    // light enhancement of the last node (core:4046-4047): its windows solved
    // again without the direct irradiance, max(0, B_SW - B_diffuse) per side
    std::vector<float> enhancement(S, 0.0f);
    {
        std::vector<float> Bdiffuse(S, 0.0f);
        fh_solve_stats ds{};
        facethip::check(fh_eb_diffuse_radiosity(ctx, Bdiffuse.data(), nullptr, &ds), ctx);
        for (size_t s = 0; s < S; ++s) enhancement[s] = std::max(0.0f, Bsw[s] - Bdiffuse[s]);
    }
    // iterations and maxDelta of the last node's radiative solution: the
    // larger of its shortwave solve and its final longwave solve
    // (combineRadiativeSolutions, core:1038-1055)
    const uint32_t radiativeIterations = std::max(last.swSweeps, last.lwFinalSweeps);
    const float radiativeMaxDelta = std::max(last.swMaxDelta, last.lwMaxDelta);

    std::ofstream out(result, std::ios::binary | std::ios::trunc);
    if (!out) throw std::runtime_error("FacetEB cannot write result: " + result);
    writeHeader(out, "HIP GPU Facet RT-EB coupled", scene);
    out << "  \"componentFile\": \".facet_steps/components.bin\",\n"
        << "  \"facetCount\": " << scene.facetCount() << ",\n  \"leafFacetCount\": " << scene.leafFacetCount
        << ",\n  \"time\": \"" << lastToken << "\",\n";
    writeVertices(out, scene);
    out << "  \"graph\": {\"directedEdges\":" << graph.directedEdgeCount << ",\"closureError\":"
        << graph.maxClosureError << ",\"fragmentOverflow\":0,\"hashOverflow\":0},\n"
        << "  \"iterations\": " << radiativeIterations << ",\n  \"maxDelta\": " << radiativeMaxDelta << ",\n"
        << "  \"couplingIterations\": " << last.couplingIterations << ",\n"
        << "  \"temperatureDelta\": " << last.maxTemperatureDelta << ",\n"
        << "  \"converged\": " << (allConverged ? "true" : "false") << ",\n"
        << "  \"energyResidual\": " << last.maxEnergyResidual << ",\n"
        << "  \"energyResidualUnits\": \"W m-2\",\n"
        << "  \"energyResidualTemperatureEquivalent\": " << last.maxResidualEquivalent << ",\n"
        << "  \"temperatureTolerance\": " << set.tolerance << ",\n"
        << "  \"iterationBudget\": " << set.couplingIterations << ",\n";
    writeArray(out, "sunlit", sunlit, true);
    writeArray(out, "radiosity", Btot, true);
    writeArray(out, "lightEnhancement", enhancement, true);
    writeArray(out, "temperature", T, true);
    writeArray(out, "netRadiation", Rn, true);
    writeArray(out, "sensibleHeat", H, true);
    writeArray(out, "latentHeat", LE, true);
    writeArray(out, "storageHeat", G, true);
    writeArray(out, "photovoltaicPower", pvPow, false);
    out << "}\n";
    out.close();
    if (!out) throw std::runtime_error("FacetEB cannot finish result output");
    index.products.push_back(std::filesystem::path(result).filename().generic_string());
    index.status = allConverged ? "completed" : "non_converged";
    index.finishedAt = utcTimestamp();
    index.write();
    std::cout << "PROGRESS\t100\t"
              << (allConverged ? "Facet radiation and energy balance converged"
                               : "Facet run finished with non-converged nodes; results kept")
              << std::endl;
    std::cout << "RESULT\t" << result << std::endl;
    return allConverged ? 0 : 2;
} catch (const std::exception& e) {
    std::cerr << "faceteb: " << e.what() << '\n';
    return 1;
}

// --solar-position Y M D h m s timezone longitude latitude pressure temperature
// facet_engine --scene-obj project.json scene.obj [factors.bin]: the
// assembled facet scene of a project (object facets, then the ground) as
// OBJ, and optionally its texture factors (float32 per facet)
int runSceneObj(int argc, char** argv) try {
    if (argc < 4) {
        std::cerr << "Usage: facet_engine --scene-obj project.json scene.obj [factors.bin]\n";
        return 2;
    }
    const facetcli::Project prj = facetcli::loadProject(argv[2]);
    facetcli::writeSceneObj(prj.scene, argv[3]);
    if (argc > 4) {
        std::vector<float> k(prj.scene.facetCount());
        for (uint32_t f = 0; f < prj.scene.facetCount(); ++f) k[f] = prj.scene.reflectanceFactorOf(f);
        std::ofstream out(argv[4], std::ios::binary | std::ios::trunc);
        out.write(reinterpret_cast<const char*>(k.data()), std::streamsize(k.size() * sizeof(float)));
        if (!out) throw std::runtime_error(std::string("cannot write ") + argv[4]);
    }
    std::cout << "facets " << prj.scene.facetCount() << " objects " << prj.scene.leafFacetCount << " materials "
              << prj.materials.size() << " raster " << prj.scene.rasterSize << " layers " << prj.scene.rasterLayers
              << std::endl;
    return 0;
} catch (const std::exception& e) {
    std::cerr << "facet_engine: " << e.what() << '\n';
    return 1;
}

// facet_engine --decode-image image out.raw: the image's samples as a
// texture read takes them (interleaved bytes), "width height channels" on
// stdout
int runDecodeImage(int argc, char** argv) try {
    if (argc != 4) {
        std::cerr << "Usage: facet_engine --decode-image image out.raw\n";
        return 2;
    }
    const facetimg::Image img = facetimg::readImage(argv[2]);
    std::ofstream out(argv[3], std::ios::binary | std::ios::trunc);
    out.write(reinterpret_cast<const char*>(img.data.data()), std::streamsize(img.data.size()));
    if (!out) throw std::runtime_error("cannot write " + std::string(argv[3]));
    std::cout << img.width << ' ' << img.height << ' ' << img.channels << std::endl;
    return 0;
} catch (const std::exception& e) {
    std::cerr << "facet_engine: " << e.what() << '\n';
    return 1;
}

// facet_engine --energy-windows width directRatio: the energy windows a
// project run solves with the spectra of the defined folder (flat without
// them): the folder (or "flat"), then per window "first last centre
// directFraction diffuseFraction", then the direct integral and the
// spectrum's values at 550 nm
int runEnergyWindows(int argc, char** argv) {
    if (argc != 4) {
        std::cerr << "Usage: facet_engine --energy-windows width directRatio\n";
        return 2;
    }
    const uint32_t width = uint32_t(std::clamp(std::atoi(argv[2]), 1, 1000));
    const float ratio = std::clamp(float(std::atof(argv[3])), 0.0f, 1.0f);
    const ShortwaveSpectrum sp = ShortwaveSpectrum::load(definedDirectory(), ratio);
    std::cout << std::setprecision(9) << (sp.source.empty() ? std::string("flat") : sp.source) << '\n';
    for (const EnergyWindow& w : energyWindows(sp, width))
        std::cout << w.first << ' ' << w.last << ' ' << w.centre() << ' ' << w.directFraction << ' '
                  << w.diffuseFraction << '\n';
    const size_t n = sp.count();
    std::cout << "direct " << (n >= 2 ? ShortwaveSpectrum::integrate(sp.direct, 0, n - 1) : 0.0f) << " at550 "
              << sp.directAt(550.0f) << ' ' << sp.diffuseAt(0.55f) << std::endl;
    return 0;
}

int runSolarPosition(int argc, char** argv) {
    if (argc != 13) {
        std::cerr << "Usage: facet_engine --solar-position Y M D h m s tz lon lat pressure temperature\n";
        return 2;
    }
    double v[11];
    for (int i = 0; i < 11; ++i) v[i] = std::atof(argv[2 + i]);
    const SunAngles a = solarPosition(int(v[0]), int(v[1]), int(v[2]), int(v[3]), int(v[4]), v[5], v[6], v[7], v[8],
                                      v[9], v[10]);
    std::cout << std::setprecision(10) << a.zenith << ' ' << a.azimuth << std::endl;
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "--solar-position") return runSolarPosition(argc, argv);
    if (mode == "--scene-obj") return runSceneObj(argc, argv);
    if (mode == "--energy-windows") return runEnergyWindows(argc, argv);
    if (mode == "--decode-image") return runDecodeImage(argc, argv);
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
    if (ext != ".obj" && ext != ".OBJ" && !isProject(input)) {
        std::cerr << (mode == "eFacetRT" ? "facetrt" : "faceteb") << ": the input must be an OBJ scene or project.json\n";
        return 1;
    }
    return mode == "eFacetRT" ? runFacetRT(input, output) : runFacetEB(input, output);
}
