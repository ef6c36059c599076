// host.cpp -- host-only parts of libtectonic: parameter defaults, the JSON
// parameter surface, the asthenosphere noise table and the headless PNG
// export.  Plain C++ (no HIP), compiled with the host compiler.
#include <dlfcn.h>
#include <zlib.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/tectonic.h"
#include "tt_math.h"

void tt_set_error(const std::string &msg);
int collision_classes(tt_world *w, size_t n, bool propagated, std::vector<uint8_t> &cls);

// KernelSettings defaults (src/cuda_compute/kernel_settings.cuh:6-71)
void tt_settings_fill_defaults(tt_kernel_settings *k)
{
    k->inelasticCollisionMultiplierContinental = 1.f;
    k->inelasticCollisionMultiplierSubduction = 0.25f;
    k->inelasticCollisionMultiplierAccretion = 0.5f;
    k->frictionCoefficientContinental = 1.f;
    k->frictionCoefficientSubduction = 0.2f;
    k->frictionCoefficientAccretion = 0.5f;
    k->environmentalDragCoefficient = 0.25f;
    k->velocitySmoothingFactor = 0.9f;
    k->directionSmoothingFactor = 0.9f;
    k->breakScoreSmoothingFactor = 0.9f;
    k->mergeDotDirectionThreshold = .998f;
    k->mergeVelocityDiffThreshold = 0.01f;
    k->mergeMinVelocity = 0.05f;
    k->mergeWaitAfterSplit = 10;
    k->minPlateSize = 10;
    k->continentalCrustThreshold = 10000.f;
    k->divergence_height_target_min = 1000.f;
    k->divergence_height_target_max = 2000.f;
    k->divergence_interpolation_factor = 0.025f;
    k->collisionTypeUpdateCooldown = 4;
    k->minCollisionSize = 50;
    k->subductionOverUpliftMultiplier = 0.005f;
    k->subductionUnderUpliftMultiplier = 0.01f;
    k->continentalUpliftMultiplier = 0.01f;
    k->minUpliftClamp = 0.f;
    k->maxUpliftClamp = 20000.f;
    k->upliftRange = 50;
    k->subductionOverRange = 20;
    k->continentalPropMaxIters = 50;
    k->continentalPropThreshold = 13000;
    k->hydrationPipeCrossSection = 1000;
    k->hydrationPipeLength = 1000;
    k->hydrationRainfall = 0.101f;
    k->hydrationEvaporation = 0.303f;
    k->gravity = 9.81f;
    k->sedimentCapacity = 0.9f;
    k->sedimentErosionRate = 0.5f;
    k->sedimentDepositionRate = 0.01f;
    k->pressureAccumulation = 0.02f;
    k->pressureDecayRate = 0.8f;
    k->pressureBlurRange = 25;
    k->pressureMultiplier = 1.0f;
    k->curlNoiseMultiplier = 30.f;
    k->stressSplitThreshold = 30.0f;
    k->targetMinimumPlateArea = 100;
    k->targetMaximumPlateArea = 100000;
    k->thermalErosionStrength = 1.0f;
    k->thermalErosionAmplitude = 0.5f;
    k->thermalCellSize = 1000.0f;
    k->thermalThresholdAngle = 0.5f;
}

// 256x256 R channel of assets/textures/noise.png, shipped as data/noise_r8.bin
// next to this library
const std::vector<uint8_t> &tt_noise_table()
{
    static std::vector<uint8_t> table;
    static bool loaded = false;
    if (loaded) return table;
    loaded = true;
    std::string dir = ".";
    Dl_info info;
    if (dladdr((void *)&tt_noise_table, &info) && info.dli_fname) {
        dir = info.dli_fname;
        const size_t p = dir.find_last_of('/');
        dir = p == std::string::npos ? "." : dir.substr(0, p);
    }
    const char *env = getenv("TECTONIC_NOISE_PATH");
    const std::string path = env ? std::string(env) : dir + "/data/noise_r8.bin";
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return table;
    table.resize(65536);
    if (fread(table.data(), 1, 65536, f) != 65536) table.clear();
    fclose(f);
    return table;
}

// ---------------------------------------------------------------------------
// JSON parameter surface (config.json, generation settings, KernelSettings)
// ---------------------------------------------------------------------------
namespace {

struct JsonReader {
    const char *p;
    bool ok = true;
    std::map<std::string, double> nums;
    std::map<std::string, std::vector<double>> arrays;

    void ws() { while (*p && isspace((unsigned char)*p)) ++p; }
    bool eat(char c)
    {
        ws();
        if (*p == c) { ++p; return true; }
        return false;
    }
    std::string str()
    {
        std::string s;
        if (!eat('"')) { ok = false; return s; }
        while (*p && *p != '"') {
            if (*p == '\\' && p[1]) ++p;
            s += *p++;
        }
        if (*p == '"') ++p; else ok = false;
        return s;
    }
    void value(const std::string &key)
    {
        ws();
        if (*p == '{') { object(key); return; }
        if (*p == '[') {
            ++p;
            std::vector<double> v;
            ws();
            if (*p == ']') { ++p; arrays[key] = v; return; }
            while (ok) {
                ws();
                char *end = nullptr;
                const double d = strtod(p, &end);
                if (end == p) { ok = false; return; }
                v.push_back(d);
                p = end;
                if (eat(',')) continue;
                if (eat(']')) break;
                ok = false;
            }
            arrays[key] = v;
            return;
        }
        if (*p == '"') { str(); return; }
        if (!strncmp(p, "true", 4)) { nums[key] = 1; p += 4; return; }
        if (!strncmp(p, "false", 5)) { nums[key] = 0; p += 5; return; }
        if (!strncmp(p, "null", 4)) { p += 4; return; }
        char *end = nullptr;
        const double d = strtod(p, &end);
        if (end == p) { ok = false; return; }
        nums[key] = d;
        p = end;
    }
    void object(const std::string &prefix)
    {
        if (!eat('{')) { ok = false; return; }
        if (eat('}')) return;
        while (ok) {
            const std::string k = str();
            if (!eat(':')) { ok = false; return; }
            value(prefix.empty() ? k : prefix + "." + k);
            if (eat(',')) continue;
            if (eat('}')) return;
            ok = false;
        }
    }
};

struct KsField { const char *name; size_t off; bool is_int; };
#define KSF(n, i) {#n, offsetof(tt_kernel_settings, n), i}
const KsField kKsFields[] = {
    KSF(inelasticCollisionMultiplierContinental, false), KSF(inelasticCollisionMultiplierSubduction, false),
    KSF(inelasticCollisionMultiplierAccretion, false), KSF(frictionCoefficientContinental, false),
    KSF(frictionCoefficientSubduction, false), KSF(frictionCoefficientAccretion, false),
    KSF(environmentalDragCoefficient, false), KSF(velocitySmoothingFactor, false),
    KSF(directionSmoothingFactor, false), KSF(breakScoreSmoothingFactor, false),
    KSF(mergeDotDirectionThreshold, false), KSF(mergeVelocityDiffThreshold, false), KSF(mergeMinVelocity, false),
    KSF(mergeWaitAfterSplit, true), KSF(minPlateSize, true), KSF(continentalCrustThreshold, false),
    KSF(divergence_height_target_min, false), KSF(divergence_height_target_max, false),
    KSF(divergence_interpolation_factor, false), KSF(collisionTypeUpdateCooldown, true),
    KSF(minCollisionSize, true), KSF(subductionOverUpliftMultiplier, false),
    KSF(subductionUnderUpliftMultiplier, false), KSF(continentalUpliftMultiplier, false),
    KSF(minUpliftClamp, false), KSF(maxUpliftClamp, false), KSF(upliftRange, true),
    KSF(subductionOverRange, true), KSF(continentalPropMaxIters, true), KSF(continentalPropThreshold, false),
    KSF(hydrationPipeCrossSection, false), KSF(hydrationPipeLength, false), KSF(hydrationRainfall, false),
    KSF(hydrationEvaporation, false), KSF(gravity, false), KSF(sedimentCapacity, false),
    KSF(sedimentErosionRate, false), KSF(sedimentDepositionRate, false), KSF(pressureAccumulation, false),
    KSF(pressureDecayRate, false), KSF(pressureBlurRange, true), KSF(pressureMultiplier, false),
    KSF(curlNoiseMultiplier, false), KSF(stressSplitThreshold, false), KSF(targetMinimumPlateArea, true),
    KSF(targetMaximumPlateArea, true), KSF(thermalErosionStrength, false), KSF(thermalErosionAmplitude, false),
    KSF(thermalCellSize, false), KSF(thermalThresholdAngle, false),
};
#undef KSF

std::string last_component(const std::string &k)
{
    const size_t p = k.find_last_of('.');
    return p == std::string::npos ? k : k.substr(p + 1);
}

}  // namespace

extern "C" int tt_config_parse_json(const char *text, tt_sim_config *cfg, tt_kernel_settings *ks)
{
    if (!text || !cfg || !ks) return TT_ERR_INVALID;
    JsonReader r{text};
    r.object("");
    if (!r.ok) return TT_ERR_INVALID;
    for (const auto &kv : r.nums) {
        const std::string &k = kv.first;
        const double v = kv.second;
        const std::string leaf = last_component(k);
        if (k == "heightmap_dimensions.x") cfg->width = (int32_t)v;
        else if (k == "heightmap_dimensions.y") cfg->height = (int32_t)v;
        else if (leaf == "seed") cfg->seed = (uint32_t)v;
        else if (leaf == "numStartingPlates") cfg->num_plates = (int32_t)v;
        else if (leaf == "iterationsForCCL") cfg->iterations_for_ccl = (int32_t)v;
        else if (leaf == "hydraulicIterations") cfg->hydraulic_iterations = (int32_t)v;
        else if (leaf == "thermalIterations") cfg->thermal_iterations = (int32_t)v;
        else if (leaf == "speedMin") cfg->speed_min = (float)v;
        else if (leaf == "speedMax") cfg->speed_max = (float)v;
        else {
            for (const KsField &f : kKsFields)
                if (leaf == f.name) {
                    char *base = (char *)ks + f.off;
                    if (f.is_int) { const int32_t iv = (int32_t)v; memcpy(base, &iv, 4); }
                    else { const float fv = (float)v; memcpy(base, &fv, 4); }
                }
        }
    }
    for (const auto &kv : r.arrays)
        if (last_component(kv.first) == "numVoronoiSeeds") {
            const size_t n = kv.second.size() < 4 ? kv.second.size() : 4;
            cfg->num_voronoi_gens = (int32_t)n;
            for (size_t i = 0; i < n; ++i) cfg->voronoi_gens[i] = (int32_t)kv.second[i];
        }
    return TT_OK;
}

// ---------------------------------------------------------------------------
// PNG export (saveFloatCudaTextureToDiskAsGray16 / saveGrayscale8BitCudaTexture
// ToDiskAsRgb, texture_save.cu:10-97,136-294): own chunk writer over zlib.
// ---------------------------------------------------------------------------
namespace {

void put32(std::vector<uint8_t> &v, uint32_t x)
{
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data)
{
    put32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    const uint32_t crc = (uint32_t)crc32(0L, out.data() + start, (uInt)(out.size() - start));
    put32(out, crc);
}

int write_png(const char *path, int w, int h, int bit_depth, int color_type, int bytes_per_px,
              const std::vector<uint8_t> &pixels, bool srgb)
{
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)w);
    put32(ihdr, (uint32_t)h);
    ihdr.push_back((uint8_t)bit_depth);
    ihdr.push_back((uint8_t)color_type);
    ihdr.push_back(0);
    ihdr.push_back(0);
    ihdr.push_back(0);
    chunk(out, "IHDR", ihdr);
    if (srgb) chunk(out, "sRGB", std::vector<uint8_t>{0});
    const size_t row = (size_t)w * bytes_per_px;
    std::vector<uint8_t> raw;
    raw.reserve((row + 1) * h);
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), pixels.begin() + (size_t)y * row, pixels.begin() + (size_t)(y + 1) * row);
    }
    uLongf clen = compressBound((uLong)raw.size());
    std::vector<uint8_t> comp(clen);
    if (compress2(comp.data(), &clen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return TT_ERR_INVALID;
    comp.resize(clen);
    chunk(out, "IDAT", comp);
    chunk(out, "IEND", std::vector<uint8_t>());
    FILE *f = fopen(path, "wb");
    if (!f) return TT_ERR_INVALID;
    const size_t wr = fwrite(out.data(), 1, out.size(), f);
    fclose(f);
    return wr == out.size() ? TT_OK : TT_ERR_INVALID;
}

// colours of the collision classes 0..5: none, divergence, subduction,
// continental with 2 / 3 / 4 plates
const uint8_t kClassColor[6][3] = {{20, 20, 24}, {70, 130, 220}, {220, 80, 60}, {240, 200, 80}, {250, 150, 40},
                                   {255, 255, 255}};

void plate_color(uint8_t val, uint8_t *rgb)
{
    if (val == 255) { rgb[0] = 230; rgb[1] = 235; rgb[2] = 240; return; }
    const float h = 0.125f * (float)(val % 8) + (float)val / 8 / 256.0f;
    const float s = 0.5f + 0.5f * tt_sinf((float)val * 0.1f);
    const float v = 0.5f + 0.5f * tt_cosf((float)val * 0.1f);
    const float c = v * s;
    const float x = c * (1.0f - tt_fabsf(tt_fmodf(h * 6.0f, 2.0f) - 1.0f));
    const float m = v - c;
    float r, g, b;
    if (h >= 0.0 && h < 1.0 / 6.0) { r = c; g = x; b = 0.0f; }
    else if (h >= 1.0 / 6.0 && h < 2.0 / 6.0) { r = x; g = c; b = 0.0f; }
    else if (h >= 2.0 / 6.0 && h < 3.0 / 6.0) { r = 0.0f; g = c; b = x; }
    else if (h >= 3.0 / 6.0 && h < 4.0 / 6.0) { r = 0.0f; g = x; b = c; }
    else if (h >= 4.0 / 6.0 && h < 5.0 / 6.0) { r = x; g = 0.0f; b = c; }
    else { r = c; g = 0.0f; b = x; }
    rgb[0] = (uint8_t)std::lround((r + m) * 255.0f);
    rgb[1] = (uint8_t)std::lround((g + m) * 255.0f);
    rgb[2] = (uint8_t)std::lround((b + m) * 255.0f);
}

}  // namespace

// world geometry and settings (defined next to the world object)
int tt_world_geometry(tt_world *w, int32_t *width, int32_t *height, tt_kernel_settings *ks);

// grey16 conversion of one height value (convertFloatToGray16Kernel,
// texture_save.cu:80-97)
static uint16_t tt_height_to_gray16(float v, float mn, float mx)
{
    float norm = (v - mn) / (mx - mn);
    norm = tt_fminf(tt_fmaxf(norm, 0.0f), 1.0f);
    return (uint16_t)(norm * 65535.0f);
}

extern "C" int tt_world_export_png(tt_world *w, int32_t what, const char *path)
{
    if (!w || !path) return TT_ERR_INVALID;
    int32_t width = 0, height = 0;
    {
        tt_kernel_settings ks;
        const int st = tt_world_geometry(w, &width, &height, &ks);
        if (st != TT_OK) return st;
        const size_t n = (size_t)width * (size_t)height;
        if (what == TT_EXPORT_HEIGHT16) {
            std::vector<float> hgt(n);
            int s2 = tt_world_download(w, TT_FIELD_HEIGHT, hgt.data(), n * 4);
            if (s2 != TT_OK) return s2;
            std::vector<uint8_t> px(n * 2);
            for (size_t i = 0; i < n; ++i) {
                const uint16_t g = tt_height_to_gray16(hgt[i], ks.minUpliftClamp, ks.maxUpliftClamp);
                px[2 * i] = (uint8_t)(g >> 8);
                px[2 * i + 1] = (uint8_t)g;
            }
            return write_png(path, width, height, 16, 0, 2, px, true);
        }
        if (what == TT_EXPORT_PLATES_RGB) {
            std::vector<uint8_t> ids(n);
            int s2 = tt_world_download(w, TT_FIELD_PLATE_IDS, ids.data(), n);
            if (s2 != TT_OK) return s2;
            std::vector<uint8_t> px(n * 3);
            for (size_t i = 0; i < n; ++i) plate_color(ids[i], &px[3 * i]);
            return write_png(path, width, height, 8, 2, 3, px, false);
        }
        if (what == TT_EXPORT_COLLISION_CLASSES || what == TT_EXPORT_PROPAGATED_CLASSES) {
            std::vector<uint8_t> cls;
            const int s2 = collision_classes(w, n, what == TT_EXPORT_PROPAGATED_CLASSES, cls);
            if (s2 != TT_OK) return s2;
            std::vector<uint8_t> px(n * 3);
            for (size_t i = 0; i < n; ++i) memcpy(&px[3 * i], kClassColor[cls[i] < 6 ? cls[i] : 0], 3);
            return write_png(path, width, height, 8, 2, 3, px, false);
        }
    }
    return TT_ERR_INVALID;
}

// ---------------------------------------------------------------------------
// diagnostics: per-plate records (copyPlateDataGuiKernel,
// plate_tectonics_kernel.cu:1220-1281), collision classes
// (createVisualizationCollisionTexture / ...PropagatedCollisionTexture,
// :3250-3313), step statistics and field summaries, as JSON
// ---------------------------------------------------------------------------
namespace {

struct Json {
    std::string s;
    bool first = true;
    void raw(const std::string &t) { s += t; }
    void key(const char *k)
    {
        if (!first) s += ", ";
        first = false;
        s += "\"";
        s += k;
        s += "\": ";
    }
    void num(const char *k, double v)
    {
        key(k);
        char b[64];
        if (std::isfinite(v)) snprintf(b, sizeof b, "%.9g", v);
        else snprintf(b, sizeof b, "null");
        s += b;
    }
    void integer(const char *k, long long v)
    {
        key(k);
        s += std::to_string(v);
    }
    void open(const char *k, char c)
    {
        key(k);
        s += c;
        first = true;
    }
    void close(char c)
    {
        s += c;
        first = false;
    }
};

unsigned tri_index(unsigned i, unsigned j) { return i * (2u * 255u - i - 1u) / 2u + (j - i - 1u); }

}  // namespace

int collision_classes(tt_world *w, size_t n, bool propagated, std::vector<uint8_t> &cls)
{
    cls.assign(n, 0);
    if (!propagated) {
        std::vector<uint32_t> word(n);
        std::vector<uint8_t> up(n * 8);
        int st = tt_world_download(w, TT_FIELD_COLLISIONS, word.data(), n * 4);
        if (st == TT_OK) st = tt_world_download(w, TT_FIELD_UPLIFT, up.data(), n * 8);
        if (st != TT_OK) return st;
        for (size_t i = 0; i < n; ++i) {
            const uint32_t p = word[i];
            const uint8_t types = up[8 * i + 4];
            uint8_t v = 0;
            if (p == 0) v = 1;
            else if (p >> 8 != 0) v = (types & 1) == 0 ? 2 : (p >> 16 == 0 ? 3 : (p >> 24 == 0 ? 4 : 5));
            cls[i] = v;
        }
    } else {
        std::vector<uint8_t> pr(n * 12);
        const int st = tt_world_download(w, TT_FIELD_PROPAGATED, pr.data(), n * 12);
        if (st != TT_OK) return st;
        for (size_t i = 0; i < n; ++i) {
            uint32_t p;
            memcpy(&p, &pr[12 * i + 4], 4);
            const uint8_t types = pr[12 * i + 8];
            uint8_t v = 0;
            if (p == 0) v = 1;
            else if (p >> 8 != 0) v = (types & 1) == 0 ? 2 : (p >> 16 == 0 ? 3 : (p >> 24 == 0 ? 4 : 5));
            cls[i] = v;
        }
    }
    return TT_OK;
}

extern "C" int tt_world_diagnostics_json(tt_world *w, char *buf, size_t cap, size_t *len)
{
    if (!w || !len) return TT_ERR_INVALID;
    int32_t width = 0, height = 0;
    tt_kernel_settings ks;
    int st = tt_world_geometry(w, &width, &height, &ks);
    if (st != TT_OK) return st;
    const size_t n = (size_t)width * (size_t)height;
    tt_step_stats ss;
    std::vector<tt_plate_data> pd(255);
    std::vector<uint8_t> types(32385);
    std::vector<float> hgt(n), water(n), sed(n);
    st = tt_world_stats(w, &ss);
    if (st == TT_OK) st = tt_world_download(w, TT_FIELD_PLATE_DATA, pd.data(), pd.size() * sizeof(tt_plate_data));
    if (st == TT_OK) st = tt_world_download(w, TT_FIELD_COLLISION_TYPES, types.data(), types.size());
    if (st == TT_OK) st = tt_world_download(w, TT_FIELD_HEIGHT, hgt.data(), n * 4);
    if (st == TT_OK) st = tt_world_download(w, TT_FIELD_WATER, water.data(), n * 4);
    if (st == TT_OK) st = tt_world_download(w, TT_FIELD_SEDIMENT, sed.data(), n * 4);
    if (st != TT_OK) return st;
    std::vector<uint8_t> cls, pcls;
    st = collision_classes(w, n, false, cls);
    if (st == TT_OK) st = collision_classes(w, n, true, pcls);
    if (st != TT_OK) return st;

    Json j;
    j.raw("{");
    j.integer("width", width);
    j.integer("height", height);
    j.integer("iteration", ss.iteration);
    j.open("stats", '{');
    j.integer("largest_value", ss.largest_value);
    j.integer("largest_plate_id", ss.largest_plate_id);
    j.num("heaviest_value", ss.heaviest_value);
    j.integer("heaviest_plate_id", ss.heaviest_plate_id);
    j.integer("number_of_plates", ss.number_of_plates);
    j.integer("overflow_cells", ss.overflow_cells);
    j.integer("num_labels", ss.num_labels);
    j.integer("unassigned_cells", ss.unassigned_cells);
    j.integer("merges", ss.merges);
    j.integer("split_plate", ss.split_plate);
    j.integer("split_new_plate", ss.split_new_plate);
    j.integer("complex_sources", ss.complex_sources);
    j.num("mass_ledger", ss.mass_ledger);
    j.close('}');
    // field summaries (fp64 sums on the host)
    double hmin = 0, hmax = 0, hsum = 0, wsum = 0, wmax = 0, ssum = 0;
    for (size_t i = 0; i < n; ++i) {
        const double h = hgt[i];
        if (i == 0 || h < hmin) hmin = h;
        if (i == 0 || h > hmax) hmax = h;
        hsum += h;
        wsum += water[i];
        wmax = std::max(wmax, (double)water[i]);
        ssum += sed[i];
    }
    j.open("fields", '{');
    j.num("height_min", hmin);
    j.num("height_max", hmax);
    j.num("height_mean", n ? hsum / (double)n : 0.0);
    j.num("water_total", wsum);
    j.num("water_max", wmax);
    j.num("sediment_total", ssum);
    j.close('}');
    const char *names[6] = {"none", "divergence", "subduction", "continental_2", "continental_3", "continental_4"};
    for (int pass = 0; pass < 2; ++pass) {
        long long cnt[6] = {0, 0, 0, 0, 0, 0};
        for (uint8_t v : pass ? pcls : cls) ++cnt[v < 6 ? v : 0];
        j.open(pass ? "propagated_classes" : "collision_classes", '{');
        for (int k = 0; k < 6; ++k) j.integer(names[k], cnt[k]);
        j.close('}');
    }
    j.open("plates", '[');
    for (int p = 0; p < 255; ++p) {
        const tt_plate_data &d = pd[(size_t)p];
        if (d.size <= 0) continue;
        if (!j.first) j.raw(", ");
        j.raw("{");
        j.first = true;
        j.integer("id", p);
        j.integer("size", d.size);
        j.num("mass", d.mass);
        j.integer("perimeter", d.perimeter);
        j.integer("has_moved", d.hasMoved);
        j.num("velocity", d.velocity);
        j.num("velocity_smoothed", d.velocitySmoothed);
        j.open("direction", '[');
        j.raw(std::to_string(d.direction[0]) + ", " + std::to_string(d.direction[1]));
        j.close(']');
        j.open("direction_smoothed", '[');
        j.raw(std::to_string(d.directionSmoothed[0]) + ", " + std::to_string(d.directionSmoothed[1]));
        j.close(']');
        j.open("center", '[');
        j.raw(std::to_string(d.geometricCenter[0]) + ", " + std::to_string(d.geometricCenter[1]));
        j.close(']');
        j.num("break_score", d.breakScore);
        j.num("circularity", d.circularity);
        j.integer("merge_wait_time", d.mergeWaitTime);
        // partners from the collision-type matrix: continental partners, and
        // subductions as +other (this plate over) or -other (this plate under)
        std::string cont, subd;
        for (int q = 0; q < 255; ++q) {
            if (q == p) continue;
            const unsigned lo = (unsigned)std::min(p, q), hi = (unsigned)std::max(p, q);
            const uint8_t t = types[tri_index(lo, hi)];
            if (!(t & 1)) continue;
            const bool continental = (t >> 1) & 1, polarity = (t >> 3) & 1;
            if (continental) {
                cont += (cont.empty() ? "" : ", ") + std::to_string(q);
            } else {
                const int v = polarity ? ((int)lo == p ? (int)hi : -(int)lo) : ((int)lo == p ? -(int)hi : (int)lo);
                subd += (subd.empty() ? "" : ", ") + std::to_string(v);
            }
        }
        j.open("continental", '[');
        j.raw(cont);
        j.close(']');
        j.open("subductions", '[');
        j.raw(subd);
        j.close(']');
        j.raw("}");
        j.first = false;
    }
    j.close(']');
    j.raw("}");
    *len = j.s.size();
    if (buf && cap > j.s.size()) {
        memcpy(buf, j.s.data(), j.s.size());
        buf[j.s.size()] = '\0';
        return TT_OK;
    }
    if (buf) {
        tt_set_error("diagnostics buffer too small: " + std::to_string(j.s.size() + 1) + " bytes needed");
        return TT_ERR_INVALID;
    }
    return TT_OK;
}
