/* tectonic.hpp -- header-only C++ facade with the public surface of the
 * reference class PlateTectonicSim (src/cuda_compute/plate_tectonic_sim.h:
 * 56-179), implemented over the C ABI of tectonic.h.  A host written against
 * the reference (main.cpp:135-154: construct, initialize, executeIteration in
 * a loop, resetSim from the GUI callback) compiles against this class with the
 * include changed; the OpenGL interop pointer must be null (headless).
 *
 * Differences from the reference, by design:
 *   - errors throw tt::Error (the reference exits the process from
 *     CUDA_ERROR_CHECK, plate_tectonic_sim.h:40-54);
 *   - settings are per world (kernelSettings is a process-wide __constant__ in
 *     the reference, kernel_settings.cuh:86-91), passed at construction or
 *     through setKernelSettings;
 *   - state access is explicit (download / upload of named fields). */
#ifndef TECTONIC_HPP
#define TECTONIC_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "tectonic.h"

namespace tt {

class Error : public std::runtime_error {
public:
    Error(const std::string &what, int status)
        : std::runtime_error(what + ": " + tt_status_string(status) + " " + tt_last_error()), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int status, const char *what)
{
    if (status != TT_OK) throw Error(what, status);
}

class PlateTectonicSim {
public:
    /* PlateTectonicSim(int width, int height, unsigned seed,
     * CudaGlInteropManager*), plate_tectonic_sim.h:111 */
    PlateTectonicSim(int width, int height, unsigned seed, void *interopManager = nullptr)
    {
        if (interopManager != nullptr) throw Error("the OpenGL interop path is not part of libtectonic", TT_ERR_UNSUPPORTED);
        check(tt_config_default(&cfg_), "tt_config_default");
        check(tt_settings_default(&ks_), "tt_settings_default");
        cfg_.width = width;
        cfg_.height = height;
        cfg_.seed = seed;
    }
    PlateTectonicSim(const tt_sim_config &cfg, const tt_kernel_settings &ks) : cfg_(cfg), ks_(ks) {}
    ~PlateTectonicSim() { release(); }
    PlateTectonicSim(const PlateTectonicSim &) = delete;
    PlateTectonicSim &operator=(const PlateTectonicSim &) = delete;

    /* initialize(int numStartingPlates, const std::vector<int>& numVoronoiSeeds),
     * plate_tectonic_sim.h:115; any number of Voronoi generations */
    void initialize(int numStartingPlates, const std::vector<int> &numVoronoiSeeds)
    {
        release();
        tt_sim_config c = cfg_;
        c.num_plates = numStartingPlates;
        c.num_voronoi_gens = 0;
        check(tt_world_create(&c, &ks_, &world_), "tt_world_create");
        reset(cfg_.seed, numStartingPlates, numVoronoiSeeds);
    }

    /* executeIteration(), plate_tectonic_sim.h:117 */
    void executeIteration() { executeIterations(1); }
    void executeIterations(int n)
    {
        requireWorld();
        check(tt_world_step(world_, n), "tt_world_step");
    }

    /* resetSim(unsigned seed, int numStartingPlates, const std::vector<int>&),
     * plate_tectonic_sim.h:119: the world keeps its device buffers */
    void resetSim(unsigned seed, int numStartingPlates, const std::vector<int> &numVoronoiSeeds)
    {
        cfg_.seed = seed;
        if (!world_) initialize(numStartingPlates, numVoronoiSeeds);
        else reset(seed, numStartingPlates, numVoronoiSeeds);
    }

    /* The static generators (plate_tectonic_sim.h:121-130).  The reference
     * passes a std::default_random_engine&; here the generator is a SplitMix64
     * state, advanced in place (a world starts from state = seed). */
    static std::vector<tt_voronoi_seed> generatePlateCenters(uint64_t &generator, int numPlates, int width, int height)
    {
        std::vector<tt_voronoi_seed> out((size_t)(numPlates > 0 ? numPlates : 0));
        check(tt_generate_plate_centers(&generator, numPlates, width, height, out.data()), "tt_generate_plate_centers");
        return out;
    }
    static std::vector<tt_voronoi_seed> generateVoronoiSeeds(uint64_t &generator,
                                                             const std::vector<tt_voronoi_seed> &centerSeeds,
                                                             int numSeeds, int width, int height)
    {
        std::vector<tt_voronoi_seed> out((size_t)(numSeeds > 0 ? numSeeds : 0));
        check(tt_generate_voronoi_seeds(&generator, centerSeeds.data(), (int32_t)centerSeeds.size(), numSeeds, width,
                                        height, out.data()), "tt_generate_voronoi_seeds");
        return out;
    }
    static std::vector<tt_plate_data> generatePlateData(uint64_t &generator,
                                                        const std::vector<tt_voronoi_seed> &plateCenters,
                                                        int numStartingPlates)
    {
        if ((size_t)numStartingPlates > plateCenters.size()) throw Error("fewer centres than plates", TT_ERR_INVALID);
        std::vector<tt_plate_data> out(TT_MAX_PLATES);
        check(tt_generate_plate_data(&generator, plateCenters.data(), numStartingPlates, out.data()),
              "tt_generate_plate_data");
        return out;
    }

    /* copyHostKernelSettingsToDevice, kernel_settings.cu:7-9 */
    void setKernelSettings(const tt_kernel_settings &ks)
    {
        ks_ = ks;
        if (world_) check(tt_world_set_settings(world_, &ks_), "tt_world_set_settings");
    }

    template <typename T>
    std::vector<T> download(int field) const
    {
        requireWorld();
        const size_t n = fieldBytes(field);
        std::vector<T> out((n + sizeof(T) - 1) / sizeof(T));
        check(tt_world_download(world_, field, out.data(), n), "tt_world_download");
        return out;
    }
    template <typename T>
    void upload(int field, const std::vector<T> &data)
    {
        requireWorld();
        check(tt_world_upload(world_, field, data.data(), data.size() * sizeof(T)), "tt_world_upload");
    }

    tt_step_stats stats() const
    {
        requireWorld();
        tt_step_stats s;
        check(tt_world_stats(world_, &s), "tt_world_stats");
        return s;
    }
    uint32_t iteration() const
    {
        requireWorld();
        uint32_t it = 0;
        check(tt_world_iteration(world_, &it), "tt_world_iteration");
        return it;
    }
    /* saveTextureCallback (plate_tectonic_sim.cu:800-803), headless */
    void exportPng(int what, const std::string &path) const
    {
        requireWorld();
        check(tt_world_export_png(world_, what, path.c_str()), "tt_world_export_png");
    }
    tt_world *handle() const { return world_; }
    const tt_sim_config &config() const { return cfg_; }
    const tt_kernel_settings &kernelSettings() const { return ks_; }

private:
    void reset(unsigned seed, int plates, const std::vector<int> &gens)
    {
        const std::vector<int32_t> g(gens.begin(), gens.end());
        check(tt_world_reset(world_, seed, plates, g.data(), (int32_t)g.size()), "tt_world_reset");
        cfg_.seed = seed;
        cfg_.num_plates = plates;
        cfg_.num_voronoi_gens = (int32_t)(g.size() < 4 ? g.size() : 4);
        for (int i = 0; i < cfg_.num_voronoi_gens; ++i) cfg_.voronoi_gens[i] = g[(size_t)i];
    }
    void requireWorld() const
    {
        if (!world_) throw Error("initialize() first", TT_ERR_STATE);
    }
    size_t fieldBytes(int field) const
    {
        const size_t n = (size_t)cfg_.width * (size_t)cfg_.height;
        switch (field) {
            case TT_FIELD_HEIGHT: case TT_FIELD_WATER: case TT_FIELD_SEDIMENT: case TT_FIELD_COLLISIONS:
            case TT_FIELD_LABELS: return n * 4;
            case TT_FIELD_PLATE_IDS: case TT_FIELD_ACCRETION: return n;
            case TT_FIELD_FLUX: return n * 16;
            case TT_FIELD_VELOCITY: case TT_FIELD_ASTHENOSPHERE: case TT_FIELD_UPLIFT: case TT_FIELD_SORTED_KEYS:
            case TT_FIELD_DISTANCE: return n * 8;
            case TT_FIELD_PROPAGATED: return n * 12;
            case TT_FIELD_PLATE_DATA: return sizeof(tt_plate_data) * 255;
            case TT_FIELD_DIVERGENCE_BITMAP: return 1013 * 4;
            case TT_FIELD_COLLISION_TYPES: return 32385;
            default: throw Error("unknown field", TT_ERR_INVALID);
        }
    }
    void release()
    {
        if (world_) {
            tt_world_destroy(world_);
            world_ = nullptr;
        }
    }

    tt_sim_config cfg_{};
    tt_kernel_settings ks_{};
    tt_world *world_ = nullptr;
};

}  // namespace tt

#endif /* TECTONIC_HPP */
