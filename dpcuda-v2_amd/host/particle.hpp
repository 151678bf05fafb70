// particle.hpp -- the particle classes of the reference's host API
// (include/particles/base/particle.h:28-717, disk/disk.h:18-80,
// rigid_bumpy/rigid_bumpy.h:19-294) over the C ABI of libdpm.so.
//
// Each object owns one dpm_system (device state on the current HIP device), created by
// the sized constructor or, as in the reference, by initializeFromConfig.  Member
// names, argument meanings and the set-up logic (setupNeighbors, setNeighborSize,
// setCellSize, getTimeUnit, setBiDispersity, ...) follow the reference; arrays are
// exchanged as std::vector<double> in the reference's static particle order.
// Errors: invalid arguments throw std::invalid_argument (as the reference does);
// failures reported by the device library throw std::runtime_error with the message
// of dpm_last_error() (the reference prints and exits on CUDA errors).
#pragma once
#include <filesystem>
#include <map>
#include <set>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/dpm.h"
#include "array_data.hpp"
#include "config_dict.hpp"

class Particle {
   public:
    virtual ~Particle();
    Particle(const Particle&) = delete;
    Particle& operator=(const Particle&) = delete;

    dpm_system* handle() const { return sys(); }
    bool allocated() const { return sys_ != nullptr; }

    // ---- public state the reference's drivers read (particle.h:120-190) ----
    long n_particles = 0;
    long n_vertices = 0;
    long n_dof = 0;
    long seed = -1;
    double e_c = 1.0, n_c = 2.0;
    std::string neighbor_list_update_method = "none";
    double neighbor_cutoff = -1.0;
    double neighbor_displacement_threshold_sq = -1.0;
    double cell_size = -1.0;
    double cell_displacement_threshold_sq = -1.0;
    long n_cells_dim = -1, n_cells = -1;
    long max_neighbors_allocated = 4;

    // the particle's configuration (particle.h:35; drivers read it, e.g. config["particle_type"])
    ConfigDict config;
    void setConfig(ConfigDict& c) { config = c; }
    ConfigDict getConfig() { return config; }
    virtual void initializeFromConfig(ConfigDict& config, bool minimize = false) = 0;
    virtual std::vector<std::string> getFundamentalValues();

    // ---- arrays (static order) ----
    void setPositions(const std::vector<double>& x, const std::vector<double>& y);
    std::pair<std::vector<double>, std::vector<double>> getPositions();
    void setVelocities(const std::vector<double>& vx, const std::vector<double>& vy);
    std::pair<std::vector<double>, std::vector<double>> getVelocities();
    std::pair<std::vector<double>, std::vector<double>> getForces();
    void setRadii(const std::vector<double>& r);
    std::vector<double> getRadii();
    void setMasses(const std::vector<double>& m);
    std::vector<double> getMasses();
    std::vector<double> getPotentialEnergy();
    std::vector<long> getStaticParticleIndex();

    // ---- box, energy scales, geometry ----
    void setBoxSize(const std::vector<double>& box);
    std::vector<double> getBoxSize();
    double getBoxArea();
    void setEnergyScale(double e, std::string which);
    double getEnergyScale(std::string which);
    void setAllEnergyScales(double e_c, double e_a, double e_b, double e_l);
    void setExponent(double n, std::string which);
    double getExponent(std::string which);
    void setAllExponents(double n_c, double n_a, double n_b, double n_l);
    double getDiameter(std::string which = "min");
    void setBiDispersity(double size_ratio, double count_ratio);
    virtual double getParticleArea();
    double getPackingFraction();
    double getNumberDensity();
    double getDensity() { return getPackingFraction(); }
    void initializeBox(double packing_fraction);
    void setMass(double mass);
    double getTimeUnit();
    void setSeed(long s) { seed = s; }
    long getSeed() { return seed; }

    // ---- random initial state ----
    virtual void setRandomPositions();
    virtual void setRandomPositions(long seed);
    virtual void setRandomVelocities(double temperature);
    void removeMeanVelocities();
    void scaleVelocitiesToTemperature(double temperature);
    double calculateTemperature();

    // ---- neighbours ----
    void setupNeighbors(ConfigDict& config);
    void setNeighborMethod(std::string method_name);
    virtual bool setNeighborSize(double neighbor_cutoff_multiplier,
                                 double neighbor_displacement_multiplier);
    bool setCellSize(double num_particles_per_cell, double cell_displacement_multiplier);
    void updateCellSize();
    virtual void initNeighborList();
    virtual void checkForNeighborUpdate();
    void updateNeighborList();
    double getMaxSquaredNeighborDisplacement();
    double getMaxSquaredCellDisplacement();

    // ---- dynamics ----
    virtual void updatePositions(double dt);
    virtual void updateVelocities(double dt);
    virtual void zeroForceAndPotentialEnergy();
    virtual void calculateForces();
    virtual void calculateKineticEnergy() {}
    double totalPotentialEnergy();
    double totalKineticEnergy();
    double totalEnergy() { return totalKineticEnergy() + totalPotentialEnergy(); }
    double getForceBalance();
    virtual double getPowerFire();
    virtual void setVelocitiesToZero();
    virtual void mixVelocitiesAndForces(double alpha);

    // ---- analytics ----
    virtual void calculateStressTensor();
    double getPressure();
    std::vector<double> getStressTensor();
    virtual void countContacts();
    double getContactCount();
    // Disk/RigidBumpy::calculateForceDistancePairs (disk.cu:228-248, rigid_bumpy.cu:1383-1470):
    // per-slot pair arrays of the current state, read through getArrayData
    void calculateForceDistancePairs();
    virtual double getVertexContactCount() { return 0.0; }

    // ---- box scaling and saved state ----
    virtual void scaleToPackingFraction(double packing_fraction);
    virtual void scaleToPackingFractionFull(double packing_fraction);
    void scaleBox(double scale_factor);
    virtual void scalePositions(double scale_factor);
    virtual void scalePositionsFull(double scale_factor);
    virtual void setLastState();
    virtual void revertToLastState();

    // ---- state files (particle.cu:45-229) ----
    // Particle::load / Disk::load / RigidBumpy::load (disk.cu:35-96, rigid_bumpy.cu:37-132):
    // the particle described by <root>/system/particle_config.json with its arrays read
    // from <root>/system/init ("init"), <root>/system/restart ("restart", step from
    // step.dat) or <root>/trajectories/t<frame> ("frame", frame -1 = the latest; arrays
    // missing there come from restart).  Returns the step (0 for init, the frame for
    // frame).  The neighbour set-up then runs as in initializeFromConfig.
    virtual long load(std::filesystem::path root_path, std::string source, long frame = -2) = 0;
    std::tuple<std::filesystem::path, std::filesystem::path, std::filesystem::path, long>
    getPaths(std::filesystem::path root_path, std::string source, long frame);
    // reads one <name>.dat into the particle (false if the file is absent or unreadable)
    virtual bool tryLoadArrayData(std::filesystem::path path);
    void tryLoadData(const std::filesystem::path& frame_path,
                     const std::filesystem::path& restart_path,
                     const std::filesystem::path& init_path, const std::string& source,
                     const std::string& file_name_without_extension);
    // every known <name><extension> file of a directory
    virtual void loadDataFromPath(std::filesystem::path root_path, std::string data_file_extension);

    // ---- named arrays and values for the I/O layer (particle.cu:485-692, particle.h:61-104) ----
    virtual ArrayData getArrayData(const std::string& array_name);
    virtual std::vector<std::string> get_reorder_arrays() { return {"static_particle_index"}; }
    // log names whose value needs a calculation before it is read
    bool is_dependent(const std::string& log_name) const;
    void calculate_dependencies(const std::string& log_name);
    void reset_dependency_status() { dependency_done_.clear(); }

   protected:
    explicit Particle(int kind) : kind_(kind) {}
    Particle(int kind, long n_particles, long n_vertices);
    // (re)creates the device system for n particles and nv vertices; an existing system of
    // that size is kept unless `fresh` (a new system: static order = upload order again)
    void allocate(long n_particles, long n_vertices, bool fresh = false);
    dpm_system* sys() const;
    void check(int status) const;
    std::vector<double> download(int field, int64_t count);
    void upload(int field, const std::vector<double>& v);
    virtual void applyNeighborConfig();
    double reduce(int what);
    virtual void calculateDependency(const std::string& calculation);
    virtual const std::map<std::string, std::string>& calculationDependencies() const;
    // arrays read by tryLoadArrayData, uploaded once the particle is set up
    std::map<std::string, std::vector<double>> staged_;
    std::set<std::string> dependency_done_;
    long pair_row_ = 0;  // row length of the last calculateForceDistancePairs
    ArrayData pairArray(const std::string& array_name);

    dpm_system* sys_ = nullptr;
    int kind_ = DPM_DISK;
    std::vector<double> box_ = {1.0, 1.0};
};

class Disk : public Particle {
   public:
    Disk() : Particle(DPM_DISK) {}
    explicit Disk(long n_particles);
    void initializeFromConfig(ConfigDict& config, bool minimize = false) override;
    long load(std::filesystem::path root_path, std::string source, long frame = -2) override;
};

class RigidBumpy : public Particle {
   public:
    RigidBumpy() : Particle(DPM_BUMPY) {}
    RigidBumpy(long n_particles, long n_vertices);
    void initializeFromConfig(ConfigDict& config, bool minimize = false) override;
    // RigidBumpy::setNeighborSize (rigid_bumpy.cu:1278-1300): vertex cutoff from the same
    // multiplier times the vertex diameter, displacement threshold from the vertex radius
    bool setNeighborSize(double neighbor_cutoff_multiplier,
                         double neighbor_displacement_multiplier) override;
    double getVertexRadius() const { return vertex_radius; }
    // RigidBumpy::getGeometryScale (rigid_bumpy.cu:1265-1269)
    double getGeometryScale() { return 2.0 * vertex_radius / getDiameter("min"); }
    std::vector<std::string> getFundamentalValues() override;

    bool rotation = true;
    double vertex_radius = 0.0;
    double segment_length_per_vertex_diameter = 1.0;
    double vertex_neighbor_cutoff = -1.0;
    double vertex_particle_neighbor_cutoff = -1.0;

    // RigidBumpy::calculateVertexRadius (rigid_bumpy.cu:570-579)
    static double calculateVertexRadius(double d_min, long n_small, double segment);
    void setVertexCounts(const std::vector<long>& n_vertices_per_particle);
    void setAngles(const std::vector<double>& theta);
    std::vector<double> getAngles();
    std::pair<std::vector<double>, std::vector<double>> getVertexPositions();
    std::vector<double> getTorques();
    double getVertexContactCount() override;

    long load(std::filesystem::path root_path, std::string source, long frame = -2) override;
    bool tryLoadArrayData(std::filesystem::path path) override;
    void loadDataFromPath(std::filesystem::path root_path, std::string data_file_extension) override;
    ArrayData getArrayData(const std::string& array_name) override;
    std::vector<std::string> get_reorder_arrays() override {
        return {"static_particle_index", "static_vertex_index"};
    }

   protected:
    void applyNeighborConfig() override;
    // uploads the staged rotation and vertex arrays after the vertex set-up
    void applyStaged();
    std::vector<long> nv_;
};

// Engine and uniform distribution of the reference's seeded fills
// (include/functors.h:17-29: thrust::default_random_engine seeded with seed + n, then
// discard(n); thrust::uniform_real_distribution maps x to (x - 1)/(m - 1)*(b - a) + a).
// Value number n of a fill started at `first` is random_uniform(seed, first + i, a, b).
double random_uniform(unsigned int seed, unsigned int n, double a, double b);
