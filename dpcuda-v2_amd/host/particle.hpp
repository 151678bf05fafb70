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
#include <string>
#include <utility>
#include <vector>

#include "../../include/dpm.h"
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

    void setConfig(ConfigDict& config) { config_ = config; }
    ConfigDict getConfig() { return config_; }
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
    virtual double getVertexContactCount() { return 0.0; }

    // ---- box scaling and saved state ----
    virtual void scaleToPackingFraction(double packing_fraction);
    virtual void scaleToPackingFractionFull(double packing_fraction);
    void scaleBox(double scale_factor);
    virtual void scalePositions(double scale_factor);
    virtual void scalePositionsFull(double scale_factor);
    virtual void setLastState();
    virtual void revertToLastState();

   protected:
    explicit Particle(int kind) : kind_(kind) {}
    Particle(int kind, long n_particles, long n_vertices);
    // (re)creates the device system for n particles and nv vertices
    void allocate(long n_particles, long n_vertices);
    dpm_system* sys() const;
    void check(int status) const;
    std::vector<double> download(int field, int64_t count);
    void upload(int field, const std::vector<double>& v);
    virtual void applyNeighborConfig();
    double reduce(int what);

    dpm_system* sys_ = nullptr;
    int kind_ = DPM_DISK;
    ConfigDict config_;
    std::vector<double> box_ = {1.0, 1.0};
};

class Disk : public Particle {
   public:
    Disk() : Particle(DPM_DISK) {}
    explicit Disk(long n_particles);
    void initializeFromConfig(ConfigDict& config, bool minimize = false) override;
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

   protected:
    void applyNeighborConfig() override;
    std::vector<long> nv_;
};

// Engine and uniform distribution of the reference's seeded fills
// (include/functors.h:17-29: thrust::default_random_engine seeded with seed + n, then
// discard(n); thrust::uniform_real_distribution maps x to (x - 1)/(m - 1)*(b - a) + a).
// Value number n of a fill started at `first` is random_uniform(seed, first + i, a, b).
double random_uniform(unsigned int seed, unsigned int n, double a, double b);
