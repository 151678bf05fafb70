// particle.cpp -- Particle / Disk / RigidBumpy over the C ABI (see particle.hpp).
#include "particle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <numeric>
#include <stdexcept>

#include "routines.hpp"

namespace {
constexpr double PI = 3.141592653589793238462643383279502884197;
constexpr uint64_t MINSTD_A = 48271, MINSTD_M = 2147483647;

// a^z mod m (the engine's discard(z) for c == 0)
uint64_t pow_mod(uint64_t a, unsigned long long z) {
    uint64_t r = 1;
    a %= MINSTD_M;
    while (z) {
        if (z & 1) r = (r * a) % MINSTD_M;
        a = (a * a) % MINSTD_M;
        z >>= 1;
    }
    return r;
}
}  // namespace

double random_uniform(unsigned int seed, unsigned int n, double a, double b) {
    // default_random_engine rng(seed + n): state = s mod m, or 1 when that is 0
    unsigned int s = seed + n;
    uint64_t x = s % MINSTD_M;
    if (x == 0) x = 1;
    x = (pow_mod(MINSTD_A, n) * x) % MINSTD_M;  // rng.discard(n)
    x = (MINSTD_A * x) % MINSTD_M;              // rng()
    double r = static_cast<double>(x - 1);       // urng() - min, min = 1
    r /= (1.0 + static_cast<double>(MINSTD_M - 1 - 1));
    return (r * (b - a)) + a;
}

// ------------------------------------------------------------------ Particle
Particle::Particle(int kind, long n, long nv) : kind_(kind) { allocate(n, nv); }

void Particle::allocate(long n, long nv, bool fresh) {
    if (n <= 0) throw std::invalid_argument("Particle: n_particles must be positive");
    if (sys_ && n == n_particles && nv == n_vertices && !fresh) return;
    if (sys_) dpm_destroy(sys_);
    sys_ = dpm_create(kind_, n, nv);
    if (!sys_) throw std::runtime_error(std::string("Particle: ") + dpm_last_error());
    n_particles = n;
    n_vertices = nv;
    n_dof = 2 * n;
    box_ = {1.0, 1.0};
}

dpm_system* Particle::sys() const {
    if (!sys_)
        throw std::logic_error("Particle: no particles allocated (call initializeFromConfig first)");
    return sys_;
}

Particle::~Particle() {
    if (sys_) dpm_destroy(sys_);
}

void Particle::check(int status) const {
    if (status != 0) throw std::runtime_error(dpm_last_error());
}

std::vector<double> Particle::download(int field, int64_t count) {
    std::vector<double> v(count);
    check(dpm_download(sys(), field, v.data(), count));
    return v;
}

void Particle::upload(int field, const std::vector<double>& v) {
    check(dpm_upload(sys(), field, v.data(), (int64_t)v.size()));
}

double Particle::reduce(int what) {
    double out = 0.0;
    check(dpm_reduce(sys(), what, &out));
    return out;
}

std::vector<std::string> Particle::getFundamentalValues() {
    return {"positions", "velocities", "box_size", "radii", "masses"};
}

void Particle::setPositions(const std::vector<double>& x, const std::vector<double>& y) {
    upload(DPM_F_X, x);
    upload(DPM_F_Y, y);
}
std::pair<std::vector<double>, std::vector<double>> Particle::getPositions() {
    return {download(DPM_F_X, n_particles), download(DPM_F_Y, n_particles)};
}
void Particle::setVelocities(const std::vector<double>& vx, const std::vector<double>& vy) {
    upload(DPM_F_VX, vx);
    upload(DPM_F_VY, vy);
}
std::pair<std::vector<double>, std::vector<double>> Particle::getVelocities() {
    return {download(DPM_F_VX, n_particles), download(DPM_F_VY, n_particles)};
}
std::pair<std::vector<double>, std::vector<double>> Particle::getForces() {
    return {download(DPM_F_FX, n_particles), download(DPM_F_FY, n_particles)};
}
void Particle::setRadii(const std::vector<double>& r) { upload(DPM_F_RAD, r); }
std::vector<double> Particle::getRadii() { return download(DPM_F_RAD, n_particles); }
void Particle::setMasses(const std::vector<double>& m) { upload(DPM_F_MASS, m); }
std::vector<double> Particle::getMasses() { return download(DPM_F_MASS, n_particles); }
std::vector<double> Particle::getPotentialEnergy() { return download(DPM_F_PE, n_particles); }
std::vector<long> Particle::getStaticParticleIndex() {
    std::vector<int64_t> v(n_particles);
    check(dpm_download(sys(), DPM_F_STATIC_INDEX, v.data(), n_particles));
    return std::vector<long>(v.begin(), v.end());
}

// Particle::setBoxSize / getBoxSize (particle.cu:694-713)
void Particle::setBoxSize(const std::vector<double>& box) {
    if (box.size() != 2) throw std::invalid_argument("Particle::setBoxSize: box must have 2 sides");
    if (!(box[0] > 0) || !(box[1] > 0))
        throw std::invalid_argument("Particle::setBoxSize: box sides must be positive");
    box_ = box;
    check(dpm_set_box(sys(), box[0], box[1]));
    if (neighbor_list_update_method == "cell") cell_size = box[0] / n_cells_dim;
}
std::vector<double> Particle::getBoxSize() {
    double lx, ly;
    check(dpm_get_box(sys(), &lx, &ly));
    return {lx, ly};
}
double Particle::getBoxArea() {
    std::vector<double> b = getBoxSize();
    return b[0] * b[1];
}

// Particle::setEnergyScale / setExponent (particle.cu:318-367).  Only the contact terms
// (c) act on this path; the others are kept for getEnergyScale/getExponent.
void Particle::setEnergyScale(double e, std::string which) {
    if (which == "c") e_c = e;
    else if (which != "a" && which != "b" && which != "l")
        throw std::invalid_argument("Particle::setEnergyScale: which must be c, a, b or l");
    check(dpm_set_energy(sys(), e_c, n_c));
}
double Particle::getEnergyScale(std::string which) {
    if (which == "c") return e_c;
    throw std::invalid_argument("Particle::getEnergyScale: only the contact scale is used");
}
void Particle::setAllEnergyScales(double ec, double, double, double) { setEnergyScale(ec, "c"); }
void Particle::setExponent(double n, std::string which) {
    if (which == "c") n_c = n;
    else if (which != "a" && which != "b" && which != "l")
        throw std::invalid_argument("Particle::setExponent: which must be c, a, b or l");
    check(dpm_set_energy(sys(), e_c, n_c));
}
double Particle::getExponent(std::string which) {
    if (which == "c") return n_c;
    throw std::invalid_argument("Particle::getExponent: only the contact exponent is used");
}
void Particle::setAllExponents(double nc, double, double, double) { setExponent(nc, "c"); }

// Particle::getDiameter (particle.cu:914-924)
double Particle::getDiameter(std::string which) {
    std::vector<double> r = getRadii();
    if (which == "min") return 2.0 * *std::min_element(r.begin(), r.end());
    if (which == "max") return 2.0 * *std::max_element(r.begin(), r.end());
    if (which == "mean") return 2.0 * std::accumulate(r.begin(), r.end(), 0.0) / r.size();
    throw std::invalid_argument("Particle::getDiameter: which must be 'min', 'max', or 'mean', not " +
                                which);
}

// Particle::setBiDispersity (particle.cu:926-944): large particles first
void Particle::setBiDispersity(double size_ratio, double count_ratio) {
    if (size_ratio < 1.0)
        throw std::invalid_argument("Particle::setBiDispersity: size_ratio must be > 1.0");
    if (count_ratio < 0.0 || count_ratio > 1.0)
        throw std::invalid_argument("Particle::setBiDispersity: count_ratio must be < 1.0 and > 0.0");
    std::vector<double> r(n_particles);
    long n_large = static_cast<long>(n_particles * count_ratio);
    for (long i = 0; i < n_particles; ++i) r[i] = (i < n_large ? size_ratio : 1.0) / 2.0;
    setRadii(r);
}

// Disk::getParticleArea (disk.cu:205-207): pi * sum r^2 (summed in index order here)
double Particle::getParticleArea() {
    std::vector<double> r = getRadii();
    double s = 0.0;
    for (double v : r) s += v * v;
    return s * PI;
}

double Particle::getPackingFraction() { return reduce(DPM_RED_PACKING_FRACTION); }
double Particle::getNumberDensity() { return n_particles / getBoxArea(); }

// Particle::initializeBox (particle.cu:841-848)
void Particle::initializeBox(double packing_fraction) {
    setBoxSize({1.0, 1.0});
    scaleToPackingFraction(packing_fraction);
}

void Particle::setMass(double mass) { setMasses(std::vector<double>(n_particles, mass)); }

// Particle::getTimeUnit (particle.cu:1343-1348)
double Particle::getTimeUnit() {
    std::vector<double> m = getMasses();
    double average_mass = std::accumulate(m.begin(), m.end(), 0.0) / n_particles;
    return getDiameter("min") * std::sqrt(average_mass / getEnergyScale("c"));
}

// Particle::setRandomPositions (particle.cu:860-878): Data2D::fillRandomUniform with
// offset 1, x from value seed + 1 + i and y from seed + 1 + N + i
void Particle::setRandomPositions() { setRandomPositions(seed); }
void Particle::setRandomPositions(long s) {
    std::vector<double> b = getBoxSize(), x(n_particles), y(n_particles);
    unsigned int us = static_cast<unsigned int>(s);
    for (long i = 0; i < n_particles; ++i) {
        x[i] = random_uniform(us, static_cast<unsigned int>(s + 1 + i), 0.0, b[0]);
        y[i] = random_uniform(us, static_cast<unsigned int>(s + 1 + n_particles + i), 0.0, b[1]);
    }
    setPositions(x, y);
}

// Particle::setRandomVelocities (particle.cu:908-912).  The reference draws its normal
// numbers with thrust::normal_distribution, whose method depends on the Thrust backend;
// here they come from the same uniform stream through Box-Muller.  The mean is then
// removed and the velocities rescaled to exactly T, as in the reference.
void Particle::setRandomVelocities(double temperature) {
    std::vector<double> vx(n_particles), vy(n_particles);
    unsigned int us = static_cast<unsigned int>(seed);
    double sd = std::sqrt(temperature);
    for (long i = 0; i < n_particles; ++i) {
        unsigned int base = static_cast<unsigned int>(seed + 1 + 4 * i);
        double u1 = random_uniform(us, base, 0.0, 1.0), u2 = random_uniform(us, base + 1, 0.0, 1.0);
        double u3 = random_uniform(us, base + 2, 0.0, 1.0), u4 = random_uniform(us, base + 3, 0.0, 1.0);
        vx[i] = sd * std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * PI * u2);
        vy[i] = sd * std::sqrt(-2.0 * std::log(1.0 - u3)) * std::cos(2.0 * PI * u4);
    }
    setVelocities(vx, vy);
    removeMeanVelocities();
    scaleVelocitiesToTemperature(temperature);
}

void Particle::removeMeanVelocities() { check(dpm_remove_mean_velocity(sys())); }
void Particle::scaleVelocitiesToTemperature(double t) { check(dpm_scale_to_temperature(sys(), t)); }
double Particle::calculateTemperature() { return reduce(DPM_RED_TEMPERATURE); }

// Particle::setupNeighbors (particle.cu:71-100), with the same fall-backs
void Particle::setupNeighbors(ConfigDict& config) {
    const ConfigValue& nc = config.at("neighbor_list_config");
    setNeighborMethod(nc.at("neighbor_list_update_method").get<std::string>());
    double cm = nc.at("neighbor_cutoff_multiplier").get<double>();
    double dm = nc.at("neighbor_displacement_multiplier").get<double>();
    setNeighborSize(cm, dm);
    if (neighbor_list_update_method == "cell") {
        if (!setCellSize(nc.at("num_particles_per_cell").get<double>(),
                         nc.at("cell_displacement_multiplier").get<double>())) {
            std::cout << "WARNING: Particle::setupNeighbors: Could not set cell size.  Attempting "
                         "to use verlet list instead."
                      << std::endl;
            setNeighborMethod("verlet");
        }
        if (!setNeighborSize(cm, dm)) {
            std::cerr << "ERROR: Particle::setupNeighbors: Could not set neighbor size for cell "
                         "list - neighbor cutoff exceeds box size.  Attempting to use all-to-all "
                         "instead."
                      << std::endl;
            setNeighborMethod("all");
        }
    }
    if (neighbor_list_update_method == "verlet") {
        if (!setNeighborSize(cm, dm)) {
            std::cout << "WARNING: Particle::setupNeighbors: Could not set neighbor size.  "
                         "Attempting to use all-to-all instead."
                      << std::endl;
            setNeighborMethod("all");
        }
    }
    initNeighborList();
    zeroForceAndPotentialEnergy();
    calculateForces();  // forces are ready before integration starts
    double force_balance = getForceBalance();
    if (force_balance / n_particles / e_c > 1e-14)
        std::cout << "WARNING: Particle::setupNeighbors: Force balance is " << force_balance
                  << ", there will be an error!" << std::endl;
}

void Particle::setNeighborMethod(std::string method_name) {
    if (method_name != "cell" && method_name != "verlet" && method_name != "all")
        throw std::invalid_argument("Particle::setNeighborMethod: unknown method " + method_name);
    neighbor_list_update_method = method_name;
}

// Particle::setNeighborSize (particle.cu:1223-1234)
bool Particle::setNeighborSize(double cutoff_multiplier, double displacement_multiplier) {
    max_neighbors_allocated = 4;
    neighbor_cutoff = cutoff_multiplier * getDiameter("max");
    neighbor_displacement_threshold_sq = std::pow(displacement_multiplier * neighbor_cutoff, 2);
    std::vector<double> b = getBoxSize();
    double diagonal = std::sqrt(b[0] * b[0] + b[1] * b[1]);
    if (neighbor_cutoff >= diagonal) {
        std::cout << "Particle::setNeighborSize: Neighbor radius exceeds the box size" << std::endl;
        return false;
    }
    return true;
}

// Particle::setCellSize (particle.cu:1242-1272).  The cells tile the box
// (cell size = L / n_cells_dim, as Particle::updateCellSize makes it on every box change).
bool Particle::setCellSize(double num_particles_per_cell, double cell_displacement_multiplier) {
    const long min_num_cells_dim = 4;
    double trial = std::sqrt(num_particles_per_cell / getNumberDensity());
    double min_cell = 2.0 * getDiameter("max");
    std::vector<double> b = getBoxSize();
    n_cells_dim = static_cast<long>(std::floor(b[0] / trial));
    if (n_cells_dim < min_num_cells_dim) n_cells_dim = min_num_cells_dim;
    cell_size = b[0] / n_cells_dim;
    if (cell_size < min_cell) {
        n_cells_dim = static_cast<long>(std::floor(b[0] / min_cell));
        if (n_cells_dim < min_num_cells_dim) {
            std::cout << "Particle::setCellSize: Failed to make cell list - fewer than "
                      << min_num_cells_dim
                      << " cells in one dimension and cell size is less than twice the maximum "
                         "particle diameter"
                      << std::endl;
            return false;
        }
        cell_size = b[0] / n_cells_dim;
    }
    n_cells = n_cells_dim * n_cells_dim;
    cell_displacement_threshold_sq = std::pow(cell_displacement_multiplier * cell_size, 2);
    return true;
}

void Particle::updateCellSize() {
    if (neighbor_list_update_method == "cell") cell_size = getBoxSize()[0] / n_cells_dim;
}

void Particle::applyNeighborConfig() {
    int method = neighbor_list_update_method == "cell"     ? DPM_NBR_CELL
                 : neighbor_list_update_method == "verlet" ? DPM_NBR_VERLET
                                                           : DPM_NBR_ALL;
    check(dpm_neighbor_config(sys(), method, neighbor_cutoff, neighbor_displacement_threshold_sq,
                              method == DPM_NBR_CELL ? n_cells_dim : 0,
                              cell_displacement_threshold_sq));
}

void Particle::initNeighborList() {
    applyNeighborConfig();
    check(dpm_neighbor_init(sys()));
    int64_t ma = 0;
    check(dpm_neighbor_stats(sys(), &ma, nullptr, nullptr, nullptr));
    max_neighbors_allocated = ma;
}

void Particle::checkForNeighborUpdate() {
    int kind = 0;
    check(dpm_neighbor_check(sys(), &kind));
    if (kind) {
        int64_t ma = 0;
        check(dpm_neighbor_stats(sys(), &ma, nullptr, nullptr, nullptr));
        max_neighbors_allocated = ma;
    }
}

void Particle::updateNeighborList() { initNeighborList(); }

double Particle::getMaxSquaredNeighborDisplacement() { return reduce(DPM_RED_MAX_NEIGH_DISP_SQ); }
double Particle::getMaxSquaredCellDisplacement() { return reduce(DPM_RED_MAX_CELL_DISP_SQ); }

void Particle::updatePositions(double dt) { check(dpm_drift(sys(), dt)); }
void Particle::updateVelocities(double dt) { check(dpm_kick(sys(), dt)); }
void Particle::zeroForceAndPotentialEnergy() { check(dpm_zero_forces(sys())); }
void Particle::calculateForces() { check(dpm_compute_forces(sys())); }
double Particle::totalPotentialEnergy() { return reduce(DPM_RED_PE); }
double Particle::totalKineticEnergy() { return reduce(DPM_RED_KE); }
double Particle::getForceBalance() { return reduce(DPM_RED_FORCE_BALANCE); }
double Particle::getPowerFire() { return reduce(DPM_RED_POWER); }
void Particle::setVelocitiesToZero() { check(dpm_zero_velocities(sys())); }
void Particle::mixVelocitiesAndForces(double alpha) { check(dpm_mix(sys(), alpha)); }

void Particle::calculateStressTensor() { check(dpm_compute_stress(sys())); }
double Particle::getPressure() { return reduce(DPM_RED_PRESSURE); }
std::vector<double> Particle::getStressTensor() {
    return {reduce(DPM_RED_STRESS_XX), reduce(DPM_RED_STRESS_XY), reduce(DPM_RED_STRESS_YX),
            reduce(DPM_RED_STRESS_YY)};
}
void Particle::countContacts() { check(dpm_count_contacts(sys())); }
double Particle::getContactCount() { return reduce(DPM_RED_CONTACTS); }

// Particle::scaleToPackingFraction (particle.cu:1037-1045)
void Particle::scaleToPackingFraction(double packing_fraction) {
    double new_side = std::pow(getParticleArea() / packing_fraction, 1.0 / 2);
    double side = std::pow(getBoxArea(), 1.0 / 2);
    scalePositions(new_side / side);
    setBoxSize({new_side, new_side});
    updateCellSize();
}
void Particle::scaleToPackingFractionFull(double phi) {
    check(dpm_scale_to_packing_fraction(sys(), phi));
    box_ = getBoxSize();
    updateCellSize();
}
void Particle::scaleBox(double f) {
    check(dpm_scale_box(sys(), f));
    box_ = getBoxSize();
    updateCellSize();
}
void Particle::scalePositions(double f) { check(dpm_scale_positions(sys(), f)); }
void Particle::scalePositionsFull(double f) { check(dpm_scale_positions(sys(), f)); }
void Particle::setLastState() { check(dpm_snapshot_save(sys())); }
void Particle::revertToLastState() {
    check(dpm_snapshot_restore(sys()));
    box_ = getBoxSize();
    updateCellSize();
}

// ------------------------------------------------------------------ Disk
Disk::Disk(long n) : Particle(DPM_DISK, n, 0) {}

// Disk::initializeFromConfig (disk.cu:98-150)
void Disk::initializeFromConfig(ConfigDict& config, bool minimize) {
    long n = config.at("n_particles").get<long>();
    allocate(n, 0);
    setConfig(config);
    setSeed(config.at("seed").get<long>());
    setBiDispersity(config.at("size_ratio").get<double>(), config.at("count_ratio").get<double>());
    n_dof = 2 * n_particles;
    setMass(config.at("mass").get<double>());
    setEnergyScale(config.at("e_c").get<double>(), "c");
    setExponent(config.at("n_c").get<double>(), "c");
    double phi = config.at("packing_fraction").get<double>();
    initializeBox(phi);
    // positions.fillRandomUniform(0, Lx, 0, Ly, 0, seed): offset 0 here
    std::vector<double> b = getBoxSize(), x(n_particles), y(n_particles);
    unsigned int us = static_cast<unsigned int>(seed);
    for (long i = 0; i < n_particles; ++i) {
        x[i] = random_uniform(us, static_cast<unsigned int>(seed + i), 0.0, b[0]);
        y[i] = random_uniform(us, static_cast<unsigned int>(seed + n_particles + i), 0.0, b[1]);
    }
    setPositions(x, y);
    scaleToPackingFraction(phi);
    setupNeighbors(config);
    if (minimize) minimizeAdam(*this);
}

// ------------------------------------------------------------------ RigidBumpy
RigidBumpy::RigidBumpy(long n, long nv) : Particle(DPM_BUMPY, n, nv) { n_dof = 3 * n; }

double RigidBumpy::calculateVertexRadius(double d_min, long n_small, double segment) {
    double ang = 2.0 * PI / static_cast<double>(n_small);
    return d_min / (1.0 + segment / std::sin(ang / 2.0)) / 2.0;
}

// RigidBumpy::getFundamentalValues (rigid_bumpy.h:32)
std::vector<std::string> RigidBumpy::getFundamentalValues() {
    return {"radii",           "positions",          "angles",        "velocities",
            "angular_velocities", "num_vertices_in_particle", "vertex_positions", "masses",
            "moments_of_inertia", "box_size",         "vertex_masses"};
}

void RigidBumpy::setVertexCounts(const std::vector<long>& nvp) {
    if ((long)nvp.size() != n_particles)
        throw std::invalid_argument("RigidBumpy::setVertexCounts: one count per particle");
    nv_ = nvp;
}

void RigidBumpy::setAngles(const std::vector<double>& t) { upload(DPM_F_ANGLE, t); }
std::vector<double> RigidBumpy::getAngles() { return download(DPM_F_ANGLE, n_particles); }
std::vector<double> RigidBumpy::getTorques() { return download(DPM_F_TORQUE, n_particles); }
std::pair<std::vector<double>, std::vector<double>> RigidBumpy::getVertexPositions() {
    return {download(DPM_F_VERTEX_X, n_vertices), download(DPM_F_VERTEX_Y, n_vertices)};
}
double RigidBumpy::getVertexContactCount() { return reduce(DPM_RED_VERTEX_CONTACTS); }

void RigidBumpy::applyNeighborConfig() {
    Particle::applyNeighborConfig();
    check(dpm_bumpy_neighbor_config(sys(), vertex_neighbor_cutoff, vertex_particle_neighbor_cutoff));
}

bool RigidBumpy::setNeighborSize(double cutoff_multiplier, double displacement_multiplier) {
    max_neighbors_allocated = 4;
    neighbor_cutoff = cutoff_multiplier * getDiameter("max");
    vertex_neighbor_cutoff = cutoff_multiplier * 2.0 * vertex_radius;
    vertex_particle_neighbor_cutoff = getDiameter("max");
    neighbor_displacement_threshold_sq =
        std::pow(displacement_multiplier * segment_length_per_vertex_diameter * vertex_radius, 2);
    std::vector<double> b = getBoxSize();
    double diagonal = std::sqrt(b[0] * b[0] + b[1] * b[1]);
    if (neighbor_cutoff >= diagonal) {
        std::cout << "Particle::setNeighborSize: Neighbor radius exceeds the box size" << std::endl;
        return false;
    }
    return true;
}

// RigidBumpy::initializeFromConfig (rigid_bumpy.cu:138-303): bidisperse radii (large
// first), n_vertices recomputed from the small/large counts, box from the disk area,
// uniform positions and angles from the seed (both fills start at value seed + i, so
// angle i is 2*pi*x_i/L_x, as in the reference), vertex rings at radius r - r_v with the
// vertex radius read from the config, masses and moments of inertia, the contact scale
// multiplied by (2 r_v / d_min)^2 (also written back into the object's config), the box
// rescaled to the packing fraction, then the neighbour set-up.
void RigidBumpy::initializeFromConfig(ConfigDict& config, bool minimize) {
    long n = config.at("n_particles").get<long>();
    long nvs = config.at("n_vertices_per_small_particle").get<long>();
    long nvl = config.at("n_vertices_per_large_particle").get<long>();
    double size_ratio = config.at("size_ratio").get<double>();
    double count_ratio = config.at("count_ratio").get<double>();
    double rv = config.at("vertex_radius").get<double>();
    double segment = config.at("segment_length_per_vertex_diameter").get<double>();
    double ec = config.at("e_c").get<double>();
    double nc = config.at("n_c").get<double>();
    double mass = config.at("mass").get<double>();
    long sd = config.at("seed").get<long>();
    double phi = config.at("packing_fraction").get<double>();
    bool rot = config.at("rotation").get<bool>();
    if (size_ratio < 1.0)
        throw std::invalid_argument("Particle::setBiDispersity: size_ratio must be > 1.0");
    if (count_ratio < 0.0 || count_ratio > 1.0)
        throw std::invalid_argument("Particle::setBiDispersity: count_ratio must be < 1.0 and > 0.0");
    if (nvs <= 0 || nvl <= 0)
        throw std::invalid_argument("RigidBumpy::initializeFromConfig: vertex counts must be positive");
    // kernelGetNumVerticesInParticles (rigid_bumpy/kernels.cu:1115-1134): a particle whose
    // radius equals the minimum is small, so with all radii equal every particle is small
    long n_large = static_cast<long>(n * count_ratio);
    bool equal = size_ratio == 1.0 || n_large == n || n_large == 0;
    std::vector<long> nvp(n);
    long total = 0;
    for (long i = 0; i < n; ++i) total += (nvp[i] = (equal || i >= n_large) ? nvs : nvl);
    allocate(n, total);
    setConfig(config);
    setSeed(sd);
    rotation = rot;
    segment_length_per_vertex_diameter = segment;
    setBiDispersity(size_ratio, count_ratio);
    n_dof = rotation ? 3 * n_particles : 2 * n_particles;
    setVertexCounts(nvp);
    initializeBox(phi);
    std::vector<double> b = getBoxSize(), x(n), y(n), th(n);
    unsigned int us = static_cast<unsigned int>(seed);
    for (long i = 0; i < n; ++i) {
        x[i] = random_uniform(us, static_cast<unsigned int>(seed + i), 0.0, b[0]);
        y[i] = random_uniform(us, static_cast<unsigned int>(seed + n + i), 0.0, b[1]);
        th[i] = random_uniform(us, static_cast<unsigned int>(seed + i), 0.0, 2 * PI);
    }
    setPositions(x, y);
    setAngles(th);
    vertex_radius = rv;
    setMass(mass);
    std::vector<int64_t> counts(nvp.begin(), nvp.end());
    check(dpm_bumpy_setup(sys(), counts.data(), vertex_radius, rotation ? 1 : 0));
    double new_ec = ec * std::pow(getGeometryScale(), 2);
    setEnergyScale(new_ec, "c");
    setExponent(nc, "c");
    this->config["e_c"] = new_ec;
    scaleToPackingFractionFull(phi);
    setupNeighbors(config);
    if (minimize) minimizeFire(*this, 1e-16, 1e-16);
}
