// state_io.cpp -- the particle side of the on-disk formats: loading a particle from the
// reference's run directories (particle.cu:45-229, disk.cu:35-96, rigid_bumpy.cu:37-132,
// :310-351, :484-556) and the named arrays the I/O layer logs (particle.cu:485-692,
// rigid_bumpy.cu:800-930).
#include <iostream>
#include <stdexcept>

#include "io.hpp"
#include "particle.hpp"

namespace {
std::string stem_of(const std::filesystem::path& p) {
    std::string f = p.filename().string();
    return f.substr(0, f.find('.'));
}

ArrayData doubles_1d(const std::string& name, std::vector<double> v, const std::string& index) {
    ArrayData a;
    a.name = name;
    a.type = DataType::Double;
    a.size = {(long)v.size(), 1};
    a.index_array_name = index;
    a.data = std::move(v);
    return a;
}
ArrayData doubles_2d(const std::string& name, std::pair<std::vector<double>, std::vector<double>> v,
                     const std::string& index) {
    ArrayData a;
    a.name = name;
    a.type = DataType::Double;
    a.size = {(long)v.first.size(), 2};
    a.index_array_name = index;
    a.data = std::move(v);
    return a;
}
ArrayData longs_1d(const std::string& name, std::vector<long> v, const std::string& index) {
    ArrayData a;
    a.name = name;
    a.type = DataType::Long;
    a.size = {(long)v.size(), 1};
    a.index_array_name = index;
    a.data = std::move(v);
    return a;
}
std::vector<long> iota_long(long n) {
    std::vector<long> v((size_t)n);
    for (long i = 0; i < n; ++i) v[(size_t)i] = i;
    return v;
}
}  // namespace

// Particle::getPaths (particle.cu:45-69).  The reference's "frame" branch declares new
// frame_path / frame_number variables inside the branch, so the returned frame path is
// empty; here the frame directory found by get_trajectory_frame_path is returned.
std::tuple<std::filesystem::path, std::filesystem::path, std::filesystem::path, long>
Particle::getPaths(std::filesystem::path root_path, std::string source, long frame) {
    std::filesystem::path system_path = root_path / "system";
    std::filesystem::path trajectory_path = root_path / "trajectories";
    std::filesystem::path init_path = system_path / "init";
    std::filesystem::path restart_path = system_path / "restart";
    std::filesystem::path frame_path;
    long frame_number = 0;
    if (source == "restart") {
        std::ifstream step_file(restart_path / "step.dat");
        if (!(step_file >> frame_number))
            throw std::runtime_error("Particle::getPaths: could not read " +
                                     (restart_path / "step.dat").string());
    } else if (source == "frame") {
        if (frame == -2) throw std::runtime_error("Frame number is undefined");
        std::tie(frame_path, frame_number) = get_trajectory_frame_path(trajectory_path, "t", frame);
    }
    return std::make_tuple(frame_path, restart_path, init_path, frame_number);
}

// Particle::tryLoadArrayData (particle.cu:102-149): positions, velocities, forces (n x 2),
// radii, masses (n), box_size (2).  The values are uploaded at once.
bool Particle::tryLoadArrayData(std::filesystem::path path) {
    if (!std::filesystem::exists(path)) {
        std::cout << "Particle::tryLoadArrayData: File " << path << " does not exist" << std::endl;
        return false;
    }
    try {
        std::cout << "Particle::tryLoadArrayData: Loading data from " << path << std::endl;
        std::string name = stem_of(path);
        const long n = n_particles;
        if (name == "positions" || name == "velocities" || name == "forces") {
            std::vector<double> d = read_array_from_file<double>(path.string(), n, 2);
            std::vector<double> x((size_t)n), y((size_t)n);
            for (long i = 0; i < n; ++i) {
                x[(size_t)i] = d[(size_t)(2 * i)];
                y[(size_t)i] = d[(size_t)(2 * i + 1)];
            }
            if (name == "positions") setPositions(x, y);
            else if (name == "velocities") setVelocities(x, y);
            else {
                upload(DPM_F_FX, x);
                upload(DPM_F_FY, y);
            }
        } else if (name == "radii") {
            setRadii(read_array_from_file<double>(path.string(), n, 1));
        } else if (name == "masses") {
            setMasses(read_array_from_file<double>(path.string(), n, 1));
        } else if (name == "box_size") {
            setBoxSize(read_array_from_file<double>(path.string(), 2, 1));
        }
        return true;
    } catch (std::exception& e) {
        std::cout << "Particle::tryLoadArrayData: Error loading data from " << path << ": "
                  << e.what() << std::endl;
        return false;
    }
}

// Particle::tryLoadData (particle.cu:151-178): the array from the chosen source; a frame
// falls back to the restart directory
void Particle::tryLoadData(const std::filesystem::path& frame_path,
                           const std::filesystem::path& restart_path,
                           const std::filesystem::path& init_path, const std::string& source,
                           const std::string& name) {
    std::filesystem::path source_path;
    if (source == "frame") source_path = frame_path;
    else if (source == "restart") source_path = restart_path;
    else if (source == "init") source_path = init_path;
    else {
        std::cout << "Particle::tryLoadData: Unsupported source: " << source
                  << ".  Results not guranteed!" << std::endl;
        source_path = restart_path.parent_path() / source;
    }
    bool ok = tryLoadArrayData(source_path / (name + ".dat"));
    if (!ok && source == "frame") {
        ok = tryLoadArrayData(restart_path / (name + ".dat"));
        if (!ok)
            std::cerr << "Particle::tryLoadData: Error loading data from " << source_path << " or "
                      << restart_path << std::endl;
    } else if (!ok && (source == "init" || source == "restart")) {
        std::cerr << "Particle::tryLoadData: Error loading data from " << source_path << " or "
                  << init_path << std::endl;
    }
}

// Particle::loadDataFromPath (particle.cu:180-229)
void Particle::loadDataFromPath(std::filesystem::path root_path, std::string ext) {
    std::cout << "Particle::loadDataFromPath: Loading data from " << root_path << " with extension "
              << ext << std::endl;
    for (const auto& entry : std::filesystem::directory_iterator(root_path)) {
        if (std::filesystem::is_directory(entry.path())) continue;
        std::string f = entry.path().filename().string();
        f = f.substr(0, f.find(ext));
        if (f == "positions" || f == "velocities" || f == "forces" || f == "radii" ||
            f == "masses" || f == "box_size")
            tryLoadArrayData(entry.path());
    }
}

// the calculation a logged name needs first (particle.h:76-102)
const std::map<std::string, std::string>& Particle::calculationDependencies() const {
    static const std::map<std::string, std::string> deps = {
        {"TE", "calculate_kinetic_energy"},      {"T", "calculate_kinetic_energy"},
        {"KE", "calculate_kinetic_energy"},      {"Zp", "count_contacts"},
        {"P", "calculate_stress_tensor"},        {"stress_tensor_x", "calculate_stress_tensor"},
        {"stress_tensor_y", "calculate_stress_tensor"},
        {"stress_tensor", "calculate_stress_tensor"},
        {"kinetic_energy", "calculate_kinetic_energy"},
        {"Zv", "count_contacts"},
        {"potential_pairs", "calculate_force_distance_pairs"},
        {"force_pairs", "calculate_force_distance_pairs"},
        {"distance_pairs", "calculate_force_distance_pairs"},
        {"pair_ids", "calculate_force_distance_pairs"},
        {"overlap_pairs", "calculate_force_distance_pairs"},
        {"radsum_pairs", "calculate_force_distance_pairs"},
        {"pair_separation_angle", "calculate_force_distance_pairs"},
        {"hessian_pairs_xx", "calculate_force_distance_pairs"},
        {"hessian_pairs_xy", "calculate_force_distance_pairs"},
        {"hessian_pairs_yx", "calculate_force_distance_pairs"},
        {"hessian_pairs_yy", "calculate_force_distance_pairs"},
        {"hessian_ii_xx", "calculate_force_distance_pairs"},
        {"hessian_ii_xy", "calculate_force_distance_pairs"},
        {"hessian_ii_yx", "calculate_force_distance_pairs"},
        {"hessian_ii_yy", "calculate_force_distance_pairs"},
        {"angle_pairs_i", "calculate_force_distance_pairs"},
        {"angle_pairs_j", "calculate_force_distance_pairs"},
        {"this_vertex_contact_counts", "calculate_force_distance_pairs"},
        {"pair_friction_coefficient", "calculate_force_distance_pairs"},
        {"pair_vertex_overlaps", "calculate_force_distance_pairs"},
        {"hessian_pairs_xt", "calculate_force_distance_pairs"},
        {"hessian_pairs_yt", "calculate_force_distance_pairs"},
        {"hessian_pairs_tt", "calculate_force_distance_pairs"},
        {"hessian_pairs_tx", "calculate_force_distance_pairs"},
        {"hessian_pairs_ty", "calculate_force_distance_pairs"},
        {"hessian_ii_xt", "calculate_force_distance_pairs"},
        {"hessian_ii_yt", "calculate_force_distance_pairs"},
        {"hessian_ii_tt", "calculate_force_distance_pairs"},
        {"hessian_ii_tx", "calculate_force_distance_pairs"},
        {"hessian_ii_ty", "calculate_force_distance_pairs"},
    };
    return deps;
}

bool Particle::is_dependent(const std::string& log_name) const {
    return calculationDependencies().count(log_name) != 0;
}

void Particle::calculate_dependencies(const std::string& log_name) {
    auto it = calculationDependencies().find(log_name);
    if (it == calculationDependencies().end()) return;
    if (dependency_done_.count(it->second)) return;
    calculateDependency(it->second);
    dependency_done_.insert(it->second);
}

// Particle::handle_calculation_for_single_dependency (particle.cu:455-471)
void Particle::calculateDependency(const std::string& calculation) {
    if (calculation == "calculate_kinetic_energy") reduce(DPM_RED_KE);
    else if (calculation == "count_contacts") countContacts();
    else if (calculation == "calculate_stress_tensor") calculateStressTensor();
    else if (calculation == "calculate_force_distance_pairs") calculateForceDistancePairs();
}

void Particle::calculateForceDistancePairs() {
    int64_t row = 0;
    check(dpm_pair_analytics(sys(), &row));
    pair_row_ = row;
    max_neighbors_allocated = row;
}

// the per-slot pair arrays (particle.cu:573-692, rigid_bumpy.cu:800-930): rows in the
// library's internal particle order, max_neighbors_allocated slots each, not reordered
// (the reference writes them unpermuted too).  Computed at the first request after
// reset_dependency_status() unless already computed for this round.
ArrayData Particle::pairArray(const std::string& name) {
    static const std::map<std::string, std::pair<int, int>> two_d = {
        {"force_pairs", {DPM_P_FORCE_X, DPM_P_FORCE_Y}},
        {"distance_pairs", {DPM_P_DIST_X, DPM_P_DIST_Y}},
        {"pair_ids", {DPM_P_ID_I, DPM_P_ID_J}},
    };
    static const std::map<std::string, int> one_d = {
        {"potential_pairs", DPM_P_POTENTIAL},
        {"overlap_pairs", DPM_P_OVERLAP},
        {"radsum_pairs", DPM_P_RADSUM},
        {"pair_separation_angle", DPM_P_ANGLE},
        {"hessian_pairs_xx", DPM_P_H_XX}, {"hessian_pairs_xy", DPM_P_H_XY},
        {"hessian_pairs_yx", DPM_P_H_YX}, {"hessian_pairs_yy", DPM_P_H_YY},
        {"hessian_ii_xx", DPM_P_HII_XX}, {"hessian_ii_xy", DPM_P_HII_XY},
        {"hessian_ii_yx", DPM_P_HII_YX}, {"hessian_ii_yy", DPM_P_HII_YY},
        {"angle_pairs_i", DPM_P_ANGLE_I}, {"angle_pairs_j", DPM_P_ANGLE_J},
        {"this_vertex_contact_counts", DPM_P_VERTEX_CONTACTS},
        {"pair_friction_coefficient", DPM_P_FRICTION},
        {"pair_vertex_overlaps", DPM_P_VERTEX_OVERLAP},
        {"hessian_pairs_xt", DPM_P_H_XT}, {"hessian_pairs_yt", DPM_P_H_YT},
        {"hessian_pairs_tx", DPM_P_H_TX}, {"hessian_pairs_ty", DPM_P_H_TY},
        {"hessian_pairs_tt", DPM_P_H_TT}, {"hessian_ii_xt", DPM_P_HII_XT},
        {"hessian_ii_yt", DPM_P_HII_YT}, {"hessian_ii_tx", DPM_P_HII_TX},
        {"hessian_ii_ty", DPM_P_HII_TY}, {"hessian_ii_tt", DPM_P_HII_TT},
    };
    ArrayData a;
    auto it2 = two_d.find(name);
    auto it1 = one_d.find(name);
    if (it2 == two_d.end() && it1 == one_d.end()) return a;  // name "NULL": not a pair array
    if (!dependency_done_.count("calculate_force_distance_pairs") || pair_row_ == 0) {
        calculateForceDistancePairs();
        dependency_done_.insert("calculate_force_distance_pairs");
    }
    const int64_t S = n_particles * pair_row_;
    auto fetch_d = [&](int f) {
        std::vector<double> v((size_t)S);
        check(dpm_pair_download(sys(), f, v.data(), S));
        return v;
    };
    auto fetch_l = [&](int f) {
        std::vector<int64_t> v((size_t)S);
        check(dpm_pair_download(sys(), f, v.data(), S));
        return std::vector<long>(v.begin(), v.end());
    };
    a.name = name;
    a.index_array_name = "";
    if (it2 != two_d.end()) {
        a.size = {(long)S, 2};
        if (name == "pair_ids") {
            a.type = DataType::Long;
            a.data = std::make_pair(fetch_l(it2->second.first), fetch_l(it2->second.second));
        } else {
            a.type = DataType::Double;
            a.data = std::make_pair(fetch_d(it2->second.first), fetch_d(it2->second.second));
        }
    } else {
        a.size = {(long)S, 1};
        if (it1->second == DPM_P_VERTEX_CONTACTS) {
            a.type = DataType::Long;
            a.data = fetch_l(it1->second);
        } else {
            a.type = DataType::Double;
            a.data = fetch_d(it1->second);
        }
    }
    return a;
}

// Particle::getArrayData (particle.cu:485-692) for the arrays of this path.  Per-particle
// arrays come in the static order, so static_particle_index is 0..N-1 (what the reference's
// file holds after its scatter by that index).  The derived arrays (kinetic_energy,
// contact_counts, stress tensors) are computed from the current state when asked for.
ArrayData Particle::getArrayData(const std::string& array_name) {
    const std::string sp = "static_particle_index";
    if (array_name == "positions" || array_name == "particlePos")
        return doubles_2d(array_name, getPositions(), sp);
    if (array_name == "velocities") return doubles_2d(array_name, getVelocities(), sp);
    if (array_name == "forces") return doubles_2d(array_name, getForces(), sp);
    if (array_name == "box_size" || array_name == "boxSize")
        return doubles_1d(array_name, getBoxSize(), "");
    if (array_name == "radii") return doubles_1d(array_name, getRadii(), sp);
    if (array_name == "masses") return doubles_1d(array_name, getMasses(), sp);
    if (array_name == "kinetic_energy") {
        reduce(DPM_RED_KE);
        return doubles_1d(array_name, download(DPM_F_KE, n_particles), sp);
    }
    if (array_name == "potential_energy") return doubles_1d(array_name, getPotentialEnergy(), sp);
    if (array_name == "static_particle_index") return longs_1d(array_name, iota_long(n_particles), sp);
    if (array_name == "num_neighbors" || array_name == "contact_counts") {
        if (array_name == "contact_counts") countContacts();
        std::vector<int64_t> v((size_t)n_particles);
        check(dpm_download(sys(), array_name == "contact_counts" ? DPM_F_CONTACTS : DPM_F_NUM_NEIGHBORS,
                           v.data(), n_particles));
        return longs_1d(array_name, std::vector<long>(v.begin(), v.end()), sp);
    }
    if (array_name == "stress_tensor_x" || array_name == "stress_tensor_y") {
        calculateStressTensor();
        bool x = array_name == "stress_tensor_x";
        return doubles_2d(array_name,
                          {download(x ? DPM_F_STRESS_XX : DPM_F_STRESS_YX, n_particles),
                           download(x ? DPM_F_STRESS_XY : DPM_F_STRESS_YY, n_particles)},
                          sp);
    }
    if (array_name == "stress_tensor") {
        calculateStressTensor();
        ArrayData a = doubles_1d(array_name, getStressTensor(), "");
        a.size = {4, 1};
        return a;
    }
    ArrayData pa = pairArray(array_name);
    if (pa.name != "NULL") return pa;
    throw std::invalid_argument("Particle::getArrayData: array_name not found: " + array_name);
}

// ------------------------------------------------------------------ Disk
// Disk::load (disk.cu:35-96).  radii, masses, positions and box_size must load (the
// reference continues with zero-filled arrays; here that is an error); velocities are
// zero when absent.
long Disk::load(std::filesystem::path root_path, std::string source, long frame) {
    auto [frame_path, restart_path, init_path, frame_number] = getPaths(root_path, source, frame);
    ConfigDict config;
    config.load(root_path / "system" / "particle_config.json");
    long n = config.at("n_particles").get<long>();
    double e = config.at("e_c").get<double>();
    double nc = config.at("n_c").get<double>();
    long sd = config.at("seed").get<long>();
    allocate(n, 0, true);  // arrays land in the file order, as in the reference
    setConfig(config);
    setSeed(sd);
    n_dof = 2 * n_particles;
    auto need = [&](const char* name) {
        std::filesystem::path src = source == "frame"     ? frame_path
                                    : source == "init"    ? init_path
                                    : source == "restart" ? restart_path
                                                          : restart_path.parent_path() / source;
        bool have = std::filesystem::exists(src / (std::string(name) + ".dat")) ||
                    (source == "frame" && std::filesystem::exists(restart_path / (std::string(name) + ".dat")));
        tryLoadData(frame_path, restart_path, init_path, source, name);
        if (!have)
            throw std::runtime_error(std::string("Disk::load: ") + name + ".dat not found under " +
                                     root_path.string());
    };
    need("radii");
    need("masses");
    setEnergyScale(e, "c");
    setExponent(nc, "c");
    need("positions");
    tryLoadData(frame_path, restart_path, init_path, source, "velocities");
    need("box_size");
    this->config["packing_fraction"] = getPackingFraction();
    setupNeighbors(config);
    return frame_number;
}

// ------------------------------------------------------------------ RigidBumpy
// RigidBumpy::tryLoadArrayData (rigid_bumpy.cu:484-556): the disk arrays as above; the
// rotation and vertex arrays are kept until the vertex set-up has run (applyStaged).
bool RigidBumpy::tryLoadArrayData(std::filesystem::path path) {
    std::string name = stem_of(path);
    const bool per_vertex = name == "vertex_positions" || name == "vertex_masses";
    const bool staged = per_vertex || name == "angles" || name == "angular_velocities" ||
                        name == "moments_of_inertia" || name == "num_vertices_in_particle" ||
                        name == "velocities";
    if (!staged) return Particle::tryLoadArrayData(path);
    if (!std::filesystem::exists(path)) {
        std::cout << "RigidBumpy::tryLoadArrayData: File " << path << " does not exist" << std::endl;
        return false;
    }
    try {
        std::cout << "RigidBumpy::tryLoadArrayData: Loading data from " << path << std::endl;
        long rows = per_vertex ? n_vertices : n_particles;
        long cols = (name == "vertex_positions" || name == "velocities") ? 2 : 1;
        if (name == "num_vertices_in_particle") {
            std::vector<long> v = read_array_from_file<long>(path.string(), rows, 1);
            staged_[name] = std::vector<double>(v.begin(), v.end());
        } else {
            staged_[name] = read_array_from_file<double>(path.string(), rows, cols);
        }
        return true;
    } catch (std::exception& e) {
        std::cout << "Particle::tryLoadArrayData: Error loading data from " << path << ": "
                  << e.what() << std::endl;
        return false;
    }
}

void RigidBumpy::loadDataFromPath(std::filesystem::path root_path, std::string ext) {
    staged_.clear();
    Particle::loadDataFromPath(root_path, ext);
    for (const auto& entry : std::filesystem::directory_iterator(root_path)) {
        if (std::filesystem::is_directory(entry.path())) continue;
        std::string f = entry.path().filename().string();
        f = f.substr(0, f.find(ext));
        if (f == "angles" || f == "angular_velocities" || f == "vertex_positions" ||
            f == "vertex_masses" || f == "moments_of_inertia" || f == "velocities")
            tryLoadArrayData(entry.path());
    }
    applyStaged();
}

void RigidBumpy::applyStaged() {
    auto split = [](const std::vector<double>& d, long n) {
        std::vector<double> x((size_t)n), y((size_t)n);
        for (long i = 0; i < n; ++i) {
            x[(size_t)i] = d[(size_t)(2 * i)];
            y[(size_t)i] = d[(size_t)(2 * i + 1)];
        }
        return std::make_pair(x, y);
    };
    if (staged_.count("vertex_positions")) {
        auto [x, y] = split(staged_["vertex_positions"], n_vertices);
        upload(DPM_F_VERTEX_X, x);
        upload(DPM_F_VERTEX_Y, y);
    }
    if (staged_.count("vertex_masses")) upload(DPM_F_VERTEX_MASS, staged_["vertex_masses"]);
    if (staged_.count("moments_of_inertia")) upload(DPM_F_MOI, staged_["moments_of_inertia"]);
    if (staged_.count("angles")) upload(DPM_F_ANGLE, staged_["angles"]);
    if (staged_.count("angular_velocities")) upload(DPM_F_OMEGA, staged_["angular_velocities"]);
    if (staged_.count("velocities")) {
        auto [x, y] = split(staged_["velocities"], n_particles);
        setVelocities(x, y);
    }
    staged_.clear();
}

// RigidBumpy::load (rigid_bumpy.cu:37-132).  The vertex rings are set up from the loaded
// positions, angles and vertex counts, then the saved vertex positions, vertex masses and
// moments of inertia replace the computed ones.  e_c in the saved config already carries
// the (2 r_v / d_min)^2 factor and is used as is.
long RigidBumpy::load(std::filesystem::path root_path, std::string source, long frame) {
    auto [frame_path, restart_path, init_path, frame_number] = getPaths(root_path, source, frame);
    ConfigDict config;
    config.load(root_path / "system" / "particle_config.json");
    long n = config.at("n_particles").get<long>();
    long nv = config.at("n_vertices").get<long>();
    {  // the vertex count follows the saved per-particle counts when they disagree
        std::filesystem::path src = source == "frame"     ? frame_path
                                    : source == "init"    ? init_path
                                    : source == "restart" ? restart_path
                                                          : restart_path.parent_path() / source;
        std::filesystem::path f = src / "num_vertices_in_particle.dat";
        if (!std::filesystem::exists(f)) f = restart_path / "num_vertices_in_particle.dat";
        if (std::filesystem::exists(f)) {
            std::vector<long> c = read_array_from_file<long>(f.string(), n, 1);
            long total = 0;
            for (long v : c) total += v;
            if (total != nv) {
                std::cout << "WARNING: RigidBumpy::load: n_vertices " << nv
                          << " in particle_config.json, " << total << " in " << f << std::endl;
                nv = total;
            }
        }
    }
    double rv = config.at("vertex_radius").get<double>();
    double segment = config.at("segment_length_per_vertex_diameter").get<double>();
    double e = config.at("e_c").get<double>();
    double nc = config.at("n_c").get<double>();
    long sd = config.at("seed").get<long>();
    bool rot = config.at("rotation").get<bool>();
    allocate(n, nv, true);
    setConfig(config);
    setSeed(sd);
    rotation = rot;
    segment_length_per_vertex_diameter = segment;
    vertex_radius = rv;
    n_dof = rotation ? 3 * n_particles : 2 * n_particles;
    staged_.clear();
    for (const char* name : {"radii", "box_size", "positions", "angles", "velocities",
                             "angular_velocities", "vertex_positions", "num_vertices_in_particle",
                             "masses", "vertex_masses", "moments_of_inertia"})
        tryLoadData(frame_path, restart_path, init_path, source, name);
    if (!staged_.count("num_vertices_in_particle"))
        throw std::runtime_error("RigidBumpy::load: num_vertices_in_particle.dat not found under " +
                                 root_path.string());
    std::vector<int64_t> counts(staged_["num_vertices_in_particle"].begin(),
                                staged_["num_vertices_in_particle"].end());
    nv_.assign(counts.begin(), counts.end());
    staged_.erase("num_vertices_in_particle");
    if (staged_.count("angles")) {
        upload(DPM_F_ANGLE, staged_["angles"]);
    }
    check(dpm_bumpy_setup(sys(), counts.data(), vertex_radius, rotation ? 1 : 0));
    applyStaged();
    setEnergyScale(e, "c");
    setExponent(nc, "c");
    this->config["packing_fraction"] = getPackingFraction();
    setupNeighbors(config);
    return frame_number;
}

// RigidBumpy::getArrayData (rigid_bumpy.cu:800-930) for the arrays of this path
ArrayData RigidBumpy::getArrayData(const std::string& array_name) {
    const std::string sp = "static_particle_index", sv = "static_vertex_index";
    if (array_name == "angles") return doubles_1d(array_name, getAngles(), sp);
    if (array_name == "angular_velocities")
        return doubles_1d(array_name, download(DPM_F_OMEGA, n_particles), sp);
    if (array_name == "torques") return doubles_1d(array_name, getTorques(), sp);
    if (array_name == "moments_of_inertia")
        return doubles_1d(array_name, download(DPM_F_MOI, n_particles), sp);
    if (array_name == "vertex_positions") return doubles_2d(array_name, getVertexPositions(), sv);
    if (array_name == "vertex_masses")
        return doubles_1d(array_name, download(DPM_F_VERTEX_MASS, n_vertices), sv);
    if (array_name == "static_vertex_index") return longs_1d(array_name, iota_long(n_vertices), sv);
    if (array_name == "area") {
        std::vector<double> a((size_t)n_particles);
        check(dpm_bumpy_area(sys(), a.data(), n_particles));
        return doubles_1d(array_name, a, sp);
    }
    if (array_name == "num_vertices_in_particle" || array_name == "vertex_contact_counts") {
        bool counts = array_name == "num_vertices_in_particle";
        if (!counts) countContacts();
        std::vector<int64_t> v((size_t)n_particles);
        check(dpm_download(sys(), counts ? DPM_F_NUM_VERTICES : DPM_F_VERTEX_CONTACTS, v.data(),
                           n_particles));
        return longs_1d(array_name, std::vector<long>(v.begin(), v.end()), sp);
    }
    return Particle::getArrayData(array_name);
}
