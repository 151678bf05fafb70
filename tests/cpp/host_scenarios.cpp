// host_scenarios.cpp -- runs the C++ host API (Particle / Disk / RigidBumpy, NVE / Fire /
// Adam, minimizeFire / minimizeAdam, the compression and jamming routines) through fixed
// scenarios and prints every result as "@ key value..." lines, doubles in hex (%a).
//
// Built twice: over libdpm.so (the HIP path) and over the CPU oracle (the same host
// sources with tests/cpp/dpm_on_oracle.h force-included).  The Python tests run both and
// compare the lines key by key.
//
//     host_scenarios <scenario>      scenarios: rng config errors disk_nve disk_fire
//                                    disk_adam disk_compress disk_jam_fire bumpy_init json
//     host_scenarios <scenario> <dir>   io_disk io_bumpy (files written under <dir>)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "factory.hpp"
#include "integrator.hpp"
#include "io.hpp"
#include "particle.hpp"
#include "routines.hpp"

namespace {

void put(const char* key, const std::vector<double>& v) {
    std::printf("@ %s", key);
    for (double x : v) std::printf(" %a", x);
    std::printf("\n");
}
void put(const char* key, double v) { put(key, std::vector<double>{v}); }
void put_int(const char* key, long v) { std::printf("@ %s %ld\n", key, v); }

void state(Particle& p, const std::string& tag) {
    auto pos = p.getPositions();
    auto vel = p.getVelocities();
    auto frc = p.getForces();
    put((tag + ".x").c_str(), pos.first);
    put((tag + ".y").c_str(), pos.second);
    put((tag + ".vx").c_str(), vel.first);
    put((tag + ".vy").c_str(), vel.second);
    put((tag + ".fx").c_str(), frc.first);
    put((tag + ".fy").c_str(), frc.second);
    put((tag + ".pe").c_str(), p.totalPotentialEnergy());
    put((tag + ".ke").c_str(), p.totalKineticEnergy());
    put((tag + ".phi").c_str(), p.getPackingFraction());
    put((tag + ".box").c_str(), p.getBoxSize());
}

template <class E, class F>
void expect_throw(const char* what, F f) {
    try {
        f();
    } catch (const E&) {
        std::printf("@ throws.%s 1\n", what);
        return;
    }
    std::printf("@ throws.%s 0\n", what);
}

std::unique_ptr<Particle> standard_disk(long n, double phi, long seed) {
    ConfigDict config = get_standard_disk_config(n, phi);
    config["seed"] = seed;
    return createParticle(config);
}

// thrust::default_random_engine is minstd_rand (a = 48271, m = 2^31 - 1); the standard
// library's std::minstd_rand is an independent implementation of the same engine (its
// discard steps one draw at a time, so n stays below 2^24 here).  Seeds that wrap or hit
// a multiple of m exercise the engine's "state 0 becomes 1" rule.
void scenario_rng() {
    long bad = 0;
    for (unsigned seed : {0u, 1u, 7u, 12345u, 2147483646u, 2147483647u, 4294967295u}) {
        for (unsigned n : {0u, 1u, 2u, 1000u, 65535u, 16777215u}) {
            std::minstd_rand e(seed + n);
            e.discard(n);
            double x = static_cast<double>(e() - std::minstd_rand::min()) /
                       (1.0 + static_cast<double>(std::minstd_rand::max() - std::minstd_rand::min()));
            double got = random_uniform(seed, n, 0.0, 1.0);
            if (got != x || !(got >= 0.0 && got < 1.0)) bad++;
            double scaled = random_uniform(seed, n, -2.0, 3.0);
            if (scaled != x * 5.0 + -2.0) bad++;
        }
    }
    put_int("rng.mismatches", bad);
}

void scenario_config() {
    ConfigDict a = get_standard_disk_config(1024, 0.8);
    ConfigDict b = a;  // deep copy
    b["neighbor_list_config"]["num_particles_per_cell"] = 4.0;
    put("config.copy_is_deep", a.at("neighbor_list_config").at("num_particles_per_cell").get<double>());
    double phi = a["packing_fraction"];
    long n = a["n_particles"];
    std::string kind = a["particle_type"];
    put("config.implicit_phi", phi);
    put_int("config.implicit_n", n);
    put_int("config.kind_is_disk", kind == "Disk");
    ConfigDict rb = get_standard_rigid_bumpy_config(10, 0.8);
    put_int("config.rb_n_vertices", rb.at("n_vertices").get<long>());
    expect_throw<std::out_of_range>("missing_key", [&] { a.at("nope"); });
    expect_throw<std::invalid_argument>("wrong_type", [&] { a.at("particle_type").get<double>(); });
}

void scenario_errors() {
    Disk d;
    expect_throw<std::logic_error>("unallocated", [&] { d.getPackingFraction(); });
    ConfigDict bad = get_standard_disk_config(16, 0.5);
    bad["particle_type"] = "Polygon";
    expect_throw<std::invalid_argument>("bad_type", [&] { createParticle(bad); });
    expect_throw<std::invalid_argument>("bad_count", [&] { Disk x(0); });
    auto p = standard_disk(64, 0.5, 3);
    expect_throw<std::invalid_argument>("size_ratio", [&] { p->setBiDispersity(0.5, 0.5); });
    expect_throw<std::invalid_argument>("count_ratio", [&] { p->setBiDispersity(1.4, 1.5); });
    expect_throw<std::invalid_argument>("diameter", [&] { p->getDiameter("median"); });
    expect_throw<std::invalid_argument>("box", [&] { p->setBoxSize({1.0}); });
    expect_throw<std::invalid_argument>("method", [&] { p->setNeighborMethod("octree"); });
    expect_throw<std::runtime_error>("size_mismatch", [&] { p->setRadii(std::vector<double>(3, 0.5)); });
}

// set-up + FIRE relaxation + Gaussian velocities + NVE through step() and run()
void scenario_disk_nve() {
    auto p = standard_disk(256, 0.85, 7);
    put_int("nve.n_cells_dim", p->n_cells_dim);
    put("nve.neighbor_cutoff", p->neighbor_cutoff);
    put("nve.cell_size", p->cell_size);
    put("nve.disp_sq", p->neighbor_displacement_threshold_sq);
    put("nve.cell_disp_sq", p->cell_displacement_threshold_sq);
    put("nve.time_unit", p->getTimeUnit());
    state(*p, "nve.init");
    minimizeFire(*p, 0.1, 1e-2, 200, 1000);
    state(*p, "nve.relaxed");
    p->setRandomVelocities(1e-4);
    put("nve.temperature", p->calculateTemperature());
    ConfigDict nc = get_nve_config_dict(0.01 / p->getTimeUnit());
    NVE nve(*p, nc);
    for (int k = 0; k < 60; ++k) nve.step();
    state(*p, "nve.step60");
    nve.run(140);
    state(*p, "nve.run200");
    p->calculateStressTensor();
    put("nve.stress", p->getStressTensor());
    put("nve.pressure", p->getPressure());
    p->countContacts();
    put("nve.contacts", p->getContactCount());
}

void scenario_disk_fire() {
    auto p = standard_disk(256, 0.9, 11);
    ConfigDict fc = get_fire_config_dict(0.1, 1e-2);
    Fire fire(*p, fc);
    for (int k = 0; k < 150; ++k) fire.step();
    state(*p, "fire.step150");
    put("fire.alpha", fire.alpha);
    put("fire.dt", fire.dt);
    put_int("fire.n_good", fire.N_good);
    put_int("fire.n_bad", fire.N_bad);
    minimizeFire(*p, 1e-16, 1e-16, 3000L);
    state(*p, "fire.min");
    long k = minimizeFireToMaxForce(*p, 1e-12, 5000);
    put_int("fire.fmax_steps", k);
    state(*p, "fire.fmax");
}

void scenario_disk_adam() {
    auto p = standard_disk(128, 0.8, 5);
    minimizeAdam(*p, 1e-3, 0.9, 0.999, 1e-8, 1e-16, 1e-16, 300, 1000);
    state(*p, "adam.min");
    ConfigDict ac = get_adam_config_dict(1e-3, 0.9, 0.999, 1e-8);
    Adam adam(*p, ac);
    jam_adam(*p, adam, 40, 200, 1e-10, 1e-14, 1e-3, 1e-5, 100.0);
    state(*p, "adam.jam");
}

void scenario_disk_compress() {
    auto p = standard_disk(128, 0.7, 9);
    ConfigDict ac = get_adam_config_dict(1e-3, 0.9, 0.999, 1e-8);
    Adam adam(*p, ac);
    compress_adam(*p, adam, 20, 150, 1e-12, 1e-14, 5e-3, 1e-6);
    state(*p, "compress.up");
    decompress_adam(*p, adam, 10, 150, 1e-12, 1e-14, 2e-3, 1e-9);
    state(*p, "compress.down");
    compress_to_phi_adam(*p, adam, 20, 150, 1e-12, 1e-14, 0.02);
    state(*p, "compress.phi");
}

void scenario_disk_jam_fire() {
    auto p = standard_disk(64, 0.82, 13);
    JamResult r = jamWithFire(*p, 1e-4, 1.5e-4, 0.995, 1e-9, 400);
    put_int("jam.status", r.status);
    put_int("jam.compression_steps", r.compression_steps);
    put_int("jam.fire_steps", r.fire_steps);
    put("jam.pressure", r.pressure);
    put("jam.phi", r.packing_fraction);
    put("jam.pe", r.potential_energy);
    put("jam.box", r.box_length);
    put_int("jam.contacts", r.contacts);
    state(*p, "jam.final");
}

void scenario_bumpy_init() {
    ConfigDict config = get_standard_rigid_bumpy_config(64, 0.7);
    config["seed"] = 21;
    config["n_vertices_per_small_particle"] = 20;
    config["n_vertices_per_large_particle"] = 28;
    config["vertex_radius"] = RigidBumpy::calculateVertexRadius(1.0, 20, 1.0);
    config["neighbor_list_config"]["neighbor_list_update_method"] = "verlet";
    auto p = createParticle(config);
    auto* rb = dynamic_cast<RigidBumpy*>(p.get());
    put_int("bumpy.n_vertices", rb->n_vertices);
    put_int("bumpy.n_dof", rb->n_dof);
    put("bumpy.vertex_radius", rb->getVertexRadius());
    put("bumpy.e_c", rb->e_c);
    put("bumpy.config_e_c", rb->getConfig().at("e_c").get<double>());
    put("bumpy.vertex_cutoff", rb->vertex_neighbor_cutoff);
    put("bumpy.disp_sq", rb->neighbor_displacement_threshold_sq);
    put("bumpy.angles", rb->getAngles());
    auto v = rb->getVertexPositions();
    put("bumpy.vx", v.first);
    put("bumpy.vy", v.second);
    state(*p, "bumpy.init");
    put("bumpy.torque", rb->getTorques());
    minimizeFire(*p, 0.1, 1e-2, 100, 1000);
    state(*p, "bumpy.fire100");
    put("bumpy.angles100", rb->getAngles());
    // rigid Adam (RigidBumpy::updatePositionsAdam): moments of force and torque
    p->scaleToPackingFractionFull(0.75);
    minimizeAdam(*p, 1e-3, 0.9, 0.999, 1e-8, 0.0, 0.0, 200, 1000);
    state(*p, "bumpy.adam200");
    put("bumpy.angles_adam", rb->getAngles());
    auto va = rb->getVertexPositions();
    put("bumpy.vx_adam", va.first);
    put("bumpy.vy_adam", va.second);
}

// JSON text of ConfigDict values: one "@ json.<k> <dump>" line per case (no spaces in the
// compact dumps), and the indented dump of a nested config written to stdout between
// markers for the test to parse
void scenario_json() {
    std::vector<std::pair<std::string, ConfigValue>> cases = {
        {"one", 1.0},          {"tenth", 0.1},        {"third", 1.0 / 3.0},
        {"tiny", 1e-7},        {"small", 1e-4},       {"smaller", 1e-5},
        {"e14", 1e14},         {"e15", 1e15},         {"e16", 1e16},
        {"big", 1.5e300},      {"neg", -2.5e-8},      {"zero", 0.0},
        {"negzero", -0.0},     {"int", 42L},          {"negint", -7L},
        {"t", true},           {"f", false},          {"str", "a\"b\\c\n\t\x01"},
        {"vec", std::vector<std::string>{"x", "y/N"}}, {"empty_vec", std::vector<double>{}},
    };
    for (auto& kv : cases) std::printf("@ json.%s %s\n", kv.first.c_str(), kv.second.dump().c_str());
    ConfigDict c = get_standard_rigid_bumpy_config(32, 0.75);
    c["log_names"] = std::vector<std::string>{"step", "PE/N"};
    c["nested"] = {{"deep", {{"x", 1e-12}}}, {"empty", ConfigValue::object_value()}};
    std::printf("=== dump4\n%s\n=== end\n", c.dump(4).c_str());
    ConfigValue back = ConfigValue::parse(c.dump(4));
    std::printf("@ json.roundtrip %d\n", back.dump(4) == c.dump(4) ? 1 : 0);
    ConfigValue spaced = ConfigValue::parse(" { \"a\" : [ 1 , 2.5e3 , -0.0 , \"\\u00e9\\ud83d\\ude00\" ] , \"b\" : null } ");
    std::printf("@ json.parsed %s\n", spaced.dump().c_str());
    expect_throw<std::invalid_argument>("json_trailing", [&] { ConfigValue::parse("{} x"); });
    expect_throw<std::invalid_argument>("json_unterminated", [&] { ConfigValue::parse("{\"a\": [1, 2"); });
    expect_throw<std::invalid_argument>("json_bad_escape", [&] { ConfigValue::parse("\"\\q\""); });
}

std::vector<std::string> with_pairs(Particle& p, bool bumpy) {
    std::vector<std::string> n = p.getFundamentalValues();
    for (const char* x : {"force_pairs", "distance_pairs", "overlap_pairs", "radsum_pairs",
                          "pair_separation_angle", "pair_ids", "potential_pairs", "contact_counts",
                          "hessian_pairs_xx", "hessian_pairs_xy", "hessian_pairs_yx",
                          "hessian_pairs_yy", "hessian_ii_xx", "hessian_ii_xy", "hessian_ii_yx",
                          "hessian_ii_yy"})
        n.push_back(x);
    if (bumpy)
        for (const char* x : {"angle_pairs_i", "angle_pairs_j", "this_vertex_contact_counts",
                              "pair_friction_coefficient", "pair_vertex_overlaps",
                              "hessian_pairs_xt", "hessian_pairs_yt", "hessian_pairs_tt",
                              "hessian_pairs_tx", "hessian_pairs_ty", "hessian_ii_xt",
                              "hessian_ii_yt", "hessian_ii_tt", "hessian_ii_tx", "hessian_ii_ty",
                              "area"})
            n.push_back(x);
    return n;
}

void compare_state(Particle& a, Particle& b, const std::string& tag) {
    auto pa = a.getPositions(), pb = b.getPositions();
    auto va = a.getVelocities(), vb = b.getVelocities();
    long flags = (pa == pb) | (va == vb) << 1 | (a.getRadii() == b.getRadii()) << 2 |
                 (a.getMasses() == b.getMasses()) << 3 | (a.getBoxSize() == b.getBoxSize()) << 4;
    auto* ra = dynamic_cast<RigidBumpy*>(&a);
    auto* rb = dynamic_cast<RigidBumpy*>(&b);
    if (ra && rb)  // orientations and vertex rings too
        flags &= ~(long)((ra->getAngles() != rb->getAngles() ||
                          ra->getVertexPositions() != rb->getVertexPositions())
                             ? 1
                             : 0);
    put_int((tag + ".same_state").c_str(), flags == 31);
    put_int((tag + ".same_flags").c_str(), flags);
}

// A run written through IOManager (energy, console, state and restart groups; NVE with the
// steps between logs run as one device loop), then read back with loadParticle from the
// restart, init and frame sources and with load_configs.
void scenario_io(const std::string& dir, bool bumpy) {
    namespace fs = std::filesystem;
    fs::path root = fs::path(dir);
    fs::path out = root / "out";
    ConfigDict pc;
    if (bumpy) {
        pc = get_standard_rigid_bumpy_config(48, 0.7);
        pc["n_vertices_per_small_particle"] = 12;
        pc["n_vertices_per_large_particle"] = 16;
        pc["vertex_radius"] = RigidBumpy::calculateVertexRadius(1.0, 12, 1.0);
        pc["n_vertices"] = 24L * 12 + 24L * 16;
        pc["neighbor_list_config"]["neighbor_list_update_method"] = "verlet";
    } else {
        pc = get_standard_disk_config(96, 0.84);
    }
    pc["seed"] = 17;
    auto p = createParticle(pc);
    minimizeFire(*p, 0.1, 1e-2, 200, 1000);
    p->setRandomVelocities(1e-3);
    ConfigDict nve_config = get_nve_config_dict(0.01 / p->getTimeUnit());
    NVE nve(*p, nve_config);
    const long num_steps = 120;
    ConfigDict console = config_from_names_lin({"step", "PE/N", "KE/N", "T"}, num_steps, 4, "console");
    ConfigDict energy = config_from_names_lin(
        {"step", "PE/N", "KE/N", "TE/N", "T", "phi", "P", "Zp", "PE/dof", "KE"}, num_steps, 12, "energy");
    std::vector<std::string> state_names = {"positions", "velocities", "forces", "radii",
                                            "kinetic_energy", "potential_energy", "contact_counts",
                                            "stress_tensor", "static_particle_index", "box_size"};
    if (bumpy)
        for (const char* x : {"angles", "angular_velocities", "torques", "vertex_positions",
                              "num_vertices_in_particle", "vertex_masses", "moments_of_inertia"})
            state_names.push_back(x);
    ConfigDict state_cfg = config_from_names_log(state_names, num_steps, 1, 10, "state");
    ConfigDict restart = config_from_names_lin_everyN(with_pairs(*p, bumpy), 50, "restart");
    {
        IOManager io({console, energy, state_cfg, restart}, *p, &nve, out.string(), 4, true);
        io.write_params();
        long step = 0, pending = 0;
        while (step < num_steps) {
            ++pending;
            if (io.prepare(step)) {
                nve.run(pending);
                pending = 0;
                io.write(step);
            }
            step++;
        }
        if (pending) nve.run(pending);
        io.log(step, true);
    }
    state(*p, "io.final");
    put_int("io.row", p->max_neighbors_allocated);
    auto [q, step_r] = loadParticle(out, "restart");
    put_int("io.restart_step", step_r);
    compare_state(*p, *q, "io.restart");
    put("io.restart.pe", q->totalPotentialEnergy());
    auto [r, step_i] = loadParticle(out, "init");
    put_int("io.init_step", step_i);
    state(*r, "io.init");
    auto [f, step_f] = loadParticle(out, "frame", -1);
    put_int("io.frame_step", step_f);
    compare_state(*p, *f, "io.frame");
    // resume through a run directory that names only the input
    fs::path run = root / "resume";
    fs::create_directories(run / "system");
    ConfigDict rc;
    rc["minimize"] = false;
    rc["input_dir"] = out.string();
    rc["output_dir"] = "";
    rc.save(run / "system" / "run_config.json");
    std::string a0 = "host_scenarios", a1 = run.string();
    char* argv[] = {a0.data(), a1.data()};
    auto [lp, lstep, lcon, len, lst, lrun] = load_configs(2, argv);
    put_int("io.resume_step", lstep);
    compare_state(*p, *lp, "io.resume");
    put_int("io.resume_energy_group", len.at("group_name").get<std::string>() == "energy");
    // a state written by write_restart_file and read back from a named source
    {
        IOManager io2({restart}, *p, nullptr, (root / "out2").string(), 1, true);
        io2.write_params();
        io2.write_restart_file(0, "last_state");
    }
    p->scaleBox(1.01);
    long before = p->n_particles;
    p->load(root / "out2", "last_state");
    put_int("io.last_state_n", p->n_particles == before);
    compare_state(*p, *q, "io.last_state");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 3) {
        std::string s = argv[1];
        try {
            if (s == "io_disk") scenario_io(argv[2], false);
            else if (s == "io_bumpy") scenario_io(argv[2], true);
            else {
                std::fprintf(stderr, "unknown scenario %s\n", s.c_str());
                return 2;
            }
        } catch (const std::exception& e) {
            std::fprintf(stderr, "scenario %s failed: %s\n", s.c_str(), e.what());
            return 1;
        }
        std::printf("@ done 1\n");
        return 0;
    }
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <scenario>\n", argv[0]);
        return 2;
    }
    std::string s = argv[1];
    try {
        if (s == "rng") scenario_rng();
        else if (s == "config") scenario_config();
        else if (s == "errors") scenario_errors();
        else if (s == "disk_nve") scenario_disk_nve();
        else if (s == "disk_fire") scenario_disk_fire();
        else if (s == "disk_adam") scenario_disk_adam();
        else if (s == "disk_compress") scenario_disk_compress();
        else if (s == "disk_jam_fire") scenario_disk_jam_fire();
        else if (s == "bumpy_init") scenario_bumpy_init();
        else if (s == "json") scenario_json();
        else {
            std::fprintf(stderr, "unknown scenario %s\n", s.c_str());
            return 2;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "scenario %s failed: %s\n", s.c_str(), e.what());
        return 1;
    }
    std::printf("@ done 1\n");
    return 0;
}
