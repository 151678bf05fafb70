// run_jam_with_fire -- scripts/runArthurJamWithFire.cpp:22-114 over the C++ host API:
// decompress until P < pressure_tolerance_low, then compress by compression_rate with
// FIRE minimisation, bisecting on the box side until the pressure lies in
// [pressure_tolerance_low, pressure_tolerance_high] with more than 2N contacts.
//
//   run_jam_with_fire <run directory>   as the reference: IOManager logs, and the last
//                                       unjammed state saved to and reloaded from
//                                       <output_dir>/system/last_state
//   run_jam_with_fire n_particles=1024 packing_fraction=0.8 compression_rate=0.999
//                     pressure_tolerance_low=1e-7 pressure_tolerance_high=1.1e-7
#include <chrono>
#include <cmath>
#include <iostream>

#include "../io.hpp"
#include "../routines.hpp"
#include "run_args.hpp"

// scripts/runArthurJamWithFire.cpp:22-114
static int run_directory(int argc, char** argv) {
    auto [particle, compression_step, console_config, energy_config, state_config, run_config] =
        load_configs(argc, argv);
    std::string particle_type = particle->getConfig().at("particle_type").get<std::string>();
    long num_compression_steps = run_config["num_compression_steps"].get<long>();
    double compression_rate = run_config["compression_rate"].get<double>();
    double pressure_tolerance_low = run_config["pressure_tolerance_low"].get<double>();
    double pressure_tolerance_high = run_config["pressure_tolerance_high"].get<double>();
    double length_tolerance = run_config["length_tolerance"].get<double>();
    std::filesystem::path output_dir = run_config["output_dir"].get<std::filesystem::path>();
    bool overwrite = true;
    long n_particles = particle->n_particles;
    while (true) {  // start with no contacts
        minimizeFire(*particle, 1e-16, 0);
        particle->calculateStressTensor();
        if (particle->getPressure() < pressure_tolerance_low) break;
        particle->scaleBox(1.01);
    }
    std::vector<std::string> init_names = particle->getFundamentalValues();
    std::vector<std::string> pair_names = {
        "force_pairs",     "distance_pairs", "overlap_pairs",    "radsum_pairs",
        "pair_separation_angle", "pair_ids",  "potential_pairs", "contact_counts",
        "hessian_pairs_xx", "hessian_pairs_xy", "hessian_pairs_yy"};
    if (particle_type == "RigidBumpy")
        for (const char* n : {"angle_pairs_i", "angle_pairs_j", "this_vertex_contact_counts",
                              "pair_friction_coefficient", "pair_vertex_overlaps"})
            pair_names.push_back(n);
    init_names.insert(init_names.end(), pair_names.begin(), pair_names.end());
    std::vector<ConfigDict> log_group_configs = {
        console_config, energy_config, state_config,
        config_from_names_lin_everyN(init_names, 1, "restart")};
    IOManager io_manager(log_group_configs, *particle, nullptr, output_dir.string(), 1, overwrite);
    io_manager.write_params();
    run_config.save(output_dir / "system" / "run_config.json");

    auto start_time = std::chrono::high_resolution_clock::now();
    double L_low = -1, L_high = -1, L_last = -1;
    while (compression_step < num_compression_steps) {
        minimizeFire(*particle);
        particle->calculateStressTensor();
        double pressure = particle->getPressure();
        particle->countContacts();
        long num_contacts = static_cast<long>(particle->getContactCount());
        double L = std::sqrt(particle->getBoxArea());
        double L_prior = L;
        io_manager.log(compression_step, true);
        if (pressure < pressure_tolerance_low) {  // unjammed
            L_low = L;
            io_manager.write_restart_file(compression_step, "last_state");
            L_last = L;
            if (L_high != -1) {
                L = (L_high + L_low) / 2;
                L_high = -1;
            } else {
                L = L * compression_rate;
            }
        } else {
            if (pressure > pressure_tolerance_high) {  // jammed: back to the last state
                L_high = L;
                particle->load(output_dir, "last_state");
                L_prior = L_last;
                L = (L_high + L_low) / 2;
            } else if (num_contacts > 2 * n_particles) {  // final state
                break;
            }
        }
        if (L_low > 0 && L_high > 0 && std::abs(L_low / L_high - 1) < length_tolerance) {
            std::cout << "Error: couldnt find jammed state" << std::endl;
            break;
        }
        particle->scaleBox(L / L_prior);
        compression_step++;
    }
    io_manager.log(compression_step + 1, true);
    auto end_time = std::chrono::high_resolution_clock::now();
    std::cout << "Time taken: "
              << std::chrono::duration_cast<std::chrono::seconds>(end_time - start_time).count()
              << " seconds" << std::endl;
    return 0;
}

int main(int argc, char** argv) {
    if (directory_mode(argc, argv)) {
        try {
            return run_directory(argc, argv);
        } catch (const std::exception& e) {
            std::cerr << "run_jam_with_fire: " << e.what() << std::endl;
            return 1;
        }
    }
    ConfigDict run_config;
    run_config["particle_type"] = "Disk";
    run_config["n_particles"] = 1024L;
    run_config["packing_fraction"] = 0.8;
    run_config["seed"] = 0L;
    run_config["num_compression_steps"] = 100000L;
    run_config["compression_rate"] = 0.999;
    run_config["pressure_tolerance_low"] = 1e-7;
    run_config["pressure_tolerance_high"] = 1.1e-7;
    run_config["length_tolerance"] = 1e-12;
    run_config["fmax_target"] = 0.0;
    run_config = parse_run_args(argc, argv, run_config);
    try {
        auto particle = particle_from_args(run_config);
        auto start_time = std::chrono::high_resolution_clock::now();
        JamResult r = jamWithFire(*particle, run_config["pressure_tolerance_low"].get<double>(),
                                  run_config["pressure_tolerance_high"].get<double>(),
                                  run_config["compression_rate"].get<double>(),
                                  run_config["length_tolerance"].get<double>(),
                                  run_config["num_compression_steps"].get<long>(),
                                  run_config["fmax_target"].get<double>());
        auto end_time = std::chrono::high_resolution_clock::now();
        std::cout.precision(17);
        std::cout << "status " << r.status << " compression_steps " << r.compression_steps
                  << " fire_steps " << r.fire_steps << " P " << r.pressure << " phi "
                  << r.packing_fraction << " PE " << r.potential_energy << " contacts "
                  << r.contacts << std::endl;
        double secs = std::chrono::duration<double>(end_time - start_time).count();
        std::cout << "Time taken: " << static_cast<long>(secs) << " seconds" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "run_jam_with_fire: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
