// run_jam -- the packing-fraction bisection of scripts/runJam.cpp:14-106 over the C++
// host API: decompress until unjammed, then FIRE-minimise at each packing fraction,
// bisecting between the last unjammed (PE/N <= tolerance) and the first jammed one until
// |phi_high/phi_low - 1| < phi_tolerance, then check stability with a small extra
// compression (is_stable_configuration).
//
//   run_jam <run directory>      as the reference (load_configs, IOManager logs)
//   run_jam n_particles=1024 packing_fraction=0.8 compression_increment=1e-3
#include <chrono>
#include <cmath>
#include <iostream>

#include "../io.hpp"
#include "../routines.hpp"
#include "run_args.hpp"

static bool is_stable_configuration(Particle& particle, double delta_phi, double avg_pe_tolerance,
                                    double avg_pe_diff_tolerance) {
    particle.setLastState();
    particle.scaleToPackingFractionFull(particle.getPackingFraction() + delta_phi);
    minimizeFire(particle, avg_pe_tolerance, avg_pe_diff_tolerance);
    double final_pe = particle.totalPotentialEnergy() / particle.n_particles;
    particle.revertToLastState();
    return final_pe > 1e-1 * std::pow(delta_phi, particle.n_c);
}

// scripts/runJam.cpp:25-113
static int run_directory(int argc, char** argv) {
    auto [particle, compression_step, console_config, energy_config, state_config, run_config] =
        load_configs(argc, argv);
    std::string particle_type = particle->getConfig().at("particle_type").get<std::string>();
    long num_compression_steps = run_config["num_compression_steps"].get<long>();
    double compression_increment = run_config["compression_increment"].get<double>();
    double avg_pe_tolerance = run_config["avg_pe_tolerance"].get<double>();
    double avg_pe_diff_tolerance = run_config["avg_pe_diff_tolerance"].get<double>();
    double phi_tolerance = run_config["phi_tolerance"].get<double>();
    double delta_phi_stability = run_config["delta_phi_stability"].get<double>();
    std::filesystem::path output_dir = run_config["output_dir"].get<std::filesystem::path>();
    bool overwrite = true;
    long n_particles = particle->n_particles;
    while (true) {  // start with no contacts
        minimizeFire(*particle, avg_pe_tolerance, avg_pe_diff_tolerance);
        if (particle->totalPotentialEnergy() / n_particles < avg_pe_tolerance) break;
        particle->scaleBox(1.1);
    }
    std::vector<std::string> init_names = particle->getFundamentalValues();
    std::vector<std::string> pair_names = {
        "force_pairs",      "distance_pairs",   "overlap_pairs",    "radsum_pairs",
        "pair_separation_angle", "pair_ids",    "potential_pairs",  "contact_counts",
        "hessian_pairs_xx", "hessian_pairs_xy", "hessian_pairs_yx", "hessian_pairs_yy",
        "hessian_ii_xx",    "hessian_ii_xy",    "hessian_ii_yx",    "hessian_ii_yy"};
    if (particle_type == "RigidBumpy")
        for (const char* n :
             {"angle_pairs_i", "angle_pairs_j", "this_vertex_contact_counts",
              "pair_friction_coefficient", "pair_vertex_overlaps", "hessian_pairs_xt",
              "hessian_pairs_yt", "hessian_pairs_tt", "hessian_pairs_tx", "hessian_pairs_ty",
              "hessian_ii_xt", "hessian_ii_yt", "hessian_ii_tt", "hessian_ii_tx",
              "hessian_ii_ty", "area"})
            pair_names.push_back(n);
    init_names.insert(init_names.end(), pair_names.begin(), pair_names.end());
    std::vector<ConfigDict> log_group_configs = {
        console_config, energy_config, state_config,
        config_from_names_lin_everyN(init_names, 1, "restart")};
    IOManager io_manager(log_group_configs, *particle, nullptr, output_dir.string(), 1, overwrite);
    io_manager.write_params();
    run_config.save(output_dir / "system" / "run_config.json");

    auto start_time = std::chrono::high_resolution_clock::now();
    particle->setLastState();
    double phi_low = particle->getPackingFraction();
    double phi;
    double phi_high = -1.0;
    while (compression_step < num_compression_steps) {
        phi = particle->getPackingFraction();
        minimizeFire(*particle, avg_pe_tolerance, avg_pe_diff_tolerance, 1000000L);
        double pe_per_particle = particle->totalPotentialEnergy() / n_particles;
        particle->countContacts();
        if (pe_per_particle > avg_pe_tolerance) {  // jammed
            phi_high = phi;
            phi = (phi_high + phi_low) / 2.0;
            particle->revertToLastState();
        } else {  // unjammed
            particle->setLastState();
            io_manager.log(compression_step, true);  // a smooth trajectory
            phi_low = phi;
            phi = phi_high > 0 ? (phi_high + phi_low) / 2.0 : phi + compression_increment;
        }
        if (std::abs(phi_high / phi_low - 1) < phi_tolerance) {  // converged
            if (is_stable_configuration(*particle, delta_phi_stability, avg_pe_tolerance,
                                        avg_pe_diff_tolerance))
                break;
            std::cout << "unstable" << std::endl;
            phi_low = phi;
            phi_high = -1.0;
            phi = particle->getPackingFraction();
            particle->setLastState();
        }
        particle->scaleToPackingFractionFull(phi);
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
            std::cerr << "run_jam: " << e.what() << std::endl;
            return 1;
        }
    }
    ConfigDict run_config;
    run_config["particle_type"] = "Disk";
    run_config["n_particles"] = 1024L;
    run_config["packing_fraction"] = 0.8;
    run_config["seed"] = 0L;
    run_config["num_compression_steps"] = 1000L;
    run_config["compression_increment"] = 1e-3;
    run_config["avg_pe_tolerance"] = 1e-16;
    run_config["avg_pe_diff_tolerance"] = 1e-16;
    run_config["phi_tolerance"] = 1e-6;
    run_config["delta_phi_stability"] = 1e-4;
    run_config = parse_run_args(argc, argv, run_config);
    try {
        auto particle = particle_from_args(run_config);
        long num_compression_steps = run_config["num_compression_steps"].get<long>();
        double compression_increment = run_config["compression_increment"].get<double>();
        double avg_pe_tolerance = run_config["avg_pe_tolerance"].get<double>();
        double avg_pe_diff_tolerance = run_config["avg_pe_diff_tolerance"].get<double>();
        double phi_tolerance = run_config["phi_tolerance"].get<double>();
        double delta_phi_stability = run_config["delta_phi_stability"].get<double>();
        long n_particles = particle->n_particles;
        while (true) {  // start with no contacts
            minimizeFire(*particle, avg_pe_tolerance, avg_pe_diff_tolerance);
            if (particle->totalPotentialEnergy() / n_particles < avg_pe_tolerance) break;
            particle->scaleBox(1.1);
        }

        auto start_time = std::chrono::high_resolution_clock::now();
        long compression_step = 0;
        particle->setLastState();
        double phi_low = particle->getPackingFraction();
        double phi;
        double phi_high = -1.0;
        while (compression_step < num_compression_steps) {
            phi = particle->getPackingFraction();
            minimizeFire(*particle, avg_pe_tolerance, avg_pe_diff_tolerance, 1000000L);
            double pe_per_particle = particle->totalPotentialEnergy() / n_particles;
            if (pe_per_particle > avg_pe_tolerance) {  // jammed
                phi_high = phi;
                phi = (phi_high + phi_low) / 2.0;
                particle->revertToLastState();
            } else {  // unjammed
                particle->setLastState();
                phi_low = phi;
                phi = phi_high > 0 ? (phi_high + phi_low) / 2.0 : phi + compression_increment;
            }
            if (std::abs(phi_high / phi_low - 1) < phi_tolerance) {  // converged
                if (is_stable_configuration(*particle, delta_phi_stability, avg_pe_tolerance,
                                            avg_pe_diff_tolerance))
                    break;
                std::cout << "unstable" << std::endl;
                phi_low = phi;
                phi_high = -1.0;
                phi = particle->getPackingFraction();
                particle->setLastState();
            }
            particle->scaleToPackingFractionFull(phi);
            compression_step++;
        }
        auto end_time = std::chrono::high_resolution_clock::now();
        particle->countContacts();
        std::cout.precision(17);
        std::cout << "compression_steps " << compression_step << " phi "
                  << particle->getPackingFraction() << " PE/N "
                  << particle->totalPotentialEnergy() / n_particles << " contacts "
                  << particle->getContactCount() << std::endl;
        double secs = std::chrono::duration<double>(end_time - start_time).count();
        std::cout << "Time taken: " << static_cast<long>(secs) << " seconds" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "run_jam: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
