// run_fire -- the flow of scripts/runFire.cpp:14-51 over the C++ host API:
// minimizeFire(particle, fire_alpha_init, fire_dt, fire_num_steps, fire_log_every_n).
//
//   run_fire <run directory>     as the reference (load_configs, IOManager logs)
//   run_fire n_particles=65536 packing_fraction=0.8 particle_type=RigidBumpy
//            n_vertices_per_small_particle=20 n_vertices_per_large_particle=28
#include <chrono>
#include <cmath>
#include <iostream>

#include "../io.hpp"
#include "../routines.hpp"
#include "run_args.hpp"

// scripts/runFire.cpp:14-51
static int run_directory(int argc, char** argv) {
    auto [particle, step, console_config, energy_config, state_config, run_config] =
        load_configs(argc, argv);
    (void)step;
    std::filesystem::path output_dir = run_config["output_dir"].get<std::filesystem::path>();
    double alpha_init = run_config["fire_alpha_init"];
    double dt = run_config["fire_dt"];
    long num_steps = run_config["fire_num_steps"];
    long log_every_n = run_config["fire_log_every_n"];
    bool overwrite = true;
    std::string particle_type = particle->getConfig().at("particle_type").get<std::string>();
    std::vector<std::string> init_names = particle->getFundamentalValues();
    std::vector<std::string> pair_names = {"force_pairs",   "distance_pairs",
                                           "overlap_pairs", "radsum_pairs",
                                           "pair_separation_angle", "pair_ids",
                                           "potential_pairs", "contact_counts"};
    if (particle_type == "RigidBumpy")
        for (const char* n : {"angle_pairs_i", "angle_pairs_j", "this_vertex_contact_counts",
                              "pair_friction_coefficient", "pair_vertex_overlaps"})
            pair_names.push_back(n);
    init_names.insert(init_names.end(), pair_names.begin(), pair_names.end());
    std::vector<ConfigDict> log_group_configs = {
        console_config, energy_config, state_config,
        config_from_names_lin_everyN(init_names, 10000, "restart")};
    IOManager io(log_group_configs, *particle, nullptr, output_dir.string(), 20, overwrite);
    io.write_params();
    run_config.save(output_dir / "system" / "run_config.json");
    auto start_time = std::chrono::high_resolution_clock::now();
    minimizeFire(*particle, alpha_init, dt, num_steps, log_every_n);
    io.log(0, true);
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
            std::cerr << "run_fire: " << e.what() << std::endl;
            return 1;
        }
    }
    ConfigDict run_config;
    run_config["particle_type"] = "Disk";
    run_config["n_particles"] = 1024L;
    run_config["packing_fraction"] = 0.85;
    run_config["seed"] = 0L;
    run_config["fire_alpha_init"] = 0.1;
    run_config["fire_dt"] = 1e-2;
    run_config["fire_num_steps"] = 100000L;
    run_config["fire_log_every_n"] = 1000L;
    run_config["fmax_target"] = 0.0;
    run_config = parse_run_args(argc, argv, run_config);
    try {
        auto particle = particle_from_args(run_config);
        double alpha_init = run_config["fire_alpha_init"];
        double dt = run_config["fire_dt"];
        long num_steps = run_config["fire_num_steps"];
        long log_every_n = run_config["fire_log_every_n"];
        double fmax = run_config["fmax_target"];

        auto start_time = std::chrono::high_resolution_clock::now();
        long steps = -1;
        if (fmax > 0) steps = minimizeFireToMaxForce(*particle, fmax, num_steps, alpha_init, dt);
        else minimizeFire(*particle, alpha_init, dt, num_steps, log_every_n);
        auto end_time = std::chrono::high_resolution_clock::now();
        particle->calculateStressTensor();
        std::cout.precision(17);
        std::cout << "PE/N " << particle->totalPotentialEnergy() / particle->n_particles
                  << " P " << particle->getPressure() << " phi " << particle->getPackingFraction();
        if (steps >= 0) std::cout << " fire_steps " << steps;
        std::cout << std::endl;
        double secs = std::chrono::duration<double>(end_time - start_time).count();
        std::cout << "Time taken: " << static_cast<long>(secs) << " seconds" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "run_fire: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
