// run_jam_with_fire -- scripts/runArthurJamWithFire.cpp:22-114 over the C++ host API:
// decompress until P < pressure_tolerance_low, then compress by compression_rate with
// FIRE minimisation, bisecting on the box side until the pressure lies in
// [pressure_tolerance_low, pressure_tolerance_high] with more than 2N contacts.
//
//   run_jam_with_fire n_particles=1024 packing_fraction=0.8 compression_rate=0.999
//                     pressure_tolerance_low=1e-7 pressure_tolerance_high=1.1e-7
#include <chrono>
#include <iostream>

#include "../routines.hpp"
#include "run_args.hpp"

int main(int argc, char** argv) {
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
