// run_jam -- the packing-fraction bisection of scripts/runJam.cpp:14-106 over the C++
// host API: decompress until unjammed, then FIRE-minimise at each packing fraction,
// bisecting between the last unjammed (PE/N <= tolerance) and the first jammed one until
// |phi_high/phi_low - 1| < phi_tolerance, then check stability with a small extra
// compression (is_stable_configuration).
//
//   run_jam n_particles=1024 packing_fraction=0.8 compression_increment=1e-3
#include <chrono>
#include <cmath>
#include <iostream>

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

int main(int argc, char** argv) {
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
