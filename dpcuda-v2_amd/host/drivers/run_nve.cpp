// run_nve -- the flow of scripts/runNVE.cpp:15-55 over the C++ host API: particle,
// optional Gaussian velocities at `temperature`, NVE with dt = dt_dimless / time unit,
// num_steps of nve.step() (or one device-resident nve.run with loop=device).
//
//   run_nve <run directory>      as the reference (load_configs, IOManager logs)
//   run_nve n_particles=262144 packing_fraction=0.85 num_steps=10000 dt_dimless=0.01
//           temperature=1e-5 seed=0 fire_steps=200 loop=device
#include <chrono>
#include <iostream>

#include "../integrator.hpp"
#include "../io.hpp"
#include "../routines.hpp"
#include "run_args.hpp"

// scripts/runNVE.cpp:15-55.  The NVE steps between two logged steps run as one
// device-resident loop (NVE::run gives the same states as that many NVE::step calls).
static int run_directory(int argc, char** argv) {
    auto [particle, step, console_config, energy_config, state_config, run_config] =
        load_configs(argc, argv);
    long num_steps = run_config["num_steps"].get<long>() + step;
    double dt_dimless = run_config["dt_dimless"].get<double>();
    double temperature = run_config["temperature"].get<double>();
    std::filesystem::path output_dir = run_config["output_dir"].get<std::filesystem::path>();
    bool overwrite = true;
    if (temperature > 0) particle->setRandomVelocities(temperature);
    ConfigDict nve_config_dict = get_nve_config_dict(dt_dimless / particle->getTimeUnit());
    NVE nve(*particle, nve_config_dict);
    std::vector<std::string> init_names = particle->getFundamentalValues();
    std::vector<ConfigDict> log_group_configs = {
        console_config, energy_config, state_config,
        config_from_names_lin_everyN(init_names, 10000, "restart")};
    auto io_ptr = std::make_unique<IOManager>(log_group_configs, *particle, &nve,
                                              output_dir.string(), 20, overwrite);
    IOManager& io = *io_ptr;
    io.write_params();
    run_config.save(output_dir / "system" / "run_config.json");
    const long first_step = step;
    auto start_time = std::chrono::high_resolution_clock::now();
    // "loop": "host" in the run config steps one NVE::step per host call, as the reference
    const bool per_step = run_config.contains("loop") && run_config["loop"].get<std::string>() == "host";
    long pending = 0;
    while (step < num_steps) {
        if (per_step) {
            nve.step();
            io.log(step);
            step++;
            continue;
        }
        ++pending;  // nve.step() of this step
        if (io.prepare(step)) {
            nve.run(pending);
            pending = 0;
            io.write(step);
        }
        step++;
    }
    if (pending) nve.run(pending);
    io.log(step, true);
    io_ptr.reset();  // queued state writes finish here
    auto end_time = std::chrono::high_resolution_clock::now();
    std::cout << "Time taken: "
              << std::chrono::duration_cast<std::chrono::seconds>(end_time - start_time).count()
              << " seconds" << std::endl;
    double secs = std::chrono::duration<double>(end_time - start_time).count();
    std::cout << "particle-steps/s " << particle->n_particles * (num_steps - first_step) / secs
              << " (logging included, " << secs << " s)" << std::endl;
    return 0;
}

int main(int argc, char** argv) {
    if (directory_mode(argc, argv)) {
        try {
            return run_directory(argc, argv);
        } catch (const std::exception& e) {
            std::cerr << "run_nve: " << e.what() << std::endl;
            return 1;
        }
    }
    ConfigDict run_config;
    run_config["particle_type"] = "Disk";
    run_config["n_particles"] = 1024L;
    run_config["packing_fraction"] = 0.8;
    run_config["seed"] = 0L;
    run_config["num_steps"] = 10000L;
    run_config["dt_dimless"] = 0.01;
    run_config["temperature"] = 1e-4;
    run_config["fire_steps"] = 200L;
    run_config["loop"] = "host";
    run_config = parse_run_args(argc, argv, run_config);
    try {
        auto particle = particle_from_args(run_config);
        long num_steps = run_config["num_steps"].get<long>();
        double dt_dimless = run_config["dt_dimless"].get<double>();
        double temperature = run_config["temperature"].get<double>();
        long fire_steps = run_config["fire_steps"].get<long>();
        if (fire_steps > 0) minimizeFire(*particle, 0.1, 1e-2, fire_steps, 1000L);
        if (temperature > 0) particle->setRandomVelocities(temperature);

        ConfigDict nve_config_dict = get_nve_config_dict(dt_dimless / particle->getTimeUnit());
        NVE nve(*particle, nve_config_dict);
        double e0 = particle->totalKineticEnergy() + particle->totalPotentialEnergy();

        auto start_time = std::chrono::high_resolution_clock::now();
        if (run_config["loop"].get<std::string>() == "device") {
            nve.run(num_steps);
        } else {
            for (long step = 0; step < num_steps; ++step) nve.step();
        }
        double e1 = particle->totalKineticEnergy() + particle->totalPotentialEnergy();
        auto end_time = std::chrono::high_resolution_clock::now();
        double secs = std::chrono::duration<double>(end_time - start_time).count();
        std::cout.precision(17);
        std::cout << "steps " << num_steps << " E0/N " << e0 / particle->n_particles << " E1/N "
                  << e1 / particle->n_particles << " rel drift " << (e1 - e0) / std::abs(e0)
                  << std::endl;
        std::cout << "particle-steps/s " << particle->n_particles * num_steps / secs << std::endl;
        std::cout << "Time taken: " << static_cast<long>(secs) << " seconds" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "run_nve: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
