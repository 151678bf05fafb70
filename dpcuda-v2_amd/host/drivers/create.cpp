// create -- scripts/create.cpp:15-31: the particle of a run directory (load_configs) with
// its parameters, run config and initial state written through IOManager
// (<output_dir>/system/{init,restart}/, the log configs, particle_config.json).
//
//   create <run directory>
#include <iostream>

#include "../io.hpp"
#include "run_args.hpp"

int main(int argc, char** argv) {
    if (!directory_mode(argc, argv)) {
        std::cerr << "usage: create <run directory>" << std::endl;
        return 2;
    }
    try {
        auto [particle, step, console_config, energy_config, state_config, run_config] =
            load_configs(argc, argv);
        (void)step;
        std::filesystem::path output_dir = run_config["output_dir"].get<std::filesystem::path>();
        std::vector<std::string> init_names = particle->getFundamentalValues();
        std::vector<ConfigDict> log_group_configs = {
            console_config, energy_config, state_config,
            config_from_names_lin_everyN(init_names, 1, "restart")};
        IOManager io_manager(log_group_configs, *particle, nullptr, output_dir.string(), 1, true);
        io_manager.write_params();
        run_config.save(output_dir / "system" / "run_config.json");
        io_manager.log(0, true);
    } catch (const std::exception& e) {
        std::cerr << "create: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
