// run_args.hpp -- the two ways a driver is started:
//   driver <run directory>      the reference's form: <run directory>/system/run_config.json
//                               read by load_configs (include/io/arg_parsing.h:7-64), logs
//                               written through IOManager as the reference's driver does;
//   driver key=value ...        a particle from the standard configuration and the run
//                               parameters on the command line, nothing written to disk.
// key=value values that parse as integers become long, other numbers double, true/false
// bool, anything else a string.
#pragma once
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>

#include "../config_dict.hpp"
#include "../factory.hpp"

// one argument that is not key=value: the reference's run-directory form
inline bool directory_mode(int argc, char** argv) {
    return argc == 2 && std::string(argv[1]).find('=') == std::string::npos;
}

inline ConfigDict parse_run_args(int argc, char** argv, ConfigDict defaults) {
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto eq = a.find('=');
        if (eq == std::string::npos) {
            std::cerr << "arguments are key=value, got " << a << std::endl;
            std::exit(2);
        }
        std::string k = a.substr(0, eq), v = a.substr(eq + 1);
        char* end = nullptr;
        long l = std::strtol(v.c_str(), &end, 10);
        if (!v.empty() && *end == '\0') {
            defaults[k] = l;
            continue;
        }
        double d = std::strtod(v.c_str(), &end);
        if (!v.empty() && *end == '\0') defaults[k] = d;
        else if (v == "true" || v == "false") defaults[k] = (v == "true");
        else defaults[k] = v;
    }
    return defaults;
}

// the particle the reference's drivers load from disk, created here from the standard
// configuration of the requested type, size, packing fraction and seed
inline std::unique_ptr<Particle> particle_from_args(ConfigDict& run_config) {
    std::string type = run_config.at("particle_type").get<std::string>();
    long n = run_config.at("n_particles").get<long>();
    double phi = run_config.at("packing_fraction").get<double>();
    ConfigDict config = type == "Disk" ? get_standard_disk_config(n, phi)
                                       : get_standard_rigid_bumpy_config(n, phi, type == "RigidBumpy");
    config["seed"] = run_config.at("seed").get<long>();
    if (type != "Disk") {
        long nvs = run_config.contains("n_vertices_per_small_particle")
                       ? run_config.at("n_vertices_per_small_particle").get<long>()
                       : config.at("n_vertices_per_small_particle").get<long>();
        long nvl = run_config.contains("n_vertices_per_large_particle")
                       ? run_config.at("n_vertices_per_large_particle").get<long>()
                       : config.at("n_vertices_per_large_particle").get<long>();
        config["n_vertices_per_small_particle"] = nvs;
        config["n_vertices_per_large_particle"] = nvl;
        config["vertex_radius"] = RigidBumpy::calculateVertexRadius(
            1.0, nvs, config.at("segment_length_per_vertex_diameter").get<double>());
    }
    return createParticle(config, false);
}
