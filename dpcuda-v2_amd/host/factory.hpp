// factory.hpp -- createParticle and loadParticle (include/particles/factory.h:10-52).
#pragma once
#include <filesystem>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>

#include "particle.hpp"
#include "standard_configs.hpp"

inline std::unique_ptr<Particle> createParticle(ConfigDict& config, bool minimize = false) {
    std::string particle_type = config.at("particle_type").get<std::string>();
    std::unique_ptr<Particle> particle;
    if (particle_type == "Disk") particle = std::make_unique<Disk>();
    else if (particle_type == "RigidBumpy") particle = std::make_unique<RigidBumpy>();
    else throw std::invalid_argument("Invalid particle type: " + particle_type);
    particle->initializeFromConfig(config, minimize);
    return particle;
}

inline std::unique_ptr<Particle> createParticle(long n_particles, double packing_fraction,
                                                std::string particle_type, bool minimize = false) {
    ConfigDict config;
    if (particle_type == "Disk") config = get_standard_disk_config(n_particles, packing_fraction);
    else if (particle_type == "RigidBumpy")
        config = get_standard_rigid_bumpy_config(n_particles, packing_fraction, true);
    else if (particle_type == "RigidBumpyNoRotation")
        config = get_standard_rigid_bumpy_config(n_particles, packing_fraction, false);
    else throw std::invalid_argument("Invalid particle type: " + particle_type);
    return createParticle(config, minimize);
}

// loadParticle (factory.h:35-52): the particle type from <path>/system/particle_config.json,
// then Particle::load from the given source ("init", "restart" or "frame")
inline std::tuple<std::unique_ptr<Particle>, long> loadParticle(std::filesystem::path path,
                                                                std::string source,
                                                                long frame = -2) {
    ConfigDict config = load_config_dict(path / "system" / "particle_config.json");
    std::string particle_type = config.at("particle_type").get<std::string>();
    std::unique_ptr<Particle> particle;
    if (particle_type == "Disk") particle = std::make_unique<Disk>();
    else if (particle_type == "RigidBumpy") particle = std::make_unique<RigidBumpy>();
    else throw std::invalid_argument("Invalid particle type: " + particle_type);
    long step = particle->load(path, source, frame);
    return std::make_tuple(std::move(particle), step);
}
