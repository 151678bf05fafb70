// factory.hpp -- createParticle (include/particles/factory.h:10-33).  Loading a saved
// particle (loadParticle) needs the reference's on-disk format and is not provided.
#pragma once
#include <memory>
#include <stdexcept>
#include <string>

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
