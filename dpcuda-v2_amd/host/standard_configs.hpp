// standard_configs.hpp -- the standard particle configurations
// (include/particles/standard_configs.h:7-68): same keys, same default values.
#pragma once
#include "config_dict.hpp"

inline ConfigDict get_standard_disk_config(long n_particles, double packing_fraction) {
    ConfigDict config;
    config["particle_type"] = "Disk";
    config["n_particles"] = n_particles;
    config["packing_fraction"] = packing_fraction;
    config["particle_dim_block"] = 256;
    config["e_c"] = 1.0;
    config["n_c"] = 2.0;
    config["mass"] = 1.0;
    config["size_ratio"] = 1.4;
    config["count_ratio"] = 0.5;
    config["seed"] = -1;
    config["neighbor_list_config"] = {{"neighbor_cutoff_multiplier", 1.5},
                                      {"neighbor_displacement_multiplier", 0.2},
                                      {"num_particles_per_cell", 8.0},
                                      {"cell_displacement_multiplier", 0.5},
                                      {"neighbor_list_update_method", "cell"}};
    return config;
}

// 26 / 36 vertices for the small / large half; the vertex radius key is 0 here, as in the
// reference, so a caller sets "vertex_radius" (RigidBumpy::calculateVertexRadius) before use.
inline ConfigDict get_standard_rigid_bumpy_config(long n_particles, double packing_fraction,
                                                  bool rotation = true) {
    const long nvs = 26, nvl = 36;
    const double count_ratio = 0.5;
    long n_small = static_cast<long>(n_particles * count_ratio);
    long n_large = n_particles - n_small;
    ConfigDict config;
    config["particle_type"] = "RigidBumpy";
    config["n_particles"] = n_particles;
    config["n_vertices"] = n_small * nvs + n_large * nvl;
    config["packing_fraction"] = packing_fraction;
    config["particle_dim_block"] = 256;
    config["vertex_dim_block"] = 256;
    config["segment_length_per_vertex_diameter"] = 1.0;
    config["vertex_radius"] = 0;
    config["n_vertices_per_small_particle"] = nvs;
    config["n_vertices_per_large_particle"] = nvl;
    config["size_ratio"] = 1.4;
    config["count_ratio"] = count_ratio;
    config["rotation"] = rotation;
    config["e_c"] = 1.0;
    config["n_c"] = 2.0;
    config["mass"] = 1.0;
    config["seed"] = -1;
    config["neighbor_list_config"] = {{"neighbor_cutoff_multiplier", 1.5},
                                      {"neighbor_displacement_multiplier", 0.2},
                                      {"num_particles_per_cell", 8.0},
                                      {"cell_displacement_multiplier", 0.5},
                                      {"vertex_neighbor_cutoff_multiplier", 1.5},
                                      {"vertex_neighbor_displacement_multiplier", 0.2},
                                      {"neighbor_list_update_method", "cell"}};
    return config;
}
