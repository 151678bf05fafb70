// integrator.hpp -- Integrator, NVE, Fire, Adam (include/integrator/*.h) over the C ABI.
#pragma once
#include "config_dict.hpp"
#include "particle.hpp"

// get_nve_config_dict (include/integrator/nve.h:6-11)
inline ConfigDict get_nve_config_dict(double dt) {
    ConfigDict config;
    config["integrator_type"] = "NVE";
    config["dt"] = dt;
    return config;
}

// get_fire_config_dict (include/integrator/fire.h:7-13)
inline ConfigDict get_fire_config_dict(double alpha_init, double dt) {
    ConfigDict config;
    config["integrator_type"] = "Fire";
    config["alpha_init"] = alpha_init;
    config["dt"] = dt;
    return config;
}

// get_adam_config_dict (include/integrator/adam.h:7-15)
inline ConfigDict get_adam_config_dict(double alpha, double beta1, double beta2, double epsilon) {
    ConfigDict config;
    config["integrator_type"] = "Adam";
    config["alpha"] = alpha;
    config["beta1"] = beta1;
    config["beta2"] = beta2;
    config["epsilon"] = epsilon;
    return config;
}

// Integrator (include/integrator/integrator.h:21-43)
class Integrator {
   protected:
    Particle& particle;

   public:
    ConfigDict& config;
    Integrator(Particle& particle, ConfigDict& config) : particle(particle), config(config) {}
    virtual ~Integrator() = default;
    virtual void step() = 0;
};

// NVE (include/integrator/nve.h:18-37; src/integrator/nve.cu:15-25)
class NVE : public Integrator {
   public:
    NVE(Particle& particle, ConfigDict& config);
    double dt;
    // one velocity-Verlet step: kick(dt/2), drift(dt), neighbour check, zero, forces,
    // kick(dt/2), issued as separate device calls exactly as NVE::step orders them
    void step() override;
    // n_steps of step() run as a device-resident loop (same results, no host sync per step)
    void run(long n_steps);
};

// Fire (include/integrator/fire.h:15-35; src/integrator/fire.cu:13-46)
class Fire : public Integrator {
   public:
    Fire(Particle& particle, ConfigDict& config);
    double alpha;
    double alpha_init;
    double dt;
    double dt_max;
    double dt_min;
    double f_inc = 1.1;
    double f_dec = 0.5;
    double f_alpha = 0.99;
    long N_min = 5;
    long N_good = 0;
    long N_bad = 0;
    long N_bad_max = 10;
    bool stopped = false;
    void step() override;
};

// Adam (include/integrator/adam.h:17-29; src/integrator/adam.cu:11-18)
class Adam : public Integrator {
   public:
    Adam(Particle& particle, ConfigDict& config);
    double alpha;
    double beta1;
    double beta2;
    double epsilon;
    // zero forces, neighbour check, forces, then the Adam position update for `step`
    void minimize(long step);
    void step() override { minimize(step_++); }

   private:
    long step_ = 0;
};
