// integrator.cpp -- NVE, Fire, Adam over the C ABI (see integrator.hpp).
#include "integrator.hpp"

#include <stdexcept>

namespace {
void check(int status) {
    if (status != 0) throw std::runtime_error(dpm_last_error());
}
}  // namespace

NVE::NVE(Particle& p, ConfigDict& c) : Integrator(p, c), dt(c.at("dt").get<double>()) {}

void NVE::step() {
    particle.updateVelocities(dt * 0.5);
    particle.updatePositions(dt);
    particle.checkForNeighborUpdate();
    particle.zeroForceAndPotentialEnergy();
    particle.calculateForces();
    particle.updateVelocities(dt * 0.5);
}

void NVE::run(long n_steps) {
    if (n_steps < 0) throw std::invalid_argument("NVE::run: negative step count");
    check(dpm_nve_run(particle.handle(), dt, n_steps));
}

// Fire::Fire (src/integrator/fire.cu:3-11): dt_max = 10 dt, dt_min = 1e-3 dt
Fire::Fire(Particle& p, ConfigDict& c) : Integrator(p, c) {
    alpha_init = c.at("alpha_init").get<double>();
    alpha = alpha_init;
    dt = c.at("dt").get<double>();
    dt_max = 10.0 * dt;
    dt_min = 0.001 * dt;
}

void Fire::step() {
    dpm_fire_state st;
    st.alpha = alpha;
    st.alpha_init = alpha_init;
    st.dt = dt;
    st.dt_max = dt_max;
    st.dt_min = dt_min;
    st.n_good = N_good;
    st.n_bad = N_bad;
    st.stopped = stopped ? 1 : 0;
    st.power = 0.0;
    check(dpm_fire_step(particle.handle(), &st));
    alpha = st.alpha;
    dt = st.dt;
    N_good = st.n_good;
    N_bad = st.n_bad;
    stopped = st.stopped != 0;
}

Adam::Adam(Particle& p, ConfigDict& c) : Integrator(p, c) {
    alpha = c.at("alpha").get<double>();
    beta1 = c.at("beta1").get<double>();
    beta2 = c.at("beta2").get<double>();
    epsilon = c.at("epsilon").get<double>();
}

void Adam::minimize(long step) {
    check(dpm_adam_step(particle.handle(), step, alpha, beta1, beta2, epsilon));
}
