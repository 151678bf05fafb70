// routines.cpp -- minimisation and jamming routines over the C ABI (see routines.hpp).
#include "routines.hpp"

#include "io.hpp"

#include <cmath>
#include <iostream>
#include <stdexcept>

namespace {
void check(int status) {
    if (status != 0) throw std::runtime_error(dpm_last_error());
}

dpm_fire_result run_fire(Particle& particle, double alpha_init, double dt, double pe_t,
                         double diff_t, double fmax_t, long num_steps) {
    dpm_fire_params p;
    p.alpha_init = alpha_init;
    p.dt = dt;
    p.avg_pe_target = pe_t;
    p.avg_pe_diff_target = diff_t;
    p.fmax_target = fmax_t;
    p.max_steps = num_steps;
    dpm_fire_result r;
    check(dpm_minimize_fire(particle.handle(), &p, &r));
    if (!r.converged && !r.stopped)
        std::cout << "Warning: minimization did not converge" << std::endl;
    return r;
}

// the inner Adam loop shared by the compression routines (compression.cu)
double adam_relax(Particle& particle, Adam& adam, long num_adam_steps, double pe_t, double diff_t,
                  long* steps_done) {
    double dof = static_cast<double>(particle.n_dof);
    double avg_pe = 0.0, last = 0.0;
    long k = 0;
    while (k < num_adam_steps) {
        adam.minimize(k);
        avg_pe = particle.totalPotentialEnergy() / dof / particle.e_c;
        double diff = std::abs(avg_pe - last);
        last = avg_pe;
        if (diff < diff_t || avg_pe < pe_t) break;
        k++;
    }
    if (steps_done) *steps_done = k;
    return avg_pe;
}
}  // namespace

void minimizeFire(Particle& particle, double alpha_init, double dt, double avg_pe_target,
                  double avg_pe_diff_target, long num_steps, long) {
    run_fire(particle, alpha_init, dt, avg_pe_target, avg_pe_diff_target, 0.0, num_steps);
}
void minimizeFire(Particle& particle, double alpha_init, double dt, long num_steps, long log_every_n) {
    minimizeFire(particle, alpha_init, dt, 1e-16, 1e-16, num_steps, log_every_n);
}
void minimizeFire(Particle& particle, long log_every_n) {
    minimizeFire(particle, 0.1, 1e-2, 100000L, log_every_n);
}
void minimizeFire(Particle& particle, double avg_pe_target, double avg_pe_diff_target) {
    minimizeFire(particle, 0.1, 1e-2, avg_pe_target, avg_pe_diff_target, 100000L, 1000L);
}
void minimizeFire(Particle& particle, double avg_pe_target, double avg_pe_diff_target, long num_steps) {
    minimizeFire(particle, 0.1, 1e-2, avg_pe_target, avg_pe_diff_target, num_steps, 1000L);
}
void minimizeFire(Particle& particle) { minimizeFire(particle, 0.1, 1e-2, 100000L, 1000L); }

long minimizeFireToMaxForce(Particle& particle, double fmax_target, long num_steps,
                            double alpha_init, double dt) {
    if (!(fmax_target > 0)) throw std::invalid_argument("minimizeFireToMaxForce: target must be > 0");
    return run_fire(particle, alpha_init, dt, 0.0, 0.0, fmax_target, num_steps).steps;
}

// minimizeAdam (minimization.cu:3-33): the loop runs through dpm_minimize_adam
void minimizeAdam(Particle& particle, double alpha, double beta1, double beta2, double epsilon,
                  double avg_pe_target, double avg_pe_diff_target, long num_steps, long) {
    int64_t steps = 0;
    check(dpm_minimize_adam(particle.handle(), alpha, beta1, beta2, epsilon, avg_pe_target,
                            avg_pe_diff_target, num_steps, &steps));
    if (steps >= num_steps) std::cout << "Warning: minimization did not converge" << std::endl;
}
void minimizeAdam(Particle& particle, double alpha, double beta1, double beta2, double epsilon,
                  long num_steps, long log_every_n) {
    minimizeAdam(particle, alpha, beta1, beta2, epsilon, 1e-16, 1e-16, num_steps, log_every_n);
}
void minimizeAdam(Particle& particle) {
    minimizeAdam(particle, 1e-4, 0.9, 0.999, 1e-8, 1e-16, 1e-16, 100000L, 1000L);
}
void minimizeAdam(Particle& particle, long log_every_n) {
    minimizeAdam(particle, 1e-5, 0.9, 0.999, 1e-8, 100000L, log_every_n);
}
void minimizeAdam(Particle& particle, double avg_pe_target, double avg_pe_diff_target) {
    minimizeAdam(particle, 1e-5, 0.9, 0.999, 1e-8, avg_pe_target, avg_pe_diff_target, 100000L, 1000L);
}

// jam_adam (compression.cu:3-59)
// The compression routines log through `io` (when given) where the reference logs: each
// compression step before the box changes, and once more, forced, at the end.
static void jam_adam_impl(Particle& particle, Adam& adam, IOManager* io, long num_compression_steps,
                          long num_adam_steps, double avg_pe_target, double avg_pe_diff_target,
                          double packing_fraction_increment,
                          double min_packing_fraction_increment, double mult, long start_step) {
    check(dpm_adam_reset(particle.handle()));
    double packing_fraction = particle.getPackingFraction();
    long compression_step = start_step;
    num_compression_steps += start_step;
    const double avg_pe_past_jamming = 1e-9;
    double avg_pe = 0.0, sign = 1.0;
    long consecutive_compressions = 0;
    while (compression_step < num_compression_steps && avg_pe < avg_pe_past_jamming) {
        avg_pe = adam_relax(particle, adam, num_adam_steps, avg_pe_target, avg_pe_diff_target, nullptr);
        if (avg_pe > avg_pe_target) {
            sign = -1.0;
            if (packing_fraction_increment / 2.0 > min_packing_fraction_increment)
                packing_fraction_increment /= 2.0;
            consecutive_compressions = 0;
        } else if (avg_pe_target * mult > avg_pe && avg_pe > avg_pe_target) {
            std::cout << "jamming complete" << std::endl;
            break;
        } else {
            sign = 1.0;
            consecutive_compressions++;
            if (consecutive_compressions > 10) {
                packing_fraction_increment *= 1.01;
                consecutive_compressions = 0;
            }
        }
        if (io) io->log(compression_step);
        particle.scaleToPackingFractionFull(packing_fraction + packing_fraction_increment * sign);
        packing_fraction = particle.getPackingFraction();
        compression_step++;
    }
    if (io) io->log(compression_step, true);
}

// compress_adam (compression.cu:61-93)
static void compress_adam_impl(Particle& particle, Adam& adam, IOManager* io,
                               long num_compression_steps, long num_adam_steps,
                               double avg_pe_target, double avg_pe_diff_target,
                               double packing_fraction_increment, double max_pe_target) {
    check(dpm_adam_reset(particle.handle()));
    double packing_fraction = particle.getPackingFraction();
    long compression_step = 0;
    double avg_pe = 0.0;
    while (compression_step < num_compression_steps && avg_pe < max_pe_target) {
        long k = 0;
        avg_pe = adam_relax(particle, adam, num_adam_steps, avg_pe_target, avg_pe_diff_target, &k);
        if (k == num_adam_steps) {
            std::cout << "Adam failed to converge" << std::endl;
            break;
        }
        if (io) io->log(compression_step);
        particle.scaleToPackingFractionFull(packing_fraction + packing_fraction_increment);
        packing_fraction = particle.getPackingFraction();
        compression_step++;
    }
    if (io) io->log(compression_step, true);
}

// decompress_adam (compression.cu:95-126)
static void decompress_adam_impl(Particle& particle, Adam& adam, IOManager* io,
                                 long num_compression_steps, long num_adam_steps,
                                 double avg_pe_target, double avg_pe_diff_target,
                                 double packing_fraction_increment, double min_pe_target) {
    check(dpm_adam_reset(particle.handle()));
    double packing_fraction = particle.getPackingFraction();
    long compression_step = 0;
    while (compression_step < num_compression_steps) {
        double avg_pe =
            adam_relax(particle, adam, num_adam_steps, avg_pe_target, avg_pe_diff_target, nullptr);
        if (io) io->log(compression_step);
        particle.scaleToPackingFractionFull(packing_fraction - packing_fraction_increment);
        packing_fraction = particle.getPackingFraction();
        compression_step++;
        if (avg_pe < min_pe_target) break;
    }
    if (io) io->log(compression_step, true);
}

// compress_to_phi_adam (compression.cu:128-159)
static void compress_to_phi_adam_impl(Particle& particle, Adam& adam, IOManager* io,
                                      long num_compression_steps, long num_adam_steps,
                                      double avg_pe_target, double avg_pe_diff_target,
                                      double delta_phi_target, long start_step) {
    check(dpm_adam_reset(particle.handle()));
    double packing_fraction = particle.getPackingFraction();
    double phi_target = packing_fraction + delta_phi_target;
    double inc = delta_phi_target / (static_cast<double>(num_compression_steps) / 2.0);
    long compression_step = start_step;
    num_compression_steps += start_step;
    while (compression_step < num_compression_steps && std::abs(packing_fraction - phi_target) > inc) {
        adam_relax(particle, adam, num_adam_steps, avg_pe_target, avg_pe_diff_target, nullptr);
        double next = packing_fraction + (phi_target > packing_fraction ? inc : -inc);
        if (io) io->log(compression_step);
        particle.scaleToPackingFractionFull(next);
        packing_fraction = particle.getPackingFraction();
        compression_step++;
    }
    if (io) io->log(compression_step, true);
}

// ---- the reference's signatures (include/routines/compression.h:8-14), logging through
// io_manager, and the same routines without logging
void jam_adam(Particle& particle, Adam& adam, IOManager& io_manager, long num_compression_steps,
              long num_adam_steps, double avg_pe_target, double avg_pe_diff_target,
              double packing_fraction_increment, double min_packing_fraction_increment,
              double mult, long start_step) {
    jam_adam_impl(particle, adam, &io_manager, num_compression_steps, num_adam_steps,
                  avg_pe_target, avg_pe_diff_target, packing_fraction_increment,
                  min_packing_fraction_increment, mult, start_step);
}
void jam_adam(Particle& particle, Adam& adam, long num_compression_steps, long num_adam_steps,
              double avg_pe_target, double avg_pe_diff_target, double packing_fraction_increment,
              double min_packing_fraction_increment, double mult, long start_step) {
    jam_adam_impl(particle, adam, nullptr, num_compression_steps, num_adam_steps, avg_pe_target,
                  avg_pe_diff_target, packing_fraction_increment, min_packing_fraction_increment,
                  mult, start_step);
}
void compress_adam(Particle& particle, Adam& adam, IOManager& io_manager,
                   long num_compression_steps, long num_adam_steps, double avg_pe_target,
                   double avg_pe_diff_target, double packing_fraction_increment,
                   double max_pe_target) {
    compress_adam_impl(particle, adam, &io_manager, num_compression_steps, num_adam_steps,
                       avg_pe_target, avg_pe_diff_target, packing_fraction_increment,
                       max_pe_target);
}
void compress_adam(Particle& particle, Adam& adam, long num_compression_steps, long num_adam_steps,
                   double avg_pe_target, double avg_pe_diff_target,
                   double packing_fraction_increment, double max_pe_target) {
    compress_adam_impl(particle, adam, nullptr, num_compression_steps, num_adam_steps,
                       avg_pe_target, avg_pe_diff_target, packing_fraction_increment,
                       max_pe_target);
}
void decompress_adam(Particle& particle, Adam& adam, IOManager& io_manager,
                     long num_compression_steps, long num_adam_steps, double avg_pe_target,
                     double avg_pe_diff_target, double packing_fraction_increment,
                     double min_pe_target) {
    decompress_adam_impl(particle, adam, &io_manager, num_compression_steps, num_adam_steps,
                         avg_pe_target, avg_pe_diff_target, packing_fraction_increment,
                         min_pe_target);
}
void decompress_adam(Particle& particle, Adam& adam, long num_compression_steps,
                     long num_adam_steps, double avg_pe_target, double avg_pe_diff_target,
                     double packing_fraction_increment, double min_pe_target) {
    decompress_adam_impl(particle, adam, nullptr, num_compression_steps, num_adam_steps,
                         avg_pe_target, avg_pe_diff_target, packing_fraction_increment,
                         min_pe_target);
}
void compress_to_phi_adam(Particle& particle, Adam& adam, IOManager& io_manager,
                          long num_compression_steps, long num_adam_steps, double avg_pe_target,
                          double avg_pe_diff_target, double delta_phi_target, long start_step) {
    compress_to_phi_adam_impl(particle, adam, &io_manager, num_compression_steps,
                              num_adam_steps, avg_pe_target, avg_pe_diff_target,
                              delta_phi_target, start_step);
}
void compress_to_phi_adam(Particle& particle, Adam& adam, long num_compression_steps,
                          long num_adam_steps, double avg_pe_target, double avg_pe_diff_target,
                          double delta_phi_target, long start_step) {
    compress_to_phi_adam_impl(particle, adam, nullptr, num_compression_steps, num_adam_steps,
                              avg_pe_target, avg_pe_diff_target, delta_phi_target, start_step);
}

JamResult jamWithFire(Particle& particle, double pressure_low, double pressure_high,
                      double compression_rate, double length_tolerance,
                      long num_compression_steps, double fmax_target,
                      double min_contacts_per_particle, double decompress_factor) {
    if (!(pressure_high >= pressure_low) || !(pressure_low > 0))
        throw std::invalid_argument("jamWithFire: need 0 < pressure_low <= pressure_high");
    dpm_jam_params p;
    p.pressure_target = pressure_low;
    p.window_rel = pressure_high / pressure_low - 1.0;
    p.strain_step = 1.0 - compression_rate;
    p.decompress_factor = decompress_factor;
    p.length_tol = length_tolerance;
    p.min_contacts_per_particle = min_contacts_per_particle;
    p.max_compression_steps = num_compression_steps;
    p.fire.alpha_init = 0.1;
    p.fire.dt = 1e-2;
    p.fire.avg_pe_target = fmax_target > 0 ? 0.0 : 1e-16;
    p.fire.avg_pe_diff_target = fmax_target > 0 ? 0.0 : 1e-16;
    p.fire.fmax_target = fmax_target;
    p.fire.max_steps = 100000;
    dpm_jam_result r;
    check(dpm_jam(particle.handle(), &p, &r));
    if (r.status == 2) std::cout << "Error: couldnt find jammed state" << std::endl;
    JamResult out;
    out.status = r.status;
    out.compression_steps = r.compression_steps;
    out.fire_steps = r.fire_steps_total;
    out.pressure = r.pressure;
    out.packing_fraction = r.packing_fraction;
    out.potential_energy = r.pe;
    out.box_length = r.box_length;
    out.contacts = r.contacts;
    particle.scaleBox(1.0);  // refresh the cached box side
    return out;
}
