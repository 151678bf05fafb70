// routines.hpp -- minimisation and jamming routines (include/routines/minimization.h,
// include/routines/compression.h, scripts/runArthurJamWithFire.cpp) over the C ABI.
//
// The compression routines come with the reference's signatures (an IOManager to log
// through) and without the IOManager (no logging).
#pragma once
#include "integrator.hpp"
#include "particle.hpp"

class IOManager;

// minimizeFire overloads (src/routines/minimization.cu:75-148).  The loop runs on the
// device (dpm_minimize_fire): Fire::step and the PE/N stop test per step, velocities
// zeroed at the start and the end.
void minimizeFire(Particle& particle, double alpha_init, double dt, double avg_pe_target,
                  double avg_pe_diff_target, long num_steps, long log_every_n);
void minimizeFire(Particle& particle, double alpha_init, double dt, long num_steps, long log_every_n);
void minimizeFire(Particle& particle, long log_every_n);
void minimizeFire(Particle& particle, double avg_pe_target, double avg_pe_diff_target);
void minimizeFire(Particle& particle, double avg_pe_target, double avg_pe_diff_target, long num_steps);
void minimizeFire(Particle& particle);

// FIRE until max_i |F_i| < fmax_target (or the reference's PE criteria, or num_steps).
// Returns the number of FIRE steps taken.
long minimizeFireToMaxForce(Particle& particle, double fmax_target, long num_steps,
                            double alpha_init = 0.1, double dt = 1e-2);

// minimizeAdam overloads (src/routines/minimization.cu:3-73)
void minimizeAdam(Particle& particle, double alpha, double beta1, double beta2, double epsilon,
                  double avg_pe_target, double avg_pe_diff_target, long num_steps, long log_every_n);
void minimizeAdam(Particle& particle, double alpha, double beta1, double beta2, double epsilon,
                  long num_steps, long log_every_n);
void minimizeAdam(Particle& particle);
void minimizeAdam(Particle& particle, long log_every_n);
void minimizeAdam(Particle& particle, double avg_pe_target, double avg_pe_diff_target);

// Adam compression routines (src/routines/compression.cu:3-159), as declared by
// include/routines/compression.h:8-14
void jam_adam(Particle& particle, Adam& adam, IOManager& io_manager, long num_compression_steps,
              long num_adam_steps, double avg_pe_target, double avg_pe_diff_target,
              double packing_fraction_increment, double min_packing_fraction_increment,
              double mult, long start_step = 0);
void compress_adam(Particle& particle, Adam& adam, IOManager& io_manager,
                   long num_compression_steps, long num_adam_steps, double avg_pe_target,
                   double avg_pe_diff_target, double packing_fraction_increment,
                   double max_pe_target);
void decompress_adam(Particle& particle, Adam& adam, IOManager& io_manager,
                     long num_compression_steps, long num_adam_steps, double avg_pe_target,
                     double avg_pe_diff_target, double packing_fraction_increment,
                     double min_pe_target);
void compress_to_phi_adam(Particle& particle, Adam& adam, IOManager& io_manager,
                          long num_compression_steps, long num_adam_steps, double avg_pe_target,
                          double avg_pe_diff_target, double delta_phi_target,
                          long start_step = 0);
// the same without logging
void jam_adam(Particle& particle, Adam& adam, long num_compression_steps, long num_adam_steps,
              double avg_pe_target, double avg_pe_diff_target, double packing_fraction_increment,
              double min_packing_fraction_increment, double mult, long start_step = 0);
void compress_adam(Particle& particle, Adam& adam, long num_compression_steps, long num_adam_steps,
                   double avg_pe_target, double avg_pe_diff_target,
                   double packing_fraction_increment, double max_pe_target);
void decompress_adam(Particle& particle, Adam& adam, long num_compression_steps,
                     long num_adam_steps, double avg_pe_target, double avg_pe_diff_target,
                     double packing_fraction_increment, double min_pe_target);
void compress_to_phi_adam(Particle& particle, Adam& adam, long num_compression_steps,
                          long num_adam_steps, double avg_pe_target, double avg_pe_diff_target,
                          double delta_phi_target, long start_step = 0);

// Pressure-bisection jamming of scripts/runArthurJamWithFire.cpp:36-106 on the device:
// decompress until P < P_low, then compress by `compression_rate` steps with FIRE
// minimisation, bisecting on the box side inside [P_low, P_high] until the pressure lies
// in the window with more than min_contacts_per_particle * N contacts.
struct JamResult {
    int status;  // 0 jammed in the window, 1 out of compression steps, 2 the box-length
                 // bracket closed below length_tolerance ("couldnt find jammed state")
    long compression_steps;
    long fire_steps;
    double pressure, packing_fraction, potential_energy, box_length;
    long contacts;
};
JamResult jamWithFire(Particle& particle, double pressure_low, double pressure_high,
                      double compression_rate, double length_tolerance,
                      long num_compression_steps, double fmax_target = 0.0,
                      double min_contacts_per_particle = 2.0, double decompress_factor = 1.01);
