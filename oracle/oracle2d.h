/* Oracle (test infrastructure only): state of the FP64 restatement of the
 * reference 2D coupled step (iceflow2d/simulator.py).  Fields the reference
 * stores in FP32 are float here too (values are rounded where the
 * reference rounds them); arithmetic is double.  Arrays are allocated by
 * the Python side (numpy) and indexed n = i * ny + j (j fastest), q-plane
 * major for populations: f[q * N + n]. */
#ifndef ORACLE2D_H
#define ORACLE2D_H

#include <stdint.h>

typedef struct o2d {
  int32_t nx, ny, boundary, body_nx, body_ny, left_kind, right_kind, bottom_kind, top_kind, advection;
  /* lattice-unit constants */
  double rho_w, rho_a, nu_w, nu_a, gx, gy, sigma, width, mobility, smagorinsky, interface_tau;
  double cutoff, threshold, dx, expansion, t_ref, linear_damping, angular_damping;
  double initial_mass, initial_inertia;
  /* thermal constants (SI) */
  double t_melt, cp_w, cp_i, k_w, k_i, latent, rho_w_si, area, velocity_scale, cell_mass;
  double t_air, left_value, right_value, bottom_value, top_value;
  /* grid fields */
  float *f, *fbuf, *h, *hbuf;          /* [9][N] */
  float *phi, *u, *a, *p;              /* u, a: [2][N] */
  float *p_ref, *rho_ref;              /* [ny] */
  int8_t *solid, *prev_solid;          /* [N] */
  float* coverage;                     /* world_body_solid_fraction [N] */
  /* thermal fields (FP64) */
  double *body_mass, *body_energy, *body_heat, *melt_mass_step, *melt_sensible_step; /* [body N] */
  int32_t* melt_target;                                                               /* [body N] */
  double *water_v, *water_e, *target_e;                                               /* [N] */
  double *flux_ex, *flux_vx;            /* [(nx+1)][ny] */
  double *flux_ey, *flux_vy;            /* [nx][ny+1] */
  int8_t* eligible;                     /* [N] */
  /* rigid body */
  float velocity[2], angle, omega, origin[2], support[4];
  int32_t projected;
  double mass, inertia, local_com[2], impulse[2], torque;
  /* scalars */
  double initial_water_volume, water_volume_current, projection_derivative;
  double cumulative_melted_momentum[2], cumulative_melted_angular;
  double carrier_momentum[2], correction_momentum[2], carrier_angular, correction_angular;
  double fluid_before[2], fluid_provisional[2], fluid_after[2], body_before[2];
  double angular_before, angular_provisional, angular_after, body_angular_before;
  double wet_weight, weight_first[2], weight_second;
  int32_t wet_target, wet_target_max;
  double volume_before, energy_before, base_capacity, full_capacity, specific_min, specific_max;
  double reconstructed_energy, cooling_capacity, heating_capacity, correction_abs;
  double volume_after, energy_after;
  double boundary_power, boundary_heat_input;
  double unassigned_mass, unassigned_energy, fallback_free, fallback_wet;
  double interval_body_melt, interval_water_melt;
  double reduced_mass, reduced_first[2], reduced_inertia;
  float max_outflow;
} o2d;

#endif
