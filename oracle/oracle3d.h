/* Oracle (test infrastructure only): FP64 CPU restatement of the reference
 * 3D solver kernels (iceflow3d/{lattice,home,phase,gradients,simulator,
 * thermal}.py).  Only tests/, the smoke check and the CPU-baseline leg of
 * bench.py may load it.  Logical layouts follow the reference's to_numpy():
 * scalar fields [N], vectors [N][3], HOME moments [N][10], D3Q7 [N][7],
 * capillary stress [N][9]; N index = (i*ny + j)*nz + k. */
#ifndef ICEFLOW_ORACLE3D_H
#define ICEFLOW_ORACLE3D_H

#include <stdint.h>

typedef struct o3d {
  int nx, ny, nz, boundary, bodies, mx, my, mz;
  int parity, check_stability, water_width, water_height, floor_depth;
  double rho_air, gravity_y, sigma, width, mobility, normal_eps, bulk_rate;
  double threshold, max_velocity, expansion, water_t0, air_t, flux_limit;
  double rho_water_si, rho_air_si, mu_air, nu_scale;
  double tm, rho_cp, cp_ice, k_water, k_ice, latent, cell_volume, dx;
  double velocity_m_s, initial_cell_mass, mass_kg, rho_ice;
  /* grid fields */
  double *home, *home_post, *h;
  double *phi, *pressure, *old_pressure, *coverage;
  double *u, *old_u, *force, *accel, *grad, *flux, *stress;
  int32_t *id, *prev_id, *cov_id;
  int8_t *flag;
  double *refill_state, *refill_next;
  int8_t *refill_valid, *refill_next_valid;
  double *ref_density, *ref_pressure;
  /* bodies */
  double *body_origin, *body_rot, *body_center, *body_v, *body_w;
  double *hydro_force, *hydro_torque;
  int32_t *material_size;
  double *body_t0;
  double *material_mass, *material_energy, *material_heat;
  double *water_volume, *water_energy, *target_energy, *sampled_ice_t;
  double *energy_flux, *volume_flux, *reduced;
  /* scalars */
  int32_t invalid_index, thermal_invalid, openings, covered, remaining;
  int32_t refill_lower[3], refill_upper[3], geo_lower[3], geo_upper[3];
  double maximum_speed, peak_speed, phase_clip_volume, remap_energy;
  double discarded_heat, max_outflow;
} o3d;

#endif
