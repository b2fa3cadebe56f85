/*
 * CPU oracle for the explicit MLS-MPM/APIC substep of warpmpm (kks32/mpm-engine).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, the smoke
 * check and the bench's cpu_baseline leg may load it. The product path (the HIP
 * library under mpm-engine_amd/csrc) never links, loads or calls it.
 *
 * Layout mirrors the reference's MPMStateStruct (src/warpmpm/kernels/warp_utils.py:87-157):
 * particle vec3 arrays are AoS float[3N], mat33 arrays are row-major float[9N], the grid
 * is dense with z fastest: node (i,j,k) -> (i*ny + j)*nz + k, grid_v_* are float[3*nodes].
 * All arithmetic is fp32, single-threaded, in particle / node order (the order Warp's
 * CPU backend runs a launch), so every function here is deterministic.
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H
#include <stdint.h>

#define O_MAX_MATERIALS 16

typedef struct OModel {
    float dx, inv_dx;
    int32_t nx, ny, nz;
    float gravity[3];
    float rpic_damping;
    float grid_v_damping_scale;
    int32_t update_cov_with_F;
    int32_t periodic_x;
    /* per-type arrays [O_MAX_MATERIALS] (mpm_solver_warp.py:238-274) */
    const float *E, *nu, *bulk, *hardening, *xi, *plastic_viscosity, *softening, *alpha;
    const float *muI_mu_s, *muI_delta_mu, *muI_I0, *muI_d, *muI_rho_s, *muI_phi_init,
        *muI_phi_chi;
    const float *eta_table;
    float eta_table_smin, eta_table_smax;
    int32_t eta_table_n;
    /* per-particle model arrays [N] (mpm_solver_warp.py:277-279), mutable by return maps */
    float *mu, *lam, *yield_stress;
} OModel;

/* CPIC thin-boundary (CDF) state (warp_utils.py:9-20 bit layout; mpm_solver_warp.py:389-436
 * allocation). Node tag: bit 2l = lane l valid, bit 2l+1 = lane l side, bits [8, 10) = owner
 * lane. Tag / distance grids are dense z-fastest like grid_m; block = 4^3 occupancy mask. */
#define O_MAX_CDF 4
#define O_CDF_OWNER_SHIFT 8
typedef struct OCdf {
    int32_t n_cdf, mask_off;
    int32_t nb[3];
    int32_t *tag, *tag_prev;
    float *d, *d_prev;
    int32_t *block;
    int32_t *ptag; /* particle_cdf_tag [N] */
    float lane_center[O_MAX_CDF][3], lane_velocity[O_MAX_CDF][3], lane_omega[O_MAX_CDF][3];
    float lane_center_prev[O_MAX_CDF][3], lane_velocity_prev[O_MAX_CDF][3], lane_omega_prev[O_MAX_CDF][3];
    float lane_friction[O_MAX_CDF];
    int32_t lane_type[O_MAX_CDF];
    float reaction_force[O_MAX_CDF][3], reaction_torque[O_MAX_CDF][3];
} OCdf;

/* one CDF collider's stamp parameters (CDFCollider, mpm_solver_warp.py:2859-2880) */
typedef struct OCdfCollider {
    const float *cdf_val, *cdf_valid; /* [res^3] */
    int32_t res;
    float origin[3];
    float cell;
    float center[3];
    float quat[4]; /* x y z w */
    float band;
    int32_t lane;
    float start_time, end_time;
} OCdfCollider;

typedef struct OState {
    int32_t n;
    float *x, *v, *F, *F_trial, *C, *stress, *L, *R;
    float *vol, *mass, *density, *Jp;
    float *cov, *init_cov;
    int32_t *selection, *material;
    float *grid_m, *grid_v_in, *grid_v_out;
    OCdf *cdf; /* NULL: no CDF colliders (n_cdf == 0) */
    int32_t *rigid_id; /* particle_rigid_id, or NULL without rigid bodies */
    float *x_ref;      /* particle_x_ref (n x 3), or NULL */
} OState;

/* collider parameter blocks (warp_utils.py collider structs; mpm_solver_warp.py:1880-2777) */
typedef struct OPlane {
    float start_time, end_time;
    float point[3], normal[3];
    int32_t surface_type;
    float friction;
} OPlane;

typedef struct OBox {
    float start_time, end_time;
    float point[3], size[3], velocity[3];
    int32_t reset;
    float *force; /* vec3 accumulator */
} OBox;

typedef struct OCup {
    float start_time, end_time;
    float point[3];
    float rot[9];
    float velocity[3], omega[3];
    float friction, outer_radius, inner_radius, half_height, inner_floor_z, fillet_radius;
    float sticky_depth, contact_band;
    float *force, *torque;
} OCup;

typedef struct OWalls {
    float start_time, end_time;
} OWalls;

typedef struct OSdf {
    const float *sdf_val;  /* [res^3] */
    const float *sdf_grad; /* [res^3 * 3] */
    int32_t res;
    float origin[3];
    float cell;
    float center[3];
    float quat[4]; /* x y z w */
    float velocity[3], omega[3];
    float band;
    int32_t surface_type;
    float friction;
    float start_time, end_time;
    float *force, *torque;
} OSdf;

typedef struct OPointCloud {
    const int32_t *occupancy; /* [nx*ny*nz] */
    float start_time, end_time;
} OPointCloud;

#endif
