/*
 * mpm_hip.h - C ABI of libmpm_hip.so, the gfx950 implementation of warpmpm's explicit
 * MLS-MPM/APIC substep.
 *
 * The reference drives its Warp kernels from Python (MPM_Simulator_WARP in
 * src/warpmpm/kernels/mpm_solver_warp.py); every kernel takes the MPMStateStruct /
 * MPMModelStruct by value (src/warpmpm/kernels/warp_utils.py:23-157). This header is the
 * replacement boundary: plain structs of device pointers and sizes, no C++ or torch types.
 * The caller (the Python Solver of this repository, or a ctypes binding a maintainer adds
 * to the reference) owns every allocation; the library never allocates persistent state.
 * Every function returns an mpm_status (0 = ok) and enqueues work on `stream` (a
 * hipStream_t passed as void*; NULL = the legacy default stream).
 *
 * Batched scenes (no counterpart in the reference, which runs one scene per solver):
 * n_scenes independent grids of the same shape live back to back in node_in / node_out
 * and one particle array holds every scene's particles, scene[p] naming the owner. One
 * launch per phase advances all scenes: block codes in block lists are global
 * (scene * nb0*nb1*nb2 + local code), collider schedules hold ncoll rows per scene
 * ([substeps][n_scenes][ncoll]), box arguments (zero box, launch box) apply to every
 * scene, and the block sort keys are scene-major so each scene stays contiguous.
 *
 * Memory layout (device):
 *  - particles are SoA planes: x[c*stride + p]; mat33 fields hold 9 planes, component
 *    (r, c) in plane 3*r + c (row-major like Warp's mat33).
 *  - the grid is dense but blocked: 4x4x4-node blocks, block (bi, bj, bk) at
 *    ((bi*nb[1] + bj)*nb[2] + bk), node (i, j, k) at block*64 + (i&3)*16 + (j&3)*4 + (k&3).
 *    node_in holds (momentum x, y, z, mass) and node_out (velocity x, y, z, 0) as float4.
 */
#ifndef MPM_HIP_H
#define MPM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_MAX_MATERIALS 16
#define MPM_ABI_VERSION 10

typedef enum mpm_status {
    MPM_OK = 0,
    MPM_ERR_ARG = 1,    /* bad argument (null pointer, size mismatch) */
    MPM_ERR_HIP = 2,    /* a HIP runtime call failed */
    MPM_ERR_SIZE = 3,   /* size beyond a documented maximum */
} mpm_status;

/* particle-pass mode bits (mpm_particle_pass) */
#define MPM_PASS_G2P 1          /* gather: g2p_particle, mpm_utils.py:944-1043 */
#define MPM_PASS_STRESS 2       /* stress_update_particle, mpm_utils.py:1056-1149 */
#define MPM_PASS_P2G 4          /* scatter: p2g_particle, mpm_utils.py:819-916 */
#define MPM_PASS_LEAN 8         /* do not store v, C, L, F_trial (dead between fused substeps) */
#define MPM_PASS_KEEP_STRESS 16 /* with LEAN: still store the stress */
#define MPM_PASS_VISCOUS_ONLY 32 /* every particle is a viscous / rigid id (6, 7, 8, 10, 12):
                                    launch the specialisation without the SVD paths */
/* launch shape (results do not depend on it; an autotuner picks it per workload) */
#define MPM_PASS_WG128 64       /* 128 threads per workgroup instead of 256 */
#define MPM_PASS_PPT2 128       /* each thread takes 2 particles in turn */
#define MPM_PASS_PPT4 256       /* each thread takes 4 particles in turn */
/* deterministic scatter: every contribution is rounded to 64-bit fixed point (node_fix,
   scales fix_scale) and accumulated with integer atomics, so the grid sums do not depend
   on the order atomics land in: runs are bit-reproducible, and identical across launch
   shapes and particle orders. Needs G->node_fix; no CDF colliders. */
#define MPM_PASS_DETERMINISTIC 512
/* paired interior pass: with MPM_PASS_DETERMINISTIC and the lean fused mode (G2P | STRESS |
   P2G | LEAN, optionally KEEP_STRESS) only, P->chunk_start holding chunks of at most 512
   particles (mpm_sort_chunks with chunk = 512) and 256-thread workgroups (no WG128 / PPT
   bits): each lane takes two particles of its chunk, sharing one stencil tile, one scatter
   arena and one flush. Bit-identical to the single-particle pass. */
#define MPM_PASS_PAIR 1024

/* grid-update flag bits (mpm_grid_update) */
#define MPM_GRID_ZERO_IN 1      /* clear node_in after it is consumed */
#define MPM_GRID_FIXED_IN 2     /* take (momentum, mass) from node_fix (deterministic scatter)
                                   instead of node_in, and clear node_fix after use */
#define MPM_GRID_WRENCH_DEFER 4 /* with MPM_GRID_FIXED_IN: add this call's integer wrench into
                                   the workspace rows and leave them there; a later
                                   mpm_grid_wrench_flush converts and clears them (one flush
                                   per tick instead of one per substep; integer sums, so the
                                   result does not depend on where the flush falls) */

/* collider kinds (the reference's grid_postprocess plugins) */
#define MPM_COLLIDER_PLANE 1      /* mpm_solver_warp.py:1929-1992 */
#define MPM_COLLIDER_BOX 2        /* mpm_solver_warp.py:2053-2088 */
#define MPM_COLLIDER_CUP 3        /* revolved SDF glass, mpm_solver_warp.py:2161-2225 */
#define MPM_COLLIDER_WALLS 4      /* mpm_solver_warp.py:2294-2359 */
#define MPM_COLLIDER_SDF 5        /* mesh SDF, mpm_solver_warp.py:2680-2736 */
#define MPM_COLLIDER_POINTCLOUD 6 /* mpm_solver_warp.py:2586-2605 */

/* Particle state (MPMStateStruct particle_* arrays, warp_utils.py:87-123, plus the
 * per-particle model arrays mu / lam / yield_stress, warp_utils.py:31-40). */
typedef struct mpm_particles {
    int32_t n;       /* particle count */
    int32_t stride;  /* plane stride in elements, >= n */
    float *x, *v;                          /* 3 planes */
    float *F, *F_trial, *C, *stress, *L;   /* 9 planes each */
    float *vol, *mass, *mu, *lam, *yield_stress, *Jp;
    int32_t *material, *selection;
    float *cov; /* 6 planes (xx, xy, xz, yy, yz, zz) or NULL when covariance is off */
    int32_t *scene; /* scene id per particle for batched scenes, or NULL (all scene 0) */
    int32_t *cdf_tag; /* CPIC particle tag (particle_cdf_tag, warp_utils.py:87-157), required
                         when the grid carries CDF colliders, else NULL */
    int32_t *rigid_id; /* rigid body per particle (particle_rigid_id; meaningful for material
                          8), or NULL when the scene has no rigid bodies */
    float *x_ref;      /* 3 planes: body-frame reference position of rigid particles
                          (particle_x_ref), or NULL */
    int32_t *modifier_mask; /* bit k: particle selected by particle modifier k, or NULL */
    const int32_t *chunk_start; /* device, *chunk_count + 1 entries (mpm_sort_chunks): chunk c of
                                   the particle pass is particles [chunk_start[c], chunk_start[c + 1]),
                                   at most 256 of them; NULL: chunks of 256 consecutive particles
                                   (also with MPM_PASS_WG128, which uses 128) */
    const int32_t *chunk_count; /* device int: chunks in chunk_start */
    int32_t chunk_capacity;     /* chunks one sweep of the launch covers (with chunk_start): a
                                   size hint, more chunks are taken by further sweeps */
    int32_t pad_;
} mpm_particles_t;

/* ---- CPIC thin-boundary (CDF) colliders ---------------------------------------------
 * Open oriented surfaces that sever particle-node transfers across themselves (the
 * reference's add_cdf_collider, mpm_solver_warp.py:2819-3155; transfer masking and ghost
 * fill, mpm_utils.py:648-1017). Node tag (warp_utils.py:9-20): bit 2l = lane l valid,
 * bit 2l+1 = lane l side, bits [8, 10) = owner lane (nearest surface). Tag and owner
 * distance grids use the blocked node layout of node_in; the occupancy mask holds one int
 * per 4x4x4 block (the grid block), OR of tag | tag_prev over its nodes. The gather half
 * of a pass reads tag_prev / dist_prev and the lanes' *_prev pose, the scatter half tag /
 * dist and the current pose: the fused pass's gather uses the grid state of the previous
 * substep, its scatter builds the next one. Single scene only. */
#define MPM_MAX_CDF 4
#define MPM_CDF_OWNER_SHIFT 8

typedef struct mpm_cdf_lanes {
    float center[MPM_MAX_CDF][3], velocity[MPM_MAX_CDF][3], omega[MPM_MAX_CDF][3];
    float center_prev[MPM_MAX_CDF][3], velocity_prev[MPM_MAX_CDF][3], omega_prev[MPM_MAX_CDF][3];
    float friction[MPM_MAX_CDF];
    int32_t type[MPM_MAX_CDF]; /* 0 sticky, 1 slip, 2 separable + Coulomb */
} mpm_cdf_lanes_t;

typedef struct mpm_cdf {
    int32_t n_cdf;        /* lanes in use; 0 = no CDF colliders (other fields unused) */
    int32_t mask_off;     /* 1 = vote for every particle (no block-occupancy pretest) */
    int32_t *tag, *tag_prev;   /* int per node */
    float *dist, *dist_prev;   /* owner distance per node (1e9 where untagged) */
    int32_t *block;            /* int per grid block */
    mpm_cdf_lanes_t *lanes;    /* device */
    float *reaction;           /* device [MPM_MAX_CDF][6]: blocked-deposit force xyz, torque xyz */
} mpm_cdf_t;

/* One collider's stamp (the reference's CDFCollider at its live pose). */
typedef struct mpm_cdf_collider {
    const float *val, *valid;  /* device [res^3], z fastest: signed distance, validity */
    int32_t res, lane;
    float origin[3], cell;     /* body-frame voxel grid */
    float center[3];
    float quat[4];             /* x, y, z, w */
    float band;
    float time, start_time, end_time; /* fp32 simulation time of the stamp; active window */
} mpm_cdf_collider_t;

typedef struct mpm_grid {
    float *node_in;       /* float4 per node: momentum xyz, mass */
    float *node_out;      /* float4 per node: velocity xyz, 0 */
    int32_t n[3];         /* logical node counts */
    int32_t nb[3];        /* blocks per axis, ceil(n / 4) */
    float dx, inv_dx;     /* fp32 like the reference's MPMModelStruct.dx / inv_dx */
    int32_t *error_flag;  /* device int, OR-ed bits: 1 = a stencil left the grid,
                             2 = a non-finite particle position, 4 = a solid material
                             met a MPM_PASS_VISCOUS_ONLY pass, 8 = a scene id outside
                             [0, n_scenes) (that particle is skipped), 64 = a fixed-point
                             contribution beyond 2^51 units (deterministic scatter: the
                             scale is too fine for this state),
                             512 = a scatter into a block without a pool slot (that
                             contribution was dropped) */
    int32_t n_scenes;     /* independent scenes sharing the buffers (>= 1): scene s owns
                             blocks [s * B, (s + 1) * B), B = nb[0] * nb[1] * nb[2];
                             block codes in every block list are these global codes */
    int32_t pad_;
    mpm_cdf_t cdf;        /* CPIC colliders (cdf.n_cdf == 0: none) */
    int64_t *node_fix;    /* deterministic scatter: 4 int64 per node (momentum xyz, mass) in
                             fixed point, zero between uses; NULL otherwise */
    float fix_scale[2];   /* units per kg (mass) and per kg m/s (momentum), powers of two */
    int32_t *block_touch; /* per grid block (global block codes), or NULL: bit 0 = a particle
                             pass scattered into the block since the last grid update (the
                             pass stores 3), bit 1 = the last grid update wrote velocities
                             into it. The grid update skips blocks with neither bit (their
                             accumulators and node_out are zero), only zeroes node_out of
                             blocks at 2, and leaves processed blocks at 2. NULL: every
                             listed block is processed. Zero-initialised with node_out. */
    const int32_t *block_slot; /* block pool, or NULL (dense storage: block code c holds nodes
                             [64 c, 64 c + 64)). Per global block code, the pool slot whose
                             64 nodes hold that block in node_in / node_out / node_fix
                             (node = 64 slot + cell), or -1: no storage this tick, gathers
                             read zero velocity there and a scatter there sets error bit 512.
                             Built by mpm_pool_assign from the active-block flags. Of the
                             entry points that address node storage, the particle pass
                             (without CPIC colliders), the grid update, mpm_grid_zero_blocks,
                             mpm_grid_wrench_flush and mpm_grid_to_dense take a pool;
                             mpm_grid_zero_box, mpm_blocks_from_box and the CPIC, rigid-body
                             and adjoint entry points return MPM_ERR_ARG for one. */
    int32_t pool_blocks;  /* slots in the node arrays when block_slot is set (slots >= it
                             are treated as -1) */
    int32_t pad2_;
} mpm_grid_t;

/* Model parameters (MPMModelStruct, warp_utils.py:23-84). Lives in device memory; the
 * per-type arrays are indexed by material id, exactly as in the reference. */
typedef struct mpm_model {
    float gravity[3];
    float rpic_damping;
    float grid_v_damping_scale;
    int32_t update_cov_with_F;
    int32_t periodic_x;
    int32_t eta_table_n;
    float eta_table_smin, eta_table_smax;
    const float *eta_table;   /* device, eta_table_n samples */
    float E[MPM_MAX_MATERIALS], nu[MPM_MAX_MATERIALS], bulk[MPM_MAX_MATERIALS];
    float hardening[MPM_MAX_MATERIALS], xi[MPM_MAX_MATERIALS];
    float plastic_viscosity[MPM_MAX_MATERIALS], softening[MPM_MAX_MATERIALS];
    float alpha[MPM_MAX_MATERIALS];
    float muI_mu_s[MPM_MAX_MATERIALS], muI_delta_mu[MPM_MAX_MATERIALS];
    float muI_I0[MPM_MAX_MATERIALS], muI_d[MPM_MAX_MATERIALS], muI_rho_s[MPM_MAX_MATERIALS];
    float muI_phi_init[MPM_MAX_MATERIALS], muI_phi_chi[MPM_MAX_MATERIALS];
} mpm_model_t;

/* One collider at one substep (the reference's per-collider param struct plus the pose
 * its host modify_bc callable integrated to, mpm_solver_warp.py:1283-1336). A tick's
 * schedule is a device table [substeps][ncoll] the host fills before the tick. */
typedef struct mpm_collider {
    int32_t kind;        /* MPM_COLLIDER_* */
    int32_t surface;     /* plane 0 sticky / 1 slip / 2 separable / 11 cut; sdf 0 / 1 / 2 */
    int32_t reset;       /* box reset flag (mpm_solver_warp.py:2086-2088) */
    int32_t slot;        /* wrench slot (force at 6*slot, torque at 6*slot + 3); -1 none */
    int32_t lo[3], hi[3];/* node box the collider acts on (its restricted launch box) */
    int32_t active;      /* 0: launch skipped this substep (empty box) */
    int32_t res;         /* sdf voxel resolution */
    float time, dt;      /* fp32 simulation time of this substep, substep dt */
    float start_time, end_time;
    float point[3], normal[3], size[3], velocity[3], omega[3];
    float rot[9];        /* cup rotation, row-major */
    float quat[4];       /* sdf orientation (x, y, z, w) */
    float friction, sticky_depth, contact_band, band;
    float outer_radius, inner_radius, half_height, inner_floor_z, fillet_radius;
    float origin[3], cell;
    const float *sdf_val;      /* device [res^3] */
    const float *sdf_grad;     /* device [res^3][3] */
    const int32_t *occupancy;  /* device dense [n0*n1*n2], z fastest */
} mpm_collider_t;

/* Device-resident per-tick control read by the grid kernels, so a captured hipGraph
 * replays with this tick's launch box without recapture. */
typedef struct mpm_tick_ctl {
    int32_t box_lo[3], box_hi[3]; /* node box of normalisation and BCs (the particle box) */
    int32_t nblocks;              /* entries valid in the block list */
    int32_t pad;
} mpm_tick_ctl_t;

/* ---- rigid bodies (material 8) -------------------------------------------------------
 * Per-body state as finalize_rigid_bodies leaves it (mpm_solver_warp.py:834-873):
 * centre of mass, velocity, angular velocity, orientation (row-major), mass and the
 * inverse body-frame inertia. A restitution plane is a surface collider registered with
 * restitution > 0 (mpm_solver_warp.py:1915-1925): point, unit normal, restitution,
 * Coulomb friction; its time window is evaluated by the caller (the reference compares
 * float64 host times) and passed per substep as a bit mask. At most 32 planes. */
typedef struct mpm_rigid_body {
    float x_cm[3], v_cm[3], omega[3];
    float R[9];
    float mass;
    float inv_inertia[9];
    float pad_[4];
} mpm_rigid_body_t;

typedef struct mpm_rigid_plane {
    float point[3], normal[3];
    float restitution, friction, start_time, end_time; /* window kept for the record */
} mpm_rigid_plane_t;

/* Device bytes of the workspace mpm_rigid_step needs for n_bodies and n_planes (no
 * counterpart in the reference, which allocates its body sums itself). */
int64_t mpm_rigid_workspace_bytes(int32_t n_bodies, int32_t n_planes);
/* The rigid step that follows the G2P pass of every split substep (_p2g2p_tail,
 * mpm_solver_warp.py:1338-1367): gather the grid velocity (node_out) at every rigid
 * particle into per-body linear and angular momentum (rigid_g2p_accumulate,
 * mpm_utils.py:1369-1418), integrate each body (rigid_body_integrate, :1421-1466: v = P/M,
 * omega = R I^-1 R^T L, x += v dt, R += skew(omega) R dt re-orthonormalised by its polar
 * factor), apply the restitution planes in order (_apply_rigid_restitution,
 * mpm_solver_warp.py:874-962, there on the host: push the body out by the deepest
 * penetration, then the normal impulse -(1 + e) v_n / (1/M + ((I^-1 (r x n)) x r) . n) at
 * the deepest particle, the lowest index on ties, and the Coulomb-capped friction impulse),
 * then set x, v, C = skew(omega) and F = I of every rigid particle (rigid_particle_update,
 * :1469-1493). `active` is a device int32: bit k set = plane k active in this substep
 * (read when the kernels run, so a captured graph replays with the live value); bodies
 * and planes are device arrays; single scene. Body momentum sums are float atomics (as
 * in the reference). */
int32_t mpm_rigid_step(const mpm_particles_t *P, const mpm_grid_t *G, mpm_rigid_body_t *bodies, int32_t n_bodies,
                       const mpm_rigid_plane_t *planes, int32_t n_planes, const int32_t *active, float dt,
                       void *workspace, void *stream);

/* ---- particle modifiers ---------------------------------------------------------------
 * Pre-P2G impulses and particle velocity modifiers (add_impulse_on_particles,
 * enforce_particle_velocity_translation / _rotation, mpm_solver_warp.py:2385-2539). Each
 * selects its particles once, when it is added (a box |x - point| < size per axis, or a
 * cylinder about `vec` through `point`: |axial offset| < half_height and radial distance
 * < radius; selection kernels mpm_utils.py:1320-1363), into bit `bit` of
 * P->modifier_mask, and acts in the substeps whose start time t has
 * start_time <= t < end_time (evaluated by the caller in float64, as the reference's host
 * loop does, and passed per substep as a bit mask over the modifiers):
 *   kind 0 impulse:      v += vec / mass * dt          (vec = force)
 *   kind 1 translation:  v = vec                       (vec = velocity)
 *   kind 2 rotation:     v = -d sin(theta) s_r axis1 + d cos(theta) s_r axis2 + s_t vec
 *                        (vec = unit axis, d = radial distance, theta the angle from axis1,
 *                        signed by axis2; s_r, s_t the rotation / translation scales) */
#define MPM_MAX_MODIFIERS 32
typedef struct mpm_particle_modifier {
    int32_t kind, bit;
    float start_time, end_time;
    float point[3], vec[3];
    float axis1[3], axis2[3];
    float size[3];
    float half_height, radius;
    float rotation_scale, translation_scale;
    float pad_[3];
} mpm_particle_modifier_t;

/* Set (selected) or clear bit mod->bit of P->modifier_mask for every particle; `mod` is a
 * host struct. Replaces the selection kernels of add_impulse_on_particles and
 * enforce_particle_velocity_* (mpm_utils.py:1320-1363, mpm_solver_warp.py:2385-2539). */
int32_t mpm_modifier_select(const mpm_particles_t *P, const mpm_particle_modifier_t *mod, void *stream);
/* Apply n_mods modifiers (device array, in the order added) to the selected particles:
 * every impulse first, then the velocity modifiers, as the reference's split substep
 * launches its pre_p2g_operations and then its particle_velocity_modifiers
 * (mpm_solver_warp.py:1181-1196); `active` is a device int32, bit j set = modifier j
 * active in this substep. */
int32_t mpm_modifiers_apply(const mpm_particles_t *P, const mpm_particle_modifier_t *mods, int32_t n_mods,
                            const int32_t *active, float dt, void *stream);

/* ---- introspection (no GPU work) ------------------------------------------------- */
int32_t mpm_abi_version(void);
const char *mpm_status_string(int32_t status);
/* sizeof of the ABI structs, so bindings can verify their mirrors */
int32_t mpm_struct_size(int32_t which); /* 0 particles, 1 grid, 2 model, 3 collider, 4 tick_ctl,
                                           5 cdf lanes, 6 cdf collider, 7 adjoint,
                                           8 rigid body, 9 rigid plane,
                                           10 particle modifier */

/* ---- particle pass --------------------------------------------------------------- */
/* One pass over all particles with selection == 0 doing the stages in `mode`:
 * G2P|STRESS|P2G is the fused interior substep g2p_stress_p2g (mpm_utils.py:1161-1178);
 * STRESS|P2G is the tick prologue (compute_stress_from_F_trial + p2g_apic_with_stress,
 * mpm_utils.py:919-923, 1152-1158); G2P alone is the tick epilogue (g2p, :1046-1052).
 * Scatter accumulates into node_in with float atomics; a stencil node outside the grid
 * is skipped and sets error bit 1. `model` is a device pointer.
 * With G->cdf.n_cdf > 0 (single scene, P->cdf_tag set, no MPM_PASS_LEAN: the ghost
 * velocity needs the stored particle velocity) the gather reconstructs each particle's
 * tag from tag_prev with persistence against cdf_tag, stores it, and substitutes the
 * ghost velocity at incompatible nodes (mpm_utils.py:971-1017); the scatter
 * reconstructs the tag from tag and sends deposits to incompatible nodes into
 * cdf.reaction instead of node_in (mpm_utils.py:851-912). */
int32_t mpm_particle_pass(const mpm_particles_t *P, const mpm_grid_t *G, const mpm_model_t *model,
                          float dt, int32_t mode, void *stream);

/* ---- CDF stamps (mpm_solver_warp.py:94-183; boxes are node boxes [lo, hi)) -------- */
/* tag = 0, dist = 1e9 over the box (cdf_zero_tags) */
int32_t mpm_cdf_zero_tags(const mpm_grid_t *G, const int32_t *lo, const int32_t *hi, void *stream);
/* tag_prev = tag, dist_prev = dist over the box (cdf_copy_prev) */
int32_t mpm_cdf_copy_prev(const mpm_grid_t *G, const int32_t *lo, const int32_t *hi, void *stream);
/* stamp one collider's bits over the box (cdf_stamp): a node is tagged when all eight
 * trilinear corners of its body-frame voxel are valid (>= 0.999) and |d| <= band; the
 * nearest lane owns it. Colliders stamp one after the other on one stream. */
int32_t mpm_cdf_stamp(const mpm_grid_t *G, const mpm_cdf_collider_t *collider, const int32_t *lo,
                      const int32_t *hi, void *stream);
/* rebuild the occupancy mask over the block box [blo, bhi) (cdf_block_mask) */
int32_t mpm_cdf_block_mask(const mpm_grid_t *G, const int32_t *blo, const int32_t *bhi, void *stream);
/* write `lanes` (host struct, copied by value into the launch) to G->cdf.lanes in stream
 * order (the reference's _sync_cdf_lane / _cdf_copy_lane_prev uploads) */
int32_t mpm_cdf_set_lanes(const mpm_grid_t *G, const mpm_cdf_lanes_t *lanes, void *stream);

/* ---- grid -------------------------------------------------------------------------- */
/* Clear node_in (which & 1) and / or node_out (which & 2) over a node box [lo, hi).
 * Replaces zero_grid / zero_grid_m_vin / zero_grid_vout (mpm_utils.py:590-598,1181-1202). */
int32_t mpm_grid_zero_box(const mpm_grid_t *G, const int32_t *lo, const int32_t *hi,
                          int32_t which, void *stream);

/* Same over a device block list (zero_grid_blocks, mpm_utils.py:1556-1564). */
int32_t mpm_grid_zero_blocks(const mpm_grid_t *G, const int32_t *block_list,
                             const int32_t *block_count, int32_t capacity, int32_t which,
                             void *stream);

/* Build the list of blocks intersecting ctl->box (ctl is device memory) in every scene
 * into block_list (capacity nb0*nb1*nb2*n_scenes) and set ctl->nblocks: the dense
 * path's launch box (Solver._update_grid_box, solver.py:485-518) as a block list. */
int32_t mpm_blocks_from_box(const mpm_grid_t *G, mpm_tick_ctl_t *ctl, int32_t *block_list,
                            void *stream);

/* Grid update of one substep over the listed blocks, restricted to ctl->box:
 * normalisation + gravity (grid_normalization_and_gravity, mpm_utils.py:927-941),
 * damping (add_damping_via_grid, :1287-1295), then every collider of `colliders`
 * (ncoll entries, device) in order, each acting inside its own box on nodes with mass
 * (the per-collider grid_postprocess launches of _apply_grid_bc, mpm_solver_warp.py:
 * 1283-1336). Reaction impulse / torque of the collider with wrench slot k is added to
 * wrench[6k .. 6k+5]. Up to 64 slots this goes through `workspace` (device floats, at
 * least mpm_grid_workspace_floats(capacity, wrench_slots)) with a fixed-order reduction,
 * so the readout is bit-reproducible; with more slots (batched scenes) per-wavefront sums
 * are added with float atomics and `workspace` may be NULL. With MPM_GRID_FIXED_IN the
 * per-wavefront sums are rounded to 64-bit fixed point and summed as integers (per
 * workgroup rows in `workspace`, or above 64 slots integer atomics into a row that is zero
 * between calls), then added to `wrench`: order-independent, for any slot count. The
 * integer rows are zero between flushes; with MPM_GRID_WRENCH_DEFER they keep
 * accumulating until mpm_grid_wrench_flush (same capacity, slots and workspace).
 * A block of scene s applies colliders[s * ncoll .. s * ncoll + ncoll). Nodes without mass above 1e-15 read velocity 0 (the
 * split-zero contract, mpm_utils.py:1193-1202). With MPM_GRID_ZERO_IN, node_in is
 * cleared afterwards. */
int32_t mpm_grid_workspace_floats(int32_t capacity, int32_t wrench_slots);
int32_t mpm_grid_update(const mpm_grid_t *G, const mpm_model_t *model, const mpm_tick_ctl_t *ctl,
                        const int32_t *block_list, const int32_t *block_count, int32_t capacity,
                        const mpm_collider_t *colliders, int32_t ncoll, float *wrench,
                        int32_t wrench_slots, float *workspace, float dt, int32_t flags,
                        void *stream);
/* Allocation-free variant of mpm_grid_update. No counterpart in the reference: it is for
 * callers that cannot allocate device memory or pass a workspace (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The wrench partial rows
 * live in a file-scope device array of MPM_NOALLOC_SCRATCH_FLOATS floats (3 MiB) that each
 * GPU the library runs on holds until the process ends; its address is looked up on the
 * first call on a device (the device current at the call; `stream` must belong to it).
 * Same arguments and results as mpm_grid_update without `workspace`; MPM_ERR_SIZE when
 * mpm_grid_workspace_floats(capacity, wrench_slots) exceeds the array, MPM_ERR_ARG with
 * MPM_GRID_WRENCH_DEFER (deferred rows cannot outlive the call in a shared array). It
 * allocates nothing. Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (other streams, other host threads) race. */
#define MPM_NOALLOC_SCRATCH_FLOATS (1024 * 64 * 12)
int32_t mpm_grid_update_noalloc(const mpm_grid_t *G, const mpm_model_t *model, const mpm_tick_ctl_t *ctl,
                                const int32_t *block_list, const int32_t *block_count, int32_t capacity,
                                const mpm_collider_t *colliders, int32_t ncoll, float *wrench,
                                int32_t wrench_slots, float dt, int32_t flags, void *stream);
/* Convert the integer wrench rows that MPM_GRID_FIXED_IN | MPM_GRID_WRENCH_DEFER calls left
 * in `workspace` into `wrench` (added) and clear them. No counterpart in the reference
 * (its wrench accumulators are float atomics, mpm_solver_warp.py:2121-2269). */
int32_t mpm_grid_wrench_flush(const mpm_grid_t *G, int32_t capacity, float *wrench, int32_t wrench_slots,
                              float *workspace, void *stream);

/* ---- sparse active blocks (mpm_utils.py:1506-1544, mpm_solver_warp.py:989-1026) ----- */
/* raw must be zero on entry; marks the stencil-base block of every particle. */
int32_t mpm_blocks_mark(const mpm_particles_t *P, const mpm_grid_t *G, int32_t *raw, void *stream);
int32_t mpm_blocks_dilate(const mpm_grid_t *G, const int32_t *raw, int32_t *cur, void *stream);
/* counts (device, 2 ints) must be zero on entry; cur_list / union_list get block codes in
 * ascending order (an order-preserving compaction, so the block lists, and the fixed-order
 * wrench reduction that follows them, are the same from run to run). */
int32_t mpm_blocks_compact(const mpm_grid_t *G, const int32_t *cur, const int32_t *prev,
                           int32_t *cur_list, int32_t *union_list, int32_t *counts, void *stream);

/* ---- guard / launch box (solver.py:485-518) ---------------------------------------- */
/* out (device, 8 floats): min x, y, z, max x, y, z, max |v| component, non-finite count.
 * Works on all particles, like the reference's host readback. */
int32_t mpm_bounds(const mpm_particles_t *P, float *out, void *stream);
/* Launch box on the device from mpm_bounds' output (device): the particle box widened by
 * pad + travel * (max |v| component) metres, as the node box [floor(lo / dx),
 * ceil(hi / dx) + 1) clamped to the grid, written to ctl->box_lo / box_hi (the whole grid
 * when a position is not finite). No counterpart in the reference, whose host computes
 * the box from a readback (solver.py:485-518): for callers that keep the tick on the
 * device (the differentiable rollout's per-segment boxes). */
int32_t mpm_box_from_bounds(const float *bounds, const mpm_grid_t *G, float pad, float travel,
                            mpm_tick_ctl_t *ctl, void *stream);

/* ---- block sort (solver.py:454-467, mpm_solver_warp.py:1548-1619) ------------------- */
/* keys[p] = scene*nb^3 + ((bx*nb + by)*nb + bz) with b = floor(x/dx - 0.5) >> 2,
 * nb = n0/4 + 2 (scene 0 without batching); vals[p] = p; *unsorted (device int) gets 1
 * if keys are not non-decreasing. MPM_ERR_SIZE when nb^3 * n_scenes exceeds int32. */
int32_t mpm_sort_keys(const mpm_particles_t *P, const mpm_grid_t *G, int32_t *keys,
                      int32_t *vals, int32_t *unsorted, void *stream);
/* Same with the key refined by the stencil base cell inside its block: (key << 6) |
 * ((cx & 3) << 4 | (cy & 3) << 2 | (cz & 3)), c = floor(x/dx - 0.5). Sorted by it, the
 * particles keep the block order of mpm_sort_keys and those sharing a stencil base cell
 * are consecutive, which the particle pass's segmented scatter reduction exploits. No
 * counterpart in the reference (its sort is by block only). */
int32_t mpm_sort_keys_cells(const mpm_particles_t *P, const mpm_grid_t *G, int32_t *keys,
                            int32_t *vals, int32_t *unsorted, void *stream);
/* temp storage needed by mpm_sort_pairs for n keys (no counterpart: the reference sorts
 * with a host argsort, solver.py:454-467) */
int64_t mpm_sort_temp_bytes(int32_t n);
/* stable radix sort of (keys, vals) pairs into (keys_out, vals_out) */
int32_t mpm_sort_pairs(const int32_t *keys, int32_t *keys_out, const int32_t *vals,
                       int32_t *vals_out, int32_t n, void *temp, int64_t temp_bytes, void *stream);
/* Chunk boundaries for the particle pass over sorted keys (mpm_sort_keys: key_shift 0;
 * mpm_sort_keys_cells: key_shift 6; nb = n0 / 4 + 2 as in the keys): the particles of each
 * group of gz consecutive blocks of one block column (scene, bx, by, bz / gz) form a run in
 * key order, and each run is cut into chunks of at most `chunk` particles. chunk_start gets
 * *count + 1 entries (capacity n + 1; the last is n), *count (device int) the number of
 * chunks. Fixed chunks of 256 consecutive particles can straddle a jump in key order, whose
 * stencil box then spans two distant clusters and overflows the pass's LDS tile; these keep
 * each box within gz blocks of one column. temp: mpm_sort_chunks_temp_bytes(n) bytes. No
 * counterpart in the reference (its particle kernels launch one thread per particle). */
int64_t mpm_sort_chunks_temp_bytes(int32_t n);
int32_t mpm_sort_chunks(const int32_t *keys, int32_t n, int32_t key_shift, int32_t nb, int32_t gz, int32_t chunk,
                        int32_t *chunk_start, int32_t *count, void *temp, int64_t temp_bytes, void *stream);

/* ---- block pool ------------------------------------------------------------------------
 * Sparse storage for the active blocks only (the batched scenes' grids: 256 scenes of
 * 64^3 would need 95M nodes densely, the dough scenes touch ~10 % of them). Replaces the
 * reference's dense grid under its sparse block lists (mpm_solver_warp.py:989-1026,
 * blocks_compact mpm_utils.py:1532-1544): from the dilated active flags `cur` (1 =
 * active, per global block code, n_codes = nb[0] nb[1] nb[2] n_scenes entries, as
 * mpm_blocks_dilate writes them) it writes slot[code] = the number of active codes below
 * `code` for an active code and -1 otherwise, list[slot] = code (the active codes in
 * increasing order: the grid update's block list) and *count = the number of active
 * codes. The caller sizes the node arrays to at least *count blocks, sets
 * G->block_slot = slot and G->pool_blocks, and clears node_in / node_out of the slots in
 * use (a slot may have held another block before). temp: mpm_pool_temp_bytes(n_codes). */
int64_t mpm_pool_temp_bytes(int32_t n_codes);
int32_t mpm_pool_assign(const mpm_grid_t *G, const int32_t *cur, int32_t *slot, int32_t *list, int32_t *count,
                        void *temp, int64_t temp_bytes, void *stream);
/* gather every particle array by perm (dst[i] = src[perm[i]]) through `scratch`
 * (>= 9 * stride floats) and copy back in place, so device pointers stay stable. */
int32_t mpm_permute(const mpm_particles_t *P, const int32_t *perm, float *scratch, void *stream);

/* ---- coupling readout (box_contact_wrench, coupling/wrench.py:15-46) -------------- */
/* Stress-integral reaction on a box end-effector from the particle state on the device.
 * box = (centre xyz, half extents xyz) on the host; layer = contact-layer thickness T (m),
 * or <= 0 for 3 * mean(vol0 |det F|)^(1/3). A particle counts when it lies under the box
 * footprint with z in (bottom - T, bottom + T/2) and, with compressive_only, sigma_zz < 0;
 * Cauchy = tau / max(|det F|, 1e-9). out (device, 8 floats): force xyz, torque xyz about
 * the centre, contact count (force and torque zero when it is below 3), count of particles
 * with det F <= 0 or non-finite. workspace: mpm_box_contact_workspace_floats() floats.
 * Fixed-order reductions: bit-reproducible. */
int32_t mpm_box_contact_workspace_floats(void);
int32_t mpm_box_contact_wrench(const mpm_particles_t *P, const float *box, float layer,
                               int32_t compressive_only, float *out, float *workspace, void *stream);

/* ---- mesh voxelisation (geometry/mesh_sdf.py:328-366, 412-481) ----------------------- */
/* Fields of a triangle mesh (verts device float[3 nv], faces device int32[3 nf]) on the
 * cubic voxel grid origin + cell * (i, j, k), i, j, k < res (origin: host float[3]), z
 * fastest (dist and aux device float[res^3]). dist = exact unsigned distance to the nearest triangle. mode 0: aux =
 * winding number (sum of solid angles / 4 pi: ~1 inside a closed outward-oriented mesh, ~0
 * outside); mode 1: aux = +1 / -1, the side of the nearest triangle's normal (for open
 * sheets). A face with a vertex index outside [0, nv) is ignored. */
int32_t mpm_mesh_field(const float *verts, int32_t nv, const int32_t *faces, int32_t nf,
                       const float *origin, float cell, int32_t res, int32_t mode, float *dist,
                       float *aux, void *stream);
/* dist = distance from each voxel to the nearest of ns segments (device float[6 ns]: a, b) */
int32_t mpm_segment_distance(const float *segments, int32_t ns, const float *origin, float cell,
                             int32_t res, float *dist, void *stream);

/* ---- exports ------------------------------------------------------------------------ */
/* compute_R_from_F (mpm_utils.py:1261-1285): R_out[9*p + ...] = (U V^T)^T from F_trial,
 * row-major AoS [n][3][3], as the reference stores it. */
int32_t mpm_export_R(const mpm_particles_t *P, float *R_out, void *stream);
/* compute_cov_from_F (mpm_utils.py:1234-1258): cov_out[6*p + ...] = F init F^T from
 * F_trial and init_cov (AoS [n][6]). */
int32_t mpm_export_cov(const mpm_particles_t *P, const float *init_cov, float *cov_out,
                       void *stream);
/* dense copies of the grid in the reference's layout (z fastest): m [n^3], v_in and
 * v_out [n^3][3], one such copy per scene back to back; any output may be NULL. */
int32_t mpm_grid_to_dense(const mpm_grid_t *G, float *m, float *v_in, float *v_out, void *stream);

/* ---- adjoint (reverse-mode gradient of one substep) ---------------------------------
 * No counterpart in the reference (its solver has no gradient). One reverse substep k is
 *   mpm_adjoint_g2p -> mpm_adjoint_grid -> mpm_adjoint_p2g_stress
 * on the forward data of substep k: P holds x_k, v_k, C_k, F_trial_k (state before the
 * substep) and F_k, stress_k (after its return map) and L_k (from the G2P that ended the
 * substep); G holds node_in = (mv, m) after
 * the scatter and node_out = grid velocity after the colliders. On entry `A` holds the
 * adjoint of S_{k+1} = (x, v, C, F_trial); on exit the adjoint of S_k, with the
 * per-particle parameter adjoints accumulated. Single scene; material 0 (fixed-corotated
 * jelly: mu, lam), material 1 (von Mises + Hencky, no hardening: mu, lam, yield_stress),
 * material 2 (Drucker-Prager sand: mu, lam), material 9 (local mu(I): mu, lam and
 * mu_s, delta_mu, I0, grain diameter as per-particle partials of the per-type
 * coefficients) and material 10 (Newtonian / Bingham / Herschel-Bulkley fluid:
 * viscosity, bulk modulus, K, n as per-particle partials of the per-type parameters, and
 * yield_stress) and the plane / box / walls / point-cloud colliders have adjoints; other
 * ids set error bits 16 / 32.
 * The viscous stress of substep k reads the velocity gradient L_k of the gather that
 * produced F_trial_k, so for mpm_adjoint_p2g_stress P->L must point at L_k (for
 * mpm_adjoint_g2p at the L its gather produced); A->L carries the adjoint of L_k from the
 * stress adjoint of substep k to the gather adjoint of substep k - 1 (zero for the
 * last state). */
typedef struct mpm_adjoint {
    float *x, *v, *C, *F_trial;      /* adjoint of the substep state (3, 3, 9, 9 planes) */
    float *F, *stress;               /* adjoint of the return-mapped F and of tau (9, 9) */
    float *mu, *lam, *yield_stress;  /* accumulated per-particle parameter adjoints */
    float *grid_v;                   /* float4 per node: adjoint of the grid velocity; zero
                                        before mpm_adjoint_g2p, consumed (zeroed) by
                                        mpm_adjoint_grid */
    float *grid_mv;                  /* float4 per node: adjoint of (momentum, mass) */
    float *L;                        /* adjoint of the velocity gradient L_k (9 planes) */
    float *visc, *bulk, *hb_K, *hb_n; /* accumulated per-particle partials of the per-type
                                        viscous parameters (material 10) */
    float *mui_mu_s, *mui_dmu, *mui_I0, *mui_d; /* accumulated per-particle partials of the
                                        per-type mu(I) coefficients (material 9): mu_s,
                                        delta_mu, I0, grain diameter */
} mpm_adjoint_t;

/* G2P adjoint: scatters the grid-velocity adjoint (float atomics), writes the adjoint
 * of x_k (partial) into A->x and of F_k into A->F. */
int32_t mpm_adjoint_g2p(const mpm_particles_t *P, const mpm_grid_t *G, const mpm_adjoint_t *A,
                        const mpm_model_t *model, float dt, void *stream);
/* Grid adjoint over the listed blocks (block_list / *block_count as mpm_blocks_from_box
 * writes them, at most `capacity`; NULL: every block): the colliders (rows of this
 * substep, ncoll <= 16) and the normalisation transposed; impulse_adj (float[6 *
 * wrench_slots]) is the adjoint of the accumulated wrench. Nodes outside the listed
 * blocks must hold no mass and no velocity adjoint. */
int32_t mpm_adjoint_grid(const mpm_grid_t *G, const mpm_model_t *model, const int32_t *block_list,
                         const int32_t *block_count, int32_t capacity, const mpm_collider_t *colliders,
                         int32_t ncoll, const float *impulse_adj, int32_t wrench_slots,
                         const mpm_adjoint_t *A, float dt, void *stream);
/* P2G adjoint (gather) fused with the return-map adjoint: completes the adjoint of x_k,
 * writes those of v_k, C_k, F_trial_k and adds the parameter adjoints. */
int32_t mpm_adjoint_p2g_stress(const mpm_particles_t *P, const mpm_grid_t *G, const mpm_adjoint_t *A,
                               const mpm_model_t *model, float dt, void *stream);

#ifdef __cplusplus
}
#endif
#endif
