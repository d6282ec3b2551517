/*
 * d3q19_abi.h -- C-ABI of the MI355X (gfx950) D3Q19 immersed-boundary shear-cell step.
 *
 * This header is the drop-in boundary that replaces the reference's OpenCL
 * context / queue / program / kernel-argument plumbing:
 *   sim_main.c:15-42        (context + two in-order queues)
 *   sim_main.c:56,115-306   (create_LB_kernels, clCreateBuffer, clEnqueueWriteBuffer, clSetKernelArg)
 *   sim_main.c:315-471      (per-step clEnqueueNDRangeKernel sequence and clFinish points)
 *   D3Q19-OpenCL.c:465-556  (create_LB_kernels)
 *   D3Q19-OpenCL.c:847-971  (analyse_platform)
 *   D3Q19-OpenCL.c:1169-1246 (error_check)
 *
 * Conventions
 *   - Every entry point returns int: 0 on success, a negative D3Q19_E* code on failure.
 *     The library never calls exit(); d3q19_strerror() and d3q19_last_error() describe failures.
 *   - Host arrays passed in are borrowed for the duration of the call only.
 *   - The library owns all device memory.  One HIP stream per simulation; calls on one
 *     simulation must come from one host thread at a time.  Kernel launches are asynchronous;
 *     d3q19_read_* and d3q19_sync() synchronise the simulation's stream.
 *   - Field arrays use the reference host layouts: node index i = x + Nx*(y + Ny*z) over the
 *     whole LatticeSize including the one-node buffer shell, populations f[q*N_C + i],
 *     velocity u[d*N_C + i] (D3Q19-OpenCL.c:148-156), particles as float4 arrays
 *     (sim_main.c:88-90, D3Q19-OpenCL.c:238-239,327-339).
 *   - No torch types, no OpenCL types, no CUDA types in any signature.
 */
#ifndef D3Q19_ABI_H
#define D3Q19_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define D3Q19_ABI_VERSION 1

/* ------------------------------------------------------------------------------------------
 * Parameter structs.  Field order, types and sizes are those of the reference's
 * int_param_struct / flp_param_struct (struct_header_host.h:4-57, duplicated for the devices
 * in GPU_program.cl:19-72 and CPU_program.cl:19-72) so that a host that fills them for the
 * reference fills them for this library unchanged: 83 x int32 = 332 bytes and
 * 44 x float32 = 176 bytes.
 * ------------------------------------------------------------------------------------------ */
typedef struct {
    int32_t MaxIterations;
    int32_t BasisVel[19][3];
    int32_t LatticeSize[3];
    int32_t SystemSize[3];
    int32_t BufferSize[3];
    int32_t BoundaryConds[3];       /* 0 periodic, 1 moving wall (velocity BC) */
    int32_t NumZones[3];
    int32_t MaintainShear;
    int32_t ViscosityModel;         /* 1 Newtonian, 2 power law, 3 Herschel-Bulkley, 4 Casson */
    int32_t NumParticles;
    int32_t ParForceModel;          /* 1 harmonic */
    int32_t MaxSurfPointsPerNode;
    int32_t InterpOrderIBM;
    int32_t PointsPerParticle;
    int32_t PointsPerWorkGroup;
    int32_t TotalSurfPoints;
    int32_t WorkGroupsPerParticle;
} int_param_struct;

typedef struct {
    float ConstBodyForce[3];
    float EqWeights[19];
    float VelUpper[3];
    float VelLower[3];
    float NewtonianTau;
    float ViscosityParams[4];
    float ZoneWidth[3];
    float ParticleDiam;
    float ParticleMass;
    float PointArea;
    float ParticleMomInertia;
    float ParForceParams[2];
    float ParticleZBuffer;
    float DirectForcingCoeff;
} flp_param_struct;

typedef struct { float x, y, z, w; } d3q19_float4;

#define D3Q19_VISC_NEWTONIAN 1
#define D3Q19_VISC_POWER_LAW 2
#define D3Q19_VISC_HB        3
#define D3Q19_VISC_CASSON    4

/* ------------------------------------------------------------------------------------------
 * Options that the reference fixes at compile time (GPU_program.cl:4-10) or that the
 * benchmark configurations add.  Zero-initialising the struct selects the reference build.
 * ------------------------------------------------------------------------------------------ */
#define D3Q19_COLL_MRT 0   /* d'Humieres MRT + Guo forcing, GPU_program.cl:396-623 */
#define D3Q19_COLL_BGK 1   /* single relaxation time with the same forcing and smoothing */

#define D3Q19_IBM_LINEAR_SMOOTHED 0  /* trilinear + 27-point force smoothing, GPU_program.cl:81-256, 356-377 */
#define D3Q19_IBM_PESKIN4         1  /* Peskin 4-point regularised delta for interpolation and spreading */
/* ibm_layout: 0 automatic (four markers per wavefront from 16384 markers, then the largest groups
 * of up to 64 markers whose LDS boxes fit in 2048 nodes while 512 groups remain, else 8); 1 one
 * marker per wavefront, groups of 8; 2 four markers per wavefront (one per 16-lane row), groups of
 * 8; 3 four per wavefront, the largest groups (up to 64 markers) whose boxes fit.  The trilinear
 * kernel always takes layout 1. */
#define D3Q19_IBM_LAYOUT_AUTO 0
#define D3Q19_IBM_LAYOUT_WAVE 1
#define D3Q19_IBM_LAYOUT_ROWS 2
#define D3Q19_IBM_LAYOUT_ROWS_LARGE 3

#define D3Q19_STORE_FP32        0
#define D3Q19_STORE_FP16_SHIFT  1    /* binary16 of (f_q - w_q)/w_q * 2^14, fp32 arithmetic */
#define D3Q19_STORE_BF16_SHIFT  2    /* bfloat16 of  f_q - w_q, fp32 arithmetic */
#define D3Q19_STORE_FP64        3    /* binary64 everywhere (GPU fp64 oracle) */

typedef struct {
    int32_t collision;          /* D3Q19_COLL_* */
    int32_t variable_viscosity; /* 0: tau = NewtonianTau everywhere (reference USE_CONSTANT_VISCOSITY);
                                   1: per-node tau from ViscosityModel (GPU_program.cl:457-564,1041-1128) */
    float   tau_min;            /* lower clamp of tau; <= 0 selects the reference's 0.505 (GPU_program.cl:6) */
    float   tau_max;            /* upper clamp of tau; <= 0 means no ceiling (GPU_program.cl:1070) */
    int32_t ibm_kernel;         /* D3Q19_IBM_* */
    int32_t storage;            /* D3Q19_STORE_* */
    int32_t calc_rho;           /* velocity BC density mode, sim_main.c:216 (reference: 0) */
    int32_t tangential_bc[3];   /* hostDat.TangentialVelBC, sim_main.c:388-404 */
    int32_t block[3];           /* collide+stream tile (threads); 0 = library default / autotuned */
    int32_t contact_detection;  /* D3Q19_CONTACT_*: which sphere pairs the contact forces visit */
    int32_t rebuild_freq;       /* hostDat.RebuildFreq: zone lists rebuilt after steps t with
                                   t % rebuild_freq == 0 (sim_main.c:429-444); <= 0 selects 10 */
    int32_t rounding;           /* D3Q19_ROUND_*: how the collide rounds fp16 / bf16 shifted stores */
    int32_t overlap_slabs;      /* single simulations with spheres: the collide runs as this many
                                   z-slabs and each slab's spheres go through the IBM pass on a second
                                   stream while the later slabs collide; 0 or 1 = one launch each (the
                                   faster on MI355X), 2..16 = that many slabs.  Results do not depend
                                   on it (bitwise). */
    int32_t ibm_layout;         /* D3Q19_IBM_LAYOUT_*: how the IBM pass maps Peskin-4 markers onto
                                   wavefronts and LDS groups (single simulations).  Results do not
                                   depend on it (bitwise). */
    int32_t reserved[1];
} d3q19_options;

/* Rounding of the collide's 16-bit stores (no counterpart in the reference, which stores fp32).
 * STOCHASTIC: unbiased stochastic rounding (gfx950 v_cvt_sr_f16_f32 for fp16, integer dither
 *   for bf16) with random bits hashed from the node index and the node's incoming populations:
 *   deterministic, the same for a cell alone and in a batch, fresh whenever the inputs change.
 *   Nearly uniform fields (the populations next to a moving wall) otherwise get the same rounding
 *   error at every node, which does not average out, and updates below half an ulp are lost.
 * NEAREST: round to nearest even. */
#define D3Q19_ROUND_STOCHASTIC 0
#define D3Q19_ROUND_NEAREST    1

/* Contact pair enumeration (particle_particle_forces, CPU_program.cl:76-184).
 * ZONES: the reference's neighbour-zone lists, quirks included (D3Q19-OpenCL.c:344-463,
 *   CPU_program.cl:280-315): a sphere's zone is (int)(x / ZoneWidth) per axis from the last
 *   rebuild (int_param_struct.NumZones, flp_param_struct.ZoneWidth), a zone's neighbour ids
 *   are formed with NumZones[2] where NumZones[1] is meant (D3Q19-OpenCL.c:450), and a
 *   neighbour zone reached twice (NumZones < 3 on a periodic axis) counts its pairs twice.
 *   Neighbour ids that fall outside the zone grid match no sphere.
 * ALL_PAIRS: every pair once. */
#define D3Q19_CONTACT_ZONES     0
#define D3Q19_CONTACT_ALL_PAIRS 1

/* Error codes */
#define D3Q19_OK               0
#define D3Q19_E_INVALID      (-1)   /* bad argument or inconsistent parameter struct */
#define D3Q19_E_DEVICE       (-2)   /* HIP runtime failure */
#define D3Q19_E_NOMEM        (-3)   /* device allocation failed */
#define D3Q19_E_UNSUPPORTED  (-4)   /* configuration not supported by this build */
#define D3Q19_E_STATE        (-5)   /* call not valid in the current state */
#define D3Q19_E_NONFINITE    (-6)   /* watchdog: non-finite or non-positive density detected */
#define D3Q19_E_CAPACITY     (-7)   /* server / batch capacity exceeded */

/* Step flags */
#define D3Q19_STEP_WRITE_U   1      /* write u on every node this step (reference writes it each step,
                                       GPU_program.cl:391-394); otherwise u is written only where the
                                       next IBM pass reads it */

typedef struct d3q19_sim d3q19_sim;

/* Device discovery; replaces analyse_platform (D3Q19-OpenCL.c:847-971). */
int d3q19_device_count(int* count);
int d3q19_device_name(int device, char* buf, size_t len);

/* Create a simulation on `device`: replaces create_LB_kernels + clCreateBuffer +
 * clEnqueueWriteBuffer of the parameter structs + clSetKernelArg (sim_main.c:56,115-306).
 * `sphere_points` holds PointsPerParticle marker offsets already scaled by the radius
 * (D3Q19-OpenCL.c:595-606); it may be NULL when NumParticles == 0.
 * Fields start as f = f_eq(rho=1,u=0), u = 0, tau = NewtonianTau, forces 0, particles 0. */
int d3q19_create(int device, const int_param_struct* intDat, const flp_param_struct* flpDat,
                 const d3q19_options* opt, const d3q19_float4* sphere_points, d3q19_sim** out);
int d3q19_destroy(d3q19_sim* sim);

/* Uploads / downloads in the reference host layouts (sim_main.c:193-205, 452-462, 478, 486-498).
 * f is the pre-collision population field: what the reference's f_s holds after
 * boundary_periodic + boundary_velocity, i.e. the input of the next collide.  Only interior
 * nodes (1..N-2 on every axis) are state; read_f fills the buffer shell with zeros. */
int d3q19_write_f(d3q19_sim* sim, const float* f);
int d3q19_read_f(d3q19_sim* sim, float* f);
int d3q19_write_f64(d3q19_sim* sim, const double* f);
int d3q19_read_f64(d3q19_sim* sim, double* f);
int d3q19_write_u(d3q19_sim* sim, const float* u);
int d3q19_read_u(d3q19_sim* sim, float* u);
int d3q19_write_tau(d3q19_sim* sim, const float* tau);
int d3q19_read_tau(d3q19_sim* sim, float* tau);
/* The fluid force that the next collide adds to ConstBodyForce, per node (3*N_C): the
 * 27-point smoothed spread force (linear IBM) or the Peskin spread force.  Reference:
 * gpf slot 0 after sum_particle_fluid_forces, smoothed as in GPU_program.cl:364-377. */
int d3q19_read_force(d3q19_sim* sim, float* g);
/* parKin float4[4*np] = [pos | vel | quat | angvel]; parForce float4[2*np] (contact force |
 * torque); parFluidForce float4[2*np*nfa] laid out as the reference's per-work-group
 * partials (GPU_program.cl:247-255); the read-back puts each sphere's total in work-group 0.
 * Any pointer may be NULL to skip that array. */
int d3q19_write_particles(d3q19_sim* sim, const d3q19_float4* parKin,
                          const d3q19_float4* parForce, const d3q19_float4* parFluidForce);
int d3q19_read_particles(d3q19_sim* sim, d3q19_float4* parKin,
                         d3q19_float4* parForce, d3q19_float4* parFluidForce);
/* FP64 particle state: pos, vel, angvel (3 doubles each per sphere) and fluid force/torque. */
int d3q19_read_particles_f64(d3q19_sim* sim, double* pos, double* vel, double* angvel,
                             double* fluid_force, double* fluid_torque);

/* One entry per reference kernel (kernel-level parity).
 * collide_stream fuses collideMRT_stream_D3Q19 (GPU_program.cl:316-624), boundary_periodic
 * (626-735) and boundary_velocity (737-865, incl. the tangential passes of sim_main.c:388-404):
 * after it, the pre-collision field (d3q19_read_f) equals the reference's f_s after those
 * three kernels. */
int d3q19_collide_stream(d3q19_sim* sim, int flags);
/* particle_dynamics, CPU_program.cl:187-277 (FP64 on the GPU). */
int d3q19_particle_dynamics(d3q19_sim* sim);
/* particle_fluid_forces_linear_stencil + sum/reset_particle_fluid_forces
 * (GPU_program.cl:81-313): interpolation, penalty force, deterministic spreading, FP64
 * per-sphere force and torque sums. */
int d3q19_particle_fluid_forces(d3q19_sim* sim);
/* particle_particle_forces, CPU_program.cl:76-184 (FP64; pairs as d3q19_options.contact_detection
 * selects, the zone lists rebuilt as in sim_main.c:428-444). */
int d3q19_particle_particle_forces(d3q19_sim* sim);

/* Checkpoint / restart (no counterpart in the reference, which writes only u at the end,
 * sim_main.c:476-481).  The image holds the whole state -- populations in their storage format,
 * u, tau, the pending spread force, brick marks and epochs, the sphere state, contact zones and
 * the step bookkeeping -- so that a simulation created with the same parameter structs, options
 * and marker offsets continues bit for bit after a load.  `size` gives the image bytes; save and
 * load move it through host memory (a server can evict and resume a cell), write and read
 * through a file.  Loads reject images of other parameters (MaxIterations aside). */
int d3q19_checkpoint_size(d3q19_sim* sim, size_t* bytes);
int d3q19_checkpoint_save(d3q19_sim* sim, void* buf, size_t bytes);
int d3q19_checkpoint_load(d3q19_sim* sim, const void* buf, size_t bytes);
int d3q19_checkpoint_write(d3q19_sim* sim, const char* path);
int d3q19_checkpoint_read(d3q19_sim* sim, const char* path);

/* One full time step t in the order of sim_main.c:315-471. */
int d3q19_step(d3q19_sim* sim, int t, int flags);
/* n_steps steps t_first .. t_first+n_steps-1; replayed from a HIP graph of one step.
 * flags apply to the last step only. */
int d3q19_run(d3q19_sim* sim, int t_first, int n_steps, int flags);
int d3q19_sync(d3q19_sim* sim);

/* Observables (FP64 device reductions).
 * plane_mean_u: mean u over interior x-y planes for every lattice z (3*Nz doubles,
 * planes 0 and Nz-1 zero), the quantity of compute_shear_stress (D3Q19-OpenCL.c:694-725);
 * needs u written on every node (D3Q19_STEP_WRITE_U on the last step). */
int d3q19_plane_mean_u(d3q19_sim* sim, double* means);
/* d3q19_plane_mean_u without any allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The plane means are staged in a device array of the library, 3 x 2048
 * doubles (48 KiB) per GPU, held until the process ends; LatticeSize[2] above 2048 is refused with
 * D3Q19_E_CAPACITY.  All calls on one GPU share that array: calls that overlap on one GPU (different
 * simulations on different streams or threads) race and must be serialised by the caller.  Output
 * and semantics are those of d3q19_plane_mean_u (bitwise). */
int d3q19_plane_mean_u_noalloc(d3q19_sim* sim, double* means);
/* Total fluid mass and momentum over interior nodes (4 doubles) from the pre-collision f. */
int d3q19_moments(d3q19_sim* sim, double* out4);
/* Watchdog: returns D3Q19_E_STATE if a sphere has moved a lattice unit or more in one step since
 * the simulation was created (its brick marks then missed part of its stencils), else
 * D3Q19_E_NONFINITE if any interior node has a non-finite or non-positive density.  Allocates
 * nothing. */
int d3q19_check_finite(d3q19_sim* sim);

/* Shear-cell observables, reduced on the device in FP64 from the stored populations.
 * Wall axis A (BoundaryConds), flow axis F = the other axis with the largest wall-velocity
 * difference.  The viscous shear stress of a node is
 *   sigma_FA = -(1 - 1/(2 tau)) (sum_q c_qF c_qA f_q - rho u_F u_A + (u_F g_A + u_A g_F)/2)
 * (g = ConstBodyForce, tau = the node's relaxation time of the F-A stress moment).  It is averaged
 * over the two planes next to the wet wall layers.  The reference's apparent-viscosity observable,
 * a finite-difference shear rate of the plane-mean velocity (compute_shear_stress,
 * D3Q19-OpenCL.c:670-750), stays available through d3q19_plane_mean_u.
 * drag_coefficient = mean over spheres of |F_fluid| / (rho/2 * U^2 * pi a^2), with
 * U = |VelUpper_F - VelLower_F| / 2, a = ParticleDiam / 2 and rho = 1. */
typedef struct {
    double wall_stress;         /* mean of the two wall-adjacent planes */
    double wall_stress_lower;
    double wall_stress_upper;
    double shear_rate;          /* nominal: (VelUpper_F - VelLower_F) / (n_A - 1), n_A interior extent */
    double apparent_viscosity;  /* wall_stress / shear_rate */
    double fluid_viscosity;     /* the fluid's viscosity (tau law and clamps) at the nominal shear rate */
    double relative_viscosity;  /* apparent_viscosity / fluid_viscosity */
    double drag_coefficient;
    double mean_fluid_force;    /* mean over spheres of |F_fluid| */
    int32_t wall_axis, flow_axis;
    double reserved[4];
} d3q19_observables;
int d3q19_observables_get(d3q19_sim* sim, d3q19_observables* out);

/* ------------------------------------------------------------------------------------------
 * Batched ensembles.  No counterpart in the reference: independent shear cells (periodic x and
 * y, moving walls normal to z, no tangential passes) of different sizes, rheologies and sphere
 * sets step together.  Each step is four persistent launches over work lists held in device
 * memory, and two steps are replayed from one HIP graph.  The lists are rewritten when cells
 * join or leave, so the graph is never re-captured for that.  All cells share the batch's
 * population storage, collision operator and IBM kernel; per-cell options (variable_viscosity,
 * tau clamps, calc_rho) come from the d3q19_options given at d3q19_batch_add.  Device memory
 * comes from one arena allocated at creation.  A cell's results do not depend on the other cells
 * of the batch: they are bitwise those of the same cell alone in a batch.
 * ------------------------------------------------------------------------------------------ */
typedef struct d3q19_batch d3q19_batch;
typedef struct {
    int32_t storage;            /* D3Q19_STORE_* */
    int32_t collision;          /* D3Q19_COLL_* */
    int32_t ibm_kernel;         /* D3Q19_IBM_* */
    int32_t max_cells;          /* 0: 1024 */
    uint64_t arena_bytes;       /* device memory for the cells' state; 0: half of the free memory */
    int32_t reserved[8];
} d3q19_batch_options;
/* Launch shape of the batched kernels, chosen by d3q19_batch_autotune or set by the caller.
 * A grid of 0 means one work item per block (collide work-list entry, sphere) or per wavefront
 * (marker), rounded up to a multiple of 1024 blocks; a positive grid is a persistent grid that
 * strides over the work.  A collide work-list entry is `span` consecutive tiles of one cell, so
 * that a block reads its cell's descriptor once for span tiles. */
typedef struct {
    int32_t tile;               /* threads = nodes per collide tile: 64, 128 or 256 */
    int32_t grid_collide;       /* blocks of the collide kernel */
    int32_t grid_sphere;        /* blocks of the sphere kernels */
    int32_t grid_ibm;           /* blocks of the IBM kernel (4 markers at a time each) */
    double collide_ms;          /* measured by the autotuner: mean collide launch */
    int32_t span;               /* tiles of one cell per work-list entry (1, 2, 4 or 8); 0 = 1 */
    int32_t reserved[3];
} d3q19_tuning;
int d3q19_batch_create(int device, const d3q19_batch_options* bo, d3q19_batch** out);
int d3q19_batch_destroy(d3q19_batch* b);
/* Adds a cell at rest (f = f_eq(1, 0), u = 0, tau = NewtonianTau) with its spheres at the
 * positions and velocities of parKin (float4[4*np] as d3q19_write_particles; NULL = all zero).
 * Returns the cell's id (>= 0) in *cell_id; D3Q19_E_CAPACITY when the arena cannot hold it. */
int d3q19_batch_add(d3q19_batch* b, const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt,
                    const d3q19_float4* sphere_points, const d3q19_float4* parKin, int* cell_id);
int d3q19_batch_remove(d3q19_batch* b, int cell_id);
/* Bytes one cell of these parameters takes from the arena; and the arena's size and use. */
int d3q19_batch_cell_bytes(const d3q19_batch* b, const int_param_struct* ip, size_t* bytes);
int d3q19_batch_memory(const d3q19_batch* b, size_t* arena_bytes, size_t* used_bytes, int* cells);
/* n_steps steps of every cell (graph replay of step pairs). */
int d3q19_batch_run(d3q19_batch* b, int n_steps, int flags);
int d3q19_batch_sync(d3q19_batch* b);
/* Observables of every cell: out[i] for the i-th id of ids[0..n-1]. */
int d3q19_batch_observables(d3q19_batch* b, const int* ids, int n, d3q19_observables* out);
/* Per-cell read-backs in the reference layouts (see the single-simulation calls). */
int d3q19_batch_read_u(d3q19_batch* b, int cell_id, float* u);
int d3q19_batch_read_f(d3q19_batch* b, int cell_id, float* f);
int d3q19_batch_read_particles_f64(d3q19_batch* b, int cell_id, double* pos, double* vel, double* angvel,
                                   double* fluid_force, double* fluid_torque);
/* Times candidate launch shapes on the batch's current cells (trial launches leave the state
 * unchanged), keeps the fastest and returns it. */
int d3q19_batch_autotune(d3q19_batch* b, d3q19_tuning* out);
int d3q19_batch_get_tuning(const d3q19_batch* b, d3q19_tuning* out);
int d3q19_batch_set_tuning(d3q19_batch* b, const d3q19_tuning* t);
/* n steps timed with HIP events on the batch's stream: total milliseconds, and the mean duration
 * of the collide launch alone (16 trial launches that leave the state unchanged). */
int d3q19_batch_time_steps(d3q19_batch* b, int n, double* total_ms, double* collide_ms);

/* ------------------------------------------------------------------------------------------
 * Serving.  No counterpart in the reference.  Shear-cell solve requests are queued as they
 * arrive and admitted in arrival order when their footprint fits the batch arena.  Every admitted
 * cell stays resident and steps in the shared batch (continuous batching): requests that arrive
 * while others run join at the next check point, and converged cells leave, freeing their
 * memory.  A request converges when the relative change of its wall stress between two checks
 * (every check_interval of its own steps) falls below tolerance, after at least min_steps.
 * Results are independent of the other requests in the batch.
 * ------------------------------------------------------------------------------------------ */
typedef struct {
    double tolerance;           /* relative change of wall_stress between checks */
    int32_t check_interval;     /* steps between checks (even) */
    int32_t min_steps;
    int32_t max_steps;          /* give up (status 1) after this many steps */
    int32_t reserved[5];
} d3q19_solve_params;
typedef struct {
    int64_t id;
    int32_t status;             /* 0 converged, 1 max_steps reached, <0 error code: D3Q19_E_NONFINITE
                                   (non-finite or non-positive density, or non-finite wall stress,
                                   at a check), D3Q19_E_STATE (a sphere moved a lattice unit or
                                   more in one step; the run is not trusted) */
    int32_t steps;
    d3q19_observables obs;
    double t_submit, t_admit, t_finish;   /* d3q19_clock() seconds */
} d3q19_result;
typedef struct d3q19_server d3q19_server;
int d3q19_server_create(int device, const d3q19_batch_options* bo, d3q19_server** out);
int d3q19_server_destroy(d3q19_server* srv);
/* Queues a request (inputs are copied); *id receives its id. */
int d3q19_server_submit(d3q19_server* srv, const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt,
                        const d3q19_float4* sphere_points, const d3q19_float4* parKin, const d3q19_solve_params* sp,
                        int64_t* id);
/* One scheduling round: admit what fits, run one check interval of steps for all resident cells,
 * check them, retire the finished ones.  Completed results are written to done[0..*n_done-1]
 * (at most cap).  Returns with *n_done = 0 and nothing run when no cell is resident. */
int d3q19_server_progress(d3q19_server* srv, d3q19_result* done, int cap, int* n_done);
int d3q19_server_status(const d3q19_server* srv, int* queued, int* resident, size_t* arena_used);
int d3q19_server_get_batch(d3q19_server* srv, d3q19_batch** b);
/* The plain sequential solve of one request: a single simulation (d3q19_create) stepped with
 * d3q19_run in check_interval chunks with the same convergence rule. */
int d3q19_solve(int device, const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt,
                const d3q19_float4* sphere_points, const d3q19_float4* parKin, const d3q19_solve_params* sp,
                d3q19_result* out);
/* Monotonic clock (seconds) used for the result time stamps. */
double d3q19_clock(void);

/* Timing on the simulation's own stream (HIP events recorded at each kernel's own start and end,
 * so dispatch gaps are not counted).
 * d3q19_time_collide runs n (1..4096) collide_stream launches -- the collide+stream kernel alone,
 * without the particle dynamics and the IBM pass -- and returns their mean duration in
 * milliseconds.  It advances the lattice state by those n collides (the spread force pending
 * before it is consumed by the first): a kernel benchmark, whose state a simulation should not
 * continue from.
 * d3q19_time_steps runs n (1..4096) real full steps t_first .. t_first+n-1 (the state advances
 * exactly as d3q19_run's would), each kernel launched directly with its own events: total_ms is
 * the time from before the first step to after the last; kernel_ms (2 doubles, may be NULL) receives
 * the mean duration of the step kernel (collide+stream with the spheres' turn) and of the IBM pass
 * (0 without spheres). */
int d3q19_time_collide(d3q19_sim* sim, int n, int flags, double* ms_per_launch);
int d3q19_time_steps(d3q19_sim* sim, int t_first, int n, double* total_ms, double* kernel_ms);

/* Attainable-bandwidth probe (no counterpart in the reference; a measurement utility).  A
 * streaming kernel with the collide+stream kernel's memory traffic and none of its arithmetic
 * on an nx*ny*nz box of fp32 planes: `streams` (19, or 1 for a plain copy) planes read and as many
 * written, `width` (1, 2, 4) consecutive nodes per lane, read at the D3Q19 pull offsets when
 * bit 0 of `pull` is set; width -64 / -256 keeps one node per lane with the 19 populations
 * interleaved in chunks of 64 / 256 nodes instead of planes, and with bit 1 of `pull` set reads
 * and writes them with non-temporal loads and stores (as the collide does on lattices beyond the
 * Infinity Cache); `lds_bytes` (0..65536) of dynamic LDS per
 * 256-thread block caps the resident blocks.  Allocates and frees its own buffers.  *gbs = bytes read + written per launch / mean
 * launch time (HIP events around `launches` launches), *ms = that mean (ms may be NULL). */
int d3q19_stream_probe(int device, int nx, int ny, int nz, int streams, int width, int pull, int lds_bytes, int launches,
                       double* gbs, double* ms);

/* Sizes of the state (bytes of device memory held by the simulation). */
int d3q19_footprint(const int_param_struct* intDat, const d3q19_options* opt, size_t* bytes);

const char* d3q19_strerror(int code);
/* Message of the most recent failure on the calling thread. */
const char* d3q19_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* D3Q19_ABI_H */
