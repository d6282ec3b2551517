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
    int32_t reserved[6];
} d3q19_options;

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
/* particle_particle_forces, CPU_program.cl:76-184 (all pairs, FP64). */
int d3q19_particle_particle_forces(d3q19_sim* sim);

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
/* Total fluid mass and momentum over interior nodes (4 doubles) from the pre-collision f. */
int d3q19_moments(d3q19_sim* sim, double* out4);
/* Watchdog: returns D3Q19_E_NONFINITE if any interior node has a non-finite or non-positive
 * density. */
int d3q19_check_finite(d3q19_sim* sim);

/* Timing on the simulation's own stream (HIP events). Runs n collide_stream launches and
 * returns the mean duration per launch in milliseconds. */
int d3q19_time_collide(d3q19_sim* sim, int n, int flags, double* ms_per_launch);
/* Runs n full steps from t_first and returns total milliseconds plus the mean collide launch
 * duration (events recorded around every collide launch on the same stream). */
int d3q19_time_steps(d3q19_sim* sim, int t_first, int n, double* total_ms, double* collide_ms);

/* Sizes of the state (bytes of device memory held by the simulation). */
int d3q19_footprint(const int_param_struct* intDat, const d3q19_options* opt, size_t* bytes);

const char* d3q19_strerror(int code);
/* Message of the most recent failure on the calling thread. */
const char* d3q19_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* D3Q19_ABI_H */
