/*
 * d3q19_host.h -- C host of the D3Q19 IBM shear-cell solver.
 *
 * Keeps the reference's input_file*.txt format (keyword table, D3Q19-OpenCL.c:31-64, substring
 * matching of D3Q19-OpenCL.c:612-668), its parameter structs (include/d3q19_abi.h) and its
 * set-up and output semantics (D3Q19-OpenCL.c:94-463, 558-610, 670-835), and drives the HIP
 * library through the C-ABI instead of OpenCL.
 */
#ifndef D3Q19_HOST_H
#define D3Q19_HOST_H
#include <stdio.h>
#include "d3q19_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HOST_WORD 128
/* largest marker set (the reference's automatic choice tops out at 1024, D3Q19-OpenCL.c:564-571);
 * callers size the marker buffer of host_sphere_discretization as 4 * HOST_MAX_MARKERS floats */
#define HOST_MAX_MARKERS 1024

/* host_param_struct, D3Q19-OpenCL_header.h:75-94 (fields not read by kernels), plus the
 * configuration-mode options that the reference fixes at compile time. */
typedef struct {
    int ConsolePrintFreq;
    char InitialDist[HOST_WORD];
    float InitialVel[3];
    int InitialParticleDistribution; /* 1 grid, 2 grid+shift, 3 single, 4 seeded random with gap check */
    float ParticleBuffer;
    int InterpOrderIBM;
    int DomainDecomp[3];
    float ParticleDensity;
    int RebuildFreq;
    int VideoFreq;
    int FluidOutputSpacing;
    int TangentialVelBC[3];
    int ShearStressFreq;
    float RandParticleShift;
    /* additions (keywords that are not substrings of, and contain none of, the reference's) */
    int MarkerCount;          /* "marker_count": 0 = automatic 2^k (D3Q19-OpenCL.c:564-571) */
    long long PlacementSeed;  /* "placement_seed" for distribution 4 */
    float PlacementGap;       /* "placement_gap": minimum surface gap for distribution 4 */
    int CollisionModel;       /* "collision_operator": 0 MRT (reference), 1 BGK */
    int VariableViscosity;    /* "local_relaxation": 0 constant tau (reference build), 1 per-node */
    float TauClamp[2];        /* "tau_clamp": min max (0 0 = reference 0.505, none) */
    int IbmKernel;            /* "ibm_kernel": 0 linear+smoothing (reference), 1 Peskin 4-point */
    int Storage;              /* "population_storage": 0 fp32, 1 fp16 shifted, 2 bf16 shifted, 3 fp64 */
    int MassModel;            /* "mass_model": 0 reference (4/3)pi rho d^3, 1 sphere (pi/6) rho d^3 */
    int ContactDetection;     /* "contact_detection": 0 zone lists (reference), 1 all pairs */
    int Rounding;             /* "population_rounding": 16-bit stores, 0 stochastic, 1 nearest */
    char DataDir[256];        /* "marker_directory": folder holding unit_sphere_partition_N.txt */
} host_param_struct;

typedef struct {
    int ShearStressCount;
    float ShearStressAvg;
    float ActualShearRate;
} output_data_struct;

/* Parser (D3Q19-OpenCL.c:6-91, 612-668).  Fills defaults, then applies the lines of `path`.
 * Returns 0, or -1 if the file cannot be opened. */
int host_initialize_data(const char* path, int_param_struct* ip, flp_param_struct* fp, host_param_struct* hp);
/* Applies one "keyword value..." line with the reference's substring semantics. */
void host_process_input_line(const char* line, int_param_struct* ip, flp_param_struct* fp, host_param_struct* hp);
/* parameter_checking, D3Q19-OpenCL.c:94-109: number of errors. */
int host_parameter_checking(const int_param_struct* ip, const flp_param_struct* fp, const host_param_struct* hp);
/* SystemSize rule, D3Q19-OpenCL.c:119-128 */
void host_system_size(int_param_struct* ip);
/* sphere_discretization, D3Q19-OpenCL.c:558-610: marker count, PointArea, scaled points.
 * `points` must hold 4 * HOST_MAX_MARKERS floats.  Returns 0, or -1 on a missing / malformed file
 * or a marker_count above HOST_MAX_MARKERS. */
int host_sphere_discretization(int_param_struct* ip, flp_param_struct* fp, const host_param_struct* hp, float* points);
/* Work-group split of the marker kernel (D3Q19-OpenCL.c:536-550) for a device whose marker
 * kernel work-group limit is `max_wg`. */
void host_workgroup_split(int_param_struct* ip, int max_wg);
/* initialize_lattice_fields (D3Q19-OpenCL.c:111-189): f (19*N_C), u (3*N_C), tau (N_C). */
void host_initialize_lattice_fields(const host_param_struct* hp, int_param_struct* ip, const flp_param_struct* fp,
                                    float* f, float* u, float* tau);
/* initialize_particle_fields (D3Q19-OpenCL.c:191-342): parKin 4*np float4, parForce 2*np,
 * parFluidForce 2*np*nfa.  Returns 0 or -1 on an unplaceable configuration. */
int host_initialize_particle_fields(const host_param_struct* hp, const int_param_struct* ip, flp_param_struct* fp,
                                    float* parKin, float* parForce, float* parFluidForce);
/* Zone geometry of initialize_particle_zones (D3Q19-OpenCL.c:348-375): NumZones, ZoneWidth. */
void host_zone_geometry(const host_param_struct* hp, int_param_struct* ip, flp_param_struct* fp);
/* f_eq(rho, vel), D3Q19-OpenCL.c:1114-1149 */
void host_equilibrium(float rho, const float* vel, float* feq);
/* compute_shear_stress (D3Q19-OpenCL.c:670-750): writes velocity_profile_z.txt into `dir`
 * (NULL = cwd) and updates od. */
void host_compute_shear_stress(output_data_struct* od, const host_param_struct* hp, const int_param_struct* ip,
                               const flp_param_struct* fp, const float* u, const char* dir, int quiet);
/* continuous_output (D3Q19-OpenCL.c:752-792) */
void host_continuous_output(const host_param_struct* hp, const int_param_struct* ip, const float* u,
                            const float* parKin, FILE* vid, int frame);
/* write_lattice_field (D3Q19-OpenCL.c:794-835): velocity_field_final.txt and
 * matlab_postproc/velocity_field_centerline.txt under `dir`. */
int host_write_lattice_field(const float* u, const int_param_struct* ip, const char* dir);
/* SplitMix64 stream used for seeded placement */
unsigned long long host_splitmix64(unsigned long long* state);
/* options struct for the library from the host parameters */
void host_fill_options(const host_param_struct* hp, d3q19_options* opt);

#ifdef __cplusplus
}
#endif
#endif
