/*
 * params.c -- input-file parser and run set-up of the C host.
 *
 * The keyword table, defaults and substring matching follow initialize_data and
 * process_input_line (D3Q19-OpenCL.c:6-91, 612-668): every line is tested against every
 * keyword with strstr, values are read after the first token, later lines override earlier
 * ones and unknown lines are ignored.  Consequences kept on purpose: "ovito_xyz_video_freq"
 * sets video_freq, "cpu_only_mode" is ignored, the keyword "fluid_ouput_spacing" keeps its
 * spelling, "particle_collision_params" reads a single float.
 *
 * One deliberate difference: "maintain_shear_rate" is typed as a 3-vector in the reference,
 * which copies 12 bytes into the 4-byte MaintainShear and overwrites ViscosityModel and
 * NumParticles with uninitialised stack values (D3Q19-OpenCL.c:45, 636-641).  That is
 * undefined behaviour; here only MaintainShear is written.  Every shipped input sets
 * viscosity_model, and NumParticles' own default follows in the table, so results agree.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "d3q19_host.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

enum { T_INT, T_FLOAT, T_INT3, T_FLOAT3, T_FLOAT4, T_STRING, T_INT1OF3, T_LL, T_FLOAT2 };

typedef struct {
    const char* key;
    int type;
    size_t off;   /* offset into the owning struct */
    int owner;    /* 0 int_param_struct, 1 flp_param_struct, 2 host_param_struct */
    const char* def;
} keyword_t;

#define IP(f) offsetof(int_param_struct, f), 0
#define FP(f) offsetof(flp_param_struct, f), 1
#define HP(f) offsetof(host_param_struct, f), 2

static const keyword_t KEYS[] = {
    /* reference keywords and defaults, D3Q19-OpenCL.c:31-64 */
    {"iterations", T_INT, IP(MaxIterations), "1000"},
    {"console_print_freq", T_INT, HP(ConsolePrintFreq), "10"},
    {"constant_body_force", T_FLOAT3, FP(ConstBodyForce), "0.0 0.0 0.0"},
    {"newtonian_tau", T_FLOAT, FP(NewtonianTau), "1.0"},
    {"viscosity_model", T_INT, IP(ViscosityModel), "0"},
    {"viscosity_params", T_FLOAT4, FP(ViscosityParams), "0.0 0.0 0.0 0.0"},
    {"total_lattice_size", T_INT3, IP(LatticeSize), "32 32 32"},
    {"domain_decomposition", T_INT3, HP(DomainDecomp), "2 2 2"},
    {"lattice_buffer_size", T_INT3, IP(BufferSize), "1 1 1"},
    {"initial_f", T_STRING, HP(InitialDist), "zero"},
    {"initial_vel", T_FLOAT3, HP(InitialVel), "0.0 0.0 0.0"},
    {"boundary_conditions_xyz", T_INT3, IP(BoundaryConds), "0 0 0"},
    {"tangential_vel_bcs", T_INT3, HP(TangentialVelBC), "0 0 0"},
    {"maintain_shear_rate", T_INT1OF3, IP(MaintainShear), "0"},
    {"velocity_bc_upper", T_FLOAT3, FP(VelUpper), "0.0 0.0 0.0"},
    {"velocity_bc_lower", T_FLOAT3, FP(VelLower), "0.0 0.0 0.0"},
    {"num_particles", T_INT, IP(NumParticles), "0"},
    {"initial_particle_distribution", T_INT, HP(InitialParticleDistribution), "1"},
    {"random_particle_shift", T_FLOAT, HP(RandParticleShift), "0.0"},
    {"initial_particle_buffer", T_FLOAT, HP(ParticleBuffer), "4.0"},
    {"z_wall_particle_buffer", T_FLOAT, FP(ParticleZBuffer), "10.0"},
    {"particle_diameter", T_FLOAT, FP(ParticleDiam), "8.0"},
    {"particle_density", T_FLOAT, HP(ParticleDensity), "1.0"},
    {"particle_collision_model", T_INT, IP(ParForceModel), "1"},
    {"particle_collision_params", T_FLOAT, FP(ParForceParams), "1.0, 0.0"},
    {"surf_point_write_atomic", T_INT, IP(MaxSurfPointsPerNode), "8"},
    {"ibm_interpolation_mode", T_INT, HP(InterpOrderIBM), "1"},
    {"direct_forcing_coeff", T_FLOAT, FP(DirectForcingCoeff), "1.0"},
    {"rebuild_neigh_list_freq", T_INT, HP(RebuildFreq), "10"},
    {"video_freq", T_INT, HP(VideoFreq), "1000"},
    {"shear_stress_freq", T_INT, HP(ShearStressFreq), "1000"},
    {"fluid_ouput_spacing", T_INT, HP(FluidOutputSpacing), "1"},
    /* additions for the benchmark configurations */
    {"marker_count", T_INT, HP(MarkerCount), "0"},
    {"placement_seed", T_LL, HP(PlacementSeed), "1"},
    {"placement_gap", T_FLOAT, HP(PlacementGap), "2.0"},
    {"collision_operator", T_INT, HP(CollisionModel), "0"},
    {"local_relaxation", T_INT, HP(VariableViscosity), "0"},
    {"tau_clamp", T_FLOAT2, HP(TauClamp), "0.0 0.0"},
    {"ibm_kernel", T_INT, HP(IbmKernel), "0"},
    {"population_storage", T_INT, HP(Storage), "0"},
    {"mass_model", T_INT, HP(MassModel), "0"},
    {"contact_detection", T_INT, HP(ContactDetection), "0"},
    {"population_rounding", T_INT, HP(Rounding), "0"},
    {"marker_directory", T_STRING, HP(DataDir), "unit_sphere_partitions"},
};
#define NKEYS ((int)(sizeof(KEYS) / sizeof(KEYS[0])))

static void* target(const keyword_t* k, int_param_struct* ip, flp_param_struct* fp, host_param_struct* hp) {
    char* base = k->owner == 0 ? (char*)ip : (k->owner == 1 ? (char*)fp : (char*)hp);
    return base + k->off;
}

void host_process_input_line(const char* line, int_param_struct* ip, flp_param_struct* fp, host_param_struct* hp) {
    for (int l = 0; l < NKEYS; l++) {
        const keyword_t* k = &KEYS[l];
        if (!strstr(line, k->key)) continue;
        void* dst = target(k, ip, fp, hp);
        switch (k->type) {
        case T_INT: {
            int v;
            if (sscanf(line, "%*s %d", &v) == 1) *(int*)dst = v;
            break;
        }
        case T_LL: {
            long long v;
            if (sscanf(line, "%*s %lld", &v) == 1) *(long long*)dst = v;
            break;
        }
        case T_FLOAT: {
            float v;
            if (sscanf(line, "%*s %f", &v) == 1) *(float*)dst = v;
            break;
        }
        case T_INT1OF3: {
            int v[3];
            if (sscanf(line, "%*s %d %d %d", v, v + 1, v + 2) >= 1) *(int*)dst = v[0];
            break;
        }
        case T_INT3: {
            int v[3];
            int n = sscanf(line, "%*s %d %d %d", v, v + 1, v + 2);
            for (int i = 0; i < n && i < 3; i++) ((int*)dst)[i] = v[i];
            break;
        }
        case T_FLOAT2: {
            float v[2];
            int n = sscanf(line, "%*s %f %f", v, v + 1);
            for (int i = 0; i < n && i < 2; i++) ((float*)dst)[i] = v[i];
            break;
        }
        case T_FLOAT3: {
            float v[3];
            int n = sscanf(line, "%*s %f %f %f", v, v + 1, v + 2);
            for (int i = 0; i < n && i < 3; i++) ((float*)dst)[i] = v[i];
            break;
        }
        case T_FLOAT4: {
            float v[4];
            int n = sscanf(line, "%*s %f %f %f %f", v, v + 1, v + 2, v + 3);
            for (int i = 0; i < n && i < 4; i++) ((float*)dst)[i] = v[i];
            break;
        }
        case T_STRING: {
            char v[HOST_WORD];
            if (sscanf(line, "%*s %127s", v) == 1) {
                size_t cap = (k->off == offsetof(host_param_struct, DataDir)) ? sizeof(hp->DataDir) : HOST_WORD;
                strncpy((char*)dst, v, cap - 1);
                ((char*)dst)[cap - 1] = 0;
            }
            break;
        }
        }
    }
}

int host_initialize_data(const char* path, int_param_struct* ip, flp_param_struct* fp, host_param_struct* hp) {
    memset(ip, 0, sizeof(*ip));
    memset(fp, 0, sizeof(*fp));
    memset(hp, 0, sizeof(*hp));
    /* velocity set and weights, D3Q19-OpenCL.c:9-25 */
    static const int c[19][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
                                 {1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {1, 0, -1}, {-1, 1, 0}, {-1, -1, 0},
                                 {-1, 0, 1}, {-1, 0, -1}, {0, 1, 1}, {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};
    for (int q = 0; q < 19; q++) {
        fp->EqWeights[q] = q == 0 ? 1.0f / 3.0f : (q < 7 ? 1.0f / 18.0f : 1.0f / 36.0f);
        for (int d = 0; d < 3; d++) ip->BasisVel[q][d] = c[q][d];
    }
    for (int l = 0; l < NKEYS; l++) {
        char line[2 * HOST_WORD + 2];
        snprintf(line, sizeof(line), "%s %s", KEYS[l].key, KEYS[l].def);
        host_process_input_line(line, ip, fp, hp);
    }
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    char line[128];
    while (fgets(line, sizeof(line), f) != NULL) host_process_input_line(line, ip, fp, hp);
    fclose(f);
    return 0;
}

int host_parameter_checking(const int_param_struct* ip, const flp_param_struct* fp, const host_param_struct* hp) {
    (void)fp;
    int errors = 0;
    /* host_sphere_discretization fills 4 floats per marker into a buffer sized for the reference's
     * largest automatic marker set (1024, D3Q19-OpenCL.c:564-571) */
    if (hp->MarkerCount < 0 || hp->MarkerCount > HOST_MAX_MARKERS) {
        fprintf(stderr, "Error: marker_count must be between 0 (automatic) and %d.\n", HOST_MAX_MARKERS);
        errors++;
    }
    if (ip->BoundaryConds[0] + ip->BoundaryConds[1] + ip->BoundaryConds[2] > 1) {
        fprintf(stderr, "Error: More than 1 pair of faces with velocity boundaries not yet supported.\n");
        errors++;
    }
    for (int d = 0; d < 3; d++)
        if (ip->BufferSize[d] != 1) {
            fprintf(stderr, "Error: lattice_buffer_size must be 1 (the periodic map assumes it).\n");
            errors++;
            break;
        }
    for (int d = 0; d < 3; d++)
        if (ip->LatticeSize[d] < 5) {
            fprintf(stderr, "Error: total_lattice_size must be at least 5 on every axis.\n");
            errors++;
            break;
        }
    return errors;
}

void host_system_size(int_param_struct* ip) {
    for (int d = 0; d < 3; d++)
        ip->SystemSize[d] = ip->BoundaryConds[d] == 0 ? ip->LatticeSize[d] - 2 * ip->BufferSize[d]
                                                      : ip->LatticeSize[d] - 2 * ip->BufferSize[d] - 1;
}

int host_sphere_discretization(int_param_struct* ip, flp_param_struct* fp, const host_param_struct* hp, float* points) {
    float d = fp->ParticleDiam;
    float area = (float)(M_PI * d * d);
    int n = 1;
    if (hp->MarkerCount > 0) {
        if (hp->MarkerCount > HOST_MAX_MARKERS) {
            fprintf(stderr, "Error: marker_count above %d\n", HOST_MAX_MARKERS);
            return -1;
        }
        n = hp->MarkerCount;
    } else {
        int p2 = (int)ceil(log2(area));
        for (int i = 0; i < p2; i++) n *= 2;
        n = n < 128 ? 128 : n;
        n = n > 1024 ? 1024 : n;
    }
    ip->PointsPerParticle = n;
    ip->TotalSurfPoints = n * ip->NumParticles;
    fp->PointArea = area / (float)n;
    char fname[512];
    snprintf(fname, sizeof(fname), "%s/unit_sphere_partition_%d.txt", hp->DataDir, n);
    FILE* fs = fopen(fname, "r");
    if (!fs) {
        fprintf(stderr, "Failed opening sphere discretization file %s\n", fname);
        return -1;
    }
    for (int m = 0; m < n; m++) {
        float px, py, pz;
        if (fscanf(fs, "%f,%f,%f\n", &px, &py, &pz) != 3) {
            fclose(fs);
            fprintf(stderr, "Error: short sphere discretization file %s\n", fname);
            return -1;
        }
        float l2 = px * px + py * py + pz * pz;
        if (!(l2 < 1.01 && l2 > 0.99)) {
            fclose(fs);
            fprintf(stderr, "Error: malformed coordinates read in sphere discretization\n");
            return -1;
        }
        points[4 * m + 0] = px * 0.5f * d;
        points[4 * m + 1] = py * 0.5f * d;
        points[4 * m + 2] = pz * 0.5f * d;
        points[4 * m + 3] = 0.0f;
    }
    fclose(fs);
    return 0;
}

void host_workgroup_split(int_param_struct* ip, int max_wg) {
    int p2 = (int)ceil(log2((double)max_wg) - 1e-6);
    int ws = 1;
    for (int i = 0; i < p2; i++) ws *= 2;
    ip->PointsPerWorkGroup = ws > ip->PointsPerParticle ? ip->PointsPerParticle : ws;
    if (ip->PointsPerWorkGroup < 1) ip->PointsPerWorkGroup = 1;
    ip->WorkGroupsPerParticle = ip->PointsPerParticle / ip->PointsPerWorkGroup;
    if (ip->WorkGroupsPerParticle < 1) ip->WorkGroupsPerParticle = 1;
}

void host_equilibrium(float rho, const float* vel, float* feq) {
    static const int c[19][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
                                 {1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {1, 0, -1}, {-1, 1, 0}, {-1, -1, 0},
                                 {-1, 0, 1}, {-1, 0, -1}, {0, 1, 1}, {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};
    float vsq = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
    feq[0] = (rho / 3.0f) * (1.0f - 1.5f * vsq);
    for (int q = 1; q < 19; q++) {
        float cu = c[q][0] * vel[0] + c[q][1] * vel[1] + c[q][2] * vel[2];
        float pre = q < 7 ? rho / 18.0f : rho / 36.0f;
        feq[q] = pre * (1.0f + 3.0f * cu + 4.5f * cu * cu - 1.5f * vsq);
    }
}

void host_initialize_lattice_fields(const host_param_struct* hp, int_param_struct* ip, const flp_param_struct* fp,
                                    float* f, float* u, float* tau) {
    host_system_size(ip);
    long nn = (long)ip->LatticeSize[0] * ip->LatticeSize[1] * ip->LatticeSize[2];
    float vel[3] = {0, 0, 0};
    if (strstr(hp->InitialDist, "constant")) {
        vel[0] = hp->InitialVel[0];
        vel[1] = hp->InitialVel[1];
        vel[2] = hp->InitialVel[2];
    }
    float feq[19];
    host_equilibrium(1.0f, vel, feq);
    for (long i = 0; i < nn; i++) {
        if (f)
            for (int q = 0; q < 19; q++) f[i + q * nn] = feq[q];
        if (u) {
            u[i] = vel[0];
            u[i + nn] = vel[1];
            u[i + 2 * nn] = vel[2];
        }
        if (tau) tau[i] = fp->NewtonianTau;
    }
}

unsigned long long host_splitmix64(unsigned long long* state) {
    unsigned long long z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double uniform01(unsigned long long* st) { return (double)(host_splitmix64(st) >> 11) * (1.0 / 9007199254740992.0); }

int host_initialize_particle_fields(const host_param_struct* hp, const int_param_struct* ip, flp_param_struct* fp,
                                    float* pk, float* pf, float* pff) {
    int np = ip->NumParticles;
    double d = fp->ParticleDiam;
    if (hp->MassModel == 1)
        fp->ParticleMass = (float)((M_PI / 6.0) * hp->ParticleDensity * d * d * d);
    else
        fp->ParticleMass = (float)((4.0f / 3.0f) * M_PI * hp->ParticleDensity * pow(fp->ParticleDiam, 3));
    fp->ParticleMomInertia = 0.1f * fp->ParticleMass * fp->ParticleDiam * fp->ParticleDiam;
#define PKS(p, k, c) pk[4 * ((p) + (k) * np) + (c)]
    if (np <= 0) return 0;
    if (hp->InitialParticleDistribution == 1) {
        /* D3Q19-OpenCL.c:207-241 */
        float pb = hp->ParticleBuffer;
        int g = (int)ceil(pow(np, 1.0 / 3.0));
        float sp[3];
        for (int i = 0; i < 3; i++) sp[i] = g > 1 ? (ip->SystemSize[i] - 2 * pb) / (float)(g - 1) : 0.0f;
        for (int p = 0; p < np; p++) {
            PKS(p, 0, 0) = pb + sp[0] * (p / (g * g));
            PKS(p, 0, 1) = pb + sp[1] * ((p / g) % g);
            PKS(p, 0, 2) = pb + sp[2] * (p % g);
            PKS(p, 0, 3) = 0;
            for (int c = 0; c < 4; c++) PKS(p, 1, c) = 0;
        }
    } else if (hp->InitialParticleDistribution == 2) {
        /* D3Q19-OpenCL.c:242-305.  The random shift uses the C library rand() as the reference
         * does (never seeded); with random_particle_shift 0 the placement is deterministic. */
        float pb = hp->ParticleBuffer;
        float wN = ip->SystemSize[0] - 2 * pb;
        float wM = ip->SystemSize[2] - 2 * (pb + fp->ParticleZBuffer);
        int mdn = (int)ceil(wM / wN);
        int npl = 0, n = 0;
        while (npl < np) {
            n++;
            npl = n * n * n * mdn;
        }
        int m = n * mdn;
        float sp[3];
        if (np > 1) {
            sp[0] = (ip->SystemSize[0] - 2.0f * pb) / (n - 1.0f);
            sp[1] = (ip->SystemSize[1] - 2.0f * pb) / (n - 1.0f);
            sp[2] = (ip->SystemSize[2] - 2.0f * (pb + fp->ParticleZBuffer)) / (n * mdn - 1.0f);
        } else {
            sp[0] = sp[1] = sp[2] = 0.0f;
        }
        for (int p = 0; p < np; p++) {
            int tries = 0;
            for (;;) {
                float px = pb + sp[0] * (p / (n * m)) + hp->RandParticleShift * (2 * rand() / (float)RAND_MAX - 1);
                float py = pb + sp[1] * ((p / m) % n) + hp->RandParticleShift * (2 * rand() / (float)RAND_MAX - 1);
                float pz = pb + fp->ParticleZBuffer + sp[2] * (p % m) + hp->RandParticleShift * (2 * rand() / (float)RAND_MAX - 1);
                int safe = 1;
                for (int p2 = 0; p2 < p; p2++) {
                    float rx = PKS(p2, 0, 0) - px, ry = PKS(p2, 0, 1) - py, rz = PKS(p2, 0, 2) - pz;
                    if (rx * rx + ry * ry + rz * rz < fp->ParticleDiam * fp->ParticleDiam) safe = 0;
                }
                if (safe) {
                    PKS(p, 0, 0) = px; PKS(p, 0, 1) = py; PKS(p, 0, 2) = pz; PKS(p, 0, 3) = 0;
                    for (int c = 0; c < 4; c++) PKS(p, 1, c) = 0;
                    break;
                }
                if (++tries > 100000) return -1;
            }
        }
    } else if (hp->InitialParticleDistribution == 3) {
        /* D3Q19-OpenCL.c:306-321: the reference writes parKin[1] (the second float4 of the
         * position block) as the velocity slot; for a single particle that is vel[0]. */
        PKS(0, 0, 0) = hp->ParticleBuffer;
        PKS(0, 0, 1) = ((float)ip->SystemSize[1]) / 2.0f;
        PKS(0, 0, 2) = ((float)ip->SystemSize[2]) / 2.0f;
        PKS(0, 0, 3) = 0;
        for (int c = 0; c < 4; c++) pk[4 * 1 + c] = 0;
    } else if (hp->InitialParticleDistribution == 4) {
        /* seeded uniform-random centres (SplitMix64), rejected if any surface gap to another
         * sphere or to a wall is below placement_gap */
        unsigned long long st = (unsigned long long)hp->PlacementSeed;
        double r = 0.5 * d, gap = hp->PlacementGap;
        for (int p = 0; p < np; p++) {
            int tries = 0;
            for (;;) {
                double x[3];
                int ok = 1;
                for (int a = 0; a < 3; a++) {
                    double L = ip->SystemSize[a];
                    if (ip->BoundaryConds[a] == 1) {
                        double lo = r + gap, hi = L - r - gap;
                        if (hi <= lo) return -1;
                        x[a] = lo + (hi - lo) * uniform01(&st);
                    } else {
                        x[a] = L * uniform01(&st);
                    }
                }
                for (int p2 = 0; p2 < p && ok; p2++) {
                    double s2 = 0;
                    for (int a = 0; a < 3; a++) {
                        double dx = x[a] - PKS(p2, 0, a);
                        double L = ip->SystemSize[a];
                        if (ip->BoundaryConds[a] == 0) dx -= L * floor(dx / L + 0.5);
                        s2 += dx * dx;
                    }
                    if (s2 < (d + gap) * (d + gap)) ok = 0;
                }
                if (ok) {
                    for (int a = 0; a < 3; a++) PKS(p, 0, a) = (float)x[a];
                    PKS(p, 0, 3) = 0;
                    for (int c = 0; c < 4; c++) PKS(p, 1, c) = 0;
                    break;
                }
                if (++tries > 200000) return -1;
            }
        }
    } else {
        fprintf(stderr, "Error: initial_particle_distribution option not a known value\n");
        return -1;
    }
    int nfa = ip->WorkGroupsPerParticle > 0 ? ip->WorkGroupsPerParticle : 1;
    for (int p = 0; p < np; p++) {
        PKS(p, 2, 0) = 0; PKS(p, 2, 1) = 0; PKS(p, 2, 2) = 0; PKS(p, 2, 3) = 1.0f;
        for (int c = 0; c < 4; c++) PKS(p, 3, c) = 0;
        if (pf)
            for (int c = 0; c < 4; c++) { pf[4 * p + c] = 0; pf[4 * (p + np) + c] = 0; }
        if (pff)
            for (int fa = 0; fa < nfa; fa++)
                for (int c = 0; c < 4; c++) {
                    pff[4 * (p + np * (2 * fa)) + c] = 0;
                    pff[4 * (p + np * (2 * fa + 1)) + c] = 0;
                }
    }
#undef PKS
    return 0;
}

void host_zone_geometry(const host_param_struct* hp, int_param_struct* ip, flp_param_struct* fp) {
    /* D3Q19-OpenCL.c:348-375 */
    float vmax = 0.05f;
    for (int i = 0; i < 3; i++) {
        vmax = vmax < fabsf(fp->VelUpper[i]) ? fabsf(fp->VelUpper[i]) : vmax;
        vmax = vmax < fabsf(fp->VelLower[i]) ? fabsf(fp->VelLower[i]) : vmax;
        float nu = (fp->NewtonianTau - 0.5f) / 3.0f;
        float vn = fp->ConstBodyForce[i] * (ip->SystemSize[i]) * (ip->SystemSize[i]) / (12.0f * nu);
        vmax = vmax < vn ? vn : vmax;
    }
    float minw = 2 * (vmax * hp->RebuildFreq) + fp->ParticleDiam;
    for (int i = 0; i < 3; i++) {
        float w = (float)ip->SystemSize[i];
        ip->NumZones[i] = floor(w / minw) > 0 ? (int)floor(w / minw) : 1;
        fp->ZoneWidth[i] = w / (float)ip->NumZones[i];
    }
}

void host_fill_options(const host_param_struct* hp, d3q19_options* opt) {
    memset(opt, 0, sizeof(*opt));
    opt->collision = hp->CollisionModel;
    opt->variable_viscosity = hp->VariableViscosity;
    opt->tau_min = hp->TauClamp[0];
    opt->tau_max = hp->TauClamp[1];
    opt->ibm_kernel = hp->IbmKernel;
    opt->storage = hp->Storage;
    opt->calc_rho = 0;
    for (int d = 0; d < 3; d++) opt->tangential_bc[d] = hp->TangentialVelBC[d];
    opt->contact_detection = hp->ContactDetection;
    opt->rebuild_freq = hp->RebuildFreq;
    opt->rounding = hp->Rounding;
}
