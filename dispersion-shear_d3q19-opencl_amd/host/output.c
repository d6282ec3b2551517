/*
 * output.c -- observables and writers of the C host: the plane-mean velocity profile and wall /
 * particle-region shear rates (compute_shear_stress, D3Q19-OpenCL.c:670-750), the OVITO XYZ
 * video frame (continuous_output, 752-792) and the final velocity field and centre plane
 * (write_lattice_field, 794-835).  File formats are kept so the reference's MATLAB scripts read
 * them unchanged.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "d3q19_host.h"

static FILE* open_in(const char* dir, const char* name, const char* mode) {
    char path[1024];
    if (dir && dir[0])
        snprintf(path, sizeof(path), "%s/%s", dir, name);
    else
        snprintf(path, sizeof(path), "%s", name);
    return fopen(path, mode);
}

void host_compute_shear_stress(output_data_struct* od, const host_param_struct* hp, const int_param_struct* ip,
                               const flp_param_struct* fp, const float* u, const char* dir, int quiet) {
    (void)hp;
    int nx = ip->LatticeSize[0], ny = ip->LatticeSize[1], nz = ip->LatticeSize[2];
    long nL = (long)nx * ny * nz;
    FILE* fp_out = open_in(dir, "velocity_profile_z.txt", "w");
    const int gradBuffer = 1;
    float nxy = (float)(nx - 2) * (ny - 2);
    float* umx = (float*)calloc((size_t)nz, sizeof(float));
    float dudz = 0.0f;
    const float dz = 4.0f;
    for (int z = 1; z < nz - 1; z++) {
        float um[3] = {0.0f, 0.0f, 0.0f};
        for (int x = 1; x < nx - 1; x++)
            for (int y = 1; y < ny - 1; y++) {
                long i = x + (long)nx * (y + (long)ny * z);
                um[0] += u[i];
                um[1] += u[i + nL];
                um[2] += u[i + 2 * nL];
                if (z == (1 + dz + gradBuffer)) {
                    long j = x + (long)nx * (y + (long)ny * (z - (int)dz));
                    dudz += (u[i] - u[j]) / dz;
                } else if (z == (nz - 2 - gradBuffer)) {
                    long j = x + (long)nx * (y + (long)ny * (z - (int)dz));
                    dudz += (u[i] - u[j]) / dz;
                }
            }
        um[0] /= nxy;
        umx[z] = um[0];
        um[1] /= nxy;
        um[2] /= nxy;
        if (fp_out) fprintf(fp_out, "%8.6f %8.6f %8.6f\n", um[0], um[1], um[2]);
    }
    dudz /= (2.0 * nxy);
    int count = ++(od->ShearStressCount);
    od->ShearStressAvg = (od->ShearStressAvg * (count - 1) + dudz) / count;
    if (!quiet) printf("Shear rate at wall = %e\n", od->ShearStressAvg);
    int ib = (int)(fp->ParticleZBuffer + 1E-8) + (int)(fp->ParticleDiam / 2);
    int zbL = 1 + ib;
    int zbU = (int)(nz - 2 - ib);
    float actual = 0.0f;
    if (zbL >= 0 && zbU < nz && zbU != zbL) actual = (umx[zbU] - umx[zbL]) / (zbU - zbL);
    od->ActualShearRate = (od->ActualShearRate * (count - 1) + actual) / count;
    if (!quiet) printf("Shear rate in particle region = %e\n", od->ActualShearRate);
    free(umx);
    if (fp_out) fclose(fp_out);
}

void host_continuous_output(const host_param_struct* hp, const int_param_struct* ip, const float* u,
                            const float* pk, FILE* vid, int frame) {
    int nx = ip->LatticeSize[0], ny = ip->LatticeSize[1], nz = ip->LatticeSize[2];
    int ns = hp->FluidOutputSpacing > 0 ? hp->FluidOutputSpacing : 1;
    long nL = (long)nx * ny * nz;
    int nfluid = (int)(1 + (nx - 3) / ns) * (1 + (ny - 3) / ns) * (1 + (nz - 3) / ns);
    int npar = ip->NumParticles;
    fprintf(vid, "%d\n", nfluid + npar);
    fprintf(vid, "D3Q19_output, frame %d\n", frame);
    for (int x = 1; x < nx - 1; x += ns)
        for (int y = 1; y < ny - 1; y += ns)
            for (int z = 1; z < nz - 1; z += ns) {
                long i = x + (long)nx * (y + (long)ny * z);
                fprintf(vid, "1 ");
                fprintf(vid, "%d %d %d ", x - 1, y - 1, z - 1);
                fprintf(vid, "%8.6f %8.6f %8.6f\n", u[i], u[i + nL], u[i + 2 * nL]);
            }
    for (int p = 0; p < npar; p++) {
        fprintf(vid, "2 ");
        fprintf(vid, "%.2f %.2f %.2f ", pk[4 * p], pk[4 * p + 1], pk[4 * p + 2]);
        fprintf(vid, "%8.6f %8.6f %8.6f\n", pk[4 * (p + npar)], pk[4 * (p + npar) + 1], pk[4 * (p + npar) + 2]);
    }
}

int host_write_lattice_field(const float* u, const int_param_struct* ip, const char* dir) {
    int nx = ip->LatticeSize[0], ny = ip->LatticeSize[1], nz = ip->LatticeSize[2];
    long nC = (long)nx * ny * nz;
    FILE* f = open_in(dir, "velocity_field_final.txt", "w");
    if (!f) return -1;
    for (int x = 1; x < nx - 1; x++)
        for (int y = 1; y < ny - 1; y++)
            for (int z = 1; z < nz - 1; z++) {
                long i = x + (long)nx * (y + (long)ny * z);
                fprintf(f, "%d %d %d ", x, y, z);
                fprintf(f, "%9.7f %9.7f %9.7f\n", u[i], u[i + nC], u[i + 2 * nC]);
            }
    fclose(f);
    FILE* f2 = open_in(dir, "matlab_postproc/velocity_field_centerline.txt", "w");
    if (!f2) return -1;
    int y = (int)floor((float)ny / 2.0f);
    for (int x = 1; x < nx - 1; x++)
        for (int z = 1; z < nz - 1; z++) {
            long i = x + (long)nx * (y + (long)ny * z);
            fprintf(f2, "%d %d %d ", x, y, z);
            fprintf(f2, "%9.7f %9.7f %9.7f\n", u[i], u[i + nC], u[i + 2 * nC]);
        }
    fclose(f2);
    return 0;
}
