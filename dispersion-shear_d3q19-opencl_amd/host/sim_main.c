/*
 * sim_main.c -- C host program of the D3Q19 IBM shear-cell solver (the reference's main,
 * sim_main.c:1-527, over the C-ABI of include/d3q19_abi.h instead of OpenCL).
 *
 * usage: d3q19_sim [input_file] [output_dir]
 *   input_file  defaults to input_file.txt (the reference reads that name, D3Q19-OpenCL.c:29)
 *   output_dir  where the reference's output files are written (default: current directory)
 *
 * Outputs as the reference: xyz_ovito_output.txt (every video_freq steps), velocity_profile_z.txt
 * and the shear-rate prints (t > 3/4 of the run, every shear_stress_freq steps),
 * velocity_field_final.txt and matlab_postproc/velocity_field_centerline.txt, and the final
 * force print of sphere 0.  Errors exit with a non-zero status (the reference exits 0).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include "d3q19_host.h"

static int check(int rc, const char* where) {
    if (rc != 0) {
        fprintf(stderr, "Error: %s failed: %s (%s)\n", where, d3q19_strerror(rc), d3q19_last_error());
        exit(EXIT_FAILURE);
    }
    return rc;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char* argv[]) {
    const char* input = argc > 1 ? argv[1] : "input_file.txt";
    const char* outdir = argc > 2 ? argv[2] : "";
    int_param_struct ip;
    flp_param_struct fp;
    host_param_struct hp;
    output_data_struct od = {0, 0.0f, 0.0f};
    printf("Int struct size: %lu\n", (unsigned long)sizeof(ip));
    printf("Flp struct size: %lu\n", (unsigned long)sizeof(fp));

    int ndev = 0;
    check(d3q19_device_count(&ndev), "d3q19_device_count");
    if (ndev == 0) {
        fprintf(stderr, "Error: No GPU found\n");
        return EXIT_FAILURE;
    }
    char name[256];
    if (d3q19_device_name(0, name, sizeof(name)) == 0) printf("Device 0: %s\n", name);

    if (host_initialize_data(input, &ip, &fp, &hp) != 0) {
        fprintf(stderr, "Error: cannot open %s\n", input);
        return EXIT_FAILURE;
    }
    printf("ViscosityModel =  %d \n", ip.ViscosityModel);
    if (host_parameter_checking(&ip, &fp, &hp) > 0) return EXIT_FAILURE;
    host_system_size(&ip);

    float* points = (float*)calloc(4 * HOST_MAX_MARKERS, sizeof(float));
    if (host_sphere_discretization(&ip, &fp, &hp, points) != 0) return EXIT_FAILURE;
    host_workgroup_split(&ip, 256);

    long nn = (long)ip.LatticeSize[0] * ip.LatticeSize[1] * ip.LatticeSize[2];
    int np = ip.NumParticles;
    int nfa = ip.WorkGroupsPerParticle > 0 ? ip.WorkGroupsPerParticle : 1;
    float* f = (float*)malloc(sizeof(float) * 19 * nn);
    float* u = (float*)malloc(sizeof(float) * 3 * nn);
    float* tau = (float*)malloc(sizeof(float) * nn);
    float* pk = (float*)calloc(16 * (size_t)(np > 0 ? np : 1), sizeof(float));
    float* pf = (float*)calloc(8 * (size_t)(np > 0 ? np : 1), sizeof(float));
    float* pff = (float*)calloc(8 * (size_t)(np > 0 ? np : 1) * nfa, sizeof(float));
    if (!f || !u || !tau || !pk || !pf || !pff) {
        fprintf(stderr, "Error: host allocation failed\n");
        return EXIT_FAILURE;
    }
    host_initialize_lattice_fields(&hp, &ip, &fp, f, u, tau);
    if (host_initialize_particle_fields(&hp, &ip, &fp, pk, pf, pff) != 0) return EXIT_FAILURE;
    host_zone_geometry(&hp, &ip, &fp);
    printf("\nThere are %d particles.\nMass = %f, moment of inertia = %f\n", np, fp.ParticleMass, fp.ParticleMomInertia);

    d3q19_options opt;
    host_fill_options(&hp, &opt);
    d3q19_sim* sim = NULL;
    check(d3q19_create(0, &ip, &fp, &opt, (const d3q19_float4*)points, &sim), "d3q19_create");
    check(d3q19_write_f(sim, f), "d3q19_write_f");
    check(d3q19_write_u(sim, u), "d3q19_write_u");
    check(d3q19_write_tau(sim, tau), "d3q19_write_tau");
    if (np > 0) check(d3q19_write_particles(sim, (d3q19_float4*)pk, (d3q19_float4*)pf, (d3q19_float4*)pff), "d3q19_write_particles");

    char path[1024];
    if (outdir[0]) {
        mkdir(outdir, 0755);
        snprintf(path, sizeof(path), "%s/matlab_postproc", outdir);
        mkdir(path, 0755);
        snprintf(path, sizeof(path), "%s/xyz_ovito_output.txt", outdir);
    } else {
        mkdir("matlab_postproc", 0755);
        snprintf(path, sizeof(path), "xyz_ovito_output.txt");
    }
    FILE* vid = fopen(path, "w");
    printf("Starting iteration 1, maximum iterations %d\n", ip.MaxIterations);
    double t0 = now_s();
    const int maxit = ip.MaxIterations;
    int t = 1;
    while (t <= maxit) {
        /* next step that needs a read-back: video frame, shear sample or the final field */
        int next = maxit;
        for (int s = t; s <= maxit; s++) {
            if ((hp.VideoFreq > 0 && s % hp.VideoFreq == 0) ||
                (s > 3 * maxit / 4 && hp.ShearStressFreq > 0 && s % hp.ShearStressFreq == 0)) {
                next = s;
                break;
            }
        }
        int nsteps = next - t + 1;
        /* console progress as sim_main.c:317-321, coarsened to the read-back points */
        if (hp.ConsolePrintFreq > 0) printf("Starting iteration %d\n", t);
        check(d3q19_run(sim, t, nsteps, D3Q19_STEP_WRITE_U), "d3q19_run");
        t = next;
        if (hp.VideoFreq > 0 && t % hp.VideoFreq == 0 && vid) {
            check(d3q19_read_u(sim, u), "d3q19_read_u");
            check(d3q19_read_particles(sim, (d3q19_float4*)pk, NULL, NULL), "d3q19_read_particles");
            host_continuous_output(&hp, &ip, u, pk, vid, t);
        }
        if (t > 3 * maxit / 4 && hp.ShearStressFreq > 0 && t % hp.ShearStressFreq == 0) {
            check(d3q19_read_u(sim, u), "d3q19_read_u");
            host_compute_shear_stress(&od, &hp, &ip, &fp, u, outdir, 0);
        }
        int fin = d3q19_check_finite(sim);
        if (fin == D3Q19_E_NONFINITE) {
            fprintf(stderr, "Error: non-finite density at step %d\n", t);
            return EXIT_FAILURE;
        }
        check(fin, "d3q19_check_finite");
        t++;
    }
    check(d3q19_sync(sim), "d3q19_sync");
    double el = now_s() - t0;
    long long lu = (long long)(ip.LatticeSize[0] - 2) * (ip.LatticeSize[1] - 2) * (ip.LatticeSize[2] - 2) * (long long)maxit;
    printf("Checkpoint: end of simulation loop (%.3f s, %.1f MLUPS)\n", el, lu / el / 1e6);
    if (vid) fclose(vid);

    check(d3q19_read_u(sim, u), "d3q19_read_u");
    host_write_lattice_field(u, &ip, outdir);
    if (np > 0) {
        check(d3q19_read_particles(sim, NULL, NULL, (d3q19_float4*)pff), "d3q19_read_particles");
        float ff[3] = {0, 0, 0};
        for (int fa = 0; fa < nfa; fa++) {
            ff[0] += pff[4 * fa];
            ff[1] += pff[4 * fa + 1];
            ff[2] += pff[4 * fa + 2];
        }
        printf("Final force on particle 1 = %f %f %f\n", ff[0], ff[1], ff[2]);
    }
    d3q19_destroy(sim);
    printf("Checkpoint: end of sim_main\n");
    free(points); free(f); free(u); free(tau); free(pk); free(pf); free(pff);
    return 0;
}
