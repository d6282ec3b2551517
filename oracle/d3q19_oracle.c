/*
 * d3q19_oracle.c -- CPU restatement of the reference step.  TEST INFRASTRUCTURE ONLY (the
 * checker for the HIP library; see d3q19_oracle.h).
 *
 * Compiled with -DORC_REAL=float -DORC_SFX=_f32 (reference arithmetic) and with
 * -DORC_REAL=double -DORC_SFX=_f64 (FP64 oracle).  Every function cites the reference lines it
 * restates.  Array layouts are the reference's: node i = x + Nx*(y + Ny*z) over LatticeSize
 * (shell included), populations f[q*N_C + i], float4 particle arrays as 4 reals.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "d3q19_oracle.h"

#ifndef ORC_REAL
#define ORC_REAL double
#endif
#ifndef ORC_SFX
#define ORC_SFX _f64
#endif
typedef ORC_REAL real;

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define ORC(name) CAT(CAT(orc_, name), ORC_SFX)

#define IS_FLOAT (sizeof(real) == sizeof(float))
static real r_sqrt(real x) { return IS_FLOAT ? (real)sqrtf((float)x) : (real)sqrt((double)x); }
static real r_pow(real x, real y) { return IS_FLOAT ? (real)powf((float)x, (float)y) : (real)pow((double)x, (double)y); }
static real r_fmod(real x, real y) { return IS_FLOAT ? (real)fmodf((float)x, (float)y) : (real)fmod((double)x, (double)y); }
static real r_floor(real x) { return IS_FLOAT ? (real)floorf((float)x) : (real)floor((double)x); }
static real r_exp(real x) { return IS_FLOAT ? (real)expf((float)x) : (real)exp((double)x); }
static real r_fabs(real x) { return x < 0 ? -x : x; }

/* D3Q19 velocity set in the reference's order (D3Q19-OpenCL.c:9-13). */
static const int C[19][3] = {{0, 0, 0},
                             {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
                             {1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {1, 0, -1},
                             {-1, 1, 0}, {-1, -1, 0}, {-1, 0, 1}, {-1, 0, -1},
                             {0, 1, 1}, {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};

struct orc_sim {
    int_param_struct ip;
    flp_param_struct fp;
    d3q19_options opt;
    long nn;          /* N_C */
    int N[3];
    int np, ppp, ppwg, nfa;
    real* arr[ORC_NARRAYS];
    long len[ORC_NARRAYS];
    real M[19][19];   /* moment matrix */
    real Minv[19][19];
    real* ffsum;      /* per-marker partials for the work-group tree reduction, 2*np*ppp*4 */
    int use_s_override;
    real s_override[19];
    /* contact zones (D3Q19-OpenCL.c:344-463): step counter and the zone lists */
    int t;            /* steps completed */
    int nzt;          /* total zones */
    int* pars_zone;   /* zone of each sphere, np */
    int* zone_members;/* nzt * np */
    int* num_in_zone; /* nzt */
    int* zone_neigh;  /* 28 per zone: count, then neighbour ids */
};
typedef struct orc_sim orc_sim;

/* ---- moment basis -------------------------------------------------------------------------
 * d'Humieres D3Q19 basis in the reference's row order [rho, e, eps, jx, qx, jy, qy, jz, qz,
 * 3pxx, 3pixx, pww, piww, pxy, pyz, pxz, mx, my, mz] (GPU_program.cl:410-449).  Rows are built
 * from their defining polynomials; rows are mutually orthogonal so M^-1 = M^T diag(1/|row|^2)
 * (the reference hard-codes the same inverse to 8 digits, GPU_program.cl:576-614). */
static void build_moments(orc_sim* s) {
    for (int q = 0; q < 19; q++) {
        int cx = C[q][0], cy = C[q][1], cz = C[q][2];
        int c2 = cx * cx + cy * cy + cz * cz;
        int row[19];
        row[0] = 1;
        row[1] = 19 * c2 - 30;
        row[2] = (21 * c2 * c2 - 53 * c2 + 24) / 2;
        row[3] = cx;
        row[4] = (5 * c2 - 9) * cx;
        row[5] = cy;
        row[6] = (5 * c2 - 9) * cy;
        row[7] = cz;
        row[8] = (5 * c2 - 9) * cz;
        row[9] = 3 * cx * cx - c2;
        row[10] = (3 * c2 - 5) * (3 * cx * cx - c2);
        row[11] = cy * cy - cz * cz;
        row[12] = (3 * c2 - 5) * (cy * cy - cz * cz);
        row[13] = cx * cy;
        row[14] = cy * cz;
        row[15] = cx * cz;
        row[16] = (cy * cy - cz * cz) * cx;
        row[17] = (cz * cz - cx * cx) * cy;
        row[18] = (cx * cx - cy * cy) * cz;
        for (int k = 0; k < 19; k++) s->M[k][q] = (real)row[k];
    }
    for (int k = 0; k < 19; k++) {
        double n2 = 0;
        for (int q = 0; q < 19; q++) n2 += (double)s->M[k][q] * (double)s->M[k][q];
        for (int q = 0; q < 19; q++) s->Minv[q][k] = (real)((double)s->M[k][q] / n2);
    }
}

/* ---- equilibrium and Guo term ---------------------------------------------------------------
 * GPU_program.cl:978-1007 (second-order f_eq) and 1010-1038 (Guo forcing before relaxation). */
static void feq_d3q19(real* feq, real rho, real ux, real uy, real uz) {
    real usq = ux * ux + uy * uy + uz * uz;
    for (int q = 0; q < 19; q++) {
        real w = (q == 0) ? (real)(1.0 / 3.0) : (q < 7 ? (real)(1.0 / 18.0) : (real)(1.0 / 36.0));
        real cu = C[q][0] * ux + C[q][1] * uy + C[q][2] * uz;
        /* the reference writes (rho/w^-1)*(1 + 3cu + 4.5cu^2 - 1.5usq) */
        real pre = (q == 0) ? rho / (real)3.0 : (q < 7 ? rho / (real)18.0 : rho / (real)36.0);
        (void)w;
        if (q == 0)
            feq[q] = pre * ((real)1.0 - (real)1.5 * usq);
        else
            feq[q] = pre * ((real)1.0 + (real)3.0 * cu + (real)4.5 * cu * cu - (real)1.5 * usq);
    }
}

static void guo_d3q19(real* fg, real ux, real uy, real uz, real gx, real gy, real gz) {
    real udg = ux * gx + uy * gy + uz * gz;
    fg[0] = -udg;
    for (int q = 1; q < 19; q++) {
        real cg = C[q][0] * gx + C[q][1] * gy + C[q][2] * gz;
        real cu = C[q][0] * ux + C[q][1] * uy + C[q][2] * uz;
        real den = (q < 7) ? (real)6.0 : (real)12.0;
        fg[q] = (cg - udg + (real)3.0 * cu * cg) / den;
    }
}

/* ---- relaxation time -------------------------------------------------------------------------
 * compute_tau, GPU_program.cl:1041-1128.  Extensions (configuration mode only): an upper clamp
 * tau_max, and Papanastasiou regularisation of Casson when ViscosityParams[2] = m > 0:
 * nu = (sqrt(eta) + sqrt(tauY/g)*(1 - exp(-sqrt(m g))))^2, which is the reference's Casson law
 * (GPU_program.cl:1085) as m -> infinity. */
static real compute_tau(const orc_sim* s, real srt) {
    const flp_param_struct* fp = &s->fp;
    real tau = (real)fp->NewtonianTau;
    real eps = (real)1e-8;
    int model = s->ip.ViscosityModel;
    real tmin = s->opt.tau_min > 0 ? (real)s->opt.tau_min : (real)0.505;
    if (model != D3Q19_VISC_POWER_LAW && model != D3Q19_VISC_CASSON && model != D3Q19_VISC_HB)
        return tau;
    srt = srt > eps ? srt : eps;
    real nu;
    if (model == D3Q19_VISC_POWER_LAW) {
        real k = (real)fp->ViscosityParams[0], n = (real)fp->ViscosityParams[1];
        if (n == (real)0.5) nu = k / r_sqrt(srt);
        else if (n == (real)1.0) nu = k;
        else if (n == (real)2.0) nu = k * srt;
        else nu = k * r_pow(srt, n - (real)1.0);
    } else if (model == D3Q19_VISC_CASSON) {
        real ty = (real)fp->ViscosityParams[0], eta = (real)fp->ViscosityParams[1];
        real m = (real)fp->ViscosityParams[2];
        real a = r_sqrt(ty / srt);
        if (m > 0) a = a * ((real)1.0 - r_exp(-r_sqrt(m * srt)));
        nu = (a + r_sqrt(eta)) * (a + r_sqrt(eta));
    } else {
        real ty = (real)fp->ViscosityParams[0], k = (real)fp->ViscosityParams[1], n = (real)fp->ViscosityParams[2];
        if (n == (real)0.5) nu = ty / srt + k / r_sqrt(srt);
        else if (n == (real)1.0) nu = ty / srt + k;
        else if (n == (real)2.0) nu = ty / srt + k * srt;
        else nu = ty / srt + k * r_pow(srt, n - (real)1.0);
    }
    tau = (real)3.0 * nu + (real)0.5;
    if (tau < tmin) tau = tmin;
    if (s->opt.tau_max > 0 && tau > (real)s->opt.tau_max) tau = (real)s->opt.tau_max;
    return tau;
}

/* relaxation of the non-equilibrium part: out = M^-1 S M v  (MRT) or v/tau (BGK) */
static void relax(const orc_sim* s, const real* sv, const real* v, real* out, int bgk, real tau) {
    if (bgk) {
        for (int q = 0; q < 19; q++) out[q] = v[q] / tau;
        return;
    }
    real m[19], sm[19];
    for (int k = 0; k < 19; k++) {
        real acc = 0;
        for (int q = 0; q < 19; q++)
            if (s->M[k][q] != 0) acc += s->M[k][q] * v[q];
        m[k] = acc;
    }
    for (int k = 0; k < 19; k++) sm[k] = sv[k] * m[k];
    for (int q = 0; q < 19; q++) {
        real acc = 0;
        for (int k = 0; k < 19; k++)
            if (s->Minv[q][k] != 0) acc += s->Minv[q][k] * sm[k];
        out[q] = acc;
    }
}

/* ---- collide + push stream -------------------------------------------------------------------
 * collideMRT_stream_D3Q19, GPU_program.cl:316-624; the BGK form is GPU_program.cl:921-945 with
 * the live pipeline's forcing (ConstBodyForce + smoothed spread force).  Variable tau follows the
 * compiled-out branch GPU_program.cl:457-564: the strain rate is formed with the previous tau. */
static void collide_node(orc_sim* s, int x, int y, int z, const real* fc, real* fs) {
    const int Nx = s->N[0], Ny = s->N[1];
    const long nn = s->nn;
    long i = x + (long)Nx * (y + (long)Ny * z);
    real f[19], rho = 0;
    for (int q = 0; q < 19; q++) {
        f[q] = fc[i + q * nn];
        rho += f[q];
    }
    real gx = (real)s->fp.ConstBodyForce[0], gy = (real)s->fp.ConstBodyForce[1], gz = (real)s->fp.ConstBodyForce[2];
    const real* gpf = s->arr[ORC_GPF];
    if (s->opt.ibm_kernel == D3Q19_IBM_LINEAR_SMOOTHED) {
        /* 27-point (1/4,1/2,1/4)^3 smoothing; shell entries of gpf are never written, so they
         * contribute zero, also across periodic faces (GPU_program.cl:360-377). */
        static const double sten[3] = {0.25, 0.5, 0.25};
        for (int sx = -1; sx <= 1; sx++)
            for (int sy = -1; sy <= 1; sy++)
                for (int sz = -1; sz <= 1; sz++) {
                    long is = (x + sx) + (long)Nx * ((y + sy) + (long)Ny * (z + sz));
                    real ws = (real)sten[sx + 1] * (real)sten[sy + 1] * (real)sten[sz + 1];
                    gx += ws * gpf[is];
                    gy += ws * gpf[is + nn];
                    gz += ws * gpf[is + 2 * nn];
                }
    } else {
        gx += gpf[i];
        gy += gpf[i + nn];
        gz += gpf[i + 2 * nn];
    }
    /* velocity with half the force (Guo), GPU_program.cl:387-389 */
    real ux = (f[1] - f[2] + f[7] + f[8] + f[9] + f[10] - f[11] - f[12] - f[13] - f[14] + (real)0.5 * gx) / rho;
    real uy = (f[3] - f[4] + f[7] - f[8] + f[11] - f[12] + f[15] + f[16] - f[17] - f[18] + (real)0.5 * gy) / rho;
    real uz = (f[5] - f[6] + f[9] - f[10] + f[13] - f[14] + f[15] - f[16] + f[17] - f[18] + (real)0.5 * gz) / rho;
    real* u = s->arr[ORC_U];
    u[i] = ux;
    u[i + nn] = uy;
    u[i + 2 * nn] = uz;

    real feq[19], n[19], fg[19];
    feq_d3q19(feq, rho, ux, uy, uz);
    for (int q = 0; q < 19; q++) n[q] = feq[q] - f[q];
    guo_d3q19(fg, ux, uy, uz, gx, gy, gz);

    const int bgk = (s->opt.collision == D3Q19_COLL_BGK);
    /* s vector, GPU_program.cl:453; 1/tau on moments 8,9,10,14,15 (GPU_program.cl:567-568) */
    real sv[19] = {0, (real)1.19, (real)1.40, 1, (real)1.20, 1, (real)1.20, 1, 1, 1, 1,
                   (real)1.20, (real)1.40, (real)1.40, 1, 1, (real)1.98, (real)1.98, (real)1.98};
    real tau = (real)s->fp.NewtonianTau;
    real* taulb = s->arr[ORC_TAU];
    real rn[19], rg[19];
    if (s->opt.variable_viscosity) {
        real told = taulb[i];
        real it = (real)1.0 / told;
        sv[8] = it; sv[9] = it; sv[10] = it; sv[14] = it; sv[15] = it;
        if (s->use_s_override) memcpy(sv, s->s_override, sizeof(sv));
        relax(s, sv, n, rn, bgk, told);
        relax(s, sv, fg, rg, bgk, told);
        real ss[19];
        for (int q = 0; q < 19; q++) ss[q] = rn[q] + fg[q] - (real)0.5 * rg[q];
        /* second moment of the collision change minus (uF+Fu), GPU_program.cl:520-557 */
        real sc[3][3];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                real acc = 0;
                for (int q = 1; q < 19; q++)
                    if (C[q][a] * C[q][b] != 0) acc += (real)(C[q][a] * C[q][b]) * ss[q];
                sc[a][b] = acc;
            }
        real uu[3] = {ux, uy, uz}, gg[3] = {gx, gy, gz};
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) sc[a][b] -= (uu[a] * gg[b] + uu[b] * gg[a]);
        real srt = 0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) srt += sc[a][b] * sc[a][b];
        srt = r_sqrt(srt) * (real)2.1213203435596424 / rho;
        tau = compute_tau(s, srt);
        taulb[i] = tau;
    }
    real it = (real)1.0 / tau;
    sv[8] = it; sv[9] = it; sv[10] = it; sv[14] = it; sv[15] = it;
    if (s->use_s_override) memcpy(sv, s->s_override, sizeof(sv));
    relax(s, sv, n, rn, bgk, tau);
    real fnew[19];
    if (bgk) {
        real pf = (real)1.0 - (real)0.5 / tau;
        for (int q = 0; q < 19; q++) fnew[q] = f[q] + rn[q] + pf * fg[q];
    } else {
        relax(s, sv, fg, rg, bgk, tau);
        for (int q = 0; q < 19; q++) fnew[q] = f[q] + rn[q] + fg[q] - (real)0.5 * rg[q];
    }
    /* push to the 19 neighbours, shell included (GPU_program.cl:616-623, 951-976) */
    for (int q = 0; q < 19; q++) {
        long t = (x + C[q][0]) + (long)Nx * ((y + C[q][1]) + (long)Ny * (z + C[q][2]));
        fs[t + q * nn] = fnew[q];
    }
}

void ORC(collide_stream)(orc_sim* s) {
    const int Nx = s->N[0], Ny = s->N[1], Nz = s->N[2];
    const real* fc = s->arr[ORC_F_C];
    real* fs = s->arr[ORC_F_S];
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 1; z < Nz - 1; z++)
        for (int y = 1; y < Ny - 1; y++)
            for (int x = 1; x < Nx - 1; x++) collide_node(s, x, y, z, fc, fs);
}

/* reset_particle_fluid_forces, GPU_program.cl:292-313 (slot 0, interior only) */
void ORC(reset_forces)(orc_sim* s) {
    const int Nx = s->N[0], Ny = s->N[1], Nz = s->N[2];
    real* g = s->arr[ORC_GPF];
    for (int z = 1; z < Nz - 1; z++)
        for (int y = 1; y < Ny - 1; y++)
            for (int x = 1; x < Nx - 1; x++) {
                long i = x + (long)Nx * (y + (long)Ny * z);
                g[i] = 0; g[i + s->nn] = 0; g[i + 2 * s->nn] = 0;
            }
}

/* ---- periodic copy ---------------------------------------------------------------------------
 * Map of create_periodic_stream_mapping (D3Q19-OpenCL.c:973-1112): face nodes of periodic axes
 * (edges excluded), edge nodes when EITHER of the edge's two axes is periodic, the 8 vertices
 * when any axis is periodic.  For each such node and each population q whose velocity points
 * inward along an axis where the node sits on the boundary, copy f_s[q] from the periodic image
 * i + ((c*m*(c+m))/2)*(N-2) (GPU_program.cl:711-734). */
static void periodic_node(orc_sim* s, real* fs, int r0, int r1, int r2, const int m[3]) {
    long nn = s->nn;
    const int* N = s->N;
    long i = r0 + (long)N[0] * (r1 + (long)N[1] * r2);
    for (int q = 1; q < 19; q++) {
        /* unknown: the velocity has a component along an inward normal of this node */
        int unknown = 0;
        for (int d = 0; d < 3; d++)
            if (m[d] != 0 && C[q][d] == m[d]) unknown = 1;
        if (!unknown) continue;
        int off[3];
        for (int d = 0; d < 3; d++) off[d] = ((C[q][d] * m[d] * (C[q][d] + m[d])) / 2) * (N[d] - 2);
        long pi = i + off[0] + (long)N[0] * (off[1] + (long)N[1] * off[2]);
        fs[q * nn + i] = fs[q * nn + pi];
    }
}

void ORC(boundary_periodic)(orc_sim* s) {
    const int* N = s->N;
    const int* bc = s->ip.BoundaryConds;
    real* fs = s->arr[ORC_F_S];
    const int b = 1;
    int lo[3] = {b, b, b}, hi[3] = {N[0] - b - 1, N[1] - b - 1, N[2] - b - 1};
    /* faces */
    for (int face = 0; face < 6; face++) {
        int ax = face / 2;
        if (bc[ax] != 0) continue;
        int m[3] = {0, 0, 0};
        m[ax] = (face % 2 == 0) ? 1 : -1;
        int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
        for (int i1 = lo[a1] + 1; i1 <= hi[a1] - 1; i1++)
            for (int i2 = lo[a2] + 1; i2 <= hi[a2] - 1; i2++) {
                int r[3];
                r[ax] = (m[ax] > 0) ? lo[ax] : hi[ax];
                r[a1] = i1;
                r[a2] = i2;
                periodic_node(s, fs, r[0], r[1], r[2], m);
            }
    }
    /* edges: fixed axes (a, b2), running axis c */
    for (int a = 0; a < 3; a++)
        for (int bb = a + 1; bb < 3; bb++) {
            if (!(bc[a] == 0 || bc[bb] == 0)) continue;
            int c = 3 - a - bb;
            for (int sa = 0; sa < 2; sa++)
                for (int sb = 0; sb < 2; sb++) {
                    int m[3] = {0, 0, 0};
                    m[a] = sa == 0 ? 1 : -1;
                    m[bb] = sb == 0 ? 1 : -1;
                    for (int ic = lo[c] + 1; ic <= hi[c] - 1; ic++) {
                        int r[3];
                        r[a] = sa == 0 ? lo[a] : hi[a];
                        r[bb] = sb == 0 ? lo[bb] : hi[bb];
                        r[c] = ic;
                        periodic_node(s, fs, r[0], r[1], r[2], m);
                    }
                }
        }
    /* vertices */
    if (bc[0] == 0 || bc[1] == 0 || bc[2] == 0)
        for (int v = 0; v < 8; v++) {
            int m[3], r[3];
            for (int d = 0; d < 3; d++) {
                int up = (v >> (2 - d)) & 1;
                m[d] = up ? -1 : 1;
                r[d] = up ? hi[d] : lo[d];
            }
            periodic_node(s, fs, r[0], r[1], r[2], m);
        }
}

/* ---- moving-wall velocity BC -----------------------------------------------------------------
 * boundary_velocity, GPU_program.cl:737-865, on the wet layers 1 and N-2 of `axis`.  Knowns are
 * the populations with zero or outward normal component; each unknown q is its opposite plus
 * rho*(c_q . U)/3 (axis-aligned) or /6 (diagonal), where U has the wall's tangential velocity and
 * the inward normal velocity u_n (computed from the knowns with rho = 1 when calc_rho == 0,
 * GPU_program.cl:826-835; the transverse-momentum correction is off, 842-843). */
static int opposite(int q) {
    for (int p = 0; p < 19; p++)
        if (C[p][0] == -C[q][0] && C[p][1] == -C[q][1] && C[p][2] == -C[q][2]) return p;
    return 0;
}

void ORC(boundary_velocity)(orc_sim* s, int axis, int calc_rho) {
    const int* N = s->N;
    real* fs = s->arr[ORC_F_S];
    long nn = s->nn;
    int a1 = (axis == 0) ? 1 : 0;
    int a2 = (axis == 2) ? 1 : 2;
    for (int side = 0; side < 2; side++) {
        int layer = side == 0 ? 1 : N[axis] - 2;
        int inward = side == 0 ? 1 : -1;
        const float* uw = side == 0 ? s->fp.VelLower : s->fp.VelUpper;
        int lo1 = 1, hi1 = N[a1] - 2, lo2 = 1, hi2 = N[a2] - 2;
        for (int j2 = lo2; j2 <= hi2; j2++)
            for (int j1 = lo1; j1 <= hi1; j1++) {
                int r[3];
                r[axis] = layer; r[a1] = j1; r[a2] = j2;
                long i = r[0] + (long)N[0] * (r[1] + (long)N[1] * r[2]);
                real sum0 = 0, sumout = 0;
                for (int q = 0; q < 19; q++) {
                    if (C[q][axis] == 0) sum0 += fs[i + q * nn];
                }
                for (int q = 0; q < 19; q++) {
                    if (C[q][axis] == -inward) sumout += fs[i + q * nn];
                }
                real un = -(real)(side == 0 ? -1 : 1) * (real)uw[axis];
                real rho;
                if (calc_rho) {
                    rho = (sum0 + (real)2 * sumout) / ((real)1.0 - un);
                } else {
                    rho = (real)1.0;
                    un = (real)1 - (sum0 + (real)2 * sumout) / rho;
                }
                real ua1 = (real)uw[a1], ua2 = (real)uw[a2];
                for (int q = 0; q < 19; q++) {
                    if (C[q][axis] != inward) continue;
                    int o = opposite(q);
                    real cu = un + (real)C[q][a1] * ua1 + (real)C[q][a2] * ua2;
                    real den = (C[q][a1] == 0 && C[q][a2] == 0) ? (real)3.0 : (real)6.0;
                    fs[i + q * nn] = fs[i + o * nn] + rho * cu / den;
                }
            }
    }
}

/* ---- immersed boundary -------------------------------------------------------------------- */
#define PK(s, p, k, c) ((s)->arr[ORC_PARKIN][4 * ((p) + (k) * (s)->np) + (c)])

static void cross3(const real* a, const real* b, real* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* Peskin's 4-point regularised delta (Peskin 2002, eq. 6.27) */
static real peskin4(real r) {
    r = r_fabs(r);
    if (r < 1) return ((real)3 - (real)2 * r + r_sqrt((real)1 + (real)4 * r - (real)4 * r * r)) / (real)8;
    if (r < 2) return ((real)5 - (real)2 * r - r_sqrt((real)-7 + (real)12 * r - (real)4 * r * r)) / (real)8;
    return 0;
}

/* particle_fluid_forces_linear_stencil (GPU_program.cl:81-256) and its Peskin variant.  Markers
 * are processed in global-id order; the spread is summed in one slot, which equals the
 * reference's slot fold (258-289) up to summation order.  Per-work-group force/torque partials
 * are reduced by the same power-of-two tree (GPU_program.cl:236-255). */
void ORC(ibm)(orc_sim* s) {
    const int np = s->np, ppp = s->ppp;
    if (np == 0) return;
    const int* N = s->N;
    const long nn = s->nn;
    real* u = s->arr[ORC_U];
    real* gpf = s->arr[ORC_GPF];
    const real* pts = s->arr[ORC_POINTS];
    real sys[3] = {(real)s->ip.SystemSize[0], (real)s->ip.SystemSize[1], (real)s->ip.SystemSize[2]};
    const int* B = s->ip.BufferSize;
    real dA = (real)s->fp.PointArea;
    long gsize = (long)np * ppp;
    real* ff = s->ffsum; /* [gid] force (4), [gid+gsize] torque (4) */
    for (int p = 0; p < np; p++) {
        for (int m = 0; m < ppp; m++) {
            long gid = (long)p * ppp + m;
            real r0[3] = {pts[4 * m], pts[4 * m + 1], pts[4 * m + 2]};
            real rpp[3];
            for (int d = 0; d < 3; d++) rpp[d] = r_fmod((PK(s, p, 0, d) + r0[d]) + sys[d], sys[d]);
            real up[3] = {0, 0, 0};
            int nodes[64];
            real wts[64];
            int nnodes = 0;
            if (s->opt.ibm_kernel == D3Q19_IBM_LINEAR_SMOOTHED) {
                int fl[3], i0[3], sh[3];
                real w[3];
                for (int d = 0; d < 3; d++) {
                    fl[d] = (int)r_floor(rpp[d]);
                    i0[d] = fl[d] + B[d];
                    sh[d] = (i0[d] == N[d] - 2) ? -(N[d] - 3) : 1;
                    w[d] = (real)1.0 - (rpp[d] - (real)fl[d]);
                }
                /* GPU_program.cl:164, 171-178: corner order and weight products */
                static const int cs[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                             {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
                for (int n = 0; n < 8; n++) {
                    real wx = cs[n][0] ? (real)1.0 - w[0] : w[0];
                    real wy = cs[n][1] ? (real)1.0 - w[1] : w[1];
                    real wz = cs[n][2] ? (real)1.0 - w[2] : w[2];
                    int xn = i0[0] + cs[n][0] * sh[0], yn = i0[1] + cs[n][1] * sh[1], zn = i0[2] + cs[n][2] * sh[2];
                    nodes[nnodes] = xn + N[0] * (yn + N[1] * zn);
                    wts[nnodes] = wx * wy * wz;
                    nnodes++;
                }
            } else {
                int base[3];
                real wd[3][4];
                int idx[3][4], ok[3][4];
                for (int d = 0; d < 3; d++) {
                    base[d] = (int)r_floor(rpp[d]);
                    int n_int = N[d] - 2; /* interior nodes along d */
                    for (int k = 0; k < 4; k++) {
                        int node = base[d] - 1 + k; /* particle coordinate of the node */
                        wd[d][k] = peskin4(rpp[d] - (real)node);
                        if (s->ip.BoundaryConds[d] == 0) {
                            node = ((node % n_int) + n_int) % n_int;
                            ok[d][k] = 1;
                        } else {
                            ok[d][k] = (node >= 0 && node < n_int);
                        }
                        idx[d][k] = node + B[d];
                    }
                }
                for (int kz = 0; kz < 4; kz++)
                    for (int ky = 0; ky < 4; ky++)
                        for (int kx = 0; kx < 4; kx++) {
                            if (!(ok[0][kx] && ok[1][ky] && ok[2][kz])) continue;
                            nodes[nnodes] = idx[0][kx] + N[0] * (idx[1][ky] + N[1] * idx[2][kz]);
                            wts[nnodes] = wd[0][kx] * wd[1][ky] * wd[2][kz];
                            nnodes++;
                        }
            }
            for (int n = 0; n < nnodes; n++) {
                up[0] += wts[n] * u[nodes[n]];
                up[1] += wts[n] * u[nodes[n] + nn];
                up[2] += wts[n] * u[nodes[n] + 2 * nn];
            }
            real av[3] = {PK(s, p, 3, 0), PK(s, p, 3, 1), PK(s, p, 3, 2)};
            real wr[3];
            cross3(av, r0, wr);
            real F[3], T[3];
            for (int d = 0; d < 3; d++) F[d] = (up[d] - (PK(s, p, 1, d) + wr[d])) * dA;
            cross3(r0, F, T);
            for (int n = 0; n < nnodes; n++) {
                gpf[nodes[n]] -= wts[n] * F[0];
                gpf[nodes[n] + nn] -= wts[n] * F[1];
                gpf[nodes[n] + 2 * nn] -= wts[n] * F[2];
            }
            for (int d = 0; d < 3; d++) {
                ff[4 * gid + d] = F[d];
                ff[4 * (gid + gsize) + d] = T[d];
            }
            ff[4 * gid + 3] = 0;
            ff[4 * (gid + gsize) + 3] = 0;
        }
    }
    /* work-group tree reduction (GPU_program.cl:236-255) */
    int L = s->ppwg;
    long ngroups = gsize / L;
    real* pff = s->arr[ORC_PARFLUID];
    for (long g = 0; g < ngroups; g++) {
        long g0 = g * L;
        for (int st = L / 2; st > 0; st >>= 1)
            for (int l = 0; l < st; l++)
                for (int c = 0; c < 4; c++) {
                    ff[4 * (g0 + l) + c] += ff[4 * (g0 + l + st) + c];
                    ff[4 * (g0 + l + gsize) + c] += ff[4 * (g0 + l + st + gsize) + c];
                }
        for (int c = 0; c < 4; c++) {
            pff[4 * g + c] = ff[4 * g0 + c];
            pff[4 * (g + ngroups) + c] = ff[4 * (g0 + gsize) + c];
        }
    }
}

/* ---- rigid-body update -----------------------------------------------------------------------
 * particle_dynamics, CPU_program.cl:187-277: explicit Newton-Euler with dt = 1, periodic wrap of
 * the position with fmod over SystemSize, the quaternion slot overwritten with dq/dt / 2 (the
 * reference's update, unused downstream), contact force/torque reset. */
void ORC(particle_dynamics)(orc_sim* s) {
    const int np = s->np, nfa = s->nfa;
    real* pff = s->arr[ORC_PARFLUID];
    real* pf = s->arr[ORC_PARFORCE];
    real w[4] = {(real)s->ip.SystemSize[0], (real)s->ip.SystemSize[1], (real)s->ip.SystemSize[2], 1};
    for (int p = 0; p < np; p++) {
        real acc[4] = {0, 0, 0, 0}, aacc[4] = {0, 0, 0, 0};
        for (int fa = 0; fa < nfa; fa++)
            for (int c = 0; c < 4; c++) {
                acc[c] += pff[4 * (p * nfa + fa) + c];
                aacc[c] += pff[4 * ((p + np) * nfa + fa) + c];
            }
        for (int c = 0; c < 4; c++) {
            acc[c] += pf[4 * p + c];
            aacc[c] += pf[4 * (p + np) + c];
        }
        for (int c = 0; c < 4; c++) {
            acc[c] /= (real)s->fp.ParticleMass;
            aacc[c] /= (real)s->fp.ParticleMomInertia;
        }
        for (int c = 0; c < 3; c++) acc[c] += (real)s->fp.ConstBodyForce[c];
        for (int c = 0; c < 4; c++) {
            real rt = PK(s, p, 0, c) + PK(s, p, 1, c) + (real)0.5 * acc[c];
            PK(s, p, 0, c) = r_fmod(rt + w[c], w[c]);
            PK(s, p, 1, c) = PK(s, p, 1, c) + acc[c];
        }
        real q[4], av[4];
        for (int c = 0; c < 4; c++) { q[c] = PK(s, p, 2, c); av[c] = PK(s, p, 3, c); }
        real dqw = -(av[0] * q[0] + av[1] * q[1] + av[2] * q[2] + av[3] * q[3]);
        real cr[3];
        cross3(av, q, cr);
        for (int c = 0; c < 3; c++) PK(s, p, 2, c) = (real)0.5 * (q[3] * av[c] + cr[c]);
        PK(s, p, 2, 3) = (real)0.5 * (q[3] * av[3] + 0) + (real)0.5 * dqw;
        for (int c = 0; c < 4; c++) PK(s, p, 3, c) = av[c] + aacc[c];
        for (int c = 0; c < 4; c++) { pf[4 * p + c] = 0; pf[4 * (p + np) + c] = 0; }
    }
}

/* ---- contact zones ------------------------------------------------------------------------------
 * The zone of a sphere: (int)(x / ZoneWidth) per axis in float, as parKin is float in the
 * reference (D3Q19-OpenCL.c:401-404, CPU_program.cl:299-302). */
static int zone_of(const orc_sim* s, int p) {
    int id[3];
    for (int d = 0; d < 3; d++) id[d] = (int)((float)PK(s, p, 0, d) / s->fp.ZoneWidth[d]);
    return id[0] + s->ip.NumZones[0] * (id[1] + s->ip.NumZones[1] * id[2]);
}

/* Membership lists from pars_zone, in sphere order: the initial assignment appends p = 0..np-1
 * (D3Q19-OpenCL.c:383-417); the rebuild runs the reference's threads, each over its own spheres
 * in increasing order, one after another (CPU_program.cl:280-315, counts reset on the host,
 * sim_main.c:429-444).  A zone id outside the grid is out of bounds in the reference (an
 * error at initialisation, an unchecked write at a rebuild); here the sphere joins no zone. */
static void assign_zones(orc_sim* s) {
    const int np = s->np;
    memset(s->num_in_zone, 0, sizeof(int) * (size_t)s->nzt);
    for (int p = 0; p < np; p++) {
        const int z = zone_of(s, p);
        s->pars_zone[p] = z;
        if (z < 0 || z >= s->nzt) continue;
        s->zone_members[(size_t)z * np + s->num_in_zone[z]++] = p;
    }
}

/* Neighbour lists, D3Q19-OpenCL.c:419-462: offsets -1..1 per axis, clipped at the zone grid's
 * faces on type-1 (velocity) boundaries, wrapped elsewhere (ls < 0 -> NumZones - 1, else
 * ls % NumZones), the neighbour id formed as lw + NumZones[0]*(mw + NumZones[2]*nw) -- NumZones[2]
 * where NumZones[1] is meant (line 450) -- and a zone reached twice listed twice. */
static void build_zone_neighbours(orc_sim* s) {
    const int* NZ = s->ip.NumZones;
    const int* BC = s->ip.BoundaryConds;
    for (int k = 0; k < NZ[2]; k++)
        for (int j = 0; j < NZ[1]; j++)
            for (int i = 0; i < NZ[0]; i++) {
                const int zone = i + NZ[0] * (j + NZ[1] * k);
                int* list = s->zone_neigh + 28 * (size_t)zone;
                list[0] = 0;
                int lm = -1, lp = 1, mm = -1, mp = 1, nm = -1, npp = 1;
                if (BC[0] == 1 && i == 0) lm = 0;
                if (BC[1] == 1 && j == 0) mm = 0;
                if (BC[2] == 1 && k == 0) nm = 0;
                if (BC[0] == 1 && i == NZ[0] - 1) lp = 0;
                if (BC[1] == 1 && j == NZ[1] - 1) mp = 0;
                if (BC[2] == 1 && k == NZ[2] - 1) npp = 0;
                for (int n = nm; n <= npp; n++)
                    for (int m = mm; m <= mp; m++)
                        for (int l = lm; l <= lp; l++) {
                            const int ls = i + l, ms = j + m, ns = k + n;
                            const int lw = ls < 0 ? NZ[0] - 1 : ls % NZ[0];
                            const int mw = ms < 0 ? NZ[1] - 1 : ms % NZ[1];
                            const int nw = ns < 0 ? NZ[2] - 1 : ns % NZ[2];
                            list[++list[0]] = lw + NZ[0] * (mw + NZ[2] * nw);
                        }
            }
}

static void init_zones(orc_sim* s) {
    const int* NZ = s->ip.NumZones;
    s->nzt = NZ[0] * NZ[1] * NZ[2];
    s->pars_zone = (int*)calloc((size_t)s->np + 1, sizeof(int));
    s->zone_members = (int*)calloc((size_t)s->nzt * s->np + 1, sizeof(int));
    s->num_in_zone = (int*)calloc((size_t)s->nzt + 1, sizeof(int));
    s->zone_neigh = (int*)calloc(28 * (size_t)s->nzt + 1, sizeof(int));
    assign_zones(s);
    build_zone_neighbours(s);
}

static int rebuild_freq(const orc_sim* s) { return s->opt.rebuild_freq > 0 ? s->opt.rebuild_freq : 10; }

/* ---- contacts ----------------------------------------------------------------------------------
 * particle_particle_forces + compute_squeeze_force (CPU_program.cl:76-184, 319-343).  With
 * D3Q19_CONTACT_ZONES every sphere pi visits the members pj < pi of the zones in its zone's
 * neighbour list, as the reference does; with D3Q19_CONTACT_ALL_PAIRS every pair pi > pj once. */
static void pair_force(orc_sim* s, int pi, int pj) {
    real* pf = s->arr[ORC_PARFORCE];
    real rp = (real)s->fp.ParticleDiam / (real)2.0;
    real w[3] = {(real)s->ip.SystemSize[0], (real)s->ip.SystemSize[1], (real)s->ip.SystemSize[2]};
    real k = (real)s->fp.ParForceParams[0];
    real rij[3], vij[3];
    for (int d = 0; d < 3; d++) {
        rij[d] = PK(s, pj, 0, d) - PK(s, pi, 0, d);
        int sg = (rij[d] < 0) ? -1 : (rij[d] > 0);
        rij[d] = (r_fabs(rij[d]) < w[d] / (real)2.0) ? rij[d] : (rij[d] - (real)sg * w[d]);
        vij[d] = PK(s, pj, 1, d) - PK(s, pi, 1, d);
    }
    real rsep = r_sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
    real minsep = rsep - 2 * rp;
    real e[3] = {rij[0] / rsep, rij[1] / rsep, rij[2] / rsep};
    real vrel = -(e[0] * vij[0] + e[1] * vij[1] + e[2] * vij[2]);
    real overlap = -minsep;
    if (overlap > 0 && s->ip.ParForceModel == 1) {
        real fm = k * overlap;
        for (int d = 0; d < 3; d++) { pf[4 * pi + d] -= e[d] * fm; pf[4 * pj + d] += e[d] * fm; }
    }
    if (minsep > 0 && minsep < (real)0.1 * rp) {
        real sq = 0;
        if (s->ip.ViscosityModel == D3Q19_VISC_NEWTONIAN) {
            real rstar = rp / (real)2.0;
            real ms = minsep < (real)0.1 ? (real)0.1 : minsep;
            real eta0 = ((real)s->fp.NewtonianTau - (real)0.5) / (real)3.0;
            sq = (real)(6.0 * 3.14159265358979323846) * eta0 * rstar * rstar * vrel / ms;
        }
        for (int d = 0; d < 3; d++) { pf[4 * pi + d] -= e[d] * sq; pf[4 * pj + d] += e[d] * sq; }
    }
}

void ORC(particle_forces)(orc_sim* s) {
    const int np = s->np;
    real* pf = s->arr[ORC_PARFORCE];
    real k = (real)s->fp.ParForceParams[0];
    const int zones = s->opt.contact_detection == D3Q19_CONTACT_ZONES;
    if (zones && !s->zone_neigh) init_zones(s);
    for (int pi = 0; pi < np; pi++) {
        if (zones) {
            const int pz = s->pars_zone[pi];
            if (pz >= 0 && pz < s->nzt) {
                const int* list = s->zone_neigh + 28 * (size_t)pz;
                for (int a = 1; a <= list[0]; a++) {
                    const int z = list[a];
                    if (z < 0 || z >= s->nzt) continue;
                    for (int j = 0; j < s->num_in_zone[z]; j++) {
                        const int pj = s->zone_members[(size_t)z * np + j];
                        if (pi > pj) pair_force(s, pi, pj);
                    }
                }
            }
        } else {
            for (int pj = 0; pj < pi; pj++) pair_force(s, pi, pj);
        }
        real lo = (real)s->fp.ParticleZBuffer - PK(s, pi, 0, 2);
        if (lo > 0) pf[4 * pi + 2] += k * lo;
        real hi = PK(s, pi, 0, 2) - ((real)s->ip.SystemSize[2] - (real)s->fp.ParticleZBuffer);
        if (hi > 0) pf[4 * pi + 2] -= k * hi;
    }
}

/* End of step t: the zone lists are rebuilt from the current positions when t % RebuildFreq == 0
 * (sim_main.c:428-444). */
void ORC(end_step)(orc_sim* s) {
    s->t++;
    if (s->np > 0 && s->opt.contact_detection == D3Q19_CONTACT_ZONES && s->t % rebuild_freq(s) == 0) {
        if (!s->zone_neigh) init_zones(s);
        assign_zones(s);
    }
}

static void swap_f(orc_sim* s) {
    real* t = s->arr[ORC_F_C];
    s->arr[ORC_F_C] = s->arr[ORC_F_S];
    s->arr[ORC_F_S] = t;
}

/* One time step in sim_main.c:315-471 order.  particle_dynamics runs on the reference's CPU
 * queue concurrently with collide; it reads nothing collide writes. */
void ORC(step)(orc_sim* s) {
    if (s->np > 0 && s->opt.contact_detection == D3Q19_CONTACT_ZONES && !s->zone_neigh) init_zones(s);
    ORC(collide_stream)(s);
    if (s->np > 0) {
        ORC(particle_dynamics)(s);
        ORC(reset_forces)(s);
    }
    ORC(boundary_periodic)(s);
    int wall = -1;
    for (int d = 0; d < 3; d++)
        if (s->ip.BoundaryConds[d] == 1) wall = d;
    if (wall >= 0) {
        ORC(boundary_velocity)(s, wall, s->opt.calc_rho);
        for (int d = 0; d < 3; d++)
            if (s->opt.tangential_bc[d] == 1) ORC(boundary_velocity)(s, d, 0);
    }
    if (s->np > 0) {
        ORC(ibm)(s);
        ORC(particle_forces)(s);
    }
    ORC(end_step)(s);
    swap_f(s);
}

void ORC(run)(orc_sim* s, int nsteps) {
    for (int t = 0; t < nsteps; t++) ORC(step)(s);
}

/* ---- lifetime and access ------------------------------------------------------------------- */
orc_sim* ORC(create)(const int_param_struct* ip, const flp_param_struct* fp, const d3q19_options* opt) {
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    s->ip = *ip;
    s->fp = *fp;
    if (opt) s->opt = *opt;
    for (int d = 0; d < 3; d++) s->N[d] = ip->LatticeSize[d];
    s->nn = (long)s->N[0] * s->N[1] * s->N[2];
    s->np = ip->NumParticles;
    s->ppp = ip->PointsPerParticle;
    s->ppwg = ip->PointsPerWorkGroup > 0 ? ip->PointsPerWorkGroup : 1;
    s->nfa = ip->WorkGroupsPerParticle > 0 ? ip->WorkGroupsPerParticle : 1;
    long np = s->np;
    s->len[ORC_F_C] = 19 * s->nn;
    s->len[ORC_F_S] = 19 * s->nn;
    s->len[ORC_U] = 3 * s->nn;
    s->len[ORC_TAU] = s->nn;
    s->len[ORC_GPF] = 3 * s->nn;
    s->len[ORC_PARKIN] = 16 * np;
    s->len[ORC_PARFORCE] = 8 * np;
    s->len[ORC_PARFLUID] = 8 * np * s->nfa;
    s->len[ORC_POINTS] = 4 * (long)s->ppp;
    for (int a = 0; a < ORC_NARRAYS; a++) s->arr[a] = (real*)calloc((size_t)(s->len[a] + 4), sizeof(real));
    s->ffsum = (real*)calloc((size_t)(8 * np * s->ppp + 8), sizeof(real));
    build_moments(s);
    /* initial state as initialize_lattice_fields (D3Q19-OpenCL.c:158-187): f_eq(1,0), tau */
    real feq[19];
    feq_d3q19(feq, 1, 0, 0, 0);
    for (int q = 0; q < 19; q++)
        for (long i = 0; i < s->nn; i++) {
            s->arr[ORC_F_C][q * s->nn + i] = feq[q];
            s->arr[ORC_F_S][q * s->nn + i] = feq[q];
        }
    for (long i = 0; i < s->nn; i++) s->arr[ORC_TAU][i] = (real)fp->NewtonianTau;
    for (long p = 0; p < np; p++) PK(s, p, 2, 3) = 1;
    return s;
}

void ORC(destroy)(orc_sim* s) {
    if (!s) return;
    for (int a = 0; a < ORC_NARRAYS; a++) free(s->arr[a]);
    free(s->ffsum);
    free(s->pars_zone);
    free(s->zone_members);
    free(s->num_in_zone);
    free(s->zone_neigh);
    free(s);
}

void* ORC(ptr)(orc_sim* s, int which) { return (which >= 0 && which < ORC_NARRAYS) ? s->arr[which] : NULL; }
long ORC(len)(orc_sim* s, int which) { return (which >= 0 && which < ORC_NARRAYS) ? s->len[which] : 0; }
void ORC(swap)(orc_sim* s) { swap_f(s); }
int ORC(real_size)(void) { return (int)sizeof(real); }

/* test hooks: moment matrices, equilibrium, Guo term, tau law, relaxation override */
void ORC(moment_matrix)(orc_sim* s, double* M, double* Minv) {
    for (int a = 0; a < 19; a++)
        for (int b = 0; b < 19; b++) {
            if (M) M[a * 19 + b] = (double)s->M[a][b];
            if (Minv) Minv[a * 19 + b] = (double)s->Minv[a][b];
        }
}
void ORC(feq)(double rho, const double* u, double* out) {
    real f[19];
    feq_d3q19(f, (real)rho, (real)u[0], (real)u[1], (real)u[2]);
    for (int q = 0; q < 19; q++) out[q] = (double)f[q];
}
void ORC(guo)(const double* u, const double* g, double* out) {
    real f[19];
    guo_d3q19(f, (real)u[0], (real)u[1], (real)u[2], (real)g[0], (real)g[1], (real)g[2]);
    for (int q = 0; q < 19; q++) out[q] = (double)f[q];
}
double ORC(tau_of)(orc_sim* s, double srt) { return (double)compute_tau(s, (real)srt); }
double ORC(peskin)(double r) { return (double)peskin4((real)r); }
/* zone lists from the current positions, copied out (arrays of np, nzt, nzt*np and 28*nzt ints) */
int ORC(zone_lists)(orc_sim* s, int* pars_zone, int* num_in_zone, int* members, int* neigh) {
    if (s->np == 0) return 0;
    if (!s->zone_neigh) init_zones(s);
    else assign_zones(s);
    memcpy(pars_zone, s->pars_zone, sizeof(int) * (size_t)s->np);
    memcpy(num_in_zone, s->num_in_zone, sizeof(int) * (size_t)s->nzt);
    memcpy(members, s->zone_members, sizeof(int) * (size_t)s->nzt * s->np);
    memcpy(neigh, s->zone_neigh, sizeof(int) * 28 * (size_t)s->nzt);
    return s->nzt;
}
int ORC(steps_done)(orc_sim* s) { return s->t; }
/* one pair's contact contribution added to the contact-force array (test hook) */
void ORC(pair_force_of)(orc_sim* s, int pi, int pj) { pair_force(s, pi, pj); }
void ORC(set_s_override)(orc_sim* s, const double* sv) {
    if (!sv) { s->use_s_override = 0; return; }
    s->use_s_override = 1;
    for (int k = 0; k < 19; k++) s->s_override[k] = (real)sv[k];
}
