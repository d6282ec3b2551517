/*
 * d3q19_oracle.h -- CPU restatement of the reference's D3Q19 IBM shear-cell step.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product path
 * (libd3q19.so and the C host) never links or calls it.
 *
 * It follows the reference step by step, in its own data layout (SoA over the whole
 * LatticeSize including the buffer shell), in the push form of the reference:
 *   collide+stream   GPU_program.cl:316-624 (MRT), 868-947 (BGK form), 1041-1128 (tau)
 *   reset forces     GPU_program.cl:292-313
 *   periodic copy    GPU_program.cl:626-735 with the map of D3Q19-OpenCL.c:973-1112
 *   wall BC          GPU_program.cl:737-865, tangential passes sim_main.c:388-404
 *   IBM              GPU_program.cl:81-256 (linear), Peskin 4-point (configuration mode)
 *   dynamics         CPU_program.cl:187-277
 *   contacts         CPU_program.cl:76-184, 319-343, over the zone lists of
 *                    D3Q19-OpenCL.c:344-463 and CPU_program.cl:280-315, or over all pairs
 *   step order       sim_main.c:315-471
 * Built twice: orc_*_f32 (float, reference arithmetic) and orc_*_f64 (double).
 */
#ifndef D3Q19_ORACLE_H
#define D3Q19_ORACLE_H
#include "../include/d3q19_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Array identifiers for orc_ptr */
enum {
    ORC_F_C = 0,      /* pre-collision populations (19*N_C) -- what collide reads */
    ORC_F_S = 1,      /* stream target (19*N_C) */
    ORC_U = 2,        /* velocity (3*N_C) */
    ORC_TAU = 3,      /* relaxation time (N_C) */
    ORC_GPF = 4,      /* spread force, one slot (3*N_C) */
    ORC_PARKIN = 5,   /* 4*np float4 as 16*np reals */
    ORC_PARFORCE = 6, /* 2*np float4 */
    ORC_PARFLUID = 7, /* 2*np*nfa float4 */
    ORC_POINTS = 8,   /* PointsPerParticle float4 */
    ORC_NARRAYS = 9
};

#ifdef __cplusplus
}
#endif
#endif
