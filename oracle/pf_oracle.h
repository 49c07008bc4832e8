/*
 * pf_oracle.h -- CPU restatement of GRAL's prognostic microscale flow solver
 * (PrognosticFlowfield.Calculate and the U/V/W/TKE column solvers it calls),
 * in the reference's deterministic single-thread mode (UseFixedRndSeedVal).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the parity checker for the HIP
 * solver in gral_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Array layout: dense, ghost-padded, index 0 present but unused as in the
 * reference's jagged arrays.
 *   2-D arrays  [nii+2][njj+2]            idx = i*(njj+2) + j
 *   3-D arrays  [nii+2][njj+2][nkk+2]     idx = (i*(njj+2) + j)*(nkk+2) + k
 */
#ifndef GRAL_PF_ORACLE_H
#define GRAL_PF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nii, njj, nkk;          /* Program.NII/NJJ/NKK */
    int kadvmax;                /* Program.KADVMAX (after MicroscaleFlat) */
    int vertcellsff;            /* Program.VertCellsFF */
    float dxk, dyk;             /* Program.DXK/DYK */
    int turbulence_model;       /* turbulence_model.txt: 0, 1 (default), 2 */
    int simd_width;             /* Vector<float>.Count semantics of the model-1 solvers: 8 */
    int iter_min, iter_max;     /* Integrationtime.txt */
    float relax, relaxp;        /* relaxation_factors.txt */
    float building_z0;          /* building_roughness.txt */
    float coriolis;             /* Program.CorolisParam */
    float potdtdz;              /* Program.PotdTdz */
    float std_dev_v, std_dev_w; /* Program.StdDeviationV/W */
    float bdlayheight;          /* Program.BdLayHeight */
    float v0_1;                 /* Program.V0[1] */
    int refp_i, refp_j;         /* Program.SubDomainRefPos */
} orc_params;

typedef struct {
    const float *dzk, *hokart;  /* [nkk+2] */
    const uint8_t *advdom;      /* [nii+2][njj+2] */
    const int16_t *kkart;       /* [nii+2][njj+2] */
    const float *cutk, *cov, *ahk;
    const float *z0_wall;       /* roughness used by the wall helpterms (PrognosticFlowfield.cs:365-381) */
    const float *z0_solver;     /* Z0 used inside the U/V/W/TKE solvers (U-prognostic-microscale_1.cs:158-177) */
    const float *ob_init, *ust_init, *z0_init;  /* TURB init (PrognosticFlowfield.cs:415-436) */
    const float *veg;           /* [nii+2][njj+2][nkk+2] or NULL */
    const uint8_t *veg_present; /* [nii+2][njj+2]: VEG[i][j] != null, or NULL */
    int istatistics;            /* Program.IStatistics: 0 MeteoZR / 3 MeteoSonic select the measured
                                   profile in the TKE initialisation (PrognosticFlowfield.cs:456-481) */
    int n_profile;              /* Program.MetProfileNumb */
    const float *meas_height, *u0, *v0;  /* [n_profile+1] MeasurementHeight, U0, V0 (index 0 unused) */
} orc_inputs;

typedef struct {
    float *uk, *vk, *wk;                    /* in/out, [nii+2][njj+2][nkk+2] */
    float *uks, *vks, *wks;                 /* out (required) */
    float *dpm, *dpmnew, *div;              /* out (required) */
    float *turb, *tdiss;                    /* out (required when turbulence_model != 1) */
} orc_fields;

typedef struct {
    int iteration_loops;        /* IterationLoops at exit (PrognosticFlowfield.cs:1245) */
    int iterations_run;         /* number of SIMPLE iterations executed */
    float delta_first, delta_finish;
    float topwind, dtime, ug, vg, ustern_factor_x;
    double frund, geschw_max;
    int unreal_u, unreal_v, unreal_w;
    int convergence_problem;    /* DeltaFinish > DeltaFirst */
    float *delta_umax_hist;     /* optional [iter_max]: DeltaUMax of each iteration */
} orc_stats;

/* Full PrognosticFlowfield.Calculate(): prologue, SIMPLE loop, epilogue checks.
 * Returns 0 on success, negative on invalid input. */
int orc_calculate(const orc_params *p, const orc_inputs *in, orc_fields *f, orc_stats *st);

/* The pressure-correction part of iter_max SIMPLE iterations after the
 * prologue: divergence (PrognosticFlowfield.cs:723-758), pressure GS-TDMA in
 * the sweep directions of each iteration (:770-901) and velocity correction
 * (:904-949), without boundary copies, averages or momentum sweeps.  With
 * quantise = 0 the corrected velocities are not rounded to 1/Frund, which
 * makes the map from the initial UK/VK/WK to the result linear (the forward
 * of the library's adjoint, gral_pf_forward_ckpt / gral_pf_vjp). */
int orc_project(const orc_params *p, const orc_inputs *in, orc_fields *f, orc_stats *st, int quantise);

/* Timing mode: the same solve with the reference's default multi-threaded
 * schedule (rows of each sweep dealt to nthreads threads updating the fields
 * in place, U and V at the same time).  Not deterministic; used only as the
 * CPU baseline of the benchmark, never as a checker. */
int orc_calculate_threads(const orc_params *p, const orc_inputs *in, orc_fields *f, orc_stats *st, int nthreads);

/* DiagnosticFlowfield.Calculate() (DiagnosticFlowfield.cs:24-436) followed by
 * the final mass adjustment of the vertical velocity (MicroscaleFlat.cs:388-462,
 * or MicroscaleTerrain.cs:862-933 with terrain = 1).
 * flow_field_level 1 (diagnostic): 5 pressure iterations, W adjusted in every
 * column; 2 (prognostic): 10 iterations, W adjusted outside the sub-domains
 * (ADVDOM == 0).  iterations > 0 overrides the count (checks only).
 * Fields are full height [nii+2][njj+2][nkk+2]: uk/vk/wk in/out, dpm out
 * (levels 1..NKK+1 of columns 1..NII x 1..NJJ are reset), div scratch. */
typedef struct {
    int nii, njj, nkk, kadvmax;
    float dxk, dyk;
    int flow_field_level;
    int iterations;
    int terrain;                /* 1: the step as MicroscaleTerrain.cs:862-933 does it (every column, DYK on
                                   the j faces); 0: as MicroscaleFlat.cs:388-462 */
} orc_diag_params;
int orc_diagnostic(const orc_diag_params *p, const float *dzk, const uint8_t *advdom, const int16_t *kkart,
                   float *uk, float *vk, float *wk, float *dpm, float *div, float *dtime_out);

/* ---- checking hooks (closed-form and textbook checks of the oracle) ---- */

/* Records, for one column (i,j) during one SIMPLE iteration, the coefficients
 * with which a mixing-length momentum sweep (comp 0 U, 1 V, 2 W) or the
 * pressure sweep (comp 3) enters its Thomas recurrence, and the column before
 * and after the update.  Momentum: a[k] x[k] = b[k] x[k+1] + c[k] x[k-1] + d[k]
 * (c dropped at k <= KKART+1).  Pressure: a = APP, b = ABP (bottom), c = ATP
 * (top), d = DIM, APP x[k] = ABP x[k-1] + ATP x[k+1] + DIM. */
typedef struct {
    int comp, i, j, iteration;
    float *a, *b, *c, *d;       /* [nkk+2] each */
    float *x_old, *x_new;       /* [nkk+2] each */
    int captured;               /* set to 1 when recorded */
    float *piece;               /* optional [ORC_NPIECE][nkk+2]: the pieces below per level, as
                                   the mixing-length solvers (comp 0-2) and the TKE sweep (comp 4,
                                   production and buoyancy) form them for the recurrence */
} orc_capture;

/* pieces of one level of a momentum equation: eddy viscosity, conductances
 * after their 1e-4 floor, face fluxes, power-law neighbour coefficients, and
 * the source terms added to DIM (pressure gradient, Coriolis/geostrophic
 * term, wall-function drag subtracted at KKART+1, eddy-viscosity cross terms,
 * vegetation drag); TKE: shear production and buoyancy */
enum {
    ORC_VIS = 0, ORC_DE, ORC_DW, ORC_DS, ORC_DN, ORC_DT, ORC_DB,
    ORC_FE, ORC_FW, ORC_FS, ORC_FN, ORC_FT, ORC_FB,
    ORC_AE, ORC_AW, ORC_AS, ORC_AN, ORC_AT, ORC_AB,
    ORC_PRESS, ORC_COR, ORC_WALL, ORC_CROSS, ORC_VEGT,
    ORC_PROD, ORC_BUOY,
    ORC_NPIECE
};
int orc_calculate_capture(const orc_params *p, const orc_inputs *in, orc_fields *f, orc_stats *st,
                          orc_capture *cap);

/* single formulas of the solver */
float orc_powlaw_vector(float d, float peclet);   /* D*Vector.Max(0,(1-0.1P)^5), U-prognostic-microscale_1.cs:348-370 */
float orc_powlaw_scalar(float d, float peclet);   /* D*FloatMax(0,(float)Pow5(1-0.1P)), U-prognostic-microscale_2.cs:180-185 */
int orc_conv_to_int(double x);                     /* ProgramFunctions.cs:792-804 */

#ifdef __cplusplus
}
#endif
#endif
