/*
 * gral_pf.h -- C ABI of the MI355X prognostic microscale flow solver.
 *
 * Drop-in for GRAL's PrognosticFlowfield.Calculate() (src/PrognosticFlowfield.cs:32),
 * the SIMPLE loop called once per weather situation from MicroscaleFlat.cs:381
 * and MicroscaleTerrain.cs:847.  The reference method takes no arguments and
 * works on Program.* statics; this ABI passes the same state explicitly as
 * blittable pointers (GRAL.csproj:23 sets DisableRuntimeMarshalling=true).
 *
 * Layout of every array argument: dense, ghost-padded, C order, index 0 kept
 * (unused) as in the reference's jagged arrays:
 *   2-D  [nii+2][njj+2]           element (i,j)   at  i*(njj+2) + j
 *   3-D  [nii+2][njj+2][nkk+2]    element (i,j,k) at (i*(njj+2) + j)*(nkk+2) + k
 * A C# caller packs Program.UK[i][j][k] etc. into pinned staging buffers of
 * this shape (missing jagged rows -> zeros).
 *
 * Conventions: every function returns 0 on success or a negative GRAL_PF_E*
 * code and never aborts; gral_pf_last_error() gives the message.  The handle
 * owns all device memory; host pointers are only read/written during the
 * call.  One handle per host thread; calls on a handle are serialised.
 */
#ifndef GRAL_PF_H
#define GRAL_PF_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRAL_PF_OK 0
#define GRAL_PF_EINVAL (-1)   /* bad argument / unsupported option */
#define GRAL_PF_EDEVICE (-2)  /* HIP runtime error (no device, launch failure) */
#define GRAL_PF_ENOMEM (-3)   /* host or device allocation failed */
#define GRAL_PF_ESTATE (-4)   /* call out of order (e.g. iterate before prepare) */

typedef struct gral_pf_handle gral_pf_handle;

/* Grid of one situation.  Replaces Program.NII/NJJ/NKK/KADVMAX/VertCellsFF/
 * DXK/DYK/DZK/HOKART (ProgramDeclarations.cs:72-92, :1011). */
typedef struct {
    int32_t nii, njj, nkk;      /* cells in x, y, z */
    int32_t kadvmax;            /* highest level the solver touches (MicroscaleFlat.cs:197) */
    int32_t vertcellsff;        /* prognostic cells above ground (micro_vert_layers.txt) */
    float dxk, dyk;             /* horizontal cell size [m] */
    const float *dzk;           /* [nkk+2] layer thickness, index 0 unused */
    const float *hokart;        /* [nkk+2] level top height, hokart[0] = 0 */
} gral_pf_grid;

/* Run options and scalars read by Calculate(): turbulence_model.txt,
 * Integrationtime.txt, relaxation_factors.txt, building_roughness.txt
 * (PrognosticFlowfield.cs:176-291), Program.CorolisParam, PotdTdz,
 * StdDeviationV/W, BdLayHeight, V0[1], SubDomainRefPos. */
typedef struct {
    int32_t turbulence_model;   /* 0 no diffusion, 1 mixing length (default), 2 k-eps */
    int32_t simd_width;         /* mixing-length solver width: 8 = Vector<float> (AVX2),
                                   16 = Vector512 solvers (PrognosticFlowfield.cs:1072-1091,
                                   U/V/W-prognostic-microscale_1_Vec512.cs) */
    int32_t iter_min, iter_max; /* IterationStepsMin/Max */
    float relax, relaxp;        /* velocity / pressure relaxation */
    float building_z0;          /* building wall roughness [m] */
    float coriolis;             /* Coriolis parameter [1/s] */
    float potdtdz;              /* potential temperature gradient [K/m] (k-eps buoyancy) */
    float std_dev_v, std_dev_w; /* StdDeviationV/W (TKE initialisation) */
    float bdlayheight;          /* boundary-layer height [m] (TKE initialisation) */
    float v0_1;                 /* Program.V0[1] (TKE initialisation) */
    int32_t refp_i, refp_j;     /* reference point (MicroscaleTerrain.GetSubDomainPoint) */
} gral_pf_params;

/* What Calculate() reports (console lines and Problemreport_GRAL.txt,
 * PrognosticFlowfield.cs:1144-1245). */
typedef struct {
    int32_t iteration_loops;    /* IterationLoops at exit ("Total number of iterations") */
    int32_t iterations_run;     /* SIMPLE iterations executed */
    float delta_first;          /* convergence measure at iteration 100 */
    float delta_finish;         /* convergence measure at the last iteration */
    int32_t convergence_problem;/* delta_finish > delta_first ("Convergence not fullfilled") */
    int32_t unreal_u, unreal_v, unreal_w; /* unrealistic wind-speed counters */
    float topwind, dtime, ug, vg;
    double frund, geschw_max;
    double last_delta;          /* last value of DeltaMean/topwind^2 tested (every 100 iterations) */
    double mass_div;            /* FP64 sum of |DIV| over the active cells at finish (gral_pf_mass_imbalance) */
    double mass_flux;           /* FP64 sum of |face flux| of the same cells; mass_div/mass_flux is the
                                   relative mass imbalance */
} gral_pf_stats;

/* Replaces the array allocations of PrognosticFlowfield.cs:45-106 (device side). */
int gral_pf_create(const gral_pf_grid *grid, gral_pf_handle **out);

/* Geometry: Program.ADVDOM, KKART, CUTK, COV, AHK, VEG (nullable columns ->
 * veg_present[i][j]=0; veg may be NULL when no vegetation exists). */
int gral_pf_set_masks(gral_pf_handle *h, const uint8_t *advdom, const int16_t *kkart,
                      const float *cutk, const float *cov, const float *ahk,
                      const float *veg, const uint8_t *veg_present);

/* Per-column roughness / stability, resolved by the caller from Z0Gral,
 * Z0Gramm, OLGral, Ob, USternGral, Ustern exactly as the reference indexes them:
 *   z0_wall   -- 'rough' of the wall helpterms     (PrognosticFlowfield.cs:365-381)
 *   z0_solver -- 'Z0' inside the U/V/W/TKE solvers  (U-prognostic-microscale_1.cs:158-177)
 *   ob_init, ust_init, z0_init -- TURB initialisation (PrognosticFlowfield.cs:415-436) */
int gral_pf_set_roughness(gral_pf_handle *h, const float *z0_wall, const float *z0_solver,
                          const float *ob_init, const float *ust_init, const float *z0_init);

/* Meteorological input of the TKE initialisation (turbulence models 0 and 2):
 * Program.IStatistics and, for IStatistics 0 (MeteoZR) or 3 (MeteoSonic),
 * MetProfileNumb and the measured standard deviations U0[]/V0[] at
 * MeasurementHeight[] (arrays [n_profile+1], index 0 unused, heights
 * increasing), which PrognosticFlowfield.cs:456-481 interpolates linearly in
 * height above ground (V0 with the slope of U0, as the reference does).
 * Other IStatistics values (4 by default, when this is never called) use the
 * wind-speed estimate (:482-487) with gral_pf_params.v0_1 as V0[1]. */
int gral_pf_set_met_profile(gral_pf_handle *h, int32_t istatistics, int32_t n_profile, const float *meas_height,
                            const float *u0, const float *v0);

/* Initial wind field Program.UK/VK/WK (the state MicroscaleFlat/MicroscaleTerrain
 * leave before the prognostic call). */
int gral_pf_upload_fields(gral_pf_handle *h, const float *uk, const float *vk, const float *wk);

/* Prologue of Calculate() (PrognosticFlowfield.cs:145-579) and transfer of the
 * solver state to the GPU. */
int gral_pf_prepare(gral_pf_handle *h, const gral_pf_params *p);

/* Runs up to n SIMPLE iterations (PrognosticFlowfield.cs:584-1178) on the
 * device-resident state.  check_convergence=1 applies the reference's stop
 * test every 100 iterations; 0 runs exactly n iterations (or until iter_max). */
int gral_pf_iterate(gral_pf_handle *h, int32_t n, int32_t check_convergence, gral_pf_stats *st);

/* Mass imbalance of the device-resident velocity field, reduced in FP64 on
 * the GPU: *sum_abs_div = sum over the active cells (ADVDOM = 1, 2 <= i <= NII-1,
 * 2 <= j <= NJJ-1, KKART < k <= VerticalIndex) of |DIV|, DIV formed in FP32 as
 * the divergence pass does (PrognosticFlowfield.cs:723-758); *sum_abs_flux =
 * sum over the same cells of the absolute fluxes through the faces that DIV
 * combines.  The reference prints no such figure; it is the FP64 residual
 * reported beside its convergence measure.  Deterministic: the same fields
 * give the same bits.  gral_pf_finish stores both in gral_pf_stats. */
int gral_pf_mass_imbalance(gral_pf_handle *h, double *sum_abs_div, double *sum_abs_flux);

/* gral_pf_mass_imbalance without any device allocation: the reduction's
 * partial sums live in a fixed file-scope device array of the library
 * (2 * 1024 + 2 doubles, 16 KiB, one per GPU, held until the process ends)
 * instead of memory the handle allocates on first use.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight loops between iterations).  Same result, same bits
 * as gral_pf_mass_imbalance.  All calls on one GPU share the array: calls that
 * overlap on one GPU (handles used from several threads) race and must be
 * serialised by the caller. */
int gral_pf_mass_imbalance_noalloc(gral_pf_handle *h, double *sum_abs_div, double *sum_abs_flux);

/* ---- Adjoint of the pressure-correction steps ----------------------------
 * The linear part of a SIMPLE iteration: the mass divergence
 * (PrognosticFlowfield.cs:723-758), the pressure-correction GS-TDMA sweep in
 * the iteration's directions (:770-901, including the row scratch reused
 * across each row) and the velocity correction (:904-949), with the
 * geometry, DTIME and the sweep directions of the prepared situation held
 * fixed.  Without the FP64 quantisation (quantise = 0) a window of n such
 * steps is a linear map of the device-resident UK, VK, WK; gral_pf_vjp gives
 * its vector-Jacobian product (hand-written transposed kernels: the sweep runs
 * backwards over the anti-diagonals with transposed Thomas recurrences).  The
 * reference has no adjoint; this is the pressure-correction slice of the
 * sensitivity of receptor winds to the inflow.
 *
 * gral_pf_forward_ckpt runs n steps from the handle's current fields (after
 * prepare, or after iterations), numbering them from the handle's next SIMPLE
 * iteration for their sweep directions, and keeps a device checkpoint of the
 * fields it started from.  It does not advance the SIMPLE iteration count. */
int gral_pf_forward_ckpt(gral_pf_handle *h, int32_t n, int32_t quantise);

/* The fields back to the checkpoint of the last gral_pf_forward_ckpt. */
int gral_pf_restore_ckpt(gral_pf_handle *h);

/* VJP of the last gral_pf_forward_ckpt window: du/dv/dw are dL/d(UK, VK, WK)
 * of its result, gu/gv/gw receive dL/d(UK, VK, WK) of its start, all
 * [nii+2][njj+2][nkk+2].  The map is the one without quantisation (with
 * quantise = 1 this is the straight-through adjoint). */
int gral_pf_vjp(gral_pf_handle *h, const float *du, const float *dv, const float *dw, float *gu, float *gv,
                float *gw);

/* Epilogue checks (PrognosticFlowfield.cs:1187-1245) and transfer back to host. */
int gral_pf_finish(gral_pf_handle *h, gral_pf_stats *st);

/* The whole of Calculate(): prepare + iterate(iter_max, check) + finish. */
int gral_pf_solve(gral_pf_handle *h, const gral_pf_params *p, gral_pf_stats *st);

/* The mass-conservation post-pass that follows the prognostic solve:
 * DiagnosticFlowfield.Calculate() (DiagnosticFlowfield.cs:24-436; 10 pressure
 * iterations for flow_field_level 2, 5 for level 1) and the caller's final
 * adjustment of the vertical velocity from the mass balance of each cell:
 * terrain = 0 as MicroscaleFlat.cs:388-462 (outside the sub-domains at level
 * 2, everywhere at level 1), terrain = 1 as MicroscaleTerrain.cs:862-933
 * (every column).  Works on the full-height fields of the handle (after
 * gral_pf_iterate/finish, or right after upload_fields for a purely
 * diagnostic run); download_fields returns the result. */
int gral_pf_diagnostic(gral_pf_handle *h, int32_t flow_field_level, int32_t terrain);

/* Replace UK, VK, WK of a prepared situation (dense [NII+2][NJJ+2][NKK+2]
 * arrays) on the device and in the handle, keeping everything the prologue and
 * the iterations so far derived (UKS/VKS/WKS, DPMNEW, TURB/TDISS, DTIME, the
 * iteration count): to continue a solve from modified velocities.  No
 * reference counterpart. */
int gral_pf_set_velocity(gral_pf_handle *h, const float *uk, const float *vk, const float *wk);

/* Adjoint of one momentum sweep, mixing-length model (turbulence model 1).
 * gral_pf_momentum_sweep runs the sweep of component comp (0 U, 1 V, 2 W) in the
 * directions of the situation's next iteration (U-, V- and W-prognostic-
 * microscale_1.cs:50-529, called from PrognosticFlowfield.cs:1079-1101) on the
 * device fields, and keeps the fields before it; gral_pf_restore_ckpt returns
 * to them.  It does not count as an iteration of the situation.
 * gral_pf_momentum_vjp returns dL/dX before the sweep from dL/dX after it (X the
 * swept component; dense [NII+2][NJJ+2][NKK+2] arrays), with UKS, VKS, WKS,
 * DPMNEW and the other components held fixed: the transposed Thomas solves and
 * Gauss-Seidel couplings, the eddy viscosity's dependence on X through every
 * power-law coefficient, and the Coriolis, vegetation, wall and cross terms.
 * It refers to the last sweep and returns GRAL_PF_ESTATE once any call has
 * changed the device fields since (iterate, set_velocity, restore_ckpt,
 * diagnostic, a batch, prepare).
 * No reference counterpart (the reference has no adjoint).  Columns up to 126
 * levels. */
int gral_pf_momentum_sweep(gral_pf_handle *h, int32_t comp);
int gral_pf_momentum_vjp(gral_pf_handle *h, const float *dx, float *gx);

/* Copy the solver state to host buffers (each may be NULL), full depth
 * [nii+2][njj+2][nkk+2]: Program.UK/VK/WK/UKS/VKS/WKS/DPM/DPMNEW/TURB/TDISS.
 * Valid after prepare, or after upload_fields (+ diagnostic) for a purely
 * diagnostic run, where the fields the solver never set come back as zeros. */
int gral_pf_download_fields(gral_pf_handle *h, float *uk, float *vk, float *wk, float *uks,
                            float *vks, float *wks, float *dpm, float *dpmnew, float *turb,
                            float *tdiss);

/* Device time (ms, HIP events on the solver's stream around the whole call)
 * and work of the most recent gral_pf_iterate / gral_pf_batch_iterate call
 * (or gral_pf_forward_ckpt / gral_pf_vjp: their steps);
 * for measurement.  No reference counterpart ("Flow field [s]" of
 * Program.cs:592 is the closest). */
typedef struct {
    double ms_total;
    int64_t launches;
    int64_t active_cells;       /* active cells of all situations (advanced per SIMPLE iteration) */
    int64_t cell_iterations;    /* active cells x SIMPLE iterations actually run, summed over situations */
    int32_t situations;         /* situations in the call (1 for gral_pf_iterate) */
    int32_t groups;             /* stream groups the batch ran its last iterations in (1 for gral_pf_iterate) */
} gral_pf_timing;
int gral_pf_last_timing(gral_pf_handle *h, gral_pf_timing *t);

/* Measurement: runs n SIMPLE iterations (continuing the solve) with HIP events
 * around every single kernel launch, and sums the device time per kernel
 * family: 0 stencil passes, 1 pressure sweep, 2 U, 3 V, 4 W momentum sweeps,
 * 5 TKE sweep, 6 the neighbour-independent coefficient pre-pass of the tiled
 * mixing-length sweeps.  cells = active cells advanced per iteration times iterations. */
typedef struct {
    double ms;
    int64_t launches;
    int64_t cells;
} gral_pf_family_time;
int gral_pf_profile_iterations(gral_pf_handle *h, int32_t n, gral_pf_family_time *out, int32_t nfam);

/* ---- Flow-field files (GFFWrite.cs) -----------------------------------
 * GFFWrite.Write() stores each situation's wind field as a zip archive
 * holding one entry of int16 values x100.  kind selects the format:
 *   0  writing mode 0 (WriteCompressedGFFFiles, GFFWrite.cs:167-250):
 *      flag -1, per column (i 1..NII+1, j 1..NJJ+1) u, v, w of each level
 *      k 1..NKK+1, clamped then truncated;
 *   1  writing mode 1 (WriteStrongCompressedGFFFiles, :256-382): flag -2,
 *      header NKK negated, per level, per row i: u(j), then v(j), then w(j);
 *   2  writing mode 2 (WriteTerrainCompressedGFFFiles, :388-537): flag -3,
 *      header NKK negated, KKART of the inner columns, then per column u, v,
 *      w from level max(KKART,1) up, each MathF.Round-ed (half to even);
 *   3  the *.tke file of the k-eps model (GFFWrite.cs:79-160): flag -1,
 *      TURB of levels 1..NKK per inner column, rounded half to even.
 * u and v of file level k are UK/VK[k-1], w is WK[k] (GFFRead.cs:181-321
 * reads them back).  The 32-byte header is NKK, NJJ, NII (int32), the wind
 * direction and speed (float), the stability class (int32), DXK (float) and
 * the flag (int32).  The payload is packed on the GPU from the fields the
 * handle holds: after prepare, the solver's device-resident levels (no
 * download of the fields) and the host levels above them; before prepare
 * (upload_fields, diagnostic), the host fields. */
typedef struct {
    float wind_dir;             /* Program.WindDirGral */
    float wind_vel;             /* Program.WindVelGral */
    int32_t stab_class;         /* Program.StabClass */
} gral_pf_gff_meta;

/* The bytes of the zip entry (header + int16 body).  out = NULL: only *len
 * (the payload size) is set.  Otherwise cap must be >= that size.  kind 3
 * needs a prepared situation with turbulence_model 2, as GFFWrite.cs:81
 * writes the *.tke file for the k-eps model only. */
int gral_pf_gff_payload(gral_pf_handle *h, int32_t kind, const gral_pf_gff_meta *meta, void *out,
                        size_t cap, size_t *len);

/* GFFWrite.Write() for one file: packs the payload and writes it to path as a
 * zip archive with one deflated entry named like the file (GFFWrite.cs:109,
 * :177).  A file that cannot be written completely is removed (:238-246).
 * No ZIP64: payloads of 4 GiB or more are refused. */
int gral_pf_gff_write(gral_pf_handle *h, int32_t kind, const gral_pf_gff_meta *meta, const char *path);

/* The zip container on its own: one deflated entry holding len bytes of
 * data.  Deflate runs on host threads in independent 4 MiB chunks. */
int gral_pf_zip_write(const char *path, const char *entry, const void *data, size_t len);

const char *gral_pf_last_error(const gral_pf_handle *h);
void gral_pf_destroy(gral_pf_handle *h);

/* ---- Set-up of a situation on flat terrain ------------------------------
 * What GRAL computes before the prognostic call when there is no terrain,
 * on the GPU.  These calls need no handle; their output feeds
 * gral_pf_create (kadvmax), gral_pf_set_masks (advdom, kkart, ahk),
 * gral_pf_upload_fields (uk, vk, wk) and gral_pf_prepare (refp_i/j).
 * On failure gral_pf_setup_last_error() gives the message. */

/* CalculateSubDomainsForPrognosticWindSimulation (ReadGeometryAndDomains.cs:321-349):
 * sets advdom[i1][j1] = 1 within (int)(factor*CUTK/DXK) columns in i and
 * (int)(factor*CUTK/DYK) in j of every column with CUTK > 0 (clipped to
 * 1..NII x 1..NJJ); other entries keep their value.  factor is
 * Program.PrognosticSubDomainFactor (15 by default). */
int gral_pf_subdomains(int32_t nii, int32_t njj, float dxk, float dyk, int32_t factor, const float *cutk,
                       uint8_t *advdom);

/* SearchReferencePoint (MicroscaleTerrainSearchRefPoint.cs:27-103) as
 * GetSubDomainPoint (MicroscaleTerrain.cs:936-963) calls it: the first
 * sub-domain cell of a spiral from the border inwards.  Without any
 * sub-domain cell it sets (1, 1) and returns GRAL_PF_ESTATE. */
int gral_pf_reference_point(int32_t nii, int32_t njj, const uint8_t *advdom, int32_t *ref_i, int32_t *ref_j);

/* Meteorological input of the initial profile: Program.MetProfileNumb,
 * MeasurementHeight, ObsWindU/V, Ob[1][1] and OLGral. */
typedef struct {
    int32_t n_profile;          /* MetProfileNumb, 1..99 */
    const float *meas_height;   /* [n_profile+1], index 0 unused */
    const float *obs_u, *obs_v; /* [n_profile+1] wind components at those heights */
    float ob_length;            /* Ob[1][1] */
    const float *ol_column;     /* [nii+2][njj+2] OLGral when AdaptiveRoughnessMax > 0, else NULL */
    int32_t near_wall;          /* damping near obstacle walls (MicroscaleFlat.cs:201-345): 0 none;
                                   1 columns outside the sub-domains (prognostic with
                                   SubDomainDistance < 10000; needs advdom); 2 every column
                                   (diagnostic flow field) */
} gral_pf_flat_met;

/* MicroscaleFlat.Calculate() up to the flow-field call (MicroscaleFlat.cs:29-345):
 * zeroes u, v, w at levels 1..NKK of columns 1..NII+1 x 1..NJJ+1, sets u, v
 * above the obstacles from the power-law / interpolated profile, KKART and
 * AHK of every inner column (and BUI_HEIGHT of the blocked ones, when
 * bui_height is not NULL), damps u, v near walls as near_wall selects, and
 * returns KADVMAX = min(NKK, max(1, max KKART over CUTK > 0) + VertCellsFF).
 * grid->kadvmax is not read.  uk/vk/wk are in/out [nii+2][njj+2][nkk+2]:
 * entries the reference does not write keep their values. */
int gral_pf_flat_init(const gral_pf_grid *grid, const gral_pf_flat_met *met, const float *cutk,
                      const uint8_t *advdom, float *uk, float *vk, float *wk, int16_t *kkart, float *ahk,
                      float *bui_height, int32_t *kadvmax);

/* The GRAMM field a terrain situation starts from: Program.NX/NY/NZ,
 * DDX[1]/DDY[1], the GRAL grid's offset in the GRAMM grid, AHE, UWIN/VWIN/WWIN,
 * ZSP, AH and ZKO[1..2] (ProgramFunctions.cs:420-440 gives the shapes). */
typedef struct {
    int32_t nx, ny, nz;         /* GRAMM cells */
    float ddx, ddy;             /* DDX[1], DDY[1] */
    int32_t delta_i, delta_j;   /* IKOOAGRAL - GrammWest, JKOOAGRAL - GrammSouth */
    const float *ahe;           /* [nx+2][ny+2][nz+2] heights of the GRAMM cell corners */
    const float *uwin, *vwin, *wwin; /* [nx+1][ny+1][nz+1] GRAMM wind */
    const float *zsp;           /* [nx+1][ny+1][nz+1] GRAMM cell-centre heights */
    const float *ah;            /* [nx+1][ny+1] GRAMM surface height */
    double zko1, zko2;          /* ZKO[1], ZKO[2] */
} gral_pf_gramm;

/* MicroscaleTerrain.Calculate() with GRAMM topography up to the flow-field
 * call (MicroscaleTerrain.cs:57-316, :600-612, damping :664-832): for every
 * inner column the GRAMM level heights at the column centre (16-point
 * inverse-distance weighting of AHE), u, v, w interpolated from the GRAMM
 * wind above terrain plus CUTK, KKART, AHK, BUI_HEIGHT (of building columns),
 * the ghost copies, KADVMAX over the sub-domain columns, the new AHMIN (in/out:
 * the value before the call goes in) and the damping near walls of u, v and w
 * (near_wall as for gral_pf_flat_init).  wind10 (may be NULL) receives the sums
 * of u and v and their count at about 10 m above ground (the mean GRAMM wind
 * near the ground, :240-249, :614-628), summed in column order.  All arrays
 * are in/out; entries the reference does not write keep their values. */
int gral_pf_terrain_init(const gral_pf_grid *grid, const gral_pf_gramm *gramm, const float *cutk,
                         const uint8_t *advdom, int32_t near_wall, float *uk, float *vk, float *wk,
                         int16_t *kkart, float *ahk, float *bui_height, float *ahmin, int32_t *kadvmax,
                         double *wind10);

/* The same set-up with the original GRAL topography (GRAL_topofile.txt),
 * MicroscaleTerrain.cs:318-598 with the common tail :600-612 and damping
 * :664-832: ahk_ori [nii+2][njj+2] is Program.AHKOri and ahk_ori_min
 * Program.AHKOriMin.  A level is blocked when its centre (float,
 * HOKART[k-1] + DZK[k]/2 + AHKOriMin) is at or below AHKOri + CUTK; the free
 * levels take the interpolated GRAMM wind as the reference's variant does
 * (absolute GRAMM heights in the interior, relative to the incoming *ahmin at
 * the GRAMM border; GRAMM level 1 with level 0 when none lies below).  AHK
 * starts at AHKOriMin; BUI_HEIGHT of building columns comes from the blocked
 * levels at or above AHKOri; the ghosts are the values the reference's serial
 * order leaves; wind10 counts the highest free level below AHKOri + 10 m +
 * DZK/2.  gramm->zsp, ah and zko are not read. */
int gral_pf_terrain_init_topo(const gral_pf_grid *grid, const gral_pf_gramm *gramm, const float *ahk_ori,
                              float ahk_ori_min, const float *cutk, const uint8_t *advdom, int32_t near_wall,
                              float *uk, float *vk, float *wk, int16_t *kkart, float *ahk, float *bui_height,
                              float *ahmin, int32_t *kadvmax, double *wind10);

const char *gral_pf_setup_last_error(void);

/* ---- The particle model's wind on the flow field ----------------------
 * Zeitschleife.IntWindCalculate (Intwind.cs:31-170) for n particles on the
 * handle's current fields (after gral_pf_iterate/solve or a batch): for each
 * particle in flow-field cell (cell_i, cell_j) = (FFCellX, FFCellY) at
 * (x, y, z) with the cell's ground height ahint (AHint), returns UXint,
 * UYint, UZint and IndexK.  topo 1: terrain (Program.Topo ==
 * TerrainAvailable; the level from BinarySearch of z - ahmin, ahmin =
 * Program.AHMIN); topo 0: flat terrain with buildings (the first level whose
 * top is at or above z).  The flat branch without buildings reads no flow
 * field and is not provided.  Cells must lie in 1..NII x 1..NJJ. */
int gral_pf_wind_at_particles(gral_pf_handle *h, int32_t topo, int32_t ikooagral, int32_t jkooagral, float ahmin,
                              int32_t n, const int32_t *cell_i, const int32_t *cell_j, const float *ahint,
                              const double *x, const double *y, const float *z, float *ux, float *uy, float *uz,
                              int32_t *cell_k);

/* ---- Batches of weather situations ------------------------------------
 * GRAL calls Calculate() once per weather situation, one after the other
 * (Program.cs main loop -> MicroscaleFlat.cs:381 / MicroscaleTerrain.cs:847);
 * a project holds hundreds of them.  A batch advances several prepared
 * situations in lock step, each kernel launch covering all of them, so the
 * GPU is filled even though one situation's Gauss-Seidel sweeps expose only
 * one anti-diagonal of columns at a time.  Each situation keeps its own
 * grid size, fields, convergence test and iteration limits, and leaves the
 * batch when the reference's stop test (PrognosticFlowfield.cs:1149-1166)
 * ends it (early retirement); its results are bit-identical to a plain
 * gral_pf_solve of the same situation.  Situations of one batch must use the
 * same turbulence model.  The handles stay owned by the caller and must
 * outlive the batch; after the batch, gral_pf_finish/download_fields work on
 * each handle as after gral_pf_iterate. */
typedef struct gral_pf_batch gral_pf_batch;

/* handles: n prepared situations (gral_pf_prepare), all at the same iteration
 * (n = 0: an empty batch, fed with gral_pf_batch_add). */
int gral_pf_batch_create(gral_pf_handle *const *handles, int32_t n, gral_pf_batch **out);

/* Up to n lock-step SIMPLE iterations; stats (may be NULL) receives one entry
 * per situation, in handle order. */
int gral_pf_batch_iterate(gral_pf_batch *b, int32_t n, int32_t check_convergence, gral_pf_stats *stats);

/* Iterates every situation to its own stop (convergence test or iter_max),
 * then runs gral_pf_finish on each; stats: one entry per situation. */
int gral_pf_batch_solve(gral_pf_batch *b, gral_pf_stats *stats);

/* A batch fed while it runs (the 256-situation workload: more situations than
 * fit in device memory at once).  gral_pf_batch_add queues a prepared
 * situation at its first iteration; gral_pf_batch_step admits the queued
 * ones when no situation is active or when the active ones stand right after
 * a multiple of 100 iterations (the newcomers then run a multiple of 100
 * iterations behind: same sweep directions, same stop-test launches), then
 * advances every active situation to the next multiple of 100 of the leading
 * one with the reference's stop test (PrognosticFlowfield.cs:1149-1166), runs
 * gral_pf_finish on those that stopped (stop test or iter_max) and removes
 * them from the batch: *n_done of them are returned in done[] with their
 * statistics in st[] (room for cap entries; st may be NULL).  A call with
 * nothing active and nothing queued returns *n_done = 0.  The caller owns the
 * handles: a returned handle can be downloaded and destroyed, freeing its
 * device memory for the next situations.  Each situation's results equal its
 * plain solve bit for bit. */
int gral_pf_batch_add(gral_pf_batch *b, gral_pf_handle *h);
int gral_pf_batch_step(gral_pf_batch *b, int32_t *n_done, gral_pf_handle **done, gral_pf_stats *st, int32_t cap);
int gral_pf_batch_size(const gral_pf_batch *b, int32_t *active, int32_t *queued);

/* As gral_pf_profile_iterations, over the whole batch. */
int gral_pf_batch_profile_iterations(gral_pf_batch *b, int32_t n, gral_pf_family_time *out, int32_t nfam);
int gral_pf_batch_last_timing(gral_pf_batch *b, gral_pf_timing *t);
const char *gral_pf_batch_last_error(const gral_pf_batch *b);
void gral_pf_batch_destroy(gral_pf_batch *b);

#ifdef __cplusplus
}
#endif
#endif
