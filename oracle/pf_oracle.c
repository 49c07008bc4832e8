/*
 * pf_oracle.c -- deterministic CPU restatement of GRAL 24.11's prognostic
 * microscale flow solver.  TEST INFRASTRUCTURE ONLY (the parity checker of
 * the HIP solver); the product library never links or calls this file.
 *
 * Semantics followed (all paths relative to the reference's src/):
 *   - PrognosticFlowfield.cs:32-1245 (prologue, SIMPLE loop, epilogue checks)
 *     in deterministic mode (UseFixedRndSeedVal: MaxDegreeOfParallelism = 1
 *     around the pressure sweep and the momentum solvers, :806-811, :1058-1136)
 *   - U/V/W-prognostic-microscale_1.cs (Vector<float>, width 8, no FMA)
 *   - U/V/W-prognostic-microscale_1_Vec512.cs (Vector512, width 16, the FMA
 *     placements of those files; selected by simd_width 16)
 *   - U/V/W-prognostic-microscale_2.cs, TKE-prognostic-microscale.cs (k-eps)
 *   - U/V/W-prognostic-microscale_0.cs (no diffusion)
 *   - ProgramFunctions.cs:768-849 (FloatMax, ConvToInt, Pow2/3/5)
 *
 * Floating point: compile with -ffp-contract=off and without -ffast-math.
 * C# evaluates float expressions in single precision and promotes to double
 * as soon as a double literal or operand appears; every such promotion is
 * written out explicitly below.  Vector.Max/Min and FloatMax have the SSE
 * maxps/minps semantics (a > b ? a : b); Math.Max/Min(float) are the
 * IEEE-754:2019 maximum/minimum of .NET Core.
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include "pf_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdbool.h>

/* ---- helpers (ProgramFunctions.cs:768-849) ---------------------------- */
static inline float vmax(float a, float b) { return a > b ? a : b; }   /* Vector.Max, FloatMax */
static inline float vmin(float a, float b) { return a < b ? a : b; }   /* Vector.Min */
static inline float mmaxf_(float a, float b) {                         /* Math.Max(float,float) */
    if (a != b) { if (!isnan(a)) return b < a ? a : b; return a; }
    return signbit(b) ? a : b;
}
static inline float mminf_(float a, float b) {                         /* Math.Min(float,float) */
    if (a != b) { if (!isnan(a)) return a < b ? a : b; return a; }
    return signbit(a) ? a : b;
}
static inline double mmaxd_(double a, double b) {                      /* Math.Max(double,double) */
    if (a != b) { if (!isnan(a)) return b < a ? a : b; return a; }
    return signbit(b) ? a : b;
}
static inline int conv_to_int_d(double x) { return x < 0 ? (int)(x - 0.5) : (int)(x + 0.5); }
static inline float p2f(float x) { return x * x; }
static inline double p2d(double x) { return x * x; }
static inline float p3f(float x) { return x * x * x; }
static inline float p5f(float x) { return x * x * x * x * x; }
static inline double p5d(double x) { return x * x * x * x * x; }

/* Planted mistakes for the oracle's own tests (tests/test_oracle_mutants.py
 * builds the file with -DORC_MUTANT=n to check that the closed-form tests
 * catch each one).  The default build has ORC_MUTANT = 0: no mistake.
 *   1 eddy viscosity without the 1/2 under the root    5 w-equation x flux from level k+1
 *   2 mixing length from the level top                 6 TKE production without (v_z + w_y)^2
 *   3 east power-law coefficient upwinded the wrong way 7 divergence with the x flux sign swapped
 *   4 wall function at z1 = dz instead of dz/2         8 boundary copy regardless of the flow sign */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

/* ---- state ---------------------------------------------------------- */
typedef struct {
    int NII, NJJ, NKK, KADVMAX, VCF, NJ2, NKP, W, model;
    float DXK, DYK, AREAxy, DTIME, relax, relaxp, building_Z0, Cor, PotdTdz;
    float UG, VG, Ustern_factorX, VISHMIN, VISVMIN;
    double Frund;
    const float *DZK, *HOKART;
    const uint8_t *ADVDOM; const int16_t *KKART_; const float *CUTK, *COV;
    const float *Z0S;                 /* solver Z0 per column */
    const float *VEG; const uint8_t *VEGP;
    float *UK, *VK, *WK, *UKS, *VKS, *WKS, *DPM, *DPMNEW, *DIV, *TURB, *TDISS;
    int *VI; float *HT, *HO;          /* VerticalIndex, Ustern terrain/obstacle helpterms */
    float *AP0;
    float *P1, *Q1, *P2, *Q2;         /* column scratch (PIMU/QIMU, PIMTDISS...) */
    orc_capture *cap;                 /* optional checking hook */
    int rec;                          /* the lane being evaluated enters the recurrence (hook only) */
    int iter, comp;                   /* current SIMPLE iteration and sweep (for the hook) */
    int nthreads;                     /* > 1: the timing mode of orc_calculate_threads */
    int proj;                         /* orc_project: divergence, pressure and correction only */
    int noquant;                      /* orc_project with quantise = 0: no FP64 quantisation */
} St;

static bool cap_on(const St *s, int comp, int i, int j) {
    const orc_capture *c = s->cap;
    return c && c->comp == comp && c->i == i && c->j == j && c->iteration == s->iter;
}

static void cap_column(St *s, int comp, int i, int j, const float *x, bool after) {
    if (!cap_on(s, comp, i, j)) return;
    float *dst = after ? s->cap->x_new : s->cap->x_old;
    memcpy(dst, x, sizeof(float) * (size_t)s->NKP);
    if (after) s->cap->captured = 1;
}


/* record piece id of level kk of the captured column */
static void piece(St *s, int i, int j, int id, int kk, float v) {
    if (!s->rec || !s->cap || !s->cap->piece || !cap_on(s, s->comp, i, j)) return;
    s->cap->piece[(size_t)id * (size_t)(s->NKK + 2) + (size_t)kk] = v;
}

#define IX2(s, i, j) ((size_t)(i) * (size_t)(s)->NJ2 + (size_t)(j))
#define IX3(s, i, j, k) (IX2(s, i, j) * (size_t)(s)->NKP + (size_t)(k))
#define COL(s, a, i, j) (&(s)->a[IX2(s, i, j) * (size_t)(s)->NKP])
#define KK(s, i, j) ((int)(s)->KKART_[IX2(s, i, j)])
#define ADV(s, i, j) ((s)->ADVDOM[IX2(s, i, j)])

/* ===================================================================== */
/*  Sub-domain boundary conditions, PrognosticFlowfield.cs:595-714         */
/* ===================================================================== */
static void boundary_conditions(St *s) {
    int NII = s->NII, NJJ = s->NJJ;
    for (int j = 2; j <= NJJ; ++j) {                 /* Parallel.For(2, NJJ + 1) */
        for (int i = 3; i < NII; ++i) {
            if (ADV(s, i, j) != 1) continue;
            int VIc = s->VI[IX2(s, i, j)];
            bool iM = (ADV(s, i - 1, j) < 1) || (i == 3);
            bool iP = (ADV(s, i + 1, j) < 1) || (i == NII - 1);
            if (iM) {
                for (int k = 1; k <= VIc; ++k) {
                    float u = s->UK[IX3(s, i - 2, j, k)];
                    if (u >= 0 || ORC_MUTANT == 8) s->UK[IX3(s, i - 1, j, k)] = u;
                    s->VK[IX3(s, i - 2, j, k)] = s->VK[IX3(s, i - 1, j, k)];
                    s->WK[IX3(s, i - 2, j, k)] = s->WK[IX3(s, i - 1, j, k)];
                }
            }
            if (iP) {
                for (int k = 1; k <= VIc; ++k) {
                    float u = s->UK[IX3(s, i + 2, j, k)];
                    if (u <= 0) s->UK[IX3(s, i + 1, j, k)] = u;
                    s->VK[IX3(s, i + 1, j, k)] = s->VK[IX3(s, i, j, k)];
                    s->WK[IX3(s, i + 1, j, k)] = s->WK[IX3(s, i, j, k)];
                }
            }
        }
    }
    for (int i = 2; i <= NII; ++i) {                 /* Parallel.For(2, NII + 1) */
        for (int j = 3; j <= NJJ - 1; ++j) {
            if (ADV(s, i, j) != 1) continue;
            int VIc = s->VI[IX2(s, i, j)];
            bool jM = (ADV(s, i, j - 1) < 1) || (j == 3);
            bool jP = (ADV(s, i, j + 1) < 1) || (j == NJJ - 1);
            if (jM) {
                for (int k = 1; k <= VIc; ++k) {
                    float v = s->VK[IX3(s, i, j - 2, k)];
                    if (v >= 0) s->VK[IX3(s, i, j - 1, k)] = v;
                    s->UK[IX3(s, i, j - 2, k)] = s->UK[IX3(s, i, j - 1, k)];
                    s->WK[IX3(s, i, j - 2, k)] = s->WK[IX3(s, i, j - 1, k)];
                }
            }
            if (jP) {
                for (int k = 1; k <= VIc; ++k) {
                    float v = s->VK[IX3(s, i, j + 2, k)];
                    if (v <= 0) s->VK[IX3(s, i, j + 1, k)] = v;
                    s->UK[IX3(s, i, j + 1, k)] = s->UK[IX3(s, i, j, k)];
                    s->WK[IX3(s, i, j + 1, k)] = s->WK[IX3(s, i, j, k)];
                }
            }
        }
    }
    for (int i = 2; i < NII; ++i) {
        for (int j = 2; j <= NJJ - 1; ++j) {
            if (ADV(s, i, j) != 1) continue;
            int k = s->VI[IX2(s, i, j)];
            s->WK[IX3(s, i, j, k + 1)] = s->WK[IX3(s, i, j, k)];
            s->UK[IX3(s, i, j, k + 1)] = s->UK[IX3(s, i, j, k)];
            s->VK[IX3(s, i, j, k + 1)] = s->VK[IX3(s, i, j, k)];
        }
    }
}

/* Mass divergence, PrognosticFlowfield.cs:723-758 */
static void divergence(St *s) {
#pragma omp parallel for schedule(static) if (s->nthreads > 1) num_threads(s->nthreads > 1 ? s->nthreads : 1)
    for (int i = 2; i < s->NII; ++i)
        for (int j = 2; j < s->NJJ; ++j) {
            if (ADV(s, i, j) != 1) continue;
            float *DIV = COL(s, DIV, i, j), *UK = COL(s, UK, i, j), *VK = COL(s, VK, i, j);
            float *WK = COL(s, WK, i, j), *DPM = COL(s, DPM, i, j);
            float *UKip = COL(s, UK, i + 1, j), *VKjp = COL(s, VK, i, j + 1);
            int KK = KK(s, i, j), VIc = s->VI[IX2(s, i, j)];
            for (int k = KK + 1; k <= VIc; ++k) {
                if (k > KK + 1)
                    DIV[k] = (UK[k] - UKip[k]) * s->DYK * s->DZK[k] + (VK[k] - VKjp[k]) * s->DXK * s->DZK[k] + (WK[k] - WK[k + 1]) * s->AREAxy;
                else
                    DIV[k] = (UK[k] - UKip[k]) * s->DYK * s->DZK[k] + (VK[k] - VKjp[k]) * s->DXK * s->DZK[k] - WK[k + 1] * s->AREAxy;
                if (ORC_MUTANT == 7) DIV[k] += 2 * (UKip[k] - UK[k]) * s->DYK * s->DZK[k];
                DPM[k] = 0;
            }
        }
}

static float pressure_row(St *s, int IS, int JS, int j1, float *PIMP, float *QIMP);

/* Pressure correction GS-TDMA, PrognosticFlowfield.cs:804-901.
 * PIMP/QIMP are per j-row (fresh zeros for each row), reused across the
 * row's columns: the stale top value PIMP[VI+1] is reproduced on purpose. */
static float pressure(St *s, int IS, int JS) {
    int NJJ = s->NJJ;
    float DeltaUMax = 0;
    float *PIMP = (float *)calloc((size_t)s->NKK + 2, sizeof(float));
    float *QIMP = (float *)calloc((size_t)s->NKK + 2, sizeof(float));
    for (int j1 = 2; j1 < NJJ; ++j1) {
        float DeltaUMaxIntern = pressure_row(s, IS, JS, j1, PIMP, QIMP);
        if (DeltaUMaxIntern > DeltaUMax) DeltaUMax = DeltaUMaxIntern;
    }
    free(PIMP); free(QIMP);
    return DeltaUMax;
}

/* one j row of the pressure sweep with fresh row arrays PIMP/QIMP */
static float pressure_row(St *s, int IS, int JS, int j1, float *PIMP, float *QIMP) {
    int NII = s->NII, NJJ = s->NJJ;
    float DXK = s->DXK, DYK = s->DYK, AREAxy = s->AREAxy, DTIME = s->DTIME;
    const float *DZK = s->DZK;
    {
        memset(PIMP, 0, sizeof(float) * ((size_t)s->NKK + 2));
        memset(QIMP, 0, sizeof(float) * ((size_t)s->NKK + 2));
        float DeltaUMaxIntern = 0;
        int j = (JS == -1) ? NJJ - j1 + 1 : j1;
        for (int i1 = 2; i1 <= NII - 1; ++i1) {
            int i = (IS == -1) ? NII - i1 + 1 : i1;
            if (ADV(s, i, j) != 1) continue;
            float *DIV = COL(s, DIV, i, j), *DPM = COL(s, DPM, i, j);
            float *DPMim = COL(s, DPM, i - 1, j), *DPMip = COL(s, DPM, i + 1, j);
            float *DPMjm = COL(s, DPM, i, j - 1), *DPMjp = COL(s, DPM, i, j + 1);
            float *DPMNEW = COL(s, DPMNEW, i, j);
            int KK = KK(s, i, j), VIc = s->VI[IX2(s, i, j)];
            for (int k = VIc; k > KK; --k) {
                float DIM = DIV[k] / DTIME + (DPMim[k] + DPMip[k]) / DXK * DYK * DZK[k] +
                            (DPMjm[k] - DPMjp[k]) / DYK * DXK * DZK[k];
                float APP = 2.0f * (DYK * DZK[k] / DXK + DXK * DZK[k] / DYK + AREAxy / DZK[k]);
                float ABP;
                if (k > 1) ABP = AREAxy / (0.5f * (DZK[k - 1] + DZK[k]));
                else ABP = AREAxy / DZK[k];
                float ATP = AREAxy / (0.5f * (DZK[k + 1] + DZK[k]));
                float TERMP = 1 / (APP - ATP * PIMP[k + 1]);
                PIMP[k] = ABP * TERMP;
                QIMP[k] = (ATP * QIMP[k + 1] + DIM) * TERMP;
                if (cap_on(s, 3, i, j)) { s->cap->a[k] = APP; s->cap->b[k] = ABP; s->cap->c[k] = ATP; s->cap->d[k] = DIM; }
            }
            cap_column(s, 3, i, j, DPM, false);
            for (int k = KK + 1; k <= VIc; ++k) {
                float temp = PIMP[k] * DPM[k - 1] + QIMP[k];
                double q = (double)conv_to_int_d((double)temp * s->Frund) / s->Frund;
                DeltaUMaxIntern = (float)mmaxd_((double)DeltaUMaxIntern, fabs((double)DPM[k] - q));
                DPM[k] = temp;
                DPMNEW[k] += s->relaxp * DPM[k];
            }
            cap_column(s, 3, i, j, DPM, true);
        }
        return DeltaUMaxIntern;
    }
}

/* Velocity correction with FP64 quantisation, PrognosticFlowfield.cs:904-949 */
static void correction(St *s) {
    double Frund = s->Frund;
#pragma omp parallel for schedule(static) if (s->nthreads > 1) num_threads(s->nthreads > 1 ? s->nthreads : 1)
    for (int i = 2; i < s->NII; ++i)
        for (int j = 2; j <= s->NJJ - 1; ++j) {
            if (!(ADV(s, i, j) == 1 || ADV(s, i - 1, j) == 1 || ADV(s, i, j - 1) == 1)) continue;
            float *DPMim = COL(s, DPM, i - 1, j), *DPMjm = COL(s, DPM, i, j - 1), *DPM = COL(s, DPM, i, j);
            float *UK = COL(s, UK, i, j), *VK = COL(s, VK, i, j), *WK = COL(s, WK, i, j);
            int KK = KK(s, i, j), VIc = s->VI[IX2(s, i, j)];
            int KKiM = KK(s, i - 1, j), KKjM = KK(s, i, j - 1);
            for (int k = 1; k <= VIc; ++k) {
                float DPMk = DPM[k];
                if (i > 2 && KK < k && KKiM < k) {
                    float e = UK[k] + (DPMim[k] - DPMk) / s->DXK * s->DTIME;
                    UK[k] = s->noquant ? e : (float)((double)conv_to_int_d((double)e * Frund) / Frund);
                }
                if (j > 2 && KK < k && KKjM < k) {
                    float e = VK[k] + (DPMjm[k] - DPMk) / s->DYK * s->DTIME;
                    VK[k] = s->noquant ? e : (float)((double)conv_to_int_d((double)e * Frund) / Frund);
                }
                if (KK + 1 < k) {
                    float e = WK[k] + (DPM[k - 1] - DPMk) / (s->DZK[k] + s->DZK[k - 1]) * 2.0f * s->DTIME;
                    WK[k] = s->noquant ? e : (float)((double)conv_to_int_d((double)e * Frund) / Frund);
                } else {
                    WK[k] = 0;
                }
            }
        }
}

/* Cell-centred averages and edge copies, PrognosticFlowfield.cs:958-1039 */
static void averages(St *s) {
    int NII = s->NII, NJJ = s->NJJ;
#pragma omp parallel for schedule(static) if (s->nthreads > 1) num_threads(s->nthreads > 1 ? s->nthreads : 1)
    for (int i = 2; i < NII; ++i)
        for (int j = 2; j <= NJJ - 1; ++j) {
            if (!(ADV(s, i, j) == 1 || ADV(s, i - 1, j) == 1 || ADV(s, i, j - 1) == 1)) continue;
            float *UK = COL(s, UK, i, j), *UKip = COL(s, UK, i + 1, j), *UKS = COL(s, UKS, i, j);
            float *VK = COL(s, VK, i, j), *VKjp = COL(s, VK, i, j + 1), *VKS = COL(s, VKS, i, j);
            float *WK = COL(s, WK, i, j), *WKS = COL(s, WKS, i, j);
            int KK = KK(s, i, j), VIc = s->VI[IX2(s, i, j)];
            int KSTART = (s->CUTK[IX2(s, i, j)] == 0) ? KK + 1 : 1;
            for (int k = KSTART; k <= VIc; ++k) {
                if (KK < k) {
                    UKS[k] = 0.5f * (UK[k] + UKip[k]);
                    VKS[k] = 0.5f * (VK[k] + VKjp[k]);
                }
                if (KK + 1 < k) WKS[k] = 0.5f * (WK[k] + WK[k + 1]);
            }
            if (j == 2) {
                float *U1 = COL(s, UKS, i, 1);
                for (int k = KSTART; k <= VIc; ++k) U1[k] = UKS[k];
            }
            if (j == NJJ - 1 && ADV(s, i, NJJ) == 1) {
                float *UN = COL(s, UKS, i, NJJ);
                for (int k = KSTART; k <= VIc; ++k) UN[k] = UKS[k];
            }
        }
    for (int j = 2; j < NJJ; ++j) {
        if (ADV(s, 1, j) == 1) {
            float *a = COL(s, VKS, 1, j), *b = COL(s, VKS, 2, j);
            for (int k = 0; k <= s->NKK; ++k) a[k] = b[k];
        }
        if (ADV(s, NII - 1, j) == 1) {
            float *a = COL(s, VKS, NII, j), *b = COL(s, VKS, NII - 1, j);
            for (int k = 0; k <= s->NKK; ++k) a[k] = b[k];
        }
    }
}

/* ===================================================================== */
/*  Mixing-length model (turbulence model 1), Vector<float> width W         */
/* ===================================================================== */
/* Common SIMD-block driver.  For each block start k (KSTART, KSTART+W, ...,
 * the last one moved down to KEND-W) the lanes kk = k..k+W-1 are evaluated
 * with the block-level conditions of the vector code, then the lanes not yet
 * consumed enter the Thomas recurrence (U-prognostic-microscale_1.cs:194-511). */
typedef void (*lane_fn)(St *s, int i, int j, int k, int kk, float *A, float *B, float *C, float *D, void *ctx);
typedef struct { float UG; } uctx;   /* UG (U) or VG (V) */
#if defined(ORC_SIMD) && ORC_MUTANT == 0
static void u1_lane(St *s, int i, int j, int k, int kk, float *pA, float *pB, float *pC, float *pD, void *vctx);
static void v1_lane(St *s, int i, int j, int k, int kk, float *pA, float *pB, float *pC, float *pD, void *vctx);
static void ml_block(St *s, int i, int j, int k, lane_fn fn, void *ctx, float *A, float *B, float *C, float *D);
static int use_blocks;
#endif

static void ml_column(St *s, int i, int j, lane_fn fn, void *ctx) {
    int W = s->W;
    int KK = KK(s, i, j), KKP1 = KK + 1;
    int KSTART = (s->CUTK[IX2(s, i, j)] == 0) ? KK + 1 : 1;
    int KEND = s->VI[IX2(s, i, j)] + 1;
    float *P = s->P1, *Q = s->Q1;
    float A[16], B[16], C[16], D[16];
    bool end = false;
    int k_real = 0;
    for (int k = KSTART; k < KEND; k += W) {
        k_real = k;
        if (k > KEND - W) { k = KEND - W; end = true; }
        const int used0 = end ? k_real - k : 0;     /* lanes below it were consumed by the previous block */
#if defined(ORC_SIMD) && ORC_MUTANT == 0
        if (use_blocks && W == 8 && !s->cap) ml_block(s, i, j, k, fn, ctx, A, B, C, D);
        else
#endif
        for (int l = 0; l < W; ++l) {
            s->rec = l >= used0 && k + l < KEND;
            fn(s, i, j, k, k + l, &A[l], &B[l], &C[l], &D[l], ctx);
        }
        s->rec = 0;
        int kint_s = 0;
        if (end) { kint_s = k_real - k; k_real -= kint_s; }
        for (int kint = kint_s; kint < W; ++kint) {
            int idx = k_real + kint;
            if (idx < KEND && cap_on(s, s->comp, i, j)) {
                s->cap->a[idx] = A[kint]; s->cap->b[idx] = B[kint]; s->cap->c[idx] = C[kint]; s->cap->d[idx] = D[kint];
            }
            if (idx < KEND) {
                if (idx > KKP1) {
                    float TERMP = 1 / (A[kint] - C[kint] * P[idx - 1]);
                    P[idx] = B[kint] * TERMP;
                    Q[idx] = (D[kint] + C[kint] * Q[idx - 1]) * TERMP;
                } else {
                    float TERMP = 1 / A[kint];
                    P[idx] = B[kint] * TERMP;
                    Q[idx] = D[kint] * TERMP;
                }
            }
        }
        if (end) k = KEND + 1;
    }
    (void)KSTART;
}

/* shared mixing-length eddy viscosity of the lane (U/V/W_1.cs, e.g. U:216-268) */
static float ml_vis(St *s, int i, int j, int k, int kk) {
    int W = s->W, KK = KK(s, i, j), KKP1 = KK + 1;
    float DZ = s->DZK[kk];
    /* the 512-bit solvers fuse the half-cell offset into the ground height
     * (U-prognostic-microscale_1_Vec512.cs:218) */
    float zs = (W == 16) ? s->HOKART[kk] - fmaf(DZ, 0.5f, s->HOKART[KK]) : s->HOKART[kk] - s->HOKART[KK] - DZ * 0.5f;
    if (ORC_MUTANT == 2) zs = s->HOKART[kk] - s->HOKART[KK];
    bool mask_v_enable = (KKP1 >= k && KKP1 < k + W);
    float mask = (kk > KKP1) ? 1.0f : 0.0f;
    float VIS = 0;
    if (k + W > KKP1) {
        float mixL = vmin(0.071f * zs, 100.0f);
        float COV = s->COV[IX2(s, i, j)];
        if (COV > 0) mixL *= (1.0f - 0.99f * COV);
        const float *UK = COL(s, UK, i, j), *VK = COL(s, VK, i, j);
        float in = (UK[kk + 1] - UK[kk - 1]) / (2.0f * DZ);
        float DUDZ = in * in;
        in = (VK[kk + 1] - VK[kk - 1]) / (2.0f * DZ);
        float DVDZ = in * in;
        float v = vmin(mixL * mixL * sqrtf((ORC_MUTANT == 1 ? 1.0f : 0.5f) * (DUDZ + DVDZ)), 15.0f);
        VIS = mask_v_enable ? v * mask : v;
    }
    return VIS;
}

/* power-law coefficient, Vector form: D*max(0,(1-0.1P)^5) */
static inline float ml_pow(float Dv, float Pv) {
    float in = 1.0f - 0.1f * Pv;
    return Dv * vmax(0.0f, in * in * in * in * in);
}

/* power-law coefficient plus its upwind flux F+ = max(+-F, 0).  Vector<float>
 * solvers: D*max(0, x*x*x*x*x) + F+ (U-prognostic-microscale_1.cs:345-367);
 * 512-bit solvers: FMA(D, max(0, x*((x*x)*(x*x))), F+), x = 1 - 0.1P
 * (U-prognostic-microscale_1_Vec512.cs:355-377) */
static inline float ml_coef(int W, float Dv, float Pv, float Fpos) {
    if (W != 16) return ml_pow(Dv, Pv) + Fpos;
    float in = 1.0f - 0.1f * Pv;
    return fmaf(Dv, vmax(0.0f, in * ((in * in) * (in * in))), Fpos);
}

/* neighbour part AW*XW + AS*XS + AE*XE + AN*XN of a source term, summed left to
 * right, or as FMA(AW,XW,AS*XS) + FMA(AE,XE,AN*XN) where the 512-bit solvers
 * fuse it (U-prognostic-microscale_1_Vec512.cs:404-422) */
static inline float ml_nb(bool fused, float aw, float xw, float as, float xs, float ae, float xe, float an, float xn) {
    if (fused) return fmaf(aw, xw, as * xs) + fmaf(ae, xe, an * xn);
    return aw * xw + as * xs + ae * xe + an * xn;
}

/* the checking hook's record of one level's conductances, fluxes and coefficients */
static void ml_pieces(St *s, int i, int j, int kk, float VIS, float DE, float DW, float DS, float DN, float DT,
                      float DB, float FE, float FW, float FS, float FN, float FT, float FB, float AE, float AW,
                      float AS, float AN, float AT, float AB) {
    if (!s->rec || !s->cap || !s->cap->piece || !cap_on(s, s->comp, i, j)) return;
    const float v[] = {VIS, DE, DW, DS, DN, DT, DB, FE, FW, FS, FN, FT, FB, AE, AW, AS, AN, AT, AB};
    for (int q = 0; q <= ORC_AB; ++q) piece(s, i, j, q, kk, v[q]);
}

/* One lane of U_PrognosticMicroscaleV1 (U-prognostic-microscale_1.cs:194-470) */
static void u1_lane(St *s, int i, int j, int k, int kk, float *pA, float *pB, float *pC, float *pD, void *vctx) {
    float UG = ((uctx *)vctx)->UG;
    int W = s->W, KK = KK(s, i, j), KKP1 = KK + 1;
    int iM1 = i - 1, iP1 = i + 1, jM1 = j - 1, jP1 = j + 1;
    float DXK = s->DXK, DYK = s->DYK, AREAxy = s->AREAxy;
    float DZ = s->DZK[kk], DXKDZK = DXK * DZ, DYKDZK = DYK * DZ;
    bool mask_v_enable = (KKP1 >= k && KKP1 < k + W);
    float mask = (kk > KKP1) ? 1.0f : 0.0f;
    const float *UK = COL(s, UK, i, j), *UKS = COL(s, UKS, i, j), *VKS = COL(s, VKS, i, j), *WKS = COL(s, WKS, i, j);
    const float *DPMNEW = COL(s, DPMNEW, i, j), *UK_W = COL(s, UK, iM1, j), *UK_E = COL(s, UK, iP1, j);
    const float *UK_S = COL(s, UK, i, jM1), *UK_N = COL(s, UK, i, jP1), *UKS_W = COL(s, UKS, iM1, j);
    const float *VKS_W = COL(s, VKS, iM1, j), *VKS_WS = COL(s, VKS, iM1, jM1), *VKS_S = COL(s, VKS, i, jM1);
    const float *VKS_WN = COL(s, VKS, iM1, jP1), *VKS_N = COL(s, VKS, i, jP1), *WKS_W = COL(s, WKS, iM1, j);
    const float *DPMNEW_W = COL(s, DPMNEW, iM1, j), *VKS_EN = COL(s, VKS, iP1, jP1), *VKS_ES = COL(s, VKS, iP1, jM1);
    const float *WKS_E = COL(s, WK, iP1, j);          /* WK, not WKS (quirk) */
    float VIS = ml_vis(s, i, j, k, kk);
    float mv = vmax(VIS, s->VISHMIN);
    float DE = mv * DYKDZK / DXK, DW = DE;
    float DS = mv * DXKDZK / DYK, DN = DS;
    float DT = mv * AREAxy / DZ;
    float DB = 0;
    if (mask_v_enable) DB = DT * mask;
    else if (k > KKP1) DB = DT;
    if ((ADV(s, i, jM1) < 1) || (j == 2)) DS = 0;
    if ((ADV(s, i, jP1) < 1) || (j == s->NJJ - 1)) DN = 0;
    if ((ADV(s, iM1, j) < 1) || (i == 2)) DW = 0;
    if ((ADV(s, iP1, j) < 1) || (i == s->NII - 1)) DE = 0;
    float FE = UKS[kk] * DYKDZK;
    float FW = UKS_W[kk] * DYKDZK;
    float FS = 0.25f * (VKS_W[kk] + VKS[kk] + VKS_WS[kk] + VKS_S[kk]) * DXKDZK;
    float FN = 0.25f * (VKS_W[kk] + VKS[kk] + VKS_WN[kk] + VKS_N[kk]) * DXKDZK;
    float FT = 0.25f * (WKS_W[kk] + WKS[kk] + WKS_W[kk + 1] + WKS[kk + 1]) * AREAxy;
    float FB = 0;
    if (mask_v_enable) FB = 0.25f * (WKS_W[kk] + WKS[kk] + WKS_W[kk - 1] + WKS[kk - 1]) * AREAxy * mask;
    else if (k > KKP1) FB = 0.25f * (WKS_W[kk] + WKS[kk] + WKS_W[kk - 1] + WKS[kk - 1]) * AREAxy;
    DE = vmax(DE, 0.0001f); DB = vmax(DB, 0.0001f); DW = vmax(DW, 0.0001f);
    DS = vmax(DS, 0.0001f); DN = vmax(DN, 0.0001f); DT = vmax(DT, 0.0001f);
    float PE = fabsf(FE / DE), PB = fabsf(FB / DB), PW = fabsf(FW / DW);
    float PS = fabsf(FS / DS), PN = fabsf(FN / DN), PT = fabsf(FT / DT);
    float BIM = ml_coef(W, DT, PT, vmax(-FT, 0.0f));
    float CIM = ml_coef(W, DB, PB, vmax(FB, 0.0f));
    float AE1 = ml_coef(W, DE, PE, vmax(ORC_MUTANT == 3 ? FE : -FE, 0.0f));
    float AW1 = ml_coef(W, DW, PW, vmax(FW, 0.0f));
    float AS1 = ml_coef(W, DS, PS, vmax(FS, 0.0f));
    float AN1 = ml_coef(W, DN, PN, vmax(-FN, 0.0f));
    float AP0 = s->AP0[kk];
    float AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + AP0;
    ml_pieces(s, i, j, kk, VIS, DE, DW, DS, DN, DT, DB, FE, FW, FS, FN, FT, FB, AE1, AW1, AS1, AN1, BIM, CIM);
    float DDPX = DPMNEW_W[kk] - DPMNEW[kk];
    float UKv = UK[kk];
    float DIM;
    piece(s, i, j, ORC_PRESS, kk, DDPX * DYKDZK);
    piece(s, i, j, ORC_COR, kk, s->Cor * (UG - UKv) * AREAxy * DZ);
    if (s->VEGP && s->VEGP[IX2(s, i, j)]) {
        float in2 = 0.5f * (VKS_W[kk] + VKS[kk]);
        float in = 0.5f * (UKS_W[kk] + UKS[kk]);
        float wh = vmax(sqrtf(in * in + in2 * in2), 0.01f);
        float VEGv = s->VEG[IX3(s, i, j, kk - 1)];
        piece(s, i, j, ORC_VEGT, kk, -VEGv * UKv * wh * AREAxy * DZ);
        DIM = ml_nb(W == 16, AW1, UK_W[kk], AS1, UK_S[kk], AE1, UK_E[kk], AN1, UK_N[kk]) +
              AP0 * 0.5f * (UKS_W[kk] + UKS[kk]) + DDPX * DYKDZK +
              (s->Cor * (UG - UKv) - VEGv * UKv * wh) * AREAxy * DZ;
    } else if (W == 16) {
        /* U-prognostic-microscale_1_Vec512.cs:415-421: the pressure term is fused too */
        DIM = ml_nb(true, AW1, UK_W[kk], AS1, UK_S[kk], AE1, UK_E[kk], AN1, UK_N[kk]) +
              AP0 * 0.5f * (UKS_W[kk] + UKS[kk]) + fmaf(DDPX, DYKDZK, s->Cor * (UG - UKv) * AREAxy * DZ);
    } else {
        DIM = AW1 * UK_W[kk] + AS1 * UK_S[kk] + AE1 * UK_E[kk] + AN1 * UK_N[kk] +
              AP0 * 0.5f * (UKS_W[kk] + UKS[kk]) + DDPX * DYKDZK + s->Cor * (UG - UKv) * AREAxy * DZ;
    }
    if ((k <= KKP1) && ((k + W) > KKP1)) {
        int ii = KKP1 - k;
        float m2 = (kk == KKP1) ? 1.0f : 0.0f;
        float in = 0.5f * (UKS_W[kk] + UKS[kk]);
        float in2 = 0.5f * (VKS_W[kk] + VKS[kk]);
        float wh = vmax(sqrtf(in * in + in2 * in2), 0.1f);
        /* scalars taken from lane ii (k + ii == KKART+1) */
        float inI = 0.5f * (UKS_W[KKP1] + UKS[KKP1]);
        float in2I = 0.5f * (VKS_W[KKP1] + VKS[KKP1]);
        float whI = vmax(sqrtf(inI * inI + in2I * in2I), 0.1f);
        float Us;
        if (s->CUTK[IX2(s, i, j)] < 1) {
            Us = s->HT[IX2(s, i, j)] * whI;
            if (s->Z0S[IX2(s, i, j)] >= s->DZK[k + ii] * 0.1f) {
                int ind = k + ii + 1;
                Us = s->HT[IX2(s, i, j)] * sqrtf(p2f(0.5f * (UKS_W[ind] + UKS[ind])) + p2f(0.5f * (VKS_W[ind] + VKS[ind])));
            }
        } else {
            Us = s->HO[IX2(s, i, j)] * whI;
        }
        DIM -= UKv / wh * (Us * Us) * AREAxy * m2;
        piece(s, i, j, ORC_WALL, kk, UKv / wh * (Us * Us) * AREAxy * m2);
    }
    if (k > KKP1) {
        float DUDXE = (UK_E[kk] - UKv) * DYKDZK;
        float DUDXW = (UKv - UK_W[kk]) * DYKDZK;
        float DVDXN = (0.5f * (VKS_EN[kk] + VKS_N[kk]) - 0.5f * (VKS_N[kk] + VKS_WN[kk])) * DXKDZK;
        float DVDXS = (0.5f * (VKS_ES[kk] + VKS_S[kk]) - 0.5f * (VKS_S[kk] + VKS_WS[kk])) * DXKDZK;
        float DWDXT = (0.5f * (WKS_E[kk] + WKS[kk]) - 0.5f * (WKS[kk] + WKS_W[kk])) * AREAxy;
        float DWDXB = (0.5f * (WKS_E[kk - 1] + WKS[kk - 1]) - 0.5f * (WKS[kk - 1] + WKS_W[kk - 1])) * AREAxy;
        float cross = VIS / DXK * (DUDXE - DUDXW + DVDXN - DVDXS + DWDXT - DWDXB);
        if (mask_v_enable) cross = cross * mask;
        piece(s, i, j, ORC_CROSS, kk, cross);
        DIM += cross;
    }
    *pA = AIM; *pB = BIM; *pC = CIM; *pD = DIM;
}

/* One lane of V_PrognosticMicroscaleV1 (V-prognostic-microscale_1.cs:194-464) */
static void v1_lane(St *s, int i, int j, int k, int kk, float *pA, float *pB, float *pC, float *pD, void *vctx) {
    float VG = ((uctx *)vctx)->UG;
    int W = s->W, KK = KK(s, i, j), KKP1 = KK + 1;
    int iM1 = i - 1, iP1 = i + 1, jM1 = j - 1, jP1 = j + 1;
    float DXK = s->DXK, DYK = s->DYK, AREAxy = s->AREAxy;
    float DZ = s->DZK[kk], DXKDZK = DXK * DZ, DYKDZK = DYK * DZ;
    bool mask_v_enable = (KKP1 >= k && KKP1 < k + W);
    float mask = (kk > KKP1) ? 1.0f : 0.0f;
    const float *VK = COL(s, VK, i, j), *UKS = COL(s, UKS, i, j), *VKS = COL(s, VKS, i, j), *WKS = COL(s, WKS, i, j);
    const float *DPMNEW = COL(s, DPMNEW, i, j), *VK_W = COL(s, VK, iM1, j), *VK_E = COL(s, VK, iP1, j);
    const float *VK_S = COL(s, VK, i, jM1), *VK_N = COL(s, VK, i, jP1);
    const float *UKS_W = COL(s, UKS, iM1, j), *UKS_E = COL(s, UKS, iP1, j), *UKS_S = COL(s, UKS, i, jM1);
    const float *UKS_ES = COL(s, UKS, iP1, jM1), *UKS_WS = COL(s, UKS, iM1, jM1);
    const float *VKS_S = COL(s, VKS, i, jM1), *VKS_W = COL(s, VKS, iM1, j), *WKS_S = COL(s, WKS, i, jM1);
    const float *DPMNEW_S = COL(s, DPMNEW, i, jM1), *UKS_EN = COL(s, UKS, iP1, jP1), *UKS_WN = COL(s, UKS, iM1, jP1);
    const float *WKS_N = COL(s, WK, i, jP1);          /* WK, not WKS (quirk) */
    float VIS = ml_vis(s, i, j, k, kk);
    float mv = vmax(VIS, s->VISHMIN);
    float DE = mv * DYKDZK / DXK, DW = DE;
    float DS = mv * DXKDZK / DYK, DN = DS;
    float DT = mv * AREAxy / DZ;
    float DB = 0;
    if (mask_v_enable) DB = DT * mask;
    else if (k > KKP1) DB = DT;
    if ((ADV(s, i, jM1) < 1) || (j == 2)) DS = 0;
    if ((ADV(s, i, jP1) < 1) || (j == s->NJJ - 1)) DN = 0;
    if ((ADV(s, iM1, j) < 1) || (i == 2)) DW = 0;
    if ((ADV(s, iP1, j) < 1) || (i == s->NII - 1)) DE = 0;
    float FE = 0.25f * (UKS[kk] + UKS_E[kk] + UKS_S[kk] + UKS_ES[kk]) * DYKDZK;
    float FW = 0.25f * (UKS[kk] + UKS_W[kk] + UKS_S[kk] + UKS_WS[kk]) * DYKDZK;
    float FS = VKS_S[kk] * DXKDZK;
    float FN = VKS[kk] * DXKDZK;
    float FT = 0.25f * (WKS_S[kk] + WKS[kk] + WKS_S[kk + 1] + WKS[kk + 1]) * AREAxy;
    float FB = 0;
    if (mask_v_enable) FB = 0.25f * (WKS_S[kk] + WKS[kk] + WKS_S[kk - 1] + WKS[kk - 1]) * AREAxy * mask;
    else if (k > KKP1) FB = 0.25f * (WKS_S[kk] + WKS[kk] + WKS_S[kk - 1] + WKS[kk - 1]) * AREAxy;
    DE = vmax(DE, 0.0001f); DB = vmax(DB, 0.0001f); DW = vmax(DW, 0.0001f);
    DS = vmax(DS, 0.0001f); DN = vmax(DN, 0.0001f); DT = vmax(DT, 0.0001f);
    float PE = fabsf(FE / DE), PB = fabsf(FB / DB), PW = fabsf(FW / DW);
    float PS = fabsf(FS / DS), PN = fabsf(FN / DN), PT = fabsf(FT / DT);
    float BIM = ml_coef(W, DT, PT, vmax(-FT, 0.0f));
    float CIM = ml_coef(W, DB, PB, vmax(FB, 0.0f));
    float AE1 = ml_coef(W, DE, PE, vmax(-FE, 0.0f));
    float AW1 = ml_coef(W, DW, PW, vmax(FW, 0.0f));
    float AS1 = ml_coef(W, DS, PS, vmax(FS, 0.0f));
    float AN1 = ml_coef(W, DN, PN, vmax(-FN, 0.0f));
    float AP0 = s->AP0[kk];
    float AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + AP0;
    ml_pieces(s, i, j, kk, VIS, DE, DW, DS, DN, DT, DB, FE, FW, FS, FN, FT, FB, AE1, AW1, AS1, AN1, BIM, CIM);
    float DDPY = DPMNEW_S[kk] - DPMNEW[kk];
    float VKv = VK[kk];
    piece(s, i, j, ORC_PRESS, kk, DDPY * DXKDZK);
    piece(s, i, j, ORC_COR, kk, s->Cor * (VG - VKv) * AREAxy * DZ);
    float DIM;
    if (s->VEGP && s->VEGP[IX2(s, i, j)]) {
        float in2 = 0.5f * (VKS_W[kk] + VKS[kk]);
        float in = 0.5f * (UKS_W[kk] + UKS[kk]);
        float wh = vmax(sqrtf(in * in + in2 * in2), 0.01f);
        float VEGv = s->VEG[IX3(s, i, j, kk - 1)];
        piece(s, i, j, ORC_VEGT, kk, -VEGv * VKv * wh * AREAxy * DZ);
        /* fused by the 512-bit solver only here (V-prognostic-microscale_1_Vec512.cs:404-411) */
        DIM = ml_nb(W == 16, AW1, VK_W[kk], AS1, VK_S[kk], AE1, VK_E[kk], AN1, VK_N[kk]) +
              AP0 * 0.5f * (VKS_S[kk] + VKS[kk]) + DDPY * DXKDZK +
              (s->Cor * (VG - VKv) - VEGv * VKv * wh) * AREAxy * DZ;
    } else {
        DIM = AW1 * VK_W[kk] + AS1 * VK_S[kk] + AE1 * VK_E[kk] + AN1 * VK_N[kk] +
              AP0 * 0.5f * (VKS_S[kk] + VKS[kk]) + DDPY * DXKDZK + s->Cor * (VG - VKv) * AREAxy * DZ;
    }
    if ((k <= KKP1) && ((k + W) > KKP1)) {
        int ii = KKP1 - k;
        float m2 = (kk == KKP1) ? 1.0f : 0.0f;
        float in = 0.5f * (UKS_W[kk] + UKS[kk]);
        float in2 = 0.5f * (VKS_W[kk] + VKS[kk]);
        float wh = vmax(sqrtf(in * in + in2 * in2), 0.1f);
        float inI = 0.5f * (UKS_W[KKP1] + UKS[KKP1]);
        float in2I = 0.5f * (VKS_W[KKP1] + VKS[KKP1]);
        float whI = vmax(sqrtf(inI * inI + in2I * in2I), 0.1f);
        float Us;
        if (s->CUTK[IX2(s, i, j)] < 1) {
            Us = s->HT[IX2(s, i, j)] * whI;
            if (s->Z0S[IX2(s, i, j)] >= s->DZK[k + ii] * 0.1f) {
                int ind = k + ii + 1;
                Us = s->HT[IX2(s, i, j)] * sqrtf(p2f(0.5f * (UKS_W[ind] + UKS[ind])) + p2f(0.5f * (VKS_W[ind] + VKS[ind])));
            }
        } else {
            Us = s->HO[IX2(s, i, j)] * whI;
        }
        DIM -= VKv / wh * (Us * Us) * AREAxy * m2;
        piece(s, i, j, ORC_WALL, kk, VKv / wh * (Us * Us) * AREAxy * m2);
    }
    if (k > KKP1) {
        float DVDYN = (VK_N[kk] - VKv) * DXKDZK;
        float DVDYS = (VKv - VK_S[kk]) * DXKDZK;
        float DUDYE = (0.5f * (UKS_EN[kk] + UKS_E[kk]) - 0.5f * (UKS_E[kk] + UKS_ES[kk])) * DXKDZK;
        float DUDYW = (0.5f * (UKS_WN[kk] + UKS_W[kk]) - 0.5f * (UKS_W[kk] + UKS_WS[kk])) * DXKDZK;
        float DWDYT = (0.5f * (WKS_N[kk] + WKS[kk]) - 0.5f * (WKS[kk] + WKS_S[kk])) * AREAxy;
        /* quirk: WKS_S at k (not k-1) in the bottom term, V-prognostic-microscale_1.cs:454 */
        float DWDYB = (0.5f * (WKS_N[kk - 1] + WKS[kk - 1]) - 0.5f * (WKS[kk - 1] + WKS_S[kk])) * AREAxy;
        float cross = VIS / DXK * (DUDYE - DUDYW + DVDYN - DVDYS + DWDYT - DWDYB);
        if (mask_v_enable) cross = cross * mask;
        piece(s, i, j, ORC_CROSS, kk, cross);
        DIM += cross;
    }
    *pA = AIM; *pB = BIM; *pC = CIM; *pD = DIM;
}

/* One lane of W_PrognosticMicroscaleV1 (W-prognostic-microscale_1.cs:170-401) */
static void w1_lane(St *s, int i, int j, int k, int kk, float *pA, float *pB, float *pC, float *pD, void *vctx) {
    (void)vctx;
    int W = s->W, KK = KK(s, i, j), KKP1 = KK + 1;
    int iM1 = i - 1, iP1 = i + 1, jM1 = j - 1, jP1 = j + 1;
    float DXK = s->DXK, DYK = s->DYK, AREAxy = s->AREAxy;
    float DZ = s->DZK[kk], DXKDZK = DXK * DZ, DYKDZK = DYK * DZ;
    bool mask_v_enable = (KKP1 >= k && KKP1 < k + W);
    float mask = (kk > KKP1) ? 1.0f : 0.0f;
    const float *WK = COL(s, WK, i, j), *WK_W = COL(s, WK, iM1, j), *WK_E = COL(s, WK, iP1, j);
    const float *WK_S = COL(s, WK, i, jM1), *WK_N = COL(s, WK, i, jP1);
    const float *UKS = COL(s, UKS, i, j), *VKS = COL(s, VKS, i, j), *WKS = COL(s, WKS, i, j);
    const float *DPMNEW = COL(s, DPMNEW, i, j);
    const float *UKS_W = COL(s, UKS, iM1, j), *UKS_E = COL(s, UKS, iP1, j);
    const float *VKS_S = COL(s, VKS, i, jM1), *VKS_N = COL(s, VKS, i, jP1);
    float VIS = ml_vis(s, i, j, k, kk);
    float mv = vmax(VIS, s->VISHMIN);
    float DE = mv * DYKDZK / DXK, DW = DE;
    float DS = mv * DXKDZK / DYK, DN = DS;
    float DT = mv * AREAxy / DZ;
    float DB = 0;
    if (mask_v_enable) DB = DT * mask;
    else if (k > KKP1) DB = DT;
    if ((ADV(s, i, jM1) < 1) || (j == 2)) DS = 0;
    if ((ADV(s, i, jP1) < 1) || (j == s->NJJ - 1)) DN = 0;
    if ((ADV(s, iM1, j) < 1) || (i == 2)) DW = 0;
    if ((ADV(s, iP1, j) < 1) || (i == s->NII - 1)) DE = 0;
    const int kw = (ORC_MUTANT == 5) ? kk + 1 : kk - 1;
    float FE = 0.25f * (UKS[kk] + UKS_E[kk] + UKS[kw] + UKS_E[kw]) * DYKDZK;
    float FW = 0.25f * (UKS[kk] + UKS_W[kk] + UKS[kk - 1] + UKS_W[kk - 1]) * DYKDZK;
    float FS = 0.25f * (VKS[kk] + VKS_S[kk] + VKS[kk - 1] + VKS_S[kk - 1]) * DXKDZK;
    float FN = 0.25f * (VKS[kk] + VKS_N[kk] + VKS[kk - 1] + VKS_N[kk - 1]) * DXKDZK;
    float FT = WKS[kk] * AREAxy;
    float FB = 0;
    if (mask_v_enable) FB = WKS[kk - 1] * AREAxy * mask;
    else if (k > KKP1) FB = WKS[kk - 1] * AREAxy;
    DE = vmax(DE, 0.0001f); DB = vmax(DB, 0.0001f); DW = vmax(DW, 0.0001f);
    DS = vmax(DS, 0.0001f); DN = vmax(DN, 0.0001f); DT = vmax(DT, 0.0001f);
    float PE = fabsf(FE / DE), PB = fabsf(FB / DB), PW = fabsf(FW / DW);
    float PS = fabsf(FS / DS), PN = fabsf(FN / DN), PT = fabsf(FT / DT);
    float BIM = ml_coef(W, DT, PT, vmax(-FT, 0.0f));
    float CIM = ml_coef(W, DB, PB, vmax(FB, 0.0f));
    float AE1 = ml_coef(W, DE, PE, vmax(-FE, 0.0f));
    float AW1 = ml_coef(W, DW, PW, vmax(FW, 0.0f));
    float AS1 = ml_coef(W, DS, PS, vmax(FS, 0.0f));
    float AN1 = ml_coef(W, DN, PN, vmax(-FN, 0.0f));
    float AP0 = s->AP0[kk];
    float AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + AP0;
    ml_pieces(s, i, j, kk, VIS, DE, DW, DS, DN, DT, DB, FE, FW, FS, FN, FT, FB, AE1, AW1, AS1, AN1, BIM, CIM);
    float DDPZ = DPMNEW[kk - 1] - DPMNEW[kk];
    float WKv = WK[kk];
    float DIM;
    piece(s, i, j, ORC_PRESS, kk, DDPZ * AREAxy);
    if (s->VEGP && s->VEGP[IX2(s, i, j)]) {
        float in2 = 0.5f * (VKS_S[kk] + VKS[kk]);
        float in = 0.5f * (UKS_W[kk] + UKS[kk]);
        float wh = vmax(sqrtf(in * in + in2 * in2), 0.01f);
        float VEGv = s->VEG[IX3(s, i, j, kk - 1)];
        piece(s, i, j, ORC_VEGT, kk, -VEGv * WKv * wh * AREAxy * DZ);
        /* W-prognostic-microscale_1_Vec512.cs:367-383 fuses both forms */
        DIM = ml_nb(W == 16, AW1, WK_W[kk], AS1, WK_S[kk], AE1, WK_E[kk], AN1, WK_N[kk]) +
              AP0 * 0.5f * (WKS[kk - 1] + WKS[kk]) + DDPZ * AREAxy - VEGv * WKv * wh * AREAxy * DZ;
    } else {
        DIM = ml_nb(W == 16, AW1, WK_W[kk], AS1, WK_S[kk], AE1, WK_E[kk], AN1, WK_N[kk]) +
              AP0 * 0.5f * (WKS[kk - 1] + WKS[kk]) + DDPZ * AREAxy;
    }
    if (k > KKP1) {
        float DUDZE = (0.5f * (UKS[kk + 1] + UKS[kk]) - 0.5f * (UKS[kk] + UKS[kk - 1])) * DYKDZK;
        float DUDZW = (0.5f * (UKS_W[kk + 1] + UKS_W[kk]) - 0.5f * (UKS_W[kk] + UKS_W[kk - 1])) * DYKDZK;
        float DVDZN = (0.5f * (VKS[kk + 1] + VKS[kk]) - 0.5f * (VKS[kk] + VKS[kk - 1])) * DXKDZK;
        float DVDZS = (0.5f * (VKS_S[kk + 1] + VKS_S[kk]) - 0.5f * (VKS_S[kk] + VKS_S[kk - 1])) * DXKDZK;
        float DWDZT = (WK[kk + 1] - WKv) * AREAxy;
        float DWDZB = (WKv - WK[kk - 1]) * AREAxy;
        float cross = VIS / DZ * (DUDZE - DUDZW + DVDZN - DVDZS + DWDZT - DWDZB);
        if (mask_v_enable) cross = cross * mask;
        piece(s, i, j, ORC_CROSS, kk, cross);
        DIM += cross;
    }
    *pA = AIM; *pB = BIM; *pC = CIM; *pD = DIM;
}

#if defined(ORC_SIMD) && ORC_MUTANT == 0
/* ---- 8-wide blocks (the Vector<float> form, for the CPU timing build) -----
 * The same lane arithmetic as u1_lane / v1_lane / w1_lane for all 8 lanes of
 * a block start k at once, in AVX2 vectors: every vector operation is the
 * lane-wise IEEE operation of the scalar form, in the same order (no FMA:
 * -ffp-contract=off), maxps/minps are Vector.Max/Min (a > b ? a : b), and the
 * division and square root are correctly rounded, so each lane is bit for bit
 * the scalar lane (tests/test_oracle_simd.py).  Width 8 only, no checking
 * hook: ml_column falls back to the lanes otherwise. */
#include <immintrin.h>
typedef float v8f __attribute__((vector_size(32)));
typedef int v8i __attribute__((vector_size(32)));
static inline v8f ld8(const float *p) { v8f v; memcpy(&v, p, sizeof v); return v; }
static inline v8f bc8(float x) { v8f v = {x, x, x, x, x, x, x, x}; return v; }
static inline v8f max8(v8f a, v8f b) { return (v8f)_mm256_max_ps((__m256)a, (__m256)b); }
static inline v8f min8(v8f a, v8f b) { return (v8f)_mm256_min_ps((__m256)a, (__m256)b); }
static inline v8f sqrt8(v8f a) { return (v8f)_mm256_sqrt_ps((__m256)a); }
static inline v8f abs8(v8f a) { return (v8f)((v8i)a & 0x7fffffff); }
static inline v8f sel8(v8i m, v8f a, v8f b) { return (v8f)((m & (v8i)a) | (~m & (v8i)b)); }
static inline v8i lanes8(int k) { v8i v = {k, k + 1, k + 2, k + 3, k + 4, k + 5, k + 6, k + 7}; return v; }
static inline v8f pow8(v8f Dv, v8f Pv) {            /* ml_pow: D*max(0,(1-0.1P)^5) */
    v8f in = 1.0f - 0.1f * Pv;
    return Dv * max8(bc8(0.0f), in * in * in * in * in);
}
static inline v8f coef8(v8f Dv, v8f Pv, v8f Fpos) { return pow8(Dv, Pv) + Fpos; }

typedef struct { v8f A, B, C, D; } blk8;

/* eddy viscosity of the block (ml_vis) */
static v8f vis8(St *s, int i, int j, int k, v8f maskf, bool mask_v_enable) {
    const int KK = KK(s, i, j), KKP1 = KK + 1;
    if (!(k + 8 > KKP1)) return bc8(0.0f);
    v8f DZ = ld8(s->DZK + k);
    v8f zs = ld8(s->HOKART + k) - s->HOKART[KK] - DZ * 0.5f;
    v8f mixL = min8(0.071f * zs, bc8(100.0f));
    float COV = s->COV[IX2(s, i, j)];
    if (COV > 0) mixL *= (1.0f - 0.99f * COV);
    const float *UK = COL(s, UK, i, j), *VK = COL(s, VK, i, j);
    v8f in = (ld8(UK + k + 1) - ld8(UK + k - 1)) / (2.0f * DZ);
    v8f DUDZ = in * in;
    in = (ld8(VK + k + 1) - ld8(VK + k - 1)) / (2.0f * DZ);
    v8f DVDZ = in * in;
    v8f v = min8(mixL * mixL * sqrt8(0.5f * (DUDZ + DVDZ)), bc8(15.0f));
    return mask_v_enable ? v * maskf : v;
}

/* diffusion conductances, fluxes and power-law coefficients shared by the
 * three components once FE..FB are known */
static void coeffs8(St *s, int i, int j, int k, v8f VIS, v8f DZ, v8f FE, v8f FW, v8f FS, v8f FN, v8f FT, v8f FB,
                    v8f maskf, bool mask_v_enable, v8f *AE1, v8f *AW1, v8f *AS1, v8f *AN1, v8f *BIM, v8f *CIM) {
    const int KKP1 = KK(s, i, j) + 1;
    const float DXK = s->DXK, DYK = s->DYK, AREAxy = s->AREAxy;
    v8f DXKDZK = DXK * DZ, DYKDZK = DYK * DZ;
    v8f mv = max8(VIS, bc8(s->VISHMIN));
    v8f DE = mv * DYKDZK / DXK, DW = DE;
    v8f DS = mv * DXKDZK / DYK, DN = DS;
    v8f DT = mv * AREAxy / DZ;
    v8f DB = bc8(0.0f);
    if (mask_v_enable) DB = DT * maskf;
    else if (k > KKP1) DB = DT;
    if ((ADV(s, i, j - 1) < 1) || (j == 2)) DS = bc8(0.0f);
    if ((ADV(s, i, j + 1) < 1) || (j == s->NJJ - 1)) DN = bc8(0.0f);
    if ((ADV(s, i - 1, j) < 1) || (i == 2)) DW = bc8(0.0f);
    if ((ADV(s, i + 1, j) < 1) || (i == s->NII - 1)) DE = bc8(0.0f);
    const v8f eps = bc8(0.0001f), zero = bc8(0.0f);
    DE = max8(DE, eps); DB = max8(DB, eps); DW = max8(DW, eps);
    DS = max8(DS, eps); DN = max8(DN, eps); DT = max8(DT, eps);
    v8f PE = abs8(FE / DE), PB = abs8(FB / DB), PW = abs8(FW / DW);
    v8f PS = abs8(FS / DS), PN = abs8(FN / DN), PT = abs8(FT / DT);
    *BIM = coef8(DT, PT, max8(-FT, zero));
    *CIM = coef8(DB, PB, max8(FB, zero));
    *AE1 = coef8(DE, PE, max8(-FE, zero));
    *AW1 = coef8(DW, PW, max8(FW, zero));
    *AS1 = coef8(DS, PS, max8(FS, zero));
    *AN1 = coef8(DN, PN, max8(-FN, zero));
}

/* the U/V wall term of the block holding KKART+1 (lane KKART+1 only) */
static v8f wall8(St *s, int i, int j, int k, v8f X, v8f AREAv, const float *UKS_W, const float *UKS, const float *VKS_W,
                 const float *VKS) {
    const int KKP1 = KK(s, i, j) + 1;
    v8f m2 = sel8(lanes8(k) == KKP1, bc8(1.0f), bc8(0.0f));
    v8f in = 0.5f * (ld8(UKS_W + k) + ld8(UKS + k));
    v8f in2 = 0.5f * (ld8(VKS_W + k) + ld8(VKS + k));
    v8f wh = max8(sqrt8(in * in + in2 * in2), bc8(0.1f));
    int ii = KKP1 - k;
    float inI = 0.5f * (UKS_W[KKP1] + UKS[KKP1]);
    float in2I = 0.5f * (VKS_W[KKP1] + VKS[KKP1]);
    float whI = vmax(sqrtf(inI * inI + in2I * in2I), 0.1f);
    float Us;
    if (s->CUTK[IX2(s, i, j)] < 1) {
        Us = s->HT[IX2(s, i, j)] * whI;
        if (s->Z0S[IX2(s, i, j)] >= s->DZK[k + ii] * 0.1f) {
            int ind = k + ii + 1;
            Us = s->HT[IX2(s, i, j)] * sqrtf(p2f(0.5f * (UKS_W[ind] + UKS[ind])) + p2f(0.5f * (VKS_W[ind] + VKS[ind])));
        }
    } else {
        Us = s->HO[IX2(s, i, j)] * whI;
    }
    return X / wh * (Us * Us) * AREAv * m2;
}

static blk8 u1_block(St *s, int i, int j, int k, float UG) {
    const int KK = KK(s, i, j), KKP1 = KK + 1;
    const int iM1 = i - 1, iP1 = i + 1, jM1 = j - 1, jP1 = j + 1;
    const float DXK = s->DXK, DYK = s->DYK, AREAxy = s->AREAxy;
    v8f DZ = ld8(s->DZK + k), DXKDZK = DXK * DZ, DYKDZK = DYK * DZ;
    const bool mask_v_enable = (KKP1 >= k && KKP1 < k + 8);
    v8f maskf = sel8(lanes8(k) > KKP1, bc8(1.0f), bc8(0.0f));
    const float *UK = COL(s, UK, i, j) + k, *UKS = COL(s, UKS, i, j) + k, *VKS = COL(s, VKS, i, j) + k;
    const float *WKS = COL(s, WKS, i, j) + k, *DPMNEW = COL(s, DPMNEW, i, j) + k;
    const float *UK_W = COL(s, UK, iM1, j) + k, *UK_E = COL(s, UK, iP1, j) + k;
    const float *UK_S = COL(s, UK, i, jM1) + k, *UK_N = COL(s, UK, i, jP1) + k, *UKS_W = COL(s, UKS, iM1, j) + k;
    const float *VKS_W = COL(s, VKS, iM1, j) + k, *VKS_WS = COL(s, VKS, iM1, jM1) + k, *VKS_S = COL(s, VKS, i, jM1) + k;
    const float *VKS_WN = COL(s, VKS, iM1, jP1) + k, *VKS_N = COL(s, VKS, i, jP1) + k, *WKS_W = COL(s, WKS, iM1, j) + k;
    const float *DPMNEW_W = COL(s, DPMNEW, iM1, j) + k, *VKS_EN = COL(s, VKS, iP1, jP1) + k;
    const float *VKS_ES = COL(s, VKS, iP1, jM1) + k;
    const float *WKS_E = COL(s, WK, iP1, j) + k;      /* WK, not WKS (quirk) */
    v8f VIS = vis8(s, i, j, k, maskf, mask_v_enable);
    v8f FE = ld8(UKS) * DYKDZK;
    v8f FW = ld8(UKS_W) * DYKDZK;
    v8f FS = 0.25f * (ld8(VKS_W) + ld8(VKS) + ld8(VKS_WS) + ld8(VKS_S)) * DXKDZK;
    v8f FN = 0.25f * (ld8(VKS_W) + ld8(VKS) + ld8(VKS_WN) + ld8(VKS_N)) * DXKDZK;
    v8f FT = 0.25f * (ld8(WKS_W) + ld8(WKS) + ld8(WKS_W + 1) + ld8(WKS + 1)) * AREAxy;
    v8f FB = bc8(0.0f);
    if (mask_v_enable) FB = 0.25f * (ld8(WKS_W) + ld8(WKS) + ld8(WKS_W - 1) + ld8(WKS - 1)) * AREAxy * maskf;
    else if (k > KKP1) FB = 0.25f * (ld8(WKS_W) + ld8(WKS) + ld8(WKS_W - 1) + ld8(WKS - 1)) * AREAxy;
    v8f AE1, AW1, AS1, AN1, BIM, CIM;
    coeffs8(s, i, j, k, VIS, DZ, FE, FW, FS, FN, FT, FB, maskf, mask_v_enable, &AE1, &AW1, &AS1, &AN1, &BIM, &CIM);
    v8f AP0 = ld8(s->AP0 + k);
    v8f AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + AP0;
    v8f DDPX = ld8(DPMNEW_W) - ld8(DPMNEW);
    v8f UKv = ld8(UK);
    v8f DIM;
    if (s->VEGP && s->VEGP[IX2(s, i, j)]) {
        v8f in2 = 0.5f * (ld8(VKS_W) + ld8(VKS));
        v8f in = 0.5f * (ld8(UKS_W) + ld8(UKS));
        v8f wh = max8(sqrt8(in * in + in2 * in2), bc8(0.01f));
        v8f VEGv = ld8(s->VEG + IX3(s, i, j, k - 1));
        DIM = AW1 * ld8(UK_W) + AS1 * ld8(UK_S) + AE1 * ld8(UK_E) + AN1 * ld8(UK_N) +
              AP0 * 0.5f * (ld8(UKS_W) + ld8(UKS)) + DDPX * DYKDZK + (s->Cor * (UG - UKv) - VEGv * UKv * wh) * AREAxy * DZ;
    } else {
        DIM = AW1 * ld8(UK_W) + AS1 * ld8(UK_S) + AE1 * ld8(UK_E) + AN1 * ld8(UK_N) +
              AP0 * 0.5f * (ld8(UKS_W) + ld8(UKS)) + DDPX * DYKDZK + s->Cor * (UG - UKv) * AREAxy * DZ;
    }
    if ((k <= KKP1) && ((k + 8) > KKP1))
        DIM -= wall8(s, i, j, k, UKv, bc8(AREAxy), UKS_W - k, UKS - k, VKS_W - k, VKS - k);
    if (k > KKP1) {
        v8f DUDXE = (ld8(UK_E) - UKv) * DYKDZK;
        v8f DUDXW = (UKv - ld8(UK_W)) * DYKDZK;
        v8f DVDXN = (0.5f * (ld8(VKS_EN) + ld8(VKS_N)) - 0.5f * (ld8(VKS_N) + ld8(VKS_WN))) * DXKDZK;
        v8f DVDXS = (0.5f * (ld8(VKS_ES) + ld8(VKS_S)) - 0.5f * (ld8(VKS_S) + ld8(VKS_WS))) * DXKDZK;
        v8f DWDXT = (0.5f * (ld8(WKS_E) + ld8(WKS)) - 0.5f * (ld8(WKS) + ld8(WKS_W))) * AREAxy;
        v8f DWDXB = (0.5f * (ld8(WKS_E - 1) + ld8(WKS - 1)) - 0.5f * (ld8(WKS - 1) + ld8(WKS_W - 1))) * AREAxy;
        v8f cross = VIS / DXK * (DUDXE - DUDXW + DVDXN - DVDXS + DWDXT - DWDXB);
        if (mask_v_enable) cross = cross * maskf;
        DIM += cross;
    }
    blk8 r = {AIM, BIM, CIM, DIM};
    return r;
}

static blk8 v1_block(St *s, int i, int j, int k, float VG) {
    const int KK = KK(s, i, j), KKP1 = KK + 1;
    const int iM1 = i - 1, iP1 = i + 1, jM1 = j - 1, jP1 = j + 1;
    const float DXK = s->DXK, DYK = s->DYK, AREAxy = s->AREAxy;
    v8f DZ = ld8(s->DZK + k), DXKDZK = DXK * DZ, DYKDZK = DYK * DZ;
    const bool mask_v_enable = (KKP1 >= k && KKP1 < k + 8);
    v8f maskf = sel8(lanes8(k) > KKP1, bc8(1.0f), bc8(0.0f));
    const float *VK = COL(s, VK, i, j) + k, *UKS = COL(s, UKS, i, j) + k, *VKS = COL(s, VKS, i, j) + k;
    const float *WKS = COL(s, WKS, i, j) + k, *DPMNEW = COL(s, DPMNEW, i, j) + k;
    const float *VK_W = COL(s, VK, iM1, j) + k, *VK_E = COL(s, VK, iP1, j) + k;
    const float *VK_S = COL(s, VK, i, jM1) + k, *VK_N = COL(s, VK, i, jP1) + k;
    const float *UKS_W = COL(s, UKS, iM1, j) + k, *UKS_E = COL(s, UKS, iP1, j) + k, *UKS_S = COL(s, UKS, i, jM1) + k;
    const float *UKS_ES = COL(s, UKS, iP1, jM1) + k, *UKS_WS = COL(s, UKS, iM1, jM1) + k;
    const float *VKS_S = COL(s, VKS, i, jM1) + k, *VKS_W = COL(s, VKS, iM1, j) + k, *WKS_S = COL(s, WKS, i, jM1) + k;
    const float *DPMNEW_S = COL(s, DPMNEW, i, jM1) + k, *UKS_EN = COL(s, UKS, iP1, jP1) + k;
    const float *UKS_WN = COL(s, UKS, iM1, jP1) + k;
    const float *WKS_N = COL(s, WK, i, jP1) + k;      /* WK, not WKS (quirk) */
    v8f VIS = vis8(s, i, j, k, maskf, mask_v_enable);
    v8f FE = 0.25f * (ld8(UKS) + ld8(UKS_E) + ld8(UKS_S) + ld8(UKS_ES)) * DYKDZK;
    v8f FW = 0.25f * (ld8(UKS) + ld8(UKS_W) + ld8(UKS_S) + ld8(UKS_WS)) * DYKDZK;
    v8f FS = ld8(VKS_S) * DXKDZK;
    v8f FN = ld8(VKS) * DXKDZK;
    v8f FT = 0.25f * (ld8(WKS_S) + ld8(WKS) + ld8(WKS_S + 1) + ld8(WKS + 1)) * AREAxy;
    v8f FB = bc8(0.0f);
    if (mask_v_enable) FB = 0.25f * (ld8(WKS_S) + ld8(WKS) + ld8(WKS_S - 1) + ld8(WKS - 1)) * AREAxy * maskf;
    else if (k > KKP1) FB = 0.25f * (ld8(WKS_S) + ld8(WKS) + ld8(WKS_S - 1) + ld8(WKS - 1)) * AREAxy;
    v8f AE1, AW1, AS1, AN1, BIM, CIM;
    coeffs8(s, i, j, k, VIS, DZ, FE, FW, FS, FN, FT, FB, maskf, mask_v_enable, &AE1, &AW1, &AS1, &AN1, &BIM, &CIM);
    v8f AP0 = ld8(s->AP0 + k);
    v8f AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + AP0;
    v8f DDPY = ld8(DPMNEW_S) - ld8(DPMNEW);
    v8f VKv = ld8(VK);
    v8f DIM;
    if (s->VEGP && s->VEGP[IX2(s, i, j)]) {
        v8f in2 = 0.5f * (ld8(VKS_W) + ld8(VKS));
        v8f in = 0.5f * (ld8(UKS_W) + ld8(UKS));
        v8f wh = max8(sqrt8(in * in + in2 * in2), bc8(0.01f));
        v8f VEGv = ld8(s->VEG + IX3(s, i, j, k - 1));
        DIM = AW1 * ld8(VK_W) + AS1 * ld8(VK_S) + AE1 * ld8(VK_E) + AN1 * ld8(VK_N) +
              AP0 * 0.5f * (ld8(VKS_S) + ld8(VKS)) + DDPY * DXKDZK + (s->Cor * (VG - VKv) - VEGv * VKv * wh) * AREAxy * DZ;
    } else {
        DIM = AW1 * ld8(VK_W) + AS1 * ld8(VK_S) + AE1 * ld8(VK_E) + AN1 * ld8(VK_N) +
              AP0 * 0.5f * (ld8(VKS_S) + ld8(VKS)) + DDPY * DXKDZK + s->Cor * (VG - VKv) * AREAxy * DZ;
    }
    if ((k <= KKP1) && ((k + 8) > KKP1))
        DIM -= wall8(s, i, j, k, VKv, bc8(AREAxy), UKS_W - k, UKS - k, VKS_W - k, VKS - k);
    if (k > KKP1) {
        v8f DVDYN = (ld8(VK_N) - VKv) * DXKDZK;
        v8f DVDYS = (VKv - ld8(VK_S)) * DXKDZK;
        v8f DUDYE = (0.5f * (ld8(UKS_EN) + ld8(UKS_E)) - 0.5f * (ld8(UKS_E) + ld8(UKS_ES))) * DXKDZK;
        v8f DUDYW = (0.5f * (ld8(UKS_WN) + ld8(UKS_W)) - 0.5f * (ld8(UKS_W) + ld8(UKS_WS))) * DXKDZK;
        v8f DWDYT = (0.5f * (ld8(WKS_N) + ld8(WKS)) - 0.5f * (ld8(WKS) + ld8(WKS_S))) * AREAxy;
        v8f DWDYB = (0.5f * (ld8(WKS_N - 1) + ld8(WKS - 1)) - 0.5f * (ld8(WKS - 1) + ld8(WKS_S))) * AREAxy;
        v8f cross = VIS / DXK * (DUDYE - DUDYW + DVDYN - DVDYS + DWDYT - DWDYB);
        if (mask_v_enable) cross = cross * maskf;
        DIM += cross;
    }
    blk8 r = {AIM, BIM, CIM, DIM};
    return r;
}

static blk8 w1_block(St *s, int i, int j, int k) {
    const int KK = KK(s, i, j), KKP1 = KK + 1;
    const int iM1 = i - 1, iP1 = i + 1, jM1 = j - 1, jP1 = j + 1;
    const float DXK = s->DXK, DYK = s->DYK, AREAxy = s->AREAxy;
    v8f DZ = ld8(s->DZK + k), DXKDZK = DXK * DZ, DYKDZK = DYK * DZ;
    const bool mask_v_enable = (KKP1 >= k && KKP1 < k + 8);
    v8f maskf = sel8(lanes8(k) > KKP1, bc8(1.0f), bc8(0.0f));
    const float *WK = COL(s, WK, i, j) + k, *WK_W = COL(s, WK, iM1, j) + k, *WK_E = COL(s, WK, iP1, j) + k;
    const float *WK_S = COL(s, WK, i, jM1) + k, *WK_N = COL(s, WK, i, jP1) + k;
    const float *UKS = COL(s, UKS, i, j) + k, *VKS = COL(s, VKS, i, j) + k, *WKS = COL(s, WKS, i, j) + k;
    const float *DPMNEW = COL(s, DPMNEW, i, j) + k;
    const float *UKS_W = COL(s, UKS, iM1, j) + k, *UKS_E = COL(s, UKS, iP1, j) + k;
    const float *VKS_S = COL(s, VKS, i, jM1) + k, *VKS_N = COL(s, VKS, i, jP1) + k;
    v8f VIS = vis8(s, i, j, k, maskf, mask_v_enable);
    v8f FE = 0.25f * (ld8(UKS) + ld8(UKS_E) + ld8(UKS - 1) + ld8(UKS_E - 1)) * DYKDZK;
    v8f FW = 0.25f * (ld8(UKS) + ld8(UKS_W) + ld8(UKS - 1) + ld8(UKS_W - 1)) * DYKDZK;
    v8f FS = 0.25f * (ld8(VKS) + ld8(VKS_S) + ld8(VKS - 1) + ld8(VKS_S - 1)) * DXKDZK;
    v8f FN = 0.25f * (ld8(VKS) + ld8(VKS_N) + ld8(VKS - 1) + ld8(VKS_N - 1)) * DXKDZK;
    v8f FT = ld8(WKS) * AREAxy;
    v8f FB = bc8(0.0f);
    if (mask_v_enable) FB = ld8(WKS - 1) * AREAxy * maskf;
    else if (k > KKP1) FB = ld8(WKS - 1) * AREAxy;
    v8f AE1, AW1, AS1, AN1, BIM, CIM;
    coeffs8(s, i, j, k, VIS, DZ, FE, FW, FS, FN, FT, FB, maskf, mask_v_enable, &AE1, &AW1, &AS1, &AN1, &BIM, &CIM);
    v8f AP0 = ld8(s->AP0 + k);
    v8f AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + AP0;
    v8f DDPZ = ld8(DPMNEW - 1) - ld8(DPMNEW);
    v8f WKv = ld8(WK);
    v8f DIM;
    if (s->VEGP && s->VEGP[IX2(s, i, j)]) {
        v8f in2 = 0.5f * (ld8(VKS_S) + ld8(VKS));
        v8f in = 0.5f * (ld8(UKS_W) + ld8(UKS));
        v8f wh = max8(sqrt8(in * in + in2 * in2), bc8(0.01f));
        v8f VEGv = ld8(s->VEG + IX3(s, i, j, k - 1));
        DIM = AW1 * ld8(WK_W) + AS1 * ld8(WK_S) + AE1 * ld8(WK_E) + AN1 * ld8(WK_N) +
              AP0 * 0.5f * (ld8(WKS - 1) + ld8(WKS)) + DDPZ * AREAxy - VEGv * WKv * wh * AREAxy * DZ;
    } else {
        DIM = AW1 * ld8(WK_W) + AS1 * ld8(WK_S) + AE1 * ld8(WK_E) + AN1 * ld8(WK_N) +
              AP0 * 0.5f * (ld8(WKS - 1) + ld8(WKS)) + DDPZ * AREAxy;
    }
    if (k > KKP1) {
        v8f DUDZE = (0.5f * (ld8(UKS + 1) + ld8(UKS)) - 0.5f * (ld8(UKS) + ld8(UKS - 1))) * DYKDZK;
        v8f DUDZW = (0.5f * (ld8(UKS_W + 1) + ld8(UKS_W)) - 0.5f * (ld8(UKS_W) + ld8(UKS_W - 1))) * DYKDZK;
        v8f DVDZN = (0.5f * (ld8(VKS + 1) + ld8(VKS)) - 0.5f * (ld8(VKS) + ld8(VKS - 1))) * DXKDZK;
        v8f DVDZS = (0.5f * (ld8(VKS_S + 1) + ld8(VKS_S)) - 0.5f * (ld8(VKS_S) + ld8(VKS_S - 1))) * DXKDZK;
        v8f DWDZT = (ld8(WK + 1) - WKv) * AREAxy;
        v8f DWDZB = (WKv - ld8(WK - 1)) * AREAxy;
        v8f cross = VIS / DZ * (DUDZE - DUDZW + DVDZN - DVDZS + DWDZT - DWDZB);
        if (mask_v_enable) cross = cross * maskf;
        DIM += cross;
    }
    blk8 r = {AIM, BIM, CIM, DIM};
    return r;
}

/* the 8 lanes of block k of the sweep whose lane function is fn */
static void ml_block(St *s, int i, int j, int k, lane_fn fn, void *ctx, float *A, float *B, float *C, float *D) {
    blk8 r;
    if (fn == u1_lane) r = u1_block(s, i, j, k, ((uctx *)ctx)->UG);
    else if (fn == v1_lane) r = v1_block(s, i, j, k, ((uctx *)ctx)->UG);
    else r = w1_block(s, i, j, k);
    memcpy(A, &r.A, sizeof r.A); memcpy(B, &r.B, sizeof r.B); memcpy(C, &r.C, sizeof r.C); memcpy(D, &r.D, sizeof r.D);
}

static int use_blocks = 1;
/* 0: the timing build evaluates lanes one by one like the default build */
void orc_simd_blocks(int on) { use_blocks = on != 0; }
#endif

/* U_PrognosticMicroscaleV1.Calculate(IS, JS): i1 in [3,NII-1], j1 in [2,NJJ-1] */
static void u1_row(St *s, int IS, int JS, int i1);
static void u1_sweep(St *s, int IS, int JS) {
    s->comp = 0;
    for (int i1 = 3; i1 < s->NII; ++i1) u1_row(s, IS, JS, i1);
}
static void u1_row(St *s, int IS, int JS, int i1) {
    uctx c = { s->UG };
    {
        int i = (IS == -1) ? s->NII - i1 + 1 : i1;
        for (int j1 = 2; j1 <= s->NJJ - 1; ++j1) {
            int j = (JS == -1) ? s->NJJ - j1 + 1 : j1;
            if (ADV(s, i, j) != 1) continue;
            ml_column(s, i, j, u1_lane, &c);
            float *UK = COL(s, UK, i, j);
            int KK = KK(s, i, j), VIc = s->VI[IX2(s, i, j)];
            int KSTART = (s->CUTK[IX2(s, i, j)] == 0) ? KK + 1 : 1;
            int KKm = KK > KK(s, i - 1, j) ? KK : KK(s, i - 1, j);
            cap_column(s, 0, i, j, UK, false);
            for (int k = VIc; k >= KSTART; --k)
                if (KKm < k) UK[k] += s->relax * (s->P1[k] * UK[k + 1] + s->Q1[k] - UK[k]);
            cap_column(s, 0, i, j, UK, true);
        }
    }
}

/* V_PrognosticMicroscaleV1.Calculate(IS, JS): i1 in [2,NII-1], j1 in [3,NJJ-1] */
static void v1_row(St *s, int IS, int JS, int i1);
static void v1_sweep(St *s, int IS, int JS) {
    s->comp = 1;
    for (int i1 = 2; i1 < s->NII; ++i1) v1_row(s, IS, JS, i1);
}
static void v1_row(St *s, int IS, int JS, int i1) {
    uctx c = { s->VG };
    {
        int i = (IS == -1) ? s->NII - i1 + 1 : i1;
        for (int j1 = 3; j1 <= s->NJJ - 1; ++j1) {
            int j = (JS == -1) ? s->NJJ - j1 + 1 : j1;
            if (ADV(s, i, j) != 1) continue;
            ml_column(s, i, j, v1_lane, &c);
            float *VK = COL(s, VK, i, j);
            int KK = KK(s, i, j), VIc = s->VI[IX2(s, i, j)];
            int KSTART = (s->CUTK[IX2(s, i, j)] == 0) ? KK + 1 : 1;
            int KKm = KK > KK(s, i, j - 1) ? KK : KK(s, i, j - 1);
            cap_column(s, 1, i, j, VK, false);
            for (int k = VIc; k >= KSTART; --k)
                if (KKm < k) VK[k] += s->relax * (s->P1[k] * VK[k + 1] + s->Q1[k] - VK[k]);
            cap_column(s, 1, i, j, VK, true);
        }
    }
}

/* W_PrognosticMicroscaleV1.Calculate(IS, JS): i1 in [2,NII-1], j1 in [2,NJJ-1] */
static void w1_row(St *s, int IS, int JS, int i1);
static void w1_sweep(St *s, int IS, int JS) {
    s->comp = 2;
    for (int i1 = 2; i1 < s->NII; ++i1) w1_row(s, IS, JS, i1);
}
static void w1_row(St *s, int IS, int JS, int i1) {
    {
        int i = (IS == -1) ? s->NII - i1 + 1 : i1;
        for (int j1 = 2; j1 < s->NJJ; ++j1) {
            int j = (JS == -1) ? s->NJJ - j1 + 1 : j1;
            if (ADV(s, i, j) != 1) continue;
            ml_column(s, i, j, w1_lane, NULL);
            float *WK = COL(s, WK, i, j);
            int KK = KK(s, i, j), VIc = s->VI[IX2(s, i, j)];
            int KSTART = (s->CUTK[IX2(s, i, j)] == 0) ? KK + 1 : 1;
            cap_column(s, 2, i, j, WK, false);
            for (int k = VIc; k >= KSTART; --k)
                WK[k] += s->relax * (s->P1[k] * WK[k + 1] + s->Q1[k] - WK[k]);
            cap_column(s, 2, i, j, WK, true);
        }
    }
}

/* ---- timing mode: the reference's default multi-threaded schedule ------
 * Not deterministic and never compared: the rows of a sweep are dealt to
 * threads that update the fields in place while their neighbours do
 * (U-prognostic-microscale_1.cs:59-61 partitions i; the pressure sweep runs
 * its j rows in parallel, PrognosticFlowfield.cs:806-811), and U and V run
 * at the same time (PrognosticFlowfield.cs:1097-1101).  Each thread has its
 * own column scratch. */
static void par_scratch(St *t, const St *s) {
    *t = *s;
    t->cap = NULL;
    t->P1 = (float *)calloc((size_t)s->NKP, sizeof(float)); t->Q1 = (float *)calloc((size_t)s->NKP, sizeof(float));
    t->P2 = (float *)calloc((size_t)s->NKP, sizeof(float)); t->Q2 = (float *)calloc((size_t)s->NKP, sizeof(float));
}
static void par_free(St *t) { free(t->P1); free(t->Q1); free(t->P2); free(t->Q2); }

static float pressure_par(St *s, int IS, int JS) {
    float DeltaUMax = 0;
#pragma omp parallel num_threads(s->nthreads)
    {
        float *PIMP = (float *)calloc((size_t)s->NKK + 2, sizeof(float));
        float *QIMP = (float *)calloc((size_t)s->NKK + 2, sizeof(float));
        float mine = 0;
#pragma omp for schedule(static)
        for (int j1 = 2; j1 < s->NJJ; ++j1) {
            float d = pressure_row(s, IS, JS, j1, PIMP, QIMP);
            if (d > mine) mine = d;
        }
#pragma omp critical
        if (mine > DeltaUMax) DeltaUMax = mine;
        free(PIMP); free(QIMP);
    }
    return DeltaUMax;
}

static void ml_sweeps_par(St *s, int IS, int JS) {
    const int nu = s->NII - 3, nv = s->NII - 2;          /* U rows i1 = 3..NII-1, V rows i1 = 2..NII-1 */
#pragma omp parallel num_threads(s->nthreads)
    {
        St t;
        par_scratch(&t, s);
        const int nmax = nu > nv ? nu : nv;
#pragma omp for schedule(dynamic, 4)
        for (int q = 0; q < 2 * nmax; ++q) {              /* U and V rows interleaved */
            const int r = q / 2;
            if (q % 2 == 0) { if (r < nu) u1_row(&t, IS, JS, 3 + r); }
            else if (r < nv) v1_row(&t, -IS, -JS, 2 + r);
        }
#pragma omp barrier
#pragma omp for schedule(static)
        for (int i1 = 2; i1 < s->NII; ++i1) w1_row(&t, IS, JS, i1);
        par_free(&t);
    }
}

/* ===================================================================== */
/*  k-eps (model 2) and no-diffusion (model 0) scalar column solvers        */
/*  U/V/W-prognostic-microscale_2.cs and _0.cs                              */
/* ===================================================================== */
static inline float pow_d(float Dv, float Pv) {    /* D*FloatMax(0,(float)Pow5(1-0.1*P)) */
    return Dv * vmax(0.0f, (float)p5d(1.0 - 0.1 * (double)Pv));
}

/* wall function of U/V _0/_2 (U-prognostic-microscale_2.cs:188-212) */
static float wall_uv_scalar(St *s, int i, int j, int k, const float *UKSim, const float *UKS,
                            const float *VKSim, const float *VKS, float vel) {
    double a = 0.5 * (double)(UKSim[k] + UKS[k]);
    double b = 0.5 * (double)(VKSim[k] + VKS[k]);
    float wh = vmax((float)sqrt(p2d(a) + p2d(b)), 0.01f);
    float Us;
    if (s->CUTK[IX2(s, i, j)] < 1) {
        Us = s->HT[IX2(s, i, j)] * wh;
        if ((double)s->Z0S[IX2(s, i, j)] >= (double)s->DZK[k] * 0.1) {
            double a1 = 0.5 * (double)(UKSim[k + 1] + UKS[k + 1]);
            double b1 = 0.5 * (double)(VKSim[k + 1] + VKS[k + 1]);
            Us = s->HT[IX2(s, i, j)] * (float)sqrt(p2d(a1) + p2d(b1));
        }
    } else {
        Us = s->HO[IX2(s, i, j)] * wh;
    }
    return vel / wh * p2f(Us) * s->AREAxy;
}

static void uv2_sweep(St *s, int IS, int JS, int is_v, int diffusion) {
    float DXK = s->DXK, DYK = s->DYK, AREAxy = s->AREAxy;
    float Cmueh = 0.09f;
    int i1a = is_v ? 2 : 3, j1a = is_v ? 3 : 2;
    for (int i1 = i1a; i1 < s->NII; ++i1) {
        for (int j1 = j1a; j1 <= s->NJJ - 1; ++j1) {
            int j = (JS == -1) ? s->NJJ - j1 + 1 : j1;
            int i = (IS == -1) ? s->NII - i1 + 1 : i1;
            if (ADV(s, i, j) != 1) continue;
            const float *TURB = diffusion ? COL(s, TURB, i, j) : NULL;
            float *X = is_v ? COL(s, VK, i, j) : COL(s, UK, i, j);
            const float *Xim = is_v ? COL(s, VK, i - 1, j) : COL(s, UK, i - 1, j);
            const float *Xip = is_v ? COL(s, VK, i + 1, j) : COL(s, UK, i + 1, j);
            const float *Xjm = is_v ? COL(s, VK, i, j - 1) : COL(s, UK, i, j - 1);
            const float *Xjp = is_v ? COL(s, VK, i, j + 1) : COL(s, UK, i, j + 1);
            const float *UKS = COL(s, UKS, i, j), *VKS = COL(s, VKS, i, j), *WKS = COL(s, WKS, i, j);
            const float *DPMNEW = COL(s, DPMNEW, i, j);
            const float *UKSim = COL(s, UKS, i - 1, j), *UKSip = COL(s, UKS, i + 1, j);
            const float *UKSjm = COL(s, UKS, i, j - 1), *UKSipjm = COL(s, UKS, i + 1, j - 1);
            const float *UKSimjm = COL(s, UKS, i - 1, j - 1), *UKSipjp = COL(s, UKS, i + 1, j + 1);
            const float *UKSimjp = COL(s, UKS, i - 1, j + 1);
            const float *VKSim = COL(s, VKS, i - 1, j), *VKSimjm = COL(s, VKS, i - 1, j - 1);
            const float *VKSjm = COL(s, VKS, i, j - 1), *VKSimjp = COL(s, VKS, i - 1, j + 1);
            const float *VKSjp = COL(s, VKS, i, j + 1), *VKSipjp = COL(s, VKS, i + 1, j + 1);
            const float *VKSipjm = COL(s, VKS, i + 1, j - 1);
            const float *WKSim = COL(s, WKS, i - 1, j), *WKSjm = COL(s, WKS, i, j - 1);
            const float *DPMNEWm = is_v ? COL(s, DPMNEW, i, j - 1) : COL(s, DPMNEW, i - 1, j);
            const float *WKip = COL(s, WK, i + 1, j), *WKjp = COL(s, WK, i, j + 1);
            float G = is_v ? s->VG : s->UG;
            int KK = KK(s, i, j), VIc = s->VI[IX2(s, i, j)];
            int KSTART = (s->CUTK[IX2(s, i, j)] == 0) ? KK + 1 : 1;
            float *P = s->P1, *Q = s->Q1;
            for (int k = KSTART; k <= VIc; ++k) {
                float DZ = s->DZK[k], DXKDZK = DXK * DZ, DYKDZK = DYK * DZ;
                float VIS = 0, DE = 0, DW = 0, DS = 0, DN = 0, DT = 0, DB = 0;
                if (diffusion) {
                    float zs = s->HOKART[k] - s->HOKART[KK] - DZ * 0.5f;
                    VIS = (float)sqrt((double)TURB[k]) * zs * Cmueh;
                    DE = vmax(VIS, s->VISHMIN) * DYKDZK / DXK; DW = DE;
                    DS = vmax(VIS, s->VISHMIN) * DXKDZK / DYK; DN = DS;
                    DT = vmax(VIS, s->VISHMIN) * AREAxy / DZ;
                    if (k > KK + 1) DB = DT;
                    if ((ADV(s, i, j - 1) < 1) || (j == 2)) DS = 0;
                    if ((ADV(s, i, j + 1) < 1) || (j == s->NJJ - 1)) DN = 0;
                    if ((ADV(s, i - 1, j) < 1) || (i == 2)) DW = 0;
                    if ((ADV(s, i + 1, j) < 1) || (i == s->NII - 1)) DE = 0;
                }
                float FE, FW, FS, FN, FT, FB = 0;
                if (!is_v) {
                    FE = UKS[k] * DYKDZK;
                    FW = UKSim[k] * DYKDZK;
                    FS = 0.25f * (VKSim[k] + VKS[k] + VKSimjm[k] + VKSjm[k]) * DXKDZK;
                    FN = 0.25f * (VKSim[k] + VKS[k] + VKSimjp[k] + VKSjp[k]) * DXKDZK;
                    FT = 0.25f * (WKSim[k] + WKS[k] + WKSim[k + 1] + WKS[k + 1]) * AREAxy;
                    if (k > KK + 1) FB = 0.25f * (WKSim[k] + WKS[k] + WKSim[k - 1] + WKS[k - 1]) * AREAxy;
                } else {
                    FE = 0.25f * (UKS[k] + UKSip[k] + UKSjm[k] + UKSipjm[k]) * DYKDZK;
                    FW = 0.25f * (UKS[k] + UKSim[k] + UKSjm[k] + UKSimjm[k]) * DYKDZK;
                    FS = VKSjm[k] * DXKDZK;
                    FN = VKS[k] * DXKDZK;
                    FT = 0.25f * (WKSjm[k] + WKS[k] + WKSjm[k + 1] + WKS[k + 1]) * AREAxy;
                    if (k > KK + 1) FB = 0.25f * (WKSjm[k] + WKS[k] + WKSjm[k - 1] + WKS[k - 1]) * AREAxy;
                }
                float BIM, CIM, AE1, AW1, AS1, AN1;
                if (diffusion) {
                    DE = vmax(DE, 0.0001f); DB = vmax(DB, 0.0001f); DW = vmax(DW, 0.0001f);
                    DS = vmax(DS, 0.0001f); DN = vmax(DN, 0.0001f); DT = vmax(DT, 0.0001f);
                    float PE = fabsf(FE / DE), PB = fabsf(FB / DB), PW = fabsf(FW / DW);
                    float PS = fabsf(FS / DS), PN = fabsf(FN / DN), PT = fabsf(FT / DT);
                    BIM = pow_d(DT, PT) + vmax(-FT, 0);
                    CIM = pow_d(DB, PB) + vmax(FB, 0);
                    AE1 = pow_d(DE, PE) + vmax(-FE, 0);
                    AW1 = pow_d(DW, PW) + vmax(FW, 0);
                    AS1 = pow_d(DS, PS) + vmax(FS, 0);
                    AN1 = pow_d(DN, PN) + vmax(-FN, 0);
                } else {
                    BIM = vmax(-FT, 0.0f); CIM = vmax(FB, 0.0f); AE1 = vmax(-FE, 0.0f);
                    AW1 = vmax(FW, 0.0f); AS1 = vmax(FS, 0.0f); AN1 = vmax(-FN, 0.0f);
                }
                float AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + s->AP0[k];
                float DDP = DPMNEWm[k] - DPMNEW[k];
                float Xv = X[k];
                float partial = AW1 * Xim[k] + AS1 * Xjm[k] + AE1 * Xip[k] + AN1 * Xjp[k];
                double t5 = (double)s->AP0[k] * 0.5 * (double)(is_v ? (VKSjm[k] + VKS[k]) : (UKSim[k] + UKS[k]));
                float t6 = DDP * (is_v ? DXKDZK : DYKDZK);
                float t7 = s->Cor * (G - Xv) * AREAxy * DZ;
                float DIMX = (float)((double)partial + t5 + (double)t6 + (double)t7);
                if (k == KK + 1) DIMX -= (float)wall_uv_scalar(s, i, j, k, UKSim, UKS, VKSim, VKS, Xv);
                if (diffusion && k > KK + 1) {
                    float ADD;
                    if (!is_v) {
                        float DUDXE = (Xip[k] - Xv) * DYKDZK;
                        float DUDXW = (Xv - Xim[k]) * DYKDZK;
                        float DVDXN = (0.5f * (VKSipjp[k] + VKSjp[k]) - 0.5f * (VKSjp[k] + VKSimjp[k])) * DXKDZK;
                        float DVDXS = (0.5f * (VKSipjm[k] + VKSjm[k]) - 0.5f * (VKSjm[k] + VKSimjm[k])) * DXKDZK;
                        float DWDXT = (0.5f * (WKip[k] + WKS[k]) - 0.5f * (WKS[k] + WKSim[k])) * AREAxy;
                        float DWDXB = (0.5f * (WKip[k - 1] + WKS[k - 1]) - 0.5f * (WKS[k - 1] + WKSim[k - 1])) * AREAxy;
                        ADD = VIS / DXK * (DUDXE - DUDXW + DVDXN - DVDXS + DWDXT - DWDXB);
                    } else {
                        float DVDYN = (Xjp[k] - Xv) * DXKDZK;
                        float DVDYS = (Xv - Xjm[k]) * DXKDZK;
                        float DUDYE = (0.5f * (UKSipjp[k] + UKSip[k]) - 0.5f * (UKSip[k] + UKSipjm[k])) * DXKDZK;
                        float DUDYW = (0.5f * (UKSimjp[k] + UKSim[k]) - 0.5f * (UKSim[k] + UKSimjm[k])) * DXKDZK;
                        float DWDYT = (0.5f * (WKjp[k] + WKS[k]) - 0.5f * (WKS[k] + WKSjm[k])) * AREAxy;
                        float DWDYB = (0.5f * (WKjp[k - 1] + WKS[k - 1]) - 0.5f * (WKS[k - 1] + WKSjm[k - 1])) * AREAxy;
                        ADD = VIS / DXK * (DUDYE - DUDYW + DVDYN - DVDYS + DWDYT - DWDYB);
                    }
                    DIMX += ADD;
                }
                if (k > KK + 1) {
                    P[k] = BIM / (AIM - CIM * P[k - 1]);
                    Q[k] = (DIMX + CIM * Q[k - 1]) / (AIM - CIM * P[k - 1]);
                } else {
                    P[k] = BIM / AIM;
                    Q[k] = DIMX / AIM;
                }
            }
            int KKn = is_v ? KK(s, i, j - 1) : KK(s, i - 1, j);
            for (int k = VIc; k >= KSTART; --k)
                if ((KK < k) && (KKn < k)) X[k] += s->relax * (P[k] * X[k + 1] + Q[k] - X[k]);
        }
    }
}

static void w2_sweep(St *s, int IS, int JS, int diffusion) {
    float DXK = s->DXK, DYK = s->DYK, AREAxy = s->AREAxy, Cmueh = 0.09f;
    for (int i1 = 2; i1 < s->NII; ++i1) {
        for (int j1 = 2; j1 <= s->NJJ - 1; ++j1) {
            int j = (JS == -1) ? s->NJJ - j1 + 1 : j1;
            int i = (IS == -1) ? s->NII - i1 + 1 : i1;
            if (ADV(s, i, j) != 1) continue;
            const float *TURB = diffusion ? COL(s, TURB, i, j) : NULL;
            float *WK = COL(s, WK, i, j);
            const float *WKim = COL(s, WK, i - 1, j), *WKip = COL(s, WK, i + 1, j);
            const float *WKjm = COL(s, WK, i, j - 1), *WKjp = COL(s, WK, i, j + 1);
            const float *UKS = COL(s, UKS, i, j), *VKS = COL(s, VKS, i, j), *WKS = COL(s, WKS, i, j);
            const float *DPMNEW = COL(s, DPMNEW, i, j);
            const float *UKSim = COL(s, UKS, i - 1, j), *UKSip = COL(s, UKS, i + 1, j);
            const float *VKSjm = COL(s, VKS, i, j - 1), *VKSjp = COL(s, VKS, i, j + 1);
            int KK = KK(s, i, j), VIc = s->VI[IX2(s, i, j)];
            int KSTART = (s->CUTK[IX2(s, i, j)] == 0) ? KK + 1 : 1;
            float *P = s->P1, *Q = s->Q1;
            for (int k = KSTART; k <= VIc; ++k) {
                float DZ = s->DZK[k], DXKDZK = DXK * DZ, DYKDZK = DYK * DZ;
                float UKS_LL = UKS[k], UKS_LL_m = UKS[k - 1], VKS_LL = VKS[k], VKS_LL_m = VKS[k - 1];
                float VIS = 0, DE = 0, DW = 0, DS = 0, DN = 0, DT = 0, DB = 0;
                if (diffusion) {
                    float zs = s->HOKART[k] - s->HOKART[KK] - DZ * 0.5f;
                    VIS = (float)sqrt((double)TURB[k]) * zs * Cmueh;
                    DE = vmax(VIS, s->VISHMIN) * DYKDZK / DXK; DW = DE;
                    DS = vmax(VIS, s->VISHMIN) * DXKDZK / DYK; DN = DS;
                    DT = vmax(VIS, s->VISHMIN) * AREAxy / DZ;
                    if (k > KK + 1) DB = DT;
                    if ((ADV(s, i, j - 1) < 1) || (j == 2)) DS = 0;
                    if ((ADV(s, i, j + 1) < 1) || (j == s->NJJ - 1)) DN = 0;
                    if ((ADV(s, i - 1, j) < 1) || (i == 2)) DW = 0;
                    if ((ADV(s, i + 1, j) < 1) || (i == s->NII - 1)) DE = 0;
                }
                float FE = 0.25f * (UKS_LL + UKSip[k] + UKS_LL_m + UKSip[k - 1]) * DYKDZK;
                float FW = 0.25f * (UKS_LL + UKSim[k] + UKS_LL_m + UKSim[k - 1]) * DYKDZK;
                float FS = 0.25f * (VKS_LL + VKSjm[k] + VKS_LL_m + VKSjm[k - 1]) * DXKDZK;
                float FN = 0.25f * (VKS_LL + VKSjp[k] + VKS_LL_m + VKSjp[k - 1]) * DXKDZK;
                float FT = WKS[k] * AREAxy;
                float FB = 0;
                if (k > KK + 1) FB = WKS[k - 1] * AREAxy;
                float BIM, CIM, AE1, AW1, AS1, AN1;
                if (diffusion) {
                    DE = vmax(DE, 0.0001f); DB = vmax(DB, 0.0001f); DW = vmax(DW, 0.0001f);
                    DS = vmax(DS, 0.0001f); DN = vmax(DN, 0.0001f); DT = vmax(DT, 0.0001f);
                    float PE = fabsf(FE / DE), PB = fabsf(FB / DB), PW = fabsf(FW / DW);
                    float PS = fabsf(FS / DS), PN = fabsf(FN / DN), PT = fabsf(FT / DT);
                    BIM = pow_d(DT, PT) + vmax(-FT, 0);
                    CIM = pow_d(DB, PB) + vmax(FB, 0);
                    AE1 = pow_d(DE, PE) + vmax(-FE, 0);
                    AW1 = pow_d(DW, PW) + vmax(FW, 0);
                    AS1 = pow_d(DS, PS) + vmax(FS, 0);
                    AN1 = pow_d(DN, PN) + vmax(-FN, 0);
                } else {
                    BIM = vmax(-FT, 0.0f); CIM = vmax(FB, 0.0f); AE1 = vmax(-FE, 0.0f);
                    AW1 = vmax(FW, 0.0f); AS1 = vmax(FS, 0.0f); AN1 = vmax(-FN, 0.0f);
                }
                float AP0 = s->AP0[k];
                float AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + AP0;
                float DDPZ = DPMNEW[k - 1] - DPMNEW[k];
                float partial = AW1 * WKim[k] + AS1 * WKjm[k] + AE1 * WKip[k] + AN1 * WKjp[k];
                float DIMW = (float)((double)partial + (double)AP0 * 0.5 * (double)(WKS[k - 1] + WKS[k]) + (double)(DDPZ * AREAxy));
                if (diffusion && k > KK + 1) {
                    float DUDZE = (0.5f * (UKS[k + 1] + UKS_LL) - 0.5f * (UKS_LL + UKS_LL_m)) * DYKDZK;
                    float DUDZW = (0.5f * (UKSim[k + 1] + UKSim[k]) - 0.5f * (UKSim[k] + UKSim[k - 1])) * DYKDZK;
                    float DVDZN = (0.5f * (VKS[k + 1] + VKS_LL) - 0.5f * (VKS_LL + VKS_LL_m)) * DXKDZK;
                    float DVDZS = (0.5f * (VKSjm[k + 1] + VKSjm[k]) - 0.5f * (VKSjm[k] + VKSjm[k - 1])) * DXKDZK;
                    float DWDZT = (WK[k + 1] - WK[k]) * AREAxy;
                    float DWDZB = (WK[k] - WK[k - 1]) * AREAxy;
                    DIMW += VIS / DZ * (DUDZE - DUDZW + DVDZN - DVDZS + DWDZT - DWDZB);
                }
                if (k > KK + 1) {
                    P[k] = BIM / (AIM - CIM * P[k - 1]);
                    Q[k] = (DIMW + CIM * Q[k - 1]) / (AIM - CIM * P[k - 1]);
                } else {
                    P[k] = BIM / AIM;
                    Q[k] = DIMW / AIM;
                }
            }
            for (int k = VIc; k >= KSTART; --k) {
                if (diffusion) WK[k] += s->relax * (P[k] * WK[k + 1] + Q[k] - WK[k]);
                else WK[k] += (float)(s->relax * (P[k] * WK[k + 1] + Q[k] - WK[k]));
            }
        }
    }
}

/* TKE_PrognosticMicroscale.Calculate (TKE-prognostic-microscale.cs:25-352) */
static void tke_sweep(St *s, int IS, int JS) {
    float Cmueh = 0.09f, Ceps = 1, Ceps1 = 1.44f, Ceps2 = 1.92f;
    s->comp = 4;
    float DXK = s->DXK, DYK = s->DYK, AREAxy = s->AREAxy;
    float DXK_Rez = 1 / DXK, DYK_Rez = 1 / DYK;
    float *PT1 = s->P1, *QT1 = s->Q1, *PT2 = s->P2, *QT2 = s->Q2;
    for (int i1 = 2; i1 < s->NII; ++i1) {
        for (int j1 = 2; j1 <= s->NJJ - 1; ++j1) {
            int j = (JS == -1) ? s->NJJ - j1 + 1 : j1;
            int i = (IS == -1) ? s->NII - i1 + 1 : i1;
            if (ADV(s, i, j) != 1) continue;
            float *TURB = COL(s, TURB, i, j), *TDISS = COL(s, TDISS, i, j);
            const float *TURBim = COL(s, TURB, i - 1, j), *TURBip = COL(s, TURB, i + 1, j);
            const float *TURBjm = COL(s, TURB, i, j - 1), *TURBjp = COL(s, TURB, i, j + 1);
            const float *TDISSim = COL(s, TDISS, i - 1, j), *TDISSip = COL(s, TDISS, i + 1, j);
            const float *TDISSjm = COL(s, TDISS, i, j - 1), *TDISSjp = COL(s, TDISS, i, j + 1);
            const float *UK = COL(s, UK, i, j), *VK = COL(s, VK, i, j), *WK = COL(s, WK, i, j);
            const float *UKip = COL(s, UK, i + 1, j), *UKjp = COL(s, UK, i, j + 1), *UKipjp = COL(s, UK, i + 1, j + 1);
            const float *UKjm = COL(s, UK, i, j - 1), *UKipjm = COL(s, UK, i + 1, j - 1);
            const float *VKjp = COL(s, VK, i, j + 1), *VKip = COL(s, VK, i + 1, j), *VKipjp = COL(s, VK, i + 1, j + 1);
            const float *VKim = COL(s, VK, i - 1, j), *VKimjp = COL(s, VK, i - 1, j + 1);
            const float *WKip = COL(s, WK, i + 1, j), *WKim = COL(s, WK, i - 1, j);
            const float *WKjp = COL(s, WK, i, j + 1), *WKjm = COL(s, WK, i, j - 1);
            float Z0 = s->Z0S[IX2(s, i, j)];
            int KK = KK(s, i, j);
            int KSTART = (s->CUTK[IX2(s, i, j)] == 0) ? KK + 1 : 1;
            int VIc = s->VI[IX2(s, i, j)];
            for (int k = KSTART; k <= VIc; ++k) {
                float DZ = s->DZK[k];
                float DXKDZK = DXK * DZ, DYKDZK = DYK * DZ, DZK_Rez = 1 / DZ;
                float zs = s->HOKART[k] - s->HOKART[KK] - DZ * 0.5f;
                float VIS = (float)sqrt((double)TURB[k]) * zs * Cmueh;
                VIS = mminf_(VIS, 15);
                float DE = mmaxf_(VIS, s->VISHMIN) * DYKDZK * DXK_Rez, DW = DE;
                float DS = mmaxf_(VIS, s->VISHMIN) * DXKDZK * DYK_Rez, DN = DS;
                float DT = mmaxf_(VIS, s->VISHMIN) * AREAxy * DZK_Rez;
                float DB = 0;
                float DTEPS = mmaxf_(VIS / Ceps, s->VISVMIN) * AREAxy * DZK_Rez;
                float DBEPS = 0;
                if (k > KK + 1) { DB = DT; DBEPS = DTEPS; }
                if ((ADV(s, i, j - 1) < 1) || (j == 2)) DS = 0;
                if ((ADV(s, i, j + 1) < 1) || (j == s->NJJ - 1)) DN = 0;
                if ((ADV(s, i - 1, j) < 1) || (i == 2)) DW = 0;
                if ((ADV(s, i + 1, j) < 1) || (i == s->NII - 1)) DE = 0;
                float FE = UKip[k] * DYKDZK, FW = UK[k] * DYKDZK;
                float FS = VK[k] * DXKDZK, FN = VKjp[k] * DXKDZK;
                float FT = WK[k + 1] * AREAxy, FB = 0;
                if (k > KK + 1) FB = WK[k] * AREAxy;
                DE = mmaxf_(DE, 0.0001f); DB = mmaxf_(DB, 0.0001f); DW = mmaxf_(DW, 0.0001f);
                DS = mmaxf_(DS, 0.0001f); DN = mmaxf_(DN, 0.0001f); DT = mmaxf_(DT, 0.0001f);
                DBEPS = mmaxf_(DBEPS, 0.0001f); DTEPS = mmaxf_(DTEPS, 0.0001f);
                float PE = fabsf(FE / DE), PB = fabsf(FB / DB), PW = fabsf(FW / DW);
                float PS = fabsf(FS / DS), PN = fabsf(FN / DN), PT = fabsf(FT / DT);
                float PBEPS = fabsf(FN / DBEPS);      /* quirk: FN, not FB */
                float PTEPS = fabsf(FT / DTEPS);
                float BIM = DT * mmaxf_(0, p5f(1 - 0.1f * PT)) + mmaxf_(-FT, 0);
                float CIM = DB * mmaxf_(0, p5f(1 - 0.1f * PB)) + mmaxf_(FB, 0);
                float AE1 = DE * mmaxf_(0, p5f(1 - 0.1f * PE)) + mmaxf_(-FE, 0);
                float AW1 = DW * mmaxf_(0, p5f(1 - 0.1f * PW)) + mmaxf_(FW, 0);
                float AS1 = DS * mmaxf_(0, p5f(1 - 0.1f * PS)) + mmaxf_(FS, 0);
                float AN1 = DN * mmaxf_(0, p5f(1 - 0.1f * PN)) + mmaxf_(-FN, 0);
                float BIMEPS = DT * mmaxf_(0, p5f(1 - 0.1f * PTEPS)) + mmaxf_(-FT, 0);   /* quirk: DT */
                float CIMEPS = DB * mmaxf_(0, p5f(1 - 0.1f * PBEPS)) + mmaxf_(FB, 0);    /* quirk: DB */
                float AP0 = s->AP0[k];
                float AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + AP0;
                float UX1 = UKip[k], UX2 = UK[k];
                float VX1 = 0.5f * (VKip[k] + VKipjp[k]), VX2 = 0.5f * (VKim[k] + VKimjp[k]);
                float WX1 = 0.5f * (WKip[k] + WKip[k + 1]), WX2 = 0.5f * (WKim[k] + WKim[k + 1]);
                float UY1 = 0.5f * (UKjp[k] + UKipjp[k]), UY2 = 0.5f * (UKjm[k] + UKipjm[k]);
                float VY1 = VKjp[k], VY2 = VK[k];
                float WY1 = 0.5f * (WKjp[k] + WKjp[k + 1]), WY2 = 0.5f * (WKjm[k] + WKjm[k + 1]);
                float DUDX = (UX1 - UX2) * DXK_Rez;
                float DVDX = (VX1 - VX2) * DXK_Rez * 0.5f;
                float DWDX = (WX1 - WX2) * DXK_Rez * 0.5f;
                float DUDY = (UY1 - UY2) * DYK_Rez * 0.5f;
                float DVDY = (VY1 - VY2) * DYK_Rez;
                float DWDY = (WY1 - WY2) * DYK_Rez * 0.5f;
                float DUDZ = UK[k + 1] * DZK_Rez * 0.5f;
                float DVDZ = VK[k + 1] * DZK_Rez * 0.5f;
                float DWDZ = WK[k + 1] * DZK_Rez * 0.5f;
                if (k > KK + 1) {
                    DUDZ = (UK[k + 1] - UK[k - 1]) * DZK_Rez * 0.5f;
                    DVDZ = (VK[k + 1] - VK[k - 1]) * DZK_Rez * 0.5f;
                    DWDZ = (WK[k + 1] - WK[k - 1]) * DZK_Rez * 0.5f;
                }
                float PROTURB = VIS * AREAxy * DZ *
                    (2.0f * p2f(DUDX) + 2.0f * p2f(DVDY) + 2.0f * p2f(DWDZ) +
                     p2f(DUDY + DVDX) + p2f(DUDZ + DWDX) + (ORC_MUTANT == 6 ? 0.0f : p2f(DVDZ + DWDY)) - TDISS[k]);
                float PROBOU = (float)((double)(-s->PotdTdz) * 9.81 / 283.0 * (double)VIS * 1.35 * (double)AREAxy * (double)DZ);
                s->rec = 1;
                piece(s, i, j, ORC_VIS, k, VIS);
                piece(s, i, j, ORC_PROD, k, PROTURB);
                piece(s, i, j, ORC_BUOY, k, PROBOU);
                s->rec = 0;
                float DIMTURB = AW1 * TURBim[k] + AS1 * TURBjm[k] + AE1 * TURBip[k] + AN1 * TURBjp[k] +
                                AP0 * TURB[k] + PROTURB + PROBOU;
                float DIMTDISS = AW1 * TDISSim[k] + AS1 * TDISSjm[k] + AE1 * TDISSip[k] + AN1 * TDISSjp[k] +
                                 AP0 * TDISS[k] + (TDISS[k] / TURB[k] * (Ceps1 * PROTURB - Ceps2 * TDISS[k]));
                if (k > KK + 1) {
                    PT1[k] = BIM / (AIM - CIM * PT1[k - 1]);
                    QT1[k] = (DIMTURB + CIM * QT1[k - 1]) / (AIM - CIM * PT1[k - 1]);
                    PT2[k] = BIMEPS / (AIM - CIMEPS * PT2[k - 1]);
                    QT2[k] = (DIMTDISS + CIMEPS * QT2[k - 1]) / (AIM - CIMEPS * PT2[k - 1]);
                } else {
                    PT1[k] = BIM / AIM;
                    QT1[k] = DIMTURB / AIM;
                    PT2[k] = BIMEPS / AIM;
                    QT2[k] = DIMTDISS / AIM;
                }
            }
            for (int k = VIc; k >= KSTART; --k) {
                if (!(KK < k)) continue;
                float Us;
                if ((k == KK + 1) && (s->CUTK[IX2(s, i, j)] == 0)) {
                    if ((double)Z0 >= (double)s->DZK[k] * 0.5) {
                        double hz = (double)s->DZK[k] + (double)s->DZK[k + 1] * 0.5;
                        Us = (float)(0.4 / log(hz / (double)Z0) *
                                     sqrt(p2d(0.5 * (double)(UK[k + 1] + UKip[k + 1])) + p2d(0.5 * (double)(VK[k + 1] + VKjp[k + 1]))));
                        TURB[k] = (float)(3.33 * (double)p2f(Us));
                        TDISS[k] = (float)((double)p3f(Us) / (hz * 0.4));
                    } else {
                        Us = (float)(0.4 / log((double)s->DZK[k] * 0.5 / (double)Z0) *
                                     sqrt(p2d(0.5 * (double)(UK[k + 1] + UKip[k])) + p2d(0.5 * (double)(VK[k] + VKjp[k]))));
                        TURB[k] = (float)(3.33 * (double)p2f(Us));
                        TDISS[k] = (float)((double)p3f(Us) / ((double)s->DZK[k] * 0.5 * 0.4));
                    }
                } else if ((k == KK + 1) && (s->CUTK[IX2(s, i, j)] > 0)) {
                    Us = s->HO[IX2(s, i, j)] * (float)sqrt(p2d(0.5 * (double)(UK[k] + UKip[k])) + p2d(0.5 * (double)(VK[k] + VKjp[k])));
                    TURB[k] = (float)(3.33 * (double)p2f(Us));
                    TDISS[k] = (float)((double)p3f(Us) / ((double)s->DZK[k] * 0.5 * 0.4));
                } else if ((k == KK(s, i + 1, j) + 1) && (s->CUTK[IX2(s, i + 1, j)] > 0)) {
                    Us = (float)((double)s->Ustern_factorX * sqrt(p2d(0.5 * (double)(WK[k] + WK[k + 1])) + p2d(0.5 * (double)(VK[k] + VKjp[k]))));
                    TURB[k] = (float)(3.33 * (double)p2f(Us));
                    TDISS[k] = (float)((double)p3f(Us) / ((double)s->building_Z0 * 0.4));
                } else if ((k == KK(s, i - 1, j) + 1) && (s->CUTK[IX2(s, i - 1, j)] > 0)) {
                    Us = (float)((double)s->Ustern_factorX * sqrt(p2d(0.5 * (double)(WK[k] + WK[k + 1])) + p2d(0.5 * (double)(VK[k] + VKjp[k]))));
                    TURB[k] = (float)(3.33 * (double)p2f(Us));
                    TDISS[k] = (float)((double)p3f(Us) / ((double)s->building_Z0 * 0.4));
                } else if ((k == KK(s, i, j + 1) + 1) && (s->CUTK[IX2(s, i, j + 1)] > 0)) {
                    Us = (float)((double)s->Ustern_factorX * sqrt(p2d(0.5 * (double)(WK[k] + WK[k + 1])) + p2d(0.5 * (double)(UK[k] + UKip[k]))));
                    TURB[k] = (float)(3.33 * (double)p2f(Us));
                    TDISS[k] = (float)((double)p3f(Us) / ((double)s->building_Z0 * 0.4));
                } else if ((k == KK(s, i, j - 1) + 1) && (s->CUTK[IX2(s, i, j - 1)] > 0)) {
                    Us = (float)((double)s->Ustern_factorX * sqrt(p2d(0.5 * (double)(WK[k] + WK[k + 1])) + p2d(0.5 * (double)(UK[k] + UKip[k]))));
                    TURB[k] = (float)(3.33 * (double)p2f(Us));
                    TDISS[k] = (float)((double)p3f(Us) / ((double)s->building_Z0 * 0.4));
                } else {
                    TURB[k] = PT1[k] * TURB[k + 1] + QT1[k];
                    TDISS[k] = PT2[k] * TDISS[k + 1] + QT2[k];
                }
                TURB[k] = (float)mmaxd_((double)TURB[k], 0.001);
                TDISS[k] = (float)mmaxd_((double)TDISS[k], 0.0000001);
            }
        }
    }
}

/* ===================================================================== */
/*  PrognosticFlowfield.Calculate()                                        */
/* ===================================================================== */
static bool dilated(const St *s, int i, int j) {
    int ip = i + 1 < s->NII + 1 ? i + 1 : s->NII + 1;
    int jp = j + 1 < s->NJJ + 1 ? j + 1 : s->NJJ + 1;
    return ADV(s, i, j) == 1 || ADV(s, i - 1, j) == 1 || ADV(s, ip, j) == 1 ||
           ADV(s, i, j - 1) == 1 || ADV(s, i - 1, j - 1) == 1 || ADV(s, ip, j - 1) == 1 ||
           ADV(s, i, jp) == 1 || ADV(s, i - 1, jp) == 1 || ADV(s, ip, jp) == 1;
}

float orc_powlaw_vector(float d, float peclet) { return ml_pow(d, peclet); }
float orc_powlaw_scalar(float d, float peclet) { return pow_d(d, peclet); }
int orc_conv_to_int(double x) { return conv_to_int_d(x); }

static int calculate(const orc_params *p, const orc_inputs *in, orc_fields *f, orc_stats *st, orc_capture *cap,
                     int nthreads, int proj, int quantise);

int orc_calculate_threads(const orc_params *p, const orc_inputs *in, orc_fields *f, orc_stats *st, int nthreads) {
    return calculate(p, in, f, st, NULL, nthreads < 1 ? 1 : nthreads, 0, 1);
}

int orc_calculate(const orc_params *p, const orc_inputs *in, orc_fields *f, orc_stats *st) {
    return calculate(p, in, f, st, NULL, 1, 0, 1);
}

int orc_project(const orc_params *p, const orc_inputs *in, orc_fields *f, orc_stats *st, int quantise) {
    return calculate(p, in, f, st, NULL, 1, 1, quantise);
}

int orc_calculate_capture(const orc_params *p, const orc_inputs *in, orc_fields *f, orc_stats *st, orc_capture *cap) {
    if (cap && (!cap->a || !cap->b || !cap->c || !cap->d || !cap->x_old || !cap->x_new)) return -7;
    if (cap) cap->captured = 0;
    return calculate(p, in, f, st, cap, 1, 0, 1);
}

static int calculate(const orc_params *p, const orc_inputs *in, orc_fields *f, orc_stats *st, orc_capture *cap,
                     int nthreads, int proj, int quantise) {
    if (!p || !in || !f || !st) return -1;
    if (p->simd_width != 8 && p->simd_width != 16) return -2;
    if (p->turbulence_model < 0 || p->turbulence_model > 2) return -3;
    if (p->turbulence_model != 1 && (!f->turb || !f->tdiss)) return -4;
    if (p->vertcellsff < p->simd_width) return -5;
    St S, *s = &S;
    memset(s, 0, sizeof(S));
    s->cap = cap;
    s->nthreads = nthreads;
    s->proj = proj;
    s->noquant = !quantise;
    s->NII = p->nii; s->NJJ = p->njj; s->NKK = p->nkk; s->KADVMAX = p->kadvmax; s->VCF = p->vertcellsff;
    s->NJ2 = p->njj + 2; s->NKP = p->nkk + 2; s->W = p->simd_width; s->model = p->turbulence_model;
    s->DXK = p->dxk; s->DYK = p->dyk; s->relax = p->relax; s->relaxp = p->relaxp;
    s->building_Z0 = p->building_z0; s->Cor = p->coriolis; s->PotdTdz = p->potdtdz;
    s->DZK = in->dzk; s->HOKART = in->hokart; s->ADVDOM = in->advdom; s->KKART_ = in->kkart;
    s->CUTK = in->cutk; s->COV = in->cov; s->Z0S = in->z0_solver; s->VEG = in->veg; s->VEGP = in->veg_present;
    if (s->VEGP && !s->VEG) return -6;
    s->UK = f->uk; s->VK = f->vk; s->WK = f->wk; s->UKS = f->uks; s->VKS = f->vks; s->WKS = f->wks;
    s->DPM = f->dpm; s->DPMNEW = f->dpmnew; s->DIV = f->div; s->TURB = f->turb; s->TDISS = f->tdiss;
    int NII = s->NII, NJJ = s->NJJ, NKK = s->NKK;
    size_t n2 = (size_t)(NII + 2) * (size_t)(NJJ + 2), n3 = n2 * (size_t)s->NKP;
    memset(s->UKS, 0, n3 * sizeof(float)); memset(s->VKS, 0, n3 * sizeof(float));
    memset(s->WKS, 0, n3 * sizeof(float)); memset(s->DPMNEW, 0, n3 * sizeof(float));
    memset(s->DIV, 0, n3 * sizeof(float));
    s->VI = (int *)calloc(n2, sizeof(int));
    s->HT = (float *)calloc(n2, sizeof(float));
    s->HO = (float *)calloc(n2, sizeof(float));
    s->AP0 = (float *)calloc((size_t)s->KADVMAX + 2, sizeof(float));
    s->P1 = (float *)calloc((size_t)s->NKP, sizeof(float)); s->Q1 = (float *)calloc((size_t)s->NKP, sizeof(float));
    s->P2 = (float *)calloc((size_t)s->NKP, sizeof(float)); s->Q2 = (float *)calloc((size_t)s->NKP, sizeof(float));

    /* --- prologue, PrognosticFlowfield.cs:145-579 --- */
    int rpi = p->refp_i, rpj = p->refp_j;
    float topwind = (float)sqrt((double)(p2f(s->UK[IX3(s, rpi, rpj, NKK - 1)]) + p2f(s->VK[IX3(s, rpi, rpj, NKK - 1)])));
    topwind = mmaxf_(topwind, 0.01f);
    float spd = sqrtf(p2f(s->UK[IX3(s, rpi, rpj, s->KADVMAX)]) + p2f(s->VK[IX3(s, rpi, rpj, s->KADVMAX)]));
    float DTIME = 0.5f * mminf_(s->DXK, s->DZK[1]) / mmaxf_(0.01f, spd);
    DTIME = mminf_(DTIME, 5);
    s->DTIME = DTIME;
    s->AREAxy = s->DXK * s->DYK;
    for (int k = 1; k <= s->KADVMAX; ++k) s->AP0[k] = s->AREAxy * s->DZK[k] / DTIME;
    for (int i = 1; i <= NII; ++i)
        for (int j = 1; j <= NJJ; ++j)
            for (int k = 1; k <= NKK + 1; ++k) s->DPM[IX3(s, i, j, k)] = 0;
    if (s->model != 1)
        for (int i = 1; i <= NII; ++i)
            for (int j = 1; j <= NJJ; ++j)
                for (int k = 1; k <= NKK + 1; ++k) {
                    s->TURB[IX3(s, i, j, k)] = 0.001f;
                    s->TDISS[IX3(s, i, j, k)] = 0.0000001f;
                }
    s->Frund = (double)(8000.0f / topwind);
    if (s->model == 2) s->Frund = (double)(10000.0f / topwind);
    double GESCHW_MAX = 0;
    for (int i = 1; i <= NII; ++i)
        for (int j = 1; j <= NJJ; ++j) {
            if (!dilated(s, i, j)) continue;
            for (int k = 1; k <= NKK - 1; ++k) {
                float uks = 0.5f * (s->UK[IX3(s, i, j, k)] + s->UK[IX3(s, i + 1, j, k)]);
                float vks = 0.5f * (s->VK[IX3(s, i, j, k)] + s->VK[IX3(s, i, j + 1, k)]);
                s->UKS[IX3(s, i, j, k)] = uks;
                s->VKS[IX3(s, i, j, k)] = vks;
                s->WKS[IX3(s, i, j, k)] = 0.5f * (s->WK[IX3(s, i, j, k)] + s->WK[IX3(s, i, j, k + 1)]);
                double g = sqrt((double)(p2f(uks) + p2f(vks)));
                GESCHW_MAX = mmaxd_(GESCHW_MAX, g);
            }
        }
    for (int i = 1; i <= NII; ++i)
        for (int j = 1; j <= NJJ; ++j) {
            int v = KK(s, i, j) + s->VCF;
            s->VI[IX2(s, i, j)] = v < NKK - 1 ? v : NKK - 1;
        }
    for (int i = 1; i <= NII; ++i)
        for (int j = 1; j <= NJJ; ++j) {
            int KK = KK(s, i, j);
            int KSTART = (s->CUTK[IX2(s, i, j)] == 0) ? KK + 1 : 1;
            float rough = in->z0_wall[IX2(s, i, j)];
            float CUTKv = s->CUTK[IX2(s, i, j)];
            for (int k = KSTART; k <= s->VI[IX2(s, i, j)]; ++k) {
                if (k != KK + 1) continue;
                if (CUTKv < 1) {
                    if ((double)rough >= (double)s->DZK[k] * 0.1)
                        s->HT[IX2(s, i, j)] = 0.4f / logf((s->DZK[k] + s->DZK[k + 1] * 0.5f) / s->DZK[k] * 0.1f);
                    else
                        s->HT[IX2(s, i, j)] = 0.4f / logf(s->DZK[k] * (ORC_MUTANT == 4 ? 1.0f : 0.5f) / rough);
                } else {
                    s->HO[IX2(s, i, j)] = 0.4f / logf(s->DZK[k] * 0.5f / s->building_Z0);
                }
            }
        }
    s->UG = s->UKS[IX3(s, rpi, rpj, NKK - 1)];
    s->VG = s->VKS[IX3(s, rpi, rpj, NKK - 1)];
    for (int i = 1; i <= NII; ++i)
        for (int j = 1; j <= NJJ; ++j) {
            float ObLength = in->ob_init[IX2(s, i, j)], ust = in->ust_init[IX2(s, i, j)], rough = in->z0_init[IX2(s, i, j)];
            for (int k = 1; k <= NKK - 1; ++k) {
                int ipo = 1;
                float V0int = p->v0_1;
                if (i == 1) s->UK[IX3(s, NII + 1, j, k)] = s->UK[IX3(s, NII, j, k)];
                if (j == 1) s->VK[IX3(s, i, NJJ + 1, k)] = s->VK[IX3(s, i, NJJ, k)];
                /* PrognosticFlowfield.cs:452 */
                float zhilf = s->HOKART[k] - (in->ahk[IX2(s, i, j)] - s->HOKART[KK(s, i, j)]);
                float U0int = 0;
                if (in->istatistics == 0 || in->istatistics == 3) {
                    /* PrognosticFlowfield.cs:456-481: measured sigma profile (MeteoZR / MeteoSonic) */
                    const float *MH = in->meas_height, *U0 = in->u0, *V0 = in->v0;
                    int NP = in->n_profile;
                    if (zhilf <= MH[1]) {
                        U0int = U0[1];
                        V0int = V0[1];
                    } else if (zhilf >= MH[NP]) {
                        U0int = U0[NP];
                        V0int = V0[NP];
                    } else {
                        for (int iprof = 1; iprof <= NP; iprof++)
                            if (zhilf > MH[iprof]) ipo = iprof + 1;
                        U0int = U0[ipo - 1] + (U0[ipo] - U0[ipo - 1]) / (MH[ipo] - MH[ipo - 1]) * (zhilf - MH[ipo - 1]);
                        /* the reference interpolates V0 with the slope of U0 (:478-479) */
                        V0int = V0[ipo - 1] + (U0[ipo] - U0[ipo - 1]) / (MH[ipo] - MH[ipo - 1]) * (zhilf - MH[ipo - 1]);
                    }
                } else {
                    double windhilf = mmaxd_(sqrt((double)(p2f(s->UK[IX3(s, i, j, k)]) + p2f(s->VK[IX3(s, i, j, k)]))), 0.01);
                    U0int = (float)(windhilf * (0.2 * pow(windhilf, -0.9) + 0.32 * (double)rough + 0.18));
                    U0int = (float)mmaxd_((double)U0int, 0.3) * p->std_dev_v;
                }
                float varw;
                if (ObLength < 0)
                    varw = p2f(ust) * p2f(1.15f + 0.1f * powf(p->bdlayheight / (-ObLength), 0.67f)) * p2f(p->std_dev_w);
                else
                    varw = p2f(ust * p->std_dev_w * 1.25f);
                if (s->model != 1)
                    s->TURB[IX3(s, i, j, k)] = (float)(0.5 * (double)(p2f(U0int) + p2f(V0int) + varw));
            }
        }
    s->VISHMIN = (s->model == 2) ? 0.01f : 0.0f;
    s->VISVMIN = (s->model == 2) ? 0.01f : 0.0f;
    s->Ustern_factorX = (float)(0.4 / log((double)s->DXK * 0.5 / (double)s->building_Z0));
    for (int i = 1; i <= NII; ++i)
        for (int j = 1; j <= NJJ; ++j) {
            if (ADV(s, i, j) != 1) continue;
            int VIc = s->VI[IX2(s, i, j)], KK = KK(s, i, j);
            for (int k = 1; k <= VIc; ++k) {
                if (KK >= k) {
                    s->UK[IX3(s, i, j, k)] = 0;
                    if (i == NII) s->UK[IX3(s, i + 1, j, k)] = 0;
                    s->VK[IX3(s, i, j, k)] = 0;
                    if (j == NJJ) s->VK[IX3(s, i, j + 1, k)] = 0;
                    s->WK[IX3(s, i, j, k)] = 0;
                    if (k == VIc) s->WK[IX3(s, i, j, k + 1)] = 0;
                    if (s->model != 1) { s->TURB[IX3(s, i, j, k)] = 0; s->TDISS[IX3(s, i, j, k)] = 0; }
                    s->UKS[IX3(s, i, j, k)] = 0; s->VKS[IX3(s, i, j, k)] = 0; s->WKS[IX3(s, i, j, k)] = 0;
                }
            }
        }

    /* --- SIMPLE loop, PrognosticFlowfield.cs:584-1178 --- */
    int IterationLoops = 1, ran = 0;
    float DeltaFinish = 0, DeltaFirst = 0;
    double DeltaMean = 0;
    while (IterationLoops <= p->iter_max && s->proj) {
        /* orc_project: the linear part of an iteration (:723-949) */
        divergence(s);
        int IS = 1, JS = 1;
        switch (IterationLoops % 4) {
            case 0: IS = -1; JS = 1; break;
            case 1: IS = -1; JS = -1; break;
            case 2: IS = 1; JS = -1; break;
            default: break;
        }
        (void)pressure(s, IS, JS);
        correction(s);
        ++ran;
        IterationLoops++;
    }
    while (IterationLoops <= p->iter_max && !s->proj) {
        s->iter = IterationLoops;
        boundary_conditions(s);
        divergence(s);
        int IS = 1, JS = 1;
        switch (IterationLoops % 4) {
            case 0: IS = -1; JS = 1; break;
            case 1: IS = -1; JS = -1; break;
            case 2: IS = 1; JS = -1; break;
            default: break;
        }
        float DeltaUMax = (s->nthreads > 1) ? pressure_par(s, IS, JS) : pressure(s, IS, JS);
        correction(s);
        averages(s);
        if (s->model == 0) {
            uv2_sweep(s, IS, JS, 0, 0);
            uv2_sweep(s, IS, JS, 1, 0);
            w2_sweep(s, IS, JS, 0);
        } else if (s->model == 1 && s->nthreads > 1) {
            ml_sweeps_par(s, IS, JS);
        } else if (s->model == 1) {
            u1_sweep(s, IS, JS);
            v1_sweep(s, -IS, -JS);
            w1_sweep(s, IS, JS);
        } else {
            uv2_sweep(s, IS, JS, 0, 1);
            uv2_sweep(s, IS, JS, 1, 1);
            w2_sweep(s, IS, JS, 1);
            tke_sweep(s, IS, JS);
        }
        if (st->delta_umax_hist) st->delta_umax_hist[IterationLoops - 1] = DeltaUMax;
        ++ran;
        if (IterationLoops >= p->iter_max) DeltaFinish = (float)(DeltaMean / (double)p2f(topwind));
        if ((IterationLoops % 100 == 0) && (IterationLoops > 1)) {
            double delta = DeltaMean / (double)p2f(topwind);
            if (IterationLoops == 100) DeltaFirst = (float)delta;
            if (delta < 0.012 && (IterationLoops >= p->iter_min)) break;
            DeltaMean = 0;
        } else {
            DeltaMean += DeltaUMax;
        }
        IterationLoops++;
    }

    /* --- epilogue checks, PrognosticFlowfield.cs:1187-1245 --- */
    st->convergence_problem = DeltaFinish > DeltaFirst;
    int nu = 0, nv = 0, nw = 0;
    float limit = (float)(10 * GESCHW_MAX);
    for (int i = 1; i <= NII; ++i)
        for (int j = 1; j <= NJJ; ++j) {
            if (ADV(s, i, j) != 1) continue;
            int kmax = KK(s, i, j) + s->VCF;
            if (kmax > NKK - 1) kmax = NKK - 1;
            for (int k = 1; k <= kmax; ++k) {
                float u = fabsf(s->UK[IX3(s, i, j, k)]), v = fabsf(s->VK[IX3(s, i, j, k)]), w = fabsf(s->WK[IX3(s, i, j, k)]);
                if (u > 55 || u > limit) ++nu;
                if (v > 55 || v > limit) ++nv;
                if (w > 20 || w > limit) ++nw;
            }
        }
    st->iteration_loops = IterationLoops;
    st->iterations_run = ran;
    st->delta_first = DeltaFirst; st->delta_finish = DeltaFinish;
    st->topwind = topwind; st->dtime = DTIME; st->ug = s->UG; st->vg = s->VG;
    st->ustern_factor_x = s->Ustern_factorX; st->frund = s->Frund; st->geschw_max = GESCHW_MAX;
    st->unreal_u = nu; st->unreal_v = nv; st->unreal_w = nw;
    free(s->VI); free(s->HT); free(s->HO); free(s->AP0);
    free(s->P1); free(s->Q1); free(s->P2); free(s->Q2);
    return 0;
}

/* ===================================================================== */
/*  DiagnosticFlowfield.Calculate, DiagnosticFlowfield.cs:24-436, and the   */
/*  vertical-velocity mass adjustment of MicroscaleFlat.cs:388-462 (or      */
/*  MicroscaleTerrain.cs:862-933)                                           */
/* ===================================================================== */
int orc_diagnostic(const orc_diag_params *p, const float *DZK, const uint8_t *ADVDOM, const int16_t *KKART,
                   float *UK, float *VK, float *WK, float *DPM, float *DIV, float *dtime_out) {
    if (!p || !DZK || !ADVDOM || !KKART || !UK || !VK || !WK || !DPM || !DIV) return -1;
    if (p->flow_field_level != 1 && p->flow_field_level != 2) return -2;
    const int NII = p->nii, NJJ = p->njj, NKK = p->nkk, NJ2 = NJJ + 2, NKH = NKK + 2;
    const float DXK = p->dxk, DYK = p->dyk;
#define D2(i, j) ((size_t)(i) * (size_t)NJ2 + (size_t)(j))
#define D3(i, j, k) (D2(i, j) * (size_t)NKH + (size_t)(k))
    /* DTIME = (float)(0.5*Math.Min(DXK,DYK)/Math.Sqrt(Pow2(UK)+Pow2(VK))) at (1,1,KADVMAX), :30-31 */
    float u = UK[D3(1, 1, p->kadvmax)], v = VK[D3(1, 1, p->kadvmax)];
    float s2 = u * u + v * v;
    float DTIME = (float)(0.5 * (double)mminf_(DXK, DYK) / sqrt((double)s2));
    DTIME = mminf_(DTIME, 5.0f);
    if (dtime_out) *dtime_out = DTIME;
    int Iterations = (p->flow_field_level == 2) ? 10 : 5;
    if (p->iterations > 0) Iterations = p->iterations;
    for (int i = 1; i <= NII; ++i)
        for (int j = 1; j <= NJJ; ++j)
            for (int k = 1; k <= NKK + 1; ++k) DPM[D3(i, j, k)] = 0;
    float *PIMP = (float *)calloc((size_t)NKK + 2, sizeof(float));
    float *QIMP = (float *)calloc((size_t)NKK + 2, sizeof(float));
    for (int it = 1; it <= Iterations; ++it) {
        /* mass divergence, :53-80 */
        for (int i = 2; i < NII; ++i)
            for (int j = 2; j < NJJ; ++j) {
                int KK = KKART[D2(i, j)];
                const float *uk = &UK[D3(i, j, 0)], *ukp = &UK[D3(i + 1, j, 0)];
                const float *vk = &VK[D3(i, j, 0)], *vkp = &VK[D3(i, j + 1, 0)], *wk = &WK[D3(i, j, 0)];
                float *dv = &DIV[D3(i, j, 0)];
                for (int k = 1; k <= NKK - 1; ++k) {
                    if (!(KK < k)) continue;
                    if (k > KK + 1)
                        dv[k] = (uk[k] - ukp[k]) * DYK * DZK[k] + (vk[k] - vkp[k]) * DXK * DZK[k] + (wk[k] - wk[k + 1]) * DXK * DYK;
                    else
                        dv[k] = (uk[k] - ukp[k]) * DYK * DZK[k] + (vk[k] - vkp[k]) * DXK * DZK[k] - wk[k + 1] * DXK * DYK;
                }
            }
        /* pressure correction, :84-341: four sweep orders by IterationLoops % 4.
         * Every column is solved over its full height, so the scratch above the
         * top level (PIMP/QIMP[NKK]) is never written: zeros. */
        int IS = 1, JS = 1;
        switch (it % 4) {
            case 0: IS = 1; JS = 1; break;
            case 1: IS = 1; JS = -1; break;
            case 2: IS = -1; JS = -1; break;
            default: IS = -1; JS = 1; break;
        }
        for (int j1 = 2; j1 < NJJ; ++j1) {
            int j = (JS == -1) ? NJJ - j1 + 1 : j1;
            for (int i1 = 2; i1 < NII; ++i1) {
                int i = (IS == -1) ? NII - i1 + 1 : i1;
                int KK = KKART[D2(i, j)];
                const float *dv = &DIV[D3(i, j, 0)];
                float *dpm = &DPM[D3(i, j, 0)];
                const float *dim_ = &DPM[D3(i - 1, j, 0)], *dip = &DPM[D3(i + 1, j, 0)];
                const float *djm = &DPM[D3(i, j - 1, 0)], *djp = &DPM[D3(i, j + 1, 0)];
                for (int k = NKK - 1; k >= 1; --k) {
                    if (!(KK < k)) continue;
                    float DIM = dv[k] / DTIME + (dim_[k] + dip[k]) / DXK * DYK * DZK[k] + (djm[k] - djp[k]) / DYK * DXK * DZK[k];
                    float APP = 2 * (DYK * DZK[k] / DXK + DXK * DZK[k] / DYK + DXK * DYK / DZK[k]);
                    float ABP = (k > KK + 1) ? DXK * DYK / (0.5f * (DZK[k - 1] + DZK[k])) : DXK * DYK / DZK[k];
                    float ATP = (DXK * DYK) / (0.5f * (DZK[k + 1] + DZK[k]));
                    float TERMP = 1 / (APP - ATP * PIMP[k + 1]);
                    PIMP[k] = ABP * TERMP;
                    QIMP[k] = (ATP * QIMP[k + 1] + DIM) * TERMP;
                }
                dpm[0] = 0;
                for (int k = 1; k <= NKK - 1; ++k)
                    dpm[k] = (KK < k) ? PIMP[k] * dpm[k - 1] + QIMP[k] : 0.0f;
            }
        }
        /* velocity correction, :345-410 (no quantisation here) */
        for (int i = 2; i < NII; ++i)
            for (int j = 2; j < NJJ; ++j) {
                int KK = KKART[D2(i, j)], KKim = KKART[D2(i - 1, j)], KKjm = KKART[D2(i, j - 1)];
                const float *dpm = &DPM[D3(i, j, 0)], *dpmi = &DPM[D3(i - 1, j, 0)], *dpmj = &DPM[D3(i, j - 1, 0)];
                float *uk = &UK[D3(i, j, 0)], *vk = &VK[D3(i, j, 0)], *wk = &WK[D3(i, j, 0)];
                for (int k = 1; k < NKK; ++k) {
                    float DDPX = dpmi[k] - dpm[k];
                    float DDPY = dpmj[k] - dpm[k];
                    float DDPZ = (k > KK + 1) ? dpm[k - 1] - dpm[k] : 0.0f;
                    if (KK < k && KKim < k) uk[k] += DDPX / DXK * DTIME; else uk[k] = 0;
                    if (KK < k && KKjm < k) vk[k] += DDPY / DYK * DTIME; else vk[k] = 0;
                    if (KK + 1 < k) wk[k] += DDPZ * 2 / (DZK[k] + DZK[k - 1]) * DTIME; else wk[k] = 0;
                }
            }
    }
    free(PIMP); free(QIMP);
    /* vertical velocity from the mass balance of each cell, MicroscaleFlat.cs:388-462 */
    for (int i = 2; i < NII; ++i)
        for (int j = 2; j < NJJ; ++j) {
            if (!p->terrain && !(p->flow_field_level <= 1 || ADVDOM[D2(i, j)] == 0)) continue;
            const float DSN = p->terrain ? DYK : DXK;   /* MicroscaleTerrain uses DYK on the j faces */
            const float *uk = &UK[D3(i, j, 0)], *uki = &UK[D3(i + 1, j, 0)];
            const float *vk = &VK[D3(i, j, 0)], *vkj = &VK[D3(i, j + 1, 0)];
            float *wk = &WK[D3(i, j, 0)];
            int KK = KKART[D2(i, j)];
            for (int k = 1; k < NKK; ++k) {
                if (!(k > KK)) continue;
                float fwo1 = (k <= KKART[D2(i - 1, j)]) ? 0.0f : DZK[k] * DYK * uk[k];
                float fwo2 = (k <= KKART[D2(i + 1, j)]) ? 0.0f : DZK[k] * DYK * uki[k];
                float fsn1 = (k <= KKART[D2(i, j - 1)]) ? 0.0f : DZK[k] * DSN * vk[k];
                float fsn2 = (k <= KKART[D2(i, j + 1)]) ? 0.0f : DZK[k] * DSN * vkj[k];
                float fbt1 = ((k <= KK + 1) || (k == 1)) ? 0.0f : DXK * DYK * wk[k];
                wk[k + 1] = (fwo1 - fwo2 + fsn1 - fsn2 + fbt1) / (DXK * DYK);
            }
        }
#undef D2
#undef D3
    return 0;
}
