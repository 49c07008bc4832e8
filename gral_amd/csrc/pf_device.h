// Device-side state and arithmetic helpers of the prognostic flow solver.
//
// Layout in HBM: every 3-D field is [nii+2][njj+2][nkp] float, k contiguous,
// with nkp >= KADVMAX+2 (the solver never touches a level above KADVMAX+1, so
// only those levels live on the device).  Per-column data are [nii+2][njj+2].
//
// Arithmetic follows the reference's C# semantics: FP32 throughout, no FMA
// contraction (built with -ffp-contract=off), correctly rounded division and
// square root, FP64 where the C# expression promotes to double.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct Dom {
    int nii, njj, nkk, nkp, nj2, kadvmax;
    int model, simd;
    float dxk, dyk, area, dtime, relax, relaxp, cor, ug, vg, vishmin, visvmin, bz0, ufx, potdtdz;
    double frund;
    // per level
    const float *dzk, *hok;          // [nkk+2]
    const float *ap0;                // [nkp]  AP0[k] = DXK*DYK*DZK[k]/DTIME
    const float *app, *abp, *atp;    // [nkp]  pressure TDMA coefficients (column independent)
    // per column
    const uint8_t *adv, *vegp;
    const int *kk, *vi;
    const float *cutk, *cov, *ht, *ho, *z0s;
    const double *tke_a, *tke_den;   // TKE terrain wall function: 0.4/ln(.) and TDISS divisor
    const int *tke_br;               // 1: Z0 >= DZK*0.5 branch, 0: otherwise
    // fields
    float *uk, *vk, *wk, *uks, *vks, *wks, *dpm, *dpmnew, *div, *turb, *tdiss;
    const float *veg;
    // scratch
    float *prow_p, *prow_q;          // pressure PIMP/QIMP per sweep row [nj2][nkp]
    unsigned *resid;                 // DeltaUMax per iteration (float bits, values >= 0)
    int resid_cap;                   // entries in resid
    // tiled sweeps
    int tiled;                       // 1: tiled wavefront kernels, 0: one launch per anti-diagonal
    float *planes;                   // [16][nii+2][njj+2][nkp] mixing-length coefficient planes
    const float *ptab[2], *ttab[2];  // pressure PIMP/TERMP per column for IS=+1 / IS=-1
    const int4 *meta;                // per column: KK, VI, KSTART | adv<<16 | vegp<<17, KKm_i | KKm_j<<16
};

__device__ __forceinline__ size_t c2(const Dom &d, int i, int j) { return (size_t)i * (size_t)d.nj2 + (size_t)j; }
__device__ __forceinline__ size_t c3(const Dom &d, int i, int j) { return c2(d, i, j) * (size_t)d.nkp; }

// Vector.Max / Vector.Min / FloatMax of the reference: SSE maxps/minps semantics.
__device__ __forceinline__ float vmaxf(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float vminf(float a, float b) { return a < b ? a : b; }
// Math.Max/Min(float) of .NET Core (IEEE 754:2019 maximum/minimum, NaN propagating).
__device__ __forceinline__ float mmaxf(float a, float b) {
    if (a != b) { if (!__builtin_isnan(a)) return b < a ? a : b; return a; }
    return __builtin_signbit(b) ? a : b;
}
__device__ __forceinline__ float mminf(float a, float b) {
    if (a != b) { if (!__builtin_isnan(a)) return a < b ? a : b; return a; }
    return __builtin_signbit(a) ? a : b;
}
__device__ __forceinline__ double mmaxd(double a, double b) {
    if (a != b) { if (!__builtin_isnan(a)) return b < a ? a : b; return a; }
    return __builtin_signbit(b) ? a : b;
}
// ConvToInt(double): round half away from zero by truncation of x +- 0.5.
__device__ __forceinline__ int conv_to_int(double x) { return x < 0 ? (int)(x - 0.5) : (int)(x + 0.5); }
// (float)(ConvToInt(x * Frund) / Frund): the FP64 velocity quantisation.
__device__ __forceinline__ float quantise(float x, double frund) {
    return (float)((double)conv_to_int((double)x * frund) / frund);
}
