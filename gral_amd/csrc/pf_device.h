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

// Pointer members are declared through a policy so that the same layout
// serves the host (generic pointers) and the kernels (global-memory pointers,
// which compile to global_* instead of flat_* memory instructions).
struct GenericPtr { template <class T> using ptr = T *; };
struct GlobalPtr { template <class T> using ptr = T __attribute__((address_space(1))) *; };

template <class PP>
struct DomT {
    template <class T> using P = typename PP::template ptr<T>;
    int nii, njj, nkk, nkp, nj2, kadvmax;
    int model, simd;
    float dxk, dyk, area, dtime, relax, relaxp, cor, ug, vg, vishmin, visvmin, bz0, ufx, potdtdz;
    double frund;
    // per level
    P<const float> dzk, hok;         // [nkk+2]
    P<const float> ap0;              // [nkp]  AP0[k] = DXK*DYK*DZK[k]/DTIME
    P<const float> app, abp, atp;    // [nkp]  pressure TDMA coefficients (column independent)
    // per column
    P<const uint8_t> adv, vegp;
    P<const int> kk, vi;
    P<const float> cutk, cov, ht, ho, z0s;
    P<const double> tke_a, tke_den;  // TKE terrain wall function: 0.4/ln(.) and TDISS divisor
    P<const int> tke_br;             // 1: Z0 >= DZK*0.5 branch, 0: otherwise
    // fields
    P<float> uk, vk, wk, uks, vks, wks, dpm, dpmnew, div, turb, tdiss;   // div holds DIV / DTIME
    P<const float> veg;
    // scratch
    P<float> prow_p, prow_q;         // pressure PIMP/QIMP per sweep row [nj2][nkp]
    P<unsigned> resid;               // DeltaUMax per iteration (float bits, values >= 0)
    int resid_cap;                   // entries in resid
    // sweep kernels
    int tiled;                       // 2: strip-pipelined sweeps, 0: one launch per anti-diagonal
    P<int> sync;                     // strip sweeps: [64] header (sticky error flag at 0), then 4 blocks
    int sync_stride;                 //   of sync_stride ints (ticket, unused, progress per strip)
    P<float> planes;                 // [cells][12 (U, V) or 8 (W)] mixing-length coefficient records
    P<float> rec3;                   // [cells][3]  C, TERMP and P of the same cells
    P<float> wcol;                   // [columns] wall term of the U / V sweep at KKART+1
    P<const float> ptab[2], ttab[2]; // pressure PIMP/TERMP per column for IS=+1 / IS=-1
    P<const int4> meta;              // per column: KK, VI, KSTART | adv<<16 | vegp<<17, KKm_i | KKm_j<<16
    int noquant;                     // 1: the velocity correction skips the FP64 quantisation (adjoint forward)
    int iter_off;                    // this situation's iteration = the launch's iteration - iter_off (a
                                     // multiple of 100: situations admitted to a running batch later)
};
using Dom = DomT<GenericPtr>;        // host side, and the descriptor array handed to the kernels
using DomG = DomT<GlobalPtr>;        // the kernels' view of the same bytes
static_assert(sizeof(Dom) == sizeof(DomG), "descriptor layouts must agree");
typedef float __attribute__((address_space(1))) gf32;
typedef int __attribute__((address_space(1))) gi32;
typedef float v4f __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef v4f __attribute__((address_space(1))) gv4f;
typedef v4i __attribute__((address_space(1))) gv4i;

// 16-byte loads and stores through global-memory pointers (p 16-byte aligned)
__device__ __forceinline__ float4 ldg4(const gf32 *p) {
    v4f v = *(const gv4f *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void stg4(gf32 *p, float a, float b, float c, float e) {
    v4f v = {a, b, c, e};
    *(gv4f *)p = v;
}
typedef float v2f __attribute__((ext_vector_type(2)));
typedef v2f __attribute__((address_space(1))) gv2f;
__device__ __forceinline__ float2 ldg2(const gf32 *p) {     // p 8-byte aligned
    v2f v = *(const gv2f *)p;
    return make_float2(v.x, v.y);
}
// 12-byte records (p 4-byte aligned): one dwordx3 access
typedef float v3fa __attribute__((ext_vector_type(3), aligned(4)));
typedef v3fa __attribute__((address_space(1))) gv3f;
__device__ __forceinline__ float3 ldg3(const gf32 *p) {
    v3fa v = *(const gv3f *)p;
    return make_float3(v.x, v.y, v.z);
}
__device__ __forceinline__ void stg3(gf32 *p, float a, float b, float c) {
    v3fa v = {a, b, c};
    *(gv3f *)p = v;
}
template <class T>
__device__ __forceinline__ int4 ldg4i(const T *p) {
    v4i v = *(const gv4i *)p;
    return make_int4(v.x, v.y, v.z, v.w);
}
// The same through the constant address space: a scalar load (counted apart
// from the vector memory operations) where the address is uniform.  Only for
// data no kernel writes while it runs.
typedef v4i __attribute__((address_space(4))) cv4i;
template <class T>
__device__ __forceinline__ int4 ldc4i(const T *p) {
    v4i v = *(const cv4i *)p;
    return make_int4(v.x, v.y, v.z, v.w);
}

// Flow-field file payload packer (GFFWrite.cs): fields u, v, w, turb, levels
// below nlo from the solver's device fields (stride nkp), the rest from a copy
// of the host's upper levels (stride nup, level nlo first).
struct GffArgs {
    const float *lo[4];
    const float *up[4];
    int nlo, nkp, nup;
    int nii, njj, nkk, nj2;
    int kind;                         // 0, 1, 2: *.gff writing modes; 3: *.tke
    const int16_t *kk;                // KKART [nii+2][njj+2]
    const long long *off;             // kind 2: int16 offset of each column record after the KKART block
    int16_t *out;                     // the payload after its 32-byte header
};

// Wind at particle positions on the solver's fields (Intwind.cs:31-170): one
// thread per particle; levels below nlo from the device fields (stride nkp),
// the ones above from the device copy of the host's upper levels (stride nup).
struct WindArgs {
    const float *lo[3];               // UK, VK, WK
    const float *up[3];
    int nlo, nkp, nup, nj2, nkk;
    int topo;                         // 1: terrain, 0: flat terrain with buildings
    float dxk, dyk, ahmin;
    int ikoo, jkoo;                   // IKOOAGRAL, JKOOAGRAL
    const float *hok, *dzk;           // HOKART, DZK [nkk+2]
    int n;
    const int32_t *ci, *cj;           // flow-field cell of each particle (FFCellX, FFCellY)
    const float *ahint, *z;
    const double *x, *y;
    float *ux, *uy, *uz;
    int32_t *ck;                      // IndexK
};

// Diagnostic post-pass (one situation per launch, full-height columns)
struct DiagArgs {
    int nii, njj, nkk, nkh, nj2;   // nkh = NKK + 2 levels per stored column
    int level, terrain;            // FlowFieldLevel (1 diagnostic, 2 prognostic); caller
    float dxk, dyk, dtime;
    const float *dzk;              // [nkh]
    const float *atp;              // [nkh] ATP by level
    const float *term, *p_in, *p_bot;  // [nkh] TERMP; PIMP above, and at, a column's bottom level
    const int *kk;                 // [nii+2][njj+2]
    const uint8_t *adv;
    float *uk, *vk, *wk, *dpm, *div;   // [nii+2][njj+2][nkh]
};

// Buffers of one step of the adjoint (pf_adjoint.hip), each [nii+2][njj+2][nkp]
// except the per-column source table.
struct AdjArgs {
    float *ub, *vb, *wb;        // adjoint of UK, VK, WK (in: of the step's output; out: of its input)
    float *pb;                  // adjoint of DPM
    float *qb;                  // adjoint of the elimination values Q read as stale tops by later columns
    float *mb;                  // adjoint of DIM per column and level
    float *db;                  // adjoint of DIV
    const float *ptab, *ttab;   // P and TERMP of every column for this step's IS
    const int *src;             // per column: i of the column whose Q the elimination starts from, or -1
};

// Buffers of the adjoint of one momentum sweep (pf_kernels.hip), each
// [nii+2][njj+2][nkp].
struct MomAdj {
    const float *xa;    // X after the sweep (the device field)
    const float *gin;   // dL/dX after
    float *gout;        // dL/dX before: the swept columns' own levels KSTART-1..VI+1
    float *gpa, *gpb;   // per column and level: adjoint for the new X of its a- / b-predecessor
    float *gna, *gnb;   // per column and level: adjoint for the old X of its a- / b-successor
    float *gres;        // dL/dX before, complete
};

#define SYNC_BLOCK0 64               // Dom::sync: sticky error flag at 0, sweep blocks from here
#define STRIP_ROWS 15                // sweep rows per strip workgroup (plus one chain wave)

__device__ __forceinline__ size_t c2(const DomG &d, int i, int j) { return (size_t)i * (size_t)d.nj2 + (size_t)j; }
__device__ __forceinline__ size_t c3(const DomG &d, int i, int j) { return c2(d, i, j) * (size_t)d.nkp; }

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
