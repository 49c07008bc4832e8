// Set-up of one weather situation on flat terrain, before the prognostic call:
//   * prognostic sub-domains around the buildings (ReadGeometryAndDomains.cs:321-349),
//   * the reference point of the sub-domains (MicroscaleTerrainSearchRefPoint.cs:27-103,
//     MicroscaleTerrain.cs:936-963),
//   * the initial wind field, KKART, AHK, BUI_HEIGHT and KADVMAX
//     (MicroscaleFlat.cs:29-197) and the damping of the initial wind near
//     obstacle walls (:199-345).
//
// The transcendental parts depend on one level and one Obukhov length only
// (the power-law profile) or on one neighbour offset only (the wall damping):
// the host evaluates them into small tables with the C library's powf, logf
// and sqrtf, as .NET's MathF does on Linux; the kernels do the per-cell work:
// masks, table look-ups, the neighbour searches and the reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gral_pf.h"

namespace {

thread_local std::string g_err;

int setfail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

inline float net_maxf(float a, float b) {  // Math.Max(float, float)
    if (a != b) { if (!std::isnan(a)) return b < a ? a : b; return a; }
    return std::signbit(b) ? a : b;
}

// device scratch of one call, released on every exit path
struct Scratch {
    std::vector<void *> p;
    hipStream_t s = nullptr;
    ~Scratch() {
        if (s) (void)hipStreamSynchronize(s);
        for (void *q : p) (void)hipFree(q);
        if (s) (void)hipStreamDestroy(s);
    }
    template <class T>
    T *alloc(size_t n) {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
        p.push_back(q);
        return static_cast<T *>(q);
    }
};

}  // namespace

#define SETCHK(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) return setfail(GRAL_PF_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// Sub-domains: every column with CUTK > 0 marks the rectangle of
// (int)(factor*CUTK/DXK) columns in i and (int)(factor*CUTK/DYK) in j around
// it (clipped to 1..NII x 1..NJJ).  Two passes give the same union without
// writing every rectangle: reach[i][j1] = the largest i-reach among the
// columns of row i whose j-range covers j1, then a forward and a backward
// scan along i marks the rows within reach.
// ---------------------------------------------------------------------------
__global__ void k_sd_reach(const float *cutk, int nii, int njj, float factor, float dxk, float dyk, int *reach) {
    const int nj2 = njj + 2;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nii * njj) return;
    const int i = 1 + (int)(t / njj), j = 1 + (int)(t % njj);
    const float c = cutk[(size_t)i * nj2 + j];
    if (!(c > 0)) return;
    const int ip = (int)(factor * c / dxk), jp = (int)(factor * c / dyk);
    const int j0 = max(1, j - jp), j1 = min(njj, j + jp);
    for (int q = j0; q <= j1; ++q) atomicMax(&reach[(size_t)i * nj2 + q], ip);
}

__global__ void k_sd_mark(const int *reach, int nii, int njj, uint8_t *adv) {
    const int nj2 = njj + 2;
    const int j = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j > njj) return;
    int r = -1;
    for (int i = 1; i <= nii; ++i) {              // a column i' <= i reaches i
        r = max(r - 1, reach[(size_t)i * nj2 + j]);
        if (r >= 0) adv[(size_t)i * nj2 + j] = 1;
    }
    r = -1;
    for (int i = nii; i >= 1; --i) {              // a column i' >= i reaches i
        r = max(r - 1, reach[(size_t)i * nj2 + j]);
        if (r >= 0) adv[(size_t)i * nj2 + j] = 1;
    }
}

extern "C" int gral_pf_subdomains(int32_t nii, int32_t njj, float dxk, float dyk, int32_t factor, const float *cutk,
                                  uint8_t *advdom) {
    if (nii < 1 || njj < 1 || !(dxk > 0) || !(dyk > 0) || factor < 0 || !cutk || !advdom)
        return setfail(GRAL_PF_EINVAL, "subdomains: bad argument");
    const size_t n2 = (size_t)(nii + 2) * (njj + 2);
    Scratch S;
    SETCHK(hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking));
    float *dc = S.alloc<float>(n2);
    int *dr = S.alloc<int>(n2);
    uint8_t *da = S.alloc<uint8_t>(n2);
    if (!dc || !dr || !da) return setfail(GRAL_PF_ENOMEM, "subdomains: hipMalloc failed");
    SETCHK(hipMemcpyAsync(dc, cutk, n2 * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(da, advdom, n2, hipMemcpyHostToDevice, S.s));
    std::vector<int> neg(n2, -1);
    SETCHK(hipMemcpyAsync(dr, neg.data(), n2 * sizeof(int), hipMemcpyHostToDevice, S.s));
    const size_t cols = (size_t)nii * njj;
    hipLaunchKernelGGL(k_sd_reach, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, S.s, dc, nii, njj, (float)factor,
                       dxk, dyk, dr);
    hipLaunchKernelGGL(k_sd_mark, dim3((unsigned)((njj + 63) / 64)), dim3(64), 0, S.s, dr, nii, njj, da);
    SETCHK(hipGetLastError());
    SETCHK(hipMemcpyAsync(advdom, da, n2, hipMemcpyDeviceToHost, S.s));
    SETCHK(hipStreamSynchronize(S.s));
    return GRAL_PF_OK;
}

// Spiral search from the border towards the centre for the first sub-domain
// cell (MicroscaleTerrainSearchRefPoint.cs:27-103); (1, 1) when there is none,
// as GetSubDomainPoint (MicroscaleTerrain.cs:936-963) sets it, and the
// return code then says so.
extern "C" int gral_pf_reference_point(int32_t nii, int32_t njj, const uint8_t *advdom, int32_t *ref_i, int32_t *ref_j) {
    if (nii < 1 || njj < 1 || !advdom || !ref_i || !ref_j) return setfail(GRAL_PF_EINVAL, "reference_point: bad argument");
    const int xmax = nii + 2, ymax = njj + 2, searchmax = std::min(xmax, ymax) / 2;
    auto A = [&](int i, int j) { return advdom[(size_t)i * ymax + j]; };
    int delta = 1, i = 1, j = 1;
    bool found = false;
    while (!found && delta < searchmax) {
        i = delta;
        j = delta;
        while (j < ymax - delta && !found) { if (A(i, j) == 1) found = true; else ++j; }
        while (i < xmax - delta && !found) { if (A(i, j) == 1) found = true; else ++i; }
        while (j > delta && !found) { if (A(i, j) == 1) found = true; else --j; }
        while (i > delta && !found) { if (A(i, j) == 1) found = true; else --i; }
        ++delta;
    }
    if (!found) {
        *ref_i = 1;
        *ref_j = 1;
        return setfail(GRAL_PF_ESTATE, "reference_point: no prognostic sub-domain");
    }
    *ref_i = i;
    *ref_j = j;
    return GRAL_PF_OK;
}

// ---------------------------------------------------------------------------
// Initial field.  Per column (i, j) of 1..NII+1 x 1..NJJ+1, levels 1..NKK.
// A level is blocked when its centre vertk = HOKART[k-1] + DZK[k]/2 is at or
// below CUTK.  The reference zeroes the levels first, then for every inner
// column sets u and v from the profile above the obstacle, zeroes the u of
// the next column in i (v of the next in j) inside it, and copies the last
// inner row and column into the ghost row NII+1 (u) and column NJJ+1 (v).
// Independent of the order the columns are visited in, this leaves:
//   u[i][j] = profile when neither (i, j) nor (i-1, j) is blocked, else 0
//   v[i][j] = profile when neither (i, j) nor (i, j-1) is blocked, else 0
//   u[NII+1][j] = u[NII][j], v[i][NJJ+1] = v[i][NJJ], everything else 0.
// ---------------------------------------------------------------------------
struct FlatArgs {
    int nii, njj, nkk, nkh, nj2;
    const float *vertk;                  // [nkh]
    const float *hok;                    // [nkh]
    const float *cutk;                   // [n2]
    const int *cls;                      // [n2] profile class of each column, or null (class 0)
    const float *ux, *uy;                // [classes][nkh] profile values per level
    float *uk, *vk, *wk;                 // [n2][nkh]
    int16_t *kkart;
    float *ahk, *bui;
    int *kad;                            // max KKART over columns with CUTK > 0
};

__device__ __forceinline__ float fl_u(const FlatArgs &a, int i, int j, int k) {
    const size_t c = (size_t)i * a.nj2 + j;
    const float v = a.vertk[k];
    if (v <= a.cutk[c]) return 0.0f;
    if (i > 1 && v <= a.cutk[c - a.nj2]) return 0.0f;
    return a.ux[(size_t)(a.cls ? a.cls[c] : 0) * a.nkh + k];
}
__device__ __forceinline__ float fl_v(const FlatArgs &a, int i, int j, int k) {
    const size_t c = (size_t)i * a.nj2 + j;
    const float v = a.vertk[k];
    if (v <= a.cutk[c]) return 0.0f;
    if (j > 1 && v <= a.cutk[c - 1]) return 0.0f;
    return a.uy[(size_t)(a.cls ? a.cls[c] : 0) * a.nkh + k];
}

// one wave per column
__global__ void __launch_bounds__(64) k_flat_column(FlatArgs a) {
    const int ncj = a.njj + 1;
    const int i = 1 + (int)(blockIdx.x / ncj), j = 1 + (int)(blockIdx.x % ncj);
    const size_t c = (size_t)i * a.nj2 + j, o = c * a.nkh;
    const bool inner_i = i <= a.nii, inner_j = j <= a.njj;
    int kb = 0;                          // highest blocked level
    float ah = 0.0f;
    for (int k = 1 + threadIdx.x; k <= a.nkk; k += 64) {
        float u = 0.0f, v = 0.0f;
        if (inner_i && inner_j) {
            u = fl_u(a, i, j, k);
            v = fl_v(a, i, j, k);
            if (a.vertk[k] <= a.cutk[c]) {
                kb = k;
                ah = fmaxf(ah, a.hok[k]);
            }
        } else if (inner_j) {            // i = NII+1: copy of row NII
            u = fl_u(a, a.nii, j, k);
        } else if (inner_i) {            // j = NJJ+1: copy of column NJJ
            v = fl_v(a, i, a.njj, k);
        }
        a.uk[o + k] = u;
        a.vk[o + k] = v;
        a.wk[o + k] = 0.0f;
    }
    if (!(inner_i && inner_j)) return;
    for (int m = 32; m >= 1; m >>= 1) {
        kb = max(kb, __shfl_xor(kb, m));
        ah = fmaxf(ah, __shfl_xor(ah, m));
    }
    if (threadIdx.x == 0) {
        a.kkart[c] = (int16_t)kb;
        a.ahk[c] = ah;
        if (kb > 0) {
            a.bui[c] = ah;
            if (a.cutk[c] > 0) atomicMax(a.kad, kb);
        }
    }
}

// Damping near walls (MicroscaleFlat.cs:201-345): for each free level above
// KKART of a column in IGEB+1..NII-IGEB x IGEB+1..NJJ-IGEB, the first
// obstacle (CUTK > 1 and vertk <= AHK) within IGEB cells in each of the eight
// directions W, E, N, S, SE, NW, NE, SW multiplies u and v by the factor of
// its distance, fac[dir class][offset] (0: along i, 1: along j, 2: diagonal;
// 1.0 where the distance exceeds 20 m).
struct DampArgs {
    int nii, njj, nkk, nkh, nj2, igeb;
    int mode;                            // 1: columns outside the sub-domains, 2: all
    const float *vertk;
    const float *cutk, *ahk;
    const int16_t *kkart;
    const uint8_t *adv;
    const float *fac;                    // [3][igeb+1]
    float *uk, *vk;
    float *wk;                           // also damped when not null (terrain)
};

__device__ __forceinline__ bool dm_hit(const DampArgs &a, int i, int j, float v) {
    const size_t c = (size_t)i * a.nj2 + j;
    return a.cutk[c] > 1.0f && v <= a.ahk[c];
}

__global__ void k_flat_damp(DampArgs a) {
    const int w = a.nii - 2 * a.igeb, hgt = a.njj - 2 * a.igeb;
    const int col = blockIdx.x;
    if (col >= w * hgt) return;
    const int i = 1 + a.igeb + col / hgt, j = 1 + a.igeb + col % hgt;
    const size_t c = (size_t)i * a.nj2 + j;
    if (a.mode == 1 && a.adv[c] != 0) return;
    const int kk = a.kkart[c];
    const int G = a.igeb;
    for (int k = max(1, kk + 1) + threadIdx.x; k <= a.nkk - 1; k += blockDim.x) {
        const float v = a.vertk[k];
        float ab = 1.0f;
        int g;
        for (g = 1; g <= G && !dm_hit(a, i - g, j, v); ++g) {}          // west
        if (g <= G) ab *= a.fac[0 * (G + 1) + g];
        for (g = 1; g <= G && !dm_hit(a, i + g, j, v); ++g) {}          // east
        if (g <= G) ab *= a.fac[0 * (G + 1) + g];
        for (g = 1; g <= G && !dm_hit(a, i, j + g, v); ++g) {}          // north
        if (g <= G) ab *= a.fac[1 * (G + 1) + g];
        for (g = 1; g <= G && !dm_hit(a, i, j - g, v); ++g) {}          // south
        if (g <= G) ab *= a.fac[1 * (G + 1) + g];
        for (g = 1; g <= G && !dm_hit(a, i + g, j + g, v); ++g) {}      // ig = i+g, jg = j+ig-i
        if (g <= G) ab *= a.fac[2 * (G + 1) + g];
        for (g = 1; g <= G && !dm_hit(a, i - g, j - g, v); ++g) {}      // ig = i-g, jg = j+ig-i
        if (g <= G) ab *= a.fac[2 * (G + 1) + g];
        for (g = 1; g <= G && !dm_hit(a, i + g, j - g, v); ++g) {}      // ig = i+g, jg = j+i-ig
        if (g <= G) ab *= a.fac[2 * (G + 1) + g];
        for (g = 1; g <= G && !dm_hit(a, i - g, j + g, v); ++g) {}      // ig = i-g, jg = j+i-ig
        if (g <= G) ab *= a.fac[2 * (G + 1) + g];
        const size_t o = c * a.nkh + k;
        a.uk[o] *= ab;
        a.vk[o] *= ab;
        if (a.wk) a.wk[o] *= ab;
    }
}

// Profile value of one level (MicroscaleFlat.cs:84-138): power law below the
// first and (with one profile level) above the last measurement height,
// linear interpolation between measurements, the top value above them.
static void profile_level(const gral_pf_flat_met *m, float vertk, float exponent, float &ux, float &uy) {
    const int n = m->n_profile;
    const float *H = m->meas_height, *U = m->obs_u, *V = m->obs_v;
    if (vertk <= H[1]) {
        const float f = powf(vertk / H[1], exponent);
        ux = U[1] * f;
        uy = V[1] * f;
    } else if (vertk > H[n]) {
        if (n == 1) {
            const float f = powf(vertk / H[n], exponent);
            ux = U[n] * f;
            uy = V[n] * f;
        } else {
            ux = U[n];
            uy = V[n];
        }
    } else {
        int ipo = 1;
        for (int q = 1; q <= n; ++q)
            if (vertk > H[q]) ipo = q + 1;
        ux = U[ipo - 1] + (U[ipo] - U[ipo - 1]) / (H[ipo] - H[ipo - 1]) * (vertk - H[ipo - 1]);
        uy = V[ipo - 1] + (V[ipo] - V[ipo - 1]) / (H[ipo] - H[ipo - 1]) * (vertk - H[ipo - 1]);
    }
}

static float profile_exponent(float ob) {
    if (ob <= 0) return net_maxf(0.35f - 0.4f * powf(std::fabs(ob), -0.15f), 0.05f);
    return 0.56f * powf(ob, -0.15f);
}

// wall damping factors 0.19*ln((entf+0.5)*10) per direction class (along i,
// along j, diagonal) and offset; 1 where the distance exceeds 20 m
static std::vector<float> damping_table(float dxk, float dyk, int igeb) {
    std::vector<float> fac(3 * (size_t)(igeb + 1), 1.0f);
    for (int o = 1; o <= igeb; ++o) {
        const float dx = (float)o * dxk, dy = (float)o * dyk;
        const float e3[3] = {std::fabs(dx), std::fabs(dy), std::sqrt(dx * dx + dy * dy)};
        for (int q = 0; q < 3; ++q)
            fac[q * (igeb + 1) + o] = (e3[q] <= 20.0f) ? 0.19f * logf((e3[q] + 0.5f) * 10.0f) : 1.0f;
    }
    return fac;
}

extern "C" int gral_pf_flat_init(const gral_pf_grid *g, const gral_pf_flat_met *m, const float *cutk,
                                 const uint8_t *advdom, float *uk, float *vk, float *wk, int16_t *kkart, float *ahk,
                                 float *bui_height, int32_t *kadvmax) {
    if (!g || !m || !cutk || !uk || !vk || !wk || !kkart || !ahk || !kadvmax || !g->dzk || !g->hokart)
        return setfail(GRAL_PF_EINVAL, "flat_init: null argument");
    if (g->nii < 1 || g->njj < 1 || g->nkk < 2 || g->vertcellsff < 0 || !(g->dxk > 0) || !(g->dyk > 0))
        return setfail(GRAL_PF_EINVAL, "flat_init: bad grid");
    if (m->n_profile < 1 || m->n_profile > 99 || !m->meas_height || !m->obs_u || !m->obs_v)
        return setfail(GRAL_PF_EINVAL, "flat_init: bad meteorological profile");
    if (m->near_wall < 0 || m->near_wall > 2 || (m->near_wall == 1 && !advdom))
        return setfail(GRAL_PF_EINVAL, "flat_init: near_wall must be 0, 1 (needs advdom) or 2");
    const int nii = g->nii, njj = g->njj, nkk = g->nkk, nkh = nkk + 2, nj2 = njj + 2;
    const size_t n2 = (size_t)(nii + 2) * nj2, n3 = n2 * nkh;
    // levels: vertk = HOKART[k-1] + DZK[k]*0.5
    std::vector<float> vertk(nkh, 0.0f);
    for (int k = 1; k <= nkk; ++k) vertk[k] = g->hokart[k - 1] + g->dzk[k] * 0.5f;
    // profile classes: one per distinct Obukhov length of the inner columns
    std::vector<float> obs;
    std::vector<int> cls;
    if (m->ol_column) {
        std::vector<uint32_t> bits;
        bits.reserve((size_t)nii * njj);
        for (int i = 1; i <= nii; ++i)
            for (int j = 1; j <= njj; ++j) {
                uint32_t b;
                std::memcpy(&b, &m->ol_column[(size_t)i * nj2 + j], 4);
                bits.push_back(b);
            }
        std::sort(bits.begin(), bits.end());
        bits.erase(std::unique(bits.begin(), bits.end()), bits.end());
        obs.resize(bits.size());
        for (size_t q = 0; q < bits.size(); ++q) std::memcpy(&obs[q], &bits[q], 4);
        cls.assign(n2, 0);
        for (int i = 1; i <= nii; ++i)
            for (int j = 1; j <= njj; ++j) {
                uint32_t b;
                std::memcpy(&b, &m->ol_column[(size_t)i * nj2 + j], 4);
                cls[(size_t)i * nj2 + j] = (int)(std::lower_bound(bits.begin(), bits.end(), b) - bits.begin());
            }
    } else {
        obs.push_back(m->ob_length);
    }
    const size_t ncls = obs.size();
    std::vector<float> ux(ncls * nkh, 0.0f), uy(ncls * nkh, 0.0f);
    for (size_t q = 0; q < ncls; ++q) {
        const float e = profile_exponent(obs[q]);
        for (int k = 1; k <= nkk; ++k) profile_level(m, vertk[k], e, ux[q * nkh + k], uy[q * nkh + k]);
    }
    const int igeb = std::max((int)(20.0f / g->dxk), 1);
    const std::vector<float> fac = damping_table(g->dxk, g->dyk, igeb);
    // device work
    Scratch S;
    SETCHK(hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking));
    float *d_vk = S.alloc<float>(nkh), *d_hok = S.alloc<float>(nkh), *d_cut = S.alloc<float>(n2);
    float *d_ux = S.alloc<float>(ux.size()), *d_uy = S.alloc<float>(uy.size()), *d_fac = S.alloc<float>(fac.size());
    float *d_u = S.alloc<float>(n3), *d_v = S.alloc<float>(n3), *d_w = S.alloc<float>(n3);
    float *d_ahk = S.alloc<float>(n2), *d_bui = S.alloc<float>(n2);
    int16_t *d_kk = S.alloc<int16_t>(n2);
    uint8_t *d_adv = S.alloc<uint8_t>(n2);
    int *d_cls = cls.empty() ? nullptr : S.alloc<int>(n2);
    int *d_kad = S.alloc<int>(1);
    if (!d_vk || !d_hok || !d_cut || !d_ux || !d_uy || !d_fac || !d_u || !d_v || !d_w || !d_ahk || !d_bui || !d_kk ||
        !d_adv || !d_kad || (!cls.empty() && !d_cls))
        return setfail(GRAL_PF_ENOMEM, "flat_init: hipMalloc failed");
    const int one = 1;
    SETCHK(hipMemcpyAsync(d_vk, vertk.data(), nkh * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_hok, g->hokart, nkh * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_cut, cutk, n2 * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_ux, ux.data(), ux.size() * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_uy, uy.data(), uy.size() * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_fac, fac.data(), fac.size() * sizeof(float), hipMemcpyHostToDevice, S.s));
    // the fields keep the caller's values outside the levels and columns the set-up writes
    SETCHK(hipMemcpyAsync(d_u, uk, n3 * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_v, vk, n3 * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_w, wk, n3 * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_kk, kkart, n2 * sizeof(int16_t), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_ahk, ahk, n2 * sizeof(float), hipMemcpyHostToDevice, S.s));
    if (bui_height) SETCHK(hipMemcpyAsync(d_bui, bui_height, n2 * sizeof(float), hipMemcpyHostToDevice, S.s));
    if (advdom) SETCHK(hipMemcpyAsync(d_adv, advdom, n2, hipMemcpyHostToDevice, S.s));
    if (d_cls) SETCHK(hipMemcpyAsync(d_cls, cls.data(), n2 * sizeof(int), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_kad, &one, sizeof(int), hipMemcpyHostToDevice, S.s));
    FlatArgs fa{nii, njj, nkk, nkh, nj2, d_vk, d_hok, d_cut, d_cls, d_ux, d_uy, d_u, d_v, d_w, d_kk, d_ahk, d_bui, d_kad};
    hipLaunchKernelGGL(k_flat_column, dim3((unsigned)((size_t)(nii + 1) * (njj + 1))), dim3(64), 0, S.s, fa);
    SETCHK(hipGetLastError());
    if (m->near_wall != 0 && nii > 2 * igeb && njj > 2 * igeb) {
        DampArgs da{nii, njj, nkk, nkh, nj2, igeb, m->near_wall, d_vk, d_cut, d_ahk, d_kk, d_adv, d_fac, d_u, d_v, nullptr};
        hipLaunchKernelGGL(k_flat_damp, dim3((unsigned)((size_t)(nii - 2 * igeb) * (njj - 2 * igeb))), dim3(64), 0, S.s, da);
        SETCHK(hipGetLastError());
    }
    int kad = 1;
    SETCHK(hipMemcpyAsync(uk, d_u, n3 * sizeof(float), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipMemcpyAsync(vk, d_v, n3 * sizeof(float), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipMemcpyAsync(wk, d_w, n3 * sizeof(float), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipMemcpyAsync(kkart, d_kk, n2 * sizeof(int16_t), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipMemcpyAsync(ahk, d_ahk, n2 * sizeof(float), hipMemcpyDeviceToHost, S.s));
    if (bui_height) SETCHK(hipMemcpyAsync(bui_height, d_bui, n2 * sizeof(float), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipMemcpyAsync(&kad, d_kad, sizeof(int), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipStreamSynchronize(S.s));
    *kadvmax = std::min(nkk, kad + g->vertcellsff);
    return GRAL_PF_OK;
}

// ---------------------------------------------------------------------------
// Terrain (MicroscaleTerrain.cs:57-316): the GRAMM topography and wind field
// interpolated onto the GRAL grid.  Per inner column: the GRAMM cell (indi,
// indj) holding the column centre and the offsets xwert, ywert in it; the
// heights of the GRAMM levels at the centre by 16-point inverse-distance
// weighting of AHE (interior GRAMM cells) or AHE of the cell itself (border);
// per GRAL level, the highest GRAMM level whose height plus CUTK lies below
// the level centre; u, v, w bilinear in the horizontal and linear in the
// vertical between that GRAMM level and the next; below it the level is
// blocked (KKART, AHK, BUI_HEIGHT).
// topo = 1: the variant with the original GRAL topography (:318-598): the cell
// index and the level centre in float (above AHKOriMin), a level blocked when
// its centre is at or below AHKOri + CUTK, the interpolated GRAMM heights
// absolute in the interior (relative to AHMIN at the GRAMM border, as the
// reference has it), GRAMM level 1 with level 0 when no GRAMM level lies
// below the centre.
// ---------------------------------------------------------------------------
struct TerrArgs {
    int nii, njj, nkk, nkh, nj2;
    int nx, ny, nz, gy2, gz2, gy1, gz1;  // GRAMM sizes and the strides of AHE [nx+2][ny+2][nz+2] and the [nx+1][ny+1][nz+1] fields
    float ddx, ddy, dxk, dyk, ahmin;
    int di, dj;
    double zlim;                         // 10 + (ZKO[2]-ZKO[1])*0.5
    const double *vertk;                 // [nkh] HOKART[k-1] + DZK[k]*0.5 + AHMIN (double)
    const float *hok;
    const float *cutk;
    const uint8_t *adv;
    const float *ahe, *uw, *vw, *ww, *zsp, *ah;
    float *uk, *vk, *wk;
    int16_t *kkart;
    float *ahk, *bui;
    float *vec;                          // [n2][2] u, v of the level counted for the 10 m wind, or NaN
    int *kad;
    int topo;
    float aom;                           // AHKOriMin (topo)
    const float *ahko;                   // [n2] AHKOri (topo)
    const float *vtf;                    // [nkh] HOKART[k-1] + DZK[k]*0.5f + AHKOriMin (float, topo)
    const float *dzk;                    // [nkh] (topo)
    int *vlev;                           // [n2] level counted for the 10 m wind, 0: none (topo)
};

__device__ __forceinline__ float tr_p2(float x) { return x * x; }
__device__ __forceinline__ float tr_w(float a, float b) { return 1.0f / tr_p2(tr_p2(a) + tr_p2(b) + 0.1f); }
__device__ __forceinline__ float tr_max(float a, float b) {   // Math.Max(float, float)
    if (a != b) { if (!__builtin_isnan(a)) return b < a ? a : b; return a; }
    return __builtin_signbit(b) ? a : b;
}

#define TERR_NZ 256
__global__ void __launch_bounds__(64) k_terrain_column(TerrArgs a) {
    __shared__ float ch[TERR_NZ + 2], chp[TERR_NZ + 2];
    const int i = 1 + (int)(blockIdx.x / a.njj), j = 1 + (int)(blockIdx.x % a.njj);
    const size_t c = (size_t)i * a.nj2 + j, o = c * a.nkh;
    const float DX = a.dxk, DY = a.dyk, X1 = a.ddx, Y1 = a.ddy, AHMIN = a.ahmin;
    int indi, indj;
    if (a.topo) {
        indi = (int)((((float)a.di + DX * (float)i) - DX * 0.5f) / X1) + 1;
        indj = (int)((((float)a.dj + DY * (float)j) - DY * 0.5f) / Y1) + 1;
    } else {
        indi = (int)(((double)((float)a.di + DX * (float)i) - (double)DX * 0.5) / (double)X1) + 1;
        indj = (int)(((double)((float)a.dj + DY * (float)j) - (double)DY * 0.5) / (double)Y1) + 1;
    }
    const float xw = ((DX * (float)i - DX * 0.5f + (float)a.di) - (float)(indi - 1) * X1);
    const float yw = ((DY * (float)j - DY * 0.5f + (float)a.dj) - (float)(indj - 1) * Y1);
    const bool interior = indi >= 2 && indi <= a.nx - 1 && indj >= 2 && indj <= a.ny - 1;
    const int ci = min(max(indi, 1), a.nx), cj = min(max(indj, 1), a.ny);
    const float *ahe_l = a.ahe + ((size_t)ci * a.gy2 + cj) * a.gz2;
    if (interior) {
        const float xd = X1 - xw, yd = Y1 - yw;
        const float r[16] = {tr_w(xw, yw), tr_w(xw, yd), tr_w(xd, yd), tr_w(xd, yw), tr_w(X1 + xw, Y1 + yw),
                             tr_w(xw, Y1 + yw), tr_w(xd, Y1 + yw), tr_w(X1 + xd, Y1 + yw), tr_w(X1 + xd, yw),
                             tr_w(X1 + xd, yd), tr_w(X1 + xd, Y1 + yd), tr_w(xd, Y1 + yd), tr_w(xw, Y1 + yd),
                             tr_w(X1 + xw, Y1 + yd), tr_w(X1 + xw, yd), tr_w(X1 + xw, yw)};
        const int di[16] = {0, 0, 1, 1, -1, 0, 1, 2, 2, 2, 2, 1, 0, -1, -1, -1};
        const int dj[16] = {0, 1, 1, 0, -1, -1, -1, -1, 0, 1, 2, 2, 2, 2, 1, 0};
        float gew = r[0];
        for (int q = 1; q < 16; ++q) gew += r[q];
        for (int ik = 1 + threadIdx.x; ik <= a.nz; ik += 64) {
            const int ikp = min(ik + 1, a.nz);
            float s = 0.0f, sp = 0.0f;
            for (int q = 0; q < 16; ++q) {
                const float *col = a.ahe + ((size_t)(indi + di[q]) * a.gy2 + (indj + dj[q])) * a.gz2;
                const float t = col[ik] * r[q], tp = col[ikp] * r[q];
                s = q ? s + t : t;
                sp = q ? sp + tp : tp;
            }
            ch[ik] = a.topo ? s / gew : s / gew - AHMIN;
            chp[ik] = a.topo ? sp / gew : sp / gew - AHMIN;
        }
    }
    __syncthreads();
    const float cut = a.cutk[c];
    const float ch1 = interior ? ch[1] : ahe_l[1] - AHMIN;
    // the highest GRAMM level below a GRAL level centre: returns ik (0: blocked),
    // with the heights of that GRAMM level and the next and the height above it
    const float ground = a.topo ? a.ahko[c] + cut : 0.0f;
    auto search = [&](int k, float &c_h, float &c_hp, float &dz) {
        if (a.topo) {                    // -1: no GRAMM level below the centre
            const float vf = a.vtf[k];
            if (vf <= ground) return 0;
            for (int ik = a.nz; ik >= 1; --ik) {
                const int ikp = min(ik + 1, a.nz);
                c_h = interior ? ch[ik] : ahe_l[ik] - AHMIN;
                c_hp = interior ? chp[ik] : ahe_l[ikp] - AHMIN;
                if (vf > c_h + cut) {
                    dz = vf - (c_h + cut);
                    return ik;
                }
            }
            dz = 0.0f;
            return -1;
        }
        const double vz = a.vertk[k] - (double)AHMIN;
        for (int ik = a.nz; ik >= 1; --ik) {
            const int ikp = min(ik + 1, a.nz);
            c_h = interior ? ch[ik] : ahe_l[ik] - AHMIN;
            c_hp = interior ? chp[ik] : ahe_l[ikp] - AHMIN;
            if (vz > (double)(c_h + cut)) {
                dz = (float)vz - (c_h + cut);
                return ik;
            }
        }
        return 0;
    };
    // pass 1: blocked levels (from the ground up), KKART and AHK
    int kb = 0;
    const float base = a.topo ? a.aom : AHMIN;
    float ahk = base;
    for (int k = 1 + threadIdx.x; k <= a.nkk; k += 64) {
        float c_h, c_hp, dz;
        if (search(k, c_h, c_hp, dz) == 0) {
            kb = max(kb, k);
            ahk = tr_max(a.hok[k] + base, ahk);
        }
    }
    for (int m = 32; m >= 1; m >>= 1) {
        kb = max(kb, __shfl_xor(kb, m));
        const float oa = __shfl_xor(ahk, m);
        ahk = oa > ahk ? oa : ahk;
    }
    // topo: BUI_HEIGHT of a building column from the blocked levels at or above
    // AHKOri (k > 1): AHK minus the bottom of the level, the largest one
    float bmax = -__builtin_inff();
    bool bany = false;
    if (a.topo && cut > 0) {
        for (int k = 2 + threadIdx.x; k <= kb; k += 64)
            if (a.vtf[k] >= a.ahko[c]) {
                const float cand = ahk - (a.hok[k - 1] + a.aom);
                bmax = bany ? tr_max(cand, bmax) : cand;
                bany = true;
            }
        for (int m = 32; m >= 1; m >>= 1) {
            const float ob = __shfl_xor(bmax, m);
            const bool oy = __shfl_xor((int)bany, m) != 0;
            if (oy) { bmax = bany ? (ob > bmax ? ob : bmax) : ob; bany = true; }
        }
    }
    // pass 2: u, v, w of the free levels; w of the first free level is zeroed by
    // the blocked level below it; the highest free level whose GRAMM level is
    // within about 10 m of the ground counts for the mean wind
    int vkk = 0;
    float vu = 0.0f, vv = 0.0f;
    for (int k = 1 + threadIdx.x; k <= a.nkk; k += 64) {
        float c_h = 0.0f, c_hp = 0.0f, dz = 0.0f;
        const int indk = search(k, c_h, c_hp, dz);
        float u = 0.0f, v = 0.0f, w = 0.0f;
        if (indk != 0) {
            const int ik0 = indk > 0 ? indk : 1;
            const int kp = indk > 0 ? min(indk + 1, a.nz) : 0;
            const int ixm = (xw < X1 * 0.5f) ? max(ci - 1, 1) : min(ci + 1, a.nx);
            const int iym = (yw < Y1 * 0.5f) ? max(cj - 1, 1) : min(cj + 1, a.ny);
            const float fx = xw - X1 * 0.5f, fy = yw - Y1 * 0.5f, sx = (float)(ixm - ci), sy = (float)(iym - cj);
            const float den = tr_max(c_hp - c_h, 0.1f);
            auto at = [&](const float *W, int x, int y, int z) { return W[((size_t)x * a.gy1 + y) * a.gz1 + z]; };
            auto interp = [&](const float *W) {
                auto hx = [&](int y, int z) {
                    const float b0 = at(W, ci, y, z);
                    return b0 + (at(W, ixm, y, z) - b0) / X1 * sx * fx;
                };
                const float x1 = hx(cj, ik0), x2 = hx(iym, ik0), x3 = hx(cj, kp), x4 = hx(iym, kp);
                const float x5 = x1 + (x3 - x1) / den * dz, x6 = x2 + (x4 - x2) / den * dz;
                return x5 + (x6 - x5) / Y1 * sy * fy;
            };
            u = interp(a.uw);
            v = interp(a.vw);
            w = (k == kb + 1 && kb > 0) ? 0.0f : interp(a.ww);
            if (a.topo) {
                if (a.vtf[k] < (a.ahko[c] + 10.0f) + 0.5f * a.dzk[k] && k > vkk) vkk = k;
            } else {
                const float zz = a.zsp[((size_t)ci * a.gy1 + cj) * a.gz1 + indk] - a.ah[(size_t)ci * a.gy1 + cj];
                if ((double)zz < a.zlim && k > vkk) { vkk = k; vu = u; vv = v; }
            }
        }
        a.uk[o + k] = u;
        a.vk[o + k] = v;
        a.wk[o + k] = w;
    }
    for (int m = 32; m >= 1; m >>= 1) {
        const int ok = __shfl_xor(vkk, m);
        const float ou = __shfl_xor(vu, m), ov = __shfl_xor(vv, m);
        if (ok > vkk) { vkk = ok; vu = ou; vv = ov; }
    }
    if (threadIdx.x == 0) {
        a.kkart[c] = (int16_t)kb;
        a.ahk[c] = ahk;
        if (a.topo) {
            if (bany) a.bui[c] = tr_max(bmax, a.bui[c]);
            a.vlev[c] = vkk;
        } else if (kb > 0 && cut > 0) {
            a.bui[c] = tr_max(ahk - AHMIN - ch1, a.bui[c]);
        }
        if (a.adv[c] > 0) atomicMax(a.kad, kb);
        a.vec[2 * c] = vkk > 0 ? vu : __builtin_nanf("");
        a.vec[2 * c + 1] = vkk > 0 ? vv : 0.0f;
    }
}

// u at levels up to KKART of the column at i-1, v up to KKART of the column
// at j-1 (the face of a blocked neighbour); after every column is set
__global__ void __launch_bounds__(64) k_terrain_faces(TerrArgs a) {
    const int i = 1 + (int)(blockIdx.x / a.njj), j = 1 + (int)(blockIdx.x % a.njj);
    const size_t c = (size_t)i * a.nj2 + j, o = c * a.nkh;
    const int ki = i > 1 ? a.kkart[c - a.nj2] : 0, kj = j > 1 ? a.kkart[c - 1] : 0;
    for (int k = 1 + threadIdx.x; k <= max(ki, kj) && k <= a.nkk; k += 64) {
        if (k <= ki) a.uk[o + k] = 0.0f;
        if (k <= kj) a.vk[o + k] = 0.0f;
    }
}

// Ghost row NII+1 of u and column NJJ+1 of v.  Every free cell of row j
// copies u[NII][j][k] into u[NII+1][j][k] when it is set: the last copy is the
// final u[NII] when column NII is free at k, else the value u[NII][j][k] had
// before the set-up (no column set it since), when any column before NII is
// free at k; else the ghost keeps its value.  Likewise for v along j.
__global__ void __launch_bounds__(64) k_terrain_ghosts(TerrArgs a, const float *u_old, const float *v_old) {
    const int q = blockIdx.x;
    const bool rowu = q < a.njj;                 // blocks 0..NJJ-1: u rows j; then v columns i
    const int j = rowu ? q + 1 : a.njj, i = rowu ? a.nii : q - a.njj + 1;
    // lowest KKART among the columns before NII (before NJJ)
    int mn = 1 << 30;
    const int nb = rowu ? a.nii - 1 : a.njj - 1;
    for (int t = threadIdx.x; t < nb; t += 64) {
        const size_t cc = rowu ? (size_t)(t + 1) * a.nj2 + j : (size_t)i * a.nj2 + (t + 1);
        mn = min(mn, (int)a.kkart[cc]);
    }
    for (int m = 32; m >= 1; m >>= 1) mn = min(mn, __shfl_xor(mn, m));
    const size_t csrc = (size_t)i * a.nj2 + j;                         // column NII (row u) or NJJ (column v)
    const size_t cdst = rowu ? (size_t)(a.nii + 1) * a.nj2 + j : (size_t)i * a.nj2 + (a.njj + 1);
    const int kb = a.kkart[csrc];
    float *f = rowu ? a.uk : a.vk;
    const float *old = rowu ? u_old + (size_t)j * a.nkh : v_old + (size_t)i * a.nkh;
    for (int k = 1 + threadIdx.x; k <= a.nkk; k += 64) {
        if (k > kb) f[cdst * a.nkh + k] = f[csrc * a.nkh + k];
        else if (k > mn) f[cdst * a.nkh + k] = old[k];
    }
}

// topo: the ghosts are copied while column i = 1 (j = 1) is set, in the serial
// order: at the levels where that column is free, u[NII+1][j] takes the value
// u[NII][j] had before the set-up (its final value when NII = 1); elsewhere the
// ghost keeps its own.  Likewise v along j.
__global__ void __launch_bounds__(64) k_topo_ghosts(TerrArgs a, const float *u_old, const float *v_old) {
    const int q = blockIdx.x;
    const bool rowu = q < a.njj;
    const int j = rowu ? q + 1 : 1, i = rowu ? 1 : q - a.njj + 1;
    const int kfree = a.kkart[(size_t)i * a.nj2 + j];                 // column (1, j) or (i, 1)
    const size_t cdst = rowu ? (size_t)(a.nii + 1) * a.nj2 + j : (size_t)i * a.nj2 + (a.njj + 1);
    const bool own = rowu ? a.nii == 1 : a.njj == 1;
    const size_t cown = rowu ? (size_t)a.nii * a.nj2 + j : (size_t)i * a.nj2 + a.njj;
    float *f = rowu ? a.uk : a.vk;
    const float *old = rowu ? u_old + (size_t)j * a.nkh : v_old + (size_t)i * a.nkh;
    for (int k = kfree + 1 + threadIdx.x; k <= a.nkk; k += 64) f[cdst * a.nkh + k] = own ? f[cown * a.nkh + k] : old[k];
}

__global__ void k_terrain_save(const float *uk, const float *vk, int nii, int njj, int nkh, float *u_old, float *v_old) {
    const int nj2 = njj + 2;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nu = (size_t)nj2 * nkh, nv = (size_t)(nii + 2) * nkh;
    if (t < nu) u_old[t] = uk[(size_t)nii * nj2 * nkh + t];                    // row NII: [j][k]
    else if (t < nu + nv) {
        const size_t q = t - nu, i = q / nkh, k = q % nkh;
        v_old[q] = vk[(i * nj2 + njj) * nkh + k];                              // column NJJ: [i][k]
    }
}

// topo: u, v of the level counted for the 10 m wind, after the faces (the
// reference counts the values its face rules left) and before any damping
__global__ void k_topo_wind(TerrArgs a) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)a.nii * a.njj) return;
    const int i = 1 + (int)(t / a.njj), j = 1 + (int)(t % a.njj);
    const size_t c = (size_t)i * a.nj2 + j;
    const int lv = a.vlev[c];
    a.vec[2 * c] = lv > 0 ? a.uk[c * a.nkh + lv] : __builtin_nanf("");
    a.vec[2 * c + 1] = lv > 0 ? a.vk[c * a.nkh + lv] : 0.0f;
}

static int terrain_run(const gral_pf_grid *g, const gral_pf_gramm *gr, int topo, const float *ahko, float aom,
                       const float *cutk, const uint8_t *advdom, int32_t near_wall, float *uk, float *vk, float *wk,
                       int16_t *kkart, float *ahk, float *bui_height, float *ahmin, int32_t *kadvmax, double *wind10) {
    if (!g || !gr || !cutk || !advdom || !uk || !vk || !wk || !kkart || !ahk || !bui_height || !ahmin || !kadvmax ||
        !g->dzk || !g->hokart || (topo && !ahko))
        return setfail(GRAL_PF_EINVAL, "terrain_init: null argument");
    if (g->nii < 1 || g->njj < 1 || g->nkk < 2 || g->vertcellsff < 0 || !(g->dxk > 0) || !(g->dyk > 0))
        return setfail(GRAL_PF_EINVAL, "terrain_init: bad grid");
    if (gr->nx < 1 || gr->ny < 1 || gr->nz < 1 || gr->nz > TERR_NZ || !(gr->ddx > 0) || !(gr->ddy > 0) || !gr->ahe ||
        !gr->uwin || !gr->vwin || !gr->wwin || (!topo && (!gr->zsp || !gr->ah)))
        return setfail(GRAL_PF_EINVAL, "terrain_init: bad GRAMM grid (1 <= nz <= 256)");
    if (near_wall < 0 || near_wall > 2) return setfail(GRAL_PF_EINVAL, "terrain_init: near_wall must be 0, 1 or 2");
    const int nii = g->nii, njj = g->njj, nkk = g->nkk, nkh = nkk + 2, nj2 = njj + 2;
    const size_t n2 = (size_t)(nii + 2) * nj2, n3 = n2 * nkh;
    const int nx = gr->nx, ny = gr->ny, nz = gr->nz;
    const size_t nahe = (size_t)(nx + 2) * (ny + 2) * (nz + 2), ngf = (size_t)(nx + 1) * (ny + 1) * (nz + 1),
                 nah = (size_t)(nx + 1) * (ny + 1);
    const float AHMIN = *ahmin;
    std::vector<double> vertk(nkh, 0.0);
    std::vector<float> vtf(nkh, 0.0f);
    for (int k = 1; k <= nkk; ++k) {
        vertk[k] = (double)g->hokart[k - 1] + (double)g->dzk[k] * 0.5 + (double)AHMIN;
        vtf[k] = (g->hokart[k - 1] + g->dzk[k] * 0.5f) + aom;
    }
    Scratch S;
    SETCHK(hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking));
    double *d_vz = S.alloc<double>(nkh);
    float *d_hok = S.alloc<float>(nkh), *d_cut = S.alloc<float>(n2), *d_ahe = S.alloc<float>(nahe);
    float *d_uw = S.alloc<float>(ngf), *d_vw = S.alloc<float>(ngf), *d_ww = S.alloc<float>(ngf);
    float *d_zsp = topo ? nullptr : S.alloc<float>(ngf), *d_ah = topo ? nullptr : S.alloc<float>(nah);
    float *d_ahko = topo ? S.alloc<float>(n2) : nullptr, *d_vtf = topo ? S.alloc<float>(nkh) : nullptr;
    float *d_dzk = topo ? S.alloc<float>(nkh) : nullptr;
    int *d_vlev = topo ? S.alloc<int>(n2) : nullptr;
    float *d_u = S.alloc<float>(n3), *d_v = S.alloc<float>(n3), *d_w = S.alloc<float>(n3);
    float *d_ahk = S.alloc<float>(n2), *d_bui = S.alloc<float>(n2), *d_vec = S.alloc<float>(2 * n2);
    float *d_uo = S.alloc<float>((size_t)nj2 * nkh), *d_vo = S.alloc<float>((size_t)(nii + 2) * nkh);
    int16_t *d_kk = S.alloc<int16_t>(n2);
    uint8_t *d_adv = S.alloc<uint8_t>(n2);
    int *d_kad = S.alloc<int>(1);
    if (!d_vz || !d_hok || !d_cut || !d_ahe || !d_uw || !d_vw || !d_ww || (!topo && (!d_zsp || !d_ah)) || !d_u || !d_v ||
        !d_w || !d_ahk || !d_bui || !d_vec || !d_uo || !d_vo || !d_kk || !d_adv || !d_kad ||
        (topo && (!d_ahko || !d_vtf || !d_dzk || !d_vlev)))
        return setfail(GRAL_PF_ENOMEM, "terrain_init: hipMalloc failed");
    const int one = 1;
    SETCHK(hipMemcpyAsync(d_vz, vertk.data(), nkh * sizeof(double), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_hok, g->hokart, nkh * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_cut, cutk, n2 * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_adv, advdom, n2, hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_ahe, gr->ahe, nahe * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_uw, gr->uwin, ngf * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_vw, gr->vwin, ngf * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_ww, gr->wwin, ngf * sizeof(float), hipMemcpyHostToDevice, S.s));
    if (topo) {
        SETCHK(hipMemcpyAsync(d_ahko, ahko, n2 * sizeof(float), hipMemcpyHostToDevice, S.s));
        SETCHK(hipMemcpyAsync(d_vtf, vtf.data(), nkh * sizeof(float), hipMemcpyHostToDevice, S.s));
        SETCHK(hipMemcpyAsync(d_dzk, g->dzk, nkh * sizeof(float), hipMemcpyHostToDevice, S.s));
    } else {
        SETCHK(hipMemcpyAsync(d_zsp, gr->zsp, ngf * sizeof(float), hipMemcpyHostToDevice, S.s));
        SETCHK(hipMemcpyAsync(d_ah, gr->ah, nah * sizeof(float), hipMemcpyHostToDevice, S.s));
    }
    SETCHK(hipMemcpyAsync(d_u, uk, n3 * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_v, vk, n3 * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_w, wk, n3 * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_kk, kkart, n2 * sizeof(int16_t), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_ahk, ahk, n2 * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_bui, bui_height, n2 * sizeof(float), hipMemcpyHostToDevice, S.s));
    SETCHK(hipMemcpyAsync(d_kad, &one, sizeof(int), hipMemcpyHostToDevice, S.s));
    TerrArgs ta{};
    ta.nii = nii; ta.njj = njj; ta.nkk = nkk; ta.nkh = nkh; ta.nj2 = nj2;
    ta.nx = nx; ta.ny = ny; ta.nz = nz; ta.gy2 = ny + 2; ta.gz2 = nz + 2; ta.gy1 = ny + 1; ta.gz1 = nz + 1;
    ta.ddx = gr->ddx; ta.ddy = gr->ddy; ta.dxk = g->dxk; ta.dyk = g->dyk; ta.ahmin = AHMIN;
    ta.di = gr->delta_i; ta.dj = gr->delta_j;
    ta.zlim = 10.0 + (gr->zko2 - gr->zko1) * (double)0.5f;
    ta.vertk = d_vz; ta.hok = d_hok; ta.cutk = d_cut; ta.adv = d_adv;
    ta.ahe = d_ahe; ta.uw = d_uw; ta.vw = d_vw; ta.ww = d_ww; ta.zsp = d_zsp; ta.ah = d_ah;
    ta.uk = d_u; ta.vk = d_v; ta.wk = d_w; ta.kkart = d_kk; ta.ahk = d_ahk; ta.bui = d_bui; ta.vec = d_vec; ta.kad = d_kad;
    ta.topo = topo; ta.aom = aom; ta.ahko = d_ahko; ta.vtf = d_vtf; ta.dzk = d_dzk; ta.vlev = d_vlev;
    const size_t nsave = (size_t)nj2 * nkh + (size_t)(nii + 2) * nkh;
    hipLaunchKernelGGL(k_terrain_save, dim3((unsigned)((nsave + 255) / 256)), dim3(256), 0, S.s, d_u, d_v, nii, njj, nkh,
                       d_uo, d_vo);
    const unsigned ncol = (unsigned)((size_t)nii * njj);
    hipLaunchKernelGGL(k_terrain_column, dim3(ncol), dim3(64), 0, S.s, ta);
    hipLaunchKernelGGL(k_terrain_faces, dim3(ncol), dim3(64), 0, S.s, ta);
    if (topo) {
        hipLaunchKernelGGL(k_topo_ghosts, dim3((unsigned)(nii + njj)), dim3(64), 0, S.s, ta, (const float *)d_uo,
                           (const float *)d_vo);
        hipLaunchKernelGGL(k_topo_wind, dim3((ncol + 255) / 256), dim3(256), 0, S.s, ta);
    } else {
        hipLaunchKernelGGL(k_terrain_ghosts, dim3((unsigned)(nii + njj)), dim3(64), 0, S.s, ta, (const float *)d_uo,
                           (const float *)d_vo);
    }
    SETCHK(hipGetLastError());
    // AHMIN of the result: the lowest AHK of the inner columns (MicroscaleTerrain.cs:605-612)
    std::vector<float> h_ahk(n2);
    SETCHK(hipMemcpyAsync(h_ahk.data(), d_ahk, n2 * sizeof(float), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipStreamSynchronize(S.s));
    float amin = 100000.0f;
    for (int i = 1; i <= nii; ++i)
        for (int j = 1; j <= njj; ++j) {
            const float x = h_ahk[(size_t)i * nj2 + j];
            if (amin != x) amin = (!std::isnan(amin)) ? (amin < x ? amin : x) : amin;
            else amin = std::signbit(amin) ? amin : x;
        }
    const int igeb = std::max((int)(20.0f / g->dxk), 1);
    if (near_wall != 0 && nii > 2 * igeb && njj > 2 * igeb) {
        std::vector<float> vf(nkh, 0.0f), fac = damping_table(g->dxk, g->dyk, igeb);
        for (int k = 1; k <= nkk; ++k) vf[k] = g->hokart[k - 1] + g->dzk[k] * 0.5f + amin;
        float *d_vf = S.alloc<float>(nkh), *d_fac = S.alloc<float>(fac.size());
        if (!d_vf || !d_fac) return setfail(GRAL_PF_ENOMEM, "terrain_init: hipMalloc failed");
        SETCHK(hipMemcpyAsync(d_vf, vf.data(), nkh * sizeof(float), hipMemcpyHostToDevice, S.s));
        SETCHK(hipMemcpyAsync(d_fac, fac.data(), fac.size() * sizeof(float), hipMemcpyHostToDevice, S.s));
        DampArgs da{nii, njj, nkk, nkh, nj2, igeb, near_wall, d_vf, d_cut, d_ahk, d_kk, d_adv, d_fac, d_u, d_v, d_w};
        hipLaunchKernelGGL(k_flat_damp, dim3((unsigned)((size_t)(nii - 2 * igeb) * (njj - 2 * igeb))), dim3(64), 0, S.s, da);
        SETCHK(hipGetLastError());
    }
    int kad = 1;
    std::vector<float> vec(2 * n2);
    SETCHK(hipMemcpyAsync(uk, d_u, n3 * sizeof(float), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipMemcpyAsync(vk, d_v, n3 * sizeof(float), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipMemcpyAsync(wk, d_w, n3 * sizeof(float), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipMemcpyAsync(kkart, d_kk, n2 * sizeof(int16_t), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipMemcpyAsync(ahk, d_ahk, n2 * sizeof(float), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipMemcpyAsync(bui_height, d_bui, n2 * sizeof(float), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipMemcpyAsync(vec.data(), d_vec, 2 * n2 * sizeof(float), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipMemcpyAsync(&kad, d_kad, sizeof(int), hipMemcpyDeviceToHost, S.s));
    SETCHK(hipStreamSynchronize(S.s));
    *kadvmax = std::min(nkk, kad + g->vertcellsff);
    *ahmin = amin;
    if (wind10) {
        // the mean wind near 10 m above ground: per row i a sum over j (u in
        // double; v in double, in float for the topography variant), the row
        // sums added in i order (MicroscaleTerrain.cs:67, :311-313, :329, :593-595)
        double sx = 0, sy = 0;
        long cnt = 0;
        for (int i = 1; i <= nii; ++i) {
            double lx = 0, ly = 0;
            float lyf = 0.0f;
            for (int j = 1; j <= njj; ++j) {
                const size_t c = (size_t)i * nj2 + j;
                if (std::isnan(vec[2 * c])) continue;
                lx += (double)vec[2 * c];
                if (topo) lyf = lyf + vec[2 * c + 1];
                else ly += (double)vec[2 * c + 1];
                ++cnt;
            }
            sx += lx;
            sy += topo ? (double)lyf : ly;
        }
        wind10[0] = sx;
        wind10[1] = sy;
        wind10[2] = (double)cnt;
    }
    return GRAL_PF_OK;
}

extern "C" int gral_pf_terrain_init(const gral_pf_grid *g, const gral_pf_gramm *gr, const float *cutk,
                                    const uint8_t *advdom, int32_t near_wall, float *uk, float *vk, float *wk,
                                    int16_t *kkart, float *ahk, float *bui_height, float *ahmin, int32_t *kadvmax,
                                    double *wind10) {
    return terrain_run(g, gr, 0, nullptr, 0.0f, cutk, advdom, near_wall, uk, vk, wk, kkart, ahk, bui_height, ahmin,
                       kadvmax, wind10);
}

extern "C" int gral_pf_terrain_init_topo(const gral_pf_grid *g, const gral_pf_gramm *gramm, const float *ahk_ori,
                                         float ahk_ori_min, const float *cutk, const uint8_t *advdom, int32_t near_wall,
                                         float *uk, float *vk, float *wk, int16_t *kkart, float *ahk, float *bui_height,
                                         float *ahmin, int32_t *kadvmax, double *wind10) {
    return terrain_run(g, gramm, 1, ahk_ori, ahk_ori_min, cutk, advdom, near_wall, uk, vk, wk, kkart, ahk, bui_height,
                       ahmin, kadvmax, wind10);
}

extern "C" const char *gral_pf_setup_last_error(void) { return g_err.c_str(); }
