// Host side of the C ABI (include/gral_pf.h): argument checking, the
// prologue and epilogue of PrognosticFlowfield.Calculate(), device memory, and
// the SIMPLE-iteration driver with the reference's convergence bookkeeping.
//
// The prologue (PrognosticFlowfield.cs:145-579) runs once per situation on the
// host: it needs the full-height columns (topwind at NKK-1, the initial UKS
// over all levels) and the init-time transcendentals (MathF.Log, Math.Pow),
// which are evaluated with the C library like .NET on Linux does.  Only levels
// 0..KADVMAX+1 -- everything the iterations touch -- are moved to the GPU.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gral_pf.h"
#include "pf_device.h"

// Levels per device column: KADVMAX + 2 rounded up to a multiple of NKP_ALIGN
// (a power of two >= 4: the cell kernels move 16-byte quads of levels).  With
// 16, every column starts on a 64-byte boundary, so a column's active levels
// touch the fewest 64-byte sectors; on the 400 x 400 x 40 city (KADVMAX 74:
// 80 levels stored instead of 76) an 8-situation iteration takes 18.18 ms
// against 18.37 with 4 and 18.35 with 32 (128-byte boundaries, 96 levels).
#ifndef NKP_ALIGN
#define NKP_ALIGN 16
#endif

struct Probe {
    hipEvent_t *ev;
    int *fam;
    int cap, n;
};
extern "C" int pf_launch_iteration(const Dom *hd, int nsit, const Dom *dd, int iter, int slot, hipStream_t s,
                                   long long *launches, Probe *probe);
extern "C" int pf_fill_u32(unsigned *p, size_t n, unsigned v, hipStream_t s);
extern "C" int pf_launch_diagnostic(const DiagArgs *ga, int iterations, hipStream_t s);
extern "C" int pf_launch_gff(const GffArgs *ga, hipStream_t s);
extern "C" int pf_launch_wind(const WindArgs *wa, hipStream_t s);
extern "C" int pf_launch_mass(const Dom *dd, int nsit, int nblk, double *partial, double *out, hipStream_t s);
extern "C" int pf_launch_projection(const Dom *hd, const Dom *dd, int iter, hipStream_t s, long long *launches);
extern "C" int pf_launch_momentum_sweep(const Dom *hd, const Dom *dd, int comp, int IS, int JS, hipStream_t s);
extern "C" int pf_launch_momentum_vjp(const Dom *hd, const Dom *dd, const MomAdj *ma, int comp, int IS, int JS,
                                      hipStream_t s);
extern "C" int pf_launch_adjoint_step(const Dom *hd, const Dom *dd, const AdjArgs *args, int iter, size_t n3,
                                      hipStream_t s);

namespace {

inline float p2f(float x) { return x * x; }
inline float fmax_net(float a, float b) {  // Math.Max(float,float)
    if (a != b) { if (!std::isnan(a)) return b < a ? a : b; return a; }
    return std::signbit(b) ? a : b;
}
inline float fmin_net(float a, float b) {  // Math.Min(float,float)
    if (a != b) { if (!std::isnan(a)) return a < b ? a : b; return a; }
    return std::signbit(a) ? a : b;
}
inline double dmax_net(double a, double b) {
    if (a != b) { if (!std::isnan(a)) return b < a ? a : b; return a; }
    return std::signbit(b) ? a : b;
}

// Host loops over columns or rows run on up to 16 threads: each element is
// computed as in a serial loop, and the only reductions (a maximum, counts)
// do not depend on the order.
unsigned host_threads() {
    static const unsigned n = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    return n;
}
template <class Fn>
void par_for(long lo, long hi, Fn fn) {            // fn(t, a, b): thread t covers [a, b)
    const long n = hi - lo;
    if (n <= 0) return;
    const long T = std::min<long>(host_threads(), n);
    if (T <= 1) { fn(0, lo, hi); return; }
    const long chunk = (n + T - 1) / T;
    std::vector<std::thread> th;
    th.reserve((size_t)T - 1);
    for (long t = 1; t < T; ++t) {
        const long a = lo + t * chunk, b = std::min(hi, a + chunk);
        if (a < b) th.emplace_back(fn, (int)t, a, b);
    }
    fn(0, lo, std::min(hi, lo + chunk));
    for (auto &x : th) x.join();
}
void par_copy(void *dst, const void *src, size_t bytes) {
    const size_t blk = size_t(1) << 22;
    par_for(0, (long)((bytes + blk - 1) / blk), [&](int, long a, long b) {
        const size_t o = (size_t)a * blk, e = std::min(bytes, (size_t)b * blk);
        std::memcpy(static_cast<char *>(dst) + o, static_cast<const char *>(src) + o, e - o);
    });
}
void par_fill(std::vector<float> &v, size_t n, float x) {
    v.resize(n);
    par_for(0, (long)((n + 1048575) >> 20), [&](int, long a, long b) {
        std::fill(v.begin() + std::min(n, (size_t)a << 20), v.begin() + std::min(n, (size_t)b << 20), x);
    });
}
void par_assign(std::vector<float> &v, const float *src, size_t n) {
    v.resize(n);
    par_copy(v.data(), src, n * sizeof(float));
}

}  // namespace

// FP64 mass-imbalance reduction: blocks of its first stage at most (one
// partial pair per block), and the device scratch of the allocation-free
// entry point gral_pf_mass_imbalance_noalloc: [2 * kMassMaxBlocks] block
// partials, then the [2] result.  Every GPU the library runs on holds its own
// copy for the life of the process; calls that overlap on one GPU race on it.
static const int kMassMaxBlocks = 1024;
static const size_t kMassScratchDoubles = 2 * (size_t)kMassMaxBlocks + 2;
static __device__ __align__(256) double scratch[kMassScratchDoubles];
static const int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch[kMaxDevices];       // device address of scratch, per device

// Looks up the device address of scratch on device dev (which must be the
// current device) once per device; 0 on success.
static int scratch_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch[dev]) return 0;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
    g_scratch[dev] = static_cast<double *>(p);
    return 0;
}
static double *scratch_of(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return g_scratch[dev];
}

// A captured chunk of iterations (hipGraph) for replay: the launches of
// `chunk` iterations of the situations in `set`, captured at an iteration
// congruent mod 4 (the sweep directions) to the chunks that replay it.  Its
// kernels address the residual history by slot within the chunk; each
// situation's Dom::iter_off, uploaded before every replay, maps that slot to
// the situation's own iteration.
struct GraphCache {
    std::vector<unsigned long long> set;   // generation numbers of the situations (prepare assigns them)
    const Dom *dd = nullptr;               // the situation table the kernels read
    int chunk = 0, it_mod4 = -1;
    long long launches = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void clear() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        set.clear();
        dd = nullptr;
        chunk = 0;
        it_mod4 = -1;
        launches = 0;
    }
    ~GraphCache() { clear(); }
};
// GRAL_PF_GRAPH=1 replays captured iteration chunks instead of launching each
// kernel (the first chunk of a set of situations runs eagerly and is captured)
static bool graphs_enabled() {
    const char *e = std::getenv("GRAL_PF_GRAPH");
    return e && e[0] == '1';
}

struct gral_pf_handle {
    unsigned long long gen = 0;     // generation: set by every prepare, unique in the process
    // grid
    int device = -1;                // the HIP device current at create
    int nii = 0, njj = 0, nkk = 0, kadvmax = 0, vcf = 0, nkp = 0, nkh = 0;
    float dxk = 0, dyk = 0;
    std::vector<float> dzk, hok;
    // host state (full depth nkh = nkk+2)
    std::vector<uint8_t> adv, vegp;
    std::vector<int16_t> kk;
    std::vector<float> cutk, cov, ahk, veg;
    std::vector<float> z0w, z0s, obi, usti, z0i;
    bool masks = false, rough = false, fields = false, has_veg = false;
    // meteorological profile of the TKE initialisation: Program.IStatistics,
    // MetProfileNumb, MeasurementHeight[], U0[], V0[] (index 0 unused)
    int istat = 4, nprof = 0;
    std::vector<float> mh, u0, v0;
    size_t dev_bytes = 0;           // device bytes allocated by prepare
    double *mass_buf = nullptr;     // FP64 mass imbalance: [mass_nblk][2] block partials, then the [2] result
    int mass_nblk = 0;
    double mass[2] = {0, 0};        // sum |DIV|, sum |face flux| at the last finish / mass_imbalance call
    // adjoint of the pressure-correction steps (gral_pf_forward_ckpt / gral_pf_vjp)
    std::vector<float> app, abp, atp;   // per-level pressure coefficients of the prologue
    float *adj = nullptr;           // ub vb wb pb qb mb db, checkpoint uk vk wk, ptab[2] ttab[2]: 15 fields
    int *adj_src = nullptr;         // stale-top source column per column, for IS = +1 and -1
    int adj_first = 0, adj_n = 0;   // iterations of the last forward
    int mom_comp = -1, mom_is = 1, mom_js = 1;   // the last gral_pf_momentum_sweep: component, directions
    bool mom_vjp_ok = false;                      // no call has changed the device fields since that sweep
    GraphCache *graphs = nullptr;   // captured iteration chunks (GRAL_PF_GRAPH=1)
    size_t dev_cap = 0;             // optional cap (GRAL_PF_DEVICE_MB), 0 = none
    std::vector<float> uk, vk, wk, uks, vks, wks, dpm, dpmnew, turb, tdiss;
    std::vector<int> vi;
    // run state
    gral_pf_params prm{};
    bool prepared = false, done = false, host_stale = false;
    int iter_loops = 1, iters_run = 0, processed = 0;
    double delta_mean = 0, last_delta = 0;
    float delta_first = 0, delta_finish = 0;
    float topwind = 0;
    double geschw_max = 0;
    int unreal[3] = {0, 0, 0};
    gral_pf_timing timing{};
    long long active = 0;           // active cells (advanced per SIMPLE iteration)
    float diag_dtime = 0;           // time step of the last diagnostic post-pass
    // device
    Dom d{};
    Dom *dd = nullptr;              // device copy of d (a one-situation launch array)
    std::vector<void *> dev;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // flow-field file packer: device copy of the host levels [gff_nlo, nkh) of
    // UK, VK, WK, TURB (the levels the solver does not hold on the device)
    void *gff_up = nullptr;
    size_t gff_bytes = 0;
    void *gff_dbuf = nullptr, *gff_hbuf = nullptr;   // packed payload: device buffer, pinned host staging
    void *pw_buf = nullptr;         // particle wind: the particles' inputs and outputs on the device
    size_t pw_bytes = 0;
    void *stage_buf = nullptr;      // pinned staging of one device field (put_field / get_field)
    size_t stage_bytes = 0;
    size_t gff_dbuf_bytes = 0, gff_hbuf_bytes = 0;
    int gff_nlo = -1;
    bool gff_ok = false;
    std::string err;

    size_t n2() const { return (size_t)(nii + 2) * (size_t)(njj + 2); }
    size_t i2(int i, int j) const { return (size_t)i * (size_t)(njj + 2) + (size_t)j; }
    size_t i3(int i, int j, int k) const { return i2(i, j) * (size_t)nkh + (size_t)k; }
};

static int fail(gral_pf_handle *h, int code, const std::string &msg) {
    if (h) h->err = msg;
    return code;
}

#define HIPCHK(h, call)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail((h), GRAL_PF_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
static int dalloc(gral_pf_handle *h, T **p, size_t n) {
    void *q = nullptr;
    const size_t bytes = n * sizeof(T) + 256;
    if (h->dev_cap && h->dev_bytes + bytes > h->dev_cap)
        return fail(h, GRAL_PF_ENOMEM, "device memory budget (GRAL_PF_DEVICE_MB) exceeded");
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(h, GRAL_PF_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    h->dev.push_back(q);
    h->dev_bytes += bytes;
    *p = (T *)q;
    return 0;
}

extern "C" int gral_pf_create(const gral_pf_grid *g, gral_pf_handle **out) {
    if (!out) return GRAL_PF_EINVAL;
    *out = nullptr;
    if (!g || !g->dzk || !g->hokart) return GRAL_PF_EINVAL;
    if (g->nii < 5 || g->njj < 5 || g->nkk < 3 || g->kadvmax < 2 || g->kadvmax > g->nkk || g->vertcellsff < 1 ||
        !(g->dxk > 0) || !(g->dyk > 0))
        return GRAL_PF_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return GRAL_PF_EDEVICE;
    gral_pf_handle *h = new (std::nothrow) gral_pf_handle();
    if (!h) return GRAL_PF_ENOMEM;
    h->nii = g->nii; h->njj = g->njj; h->nkk = g->nkk; h->kadvmax = g->kadvmax; h->vcf = g->vertcellsff;
    h->dxk = g->dxk; h->dyk = g->dyk;
    h->nkh = g->nkk + 2;
    h->nkp = ((g->kadvmax + 2) + NKP_ALIGN - 1) & ~(NKP_ALIGN - 1);
    h->dzk.assign(g->dzk, g->dzk + h->nkh);
    h->hok.assign(g->hokart, g->hokart + h->nkh);
    // optional per-handle budget of device memory in MiB (prepare fails with
    // GRAL_PF_ENOMEM beyond it), e.g. to keep room for other situations
    if (const char *cap = std::getenv("GRAL_PF_DEVICE_MB")) h->dev_cap = (size_t)std::strtoull(cap, nullptr, 10) << 20;
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return GRAL_PF_EDEVICE;
    }
    // the allocation-free mass-imbalance entry point finds its scratch here;
    // a failure only disables that entry point (it reports it)
    (void)scratch_setup(h->device);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
        gral_pf_destroy(h);
        return GRAL_PF_EDEVICE;
    }
    *out = h;
    return GRAL_PF_OK;
}

extern "C" int gral_pf_set_masks(gral_pf_handle *h, const uint8_t *advdom, const int16_t *kkart, const float *cutk,
                                 const float *cov, const float *ahk, const float *veg, const uint8_t *veg_present) {
    if (!h) return GRAL_PF_EINVAL;
    h->prepared = false;
    if (!advdom || !kkart || !cutk || !cov || !ahk) return fail(h, GRAL_PF_EINVAL, "set_masks: null mask array");
    if ((veg == nullptr) != (veg_present == nullptr))
        return fail(h, GRAL_PF_EINVAL, "set_masks: veg and veg_present must both be given or both be NULL");
    size_t n2 = h->n2();
    h->adv.assign(advdom, advdom + n2);
    h->kk.assign(kkart, kkart + n2);
    h->cutk.assign(cutk, cutk + n2);
    h->cov.assign(cov, cov + n2);
    h->ahk.assign(ahk, ahk + n2);
    for (size_t c = 0; c < n2; ++c) {
        if (h->adv[c] > 1) return fail(h, GRAL_PF_EINVAL, "set_masks: ADVDOM must be 0 or 1");
        if (h->kk[c] < 0 || h->kk[c] + 1 > h->nkk - 1) return fail(h, GRAL_PF_EINVAL, "set_masks: KKART out of range");
    }
    h->has_veg = veg != nullptr;
    if (h->has_veg) {
        h->veg.assign(veg, veg + n2 * (size_t)h->nkh);
        h->vegp.assign(veg_present, veg_present + n2);
    } else {
        h->veg.clear();
        h->vegp.clear();
    }
    h->masks = true;
    h->prepared = false;
    return GRAL_PF_OK;
}

extern "C" int gral_pf_set_roughness(gral_pf_handle *h, const float *z0_wall, const float *z0_solver, const float *ob_init,
                                     const float *ust_init, const float *z0_init) {
    if (!h) return GRAL_PF_EINVAL;
    h->prepared = false;
    if (!z0_wall || !z0_solver || !ob_init || !ust_init || !z0_init) return fail(h, GRAL_PF_EINVAL, "set_roughness: null array");
    size_t n2 = h->n2();
    h->z0w.assign(z0_wall, z0_wall + n2);
    h->z0s.assign(z0_solver, z0_solver + n2);
    h->obi.assign(ob_init, ob_init + n2);
    h->usti.assign(ust_init, ust_init + n2);
    h->z0i.assign(z0_init, z0_init + n2);
    h->rough = true;
    h->prepared = false;
    return GRAL_PF_OK;
}

extern "C" int gral_pf_set_met_profile(gral_pf_handle *h, int32_t istatistics, int32_t n_profile,
                                       const float *meas_height, const float *u0, const float *v0) {
    if (!h) return GRAL_PF_EINVAL;
    h->prepared = false;
    if (istatistics < 0 || istatistics > 4) return fail(h, GRAL_PF_EINVAL, "set_met_profile: IStatistics must be 0..4");
    if (istatistics == 0 || istatistics == 3) {
        if (n_profile < 1 || n_profile > 50) return fail(h, GRAL_PF_EINVAL, "set_met_profile: n_profile must be 1..50");
        if (!meas_height || !u0 || !v0) return fail(h, GRAL_PF_EINVAL, "set_met_profile: null profile array");
        for (int q = 2; q <= n_profile; ++q)
            if (!(meas_height[q] > meas_height[q - 1]))
                return fail(h, GRAL_PF_EINVAL, "set_met_profile: measurement heights must increase");
        h->mh.assign(meas_height, meas_height + n_profile + 1);
        h->u0.assign(u0, u0 + n_profile + 1);
        h->v0.assign(v0, v0 + n_profile + 1);
        h->nprof = n_profile;
    } else {
        h->mh.clear(); h->u0.clear(); h->v0.clear();
        h->nprof = 0;
    }
    h->istat = istatistics;
    return GRAL_PF_OK;
}

extern "C" int gral_pf_upload_fields(gral_pf_handle *h, const float *uk, const float *vk, const float *wk) {
    if (!h) return GRAL_PF_EINVAL;
    h->prepared = false;
    if (!uk || !vk || !wk) return fail(h, GRAL_PF_EINVAL, "upload_fields: null field");
    size_t n3 = h->n2() * (size_t)h->nkh;
    par_assign(h->uk, uk, n3);
    par_assign(h->vk, vk, n3);
    par_assign(h->wk, wk, n3);
    h->fields = true;
    h->prepared = false;
    h->gff_ok = false;
    return GRAL_PF_OK;
}

static void free_device(gral_pf_handle *h) {
    for (void *p : h->dev) (void)hipFree(p);
    h->dev.clear();
    h->dev_bytes = 0;
    h->dd = nullptr;
    h->d = Dom{};
    h->mass_buf = nullptr;
    if (h->graphs) h->graphs->clear();
    h->adj = nullptr;
    h->adj_src = nullptr;
    h->adj_n = 0;
    h->mom_comp = -1;
    h->mom_vjp_ok = false;
}

// ---------------------------------------------------------------------------
// prologue, PrognosticFlowfield.cs:145-579
// ---------------------------------------------------------------------------
static int prologue(gral_pf_handle *h, const gral_pf_params *p, std::vector<float> &ap0, std::vector<float> &ht,
                    std::vector<float> &ho, std::vector<double> &ta, std::vector<double> &tden, std::vector<int> &tbr,
                    float &dtime, double &frund, float &ug, float &vg, float &ufx) {
    const int NII = h->nii, NJJ = h->njj, NKK = h->nkk;
    const int model = p->turbulence_model;
    const size_t n3 = h->n2() * (size_t)h->nkh;
    const float *DZK = h->dzk.data();
    auto &UK = h->uk, &VK = h->vk, &WK = h->wk;
    par_fill(h->uks, n3, 0.0f); par_fill(h->vks, n3, 0.0f); par_fill(h->wks, n3, 0.0f);
    par_fill(h->dpm, n3, 0.0f); par_fill(h->dpmnew, n3, 0.0f);
    if (model != 1) { par_fill(h->turb, n3, 0.0f); par_fill(h->tdiss, n3, 0.0f); }
    else { h->turb.clear(); h->tdiss.clear(); }
    int ri = p->refp_i, rj = p->refp_j;
    float topwind = std::sqrt(p2f(UK[h->i3(ri, rj, NKK - 1)]) + p2f(VK[h->i3(ri, rj, NKK - 1)]));
    topwind = fmax_net(topwind, 0.01f);
    float spd = std::sqrt(p2f(UK[h->i3(ri, rj, h->kadvmax)]) + p2f(VK[h->i3(ri, rj, h->kadvmax)]));
    dtime = 0.5f * fmin_net(h->dxk, DZK[1]) / fmax_net(0.01f, spd);
    dtime = fmin_net(dtime, 5.0f);
    const float area = h->dxk * h->dyk;
    ap0.assign((size_t)h->nkp, 0.0f);
    for (int k = 1; k <= h->kadvmax; ++k) ap0[k] = area * DZK[k] / dtime;
    if (model != 1)
        par_for(1, NII + 1, [&](int, long a, long b) {
            for (int i = (int)a; i < b; ++i)
                for (int j = 1; j <= NJJ; ++j)
                    for (int k = 1; k <= NKK + 1; ++k) { h->turb[h->i3(i, j, k)] = 0.001f; h->tdiss[h->i3(i, j, k)] = 0.0000001f; }
        });
    frund = (double)((model == 2 ? 10000.0f : 8000.0f) / topwind);
    // initial cell averages on the 3x3-dilated sub-domain footprint, and GESCHW_MAX
    auto A = [&](int i, int j) { return (int)h->adv[h->i2(i, j)]; };
    std::vector<double> gpart(host_threads(), 0.0);
    par_for(1, NII + 1, [&](int t, long a, long b) {
        double gmax = 0;
        for (int i = (int)a; i < b; ++i)
            for (int j = 1; j <= NJJ; ++j) {
                int ip = std::min(i + 1, NII + 1), jp = std::min(j + 1, NJJ + 1);
                bool dil = A(i, j) == 1 || A(i - 1, j) == 1 || A(ip, j) == 1 || A(i, j - 1) == 1 || A(i - 1, j - 1) == 1 ||
                           A(ip, j - 1) == 1 || A(i, jp) == 1 || A(i - 1, jp) == 1 || A(ip, jp) == 1;
                if (!dil) continue;
                const size_t o = h->i3(i, j, 0), oip = h->i3(i + 1, j, 0), ojp = h->i3(i, j + 1, 0);
                for (int k = 1; k <= NKK - 1; ++k) {
                    float u = 0.5f * (UK[o + k] + UK[oip + k]);
                    float v = 0.5f * (VK[o + k] + VK[ojp + k]);
                    h->uks[o + k] = u;
                    h->vks[o + k] = v;
                    h->wks[o + k] = 0.5f * (WK[o + k] + WK[o + k + 1]);
                    gmax = dmax_net(gmax, std::sqrt((double)(p2f(u) + p2f(v))));
                }
            }
        gpart[t] = gmax;
    });
    double gmax = 0;
    for (double g : gpart) gmax = dmax_net(gmax, g);
    h->geschw_max = gmax;
    h->vi.assign(h->n2(), 0);
    for (int i = 1; i <= NII; ++i)
        for (int j = 1; j <= NJJ; ++j) h->vi[h->i2(i, j)] = std::min(NKK - 1, h->kk[h->i2(i, j)] + h->vcf);
    // wall-function helpterms at k = KKART+1
    ht.assign(h->n2(), 0.0f); ho.assign(h->n2(), 0.0f);
    ta.assign(h->n2(), 0.0); tden.assign(h->n2(), 1.0); tbr.assign(h->n2(), 0);
    for (int i = 1; i <= NII; ++i)
        for (int j = 1; j <= NJJ; ++j) {
            size_t c = h->i2(i, j);
            int KK = h->kk[c];
            int KSTART = (h->cutk[c] == 0) ? KK + 1 : 1;
            int k = KK + 1;
            if (k < KSTART || k > h->vi[c]) continue;
            float rough = h->z0w[c];
            if (h->cutk[c] < 1) {
                if ((double)rough >= (double)DZK[k] * 0.1)
                    ht[c] = 0.4f / logf((DZK[k] + DZK[k + 1] * 0.5f) / DZK[k] * 0.1f);
                else
                    ht[c] = 0.4f / logf(DZK[k] * 0.5f / rough);
            } else {
                ho[c] = 0.4f / logf(DZK[k] * 0.5f / p->building_z0);
            }
            // TKE terrain wall function constants (TKE-prognostic-microscale.cs:291-306)
            double z0 = (double)h->z0s[c];
            if (z0 >= (double)DZK[k] * 0.5) {
                double hz = (double)DZK[k] + (double)DZK[k + 1] * 0.5;
                tbr[c] = 1; ta[c] = 0.4 / std::log(hz / z0); tden[c] = hz * 0.4;
            } else {
                tbr[c] = 0; ta[c] = 0.4 / std::log((double)DZK[k] * 0.5 / z0); tden[c] = (double)DZK[k] * 0.5 * 0.4;
            }
        }
    ug = h->uks[h->i3(ri, rj, NKK - 1)];
    vg = h->vks[h->i3(ri, rj, NKK - 1)];
    // TKE initialisation and the ghost-row copies done in the same loop
    par_for(1, NII + 1, [&](int, long ia, long ib) {
    for (int i = (int)ia; i < ib; ++i)
        for (int j = 1; j <= NJJ; ++j) {
            size_t c = h->i2(i, j);
            float ob = h->obi[c], ust = h->usti[c], rough = h->z0i[c];
            const float ground = h->ahk[c] - h->hok[h->kk[c]];
            for (int k = 1; k <= NKK - 1; ++k) {
                if (i == 1) UK[h->i3(NII + 1, j, k)] = UK[h->i3(NII, j, k)];
                if (j == 1) VK[h->i3(i, NJJ + 1, k)] = VK[h->i3(i, NJJ, k)];
                if (model == 1) continue;
                float V0int = p->v0_1, U0int;
                if (h->istat == 0 || h->istat == 3) {
                    // observed sigma_u, sigma_v over the measurement heights (IStatistics
                    // MeteoZR / MeteoSonic, PrognosticFlowfield.cs:456-481); V0int takes the
                    // slope of U0 between the two heights, as the reference does
                    const float zh = h->hok[k] - ground;
                    const float *MH = h->mh.data(), *U0 = h->u0.data(), *V0 = h->v0.data();
                    const int np = h->nprof;
                    if (zh <= MH[1]) { U0int = U0[1]; V0int = V0[1]; }
                    else if (zh >= MH[np]) { U0int = U0[np]; V0int = V0[np]; }
                    else {
                        int ipo = 1;
                        for (int q = 1; q <= np; ++q)
                            if (zh > MH[q]) ipo = q + 1;
                        const float slope = (U0[ipo] - U0[ipo - 1]) / (MH[ipo] - MH[ipo - 1]);
                        U0int = U0[ipo - 1] + slope * (zh - MH[ipo - 1]);
                        V0int = V0[ipo - 1] + slope * (zh - MH[ipo - 1]);
                    }
                } else {
                    double wh = dmax_net(std::sqrt((double)(p2f(UK[h->i3(i, j, k)]) + p2f(VK[h->i3(i, j, k)]))), 0.01);
                    U0int = (float)(wh * (0.2 * std::pow(wh, -0.9) + 0.32 * (double)rough + 0.18));
                    U0int = (float)dmax_net((double)U0int, 0.3) * p->std_dev_v;
                }
                float varw;
                if (ob < 0) varw = p2f(ust) * p2f(1.15f + 0.1f * powf(p->bdlayheight / (-ob), 0.67f)) * p2f(p->std_dev_w);
                else varw = p2f(ust * p->std_dev_w * 1.25f);
                h->turb[h->i3(i, j, k)] = (float)(0.5 * (double)(p2f(U0int) + p2f(V0int) + varw));
            }
        }
    });
    ufx = (float)(0.4 / std::log((double)h->dxk * 0.5 / (double)p->building_z0));
    // zero velocities (and k, eps) inside buildings
    par_for(1, NII + 1, [&](int, long ia, long ib) {
    for (int i = (int)ia; i < ib; ++i)
        for (int j = 1; j <= NJJ; ++j) {
            size_t c = h->i2(i, j);
            if (h->adv[c] != 1) continue;
            int VI = h->vi[c], KK = h->kk[c];
            for (int k = 1; k <= VI && k <= KK; ++k) {
                UK[h->i3(i, j, k)] = 0;
                if (i == NII) UK[h->i3(i + 1, j, k)] = 0;
                VK[h->i3(i, j, k)] = 0;
                if (j == NJJ) VK[h->i3(i, j + 1, k)] = 0;
                WK[h->i3(i, j, k)] = 0;
                if (k == VI) WK[h->i3(i, j, k + 1)] = 0;
                if (model != 1) { h->turb[h->i3(i, j, k)] = 0; h->tdiss[h->i3(i, j, k)] = 0; }
                h->uks[h->i3(i, j, k)] = 0; h->vks[h->i3(i, j, k)] = 0; h->wks[h->i3(i, j, k)] = 0;
            }
        }
    });
    h->topwind = topwind;
    return 0;
}

// blocks of the first stage of the mass-imbalance reduction: about 16 columns
// per block, at most MASS_MAX_BLOCKS; fixed by the grid, so repeated calls
// reduce in the same order
static int mass_blocks(int nii, int njj) {
    const long ncols = (long)(nii - 2) * (long)(njj - 2);
    return (int)std::max(1L, std::min<long>(kMassMaxBlocks, (ncols + 15) / 16));
}

// pinned host staging of one device field [n2][nkp], kept by the handle
static float *stage(gral_pf_handle *h) {
    const size_t n = h->n2() * (size_t)h->nkp * sizeof(float);
    if (h->stage_bytes < n) {
        if (h->stage_buf) (void)hipHostFree(h->stage_buf);
        h->stage_buf = nullptr;
        h->stage_bytes = 0;
        if (hipHostMalloc(&h->stage_buf, n, hipHostMallocDefault) != hipSuccess) { h->stage_buf = nullptr; return nullptr; }
        h->stage_bytes = n;
    }
    return static_cast<float *>(h->stage_buf);
}

// copy levels [0, nkp) of a host full-depth field into a device [.,.,nkp] field
static int put_raw(gral_pf_handle *h, float *dptr, const float *host);
static int put_field(gral_pf_handle *h, float *dptr, const std::vector<float> &host) { return put_raw(h, dptr, host.data()); }
static int put_raw(gral_pf_handle *h, float *dptr, const float *host) {
    const size_t n2 = h->n2(), n = n2 * (size_t)h->nkp;
    float *tmp = stage(h);
    if (!tmp) return fail(h, GRAL_PF_ENOMEM, "hipHostMalloc of the staging buffer failed");
    const int kc = std::min(h->nkp, h->nkh);
    par_for(0, (long)n2, [&](int, long a, long b) {
        for (size_t c = (size_t)a; c < (size_t)b; ++c) {
            std::memcpy(tmp + c * h->nkp, &host[c * h->nkh], sizeof(float) * kc);
            std::fill(tmp + c * h->nkp + kc, tmp + (c + 1) * h->nkp, 0.0f);
        }
    });
    HIPCHK(h, hipMemcpyAsync(dptr, tmp, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}
static int get_raw(gral_pf_handle *h, const float *dptr, float *host);
static int get_field(gral_pf_handle *h, const float *dptr, std::vector<float> &host) { return get_raw(h, dptr, host.data()); }
static int get_raw(gral_pf_handle *h, const float *dptr, float *host) {
    const size_t n2 = h->n2(), n = n2 * (size_t)h->nkp;
    float *tmp = stage(h);
    if (!tmp) return fail(h, GRAL_PF_ENOMEM, "hipHostMalloc of the staging buffer failed");
    HIPCHK(h, hipMemcpyAsync(tmp, dptr, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const int kc = std::min(h->nkp, h->nkh);
    par_for(0, (long)n2, [&](int, long a, long b) {
        for (size_t c = (size_t)a; c < (size_t)b; ++c) std::memcpy(&host[c * h->nkh], tmp + c * h->nkp, sizeof(float) * kc);
    });
    return 0;
}
template <class T>
static int put_vec(gral_pf_handle *h, T *dptr, const std::vector<T> &v) {
    HIPCHK(h, hipMemcpyAsync(dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int gral_pf_prepare(gral_pf_handle *h, const gral_pf_params *p) {
    if (!h) return GRAL_PF_EINVAL;
    // a failed prepare leaves the handle unprepared: the device state of an
    // earlier prepare is released below, so nothing may run on it afterwards
    h->prepared = false;
    h->done = false;
    h->host_stale = false;
    h->gff_ok = false;
    if (!p) return fail(h, GRAL_PF_EINVAL, "prepare: null params");
    if (!h->masks || !h->rough || !h->fields) return fail(h, GRAL_PF_ESTATE, "prepare: masks, roughness and fields must be set first");
    if (p->turbulence_model < 0 || p->turbulence_model > 2) return fail(h, GRAL_PF_EINVAL, "prepare: turbulence_model must be 0, 1 or 2");
    if (p->simd_width != 8 && p->simd_width != 16) return fail(h, GRAL_PF_EINVAL, "prepare: simd_width must be 8 or 16");
    if (h->vcf < p->simd_width) return fail(h, GRAL_PF_EINVAL, "prepare: vertcellsff must be >= simd_width");
    if (p->iter_max < 1 || p->iter_min < 0 || p->iter_max > 10000000) return fail(h, GRAL_PF_EINVAL, "prepare: bad iteration limits");
    if (p->refp_i < 1 || p->refp_i > h->nii || p->refp_j < 1 || p->refp_j > h->njj) return fail(h, GRAL_PF_EINVAL, "prepare: reference point outside the grid");
    // the solver reads KKART+1 <= VerticalIndex <= KADVMAX; check the grid agrees
    for (int i = 1; i <= h->nii; ++i)
        for (int j = 1; j <= h->njj; ++j) {
            size_t c = h->i2(i, j);
            int v = std::min(h->nkk - 1, h->kk[c] + h->vcf);
            if (h->adv[c] == 1 && v > h->kadvmax) return fail(h, GRAL_PF_EINVAL, "prepare: KKART+VertCellsFF exceeds KADVMAX in a sub-domain column");
        }
    h->prm = *p;
    std::vector<float> ap0, ht, ho;
    std::vector<double> ta, tden;
    std::vector<int> tbr;
    float dtime, ug, vg, ufx;
    double frund;
    prologue(h, p, ap0, ht, ho, ta, tden, tbr, dtime, frund, ug, vg, ufx);

    free_device(h);
    Dom &d = h->d;
    d = Dom{};
    d.nii = h->nii; d.njj = h->njj; d.nkk = h->nkk; d.nkp = h->nkp; d.nj2 = h->njj + 2; d.kadvmax = h->kadvmax;
    d.model = p->turbulence_model; d.simd = p->simd_width;
    d.dxk = h->dxk; d.dyk = h->dyk; d.area = h->dxk * h->dyk; d.dtime = dtime; d.relax = p->relax; d.relaxp = p->relaxp;
    d.cor = p->coriolis; d.ug = ug; d.vg = vg;
    d.vishmin = (p->turbulence_model == 2) ? 0.01f : 0.0f;
    d.visvmin = (p->turbulence_model == 2) ? 0.01f : 0.0f;
    d.bz0 = p->building_z0; d.ufx = ufx; d.potdtdz = p->potdtdz; d.frund = frund;
    const size_t n2 = h->n2(), n3 = n2 * (size_t)h->nkp;
    int rc = 0;
    float *fdzk, *fhok, *fap0, *fapp, *fabp, *fatp, *fcutk, *fcov, *fht, *fho, *fz0s;
    uint8_t *fadv, *fvegp = nullptr;
    int *fkk, *fvi, *ftbr;
    double *fta, *ftden;
    if ((rc = dalloc(h, &fdzk, h->nkh)) || (rc = dalloc(h, &fhok, h->nkh)) || (rc = dalloc(h, &fap0, h->nkp)) ||
        (rc = dalloc(h, &fapp, h->nkp)) || (rc = dalloc(h, &fabp, h->nkp)) || (rc = dalloc(h, &fatp, h->nkp)) ||
        (rc = dalloc(h, &fadv, n2)) || (rc = dalloc(h, &fkk, n2)) || (rc = dalloc(h, &fvi, n2)) ||
        (rc = dalloc(h, &fcutk, n2)) || (rc = dalloc(h, &fcov, n2)) || (rc = dalloc(h, &fht, n2)) ||
        (rc = dalloc(h, &fho, n2)) || (rc = dalloc(h, &fz0s, n2)) || (rc = dalloc(h, &fta, n2)) ||
        (rc = dalloc(h, &ftden, n2)) || (rc = dalloc(h, &ftbr, n2)))
        return rc;
    float *F[12];
    for (int f = 0; f < 12; ++f)
        if ((rc = dalloc(h, &F[f], n3))) return rc;
    d.uk = F[0]; d.vk = F[1]; d.wk = F[2]; d.uks = F[3]; d.vks = F[4]; d.wks = F[5]; d.dpm = F[6]; d.dpmnew = F[7];
    d.div = F[8]; d.turb = F[9]; d.tdiss = F[10];
    float *fveg = F[11];
    if ((rc = dalloc(h, &d.prow_p, (size_t)d.nj2 * d.nkp)) || (rc = dalloc(h, &d.prow_q, (size_t)d.nj2 * d.nkp))) return rc;
    unsigned *fres;
    if ((rc = dalloc(h, &fres, (size_t)p->iter_max + 1))) return rc;
    d.resid = fres;
    d.resid_cap = p->iter_max + 1;
    h->mass_nblk = mass_blocks(h->nii, h->njj);
    h->mass[0] = h->mass[1] = 0;
    // per-level pressure coefficients (PrognosticFlowfield.cs:855-866), column independent
    std::vector<float> app(h->nkp, 0.0f), abp(h->nkp, 0.0f), atp(h->nkp, 0.0f);
    const float *DZK = h->dzk.data();
    float area = d.area;
    for (int k = 1; k <= h->kadvmax && k + 1 < h->nkh; ++k) {
        app[k] = 2.0f * (h->dyk * DZK[k] / h->dxk + h->dxk * DZK[k] / h->dyk + area / DZK[k]);
        abp[k] = (k > 1) ? area / (0.5f * (DZK[k - 1] + DZK[k])) : area / DZK[k];
        atp[k] = area / (0.5f * (DZK[k + 1] + DZK[k]));
    }
    h->app = app; h->abp = abp; h->atp = atp;
    // sweep kernels: strip-pipelined (default) or one launch per anti-diagonal
    // (GRAL_PF_SWEEP = strip | diag)
    const char *sw = std::getenv("GRAL_PF_SWEEP");
    d.tiled = 2;
    if (sw && std::strcmp(sw, "diag") == 0) d.tiled = 0;
    if (h->kadvmax + 2 > 128) d.tiled = 0;   // the strip kernels hold at most 128 levels per column
    {
        // strip synchronisation: header + one block per sweep kind (pressure, U, V, W)
        const int rows = STRIP_ROWS;
        int nstrip = (std::max(h->nii, h->njj) + rows - 1) / rows + 1;
        d.sync_stride = ((2 + nstrip) + 63) & ~63;
        int *fs;
        if ((rc = dalloc(h, &fs, (size_t)SYNC_BLOCK0 + 5 * (size_t)d.sync_stride))) return rc;
        HIPCHK(h, hipMemsetAsync(fs, 0, sizeof(int) * ((size_t)SYNC_BLOCK0 + 5 * (size_t)d.sync_stride), h->stream));
        d.sync = fs;
    }
    d.planes = nullptr;
    d.rec3 = nullptr;
    d.wcol = nullptr;
    d.ptab[0] = d.ptab[1] = d.ttab[0] = d.ttab[1] = nullptr;
    if (d.tiled) {
        if (p->turbulence_model == 2) {
            // the momentum sweeps' records (12 or 8 floats per cell, then
            // {C, denominator, P} at 12..14) and, in turn, the TKE sweep's 20
            // floats per cell share the planes; the U/V wall terms follow
            float *pl;
            if ((rc = dalloc(h, &pl, 20 * n3 + n2))) return rc;
            d.planes = pl;
            d.rec3 = pl + 12 * n3;
            d.wcol = pl + 20 * n3;
        }
        if (p->turbulence_model != 2) {          // mixing length and no diffusion
            float *pl;
            if ((rc = dalloc(h, &pl, 15 * n3 + n2))) return rc;
            d.planes = pl;                       // 12 (U, V) or 8 (W) floats per cell
            d.rec3 = pl + 12 * n3;               // then 3 per cell
            d.wcol = pl + 15 * n3;               // then 1 per column
        }
        // PIMP/TERMP of the pressure TDMA for every column, in the reference's
        // row order (fresh row arrays per j row, columns in i1 order for IS=+1 and
        // IS=-1), so the stale top value PIMP[VI+1] is part of the table
        std::vector<float> tp, tt;
        for (int v = 0; v < 2; ++v) {
            par_fill(tp, n3, 0.0f);
            par_fill(tt, n3, 0.0f);
            // rows are independent: each has its own row arrays
            par_for(2, h->njj, [&](int, long ja, long jb) {
                std::vector<float> prow((size_t)h->nkh + 1, 0.0f);
                for (int j = (int)ja; j < jb; ++j) {
                    std::fill(prow.begin(), prow.end(), 0.0f);
                    for (int i1 = 2; i1 <= h->nii - 1; ++i1) {
                        int i = (v == 1) ? h->nii - i1 + 1 : i1;
                        size_t c = h->i2(i, j);
                        if (h->adv[c] != 1) continue;
                        int KK = h->kk[c], VI = h->vi[c];
                        size_t o = c * (size_t)h->nkp;
                        for (int k = VI; k > KK; --k) {
                            float TERMP = 1 / (app[k] - atp[k] * prow[k + 1]);
                            float P = abp[k] * TERMP;
                            prow[k] = P;
                            tp[o + k] = P;
                            tt[o + k] = TERMP;
                        }
                    }
                }
            });
            float *dp, *dt;
            if ((rc = dalloc(h, &dp, n3)) || (rc = dalloc(h, &dt, n3)) || (rc = put_vec(h, dp, tp)) || (rc = put_vec(h, dt, tt)))
                return rc;
            d.ptab[v] = dp;
            d.ttab[v] = dt;
        }
    }
    // per-column metadata of the tiled kernels: KK, VI, KSTART | adv<<16 | vegp<<17,
    // max(KK, KK[i-1]) | max(KK, KK[j-1])<<16
    std::vector<int4> meta(n2, make_int4(0, -1, 0, 0));
    for (int i = 1; i <= h->nii; ++i)
        for (int j = 1; j <= h->njj; ++j) {
            size_t c = h->i2(i, j);
            int KK = h->kk[c];
            int KSTART = (h->cutk[c] == 0) ? KK + 1 : 1;
            int vegp = (h->has_veg && h->vegp[c]) ? 1 : 0;
            int kmi = std::max(KK, (int)h->kk[h->i2(i - 1, j)]), kmj = std::max(KK, (int)h->kk[h->i2(i, j - 1)]);
            meta[c] = make_int4(KK, h->vi[c], KSTART | ((int)h->adv[c] << 16) | (vegp << 17), kmi | (kmj << 16));
        }
    int4 *fmeta;
    if ((rc = dalloc(h, &fmeta, n2)) || (rc = put_vec(h, fmeta, meta))) return rc;
    d.meta = fmeta;
    std::vector<int> kk32(n2);
    for (size_t c = 0; c < n2; ++c) kk32[c] = h->kk[c];
    if ((rc = put_vec(h, fdzk, h->dzk)) || (rc = put_vec(h, fhok, h->hok)) || (rc = put_vec(h, fap0, ap0)) ||
        (rc = put_vec(h, fapp, app)) || (rc = put_vec(h, fabp, abp)) || (rc = put_vec(h, fatp, atp)) ||
        (rc = put_vec(h, fadv, h->adv)) || (rc = put_vec(h, fkk, kk32)) || (rc = put_vec(h, fvi, h->vi)) ||
        (rc = put_vec(h, fcutk, h->cutk)) || (rc = put_vec(h, fcov, h->cov)) || (rc = put_vec(h, fht, ht)) ||
        (rc = put_vec(h, fho, ho)) || (rc = put_vec(h, fz0s, h->z0s)) || (rc = put_vec(h, fta, ta)) ||
        (rc = put_vec(h, ftden, tden)) || (rc = put_vec(h, ftbr, tbr)))
        return rc;
    d.dzk = fdzk; d.hok = fhok; d.ap0 = fap0; d.app = fapp; d.abp = fabp; d.atp = fatp;
    d.adv = fadv; d.kk = fkk; d.vi = fvi; d.cutk = fcutk; d.cov = fcov; d.ht = fht; d.ho = fho; d.z0s = fz0s;
    d.tke_a = fta; d.tke_den = ftden; d.tke_br = ftbr;
    if (h->has_veg) {
        if ((rc = dalloc(h, &fvegp, n2)) || (rc = put_vec(h, fvegp, h->vegp)) || (rc = put_field(h, fveg, h->veg))) return rc;
        d.vegp = fvegp; d.veg = fveg;
    } else {
        d.vegp = nullptr; d.veg = nullptr;
    }
    std::vector<float> zeros;
    if ((rc = put_field(h, d.uk, h->uk)) || (rc = put_field(h, d.vk, h->vk)) || (rc = put_field(h, d.wk, h->wk)) ||
        (rc = put_field(h, d.uks, h->uks)) || (rc = put_field(h, d.vks, h->vks)) || (rc = put_field(h, d.wks, h->wks)) ||
        (rc = put_field(h, d.dpm, h->dpm)) || (rc = put_field(h, d.dpmnew, h->dpmnew)))
        return rc;
    HIPCHK(h, hipMemsetAsync(d.div, 0, n3 * sizeof(float), h->stream));
    if (p->turbulence_model != 1) {
        if ((rc = put_field(h, d.turb, h->turb)) || (rc = put_field(h, d.tdiss, h->tdiss))) return rc;
    } else {
        HIPCHK(h, hipMemsetAsync(d.turb, 0, n3 * sizeof(float), h->stream));
        HIPCHK(h, hipMemsetAsync(d.tdiss, 0, n3 * sizeof(float), h->stream));
    }
    if (pf_fill_u32(d.resid, (size_t)p->iter_max + 1, 0u, h->stream)) return fail(h, GRAL_PF_EDEVICE, "fill resid");
    if ((rc = dalloc(h, &h->dd, 1))) return rc;
    HIPCHK(h, hipMemcpyAsync(h->dd, &h->d, sizeof(Dom), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->active = 0;
    for (int i = 2; i <= h->nii - 1; ++i)
        for (int j = 2; j <= h->njj - 1; ++j) {
            size_t c = h->i2(i, j);
            if (h->adv[c] == 1) h->active += h->vi[c] - h->kk[c];
        }
    static std::atomic<unsigned long long> next_gen{1};
    h->gen = next_gen++;
    h->prepared = true;
    h->gff_ok = false;
    h->done = false;
    h->host_stale = false;
    h->iter_loops = 1;
    h->iters_run = 0;
    h->processed = 0;
    h->delta_mean = 0;
    h->last_delta = 0;
    h->delta_first = 0;
    h->delta_finish = 0;
    h->unreal[0] = h->unreal[1] = h->unreal[2] = 0;
    h->timing = gral_pf_timing{};
    return GRAL_PF_OK;
}

static void fill_stats(gral_pf_handle *h, gral_pf_stats *st) {
    if (!st) return;
    st->iteration_loops = h->iter_loops;
    st->iterations_run = h->iters_run;
    st->delta_first = h->delta_first;
    st->delta_finish = h->delta_finish;
    st->convergence_problem = h->delta_finish > h->delta_first;
    st->unreal_u = h->unreal[0]; st->unreal_v = h->unreal[1]; st->unreal_w = h->unreal[2];
    st->topwind = h->topwind; st->dtime = h->d.dtime; st->ug = h->d.ug; st->vg = h->d.vg;
    st->frund = h->d.frund; st->geschw_max = h->geschw_max;
    st->last_delta = h->last_delta;
    st->mass_div = h->mass[0];
    st->mass_flux = h->mass[1];
}

// FP64 sums of |DIV| and of the absolute face fluxes over the active cells of
// the device-resident fields; partial is scratch of 2*mass_nblk doubles, out
// of 2 (device memory)
static int mass_run(gral_pf_handle *h, double *partial, double *out) {
    if (pf_launch_mass(h->dd, 1, h->mass_nblk, partial, out, h->stream))
        return fail(h, GRAL_PF_EDEVICE, "mass imbalance: kernel launch failed");
    HIPCHK(h, hipMemcpyAsync(h->mass, out, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

// the handle's scratch of the reduction, allocated on first use
static int mass_handle_run(gral_pf_handle *h) {
    int rc;
    if (!h->mass_buf && (rc = dalloc(h, &h->mass_buf, 2 * (size_t)h->mass_nblk + 2))) return rc;
    return mass_run(h, h->mass_buf, h->mass_buf + 2 * (size_t)h->mass_nblk);
}

// The same reduction with the file-scope device scratch instead of memory
// the handle allocates: no hipMalloc, hipFree or allocating call on any path.
// Nothing in the scratch needs to start at zero: stage 1 writes every partial
// stage 2 reads.
extern "C" int gral_pf_mass_imbalance_noalloc(gral_pf_handle *h, double *sum_abs_div, double *sum_abs_flux) {
    if (!h) return GRAL_PF_EINVAL;
    if (!h->prepared) return fail(h, GRAL_PF_ESTATE, "mass_imbalance_noalloc: call prepare first");
    const size_t need = 2 * (size_t)h->mass_nblk + 2;
    if (need > kMassScratchDoubles)
        return fail(h, GRAL_PF_EINVAL, "mass_imbalance_noalloc: " + std::to_string(need) + " doubles of scratch needed, " +
                                           std::to_string(kMassScratchDoubles) + " available");
    double *p = scratch_of(h->device);
    if (!p) return fail(h, GRAL_PF_EDEVICE, "mass_imbalance_noalloc: device address of the scratch array unavailable");
    int rc = mass_run(h, p, p + 2 * (size_t)h->mass_nblk);
    if (rc) return rc;
    if (sum_abs_div) *sum_abs_div = h->mass[0];
    if (sum_abs_flux) *sum_abs_flux = h->mass[1];
    return GRAL_PF_OK;
}

extern "C" int gral_pf_mass_imbalance(gral_pf_handle *h, double *sum_abs_div, double *sum_abs_flux) {
    if (!h) return GRAL_PF_EINVAL;
    if (!h->prepared) return fail(h, GRAL_PF_ESTATE, "mass_imbalance: call prepare first");
    int rc = mass_handle_run(h);
    if (rc) return rc;
    if (sum_abs_div) *sum_abs_div = h->mass[0];
    if (sum_abs_flux) *sum_abs_flux = h->mass[1];
    return GRAL_PF_OK;
}

// Convergence bookkeeping of one finished iteration (PrognosticFlowfield.cs:1144-1178).
// Returns true when the reference would leave the loop here.
// iterations launched between two host synchronisations
static const int kMaxChunk = 20;

static bool book_iteration(gral_pf_handle *h, int it, float dumax, bool check) {
    const gral_pf_params &p = h->prm;
    double tw2 = (double)(h->topwind * h->topwind);
    if (it >= p.iter_max) h->delta_finish = (float)(h->delta_mean / tw2);
    if ((it % 100 == 0) && (it > 1)) {
        double delta = h->delta_mean / tw2;
        if (it == 100) h->delta_first = (float)delta;
        h->last_delta = delta;
        if (check && delta < 0.012 && it >= p.iter_min) return true;
        h->delta_mean = 0;
    } else {
        h->delta_mean += (double)dumax;
    }
    return false;
}

// Lock-step driver shared by single and batched solves.  Runs up to n SIMPLE
// iterations of the situations in hs that are not finished; all of them must
// stand at the same iteration.  Between host synchronisations it launches up
// to kMaxChunk iterations, never past a multiple of 100 when the stop test
// applies and never past any situation's iter_max (the size of its residual
// history); after each chunk it books the residuals of every situation and
// drops the finished ones (early retirement).  dd holds room for hs.size()
// Dom entries in device memory.
// Extra streams of a batch: its situations split into groups that run their
// iterations on streams of their own, so that one group's latency-bound
// sweeps overlap the bandwidth-bound passes of another.  The groups never
// share a buffer; stream s (the batch's own) forks them per chunk and joins
// them before the chunk's host synchronisation.
static const int kMaxGroups = 4;
struct StreamGroup {
    hipStream_t s[kMaxGroups - 1] = {};
    hipEvent_t join[kMaxGroups - 1] = {};
    hipEvent_t fork = nullptr;
    int n = 0;                               // extra streams created
    // online choice of the group count: for the set of candidate counts a batch
    // of its size allows (tuned_for: how many), one timed chunk of
    // kTrialIters iterations (one cycle of the four sweep directions) per
    // candidate, then the fastest until that set changes
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int tuned_for = -1, best = 0, trial = 0;
    float trial_ms[kMaxGroups + 1] = {};
    void clear() {
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        t0 = t1 = nullptr;
        tuned_for = -1;
        best = trial = 0;
        for (int g = 0; g < n; ++g) {
            if (s[g]) (void)hipStreamSynchronize(s[g]);
            if (join[g]) (void)hipEventDestroy(join[g]);
            if (s[g]) (void)hipStreamDestroy(s[g]);
            s[g] = nullptr; join[g] = nullptr;
        }
        if (fork) (void)hipEventDestroy(fork);
        fork = nullptr;
        n = 0;
    }
    // at least k extra streams (k < kMaxGroups); false if creation failed
    bool ensure(int k) {
        if (!fork && hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess) return false;
        while (n < k) {
            if (hipStreamCreateWithFlags(&s[n], hipStreamNonBlocking) != hipSuccess) return false;
            if (hipEventCreateWithFlags(&join[n], hipEventDisableTiming) != hipSuccess) {
                (void)hipStreamDestroy(s[n]);
                s[n] = nullptr;
                return false;
            }
            ++n;
        }
        return true;
    }
};

// groups a lock-step batch of n situations may run in: at most n / 2 and 4
// (one stream per hardware queue).  Measured on the 8-direction city batch:
// 24.30 ms per iteration in one group, 23.28 in two, 23.09 in four.
static int max_groups(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>(n / 2, (size_t)kMaxGroups)); }
// GRAL_PF_GROUPS (1 to 4) fixes the count; 0 if unset
static int forced_groups() {
    const char *e = std::getenv("GRAL_PF_GROUPS");
    if (!e) return 0;
    return std::max(1, std::min(std::atoi(e), kMaxGroups));
}
// iterations of one trial of the group tuner: a whole cycle of the sweep
// directions (iteration mod 4), whose sweeps differ in cost
static const int kTrialIters = 4;
// candidate counts the tuner tries for n situations: 1, 2, 4 up to max_groups(n)
static int group_candidates(size_t n, int *c) {
    int k = 0;
    for (int g = 1; g <= max_groups(n); g *= 2) c[k++] = g;
    return k;
}

struct RunResult {
    long long launches = 0, cell_iters = 0;
    float ms = 0;
    int groups = 1;                          // stream groups of the last chunk
    std::string err;
    int code = 0;
};


static void run_lockstep(const std::vector<gral_pf_handle *> &hs, Dom *dd, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                         int n, bool check, Probe *probe, RunResult &R, GraphCache *gc = nullptr,
                         StreamGroup *sg = nullptr) {
    auto bad = [&](int code, const std::string &m) { R.code = code; R.err = m; };
    if (probe || !graphs_enabled()) gc = nullptr;
    if (probe || gc) sg = nullptr;                  // profiling and graph capture run on the one stream
    std::vector<gral_pf_handle *> act;
    for (auto *h : hs)
        if (!h->done) act.push_back(h);
    if (hipEventRecord(e0, s) != hipSuccess) return bad(GRAL_PF_EDEVICE, "hipEventRecord failed");
    std::vector<Dom> hd;
    std::vector<unsigned> bits;
    bool upload = true;
    int ng = 1, goff[kMaxGroups] = {0}, gcnt[kMaxGroups] = {0};
    int remaining = n;
    while (remaining > 0 && !act.empty()) {
        // the launches run at act[0]'s iteration; the others stand at the same
        // iteration or one a multiple of 100 earlier (admitted later), so all
        // share the sweep directions and reach the stop test together
        const int it0 = act[0]->iter_loops;
        int chunk = std::min(remaining, kMaxChunk);
        if (check) chunk = std::min(chunk, 100 - ((it0 - 1) % 100));
        for (auto *h : act) {
            if ((it0 - h->iter_loops) % 100 != 0)
                return bad(GRAL_PF_ESTATE, "iterate: situations of a batch must stand at iterations a multiple of 100 apart");
            chunk = std::min(chunk, h->prm.iter_max - h->iter_loops + 1);
        }
        if (it0 < 1 || chunk < 1) return bad(GRAL_PF_ESTATE, "iterate: iteration counter out of range");
        // the group count: forced, tuned, or on trial for this many situations
        int want = 1;
        bool trial = false;
        if (sg) {
            int cand[kMaxGroups + 1];
            const int nc = group_candidates(act.size(), cand);
            if (const int f = forced_groups()) {
                want = std::min(f, max_groups(act.size()));
            } else if (nc == 1) {
                want = 1;
            } else {
                if (sg->tuned_for != nc) {
                    sg->tuned_for = nc;
                    sg->best = sg->trial = 0;
                }
                if (sg->best) {
                    want = sg->best;
                } else if (chunk >= kTrialIters) {       // the next candidate's timed cycle
                    want = cand[sg->trial];
                    trial = true;
                    chunk = kTrialIters;
                    if ((!sg->t0 && hipEventCreate(&sg->t0) != hipSuccess) || (!sg->t1 && hipEventCreate(&sg->t1) != hipSuccess))
                        return bad(GRAL_PF_EDEVICE, "hipEventCreate failed");
                } else {                                  // too short for a trial: the largest count
                    want = cand[nc - 1];
                }
            }
        }
        if (want != ng) upload = true;
        // slots: the launch's residual slot minus iter_off is each situation's
        // own iteration - 1; with graphs the slot counts from the chunk's start
        if (upload || gc) {
            // group g takes situations g, g+G, g+2G, ... of act (act is ordered
            // largest first by the callers that care), stored contiguously
            ng = want;
            if (ng > 1 && !sg->ensure(ng - 1)) return bad(GRAL_PF_EDEVICE, "stream creation failed");
            hd.resize(act.size());
            size_t pos = 0;
            for (int g = 0; g < ng; ++g) {
                goff[g] = (int)pos;
                for (size_t q = (size_t)g; q < act.size(); q += (size_t)ng) {
                    hd[pos] = act[q]->d;
                    hd[pos].iter_off = gc ? 1 - act[q]->iter_loops : it0 - act[q]->iter_loops;
                    ++pos;
                }
                gcnt[g] = (int)pos - goff[g];
            }
            if (hipMemcpyAsync(dd, hd.data(), sizeof(Dom) * hd.size(), hipMemcpyHostToDevice, s) != hipSuccess)
                return bad(GRAL_PF_EDEVICE, "upload of the situation table failed");
            upload = false;
        }
        auto launch_chunk = [&](long long *launches) -> int {
            if (ng > 1) {                                 // fork the group streams off s
                if (hipEventRecord(sg->fork, s) != hipSuccess) return bad(GRAL_PF_EDEVICE, "hipEventRecord failed"), 1;
                for (int g = 1; g < ng; ++g)
                    if (hipStreamWaitEvent(sg->s[g - 1], sg->fork, 0) != hipSuccess)
                        return bad(GRAL_PF_EDEVICE, "hipStreamWaitEvent failed"), 1;
            }
            for (int c = 0; c < chunk; ++c) {
                const int it = it0 + c;
                for (int g = 0; g < ng; ++g) {
                    hipStream_t sgs = g == 0 ? s : sg->s[g - 1];
                    int e = pf_launch_iteration(hd.data() + goff[g], gcnt[g], dd + goff[g], it, gc ? c : it - 1, sgs,
                                                launches, probe);
                    if (e) {
                        bad(GRAL_PF_EDEVICE, std::string("kernel launch failed at iteration ") + std::to_string(it) + ": " +
                                                 hipGetErrorString((hipError_t)e));
                        return e;
                    }
                }
            }
            for (int g = 1; g < ng; ++g)                  // and join them back
                if (hipEventRecord(sg->join[g - 1], sg->s[g - 1]) != hipSuccess ||
                    hipStreamWaitEvent(s, sg->join[g - 1], 0) != hipSuccess)
                    return bad(GRAL_PF_EDEVICE, "stream join failed"), 1;
            return 0;
        };
        std::vector<unsigned long long> key;
        if (gc)
            for (auto *h : act) key.push_back(h->gen);
        R.groups = ng;
        if (trial && hipEventRecord(sg->t0, s) != hipSuccess) return bad(GRAL_PF_EDEVICE, "hipEventRecord failed");
        if (gc && gc->exec && gc->set == key && gc->dd == dd && gc->chunk == chunk && gc->it_mod4 == it0 % 4) {
            if (hipGraphLaunch(gc->exec, s) != hipSuccess) return bad(GRAL_PF_EDEVICE, "hipGraphLaunch failed");
            R.launches += gc->launches;
        } else {
            if (launch_chunk(&R.launches)) return;
            if (gc) {
                // capture the same chunk for the next ones (capturing does not run it)
                gc->clear();
                long long nl = 0;
                if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
                    return bad(GRAL_PF_EDEVICE, "hipStreamBeginCapture failed");
                const int e = launch_chunk(&nl);
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(s, &g);
                if (e) return;
                if (ec != hipSuccess || !g) return bad(GRAL_PF_EDEVICE, "hipStreamEndCapture failed");
                gc->graph = g;
                if (hipGraphInstantiate(&gc->exec, g, nullptr, nullptr, 0) != hipSuccess) {
                    gc->clear();
                    return bad(GRAL_PF_EDEVICE, "hipGraphInstantiate failed");
                }
                gc->set = key;
                gc->dd = dd;
                gc->chunk = chunk;
                gc->it_mod4 = it0 % 4;
                gc->launches = nl;
            }
        }
        if (trial) {                                      // the trial's time; after the last, the fastest
            float ms = 0;
            if (hipEventRecord(sg->t1, s) != hipSuccess || hipEventSynchronize(sg->t1) != hipSuccess ||
                hipEventElapsedTime(&ms, sg->t0, sg->t1) != hipSuccess)
                return bad(GRAL_PF_EDEVICE, "event timing failed");
            int cand[kMaxGroups + 1];
            const int nc = group_candidates(act.size(), cand);
            sg->trial_ms[sg->trial] = ms;
            if (++sg->trial == nc) {
                // the largest count within 2 % of the fastest: overlap is the
                // rule, the trials guard against batches it slows down (many
                // situations of unlike sizes); closer timings are within the
                // spread of one cycle
                float fastest = sg->trial_ms[0];
                for (int q = 1; q < nc; ++q) fastest = std::min(fastest, sg->trial_ms[q]);
                int b = 0;
                for (int q = 0; q < nc; ++q)
                    if (sg->trial_ms[q] <= 1.02f * fastest) b = q;
                sg->best = cand[b];
            }
        }
        remaining -= chunk;
        if (hipStreamSynchronize(s) != hipSuccess) return bad(GRAL_PF_EDEVICE, "stream synchronisation failed");
        hipError_t le = hipGetLastError();
        if (le != hipSuccess)
            return bad(GRAL_PF_EDEVICE, std::string("pending HIP error after iterations ") + std::to_string(it0) + ".." +
                                            std::to_string(it0 + chunk - 1) + ": " + hipGetErrorString(le));
        // bookkeeping of the iterations just run, situation by situation
        std::vector<gral_pf_handle *> keep;
        bits.resize((size_t)chunk);
        for (auto *h : act) {
            int serr = 0;
            const int hit0 = h->iter_loops;              // this situation's first iteration of the chunk
            if (hipMemcpy(bits.data(), h->d.resid + (hit0 - 1), sizeof(unsigned) * chunk, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(&serr, h->d.sync, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
                return bad(GRAL_PF_EDEVICE, "residual download failed");
            if (serr)
                return bad(GRAL_PF_EDEVICE, "strip sweep: a wait for the previous strip timed out (results invalid)");
            for (int c = 0; c < chunk; ++c) {
                float du;
                std::memcpy(&du, &bits[c], 4);
                int it = hit0 + c;
                h->iters_run++;
                R.cell_iters += h->active;
                if (book_iteration(h, it, du, check)) {
                    h->iter_loops = it;      // the reference breaks before IterationLoops++
                    h->done = true;
                    break;
                }
                h->iter_loops = it + 1;
            }
            if (h->iter_loops > h->prm.iter_max) h->done = true;
            h->host_stale = true;
            h->mom_vjp_ok = false;
            if (!h->done) keep.push_back(h);
        }
        if (keep.size() != act.size()) {
            act.swap(keep);
            upload = true;
        }
    }
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&R.ms, e0, e1) != hipSuccess)
        return bad(GRAL_PF_EDEVICE, "event timing failed");
}

extern "C" int gral_pf_iterate(gral_pf_handle *h, int32_t n, int32_t check, gral_pf_stats *st) {
    if (!h) return GRAL_PF_EINVAL;
    if (!h->prepared) return fail(h, GRAL_PF_ESTATE, "iterate: call prepare first");
    if (n < 0) return fail(h, GRAL_PF_EINVAL, "iterate: n < 0");
    h->timing = gral_pf_timing{};
    RunResult R;
    if (!h->graphs) h->graphs = new (std::nothrow) GraphCache();
    run_lockstep({h}, h->dd, h->stream, h->ev0, h->ev1, n, check != 0, nullptr, R, h->graphs);
    if (R.code) return fail(h, R.code, R.err);
    h->timing.ms_total = R.ms;
    h->timing.launches = R.launches;
    h->timing.active_cells = h->active;
    h->timing.cell_iterations = R.cell_iters;
    h->timing.situations = 1;
    h->timing.groups = 1;
    fill_stats(h, st);
    return GRAL_PF_OK;
}

static int sync_host(gral_pf_handle *h) {
    if (!h->host_stale) return 0;
    int rc;
    if ((rc = get_field(h, h->d.uk, h->uk)) || (rc = get_field(h, h->d.vk, h->vk)) || (rc = get_field(h, h->d.wk, h->wk)) ||
        (rc = get_field(h, h->d.uks, h->uks)) || (rc = get_field(h, h->d.vks, h->vks)) ||
        (rc = get_field(h, h->d.wks, h->wks)) || (rc = get_field(h, h->d.dpm, h->dpm)) ||
        (rc = get_field(h, h->d.dpmnew, h->dpmnew)))
        return rc;
    if (h->prm.turbulence_model != 1)
        if ((rc = get_field(h, h->d.turb, h->turb)) || (rc = get_field(h, h->d.tdiss, h->tdiss))) return rc;
    // VKS edge-row copies above the device levels (PrognosticFlowfield.cs:1019-1039)
    if (h->iters_run > 0) {
        for (int j = 2; j < h->njj; ++j) {
            if (h->adv[h->i2(1, j)] == 1)
                for (int k = h->nkp; k <= h->nkk; ++k) h->vks[h->i3(1, j, k)] = h->vks[h->i3(2, j, k)];
            if (h->adv[h->i2(h->nii - 1, j)] == 1)
                for (int k = h->nkp; k <= h->nkk; ++k) h->vks[h->i3(h->nii, j, k)] = h->vks[h->i3(h->nii - 1, j, k)];
        }
    }
    h->host_stale = false;
    return 0;
}

extern "C" int gral_pf_finish(gral_pf_handle *h, gral_pf_stats *st) {
    if (!h) return GRAL_PF_EINVAL;
    if (!h->prepared) return fail(h, GRAL_PF_ESTATE, "finish: call prepare first");
    int rc = mass_handle_run(h);
    if (rc) return rc;
    rc = sync_host(h);
    if (rc) return rc;
    // unrealistic wind speeds (PrognosticFlowfield.cs:1192-1243)
    float limit = (float)(10 * h->geschw_max);
    std::vector<int> cnt(3 * (size_t)host_threads(), 0);
    par_for(1, h->nii + 1, [&](int t, long a, long b) {
        int nu = 0, nv = 0, nw = 0;
        for (int i = (int)a; i < b; ++i)
            for (int j = 1; j <= h->njj; ++j) {
                size_t c = h->i2(i, j);
                if (h->adv[c] != 1) continue;
                int kmax = std::min(h->nkk - 1, h->kk[c] + h->vcf);
                for (int k = 1; k <= kmax; ++k) {
                    float u = std::fabs(h->uk[h->i3(i, j, k)]), v = std::fabs(h->vk[h->i3(i, j, k)]), w = std::fabs(h->wk[h->i3(i, j, k)]);
                    if (u > 55 || u > limit) ++nu;
                    if (v > 55 || v > limit) ++nv;
                    if (w > 20 || w > limit) ++nw;
                }
            }
        cnt[3 * t] = nu; cnt[3 * t + 1] = nv; cnt[3 * t + 2] = nw;
    });
    h->unreal[0] = h->unreal[1] = h->unreal[2] = 0;
    for (size_t t = 0; t < cnt.size(); ++t) h->unreal[t % 3] += cnt[t];
    fill_stats(h, st);
    return GRAL_PF_OK;
}

extern "C" int gral_pf_solve(gral_pf_handle *h, const gral_pf_params *p, gral_pf_stats *st) {
    int rc = gral_pf_prepare(h, p);
    if (rc) return rc;
    rc = gral_pf_iterate(h, p->iter_max, 1, st);
    if (rc) return rc;
    return gral_pf_finish(h, st);
}

// ---------------------------------------------------------------------------
// Adjoint of the pressure-correction steps (pf_adjoint.hip)
// ---------------------------------------------------------------------------
// P, TERMP and the stale-top source of every column's pressure elimination in
// the reference's row order for IS = +1 (v = 0) and -1 (v = 1): the row
// scratch starts at zero for every row and each column leaves its P, Q at its
// levels KK+1..VI; the next column starts from whatever level VI+1 holds.
static void pressure_tables(gral_pf_handle *h, int v, std::vector<float> &tp, std::vector<float> &tt,
                            std::vector<int> &src) {
    const size_t n3 = h->n2() * (size_t)h->nkp;
    tp.assign(n3, 0.0f);
    tt.assign(n3, 0.0f);
    src.assign(h->n2(), -1);
    par_for(2, h->njj, [&](int, long ja, long jb) {
        std::vector<float> prow((size_t)h->nkh + 1, 0.0f);
        std::vector<int> writer((size_t)h->nkh + 1, -1);
        for (int j = (int)ja; j < jb; ++j) {
            std::fill(prow.begin(), prow.end(), 0.0f);
            std::fill(writer.begin(), writer.end(), -1);
            for (int i1 = 2; i1 <= h->nii - 1; ++i1) {
                const int i = (v == 1) ? h->nii - i1 + 1 : i1;
                const size_t c = h->i2(i, j);
                if (h->adv[c] != 1) continue;
                const int KK = h->kk[c], VI = h->vi[c];
                const size_t o = c * (size_t)h->nkp;
                src[c] = writer[VI + 1];
                for (int k = VI; k > KK; --k) {
                    const float TERMP = 1 / (h->app[k] - h->atp[k] * prow[k + 1]);
                    prow[k] = h->abp[k] * TERMP;
                    tp[o + k] = prow[k];
                    tt[o + k] = TERMP;
                    writer[k] = i;
                }
            }
        }
    });
}

static int adj_alloc(gral_pf_handle *h) {
    if (h->adj) return 0;
    const size_t n2 = h->n2(), n3 = n2 * (size_t)h->nkp;
    float *a;
    int *src;
    int rc;
    if ((rc = dalloc(h, &a, 15 * n3)) || (rc = dalloc(h, &src, 2 * n2))) return rc;
    std::vector<float> tp, tt;
    std::vector<int> sv;
    for (int v = 0; v < 2; ++v) {
        pressure_tables(h, v, tp, tt, sv);
        HIPCHK(h, hipMemcpyAsync(a + (size_t)(10 + v) * n3, tp.data(), n3 * sizeof(float), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(a + (size_t)(12 + v) * n3, tt.data(), n3 * sizeof(float), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(src + (size_t)v * n2, sv.data(), n2 * sizeof(int), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    h->adj = a;
    h->adj_src = src;
    return 0;
}

extern "C" int gral_pf_forward_ckpt(gral_pf_handle *h, int32_t n, int32_t quantise) {
    if (!h) return GRAL_PF_EINVAL;
    if (!h->prepared) return fail(h, GRAL_PF_ESTATE, "forward_ckpt: call prepare first");
    if (n < 1 || n > 100000) return fail(h, GRAL_PF_EINVAL, "forward_ckpt: n must be 1..100000");
    if (quantise != 0 && quantise != 1) return fail(h, GRAL_PF_EINVAL, "forward_ckpt: quantise must be 0 or 1");
    int rc = adj_alloc(h);
    if (rc) return rc;
    const size_t n3 = h->n2() * (size_t)h->nkp;
    float *ck = h->adj + 7 * n3;
    HIPCHK(h, hipMemcpyAsync(ck, h->d.uk, n3 * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(ck + n3, h->d.vk, n3 * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(ck + 2 * n3, h->d.wk, n3 * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    h->d.noquant = quantise ? 0 : 1;
    HIPCHK(h, hipMemcpyAsync(h->dd, &h->d, sizeof(Dom), hipMemcpyHostToDevice, h->stream));
    long long launches = 0;
    int e = 0;
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    for (int q = 0; q < n && !e; ++q) e = pf_launch_projection(&h->d, h->dd, h->iter_loops + q, h->stream, &launches);
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    h->d.noquant = 0;
    HIPCHK(h, hipMemcpyAsync(h->dd, &h->d, sizeof(Dom), hipMemcpyHostToDevice, h->stream));
    if (e) return fail(h, GRAL_PF_EDEVICE, std::string("forward_ckpt: launch failed: ") + hipGetErrorString((hipError_t)e));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int serr = 0;
    HIPCHK(h, hipMemcpy(&serr, h->d.sync, sizeof(int), hipMemcpyDeviceToHost));
    if (serr) return fail(h, GRAL_PF_EDEVICE, "forward_ckpt: a strip wait timed out (results invalid)");
    float ms = 0;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->timing = gral_pf_timing{};
    h->timing.ms_total = ms;
    h->timing.launches = launches;
    h->timing.active_cells = h->active;
    h->timing.cell_iterations = h->active * (long long)n;
    h->timing.situations = 1;
    h->timing.groups = 1;
    h->adj_first = h->iter_loops;
    h->adj_n = n;
    h->mom_comp = -1;
    h->mom_vjp_ok = false;
    h->host_stale = true;
    h->gff_ok = false;
    return GRAL_PF_OK;
}

extern "C" int gral_pf_restore_ckpt(gral_pf_handle *h) {
    if (!h) return GRAL_PF_EINVAL;
    if (!h->prepared || (h->adj_n < 1 && h->mom_comp < 0))
        return fail(h, GRAL_PF_ESTATE, "restore_ckpt: no forward_ckpt or momentum_sweep since prepare");
    const size_t n3 = h->n2() * (size_t)h->nkp;
    const float *ck = h->adj + 7 * n3;
    HIPCHK(h, hipMemcpyAsync(h->d.uk, ck, n3 * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d.vk, ck + n3, n3 * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d.wk, ck + 2 * n3, n3 * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->host_stale = true;
    h->gff_ok = false;
    h->mom_vjp_ok = false;
    return GRAL_PF_OK;
}

extern "C" int gral_pf_vjp(gral_pf_handle *h, const float *du, const float *dv, const float *dw, float *gu, float *gv,
                           float *gw) {
    if (!h) return GRAL_PF_EINVAL;
    if (!h->prepared || h->adj_n < 1) return fail(h, GRAL_PF_ESTATE, "vjp: call forward_ckpt first");
    if (!du || !dv || !dw || !gu || !gv || !gw) return fail(h, GRAL_PF_EINVAL, "vjp: null argument");
    const size_t n2 = h->n2(), n3 = n2 * (size_t)h->nkp;
    float *a = h->adj;
    AdjArgs g{};
    g.ub = a; g.vb = a + n3; g.wb = a + 2 * n3; g.pb = a + 3 * n3; g.qb = a + 4 * n3; g.mb = a + 5 * n3; g.db = a + 6 * n3;
    int rc;
    if ((rc = put_raw(h, g.ub, du)) || (rc = put_raw(h, g.vb, dv)) || (rc = put_raw(h, g.wb, dw))) return rc;
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    for (int q = h->adj_n - 1; q >= 0; --q) {
        const int it = h->adj_first + q;
        int IS = 1;
        switch (it % 4) {
            case 0: case 1: IS = -1; break;
            default: break;
        }
        const int v = IS == -1 ? 1 : 0;
        g.ptab = a + (size_t)(10 + v) * n3;
        g.ttab = a + (size_t)(12 + v) * n3;
        g.src = h->adj_src + (size_t)v * n2;
        int e = pf_launch_adjoint_step(&h->d, h->dd, &g, it, n3, h->stream);
        if (e) return fail(h, GRAL_PF_EDEVICE, std::string("vjp: launch failed: ") + hipGetErrorString((hipError_t)e));
    }
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    HIPCHK(h, hipEventSynchronize(h->ev1));
    float ms = 0;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->timing = gral_pf_timing{};
    h->timing.ms_total = ms;
    h->timing.active_cells = h->active;
    h->timing.cell_iterations = h->active * (long long)h->adj_n;
    h->timing.situations = 1;
    h->timing.groups = 1;
    // levels the device does not hold pass through unchanged (the steps never touch them)
    const size_t nh = n2 * (size_t)h->nkh;
    par_copy(gu, du, nh * sizeof(float));
    par_copy(gv, dv, nh * sizeof(float));
    par_copy(gw, dw, nh * sizeof(float));
    if ((rc = get_raw(h, g.ub, gu)) || (rc = get_raw(h, g.vb, gv)) || (rc = get_raw(h, g.wb, gw))) return rc;
    return GRAL_PF_OK;
}

extern "C" int gral_pf_set_velocity(gral_pf_handle *h, const float *uk, const float *vk, const float *wk) {
    if (!h) return GRAL_PF_EINVAL;
    if (!h->prepared) return fail(h, GRAL_PF_ESTATE, "set_velocity: call prepare first");
    if (!uk || !vk || !wk) return fail(h, GRAL_PF_EINVAL, "set_velocity: null field");
    int rc = sync_host(h);                                // the other fields' host copies stay current
    if (rc) return rc;
    const size_t n3 = h->n2() * (size_t)h->nkh;
    par_assign(h->uk, uk, n3);
    par_assign(h->vk, vk, n3);
    par_assign(h->wk, wk, n3);
    h->mom_vjp_ok = false;
    if ((rc = put_field(h, h->d.uk, h->uk)) || (rc = put_field(h, h->d.vk, h->vk)) || (rc = put_field(h, h->d.wk, h->wk)))
        return rc;
    h->gff_ok = false;
    return GRAL_PF_OK;
}

// sweep directions of momentum component comp in iteration it (PrognosticFlowfield.cs:1058-1101):
// U and W sweep (IS, JS), the mixing-length V solver (-IS, -JS)
static void momentum_dirs(int it, int comp, int &IS, int &JS) {
    IS = 1; JS = 1;
    switch (it % 4) {
        case 0: IS = -1; JS = 1; break;
        case 1: IS = -1; JS = -1; break;
        case 2: IS = 1; JS = -1; break;
        default: break;
    }
    if (comp == 1) { IS = -IS; JS = -JS; }
}

extern "C" int gral_pf_momentum_sweep(gral_pf_handle *h, int32_t comp) {
    if (!h) return GRAL_PF_EINVAL;
    if (!h->prepared) return fail(h, GRAL_PF_ESTATE, "momentum_sweep: call prepare first");
    if (comp < 0 || comp > 2) return fail(h, GRAL_PF_EINVAL, "momentum_sweep: comp must be 0 (U), 1 (V) or 2 (W)");
    if (h->d.model != 1) return fail(h, GRAL_PF_EINVAL, "momentum_sweep: the mixing-length model (turbulence model 1) only");
    if (h->d.kadvmax + 2 > 128) return fail(h, GRAL_PF_EINVAL, "momentum_sweep: columns above 126 levels");
    int rc = adj_alloc(h);
    if (rc) return rc;
    const size_t n3 = h->n2() * (size_t)h->nkp;
    float *ck = h->adj + 7 * n3;
    HIPCHK(h, hipMemcpyAsync(ck, h->d.uk, n3 * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(ck + n3, h->d.vk, n3 * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(ck + 2 * n3, h->d.wk, n3 * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    int IS, JS;
    momentum_dirs(h->iter_loops, comp, IS, JS);
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    const int e = pf_launch_momentum_sweep(&h->d, h->dd, comp, IS, JS, h->stream);
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    if (e) return fail(h, GRAL_PF_EDEVICE, std::string("momentum_sweep: launch failed: ") + hipGetErrorString((hipError_t)e));
    HIPCHK(h, hipEventSynchronize(h->ev1));
    float ms = 0;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->timing = gral_pf_timing{};
    h->timing.ms_total = ms;
    h->timing.active_cells = h->active;
    h->timing.situations = 1;
    h->timing.groups = 1;
    h->mom_comp = comp;
    h->mom_is = IS;
    h->mom_js = JS;
    h->mom_vjp_ok = true;
    h->adj_n = 0;
    h->host_stale = true;
    h->gff_ok = false;
    return GRAL_PF_OK;
}

extern "C" int gral_pf_momentum_vjp(gral_pf_handle *h, const float *dx, float *gx) {
    if (!h) return GRAL_PF_EINVAL;
    if (!h->prepared || h->mom_comp < 0 || !h->mom_vjp_ok)
        return fail(h, GRAL_PF_ESTATE, "momentum_vjp: no momentum_sweep since the device fields last changed");
    if (!dx || !gx) return fail(h, GRAL_PF_EINVAL, "momentum_vjp: null argument");
    const size_t n2 = h->n2(), n3 = n2 * (size_t)h->nkp;
    float *a = h->adj;
    const int comp = h->mom_comp;
    // the situation's descriptor with the swept field at its value before the sweep (the checkpoint)
    Dom hb = h->d;
    float *xb = a + (size_t)(7 + comp) * n3;
    float *xa = comp == 0 ? h->d.uk : comp == 1 ? h->d.vk : h->d.wk;
    if (comp == 0) hb.uk = xb; else if (comp == 1) hb.vk = xb; else hb.wk = xb;
    Dom *db = reinterpret_cast<Dom *>(a + 14 * n3);      // the spare field of the adjoint buffers
    HIPCHK(h, hipMemcpyAsync(db, &hb, sizeof(Dom), hipMemcpyHostToDevice, h->stream));
    MomAdj m{};
    m.xa = xa;
    m.gin = a;
    m.gout = a + n3; m.gpa = a + 2 * n3; m.gpb = a + 3 * n3; m.gna = a + 4 * n3; m.gnb = a + 5 * n3; m.gres = a + 6 * n3;
    int rc;
    if ((rc = put_raw(h, a, dx))) return rc;
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    const int e = pf_launch_momentum_vjp(&hb, db, &m, comp, h->mom_is, h->mom_js, h->stream);
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    if (e) return fail(h, GRAL_PF_EDEVICE, std::string("momentum_vjp: launch failed: ") + hipGetErrorString((hipError_t)e));
    HIPCHK(h, hipEventSynchronize(h->ev1));
    float ms = 0;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->timing = gral_pf_timing{};
    h->timing.ms_total = ms;
    h->timing.active_cells = h->active;
    h->timing.situations = 1;
    h->timing.groups = 1;
    // levels the device does not hold pass through unchanged (the sweep never touches them)
    par_copy(gx, dx, n2 * (size_t)h->nkh * sizeof(float));
    return get_raw(h, m.gres, gx);
}

extern "C" int gral_pf_download_fields(gral_pf_handle *h, float *uk, float *vk, float *wk, float *uks, float *vks,
                                       float *wks, float *dpm, float *dpmnew, float *turb, float *tdiss) {
    if (!h) return GRAL_PF_EINVAL;
    if (!h->prepared && !h->fields) return fail(h, GRAL_PF_ESTATE, "download: upload fields first");
    int rc = sync_host(h);
    if (rc) return rc;
    size_t n3 = h->n2() * (size_t)h->nkh;
    auto cp = [&](float *dst, const std::vector<float> &src) {
        if (!dst) return;
        if (src.size() == n3) par_copy(dst, src.data(), n3 * sizeof(float));
        else std::memset(dst, 0, n3 * sizeof(float));
    };
    cp(uk, h->uk); cp(vk, h->vk); cp(wk, h->wk); cp(uks, h->uks); cp(vks, h->vks); cp(wks, h->wks);
    cp(dpm, h->dpm); cp(dpmnew, h->dpmnew); cp(turb, h->turb); cp(tdiss, h->tdiss);
    return GRAL_PF_OK;
}

// Runs n iterations one at a time with HIP events around every launch and
// sums the device time per kernel family.
static void profile_group(const std::vector<gral_pf_handle *> &hs, Dom *dd, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                          int n, gral_pf_family_time *out, int nfam, RunResult &R) {
    for (int f = 0; f < nfam; ++f) out[f] = gral_pf_family_time{};
    int nmax = 0;
    for (auto *h : hs) nmax = std::max(nmax, h->nii + h->njj);
    const int cap = 8 * nmax + 64;
    std::vector<hipEvent_t> ev(2 * (size_t)cap, nullptr);
    std::vector<int> fam(cap);
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) { R.code = GRAL_PF_EDEVICE; R.err = "hipEventCreate failed"; break; }
    for (int r = 0; r < n && !R.code; ++r) {
        Probe pr{ev.data(), fam.data(), cap, 0};
        RunResult one;
        run_lockstep(hs, dd, s, e0, e1, 1, false, &pr, one);
        if (one.code) { R.code = one.code; R.err = one.err; break; }
        for (int q = 0; q < pr.n; ++q) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev[2 * q], ev[2 * q + 1]);
            if (fam[q] < nfam) { out[fam[q]].ms += ms; out[fam[q]].launches += 1; }
        }
        for (int f = 0; f < nfam; ++f) out[f].cells += one.cell_iters;
        R.launches += one.launches;
        R.cell_iters += one.cell_iters;
        R.ms += one.ms;
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
}

extern "C" int gral_pf_profile_iterations(gral_pf_handle *h, int32_t n, gral_pf_family_time *out, int32_t nfam) {
    if (!h || !out || nfam < 1 || n < 0) return GRAL_PF_EINVAL;
    if (!h->prepared) return fail(h, GRAL_PF_ESTATE, "profile: call prepare first");
    RunResult R;
    profile_group({h}, h->dd, h->stream, h->ev0, h->ev1, n, out, nfam, R);
    if (R.code) return fail(h, R.code, R.err);
    return GRAL_PF_OK;
}

// ------------------------------------------------------------------------
// Batches: several prepared situations advanced in lock step, every launch
// covering all of them (grid slice z = situation).
// ------------------------------------------------------------------------
struct gral_pf_batch {
    std::vector<gral_pf_handle *> hs;
    std::vector<gral_pf_handle *> pending;   // added, waiting for an admission point (gral_pf_batch_step)
    GraphCache graphs;                       // captured iteration chunks (GRAL_PF_GRAPH=1)
    StreamGroup groups;                      // streams of the situation groups after the first
    Dom *dd = nullptr;
    size_t dd_cap = 0;                       // Dom entries dd holds
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    gral_pf_timing timing{};
    std::string err;
};

static int bfail(gral_pf_batch *b, int code, const std::string &msg) {
    if (b) b->err = msg;
    return code;
}

extern "C" int gral_pf_batch_create(gral_pf_handle *const *handles, int32_t n, gral_pf_batch **out) {
    if (!out) return GRAL_PF_EINVAL;
    *out = nullptr;
    if (n < 0 || (n > 0 && !handles)) return GRAL_PF_EINVAL;
    for (int q = 0; q < n; ++q) {
        if (!handles[q] || !handles[q]->prepared) return GRAL_PF_ESTATE;
        for (int r = 0; r < q; ++r)
            if (handles[r] == handles[q]) return GRAL_PF_EINVAL;
        const Dom &a = handles[0]->d, &c = handles[q]->d;
        if (a.model != c.model || a.tiled != c.tiled) return GRAL_PF_EINVAL;
    }
    gral_pf_batch *b = new (std::nothrow) gral_pf_batch();
    if (!b) return GRAL_PF_ENOMEM;
    if (n > 0) b->hs.assign(handles, handles + n);
    const size_t cap = std::max<size_t>(1, (size_t)n);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
        hipEventCreate(&b->ev1) != hipSuccess || hipMalloc((void **)&b->dd, sizeof(Dom) * cap) != hipSuccess) {
        gral_pf_batch_destroy(b);
        return GRAL_PF_EDEVICE;
    }
    b->dd_cap = cap;
    *out = b;
    return GRAL_PF_OK;
}

// ---- a batch fed while it runs ------------------------------------------
// Situations join a running batch at an admission point: when no situation is
// active, or when the active ones stand right after a multiple of 100
// iterations (where the stop test is applied).  A newcomer starts at its
// iteration 1, which then lies a multiple of 100 iterations behind the
// others: the sweep directions (iteration mod 4) agree, and every situation
// meets its stop test at the same launch.
extern "C" int gral_pf_batch_add(gral_pf_batch *b, gral_pf_handle *h) {
    if (!b) return GRAL_PF_EINVAL;
    if (!h || !h->prepared) return bfail(b, GRAL_PF_ESTATE, "batch add: the situation must be prepared");
    if (h->iter_loops != 1 || h->done) return bfail(b, GRAL_PF_ESTATE, "batch add: the situation must stand at its first iteration");
    for (auto *x : b->hs)
        if (x == h) return bfail(b, GRAL_PF_EINVAL, "batch add: the situation is already in the batch");
    for (auto *x : b->pending)
        if (x == h) return bfail(b, GRAL_PF_EINVAL, "batch add: the situation is already queued");
    const gral_pf_handle *ref = !b->hs.empty() ? b->hs[0] : (!b->pending.empty() ? b->pending[0] : nullptr);
    if (ref && (ref->d.model != h->d.model || ref->d.tiled != h->d.tiled))
        return bfail(b, GRAL_PF_EINVAL, "batch add: situations of a batch must share the turbulence model and sweep kind");
    const size_t need = b->hs.size() + b->pending.size() + 1;
    if (need > b->dd_cap) {
        const size_t cap = std::max(need, 2 * b->dd_cap);
        Dom *nd = nullptr;
        if (hipMalloc((void **)&nd, sizeof(Dom) * cap) != hipSuccess) return bfail(b, GRAL_PF_ENOMEM, "batch add: hipMalloc failed");
        if (hipStreamSynchronize(b->stream) != hipSuccess) return bfail(b, GRAL_PF_EDEVICE, "batch add: stream failed");
        b->graphs.clear();
        (void)hipFree(b->dd);
        b->dd = nd;
        b->dd_cap = cap;
    }
    b->pending.push_back(h);
    return GRAL_PF_OK;
}

extern "C" int gral_pf_batch_step(gral_pf_batch *b, int32_t *n_done, gral_pf_handle **done, gral_pf_stats *st,
                                  int32_t cap) {
    if (!b || !n_done) return GRAL_PF_EINVAL;
    *n_done = 0;
    if (cap < 0 || (cap > 0 && !done)) return bfail(b, GRAL_PF_EINVAL, "batch step: bad output arrays");
    std::vector<gral_pf_handle *> act;
    for (auto *h : b->hs)
        if (!h->done) act.push_back(h);
    bool admit = act.empty();
    if (!admit) {
        admit = true;
        for (auto *h : act) admit = admit && (h->iter_loops % 100 == 1);
    }
    if (admit && !b->pending.empty()) {
        for (auto *h : b->pending) {
            if (!h->prepared || h->iter_loops != 1) return bfail(b, GRAL_PF_ESTATE, "batch step: a queued situation was changed");
            b->hs.push_back(h);
            act.push_back(h);
        }
        b->pending.clear();
    }
    if (act.empty()) return GRAL_PF_OK;
    // run to the next multiple of 100 of the leading situation, where
    // situations stop and newcomers may join
    int lead = act[0]->iter_loops;
    for (auto *h : act) lead = std::max(lead, h->iter_loops);
    const int n = 100 - ((lead - 1) % 100);
    RunResult R;
    std::vector<gral_pf_handle *> order(act);
    std::stable_sort(order.begin(), order.end(),
                     [](const gral_pf_handle *x, const gral_pf_handle *y) { return x->iter_loops > y->iter_loops; });
    run_lockstep(order, b->dd, b->stream, b->ev0, b->ev1, n, true, nullptr, R, &b->graphs, &b->groups);
    if (R.code) return bfail(b, R.code, R.err);
    b->timing = gral_pf_timing{};
    b->timing.ms_total = R.ms;
    b->timing.launches = R.launches;
    for (auto *h : act) b->timing.active_cells += h->active;
    b->timing.cell_iterations = R.cell_iters;
    b->timing.situations = (int32_t)act.size();
    b->timing.groups = R.groups;
    // finished situations: epilogue, then out of the batch
    std::vector<gral_pf_handle *> keep;
    for (auto *h : b->hs) {
        if (!h->done) { keep.push_back(h); continue; }
        if (*n_done >= cap) return bfail(b, GRAL_PF_EINVAL, "batch step: more finished situations than output room");
        int rc = gral_pf_finish(h, st ? st + *n_done : nullptr);
        if (rc) return bfail(b, rc, h->err);
        done[(*n_done)++] = h;
    }
    b->hs.swap(keep);
    return GRAL_PF_OK;
}

extern "C" int gral_pf_batch_size(const gral_pf_batch *b, int32_t *active, int32_t *queued) {
    if (!b) return GRAL_PF_EINVAL;
    int a = 0;
    for (auto *h : b->hs) a += h->done ? 0 : 1;
    if (active) *active = a;
    if (queued) *queued = (int32_t)b->pending.size();
    return GRAL_PF_OK;
}

static int batch_ready(gral_pf_batch *b) {
    for (auto *h : b->hs)
        if (!h->prepared) return bfail(b, GRAL_PF_ESTATE, "batch: a situation was not prepared");
    return 0;
}

static void batch_stats(gral_pf_batch *b, gral_pf_stats *st) {
    if (!st) return;
    for (size_t q = 0; q < b->hs.size(); ++q) fill_stats(b->hs[q], st + q);
}

extern "C" int gral_pf_batch_iterate(gral_pf_batch *b, int32_t n, int32_t check, gral_pf_stats *st) {
    if (!b) return GRAL_PF_EINVAL;
    if (n < 0) return bfail(b, GRAL_PF_EINVAL, "batch iterate: n < 0");
    int rc = batch_ready(b);
    if (rc) return rc;
    b->timing = gral_pf_timing{};
    RunResult R;
    run_lockstep(b->hs, b->dd, b->stream, b->ev0, b->ev1, n, check != 0, nullptr, R, &b->graphs, &b->groups);
    if (R.code) return bfail(b, R.code, R.err);
    b->timing.ms_total = R.ms;
    b->timing.launches = R.launches;
    for (auto *h : b->hs) b->timing.active_cells += h->active;
    b->timing.cell_iterations = R.cell_iters;
    b->timing.situations = (int32_t)b->hs.size();
    b->timing.groups = R.groups;
    batch_stats(b, st);
    return GRAL_PF_OK;
}

extern "C" int gral_pf_batch_solve(gral_pf_batch *b, gral_pf_stats *st) {
    if (!b) return GRAL_PF_EINVAL;
    int rc = batch_ready(b);
    if (rc) return rc;
    int imax = 0;
    for (auto *h : b->hs) imax = std::max(imax, h->prm.iter_max);
    if ((rc = gral_pf_batch_iterate(b, imax, 1, nullptr))) return rc;
    for (size_t q = 0; q < b->hs.size(); ++q)
        if ((rc = gral_pf_finish(b->hs[q], st ? st + q : nullptr))) return bfail(b, rc, b->hs[q]->err);
    return GRAL_PF_OK;
}

extern "C" int gral_pf_batch_profile_iterations(gral_pf_batch *b, int32_t n, gral_pf_family_time *out, int32_t nfam) {
    if (!b || !out || nfam < 1 || n < 0) return GRAL_PF_EINVAL;
    int rc = batch_ready(b);
    if (rc) return rc;
    RunResult R;
    profile_group(b->hs, b->dd, b->stream, b->ev0, b->ev1, n, out, nfam, R);
    if (R.code) return bfail(b, R.code, R.err);
    return GRAL_PF_OK;
}

extern "C" int gral_pf_batch_last_timing(gral_pf_batch *b, gral_pf_timing *t) {
    if (!b || !t) return GRAL_PF_EINVAL;
    *t = b->timing;
    return GRAL_PF_OK;
}

extern "C" const char *gral_pf_batch_last_error(const gral_pf_batch *b) {
    if (!b) return "null batch";
    return b->err.c_str();
}

extern "C" void gral_pf_batch_destroy(gral_pf_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->graphs.clear();
    b->groups.clear();
    if (b->dd) (void)hipFree(b->dd);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// DiagnosticFlowfield.Calculate (DiagnosticFlowfield.cs:24-436) and the final
// W mass adjustment of the caller (MicroscaleFlat.cs:388-462 or
// MicroscaleTerrain.cs:862-933) on the handle's current fields, full height.
// Temporary device arrays hold the full columns; afterwards the solver's own
// device fields are refreshed from the result.
extern "C" int gral_pf_diagnostic(gral_pf_handle *h, int32_t flow_field_level, int32_t terrain) {
    if (!h) return GRAL_PF_EINVAL;
    if (flow_field_level != 1 && flow_field_level != 2) return fail(h, GRAL_PF_EINVAL, "diagnostic: flow_field_level must be 1 or 2");
    if (!h->masks || !h->fields) return fail(h, GRAL_PF_ESTATE, "diagnostic: masks and fields must be set first");
    h->mom_vjp_ok = false;
    int rc = 0;
    if (h->prepared && (rc = sync_host(h))) return rc;
    const size_t n2 = h->n2(), nkh = (size_t)h->nkh, n3 = n2 * nkh;
    if (h->dpm.size() != n3) h->dpm.assign(n3, 0.0f);
    // DTIME = (float)(0.5*Math.Min(DXK,DYK)/Math.Sqrt(Pow2(UK)+Pow2(VK))) at (1,1,KADVMAX), :30-31
    float u = h->uk[h->i3(1, 1, h->kadvmax)], v = h->vk[h->i3(1, 1, h->kadvmax)];
    float s2 = u * u + v * v;
    float dt = (float)(0.5 * (double)fmin_net(h->dxk, h->dyk) / std::sqrt((double)s2));
    dt = fmin_net(dt, 5.0f);
    std::vector<float> app(nkh, 0.0f), atp(nkh, 0.0f), abi(nkh, 0.0f), abb(nkh, 0.0f);
    const float DXK = h->dxk, DYK = h->dyk;
    const float *DZK = h->dzk.data();
    for (int k = 1; k <= h->nkk - 1; ++k) {
        app[k] = 2 * (DYK * DZK[k] / DXK + DXK * DZK[k] / DYK + DXK * DYK / DZK[k]);
        abi[k] = DXK * DYK / (0.5f * (DZK[k - 1] + DZK[k]));
        abb[k] = DXK * DYK / DZK[k];
        atp[k] = (DXK * DYK) / (0.5f * (DZK[k + 1] + DZK[k]));
    }
    // TERMP and PIMP of the column recurrence (:124-127): every column runs it
    // from NKK-1 down with PIMP[NKK] = 0, so above a column's bottom level the
    // values are the same in every column; at the bottom level ABP differs.
    std::vector<float> term(nkh, 0.0f), pin(nkh, 0.0f), pbot(nkh, 0.0f);
    {
        float Pn = 0.0f;
        for (int k = h->nkk - 1; k >= 1; --k) {
            const float T = 1 / (app[k] - atp[k] * Pn);
            term[k] = T;
            pin[k] = abi[k] * T;
            pbot[k] = abb[k] * T;
            Pn = pin[k];
        }
    }
    std::vector<int> kk32(n2);
    for (size_t c = 0; c < n2; ++c) kk32[c] = h->kk[c];
    std::vector<void *> tmp;
    auto talloc = [&](void **p, size_t bytes) -> int {
        if (hipMalloc(p, bytes + 256) != hipSuccess) return fail(h, GRAL_PF_ENOMEM, "diagnostic: hipMalloc failed");
        tmp.push_back(*p);
        return 0;
    };
    auto cleanup = [&]() { for (void *p : tmp) (void)hipFree(p); };
    DiagArgs g{};
    g.nii = h->nii; g.njj = h->njj; g.nkk = h->nkk; g.nkh = (int)nkh; g.nj2 = h->njj + 2;
    g.level = flow_field_level; g.terrain = terrain ? 1 : 0;
    g.dxk = DXK; g.dyk = DYK; g.dtime = dt;
    void *p[13];
    size_t sz[13] = {nkh * 4, nkh * 4, nkh * 4, nkh * 4, nkh * 4, n2 * 4, n2, n3 * 4, n3 * 4, n3 * 4, n3 * 4, n3 * 4, 0};
    for (int q = 0; q < 12; ++q)
        if ((rc = talloc(&p[q], sz[q]))) { cleanup(); return rc; }
    const void *src[12] = {h->dzk.data(), term.data(), atp.data(), pin.data(), pbot.data(), kk32.data(), h->adv.data(),
                           h->uk.data(), h->vk.data(), h->wk.data(), h->dpm.data(), nullptr};
    for (int q = 0; q < 11; ++q)
        if (hipMemcpyAsync(p[q], src[q], sz[q], hipMemcpyHostToDevice, h->stream) != hipSuccess) {
            cleanup();
            return fail(h, GRAL_PF_EDEVICE, "diagnostic: upload failed");
        }
    if (hipMemsetAsync(p[11], 0, sz[11], h->stream) != hipSuccess) { cleanup(); return fail(h, GRAL_PF_EDEVICE, "diagnostic: memset failed"); }
    g.dzk = (const float *)p[0]; g.term = (const float *)p[1]; g.atp = (const float *)p[2];
    g.p_in = (const float *)p[3]; g.p_bot = (const float *)p[4];
    g.kk = (const int *)p[5]; g.adv = (const uint8_t *)p[6];
    g.uk = (float *)p[7]; g.vk = (float *)p[8]; g.wk = (float *)p[9]; g.dpm = (float *)p[10]; g.div = (float *)p[11];
    int iters = (flow_field_level == 2) ? 10 : 5;
    int e = pf_launch_diagnostic(&g, iters, h->stream);
    if (e || hipStreamSynchronize(h->stream) != hipSuccess) {
        cleanup();
        return fail(h, GRAL_PF_EDEVICE, "diagnostic: kernel launch failed");
    }
    float *dst[4] = {h->uk.data(), h->vk.data(), h->wk.data(), h->dpm.data()};
    for (int q = 0; q < 4; ++q)
        if (hipMemcpy(dst[q], p[7 + q], n3 * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            cleanup();
            return fail(h, GRAL_PF_EDEVICE, "diagnostic: download failed");
        }
    cleanup();
    h->diag_dtime = dt;
    h->host_stale = false;
    h->gff_ok = false;
    if (h->prepared) {
        // keep the solver's device state (levels below nkp) equal to the host's
        if ((rc = put_field(h, h->d.uk, h->uk)) || (rc = put_field(h, h->d.vk, h->vk)) || (rc = put_field(h, h->d.wk, h->wk)) ||
            (rc = put_field(h, h->d.dpm, h->dpm)))
            return rc;
    }
    return GRAL_PF_OK;
}

// ---------------------------------------------------------------------------
// Flow-field files, GFFWrite.cs: the int16 payload is packed on the device
// from the solver's resident levels (no download of the fields) and the host
// copy of the levels above them; the host adds the 32-byte header.
// ---------------------------------------------------------------------------
static size_t gff_columns(const gral_pf_handle *h, int kind) {
    return kind == 3 ? (size_t)h->nii * h->njj : (size_t)(h->nii + 1) * (size_t)(h->njj + 1);
}

static size_t gff_size(const gral_pf_handle *h, int kind) {
    const size_t nkk = (size_t)h->nkk;
    if (kind == 3) return 32 + gff_columns(h, 3) * nkk * 2;
    if (kind == 0 || kind == 1) return 32 + gff_columns(h, kind) * (nkk + 1) * 6;
    size_t n = 32 + (size_t)h->nii * h->njj * 2;
    for (int i = 1; i <= h->nii + 1; ++i)
        for (int j = 1; j <= h->njj + 1; ++j) n += 6 * (size_t)(h->nkk + 2 - std::max<int>(h->kk[h->i2(i, j)], 1));
    return n;
}

// device copy of the host's levels [nlo, nkh) of UK, VK, WK, TURB
static int gff_upper(gral_pf_handle *h, int nlo) {
    const int nup = h->nkh - nlo;
    if (h->gff_ok && h->gff_nlo == nlo) return 0;
    if (nup <= 0) { h->gff_ok = true; h->gff_nlo = nlo; return 0; }
    const size_t n2 = h->n2(), per = n2 * (size_t)nup, bytes = 4 * per * sizeof(float);
    if (bytes > h->gff_bytes) {
        if (h->gff_up) (void)hipFree(h->gff_up);
        h->gff_up = nullptr;
        h->gff_bytes = 0;
        if (hipMalloc(&h->gff_up, bytes) != hipSuccess) { h->gff_up = nullptr; return fail(h, GRAL_PF_ENOMEM, "gff: hipMalloc failed"); }
        h->gff_bytes = bytes;
    }
    std::vector<float> tmp(4 * per, 0.0f);
    const std::vector<float> *src[4] = {&h->uk, &h->vk, &h->wk, &h->turb};
    const size_t n3 = n2 * (size_t)h->nkh;
    for (int f = 0; f < 4; ++f) {
        if (src[f]->size() != n3) continue;
        for (size_t c = 0; c < n2; ++c)
            std::memcpy(&tmp[f * per + c * nup], src[f]->data() + c * h->nkh + nlo, sizeof(float) * nup);
    }
    HIPCHK(h, hipMemcpyAsync(h->gff_up, tmp.data(), bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->gff_ok = true;
    h->gff_nlo = nlo;
    return 0;
}

// Packs the payload of one file on the device: the 32-byte header into hdr,
// the int16 body into the handle's device buffer (h->gff_dbuf); returns its
// size in bytes through body.
static int gff_pack(gral_pf_handle *h, int kind, const gral_pf_gff_meta *meta, uint8_t *hdr, size_t &body) {
    const size_t total = gff_size(h, kind);
    const int nlo = h->prepared ? std::min(h->nkp, h->nkh) : 0;
    int rc = gff_upper(h, nlo);
    if (rc) return rc;
    // header: NKK (negative for modes 1 and 2), NJJ, NII, direction, speed, class, DXK, format flag
    const int32_t hn[3] = {(kind == 1 || kind == 2) ? -h->nkk : h->nkk, h->njj, h->nii};
    const int32_t flag = (kind == 1) ? -2 : (kind == 2) ? -3 : -1;
    std::memcpy(hdr, hn, 12);
    std::memcpy(hdr + 12, &meta->wind_dir, 4);
    std::memcpy(hdr + 16, &meta->wind_vel, 4);
    std::memcpy(hdr + 20, &meta->stab_class, 4);
    std::memcpy(hdr + 24, &h->dxk, 4);
    std::memcpy(hdr + 28, &flag, 4);
    // device buffer: body, KKART, kind-2 column offsets; kept by the handle
    body = total - 32;
    const size_t ncol = gff_columns(h, kind), n2 = h->n2();
    const size_t o_kk = (body + 255) & ~size_t(255), o_off = (o_kk + n2 * 2 + 255) & ~size_t(255);
    const size_t bytes = o_off + (kind == 2 ? ncol * sizeof(long long) : 0);
    if (bytes > h->gff_dbuf_bytes) {
        if (h->gff_dbuf) (void)hipFree(h->gff_dbuf);
        h->gff_dbuf = nullptr;
        h->gff_dbuf_bytes = 0;
        if (hipMalloc(&h->gff_dbuf, bytes) != hipSuccess) { h->gff_dbuf = nullptr; return fail(h, GRAL_PF_ENOMEM, "gff: hipMalloc failed"); }
        h->gff_dbuf_bytes = bytes;
    }
    char *db = static_cast<char *>(h->gff_dbuf);
    std::vector<long long> off;
    if (kind == 2) {
        off.resize(ncol);
        long long acc = 0;
        size_t q = 0;
        for (int i = 1; i <= h->nii + 1; ++i)
            for (int j = 1; j <= h->njj + 1; ++j) {
                off[q++] = acc;
                acc += 3LL * (h->nkk + 2 - std::max<int>(h->kk[h->i2(i, j)], 1));
            }
        HIPCHK(h, hipMemcpyAsync(db + o_off, off.data(), ncol * sizeof(long long), hipMemcpyHostToDevice, h->stream));
    }
    HIPCHK(h, hipMemcpyAsync(db + o_kk, h->kk.data(), n2 * 2, hipMemcpyHostToDevice, h->stream));
    GffArgs g{};
    const float *lo[4] = {h->d.uk, h->d.vk, h->d.wk, h->d.turb};
    const int nup = h->nkh - nlo;
    for (int f = 0; f < 4; ++f) {
        g.lo[f] = nlo > 0 ? lo[f] : nullptr;
        g.up[f] = nup > 0 ? static_cast<const float *>(h->gff_up) + (size_t)f * n2 * nup : nullptr;
    }
    g.nlo = nlo; g.nkp = h->nkp; g.nup = nup;
    g.nii = h->nii; g.njj = h->njj; g.nkk = h->nkk; g.nj2 = h->njj + 2;
    g.kind = kind;
    g.kk = reinterpret_cast<const int16_t *>(db + o_kk);
    g.off = kind == 2 ? reinterpret_cast<const long long *>(db + o_off) : nullptr;
    g.out = reinterpret_cast<int16_t *>(db);
    if (pf_launch_gff(&g, h->stream)) return fail(h, GRAL_PF_EDEVICE, "gff: kernel launch failed");
    // the upload sources (off, h->kk) must stay alive until the copies ran
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

// pinned host staging buffer of the handle, at least n bytes
static int gff_host(gral_pf_handle *h, size_t n) {
    if (n <= h->gff_hbuf_bytes) return 0;
    if (h->gff_hbuf) (void)hipHostFree(h->gff_hbuf);
    h->gff_hbuf = nullptr;
    h->gff_hbuf_bytes = 0;
    if (hipHostMalloc(&h->gff_hbuf, n, hipHostMallocDefault) != hipSuccess) {
        h->gff_hbuf = nullptr;
        return fail(h, GRAL_PF_ENOMEM, "gff: hipHostMalloc failed");
    }
    h->gff_hbuf_bytes = n;
    return 0;
}

static void gff_free(gral_pf_handle *h) {
    if (h->gff_up) (void)hipFree(h->gff_up);
    if (h->gff_dbuf) (void)hipFree(h->gff_dbuf);
    if (h->gff_hbuf) (void)hipHostFree(h->gff_hbuf);
    h->gff_up = h->gff_dbuf = h->gff_hbuf = nullptr;
    h->gff_bytes = h->gff_dbuf_bytes = h->gff_hbuf_bytes = 0;
    if (h->pw_buf) (void)hipFree(h->pw_buf);
    h->pw_buf = nullptr;
    h->pw_bytes = 0;
}

// The particle model's wind at n particle positions on the current fields
// (Zeitschleife.IntWindCalculate, Intwind.cs:31-170).
extern "C" int gral_pf_wind_at_particles(gral_pf_handle *h, int32_t topo, int32_t ikooagral, int32_t jkooagral,
                                         float ahmin, int32_t n, const int32_t *cell_i, const int32_t *cell_j,
                                         const float *ahint, const double *x, const double *y, const float *z,
                                         float *ux, float *uy, float *uz, int32_t *cell_k) {
    if (!h) return GRAL_PF_EINVAL;
    if (!h->prepared) return fail(h, GRAL_PF_ESTATE, "wind_at_particles: call prepare first");
    if (topo != 0 && topo != 1) return fail(h, GRAL_PF_EINVAL, "wind_at_particles: topo must be 0 (flat, buildings) or 1 (terrain)");
    if (n < 0) return fail(h, GRAL_PF_EINVAL, "wind_at_particles: n < 0");
    if (n == 0) return GRAL_PF_OK;
    if (!cell_i || !cell_j || !ahint || !x || !y || !z || !ux || !uy || !uz || !cell_k)
        return fail(h, GRAL_PF_EINVAL, "wind_at_particles: null argument");
    for (int32_t p = 0; p < n; ++p)
        if (cell_i[p] < 1 || cell_i[p] > h->nii || cell_j[p] < 1 || cell_j[p] > h->njj)
            return fail(h, GRAL_PF_EINVAL, "wind_at_particles: particle " + std::to_string(p) + " outside the cells 1..NII x 1..NJJ");
    const int nlo = std::min(h->nkp, h->nkh);
    int rc = gff_upper(h, nlo);
    if (rc) return rc;
    const size_t nn = (size_t)n, nk = (size_t)h->nkh;
    // layout: x, y (double), then HOKART, DZK, ahint, z, ux, uy, uz (float), cell_i, cell_j, cell_k (int)
    const size_t o_y = 8 * nn, o_hok = 16 * nn, o_dzk = o_hok + 4 * nk, o_ah = o_dzk + 4 * nk, o_z = o_ah + 4 * nn,
                 o_ux = o_z + 4 * nn, o_uy = o_ux + 4 * nn, o_uz = o_uy + 4 * nn, o_ci = o_uz + 4 * nn,
                 o_cj = o_ci + 4 * nn, o_ck = o_cj + 4 * nn, bytes = o_ck + 4 * nn;
    if (bytes > h->pw_bytes) {
        if (h->pw_buf) (void)hipFree(h->pw_buf);
        h->pw_buf = nullptr;
        h->pw_bytes = 0;
        if (hipMalloc(&h->pw_buf, bytes) != hipSuccess) { h->pw_buf = nullptr; return fail(h, GRAL_PF_ENOMEM, "wind_at_particles: hipMalloc failed"); }
        h->pw_bytes = bytes;
    }
    char *db = static_cast<char *>(h->pw_buf);
    HIPCHK(h, hipMemcpyAsync(db, x, 8 * nn, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(db + o_y, y, 8 * nn, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(db + o_hok, h->hok.data(), 4 * nk, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(db + o_dzk, h->dzk.data(), 4 * nk, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(db + o_ah, ahint, 4 * nn, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(db + o_z, z, 4 * nn, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(db + o_ci, cell_i, 4 * nn, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(db + o_cj, cell_j, 4 * nn, hipMemcpyHostToDevice, h->stream));
    WindArgs a{};
    const float *lo[3] = {h->d.uk, h->d.vk, h->d.wk};
    const int nup = h->nkh - nlo;
    const size_t n2 = h->n2();
    for (int f = 0; f < 3; ++f) {
        a.lo[f] = lo[f];
        a.up[f] = nup > 0 ? static_cast<const float *>(h->gff_up) + (size_t)f * n2 * nup : nullptr;
    }
    a.nlo = nlo; a.nkp = h->nkp; a.nup = nup; a.nj2 = h->njj + 2; a.nkk = h->nkk;
    a.topo = topo; a.dxk = h->dxk; a.dyk = h->dyk; a.ahmin = ahmin; a.ikoo = ikooagral; a.jkoo = jkooagral;
    a.hok = reinterpret_cast<const float *>(db + o_hok);
    a.dzk = reinterpret_cast<const float *>(db + o_dzk);
    a.n = n;
    a.ci = reinterpret_cast<const int32_t *>(db + o_ci);
    a.cj = reinterpret_cast<const int32_t *>(db + o_cj);
    a.ahint = reinterpret_cast<const float *>(db + o_ah);
    a.z = reinterpret_cast<const float *>(db + o_z);
    a.x = reinterpret_cast<const double *>(db);
    a.y = reinterpret_cast<const double *>(db + o_y);
    a.ux = reinterpret_cast<float *>(db + o_ux);
    a.uy = reinterpret_cast<float *>(db + o_uy);
    a.uz = reinterpret_cast<float *>(db + o_uz);
    a.ck = reinterpret_cast<int32_t *>(db + o_ck);
    if (pf_launch_wind(&a, h->stream)) return fail(h, GRAL_PF_EDEVICE, "wind_at_particles: kernel launch failed");
    HIPCHK(h, hipMemcpyAsync(ux, db + o_ux, 4 * nn, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(uy, db + o_uy, 4 * nn, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(uz, db + o_uz, 4 * nn, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(cell_k, db + o_ck, 4 * nn, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return GRAL_PF_OK;
}

extern "C" int gral_pf_gff_payload(gral_pf_handle *h, int32_t kind, const gral_pf_gff_meta *meta, void *out,
                                   size_t cap, size_t *len) {
    if (!h) return GRAL_PF_EINVAL;
    if (kind < 0 || kind > 3) return fail(h, GRAL_PF_EINVAL, "gff: kind must be 0, 1, 2 (writing modes) or 3 (tke)");
    if (!meta || !len) return fail(h, GRAL_PF_EINVAL, "gff: null meta or len");
    if (!h->masks || !h->fields) return fail(h, GRAL_PF_ESTATE, "gff: masks and fields must be set first");
    if (kind == 3 && !(h->prepared && h->prm.turbulence_model == 2))
        return fail(h, GRAL_PF_ESTATE, "gff: the tke file needs a prepared k-eps (turbulence_model 2) situation");
    *len = gff_size(h, kind);
    if (!out) return GRAL_PF_OK;
    if (cap < *len) return fail(h, GRAL_PF_EINVAL, "gff: output buffer too small");
    uint8_t *o8 = static_cast<uint8_t *>(out);
    size_t body = 0;
    int rc = gff_pack(h, kind, meta, o8, body);
    if (rc) return rc;
    // device -> pinned staging in chunks, each copied on to the caller's buffer
    // as soon as its transfer has completed (the two copies overlap)
    constexpr size_t kCh = size_t(8) << 20;
    const size_t stage = std::min(body, 4 * kCh);
    if ((rc = gff_host(h, std::max<size_t>(stage, 1)))) return rc;
    const size_t nch = (body + kCh - 1) / kCh;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int q = 0; q < 4; ++q)
        if (hipEventCreateWithFlags(&ev[q], hipEventDisableTiming) != hipSuccess) {
            for (int r = 0; r < q; ++r) (void)hipEventDestroy(ev[r]);
            return fail(h, GRAL_PF_EDEVICE, "gff: event create failed");
        }
    const char *db = static_cast<const char *>(h->gff_dbuf);
    char *hb = static_cast<char *>(h->gff_hbuf);
    bool ok = true;
    auto issue = [&](size_t c) {
        const size_t o = c * kCh, n = std::min(kCh, body - o);
        ok = ok && hipMemcpyAsync(hb + (c % 4) * kCh, db + o, n, hipMemcpyDeviceToHost, h->stream) == hipSuccess &&
             hipEventRecord(ev[c % 4], h->stream) == hipSuccess;
    };
    for (size_t c = 0; c < std::min<size_t>(nch, 4); ++c) issue(c);
    for (size_t c = 0; ok && c < nch; ++c) {
        ok = hipEventSynchronize(ev[c % 4]) == hipSuccess;
        if (!ok) break;
        const size_t o = c * kCh, n = std::min(kCh, body - o);
        std::memcpy(o8 + 32 + o, hb + (c % 4) * kCh, n);
        if (c + 4 < nch) issue(c + 4);
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess) ok = false;
    for (int q = 0; q < 4; ++q) (void)hipEventDestroy(ev[q]);
    if (!ok) return fail(h, GRAL_PF_EDEVICE, "gff: download failed");
    return GRAL_PF_OK;
}

extern "C" int gral_pf_gff_write(gral_pf_handle *h, int32_t kind, const gral_pf_gff_meta *meta, const char *path) {
    if (!h) return GRAL_PF_EINVAL;
    if (!path) return fail(h, GRAL_PF_EINVAL, "gff: null path");
    size_t len = 0;
    int rc = gral_pf_gff_payload(h, kind, meta, nullptr, 0, &len);
    if (rc) return rc;
    if (len >= 0xffffffffull) return fail(h, GRAL_PF_EINVAL, "gff: payload of 4 GiB or more (zip without ZIP64)");
    // whole payload into the pinned buffer, deflated from there
    if ((rc = gff_host(h, len))) return rc;
    uint8_t *hb = static_cast<uint8_t *>(h->gff_hbuf);
    size_t body = 0;
    if ((rc = gff_pack(h, kind, meta, hb, body))) return rc;
    HIPCHK(h, hipMemcpyAsync(hb + 32, h->gff_dbuf, body, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // the zip entry carries the file's own name (GFFWrite.cs:109, :177)
    std::string entry(path);
    const size_t slash = entry.find_last_of("/\\");
    if (slash != std::string::npos) entry = entry.substr(slash + 1);
    if ((rc = gral_pf_zip_write(path, entry.c_str(), hb, len))) return fail(h, rc, "gff: writing the zip file failed");
    return GRAL_PF_OK;
}

extern "C" int gral_pf_last_timing(gral_pf_handle *h, gral_pf_timing *t) {
    if (!h || !t) return GRAL_PF_EINVAL;
    *t = h->timing;
    return GRAL_PF_OK;
}

extern "C" const char *gral_pf_last_error(const gral_pf_handle *h) {
    if (!h) return "null handle";
    return h->err.c_str();
}

extern "C" void gral_pf_destroy(gral_pf_handle *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_device(h);
    delete h->graphs;
    gff_free(h);
    if (h->stage_buf) (void)hipHostFree(h->stage_buf);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}
