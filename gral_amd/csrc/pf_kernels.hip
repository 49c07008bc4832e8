// HIP kernels of the prognostic microscale flow solver (gfx950).
//
// One SIMPLE iteration (PrognosticFlowfield.cs:584-1178) is a sequence of
//   stencil passes   - sub-domain BCs, divergence, velocity correction, averages:
//                      one thread per (column, level), k contiguous -> coalesced;
//   Gauss-Seidel passes - pressure TDMA, U/V/W momentum, TKE/eps:
//                      the reference updates columns in place in lexicographic
//                      sweep order (a outer, b inner).  A column depends on the
//                      new values of (a-1,b) and (a,b-1) only, so all columns of
//                      one anti-diagonal a+b=t are independent.  Each pass is a
//                      sequence of per-diagonal launches; one 64-lane wavefront
//                      per column computes the coefficients of all levels in
//                      parallel (lanes over k), stages them in LDS, and one lane
//                      runs the Thomas recurrences, which are serial in k.
// This replays the reference's deterministic (single-thread) order exactly.
#include <mutex>
#include "pf_device.h"

#define WAVE 64

// ------------------------------------------------------------------------
// diagonal -> column mapping
// ------------------------------------------------------------------------
// Sweep indices a in [oa, ob] (outer), b in [ia, ib] (inner); diagonal t
// holds the columns with (a-oa)+(b-ia) == t, enumerated by a.
__device__ __forceinline__ bool diag_col(int t, int oa, int ob, int ia, int ib, int idx, int &a, int &b) {
    int nb = ib - ia;
    int amin = max(oa, oa + t - nb);
    int amax = min(ob, oa + t);
    a = amin + idx;
    if (a > amax) return false;
    b = ia + (t - (a - oa));
    return true;
}

// ------------------------------------------------------------------------
// utility
// ------------------------------------------------------------------------
// Workgroups are dealt to the 8 XCDs (each with its own L2) round robin by
// their position in the grid.  For kernels whose neighbouring blocks read each
// other's columns, xcd_block() renumbers the blocks of a launch whose x extent is
// a multiple of 8 so that every XCD takes one contiguous range of logical
// blocks; it is a bijection on [0, n) and returns -1 for the padding blocks.
#define NXCD 8
__device__ __forceinline__ int xcd_block(int n) {
    int b = blockIdx.x, x = b % NXCD, idx = b / NXCD;
    int per = n / NXCD, rem = n % NXCD;
    int cnt = per + (x < rem ? 1 : 0);
    if (idx >= cnt) return -1;
    return x * per + min(x, rem) + idx;
}
// column metadata from the per-column table (see gral_pf_prepare)
struct Meta {
    int KK, VI, KSTART, KKmi, KKmj;
    bool adv, vegp;
};
__device__ __forceinline__ Meta unpack_meta(int4 m) {
    Meta r;
    r.KK = m.x; r.VI = m.y; r.KSTART = m.z & 0xffff; r.adv = (m.z >> 16) & 1; r.vegp = (m.z >> 17) & 1;
    r.KKmi = m.w & 0xffff; r.KKmj = (m.w >> 16) & 0xffff;
    return r;
}

// The cell kernels (divergence, velocity correction, averages) walk the
// interior columns of one sweep row i (j = 2..NJJ-1 are contiguous in memory,
// nkp levels each) as one flat run of floats: a block per row, each thread
// four consecutive levels of one column per step (16-byte accesses; nkp is a
// multiple of 4), so every load of a wave is contiguous and a block keeps
// many independent loads in flight.  Blocks take rows in XCD order.  With
// blocks of 1024 threads (16 waves; about 8 steps per thread on a 400-column
// row) the cell and boundary kernels of the 8-situation city batch take
// 2.01 ms per iteration, against 2.06 with 512, 2.10 with 256, 2.19 with 128.
#ifndef CELL_THREADS
#define CELL_THREADS 1024
#endif

__global__ void k_fill_u32(unsigned *p, size_t n, unsigned v) {
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) p[g] = v;
}
// zero the pressure row scratch of every situation (PIMP/QIMP rows are fresh
// arrays at the start of each pressure pass)
// and the strip synchronisation blocks of the iteration's sweeps
__global__ void k_zero_rows(const Dom *__restrict__ ds, int both) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];
    size_t n = (size_t)d.nj2 * d.nkp;
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) {
        d.prow_q[g] = 0.0f;
        if (both) d.prow_p[g] = 0.0f;
    }
    if (d.sync && g < (size_t)5 * d.sync_stride) d.sync[SYNC_BLOCK0 + g] = 0;
}

// ------------------------------------------------------------------------
// Sub-domain boundary conditions, PrognosticFlowfield.cs:595-714
// ------------------------------------------------------------------------
// E/W: each row j is independent; within a row the reference scans i upwards
// and later i read what earlier i wrote, so one thread per (j,k) scans i.  One
// block per row: the row's ADVDOM and VerticalIndex are staged in LDS first,
// and one wave lists the row's edge cells (ADVDOM = 1 next to a cell outside
// the sub-domains, or at the first / last interior index) in increasing order,
// so the scan visits only the cells where the reference's loop does anything,
// in the reference's order.
// LDS per block: [n] VI, [n] edge list, its length, then [n] ADVDOM bytes.
__host__ __device__ inline size_t bc_lds_bytes(int n) { return (size_t)(2 * n + 1) * sizeof(int) + n; }
// the indices x in [3, last] with adv[x] = 1 and at least one of adv[x-1] < 1,
// adv[x+1] < 1, x = 3, x = last, ascending, into edges[]; returns their count
// (to every thread of the block)
__device__ __forceinline__ int bc_edges(const uint8_t *adv, int last, int *edges, int *count) {
    if (threadIdx.x < WAVE) {
        const int lane = threadIdx.x;
        int cnt = 0;
        for (int base = 3; base <= last; base += WAVE) {
            const int x = base + lane;
            const bool e = x <= last && adv[x] == 1 && (adv[x - 1] < 1 || x == 3 || adv[x + 1] < 1 || x == last);
            const unsigned long long m = __ballot(e);
            if (e) edges[cnt + __popcll(m & ((1ull << lane) - 1))] = x;
            cnt += __popcll(m);
        }
        if (lane == 0) *count = cnt;
    }
    __syncthreads();
    return *count;
}
__global__ void k_bc_ew(const Dom *__restrict__ ds) {
    extern __shared__ int bc_row[];
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice
    if ((int)blockIdx.y >= d.njj - 1) return;

    int j = blockIdx.y + 2;                   // 2..NJJ
    const int n = d.nii + 2;
    int *svi = bc_row, *sedge = bc_row + n;
    uint8_t *sadv = reinterpret_cast<uint8_t *>(bc_row + 2 * n + 1);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const size_t c = c2(d, i, j);
        svi[i] = d.vi[c];
        sadv[i] = d.adv[c];
    }
    __syncthreads();
    const int nedge = bc_edges(sadv, d.nii - 1, sedge, bc_row + 2 * n);
    for (int k = threadIdx.x + 1; k <= d.kadvmax; k += blockDim.x)
    for (int q = 0; q < nedge; ++q) {
        const int i = sedge[q];
        if (k > svi[i]) continue;
        bool iM = (sadv[i - 1] < 1) || (i == 3);
        bool iP = (sadv[i + 1] < 1) || (i == d.nii - 1);
        if (iM) {
            float u = d.uk[c3(d, i - 2, j) + k];
            if (u >= 0) d.uk[c3(d, i - 1, j) + k] = u;
            d.vk[c3(d, i - 2, j) + k] = d.vk[c3(d, i - 1, j) + k];
            d.wk[c3(d, i - 2, j) + k] = d.wk[c3(d, i - 1, j) + k];
        }
        if (iP) {
            float u = d.uk[c3(d, i + 2, j) + k];
            if (u <= 0) d.uk[c3(d, i + 1, j) + k] = u;
            d.vk[c3(d, i + 1, j) + k] = d.vk[c3(d, i, j) + k];
            d.wk[c3(d, i + 1, j) + k] = d.wk[c3(d, i, j) + k];
        }
    }
}

// N/S: as E/W with the roles of i and j exchanged (the edge list over j)
__global__ void k_bc_ns(const Dom *__restrict__ ds) {
    extern __shared__ int bc_row[];
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice
    if ((int)blockIdx.y >= d.nii - 1) return;

    int i = blockIdx.y + 2;                   // 2..NII
    const int n = d.njj + 2;
    int *svi = bc_row, *sedge = bc_row + n;
    uint8_t *sadv = reinterpret_cast<uint8_t *>(bc_row + 2 * n + 1);
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const size_t c = c2(d, i, j);
        svi[j] = d.vi[c];
        sadv[j] = d.adv[c];
    }
    __syncthreads();
    const int nedge = bc_edges(sadv, d.njj - 1, sedge, bc_row + 2 * n);
    for (int k = threadIdx.x + 1; k <= d.kadvmax; k += blockDim.x)
    for (int q = 0; q < nedge; ++q) {
        const int j = sedge[q];
        if (k > svi[j]) continue;
        bool jM = (sadv[j - 1] < 1) || (j == 3);
        bool jP = (sadv[j + 1] < 1) || (j == d.njj - 1);
        if (jM) {
            float v = d.vk[c3(d, i, j - 2) + k];
            if (v >= 0) d.vk[c3(d, i, j - 1) + k] = v;
            d.uk[c3(d, i, j - 2) + k] = d.uk[c3(d, i, j - 1) + k];
            d.wk[c3(d, i, j - 2) + k] = d.wk[c3(d, i, j - 1) + k];
        }
        if (jP) {
            float v = d.vk[c3(d, i, j + 2) + k];
            if (v <= 0) d.vk[c3(d, i, j + 1) + k] = v;
            d.uk[c3(d, i, j + 1) + k] = d.uk[c3(d, i, j) + k];
            d.wk[c3(d, i, j + 1) + k] = d.wk[c3(d, i, j) + k];
        }
    }
    // top of the sub-domains, VI -> VI+1 (PrognosticFlowfield.cs:698-714): the
    // N/S pass of this row has written all it writes of the row's columns, and
    // no other row writes them
    if (i > d.nii - 1) return;
    __syncthreads();
    for (int j = 2 + threadIdx.x; j <= d.njj - 1; j += blockDim.x) {
        if (sadv[j] != 1) continue;
        const int k = svi[j];
        const size_t o = c3(d, i, j);
        d.wk[o + k + 1] = d.wk[o + k];
        d.uk[o + k + 1] = d.uk[o + k];
        d.vk[o + k + 1] = d.vk[o + k];
    }
}

// ------------------------------------------------------------------------
// Mass divergence, PrognosticFlowfield.cs:723-758, over the active cells
// (ADVDOM = 1, KKART < k <= VI); DPM = 0 there.  The array keeps DIV / DTIME,
// the only form the pressure sweep reads (its source term, :853), divided here
// in the bandwidth-bound pass rather than on the sweep's serial path (the same
// correctly rounded division, so the same bits).
// ------------------------------------------------------------------------
__global__ __launch_bounds__(CELL_THREADS) void k_divergence(const Dom *__restrict__ ds) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice
    const int blk = xcd_block(d.nii - 2);
    if (blk < 0) return;
    const int i = 2 + blk, nkp = d.nkp, n4 = ((d.njj - 2) * nkp) >> 2;
    const size_t row = c3(d, i, 2), ip = (size_t)d.nj2 * nkp;
    for (int q = threadIdx.x; q < n4; q += CELL_THREADS) {
        const int e = q << 2, jo = e / nkp, k0 = e - jo * nkp;
        const Meta m = unpack_meta(ldg4i(d.meta + c2(d, i, 2 + jo)));
        if (!m.adv || k0 + 3 <= m.KK || k0 > m.VI) continue;
        const size_t o = row + e;
        const float4 u = ldg4(d.uk + o), ue = ldg4(d.uk + o + ip), v = ldg4(d.vk + o), vn = ldg4(d.vk + o + nkp);
        const float4 w = ldg4(d.wk + o);
        const float uu[4] = {u.x, u.y, u.z, u.w}, uue[4] = {ue.x, ue.y, ue.z, ue.w};
        const float vv[4] = {v.x, v.y, v.z, v.w}, vvn[4] = {vn.x, vn.y, vn.z, vn.w};
        const float ww[5] = {w.x, w.y, w.z, w.w, d.wk[o + 4]};
        float r[4];
        bool on[4];
#pragma unroll
        for (int u4 = 0; u4 < 4; ++u4) {
            const int k = k0 + u4;
            on[u4] = m.KK < k && k <= m.VI;
            r[u4] = 0.0f;
            if (on[u4]) {
                const float dz = d.dzk[k];
                const float t1 = (uu[u4] - uue[u4]) * d.dyk * dz;
                const float t2 = (vv[u4] - vvn[u4]) * d.dxk * dz;
                r[u4] = (k > m.KK + 1) ? t1 + t2 + (ww[u4] - ww[u4 + 1]) * d.area : t1 + t2 - ww[u4 + 1] * d.area;
                r[u4] = r[u4] / d.dtime;
            }
        }
        if (on[0] && on[1] && on[2] && on[3]) {
            stg4(d.div + o, r[0], r[1], r[2], r[3]);
            stg4(d.dpm + o, 0.0f, 0.0f, 0.0f, 0.0f);
        } else {
#pragma unroll
            for (int u4 = 0; u4 < 4; ++u4)
                if (on[u4]) { d.div[o + u4] = r[u4]; d.dpm[o + u4] = 0.0f; }
        }
    }
}

// ------------------------------------------------------------------------
// Pressure-correction GS-TDMA, PrognosticFlowfield.cs:770-901.
// Sweep: a = j1 in [2,NJJ-1] (outer), b = i1 in [2,NII-1] (inner).
// PIMP/QIMP live per sweep row in global scratch, zeroed at the start of the
// pass, so the top coupling PIMP[VI+1] sees what the last earlier column of
// the same row left there, as in the reference.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(WAVE) void k_pressure(const Dom *__restrict__ ds, int t, int IS, int JS, int slot) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice

    extern __shared__ float sm[];
    int a, b;
    if (!diag_col(t, 2, d.njj - 1, 2, d.nii - 1, blockIdx.x, a, b)) return;
    int j = (JS == -1) ? d.njj - a + 1 : a;
    int i = (IS == -1) ? d.nii - b + 1 : b;
    size_t c = c2(d, i, j);
    if (d.adv[c] != 1) return;
    const int nkp = d.nkp, tid = threadIdx.x;
    float *sDIM = sm, *sP = sDIM + nkp, *sQ = sP + nkp, *sN = sQ + nkp, *sO = sN + nkp;
    int KK = d.kk[c], VI = d.vi[c];
    size_t o = c3(d, i, j), oim = c3(d, i - 1, j), oip = c3(d, i + 1, j), ojm = c3(d, i, j - 1), ojp = c3(d, i, j + 1);
    float DXK = d.dxk, DYK = d.dyk;
    for (int k = KK + 1 + tid; k <= VI; k += WAVE) {
        float dz = d.dzk[k];
        float DIM = d.div[o + k] + (d.dpm[oim + k] + d.dpm[oip + k]) / DXK * DYK * dz +
                    (d.dpm[ojm + k] - d.dpm[ojp + k]) / DYK * DXK * dz;
        sDIM[k] = DIM;
    }
    for (int k = KK + tid; k <= VI; k += WAVE) sO[k] = d.dpm[o + k];
    size_t row = (size_t)j * nkp;
    __syncthreads();
    if (tid == 0) {
        float Pn = d.prow_p[row + VI + 1], Qn = d.prow_q[row + VI + 1];
        for (int k = VI; k > KK; --k) {
            float ATP = d.atp[k];
            float TERMP = 1 / (d.app[k] - ATP * Pn);
            float P = d.abp[k] * TERMP;
            float Q = (ATP * Qn + sDIM[k]) * TERMP;
            sP[k] = P; sQ[k] = Q;
            Pn = P; Qn = Q;
        }
        float dprev = sO[KK];
        for (int k = KK + 1; k <= VI; ++k) {
            float temp = sP[k] * dprev + sQ[k];
            sN[k] = temp;
            dprev = temp;
        }
    }
    __syncthreads();
    float rmax = 0.0f;
    for (int k = KK + 1 + tid; k <= VI; k += WAVE) {
        float temp = sN[k];
        double q = (double)conv_to_int((double)temp * d.frund) / d.frund;
        float r = (float)fabs((double)sO[k] - q);
        rmax = r > rmax ? r : rmax;
        d.dpm[o + k] = temp;
        d.dpmnew[o + k] = d.dpmnew[o + k] + d.relaxp * temp;
        d.prow_p[row + k] = sP[k];
        d.prow_q[row + k] = sQ[k];
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        float other = __shfl_xor(rmax, off);
        rmax = other > rmax ? other : rmax;
    }
    const int rs = slot - d.iter_off;                 // this situation's residual slot
    if (tid == 0 && rmax > 0.0f && slot >= 0 && rs >= 0 && rs < d.resid_cap) atomicMax((unsigned *)&d.resid[rs], __float_as_uint(rmax));
}

// ------------------------------------------------------------------------
// Velocity correction with FP64 quantisation, PrognosticFlowfield.cs:904-949,
// over the columns with ADVDOM = 1 at (i,j), (i-1,j) or (i,j-1), levels 1..VI.
// ------------------------------------------------------------------------
__device__ __forceinline__ bool corr_column(const DomG &d, int i, int j, bool adv) {
    return adv || d.adv[c2(d, i - 1, j)] == 1 || d.adv[c2(d, i, j - 1)] == 1;
}
__global__ __launch_bounds__(CELL_THREADS) void k_correction(const Dom *__restrict__ ds) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice
    const int blk = xcd_block(d.nii - 2);
    if (blk < 0) return;
    const int i = 2 + blk, nkp = d.nkp, n4 = ((d.njj - 2) * nkp) >> 2;
    const size_t row = c3(d, i, 2), ip = (size_t)d.nj2 * nkp;
    const bool q = !d.noquant;
    for (int t = threadIdx.x; t < n4; t += CELL_THREADS) {
        const int e = t << 2, jo = e / nkp, k0 = e - jo * nkp, j = 2 + jo;
        const Meta m = unpack_meta(ldg4i(d.meta + c2(d, i, j)));
        if (k0 > m.VI || !corr_column(d, i, j, m.adv)) continue;
        const size_t o = row + e;
        const float4 p = ldg4(d.dpm + o), pw = ldg4(d.dpm + o - ip), ps = ldg4(d.dpm + o - nkp);
        const float4 u = ldg4(d.uk + o), v = ldg4(d.vk + o), w = ldg4(d.wk + o);
        const float pp[5] = {d.dpm[o - 1], p.x, p.y, p.z, p.w};              // pp[u4 + 1]: level k0 + u4
        const float ppw[4] = {pw.x, pw.y, pw.z, pw.w}, pps[4] = {ps.x, ps.y, ps.z, ps.w};
        float uu[4] = {u.x, u.y, u.z, u.w}, vv[4] = {v.x, v.y, v.z, v.w}, ww[4] = {w.x, w.y, w.z, w.w};
        bool onu[4], onv[4], on[4];
#pragma unroll
        for (int u4 = 0; u4 < 4; ++u4) {
            const int k = k0 + u4;
            on[u4] = k >= 1 && k <= m.VI;
            onu[u4] = on[u4] && i > 2 && m.KKmi < k;              // KKART(i,j) < k and KKART(i-1,j) < k
            onv[u4] = on[u4] && j > 2 && m.KKmj < k;
            const float dpmk = pp[u4 + 1];
            if (onu[u4]) {
                const float x = uu[u4] + (ppw[u4] - dpmk) / d.dxk * d.dtime;
                uu[u4] = q ? quantise(x, d.frund) : x;
            }
            if (onv[u4]) {
                const float x = vv[u4] + (pps[u4] - dpmk) / d.dyk * d.dtime;
                vv[u4] = q ? quantise(x, d.frund) : x;
            }
            if (on[u4]) {
                if (m.KK + 1 < k) {
                    const float x = ww[u4] + (pp[u4] - dpmk) / (d.dzk[k] + d.dzk[k - 1]) * 2.0f * d.dtime;
                    ww[u4] = q ? quantise(x, d.frund) : x;
                } else {
                    ww[u4] = 0.0f;
                }
            }
        }
        if (onu[0] && onu[1] && onu[2] && onu[3]) stg4(d.uk + o, uu[0], uu[1], uu[2], uu[3]);
        else
#pragma unroll
            for (int u4 = 0; u4 < 4; ++u4) if (onu[u4]) d.uk[o + u4] = uu[u4];
        if (onv[0] && onv[1] && onv[2] && onv[3]) stg4(d.vk + o, vv[0], vv[1], vv[2], vv[3]);
        else
#pragma unroll
            for (int u4 = 0; u4 < 4; ++u4) if (onv[u4]) d.vk[o + u4] = vv[u4];
        if (on[0] && on[1] && on[2] && on[3]) stg4(d.wk + o, ww[0], ww[1], ww[2], ww[3]);
        else
#pragma unroll
            for (int u4 = 0; u4 < 4; ++u4) if (on[u4]) d.wk[o + u4] = ww[u4];
    }
}

// ------------------------------------------------------------------------
// Cell-centred averages, PrognosticFlowfield.cs:958-1014 (+ UKS row copies),
// over the correction's columns, levels KSTART..VI.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(CELL_THREADS) void k_averages(const Dom *__restrict__ ds) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice
    const int blk = xcd_block(d.nii - 2);
    if (blk < 0) return;
    const int i = 2 + blk, nkp = d.nkp, n4 = ((d.njj - 2) * nkp) >> 2;
    const size_t row = c3(d, i, 2), ip = (size_t)d.nj2 * nkp;
    const bool copy_n = d.adv[c2(d, i, d.njj)] == 1;
    for (int t = threadIdx.x; t < n4; t += CELL_THREADS) {
        const int e = t << 2, jo = e / nkp, k0 = e - jo * nkp, j = 2 + jo;
        const Meta m = unpack_meta(ldg4i(d.meta + c2(d, i, j)));
        if (k0 + 3 < m.KSTART || k0 > m.VI || !corr_column(d, i, j, m.adv)) continue;
        const size_t o = row + e;
        const float4 u = ldg4(d.uk + o), ue = ldg4(d.uk + o + ip), v = ldg4(d.vk + o), vn = ldg4(d.vk + o + nkp);
        const float4 w = ldg4(d.wk + o);
        const float uu[4] = {u.x, u.y, u.z, u.w}, uue[4] = {ue.x, ue.y, ue.z, ue.w};
        const float vv[4] = {v.x, v.y, v.z, v.w}, vvn[4] = {vn.x, vn.y, vn.z, vn.w};
        const float ww[5] = {w.x, w.y, w.z, w.w, d.wk[o + 4]};
        float ru[4], rv[4], rw[4];
        bool on[4], onuv[4], onw[4];
#pragma unroll
        for (int u4 = 0; u4 < 4; ++u4) {
            const int k = k0 + u4;
            on[u4] = k >= m.KSTART && k <= m.VI;
            onuv[u4] = on[u4] && m.KK < k;
            onw[u4] = on[u4] && m.KK + 1 < k;
            ru[u4] = 0.5f * (uu[u4] + uue[u4]);
            rv[u4] = 0.5f * (vv[u4] + vvn[u4]);
            rw[u4] = 0.5f * (ww[u4] + ww[u4 + 1]);
        }
        if (onuv[0] && onuv[1] && onuv[2] && onuv[3]) {
            stg4(d.uks + o, ru[0], ru[1], ru[2], ru[3]);
            stg4(d.vks + o, rv[0], rv[1], rv[2], rv[3]);
        } else {
#pragma unroll
            for (int u4 = 0; u4 < 4; ++u4)
                if (onuv[u4]) { d.uks[o + u4] = ru[u4]; d.vks[o + u4] = rv[u4]; }
        }
        if (onw[0] && onw[1] && onw[2] && onw[3]) stg4(d.wks + o, rw[0], rw[1], rw[2], rw[3]);
        else
#pragma unroll
            for (int u4 = 0; u4 < 4; ++u4) if (onw[u4]) d.wks[o + u4] = rw[u4];
        // UKS[i][1] = UKS[i][2] and UKS[i][NJJ] = UKS[i][NJJ-1] (if ADVDOM there) over
        // KSTART..VI: the new value where it was just written, else the old one
        const bool cs = (j == 2), cn = (j == d.njj - 1) && copy_n;
        if (cs || cn) {
            const size_t oc = cs ? c3(d, i, 1) : c3(d, i, d.njj);
#pragma unroll
            for (int u4 = 0; u4 < 4; ++u4)
                if (on[u4]) d.uks[oc + k0 + u4] = onuv[u4] ? ru[u4] : d.uks[o + u4];
        }
    }
}

// VKS[1][j] = VKS[2][j], VKS[NII][j] = VKS[NII-1][j] (PrognosticFlowfield.cs:1019-1039)
__global__ void k_avg_edges(const Dom *__restrict__ ds) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice
    if ((int)blockIdx.x >= d.njj - 2) return;

    int j = blockIdx.x + 2;                  // 2..NJJ-1
    bool w = d.adv[c2(d, 1, j)] == 1, e = d.adv[c2(d, d.nii - 1, j)] == 1;
    size_t o1 = c3(d, 1, j), o2 = c3(d, 2, j), on = c3(d, d.nii, j), on1 = c3(d, d.nii - 1, j);
    for (int k = threadIdx.x; k < d.nkp; k += blockDim.x) {
        if (w) d.vks[o1 + k] = d.vks[o2 + k];
        if (e) d.vks[on + k] = d.vks[on1 + k];
    }
}

// ------------------------------------------------------------------------
// Mixing-length momentum (turbulence model 1): U/V/W-prognostic-microscale_1.cs
// ------------------------------------------------------------------------
// The reference evaluates a column in vector blocks of W lanes starting at
// KSTART, KSTART+W, ..., the last block moved down to KEND-W.  Every lane
// quantity is lane-local except the gate of the eddy-viscosity cross terms,
// which tests the block start (k_block > KKART+1); block_base() returns the
// start of the block whose lane k enters the recurrence.
// W = 8: Vector<float> solvers (U/V/W-prognostic-microscale_1.cs, AVX2 width);
// W = 16: the 512-bit solvers (U/V/W-prognostic-microscale_1_Vec512.cs), which
// also fuse some multiply-adds (see v512 below).  The host accepts only 8 or 16.
__device__ __forceinline__ int block_base(int k, int KSTART, int KEND, int W) {
    const int sh = (W == 16) ? 4 : 3;
    int base = KSTART + (((k - KSTART) >> sh) << sh);
    if (base > KEND - W) base = KEND - W;
    return base;
}

struct ColInfo {
    int i, j, KK, KSTART, VI, W;
    bool as, an, aw, ae, vegp;
    float cutk, cov, ht, ho, z0s;
    __device__ __forceinline__ bool v512() const { return W == 16; }
};

__device__ __forceinline__ ColInfo col_info(const DomG &d, int i, int j) {
    ColInfo q;
    size_t c = c2(d, i, j);
    q.i = i; q.j = j; q.KK = d.kk[c]; q.VI = d.vi[c];
    q.cutk = d.cutk[c]; q.cov = d.cov[c]; q.ht = d.ht[c]; q.ho = d.ho[c]; q.z0s = d.z0s[c];
    q.KSTART = (q.cutk == 0) ? q.KK + 1 : 1;
    q.vegp = d.vegp != nullptr && d.vegp[c] != 0;
    q.as = (d.adv[c2(d, i, j - 1)] < 1) || (j == 2);
    q.an = (d.adv[c2(d, i, j + 1)] < 1) || (j == d.njj - 1);
    q.aw = (d.adv[c2(d, i - 1, j)] < 1) || (i == 2);
    q.ae = (d.adv[c2(d, i + 1, j)] < 1) || (i == d.nii - 1);
    q.W = d.simd;
    return q;
}

// mixing-length eddy viscosity (U-prognostic-microscale_1.cs:216-268), lane-local;
// the 512-bit solvers fuse the half-cell offset into the ground height
// (U-prognostic-microscale_1_Vec512.cs:218)
__device__ __forceinline__ float ml_visc(const DomG &d, const ColInfo &q, int k, const gf32 *UK, const gf32 *VK) {
    if (!(k > q.KK + 1)) return 0.0f;
    float dz = d.dzk[k];
    float zs = q.v512() ? d.hok[k] - fmaf(dz, 0.5f, d.hok[q.KK]) : d.hok[k] - d.hok[q.KK] - dz * 0.5f;
    float mixL = vminf(0.071f * zs, 100.0f);
    if (q.cov > 0) mixL *= (1.0f - 0.99f * q.cov);
    float in = (UK[k + 1] - UK[k - 1]) / (2.0f * dz);
    float dudz = in * in;
    in = (VK[k + 1] - VK[k - 1]) / (2.0f * dz);
    float dvdz = in * in;
    return vminf(mixL * mixL * sqrtf(0.5f * (dudz + dvdz)), 15.0f);
}

// power-law weight D*max(0,(1-0.1P)^5) in float (Vector path)
__device__ __forceinline__ float plaw_f(float D, float P) {
    float in = 1.0f - 0.1f * P;
    return D * vmaxf(0.0f, in * in * in * in * in);
}

// power-law coefficient plus its upwind flux Fp = max(+-F, 0):
// D*max(0, x^5) + Fp (U-prognostic-microscale_1.cs:345-367), or, in the 512-bit
// solvers, FMA(D, max(0, x*((x*x)*(x*x))), Fp) (U-prognostic-microscale_1_Vec512.cs:355-377)
__device__ __forceinline__ float plaw_coef(bool v512, float D, float P, float Fp) {
    if (!v512) return plaw_f(D, P) + Fp;
    float in = 1.0f - 0.1f * P;
    return fmaf(D, vmaxf(0.0f, in * ((in * in) * (in * in))), Fp);
}

struct Coef { float A, B, C, D; };

// Diffusion, advection and power-law coefficients shared by the three
// mixing-length solvers.  F* are the face fluxes of the component.
__device__ __forceinline__ void ml_coeffs(const DomG &d, const ColInfo &q, int k, float VIS, float FE, float FW,
                                          float FS, float FN, float FT, float FB, float &AW1, float &AS1,
                                          float &AE1, float &AN1, float &BIM, float &CIM, float &AIM) {
    float dz = d.dzk[k], dxkdzk = d.dxk * dz, dykdzk = d.dyk * dz;
    float mv = vmaxf(VIS, d.vishmin);
    float DE = mv * dykdzk / d.dxk, DW = DE;
    float DS = mv * dxkdzk / d.dyk, DN = DS;
    float DT = mv * d.area / dz;
    float DB = (k > q.KK + 1) ? DT : 0.0f;
    if (q.as) DS = 0; if (q.an) DN = 0; if (q.aw) DW = 0; if (q.ae) DE = 0;
    DE = vmaxf(DE, 0.0001f); DB = vmaxf(DB, 0.0001f); DW = vmaxf(DW, 0.0001f);
    DS = vmaxf(DS, 0.0001f); DN = vmaxf(DN, 0.0001f); DT = vmaxf(DT, 0.0001f);
    float PE = fabsf(FE / DE), PB = fabsf(FB / DB), PW = fabsf(FW / DW);
    float PS = fabsf(FS / DS), PN = fabsf(FN / DN), PT = fabsf(FT / DT);
    const bool v = q.v512();
    BIM = plaw_coef(v, DT, PT, vmaxf(-FT, 0.0f));
    CIM = plaw_coef(v, DB, PB, vmaxf(FB, 0.0f));
    AE1 = plaw_coef(v, DE, PE, vmaxf(-FE, 0.0f));
    AW1 = plaw_coef(v, DW, PW, vmaxf(FW, 0.0f));
    AS1 = plaw_coef(v, DS, PS, vmaxf(FS, 0.0f));
    AN1 = plaw_coef(v, DN, PN, vmaxf(-FN, 0.0f));
    AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + d.ap0[k];
}

// wall function at k = KKART+1 for U and V (U-prognostic-microscale_1.cs:419-450)
__device__ __forceinline__ float ml_wall_uv(const DomG &d, const ColInfo &q, int k, float vel, const gf32 *UKS,
                                            const gf32 *UKS_W, const gf32 *VKS, const gf32 *VKS_W) {
    float in = 0.5f * (UKS_W[k] + UKS[k]);
    float in2 = 0.5f * (VKS_W[k] + VKS[k]);
    float wh = vmaxf(sqrtf(in * in + in2 * in2), 0.1f);
    float us;
    if (q.cutk < 1) {
        us = q.ht * wh;
        if (q.z0s >= d.dzk[k] * 0.1f) {
            int ind = k + 1;
            float a = 0.5f * (UKS_W[ind] + UKS[ind]);
            float b = 0.5f * (VKS_W[ind] + VKS[ind]);
            us = q.ht * sqrtf(a * a + b * b);
        }
    } else {
        us = q.ho * wh;
    }
    return vel / wh * (us * us) * d.area;
}

// Pieces of the mixing-length source term D of one cell.  Everything here is
// fixed during a sweep (it reads only arrays the sweep does not write, plus the
// cell's own column, which is updated only when its own turn comes), except
// the four neighbour values of the swept component, which enter as
//   D = AW1*X_W + AS1*X_S + AE1*X_E + AN1*X_N + e + f + g  [- w]  [+ cross]
// summed left to right exactly as the reference does.  The 512-bit solvers
// form the neighbour part as FMA(AW1,X_W,AS1*X_S) + FMA(AE1,X_E,AN1*X_N)
// (not V without vegetation), and U without vegetation folds g into
// f = FMA(DDPX, DYKDZK, g) (U-prognostic-microscale_1_Vec512.cs:404-421).
struct MlPieces {
    float AW1, AS1, AE1, AN1, A, B, C;
    float e, f, g, w;          // g: Coriolis(+vegetation) term of U/V; vegetation drag (subtracted) of W
    float c0, X1, X2, X3, X4;  // cross terms: c0 = VIS/DXK, X* the neighbour-free gradients (W: X1 = whole term)
};

__device__ __forceinline__ bool ml_cross_on(const ColInfo &q, int k) {
    return block_base(k, q.KSTART, q.VI + 1, q.W) > q.KK + 1;
}

// face fluxes of the component's control volume at level k (U-prognostic-
// microscale_1.cs:313-330, V-...:317-331, W-...:286-300), in the reference's order
template <int COMP>
__device__ __forceinline__ void ml_fluxes(const DomG &d, const ColInfo &q, int k, float &FE, float &FW, float &FS,
                                          float &FN, float &FT, float &FB) {
    const int i = q.i, j = q.j;
    const float dz = d.dzk[k], dxkdzk = d.dxk * dz, dykdzk = d.dyk * dz;
    const gf32 *UKS = d.uks + c3(d, i, j), *VKS = d.vks + c3(d, i, j), *WKS = d.wks + c3(d, i, j);
    if (COMP == 0) {
        const gf32 *UKS_W = d.uks + c3(d, i - 1, j);
        const gf32 *VKS_W = d.vks + c3(d, i - 1, j), *VKS_WS = d.vks + c3(d, i - 1, j - 1);
        const gf32 *VKS_S = d.vks + c3(d, i, j - 1), *VKS_WN = d.vks + c3(d, i - 1, j + 1);
        const gf32 *VKS_N = d.vks + c3(d, i, j + 1), *WKS_W = d.wks + c3(d, i - 1, j);
        FE = UKS[k] * dykdzk;
        FW = UKS_W[k] * dykdzk;
        FS = 0.25f * (VKS_W[k] + VKS[k] + VKS_WS[k] + VKS_S[k]) * dxkdzk;
        FN = 0.25f * (VKS_W[k] + VKS[k] + VKS_WN[k] + VKS_N[k]) * dxkdzk;
        FT = 0.25f * (WKS_W[k] + WKS[k] + WKS_W[k + 1] + WKS[k + 1]) * d.area;
        FB = (k > q.KK + 1) ? 0.25f * (WKS_W[k] + WKS[k] + WKS_W[k - 1] + WKS[k - 1]) * d.area : 0.0f;
    } else if (COMP == 1) {
        const gf32 *UKS_W = d.uks + c3(d, i - 1, j), *UKS_E = d.uks + c3(d, i + 1, j);
        const gf32 *UKS_S = d.uks + c3(d, i, j - 1), *UKS_ES = d.uks + c3(d, i + 1, j - 1);
        const gf32 *UKS_WS = d.uks + c3(d, i - 1, j - 1);
        const gf32 *VKS_S = d.vks + c3(d, i, j - 1), *WKS_S = d.wks + c3(d, i, j - 1);
        FE = 0.25f * (UKS[k] + UKS_E[k] + UKS_S[k] + UKS_ES[k]) * dykdzk;
        FW = 0.25f * (UKS[k] + UKS_W[k] + UKS_S[k] + UKS_WS[k]) * dykdzk;
        FS = VKS_S[k] * dxkdzk;
        FN = VKS[k] * dxkdzk;
        FT = 0.25f * (WKS_S[k] + WKS[k] + WKS_S[k + 1] + WKS[k + 1]) * d.area;
        FB = (k > q.KK + 1) ? 0.25f * (WKS_S[k] + WKS[k] + WKS_S[k - 1] + WKS[k - 1]) * d.area : 0.0f;
    } else {
        const gf32 *UKS_W = d.uks + c3(d, i - 1, j), *UKS_E = d.uks + c3(d, i + 1, j);
        const gf32 *VKS_S = d.vks + c3(d, i, j - 1), *VKS_N = d.vks + c3(d, i, j + 1);
        FE = 0.25f * (UKS[k] + UKS_E[k] + UKS[k - 1] + UKS_E[k - 1]) * dykdzk;
        FW = 0.25f * (UKS[k] + UKS_W[k] + UKS[k - 1] + UKS_W[k - 1]) * dykdzk;
        FS = 0.25f * (VKS[k] + VKS_S[k] + VKS[k - 1] + VKS_S[k - 1]) * dxkdzk;
        FN = 0.25f * (VKS[k] + VKS_N[k] + VKS[k - 1] + VKS_N[k - 1]) * dxkdzk;
        FT = WKS[k] * d.area;
        FB = (k > q.KK + 1) ? WKS[k - 1] * d.area : 0.0f;
    }
}

__device__ MlPieces u1_pieces(const DomG &d, const ColInfo &q, int k) {
    int i = q.i, j = q.j;
    const gf32 *UK = d.uk + c3(d, i, j), *VK = d.vk + c3(d, i, j);
    const gf32 *UKS = d.uks + c3(d, i, j), *VKS = d.vks + c3(d, i, j), *WKS = d.wks + c3(d, i, j);
    const gf32 *UKS_W = d.uks + c3(d, i - 1, j);
    const gf32 *VKS_W = d.vks + c3(d, i - 1, j), *VKS_WS = d.vks + c3(d, i - 1, j - 1);
    const gf32 *VKS_S = d.vks + c3(d, i, j - 1), *VKS_WN = d.vks + c3(d, i - 1, j + 1);
    const gf32 *VKS_N = d.vks + c3(d, i, j + 1), *VKS_EN = d.vks + c3(d, i + 1, j + 1);
    const gf32 *VKS_ES = d.vks + c3(d, i + 1, j - 1);
    const gf32 *WKS_W = d.wks + c3(d, i - 1, j), *WK_E = d.wk + c3(d, i + 1, j);
    const gf32 *DPMNEW = d.dpmnew + c3(d, i, j), *DPMNEW_W = d.dpmnew + c3(d, i - 1, j);
    float dz = d.dzk[k], dxkdzk = d.dxk * dz, dykdzk = d.dyk * dz;
    float VIS = ml_visc(d, q, k, UK, VK);
    float FE, FW, FS, FN, FT, FB;
    ml_fluxes<0>(d, q, k, FE, FW, FS, FN, FT, FB);
    MlPieces o;
    ml_coeffs(d, q, k, VIS, FE, FW, FS, FN, FT, FB, o.AW1, o.AS1, o.AE1, o.AN1, o.B, o.C, o.A);
    float ap0 = d.ap0[k];
    float u = UK[k];
    o.e = ap0 * 0.5f * (UKS_W[k] + UKS[k]);
    o.f = (DPMNEW_W[k] - DPMNEW[k]) * dykdzk;
    if (q.vegp) {
        float in2 = 0.5f * (VKS_W[k] + VKS[k]);
        float in = 0.5f * (UKS_W[k] + UKS[k]);
        float wh = vmaxf(sqrtf(in * in + in2 * in2), 0.01f);
        float veg = d.veg[c3(d, i, j) + k - 1];
        o.g = (d.cor * (d.ug - u) - veg * u * wh) * d.area * dz;
    } else {
        o.g = d.cor * (d.ug - u) * d.area * dz;
        if (q.v512()) o.f = fmaf(DPMNEW_W[k] - DPMNEW[k], dykdzk, o.g);
    }
    o.w = (k == q.KK + 1) ? ml_wall_uv(d, q, k, u, UKS, UKS_W, VKS, VKS_W) : 0.0f;
    o.c0 = VIS / d.dxk;
    o.X1 = (0.5f * (VKS_EN[k] + VKS_N[k]) - 0.5f * (VKS_N[k] + VKS_WN[k])) * dxkdzk;        // DVDXN
    o.X2 = (0.5f * (VKS_ES[k] + VKS_S[k]) - 0.5f * (VKS_S[k] + VKS_WS[k])) * dxkdzk;        // DVDXS
    o.X3 = (0.5f * (WK_E[k] + WKS[k]) - 0.5f * (WKS[k] + WKS_W[k])) * d.area;                // DWDXT
    o.X4 = (0.5f * (WK_E[k - 1] + WKS[k - 1]) - 0.5f * (WKS[k - 1] + WKS_W[k - 1])) * d.area; // DWDXB
    return o;
}

__device__ MlPieces v1_pieces(const DomG &d, const ColInfo &q, int k) {
    int i = q.i, j = q.j;
    const gf32 *UK = d.uk + c3(d, i, j), *VK = d.vk + c3(d, i, j);
    const gf32 *UKS = d.uks + c3(d, i, j), *VKS = d.vks + c3(d, i, j), *WKS = d.wks + c3(d, i, j);
    const gf32 *UKS_W = d.uks + c3(d, i - 1, j), *UKS_E = d.uks + c3(d, i + 1, j);
    const gf32 *UKS_S = d.uks + c3(d, i, j - 1), *UKS_ES = d.uks + c3(d, i + 1, j - 1);
    const gf32 *UKS_WS = d.uks + c3(d, i - 1, j - 1), *UKS_EN = d.uks + c3(d, i + 1, j + 1);
    const gf32 *UKS_WN = d.uks + c3(d, i - 1, j + 1);
    const gf32 *VKS_S = d.vks + c3(d, i, j - 1), *VKS_W = d.vks + c3(d, i - 1, j);
    const gf32 *WKS_S = d.wks + c3(d, i, j - 1), *WK_N = d.wk + c3(d, i, j + 1);
    const gf32 *DPMNEW = d.dpmnew + c3(d, i, j), *DPMNEW_S = d.dpmnew + c3(d, i, j - 1);
    float dz = d.dzk[k], dxkdzk = d.dxk * dz, dykdzk = d.dyk * dz;
    float VIS = ml_visc(d, q, k, UK, VK);
    float FE, FW, FS, FN, FT, FB;
    ml_fluxes<1>(d, q, k, FE, FW, FS, FN, FT, FB);
    MlPieces o;
    ml_coeffs(d, q, k, VIS, FE, FW, FS, FN, FT, FB, o.AW1, o.AS1, o.AE1, o.AN1, o.B, o.C, o.A);
    float ap0 = d.ap0[k];
    float v = VK[k];
    o.e = ap0 * 0.5f * (VKS_S[k] + VKS[k]);
    o.f = (DPMNEW_S[k] - DPMNEW[k]) * dxkdzk;
    if (q.vegp) {
        float in2 = 0.5f * (VKS_W[k] + VKS[k]);
        float in = 0.5f * (UKS_W[k] + UKS[k]);
        float wh = vmaxf(sqrtf(in * in + in2 * in2), 0.01f);
        float veg = d.veg[c3(d, i, j) + k - 1];
        o.g = (d.cor * (d.vg - v) - veg * v * wh) * d.area * dz;
    } else {
        o.g = d.cor * (d.vg - v) * d.area * dz;
    }
    o.w = (k == q.KK + 1) ? ml_wall_uv(d, q, k, v, UKS, UKS_W, VKS, VKS_W) : 0.0f;
    o.c0 = VIS / d.dxk;
    o.X1 = (0.5f * (UKS_EN[k] + UKS_E[k]) - 0.5f * (UKS_E[k] + UKS_ES[k])) * dxkdzk;        // DUDYE
    o.X2 = (0.5f * (UKS_WN[k] + UKS_W[k]) - 0.5f * (UKS_W[k] + UKS_WS[k])) * dxkdzk;        // DUDYW
    o.X3 = (0.5f * (WK_N[k] + WKS[k]) - 0.5f * (WKS[k] + WKS_S[k])) * d.area;                // DWDYT
    // the reference takes WKS of the southern column at k, not k-1, here
    o.X4 = (0.5f * (WK_N[k - 1] + WKS[k - 1]) - 0.5f * (WKS[k - 1] + WKS_S[k])) * d.area;    // DWDYB
    return o;
}

__device__ MlPieces w1_pieces(const DomG &d, const ColInfo &q, int k) {
    int i = q.i, j = q.j;
    const gf32 *UK = d.uk + c3(d, i, j), *VK = d.vk + c3(d, i, j), *WK = d.wk + c3(d, i, j);
    const gf32 *UKS = d.uks + c3(d, i, j), *VKS = d.vks + c3(d, i, j), *WKS = d.wks + c3(d, i, j);
    const gf32 *UKS_W = d.uks + c3(d, i - 1, j), *UKS_E = d.uks + c3(d, i + 1, j);
    const gf32 *VKS_S = d.vks + c3(d, i, j - 1), *VKS_N = d.vks + c3(d, i, j + 1);
    const gf32 *DPMNEW = d.dpmnew + c3(d, i, j);
    float dz = d.dzk[k], dxkdzk = d.dxk * dz, dykdzk = d.dyk * dz;
    float VIS = ml_visc(d, q, k, UK, VK);
    float FE, FW, FS, FN, FT, FB;
    ml_fluxes<2>(d, q, k, FE, FW, FS, FN, FT, FB);
    MlPieces o;
    ml_coeffs(d, q, k, VIS, FE, FW, FS, FN, FT, FB, o.AW1, o.AS1, o.AE1, o.AN1, o.B, o.C, o.A);
    float ap0 = d.ap0[k];
    float w = WK[k];
    o.e = ap0 * 0.5f * (WKS[k - 1] + WKS[k]);
    o.f = (DPMNEW[k - 1] - DPMNEW[k]) * d.area;
    if (q.vegp) {
        float in2 = 0.5f * (VKS_S[k] + VKS[k]);
        float in = 0.5f * (UKS_W[k] + UKS[k]);
        float wh = vmaxf(sqrtf(in * in + in2 * in2), 0.01f);
        float veg = d.veg[c3(d, i, j) + k - 1];
        o.g = veg * w * wh * d.area * dz;
    } else {
        o.g = 0.0f;
    }
    o.w = 0.0f;
    float DUDZE = (0.5f * (UKS[k + 1] + UKS[k]) - 0.5f * (UKS[k] + UKS[k - 1])) * dykdzk;
    float DUDZW = (0.5f * (UKS_W[k + 1] + UKS_W[k]) - 0.5f * (UKS_W[k] + UKS_W[k - 1])) * dykdzk;
    float DVDZN = (0.5f * (VKS[k + 1] + VKS[k]) - 0.5f * (VKS[k] + VKS[k - 1])) * dxkdzk;
    float DVDZS = (0.5f * (VKS_S[k + 1] + VKS_S[k]) - 0.5f * (VKS_S[k] + VKS_S[k - 1])) * dxkdzk;
    float DWDZT = (WK[k + 1] - w) * d.area;
    float DWDZB = (w - WK[k - 1]) * d.area;
    o.c0 = 0.0f;
    o.X1 = VIS / dz * (DUDZE - DUDZW + DVDZN - DVDZS + DWDZT - DWDZB);
    o.X2 = o.X3 = o.X4 = 0.0f;
    return o;
}

// D of one cell from its pieces and the four neighbour values (and the cell's
// own old value x), in the reference's order of operations.
template <int COMP>
__device__ __forceinline__ float ml_source(const DomG &d, const ColInfo &q, int k, float dz, const MlPieces &o, float xw,
                                           float xs, float xe, float xn, float x) {
    float D;
    if (q.v512() && (COMP != 1 || q.vegp))
        D = fmaf(o.AW1, xw, o.AS1 * xs) + fmaf(o.AE1, xe, o.AN1 * xn);
    else
        D = o.AW1 * xw + o.AS1 * xs + o.AE1 * xe + o.AN1 * xn;
    D = D + o.e + o.f;
    if (COMP != 2) {
        if (COMP != 0 || !q.v512() || q.vegp) D = D + o.g;
        if (k == q.KK + 1) D -= o.w;
        if (ml_cross_on(q, k)) {
            if (COMP == 0) {
                float dykdzk = d.dyk * dz;
                float DUDXE = (xe - x) * dykdzk;
                float DUDXW = (x - xw) * dykdzk;
                D += o.c0 * (DUDXE - DUDXW + o.X1 - o.X2 + o.X3 - o.X4);
            } else {
                float dxkdzk = d.dxk * dz;
                float DVDYN = (xn - x) * dxkdzk;
                float DVDYS = (x - xs) * dxkdzk;
                D += o.c0 * (o.X1 - o.X2 + DVDYN - DVDYS + o.X3 - o.X4);
            }
        }
    } else {
        if (q.vegp) D = D - o.g;
        if (ml_cross_on(q, k)) D += o.X1;
    }
    return D;
}

template <int COMP>
__device__ __forceinline__ MlPieces ml_pieces(const DomG &d, const ColInfo &q, int k) {
    if (COMP == 0) return u1_pieces(d, q, k);
    if (COMP == 1) return v1_pieces(d, q, k);
    return w1_pieces(d, q, k);
}

// neighbour values X_W, X_S, X_E, X_N of the swept component at level k
template <int COMP>
__device__ __forceinline__ void ml_neighbours(const DomG &d, int i, int j, int k, float &xw, float &xs, float &xe,
                                              float &xn, float &x) {
    const gf32 *X = COMP == 0 ? d.uk : COMP == 1 ? d.vk : d.wk;
    xw = X[c3(d, i - 1, j) + k];
    xe = X[c3(d, i + 1, j) + k];
    xs = X[c3(d, i, j - 1) + k];
    xn = X[c3(d, i, j + 1) + k];
    x = X[c3(d, i, j) + k];
}

template <int COMP>
__device__ Coef ml_cell(const DomG &d, const ColInfo &q, int k) {
    MlPieces o = ml_pieces<COMP>(d, q, k);
    float xw, xs, xe, xn, x;
    ml_neighbours<COMP>(d, q.i, q.j, k, xw, xs, xe, xn, x);
    return Coef{o.A, o.B, o.C, ml_source<COMP>(d, q, k, d.dzk[k], o, xw, xs, xe, xn, x)};
}
__device__ Coef u1_cell(const DomG &d, const ColInfo &q, int k) { return ml_cell<0>(d, q, k); }
__device__ Coef v1_cell(const DomG &d, const ColInfo &q, int k) { return ml_cell<1>(d, q, k); }
__device__ Coef w1_cell(const DomG &d, const ColInfo &q, int k) { return ml_cell<2>(d, q, k); }

// ------------------------------------------------------------------------
// k-eps (model 2) and no-diffusion (model 0) momentum, _2.cs / _0.cs
// ------------------------------------------------------------------------
__device__ __forceinline__ float plaw_d(float D, float P) {   // D*FloatMax(0,(float)Pow5(1-0.1*P))
    double x = 1.0 - 0.1 * (double)P;
    return D * vmaxf(0.0f, (float)(x * x * x * x * x));
}

// wall function of the scalar U/V solvers (U-prognostic-microscale_2.cs:188-212)
__device__ __forceinline__ float ke_wall_uv(const DomG &d, const ColInfo &q, int k, float vel, const gf32 *UKSim,
                                            const gf32 *UKS, const gf32 *VKSim, const gf32 *VKS) {
    double a = 0.5 * (double)(UKSim[k] + UKS[k]);
    double b = 0.5 * (double)(VKSim[k] + VKS[k]);
    float wh = vmaxf((float)sqrt(a * a + b * b), 0.01f);
    float us;
    if (q.cutk < 1) {
        us = q.ht * wh;
        if ((double)q.z0s >= (double)d.dzk[k] * 0.1) {
            double a1 = 0.5 * (double)(UKSim[k + 1] + UKS[k + 1]);
            double b1 = 0.5 * (double)(VKSim[k + 1] + VKS[k + 1]);
            us = q.ht * (float)sqrt(a1 * a1 + b1 * b1);
        }
    } else {
        us = q.ho * wh;
    }
    return vel / wh * (us * us) * d.area;
}

// Everything of one k-eps (DIFF) or no-diffusion cell of U/V but the four
// neighbour values of the swept component (U-prognostic-microscale_2.cs:100-260,
// V-prognostic-microscale_2.cs, _0.cs): the coefficients, e = the old-velocity
// sum of the transient term (its double product with AP0 is formed with the
// neighbour part), f = the pressure term, g = the Coriolis term, w = the wall
// term at KKART+1, c0 = VIS/DXK and X1..X4 the neighbour-free gradients of the
// cross-diffusion term above KKART+1.
template <bool IS_V, bool DIFF>
__device__ MlPieces ke_uv_pieces(const DomG &d, const ColInfo &q, int k) {
    int i = q.i, j = q.j;
    const gf32 *X = (IS_V ? d.vk : d.uk) + c3(d, i, j);
    const gf32 *UKS = d.uks + c3(d, i, j), *VKS = d.vks + c3(d, i, j), *WKS = d.wks + c3(d, i, j);
    const gf32 *UKSim = d.uks + c3(d, i - 1, j), *VKSim = d.vks + c3(d, i - 1, j);
    float dz = d.dzk[k], dxkdzk = d.dxk * dz, dykdzk = d.dyk * dz;
    float VIS = 0, DE = 0, DW = 0, DS = 0, DN = 0, DT = 0, DB = 0;
    if (DIFF) {
        float zs = d.hok[k] - d.hok[q.KK] - dz * 0.5f;
        VIS = sqrtf(d.turb[c3(d, i, j) + k]) * zs * 0.09f;
        DE = vmaxf(VIS, d.vishmin) * dykdzk / d.dxk; DW = DE;
        DS = vmaxf(VIS, d.vishmin) * dxkdzk / d.dyk; DN = DS;
        DT = vmaxf(VIS, d.vishmin) * d.area / dz;
        if (k > q.KK + 1) DB = DT;
        if (q.as) DS = 0; if (q.an) DN = 0; if (q.aw) DW = 0; if (q.ae) DE = 0;
    }
    float FE, FW, FS, FN, FT, FB = 0;
    if (!IS_V) {
        const gf32 *VKSimjm = d.vks + c3(d, i - 1, j - 1), *VKSjm = d.vks + c3(d, i, j - 1);
        const gf32 *VKSimjp = d.vks + c3(d, i - 1, j + 1), *VKSjp = d.vks + c3(d, i, j + 1);
        const gf32 *WKSim = d.wks + c3(d, i - 1, j);
        FE = UKS[k] * dykdzk;
        FW = UKSim[k] * dykdzk;
        FS = 0.25f * (VKSim[k] + VKS[k] + VKSimjm[k] + VKSjm[k]) * dxkdzk;
        FN = 0.25f * (VKSim[k] + VKS[k] + VKSimjp[k] + VKSjp[k]) * dxkdzk;
        FT = 0.25f * (WKSim[k] + WKS[k] + WKSim[k + 1] + WKS[k + 1]) * d.area;
        if (k > q.KK + 1) FB = 0.25f * (WKSim[k] + WKS[k] + WKSim[k - 1] + WKS[k - 1]) * d.area;
    } else {
        const gf32 *UKSip = d.uks + c3(d, i + 1, j), *UKSjm = d.uks + c3(d, i, j - 1);
        const gf32 *UKSipjm = d.uks + c3(d, i + 1, j - 1), *UKSimjm = d.uks + c3(d, i - 1, j - 1);
        const gf32 *VKSjm = d.vks + c3(d, i, j - 1), *WKSjm = d.wks + c3(d, i, j - 1);
        FE = 0.25f * (UKS[k] + UKSip[k] + UKSjm[k] + UKSipjm[k]) * dykdzk;
        FW = 0.25f * (UKS[k] + UKSim[k] + UKSjm[k] + UKSimjm[k]) * dykdzk;
        FS = VKSjm[k] * dxkdzk;
        FN = VKS[k] * dxkdzk;
        FT = 0.25f * (WKSjm[k] + WKS[k] + WKSjm[k + 1] + WKS[k + 1]) * d.area;
        if (k > q.KK + 1) FB = 0.25f * (WKSjm[k] + WKS[k] + WKSjm[k - 1] + WKS[k - 1]) * d.area;
    }
    MlPieces o;
    float BIM, CIM, AE1, AW1, AS1, AN1;
    if (DIFF) {
        DE = vmaxf(DE, 0.0001f); DB = vmaxf(DB, 0.0001f); DW = vmaxf(DW, 0.0001f);
        DS = vmaxf(DS, 0.0001f); DN = vmaxf(DN, 0.0001f); DT = vmaxf(DT, 0.0001f);
        float PE = fabsf(FE / DE), PB = fabsf(FB / DB), PW = fabsf(FW / DW);
        float PS = fabsf(FS / DS), PN = fabsf(FN / DN), PT = fabsf(FT / DT);
        BIM = plaw_d(DT, PT) + vmaxf(-FT, 0.0f);
        CIM = plaw_d(DB, PB) + vmaxf(FB, 0.0f);
        AE1 = plaw_d(DE, PE) + vmaxf(-FE, 0.0f);
        AW1 = plaw_d(DW, PW) + vmaxf(FW, 0.0f);
        AS1 = plaw_d(DS, PS) + vmaxf(FS, 0.0f);
        AN1 = plaw_d(DN, PN) + vmaxf(-FN, 0.0f);
    } else {
        BIM = vmaxf(-FT, 0.0f); CIM = vmaxf(FB, 0.0f); AE1 = vmaxf(-FE, 0.0f);
        AW1 = vmaxf(FW, 0.0f); AS1 = vmaxf(FS, 0.0f); AN1 = vmaxf(-FN, 0.0f);
    }
    float ap0 = d.ap0[k];
    o.A = BIM + CIM + AW1 + AS1 + AE1 + AN1 + ap0;
    o.B = BIM; o.C = CIM; o.AW1 = AW1; o.AS1 = AS1; o.AE1 = AE1; o.AN1 = AN1;
    const gf32 *DPMNEW = d.dpmnew + c3(d, i, j);
    const gf32 *DPMNEWm = d.dpmnew + (IS_V ? c3(d, i, j - 1) : c3(d, i - 1, j));
    float ddp = DPMNEWm[k] - DPMNEW[k];
    float x = X[k];
    o.e = IS_V ? (d.vks[c3(d, i, j - 1) + k] + VKS[k]) : (UKSim[k] + UKS[k]);
    o.f = ddp * (IS_V ? dxkdzk : dykdzk);
    o.g = d.cor * ((IS_V ? d.vg : d.ug) - x) * d.area * dz;
    o.w = (k == q.KK + 1) ? ke_wall_uv(d, q, k, x, UKSim, UKS, VKSim, VKS) : 0.0f;
    o.c0 = o.X1 = o.X2 = o.X3 = o.X4 = 0.0f;
    if (DIFF && k > q.KK + 1) {
        o.c0 = VIS / d.dxk;
        if (!IS_V) {
            const gf32 *VKSipjp = d.vks + c3(d, i + 1, j + 1), *VKSjp = d.vks + c3(d, i, j + 1);
            const gf32 *VKSimjp = d.vks + c3(d, i - 1, j + 1), *VKSipjm = d.vks + c3(d, i + 1, j - 1);
            const gf32 *VKSjm = d.vks + c3(d, i, j - 1), *VKSimjm = d.vks + c3(d, i - 1, j - 1);
            const gf32 *WKip = d.wk + c3(d, i + 1, j), *WKSim = d.wks + c3(d, i - 1, j);
            o.X1 = (0.5f * (VKSipjp[k] + VKSjp[k]) - 0.5f * (VKSjp[k] + VKSimjp[k])) * dxkdzk;          // DVDXN
            o.X2 = (0.5f * (VKSipjm[k] + VKSjm[k]) - 0.5f * (VKSjm[k] + VKSimjm[k])) * dxkdzk;          // DVDXS
            o.X3 = (0.5f * (WKip[k] + WKS[k]) - 0.5f * (WKS[k] + WKSim[k])) * d.area;                   // DWDXT
            o.X4 = (0.5f * (WKip[k - 1] + WKS[k - 1]) - 0.5f * (WKS[k - 1] + WKSim[k - 1])) * d.area;   // DWDXB
        } else {
            const gf32 *UKSipjp = d.uks + c3(d, i + 1, j + 1), *UKSip = d.uks + c3(d, i + 1, j);
            const gf32 *UKSipjm = d.uks + c3(d, i + 1, j - 1), *UKSimjp = d.uks + c3(d, i - 1, j + 1);
            const gf32 *UKSimjm = d.uks + c3(d, i - 1, j - 1);
            const gf32 *WKjp = d.wk + c3(d, i, j + 1), *WKSjm = d.wks + c3(d, i, j - 1);
            o.X1 = (0.5f * (UKSipjp[k] + UKSip[k]) - 0.5f * (UKSip[k] + UKSipjm[k])) * dxkdzk;          // DUDYE
            o.X2 = (0.5f * (UKSimjp[k] + UKSim[k]) - 0.5f * (UKSim[k] + UKSimjm[k])) * dxkdzk;          // DUDYW
            o.X3 = (0.5f * (WKjp[k] + WKS[k]) - 0.5f * (WKS[k] + WKSjm[k])) * d.area;                   // DWDYT
            o.X4 = (0.5f * (WKjp[k - 1] + WKS[k - 1]) - 0.5f * (WKS[k - 1] + WKSjm[k - 1])) * d.area;   // DWDYB
        }
    }
    return o;
}

// W of the k-eps / no-diffusion solvers (W-prognostic-microscale_2.cs, _0.cs):
// e = the old-velocity sum of the transient term, f = the pressure term, X1 =
// the whole cross-diffusion term above KKART+1 (it reads no neighbour of W).
template <bool DIFF>
__device__ MlPieces ke_w_pieces(const DomG &d, const ColInfo &q, int k) {
    int i = q.i, j = q.j;
    const gf32 *WK = d.wk + c3(d, i, j);
    const gf32 *UKS = d.uks + c3(d, i, j), *VKS = d.vks + c3(d, i, j), *WKS = d.wks + c3(d, i, j);
    const gf32 *UKSim = d.uks + c3(d, i - 1, j), *UKSip = d.uks + c3(d, i + 1, j);
    const gf32 *VKSjm = d.vks + c3(d, i, j - 1), *VKSjp = d.vks + c3(d, i, j + 1);
    const gf32 *DPMNEW = d.dpmnew + c3(d, i, j);
    float dz = d.dzk[k], dxkdzk = d.dxk * dz, dykdzk = d.dyk * dz;
    float uks = UKS[k], uks_m = UKS[k - 1], vks = VKS[k], vks_m = VKS[k - 1];
    float VIS = 0, DE = 0, DW = 0, DS = 0, DN = 0, DT = 0, DB = 0;
    if (DIFF) {
        float zs = d.hok[k] - d.hok[q.KK] - dz * 0.5f;
        VIS = sqrtf(d.turb[c3(d, i, j) + k]) * zs * 0.09f;
        DE = vmaxf(VIS, d.vishmin) * dykdzk / d.dxk; DW = DE;
        DS = vmaxf(VIS, d.vishmin) * dxkdzk / d.dyk; DN = DS;
        DT = vmaxf(VIS, d.vishmin) * d.area / dz;
        if (k > q.KK + 1) DB = DT;
        if (q.as) DS = 0; if (q.an) DN = 0; if (q.aw) DW = 0; if (q.ae) DE = 0;
    }
    float FE = 0.25f * (uks + UKSip[k] + uks_m + UKSip[k - 1]) * dykdzk;
    float FW = 0.25f * (uks + UKSim[k] + uks_m + UKSim[k - 1]) * dykdzk;
    float FS = 0.25f * (vks + VKSjm[k] + vks_m + VKSjm[k - 1]) * dxkdzk;
    float FN = 0.25f * (vks + VKSjp[k] + vks_m + VKSjp[k - 1]) * dxkdzk;
    float FT = WKS[k] * d.area;
    float FB = (k > q.KK + 1) ? WKS[k - 1] * d.area : 0.0f;
    float BIM, CIM, AE1, AW1, AS1, AN1;
    if (DIFF) {
        DE = vmaxf(DE, 0.0001f); DB = vmaxf(DB, 0.0001f); DW = vmaxf(DW, 0.0001f);
        DS = vmaxf(DS, 0.0001f); DN = vmaxf(DN, 0.0001f); DT = vmaxf(DT, 0.0001f);
        float PE = fabsf(FE / DE), PB = fabsf(FB / DB), PW = fabsf(FW / DW);
        float PS = fabsf(FS / DS), PN = fabsf(FN / DN), PT = fabsf(FT / DT);
        BIM = plaw_d(DT, PT) + vmaxf(-FT, 0.0f);
        CIM = plaw_d(DB, PB) + vmaxf(FB, 0.0f);
        AE1 = plaw_d(DE, PE) + vmaxf(-FE, 0.0f);
        AW1 = plaw_d(DW, PW) + vmaxf(FW, 0.0f);
        AS1 = plaw_d(DS, PS) + vmaxf(FS, 0.0f);
        AN1 = plaw_d(DN, PN) + vmaxf(-FN, 0.0f);
    } else {
        BIM = vmaxf(-FT, 0.0f); CIM = vmaxf(FB, 0.0f); AE1 = vmaxf(-FE, 0.0f);
        AW1 = vmaxf(FW, 0.0f); AS1 = vmaxf(FS, 0.0f); AN1 = vmaxf(-FN, 0.0f);
    }
    MlPieces o;
    float ap0 = d.ap0[k];
    o.A = BIM + CIM + AW1 + AS1 + AE1 + AN1 + ap0;
    o.B = BIM; o.C = CIM; o.AW1 = AW1; o.AS1 = AS1; o.AE1 = AE1; o.AN1 = AN1;
    o.e = WKS[k - 1] + WKS[k];
    o.f = (DPMNEW[k - 1] - DPMNEW[k]) * d.area;
    o.g = o.w = o.c0 = o.X1 = o.X2 = o.X3 = o.X4 = 0.0f;
    if (DIFF && k > q.KK + 1) {
        float DUDZE = (0.5f * (UKS[k + 1] + uks) - 0.5f * (uks + uks_m)) * dykdzk;
        float DUDZW = (0.5f * (UKSim[k + 1] + UKSim[k]) - 0.5f * (UKSim[k] + UKSim[k - 1])) * dykdzk;
        float DVDZN = (0.5f * (VKS[k + 1] + vks) - 0.5f * (vks + vks_m)) * dxkdzk;
        float DVDZS = (0.5f * (VKSjm[k + 1] + VKSjm[k]) - 0.5f * (VKSjm[k] + VKSjm[k - 1])) * dxkdzk;
        float DWDZT = (WK[k + 1] - WK[k]) * d.area;
        float DWDZB = (WK[k] - WK[k - 1]) * d.area;
        o.X1 = VIS / dz * (DUDZE - DUDZW + DVDZN - DVDZS + DWDZT - DWDZB);
    }
    return o;
}

// D of a k-eps / no-diffusion cell from its pieces, the four neighbour values
// and the cell's own old value x: the neighbour sum in float, then the
// transient, pressure and Coriolis terms added in double in the reference's
// order, rounded to float; the wall term (U, V at KKART+1) and the
// cross-diffusion term follow in float.
template <int COMP, bool DIFF>
__device__ __forceinline__ float ke_source(const DomG &d, const ColInfo &q, int k, float dz, float ap0, const MlPieces &o,
                                           float xw, float xs, float xe, float xn, float x) {
    float partial = o.AW1 * xw + o.AS1 * xs + o.AE1 * xe + o.AN1 * xn;
    double t5 = (double)ap0 * 0.5 * (double)o.e;
    if (COMP == 2) {
        float D = (float)((double)partial + t5 + (double)o.f);
        if (DIFF && k > q.KK + 1) D += o.X1;
        return D;
    }
    float D = (float)((double)partial + t5 + (double)o.f + (double)o.g);
    if (k == q.KK + 1) D -= o.w;
    if (DIFF && k > q.KK + 1) {
        if (COMP == 0) {
            float dykdzk = d.dyk * dz;
            float DUDXE = (xe - x) * dykdzk;
            float DUDXW = (x - xw) * dykdzk;
            D += o.c0 * (DUDXE - DUDXW + o.X1 - o.X2 + o.X3 - o.X4);
        } else {
            float dxkdzk = d.dxk * dz;
            float DVDYN = (xn - x) * dxkdzk;
            float DVDYS = (x - xs) * dxkdzk;
            D += o.c0 * (o.X1 - o.X2 + DVDYN - DVDYS + o.X3 - o.X4);
        }
    }
    return D;
}

template <int COMP, bool DIFF>
__device__ Coef ke_cell(const DomG &d, const ColInfo &q, int k) {
    MlPieces o = (COMP == 2) ? ke_w_pieces<DIFF>(d, q, k) : ke_uv_pieces<COMP == 1, DIFF>(d, q, k);
    float xw, xs, xe, xn, x;
    ml_neighbours<COMP>(d, q.i, q.j, k, xw, xs, xe, xn, x);
    return Coef{o.A, o.B, o.C, ke_source<COMP, DIFF>(d, q, k, d.dzk[k], d.ap0[k], o, xw, xs, xe, xn, x)};
}

// ------------------------------------------------------------------------
// Momentum column kernel: one wavefront per column of diagonal t.
//   COMP 0:U 1:V 2:W;  KIND 0: mixing length (Vector<float> path, TERMP form)
//   1: k-eps, 2: no diffusion (scalar path, division form)
// ------------------------------------------------------------------------
template <int COMP, int KIND>
__global__ __launch_bounds__(WAVE) void k_momentum(const Dom *__restrict__ ds, int t, int IS, int JS) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice

    extern __shared__ float sm[];
    const int nkp = d.nkp, tid = threadIdx.x;
    // sweep ranges (outer i1, inner j1): U i1 from 3; V j1 from 3
    const int oa = (COMP == 0) ? 3 : 2, ia = (COMP == 1) ? 3 : 2;
    int a, b;
    if (!diag_col(t, oa, d.nii - 1, ia, d.njj - 1, blockIdx.x, a, b)) return;
    int i = (IS == -1) ? d.nii - a + 1 : a;
    int j = (JS == -1) ? d.njj - b + 1 : b;
    if (d.adv[c2(d, i, j)] != 1) return;
    ColInfo q = col_info(d, i, j);
    float *sA = sm, *sB = sA + nkp, *sC = sB + nkp, *sD = sC + nkp, *sX = sD + nkp;
    gf32 *X = (COMP == 0 ? d.uk : COMP == 1 ? d.vk : d.wk) + c3(d, i, j);
    for (int k = q.KSTART + tid; k <= q.VI; k += WAVE) {
        Coef cf;
        if (KIND == 0) cf = (COMP == 0) ? u1_cell(d, q, k) : (COMP == 1) ? v1_cell(d, q, k) : w1_cell(d, q, k);
        else cf = ke_cell<COMP, KIND == 1>(d, q, k);
        sA[k] = cf.A; sB[k] = cf.B; sC[k] = cf.C; sD[k] = cf.D;
    }
    for (int k = q.KSTART + tid; k <= q.VI + 1; k += WAVE) sX[k] = X[k];
    __syncthreads();
    if (tid == 0) {
        const int KKP1 = q.KK + 1;
        float Pp = 0, Qp = 0;
        for (int k = q.KSTART; k <= q.VI; ++k) {
            float A = sA[k], B = sB[k], Cc = sC[k], Dd = sD[k], P, Q;
            if (KIND == 0) {
                if (k > KKP1) { float T = 1 / (A - Cc * Pp); P = B * T; Q = (Dd + Cc * Qp) * T; }
                else { float T = 1 / A; P = B * T; Q = Dd * T; }
            } else {
                if (k > KKP1) { P = B / (A - Cc * Pp); Q = (Dd + Cc * Qp) / (A - Cc * Pp); }
                else { P = B / A; Q = Dd / A; }
            }
            sA[k] = P; sB[k] = Q;
            Pp = P; Qp = Q;
        }
        // relaxed back-substitution; U/V skip levels inside the building or its
        // upstream neighbour, W updates every level (W-prognostic-microscale_1.cs:445-448)
        int KKm = q.KK;
        if (COMP == 0) KKm = max(q.KK, d.kk[c2(d, i - 1, j)]);
        if (COMP == 1) KKm = max(q.KK, d.kk[c2(d, i, j - 1)]);
        float up = sX[q.VI + 1];
        for (int k = q.VI; k >= q.KSTART; --k) {
            float x = sX[k];
            if (COMP == 2 || KKm < k) {
                x = x + d.relax * (sA[k] * up + sB[k] - x);
                sX[k] = x;
            }
            up = x;
        }
    }
    __syncthreads();
    for (int k = q.KSTART + tid; k <= q.VI; k += WAVE) X[k] = sX[k];
}

// ------------------------------------------------------------------------
// TKE / dissipation (k-eps), TKE-prognostic-microscale.cs:25-352
// ------------------------------------------------------------------------
// Per-column pointers and neighbour data of the TKE sweep, and the per-level
// coefficients (TKE-prognostic-microscale.cs:100-337) that the sweep's
// in-place neighbours (TURB, TDISS of the four neighbour columns) do not touch.
struct TkeCol {
    size_t o, c;
    const gf32 *TURB, *TDISS, *UK, *VK, *WK;
    const gf32 *UKip, *UKjp, *UKipjp, *UKjm, *UKipjm, *VKjp, *VKip, *VKipjp, *VKim, *VKimjp, *WKip, *WKim, *WKjp, *WKjm;
    int kkip, kkim, kkjp, kkjm;
    bool cip, cim, cjp, cjm;
};
__device__ __forceinline__ TkeCol tke_col_ptrs(const DomG &d, int i, int j) {
    TkeCol t;
    t.c = c2(d, i, j);
    t.o = c3(d, i, j);
    t.TURB = d.turb + t.o; t.TDISS = d.tdiss + t.o;
    t.UK = d.uk + t.o; t.VK = d.vk + t.o; t.WK = d.wk + t.o;
    t.UKip = d.uk + c3(d, i + 1, j); t.UKjp = d.uk + c3(d, i, j + 1); t.UKipjp = d.uk + c3(d, i + 1, j + 1);
    t.UKjm = d.uk + c3(d, i, j - 1); t.UKipjm = d.uk + c3(d, i + 1, j - 1);
    t.VKjp = d.vk + c3(d, i, j + 1); t.VKip = d.vk + c3(d, i + 1, j); t.VKipjp = d.vk + c3(d, i + 1, j + 1);
    t.VKim = d.vk + c3(d, i - 1, j); t.VKimjp = d.vk + c3(d, i - 1, j + 1);
    t.WKip = d.wk + c3(d, i + 1, j); t.WKim = d.wk + c3(d, i - 1, j);
    t.WKjp = d.wk + c3(d, i, j + 1); t.WKjm = d.wk + c3(d, i, j - 1);
    return t;
}
__device__ __forceinline__ TkeCol tke_col(const DomG &d, int i, int j) {
    TkeCol t = tke_col_ptrs(d, i, j);
    t.kkip = d.kk[c2(d, i + 1, j)]; t.kkim = d.kk[c2(d, i - 1, j)];
    t.kkjp = d.kk[c2(d, i, j + 1)]; t.kkjm = d.kk[c2(d, i, j - 1)];
    t.cip = d.cutk[c2(d, i + 1, j)] > 0; t.cim = d.cutk[c2(d, i - 1, j)] > 0;
    t.cjp = d.cutk[c2(d, i, j + 1)] > 0; t.cjm = d.cutk[c2(d, i, j - 1)] > 0;
    return t;
}

struct TkeCell {
    float A, B1, C1, B2, C2, AW1, AS1, AE1, AN1;
    float w1, PRO, PBO, w2, EPS;      // ap0*TURB, production, buoyancy, ap0*TDISS, eps source
    int wf;                           // wall function replaces the back-substitution
    float wt, we;
};

__device__ __forceinline__ TkeCell tke_cell(const DomG &d, const ColInfo &q, const TkeCol &t, int k) {
    const float DXK = d.dxk, DYK = d.dyk, AREA = d.area, DXK_Rez = 1 / DXK, DYK_Rez = 1 / DYK;
    const float Ceps = 1, Ceps1 = 1.44f, Ceps2 = 1.92f;
    const int KK = q.KK;
    const gf32 *TURB = t.TURB, *TDISS = t.TDISS, *UK = t.UK, *VK = t.VK, *WK = t.WK;
    const gf32 *UKip = t.UKip, *UKjp = t.UKjp, *UKipjp = t.UKipjp, *UKjm = t.UKjm, *UKipjm = t.UKipjm;
    const gf32 *VKjp = t.VKjp, *VKip = t.VKip, *VKipjp = t.VKipjp, *VKim = t.VKim, *VKimjp = t.VKimjp;
    const gf32 *WKip = t.WKip, *WKim = t.WKim, *WKjp = t.WKjp, *WKjm = t.WKjm;
    TkeCell r;
    float dz = d.dzk[k];
    float dxkdzk = DXK * dz, dykdzk = DYK * dz, DZK_Rez = 1 / dz;
    float zs = d.hok[k] - d.hok[KK] - dz * 0.5f;
    float VIS = sqrtf(TURB[k]) * zs * 0.09f;
    VIS = mminf(VIS, 15.0f);
    float DE = mmaxf(VIS, d.vishmin) * dykdzk * DXK_Rez, DW = DE;
    float DS = mmaxf(VIS, d.vishmin) * dxkdzk * DYK_Rez, DN = DS;
    float DT = mmaxf(VIS, d.vishmin) * AREA * DZK_Rez, DB = 0;
    float DTEPS = mmaxf(VIS / Ceps, d.visvmin) * AREA * DZK_Rez, DBEPS = 0;
    if (k > KK + 1) { DB = DT; DBEPS = DTEPS; }
    if (q.as) DS = 0; if (q.an) DN = 0; if (q.aw) DW = 0; if (q.ae) DE = 0;
    float FE = UKip[k] * dykdzk, FW = UK[k] * dykdzk, FS = VK[k] * dxkdzk, FN = VKjp[k] * dxkdzk;
    float FT = WK[k + 1] * AREA, FB = 0;
    if (k > KK + 1) FB = WK[k] * AREA;
    DE = mmaxf(DE, 0.0001f); DB = mmaxf(DB, 0.0001f); DW = mmaxf(DW, 0.0001f);
    DS = mmaxf(DS, 0.0001f); DN = mmaxf(DN, 0.0001f); DT = mmaxf(DT, 0.0001f);
    DBEPS = mmaxf(DBEPS, 0.0001f); DTEPS = mmaxf(DTEPS, 0.0001f);
    float PE = fabsf(FE / DE), PB = fabsf(FB / DB), PW = fabsf(FW / DW);
    float PS = fabsf(FS / DS), PN = fabsf(FN / DN), PT = fabsf(FT / DT);
    float PBEPS = fabsf(FN / DBEPS);   // the reference uses FN here
    float PTEPS = fabsf(FT / DTEPS);
    float x;
    x = 1 - 0.1f * PT; float BIM = DT * mmaxf(0, x * x * x * x * x) + mmaxf(-FT, 0);
    x = 1 - 0.1f * PB; float CIM = DB * mmaxf(0, x * x * x * x * x) + mmaxf(FB, 0);
    x = 1 - 0.1f * PE; float AE1 = DE * mmaxf(0, x * x * x * x * x) + mmaxf(-FE, 0);
    x = 1 - 0.1f * PW; float AW1 = DW * mmaxf(0, x * x * x * x * x) + mmaxf(FW, 0);
    x = 1 - 0.1f * PS; float AS1 = DS * mmaxf(0, x * x * x * x * x) + mmaxf(FS, 0);
    x = 1 - 0.1f * PN; float AN1 = DN * mmaxf(0, x * x * x * x * x) + mmaxf(-FN, 0);
    // the reference weights the eps system with DT/DB, not DTEPS/DBEPS
    x = 1 - 0.1f * PTEPS; float BIMEPS = DT * mmaxf(0, x * x * x * x * x) + mmaxf(-FT, 0);
    x = 1 - 0.1f * PBEPS; float CIMEPS = DB * mmaxf(0, x * x * x * x * x) + mmaxf(FB, 0);
    float ap0 = d.ap0[k];
    float AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + ap0;
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
    float s1 = DUDY + DVDX, s2 = DUDZ + DWDX, s3 = DVDZ + DWDY;
    float PROTURB = VIS * AREA * dz *
                    (2.0f * (DUDX * DUDX) + 2.0f * (DVDY * DVDY) + 2.0f * (DWDZ * DWDZ) + s1 * s1 + s2 * s2 + s3 * s3 - TDISS[k]);
    float PROBOU = (float)((double)(-d.potdtdz) * 9.81 / 283.0 * (double)VIS * 1.35 * (double)AREA * (double)dz);
    r.A = AIM; r.B1 = BIM; r.C1 = CIM; r.B2 = BIMEPS; r.C2 = CIMEPS;
    r.AW1 = AW1; r.AS1 = AS1; r.AE1 = AE1; r.AN1 = AN1;
    r.w1 = ap0 * TURB[k]; r.PRO = PROTURB; r.PBO = PROBOU;
    r.w2 = ap0 * TDISS[k];
    r.EPS = TDISS[k] / TURB[k] * (Ceps1 * PROTURB - Ceps2 * TDISS[k]);
    // wall functions of the back-substitution (they read only UK/VK/WK)
    int wf = 0; float wt = 0, we = 0;
    if (KK < k) {
        float us;
        if ((k == KK + 1) && (q.cutk == 0)) {
            wf = 1;
            const size_t cc = t.c;
            if (d.tke_br[cc]) {
                double s = sqrt((0.5 * (double)(UK[k + 1] + UKip[k + 1])) * (0.5 * (double)(UK[k + 1] + UKip[k + 1])) +
                                (0.5 * (double)(VK[k + 1] + VKjp[k + 1])) * (0.5 * (double)(VK[k + 1] + VKjp[k + 1])));
                us = (float)(d.tke_a[cc] * s);
            } else {
                double s = sqrt((0.5 * (double)(UK[k + 1] + UKip[k])) * (0.5 * (double)(UK[k + 1] + UKip[k])) +
                                (0.5 * (double)(VK[k] + VKjp[k])) * (0.5 * (double)(VK[k] + VKjp[k])));
                us = (float)(d.tke_a[cc] * s);
            }
            wt = (float)(3.33 * (double)(us * us));
            we = (float)((double)(us * us * us) / d.tke_den[cc]);
        } else if ((k == KK + 1) && (q.cutk > 0)) {
            wf = 1;
            double a1 = 0.5 * (double)(UK[k] + UKip[k]), b1 = 0.5 * (double)(VK[k] + VKjp[k]);
            us = q.ho * (float)sqrt(a1 * a1 + b1 * b1);
            wt = (float)(3.33 * (double)(us * us));
            we = (float)((double)(us * us * us) / ((double)dz * 0.5 * 0.4));
        } else if (((k == t.kkip + 1) && t.cip) || ((k == t.kkim + 1) && t.cim)) {
            wf = 1;
            double a1 = 0.5 * (double)(WK[k] + WK[k + 1]), b1 = 0.5 * (double)(VK[k] + VKjp[k]);
            us = (float)((double)d.ufx * sqrt(a1 * a1 + b1 * b1));
            wt = (float)(3.33 * (double)(us * us));
            we = (float)((double)(us * us * us) / ((double)d.bz0 * 0.4));
        } else if (((k == t.kkjp + 1) && t.cjp) || ((k == t.kkjm + 1) && t.cjm)) {
            wf = 1;
            double a1 = 0.5 * (double)(WK[k] + WK[k + 1]), b1 = 0.5 * (double)(UK[k] + UKip[k]);
            us = (float)((double)d.ufx * sqrt(a1 * a1 + b1 * b1));
            wt = (float)(3.33 * (double)(us * us));
            we = (float)((double)(us * us * us) / ((double)d.bz0 * 0.4));
        }
    }
    r.wf = wf; r.wt = wt; r.we = we;
    return r;
}

__global__ __launch_bounds__(WAVE) void k_tke(const Dom *__restrict__ ds, int t, int IS, int JS) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice

    extern __shared__ float sm[];
    const int nkp = d.nkp, tid = threadIdx.x;
    int a, b;
    if (!diag_col(t, 2, d.nii - 1, 2, d.njj - 1, blockIdx.x, a, b)) return;
    int i = (IS == -1) ? d.nii - a + 1 : a;
    int j = (JS == -1) ? d.njj - b + 1 : b;
    size_t c = c2(d, i, j);
    if (d.adv[c] != 1) return;
    ColInfo q = col_info(d, i, j);
    float *sA = sm, *sB1 = sA + nkp, *sC1 = sB1 + nkp, *sD1 = sC1 + nkp, *sB2 = sD1 + nkp, *sC2 = sB2 + nkp,
          *sD2 = sC2 + nkp, *sT = sD2 + nkp, *sE = sT + nkp, *sWT = sE + nkp, *sWE = sWT + nkp;
    int *sWF = (int *)(sWE + nkp);
    const TkeCol tc = tke_col(d, i, j);
    size_t o = tc.o;
    const gf32 *TURB = tc.TURB, *TDISS = tc.TDISS;
    const gf32 *TURBim = d.turb + c3(d, i - 1, j), *TURBip = d.turb + c3(d, i + 1, j);
    const gf32 *TURBjm = d.turb + c3(d, i, j - 1), *TURBjp = d.turb + c3(d, i, j + 1);
    const gf32 *TDISSim = d.tdiss + c3(d, i - 1, j), *TDISSip = d.tdiss + c3(d, i + 1, j);
    const gf32 *TDISSjm = d.tdiss + c3(d, i, j - 1), *TDISSjp = d.tdiss + c3(d, i, j + 1);
    int KK = q.KK;
    for (int k = q.KSTART + tid; k <= q.VI; k += WAVE) {
        const TkeCell e = tke_cell(d, q, tc, k);
        float DIMTURB = e.AW1 * TURBim[k] + e.AS1 * TURBjm[k] + e.AE1 * TURBip[k] + e.AN1 * TURBjp[k] + e.w1 + e.PRO + e.PBO;
        float DIMTDISS = e.AW1 * TDISSim[k] + e.AS1 * TDISSjm[k] + e.AE1 * TDISSip[k] + e.AN1 * TDISSjp[k] + e.w2 + e.EPS;
        sA[k] = e.A; sB1[k] = e.B1; sC1[k] = e.C1; sD1[k] = DIMTURB; sB2[k] = e.B2; sC2[k] = e.C2; sD2[k] = DIMTDISS;
        sWF[k] = e.wf; sWT[k] = e.wt; sWE[k] = e.we;
    }
    for (int k = q.KSTART + tid; k <= q.VI + 1; k += WAVE) { sT[k] = TURB[k]; sE[k] = TDISS[k]; }
    __syncthreads();
    if (tid == 0) {
        float P1 = 0, Q1 = 0, P2 = 0, Q2 = 0;
        for (int k = q.KSTART; k <= q.VI; ++k) {
            float A = sA[k];
            float p1, q1, p2, q2;
            if (k > KK + 1) {
                p1 = sB1[k] / (A - sC1[k] * P1);
                q1 = (sD1[k] + sC1[k] * Q1) / (A - sC1[k] * P1);
                p2 = sB2[k] / (A - sC2[k] * P2);
                q2 = (sD2[k] + sC2[k] * Q2) / (A - sC2[k] * P2);
            } else {
                p1 = sB1[k] / A; q1 = sD1[k] / A; p2 = sB2[k] / A; q2 = sD2[k] / A;
            }
            sB1[k] = p1; sD1[k] = q1; sB2[k] = p2; sD2[k] = q2;
            P1 = p1; Q1 = q1; P2 = p2; Q2 = q2;
        }
        float tup = sT[q.VI + 1], eup = sE[q.VI + 1];
        for (int k = q.VI; k >= q.KSTART; --k) {
            float tv = sT[k], ev = sE[k];
            if (KK < k) {
                if (sWF[k]) { tv = sWT[k]; ev = sWE[k]; }
                else { tv = sB1[k] * tup + sD1[k]; ev = sB2[k] * eup + sD2[k]; }
                tv = (float)mmaxd((double)tv, 0.001);
                ev = (float)mmaxd((double)ev, 0.0000001);
                sT[k] = tv; sE[k] = ev;
            }
            tup = tv; eup = ev;
        }
    }
    __syncthreads();
    gf32 *TW = d.turb + o, *EW = d.tdiss + o;
    for (int k = q.KSTART + tid; k <= q.VI; k += WAVE) { TW[k] = sT[k]; EW[k] = sE[k]; }
}

// ------------------------------------------------------------------------
// Strip wavefront path (mixing-length model and pressure)
// ------------------------------------------------------------------------
// A workgroup owns ROWS consecutive sweep rows (b) with one wave per row plus
// a chain wave.  Row wave r owns row b0 + r and at local step s handles column
// a = oa + s - r, so every local step is one anti-diagonal of the strip.  Each
// step is three phases separated by workgroup barriers: (1) every active row
// wave assembles its column's source terms (lanes over k) into LDS, (2) the
// chain wave runs the serial Thomas recurrences of all the strip's columns
// (lane = row) while the row waves prefetch their next columns, (3) the row
// waves write their columns back.  Neighbours updated in the previous step are
// exchanged through LDS column buffers.
#define ROWS STRIP_ROWS
#define STRIP_WAVES (STRIP_ROWS + 1)

// coefficient records written by k_ml_coef and read by the strip sweeps; per
// cell, two records, each written whole, in one pass:
//   planes: U, V {AW, AS, AE, AN | f, g, e, c0 | X1, X2, X3, X4} (REC_UV floats),
//           W    {AW, AS, AE, AN | f, g, e, X1}                   (REC_W floats),
//   written by the first pass of k_ml_coef;
//   rec3 = {C, T, P}, written after the P/TERMP recurrence.
// The U/V wall term exists only at KKART+1 and is kept per column (wcol).
#define REC_UV 12
#define REC_W 8
template <int COMP> __device__ __forceinline__ constexpr int rec_floats() { return COMP == 2 ? REC_W : REC_UV; }

__device__ __forceinline__ void sweep_to_ij(const DomG &d, int comp_is_pressure, int a, int b, int IS, int JS, int &i, int &j) {
    if (comp_is_pressure) {   // a = j1, b = i1
        j = (JS == -1) ? d.njj - a + 1 : a;
        i = (IS == -1) ? d.nii - b + 1 : b;
    } else {                  // a = i1, b = j1
        i = (IS == -1) ? d.nii - a + 1 : a;
        j = (JS == -1) ? d.njj - b + 1 : b;
    }
}

// workgroup barrier that orders LDS only; global loads stay in flight across it
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// odd LDS row stride: lanes of one wave that walk different columns at the
// same relative level hit different banks
__device__ __forceinline__ int lds_stride(int nkp) { return nkp | 1; }

#define COEF_COLS 16
#define SCOL 16           // words of column data per column in k_ml_coef
// All neighbour-independent work of one mixing-length sweep for every column
// at once.  A block of 4 waves takes COEF_COLS consecutive columns of the
// sweep; their active levels are laid out one after another and dealt to the
// block's 256 lanes (a lane per level, whatever the column heights), which
// compute the coefficients.  One wave then runs the P/TERMP recurrence of all
// COEF_COLS columns at once (lane = column), and the lanes write the last part
// of each cell's record.  Each record is written once, as float4 / float3 stores.
// WFIX: the solver width when every situation of the launch has it (the
// width-dependent arithmetic then folds at compile time), 0 to read it per
// situation.  KIND 0: mixing length (rec3 = {C, TERMP, P}); KIND 1: k-eps and
// KIND 2: no diffusion, whose solvers divide by the TDMA denominator instead
// of multiplying by its reciprocal (rec3 = {C, denominator, P}).
template <int COMP, int WFIX, int KIND = 0>
__global__ __launch_bounds__(4 * WAVE) void k_ml_coef(const Dom *__restrict__ ds, int IS, int JS) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice
    gf32 *planes = d.planes;

    extern __shared__ float sm[];
    const int nkp = d.nkp, nkl = lds_stride(nkp);
    const int lane = threadIdx.x % WAVE;
    const int oa = (COMP == 0) ? 3 : 2, ia = (COMP == 1) ? 3 : 2;
    const int na = d.nii - 1 - oa + 1, nb = d.njj - 1 - ia + 1;
    // per column c: A, B, C over k at sm + c*3*nkl, and the column's data
    // (read once per column here, not once per level)
    int *scol = (int *)(sm + 3 * COEF_COLS * nkl);         // [COEF_COLS][SCOL]
    int *stot = scol + SCOL * COEF_COLS;
    const int blk = xcd_block((na * nb + COEF_COLS - 1) / COEF_COLS);
    if (blk < 0) return;                                      // uniform over the block
    const int col0 = blk * COEF_COLS;
    if (threadIdx.x < WAVE) {
        int c = lane, n = 0;
        if (c < COEF_COLS) {
            int col = col0 + c;
            int i = 0, j = 0;
            Meta m;
            m.adv = false;
            if (col < na * nb) {
                sweep_to_ij(d, 0, oa + col / nb, ia + col % nb, IS, JS, i, j);
                m = unpack_meta(ldg4i(d.meta + c2(d, i, j)));
            }
            if (m.adv) {
                n = m.VI - m.KSTART + 1;
                const ColInfo q = col_info(d, i, j);
                int *sc = scol + c * SCOL;
                sc[7] = (q.as ? 1 : 0) | (q.an ? 2 : 0) | (q.aw ? 4 : 0) | (q.ae ? 8 : 0) | (q.vegp ? 16 : 0);
                sc[8] = __float_as_int(q.cutk); sc[9] = __float_as_int(q.cov); sc[10] = __float_as_int(q.ht);
                sc[11] = __float_as_int(q.ho); sc[12] = __float_as_int(q.z0s);
            }
            scol[c * SCOL + 0] = m.KSTART; scol[c * SCOL + 1] = m.KK; scol[c * SCOL + 2] = m.VI;
            scol[c * SCOL + 3] = i; scol[c * SCOL + 4] = j; scol[c * SCOL + 6] = m.adv ? 1 : 0;
        }
        // inclusive prefix sum of the level counts over the lanes
        int incl = n;
        for (int off = 1; off < WAVE; off <<= 1) {
            int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        if (c < COEF_COLS) scol[c * SCOL + 5] = incl - n;
        if (lane == WAVE - 1) stot[0] = incl;
    }
    lds_barrier();
    const int total = stot[0];
    // the column holding item t is the last c with first[c] <= t: first[] does
    // not decrease, and every later column starts beyond t
    auto column_of = [&](int t) {
        int lo = 0;
        for (int step = COEF_COLS / 2; step > 0; step >>= 1)
            if (scol[(lo + step) * SCOL + 5] <= t) lo += step;
        return lo;
    };
    for (int t = threadIdx.x; t < total; t += 4 * WAVE) {
        const int c = column_of(t);
        const int *sc = scol + c * SCOL;
        const int k = sc[0] + (t - sc[5]);
        ColInfo q;
        q.KSTART = sc[0]; q.KK = sc[1]; q.VI = sc[2]; q.i = sc[3]; q.j = sc[4];
        const int fl = sc[7];
        q.as = fl & 1; q.an = (fl >> 1) & 1; q.aw = (fl >> 2) & 1; q.ae = (fl >> 3) & 1; q.vegp = (fl >> 4) & 1;
        q.cutk = __int_as_float(sc[8]); q.cov = __int_as_float(sc[9]); q.ht = __int_as_float(sc[10]);
        q.ho = __int_as_float(sc[11]); q.z0s = __int_as_float(sc[12]);
        q.W = WFIX ? WFIX : d.simd;
        MlPieces p;
        if constexpr (KIND != 0)
            p = (COMP == 2) ? ke_w_pieces<KIND == 1>(d, q, k) : ke_uv_pieces<COMP == 1, KIND == 1>(d, q, k);
        else p = ml_pieces<COMP>(d, q, k);
        float *sA = sm + (size_t)c * 3 * nkl;
        sA[k] = p.A; sA[nkl + k] = p.B; sA[2 * nkl + k] = p.C;
        gf32 *P = planes + (c3(d, q.i, q.j) + k) * rec_floats<COMP>();
        stg4(P, p.AW1, p.AS1, p.AE1, p.AN1);
        if (COMP == 2) {
            stg4(P + 4, p.f, p.g, p.e, p.X1);
        } else {
            stg4(P + 4, p.f, p.g, p.e, p.c0);
            stg4(P + 8, p.X1, p.X2, p.X3, p.X4);
            if (k == q.KK + 1) d.wcol[c2(d, q.i, q.j)] = p.w;
        }
    }
    lds_barrier();
    if (threadIdx.x < COEF_COLS && scol[threadIdx.x * SCOL + 6]) {
        const int c = threadIdx.x;
        float *__restrict__ A = sm + (size_t)c * 3 * nkl;
        float *__restrict__ Bq = A + nkl;
        float *__restrict__ Cq = Bq + nkl;
        const int KSTART = scol[c * SCOL + 0], KKP1 = scol[c * SCOL + 1] + 1, VI = scol[c * SCOL + 2];
        float Pp = 0;
        for (int k = KSTART; k <= VI; ++k) {
            float T, P;
            if (KIND != 0) {
                T = (k > KKP1) ? A[k] - Cq[k] * Pp : A[k];
                P = Bq[k] / T;
            } else {
                T = (k > KKP1) ? 1 / (A[k] - Cq[k] * Pp) : 1 / A[k];
                P = Bq[k] * T;
            }
            A[k] = T;          // T replaces A, P replaces B
            Bq[k] = P;
            Pp = P;
        }
    }
    lds_barrier();
    for (int t = threadIdx.x; t < total; t += 4 * WAVE) {
        const int c = column_of(t);
        const int k = scol[c * SCOL + 0] + (t - scol[c * SCOL + 5]);
        const float *sA = sm + (size_t)c * 3 * nkl;
        gf32 *P = d.rec3 + (c3(d, scol[c * SCOL + 3], scol[c * SCOL + 4]) + k) * 3;
        stg3(P, sA[2 * nkl + k], sA[k], sA[nkl + k]);                   // C, T (or the denominator), P
    }
}


// Serial Thomas recurrences of one column (one lane), software-pipelined: the
// operands of the next CH levels are read from LDS while the current CH levels
// are computed.  Pointers are relative to the first level of the recurrence and
// the loops run over n levels; the LDS slots carry 2*CH floats of padding on
// both sides, so the look-ahead reads need no clamping (their values are unused).
#define CH 8
#define SLOT_PAD (2 * CH)
__device__ __forceinline__ int lds_slot(int nkp) { return (nkp + 2 * SLOT_PAD) | 1; }

// The chunk loops below are unrolled over the largest column the strip
// kernels hold (MAXCH chunks of CH levels) and leave early at the
// wave-wide longest chain (nmax): straight-line code lets the compiler wait
// for exactly the LDS reads a chunk needs, while the reads of the next chunk
// are in flight.  Reads stay within CH levels past a lane's own range (the
// slot padding); a lane that is done re-reads its first chunk.  A lane's
// last chunk stores up to CH-1 levels past its range (into the slot padding or
// levels nobody reads), and chunks after its end go to a per-lane sink of CH
// words, so the chunks are branch-free.  The back-substitution writes its
// result to a separate slot (the old column stays readable and untouched below
// the range).
#define MAXCH (128 / CH)

// forward: Q[t] = (D[t] + C[t]*Q[t-1]) * T[t], t = 0..n-1, Q[-1] = q0
template <bool DIV>
__device__ __forceinline__ void chain_fwd(const float *__restrict__ D, const float *__restrict__ C,
                                          const float *__restrict__ T, float *__restrict__ Q, int n, int nmax, float q0,
                                          float *sink) {
    float Qp = q0;
    float cd[CH], cc[CH], ct[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) { cd[u] = D[u]; cc[u] = C[u]; ct[u] = T[u]; }
#pragma unroll
    for (int h = 0; h < MAXCH; ++h) {
        const int t = h * CH;
        if (t >= nmax) break;
        float nd[CH], nc[CH], nt[CH];
        const int o = (t + CH < n) ? t + CH : 0;          // lanes already past their end re-read level 0
#pragma unroll
        for (int u = 0; u < CH; ++u) { nd[u] = D[o + u]; nc[u] = C[o + u]; nt[u] = T[o + u]; }
        __builtin_amdgcn_sched_barrier(0);                // keep the next chunk's reads ahead of this chunk
        float *Qw = (t < n) ? Q : sink - t;              // a lane past its end stores to its sink
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            float q = DIV ? (cd[u] + cc[u] * Qp) / ct[u] : (cd[u] + cc[u] * Qp) * ct[u];
            Qw[t + u] = q;
            Qp = q;
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) { cd[u] = nd[u]; cc[u] = nc[u]; ct[u] = nt[u]; }
    }
}

// relaxed back-substitution downwards: X[-t] += relax*(P[-t]*X[-t+1] + Q[-t] - X[-t]), t = 0..n-1
__device__ __forceinline__ void chain_back(const float *__restrict__ P, const float *__restrict__ Q,
                                           const float *__restrict__ X, float *__restrict__ Xn, float relax, int n,
                                           int nmax, float *sink) {
    float up = X[1];
    float cp[CH], cq[CH], cx[CH];
    // each chunk addresses its CH levels from the lowest one upwards, so that
    // the accesses pair up with positive offsets (ds_read2/ds_write2 from one
    // base) and few enough reads are in flight for a partial LDS wait
    {
        const float *Pb = P - (CH - 1), *Qb = Q - (CH - 1), *Xb = X - (CH - 1);
#pragma unroll
        for (int u = 0; u < CH; ++u) { cp[u] = Pb[CH - 1 - u]; cq[u] = Qb[CH - 1 - u]; cx[u] = Xb[CH - 1 - u]; }
    }
#pragma unroll
    for (int h = 0; h < MAXCH; ++h) {
        const int t = h * CH;
        if (t >= nmax) break;
        float np[CH], nq[CH], nx[CH];
        const int o = (t + CH < n) ? t + CH : 0;
        const float *Pb = P - (o + CH - 1), *Qb = Q - (o + CH - 1), *Xb = X - (o + CH - 1);
#pragma unroll
        for (int u = 0; u < CH; ++u) { np[u] = Pb[CH - 1 - u]; nq[u] = Qb[CH - 1 - u]; nx[u] = Xb[CH - 1 - u]; }
        __builtin_amdgcn_sched_barrier(0);                // keep the next chunk's reads ahead of this chunk
        float *Xw = (t < n) ? Xn - (t + CH - 1) : sink;   // a lane past its end stores to its sink
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            float x = cx[u] + relax * (cp[u] * up + cq[u] - cx[u]);
            Xw[CH - 1 - u] = x;
            up = x;
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) { cp[u] = np[u]; cq[u] = nq[u]; cx[u] = nx[u]; }
    }
}

// pressure, downwards from the top: Q[-t] = (A[-t]*Q[-t+1] + D[-t]) * T[-t] with
// A (= ATP) held in Q on entry and Q[1] = qtop, t = 0..n-1
__device__ __forceinline__ void chain_pdown(const float *__restrict__ D, const float *__restrict__ T, float *__restrict__ Q,
                                            float qtop, int n, int nmax, float *sink) {
    float Qn = qtop;
    float ca[CH], cd[CH], ct[CH];
    // chunks addressed from their lowest level upwards, as in chain_back
    {
        const float *Qb = Q - (CH - 1), *Db = D - (CH - 1), *Tb = T - (CH - 1);
#pragma unroll
        for (int u = 0; u < CH; ++u) { ca[u] = Qb[CH - 1 - u]; cd[u] = Db[CH - 1 - u]; ct[u] = Tb[CH - 1 - u]; }
    }
#pragma unroll
    for (int h = 0; h < MAXCH; ++h) {
        const int t = h * CH;
        if (t >= nmax) break;
        float na[CH], nd[CH], nt[CH];
        const int o = (t + CH < n) ? t + CH : 0;
        const float *Qb = Q - (o + CH - 1), *Db = D - (o + CH - 1), *Tb = T - (o + CH - 1);
#pragma unroll
        for (int u = 0; u < CH; ++u) { na[u] = Qb[CH - 1 - u]; nd[u] = Db[CH - 1 - u]; nt[u] = Tb[CH - 1 - u]; }
        __builtin_amdgcn_sched_barrier(0);                // keep the next chunk's reads ahead of this chunk
        float *Qw = (t < n) ? Q - (t + CH - 1) : sink;
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            float q = (ca[u] * Qn + cd[u]) * ct[u];
            Qw[CH - 1 - u] = q;
            Qn = q;
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) { ca[u] = na[u]; cd[u] = nd[u]; ct[u] = nt[u]; }
    }
}

// pressure, upwards: N[t] = P[t]*N[t-1] + Q[t], t = 0..n-1, N[-1] = bottom
__device__ __forceinline__ void chain_pup(const float *__restrict__ P, const float *__restrict__ Q, float *__restrict__ N,
                                          float bottom, int n, int nmax, float *sink) {
    float dp = bottom;
    float cp[CH], cq[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) { cp[u] = P[u]; cq[u] = Q[u]; }
#pragma unroll
    for (int h = 0; h < MAXCH; ++h) {
        const int t = h * CH;
        if (t >= nmax) break;
        float np[CH], nq[CH];
        const int o = (t + CH < n) ? t + CH : 0;
#pragma unroll
        for (int u = 0; u < CH; ++u) { np[u] = P[o + u]; nq[u] = Q[o + u]; }
        __builtin_amdgcn_sched_barrier(0);                // keep the next chunk's reads ahead of this chunk
        float *Nw = (t < n) ? N : sink - t;
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            float x = cp[u] * dp + cq[u];
            Nw[t + u] = x;
            dp = x;
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) { cp[u] = np[u]; cq[u] = nq[u]; }
    }
}

// largest value of v over the lanes of the wave that pass `on` (0 if none)
__device__ __forceinline__ int wave_max(int v, bool on) {
    int m = on ? v : 0;
    for (int off = WAVE / 2; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    return __builtin_amdgcn_readfirstlane(m);
}

// per-level values of one column that do not change during the sweep, loaded
// one local step ahead
struct MlPre {
    float aw, as, ae, an, e, f, g, c0, x1, x2, x3, x4, c, t, p;
    float x, xa, xb;      // own old value; old neighbours in the +IS (xa) and +JS (xb) directions
    float ni, nj;         // -IS / -JS neighbours, prefetched only when another strip or the boundary holds them
    float ap0;            // transient coefficient of the level (k-eps)
};

// p + byte_off with a 32-bit byte offset: with a uniform base the access can
// take the scalar-base form (base in scalar registers, 32-bit lane offset), so
// no 64-bit address arithmetic per lane
__device__ __forceinline__ const gf32 *at_bytes(const gf32 *base, unsigned byte_off) {
    typedef const char __attribute__((address_space(1))) gc8;
    return (const gf32 *)((const gc8 *)base + byte_off);
}
template <int COMP, int KIND>
__device__ __forceinline__ void ml_prefetch(const DomG &d, const gf32 *Pc, const gf32 *R3, const gf32 *Xc,
                                            const gf32 *Xa, const gf32 *Xb, unsigned k, MlPre &r) {
    // Pc, R3, Xc, Xa, Xb: the column's records and values and its +IS / +JS
    // neighbours (uniform bases, so the level's byte offset is the only
    // per-lane part of each address)
    const gf32 *P = at_bytes(Pc, k * (unsigned)(4 * rec_floats<COMP>()));
    float4 v0 = ldg4(P), v2 = ldg4(P + 4);
    float3 v1 = ldg3(at_bytes(R3, k * 12u));
    r.aw = v0.x; r.as = v0.y; r.ae = v0.z; r.an = v0.w;
    r.c = v1.x; r.t = v1.y; r.p = v1.z;
    r.f = v2.x; r.g = v2.y; r.e = v2.z;
    if (COMP == 2) {
        r.x1 = v2.w;
        r.c0 = r.x2 = r.x3 = r.x4 = 0.0f;
    } else {
        float4 v3 = ldg4(P + 8);
        r.c0 = v2.w;
        r.x1 = v3.x; r.x2 = v3.y; r.x3 = v3.z; r.x4 = v3.w;
    }
    r.x = *at_bytes(Xc, 4u * k);
    r.xa = *at_bytes(Xa, 4u * k);
    r.xb = *at_bytes(Xb, 4u * k);
    if (KIND != 0) r.ap0 = d.ap0[k];
}

__device__ __forceinline__ ColInfo meta_col(const Meta &m, int i, int j, int W) {
    ColInfo q;
    q.i = i; q.j = j; q.KK = m.KK; q.KSTART = m.KSTART; q.VI = m.VI; q.W = W; q.vegp = m.vegp;
    q.as = q.an = q.aw = q.ae = false; q.cutk = q.cov = q.ht = q.ho = q.z0s = 0.0f;
    return q;
}

// prefetched per-level values of one column of the pressure sweep
struct PrPre {
    float div, t, p, pa, pb;       // DIV / DTIME, TERMP, PIMP, old DPM in the +IS (i) and +JS (j) directions
    float dn;                      // DPMNEW of the cell (read-modify-written one step later)
    float ni, nj;                  // -IS / -JS neighbours when another strip or the boundary holds them
};

// ------------------------------------------------------------------------
// Strip-pipelined sweeps.  Strip B reads the last row of strip B-1 (its new
// values, and for the pressure sweep the QIMP row scratch) and nothing else
// from other strips, so it only has to stay ROWS steps behind strip B-1: one
// launch holds the whole sweep, whose serial path is about (a range) +
// ROWS * (number of strips) local steps.
//
// Synchronisation, per situation and sweep kind, in int sync[...]:
//   [0] ticket counter: strips are numbered in the order their workgroups
//       start, so the strip a workgroup waits for is held by a workgroup
//       that is already running (no dependence on dispatch order);
//   [2 + B] steps of strip B whose hand-off stores are complete;
// the blocks are cleared at the start of every iteration.  d.sync[0] is the
// situation's sticky error flag: set when a wait exceeded the time limit.
// The wave that stores the last row's results (the chain wave in the momentum
// strips, the last row wave otherwise) publishes the strip's progress after
// waiting for those write-through stores; row 0 of the next strip polls it and
// then reads the handed-over values with agent-scope loads.  Every
// wait is bounded (STRIP_TIMEOUT_TICKS of the constant 100 MHz clock): on
// timeout the strip sets the error flag and goes on without waiting, so the
// kernel always ends; the host reports the error.
// ------------------------------------------------------------------------
#define SYNC_HDR 2
// a strip publishes its progress every STRIP_PUBLISH steps (and at its end); a
// longer interval adds up to STRIP_PUBLISH-1 steps of lag per strip.  The
// publish is a wait for the strip's own write-through stores, and the reader
// takes the hand-off values with agent-scope loads, so neither side writes back
// or invalidates an L2.  Measured on the 8-direction city batch (ms per launch,
// publishing every 1 / 2 / 3 / 4 steps): U and V sweeps 2.8 / 2.9 / 2.95 / 3.0,
// W 2.77 / 2.84 / 2.96 / 3.06, pressure 2.62 / 2.78 / 2.81 / 2.95.  (With an
// agent-scope release fence, which writes back the whole L2 of the XCD, every
// publish cost about 4 us: 7.6-8.2 ms for U and V at every step, 4.1 at every 6th.)
#ifndef STRIP_PUBLISH
#define STRIP_PUBLISH 1
#endif
#ifndef ML_PUBLISH_UV
#define ML_PUBLISH_UV 1
#endif
#ifndef ML_PUBLISH_W
#define ML_PUBLISH_W 1
#endif
template <int COMP> __device__ __forceinline__ constexpr int ml_publish() { return COMP == 2 ? ML_PUBLISH_W : ML_PUBLISH_UV; }
#define STRIP_TIMEOUT_TICKS 200000000ull

__device__ __forceinline__ int load_agent(const gi32 *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent(gi32 *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The values a strip hands to the next one (the new last row, the pressure row
// scratch) are read with agent-scope loads: they are served coherently across
// the XCDs' L2 caches, so the reader needs no L2 invalidation (an agent-scope
// acquire fence, which would also drop every other workgroup's cached lines on
// that XCD) - only the order after its progress poll, which the poll's
// control dependency and a compiler fence give.
__device__ __forceinline__ float ld_handoff(const gf32 *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct StripWait {
    const gi32 *prev;              // progress of the previous strip (null for strip 0)
    gi32 *error;
    int seen;                      // progress of the previous strip known so far
    int cap;                       // steps of a strip
    unsigned long long deadline;
    // called by every lane of one wave; returns once the previous strip has
    // completed min(target, cap) steps (or the time limit passed)
    __device__ __forceinline__ void wait(int target) {
        target = min(target, cap);
        if (seen >= target) return;
        for (int spins = 0;; ++spins) {
            int p = __builtin_amdgcn_readfirstlane(load_agent(prev));
            if (p >= target) { seen = p; break; }
            if (wall_clock64() > deadline || spins > (1 << 22)) {
                store_agent(error, 1);
                seen = 0x7fffffff;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        // the hand-off loads that follow are agent-scope loads (ld_handoff):
        // keep the compiler from moving them above the poll
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};

// The sweeps' own stores are agent-scope (sc1) stores, which the L2 writes
// through, so the L2 of the XCD holds no dirty lines from them when a strip
// publishes (an agent-scope release writes back every dirty line of the L2).
__device__ __forceinline__ void st_through(gf32 *p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// called by every lane of the publishing wave: progress v (steps whose
// stores this wave issued) becomes visible with the stores before it.  Those
// stores are agent-scope write-through stores: once complete they are visible
// to the agent-scope loads of every XCD, so the release is a wait for them
// (no write-back of the whole L2, as an agent-scope release fence would do)
// followed by the progress store.
__device__ __forceinline__ void strip_publish(gi32 *mine, int v, int lane) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (lane == 0) store_agent(mine, v);
}

// KIND 0: mixing length; KIND 1: k-eps, KIND 2: no diffusion (sources by
// ke_source, the TDMA divides by the denominators k_ml_coef<.., KIND> stored).
template <int COMP, int KIND = 0>
__global__ __launch_bounds__(STRIP_WAVES * WAVE) void k_ml_strip(const Dom *__restrict__ ds, int IS, int JS) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice
    const gf32 *planes = d.planes;
    extern __shared__ float sm[];
    const int nkp = d.nkp, nkl = lds_slot(nkp);
    const int oa = (COMP == 0) ? 3 : 2, ia = (COMP == 1) ? 3 : 2, ob = d.nii - 1, ib = d.njj - 1;
    const int nstrip = (ib - ia + ROWS) / ROWS, nsteps = (ob - oa + 1) + ROWS - 1;
    const int r = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE), lane = threadIdx.x % WAVE;
    const bool chainw = r == ROWS;                            // the chain wave owns no row
    gi32 *sync = d.sync + SYNC_BLOCK0 + (size_t)(1 + COMP) * d.sync_stride;
    // per row: D, C, T, P, Q and the back-substitution results Xn of even and
    // odd steps (a row stores its previous column while the next one is solved)
    float *sD = sm + (size_t)r * 7 * nkl + SLOT_PAD, *sC = sD + nkl, *sT = sC + nkl, *sP = sT + nkl, *sQ = sP + nkl;
    float *sXn0 = sQ + nkl;                                   // [2][nkl] by step parity
    float *xbuf = sm + (size_t)ROWS * 7 * nkl + SLOT_PAD;      // [2][ROWS] slots of whole columns (old values)
    float *sdzk = sm + (size_t)ROWS * 9 * nkl;                // [nkp]
    int *sinfo = (int *)(sdzk + nkp);                         // [2][ROWS] per row: active, KSTART, KK, VI, KKm
    int *sctl = sinfo + 10 * ROWS;                            // strip index of this workgroup
    float *sdump = (float *)(sctl + 4);                       // [WAVE][CH] sinks of the chains
    if (threadIdx.x == 0) sctl[0] = __hip_atomic_fetch_add(&sync[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int k = threadIdx.x; k < nkp; k += STRIP_WAVES * WAVE) sdzk[k] = d.dzk[k];
    __syncthreads();
    const int B = sctl[0];
    if (B >= nstrip) return;                                  // uniform over the workgroup
    gi32 *mine = sync + SYNC_HDR + B;
    gf32 *X = (COMP == 0 ? d.uk : COMP == 1 ? d.vk : d.wk);
    if (chainw) {
        // The chain wave runs its own loop with the same three barriers per
        // step as the row waves (s_barrier counts waves, wherever they are in
        // the code), so none of the row waves' loop-carried prefetch registers
        // is live in the recurrences.  It issues ahead of the row waves on its
        // SIMD (wave priority): the recurrences are the step's serial path,
        // while the rows' look-ahead that overlaps them has slack.
        __builtin_amdgcn_s_setprio(3);
        // The rows publish the column data of step s one phase early (before
        // the barrier that ends the recurrences of step s-1), so the chain wave
        // reads them and takes its wave-wide maxima while the rows assemble.
        const int cl = lane < ROWS ? lane : 0;
        bool on = false;
        int KK = 0, VI = 0, nf = 0, nb = 0, nfmax = 0, nbmax = 0;
        auto info = [&](int s) {
            const int *si = sinfo + (s & 1) * 5 * ROWS;
            on = lane < ROWS && si[cl * 5];
            const int KSTART = si[cl * 5 + 1], KKm = si[cl * 5 + 4];
            KK = si[cl * 5 + 2]; VI = si[cl * 5 + 3];
            // first level the back-substitution updates (U/V skip levels inside
            // the building or its upstream neighbour; W updates every level)
            const int kend = (COMP == 2) ? KSTART : max(KSTART, KKm + 1);
            nf = VI - KK - 1; nb = VI - kend + 1;
            nfmax = wave_max(nf, on); nbmax = wave_max(nb, on);
        };
        // The chain wave also stores the last row's results (the only column
        // another strip reads) and publishes the strip's progress, while the
        // rows assemble the next step: these stores are its only vector-memory
        // operations, so the wait before a publish finds those of the step
        // before complete, and the last row's own second phase is no longer
        // than the others'.
        const int blast = ia + B * ROWS + ROWS - 1;
        auto store_last = [&](int s) {
            const int l = s - (ROWS - 1);
            if (__builtin_amdgcn_readlane((int)on, ROWS - 1) && blast <= ib && l >= 0 && oa + l <= ob) {
                const int vil = __builtin_amdgcn_readlane(VI, ROWS - 1);
                const int kendl = vil - __builtin_amdgcn_readlane(nb, ROWS - 1) + 1;
                int i, j;
                sweep_to_ij(d, 0, oa + l, blast, IS, JS, i, j);
                gf32 *Xp = X + c3(d, i, j);
                const float *xn = sm + (size_t)(ROWS - 1) * 7 * nkl + SLOT_PAD + (size_t)(5 + (s & 1)) * nkl;
                for (int k = kendl + lane; k <= vil; k += WAVE) st_through(Xp + (unsigned)k, xn[k]);
            }
        };
        for (int s = 0; s < nsteps; ++s) {
            lds_barrier();                                    // the rows' sources are assembled
            if (s == 0) info(0);
            if (on) {
                float *D = sm + (size_t)cl * 7 * nkl + SLOT_PAD;
                float *Cc = D + nkl, *T = Cc + nkl, *P = T + nkl, *Q = P + nkl, *Xn = Q + nkl + (size_t)(s & 1) * nkl;
                const float *Xs = xbuf + (size_t)((s & 1) * ROWS + cl) * nkl;
                chain_fwd<KIND != 0>(D + KK + 2, Cc + KK + 2, T + KK + 2, Q + KK + 2, nf, nfmax, Q[KK + 1], sdump + lane * CH);
                chain_back(P + VI, Q + VI, Xs + VI, Xn + VI, d.relax, nb, nbmax, sdump + lane * CH);
            }
            lds_barrier();                                    // the recurrences are done
            // progress s: the last row's results of steps < s are stored (those
            // of step s-1 one step ago)
            if (s > 0 && s % ml_publish<COMP>() == 0) strip_publish(mine, s, lane);
            store_last(s);
            if (s + 1 < nsteps) info(s + 1);
        }
        strip_publish(mine, nsteps, lane);
        return;
    }
    const int b = ia + B * ROWS + r;
    const bool row_ok = b <= ib;
    StripWait sw{B > 0 ? sync + SYNC_HDR + B - 1 : nullptr, d.sync, B > 0 ? 0 : 0x7fffffff, nsteps,
                 wall_clock64() + STRIP_TIMEOUT_TICKS};
    MlPre pre[2];
    float xfull[2];
    float wn = 0.0f;                                          // wall term of the located column (U, V)
    Meta mn;
    int in_ = 0, jn_ = 0;
    bool inr_n = false;
    // the metadata of a column is requested one step before its data (the data
    // loads need it for their addresses)
    auto meta_at = [&](int s) {
        int l = s - r, i, j;
        if (row_ok && l >= 0 && oa + l <= ob) {
            sweep_to_ij(d, 0, oa + l, b, IS, JS, i, j);
            return ldc4i(d.meta + c2(d, i, j));
        }
        return make_int4(0, -1, 0, 0);
    };
    int4 mq = meta_at(0);
    auto locate = [&](int s) {
        int l = s - r;                                        // column a = oa + l
        inr_n = row_ok && l >= 0 && oa + l <= ob;
        const int4 m4 = mq;
        mq = meta_at(s + 1);
        mn.adv = false;
        if (inr_n) {
            sweep_to_ij(d, 0, oa + l, b, IS, JS, in_, jn_);
            in_ = __builtin_amdgcn_readfirstlane(in_);        // uniform: column bases in scalar registers
            jn_ = __builtin_amdgcn_readfirstlane(jn_);
            mn = unpack_meta(m4);
            const gf32 *Xc = X + c3(d, in_, jn_);
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const unsigned k = lane + WAVE * m;
                if (k < (unsigned)nkp) xfull[m] = *at_bytes(Xc, 4u * k);
            }
        }
        if (mn.adv) {
            if (COMP != 2) wn = d.wcol[c2(d, in_, jn_)];
            const size_t o = c3(d, in_, jn_);
            const gf32 *Pc = planes + o * rec_floats<COMP>(), *R3 = d.rec3 + o * 3, *Xc = X + o;
            const gf32 *Xa = X + c3(d, in_ + IS, jn_), *Xb = X + c3(d, in_, jn_ + JS), *Xi = X + c3(d, in_ - IS, jn_);
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const int k = mn.KSTART + lane + WAVE * m;
                if (k <= mn.VI) {
                    ml_prefetch<COMP, KIND>(d, Pc, R3, Xc, Xa, Xb, (unsigned)k, pre[m]);
                    if (l == 0) pre[m].ni = *at_bytes(Xi, 4u * (unsigned)k);
                }
            }
        }
    };
    // row 0's -JS neighbours lie in the previous strip's last row: row 0 issues
    // its other look-ahead loads first, then waits for that strip, then reads them
    auto locate_prev = [&]() {
        if (mn.adv) {
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                int k = mn.KSTART + lane + WAVE * m;
                if (k <= mn.VI) pre[m].nj = ld_handoff(X + c3(d, in_, jn_ - JS) + (unsigned)k);
            }
        }
    };
    // the chain wave's data of the column located for step s (one lane writes)
    auto put_info = [&](int s) {
        if (lane == 0) {
            int *si = sinfo + (s & 1) * 5 * ROWS;
            si[r * 5] = mn.adv;
            if (mn.adv) {
                const int KKm = (COMP == 0) ? mn.KKmi : (COMP == 1) ? mn.KKmj : mn.KK;
                si[r * 5 + 1] = mn.KSTART; si[r * 5 + 2] = mn.KK; si[r * 5 + 3] = mn.VI; si[r * 5 + 4] = KKm;
            }
        }
    };
    locate(0);
    put_info(0);
    if (r == 0) {
        sw.wait(ROWS);
        locate_prev();
    }
    // The results of a step are merged on use: level k of a row's previous
    // column is the recurrence output (Xn of that step's parity) where the
    // back-substitution updated it (kend..VI), else the old value in the column
    // buffer.  Each row but the last (whose column the chain wave stores)
    // stores the updated levels of its previous column in the next step's
    // second phase, after issuing that phase's look-ahead loads: vector-memory
    // waits count loads and stores in one ordered counter, so a store issued
    // before those loads would make every wait for them (in the look-ahead
    // itself and in the next assembly) wait for the write as well.
    int pkend = 1 << 30, pvi = -1;                            // own previous column: updated levels
    size_t po = 0;
    auto store_prev = [&](const float *xn) {
        gf32 *Xp = X + po;
        for (int k = pkend + lane; k <= pvi; k += WAVE) st_through(Xp + (unsigned)k, xn[k]);
    };
    for (int s = 0; s < nsteps; ++s) {
        const bool act = mn.adv, inr = inr_n;
        const Meta mt = mn;
        const int i = in_, j = jn_;
        const float wcur = wn;
        const ColInfo q = meta_col(mt, i, j, d.simd);
        float *xcur = xbuf + (size_t)((s & 1) * ROWS + r) * nkl;
        const float *xw_lds = xbuf + (size_t)(((s - 1) & 1) * ROWS + r) * nkl;
        const float *xs_lds = xbuf + (size_t)(((s - 1) & 1) * ROWS + r - 1) * nkl;
        const float *xprev = sXn0 + (size_t)((s - 1) & 1) * nkl;   // recurrence output of step s-1
        const float *xs_new = xprev - (size_t)7 * nkl;       // row r-1's
        const int *sp = sinfo + ((s - 1) & 1) * 5 * ROWS;
        const bool w_in = (s - r >= 1), s_in = (r >= 1);
        if (inr) {
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                int k = lane + WAVE * m;
                if (k < nkp) xcur[k] = xfull[m];
            }
        }
        if (act) {
            const int KKP1 = q.KK + 1;
            // updated levels of the southern neighbour's column (row r-1, previous step)
            int skend = 1 << 30, svi = -1;
            if (s_in && sp[(r - 1) * 5]) {
                const int sks = sp[(r - 1) * 5 + 1], skm = sp[(r - 1) * 5 + 4];
                skend = (COMP == 2) ? sks : max(sks, skm + 1);
                svi = sp[(r - 1) * 5 + 3];
            }
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                int k = q.KSTART + lane + WAVE * m;
                if (k <= q.VI) {
                    const MlPre &r_ = pre[m];
                    float xnew_i = w_in ? ((k >= pkend && k <= pvi) ? xprev[k] : xw_lds[k]) : r_.ni;
                    float xnew_j = s_in ? ((k >= skend && k <= svi) ? xs_new[k] : xs_lds[k]) : r_.nj;
                    float xw = (IS == 1) ? xnew_i : r_.xa, xe = (IS == 1) ? r_.xa : xnew_i;
                    float xs = (JS == 1) ? xnew_j : r_.xb, xn = (JS == 1) ? r_.xb : xnew_j;
                    MlPieces pc;
                    pc.AW1 = r_.aw; pc.AS1 = r_.as; pc.AE1 = r_.ae; pc.AN1 = r_.an; pc.e = r_.e; pc.f = r_.f;
                    pc.g = r_.g; pc.w = wcur; pc.c0 = r_.c0; pc.X1 = r_.x1; pc.X2 = r_.x2; pc.X3 = r_.x3; pc.X4 = r_.x4;
                    float D;
                    if constexpr (KIND != 0) D = ke_source<COMP, KIND == 1>(d, q, k, sdzk[k], r_.ap0, pc, xw, xs, xe, xn, r_.x);
                    else D = ml_source<COMP>(d, q, k, sdzk[k], pc, xw, xs, xe, xn, r_.x);
                    sD[k] = D; sC[k] = r_.c; sT[k] = r_.t; sP[k] = r_.p;
                    if (k <= KKP1) sQ[k] = (KIND != 0) ? D / r_.t : D * r_.t;
                }
            }
        }
        const int kend = (COMP == 2) ? q.KSTART : max(q.KSTART, ((COMP == 0) ? mt.KKmi : mt.KKmj) + 1);
        lds_barrier();
        // while the chain wave runs the recurrences every row prefetches its
        // next column, then stores its previous one.  The look-ahead of row r
        // reads columns that rows r and r+1 write two steps later at the
        // earliest (for the last row, the next strip, which waits for this one);
        // row 0 also reads the previous strip's last row, after its wait.
        locate(s + 1);
        if (s + 1 < nsteps) put_info(s + 1);
        if (r == 0) {
            sw.wait(s + 1 + ROWS);
            locate_prev();
        }
        if (r != ROWS - 1) store_prev(xprev);
        pkend = act ? kend : 1 << 30;
        pvi = act ? q.VI : -1;
        po = c3(d, i, j);
        lds_barrier();                                        // the recurrence outputs are ready
    }
    if (r != ROWS - 1) store_prev(sXn0 + (size_t)((nsteps - 1) & 1) * nkl);
}

// Pressure sweep on strips.  The QIMP row scratch of sweep row a travels
// with the wavefront through LDS (ring slot (a - oa) mod RING; the ROWS rows of
// a step hold ROWS consecutive sweep rows): row 0 loads it from global memory
// (as strip B-1 left it), each row updates it with its column, and the last
// row stores it back for strip B+1.
#define RING 16
__global__ __launch_bounds__(STRIP_WAVES * WAVE) void k_pressure_strip(const Dom *__restrict__ ds, int IS, int JS, int slot) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];
    const gf32 *ptab = d.ptab[IS == 1 ? 0 : 1], *ttab = d.ttab[IS == 1 ? 0 : 1];
    extern __shared__ float sm[];
    const int nkp = d.nkp, nkl = lds_slot(nkp);
    const int oa = 2, ia = 2, ob = d.njj - 1, ib = d.nii - 1;
    const int nstrip = (ib - ia + ROWS) / ROWS, nsteps = (ob - oa + 1) + ROWS - 1;
    const int r = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE), lane = threadIdx.x % WAVE;
    const bool chainw = r == ROWS;                            // the chain wave owns no row
    gi32 *sync = d.sync + SYNC_BLOCK0;
    // per row: D, Q, T, P and the recurrence output N (the new DPM) of even and
    // odd steps (a row stores its previous column while the next one is solved)
    float *sD = sm + (size_t)r * 6 * nkl + SLOT_PAD, *sQ = sD + nkl, *sT = sQ + nkl, *sP = sT + nkl, *sN0 = sP + nkl;
    float *xbuf = sm + (size_t)ROWS * 6 * nkl + SLOT_PAD;      // [2][ROWS] slots of whole DPM columns (old values)
    float *ring = sm + (size_t)ROWS * 8 * nkl + SLOT_PAD;      // [RING] QIMP row scratch slots
    float *sdzk = sm + (size_t)(ROWS * 8 + RING) * nkl, *satp = sdzk + nkp;
    int *sinfo = (int *)(satp + nkp);                         // [2][ROWS] per row: active, KK, VI, ring slot
    int *sctl = sinfo + 8 * ROWS;
    float *sdump = (float *)(sctl + 4);                       // [WAVE][CH] sinks of the chains
    if (threadIdx.x == 0) sctl[0] = __hip_atomic_fetch_add(&sync[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int k = threadIdx.x; k < nkp; k += STRIP_WAVES * WAVE) { sdzk[k] = d.dzk[k]; satp[k] = d.atp[k]; }
    __syncthreads();
    const int B = sctl[0];
    if (B >= nstrip) return;
    if (chainw) {                                             // own loop, as in k_ml_strip
        __builtin_amdgcn_s_setprio(3);                        // ahead of the row waves, as in k_ml_strip
        // column data of step s published by the rows one phase early, read
        // while they assemble (as in k_ml_strip)
        const int cl = lane < ROWS ? lane : 0;
        bool on = false;
        int KK2 = 0, VI2 = 0, rs2 = 0, nmax = 0;
        auto info = [&](int s) {
            const int *si = sinfo + (s & 1) * 4 * ROWS;
            on = lane < ROWS && si[cl * 4];
            KK2 = si[cl * 4 + 1]; VI2 = si[cl * 4 + 2]; rs2 = si[cl * 4 + 3];
            nmax = wave_max(VI2 - KK2, on);
        };
        for (int s = 0; s < nsteps; ++s) {
            lds_barrier();                                    // the rows' sources are assembled
            if (s == 0) info(0);
            if (on) {
                float *Dd = sm + (size_t)cl * 6 * nkl + SLOT_PAD;
                float *Q = Dd + nkl, *T = Q + nkl, *P = T + nkl, *N = P + nkl + (size_t)(s & 1) * nkl;
                const float *O = xbuf + (size_t)((s & 1) * ROWS + cl) * nkl;
                // Q[k] holds ATP[k] on entry; the new DPM goes to N
                chain_pdown(Dd + VI2, T + VI2, Q + VI2, ring[(size_t)rs2 * nkl + VI2 + 1], VI2 - KK2, nmax, sdump + lane * CH);
                chain_pup(P + KK2 + 1, Q + KK2 + 1, N + KK2 + 1, O[KK2], VI2 - KK2, nmax, sdump + lane * CH);
            }
            lds_barrier();                                    // the recurrences are done
            if (s + 1 < nsteps) info(s + 1);
        }
        return;
    }
    const int b = ia + B * ROWS + r;
    const bool row_ok = b <= ib;
    StripWait sw{B > 0 ? sync + SYNC_HDR + B - 1 : nullptr, d.sync, B > 0 ? 0 : 0x7fffffff, nsteps,
                 wall_clock64() + STRIP_TIMEOUT_TICKS};
    gi32 *mine = sync + SYNC_HDR + B;
    const float DXK = d.dxk, DYK = d.dyk;
    float rmax = 0.0f;
    PrPre pre[2];
    float xfull[2], qfull[2];
    int in_ = 0, jn_ = 0, KKn = 0, VIn = 0;
    bool inr_n = false, actn = false;
    auto meta_at = [&](int s) {
        int l = s - r, i, j;
        if (row_ok && l >= 0 && oa + l <= ob) {
            sweep_to_ij(d, 1, oa + l, b, IS, JS, i, j);
            return ldc4i(d.meta + c2(d, i, j));
        }
        return make_int4(0, -1, 0, 0);
    };
    int4 mq = meta_at(0);
    auto locate = [&](int s) {
        int l = s - r;                                        // sweep row a = oa + l
        inr_n = row_ok && l >= 0 && oa + l <= ob;
        actn = false;
        const int4 m4 = mq;
        mq = meta_at(s + 1);
        if (inr_n) {
            sweep_to_ij(d, 1, oa + l, b, IS, JS, in_, jn_);
            Meta m = unpack_meta(m4);
            actn = m.adv;
            KKn = m.KK; VIn = m.VI;
            const size_t o = c3(d, in_, jn_);
#pragma unroll
            for (int m2 = 0; m2 < 2; ++m2) {
                int k = lane + WAVE * m2;
                if (k < nkp) xfull[m2] = d.dpm[o + k];
            }
        }
        if (actn) {
            size_t o = c3(d, in_, jn_);
#pragma unroll
            for (int m2 = 0; m2 < 2; ++m2) {
                int k = KKn + 1 + lane + WAVE * m2;
                if (k <= VIn) {
                    pre[m2].div = d.div[o + k];
                    pre[m2].t = ttab[o + k];
                    pre[m2].p = ptab[o + k];
                    pre[m2].pa = d.dpm[c3(d, in_ + IS, jn_) + k];
                    pre[m2].pb = d.dpm[c3(d, in_, jn_ + JS) + k];
                    pre[m2].dn = d.dpmnew[o + k];
                    if (l == 0) pre[m2].nj = d.dpm[c3(d, in_, jn_ - JS) + k];
                }
            }
        }
    };
    // row 0 reads the previous strip's last row (-IS neighbours) and the row
    // scratch as that strip left it: after its other look-ahead loads and its wait
    auto locate_prev = [&]() {
        if (inr_n) {
#pragma unroll
            for (int m2 = 0; m2 < 2; ++m2) {
                int k = lane + WAVE * m2;
                if (k < nkp) qfull[m2] = ld_handoff(d.prow_q + (size_t)jn_ * nkp + k);
            }
        }
        if (actn) {
            size_t o = c3(d, in_ - IS, jn_);
#pragma unroll
            for (int m2 = 0; m2 < 2; ++m2) {
                int k = KKn + 1 + lane + WAVE * m2;
                if (k <= VIn) pre[m2].ni = ld_handoff(d.dpm + o + k);
            }
        }
    };
    // the chain wave's data of the column located for step s (one lane writes)
    auto put_info = [&](int s) {
        if (lane == 0) {
            int *si = sinfo + (s & 1) * 4 * ROWS;
            si[r * 4] = actn;
            if (actn) { si[r * 4 + 1] = KKn; si[r * 4 + 2] = VIn; si[r * 4 + 3] = (s - r) & (RING - 1); }
        }
    };
    locate(0);
    put_info(0);
    if (r == 0) {
        sw.wait(ROWS);
        locate_prev();
    }
    // The results of a step are merged on use, as in k_ml_strip: level k of a
    // row's previous column is the new DPM (N of that step's parity) in
    // KK+1..VI, else the old value.  Each row finishes its previous column in
    // the next step: the row scratch update before the new sources replace sQ;
    // then, after issuing the look-ahead loads of the second phase (whose waits
    // would otherwise include these stores, as in k_ml_strip), the last row's
    // hand-off of the row scratch to strip B+1, the residual and the DPM/DPMNEW
    // stores (DPMNEW prefetched one step ahead).  The stores of step s-1 are
    // issued in step s, so the progress a strip publishes in step s is s-1.
    int pkk = 0, pvi = -1, prs = 0, pj = 0;                  // own previous column (pvi < 0: none)
    bool pinr = false;
    size_t po = 0;
    float dnp[2] = {0.0f, 0.0f};
    auto finish_ring = [&]() {
        float *qring = ring + (size_t)prs * nkl;
        for (int k = pkk + 1 + lane; k <= pvi; k += WAVE) qring[k] = sQ[k];
    };
    auto hand_off = [&]() {
        if (r == ROWS - 1 && pinr) {                          // hand the row scratch to strip B+1
            const float *qring = ring + (size_t)prs * nkl;
            gf32 *qrow = d.prow_q + (size_t)pj * nkp;
            for (int k = lane; k < nkp; k += WAVE) st_through(qrow + k, qring[k]);
        }
    };
    auto finish_store = [&](const float *xold, const float *sN) {
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            int k = pkk + 1 + lane + WAVE * m;
            if (k <= pvi) {
                float temp = sN[k];
                double qv = (double)conv_to_int((double)temp * d.frund) / d.frund;
                float rr = (float)fabs((double)xold[k] - qv);
                rmax = rr > rmax ? rr : rmax;
                st_through(d.dpm + po + k, temp);
                st_through(d.dpmnew + po + k, dnp[m] + d.relaxp * temp);
            }
        }
    };
    for (int s = 0; s < nsteps; ++s) {
        const bool act = actn, inr = inr_n;
        const int i = in_, j = jn_, KK = KKn, VI = VIn;
        const int rs = (s - r) & (RING - 1);                  // ring slot of this wave's sweep row
        float *xcur = xbuf + (size_t)((s & 1) * ROWS + r) * nkl;
        const float *xw_lds = xbuf + (size_t)(((s - 1) & 1) * ROWS + r) * nkl;
        const float *xs_lds = xbuf + (size_t)(((s - 1) & 1) * ROWS + r - 1) * nkl;
        const float *snprev = sN0 + (size_t)((s - 1) & 1) * nkl;  // new DPM of the previous step
        const float *xs_new = snprev - (size_t)6 * nkl;      // row r-1's
        const int *sp = sinfo + ((s - 1) & 1) * 4 * ROWS;
        const bool w_in = (s - r >= 1), s_in = (r >= 1);
        finish_ring();
        if (inr) {
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                int k = lane + WAVE * m;
                if (k < nkp) {
                    xcur[k] = xfull[m];
                    if (r == 0) ring[(size_t)rs * nkl + k] = qfull[m];
                }
            }
        }
        if (act) {
            // updated levels of the neighbouring row's column (row r-1, previous step)
            int slo = 1 << 30, shi = -1;
            if (s_in && sp[(r - 1) * 4]) {
                slo = sp[(r - 1) * 4 + 1] + 1;
                shi = sp[(r - 1) * 4 + 2];
            }
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                int k = KK + 1 + lane + WAVE * m;
                if (k <= VI) {
                    const PrPre &r_ = pre[m];
                    float newj = w_in ? ((k > pkk && k <= pvi) ? snprev[k] : xw_lds[k]) : r_.nj;
                    float newi = s_in ? ((k >= slo && k <= shi) ? xs_new[k] : xs_lds[k]) : r_.ni;
                    float dim_ = (IS == 1) ? newi : r_.pa, dip = (IS == 1) ? r_.pa : newi;
                    float djm = (JS == 1) ? newj : r_.pb, djp = (JS == 1) ? r_.pb : newj;
                    float dz = sdzk[k];
                    sD[k] = r_.div + (dim_ + dip) / DXK * DYK * dz + (djm - djp) / DYK * DXK * dz;
                    sT[k] = r_.t;
                    sP[k] = r_.p;
                    sQ[k] = satp[k];                  // consumed and replaced by QIMP in phase 2
                }
            }
        }
        const float dn0 = pre[0].dn, dn1 = pre[1].dn;         // this column's DPMNEW, before the look-ahead
        lds_barrier();
        // while the chain wave runs the recurrences: publish and prefetch as in
        // k_ml_strip; row 0 reads the previous strip's row scratch and last row
        // after its wait; then the previous column's stores
        if (r == ROWS - 1 && s > 1 && (s - 1) % STRIP_PUBLISH == 0) strip_publish(mine, s - 1, lane);
        locate(s + 1);
        if (s + 1 < nsteps) put_info(s + 1);
        if (r == 0) {
            sw.wait(s + 1 + ROWS);
            locate_prev();
        }
        hand_off();
        finish_store(xw_lds, snprev);
        pkk = KK; pvi = act ? VI : -1; prs = rs; pj = j; pinr = inr; po = c3(d, i, j);
        dnp[0] = dn0; dnp[1] = dn1;
        lds_barrier();                                        // the recurrence outputs are ready
    }
    finish_ring();
    hand_off();
    finish_store(xbuf + (size_t)(((nsteps - 1) & 1) * ROWS + r) * nkl, sN0 + (size_t)((nsteps - 1) & 1) * nkl);
    if (r == ROWS - 1) strip_publish(mine, nsteps, lane);
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        float other = __shfl_xor(rmax, off);
        rmax = other > rmax ? other : rmax;
    }
    const int rs = slot - d.iter_off;                 // this situation's residual slot
    if (lane == 0 && rmax > 0.0f && slot >= 0 && rs >= 0 && rs < d.resid_cap) atomicMax((unsigned *)&d.resid[rs], __float_as_uint(rmax));
}

static inline int ndiag(int oa, int ob, int ia, int ib) { return (ob - oa) + (ib - ia) + 1; }
static inline int diag_len(int t, int oa, int ob, int ia, int ib) {
    int nb = ib - ia;
    int amin = oa + t - nb > oa ? oa + t - nb : oa;
    int amax = ob < oa + t ? ob : oa + t;
    return amax - amin + 1;
}

// ------------------------------------------------------------------------
// Diagnostic post-pass: DiagnosticFlowfield.Calculate (DiagnosticFlowfield.cs:
// 24-436) and the final vertical-velocity mass adjustment of MicroscaleFlat
// (MicroscaleFlat.cs:388-462) or MicroscaleTerrain (MicroscaleTerrain.cs:
// 862-933).  It works on full-height columns (levels 0..NKK+1), in its own
// arrays; one situation per launch, passed by value.
// ------------------------------------------------------------------------
__device__ __forceinline__ size_t dg3(const DiagArgs &g, int i, int j) { return ((size_t)i * g.nj2 + j) * (size_t)g.nkh; }

// DPM = 0 for i in [1,NII], j in [1,NJJ], k in [1,NKK+1] (DiagnosticFlowfield.cs:40-49)
__global__ void k_dg_zero(DiagArgs g) {
    int c = blockIdx.x, i = 1 + c / g.njj, j = 1 + c % g.njj;
    if (i > g.nii) return;
    size_t o = dg3(g, i, j);
    for (int k = 1 + threadIdx.x; k <= g.nkk + 1; k += blockDim.x) g.dpm[o + k] = 0.0f;
}

// mass divergence of every interior column (:53-80); block = one column
__global__ void k_dg_divergence(DiagArgs g) {
    int nj = g.njj - 2, i = 2 + blockIdx.x / nj, j = 2 + blockIdx.x % nj;
    int KK = g.kk[(size_t)i * g.nj2 + j];
    size_t o = dg3(g, i, j), oip = dg3(g, i + 1, j), ojp = dg3(g, i, j + 1);
    const float DXK = g.dxk, DYK = g.dyk;
    for (int k = KK + 1 + threadIdx.x; k <= g.nkk - 1; k += blockDim.x) {
        float dz = g.dzk[k];
        float t = (g.uk[o + k] - g.uk[oip + k]) * DYK * dz + (g.vk[o + k] - g.vk[ojp + k]) * DXK * dz;
        g.div[o + k] = (k > KK + 1) ? t + (g.wk[o + k] - g.wk[o + k + 1]) * DXK * DYK : t - g.wk[o + k + 1] * DXK * DYK;
    }
}

// pressure correction (:84-341), one anti-diagonal of the sweep per launch,
// one wave per column.  Every column is solved from NKK-1 down to KKART+1 and
// PIMP[NKK] is never written, so TERMP and PIMP at level k are the same in
// every column whose bottom lies below k; at the bottom level only ABP (and so
// PIMP) differs.  The host runs that recurrence once (term, p_in, p_bot);
// here the lanes build DIM and stage the tables in LDS, and lane 0 runs the
// two recurrences that depend on the column (QIMP top-down from QIMP[NKK] = 0,
// then the new DPM bottom-up from DPM[KKART] = 0) over LDS.
#define DG_UNROLL 8
__global__ __launch_bounds__(WAVE) void k_dg_pressure(DiagArgs g, int t, int IS, int JS) {
    extern __shared__ float sm[];
    int a, b;
    if (!diag_col(t, 2, g.njj - 1, 2, g.nii - 1, blockIdx.x, a, b)) return;
    int j = (JS == -1) ? g.njj - a + 1 : a;
    int i = (IS == -1) ? g.nii - b + 1 : b;
    const int KK = g.kk[(size_t)i * g.nj2 + j];
    const int nkh = g.nkh, top = g.nkk - 1;
    // LDS: DIM, ATP, TERMP, P (then the new DPM) and Q, nkh levels each
    float *sDIM = sm, *sA = sDIM + nkh, *sT = sA + nkh, *sP = sT + nkh, *sQ = sP + nkh;
    size_t o = dg3(g, i, j), oim = dg3(g, i - 1, j), oip = dg3(g, i + 1, j), ojm = dg3(g, i, j - 1), ojp = dg3(g, i, j + 1);
    const float DXK = g.dxk, DYK = g.dyk;
    for (int k = KK + 1 + threadIdx.x; k <= top; k += WAVE) {
        float dz = g.dzk[k];
        sDIM[k] = g.div[o + k] / g.dtime + (g.dpm[oim + k] + g.dpm[oip + k]) / DXK * DYK * dz +
                  (g.dpm[ojm + k] - g.dpm[ojp + k]) / DYK * DXK * dz;
        sA[k] = g.atp[k];
        sT[k] = g.term[k];
        sP[k] = (k > KK + 1) ? g.p_in[k] : g.p_bot[k];
    }
    __syncthreads();
    if (threadIdx.x == 0 && top > KK) {
        // QIMP[k] = (ATP[k] QIMP[k+1] + DIM[k]) TERMP[k], k = top .. KK+1, in blocks
        // of DG_UNROLL levels whose operands are read before the block's chain
        float Qn = 0.0f;
        int k = top;
        for (; k - DG_UNROLL + 1 > KK; k -= DG_UNROLL) {
            float ca[DG_UNROLL], cd[DG_UNROLL], ct[DG_UNROLL];
#pragma unroll
            for (int u = 0; u < DG_UNROLL; ++u) { ca[u] = sA[k - u]; cd[u] = sDIM[k - u]; ct[u] = sT[k - u]; }
#pragma unroll
            for (int u = 0; u < DG_UNROLL; ++u) {
                Qn = (ca[u] * Qn + cd[u]) * ct[u];
                sQ[k - u] = Qn;
            }
        }
        for (; k > KK; --k) {
            Qn = (sA[k] * Qn + sDIM[k]) * sT[k];
            sQ[k] = Qn;
        }
        // DPM[k] = PIMP[k] DPM[k-1] + QIMP[k], k = KK+1 .. top; written over P
        float dprev = 0.0f;
        k = KK + 1;
        for (; k + DG_UNROLL - 1 <= top; k += DG_UNROLL) {
            float cp[DG_UNROLL], cq[DG_UNROLL];
#pragma unroll
            for (int u = 0; u < DG_UNROLL; ++u) { cp[u] = sP[k + u]; cq[u] = sQ[k + u]; }
#pragma unroll
            for (int u = 0; u < DG_UNROLL; ++u) {
                dprev = cp[u] * dprev + cq[u];
                sP[k + u] = dprev;
            }
        }
        for (; k <= top; ++k) {
            dprev = sP[k] * dprev + sQ[k];
            sP[k] = dprev;
        }
    }
    __syncthreads();
    // DPM[0] = 0 and DPM[k] = 0 up to KKART (the reference writes them each sweep)
    for (int k = threadIdx.x; k <= top; k += WAVE) g.dpm[o + k] = (k > KK) ? sP[k] : 0.0f;
}

// velocity correction without quantisation (:345-410); thread per (column, level)
__global__ void k_dg_correction(DiagArgs g) {
    int nj = g.njj - 2, i = 2 + blockIdx.x / nj, j = 2 + blockIdx.x % nj;
    size_t c = (size_t)i * g.nj2 + j;
    int KK = g.kk[c], KKim = g.kk[c - g.nj2], KKjm = g.kk[c - 1];
    size_t o = dg3(g, i, j), oim = dg3(g, i - 1, j), ojm = dg3(g, i, j - 1);
    const float DXK = g.dxk, DYK = g.dyk, DT = g.dtime;
    for (int k = 1 + threadIdx.x; k < g.nkk; k += blockDim.x) {
        float d0 = g.dpm[o + k];
        float DDPX = g.dpm[oim + k] - d0;
        float DDPY = g.dpm[ojm + k] - d0;
        float DDPZ = (k > KK + 1) ? g.dpm[o + k - 1] - d0 : 0.0f;
        g.uk[o + k] = (KK < k && KKim < k) ? g.uk[o + k] + DDPX / DXK * DT : 0.0f;
        g.vk[o + k] = (KK < k && KKjm < k) ? g.vk[o + k] + DDPY / DYK * DT : 0.0f;
        g.wk[o + k] = (KK + 1 < k) ? g.wk[o + k] + DDPZ * 2 / (g.dzk[k] + g.dzk[k - 1]) * DT : 0.0f;
    }
}

// vertical velocity from the mass balance of each cell, serial in k; thread per column
__global__ void k_dg_wmass(DiagArgs g) {
    int nj = g.njj - 2, ncol = (g.nii - 2) * nj;
    int cidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (cidx >= ncol) return;
    int i = 2 + cidx / nj, j = 2 + cidx % nj;
    size_t c = (size_t)i * g.nj2 + j;
    if (!g.terrain && !(g.level <= 1 || g.adv[c] == 0)) return;
    const float DXK = g.dxk, DYK = g.dyk, DSN = g.terrain ? g.dyk : g.dxk;
    int KK = g.kk[c], KKim = g.kk[c - g.nj2], KKip = g.kk[c + g.nj2], KKjm = g.kk[c - 1], KKjp = g.kk[c + 1];
    size_t o = dg3(g, i, j), oip = dg3(g, i + 1, j), ojp = dg3(g, i, j + 1);
    for (int k = KK + 1; k < g.nkk; ++k) {
        float dz = g.dzk[k];
        float fwo1 = (k <= KKim) ? 0.0f : dz * DYK * g.uk[o + k];
        float fwo2 = (k <= KKip) ? 0.0f : dz * DYK * g.uk[oip + k];
        float fsn1 = (k <= KKjm) ? 0.0f : dz * DSN * g.vk[o + k];
        float fsn2 = (k <= KKjp) ? 0.0f : dz * DSN * g.vk[ojp + k];
        float fbt1 = ((k <= KK + 1) || (k == 1)) ? 0.0f : DXK * DYK * g.wk[o + k];
        g.wk[o + k + 1] = (fwo1 - fwo2 + fsn1 - fsn2 + fbt1) / (DXK * DYK);
    }
}

extern "C" int pf_launch_diagnostic(const DiagArgs *ga, int iterations, hipStream_t s) {
    const DiagArgs &g = *ga;
    int ncol = (g.nii - 2) * (g.njj - 2);
    hipLaunchKernelGGL(k_dg_zero, dim3(g.nii * g.njj), dim3(64), 0, s, g);
    for (int it = 1; it <= iterations; ++it) {
        hipLaunchKernelGGL(k_dg_divergence, dim3(ncol), dim3(64), 0, s, g);
        int IS = 1, JS = 1;
        switch (it % 4) {
            case 0: IS = 1; JS = 1; break;
            case 1: IS = 1; JS = -1; break;
            case 2: IS = -1; JS = -1; break;
            default: IS = -1; JS = 1; break;
        }
        int oa = 2, ob = g.njj - 1, ia = 2, ib = g.nii - 1;
        size_t shm = sizeof(float) * 5 * (size_t)g.nkh;
        for (int t = 0; t < ndiag(oa, ob, ia, ib); ++t)
            hipLaunchKernelGGL(k_dg_pressure, dim3(diag_len(t, oa, ob, ia, ib)), dim3(WAVE), shm, s, g, t, IS, JS);
        hipLaunchKernelGGL(k_dg_correction, dim3(ncol), dim3(64), 0, s, g);
    }
    hipLaunchKernelGGL(k_dg_wmass, dim3((ncol + 63) / 64), dim3(64), 0, s, g);
    return (int)hipGetLastError();
}

// ------------------------------------------------------------------------
// host-side launchers
// ------------------------------------------------------------------------
// ------------------------------------------------------------------------
// Strip-pipelined TKE sweep (TKE-prognostic-microscale.cs:25-352).
// k_tke_coef writes, for every active cell above KKART, what the sweep's
// in-place neighbours do not touch: the neighbour coefficients AW..AN, the
// remaining source terms, the halves of the two Thomas recurrences that need
// no source (P and its denominator, per level), and the wall function.  The
// strip kernel then only assembles the neighbour terms, runs the two Q
// recurrences and the back-substitution, and writes back.  Record per cell
// (TKE_REC floats, written whole by the first pass): {AW, AS, AE, AN |
// ap0*TURB, PROTURB, PROBOU, ap0*TDISS | EPS source, C1, C2, wall flag | wall
// TURB, wall TDISS, 0, 0}; after the records (at TKE_REC*n3 floats), per cell
// {den1, P1, den2, P2}, written whole by the last pass.
// ------------------------------------------------------------------------
#define TKE_REC 16
__device__ __forceinline__ size_t tke_rec4(const DomG &d) { return (size_t)TKE_REC * (d.nii + 2) * d.nj2 * d.nkp; }
#define TKE_SCOL 16       // ints of column data per column in k_tke_coef

// A block of 4 waves takes COEF_COLS consecutive columns (j fastest); their
// levels KKART+1..VI are laid out one after another and dealt to the 256
// lanes, which compute the cells (tke_cell) and write the first three float4
// of each record and the wall values.  One wave then runs the two P
// recurrences of all COEF_COLS columns at once (lane = column), and the lanes
// write the fourth float4.
__global__ __launch_bounds__(4 * WAVE) void k_tke_coef(const Dom *__restrict__ ds) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice
    extern __shared__ float sm[];
    const int nkp = d.nkp, nkl = lds_stride(nkp);
    const int lane = threadIdx.x % WAVE;
    const int na = d.nii - 2, nb = d.njj - 2;
    // per column c: A, B1, C1, B2, C2 over k at sm + c*5*nkl, and the column's data
    int *scol = (int *)(sm + 5 * COEF_COLS * nkl);        // [COEF_COLS][TKE_SCOL]
    int *stot = scol + TKE_SCOL * COEF_COLS;
    const int blk = xcd_block((na * nb + COEF_COLS - 1) / COEF_COLS);
    if (blk < 0) return;                                      // uniform over the block
    const int col0 = blk * COEF_COLS;
    if (threadIdx.x < WAVE) {
        int c = lane, n = 0;
        if (c < COEF_COLS) {
            int *sc = scol + c * TKE_SCOL;
            const int col = col0 + c;
            int act = 0;
            if (col < na * nb) {
                const int i = 2 + col / nb, j = 2 + col % nb;
                if (d.adv[c2(d, i, j)] == 1) {
                    const ColInfo q = col_info(d, i, j);
                    const TkeCol t = tke_col(d, i, j);
                    act = 1;
                    n = q.VI - q.KK;
                    sc[0] = q.KK; sc[1] = q.VI; sc[2] = i; sc[3] = j;
                    sc[6] = (q.as ? 1 : 0) | (q.an ? 2 : 0) | (q.aw ? 4 : 0) | (q.ae ? 8 : 0) | (t.cip ? 16 : 0) |
                            (t.cim ? 32 : 0) | (t.cjp ? 64 : 0) | (t.cjm ? 128 : 0);
                    sc[7] = t.kkip; sc[8] = t.kkim; sc[9] = t.kkjp; sc[10] = t.kkjm;
                    sc[11] = __float_as_int(q.cutk); sc[12] = __float_as_int(q.ho);
                }
            }
            sc[5] = act;
        }
        // inclusive prefix sum of the level counts over the lanes
        int incl = n;
        for (int off = 1; off < WAVE; off <<= 1) {
            int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        if (c < COEF_COLS) scol[c * TKE_SCOL + 4] = incl - n;
        if (lane == WAVE - 1) stot[0] = incl;
    }
    lds_barrier();
    const int total = stot[0];
    auto column_of = [&](int t) {
        int lo = 0;
        for (int step = COEF_COLS / 2; step > 0; step >>= 1)
            if (scol[(lo + step) * TKE_SCOL + 4] <= t) lo += step;
        return lo;
    };
    for (int t = threadIdx.x; t < total; t += 4 * WAVE) {
        const int c = column_of(t);
        const int *sc = scol + c * TKE_SCOL;
        const int k = sc[0] + 1 + (t - sc[4]);
        ColInfo q;
        q.KK = sc[0]; q.VI = sc[1]; q.i = sc[2]; q.j = sc[3]; q.KSTART = q.KK + 1; q.W = d.simd;
        const int fl = sc[6];
        q.as = fl & 1; q.an = (fl >> 1) & 1; q.aw = (fl >> 2) & 1; q.ae = (fl >> 3) & 1; q.vegp = false;
        q.cutk = __int_as_float(sc[11]); q.ho = __int_as_float(sc[12]); q.cov = q.ht = q.z0s = 0.0f;
        TkeCol tc = tke_col_ptrs(d, q.i, q.j);
        tc.cip = (fl >> 4) & 1; tc.cim = (fl >> 5) & 1; tc.cjp = (fl >> 6) & 1; tc.cjm = (fl >> 7) & 1;
        tc.kkip = sc[7]; tc.kkim = sc[8]; tc.kkjp = sc[9]; tc.kkjm = sc[10];
        const TkeCell e = tke_cell(d, q, tc, k);
        gf32 *r = d.planes + (tc.o + k) * TKE_REC;
        stg4(r, e.AW1, e.AS1, e.AE1, e.AN1);
        stg4(r + 4, e.w1, e.PRO, e.PBO, e.w2);
        stg4(r + 8, e.EPS, e.C1, e.C2, e.wf ? 1.0f : 0.0f);
        stg4(r + 12, e.wt, e.we, 0.0f, 0.0f);
        float *sA = sm + (size_t)c * 5 * nkl;
        sA[k] = e.A; sA[nkl + k] = e.B1; sA[2 * nkl + k] = e.C1; sA[3 * nkl + k] = e.B2; sA[4 * nkl + k] = e.C2;
    }
    lds_barrier();
    if (threadIdx.x < COEF_COLS && scol[threadIdx.x * TKE_SCOL + 5]) {
        // P = B/(A - C*P_prev) above KKART+1, B/A at KKART+1 (the recurrence
        // restarts there; the levels below it do not reach the ones above);
        // den1, P1, den2, P2 replace A, B1, C1, B2
        const int c = threadIdx.x;
        float *__restrict__ A = sm + (size_t)c * 5 * nkl;
        float *__restrict__ B1 = A + nkl, *__restrict__ C1 = B1 + nkl, *__restrict__ B2 = C1 + nkl;
        float *__restrict__ C2 = B2 + nkl;
        const int lo = scol[c * TKE_SCOL + 0] + 1, VI = scol[c * TKE_SCOL + 1];
        float P1 = 0, P2 = 0;
        for (int k = lo; k <= VI; ++k) {
            const float a = A[k];
            float den1 = a, den2 = a;
            if (k > lo) { den1 = a - C1[k] * P1; den2 = a - C2[k] * P2; }
            P1 = B1[k] / den1;
            P2 = B2[k] / den2;
            A[k] = den1; B1[k] = P1; C1[k] = den2; B2[k] = P2;
        }
    }
    lds_barrier();
    for (int t = threadIdx.x; t < total; t += 4 * WAVE) {
        const int c = column_of(t);
        const int *sc = scol + c * TKE_SCOL;
        const int k = sc[0] + 1 + (t - sc[4]);
        const float *sA = sm + (size_t)c * 5 * nkl;
        stg4(d.planes + tke_rec4(d) + (c3(d, sc[2], sc[3]) + k) * 4, sA[k], sA[nkl + k], sA[2 * nkl + k], sA[3 * nkl + k]);
    }
}

// Relaxation-free back-substitution of the TKE sweep, downwards from level
// VI (pointers at VI): X[-t] = max(P*X[-t+1] + Q, floor), or the wall value
// where the wall flag is set, t = 0..n-1, X[1] = top (the old value above VI).
// The clamp is the reference's double-precision Math.Max.
__device__ __forceinline__ void chain_tke_back(const float *__restrict__ P, const float *__restrict__ Q,
                                               const float *__restrict__ WF, const float *__restrict__ WV,
                                               float *__restrict__ X, float top, double floor_, int n, int nmax,
                                               float *sink) {
    float up = top;
    float cp[CH], cq[CH], cf[CH], cv[CH];
    // chunks addressed from their lowest level upwards, as in chain_back
    {
        const float *Pb = P - (CH - 1), *Qb = Q - (CH - 1), *Fb = WF - (CH - 1), *Vb = WV - (CH - 1);
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            cp[u] = Pb[CH - 1 - u]; cq[u] = Qb[CH - 1 - u]; cf[u] = Fb[CH - 1 - u]; cv[u] = Vb[CH - 1 - u];
        }
    }
#pragma unroll
    for (int h = 0; h < MAXCH; ++h) {
        const int t = h * CH;
        if (t >= nmax) break;
        float np[CH], nq[CH], nf[CH], nv[CH];
        const int o = (t + CH < n) ? t + CH : 0;
        const float *Pb = P - (o + CH - 1), *Qb = Q - (o + CH - 1), *Fb = WF - (o + CH - 1), *Vb = WV - (o + CH - 1);
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            np[u] = Pb[CH - 1 - u]; nq[u] = Qb[CH - 1 - u]; nf[u] = Fb[CH - 1 - u]; nv[u] = Vb[CH - 1 - u];
        }
        __builtin_amdgcn_sched_barrier(0);                // keep the next chunk's reads ahead of this chunk
        float *Xw = (t < n) ? X - (t + CH - 1) : sink;
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            float x = (cf[u] != 0.0f) ? cv[u] : cp[u] * up + cq[u];
            x = (float)mmaxd((double)x, floor_);
            Xw[CH - 1 - u] = x;
            up = x;
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) { cp[u] = np[u]; cq[u] = nq[u]; cf[u] = nf[u]; cv[u] = nv[u]; }
    }
}

// per-level values of one column of the TKE sweep, loaded one local step ahead
struct TkePre {
    float4 c0, c1, c2;              // record words 0..11
    float4 c3;                      // den1, P1, den2, P2
    float2 c4;                      // wall TURB, wall TDISS
    float ta, ea, tb, eb;           // old TURB/TDISS in the +IS (a) and +JS (b) directions
    float nit, nie, njt, nje;       // -IS / -JS neighbours as memory holds them
};

// Strip sweep of TURB and TDISS, laid out as k_ml_strip: ROWS row waves and a
// chain wave.  Row wave r owns sweep row b0 + r (j direction) and at step s
// handles column a = 2 + s - r: it assembles the column's two sources from the
// record (k_tke_coef) and the four neighbours (lanes over levels) and lays the
// recurrence operands out in LDS; the chain wave then runs the Q recurrences and
// the back-substitution of every row at once (lane r: TURB of row r, lane
// ROWS + r: TDISS of row r), while the row waves load the next column.  The
// neighbours updated earlier in the sweep (the previous column of the same
// row, the previous row's column of the previous step) are merged on use from
// the recurrence outputs in LDS over the levels they updated; their other
// levels, and row 0's neighbours in the previous strip, come from memory.
// LDS per row: D1, D2, C1, den1, P1, C2, den2, P2, wall flag, wall TURB,
// wall TDISS, Q1, Q2, new TURB, new TDISS.
#define TKE_SLOTS 15
__global__ __launch_bounds__(STRIP_WAVES * WAVE) void k_tke_strip(const Dom *__restrict__ ds, int IS, int JS) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];                   // situation of this launch slice
    extern __shared__ float sm[];
    const int nkp = d.nkp, nkl = lds_slot(nkp);
    const int oa = 2, ia = 2, ob = d.nii - 1, ib = d.njj - 1;
    const int nstrip = (ib - ia + ROWS) / ROWS, nsteps = (ob - oa + 1) + ROWS - 1;
    const int r = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE), lane = threadIdx.x % WAVE;
    const bool chainw = r == ROWS;
    gi32 *sync = d.sync + SYNC_BLOCK0 + (size_t)4 * d.sync_stride;
    float *rowb = sm + (size_t)r * TKE_SLOTS * nkl + SLOT_PAD;
    float *sD1 = rowb, *sD2 = rowb + nkl, *sC1 = rowb + 2 * nkl, *sN1 = rowb + 3 * nkl, *sP1 = rowb + 4 * nkl;
    float *sC2 = rowb + 5 * nkl, *sN2 = rowb + 6 * nkl, *sP2 = rowb + 7 * nkl, *sWF = rowb + 8 * nkl;
    float *sWT = rowb + 9 * nkl, *sWE = rowb + 10 * nkl, *sQ1 = rowb + 11 * nkl, *sQ2 = rowb + 12 * nkl;
    float *sOT = rowb + 13 * nkl, *sOE = rowb + 14 * nkl;
    float *sdump = sm + (size_t)ROWS * TKE_SLOTS * nkl;          // [WAVE][CH] sinks of the chains
    int *sinfo = (int *)(sdump + WAVE * CH);                    // [2][ROWS] per row: active, lo, VI, top TURB, top TDISS
    int *sctl = sinfo + 10 * ROWS;
    if (threadIdx.x == 0) sctl[0] = __hip_atomic_fetch_add(&sync[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int B = sctl[0];
    if (B >= nstrip) return;                                  // uniform over the workgroup
    if (chainw) {
        __builtin_amdgcn_s_setprio(3);                        // ahead of the row waves, as in k_ml_strip
        // the rows publish the levels of step s one phase early (as in
        // k_ml_strip) and the values above the top with their sources
        const int f = lane >= ROWS ? 1 : 0, cl = lane < 2 * ROWS ? lane - f * ROWS : 0;
        bool on = false;
        int lo = 0, VI = 0, nfmax = 0, nbmax = 0;
        auto info = [&](int s) {
            const int *si = sinfo + (s & 1) * 5 * ROWS;
            on = lane < 2 * ROWS && si[cl * 5];
            lo = si[cl * 5 + 1]; VI = si[cl * 5 + 2];
            nfmax = wave_max(VI - lo, on); nbmax = wave_max(VI - lo + 1, on);
        };
        for (int s = 0; s < nsteps; ++s) {
            lds_barrier();                                    // the rows' sources are assembled
            if (s == 0) info(0);
            const int *si = sinfo + (s & 1) * 5 * ROWS;
            const float top = __int_as_float(si[cl * 5 + 3 + f]);
            if (on) {
                float *R = sm + (size_t)cl * TKE_SLOTS * nkl + SLOT_PAD;
                float *Dv = R + (size_t)f * nkl, *Cv = R + (size_t)(2 + 3 * f) * nkl, *Nv = R + (size_t)(3 + 3 * f) * nkl;
                float *Pv = R + (size_t)(4 + 3 * f) * nkl, *WF = R + 8 * (size_t)nkl, *WV = R + (size_t)(9 + f) * nkl;
                float *Qv = R + (size_t)(11 + f) * nkl, *Ov = R + (size_t)(13 + f) * nkl;
                chain_fwd<true>(Dv + lo + 1, Cv + lo + 1, Nv + lo + 1, Qv + lo + 1, VI - lo, nfmax, Qv[lo], sdump + lane * CH);
                chain_tke_back(Pv + VI, Qv + VI, WF + VI, WV + VI, Ov + VI, top, f ? 0.0000001 : 0.001, VI - lo + 1,
                               nbmax, sdump + lane * CH);
            }
            lds_barrier();                                    // the recurrences are done
            if (s + 1 < nsteps) info(s + 1);
        }
        return;
    }
    const int b = ia + B * ROWS + r;
    const bool row_ok = b <= ib;
    StripWait sw{B > 0 ? sync + SYNC_HDR + B - 1 : nullptr, d.sync, B > 0 ? 0 : 0x7fffffff, nsteps,
                 wall_clock64() + STRIP_TIMEOUT_TICKS};
    gi32 *mine = sync + SYNC_HDR + B;
    TkePre pre[2];
    Meta mn;
    int in_ = 0, jn_ = 0;
    bool inr_n = false;
    float topt = 0.0f, tope = 0.0f;
    auto meta_at = [&](int s) {
        int l = s - r, i, j;
        if (row_ok && l >= 0 && oa + l <= ob) {
            sweep_to_ij(d, 0, oa + l, b, IS, JS, i, j);
            return ldc4i(d.meta + c2(d, i, j));
        }
        return make_int4(0, -1, 0, 0);
    };
    int4 mq = meta_at(0);
    auto locate = [&](int s) {
        int l = s - r;
        inr_n = row_ok && l >= 0 && oa + l <= ob;
        const int4 m4 = mq;
        mq = meta_at(s + 1);
        mn.adv = false;
        if (inr_n) {
            sweep_to_ij(d, 0, oa + l, b, IS, JS, in_, jn_);
            mn = unpack_meta(m4);
        }
        if (mn.adv) {
            const size_t o = c3(d, in_, jn_);
            topt = d.turb[o + mn.VI + 1];
            tope = d.tdiss[o + mn.VI + 1];
            const size_t oa_ = c3(d, in_ + IS, jn_), ob_ = c3(d, in_, jn_ + JS), oi = c3(d, in_ - IS, jn_);
            const size_t oj = c3(d, in_, jn_ - JS);
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                int k = mn.KK + 1 + lane + WAVE * m;
                if (k <= mn.VI) {
                    TkePre &p = pre[m];
                    const gf32 *rr = d.planes + (o + k) * TKE_REC;
                    p.c0 = ldg4(rr); p.c1 = ldg4(rr + 4); p.c2 = ldg4(rr + 8); p.c4 = ldg2(rr + 12);
                    p.c3 = ldg4(d.planes + tke_rec4(d) + (o + k) * 4);
                    p.ta = d.turb[oa_ + k]; p.ea = d.tdiss[oa_ + k];
                    p.tb = d.turb[ob_ + k]; p.eb = d.tdiss[ob_ + k];
                    p.nit = d.turb[oi + k]; p.nie = d.tdiss[oi + k];
                    if (r > 0) { p.njt = d.turb[oj + k]; p.nje = d.tdiss[oj + k]; }
                }
            }
        }
    };
    // row 0's -JS neighbours lie in the previous strip's last row: read after the wait
    auto locate_prev = [&]() {
        if (mn.adv) {
            const size_t oj = c3(d, in_, jn_ - JS);
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                int k = mn.KK + 1 + lane + WAVE * m;
                if (k <= mn.VI) { pre[m].njt = ld_handoff(d.turb + oj + k); pre[m].nje = ld_handoff(d.tdiss + oj + k); }
            }
        }
    };
    // the chain wave's levels of the column located for step s (one lane writes)
    auto put_info = [&](int s) {
        if (lane == 0) {
            int *si = sinfo + (s & 1) * 5 * ROWS;
            si[r * 5] = mn.adv;
            if (mn.adv) { si[r * 5 + 1] = mn.KK + 1; si[r * 5 + 2] = mn.VI; }
        }
    };
    locate(0);
    put_info(0);
    if (r == 0) {
        sw.wait(ROWS);
        locate_prev();
    }
    // the outputs of a row's previous column (levels plo..pvi) are stored during
    // the next step's assembly, before that step's publish
    int plo = 1 << 30, pvi = -1;
    size_t po = 0;
    auto store_prev = [&]() {
        gf32 *TW = d.turb + po, *EW = d.tdiss + po;
        for (int k = plo + lane; k <= pvi; k += WAVE) { st_through(TW + k, sOT[k]); st_through(EW + k, sOE[k]); }
    };
    for (int s = 0; s < nsteps; ++s) {
        const bool act = mn.adv;
        const Meta mt = mn;
        const int i = in_, j = jn_;
        int *si = sinfo + (s & 1) * 5 * ROWS;
        const int *sp = sinfo + ((s - 1) & 1) * 5 * ROWS;
        if (act) {
            const int lo = mt.KK + 1, VI = mt.VI;
            // levels of the two new neighbours that hold this sweep's values
            const bool w_in = s - r >= 1;
            int slo = 1 << 30, svi = -1;
            if (r >= 1 && sp[(r - 1) * 5]) { slo = sp[(r - 1) * 5 + 1]; svi = sp[(r - 1) * 5 + 2]; }
            const float *sOTs = sOT - (size_t)TKE_SLOTS * nkl, *sOEs = sOE - (size_t)TKE_SLOTS * nkl;
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                int k = lo + lane + WAVE * m;
                if (k <= VI) {
                    const TkePre &p = pre[m];
                    float tA = p.nit, eA = p.nie, tB = p.njt, eB = p.nje;
                    if (w_in && k >= plo && k <= pvi) { tA = sOT[k]; eA = sOE[k]; }
                    if (k >= slo && k <= svi) { tB = sOTs[k]; eB = sOEs[k]; }
                    const float tW = (IS == 1) ? tA : p.ta, tE = (IS == 1) ? p.ta : tA;
                    const float tS = (JS == 1) ? tB : p.tb, tN = (JS == 1) ? p.tb : tB;
                    const float eW = (IS == 1) ? eA : p.ea, eE = (IS == 1) ? p.ea : eA;
                    const float eS = (JS == 1) ? eB : p.eb, eN = (JS == 1) ? p.eb : eB;
                    const float4 c0 = p.c0, c1 = p.c1, c2v = p.c2, c3v = p.c3;
                    const float D1 = c0.x * tW + c0.y * tS + c0.z * tE + c0.w * tN + c1.x + c1.y + c1.z;
                    const float D2 = c0.x * eW + c0.y * eS + c0.z * eE + c0.w * eN + c1.w + c2v.x;
                    sD1[k] = D1; sD2[k] = D2;
                    sC1[k] = c2v.y; sN1[k] = c3v.x; sP1[k] = c3v.y;
                    sC2[k] = c2v.z; sN2[k] = c3v.z; sP2[k] = c3v.w; sWF[k] = c2v.w;
                    sWT[k] = p.c4.x; sWE[k] = p.c4.y;
                    if (k == lo) { sQ1[k] = D1 / c3v.x; sQ2[k] = D2 / c3v.z; }
                }
            }
            if (lane == 0) { si[r * 5 + 3] = __float_as_int(topt); si[r * 5 + 4] = __float_as_int(tope); }
        }
        store_prev();
        plo = act ? mt.KK + 1 : 1 << 30;
        pvi = act ? mt.VI : -1;
        po = c3(d, i, j);
        lds_barrier();
        if (r == ROWS - 1 && s > 0 && s % STRIP_PUBLISH == 0) strip_publish(mine, s, lane);
        locate(s + 1);
        if (s + 1 < nsteps) put_info(s + 1);
        if (r == 0) {
            sw.wait(s + 1 + ROWS);
            locate_prev();
        }
        lds_barrier();                                        // the recurrence outputs are ready
    }
    store_prev();
    if (r == ROWS - 1) strip_publish(mine, nsteps, lane);
}

static inline int lds_slot_h(int nkp) { return (nkp + 2 * SLOT_PAD) | 1; }
// raises a kernel's dynamic LDS limit to bytes (above the default 64 KB; the
// device's is 160 KB) the first time a launch needs it; 1 on failure
static inline int lds_limit(const void *kern, size_t bytes) {
    static std::mutex mu;
    static const void *keys[32];
    static size_t set[32];
    static int n = 0;
    if (bytes <= 65536) return 0;
    std::lock_guard<std::mutex> g(mu);
    int q = 0;
    while (q < n && keys[q] != kern) ++q;
    if (q < n && bytes <= set[q]) return 0;
    if (q == n && n == 32) return 1;
    if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) return 1;
    keys[q] = kern;
    set[q] = bytes;
    if (q == n) ++n;
    return 0;
}
// longest diagonal t over the situations; the pressure sweep runs a over j
// (outer) and b over i, the momentum sweeps a over i and b over j
static inline int diag_len_max(const Dom *hd, int nsit, int t, int oa, int ia, bool pressure) {
    int m = 0;
    for (int q = 0; q < nsit; ++q) {
        int ob = pressure ? hd[q].njj - 1 : hd[q].nii - 1, ib = pressure ? hd[q].nii - 1 : hd[q].njj - 1;
        if (t < ndiag(oa, ob, ia, ib)) m = max(m, diag_len(t, oa, ob, ia, ib));
    }
    return m;
}

// Optional per-launch timing: events recorded on the solver's stream right
// before and after every launch, tagged with the kernel family.
struct Probe {
    hipEvent_t *ev;      // 2*cap events
    int *fam;            // cap family tags
    int cap, n;
};
enum { FAM_STENCIL = 0, FAM_PRESSURE = 1, FAM_U = 2, FAM_V = 3, FAM_W = 4, FAM_TKE = 5, FAM_COEF = 6 };

struct Launcher {
    hipStream_t s;
    Probe *pr;
    long long n = 0;
    template <class F>
    void go(int fam, F &&f) {
        bool rec = pr && pr->n < pr->cap;
        if (rec) { (void)hipEventRecord(pr->ev[2 * pr->n], s); }
        f();
        if (rec) { (void)hipEventRecord(pr->ev[2 * pr->n + 1], s); pr->fam[pr->n] = fam; pr->n++; }
        ++n;
    }
};

// Launches one SIMPLE iteration of nsit situations that advance in lock step
// (same iteration number, hence the same sweep directions).  hd: host copies
// of the situations' Dom, dd: the same array in device memory.  Grid slice z
// of every launch is situation z; grid x/y are sized for the largest
// situation and each kernel skips what lies outside its own grid.  All
// situations share the turbulence model and the sweep kind.
static int launch_iteration(const Dom *hd, int nsit, const Dom *dd, int iter, int slot, hipStream_t s,
                            long long *launches, Probe *probe, bool proj);
extern "C" int pf_launch_iteration(const Dom *hd, int nsit, const Dom *dd, int iter, int slot, hipStream_t s,
                                   long long *launches, Probe *probe) {
    return launch_iteration(hd, nsit, dd, iter, slot, s, launches, probe, false);
}
// The pressure-correction part of iteration iter only (divergence, pressure
// sweep, velocity correction): the forward of the adjoint.  No residual slot.
extern "C" int pf_launch_projection(const Dom *hd, const Dom *dd, int iter, hipStream_t s, long long *launches) {
    return launch_iteration(hd, 1, dd, iter, -1, s, launches, nullptr, true);
}
static int launch_iteration(const Dom *hd, int nsit, const Dom *dd, int iter, int slot, hipStream_t s,
                            long long *launches, Probe *probe, bool proj) {
    if (nsit < 1) return 0;
    Launcher L{s, probe};
    const Dom &d0 = hd[0];
    int nii = 0, njj = 0, nk = 0, nkp = 0, nj2 = 0;
    for (int q = 0; q < nsit; ++q) {
        nii = max(nii, hd[q].nii); njj = max(njj, hd[q].njj); nk = max(nk, hd[q].kadvmax);
        nkp = max(nkp, hd[q].nkp); nj2 = max(nj2, hd[q].nj2);
    }
    const unsigned Z = (unsigned)nsit;
    int kb = (nk + 63) / 64;
    int IS = 1, JS = 1;
    switch (iter % 4) {
        case 0: IS = -1; JS = 1; break;
        case 1: IS = -1; JS = -1; break;
        case 2: IS = 1; JS = -1; break;
        default: break;
    }
    // every strip kernel this iteration launches gets its dynamic-LDS limit
    // before the first launch, so a refusal leaves the fields untouched
    if (d0.tiled == 2) {
        const size_t shm_p = sizeof(float) * ((size_t)(8 * ROWS + RING) * lds_slot_h(nkp) + 2 * nkp + 8 + WAVE * CH) +
                             sizeof(int) * (9 * ROWS + 4);
        const size_t shm_m = sizeof(float) * ((size_t)9 * lds_slot_h(nkp) * ROWS + nkp + 8 + WAVE * CH) +
                             sizeof(int) * (10 * ROWS + 4);
        int bad = lds_limit((const void *)k_pressure_strip, shm_p);
        if (d0.model == 1) {
            bad |= lds_limit((const void *)k_ml_strip<0>, shm_m) | lds_limit((const void *)k_ml_strip<1>, shm_m) |
                   lds_limit((const void *)k_ml_strip<2>, shm_m);
        } else if (d0.model == 2) {
            bad |= lds_limit((const void *)k_ml_strip<0, 1>, shm_m) | lds_limit((const void *)k_ml_strip<1, 1>, shm_m) |
                   lds_limit((const void *)k_ml_strip<2, 1>, shm_m);
            bad |= lds_limit((const void *)k_tke_coef,
                             sizeof(float) * 5 * (nkp | 1) * COEF_COLS + sizeof(int) * (TKE_SCOL * COEF_COLS + 4));
            bad |= lds_limit((const void *)k_tke_strip, sizeof(float) * ((size_t)TKE_SLOTS * ROWS * lds_slot_h(nkp) + WAVE * CH) +
                                                            sizeof(int) * (10 * ROWS + 4));
        } else {
            bad |= lds_limit((const void *)k_ml_strip<0, 2>, shm_m) | lds_limit((const void *)k_ml_strip<1, 2>, shm_m) |
                   lds_limit((const void *)k_ml_strip<2, 2>, shm_m);
        }
        if (bad) return (int)hipErrorInvalidValue;
    }
    int ncols = (nii - 2) * (njj - 2);
    const size_t bc_shm = bc_lds_bytes(max(nii, njj) + 2);
    const unsigned bc_threads = (unsigned)min(64 * kb, 1024);
    if (!proj) {
        L.go(FAM_STENCIL, [&] { hipLaunchKernelGGL(k_bc_ew, dim3(1, njj - 1, Z), dim3(bc_threads), bc_shm, s, dd); });
        L.go(FAM_STENCIL, [&] { hipLaunchKernelGGL(k_bc_ns, dim3(1, nii - 1, Z), dim3(bc_threads), bc_shm, s, dd); });
    }
    const int nrow8 = (nii - 2 + NXCD - 1) / NXCD * NXCD;                 // x extent of the cell kernels (rows)
    L.go(FAM_STENCIL, [&] { hipLaunchKernelGGL(k_divergence, dim3(nrow8, 1, Z), dim3(CELL_THREADS), 0, s, dd); });
    size_t nrow = (size_t)nj2 * nkp;
    for (int q = 0; q < nsit; ++q) nrow = max(nrow, (size_t)5 * hd[q].sync_stride);
    L.go(FAM_STENCIL, [&] {
        hipLaunchKernelGGL(k_zero_rows, dim3((unsigned)((nrow + 255) / 256), 1, Z), dim3(256), 0, s, dd, d0.tiled ? 0 : 1);
    });
    if (d0.tiled == 2) {
        int nst = 0;
        for (int q = 0; q < nsit; ++q) nst = max(nst, (hd[q].nii - 1 - 2 + ROWS) / ROWS);
        size_t shm = sizeof(float) * ((size_t)(8 * ROWS + RING) * lds_slot_h(nkp) + 2 * nkp + 8 + WAVE * CH) +
                     sizeof(int) * (9 * ROWS + 4);
        if (lds_limit((const void *)k_pressure_strip, shm)) return (int)hipErrorInvalidValue;
        L.go(FAM_PRESSURE, [&] {
            hipLaunchKernelGGL(k_pressure_strip, dim3(nst, 1, Z), dim3(STRIP_WAVES * WAVE), shm, s, dd, IS, JS, slot);
        });
    } else {
        int oa = 2, ob = njj - 1, ia = 2, ib = nii - 1;
        size_t shm = sizeof(float) * 5 * nkp;
        for (int t = 0; t < ndiag(oa, ob, ia, ib); ++t) {
            int len = diag_len_max(hd, nsit, t, 2, 2, true);
            if (len > 0)
                L.go(FAM_PRESSURE, [&] {
                    hipLaunchKernelGGL(k_pressure, dim3(len, 1, Z), dim3(WAVE), shm, s, dd, t, IS, JS, slot);
                });
        }
    }
    L.go(FAM_STENCIL, [&] { hipLaunchKernelGGL(k_correction, dim3(nrow8, 1, Z), dim3(CELL_THREADS), 0, s, dd); });
    if (proj) {
        if (launches) *launches += L.n;
        return (int)hipGetLastError();
    }
    L.go(FAM_STENCIL, [&] { hipLaunchKernelGGL(k_averages, dim3(nrow8, 1, Z), dim3(CELL_THREADS), 0, s, dd); });
    L.go(FAM_STENCIL, [&] { hipLaunchKernelGGL(k_avg_edges, dim3(njj - 2, 1, Z), dim3(64), 0, s, dd); });
    size_t shm = sizeof(float) * 5 * nkp;
    // per-diagonal sweeps: sweep ranges a in [oa, nii-1] (outer), b in [ia, njj-1]
    auto sweep = [&](int fam, auto kern, int oa, int ia, int is, int js) {
        int nd = ndiag(oa, nii - 1, ia, njj - 1);
        for (int t = 0; t < nd; ++t) {
            int len = diag_len_max(hd, nsit, t, oa, ia, false);
            if (len > 0)
                L.go(fam, [&] { hipLaunchKernelGGL(kern, dim3(len, 1, Z), dim3(WAVE), shm, s, dd, t, is, js); });
        }
    };
    int lds_err = 0;
    auto strip_sweep = [&](int fam, auto coef, auto strip, int oa, int ia, int is, int js) {
        int nblk = 0, nst = 0;
        for (int q = 0; q < nsit; ++q) {
            int ob = hd[q].nii - 1, ib = hd[q].njj - 1;
            int na = ob - oa + 1, nb = ib - ia + 1;
            nblk = max(nblk, ((na * nb + COEF_COLS - 1) / COEF_COLS + NXCD - 1) / NXCD * NXCD);
            nst = max(nst, (ib - ia + ROWS) / ROWS);
        }
        size_t shm_c = sizeof(float) * 3 * (nkp | 1) * COEF_COLS + sizeof(int) * (SCOL * COEF_COLS + 4);
        L.go(FAM_COEF, [&] { hipLaunchKernelGGL(coef, dim3(nblk, 1, Z), dim3(4 * WAVE), shm_c, s, dd, is, js); });
        size_t shm_t = sizeof(float) * ((size_t)9 * lds_slot_h(nkp) * ROWS + nkp + 8 + WAVE * CH) + sizeof(int) * (10 * ROWS + 4);
        if (lds_limit((const void *)strip, shm_t)) { lds_err = 1; return; }
        L.go(fam, [&] { hipLaunchKernelGGL(strip, dim3(nst, 1, Z), dim3(STRIP_WAVES * WAVE), shm_t, s, dd, is, js); });
    };
    bool all_w8 = true;
    for (int q = 0; q < nsit; ++q) all_w8 = all_w8 && hd[q].simd == 8;
    if (d0.model == 1 && d0.tiled == 2 && all_w8) {
        strip_sweep(FAM_U, k_ml_coef<0, 8>, k_ml_strip<0>, 3, 2, IS, JS);
        strip_sweep(FAM_V, k_ml_coef<1, 8>, k_ml_strip<1>, 2, 3, -IS, -JS);
        strip_sweep(FAM_W, k_ml_coef<2, 8>, k_ml_strip<2>, 2, 2, IS, JS);
    } else if (d0.model == 1 && d0.tiled == 2) {
        strip_sweep(FAM_U, k_ml_coef<0, 0>, k_ml_strip<0>, 3, 2, IS, JS);
        strip_sweep(FAM_V, k_ml_coef<1, 0>, k_ml_strip<1>, 2, 3, -IS, -JS);
        strip_sweep(FAM_W, k_ml_coef<2, 0>, k_ml_strip<2>, 2, 2, IS, JS);
    } else if (d0.model == 1) {
        sweep(FAM_U, k_momentum<0, 0>, 3, 2, IS, JS);
        sweep(FAM_V, k_momentum<1, 0>, 2, 3, -IS, -JS);
        sweep(FAM_W, k_momentum<2, 0>, 2, 2, IS, JS);
    } else if (d0.model == 2) {
        if (d0.tiled == 2) {
            // the k-eps V solver sweeps in the same directions as U and W
            strip_sweep(FAM_U, k_ml_coef<0, 0, 1>, k_ml_strip<0, 1>, 3, 2, IS, JS);
            strip_sweep(FAM_V, k_ml_coef<1, 0, 1>, k_ml_strip<1, 1>, 2, 3, IS, JS);
            strip_sweep(FAM_W, k_ml_coef<2, 0, 1>, k_ml_strip<2, 1>, 2, 2, IS, JS);
        } else {
            sweep(FAM_U, k_momentum<0, 1>, 3, 2, IS, JS);
            sweep(FAM_V, k_momentum<1, 1>, 2, 3, IS, JS);
            sweep(FAM_W, k_momentum<2, 1>, 2, 2, IS, JS);
        }
        if (d0.tiled == 2) {
            // strip-pipelined TKE: the neighbour-independent pre-pass, then the strips
            int nblk_t = 0;
            for (int q = 0; q < nsit; ++q)
                nblk_t = max(nblk_t, (((hd[q].nii - 2) * (hd[q].njj - 2) + COEF_COLS - 1) / COEF_COLS + NXCD - 1) / NXCD * NXCD);
            const size_t shm_c = sizeof(float) * 5 * (nkp | 1) * COEF_COLS + sizeof(int) * (TKE_SCOL * COEF_COLS + 4);
            if (lds_limit((const void *)k_tke_coef, shm_c)) return (int)hipErrorInvalidValue;
            L.go(FAM_COEF, [&] { hipLaunchKernelGGL(k_tke_coef, dim3(nblk_t, 1, Z), dim3(4 * WAVE), shm_c, s, dd); });
            int nst = 0;
            for (int q = 0; q < nsit; ++q) nst = max(nst, (hd[q].njj - 1 - 2 + ROWS) / ROWS);
            const size_t shm_k = sizeof(float) * ((size_t)TKE_SLOTS * ROWS * lds_slot_h(nkp) + WAVE * CH) + sizeof(int) * (10 * ROWS + 4);
            if (lds_limit((const void *)k_tke_strip, shm_k)) return (int)hipErrorInvalidValue;
            L.go(FAM_TKE, [&] {
                hipLaunchKernelGGL(k_tke_strip, dim3(nst, 1, Z), dim3(STRIP_WAVES * WAVE), shm_k, s, dd, IS, JS);
            });
        } else {
            size_t shm_t = sizeof(float) * 12 * nkp;
            int nd = ndiag(2, nii - 1, 2, njj - 1);
            for (int t = 0; t < nd; ++t) {
                int len = diag_len_max(hd, nsit, t, 2, 2, false);
                if (len > 0)
                    L.go(FAM_TKE, [&] { hipLaunchKernelGGL(k_tke, dim3(len, 1, Z), dim3(WAVE), shm_t, s, dd, t, IS, JS); });
            }
        }
    } else if (d0.tiled == 2) {
        strip_sweep(FAM_U, k_ml_coef<0, 0, 2>, k_ml_strip<0, 2>, 3, 2, IS, JS);
        strip_sweep(FAM_V, k_ml_coef<1, 0, 2>, k_ml_strip<1, 2>, 2, 3, IS, JS);
        strip_sweep(FAM_W, k_ml_coef<2, 0, 2>, k_ml_strip<2, 2>, 2, 2, IS, JS);
    } else {
        sweep(FAM_U, k_momentum<0, 2>, 3, 2, IS, JS);
        sweep(FAM_V, k_momentum<1, 2>, 2, 3, IS, JS);
        sweep(FAM_W, k_momentum<2, 2>, 2, 2, IS, JS);
    }
    if (launches) *launches += L.n;
    if (lds_err) return (int)hipErrorInvalidValue;
    return (int)hipGetLastError();
}

extern "C" int pf_fill_u32(unsigned *p, size_t n, unsigned v, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return (int)hipGetLastError();
}

// ------------------------------------------------------------------------
// Flow-field file payloads, GFFWrite.cs: int16 of value*100 (FP32 product)
// kind 0 (:167-225): column records (u, v, w per level), clamp then truncate;
// kind 1 (:256-360): level-major rows u(j), v(j), w(j), clamp then truncate;
// kind 2 (:388-510): KKART, then per column u, v, w from max(KKART,1) up,
//                    MathF.Round (half to even) then clamp;
// kind 3 (:79-133, *.tke): TURB, clamp then Convert.ToInt16 (half to even).
// u and v of file level k are UK/VK[k-1], w is WK[k].  A NaN packs to 0.
// ------------------------------------------------------------------------
__device__ __forceinline__ float gff_at(const GffArgs &g, int f, size_t col, int k) {
    return k < g.nlo ? g.lo[f][col * (size_t)g.nkp + k] : g.up[f][col * (size_t)g.nup + (k - g.nlo)];
}
__device__ __forceinline__ int16_t gff_trunc(float v) {
    if (v != v) return 0;                                     // NaN
    if (v > 32767.0f) v = 32767.0f;
    else if (v < -32768.0f) v = -32768.0f;
    return (int16_t)(int)v;                                   // truncation toward zero
}
__device__ __forceinline__ int16_t gff_round(float v) {
    if (v != v) return 0;
    v = rintf(v);                                             // half to even
    if (v > 32767.0f) v = 32767.0f;
    else if (v < -32768.0f) v = -32768.0f;
    return (int16_t)(int)v;
}

// kinds 0 and 3: one block per column (a, b), threads over levels
__global__ void k_gff_column(GffArgs g) {
    const int nb = (g.kind == 3) ? g.njj : g.njj + 1;
    const int a = blockIdx.x / nb, b = blockIdx.x % nb;
    const size_t col = (size_t)(a + 1) * g.nj2 + (b + 1);
    if (g.kind == 3) {
        int16_t *o = g.out + ((size_t)a * g.njj + b) * g.nkk;
        for (int k = 1 + threadIdx.x; k <= g.nkk; k += blockDim.x) o[k - 1] = gff_round(gff_at(g, 3, col, k) * 100.0f);
        return;
    }
    int16_t *o = g.out + ((size_t)a * (g.njj + 1) + b) * (size_t)(g.nkk + 1) * 3;
    for (int kq = threadIdx.x; kq <= g.nkk; kq += blockDim.x) {
        o[3 * kq] = gff_trunc(gff_at(g, 0, col, kq) * 100.0f);
        o[3 * kq + 1] = gff_trunc(gff_at(g, 1, col, kq) * 100.0f);
        o[3 * kq + 2] = gff_trunc(gff_at(g, 2, col, kq + 1) * 100.0f);
    }
}

// kind 1: block (a, 64 columns b), levels in tiles of 32: the column reads
// (contiguous in k) are transposed through LDS into the rows of the file
// (contiguous in b)
#define GFF_BJ 64
#define GFF_KT 32
__global__ void __launch_bounds__(256) k_gff_rows(GffArgs g) {
    __shared__ int16_t tile[3][GFF_KT][GFF_BJ + 2];
    const int a = blockIdx.x, b0 = blockIdx.y * GFF_BJ;
    const int n = g.njj + 1, nb = min(GFF_BJ, n - b0), nk = g.nkk + 1;
    for (int k0 = 0; k0 < nk; k0 += GFF_KT) {
        const int kt = min(GFF_KT, nk - k0);
        for (int t = threadIdx.x; t < GFF_BJ * GFF_KT; t += blockDim.x) {
            const int cb = t / GFF_KT, kk = t % GFF_KT;
            if (cb < nb && kk < kt) {
                const size_t col = (size_t)(a + 1) * g.nj2 + (b0 + cb + 1);
                const int kq = k0 + kk;
                tile[0][kk][cb] = gff_trunc(gff_at(g, 0, col, kq) * 100.0f);
                tile[1][kk][cb] = gff_trunc(gff_at(g, 1, col, kq) * 100.0f);
                tile[2][kk][cb] = gff_trunc(gff_at(g, 2, col, kq + 1) * 100.0f);
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < GFF_KT * 3 * GFF_BJ; t += blockDim.x) {
            const int cb = t % GFF_BJ, cp = (t / GFF_BJ) % 3, kk = t / (3 * GFF_BJ);
            if (cb < nb && kk < kt) {
                const size_t row = ((size_t)(k0 + kk) * (g.nii + 1) + a) * 3 + cp;
                g.out[row * n + b0 + cb] = tile[cp][kk][cb];
            }
        }
        __syncthreads();
    }
}

// kind 2: block per column (a, b) of the record part; block 0 also writes KKART
__global__ void k_gff_terrain(GffArgs g) {
    const int a = blockIdx.x / (g.njj + 1), b = blockIdx.x % (g.njj + 1);
    if (blockIdx.x == 0)
        for (size_t t = threadIdx.x; t < (size_t)g.nii * g.njj; t += blockDim.x) {
            const int ia = (int)(t / g.njj), jb = (int)(t % g.njj);
            g.out[t] = g.kk[(size_t)(ia + 1) * g.nj2 + (jb + 1)];
        }
    const size_t col = (size_t)(a + 1) * g.nj2 + (b + 1);
    const int k0 = max((int)g.kk[col], 1);
    const int n = g.nkk + 2 - k0;
    int16_t *o = g.out + (size_t)g.nii * g.njj + g.off[blockIdx.x];
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        o[t] = gff_round(gff_at(g, 0, col, k0 - 1 + t) * 100.0f);
        o[n + t] = gff_round(gff_at(g, 1, col, k0 - 1 + t) * 100.0f);
        o[2 * n + t] = gff_round(gff_at(g, 2, col, k0 + t) * 100.0f);
    }
}

// ------------------------------------------------------------------------
// The particle model's wind at a particle (Zeitschleife.IntWindCalculate,
// Intwind.cs:31-170): the level from HOKART (terrain: BinarySearch of the
// height above AHMIN, Intwind.cs:174-210; flat terrain with buildings: the
// first level whose top is at or above the particle), u and v linear in x and
// y across the cell faces (in double: the particle coordinate is a double),
// w linear in z between the level's bottom and top faces.
// ------------------------------------------------------------------------
__device__ __forceinline__ float wind_field(const WindArgs &a, int f, int i, int j, int k) {
    const size_t c = (size_t)i * a.nj2 + j;
    return k < a.nlo ? a.lo[f][c * a.nkp + k] : a.up[f][c * a.nup + (k - a.nlo)];
}

__global__ __launch_bounds__(256) void k_wind_at_particles(WindArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const int i = a.ci[p], j = a.cj[p];
    const float z = a.z[p];
    float f1 = 0.0f;
    int k;
    if (a.topo) {
        f1 = z - a.ahmin;
        int low = 1, high = a.nkk;
        const int q = high >> 3;
        if (a.hok[q] >= f1) high = q;
        else low = q;
        while (low <= high) {
            const int mid = (low + high) >> 1;
            if (a.hok[mid] >= f1) high = mid - 1;
            else low = mid + 1;
        }
        k = min(low, a.nkk);
    } else {
        k = 1;
        while (k <= a.nkk && !(z <= a.hok[k])) ++k;
        k = min(k, a.nkk);
    }
    const float DXK = a.dxk, DYK = a.dyk;
    const float u1 = wind_field(a, 0, i, j, k), u2 = wind_field(a, 0, i + 1, j, k);
    const float ox = (float)a.ikoo + (float)(i - 1) * DXK;
    a.ux[p] = (float)((double)u1 + (double)((u2 - u1) / DXK) * (a.x[p] - (double)ox));
    const float v1 = wind_field(a, 1, i, j, k), v2 = wind_field(a, 1, i, j + 1, k);
    const float oy = (float)a.jkoo + (float)(j - 1) * DYK;
    a.uy[p] = (float)((double)v1 + (double)((v2 - v1) / DYK) * (a.y[p] - (double)oy));
    const float w1 = wind_field(a, 2, i, j, k), w2 = wind_field(a, 2, i, j, k + 1);
    const float dz = a.topo ? f1 - a.hok[k - 1] : (z + a.ahint[p]) - a.hok[k - 1];
    a.uz[p] = w1 + ((w2 - w1) / a.dzk[k]) * dz;
    a.ck[p] = k;
}

extern "C" int pf_launch_wind(const WindArgs *wa, hipStream_t s) {
    if (wa->n <= 0) return 0;
    hipLaunchKernelGGL(k_wind_at_particles, dim3((unsigned)((wa->n + 255) / 256)), dim3(256), 0, s, *wa);
    return (int)hipGetLastError();
}

extern "C" int pf_launch_gff(const GffArgs *ga, hipStream_t s) {
    const GffArgs g = *ga;
    if (g.kind == 0)
        hipLaunchKernelGGL(k_gff_column, dim3((unsigned)((g.nii + 1) * (g.njj + 1))), dim3(256), 0, s, g);
    else if (g.kind == 3)
        hipLaunchKernelGGL(k_gff_column, dim3((unsigned)(g.nii * g.njj)), dim3(256), 0, s, g);
    else if (g.kind == 1)
        hipLaunchKernelGGL(k_gff_rows, dim3((unsigned)(g.nii + 1), (unsigned)((g.njj + GFF_BJ) / GFF_BJ)), dim3(256), 0, s, g);
    else
        hipLaunchKernelGGL(k_gff_terrain, dim3((unsigned)((g.nii + 1) * (g.njj + 1))), dim3(256), 0, s, g);
    return (int)hipGetLastError();
}

// ------------------------------------------------------------------------
// FP64 mass imbalance of the current velocity field: over the active cells
// (ADVDOM = 1, 2 <= i <= NII-1, 2 <= j <= NJJ-1, KKART < k <= VerticalIndex)
// the sum of |DIV|, with DIV evaluated in FP32 exactly as the divergence pass
// forms it (PrognosticFlowfield.cs:723-758, the bottom flux dropped at
// KKART+1), and the sum of the absolute fluxes through the cell faces the
// same expression uses.  Each FP32 term is widened to double before it is
// accumulated.  Stage 1: a fixed range of columns per block, one column per
// wave with the lanes over k (coalesced), wavefront butterfly reductions in
// double, one partial pair per block.  Stage 2: one block per situation sums
// the partials.  The partition and the reduction trees are fixed, so the
// result is the same bits on every run.
// ------------------------------------------------------------------------
#define MASS_THREADS 256
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}
__device__ __forceinline__ void block_sum2_f64(double &a, double &b) {
    __shared__ double red[MASS_THREADS / WAVE][2];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    a = wave_sum_f64(a);
    b = wave_sum_f64(b);
    if (lane == 0) { red[w][0] = a; red[w][1] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = 0; b = 0;
        for (int q = 0; q < MASS_THREADS / WAVE; ++q) { a += red[q][0]; b += red[q][1]; }
    }
}
__global__ __launch_bounds__(MASS_THREADS) void k_mass_partial(const Dom *__restrict__ ds, double *__restrict__ partial) {
    const DomG &d = reinterpret_cast<const DomG *>(ds)[blockIdx.z];
    const int nj = d.njj - 2, ncols = (d.nii - 2) * nj;
    const int per = (ncols + (int)gridDim.x - 1) / (int)gridDim.x;
    const int c0 = (int)blockIdx.x * per, c1 = min(ncols, c0 + per);
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    double sd = 0, sf = 0;
    for (int col = c0 + w; col < c1; col += MASS_THREADS / WAVE) {
        const int i = 2 + col / nj, j = 2 + col % nj;
        const size_t c = c2(d, i, j);
        if (d.adv[c] != 1) continue;
        const int KK = d.kk[c], VI = d.vi[c];
        const size_t o = c3(d, i, j), oip = c3(d, i + 1, j), ojp = c3(d, i, j + 1);
        for (int k = KK + 1 + lane; k <= VI; k += WAVE) {
            const float dz = d.dzk[k];
            const float u0 = d.uk[o + k], u1 = d.uk[oip + k], v0 = d.vk[o + k], v1 = d.vk[ojp + k];
            const float w0 = d.wk[o + k], w1 = d.wk[o + k + 1];
            const float t1 = (u0 - u1) * d.dyk * dz;
            const float t2 = (v0 - v1) * d.dxk * dz;
            const float dv = (k > KK + 1) ? t1 + t2 + (w0 - w1) * d.area : t1 + t2 - w1 * d.area;
            sd += (double)fabsf(dv);
            double f = (double)fabsf(u0 * d.dyk * dz) + (double)fabsf(u1 * d.dyk * dz) +
                       (double)fabsf(v0 * d.dxk * dz) + (double)fabsf(v1 * d.dxk * dz) + (double)fabsf(w1 * d.area);
            if (k > KK + 1) f += (double)fabsf(w0 * d.area);
            sf += f;
        }
    }
    block_sum2_f64(sd, sf);
    if (threadIdx.x == 0) {
        double *p = partial + 2 * ((size_t)blockIdx.z * gridDim.x + blockIdx.x);
        p[0] = sd;
        p[1] = sf;
    }
}
__global__ __launch_bounds__(MASS_THREADS) void k_mass_final(const double *__restrict__ partial, int nblk,
                                                             double *__restrict__ out) {
    const double *p = partial + 2 * (size_t)blockIdx.x * nblk;
    double sd = 0, sf = 0;
    for (int b = threadIdx.x; b < nblk; b += MASS_THREADS) { sd += p[2 * b]; sf += p[2 * b + 1]; }
    block_sum2_f64(sd, sf);
    if (threadIdx.x == 0) { out[2 * blockIdx.x] = sd; out[2 * blockIdx.x + 1] = sf; }
}
// partial: [nsit][nblk][2] doubles of scratch; out: [nsit][2] (sum |DIV|, sum |face flux|)
extern "C" int pf_launch_mass(const Dom *dd, int nsit, int nblk, double *partial, double *out, hipStream_t s) {
    if (nsit < 1 || nblk < 1) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mass_partial, dim3((unsigned)nblk, 1, (unsigned)nsit), dim3(MASS_THREADS), 0, s, dd, partial);
    hipLaunchKernelGGL(k_mass_final, dim3((unsigned)nsit), dim3(MASS_THREADS), 0, s, partial, nblk, out);
    return (int)hipGetLastError();
}

// ------------------------------------------------------------------------
// Adjoint of one mixing-length momentum sweep (U/V/W-prognostic-microscale_1.cs:
// coefficients :194-470, Thomas forward :473-504, relaxed back-substitution
// :514-524) with respect to the swept component X, with everything else the
// sweep reads held fixed (UKS, VKS, WKS, DPMNEW, the other components).  X
// enters a column's equations through the eddy viscosity (own column, levels
// k-1 and k+1: the conductances, hence every power-law coefficient, and the
// cross-term factor), the Coriolis/vegetation and wall terms (own level), the
// cross terms (own level and the two neighbours along the component), the four
// neighbour couplings (new values of the two columns visited before it, old
// values of the two after it) and the relaxation.  The transpose runs the
// sweep's anti-diagonals from the last to the first, one wave per column: the
// lanes recompute the column's coefficients from the checkpointed fields, lane
// 0 runs the Thomas recurrences forward and then both recurrences transposed,
// and the lanes push the adjoints back through the coefficient assembly.  The
// adjoint handed to another column has one writer (buffers indexed by the
// handing column, per direction) and is gathered by its receiver, so the
// result is deterministic.
// ------------------------------------------------------------------------

// d VIS / d X[k+1] of ml_visc (d VIS / d X[k-1] is its negative) for the
// component whose vertical shear enters VIS (U: UK, V: VK; W's VIS holds no WK)
template <int COMP>
__device__ __forceinline__ float ml_visc_dx(const DomG &d, const ColInfo &q, int k, const gf32 *UK, const gf32 *VK) {
    if (COMP == 2 || !(k > q.KK + 1)) return 0.0f;
    const float dz = d.dzk[k];
    const float zs = q.v512() ? d.hok[k] - fmaf(dz, 0.5f, d.hok[q.KK]) : d.hok[k] - d.hok[q.KK] - dz * 0.5f;
    float mixL = vminf(0.071f * zs, 100.0f);
    if (q.cov > 0) mixL *= (1.0f - 0.99f * q.cov);
    const float iu = (UK[k + 1] - UK[k - 1]) / (2.0f * dz), iv = (VK[k + 1] - VK[k - 1]) / (2.0f * dz);
    const float r = sqrtf(0.5f * (iu * iu + iv * iv));
    if (!(mixL * mixL * r < 15.0f) || !(r > 0.0f)) return 0.0f;      // at the cap of 15, or no shear
    return mixL * mixL * 0.5f * ((COMP == 0) ? iu : iv) / r / (2.0f * dz);
}

// d A / d D of a power-law coefficient A = D max(0, (1 - 0.1 P)^5) + F+ with
// P = |F / D| and F fixed
__device__ __forceinline__ float plaw_dd(float P) {
    const float x = 1.0f - 0.1f * P;
    if (!(x > 0.0f)) return 0.0f;
    const float x4 = (x * x) * (x * x);
    return x4 * x + 0.5f * x4 * P;
}

// d A / d VIS of the six coefficients of ml_coeffs (AE1, AW1, AS1, AN1, BIM, CIM)
__device__ __forceinline__ void ml_coeffs_dvis(const DomG &d, const ColInfo &q, int k, float VIS, float FE, float FW,
                                               float FS, float FN, float FT, float FB, float *g) {
    const float dz = d.dzk[k], dxkdzk = d.dxk * dz, dykdzk = d.dyk * dz;
    const float mv = vmaxf(VIS, d.vishmin), dmv = VIS > d.vishmin ? 1.0f : 0.0f;
    auto face = [&](bool off, float geo, float F) {           // conductance mv*geo, floored at 1e-4
        if (off) return 0.0f;
        const float Dr = mv * geo;
        if (!(Dr > 0.0001f)) return 0.0f;
        return dmv * geo * plaw_dd(fabsf(F / Dr));
    };
    const float gE = dykdzk / d.dxk, gS = dxkdzk / d.dyk, gT = d.area / dz;
    g[0] = face(q.ae, gE, FE);
    g[1] = face(q.aw, gE, FW);
    g[2] = face(q.as, gS, FS);
    g[3] = face(q.an, gS, FN);
    g[4] = face(false, gT, FT);
    g[5] = face(!(k > q.KK + 1), gT, FB);
}

#define MA_SLOTS 18
template <int COMP>
__global__ __launch_bounds__(WAVE) void k_adj_mom(const Dom *__restrict__ ds, MomAdj m, int t, int IS, int JS) {
    const DomG &d = *reinterpret_cast<const DomG *>(ds);       // its swept field is X before the sweep
    extern __shared__ float sm[];
    const int oa = (COMP == 0) ? 3 : 2, ia = (COMP == 1) ? 3 : 2;
    int a, b;
    if (!diag_col(t, oa, d.nii - 1, ia, d.njj - 1, blockIdx.x, a, b)) return;
    int i, j;
    sweep_to_ij(d, 0, a, b, IS, JS, i, j);
    if (d.adv[c2(d, i, j)] != 1) return;
    const ColInfo q = col_info(d, i, j);
    const int nkp = d.nkp, lane = threadIdx.x;
    const int KK = q.KK, KKP1 = KK + 1, KSTART = q.KSTART, VI = q.VI;
    int KKm = KK;
    if (COMP == 0) KKm = max(KK, d.kk[c2(d, i - 1, j)]);
    if (COMP == 1) KKm = max(KK, d.kk[c2(d, i, j - 1)]);
    const gf32 *XB = (COMP == 0) ? d.uk : (COMP == 1) ? d.vk : d.wk;
    const gf32 *XA = (const gf32 *)m.xa;
    const size_t o = c3(d, i, j), opa = c3(d, i - IS, j), ona = c3(d, i + IS, j), opb = c3(d, i, j - JS),
                 onb = c3(d, i, j + JS);
    float *sA = sm, *sB = sA + nkp, *sC = sB + nkp, *sD = sC + nkp, *sT = sD + nkp, *sP = sT + nkp, *sQ = sP + nkp;
    float *sG = sQ + nkp, *sXa = sG + nkp, *sAb = sXa + nkp, *sBb = sAb + nkp, *sCb = sBb + nkp, *sDb = sCb + nkp;
    // sR: own X before from the relaxation; sUp / sDn / sMid: from the assembly, for own X before at k+1 /
    // k-1 / k (sUp first holds the adjoint of P between the two transposed recurrences); sQb: adjoint of Q
    float *sR = sDb + nkp, *sUp = sR + nkp, *sDn = sUp + nkp, *sMid = sDn + nkp, *sQb = sMid + nkp;
    const float dyk = d.dyk, dxk = d.dxk;
    // 1. the column's equations from the checkpointed fields, and dL/dX after
    for (int k = KSTART + lane; k <= VI; k += WAVE) {
        const MlPieces pc = ml_pieces<COMP>(d, q, k);
        const float xpa = XA[opa + k], xna = XB[ona + k], xpb = XA[opb + k], xnb = XB[onb + k], x = XB[o + k];
        const float xw = (IS == 1) ? xpa : xna, xe = (IS == 1) ? xna : xpa;
        const float xs = (JS == 1) ? xpb : xnb, xn = (JS == 1) ? xnb : xpb;
        sA[k] = pc.A; sB[k] = pc.B; sC[k] = pc.C;
        sD[k] = ml_source<COMP>(d, q, k, d.dzk[k], pc, xw, xs, xe, xn, x);
        sG[k] = m.gin[o + k] + m.gpa[c3(d, i + IS, j) + k] + m.gpb[c3(d, i, j + JS) + k];
    }
    for (int k = KSTART + lane; k <= VI + 1; k += WAVE) sXa[k] = XA[o + k];
    __syncthreads();
    if (lane == 0) {
        // Thomas forward, as the sweep ran it
        float Pp = 0.0f, Qp = 0.0f;
        for (int k = KSTART; k <= VI; ++k) {
            float T, P, Q;
            if (k > KKP1) { T = 1 / (sA[k] - sC[k] * Pp); P = sB[k] * T; Q = (sD[k] + sC[k] * Qp) * T; }
            else { T = 1 / sA[k]; P = sB[k] * T; Q = sD[k] * T; }
            sT[k] = T; sP[k] = P; sQ[k] = Q;
            Pp = P; Qp = Q;
        }
        // relaxed back-substitution X_k += relax (P_k X_{k+1} + Q_k - X_k) from VI down, transposed
        const float rl = d.relax;
        float ub = 0.0f;
        for (int k = KSTART; k <= VI; ++k) {
            const float g = sG[k] + ub;
            if (COMP == 2 || KKm < k) {
                sR[k] = g * (1.0f - rl);
                sQb[k] = g * rl;
                sUp[k] = g * rl * sXa[k + 1];
                ub = g * rl * sP[k];
            } else {
                sR[k] = g;
                sQb[k] = 0.0f;
                sUp[k] = 0.0f;
                ub = 0.0f;
            }
        }
        sR[VI + 1] = ub;                                  // X[VI+1] is the top value the first level read
        // Thomas forward transposed (VI down to KSTART)
        float Pc = 0.0f, Qc = 0.0f;
        for (int k = VI; k >= KSTART; --k) {
            const float Qt = sQb[k] + Qc, Pt = sUp[k] + Pc;
            const float B = sB[k], C = sC[k], D = sD[k], T = sT[k];
            float Ab, Bb, Cb, Db;
            if (k > KKP1) {
                const float Pm = sP[k - 1], Qm = sQ[k - 1];
                Db = Qt * T;
                const float Tb = Qt * (D + C * Qm) + Pt * B;
                Bb = Pt * T;
                Cb = Qt * Qm * T;
                Qc = Qt * C * T;
                const float denb = -Tb * T * T;
                Ab = denb;
                Cb += -denb * Pm;
                Pc = -denb * C;
            } else {
                Db = Qt * T;
                Bb = Pt * T;
                Ab = -(Qt * D + Pt * B) * T * T;
                Cb = 0.0f;
                Pc = 0.0f;
                Qc = 0.0f;
            }
            sAb[k] = Ab; sBb[k] = Bb; sCb[k] = Cb; sDb[k] = Db;
        }
    }
    __syncthreads();
    // 3. the assembly transposed, lanes over levels
    for (int k = KSTART + lane; k <= VI; k += WAVE) {
        const MlPieces pc = ml_pieces<COMP>(d, q, k);
        const float xpa = XA[opa + k], xna = XB[ona + k], xpb = XA[opb + k], xnb = XB[onb + k], x = XB[o + k];
        const float xw = (IS == 1) ? xpa : xna, xe = (IS == 1) ? xna : xpa;
        const float xs = (JS == 1) ? xpb : xnb, xn = (JS == 1) ? xnb : xpb;
        const float Ab = sAb[k], Bb = sBb[k], Cb = sCb[k], Db = sDb[k];
        const float dz = d.dzk[k];
        // AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + AP0, D = AW1 X_W + AS1 X_S + AE1 X_E + AN1 X_N + ...
        const float awb = Ab + Db * xw, asb = Ab + Db * xs, aeb = Ab + Db * xe, anb = Ab + Db * xn;
        const float bimb = Bb + Ab, cimb = Cb + Ab;
        float xwb = Db * pc.AW1, xsb = Db * pc.AS1, xeb = Db * pc.AE1, xnb_ = Db * pc.AN1, xb = 0.0f;
        float visb = 0.0f, up = 0.0f, dn = 0.0f;
        const bool cross = ml_cross_on(q, k);
        if (COMP != 2) {
            const gf32 *UKS = d.uks + o, *VKS = d.vks + o;
            const gf32 *UKS_W = d.uks + c3(d, i - 1, j), *VKS_W = d.vks + c3(d, i - 1, j);
            float gfac = -d.cor;                              // g = (Cor (UG - x) [- VEG x wh]) AREA DZ
            if (q.vegp) {
                const float in2 = 0.5f * (VKS_W[k] + VKS[k]), in = 0.5f * (UKS_W[k] + UKS[k]);
                const float wh = vmaxf(sqrtf(in * in + in2 * in2), 0.01f);
                gfac = -d.cor - d.veg[o + k - 1] * wh;
            }
            xb += Db * gfac * d.area * dz;
            if (k == KKP1) xb -= Db * ml_wall_uv(d, q, k, 1.0f, UKS, UKS_W, VKS, VKS_W);   // w = x / wh us^2 AREA
            if (cross) {
                if (COMP == 0) {
                    const float dk = dyk * dz;
                    const float S = (xe - x) * dk - (x - xw) * dk + pc.X1 - pc.X2 + pc.X3 - pc.X4;
                    const float Sb = Db * pc.c0;
                    xeb += Sb * dk; xwb += Sb * dk; xb -= 2.0f * Sb * dk;
                    visb += Db * S / dxk;                    // c0 = VIS / DXK
                } else {
                    const float dk = dxk * dz;
                    const float S = pc.X1 - pc.X2 + (xn - x) * dk - (x - xs) * dk + pc.X3 - pc.X4;
                    const float Sb = Db * pc.c0;
                    xnb_ += Sb * dk; xsb += Sb * dk; xb -= 2.0f * Sb * dk;
                    visb += Db * S / dxk;
                }
            }
            float FE, FW, FS, FN, FT, FB, g6[6];
            ml_fluxes<COMP>(d, q, k, FE, FW, FS, FN, FT, FB);
            const gf32 *UK = d.uk + o, *VK = d.vk + o;
            const float VIS = ml_visc(d, q, k, UK, VK);
            ml_coeffs_dvis(d, q, k, VIS, FE, FW, FS, FN, FT, FB, g6);
            visb += aeb * g6[0] + awb * g6[1] + asb * g6[2] + anb * g6[3] + bimb * g6[4] + cimb * g6[5];
            const float dvx = ml_visc_dx<COMP>(d, q, k, UK, VK);
            up = visb * dvx;
            dn = -visb * dvx;
        } else {
            if (q.vegp) {                                     // D -= VEG x wh AREA DZ
                const gf32 *UKS = d.uks + o, *VKS = d.vks + o, *UKS_W = d.uks + c3(d, i - 1, j);
                const gf32 *VKS_S = d.vks + c3(d, i, j - 1);
                const float in2 = 0.5f * (VKS_S[k] + VKS[k]), in = 0.5f * (UKS_W[k] + UKS[k]);
                const float wh = vmaxf(sqrtf(in * in + in2 * in2), 0.01f);
                xb -= Db * d.veg[o + k - 1] * wh * d.area * dz;
            }
            if (cross) {                                      // X1 = VIS/DZ (... + (X_{k+1} - x) AREA - (x - X_{k-1}) AREA)
                const float VIS = ml_visc(d, q, k, d.uk + o, d.vk + o);
                const float c = Db * (VIS / dz) * d.area;
                up = c; dn = c; xb -= 2.0f * c;
            }
        }
        sUp[k] = up; sDn[k] = dn; sMid[k] = xb;
        // the four neighbours' adjoints, handed on by direction
        m.gpa[o + k] = (IS == 1) ? xwb : xeb;
        m.gna[o + k] = (IS == 1) ? xeb : xwb;
        m.gpb[o + k] = (JS == 1) ? xsb : xnb_;
        m.gnb[o + k] = (JS == 1) ? xnb_ : xsb;
    }
    __syncthreads();
    // 4. own X before, levels KSTART-1 .. VI+1, summed in a fixed order
    for (int k = KSTART - 1 + lane; k <= VI + 1; k += WAVE) {
        float v = 0.0f;
        if (k >= KSTART && k <= VI) v = sR[k] + sMid[k];
        if (k == VI + 1) v = sR[VI + 1];
        if (k - 1 >= KSTART && k - 1 <= VI) v += sUp[k - 1];
        if (k + 1 >= KSTART && k + 1 <= VI) v += sDn[k + 1];
        m.gout[o + k] = v;
    }
}

// dL/dX before of every cell: its own column's part, the adjoints its
// predecessors handed to its old value, and, where the sweep left X as it was
// (columns it does not visit, levels outside KSTART..VI), dL/dX after with the
// adjoints its successors handed to its new value.  Thread per cell and level.
template <int COMP>
__global__ void k_adj_mom_combine(const Dom *__restrict__ ds, MomAdj m, int IS, int JS) {
    const DomG &d = *reinterpret_cast<const DomG *>(ds);
    const size_t n3 = (size_t)(d.nii + 2) * d.nj2 * d.nkp;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n3) return;
    const int k = (int)(g % d.nkp);
    const size_t c = g / d.nkp;
    const int i = (int)(c / d.nj2), j = (int)(c % d.nj2);
    const int oa = (COMP == 0) ? 3 : 2, ia = (COMP == 1) ? 3 : 2;
    auto inside = [&](int ii, int jj) { return ii >= 0 && ii <= d.nii + 1 && jj >= 0 && jj <= d.njj + 1; };
    float v = m.gout[g];
    if (inside(i - IS, j)) v += m.gna[c3(d, i - IS, j) + k];
    if (inside(i, j - JS)) v += m.gnb[c3(d, i, j - JS) + k];
    bool done = false;
    if (i >= 1 && i <= d.nii && j >= 1 && j <= d.njj && d.adv[c] == 1) {
        const int a = (IS == -1) ? d.nii - i + 1 : i, b = (JS == -1) ? d.njj - j + 1 : j;
        if (a >= oa && a <= d.nii - 1 && b >= ia && b <= d.njj - 1) {
            const int KK = d.kk[c];
            const int KSTART = (d.cutk[c] == 0) ? KK + 1 : 1;
            done = k >= KSTART && k <= d.vi[c];
        }
    }
    if (!done) {
        v += m.gin[g];
        if (inside(i + IS, j)) v += m.gpa[c3(d, i + IS, j) + k];
        if (inside(i, j + JS)) v += m.gpb[c3(d, i, j + JS) + k];
    }
    m.gres[g] = v;
}

__global__ void k_mom_fill(float *p, size_t n) {
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) p[g] = 0.0f;
}

// One sweep of momentum component comp (0 U, 1 V, 2 W) of the mixing-length
// model in the directions (IS, JS) the iteration gives it, one launch per
// anti-diagonal (bit for bit the strip sweep's result).
extern "C" int pf_launch_momentum_sweep(const Dom *hd, const Dom *dd, int comp, int IS, int JS, hipStream_t s) {
    const Dom &h = *hd;
    if (h.model != 1 || comp < 0 || comp > 2) return (int)hipErrorInvalidValue;
    const int oa = (comp == 0) ? 3 : 2, ia = (comp == 1) ? 3 : 2, ob = h.nii - 1, ib = h.njj - 1;
    const size_t shm = sizeof(float) * 5 * h.nkp;
    for (int t = 0; t < ndiag(oa, ob, ia, ib); ++t) {
        const unsigned len = (unsigned)diag_len(t, oa, ob, ia, ib);
        if (comp == 0) hipLaunchKernelGGL((k_momentum<0, 0>), dim3(len), dim3(WAVE), shm, s, dd, t, IS, JS);
        else if (comp == 1) hipLaunchKernelGGL((k_momentum<1, 0>), dim3(len), dim3(WAVE), shm, s, dd, t, IS, JS);
        else hipLaunchKernelGGL((k_momentum<2, 0>), dim3(len), dim3(WAVE), shm, s, dd, t, IS, JS);
    }
    return (int)hipGetLastError();
}

// The transpose of that sweep: dd is the situation's descriptor with the swept
// field pointing at X before the sweep.
extern "C" int pf_launch_momentum_vjp(const Dom *hd, const Dom *dd, const MomAdj *ma, int comp, int IS, int JS,
                                      hipStream_t s) {
    const Dom &h = *hd;
    if (h.model != 1 || comp < 0 || comp > 2) return (int)hipErrorInvalidValue;
    const MomAdj m = *ma;
    const size_t n3 = (size_t)(h.nii + 2) * h.nj2 * h.nkp;
    const unsigned fb = (unsigned)((n3 + 255) / 256);
    float *z[5] = {m.gout, m.gpa, m.gpb, m.gna, m.gnb};
    for (float *p : z) hipLaunchKernelGGL(k_mom_fill, dim3(fb), dim3(256), 0, s, p, n3);
    const int oa = (comp == 0) ? 3 : 2, ia = (comp == 1) ? 3 : 2, ob = h.nii - 1, ib = h.njj - 1;
    const size_t shm = sizeof(float) * MA_SLOTS * h.nkp;
    for (int t = ndiag(oa, ob, ia, ib) - 1; t >= 0; --t) {
        const unsigned len = (unsigned)diag_len(t, oa, ob, ia, ib);
        if (comp == 0) hipLaunchKernelGGL(k_adj_mom<0>, dim3(len), dim3(WAVE), shm, s, dd, m, t, IS, JS);
        else if (comp == 1) hipLaunchKernelGGL(k_adj_mom<1>, dim3(len), dim3(WAVE), shm, s, dd, m, t, IS, JS);
        else hipLaunchKernelGGL(k_adj_mom<2>, dim3(len), dim3(WAVE), shm, s, dd, m, t, IS, JS);
    }
    if (comp == 0) hipLaunchKernelGGL(k_adj_mom_combine<0>, dim3(fb), dim3(256), 0, s, dd, m, IS, JS);
    else if (comp == 1) hipLaunchKernelGGL(k_adj_mom_combine<1>, dim3(fb), dim3(256), 0, s, dd, m, IS, JS);
    else hipLaunchKernelGGL(k_adj_mom_combine<2>, dim3(fb), dim3(256), 0, s, dd, m, IS, JS);
    return (int)hipGetLastError();
}
