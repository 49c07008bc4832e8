// Adjoint (vector-Jacobian product) of the pressure-correction part of the
// SIMPLE iteration: divergence (PrognosticFlowfield.cs:723-758), the pressure
// GS-TDMA sweep (:770-901) and the velocity correction (:904-949), without the
// FP64 quantisation.  With the geometry, DTIME and the sweep directions fixed,
// one such step is linear in (UK, VK, WK):
//
//   DIV = D u,   DPM = S DIV,   u' = u + G DPM      (S: the Gauss-Seidel sweep)
//
// so its VJP is  u_bar = g + D^T S^T G^T g  (with the correction's masks), and
// an unrolled window of n steps is transposed step by step in reverse.  The
// sweep S is a block lower-triangular solve in sweep order: each column solves
// its tridiagonal system (Thomas, top-down elimination, bottom-up
// substitution) with a right-hand side that reads the new DPM of the two
// columns visited before it, and its elimination starts from the Q value the
// last earlier column of the same sweep row left at level VI+1 (the row
// scratch of the reference, reused across the row).  S^T is therefore a
// block upper-triangular solve in reverse sweep order: anti-diagonals from
// the last to the first, each column running the transposed Thomas
// recurrences (substitution transposed top-down, elimination transposed
// bottom-up), handing DIM_bar to the neighbours it read and the adjoint of its
// top Q value to the column that wrote it.  Every accumulation is a gather or
// a single writer, so the result is deterministic.
#include "pf_device.h"

#define AWAVE 64

__device__ __forceinline__ bool adj_diag_col(int t, int oa, int ob, int ia, int ib, int idx, int &a, int &b) {
    int nb = ib - ia;
    int amin = max(oa, oa + t - nb);
    int amax = min(ob, oa + t);
    a = amin + idx;
    if (a > amax) return false;
    b = ia + (t - (a - oa));
    return true;
}

__global__ void k_adj_fill(float *p, size_t n) {
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) p[g] = 0.0f;
}

// G^T: the adjoint of DPM from the adjoint velocities, for the cells the
// correction reads DPM of (gathered from the faces each cell's DPM enters),
// and W_bar = 0 where the correction sets WK = 0 (k <= KKART+1).  One block
// per column (i, j) of 2..NII-1 x 2..NJJ-1, threads over levels.
__device__ __forceinline__ bool corr_cell(const DomG &d, int i, int j) {
    if (i < 2 || i > d.nii - 1 || j < 2 || j > d.njj - 1) return false;
    return d.adv[c2(d, i, j)] == 1 || d.adv[c2(d, i - 1, j)] == 1 || d.adv[c2(d, i, j - 1)] == 1;
}
__global__ void k_adj_correction(const Dom *__restrict__ ds, AdjArgs a) {
    const DomG &d = *reinterpret_cast<const DomG *>(ds);
    const int nj = d.njj - 2;
    const int i = 2 + (int)blockIdx.x / nj, j = 2 + (int)blockIdx.x % nj;
    const size_t c = c2(d, i, j);
    const size_t o = c3(d, i, j), oip = c3(d, i + 1, j), ojp = c3(d, i, j + 1);
    const int KK = d.kk[c], VI = d.vi[c];
    const bool el = corr_cell(d, i, j), eli = corr_cell(d, i + 1, j), elj = corr_cell(d, i, j + 1);
    const int KKim = d.kk[c2(d, i - 1, j)], KKjm = d.kk[c2(d, i, j - 1)];
    const int KKip = d.kk[c2(d, i + 1, j)], VIip = d.vi[c2(d, i + 1, j)];
    const int KKjp = d.kk[c2(d, i, j + 1)], VIjp = d.vi[c2(d, i, j + 1)];
    const bool active = d.adv[c] == 1;
    const float ux = d.dtime / d.dxk, vy = d.dtime / d.dyk;
    for (int k = 1 + threadIdx.x; k <= d.kadvmax; k += blockDim.x) {
        if (active && k > KK && k <= VI) {
            float g = 0;
            if (el && i > 2 && KK < k && KKim < k) g -= a.ub[o + k] * ux;           // UK[i]   = .. - DPM[i]
            if (eli && KKip < k && KK < k && k <= VIip) g += a.ub[oip + k] * ux;    // UK[i+1] = .. + DPM[i]
            if (el && j > 2 && KK < k && KKjm < k) g -= a.vb[o + k] * vy;
            if (elj && KKjp < k && KK < k && k <= VIjp) g += a.vb[ojp + k] * vy;
            if (el && KK + 1 < k) g -= a.wb[o + k] * (2.0f * d.dtime / (d.dzk[k] + d.dzk[k - 1]));
            if (el && KK + 1 < k + 1 && k + 1 <= VI) g += a.wb[o + k + 1] * (2.0f * d.dtime / (d.dzk[k + 1] + d.dzk[k]));
            a.pb[o + k] = g;
        }
    }
    __syncthreads();
    if (el)
        for (int k = 1 + threadIdx.x; k <= min(KK + 1, VI); k += blockDim.x) a.wb[o + k] = 0.0f;
}

// S^T on anti-diagonal t of the sweep (a = j1 outer, b = i1 inner, mirrored by
// IS/JS as in the forward), launched from the last diagonal to the first.
// One wave per column: the lanes stage the column's vectors and its P, TERMP
// and ATP in LDS, lane 0 runs the transposed recurrences over LDS in blocks of
// AU levels whose operands are read before the block's chain.
#define AU 8
__global__ __launch_bounds__(AWAVE) void k_adj_pressure(const Dom *__restrict__ ds, AdjArgs a, int t, int IS, int JS) {
    const DomG &d = *reinterpret_cast<const DomG *>(ds);
    extern __shared__ float sm[];
    int aa, bb;
    if (!adj_diag_col(t, 2, d.njj - 1, 2, d.nii - 1, blockIdx.x, aa, bb)) return;
    const int j = (JS == -1) ? d.njj - aa + 1 : aa;
    const int i = (IS == -1) ? d.nii - bb + 1 : bb;
    const size_t c = c2(d, i, j);
    if (d.adv[c] != 1) return;
    const int nkp = d.nkp, tid = threadIdx.x;
    float *sx = sm, *sq = sx + nkp, *sm_ = sq + nkp, *sp = sm_ + nkp, *st = sp + nkp, *sat = st + nkp;
    const int KK = d.kk[c], VI = d.vi[c];
    const size_t o = c3(d, i, j), onb_i = c3(d, i + IS, j), onb_j = c3(d, i, j + JS);
    // the two columns visited after this one read its new DPM in their DIM
    const float sgn_j = (JS == 1) ? 1.0f : -1.0f;
    for (int k = KK + 1 + tid; k <= VI; k += AWAVE) {
        const float dz = d.dzk[k];
        sx[k] = a.pb[o + k] + a.mb[onb_i + k] * (d.dyk * dz / d.dxk) + sgn_j * a.mb[onb_j + k] * (d.dxk * dz / d.dyk);
        sq[k] = a.qb[o + k];
        sp[k] = a.ptab[o + k];
        st[k] = a.ttab[o + k];
        sat[k] = d.atp[k] * a.ttab[o + k];
    }
    __syncthreads();
    if (tid == 0 && VI > KK) {
        // x_k = P_k x_{k-1} + Q_k (k = KK+1 .. VI), transposed top-down:
        // Q_bar[k] += x_bar[k]; x_bar[k-1] += P_k x_bar[k]
        float xk = sx[VI];
        int k = VI;
        for (; k - AU + 1 > KK; k -= AU) {
            float cx[AU], cp[AU], cq[AU];
#pragma unroll
            for (int u = 0; u < AU; ++u) { cp[u] = sp[k - u]; cq[u] = sq[k - u]; cx[u] = sx[k - u - 1]; }
#pragma unroll
            for (int u = 0; u < AU; ++u) {
                sq[k - u] = cq[u] + xk;
                if (k - u - 1 > KK) xk = cx[u] + cp[u] * xk;
            }
        }
        for (; k > KK; --k) {
            sq[k] = sq[k] + xk;
            if (k - 1 > KK) xk = sx[k - 1] + sp[k] * xk;
        }
        // Q_k = (ATP_k Q_{k+1} + DIM_k) TERMP_k (k = VI .. KK+1), transposed bottom-up
        float carry = 0.0f;
        k = KK + 1;
        for (; k + AU - 1 <= VI; k += AU) {
            float cq[AU], ct[AU], ca[AU];
#pragma unroll
            for (int u = 0; u < AU; ++u) { cq[u] = sq[k + u]; ct[u] = st[k + u]; ca[u] = sat[k + u]; }
#pragma unroll
            for (int u = 0; u < AU; ++u) {
                const float qt = cq[u] + carry;
                sm_[k + u] = ct[u] * qt;
                carry = ca[u] * qt;
            }
        }
        for (; k <= VI; ++k) {
            const float qt = sq[k] + carry;
            sm_[k] = st[k] * qt;
            carry = sat[k] * qt;
        }
        const int s = a.src[c];
        if (s >= 0) a.qb[c3(d, s, j) + VI + 1] += carry;   // the stale top's writer (same row, earlier)
    }
    __syncthreads();
    for (int k = KK + 1 + tid; k <= VI; k += AWAVE) {
        a.mb[o + k] = sm_[k];
        a.db[o + k] = sm_[k] / d.dtime;                    // DIM = DIV / DTIME + ...
    }
}

// D^T: the adjoint velocities gain the divergence adjoint of the cells whose
// DIV each face enters.  One block per column (i, j) of 2..NII x 2..NJJ.
__global__ void k_adj_divergence(const Dom *__restrict__ ds, AdjArgs a) {
    const DomG &d = *reinterpret_cast<const DomG *>(ds);
    const int nj = d.njj - 1;
    const int i = 2 + (int)blockIdx.x / nj, j = 2 + (int)blockIdx.x % nj;
    const size_t o = c3(d, i, j), oim = c3(d, i - 1, j), ojm = c3(d, i, j - 1);
    const int KK = d.kk[c2(d, i, j)];
    for (int k = 1 + threadIdx.x; k <= d.kadvmax; k += blockDim.x) {
        const float dz = d.dzk[k];
        if (j <= d.njj - 1) a.ub[o + k] += d.dyk * dz * (a.db[o + k] - a.db[oim + k]);
        if (i <= d.nii - 1) a.vb[o + k] += d.dxk * dz * (a.db[o + k] - a.db[ojm + k]);
        if (i <= d.nii - 1 && j <= d.njj - 1)
            a.wb[o + k] += d.area * ((k > KK + 1 ? a.db[o + k] : 0.0f) - a.db[o + k - 1]);
    }
}

// One step of the VJP, for forward iteration iter (its sweep directions).
// n3: floats of one [n2][nkp] field.
extern "C" int pf_launch_adjoint_step(const Dom *hd, const Dom *dd, const AdjArgs *args, int iter, size_t n3,
                                      hipStream_t s) {
    const Dom &h = *hd;
    int IS = 1, JS = 1;
    switch (iter % 4) {
        case 0: IS = -1; JS = 1; break;
        case 1: IS = -1; JS = -1; break;
        case 2: IS = 1; JS = -1; break;
        default: break;
    }
    AdjArgs a = *args;
    const unsigned fb = (unsigned)((n3 + 255) / 256);
    hipLaunchKernelGGL(k_adj_fill, dim3(fb), dim3(256), 0, s, a.pb, n3);
    hipLaunchKernelGGL(k_adj_fill, dim3(fb), dim3(256), 0, s, a.qb, n3);
    hipLaunchKernelGGL(k_adj_fill, dim3(fb), dim3(256), 0, s, a.mb, n3);
    hipLaunchKernelGGL(k_adj_fill, dim3(fb), dim3(256), 0, s, a.db, n3);
    hipLaunchKernelGGL(k_adj_correction, dim3((unsigned)((h.nii - 2) * (h.njj - 2))), dim3(64), 0, s, dd, a);
    const int oa = 2, ob = h.njj - 1, ia = 2, ib = h.nii - 1;
    const int nd = (ob - oa) + (ib - ia) + 1;
    const size_t shm = sizeof(float) * 6 * (size_t)h.nkp;
    for (int t = nd - 1; t >= 0; --t) {
        const int amin = max(oa, oa + t - (ib - ia)), amax = min(ob, oa + t);
        if (amax >= amin)
            hipLaunchKernelGGL(k_adj_pressure, dim3((unsigned)(amax - amin + 1)), dim3(AWAVE), shm, s, dd, a, t, IS, JS);
    }
    hipLaunchKernelGGL(k_adj_divergence, dim3((unsigned)((h.nii - 1) * (h.njj - 1))), dim3(64), 0, s, dd, a);
    return (int)hipGetLastError();
}
