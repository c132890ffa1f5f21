// Reverse-mode (adjoint) kernels: the derivative of the scored recombined
// charge with respect to the parameters of one time step. Included once, by
// pic_hip.hip, after pic_kernels.hpp. FP64 fields only.
//
// With S_t the state at the start of step t, X_t = S_t + D_t the swept input
// (D_t the step's deposition on the gap layers) and S_{t+1} = H(X_t) the
// sweep plus the wall, the adjoint state mu_t = dJ/dX_t of
// J = sum_t alpha_dt sum_scored p n obeys
//
//     mu_t = grad r(X_t) + (dH/dX_t)^T mu_{t+1},        mu_T = 0.
//
// (dH/dX)^T is the stencil with its z weights exchanged (a voxel feeds its
// k+1 neighbour with the z_minus weight, so it receives that neighbour's
// adjoint with the same weight), restricted to interior targets, plus the
// bilinear recombination term; the absorbing ring and the k end planes are
// carried unchanged by H, so their adjoint is carried back; under a
// reflecting wall the ring of S_{t+1} is a copy of its nearest interior
// column (state.py:148-154), so the ring's adjoint belongs to that column:
// it is folded in as it is computed (the transpose of the copy), and the
// stored adjoint is the folded one. The kernel also accumulates the step's parameter
// sensitivities mu_{t+1} . dH/dw for the eight stencil weights and alpha_dt,
// and mu_t . dD_t/d(gf, h2b2) for the deposition.
#pragma once
#include "pic_kernels.hpp"

namespace pic {

constexpr int kAdjNV = 11;  // p_lat p_zm p_zp p_cen n_lat n_zm n_zp n_cen alpha_dt gf h2b2

// Compensated fold of an [n][stride] table of block partials (first NV
// columns) over the block in a fixed order; thread 0 gets the NV totals.
template <int NV>
__device__ void block_fold(const double *part, int n, int stride, double out[NV]) {
    __shared__ double s_s[NV][16], s_c[NV][16];
    Comp q[NV];
    constexpr int U = 2;  // NV x U loads in flight per thread; more costs registers in the kernel that inlines this
    const int B = blockDim.x;
    for (int b0 = threadIdx.x; b0 < n; b0 += U * B) {
        double v[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t b = min(b0 + u * B, n - 1);
#pragma unroll
            for (int c = 0; c < NV; ++c) v[u][c] = part[b * stride + c];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (b0 + u * B < n)
#pragma unroll
                for (int c = 0; c < NV; ++c) q[c].add(v[u][c]);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        q[c] = wave_comp(q[c]);
        if (lane == 0) {
            s_s[c][wave] = q[c].s;
            s_c[c][wave] = q[c].c;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + kWave - 1) / kWave;
        for (int c = 0; c < NV; ++c) {
            Comp t{s_s[c][0], s_c[c][0]};
            for (int w = 1; w < nw; ++w) t.add(Comp{s_s[c][w], s_c[c][w]});
            out[c] = t.value();
        }
    }
}

// One block per row: fold rows [0, gridDim.x) of a table of per-row block
// partials ([row][n][stride]) into out[row][NV].
template <int NV>
__global__ void __launch_bounds__(256) fold_rows_kernel(const double *part, int n, int stride, double *out) {
    double v[NV];
    block_fold<NV>(part + (int64_t)blockIdx.x * n * stride, n, stride, v);
    if (threadIdx.x == 0)
        for (int c = 0; c < NV; ++c) out[(int64_t)blockIdx.x * NV + c] = v[c];
}

struct AdjointArgs {
    const double *sp, *sn;   // S_t (under a reflecting wall the stored ring is not read)
    const double *dens;      // d_t (zeros when the step deposits nothing)
    const double *dgf, *dh;  // dd_t/dgf and dd_t/dh2b2 per column (zeros when the step deposits nothing)
    const double *mp, *mn;   // mu_{t+1}
    double *op, *on;         // mu_t
    double *acc;             // [nblk][2 * kAdjNV]: each block's running (sum, compensation) over the steps
    int no_xy, nzb, nze, no_z, pitch;
    int n_kb, n_jb;  // 64-wide k blocks and 4-row j blocks per i plane
    double mid, rsq;
    double p_lat, p_zm, p_zp, p_cen, n_lat, n_zm, n_zp, n_cen, adt;
};

// One thread per voxel (i, j, k) of the whole grid, ring included: 64 k by 4
// j rows per block, one i plane per block row. Neighbours come straight from
// memory (the grids this is for -- tens of columns a side -- stay in cache),
// all loads issued before the first use. Only the run's totals of the
// sensitivities are wanted, so each block adds its step sums to its own
// compensated running sum (the same block, the same slot, in step order:
// deterministic and with no dependence between blocks); the blocks' sums are
// folded once, after the last step, by fold_accumulators_kernel.
template <bool REFLECT>
__global__ void __launch_bounds__(256) adjoint_kernel(AdjointArgs a) {
    const int b = blockIdx.x;
    const int kb = b % a.n_kb;
    const int rest = b / a.n_kb;
    const int jb = rest % a.n_jb;
    const int i = rest / a.n_jb;
    const int k = kb * kWave + (threadIdx.x & 63);
    const int j = jb * 4 + (threadIdx.x >> 6);
    const int n = a.no_xy, nzb = a.nzb;
    double g[kAdjNV];
#pragma unroll
    for (int c = 0; c < kAdjNV; ++c) g[c] = 0.0;

    if (k < nzb && j < n) {
        // Every load below is unconditional (indices clamped into the grid) and
        // every contribution is selected by its predicate, so the compiler can
        // keep all of a thread's loads in flight at once; a skipped term is an
        // added +0.0, which leaves each sum bit-identical.
        auto at = [&](int ii, int jj, int kk) -> int64_t {
            ii = min(max(ii, 0), n - 1);
            jj = min(max(jj, 0), n - 1);
            kk = min(max(kk, 0), nzb - 1);
            return ((int64_t)ii * n + jj) * a.pitch + kk;
        };
        auto in_gap = [&](int kk) { return kk >= a.nze && kk < a.nze + a.no_z; };
        auto interior = [&](int ii, int jj, int kk) {
            return ii >= 1 && ii <= n - 2 && jj >= 1 && jj <= n - 2 && kk >= 1 && kk <= nzb - 2;
        };
        // X_t = S_t + d_t, the swept input (a reflecting ring reads its mirror
        // column and adds its own column's deposition, as the sweep does)
        auto X = [&](const double *S, int ii, int jj, int kk) -> double {
            const int si = REFLECT ? min(max(ii, 1), n - 2) : ii;
            const int sj = REFLECT ? min(max(jj, 1), n - 2) : jj;
            const double v = S[at(si, sj, kk)];
            const double d = a.dens[(int64_t)min(max(ii, 0), n - 1) * n + min(max(jj, 0), n - 1)];
            return v + (in_gap(kk) ? d : 0.0);
        };
        // mu_{t+1} at interior (ii, jj, kk); under a reflecting wall it is
        // stored with the ring already folded in (below)
        auto A = [&](const double *M, int ii, int jj, int kk) -> double { return M[at(ii, jj, kk)]; };
        const bool u_int = interior(i, j, k);
        double op = 0.0, on = 0.0;
        if (!REFLECT) {  // carried ring and k end planes
            const double cp = a.mp[at(i, j, k)], cn = a.mn[at(i, j, k)];
            op = u_int ? 0.0 : cp;
            on = u_int ? 0.0 : cn;
        }
        // transposed stencil: v = u + offset must be interior
        const bool zp1 = interior(i, j, k + 1), zm1 = interior(i, j, k - 1);
        const bool jp1 = interior(i, j + 1, k), jm1 = interior(i, j - 1, k);
        const bool ip1 = interior(i + 1, j, k), im1 = interior(i - 1, j, k);
        const double ap_zp = A(a.mp, i, j, k + 1), ap_zm = A(a.mp, i, j, k - 1);
        const double ap_jp = A(a.mp, i, j + 1, k), ap_jm = A(a.mp, i, j - 1, k);
        const double ap_ip = A(a.mp, i + 1, j, k), ap_im = A(a.mp, i - 1, j, k);
        const double an_zp = A(a.mn, i, j, k + 1), an_zm = A(a.mn, i, j, k - 1);
        const double an_jp = A(a.mn, i, j + 1, k), an_jm = A(a.mn, i, j - 1, k);
        const double an_ip = A(a.mn, i + 1, j, k), an_im = A(a.mn, i - 1, j, k);
        const double ap = A(a.mp, i, j, k), an = A(a.mn, i, j, k);
        const double xp = X(a.sp, i, j, k), xn = X(a.sn, i, j, k);
        const double xp_jm = X(a.sp, i, j - 1, k), xp_jp = X(a.sp, i, j + 1, k);
        const double xp_im = X(a.sp, i - 1, j, k), xp_ip = X(a.sp, i + 1, j, k);
        const double xp_km = X(a.sp, i, j, k - 1), xp_kp = X(a.sp, i, j, k + 1);
        const double xn_jm = X(a.sn, i, j - 1, k), xn_jp = X(a.sn, i, j + 1, k);
        const double xn_im = X(a.sn, i - 1, j, k), xn_ip = X(a.sn, i + 1, j, k);
        const double xn_km = X(a.sn, i, j, k - 1), xn_kp = X(a.sn, i, j, k + 1);
        const int64_t col = (int64_t)i * n + j;
        const double dgf = a.dgf[col], dh = a.dh[col];
        op += zp1 ? a.p_zm * ap_zp : 0.0;
        op += zm1 ? a.p_zp * ap_zm : 0.0;
        op += jp1 ? a.p_lat * ap_jp : 0.0;
        op += jm1 ? a.p_lat * ap_jm : 0.0;
        op += ip1 ? a.p_lat * ap_ip : 0.0;
        op += im1 ? a.p_lat * ap_im : 0.0;
        on += zm1 ? a.n_zm * an_zm : 0.0;
        on += zp1 ? a.n_zp * an_zp : 0.0;
        on += jp1 ? a.n_lat * an_jp : 0.0;
        on += jm1 ? a.n_lat * an_jm : 0.0;
        on += ip1 ? a.n_lat * an_ip : 0.0;
        on += im1 ? a.n_lat * an_im : 0.0;
        if (u_int) {
            const double s = ap + an;
            op += a.p_cen * ap;
            on += a.n_cen * an;
            op -= a.adt * xn * s;
            on -= a.adt * xp * s;
            const double di = (double)i - a.mid, dj = (double)j - a.mid;
            const bool scored = in_gap(k) && di * di + dj * dj < a.rsq;
            op += scored ? a.adt * xn : 0.0;
            on += scored ? a.adt * xp : 0.0;
            // sensitivities of H (and r) at this interior voxel
            g[0] = ap * (((xp_jm + xp_jp) + xp_im) + xp_ip);
            g[1] = ap * xp_km;
            g[2] = ap * xp_kp;
            g[3] = ap * xp;
            g[4] = an * (((xn_jm + xn_jp) + xn_im) + xn_ip);
            g[5] = an * xn_kp;
            g[6] = an * xn_km;
            g[7] = an * xn;
            const double pn = xp * xn;
            g[8] = (scored ? pn : 0.0) - s * pn;
        }
        if (in_gap(k) && (!REFLECT || u_int)) {  // (a reflecting ring's terms are taken by its partner)
            g[9] = (op + on) * dgf;
            g[10] = (op + on) * dh;
        }
        if constexpr (REFLECT) {
            // A ring voxel next to this one reads this column (the ring is a
            // copy of its neighbour), and its only interior consumer is this
            // voxel, through the lateral weight: its adjoint is lat * mu(u).
            // Its deposition term is taken here, with the ring column's
            // dd/dgf and dd/dh2b2, and its adjoint is folded into this voxel's
            // -- the transposed copy -- so the ring itself is never stored.
            if (u_int) {
                const double rp = a.p_lat * ap, rn = a.n_lat * an;
                const int pi[4] = {0, n - 1, i, i}, pj[4] = {j, j, 0, n - 1};
                const bool has[4] = {i == 1, i == n - 2, j == 1, j == n - 2};
                double fp = op, fn = on;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int64_t rcol = (int64_t)pi[q] * n + pj[q];
                    const double rgf = a.dgf[rcol], rh = a.dh[rcol];
                    fp += has[q] ? rp : 0.0;
                    fn += has[q] ? rn : 0.0;
                    g[9] += (has[q] && in_gap(k)) ? (rp + rn) * rgf : 0.0;
                    g[10] += (has[q] && in_gap(k)) ? (rp + rn) * rh : 0.0;
                }
                a.op[at(i, j, k)] = fp;
                a.on[at(i, j, k)] = fn;
            }
        } else {
            a.op[at(i, j, k)] = op;
            a.on[at(i, j, k)] = on;
        }
    }
    // block sums of the eleven sensitivities, fixed order
    __shared__ double s_g[kAdjNV][4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < kAdjNV; ++c) {
        const double v = wave_sum(g[c]);
        if (lane == 0) s_g[c][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < kAdjNV) {
        const int c = threadIdx.x;
        double *slot = a.acc + (int64_t)b * 2 * kAdjNV + 2 * c;
        Comp q{slot[0], slot[1]};
        q.add(((s_g[c][0] + s_g[c][1]) + s_g[c][2]) + s_g[c][3]);
        slot[0] = q.s;
        slot[1] = q.c;
    }
}

// One block: thread c < NV folds column c of nblk (sum, compensation) pairs,
// block by block in index order, into out[c].
template <int NV>
__global__ void __launch_bounds__(64) fold_accumulators_kernel(const double *acc, int nblk, double *out) {
    const int c = threadIdx.x;
    if (c >= NV) return;
    Comp t;
    for (int b = 0; b < nblk; ++b) t.add(Comp{acc[(int64_t)b * 2 * NV + 2 * c], acc[(int64_t)b * 2 * NV + 2 * c + 1]});
    out[c] = t.value();
}

// The derivative variant of the deposition gather (dd/dgf, dd/dh2b2).
__global__ void __launch_bounds__(256) gather_deriv_kernel(GatherArgs a) { gather_body<true>(a, blockIdx.x, blockIdx.y); }

}  // namespace pic
