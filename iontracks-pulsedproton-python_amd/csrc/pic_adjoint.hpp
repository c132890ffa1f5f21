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
// column, so the ring's adjoint is folded into that column (the transpose of
// the copy: j copies first, then i copies, the reverse of state.py:148-154)
// before it is used. The kernel also accumulates the step's parameter
// sensitivities mu_{t+1} . dH/dw for the eight stencil weights and alpha_dt,
// and mu_t . dD_t/d(gf, h2b2) for the deposition.
#pragma once
#include "pic_kernels.hpp"

namespace pic {

constexpr int kAdjNV = 11;      // p_lat p_zm p_zp p_cen n_lat n_zm n_zp n_cen alpha_dt gf h2b2
constexpr int kAdjStride = 12;  // doubles per block partial

// Compensated fold of an [n][stride] table of block partials (first NV
// columns) over the block in a fixed order; thread 0 gets the NV totals.
template <int NV>
__device__ void block_fold(const double *part, int n, int stride, double out[NV]) {
    __shared__ double s_s[NV][16], s_c[NV][16];
    Comp q[NV];
    constexpr int U = 4;
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
    const double *dens;      // d_t, or null when the step deposits nothing
    const double *dgf, *dh;  // dd_t/dgf and dd_t/dh2b2 per column, or null
    const double *mp, *mn;   // mu_{t+1}
    double *op, *on;         // mu_t
    double *partials;        // [nblk][kAdjStride]
    // the previous launch's (step t+1's) partials, folded by block 0
    const double *prev_part;
    int prev_n;
    double *prev_out;  // [kAdjNV], or null
    int no_xy, nzb, nze, no_z, pitch;
    int n_kb, n_jb;  // 64-wide k blocks and 4-row j blocks per i plane
    double mid, rsq;
    double p_lat, p_zm, p_zp, p_cen, n_lat, n_zm, n_zp, n_cen, adt;
};

// One thread per voxel (i, j, k) of the whole grid, ring included: 64 k by 4
// j rows per block, one i plane per block row. Block 0 folds the previous
// launch's partials when prev_out is set. Neighbours come straight from
// memory (the grids this is for -- tens of columns a side -- stay in cache).
template <bool REFLECT>
__global__ void __launch_bounds__(256) adjoint_kernel(AdjointArgs a) {
    const int nf = a.prev_out ? 1 : 0;
    if ((int)blockIdx.x < nf) {
        double out[kAdjNV];
        block_fold<kAdjNV>(a.prev_part, a.prev_n, kAdjStride, out);
        if (threadIdx.x == 0)
            for (int c = 0; c < kAdjNV; ++c) a.prev_out[c] = out[c];
        return;
    }
    const int b = blockIdx.x - nf;
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
        auto at = [&](int ii, int jj, int kk) -> int64_t { return ((int64_t)ii * n + jj) * a.pitch + kk; };
        auto in_gap = [&](int kk) { return kk >= a.nze && kk < a.nze + a.no_z; };
        auto interior = [&](int ii, int jj, int kk) {
            return ii >= 1 && ii <= n - 2 && jj >= 1 && jj <= n - 2 && kk >= 1 && kk <= nzb - 2;
        };
        // X_t = S_t + d_t, the swept input (a reflecting ring reads its mirror
        // column and adds its own column's deposition, as the sweep does)
        auto X = [&](const double *S, int ii, int jj, int kk) -> double {
            const int si = REFLECT ? min(max(ii, 1), n - 2) : ii;
            const int sj = REFLECT ? min(max(jj, 1), n - 2) : jj;
            double v = S[at(si, sj, kk)];
            if (a.dens && in_gap(kk)) v = v + a.dens[(int64_t)ii * n + jj];
            return v;
        };
        // mu_{t+1} at interior (ii, jj, kk), the reflecting ring folded in
        auto A = [&](const double *M, int ii, int jj, int kk) -> double {
            if (!REFLECT) return M[at(ii, jj, kk)];
            auto row = [&](int ri) {
                double v = M[at(ri, jj, kk)];
                if (jj == 1) v += M[at(ri, 0, kk)];
                if (jj == n - 2) v += M[at(ri, n - 1, kk)];
                return v;
            };
            double v = row(ii);
            if (ii == 1) v += row(0);
            if (ii == n - 2) v += row(n - 1);
            return v;
        };
        const bool u_int = interior(i, j, k);
        double op = 0.0, on = 0.0;
        if (!REFLECT && !u_int) {  // carried ring and k end planes
            op = a.mp[at(i, j, k)];
            on = a.mn[at(i, j, k)];
        }
        // transposed stencil: v = u + offset must be interior
        if (interior(i, j, k + 1)) op += a.p_zm * A(a.mp, i, j, k + 1);
        if (interior(i, j, k - 1)) op += a.p_zp * A(a.mp, i, j, k - 1);
        if (interior(i, j + 1, k)) op += a.p_lat * A(a.mp, i, j + 1, k);
        if (interior(i, j - 1, k)) op += a.p_lat * A(a.mp, i, j - 1, k);
        if (interior(i + 1, j, k)) op += a.p_lat * A(a.mp, i + 1, j, k);
        if (interior(i - 1, j, k)) op += a.p_lat * A(a.mp, i - 1, j, k);
        if (interior(i, j, k - 1)) on += a.n_zm * A(a.mn, i, j, k - 1);
        if (interior(i, j, k + 1)) on += a.n_zp * A(a.mn, i, j, k + 1);
        if (interior(i, j + 1, k)) on += a.n_lat * A(a.mn, i, j + 1, k);
        if (interior(i, j - 1, k)) on += a.n_lat * A(a.mn, i, j - 1, k);
        if (interior(i + 1, j, k)) on += a.n_lat * A(a.mn, i + 1, j, k);
        if (interior(i - 1, j, k)) on += a.n_lat * A(a.mn, i - 1, j, k);
        if (u_int) {
            const double ap = A(a.mp, i, j, k), an = A(a.mn, i, j, k);
            const double xp = X(a.sp, i, j, k), xn = X(a.sn, i, j, k);
            const double s = ap + an;
            op += a.p_cen * ap;
            on += a.n_cen * an;
            op -= a.adt * xn * s;
            on -= a.adt * xp * s;
            const double di = (double)i - a.mid, dj = (double)j - a.mid;
            const bool scored = in_gap(k) && di * di + dj * dj < a.rsq;
            if (scored) {
                op += a.adt * xn;
                on += a.adt * xp;
            }
            // sensitivities of H (and r) at this interior voxel
            g[0] = ap * (((X(a.sp, i, j - 1, k) + X(a.sp, i, j + 1, k)) + X(a.sp, i - 1, j, k)) + X(a.sp, i + 1, j, k));
            g[1] = ap * X(a.sp, i, j, k - 1);
            g[2] = ap * X(a.sp, i, j, k + 1);
            g[3] = ap * xp;
            g[4] = an * (((X(a.sn, i, j - 1, k) + X(a.sn, i, j + 1, k)) + X(a.sn, i - 1, j, k)) + X(a.sn, i + 1, j, k));
            g[5] = an * X(a.sn, i, j, k + 1);
            g[6] = an * X(a.sn, i, j, k - 1);
            g[7] = an * xn;
            const double pn = xp * xn;
            g[8] = (scored ? pn : 0.0) - s * pn;
        }
        if (a.dgf && in_gap(k)) {
            const int64_t col = (int64_t)i * n + j;
            g[9] = (op + on) * a.dgf[col];
            g[10] = (op + on) * a.dh[col];
        }
        a.op[at(i, j, k)] = op;
        a.on[at(i, j, k)] = on;
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
        a.partials[(int64_t)b * kAdjStride + c] = ((s_g[c][0] + s_g[c][1]) + s_g[c][2]) + s_g[c][3];
    }
}

// The derivative variant of the deposition gather (dd/dgf, dd/dh2b2).
__global__ void __launch_bounds__(256) gather_deriv_kernel(GatherArgs a) { gather_body<true>(a, blockIdx.x, blockIdx.y); }

}  // namespace pic
