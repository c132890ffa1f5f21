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
    int n_kb, n_jb;  // 64-wide k blocks and 4-row j blocks per i chunk
    int chunk;       // i planes each block marches through
    double mid, rsq;
    double p_lat, p_zm, p_zp, p_cen, n_lat, n_zm, n_zp, n_cen, adt;
};

// One thread per (j, k) column of the whole grid, ring included (64 k by 4 j
// per block), marching through a chunk of i planes, as the forward sweep
// does: the thread's own column at planes i-1, i, i+1 (mu_{t+1} and X_t)
// stays in registers and moves down the march; the j and k neighbours of
// plane i are loaded a plane ahead, as is the own column of plane i+2, so a
// step waits only for loads issued one step earlier. Every load is
// unconditional (indices clamped into the grid) and every contribution is
// selected by its predicate (a skipped term is an added +0.0), so each
// voxel's value is the one of the per-voxel form term for term.
// Only the run's totals of the sensitivities are wanted: each thread sums
// its chunk's voxels, each block adds its step sum to its own compensated
// running sum (the same block, the same slot, in step order: deterministic
// and with no dependence between blocks), and fold_accumulators_kernel folds
// the blocks' sums once, after the last step.
template <bool REFLECT>
__global__ void __launch_bounds__(256) adjoint_kernel(AdjointArgs a) {
    const int b = blockIdx.x;
    const int kb = b % a.n_kb;
    const int rest = b / a.n_kb;
    const int jb = rest % a.n_jb;
    const int ic = rest / a.n_jb;
    const int k = kb * kWave + (threadIdx.x & 63);
    // the row index is uniform across a wave; saying so keeps the row parts of
    // the addresses in scalar registers
    const int j = jb * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int n = a.no_xy, nzb = a.nzb;
    double g[kAdjNV];
#pragma unroll
    for (int c = 0; c < kAdjNV; ++c) g[c] = 0.0;

    if (k < nzb && j < n) {
        const int i0 = ic * a.chunk, i1 = min(i0 + a.chunk, n);
        const int64_t rowp = a.pitch, plane = (int64_t)n * a.pitch;
        auto clampi = [](int x, int lo, int hi) { return min(max(x, lo), hi); };
        auto in_gap = [&](int kk) { return kk >= a.nze && kk < a.nze + a.no_z; };
        auto interior = [&](int ii, int jj, int kk) {
            return ii >= 1 && ii <= n - 2 && jj >= 1 && jj <= n - 2 && kk >= 1 && kk <= nzb - 2;
        };
        // the stored column of (ii, jj) (clamped into the grid) and the column
        // a swept input is read from: under a reflecting wall a ring column
        // reads its mirror, as the sweep does
        auto acol = [&](int ii, int jj) -> int64_t {
            return (int64_t)clampi(ii, 0, n - 1) * plane + (int64_t)clampi(jj, 0, n - 1) * rowp;
        };
        auto scol = [&](int ii, int jj) -> int64_t {
            return REFLECT ? (int64_t)clampi(ii, 1, n - 2) * plane + (int64_t)clampi(jj, 1, n - 2) * rowp : acol(ii, jj);
        };
        auto xycol = [&](int ii, int jj) -> int64_t { return (int64_t)clampi(ii, 0, n - 1) * n + clampi(jj, 0, n - 1); };
        const int km = max(k - 1, 0), kp = min(k + 1, nzb - 1);
        // The four state arrays through buffer resources: the (wave-uniform)
        // column offset is the scalar offset and the lane's k a 32-bit vector
        // offset, so no load spends vector instructions on a 64-bit address
        // (the host keeps an array below 2 GB).
        const int arr_bytes = (int)((int64_t)n * n * a.pitch * 8);
        auto rsrc = [&](const double *q) {
            return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(q), 0, arr_bytes, kBufferWord3);
        };
        const auto r_mp = rsrc(a.mp), r_mn = rsrc(a.mn), r_sp = rsrc(a.sp), r_sn = rsrc(a.sn);
        auto ld = [](__amdgpu_buffer_rsrc_t r, int64_t col, int kk) -> double {
            return __builtin_bit_cast(double,
                                      __builtin_amdgcn_raw_buffer_load_b64(r, (unsigned)kk * 8u, (int)(col * 8), 0));
        };
        const bool gap_k = in_gap(k), gap_km = in_gap(k - 1), gap_kp = in_gap(k + 1);

        // the own column at one plane: mu_{t+1}, the stored S_t and the density
        struct Col {
            double ap, an, sp, sn, d;
        };
        auto load_col = [&](int ii) {
            Col c;
            const int64_t ao = acol(ii, j), so = scol(ii, j);
            c.ap = ld(r_mp, ao, k);
            c.an = ld(r_mn, ao, k);
            c.sp = ld(r_sp, so, k);
            c.sn = ld(r_sn, so, k);
            c.d = a.dens[xycol(ii, j)];
            return c;
        };
        // the j and k neighbours at one plane, the densities of the j
        // neighbour columns, the column's deposition derivatives and (reflecting
        // wall) those of its four possible ring partners
        struct Nb {
            double ap_jm, ap_jp, ap_km, ap_kp, an_jm, an_jp, an_km, an_kp;
            double sp_jm, sp_jp, sp_km, sp_kp, sn_jm, sn_jp, sn_km, sn_kp;
            double d_jm, d_jp, dgf, dh;
            double rgf[REFLECT ? 4 : 1], rh[REFLECT ? 4 : 1];
        };
        auto load_nb = [&](int ii) {
            Nb q;
            const int64_t a_jm = acol(ii, j - 1), a_jp = acol(ii, j + 1), a0 = acol(ii, j);
            const int64_t s_jm = scol(ii, j - 1), s_jp = scol(ii, j + 1), s0 = scol(ii, j);
            q.ap_jm = ld(r_mp, a_jm, k);
            q.ap_jp = ld(r_mp, a_jp, k);
            q.ap_km = ld(r_mp, a0, km);
            q.ap_kp = ld(r_mp, a0, kp);
            q.an_jm = ld(r_mn, a_jm, k);
            q.an_jp = ld(r_mn, a_jp, k);
            q.an_km = ld(r_mn, a0, km);
            q.an_kp = ld(r_mn, a0, kp);
            q.sp_jm = ld(r_sp, s_jm, k);
            q.sp_jp = ld(r_sp, s_jp, k);
            q.sp_km = ld(r_sp, s0, km);
            q.sp_kp = ld(r_sp, s0, kp);
            q.sn_jm = ld(r_sn, s_jm, k);
            q.sn_jp = ld(r_sn, s_jp, k);
            q.sn_km = ld(r_sn, s0, km);
            q.sn_kp = ld(r_sn, s0, kp);
            q.d_jm = a.dens[xycol(ii, j - 1)];
            q.d_jp = a.dens[xycol(ii, j + 1)];
            const int64_t col = xycol(ii, j);
            q.dgf = a.dgf[col];
            q.dh = a.dh[col];
            if constexpr (REFLECT) {
                const int64_t rc[4] = {xycol(0, j), xycol(n - 1, j), xycol(ii, 0), xycol(ii, n - 1)};
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    q.rgf[r] = a.dgf[rc[r]];
                    q.rh[r] = a.dh[rc[r]];
                }
            }
            return q;
        };

        // mu_t at voxel (i, j, k); adds the voxel's sensitivities to g
        auto voxel = [&](int i, const Col &cm, const Col &c0, const Col &cpl, const Nb &q) {
            const bool u_int = interior(i, j, k);
            const double ap = c0.ap, an = c0.an;
            // X_t = S_t + d_t on the gap layers
            const double xp = c0.sp + (gap_k ? c0.d : 0.0), xn = c0.sn + (gap_k ? c0.d : 0.0);
            const double xp_im = cm.sp + (gap_k ? cm.d : 0.0), xn_im = cm.sn + (gap_k ? cm.d : 0.0);
            const double xp_ip = cpl.sp + (gap_k ? cpl.d : 0.0), xn_ip = cpl.sn + (gap_k ? cpl.d : 0.0);
            const double xp_jm = q.sp_jm + (gap_k ? q.d_jm : 0.0), xn_jm = q.sn_jm + (gap_k ? q.d_jm : 0.0);
            const double xp_jp = q.sp_jp + (gap_k ? q.d_jp : 0.0), xn_jp = q.sn_jp + (gap_k ? q.d_jp : 0.0);
            const double xp_km = q.sp_km + (gap_km ? c0.d : 0.0), xn_km = q.sn_km + (gap_km ? c0.d : 0.0);
            const double xp_kp = q.sp_kp + (gap_kp ? c0.d : 0.0), xn_kp = q.sn_kp + (gap_kp ? c0.d : 0.0);
            double op = 0.0, on = 0.0;
            if (!REFLECT) {  // carried ring and k end planes
                op = u_int ? 0.0 : ap;
                on = u_int ? 0.0 : an;
            }
            // transposed stencil: v = u + offset must be interior
            const bool zp1 = interior(i, j, k + 1), zm1 = interior(i, j, k - 1);
            const bool jp1 = interior(i, j + 1, k), jm1 = interior(i, j - 1, k);
            const bool ip1 = interior(i + 1, j, k), im1 = interior(i - 1, j, k);
            op += zp1 ? a.p_zm * q.ap_kp : 0.0;
            op += zm1 ? a.p_zp * q.ap_km : 0.0;
            op += jp1 ? a.p_lat * q.ap_jp : 0.0;
            op += jm1 ? a.p_lat * q.ap_jm : 0.0;
            op += ip1 ? a.p_lat * cpl.ap : 0.0;
            op += im1 ? a.p_lat * cm.ap : 0.0;
            on += zm1 ? a.n_zm * q.an_km : 0.0;
            on += zp1 ? a.n_zp * q.an_kp : 0.0;
            on += jp1 ? a.n_lat * q.an_jp : 0.0;
            on += jm1 ? a.n_lat * q.an_jm : 0.0;
            on += ip1 ? a.n_lat * cpl.an : 0.0;
            on += im1 ? a.n_lat * cm.an : 0.0;
            double v[kAdjNV];
#pragma unroll
            for (int c = 0; c < kAdjNV; ++c) v[c] = 0.0;
            if (u_int) {
                const double s = ap + an;
                op += a.p_cen * ap;
                on += a.n_cen * an;
                op -= a.adt * xn * s;
                on -= a.adt * xp * s;
                const double di = (double)i - a.mid, dj = (double)j - a.mid;
                const bool scored = gap_k && di * di + dj * dj < a.rsq;
                op += scored ? a.adt * xn : 0.0;
                on += scored ? a.adt * xp : 0.0;
                // sensitivities of H (and r) at this interior voxel
                v[0] = ap * (((xp_jm + xp_jp) + xp_im) + xp_ip);
                v[1] = ap * xp_km;
                v[2] = ap * xp_kp;
                v[3] = ap * xp;
                v[4] = an * (((xn_jm + xn_jp) + xn_im) + xn_ip);
                v[5] = an * xn_kp;
                v[6] = an * xn_km;
                v[7] = an * xn;
                const double pn = xp * xn;
                v[8] = (scored ? pn : 0.0) - s * pn;
            }
            if (gap_k && (!REFLECT || u_int)) {  // (a reflecting ring's terms are taken by its partner)
                v[9] = (op + on) * q.dgf;
                v[10] = (op + on) * q.dh;
            }
            const int64_t out = acol(i, j) + k;
            if constexpr (REFLECT) {
                // A ring voxel next to this one reads this column (the ring is a
                // copy of its neighbour), and its only interior consumer is this
                // voxel, through the lateral weight: its adjoint is lat * mu(u).
                // Its deposition term is taken here, with the ring column's
                // dd/dgf and dd/dh2b2, and its adjoint is folded into this voxel's
                // -- the transposed copy -- so the ring itself is never stored.
                if (u_int) {
                    const double rp = a.p_lat * ap, rn = a.n_lat * an;
                    const bool has[4] = {i == 1, i == n - 2, j == 1, j == n - 2};
                    double fp = op, fn = on;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        fp += has[r] ? rp : 0.0;
                        fn += has[r] ? rn : 0.0;
                        v[9] += (has[r] && gap_k) ? (rp + rn) * q.rgf[r] : 0.0;
                        v[10] += (has[r] && gap_k) ? (rp + rn) * q.rh[r] : 0.0;
                    }
                    a.op[out] = fp;
                    a.on[out] = fn;
                }
            } else {
                a.op[out] = op;
                a.on[out] = on;
            }
#pragma unroll
            for (int c = 0; c < kAdjNV; ++c) g[c] += v[c];
        };

        Col cm = load_col(i0 - 1), c0 = load_col(i0), cpl = load_col(i0 + 1);
        Nb q = load_nb(i0);
        for (int i = i0; i < i1; ++i) {
            const Col cn = load_col(i + 2);  // planes past the grid repeat its last one (never used)
            const Nb qn = load_nb(i + 1);
            voxel(i, cm, c0, cpl, q);
            cm = c0;
            c0 = cpl;
            cpl = cn;
            q = qn;
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
