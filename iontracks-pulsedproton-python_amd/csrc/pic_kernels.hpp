// Device kernels of the gfx950 backend. Included once, by pic_hip.hip.
//
// Grid layout: each species is (no_xy, no_xy, nzb), k fastest, stored as S
// (double or float). Interior voxels are 1 <= i,j <= no_xy-2, 1 <= k <= nzb-2.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pic {

constexpr int kWave = 64;
constexpr int kOut = 62;  // output lanes per wave window (lanes 1..62)
constexpr int kXcds = 8;  // MI355X: blocks are dispatched round-robin over 8 XCDs

// Map the hardware block index to a logical tile index so that each XCD
// (hence each L2) receives a contiguous run of logical tiles: neighbouring
// j-tiles of one i-chunk then share an L2 and their halo rows hit in it.
__device__ __forceinline__ int xcd_logical_block(int b, int nblk) {
    const int per = nblk / kXcds, rem = nblk % kXcds;
    const int xcd = b % kXcds, slot = b / kXcds;
    const int start = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    return start + slot;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum three per-thread values over the block; thread 0 returns them.
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c) {
    __shared__ double s[3][16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    if (lane == 0) {
        s[0][wave] = a;
        s[1][wave] = b;
        s[2][wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + kWave - 1) / kWave;
        a = s[0][0];
        b = s[1][0];
        c = s[2][0];
        for (int w = 1; w < nw; ++w) {
            a += s[0][w];
            b += s[1][w];
            c += s[2][w];
        }
    }
}

struct SweepArgs {
    const void *pos;
    const void *neg;
    void *pos_out;
    void *neg_out;
    const double *dens;  // gf-scaled 2D density of this step, or null
    double *partials;    // [gridDim.x][3]: recombined, total+, total-
    int no_xy, nzb, nze, no_z;
    int n_windows;  // k windows of kOut outputs covering k = 1..nzb-2
    int wpb;        // windows (waves) per block
    int n_kg, n_jt, n_ic;
    int chunk;  // planes per block
    double mid, rsq;
    double p_lat, p_zm, p_zp, p_cen;
    double n_lat, n_zm, n_zp, n_cen;
    double alpha_dt;
};

typedef float f32x2 __attribute__((ext_vector_type(2)));

// One Lax-Wendroff step of both species, fused with the gap-layer deposition
// of this step (DEP) and the scored reduction.
//
// Work decomposition: a wave owns a k window, TJ consecutive interior rows j of
// one tile, and marches through a chunk of i planes. Each lane holds V
// consecutive k values (V = 2 for FP32 fields, one 8-byte load per row, so a
// wave moves 512 B per load as the FP64 path does with V = 1); a window of
// 64*V values writes all but its first and last one, and consecutive windows
// overlap by two so those edge values are supplied by the neighbour window.
// Per lane it keeps, in registers, the tile's rows of plane i-1, plane i (with
// the two halo rows j0-1 and j0+TJ), plane i+1, and plane i+2 which is already
// being loaded while plane i is computed, so memory latency overlaps the
// arithmetic of a whole plane. The j and i neighbours are register reads and
// every value is fetched from memory once per plane (plus the halo rows, which
// the neighbouring tile on the same XCD has usually just brought into L2). The
// k neighbours come from the adjacent lanes by wave shuffles. No LDS staging
// and no barriers in the march: waves are independent.
//
// The arithmetic is the reference's, term for term and in the same order
// (solver_numba_parallel.py:374-405); with S = A = double and contraction off
// a voxel update is bit-identical to the reference's FP64 sweep. With
// S = A = float the stencil runs in FP32 while the recombination product, the
// deposited density and every scored sum stay FP64.
template <typename S, typename A, int TJ, bool DEP, int V>
__global__ void __launch_bounds__(512) sweep_kernel(SweepArgs a) {
    static_assert(V == 1 || (V == 2 && sizeof(S) == 4), "two values per lane is the FP32 layout");
    constexpr int WIN = kWave * V - 2;  // outputs per window
    const int logical = xcd_logical_block(blockIdx.x, gridDim.x);
    const int kg = logical % a.n_kg;
    const int rest = logical / a.n_kg;
    const int jt = rest % a.n_jt;
    const int ic = rest / a.n_jt;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int win = kg * a.wpb + wave;

    double acc_r = 0.0, acc_p = 0.0, acc_n = 0.0;
    if (win < a.n_windows) {
        const int no_xy = a.no_xy, nzb = a.nzb;
        const int k0 = win * WIN + V * lane;
        const bool k_in = k0 < nzb;  // with V = 2, nzb is even so k0+1 < nzb too
        bool k_out[V], k_gap[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int q = V * lane + v, k = k0 + v;
            k_out[v] = q >= 1 && q <= kWave * V - 2 && k >= 1 && k <= nzb - 2;
            k_gap[v] = k >= a.nze && k < a.nze + a.no_z;
        }
        const int j0 = 1 + jt * TJ;
        const int nrows = min(TJ, no_xy - 1 - j0);  // valid own rows
        const int i0 = 1 + ic * a.chunk;
        const int i1 = min(i0 + a.chunk, no_xy - 1);
        const S *P = static_cast<const S *>(a.pos);
        const S *N = static_cast<const S *>(a.neg);
        S *PO = static_cast<S *>(a.pos_out);
        S *NO = static_cast<S *>(a.neg_out);
        const int64_t plane = (int64_t)no_xy * nzb;

        // values of (i, j, k0..k0+V-1) with this step's deposition added in the
        // gap layers; the row base is wave-uniform, so each load is one scalar
        // base plus the lane's k offset
        auto load2 = [&](int i, int j, A (&vp)[V], A (&vn)[V]) {
            const int64_t row = (int64_t)i * plane + (int64_t)j * nzb;
            S sp[V], sn[V];
#pragma unroll
            for (int v = 0; v < V; ++v) sp[v] = sn[v] = (S)0;
            if (k_in) {
                if constexpr (V == 2) {
                    const f32x2 tp = *reinterpret_cast<const f32x2 *>(P + row + k0);
                    const f32x2 tn = *reinterpret_cast<const f32x2 *>(N + row + k0);
                    sp[0] = tp.x;
                    sp[1] = tp.y;
                    sn[0] = tn.x;
                    sn[1] = tn.y;
                } else {
                    sp[0] = (P + row)[k0];
                    sn[0] = (N + row)[k0];
                }
            }
            double d = 0.0;
            if (DEP) d = a.dens[(int64_t)i * no_xy + j];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                if (DEP && k_gap[v]) {
                    vp[v] = (A)((double)sp[v] + d);
                    vn[v] = (A)((double)sn[v] + d);
                } else {
                    vp[v] = (A)sp[v];
                    vn[v] = (A)sn[v];
                }
            }
        };
        auto load_rows = [&](int i, A (&rp)[TJ + 2][V], A (&rn)[TJ + 2][V]) {
#pragma unroll
            for (int r = 0; r < TJ + 2; ++r) {
#pragma unroll
                for (int v = 0; v < V; ++v) rp[r][v] = rn[r][v] = (A)0;
                if (r <= nrows + 1) load2(i, j0 + r - 1, rp[r], rn[r]);
            }
        };

        A pp[TJ][V], pn[TJ][V];          // plane i-1, own rows
        A cp[TJ + 2][V], cn[TJ + 2][V];  // plane i, rows j0-1 .. j0+TJ
        A xp[TJ + 2][V], xn[TJ + 2][V];  // plane i+1
        A fp[TJ + 2][V], fn[TJ + 2][V];  // plane i+2, in flight
#pragma unroll
        for (int r = 0; r < TJ; ++r) {
#pragma unroll
            for (int v = 0; v < V; ++v) pp[r][v] = pn[r][v] = (A)0;
            if (r < nrows) load2(i0 - 1, j0 + r, pp[r], pn[r]);
        }
        load_rows(i0, cp, cn);
        load_rows(i0 + 1, xp, xn);

        for (int i = i0; i < i1; ++i) {
            if (i + 1 < i1) load_rows(i + 2, fp, fn);
            const double di = (double)i - a.mid;
            const double di2 = di * di;
#pragma unroll
            for (int r = 0; r < TJ; ++r) {
                A p_km[V], p_kp[V], n_km[V], n_kp[V];
                if constexpr (V == 2) {
                    p_km[0] = __shfl_up(cp[r + 1][1], 1, kWave);
                    p_kp[0] = cp[r + 1][1];
                    p_km[1] = cp[r + 1][0];
                    p_kp[1] = __shfl_down(cp[r + 1][0], 1, kWave);
                    n_km[0] = __shfl_up(cn[r + 1][1], 1, kWave);
                    n_kp[0] = cn[r + 1][1];
                    n_km[1] = cn[r + 1][0];
                    n_kp[1] = __shfl_down(cn[r + 1][0], 1, kWave);
                } else {
                    p_km[0] = __shfl_up(cp[r + 1][0], 1, kWave);
                    p_kp[0] = __shfl_down(cp[r + 1][0], 1, kWave);
                    n_km[0] = __shfl_up(cn[r + 1][0], 1, kWave);
                    n_kp[0] = __shfl_down(cn[r + 1][0], 1, kWave);
                }
                if (r < nrows) {
                    const int j = j0 + r;
                    const double dj = (double)j - a.mid;
                    const bool inside = di2 + dj * dj < a.rsq;
                    S po[V], no[V];
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        const A p = cp[r + 1][v], n = cn[r + 1][v];
                        const A p_new = (A)a.p_zm * p_km[v] + (A)a.p_zp * p_kp[v] + (A)a.p_lat * cp[r][v] +
                                        (A)a.p_lat * cp[r + 2][v] + (A)a.p_lat * pp[r][v] +
                                        (A)a.p_lat * xp[r + 1][v] + (A)a.p_cen * p;
                        const A n_new = (A)a.n_zm * n_kp[v] + (A)a.n_zp * n_km[v] + (A)a.n_lat * cn[r][v] +
                                        (A)a.n_lat * cn[r + 2][v] + (A)a.n_lat * pn[r][v] +
                                        (A)a.n_lat * xn[r + 1][v] + (A)a.n_cen * n;
                        const double rec = a.alpha_dt * (double)p * (double)n;
                        po[v] = (S)((double)p_new - rec);
                        no[v] = (S)((double)n_new - rec);
                        if (k_out[v] && k_gap[v] && inside) {
                            acc_r += rec;
                            acc_p += (double)po[v];
                            acc_n += (double)no[v];
                        }
                    }
                    const int64_t row = (int64_t)i * plane + (int64_t)j * nzb;
                    if constexpr (V == 2) {
                        if (k_out[0] && k_out[1]) {
                            *reinterpret_cast<f32x2 *>(PO + row + k0) = f32x2{po[0], po[1]};
                            *reinterpret_cast<f32x2 *>(NO + row + k0) = f32x2{no[0], no[1]};
                        } else {
#pragma unroll
                            for (int v = 0; v < V; ++v)
                                if (k_out[v]) {
                                    (PO + row)[k0 + v] = po[v];
                                    (NO + row)[k0 + v] = no[v];
                                }
                        }
                    } else {
                        if (k_out[0]) {
                            (PO + row)[k0] = po[0];
                            (NO + row)[k0] = no[0];
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < TJ; ++r)
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    pp[r][v] = cp[r + 1][v];
                    pn[r][v] = cn[r + 1][v];
                }
#pragma unroll
            for (int r = 0; r < TJ + 2; ++r)
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    cp[r][v] = xp[r][v];
                    cn[r][v] = xn[r][v];
                    xp[r][v] = fp[r][v];
                    xn[r][v] = fn[r][v];
                }
        }
    }
    block_sum3(acc_r, acc_p, acc_n);
    if (threadIdx.x == 0) {
        a.partials[3 * (int64_t)blockIdx.x + 0] = acc_r;
        a.partials[3 * (int64_t)blockIdx.x + 1] = acc_p;
        a.partials[3 * (int64_t)blockIdx.x + 2] = acc_n;
    }
}

struct GatherArgs {
    const double *xs;
    const double *ys;
    int64_t m;
    double *dens;      // no_xy^2, gf-scaled summed density
    double *partials;  // [gridDim.x] scored injected charge
    int no_xy, no_z, ntj;
    double cutoff, h2b2, gf, mid, rsq;
};

constexpr int kGT = 16;  // gather tile edge

__device__ __forceinline__ void track_box(double x, double y, double c, double nx, int &li, int &hi, int &lj,
                                          int &hj) {
    li = (int)fmax(ceil(x - c), 0.0);
    hi = (int)fmin(floor(x + c) + 1.0, nx);
    lj = (int)fmax(ceil(y - c), 0.0);
    hj = (int)fmin(floor(y + c) + 1.0, nx);
}

// Deposition of one step's tracks as a gather: each thread owns one (i, j)
// column of a 16x16 tile and sums, in ascending track order, gi(t)*gj(t) over
// the tracks whose stencil box covers it -- the same terms in the same order
// as the reference's CSR row accumulation (solver_numba_parallel.py:183-219),
// so no atomics and a deterministic result. The block first compacts, in
// order, the tracks whose box meets the tile (ballot + prefix count), then
// evaluates each listed track's 16 row and 16 column factors once into LDS.
// Output: dens = sum * gaussian_factor, and per block the scored injected
// charge sum(dens * no_z) over interior columns inside the scoring disc
// (solver_numba_parallel.py:250-277).
__global__ void __launch_bounds__(256) gather_kernel(GatherArgs a) {
    __shared__ int s_list[256];
    __shared__ int s_wcnt[4];
    __shared__ int s_box[8][4];
    __shared__ double s_g[8][2 * kGT];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ti0 = (blockIdx.x / a.ntj) * kGT, tj0 = (blockIdx.x % a.ntj) * kGT;
    const int ty = tid / kGT, tx = tid % kGT;
    const int i = ti0 + ty, j = tj0 + tx;
    const double nx = (double)a.no_xy;
    double acc = 0.0;
    for (int64_t base = 0; base < a.m; base += 256) {
        const int64_t t = base + tid;
        bool hit = false;
        if (t < a.m) {
            int li, hi, lj, hj;
            track_box(a.xs[t], a.ys[t], a.cutoff, nx, li, hi, lj, hj);
            hit = li < ti0 + kGT && hi > ti0 && lj < tj0 + kGT && hj > tj0;
        }
        const unsigned long long bal = __ballot(hit);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) s_wcnt[wave] = __popcll(bal);
        __syncthreads();
        int off = 0, nlist = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wave) off += s_wcnt[w];
            nlist += s_wcnt[w];
        }
        if (hit) s_list[off + before] = tid;
        __syncthreads();
        for (int sb = 0; sb < nlist; sb += 8) {
            const int q = tid >> 5, f = tid & 31;
            if (sb + q < nlist) {
                const int64_t tt = base + s_list[sb + q];
                const double x = a.xs[tt], y = a.ys[tt];
                const double d = f < kGT ? (double)(ti0 + f) - x : (double)(tj0 + f - kGT) - y;
                s_g[q][f] = exp(-(d * d) * a.h2b2);
                if (f == 0) {
                    int li, hi, lj, hj;
                    track_box(x, y, a.cutoff, nx, li, hi, lj, hj);
                    s_box[q][0] = li;
                    s_box[q][1] = hi;
                    s_box[q][2] = lj;
                    s_box[q][3] = hj;
                }
            }
            __syncthreads();
            const int nq = min(8, nlist - sb);
            for (int q2 = 0; q2 < nq; ++q2) {
                if (i >= s_box[q2][0] && i < s_box[q2][1] && j >= s_box[q2][2] && j < s_box[q2][3])
                    acc += s_g[q2][ty] * s_g[q2][kGT + tx];
            }
            __syncthreads();
        }
    }
    double inj = 0.0, z1 = 0.0, z2 = 0.0;
    if (i < a.no_xy && j < a.no_xy) {
        const double d = acc * a.gf;
        a.dens[(int64_t)i * a.no_xy + j] = d;
        const double di = (double)i - a.mid, dj = (double)j - a.mid;
        if (acc != 0.0 && i >= 1 && i <= a.no_xy - 2 && j >= 1 && j <= a.no_xy - 2 && di * di + dj * dj < a.rsq)
            inj = d * (double)a.no_z;
    }
    block_sum3(inj, z1, z2);
    if (tid == 0) a.partials[blockIdx.x] = inj;
}

// Explicit deposition into the current arrays (kernel-level API): add the
// step's 2D density to every gap layer (solver_numba_parallel.py:253-262).
template <typename S>
__global__ void __launch_bounds__(256) broadcast_kernel(S *pos, S *neg, const double *dens, int no_xy, int nzb,
                                                        int nze, int no_z) {
    const int64_t n = (int64_t)no_xy * no_xy * no_z;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int kz = (int)(e % no_z);
        const int64_t col = e / no_z;
        const double d = dens[col];
        if (d == 0.0) continue;
        const int64_t idx = col * nzb + nze + kz;
        pos[idx] = (S)((double)pos[idx] + d);
        neg[idx] = (S)((double)neg[idx] + d);
    }
}

// Lateral ring of the next arrays. The sweep writes only the interior, so
// after it the next buffers' outer ring is stale:
//  - absorbing: the ring is carried over from the current arrays, with this
//    step's deposition (which the reference adds to the ring too) included
//    (state.py:157-176, solver_numba_parallel.py:691-695);
//  - reflecting: every ring voxel copies the nearest interior column of the
//    next arrays, i.e. i planes mirrored first, then j planes, corners taking
//    the diagonal interior column (state.py:113-154).
template <typename S, bool REFLECT, bool DEP>
__global__ void __launch_bounds__(256) ring_kernel(S *pos_next, S *neg_next, const S *pos, const S *neg,
                                                   const double *dens, int no_xy, int nzb, int nze, int no_z) {
    const int64_t face = (int64_t)no_xy * nzb;           // one i plane
    const int64_t side = (int64_t)(no_xy - 2) * nzb;     // one j column set, i = 1..no_xy-2
    const int64_t n = 2 * face + 2 * side;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        if (e < 2 * face) {
            i = e < face ? 0 : no_xy - 1;
            const int64_t r = e % face;
            j = (int)(r / nzb);
            k = (int)(r % nzb);
        } else {
            const int64_t r = e - 2 * face;
            j = r < side ? 0 : no_xy - 1;
            const int64_t q = r % side;
            i = 1 + (int)(q / nzb);
            k = (int)(q % nzb);
        }
        const int64_t idx = ((int64_t)i * no_xy + j) * nzb + k;
        if (REFLECT) {
            const int ci = min(max(i, 1), no_xy - 2), cj = min(max(j, 1), no_xy - 2);
            const int64_t src = ((int64_t)ci * no_xy + cj) * nzb + k;
            pos_next[idx] = pos_next[src];
            neg_next[idx] = neg_next[src];
        } else {
            double vp = (double)pos[idx], vn = (double)neg[idx];
            if (DEP && k >= nze && k < nze + no_z) {
                const double d = dens[(int64_t)i * no_xy + j];
                vp = vp + d;
                vn = vn + d;
            }
            pos_next[idx] = (S)vp;
            neg_next[idx] = (S)vn;
        }
    }
}

// Neumaier-compensated accumulation of one term.
struct Comp {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double x) {
        const double t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ void add(const Comp &o) {
        add(o.s);
        c += o.c;
    }
    __device__ __forceinline__ double value() const { return s + c; }
};

// Per-step scalars from the block partials, in a fixed order with compensated
// summation: out4 = (injected, recombined, total+, total-). One block.
__global__ void __launch_bounds__(256) reduce_kernel(const double *sweep_part, int n_sweep, const double *dep_part,
                                                     int n_dep, double *out4) {
    __shared__ double s_s[4][256];
    __shared__ double s_c[4][256];
    Comp q[4];
    for (int b = threadIdx.x; b < n_dep; b += 256) q[0].add(dep_part[b]);
    for (int b = threadIdx.x; b < n_sweep; b += 256) {
        q[1].add(sweep_part[3 * (int64_t)b + 0]);
        q[2].add(sweep_part[3 * (int64_t)b + 1]);
        q[3].add(sweep_part[3 * (int64_t)b + 2]);
    }
    for (int v = 0; v < 4; ++v) {
        s_s[v][threadIdx.x] = q[v].s;
        s_c[v][threadIdx.x] = q[v].c;
    }
    __syncthreads();
    for (int stride = 128; stride >= 1; stride >>= 1) {
        if (threadIdx.x < stride) {
            for (int v = 0; v < 4; ++v) {
                Comp x{s_s[v][threadIdx.x], s_c[v][threadIdx.x]};
                Comp y{s_s[v][threadIdx.x + stride], s_c[v][threadIdx.x + stride]};
                x.add(y);
                s_s[v][threadIdx.x] = x.s;
                s_c[v][threadIdx.x] = x.c;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int v = 0; v < 4; ++v) out4[v] = s_s[v][0] + s_c[v][0];
}

template <typename S>
__global__ void __launch_bounds__(256) fill_kernel(S *p, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        p[e] = (S)0;
}

}  // namespace pic
