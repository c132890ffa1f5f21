// Device kernels of the gfx950 backend. Included once, by pic_hip.hip.
//
// Grid layout: each species is (no_xy, no_xy, nzb), k fastest, stored as S
// (double or float) with a row pitch of `pitch` >= nzb elements (FP32 rows are
// padded to an even length, so every lane's first value is 8-byte aligned;
// the padding is zero and only ever written with zeros), followed by
// kFieldSlack zero bytes. Interior voxels are
// 1 <= i,j <= no_xy-2,
// 1 <= k <= nzb-2.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

#include <type_traits>
#include <utility>

namespace pic {

constexpr int kWave = 64;
constexpr int kOut = 62;  // output lanes per wave window (lanes 1..62)

// Zero bytes after each carrier array, and past each i plane in the sweep's
// buffer resources: with four values per lane on rows whose length is 2 mod
// 4, the lane holding a row's last two values loads 16 bytes, 8 of them past
// the row (in the next row, or past the plane and the array on their last
// row). Those two values are never used; the slack keeps the load inside the
// allocation and the resource's range.
constexpr int kFieldSlack = 16;

// k values a wave window of v values per lane writes: those of lanes 1..62,
// so windows step by 62 v and overlap by 2 v, and every storing lane writes
// all of its values at once (one 4-, 8- or 16-byte store)
constexpr int window_outputs(int v) { return v * (kWave - 2); }

// k step between the windows of a row of n interior values: window_outputs,
// or for two values per lane 126 where that takes fewer windows (lane 0's
// second and lane 63's first value are then written too, by single-value
// stores; 1008 values are eight such windows but nine of 124)
__host__ __device__ inline int window_step(int v, long long n) {
    const int narrow = window_outputs(v), wide = 2 * kWave - 2;
    if (v != 2) return narrow;
    return (n + wide - 1) / wide < (n + narrow - 1) / narrow ? wide : narrow;
}
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

// Call f(std::integral_constant<int, q>) for q = 0 .. N-1 (compile-time loop).
template <typename F, int... Q>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, Q...>) {
    (f(std::integral_constant<int, Q>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Planes a sweep wave keeps in flight ahead of the one it computes. Each
// buffered plane costs 2 (TJ + 2) V registers. Measured on MI355X: one plane
// ahead (its load issued a full step before its first use) is best for every
// tile shape; two to four cost occupancy and gained nothing, from the 16^2 x
// 106 grids of a saturation curve to the 1066^2 x 410 full electrode. Two
// planes for the 2-row FP32 tile, which keeps four waves per SIMD at 124
// registers, still lost: 32-point curve 36.9 against 33.9 ms batched, 202
// against 197 ms sequential.
template <typename S, int TJ, int V>
constexpr int lookahead_planes() {
    return 1;
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

// Fold (s, c) pairs across the wave by shuffles, fixed tree order.
__device__ __forceinline__ Comp wave_comp(Comp x) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        Comp y{__shfl_xor(x.s, off, kWave), __shfl_xor(x.c, off, kWave)};
        x.add(y);
    }
    return x;
}

// Sum a [n3][3] table of block partials and, if given, an [n_dep] table,
// over the whole block in a fixed order with compensated summation; thread 0
// receives the four totals. Used by the fold block of the next sweep (and by
// reduce_kernel) to turn block partials into the step's scalars.
__device__ void block_final4(const double *part3, int n3, const double *dep, int n_dep, double out[4]) {
    __shared__ double s_s[4][16], s_c[4][16];
    Comp q[4];
    // Thread t adds entries t, t + B, t + 2B, ... (B = blockDim.x) in that
    // order. The loads of U entries are issued together before the adds (index
    // clamped, so no load sits behind a branch): a compensated sum is a
    // dependent chain, and loading inside it would cost one memory latency
    // per entry.
    constexpr int U = 8;
    const int B = blockDim.x;
    for (int b0 = threadIdx.x; b0 < n_dep; b0 += U * B) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = dep[min(b0 + u * B, n_dep - 1)];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (b0 + u * B < n_dep) q[0].add(v[u]);
    }
    for (int b0 = threadIdx.x; b0 < n3; b0 += U * B) {
        double v[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t b = min(b0 + u * B, n3 - 1);
            v[u][0] = part3[3 * b + 0];
            v[u][1] = part3[3 * b + 1];
            v[u][2] = part3[3 * b + 2];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (b0 + u * B < n3) {
                q[1].add(v[u][0]);
                q[2].add(v[u][1]);
                q[3].add(v[u][2]);
            }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        q[v] = wave_comp(q[v]);
        if (lane == 0) {
            s_s[v][wave] = q[v].s;
            s_c[v][wave] = q[v].c;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + kWave - 1) / kWave;
        for (int v = 0; v < 4; ++v) {
            Comp t{s_s[v][0], s_c[v][0]};
            for (int w = 1; w < nw; ++w) t.add(Comp{s_s[v][w], s_c[v][w]});
            out[v] = t.value();
        }
    }
}

struct SweepArgs {
    const void *pos;
    const void *neg;
    void *pos_out;
    void *neg_out;
    const double *dens;      // gf-scaled 2D density of this step, or null
    double *partials;        // [gridDim.x][3]: recombined, total+, total-
    // the previous step's partials, folded into its record by a block of its own
    const double *prev_part;      // [prev_n][3], or null
    int prev_n;
    const double *prev_dep_part;  // [prev_n_dep], or null
    int prev_n_dep;
    double *prev_out4;            // (injected, recombined, total+, total-), or null
    int no_xy, nzb, nze, no_z;
    int pitch;      // elements from one (i, j) row to the next
    int n_windows;  // k windows covering k = 1..nzb-2
    int win_step;   // k values between windows (window_step)
    int wpb;        // windows per tile of a block
    int tpb;        // consecutive j tiles per block (0 or 1: one); waves = wpb * tpb
    int n_kg, n_jt, n_ic;
    int chunk;  // planes per block
    double mid, rsq;
    double p_lat, p_zm, p_zp, p_cen;
    double n_lat, n_zm, n_zp, n_cen;
    double alpha_dt;
};

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Fourth word of a raw buffer resource on gfx9-family devices (gfx950
// included): 32-bit data format, no swizzle, no index stride.
constexpr int kBufferWord3 = 0x00020000;

// The FP32 voxel update (flux form): x + sum_w w (x_w - x), the two drift
// terms as fused multiply-adds over the lateral sum. The weights sum to one,
// so this is the same update as the reference's weighted sum, and written as
// fluxes it moves charge between neighbours without creating any, whatever
// the rounding of the FP32 weights. The recombination product and the final
// x + flux - rec are FP64 (x + flux is exact there), so the stored value is
// rounded to FP32 once: on the strongest-recombination point of a saturation
// curve (50 V, 40 Gy/s) that keeps |dk_s| at 2.3e-7 where two FP32 roundings
// reach 9.5e-7. n's drift runs the other way (km and kp swapped), as in the
// reference (solver_numba_parallel.py:374-405).
__device__ __forceinline__ void fp32_update(float p, float n, float p_km, float p_kp, float n_km, float n_kp,
                                            float p_jm, float p_jp, float p_im, float p_ip, float n_jm, float n_jp,
                                            float n_im, float n_ip, float f_pzm, float f_pzp, float f_plat,
                                            float f_nzm, float f_nzp, float f_nlat, double alpha_dt, float &po,
                                            float &no, double &rec) {
    const float lp = ((p_jm - p) + (p_jp - p)) + ((p_im - p) + (p_ip - p));
    const float ln = ((n_jm - n) + (n_jp - n)) + ((n_im - n) + (n_ip - n));
    const float tp = __builtin_fmaf(f_pzm, p_km - p, __builtin_fmaf(f_pzp, p_kp - p, f_plat * lp));
    const float tn = __builtin_fmaf(f_nzm, n_kp - n, __builtin_fmaf(f_nzp, n_km - n, f_nlat * ln));
    rec = alpha_dt * (double)p * (double)n;
    po = (float)(((double)p + (double)tp) - rec);
    no = (float)(((double)n + (double)tn) - rec);
}

// Scored sums of a lane's four FP32 values, in FP64: the scored ones of the
// four summed first, then added to the lane's sums.
__device__ __forceinline__ void score4(const bool (&k_score)[4], const double (&rec)[4], const float (&po)[4],
                                       const float (&no)[4], double &sr, double &sp, double &sn) {
    double r = k_score[0] ? rec[0] : 0.0;
    double pp = k_score[0] ? (double)po[0] : 0.0, nn = k_score[0] ? (double)no[0] : 0.0;
#pragma unroll
    for (int v = 1; v < 4; ++v) {
        r += k_score[v] ? rec[v] : 0.0;
        pp += k_score[v] ? (double)po[v] : 0.0;
        nn += k_score[v] ? (double)no[v] : 0.0;
    }
    sr += r;
    sp += pp;
    sn += nn;
}

// One Lax-Wendroff step of both species, fused with the gap-layer deposition
// of this step (DEP), the lateral-ring update of the next arrays and the
// scored reduction.
//
// Work decomposition: a wave owns a k window, TJ consecutive interior rows j of
// one tile, and marches through a chunk of i planes. Each lane holds V
// consecutive k values (V = 2 for FP32 fields, one 8-byte load per row, so a
// wave moves 512 B per load as the FP64 path does with V = 1). The values of
// lanes 0 and 63 only feed their neighbours' stencils, so the window writes
// lanes 1..62 (window_outputs) and consecutive windows overlap by 2V, the
// first window's lane 0 also writing k = 1. With V = 2 each storing lane then
// writes a whole pair -- one 8-byte store per row and array, no partial-lane
// stores -- and the one value of a pair it does not own is a k end plane,
// which stays zero.
// Per lane it keeps, in registers, the tile's rows of plane i-1, plane i (with
// the two halo rows j0-1 and j0+TJ), plane i+1, and plane i+2 which is already
// being loaded while plane i is computed, so memory latency overlaps the
// arithmetic of a whole plane. The j and i neighbours are register reads and
// every value is fetched from memory once per plane (plus the halo rows, which
// the neighbouring tile on the same XCD has usually just brought into L2). The
// k neighbours come from the adjacent lanes by wave shuffles. No LDS staging
// and no barriers in the march: waves are independent.
//
// Lateral ring (the sweep proper writes only the interior):
//  - REFLECT: the ring is never stored. A zero-gradient wall makes every ring
//    voxel equal the nearest interior column -- i planes mirrored first, then
//    j planes, a corner taking the diagonal column (state.py:113-154) -- so a
//    load of ring voxel (i', j') reads interior column (clamp(i'), clamp(j'))
//    and adds the deposition at (i', j') itself, which the reference deposits
//    into the ring after mirroring. The stored ring is filled in only when the
//    fields are read back (ring_mirror_kernel).
//  - absorbing: the tiles and chunks at the grid edge carry the ring values
//    they load -- current value plus this step's deposition, which the
//    reference also adds to the ring -- into the next arrays
//    (state.py:157-176, solver_numba_parallel.py:691-695).
// The k end planes are zero in both buffers and are never written.
//
// Reduction: each block writes its three scored sums to its slot; a block of
// its own in the NEXT step's launch folds them (with that step's deposition
// partials) in a fixed order with compensated FP64 summation into the step's
// record, so the kernel boundary orders the partials and no device-wide fence
// is needed. The last step of a run is folded by reduce_kernel.
//
// The arithmetic is the reference's, term for term and in the same order
// (solver_numba_parallel.py:374-405); with S = A = double and contraction off
// a voxel update is bit-identical to the reference's FP64 sweep. With FP32
// fields (S = float) the stencil runs in FP32 in flux form (see the update):
// the reference's weighted sum with FP32-rounded weights would not conserve
// charge -- the weights no longer sum to one -- and that per-step drift, not
// the FP32 rounding of each value, is what would move k_s. The recombination
// product and every scored sum stay FP64.
//
// sweep_body runs one block's share: `logical` is the block's index among the
// `nblk` blocks of this configuration (already XCD-remapped), DEP says whether
// this step deposits (a runtime flag in the multi-configuration kernel). CH > 0
// compiles a march of exactly CH planes (chunks of at most CH planes): no
// step past the chunk runs, and the last step loads nothing ahead. CH = 0
// marches any chunk in groups of the plane-buffer count.
template <typename S, typename A, int TJ, int V, bool REFLECT, int CH = 0>
__device__ __forceinline__ void sweep_body(const SweepArgs &a, const int logical, const bool DEP) {
    static_assert(V == 1 || V == 2 || (V == 4 && sizeof(S) == 4), "four values per lane are an FP32 layout");
    constexpr int WIN = window_outputs(V);  // k step between windows (V = 2: a.win_step, 124 or 126)
    // a block holds tpb consecutive j tiles of one chunk, each with wpb
    // windows: waves [q * wpb, (q + 1) * wpb) march tile q of the group, so
    // neighbouring tiles load their shared halo rows from one CU at the same
    // time
    const int tpb = a.tpb > 1 ? a.tpb : 1;
    const int n_jg = (a.n_jt + tpb - 1) / tpb;
    const int kg = logical % a.n_kg;
    const int rest = logical / a.n_kg;
    const int jg = rest % n_jg;
    const int ic = rest / n_jg;
    // the wave index is uniform across the wave; saying so lets the compiler
    // keep the whole march in scalar control flow and scalar loads
    const int wave_id = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int wave = wave_id % a.wpb;
    const int jt = jg * tpb + wave_id / a.wpb;
    const int win = kg * a.wpb + wave;

    double acc_r = 0.0, acc_p = 0.0, acc_n = 0.0;
    // The tiles of a block march in lockstep, one barrier per plane, when
    // every wave of the block marches: the block is full-sized, its window
    // group holds wpb windows and its tile group tpb tiles (a block-uniform
    // condition: all its waves then run the same planes and reach every
    // barrier). Left to
    // themselves, neighbouring tiles drift apart along i and the halo rows
    // one loads are gone from the L2 by the time the other needs them:
    // measured on the 5 um full electrode, two tiles per block of 32 planes
    // take 1.58 ms per step in lockstep, 1.70 ms free-running.
    const bool lockstep = tpb > 1 && (int)(blockDim.x >> 6) == a.wpb * tpb && (kg + 1) * a.wpb <= a.n_windows &&
                          (jg + 1) * tpb <= a.n_jt;
    if (wave_id < a.wpb * tpb && win < a.n_windows && jt < a.n_jt) {
        const int no_xy = a.no_xy, nzb = a.nzb, pitch = a.pitch;
        const int step = V == 2 ? a.win_step : WIN;
        const int k0 = win * step + V * lane;
        // k0 is a multiple of V; the pitch is a multiple of two (FP32), so a
        // lane with k0 < nzb reads at most the row's padding or, with four
        // values on a row of 2 mod 4 values, two values past the row; with two
        // FP64 values on a row of odd length (FP64 rows are not padded), one
        // value past the row -- the next row's first, or slack past the array
        const bool k_in = k0 < nzb;
        bool k_out[V], k_gap[V], k_score[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int q = V * lane + v, k = k0 + v;
            // owned: lanes 1..62 (and k = 1 .. V-1 in the first window), inside
            // the k range; 126-value windows also own q = 1 and q = 126
            const bool own_lane = step == 2 * kWave - 2 ? q >= 1 && q <= 2 * kWave - 2
                                                        : (q >= V || (win == 0 && v >= 1)) && q < V * (kWave - 1);
            k_out[v] = own_lane && k >= 1 && k <= nzb - 2;
            k_gap[v] = k >= a.nze && k < a.nze + a.no_z;
            k_score[v] = k_out[v] && k_gap[v];
        }
        // two values per lane: a lane writes its pair in one store when the
        // value it does not own (if any) is a k end plane; only the edge lanes
        // of 126-value windows, whose other value the neighbour window owns,
        // write single values
        bool pair_store = false;
        if constexpr (V == 2) {
            const bool z0 = k0 == 0 || k0 == nzb - 1, z1 = k0 + 1 == nzb - 1;
            pair_store = (k_out[0] || k_out[1]) && (k_out[0] || z0) && (k_out[1] || z1);
        }
        const int j0 = 1 + jt * TJ;
        const int nrows = min(TJ, no_xy - 1 - j0);  // valid own rows
        const int i0 = 1 + ic * a.chunk;
        const int i1 = min(i0 + a.chunk, no_xy - 1);
        const bool first_tile = j0 == 1, last_tile = j0 + nrows == no_xy - 1;
        // four distinct arrays: restrict lets the wave-uniform density reads
        // go through the scalar cache
        const S *__restrict__ P = static_cast<const S *>(a.pos);
        const S *__restrict__ N = static_cast<const S *>(a.neg);
        S *__restrict__ PO = static_cast<S *>(a.pos_out);
        S *__restrict__ NO = static_cast<S *>(a.neg_out);
        const double *__restrict__ dens = a.dens;
        const int64_t plane = (int64_t)no_xy * pitch;
        const float f_pzm = (float)a.p_zm, f_pzp = (float)a.p_zp, f_plat = (float)a.p_lat;
        const float f_nzm = (float)a.n_zm, f_nzp = (float)a.n_zp, f_nlat = (float)a.n_lat;

        // raw values of (i, j, k0..k0+V-1); the row base is wave-uniform, so
        // each load is one scalar base plus the lane's k offset. Under a
        // reflecting wall a ring row reads its mirror column.
        //
        // No load of the march sits behind a branch: lanes past the end of the
        // row read its last values instead (they only ever feed lanes that
        // store nothing and sums that are masked), and callers clamp rows and
        // planes into the grid. With straight-line loads the compiler knows
        // exactly how many memory operations were issued after each one, so a
        // wave waits only for the plane it is about to use while the planes
        // behind it stay in flight; a conditional load would force it to drain
        // every load at each use.
        //
        // Every field access is a buffer access over one i plane: a resource
        // whose base is the plane (wave-uniform, scalar registers), the row
        // as the scalar offset and the lane's k as a 32-bit vector offset, so
        // no access spends vector instructions on a 64-bit address (the host
        // keeps a plane below 2 GB).
        const int kc = k_in ? k0 : pitch - V;
        const unsigned vo_load = (unsigned)kc * (unsigned)sizeof(S), vo_store = (unsigned)k0 * (unsigned)sizeof(S);
        const int plane_bytes = (int)(plane * (int64_t)sizeof(S));
        auto plane_rsrc = [&](const void *arr, int i) {
            return __builtin_amdgcn_make_buffer_rsrc(
                const_cast<void *>(static_cast<const void *>(static_cast<const S *>(arr) + (int64_t)i * plane)), 0,
                plane_bytes + kFieldSlack, kBufferWord3);
        };
        auto row_off = [&](int j) -> int { return (int)((int64_t)j * pitch * (int64_t)sizeof(S)); };
        auto load_raw = [&](int i, int j, S (&vp)[V], S (&vn)[V]) {
            const int si = REFLECT ? min(max(i, 1), no_xy - 2) : i;
            const int sj = REFLECT ? min(max(j, 1), no_xy - 2) : j;
            const auto rp = plane_rsrc(P, si), rn = plane_rsrc(N, si);
            const int so = row_off(sj);
            if constexpr (V == 4) {
                const f32x4 tp = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rp, vo_load, so, 0));
                const f32x4 tn = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rn, vo_load, so, 0));
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    vp[v] = tp[v];
                    vn[v] = tn[v];
                }
            } else if constexpr (V == 2 && sizeof(S) == 8) {
                // two FP64 values: 16 bytes per lane
                const f64x2 tp = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(rp, vo_load, so, 0));
                const f64x2 tn = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(rn, vo_load, so, 0));
                vp[0] = tp.x;
                vp[1] = tp.y;
                vn[0] = tn.x;
                vn[1] = tn.y;
            } else if constexpr (V == 2) {
                const f32x2 tp = __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(rp, vo_load, so, 0));
                const f32x2 tn = __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(rn, vo_load, so, 0));
                vp[0] = tp.x;
                vp[1] = tp.y;
                vn[0] = tn.x;
                vn[1] = tn.y;
            } else if constexpr (sizeof(S) == 8) {
                vp[0] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rp, vo_load, so, 0));
                vn[0] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rn, vo_load, so, 0));
            } else {
                vp[0] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rp, vo_load, so, 0));
                vn[0] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rn, vo_load, so, 0));
            }
        };
        // This step's deposition of the TJ+2 rows of a plane: one vector load in
        // which lane r holds row j0-1+r (the rows are contiguous in the 2D
        // density), read back per row with readlane into scalar registers.
        static_assert(TJ + 2 <= kWave, "one lane per row");
        auto load_dens = [&](int i) -> double {
            if (!DEP) return 0.0;
            const int r = min(lane, nrows + 1);
            return dens[(int64_t)i * no_xy + (j0 - 1 + r)];
        };
        auto row_dens = [&](double dv_lanes, int r) -> double {
            const int lo = __builtin_amdgcn_readlane(__double2loint(dv_lanes), r);
            const int hi = __builtin_amdgcn_readlane(__double2hiint(dv_lanes), r);
            return __hiloint2double(hi, lo);
        };
        // branch-free: x + 0 == x exactly, so the non-gap layers are unchanged
        auto add_dep = [&](double d, S (&vp)[V], S (&vn)[V]) {
            if (DEP) {
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    if constexpr (sizeof(S) == 8) {
                        const double dv = k_gap[v] ? d : 0.0;
                        vp[v] = vp[v] + dv;
                        vn[v] = vn[v] + dv;
                    } else {
                        // FP32 fields: the density is rounded once, then added in FP32
                        const float dv = k_gap[v] ? (float)d : 0.0f;
                        vp[v] = vp[v] + dv;
                        vn[v] = vn[v] + dv;
                    }
                }
            }
        };
        // rows j0-1 .. j0+nrows of plane i; the deposition is added later, when
        // the plane is consumed, so the loads stay in flight meanwhile
        // (rows past the tile's last halo row repeat it and are never used)
        auto load_rows = [&](int i, S (&rp)[TJ + 2][V], S (&rn)[TJ + 2][V], double &rd) {
#pragma unroll
            for (int r = 0; r < TJ + 2; ++r) load_raw(i, j0 + min(r, nrows + 1) - 1, rp[r], rn[r]);
            rd = load_dens(i);
        };
        auto add_dep_rows = [&](S (&rp)[TJ + 2][V], S (&rn)[TJ + 2][V], double rd) {
            if (DEP) {
#pragma unroll
                for (int r = 0; r < TJ + 2; ++r) add_dep(row_dens(rd, r), rp[r], rn[r]);
            }
        };
        // store V values of (i, j) to the next arrays, lanes/values that own k
        auto store2 = [&](int i, int j, const S (&vp)[V], const S (&vn)[V]) {
            const auto rp = plane_rsrc(PO, i), rn = plane_rsrc(NO, i);
            const int so = row_off(j);
            if constexpr (V == 4) {
                // all four values: one the lane does not own is k = 0, a k end
                // plane or the row padding, all of which stay zero; the lane
                // holding a row's last two values (rows of 2 mod 4 values)
                // writes those two only
                if (k_out[0] || k_out[1] || k_out[2] || k_out[3]) {
                    f32x4 wp, wn;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        wp[v] = k_out[v] ? vp[v] : 0.0f;
                        wn[v] = k_out[v] ? vn[v] : 0.0f;
                    }
                    if (k0 + 4 <= pitch) {
                        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, wp), rp, vo_store, so, 0);
                        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, wn), rn, vo_store, so, 0);
                    } else {
                        const f32x2 hp = {wp[0], wp[1]}, hn = {wn[0], wn[1]};
                        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, hp), rp, vo_store, so, 0);
                        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, hn), rn, vo_store, so, 0);
                    }
                }
                return;
            }
            if constexpr (V == 2) {
                // the whole pair where the value the lane does not own is k = 0
                // or nzb-1, a k end plane, which must stay zero; single values
                // only at the edges of 126-value windows
                if (pair_store) {
                    if constexpr (sizeof(S) == 8) {
                        const f64x2 wp = {k_out[0] ? vp[0] : 0.0, k_out[1] ? vp[1] : 0.0};
                        const f64x2 wn = {k_out[0] ? vn[0] : 0.0, k_out[1] ? vn[1] : 0.0};
                        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, wp), rp, vo_store, so, 0);
                        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, wn), rn, vo_store, so, 0);
                    } else {
                        const f32x2 wp = {k_out[0] ? vp[0] : 0.0f, k_out[1] ? vp[1] : 0.0f};
                        const f32x2 wn = {k_out[0] ? vn[0] : 0.0f, k_out[1] ? vn[1] : 0.0f};
                        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, wp), rp, vo_store, so, 0);
                        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, wn), rn, vo_store, so, 0);
                    }
                    return;
                }
            }
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (k_out[v]) {
                    const unsigned vo = vo_store + (unsigned)(v * sizeof(S));
                    if constexpr (sizeof(S) == 8) {
                        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, vp[v]), rp, vo, so, 0);
                        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, vn[v]), rn, vo, so, 0);
                    } else {
                        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, vp[v]), rp, vo, so, 0);
                        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, vn[v]), rn, vo, so, 0);
                    }
                }
        };
        // absorbing wall: carry the rows of a loaded plane that lie on the ring
        // (every own row too when the whole plane is ring). Every buffer index
        // is a constant: the row after the last own one, j0 + nrows, is always
        // in slot TJ + 1 (load_rows repeats it past the tile), so no register
        // array is indexed at run time -- which would move it to scratch memory.
        auto carry_ring_rows = [&](int i, const S (&rp)[TJ + 2][V], const S (&rn)[TJ + 2][V], bool whole_plane) {
            if (whole_plane) {
#pragma unroll
                for (int r = 1; r <= TJ; ++r)
                    if (r <= nrows) store2(i, j0 + r - 1, rp[r], rn[r]);
            }
            if (first_tile) store2(i, j0 - 1, rp[0], rn[0]);
            if (last_tile) store2(i, j0 + nrows, rp[TJ + 1], rn[TJ + 1]);
        };

        // Plane buffers in rotation, NB = D + 3 of them for a lookahead of D
        // planes: entering plane i, buffer (i - i0) % NB holds plane i-1 and
        // the next one plane i (both ready: deposition added), and the D+1
        // after them planes i+1 .. i+1+D, loaded in that order. A step issues
        // the load of plane i+1+D, then adds the deposition to plane i+1 --
        // issued D steps earlier, and with D planes of loads behind it, so the
        // wait leaves those in flight -- then computes plane i. The march is
        // unrolled by NB so the rotation is a renaming of registers, not a
        // copy. Small tiles (few rows, so few registers per plane) look
        // further ahead: on small grids there are too few waves to hide the
        // memory latency of a plane by occupancy, so each wave has to keep
        // several planes in flight.
        constexpr int D = lookahead_planes<S, TJ, V>();
        constexpr int NB = D + 3;
        S bp[NB][TJ + 2][V], bn[NB][TJ + 2][V];
        double fd[NB];  // deposition of each buffered plane's rows, lane r = row r
        // every prologue load is issued before the first wait (planes past the
        // chunk's last one, i1, repeat it and are never used)
        static_for<D + 2>([&](auto Qc) {
            constexpr int q = decltype(Qc)::value;
            load_rows(min(i0 - 1 + q, i1), bp[q], bn[q], fd[q]);
        });
        static_for<2>([&](auto Qc) {
            constexpr int q = decltype(Qc)::value;
            add_dep_rows(bp[q], bn[q], fd[q]);
        });
        if (!REFLECT && i0 == 1) carry_ring_rows(0, bp[0], bn[0], true);

        // per-lane sums of the scored voxels; lanes whose k is never scored are
        // masked once at the end (k_score is fixed for the whole march). With
        // four values per lane the lane's values are masked as they are added,
        // into one sum per quantity, which saves 18 registers.
        constexpr bool kMerged = V > 2;
        constexpr int NS = kMerged ? 1 : V;
        double sr[NS], sp[NS], sn[NS];
#pragma unroll
        for (int v = 0; v < NS; ++v) sr[v] = sp[v] = sn[v] = 0.0;

        auto plane_step = [&](int i, auto Oc) {
            constexpr int O = decltype(Oc)::value;  // (i - i0) % NB
            constexpr int PB = O % NB, CB = (O + 1) % NB, XB = (O + 2) % NB;
            constexpr int LB = (O + 2 + D) % NB;  // plane i+1+D: loaded now
            // every step of a group runs; one past the chunk (`live` false)
            // loads a repeat of plane i1, computes, and stores nothing
            const bool live = i < i1;
            if (lockstep) __syncthreads();
            if constexpr (CH == 0 || O + 1 < CH) load_rows(min(i + 1 + D, i1), bp[LB], bn[LB], fd[LB]);
            add_dep_rows(bp[XB], bn[XB], fd[XB]);
            if (!REFLECT && live) {
                carry_ring_rows(i, bp[CB], bn[CB], false);
                if (i == no_xy - 2) carry_ring_rows(no_xy - 1, bp[XB], bn[XB], true);
            }
            const double di = (double)i - a.mid;
            const double di2 = di * di;
#pragma unroll
            for (int r = 0; r < TJ; ++r) {
                const S(&cp)[TJ + 2][V] = bp[CB];
                const S(&cn)[TJ + 2][V] = bn[CB];
                S p_km[V], p_kp[V], n_km[V], n_kp[V];
                if constexpr (V >= 2) {
                    // inner neighbours are the lane's own values; the outer
                    // ones come from the adjacent lanes
                    p_km[0] = __shfl_up(cp[r + 1][V - 1], 1, kWave);
                    p_kp[V - 1] = __shfl_down(cp[r + 1][0], 1, kWave);
                    n_km[0] = __shfl_up(cn[r + 1][V - 1], 1, kWave);
                    n_kp[V - 1] = __shfl_down(cn[r + 1][0], 1, kWave);
#pragma unroll
                    for (int v = 1; v < V; ++v) {
                        p_km[v] = cp[r + 1][v - 1];
                        p_kp[v - 1] = cp[r + 1][v];
                        n_km[v] = cn[r + 1][v - 1];
                        n_kp[v - 1] = cn[r + 1][v];
                    }
                } else {
                    p_km[0] = __shfl_up(cp[r + 1][0], 1, kWave);
                    p_kp[0] = __shfl_down(cp[r + 1][0], 1, kWave);
                    n_km[0] = __shfl_up(cn[r + 1][0], 1, kWave);
                    n_kp[0] = __shfl_down(cn[r + 1][0], 1, kWave);
                }
                if (live && r < nrows) {
                    const int j = j0 + r;
                    const double dj = (double)j - a.mid;
                    const bool inside = di2 + dj * dj < a.rsq;  // wave-uniform
                    S po[V], no[V];
                    double rec[V];
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        const S im_p = bp[PB][r + 1][v], ip_p = bp[XB][r + 1][v];
                        const S im_n = bn[PB][r + 1][v], ip_n = bn[XB][r + 1][v];
                        if constexpr (sizeof(S) == 8) {
                            // the reference's expression, term by term
                            const A p = (A)cp[r + 1][v], n = (A)cn[r + 1][v];
                            const A p_new = (A)a.p_zm * (A)p_km[v] + (A)a.p_zp * (A)p_kp[v] +
                                            (A)a.p_lat * (A)cp[r][v] + (A)a.p_lat * (A)cp[r + 2][v] +
                                            (A)a.p_lat * (A)im_p + (A)a.p_lat * (A)ip_p + (A)a.p_cen * p;
                            const A n_new = (A)a.n_zm * (A)n_kp[v] + (A)a.n_zp * (A)n_km[v] +
                                            (A)a.n_lat * (A)cn[r][v] + (A)a.n_lat * (A)cn[r + 2][v] +
                                            (A)a.n_lat * (A)im_n + (A)a.n_lat * (A)ip_n + (A)a.n_cen * n;
                            rec[v] = a.alpha_dt * (double)p * (double)n;
                            po[v] = (S)((double)p_new - rec[v]);
                            no[v] = (S)((double)n_new - rec[v]);
                        } else {
                            // FP32 fields: the flux form (fp32_update)
                            const float p = cp[r + 1][v], n = cn[r + 1][v];
                            fp32_update(p, n, p_km[v], p_kp[v], n_km[v], n_kp[v], cp[r][v], cp[r + 2][v], im_p,
                                        ip_p, cn[r][v], cn[r + 2][v], im_n, ip_n, f_pzm, f_pzp, f_plat, f_nzm, f_nzp,
                                        f_nlat, a.alpha_dt, po[v], no[v], rec[v]);
                        }
                    }
                    if (inside) {
                        if constexpr (kMerged) {
                            score4(k_score, rec, po, no, sr[0], sp[0], sn[0]);
                        } else {
#pragma unroll
                            for (int v = 0; v < V; ++v) {
                                sr[v] += rec[v];
                                sp[v] += (double)po[v];
                                sn[v] += (double)no[v];
                            }
                        }
                    }
                    store2(i, j, po, no);
                }
            }
        };
        if constexpr (CH > 0) {
            static_assert(CH <= NB, "a short march is one group of plane buffers");
            static_for<CH>([&](auto Oc) { plane_step(i0 + decltype(Oc)::value, Oc); });
        } else {
            for (int i = i0; i < i1; i += NB) {
                static_for<NB>([&](auto Oc) { plane_step(i + decltype(Oc)::value, Oc); });
            }
        }
        if constexpr (kMerged) {
            acc_r += sr[0];
            acc_p += sp[0];
            acc_n += sn[0];
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                acc_r += k_score[v] ? sr[v] : 0.0;
                acc_p += k_score[v] ? sp[v] : 0.0;
                acc_n += k_score[v] ? sn[v] : 0.0;
            }
        }
    }
    block_sum3(acc_r, acc_p, acc_n);
    if (threadIdx.x == 0) {
        a.partials[3 * (int64_t)logical + 0] = acc_r;
        a.partials[3 * (int64_t)logical + 1] = acc_p;
        a.partials[3 * (int64_t)logical + 2] = acc_n;
    }
}

// Fold the previous step's partials into its record (a block of its own, so
// the fold's loads and its reduction run beside the march instead of ahead of
// one of the marching blocks).
__device__ __forceinline__ void fold_previous(const SweepArgs &a) {
    double out[4];
    block_final4(a.prev_part, a.prev_n, a.prev_dep_part, a.prev_n_dep, out);
    if (threadIdx.x == 0)
        for (int v = 0; v < 4; ++v) a.prev_out4[v] = out[v];
}

// One configuration, one step. Block 0 folds the previous step when one is
// pending (prev_out4 set), the other blocks sweep. Shifting the sweep blocks
// by one keeps each XCD's run of logical tiles on a single XCD.
template <typename S, typename A, int TJ, bool DEP, int V, bool REFLECT, int CH = 0>
__global__ void __launch_bounds__(512) sweep_kernel(SweepArgs a) {
    const int nf = a.prev_out4 ? 1 : 0;
    if ((int)blockIdx.x < nf) {
        fold_previous(a);
        return;
    }
    sweep_body<S, A, TJ, V, REFLECT, CH>(a, xcd_logical_block(blockIdx.x - nf, gridDim.x - nf), DEP);
}

// The one argument of a batched launch: per-configuration descriptors, the
// first block of each configuration (offs[c]; offs[n_cfg] = all blocks) and
// the number of configurations. A single by-value argument lets a launch be
// held as data and replayed as a graph node. (Carrying the offsets in the
// argument itself measured slower on MI355X than one vector load of this
// table: kernel arguments are fetched by dependent scalar loads.)
template <typename D>
struct BatchTable {
    const D *descs;
    const int *offs;
    int n_cfg;
};

// The configuration c of batched block g (offs[c] <= g < offs[c+1]) and its
// first block offs[c]: c = the number of offs[1..n_cfg-1] <= g, one load per
// lane instead of a chain of dependent scalar loads.
template <typename D>
__device__ __forceinline__ void find_config(const BatchTable<D> &t, int g, int &c, int &first) {
    const int lane = threadIdx.x & 63;
    int cc = 0;
    for (int b = 1; b < t.n_cfg; b += kWave) {
        const int q = b + lane;
        cc += __popcll(__ballot(q < t.n_cfg && t.offs[q] <= g));
    }
    c = __builtin_amdgcn_readfirstlane(cc);
    first = t.offs[c];
}

// Many configurations, one step each, in one launch. Blocks [0, n_cfg) fold
// configuration b's previous step (if pending); sweep block g (after the XCD
// remap) belongs to configuration c with offs[c] <= g < offs[c+1] and runs it
// with that configuration's arguments descs[c] -- its arrays, weights,
// tiling and this step's deposition. Every configuration takes the
// deposition path (one compiled path keeps the register count of the
// single-configuration kernel): one that deposits nothing this step points
// dens at zeros, and x + 0.0 == x. Configurations are ordered so the active
// ones are a prefix; the grid covers exactly those.
template <typename S, typename A, int TJ, int V, bool REFLECT, int CH = 0>
__global__ void __launch_bounds__(512) batch_sweep_kernel(BatchTable<SweepArgs> t) {
    const SweepArgs *__restrict__ descs = t.descs;
    const int n_cfg = t.n_cfg;
    if ((int)blockIdx.x < n_cfg) {
        const SweepArgs a = descs[blockIdx.x];
        if (a.prev_out4) fold_previous(a);
        return;
    }
    const int g = xcd_logical_block(blockIdx.x - n_cfg, gridDim.x - n_cfg);
    int c, first;
    find_config(t, g, c, first);
    const SweepArgs a = descs[c];
    sweep_body<S, A, TJ, V, REFLECT, CH>(a, g - first, true);
}

// Tracks binned by gather tile, for one launch's steps (built on the device
// by bin_count_kernel, bin_scan_kernel and bin_fill_kernel, in that order on
// one stream): per (step y of the launch, tile) the number of tracks whose
// stencil box meets the tile, `cnt`, and where their entries start among the
// step's, `off`; per step its number of entries, `tot`; the entries are the
// tracks' indices within their step, in `list` (each bin in no particular
// order: the gather sorts it).
struct BinView {
    const int *cnt;   // [steps][tiles]
    const int *off;   // [steps][tiles]
    const int *tot;   // [steps]
    const int *list;  // entries of step y start at sum(tot[0..y))
};

struct GatherArgs {
    const double *xs;
    const double *ys;
    const int64_t *offsets;  // device step offsets into xs/ys (n_steps + 1)
    int64_t first_step;      // step of blockIdx.y == 0
    int64_t n_slots;         // step s uses slot s % n_slots of dens and partials
    double *dens;            // [n_slots][no_xy^2], gf-scaled summed density
    double *partials;        // [n_slots][gridDim.x], scored injected charge
    int no_xy, no_z, ntj;
    double cutoff, h2b2, gf, mid, rsq;
    // derivative variant only: per step y = blockIdx.y, the column densities'
    // derivatives with respect to gf and h2b2 ([gridDim.y][no_xy^2] each) and
    // the per-block scored injected-charge derivatives ([gridDim.y][gridDim.x][2])
    double *dgf, *dh, *dpart;
    // binned tracks of this launch's steps (gather_kernel only), or all null:
    // then every block scans all of its step's tracks
    BinView bins;
};

constexpr int kGT = 16;  // gather tile edge
// Tracks whose factors are evaluated per LDS batch. The factor table takes
// kGB x 32 doubles of LDS, which sets how many gather blocks a CU holds: 128
// (32 KB, three 256-thread blocks per CU) against 32 (8 KB, seven waves per
// SIMD) measured 0.143 against 0.042 ms of deposition per step on the 5 um
// full electrode, whose binned tiles list ~30 tracks each; the order of the
// sums, and so every bit, is the same.
constexpr int kGB = 32;
constexpr int kBinCap = 2048;  // largest bin a block sorts in LDS; larger bins fall back to the scan

__device__ __forceinline__ void track_box(double x, double y, double c, double nx, int &li, int &hi, int &lj,
                                          int &hj) {
    li = (int)fmax(ceil(x - c), 0.0);
    hi = (int)fmin(floor(x + c) + 1.0, nx);
    lj = (int)fmax(ceil(y - c), 0.0);
    hj = (int)fmin(floor(y + c) + 1.0, nx);
}

// The gather tiles a track's stencil box meets, [ti0, ti1] x [tj0, tj1];
// false for an empty box (which adds nothing anywhere). A tile meets the box
// exactly when the scan's test in gather_body says so.
__device__ __forceinline__ bool track_tiles(double x, double y, double c, double nx, int &ti0, int &ti1, int &tj0,
                                            int &tj1) {
    int li, hi, lj, hj;
    track_box(x, y, c, nx, li, hi, lj, hj);
    if (hi <= li || hj <= lj) return false;
    ti0 = li / kGT;
    ti1 = (hi - 1) / kGT;
    tj0 = lj / kGT;
    tj1 = (hj - 1) / kGT;
    return true;
}

// Sum of tot[0..y) (y <= a launch's steps, at most a few hundred): the first
// entry of step y in the bin list. Every thread of the block returns it.
__device__ __forceinline__ int bin_step_base(const int *tot, int y) {
    __shared__ int s_base;
    if (threadIdx.x < kWave) {
        int v = 0;
        for (int q = threadIdx.x; q < y; q += kWave) v += tot[q];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
        if (threadIdx.x == 0) s_base = v;
    }
    __syncthreads();
    return s_base;
}

// Deposition of a block of steps as a gather. Block (x, y) owns a 16x16 tile
// of columns for step first_step + y; each thread owns one (i, j) column and
// sums, in ascending track order, gi(t)*gj(t) over the tracks whose stencil
// box covers it -- the same terms in the same order as the reference's CSR
// row accumulation (solver_numba_parallel.py:183-219), so no atomics and a
// deterministic result.
//
// The block's list of tracks -- those whose box meets the tile, ascending --
// comes from one of two places, with the same result:
//  - binned (a.bins set, bin of at most kBinCap entries): the tile's bin,
//    sorted in LDS (bitonic), so the block touches only the ~tens of tracks
//    that reach it;
//  - scan: the block tests every track of its step against the tile and
//    compacts the hits in arrival order (ballot + prefix count), 256 at a
//    time.
// It then evaluates each listed track's 16 row and 16 column factors once
// into LDS, kGB tracks at a time (two barriers per batch; with 32 tracks,
// four independent exp evaluations per thread). A factor whose row or column lies outside the
// track's box is stored as zero, so the sum needs no box test: a skipped
// term and an added +0.0 leave the (non-negative) sum bit-identical.
// Output: dens = sum * gaussian_factor, and per block the scored injected
// charge sum(dens * no_z) over interior columns inside the scoring disc
// (solver_numba_parallel.py:250-277). Steps without tracks are skipped.
//
// DERIV (adjoint only) also forms, per column, dd/dgf = sum gi gj and
// dd/dh2b2 = -gf sum gi gj (di^2 + dj^2), di = i - x, the factors being
// exp(-di^2 h2b2) and exp(-dj^2 h2b2).
template <bool DERIV>
__device__ __forceinline__ void gather_body(const GatherArgs &a, const int tile, const int ystep) {
    __shared__ double s_x[256], s_y[256];  // centres of listed tracks
    __shared__ int s_wcnt[4];
    __shared__ double s_g[kGB][2 * kGT];
    __shared__ double s_d2[DERIV ? kGB : 1][2 * kGT];  // squared offsets (DERIV)
    __shared__ int s_list[DERIV ? 1 : kBinCap];        // the tile's bin, sorted
    const int n_tiles = a.ntj * a.ntj;
    const int64_t step = a.first_step + ystep;
    const int64_t t_begin = a.offsets[step];
    const int64_t m = a.offsets[step + 1] - t_begin;
    if (!DERIV && m <= 0) return;  // (the derivative variant writes its zeros)
    const double *xs = a.xs + t_begin;
    const double *ys = a.ys + t_begin;
    const int64_t slot = step % a.n_slots;
    double *dens = a.dens + slot * a.no_xy * a.no_xy;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ti0 = (tile / a.ntj) * kGT, tj0 = (tile % a.ntj) * kGT;
    const int ty = tid / kGT, tx = tid % kGT;
    const int i = ti0 + ty, j = tj0 + tx;
    const double nx = (double)a.no_xy;
    double acc = 0.0, acc2 = 0.0;
    // factors of the nq listed centres at s_x/s_y[at ...] into s_g, then the
    // column sums; two barriers
    auto add_listed = [&](int at, int nq) {
        for (int e = tid; e < nq * 2 * kGT; e += 256) {
            const int q = e / (2 * kGT), f = e % (2 * kGT);
            const double x = s_x[at + q], y = s_y[at + q];
            int li, hi, lj, hj;
            track_box(x, y, a.cutoff, nx, li, hi, lj, hj);
            const int cell = f < kGT ? ti0 + f : tj0 + f - kGT;
            const bool in_box = f < kGT ? (cell >= li && cell < hi) : (cell >= lj && cell < hj);
            const double d = f < kGT ? (double)cell - x : (double)cell - y;
            s_g[q][f] = in_box ? exp(-(d * d) * a.h2b2) : 0.0;
            if (DERIV) s_d2[q][f] = d * d;
        }
        __syncthreads();
        for (int q2 = 0; q2 < nq; ++q2) acc += s_g[q2][ty] * s_g[q2][kGT + tx];
        if (DERIV)
            for (int q2 = 0; q2 < nq; ++q2)
                acc2 += (s_g[q2][ty] * s_g[q2][kGT + tx]) * (s_d2[q2][ty] + s_d2[q2][kGT + tx]);
        __syncthreads();
    };
    bool binned = false;
    if constexpr (!DERIV) {
        if (a.bins.list) {
            const int cell = ystep * n_tiles + tile;
            const int nb = a.bins.cnt[cell];
            if (nb <= kBinCap) {
                binned = true;
                const int base = bin_step_base(a.bins.tot, ystep) + a.bins.off[cell];
                // sort ascending: the bin was filled in no particular order
                if (nb <= 256) {
                    // one entry per thread, placed at its rank (the entries are
                    // distinct track indices): two barriers
                    const int v = tid < nb ? a.bins.list[base + tid] : 0;
                    if (tid < nb) s_list[kBinCap - 256 + tid] = v;
                    __syncthreads();
                    if (tid < nb) {
                        int rank = 0;
                        for (int u = 0; u < nb; ++u) rank += s_list[kBinCap - 256 + u] < v;
                        s_list[rank] = v;
                    }
                    __syncthreads();
                } else {
                    int p2 = 2;
                    while (p2 < nb) p2 <<= 1;
                    for (int e = tid; e < p2; e += 256) s_list[e] = e < nb ? a.bins.list[base + e] : INT_MAX;
                    __syncthreads();
                    // bitonic
                    for (int k = 2; k <= p2; k <<= 1) {
                        for (int h = k >> 1; h > 0; h >>= 1) {
                            for (int e = tid; e < p2; e += 256) {
                                const int o = e ^ h;
                                if (o > e) {
                                    const int u = s_list[e], w = s_list[o];
                                    if ((u > w) == ((e & k) == 0)) {
                                        s_list[e] = w;
                                        s_list[o] = u;
                                    }
                                }
                            }
                            __syncthreads();
                        }
                    }
                }
                for (int sb = 0; sb < nb; sb += kGB) {
                    const int nq = min(kGB, nb - sb);
                    if (tid < nq) {
                        const int t = s_list[sb + tid];
                        s_x[tid] = xs[t];
                        s_y[tid] = ys[t];
                    }
                    __syncthreads();
                    add_listed(0, nq);
                }
            }
        }
    }
    if (!binned) {
        for (int64_t base = 0; base < m; base += 256) {
            const int64_t t = base + tid;
            bool hit = false;
            const double tx_c = xs[min(t, m - 1)], ty_c = ys[min(t, m - 1)];
            if (t < m) {
                int li, hi, lj, hj;
                track_box(tx_c, ty_c, a.cutoff, nx, li, hi, lj, hj);
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
            // the list keeps the tracks' arrival order, so every voxel still adds
            // them in ascending order
            if (hit) {
                s_x[off + before] = tx_c;
                s_y[off + before] = ty_c;
            }
            __syncthreads();
            for (int sb = 0; sb < nlist; sb += kGB) add_listed(sb, min(kGB, nlist - sb));
        }
    }
    double inj = 0.0, z1 = 0.0, z2 = 0.0;
    if (!DERIV && i < a.no_xy && j < a.no_xy) {
        const double d = acc * a.gf;
        dens[(int64_t)i * a.no_xy + j] = d;
        const double di = (double)i - a.mid, dj = (double)j - a.mid;
        if (acc != 0.0 && i >= 1 && i <= a.no_xy - 2 && j >= 1 && j <= a.no_xy - 2 && di * di + dj * dj < a.rsq)
            inj = d * (double)a.no_z;
    }
    if constexpr (DERIV) {
        double e_gf = 0.0, e_h = 0.0, z3 = 0.0;
        if (i < a.no_xy && j < a.no_xy) {
            const int64_t col = (int64_t)ystep * a.no_xy * a.no_xy + (int64_t)i * a.no_xy + j;
            const double dh = -a.gf * acc2;
            a.dgf[col] = acc;
            a.dh[col] = dh;
            const double di = (double)i - a.mid, dj = (double)j - a.mid;
            if (i >= 1 && i <= a.no_xy - 2 && j >= 1 && j <= a.no_xy - 2 && di * di + dj * dj < a.rsq) {
                e_gf = acc * (double)a.no_z;
                e_h = dh * (double)a.no_z;
            }
        }
        block_sum3(e_gf, e_h, z3);
        if (tid == 0) {
            a.dpart[2 * ((int64_t)ystep * n_tiles + tile) + 0] = e_gf;
            a.dpart[2 * ((int64_t)ystep * n_tiles + tile) + 1] = e_h;
        }
    } else {
        block_sum3(inj, z1, z2);
        if (tid == 0) a.partials[slot * n_tiles + tile] = inj;
    }
}

__global__ void __launch_bounds__(256) gather_kernel(GatherArgs a) { gather_body<false>(a, blockIdx.x, blockIdx.y); }

// The three passes that bin one launch's tracks by gather tile (BinView).
struct BinArgs {
    const double *xs, *ys;
    const int64_t *offsets;  // device step offsets
    int64_t first_step;      // step of blockIdx.y == 0
    int *cnt, *off, *cur, *tot, *list;
    int64_t n_cnt;           // counters bin_clear_kernel zeroes (steps x tiles)
    int no_xy, ntj;
    double cutoff;
};

// Pass 0: zero the counters of the launch's steps (a kernel rather than a
// memset, so the pass can be replayed in a graph).
__global__ void __launch_bounds__(256) bin_clear_kernel(BinArgs a) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < a.n_cnt; e += (int64_t)gridDim.x * blockDim.x)
        a.cnt[e] = 0;
}

// Pass 1, grid (any, steps): count the tracks meeting each tile. cnt starts
// at zero (the host clears it, with a kernel, before the pass).
__global__ void __launch_bounds__(256) bin_count_kernel(BinArgs a) {
    const int y = blockIdx.y, nt = a.ntj * a.ntj;
    const int64_t t0 = a.offsets[a.first_step + y], m = a.offsets[a.first_step + y + 1] - t0;
    const double nx = (double)a.no_xy;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m; t += (int64_t)gridDim.x * blockDim.x) {
        int ti0, ti1, tj0, tj1;
        if (!track_tiles(a.xs[t0 + t], a.ys[t0 + t], a.cutoff, nx, ti0, ti1, tj0, tj1)) continue;
        for (int ti = ti0; ti <= ti1; ++ti)
            for (int tj = tj0; tj <= tj1; ++tj) atomicAdd(&a.cnt[(int64_t)y * nt + ti * a.ntj + tj], 1);
    }
}

// Pass 2, one block of 256 per step: exclusive scan of the step's counts into
// off, the step's total into tot, the fill cursors cleared.
__global__ void __launch_bounds__(256) bin_scan_kernel(BinArgs a) {
    __shared__ int s_wsum[4];
    const int y = blockIdx.x, nt = a.ntj * a.ntj;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t row = (int64_t)y * nt;
    int carry = 0;
    for (int b = 0; b < nt; b += 256) {
        const int e = b + threadIdx.x;
        const int v = e < nt ? a.cnt[row + e] : 0;
        int incl = v;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int u = __shfl_up(incl, o, kWave);
            if (lane >= o) incl += u;
        }
        if (lane == kWave - 1) s_wsum[wave] = incl;
        __syncthreads();
        int before = carry;
        for (int w = 0; w < wave; ++w) before += s_wsum[w];
        if (e < nt) {
            a.off[row + e] = before + incl - v;
            a.cur[row + e] = 0;
        }
        carry += s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.tot[y] = carry;
}

// Pass 3, grid as pass 1: each track's index (within its step) into every
// bin it meets, at a slot taken from the bin's cursor.
__global__ void __launch_bounds__(256) bin_fill_kernel(BinArgs a) {
    const int y = blockIdx.y, nt = a.ntj * a.ntj;
    const int base = bin_step_base(a.tot, y);
    const int64_t t0 = a.offsets[a.first_step + y], m = a.offsets[a.first_step + y + 1] - t0;
    const double nx = (double)a.no_xy;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m; t += (int64_t)gridDim.x * blockDim.x) {
        int ti0, ti1, tj0, tj1;
        if (!track_tiles(a.xs[t0 + t], a.ys[t0 + t], a.cutoff, nx, ti0, ti1, tj0, tj1)) continue;
        for (int ti = ti0; ti <= ti1; ++ti)
            for (int tj = tj0; tj <= tj1; ++tj) {
                const int64_t cell = (int64_t)y * nt + ti * a.ntj + tj;
                const int slot = atomicAdd(&a.cur[cell], 1);
                a.list[base + a.off[cell] + slot] = (int)t;
            }
    }
}

// The gathers of many configurations in one launch: block g belongs to
// configuration c with offs[c] <= g < offs[c+1] and runs tile
// (g - offs[c]) % tiles of step (g - offs[c]) / tiles of that
// configuration's arguments descs[c].
__global__ void __launch_bounds__(256) batch_gather_kernel(BatchTable<GatherArgs> t) {
    int c, first;
    find_config(t, (int)blockIdx.x, c, first);
    const GatherArgs a = t.descs[c];
    const int local = (int)blockIdx.x - first;
    const int tiles = a.ntj * a.ntj;
    gather_body<false>(a, local % tiles, local / tiles);
}

// Explicit deposition into the current arrays (kernel-level API): add the
// step's 2D density to every gap layer (solver_numba_parallel.py:253-262).
template <typename S>
__global__ void __launch_bounds__(256) broadcast_kernel(S *pos, S *neg, const double *dens, int no_xy, int pitch,
                                                        int nze, int no_z) {
    const int64_t n = (int64_t)no_xy * no_xy * no_z;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int kz = (int)(e % no_z);
        const int64_t col = e / no_z;
        const double d = dens[col];
        if (d == 0.0) continue;
        const int64_t idx = col * pitch + nze + kz;
        pos[idx] = (S)((double)pos[idx] + d);
        neg[idx] = (S)((double)neg[idx] + d);
    }
}

// Per-step scalars from the block partials of one step, in the same fixed
// order and compensated summation as the fold the next sweep performs:
// out4 = (injected, recombined, total+, total-). One block.
__global__ void __launch_bounds__(256) reduce_kernel(const double *sweep_part, int n_sweep, const double *dep_part,
                                                     int n_dep, double *out4) {
    double out[4];
    block_final4(sweep_part, n_sweep, dep_part, n_dep, out);
    if (threadIdx.x == 0)
        for (int v = 0; v < 4; ++v) out4[v] = out[v];
}

// Reflecting wall, for read-back only: write the lateral ring of an array
// pair as the copy of the nearest interior column (state.py:113-154); whole
// rows are copied, row padding (zero) included.
template <typename S>
__global__ void __launch_bounds__(256) ring_mirror_kernel(S *pos, S *neg, int no_xy, int pitch) {
    const int64_t face = (int64_t)no_xy * pitch;        // one i plane
    const int64_t side = (int64_t)(no_xy - 2) * pitch;  // j = 0 or no_xy-1, i = 1..no_xy-2
    const int64_t n = 2 * face + 2 * side;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        if (e < 2 * face) {
            i = e < face ? 0 : no_xy - 1;
            const int64_t r = e % face;
            j = (int)(r / pitch);
            k = (int)(r % pitch);
        } else {
            const int64_t r = e - 2 * face;
            j = r < side ? 0 : no_xy - 1;
            const int64_t q = r % side;
            i = 1 + (int)(q / pitch);
            k = (int)(q % pitch);
        }
        const int ci = min(max(i, 1), no_xy - 2), cj = min(max(j, 1), no_xy - 2);
        const int64_t dst = ((int64_t)i * no_xy + j) * pitch + k;
        const int64_t src = ((int64_t)ci * no_xy + cj) * pitch + k;
        pos[dst] = pos[src];
        neg[dst] = neg[src];
    }
}

// The last node of every graph chain: stores the chain's launch number to
// pinned host memory once every node before it has run, so the host can tell
// a finished chain from one in flight by reading memory, with no runtime call.
// The store is relaxed: the host needs to know that the nodes have finished,
// not to see their data, and a release here would write back the whole L2.
struct ChainMark {
    unsigned long long *flag;
    unsigned long long seq;
};
__global__ void __launch_bounds__(64) chain_mark_kernel(ChainMark m) {
    if (threadIdx.x == 0) __hip_atomic_store(m.flag, m.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <typename S>
__global__ void __launch_bounds__(256) fill_kernel(S *p, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        p[e] = (S)0;
}

}  // namespace pic
