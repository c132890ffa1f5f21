// Two time steps per pass over FP32 fields, either lateral wall:
// the field traffic of one step for two. Included once, by pic_hip.hip, after
// pic_kernels.hpp.
//
// A block is 16 waves, one row each: rows j0-2 .. j0+13 of one k window of 64
// lanes x 4 values, marching through a chunk of i planes. Per plane it
//   1. shares its rows of the step-t field (deposition of step t+1 added)
//      through LDS and computes step t+1 (T1) on rows j0-1 .. j0+12;
//   2. shares T1 (deposition of step t+2 added) through LDS and computes step
//      t+2 (T2) on the 12 output rows j0 .. j0+11, which it stores.
// The i neighbours stay in registers (the march), the k neighbours are lane
// shuffles. T1 is computed on one halo row each side and on lanes 1..62, so
// T2 is right on lanes 2..61 (windows of 240 values; the first window also
// on lanes 0..1, whose left neighbour is the zero k = 0 plane); per output
// value the pass does 14/12 evaluations of the first step and one of the
// second: 1.08 times the arithmetic of two single steps.
//
// Every voxel update is the single-step sweep's FP32 arithmetic
// (fp32_update), so the fields equal two single steps bit for bit. The
// lateral ring follows the single-step kernel's rules. Reflecting: a ring
// voxel reads its nearest interior column plus the deposition at the ring
// position -- for T1 too: the ring row's T1 entry in LDS is written by the
// interior row next to it, the ring plane's from the raw T1 of the plane
// next to it. Absorbing: a ring voxel is carried, its value plus each step's
// deposition (T1 = step-t value + both depositions, which is also step t+2);
// the ring row's own wave writes its T1 entry, the block holding row 1 or
// no_xy-2 stores the ring row, the chunk holding plane 1 or no_xy-2 the ring
// plane.
#pragma once

namespace pic {

constexpr int kS2Rows = 12;               // output rows per block
constexpr int kS2Waves = kS2Rows + 4;     // one row per wave: j0-2 .. j0+13
constexpr int kS2Step = 4 * (kWave - 4);  // k values between windows (lanes 2..61 of four values)

// The fold of one step's partials into its record (block_final4), as
// reduce_kernel does; pointers first, so two fit a launch record beside the
// pass's own arguments.
struct FoldArgs {
    const double *part;
    const double *dep;
    double *out4;
    int n;
    int n_dep;
};
struct Fold2Args {
    FoldArgs step[2];
};

__device__ __forceinline__ void fold_one(const FoldArgs &f) {
    double out[4];
    block_final4(f.part, f.n, f.dep, f.n_dep, out);
    if (threadIdx.x == 0)
        for (int v = 0; v < 4; ++v) f.out4[v] = out[v];
}

// the folds of a pass's two steps (one block each), for the end of a run
__global__ void __launch_bounds__(256) fold2_kernel(Fold2Args a) { fold_one(a.step[blockIdx.x]); }

struct Sweep2Args {
    const float *pos;
    const float *neg;
    float *pos_out;
    float *neg_out;
    const double *dens1;  // deposition density of step t+1 (gf-scaled, 2D; zeros without tracks)
    const double *dens2;  // of step t+2
    double *part1;        // [blocks][3] scored partials of step t+1: recombined, total+, total-
    double *part2;        // of step t+2
    int no_xy, nzb, nze, no_z;
    int pitch;
    int n_windows, n_jg, n_ic;
    int chunk;  // planes per block
    double mid, rsq;
    double p_lat, p_zm, p_zp;
    double n_lat, n_zm, n_zp;
    double alpha_dt;
    // the steps of the previous launch still to fold into their records, one
    // block each ahead of the sweep blocks (0, 1 or 2)
    FoldArgs prev[2];
    int n_prev;
    int reflect;  // lateral wall: 1 reflecting, 0 absorbing
};

// number of k windows of the two-step pass over nzb values per row
__host__ __device__ inline int sweep2_windows(int nzb) { return nzb <= 248 ? 1 : 1 + (nzb - 248 + kS2Step - 1) / kS2Step; }

__global__ void __launch_bounds__(1024) sweep2_kernel(Sweep2Args a) {
    // rows of one plane shared by the block, double-buffered by plane parity:
    // the step-t field (with step t+1's deposition) and T1 (with step t+2's)
    __shared__ f32x4 tx[2][kS2Waves][2][kWave];
    __shared__ f32x4 ux[2][kS2Waves][2][kWave];

    if ((int)blockIdx.x < a.n_prev) {
        fold_one(a.prev[blockIdx.x]);
        return;
    }
    const int logical = xcd_logical_block(blockIdx.x - a.n_prev, gridDim.x - a.n_prev);
    const int kw = logical % a.n_windows;
    const int rest = logical / a.n_windows;
    const int jg = rest % a.n_jg;
    const int ic = rest / a.n_jg;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int no_xy = a.no_xy, nzb = a.nzb, pitch = a.pitch;
    const int j0 = 1 + kS2Rows * jg;
    const int r = j0 - 2 + w;                  // this wave's row; may lie past the ring
    const int rl = min(max(r, 0), no_xy - 1);  // the row whose values it loads
    const bool interior_row = r >= 1 && r <= no_xy - 2;
    const bool t1_row = w >= 1 && w <= kS2Waves - 2 && interior_row;
    const bool out_row = w >= 2 && w <= kS2Waves - 3 && interior_row;
    const bool reflect = a.reflect != 0;
    // absorbing wall: a ring row's wave carries its values through both steps
    // (t-level value + both depositions); the block whose output rows reach
    // the ring stores them
    const bool ring_row = (r == 0 || r == no_xy - 1) && w >= 1 && w <= kS2Waves - 2;
    const bool ring_store = !reflect && ring_row && ((r == 0 && jg == 0) || (r == no_xy - 1 && j0 + kS2Rows - 1 >= no_xy - 2));
    const int i0 = 1 + ic * a.chunk;
    const int i1 = min(i0 + a.chunk, no_xy - 1);
    const int k0 = kS2Step * kw + 4 * lane;

    bool k_in[4], k_gap[4], k_score[4], k_out[4];
    {
        const bool own = kw == 0 ? lane <= kWave - 3 : lane >= 2 && lane <= kWave - 3;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int k = k0 + v;
            k_in[v] = k >= 1 && k <= nzb - 2;
            k_gap[v] = k >= a.nze && k < a.nze + a.no_z;
            k_out[v] = own && k_in[v];
            k_score[v] = k_out[v] && k_gap[v];
        }
    }
    const float f_pzm = (float)a.p_zm, f_pzp = (float)a.p_zp, f_plat = (float)a.p_lat;
    const float f_nzm = (float)a.n_zm, f_nzp = (float)a.n_zp, f_nlat = (float)a.n_lat;
    const int64_t plane = (int64_t)no_xy * pitch;
    const int plane_bytes = (int)(plane * 4);
    const int kc = k0 < nzb ? k0 : pitch - 4;
    const unsigned vo_load = (unsigned)kc * 4u, vo_store = (unsigned)k0 * 4u;
    auto rsrc = [&](const void *arr, int i) {
        return __builtin_amdgcn_make_buffer_rsrc(
            const_cast<void *>(static_cast<const void *>(static_cast<const float *>(arr) + (int64_t)i * plane)), 0,
            plane_bytes + kFieldSlack, kBufferWord3);
    };
    auto row_off = [&](int j) -> int { return (int)((int64_t)j * pitch * 4); };
    auto clamp_plane = [&](int i) { return min(max(i, 0), no_xy - 1); };

    // step-t values of plane i, this wave's row: the nearest interior column
    // (reflecting wall), the deposition of step t+1 at (i, rl) added later
    auto load_t = [&](int i, f32x4 &vp, f32x4 &vn) {
        const int si = reflect ? min(max(i, 1), no_xy - 2) : min(max(i, 0), no_xy - 1);
        const int sj = reflect ? min(max(rl, 1), no_xy - 2) : rl;
        vp = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rsrc(a.pos, si), vo_load, row_off(sj), 0));
        vn = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rsrc(a.neg, si), vo_load, row_off(sj), 0));
    };
    auto dens_at = [&](const double *d, int i, int j) -> double { return d[(int64_t)clamp_plane(i) * no_xy + j]; };
    auto add_gap = [&](f32x4 &vp, f32x4 &vn, double d) {
        const float df = (float)d;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float dv = k_gap[v] ? df : 0.0f;
            vp[v] = vp[v] + dv;
            vn[v] = vn[v] + dv;
        }
    };

    // One step of the sweep's FP32 update on this wave's row: centre c, rows
    // j-1 / j+1 (jm, jp), planes i-1 / i+1 (im, ip). Writes the new values
    // and the recombination of each value.
    auto update = [&](const f32x4 &cp, const f32x4 &cn, const f32x4 &jmp, const f32x4 &jmn, const f32x4 &jpp,
                      const f32x4 &jpn, const f32x4 &imp, const f32x4 &imn, const f32x4 &ipp, const f32x4 &ipn,
                      f32x4 &op, f32x4 &on, double (&rec)[4]) {
        float p_km[4], p_kp[4], n_km[4], n_kp[4];
        p_km[0] = __shfl_up(cp[3], 1, kWave);
        p_kp[3] = __shfl_down(cp[0], 1, kWave);
        n_km[0] = __shfl_up(cn[3], 1, kWave);
        n_kp[3] = __shfl_down(cn[0], 1, kWave);
#pragma unroll
        for (int v = 1; v < 4; ++v) {
            p_km[v] = cp[v - 1];
            p_kp[v - 1] = cp[v];
            n_km[v] = cn[v - 1];
            n_kp[v - 1] = cn[v];
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            float po, no;
            fp32_update(cp[v], cn[v], p_km[v], p_kp[v], n_km[v], n_kp[v], jmp[v], jpp[v], imp[v], ipp[v], jmn[v],
                        jpn[v], imn[v], ipn[v], f_pzm, f_pzp, f_plat, f_nzm, f_nzp, f_nlat, a.alpha_dt, po, no, rec[v]);
            op[v] = po;
            on[v] = no;
        }
    };
    auto inside = [&](int i) {
        const double di = (double)i - a.mid, dj = (double)r - a.mid;
        return di * di + dj * dj < a.rsq;
    };
    double s1r = 0.0, s1p = 0.0, s1n = 0.0, s2r = 0.0, s2p = 0.0, s2n = 0.0;
    auto score = [&](const f32x4 &op, const f32x4 &on, const double (&rec)[4], double &sr, double &sp, double &sn) {
        const float po[4] = {op[0], op[1], op[2], op[3]}, no[4] = {on[0], on[1], on[2], on[3]};
        score4(k_score, rec, po, no, sr, sp, sn);
    };

    // step-t planes in four slots (plane q in slot (q - i0 + 2) & 3); T1 of
    // planes p-1, p, p+1 (deposition of step t+2 added); the raw T1 of the
    // latest plane, for a ring plane's mirror
    f32x4 tp[4], tn[4];
    f32x4 u_prev_p, u_prev_n, u_cur_p, u_cur_n, u_next_p, u_next_n, raw_p, raw_n;

    // T1 of plane q on this wave's row (t-level planes q-1, q, q+1 in slots
    // A, B, C; rows j-/+1 of plane q from LDS buffer b): into (up, un) with
    // step t+2's deposition, raw into (raw_p, raw_n), the ring rows' T1 entries
    // and this row's entry into ux[b]; scored for step t+1 when `scored`.
    auto t1_plane = [&](int q, const f32x4 &ap, const f32x4 &an, const f32x4 &bp, const f32x4 &bn, const f32x4 &cp_,
                        const f32x4 &cn_, int b, bool scored, f32x4 &up, f32x4 &un) {
        if (t1_row) {
            const f32x4 jmp = tx[b][w - 1][0][lane], jmn = tx[b][w - 1][1][lane];
            const f32x4 jpp = tx[b][w + 1][0][lane], jpn = tx[b][w + 1][1][lane];
            f32x4 op, on;
            double rec[4];
            update(bp, bn, jmp, jmn, jpp, jpn, ap, an, cp_, cn_, op, on, rec);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                op[v] = k_in[v] ? op[v] : 0.0f;  // the k end planes and row padding stay zero
                on[v] = k_in[v] ? on[v] : 0.0f;
            }
            if (scored && out_row && inside(q)) score(op, on, rec, s1r, s1p, s1n);
            raw_p = op;
            raw_n = on;
            up = op;
            un = on;
            add_gap(up, un, dens_at(a.dens2, q, r));
            ux[b][w][0][lane] = up;
            ux[b][w][1][lane] = un;
            if (reflect && (r == 1 || r == no_xy - 2)) {
                // the ring row next to this one: its T1 is this row's, plus the
                // deposition at the ring position
                const int ring = r == 1 ? 0 : no_xy - 1;
                const int wr = r == 1 ? w - 1 : w + 1;
                f32x4 mp = op, mn = on;
                add_gap(mp, mn, dens_at(a.dens2, q, ring));
                ux[b][wr][0][lane] = mp;
                ux[b][wr][1][lane] = mn;
            }
        } else if (!reflect && ring_row) {
            // absorbing: the ring row is carried, not swept
            up = bp;
            un = bn;
            add_gap(up, un, dens_at(a.dens2, q, r));
            ux[b][w][0][lane] = up;
            ux[b][w][1][lane] = un;
        }
    };
    // T1 of a ring plane (0 or no_xy-1), plus the deposition at the ring
    // plane: reflecting, the raw T1 of the plane next to it; absorbing, the
    // plane's own step-t value (deposition of t+1 added), carried
    auto t1_ring_plane = [&](int q, const f32x4 &tq_p, const f32x4 &tq_n, f32x4 &up, f32x4 &un) {
        up = reflect ? raw_p : tq_p;
        un = reflect ? raw_n : tq_n;
        add_gap(up, un, dens_at(a.dens2, q, rl));
    };
    // store the values of this wave's row at plane p (own lanes of the window)
    auto store_row = [&](int p, const f32x4 &op, const f32x4 &on) {
        if (k_out[0] || k_out[1] || k_out[2] || k_out[3]) {
            f32x4 wp, wn;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                wp[v] = k_out[v] ? op[v] : 0.0f;
                wn[v] = k_out[v] ? on[v] : 0.0f;
            }
            const auto rp = rsrc(a.pos_out, p), rn = rsrc(a.neg_out, p);
            const int so = row_off(r);
            if (k0 + 4 <= pitch) {
                __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, wp), rp, vo_store, so, 0);
                __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, wn), rn, vo_store, so, 0);
            } else {
                const f32x2 hp = {wp[0], wp[1]}, hn = {wn[0], wn[1]};
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, hp), rp, vo_store, so, 0);
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, hn), rn, vo_store, so, 0);
            }
        }
    };
    // T2 of output plane p from T1 in registers and LDS buffer b; stored
    auto t2_plane = [&](int p, int b) {
        if (out_row) {
            const f32x4 jmp = ux[b][w - 1][0][lane], jmn = ux[b][w - 1][1][lane];
            const f32x4 jpp = ux[b][w + 1][0][lane], jpn = ux[b][w + 1][1][lane];
            f32x4 op, on;
            double rec[4];
            update(u_cur_p, u_cur_n, jmp, jmn, jpp, jpn, u_prev_p, u_prev_n, u_next_p, u_next_n, op, on, rec);
            if (inside(p)) score(op, on, rec, s2r, s2p, s2n);
            store_row(p, op, on);
        } else if (ring_store) {
            store_row(p, u_cur_p, u_cur_n);  // the carried ring row at step t+2
        }
    };

    // prologue: step-t planes i0-2 .. i0+1, T1 of planes i0-1 and i0
    static_for<4>([&](auto Qc) {
        constexpr int q = decltype(Qc)::value;
        load_t(i0 - 2 + q, tp[q], tn[q]);
    });
    static_for<4>([&](auto Qc) {
        constexpr int q = decltype(Qc)::value;
        add_gap(tp[q], tn[q], dens_at(a.dens1, i0 - 2 + q, rl));
    });
    tx[(i0 - 1) & 1][w][0][lane] = tp[1];
    tx[(i0 - 1) & 1][w][1][lane] = tn[1];
    __syncthreads();
    if (i0 - 1 >= 1) t1_plane(i0 - 1, tp[0], tn[0], tp[1], tn[1], tp[2], tn[2], (i0 - 1) & 1, false, u_cur_p, u_cur_n);
    tx[i0 & 1][w][0][lane] = tp[2];
    tx[i0 & 1][w][1][lane] = tn[2];
    load_t(min(i0 + 2, no_xy), tp[0], tn[0]);  // plane i0+2 into the slot of i0-2
    __syncthreads();
    u_prev_p = u_cur_p;
    u_prev_n = u_cur_n;
    t1_plane(i0, tp[1], tn[1], tp[2], tn[2], tp[3], tn[3], i0 & 1, true, u_cur_p, u_cur_n);
    if (i0 - 1 == 0 && (t1_row || (!reflect && ring_row))) {
        t1_ring_plane(0, tp[1], tn[1], u_prev_p, u_prev_n);
        if (!reflect && (out_row || ring_store)) store_row(0, u_prev_p, u_prev_n);  // the carried ring plane
    }

    // the march: output plane p, step-t planes p .. p+3 in slots o+2 .. o+5 (mod 4)
    auto step = [&](int p, auto Oc) {
        constexpr int o = decltype(Oc)::value;
        constexpr int S0 = (o + 2) & 3, S1 = (o + 3) & 3, S2 = o & 3, S3 = (o + 1) & 3;
        if (p >= i1) return;  // block-uniform: every wave skips the same steps
        add_gap(tp[S2], tn[S2], dens_at(a.dens1, p + 2, rl));        // plane p+2, loaded a step ago
        tx[(p + 1) & 1][w][0][lane] = tp[S1];
        tx[(p + 1) & 1][w][1][lane] = tn[S1];
        load_t(min(p + 3, no_xy), tp[S3], tn[S3]);                   // plane p+3 into the slot of p-1
        __syncthreads();
        if (p + 1 <= no_xy - 2)
            t1_plane(p + 1, tp[S0], tn[S0], tp[S1], tn[S1], tp[S2], tn[S2], (p + 1) & 1, p + 1 < i1, u_next_p, u_next_n);
        else if (t1_row || (!reflect && ring_row)) {
            t1_ring_plane(p + 1, tp[S1], tn[S1], u_next_p, u_next_n);
            if (!reflect && (out_row || ring_store)) store_row(p + 1, u_next_p, u_next_n);  // the carried ring plane
        }
        t2_plane(p, p & 1);
        u_prev_p = u_cur_p;
        u_prev_n = u_cur_n;
        u_cur_p = u_next_p;
        u_cur_n = u_next_n;
    };
    for (int p = i0; p < i1; p += 4) static_for<4>([&](auto Oc) { step(p + decltype(Oc)::value, Oc); });

    double r1 = s1r, p1 = s1p, n1 = s1n;
    block_sum3(r1, p1, n1);
    if (threadIdx.x == 0) {
        a.part1[3 * (int64_t)logical + 0] = r1;
        a.part1[3 * (int64_t)logical + 1] = p1;
        a.part1[3 * (int64_t)logical + 2] = n1;
    }
    __syncthreads();
    double r2 = s2r, p2 = s2p, n2 = s2n;
    block_sum3(r2, p2, n2);
    if (threadIdx.x == 0) {
        a.part2[3 * (int64_t)logical + 0] = r2;
        a.part2[3 * (int64_t)logical + 1] = p2;
        a.part2[3 * (int64_t)logical + 2] = n2;
    }
}

}  // namespace pic
