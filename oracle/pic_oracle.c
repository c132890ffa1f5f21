/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A C/OpenMP restatement of the reference's batched multi-core backend,
 * pulsed_ion_chamber/solver_numba_parallel.py, used as
 *   - the checker the HIP backend is compared against (tests, smoke), and
 *   - the timed CPU baseline in bench.py.
 * Nothing in the shipped library links or calls this file.
 *
 * What is restated, and from where:
 *   orc_gaussian_factors  _precompute_track_gaussians   solver_numba_parallel.py:421-457
 *   orc_row_index         _build_row_index              solver_numba_parallel.py:140-180
 *   orc_accumulate        _accumulate_track_density...  solver_numba_parallel.py:183-219
 *   orc_broadcast         _broadcast_density...         solver_numba_parallel.py:222-277
 *   orc_lw_step           _lax_wendroff_step...         solver_numba_parallel.py:280-418
 *   orc_swap_boundary     swap + carry / mirror         solver_numba_parallel.py:686-695,
 *                                                       state.py:113-176
 *   orc_run               the time-step loop            solver_numba_parallel.py:600-707
 *
 * Arithmetic is FP64 throughout and must be built with -ffp-contract=off so
 * no a*b+c is fused: the reference's Numba kernels are compiled without
 * fast-math, so each product and sum is rounded separately, and keeping that
 * makes a single sweep bit-identical to the reference's.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int64_t no_xy, nzb, nze, no_z;
    double mid_xy, scoring_radius_sq;
    double p_lat, p_zm, p_zp, p_cen;
    double n_lat, n_zm, n_zp, n_cen;
    double alpha_dt;
    double gaussian_factor, h2_over_b2, cutoff_voxels;
    int32_t reflecting;
    int32_t pad;
} orc_params;

#define AT(a, i, j, k) (a)[((i) * no_xy + (j)) * nzb + (k)]

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* solver_numba_parallel.py:280-418. Returns recombined, total+, total- over the
 * scored region (disc and gap layers nze <= k < nze+no_z). */
void orc_lw_step(const orc_params *q, const double *pos, const double *neg,
                 double *pos_next, double *neg_next, double out3[3]) {
    const int64_t no_xy = q->no_xy, nzb = q->nzb, nze = q->nze, no_z = q->no_z;
    const int64_t n_inner = no_xy - 2;
    const double p_lat = q->p_lat, p_zm = q->p_zm, p_zp = q->p_zp, p_cen = q->p_cen;
    const double n_lat = q->n_lat, n_zm = q->n_zm, n_zp = q->n_zp, n_cen = q->n_cen;
    const double alpha_dt = q->alpha_dt, mid = q->mid_xy, rsq = q->scoring_radius_sq;
    double rec = 0.0, tp = 0.0, tn = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rec, tp, tn)
    for (int64_t idx = 0; idx < n_inner * n_inner; ++idx) {
        const int64_t i = 1 + idx / n_inner;
        const int64_t j = 1 + idx % n_inner;
        const double di = (double)i - mid, dj = (double)j - mid;
        const int inside = di * di + dj * dj < rsq;
        for (int64_t k = 1; k < nzb - 1; ++k) {
            const double p = AT(pos, i, j, k);
            const double n = AT(neg, i, j, k);
            const double p_new = p_zm * AT(pos, i, j, k - 1) + p_zp * AT(pos, i, j, k + 1) +
                                 p_lat * AT(pos, i, j - 1, k) + p_lat * AT(pos, i, j + 1, k) +
                                 p_lat * AT(pos, i - 1, j, k) + p_lat * AT(pos, i + 1, j, k) + p_cen * p;
            const double n_new = n_zm * AT(neg, i, j, k + 1) + n_zp * AT(neg, i, j, k - 1) +
                                 n_lat * AT(neg, i, j - 1, k) + n_lat * AT(neg, i, j + 1, k) +
                                 n_lat * AT(neg, i - 1, j, k) + n_lat * AT(neg, i + 1, j, k) + n_cen * n;
            const double r = alpha_dt * p * n;
            const double po = p_new - r, no = n_new - r;
            AT(pos_next, i, j, k) = po;
            AT(neg_next, i, j, k) = no;
            if (inside && nze <= k && k < no_z + nze) {
                rec += r;
                tp += po;
                tn += no;
            }
        }
    }
    out3[0] = rec;
    out3[1] = tp;
    out3[2] = tn;
}

/* solver_numba_parallel.py:445-457: stencil box and 1D Gaussian factors per
 * track. gauss_* are (m, width) with width = max box extent. */
int64_t orc_box_width(const orc_params *q, const double *xs, const double *ys, int64_t m) {
    int64_t width = 1;
    for (int64_t t = 0; t < m; ++t) {
        const double c = q->cutoff_voxels, nx = (double)q->no_xy;
        const int64_t li = (int64_t)fmax(ceil(xs[t] - c), 0.0), hi = (int64_t)fmin(floor(xs[t] + c) + 1.0, nx);
        const int64_t lj = (int64_t)fmax(ceil(ys[t] - c), 0.0), hj = (int64_t)fmin(floor(ys[t] + c) + 1.0, nx);
        if (hi - li > width) width = hi - li;
        if (hj - lj > width) width = hj - lj;
    }
    return width;
}

static void gaussian_factors(const orc_params *q, const double *xs, const double *ys, int64_t m,
                             int64_t width, double *gi, double *gj, int64_t *lo_i, int64_t *hi_i,
                             int64_t *lo_j, int64_t *hi_j) {
    const double c = q->cutoff_voxels, nx = (double)q->no_xy, s = q->h2_over_b2;
    for (int64_t t = 0; t < m; ++t) {
        lo_i[t] = (int64_t)fmax(ceil(xs[t] - c), 0.0);
        hi_i[t] = (int64_t)fmin(floor(xs[t] + c) + 1.0, nx);
        lo_j[t] = (int64_t)fmax(ceil(ys[t] - c), 0.0);
        hi_j[t] = (int64_t)fmin(floor(ys[t] + c) + 1.0, nx);
        for (int64_t o = 0; o < width; ++o) {
            const double di = (double)(lo_i[t] + o) - xs[t];
            const double dj = (double)(lo_j[t] + o) - ys[t];
            gi[t * width + o] = exp(-(di * di) * s);
            gj[t * width + o] = exp(-(dj * dj) * s);
        }
    }
}

/* Batched deposition of one step's m tracks into pos/neg (restates
 * solver_numba_parallel.py:140-277): 2D sum per row in ascending track order,
 * then one broadcast down the gap layers. Returns the scored injected charge.
 * total (no_xy^2) is caller scratch. */
double orc_deposit(const orc_params *q, double *pos, double *neg, const double *xs, const double *ys,
                   int64_t m, double *total) {
    const int64_t no_xy = q->no_xy, nzb = q->nzb, nze = q->nze, no_z = q->no_z;
    if (m <= 0) return 0.0;
    const int64_t width = orc_box_width(q, xs, ys, m);
    double *gi = (double *)malloc(sizeof(double) * m * width);
    double *gj = (double *)malloc(sizeof(double) * m * width);
    int64_t *lo_i = (int64_t *)malloc(sizeof(int64_t) * m), *hi_i = (int64_t *)malloc(sizeof(int64_t) * m);
    int64_t *lo_j = (int64_t *)malloc(sizeof(int64_t) * m), *hi_j = (int64_t *)malloc(sizeof(int64_t) * m);
    gaussian_factors(q, xs, ys, m, width, gi, gj, lo_i, hi_i, lo_j, hi_j);

    /* CSR row index, counting sort (solver_numba_parallel.py:140-180) */
    int64_t *offsets = (int64_t *)calloc(no_xy + 1, sizeof(int64_t));
    for (int64_t t = 0; t < m; ++t)
        for (int64_t i = lo_i[t]; i < hi_i[t]; ++i) offsets[i + 1] += 1;
    for (int64_t i = 0; i < no_xy; ++i) offsets[i + 1] += offsets[i];
    int64_t *cursor = (int64_t *)malloc(sizeof(int64_t) * (no_xy + 1));
    memcpy(cursor, offsets, sizeof(int64_t) * no_xy);
    int64_t *ids = (int64_t *)malloc(sizeof(int64_t) * (offsets[no_xy] > 0 ? offsets[no_xy] : 1));
    for (int64_t t = 0; t < m; ++t)
        for (int64_t i = lo_i[t]; i < hi_i[t]; ++i) ids[cursor[i]++] = t;

    memset(total, 0, sizeof(double) * no_xy * no_xy);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < no_xy; ++i) {
        for (int64_t p = offsets[i]; p < offsets[i + 1]; ++p) {
            const int64_t t = ids[p];
            const double g = gi[t * width + (i - lo_i[t])];
            for (int64_t j = lo_j[t]; j < hi_j[t]; ++j) total[i * no_xy + j] += g * gj[t * width + (j - lo_j[t])];
        }
    }

    const double gf = q->gaussian_factor, mid = q->mid_xy, rsq = q->scoring_radius_sq;
    double inserted = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : inserted)
    for (int64_t idx = 0; idx < no_xy * no_xy; ++idx) {
        const int64_t i = idx / no_xy, j = idx % no_xy;
        double d = total[idx];
        if (d == 0.0) continue;
        d *= gf;
        for (int64_t k = nze; k < nze + no_z; ++k) {
            AT(pos, i, j, k) += d;
            AT(neg, i, j, k) += d;
        }
        const double di = (double)i - mid, dj = (double)j - mid;
        if (1 <= i && i <= no_xy - 2 && 1 <= j && j <= no_xy - 2 && di * di + dj * dj < rsq) inserted += d * (double)no_z;
    }
    free(gi); free(gj); free(lo_i); free(hi_i); free(lo_j); free(hi_j);
    free(offsets); free(cursor); free(ids);
    return inserted;
}

/* state.py:113-154 (reflecting mirror, i planes then j planes) */
static void mirror_ring(const orc_params *q, double *a) {
    const int64_t no_xy = q->no_xy, nzb = q->nzb, plane = no_xy * nzb;
    memcpy(a, a + plane, sizeof(double) * plane);
    memcpy(a + (no_xy - 1) * plane, a + (no_xy - 2) * plane, sizeof(double) * plane);
    for (int64_t i = 0; i < no_xy; ++i) {
        memcpy(&AT(a, i, 0, 0), &AT(a, i, 1, 0), sizeof(double) * nzb);
        memcpy(&AT(a, i, no_xy - 1, 0), &AT(a, i, no_xy - 2, 0), sizeof(double) * nzb);
    }
}

/* state.py:157-176 (absorbing: carry the four ring planes src -> dst) */
static void carry_ring(const orc_params *q, double *dst, const double *src) {
    const int64_t no_xy = q->no_xy, nzb = q->nzb, plane = no_xy * nzb;
    memcpy(dst, src, sizeof(double) * plane);
    memcpy(dst + (no_xy - 1) * plane, src + (no_xy - 1) * plane, sizeof(double) * plane);
    for (int64_t i = 0; i < no_xy; ++i) {
        memcpy(&AT(dst, i, 0, 0), &AT(src, i, 0, 0), sizeof(double) * nzb);
        memcpy(&AT(dst, i, no_xy - 1, 0), &AT(src, i, no_xy - 1, 0), sizeof(double) * nzb);
    }
}

static void first_touch(const orc_params *q, double *a) {
    const int64_t no_xy = q->no_xy, nzb = q->nzb;
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < no_xy * no_xy; ++idx) memset(a + idx * nzb, 0, sizeof(double) * nzb);
}

/* The time-step loop (solver_numba_parallel.py:600-707). offsets has
 * n_total_steps+1 entries; the tracks of step s are xs[offsets[s]:offsets[s+1]].
 * Runs steps [0, n_steps). per_step is (n_steps, 4): injected, recombined,
 * total+, total-. pos_out/neg_out (may be NULL) receive the final current
 * arrays. Returns 0, or -1 if memory could not be allocated. */
static double g_loop_seconds = 0.0;

static double wall_seconds(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

int orc_run(const orc_params *q, const double *xs, const double *ys, const int64_t *offsets,
            int64_t n_steps, double *per_step, double *pos_out, double *neg_out, int nthreads) {
    set_threads(nthreads);
    const int64_t n = q->no_xy * q->no_xy * q->nzb;
    double *a[4];
    for (int s = 0; s < 4; ++s) {
        a[s] = (double *)malloc(sizeof(double) * n);
        if (!a[s]) {
            for (int u = 0; u < s; ++u) free(a[u]);
            return -1;
        }
        first_touch(q, a[s]);
    }
    double *total = (double *)calloc(q->no_xy * q->no_xy, sizeof(double));
    double *pos = a[0], *neg = a[1], *pos_next = a[2], *neg_next = a[3];
    const double t0 = wall_seconds();
    for (int64_t step = 0; step < n_steps; ++step) {
        const int64_t m = offsets[step + 1] - offsets[step];
        double injected = 0.0;
        if (m > 0) injected = orc_deposit(q, pos, neg, xs + offsets[step], ys + offsets[step], m, total);
        double out3[3];
        orc_lw_step(q, pos, neg, pos_next, neg_next, out3);
        per_step[step * 4 + 0] = injected;
        per_step[step * 4 + 1] = out3[0];
        per_step[step * 4 + 2] = out3[1];
        per_step[step * 4 + 3] = out3[2];
        double *tmp;
        if (q->reflecting) {
            tmp = pos; pos = pos_next; pos_next = tmp;
            tmp = neg; neg = neg_next; neg_next = tmp;
            mirror_ring(q, pos);
            mirror_ring(q, neg);
        } else {
            carry_ring(q, pos_next, pos);
            carry_ring(q, neg_next, neg);
            tmp = pos; pos = pos_next; pos_next = tmp;
            tmp = neg; neg = neg_next; neg_next = tmp;
        }
    }
    g_loop_seconds = wall_seconds() - t0;
    if (pos_out) memcpy(pos_out, pos, sizeof(double) * n);
    if (neg_out) memcpy(neg_out, neg, sizeof(double) * n);
    for (int s = 0; s < 4; ++s) free(a[s]);
    free(total);
    return 0;
}

/* Wall time of the last orc_run's step loop: allocation and first touch of
 * the four arrays and the final copy-out excluded. */
double orc_last_loop_seconds(void) { return g_loop_seconds; }

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
