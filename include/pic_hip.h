/*
 * pic_hip.h -- C ABI of the MI355X (gfx950) time-stepping backend for the
 * pulsed ion chamber: Gaussian track deposition, the Lax-Wendroff drift +
 * diffusion + recombination sweep of both carrier species, and the scored
 * charge reduction.
 *
 * Plain C types only: opaque handles, plain pointers, sizes. Every entry point
 * returns 0 on success or a negative PIC_E_* code; pic_last_error() returns
 * the message of the last failure on the calling thread. Host pointers are
 * copied during the call and never retained. One handle owns one HIP stream
 * and its device arrays; a handle must not be used from two threads at once.
 *
 * Grid layout (same as the reference): each species is a C-contiguous
 * (no_xy, no_xy, nzb) array, k (the field axis) fastest. Fields are stored in
 * FP64 (PIC_FP64) or FP32 (PIC_FP32); every scored sum, the recombination
 * product and the deposited density are FP64 in both modes.
 *
 * The reference binds its backends as Python functions; the entry points
 * below replace the pieces of those functions named in each comment
 * (paths relative to the reference checkout).
 */
#ifndef PIC_HIP_H
#define PIC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIC_OK 0
#define PIC_E_INVALID -1  /* bad argument or state */
#define PIC_E_NOMEM -2    /* device allocation or size check failed */
#define PIC_E_HIP -3      /* HIP runtime error */
#define PIC_E_NODEVICE -4 /* no usable GPU */

#define PIC_FP64 0
#define PIC_FP32 1

#define PIC_WALL_ABSORBING 0
#define PIC_WALL_REFLECTING 1

#define PIC_GATHER_AUTO 0   /* bin tracks by tile on grids of >= 256 gather tiles */
#define PIC_GATHER_SCAN 1   /* every tile tests every track of its step */
#define PIC_GATHER_BINNED 2 /* tracks binned by tile first, on any grid */

typedef struct pic_sim pic_sim;
typedef struct pic_batch pic_batch;
typedef struct pic_adjoint pic_adjoint;

/* Everything the loop reads from SimulationConfig (config.py:180-401,
 * scheme_coefficients config.py:466-482, alpha_dt solver_numba.py:273-274). */
typedef struct pic_params {
    int64_t no_xy;            /* lateral voxels per side */
    int64_t nzb;              /* no_z_with_buffer: voxels along the field axis */
    int64_t nze;              /* no_z_electrode: first gap layer */
    int64_t no_z;             /* gap layers */
    double mid_xy;            /* grid centre, no_xy/2.0 */
    double scoring_radius_sq; /* (i-mid)^2 + (j-mid)^2 < this is scored */
    double p_lat, p_zm, p_zp, p_cen; /* positive-ion stencil weights */
    double n_lat, n_zm, n_zp, n_cen; /* negative-ion stencil weights */
    double alpha_dt;          /* recombination coefficient times dt */
    double gaussian_factor;   /* N0 / (pi b^2) */
    double h2_over_b2;        /* voxel area over squared track radius */
    double cutoff_voxels;     /* deposition box half width */
    int32_t wall;             /* PIC_WALL_* */
    int32_t reserved;         /* must be 0 */
} pic_params;

/* Message of the last failed call on this thread (always NUL-terminated). */
int pic_last_error(char *buf, size_t len);

/* Device name and free/total memory of the current device
 * (replaces the device checks of solver_cuda.py:141-172, 589-642). */
int pic_device_info(char *name, size_t name_len, size_t *free_bytes, size_t *total_bytes);

/* Device bytes a handle needs for these params and precision, before any
 * tracks are uploaded (carrier arrays solver_cuda.py:420-423 plus scratch). */
int pic_mem_required(const pic_params *p, int precision, size_t *bytes);

/* Allocate a zeroed state on the current device (solver_cuda.py:742-772). */
int pic_create(const pic_params *p, int precision, pic_sim **out);
int pic_destroy(pic_sim *sim);

/* Upload every track of a run once: xs/ys of n_tracks centres in arrival
 * order, and n_steps+1 offsets so that step s deposits
 * xs[offsets[s]:offsets[s+1]] (solver_cuda.py:791-816 does this per step).
 * Replaces any earlier upload; per-step results are reset to zero. */
int pic_upload_tracks(pic_sim *sim, const double *xs, const double *ys, int64_t n_tracks,
                      const int64_t *offsets, int64_t n_steps);

/* Run time steps [first, first+count) of the uploaded schedule: per step,
 * gather-deposit that step's tracks, sweep, wall update, reduce, swap
 * (the loop body of solver_cuda.py:791-865). Asynchronous on the handle's
 * stream; results land in a device (n_steps, 4) array. */
int pic_run_steps(pic_sim *sim, int64_t first, int64_t count);

/* Copy per-step (injected, recombined, total+, total-) rows [first,
 * first+count) to host (n x 4 doubles). Synchronises the stream. */
int pic_read_per_step(pic_sim *sim, int64_t first, int64_t count, double *out);

/* A whole run (or one pulse and its clearance) in one call: upload the track
 * centres of n_steps steps (step s has tracks [step_offsets[s],
 * step_offsets[s+1]) of xs/ys), run steps [0, n_steps) from the handle's
 * current fields, and copy the (n_steps, 4) per-step record into
 * per_step_out. Synchronous. The run-level form of the time-step loop of
 * solver_cuda.py:791-865 with its per-step track uploads
 * (solver_cuda.py:794-797) done once; equal to pic_upload_tracks +
 * pic_run_steps + pic_read_per_step. */
int pic_run_pulse(pic_sim *sim, const double *xs, const double *ys, const int64_t *step_offsets, int64_t n_steps,
                  double *per_step_out);

/* pic_run_pulse without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc). It has no counterpart in the reference. Same
 * arguments, results and errors as pic_run_pulse, bit for bit, except:
 *  - the device copies of the offsets and track centres and the device
 *    per-step record are staged in a static device array of 64 MiB that the
 *    library holds on every GPU it runs on, so a call fails with
 *    PIC_E_INVALID when 32*n_steps + 8*(n_steps+1) + 16*n_tracks bytes
 *    (each region rounded up to 256) exceed 64 MiB -- about four million
 *    tracks;
 *  - the handle's uploaded tracks (pic_upload_tracks) are left as they were;
 *  - launches are direct: graph replay and timing do not apply.
 * The density slots and partials are the handle's own, allocated by
 * pic_create. Every call on one GPU shares that GPU's array: calls that
 * overlap in time on the same GPU (two threads, or two handles) race. */
int pic_run_pulse_static(pic_sim *sim, const double *xs, const double *ys, const int64_t *step_offsets,
                         int64_t n_steps, double *per_step_out);

/* Kernel-level calls (the granularity tests/test_charge_conservation.py
 * drives _insert_track_numba / _lax_wendroff_step_numba at):
 * deposit m tracks into the current arrays and return the scored injected
 * charge (solver_numba_parallel.py:140-277 semantics, ascending track order
 * per voxel) ... */
int pic_deposit(pic_sim *sim, const double *xs, const double *ys, int64_t m, double *inserted);
/* ... and sweep once with the wall update and buffer swap, returning
 * (recombined, total+, total-) (solver_numba_parallel.py:280-418, 686-695). */
int pic_step(pic_sim *sim, double out3[3]);

/* Current carrier arrays to host as FP64, no_xy*no_xy*nzb each (either may be
 * NULL). Synchronises. (solver_cuda.py:882-887) */
int pic_read_fields(pic_sim *sim, double *pos, double *neg);
/* Overwrite the current (and next) arrays from host FP64 data (test hook). */
int pic_write_fields(pic_sim *sim, const double *pos, const double *neg);

/* Zero the four carrier arrays and the per-step record, keeping the uploaded
 * tracks: the state of a fresh handle (a rerun of the same schedule). */
int pic_reset(pic_sim *sim);

/* How the deposition finds the tracks that reach each 16x16 tile of columns
 * (the role of the CSR row index of solver_numba_parallel.py:140-180): every
 * tile testing every track of its step, or the tracks binned by tile on the
 * device first (count, scan, fill), each tile then sorting and reading only
 * its own bin. Both add the same terms in the same ascending track order, so
 * the results are bit-identical; binning is faster on large grids. The mode
 * applies to the uploaded tracks (set before or after pic_upload_tracks);
 * get_gather_mode reports whether bins are in use. */
int pic_set_gather_mode(pic_sim *sim, int mode);
int pic_get_gather_mode(pic_sim *sim, int *binned);

/* Block until all work queued on the handle's stream is done. */
int pic_synchronize(pic_sim *sim);

/* Sweep launch shape, for tuning: rows per tile (1, 2, 3, 4, 8), planes per
 * block chunk, and k values per lane (1 or 2 for FP64 fields, two with rows
 * per tile 1 or 2 only; 1, 2 or 4 for FP32 fields, four with rows per tile 1,
 * 2 or 3 only; three rows with four FP32 values only, eight rows with one
 * FP32 value only); 0 = automatic. Takes effect at the next launch. */
int pic_set_tiling(pic_sim *sim, int rows_per_tile, int planes_per_chunk, int values_per_lane);

/* Consecutive j tiles per sweep block (0 = automatic, at most 8 waves per
 * block): neighbouring tiles then run on one CU side by side. */
int pic_set_tile_group(pic_sim *sim, int tiles_per_block);

/* Time steps per sweep launch: 1, or 2 for FP32 fields, either lateral
 * wall (pairs of steps in one pass over the fields, bit-identical to single
 * steps; a run's odd step and a gather segment's odd step take one); 0 = the
 * library's choice. planes_per_chunk (0 = keep) sets the two-step pass's
 * planes per block. No counterpart in the reference, whose backends take one
 * step per sweep (solver_cuda.py:834-863); a performance knob. */
int pic_set_steps_per_pass(pic_sim *sim, int steps, int planes_per_chunk);

/* The tiling the next sweep launch uses: the rule's choice for every knob
 * pic_set_tiling left at 0 (replaces reading back the threads per block the
 * reference's set_threads_per_block stores, solver_cuda.py:121-132). */
int pic_get_tiling(pic_sim *sim, int *rows_per_tile, int *planes_per_chunk, int *values_per_lane);

/* Kernel timing of the last pic_run_steps call when timing is enabled:
 * summed milliseconds of the sweep and of the deposition kernels (with their
 * launch counts) and of the ring + reduce kernels, measured with HIP events
 * recorded on the handle's stream. enable = 0 turns it off, 1 brackets each
 * launch, n > 1 brackets runs of up to n back-to-back sweep launches with one
 * pair (and each other launch): a pair costs the stream a few microseconds,
 * so grouping keeps a timed run as fast as an untimed one while every sweep
 * is still timed (its time then includes the ~1 us between launches). The
 * launch counts are those of the launches bracketed. */
int pic_enable_timing(pic_sim *sim, int enable);
int pic_get_timing(pic_sim *sim, double *sweep_ms, int64_t *sweep_launches, double *deposit_ms,
                   int64_t *deposit_launches, double *other_ms);

/* Graph replay of pic_run_steps (off by default): each gather batch's
 * launches are collected and run through graphs of 32 kernel nodes,
 * re-parameterised per use, instead of one launch at a time; the kernels and
 * their results are the same. It frees the host thread (about 1 us of host
 * time per kernel instead of 4); on MI355X under ROCm 7 the device runs the
 * replayed kernels slightly slower (+5% on small-grid runs, whose kernels take
 * ~7 us), so it is not the default. Timing (pic_enable_timing) launches one
 * at a time regardless. Replaces the per-step launch loop of
 * solver_cuda.py:791-865. */
int pic_set_graphs(pic_sim *sim, int enable);

/* Batched runs: many handles (a voltage / dose-rate / seed sweep with uneven
 * grids and step counts) advanced together, one sweep launch per time step
 * covering every handle still running. Replaces the reference's answer to a
 * parameter study -- one process per configuration, each running the loop of
 * solver_numba_parallel.py:526-723 (docs/PERFORMANCE.md:211-218) -- with one
 * device stream. The handles must share precision, wall type and device;
 * each keeps its own arrays, uploaded tracks and per-step record, which
 * pic_read_per_step / pic_read_fields read as after pic_run_steps. A handle
 * must not be used on its own while its batch runs; the batch does not own
 * the handles (destroy it first). */
int pic_batch_create(pic_sim *const *sims, int n, pic_batch **out);
int pic_batch_destroy(pic_batch *batch);
/* Sweep tiling of every step (rows per tile 1, 2 or 4 -- FP64: 1 or 2 -- and
 * planes per block chunk), or 0, 0 for the automatic choice, made per step
 * from the configurations still running; get_tiling reports the preferred
 * (first) tiling, the value layout and the waves per block. */
int pic_batch_set_tiling(pic_batch *batch, int rows_per_tile, int planes_per_chunk);
int pic_batch_get_tiling(pic_batch *batch, int *rows_per_tile, int *planes_per_chunk, int *values_per_lane,
                         int *windows_per_block);
/* Steps [first, first+count) of every handle, each handle stopping at its
 * own uploaded step count (the loop body of solver_cuda.py:791-865 for all of
 * them). Asynchronous on the batch's stream; the handles' own streams are
 * ordered after it. */
int pic_batch_run(pic_batch *batch, int64_t first, int64_t count);
int pic_batch_synchronize(pic_batch *batch);
/* Kernel timing of the last pic_batch_run (as pic_get_timing): summed
 * milliseconds and launch counts of the batched sweeps and of the gathers. */
int pic_batch_enable_timing(pic_batch *batch, int enable);
int pic_batch_get_timing(pic_batch *batch, double *sweep_ms, int64_t *sweep_launches, double *gather_ms,
                         int64_t *gather_launches);
/* Graph replay of pic_batch_run's segments (off by default), as pic_set_graphs. */
int pic_batch_set_graphs(pic_batch *batch, int enable);

/* Reverse-mode derivative of one FP64 handle's run (no counterpart in the
 * reference, which has no derivatives; it differentiates the loop the
 * reference runs, solver_numba.py:237-370). The forward pass keeps a
 * checkpoint of the state every checkpoint_every steps (capped at the
 * handle's gather batch); the reverse pass recomputes each segment from its
 * checkpoint and runs the adjoint steps backwards through it.
 * pic_adjoint_run resets the handle and runs the uploaded schedule:
 *   sums  = (sum of injected, sum of recombined) charge over the run,
 *   d_rec = d(sum recombined) / d(p_lat, p_zm, p_zp, p_cen, n_lat, n_zm, n_zp,
 *           n_cen, alpha_dt, gaussian_factor, h2_over_b2) -- the pic_params
 *           fields, the deposition box held fixed,
 *   d_inj = d(sum injected) / d(gaussian_factor, h2_over_b2).
 * k_s = sum_inj / (sum_inj - sum_rec) and its derivatives follow on the host.
 * Afterwards the handle's per-step record is the run's; its fields are not
 * the final state (the reverse pass reruns parts of the run in them). */
int pic_adjoint_create(pic_sim *sim, int64_t checkpoint_every, pic_adjoint **out);
/* As pic_adjoint_create, with the checkpoints in two levels: the forward pass
 * keeps the state at the start of every segments_per_outer-th segment only,
 * and the reverse pass reruns one such span of segments forward (keeping each
 * segment's start) before differentiating through it -- one more forward
 * pass, for ceil(segments / M) + M states held instead of one per segment.
 * 0 picks M = ceil(sqrt(segments)) when there are more than 64 segments and
 * 1 (one level, as pic_adjoint_create) otherwise. The derivatives do not
 * depend on the choice: the recomputed states are the forward's, bit for bit. */
int pic_adjoint_create_levels(pic_sim *sim, int64_t checkpoint_every, int64_t segments_per_outer,
                              pic_adjoint **out);
int pic_adjoint_destroy(pic_adjoint *adj);
int pic_adjoint_run(pic_adjoint *adj, double sums[2], double d_rec[11], double d_inj[2]);
/* Segment length in use, device bytes held, and the last run's forward and
 * reverse times (HIP events on the handle's stream). */
int pic_adjoint_get_info(pic_adjoint *adj, int64_t *checkpoint_every, size_t *device_bytes, double *forward_ms,
                         double *backward_ms);
/* Segments per outer checkpoint (M), outer checkpoints held, and segments. */
int pic_adjoint_get_levels(pic_adjoint *adj, int64_t *segments_per_outer, int64_t *outer_checkpoints,
                           int64_t *segments);

#ifdef __cplusplus
}
#endif
#endif /* PIC_HIP_H */
