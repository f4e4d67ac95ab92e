/*
 * sofi_hip.h - C-ABI of libsofi_hip.so: SOFI2D's velocity-stress time loop
 * on one MI355X (gfx950).
 *
 * The reference's replaceable unit is
 *     void time_loop(int iter, int ishot, int snapcheck, float *hc,
 *                    float **srcpos_loc, float **signalx, float **signaly,
 *                    int nsrc, int sw, AcqVar *acq, MemModel *mpm,
 *                    MemWavefield *mpw, MemInv *minv, GlobVar *gv,
 *                    GlobVarInv *vinv, Perform *perf)
 * (src/fd.h:335-337, src/time_loop.c:27-326), called per shot from
 * src/sofi2D.c:369 (forward) and src/inversion.c:222 (adjoint).  Its
 * arguments are split here into plain-C calls: the grid/scheme descriptor
 * (GlobVar scalars), the material arrays (MemModel), the sources and
 * receivers (AcqVar, srcpos_loc, signals), and the outputs (GlobVar
 * SECTION* buffers, snapshot files, MemInv gradients).  A handle owns the
 * device memory of one shot's state (MemWavefield) and one HIP stream; a
 * handle is not thread-safe, different handles may be driven from different
 * threads at once.  Every call returns 0 on success or a negative code,
 * with the message in sofi_last_error() (the reference aborts through
 * log_fatal, src/logging.c:148-210).
 *
 * Arrays crossing this boundary are host pointers in the padded layout of
 * sofi_types.h.
 */
#ifndef SOFI_HIP_H
#define SOFI_HIP_H

#include "sofi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sofi_dev sofi_dev;

/* Last error message of the calling thread. */
const char *sofi_last_error(void);

/* Allocate the device state for one grid/rheology; replaces initmem()
 * (src/initmem.c:29-170) and initfd() (src/initfd.c:363-371) for the path. */
int sofi_open(const sofi_desc *desc, sofi_dev **out);

/* A batch: nshots shot slots sharing the handle's medium, advanced together
 * by every sofi_run (the reference runs its shots one after the other,
 * src/sofi2D.c:320-395).  sofi_select_shot picks the slot that
 * sofi_set_sources / sofi_set_receivers / sofi_get_section / sofi_get_field
 * address (slot 0 after opening); sofi_reset clears every slot.  Snapshots
 * need one-slot handles; under MODE=FWI each slot keeps its own forward
 * store and gradient accumulators, and the residual and gradient calls
 * address the selected slot.  The update kernels run over groups of
 * `group` slots per launch and `spt` slots per thread (the CELL shape reads
 * a cell's material once for them); sofi_set_batch_launch fixes them (0:
 * measured on the first steps of the next run, the default; the results do
 * not depend on them).  No counterpart in the reference. */
int sofi_open_batch(const sofi_desc *desc, int nshots, sofi_dev **out);
int sofi_select_shot(sofi_dev *h, int k);
int sofi_set_batch_launch(sofi_dev *h, int group, int spt);
int sofi_get_batch_launch(const sofi_dev *h, int *group, int *spt);
/* Slot blocks of a batch: with slots > 0 the update kernels of every step
 * cover all slots of the (resident) group in one launch each, one thread per
 * cell for `slots` consecutive slots (1, 2 or 4), the cell's material
 * coefficients read once for them and every slot's stencil loaded before any
 * slot is updated; forces, free surface and sampling run as launches of
 * their own, and so do explosive sources beyond 8 per group (fewer ride in
 * the stress launch).  TTI stress updates then compute the velocity
 * derivatives in the kernel for the same slot blocks.  0 (the default) keeps
 * the slot-loop kernels of sofi_set_batch_launch.  Not for AC_ISO or 4th
 * order in time (the slot-loop kernels run).  The slot-block kernels are code of their
 * own: their results agree with the slot-loop kernels' to rounding, not
 * always bit for bit.  No counterpart in the reference. */
int sofi_set_batch_block(sofi_dev *h, int slots);
int sofi_get_batch_block(const sofi_dev *h);
/* Resident groups of a batch: sofi_run advances `slots` slots through its
 * whole step range before the next group (the groups that run at once on the
 * batch's streams in turns of 32 steps), so that a group's wavefields and
 * the shared medium stay in the 256 MB memory-side cache from one kernel to
 * the next (0, the default: the largest group such that the medium and the
 * groups that run at once on the batch's streams fit 240 MB, at least one
 * slot; all slots when the medium and one slot do not fit).  The results do
 * not depend on it.
 * sofi_get_batch_resident returns the group size the next run uses.  No
 * counterpart in the reference. */
int sofi_set_batch_resident(sofi_dev *h, int slots);
int sofi_get_batch_resident(const sofi_dev *h);
/* Streams a batch's resident groups run on at once when it has several
 * groups (0, the default: 2; at most 8): each group's steps go to one of
 * them, joined to the handle's stream before sofi_run returns.  The results
 * do not depend on it.  No counterpart in the reference. */
int sofi_set_batch_streams(sofi_dev *h, int streams);
/* The autotuner of a batch's launch (on by default): the first sofi_run of
 * at least 2 x 5 steps per candidate measures the candidates on 5 steps
 * each and keeps the fastest for the handle and for every later batch of
 * the same rheology, FD order, grid, slot count and boundaries in the
 * process.  The candidates are slot blocks of 1, 2 or 4 slots per thread
 * over all slots or over cache-sized resident groups on two streams; for
 * AC_ISO, which has no slot blocks, the slot-loop kernels with one slot per
 * launch in cache-sized groups on two streams, or all slots per launch.
 * They are one family of kernels whose results are the same bits for every
 * launch, so the choice changes no bit; a shorter sofi_run of a batch not
 * yet tuned runs the first candidate.  Setting sofi_set_batch_block, or a
 * non-zero value through sofi_set_batch_launch / _resident / _streams,
 * fixes the launch and turns the tuner off (the default launch then is the
 * slot-loop kernels).  sofi_get_batch_config returns
 * {slots per thread of the slot blocks (0: slot-loop kernels), resident
 * group, streams, slots per launch, slots per thread, tuned (0 / 1)}.  No
 * counterpart in the reference. */
int sofi_set_batch_tuning(sofi_dev *h, int on);
int sofi_get_batch_config(const sofi_dev *h, int *cfg);

/* Upload the material arrays (enum sofi_array, NULL where unused) and
 * profiles (enum sofi_profile); replaces the MemModel the stress and
 * velocity kernels read (src/memm_struct.h:28-126) after prepmod()
 * (src/prepmod.c:28-194) and PML_pro()/absorb(). */
int sofi_set_medium(sofi_dev *h, const float *const *arrays, const float *const *profiles);

/* Sources of the shot: grid indices (1-based), per-source type
 * (srcpos_loc[8]), azimuth/dip in degrees (srcpos_loc[7]), amplitude
 * (srcpos_loc[6]) and signals[nsrc][nt]; replaces srcpos_loc/signals of
 * time_loop (src/time_loop.c:27-29).  source_type is the SOURCE_TYPE that
 * selects psource/eqsource (src/time_loop.c:248-253). */
int sofi_set_sources(sofi_dev *h, int nsrc, const int *src_i, const int *src_j, const int *src_type,
                     const float *src_azi, const float *src_amp, const float *signals, int source_type);

/* Receivers (1-based grid indices); replaces acq->recpos_loc read by
 * seismo_ssg (src/seismo_ssg.c:28-67).  Call after sofi_set_medium. */
int sofi_set_receivers(sofi_dev *h, int ntr, const int *rec_i, const int *rec_j);

/* Reserve device space for up to nsnap snapshots of the shot (snap(),
 * src/snap_ssg.c:28-209). */
int sofi_set_snapshots(sofi_dev *h, int nsnap);

/* Zero the wavefield, memory variables, CPML memory, sections and the
 * snapshot counter: zero_wavefield() (src/zero_wavefield.c:28-130). */
int sofi_reset(sofi_dev *h);

/* Advance time steps nt_first..nt_last (1-based, inclusive) of the forward
 * loop; enqueues the work on the handle's stream and waits for it. */
int sofi_run(sofi_dev *h, int nt_first, int nt_last);

/* Shots on grids of uneven size advanced together: handle i through its
 * steps nt_first[i]..nt_last[i], n one-shot forward handles of one device,
 * FD order and rheology (EL_ISO, VEL_ISO, EL_VTI, VEL_VTI or AC_ISO; not 4th
 * order in time, not MODE=FWI), each with its own grid, medium, sources and
 * receivers.  Every step's velocity update of all handles still advancing is
 * one launch over a work-list of (handle, tile) items, and so is the stress
 * update; sampling, snapshots and the sources and free surface that do not
 * ride in the update kernels are launches per handle.  A block takes item b,
 * b + blocks, ... (a persistent grid when blocks is below the item count),
 * and an item covers `rows` tiles of 64 x `waves` cells (`waves` wavefronts
 * per block) down a tile column.  launch = {rows (1..64), blocks (0: one
 * block per item), waves (2, 4 or 8)} fixes the launch; NULL lets the first
 * call of a rheology, FD order, handle count and size measure waves 2, 4, 8
 * x rows 1, 2, 4 x blocks 0, 2048 on two steps each (after one untimed step;
 * the steps are part of the run) and keep the fastest in the process.  Each handle's results equal those of sofi_run with its update
 * kernels set to one thread per cell (sofi_set_launch), bit for bit, for any
 * launch, and the handle can go on with sofi_run afterwards.  The work runs
 * on the first handle's stream; the call waits for it and sets every
 * handle's sofi_last_run_ms to the whole run's time (sofi_set_kernel_timing
 * does not sample its steps).  No counterpart in the reference, which runs
 * one shot after another (src/sofi2D.c:325-395). */
int sofi_run_many(sofi_dev *const *h, int n, const int *nt_first, const int *nt_last, const int *launch);
/* The launch of this thread's last sofi_run_many and whether it came from
 * the measurement or the cache (1) or was fixed (0). */
int sofi_get_many_config(int *rows, int *blocks, int *waves, int *tuned);

/* Device time of the last sofi_run in milliseconds (HIP events on the
 * handle's stream). */
double sofi_last_run_ms(const sofi_dev *h);

/* Seismogram section of component comp (enum sofi_comp) as [ntr][ns]
 * floats; replaces gv->SECTIONVX..SECTIONDIV. */
int sofi_get_section(sofi_dev *h, int comp, float *out);

/* Snapshots taken so far and one of them: component comp, x-major
 * [nxs][nys] floats as written by snap() and merged by snapmerge. */
int sofi_snapshot_count(const sofi_dev *h);
int sofi_get_snapshot(sofi_dev *h, int k, int comp, float *out);

/* Copy a wavefield array (0 vx, 1 vy, 2 sxx, 3 syy, 4 sxy) in the padded
 * layout; mirrors reading mpw->pvx... after time_loop. */
int sofi_get_field(sofi_dev *h, int id, float *out);

/* Free the handle and its device memory. */
int sofi_close(sofi_dev *h);

/* Device bytes the handle holds (HBM footprint of one resident shot). */
size_t sofi_footprint(const sofi_dev *h);

/* The bytes a handle for desc with nshots slots, ntr receivers per slot and
 * nsnap snapshots will hold, without allocating (admission control of a
 * serving layer; no counterpart in the reference).  An upper bound of
 * sofi_footprint after the same calls, by at most 1 MiB of small buffers. */
size_t sofi_estimate_footprint(const sofi_desc *desc, int nshots, int ntr, int nsnap);

/* FWI gradient path (MODE=FWI; EL_ISO, VEL_ISO, EL_VTI, VEL_VTI, EL_TTI,
 * VEL_TTI; one-slot handles or batches): what time_loop does with sw = 0 / 1 and
 * MemInv (src/time_loop.c:27-326 with src/snap_store.c:27-47,
 * src/calc_conv.c:30-73), driven per shot as src/inversion.c:27-258 does.
 * The viscoelastic and anisotropic stress updates write the strain rates
 * ux / uy / uxy the store reads with VELOCITY = 1, as
 * src/wavefield_update_s_visc.c:60-66 and its siblings do.
 *   sofi_fwi_set_model    the iteration's model Vp0, Vs0, Rho0 (padded planes)
 *   sofi_fwi_set_pass     sw = 0: the next sofi_run is the forward pass and
 *                         stores the forward fields every DTINV steps;
 *                         sw = 1: the adjoint pass, correlating with them
 *   sofi_fwi_set_adjoint  adjoint sources: the receivers and the time-reversed
 *                         residuals [n][nt] in x and y (ADJOINT_TYPE picks;
 *                         one-slot handles)
 *   sofi_fwi_gradient     the selected slot's vp / vs / rho' gradients (padded
 *                         planes, NULL to skip), added to the sums; clears the
 *                         slot's accumulators (src/inversion.c:226-249)
 *   sofi_fwi_gradient_sum the sums over shots; reset != 0 clears them
 *                         (init_inv, src/init_inv.c:20-25)
 * At DTINV = 1 the store and correlation steps are the reference's; with
 * DTINV > 1 the reference indexes its store past its end, and the forward
 * step s is kept in store slot (s - 1) / DTINV.  In a batch every shot slot
 * runs both passes in the same launches (the reference runs its shots one
 * after the other, src/sofi2D.c:320-395). */
int sofi_fwi_set_model(sofi_dev *h, const float *vp0, const float *vs0, const float *rho0);
int sofi_fwi_set_pass(sofi_dev *h, int sw);
int sofi_fwi_set_adjoint(sofi_dev *h, int n, const int *i, const int *j, const float *sig_x, const float *sig_y);
int sofi_fwi_gradient(sofi_dev *h, float *grad_vp, float *grad_vs, float *grad_rho);
int sofi_fwi_gradient_sum(sofi_dev *h, float *grad_vp, float *grad_vs, float *grad_rho, int reset);

/* The residuals of calc_res on the device (src/calc_res.c:127-301, every
 * LNORM it has: 1-5, 7, 8), from the selected slot's synthetic vx / vy sections as the
 * forward pass left them in device memory and the measured sections obs_x /
 * obs_y ([ntr][ns] host arrays in receiver order; NULL for a component that
 * ADJOINT_TYPE does not use), time-reversed into the adjoint sources at the
 * slot's receivers: what inversion.c hands the adjoint pass
 * (src/inversion.c:86-110).  With time_filt set both sections are first
 * filtered as timedomain_filt does (src/timedomain_filt.c:28-113: zero-phase
 * Butterworth low-pass, then high-pass when f_high_pass > 0).  *misfit (may
 * be NULL) receives the misfit calc_res accumulates for a test shot, the
 * components' float sums added in double; the device sums each trace's share
 * and then the traces in order, so it agrees with the reference's single
 * float accumulator to float rounding.  The adjoint sources set by
 * sofi_fwi_set_adjoint are replaced.  No counterpart as a separate call in
 * the reference, whose calc_res runs on the host sections. */
typedef struct sofi_residual_opts {
    int lnorm;               /* LNORM */
    int time_filt;           /* TIME_FILT: 0 none, else filter both sections first */
    float f_low_pass;        /* low-pass corner [Hz] (F_LOW_PASS_START or FREQ_FILE's frequency) */
    float f_high_pass;       /* high-pass corner [Hz], 0: none */
    int order;               /* ORDER of the Butterworth filter (1..17) */
    float waterlevel_lnorm8; /* WATERLEVEL_LNORM8 */
} sofi_residual_opts;
int sofi_fwi_residual(sofi_dev *h, const sofi_residual_opts *opts, const float *obs_x, const float *obs_y,
                      double *misfit);
/* The adjoint sources the next adjoint pass injects: component comp (0 vx,
 * 1 vy) as [n][NT] floats, n = sofi_fwi_adjoint_count. */
int sofi_fwi_get_adjoint(sofi_dev *h, int comp, float *out);
int sofi_fwi_adjoint_count(const sofi_dev *h);

/* sofi_fwi_residual without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * measured sections obs_x / obs_y are device arrays [ntr][ns] (NULL for a
 * component ADJOINT_TYPE does not use); misfit (may be NULL) is a device
 * array of 2 floats that receives the vx and vy misfits.  The workspace is
 * a static device array of 4 Mi floats (16 MiB) per GPU, held until the
 * process ends: 3 ntr ns + ntr + 2 floats must fit, else the call fails.
 * The adjoint arrays must be those of the handle's receivers (set up by the
 * first forward pass, or by sofi_fwi_residual).  Calls on one GPU share the
 * workspace and are serialised; each returns when its work is done.  The
 * results are those of sofi_fwi_residual bit for bit. */
int sofi_fwi_residual_noalloc(sofi_dev *h, const sofi_residual_opts *opts, const float *obs_x, const float *obs_y,
                              float *misfit);

/* Shapes of the isotropic/VTI velocity and stress update kernels.  No counterpart in the
 * reference (a launch-tuning knob): SOFI_LAUNCH_CELL runs one thread per
 * cell in 64 x 4 tiles, SOFI_LAUNCH_VEC2 / SOFI_LAUNCH_VEC4 one thread per
 * two / four consecutive cells of a row (vector loads and stores) in
 * 128 x 4 / 256 x 4 tiles, SOFI_LAUNCH_AUTO (the default) the shape
 * measured fastest for the rheology (VEC2 for EL_ISO / EL_VTI, CELL
 * otherwise), for one-shot handles and batches alike.  Results agree to
 * rounding between shapes and bit for bit between a batch slot and a
 * one-shot handle of the same shape. */
enum { SOFI_LAUNCH_AUTO = 0, SOFI_LAUNCH_CELL = 1, SOFI_LAUNCH_VEC2 = 2, SOFI_LAUNCH_VEC4 = 3 };
int sofi_set_launch(sofi_dev *h, int velocity_shape, int stress_shape);

/* Captured blocks of time steps.  The reference's time loop launches every
 * step's updates one after the other (src/time_loop.c:40-325); here sofi_run
 * records `steps` consecutive steps (rounded up to a multiple of NDT, and of
 * DTINV for FWI; 0: 64) once into a HIP graph whose kernels read the block's
 * base step from device memory, and replays it for every block that starts
 * on a sampling (and FWI store) boundary and holds no snapshot and not the
 * loop's last step; the other steps run one by one.  SOFI_GRAPH_AUTO (the
 * default) does so for grids of at most 2^20 cells (all slots of a batch
 * together), where launch overhead is a large part of a step; timed runs
 * (sofi_set_kernel_timing) and batches advanced by resident groups run step
 * by step.  Results are identical with and without graphs. */
enum { SOFI_GRAPH_AUTO = 0, SOFI_GRAPH_OFF = 1, SOFI_GRAPH_ON = 2 };
int sofi_set_graph(sofi_dev *h, int mode, int steps);
/* Blocks replayed from a captured graph since the handle was opened. */
long long sofi_graph_blocks(const sofi_dev *h);

/* Sampled kernel timing on the handle's stream: every `every`-th step of
 * the following sofi_run calls records HIP events around the velocity
 * kernel and around the stress-update kernels (at most max_samples steps);
 * the reference's counterpart is the Perform timers of time_loop
 * (src/perform_struct.h:29-36, src/time_loop.c:49-116).  every = 0 turns it off. */
int sofi_set_kernel_timing(sofi_dev *h, int every, int max_samples);
/* Average device time per timed step of the velocity and stress kernels. */
int sofi_get_kernel_timing(const sofi_dev *h, double *velocity_ms, double *stress_ms, int *samples);

#ifdef __cplusplus
}
#endif
#endif /* SOFI_HIP_H */
