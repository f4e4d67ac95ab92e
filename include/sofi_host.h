/*
 * sofi_host.h - host side of the SOFI2D-compatible driver (libsofi_host.so).
 *
 * It reads the reference's input files (pseudo-JSON parameter file, model
 * files, source and receiver files), builds every array the time loop
 * reads, and writes the reference's output files (SU seismograms, merged
 * binary snapshots, source-signal SU file).  It does no time stepping: the
 * time loop runs on the GPU through sofi_hip.h.
 *
 * All functions return 0 on success and a non-zero code on failure, with a
 * message available from sofi_host_last_error(); this replaces the
 * reference's log_fatal() -> MPI_Abort (src/logging.c:148-210).
 */
#ifndef SOFI_HOST_H
#define SOFI_HOST_H

#include "sofi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SOFI_STR 256
#define SOFI_NSPAR 12

/* Parameters read from the JSON file: the subset of GlobVar/GlobVarInv the
 * forward and gradient paths use (src/read_par_json.c:108-620,
 * src/read_par_json_fwi.c). */
typedef struct sofi_params {
    int mode, weq;
    int nprocx, nprocy;          /* read for compatibility; one device runs the whole grid */
    int fdorder, fdorder_time, maxrelerror;
    int nxg, nyg;
    float dh, time, dt;
    int nt, ns;
    int source_type, source_shape, sigout, sigout_format, srcrec, run_multiple_shots, source_topo;
    char sigout_file[SOFI_STR], signal_file[SOFI_STR], source_file[SOFI_STR];
    float plane_wave_depth, plane_wave_angle, ts;
    int free_surf, boundary, fw, abs_type;
    float damping, npower, k_max_cpml, fpml, vppml;
    int snap, snap_format, snapshot_start, snapshot_end, snapshot_incr, idx, idy;
    float tsnap1, tsnap2, tsnapinc;
    char snap_file[SOFI_STR];
    int seismo, readrec, rec_topo, ndt, seis_format, rec_array, drx;
    char rec_file[SOFI_STR], seis_file[SOFI_STR];
    float xrec1, yrec1, xrec2, yrec2, refrec[4], ngeoph, rec_array_depth, rec_array_dist;
    char mfile[SOFI_STR];
    int write_modelfiles, log, readmod, outntimestepinfo;
    char log_file[SOFI_STR], log_verbosity[SOFI_STR];
    float tau;
    int L;
    float f_ref, fl[SOFI_MAXL + 1];
    /* FWI (src/read_par_json_fwi.c) */
    int itermax, dtinv, adjoint_type, velocity, lnorm, time_filt, eprecond;
    float waterlevel_lnorm8;
    float f_low_pass, f_high_pass;   /* TIME_FILT corner frequencies (F_LOW_PASS_START or FREQ_FILE's first) */
    int filt_order;
    int use_workflow;                /* USE_WORKFLOW: stage 1 of FILE_WORKFLOW sets TIME_FILT, the corners, EPRECOND */
    int trkill, timewin;             /* TRKILL, TIMEWIN: read to refuse them */
    char file_workflow[SOFI_STR];
    char data_dir[SOFI_STR], gradient_file[SOFI_STR];
    /* derived */
    int nd;
    float dtdh;
    /* the directory relative paths of the parameter file are resolved
     * against ("" : the working directory, as the reference does) */
    char base_dir[SOFI_STR];
} sofi_params;

/* Model: every per-cell array of enum sofi_array that the rheology uses, in
 * the padded layout, plus the 1-D profiles of enum sofi_profile. */
typedef struct sofi_model {
    int nx, ny, pitch, rows, L, weq;
    float *a[SOFI_A_NARRAYS];        /* NULL when the rheology does not use it */
    float *prof[SOFI_P_NPROFILES];   /* CPML: 2FW+1 entries; ETA/BIP/CIP: L+1 entries */
    float *prho;                     /* density (kept for FWI) */
    float *vp0, *vs0, *rho0;         /* model as read (kept for FWI gradient assembly) */
    float vpmin, vpmax, vsmin, vsmax;
} sofi_model;

/* Acquisition read once per run (src/acq_read.c:25-80). */
typedef struct sofi_acq {
    int nsrc, nshots;
    float *srcpos;                   /* [SOFI_NSPAR+1][nsrc+1], reference srcpos[row][l], row 1..12 */
    int ntr;
    int *recpos;                     /* [4][ntr+1]: rows 1..3 = i, j, global trace number */
} sofi_acq;

/* One shot's local sources with their signals (src/initsrc.c, src/wavelet.c). */
typedef struct sofi_shot {
    int nsrc;
    int *src_i, *src_j, *src_type;   /* grid indices (1-based) and per-source type */
    float *src_azi;                  /* srcpos[7]: azimuth / dip [deg] */
    float *src_amp;                  /* srcpos[6]: amplitude (eqsource) */
    float *signals;                  /* [nsrc][nt]: signals[k*nt + (nt_step-1)] */
    int snapcheck;                   /* snapshots wanted for this shot */
    int source_type;                 /* SOURCE_TYPE in force after the first velocity update */
} sofi_shot;

const char *sofi_host_last_error(void);

int sofi_read_params(const char *json_file, sofi_params *p);
/* sofi_read_params with every relative file name of the parameter file (the
 * parameter file's own included) taken relative to base_dir instead of the
 * working directory, so that several runs can be set up at once from
 * different directories in one process. */
int sofi_read_params_in(const char *json_file, const char *base_dir, sofi_params *p);
int sofi_check_supported(const sofi_params *p);
int sofi_build_model(const sofi_params *p, sofi_model *m);
void sofi_free_model(sofi_model *m);
int sofi_read_acq(sofi_params *p, sofi_acq *a);
void sofi_free_acq(sofi_acq *a);
int sofi_init_shot(const sofi_params *p, const sofi_acq *a, int ishot, int prev_source_type, sofi_shot *s);
void sofi_free_shot(sofi_shot *s);
int sofi_checkfd(sofi_params *p, const sofi_model *m, const sofi_acq *a);
void sofi_fill_desc(const sofi_params *p, const sofi_shot *s, sofi_desc *d);
void sofi_holberg(int fdorder, int maxrelerror, float hc[7]);

/* Outputs (src/saveseis_glob.c, src/outseis_glob.c, src/write_su.c, src/merge.c). */
int sofi_write_section(const sofi_params *p, const sofi_acq *a, int ishot, int comp, const float *section);
int sofi_write_signal(const sofi_params *p, const sofi_acq *a, const sofi_shot *s, int ishot);
int sofi_count_snapshots(const sofi_params *p);
/* Appends one snapshot (x-major over the decimated grid) to the merged file
 * SNAP_FILE<ext>.<comp> (what snapmerge makes, src/merge.c:85-111) and to the
 * per-process files SNAP_FILE<ext>.<comp>.<px>.<py> the reference's
 * NPROCX x NPROCY processes write (src/snap_ssg.c:62-67); first != 0
 * truncates them. */
int sofi_write_snapshot(const sofi_params *p, int comp, int first, const float *data, int n);
int sofi_snapshot_dims(const sofi_params *p, int *nxs, int *nys);

/* One-call setup of a run (src/sofi2D.c:142-352 for one process) with plain
 * getters; 0-based source/receiver arrays hold 1-based grid indices. */
typedef struct sofi_problem sofi_problem;
sofi_problem *sofi_problem_load(const char *json);
/* sofi_problem_load with relative file names resolved against base_dir
 * (sofi_read_params_in); NULL or "" behaves as sofi_problem_load */
sofi_problem *sofi_problem_load_in(const char *json, const char *base_dir);
void sofi_problem_free(sofi_problem *P);
int sofi_problem_nshots(const sofi_problem *P);
int sofi_problem_set_shot(sofi_problem *P, int ishot);
const sofi_desc *sofi_problem_desc(const sofi_problem *P);
const sofi_params *sofi_problem_params(const sofi_problem *P);
const float *sofi_problem_array(const sofi_problem *P, int id);
const float *sofi_problem_profile(const sofi_problem *P, int id);
const float *const *sofi_problem_arrays(const sofi_problem *P);
const float *const *sofi_problem_profiles(const sofi_problem *P);
const float *sofi_problem_model(const sofi_problem *P, int which);
int sofi_problem_nsrc(const sofi_problem *P);
const int *sofi_problem_src_i(const sofi_problem *P);
const int *sofi_problem_src_j(const sofi_problem *P);
const int *sofi_problem_src_type(const sofi_problem *P);
const float *sofi_problem_src_azi(const sofi_problem *P);
const float *sofi_problem_src_amp(const sofi_problem *P);
const float *sofi_problem_signals(const sofi_problem *P);
int sofi_problem_source_type(const sofi_problem *P);
int sofi_problem_snapcheck(const sofi_problem *P);
int sofi_problem_ntr(const sofi_problem *P);
const int *sofi_problem_rec_i(const sofi_problem *P);
const int *sofi_problem_rec_j(const sofi_problem *P);
int sofi_problem_nsnap(const sofi_problem *P);
int sofi_problem_itermax(const sofi_problem *P);     /* FWI iterations (ITERMAX) */
/* FWI residual options: ivals {LNORM, TIME_FILT, ORDER}, fvals {F_LOW_PASS,
 * F_HIGH_PASS, WATERLEVEL_LNORM8} (the values calc_res / timedomain_filt use) */
int sofi_problem_residual_opts(const sofi_problem *P, int *ivals, float *fvals);
int sofi_problem_write_section(const sofi_problem *P, int comp, const float *section);
int sofi_problem_write_signal(const sofi_problem *P);
int sofi_problem_write_snapshot(const sofi_problem *P, int comp, int first, const float *data, int n);
int sofi_problem_seismo_comps(const sofi_problem *P, int *comps);
int sofi_problem_snap_comps(const sofi_problem *P, int *comps);

/* Low-level writer shared with tests: one SU trace set (outseis_glob, SEIS_FORMAT 1/2/3). */
/* FWI gradient path (src/inversion.c:27-258): measured data of a shot
 * (inseis, src/inseis.c:28-80) as section[ntr][ns]; the adjoint sources of
 * LNORM 1-5, 7 and 8 (calc_res, src/calc_res.c:26-324), time-reversed, with
 * *l2 the misfit calc_res accumulates for a test shot (LNORM 2: sum of
 * squares; 5 / 7: minus the normalised correlation; otherwise 0), both
 * sections filtered by sofi_time_filter first when TIME_FILT is set;
 * the gradient files (output_gradient, src/output_gradient.c:25-68), per shot
 * (ishot > 0) or summed (ishot = 0), from padded planes. */
int sofi_read_data(const sofi_params *p, int ishot, int comp, int ntr, float *section);
int sofi_residual(const sofi_params *p, int ntr, const float *data, const float *calc, float *diff, double *l2);
/* timedomain_filt (src/timedomain_filt.c:28-113): with TIME_FILT the
 * measured and synthetic traces [ntr][NS] are filtered in place before the
 * residual, a zero-phase (forward, then reversed) Butterworth low-pass of
 * order ORDER at F_LOW_PASS, then the same high-pass at F_HIGH_PASS when it
 * is set; no-op without TIME_FILT. */
int sofi_time_filter(const sofi_params *p, int ntr, float *traces);
int sofi_write_gradient(const sofi_params *p, int ishot, const float *gvp, const float *gvs, const float *grho);

int sofi_outseis(const char *path, const float *section, const int *recpos, int ntr,
                 const float *srcpos, int nsrc, int ns, int seis_form, int ishot, float dh,
                 float dt, int ndt, const float *refrec);

#ifdef __cplusplus
}
#endif
#endif /* SOFI_HOST_H */
