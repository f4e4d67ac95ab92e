/*
 * sofi_problem.c - one-call setup of a run from its JSON file, with plain
 * getters (used by the sofi2D driver and by the Python layer through ctypes).
 *
 * Sequence of src/sofi2D.c:142-352 for one process: read_par_json,
 * acq_read, readmod, checkfd, PML_pro/absorb, prepmod, then per shot
 * initsrc + subgrid_bounds.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sofi_host_internal.h"

typedef struct sofi_problem {
    sofi_params p;
    sofi_model m;
    sofi_acq a;
    sofi_shot s;
    sofi_desc d;
    int ishot;
    int source_type;     /* SOURCE_TYPE carried from shot to shot (update_v_interior side effect) */
    int *rec_i, *rec_j;  /* 0-based arrays of receiver grid indices */
} sofi_problem;

sofi_problem *sofi_problem_load(const char *json) { return sofi_problem_load_in(json, NULL); }

sofi_problem *sofi_problem_load_in(const char *json, const char *base_dir)
{
    sofi_problem *P = (sofi_problem *)calloc(1, sizeof(sofi_problem));
    if (!P) { sofi_fail("out of memory"); return NULL; }
    if (sofi_read_params_in(json, base_dir, &P->p) || sofi_check_supported(&P->p) || sofi_read_acq(&P->p, &P->a) ||
        sofi_build_model(&P->p, &P->m) || sofi_checkfd(&P->p, &P->m, &P->a)) {
        sofi_free_model(&P->m);
        sofi_free_acq(&P->a);
        free(P);
        return NULL;
    }
    P->source_type = P->p.source_type;
    P->rec_i = (int *)calloc(P->a.ntr + 1, sizeof(int));
    P->rec_j = (int *)calloc(P->a.ntr + 1, sizeof(int));
    for (int k = 0; k < P->a.ntr; ++k) {
        P->rec_i[k] = P->a.recpos[(size_t)1 * (P->a.ntr + 1) + k + 1];
        P->rec_j[k] = P->a.recpos[(size_t)2 * (P->a.ntr + 1) + k + 1];
    }
    sofi_fill_desc(&P->p, NULL, &P->d);
    return P;
}

void sofi_problem_free(sofi_problem *P)
{
    if (!P) return;
    sofi_free_shot(&P->s);
    sofi_free_model(&P->m);
    sofi_free_acq(&P->a);
    free(P->rec_i);
    free(P->rec_j);
    free(P);
}

int sofi_problem_nshots(const sofi_problem *P) { return P->a.nshots; }

/* initsrc for shot ishot (1-based) and the descriptor that goes with it */
int sofi_problem_set_shot(sofi_problem *P, int ishot)
{
    if (ishot < 1 || ishot > P->a.nshots) return sofi_fail("shot %d out of range 1..%d", ishot, P->a.nshots);
    sofi_free_shot(&P->s);
    if (sofi_init_shot(&P->p, &P->a, ishot, P->source_type, &P->s)) return -1;
    P->source_type = P->s.source_type;
    P->ishot = ishot;
    sofi_fill_desc(&P->p, &P->s, &P->d);
    return 0;
}

const sofi_desc *sofi_problem_desc(const sofi_problem *P) { return &P->d; }
const sofi_params *sofi_problem_params(const sofi_problem *P) { return &P->p; }
const float *sofi_problem_array(const sofi_problem *P, int id) { return (id >= 0 && id < SOFI_A_NARRAYS) ? P->m.a[id] : NULL; }
const float *sofi_problem_profile(const sofi_problem *P, int id) { return (id >= 0 && id < SOFI_P_NPROFILES) ? P->m.prof[id] : NULL; }
const float *const *sofi_problem_arrays(const sofi_problem *P) { return (const float *const *)P->m.a; }
const float *const *sofi_problem_profiles(const sofi_problem *P) { return (const float *const *)P->m.prof; }
const float *sofi_problem_model(const sofi_problem *P, int which)
{
    switch (which) {
      case 0: return P->m.vp0;
      case 1: return P->m.vs0;
      case 2: return P->m.rho0;
      case 3: return P->m.prho;
    }
    return NULL;
}

int sofi_problem_nsrc(const sofi_problem *P) { return P->s.nsrc; }
const int *sofi_problem_src_i(const sofi_problem *P) { return P->s.src_i; }
const int *sofi_problem_src_j(const sofi_problem *P) { return P->s.src_j; }
const int *sofi_problem_src_type(const sofi_problem *P) { return P->s.src_type; }
const float *sofi_problem_src_azi(const sofi_problem *P) { return P->s.src_azi; }
const float *sofi_problem_src_amp(const sofi_problem *P) { return P->s.src_amp; }
const float *sofi_problem_signals(const sofi_problem *P) { return P->s.signals; }
int sofi_problem_source_type(const sofi_problem *P) { return P->s.source_type; }
int sofi_problem_snapcheck(const sofi_problem *P) { return P->s.snapcheck; }

int sofi_problem_ntr(const sofi_problem *P) { return P->a.ntr; }
const int *sofi_problem_rec_i(const sofi_problem *P) { return P->rec_i; }
const int *sofi_problem_rec_j(const sofi_problem *P) { return P->rec_j; }
int sofi_problem_itermax(const sofi_problem *P) { return P->p.itermax; }

/* the residual options calc_res and timedomain_filt read (LNORM, TIME_FILT
 * and its corners and ORDER, WATERLEVEL_LNORM8; src/calc_res.c,
 * src/timedomain_filt.c), after the workflow's stage 1 applied them:
 * ivals = {lnorm, time_filt, order}, fvals = {f_low_pass, f_high_pass,
 * waterlevel_lnorm8} */
int sofi_problem_residual_opts(const sofi_problem *P, int *ivals, float *fvals)
{
    ivals[0] = P->p.lnorm;
    ivals[1] = P->p.time_filt;
    ivals[2] = P->p.filt_order;
    fvals[0] = P->p.f_low_pass;
    fvals[1] = P->p.f_high_pass;
    fvals[2] = P->p.waterlevel_lnorm8;
    return 0;
}

int sofi_problem_nsnap(const sofi_problem *P) { return P->p.snap ? sofi_count_snapshots(&P->p) : 0; }

int sofi_problem_write_section(const sofi_problem *P, int comp, const float *section)
{
    return sofi_write_section(&P->p, &P->a, P->ishot, comp, section);
}

int sofi_problem_write_signal(const sofi_problem *P) { return sofi_write_signal(&P->p, &P->a, &P->s, P->ishot); }

int sofi_problem_write_snapshot(const sofi_problem *P, int comp, int first, const float *data, int n)
{
    return sofi_write_snapshot(&P->p, comp, first, data, n);
}

/* which section components SEISMO writes, in saveseis order (src/saveseis.c) */
int sofi_problem_seismo_comps(const sofi_problem *P, int *comps)
{
    int n = 0;
    switch (P->p.seismo) {
      case 1: comps[n++] = SOFI_C_VX; comps[n++] = SOFI_C_VY; break;
      case 2: comps[n++] = SOFI_C_P; break;
      case 3: comps[n++] = SOFI_C_DIV; comps[n++] = SOFI_C_CURL; break;
      case 4: comps[n++] = SOFI_C_VX; comps[n++] = SOFI_C_VY; comps[n++] = SOFI_C_P;
              comps[n++] = SOFI_C_DIV; comps[n++] = SOFI_C_CURL; break;
    }
    return n;
}

/* which snapshot components SNAP writes (src/snap_ssg.c:86-205) */
int sofi_problem_snap_comps(const sofi_problem *P, int *comps)
{
    int n = 0;
    switch (P->p.snap) {
      case 1: comps[n++] = SOFI_C_VX; comps[n++] = SOFI_C_VY; break;
      case 2: comps[n++] = SOFI_C_P; break;
      case 3: comps[n++] = SOFI_C_DIV; comps[n++] = SOFI_C_CURL; break;
      case 4: comps[n++] = SOFI_C_VX; comps[n++] = SOFI_C_VY; comps[n++] = SOFI_C_P;
              comps[n++] = SOFI_C_DIV; comps[n++] = SOFI_C_CURL; break;
    }
    return n;
}
