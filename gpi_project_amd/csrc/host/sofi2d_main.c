/*
 * sofi2d_main.c - the sofi2D command line on one MI355X.
 *
 *     sofi2D <file.json>
 *
 * Same argument check, input files and output files as the reference's
 * main (src/sofi2D.c:54-442): reads the parameter file and the
 * model/source/receiver files, runs every shot's time loop on the GPU
 * through libsofi_hip, writes SEIS_FILE_{vx,vy,p,div,curl}.su, the
 * SIGOUT_FILE source signal and the snapshots (already in the merged
 * layout that snapmerge would produce).  MODE=FWI adds, per shot, the
 * residuals against DATA_DIR_{vx,vy}.su.shot<n>, the adjoint pass and the
 * gradient files GRADIENT_gradient.{vp,vs,rho}.shot<n>, and the summed
 * GRADIENT_gradient.{vp,vs,rho} after the last shot (src/inversion.c,
 * src/output_gradient.c).  Errors print a message and exit with status 1,
 * as log_fatal does.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "sofi_hip.h"
#include "sofi_host.h"

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

static int die_host(void)
{
    fprintf(stderr, "FATAL: %s\n", sofi_host_last_error());
    return 1;
}

static int die_dev(void)
{
    fprintf(stderr, "FATAL: %s\n", sofi_last_error());
    return 1;
}

typedef struct driver {
    sofi_problem *P;
    const sofi_params *p;
    sofi_dev *h;
    int ntr, nsnap, nxs, nys, snaps_written;
    float *sec, *snapb;
} driver;

/* one forward pass of the current shot: sources, time loop, seismograms and
 * snapshots written as the reference writes them */
static int forward_shot(driver *D, int ishot, double *loop_s)
{
    sofi_problem *P = D->P;
    const sofi_params *p = D->p;
    sofi_dev *h = D->h;
    if (sofi_problem_write_signal(P)) return die_host();
    if (sofi_reset(h)) return die_dev();
    if (sofi_set_sources(h, sofi_problem_nsrc(P), sofi_problem_src_i(P), sofi_problem_src_j(P),
                         sofi_problem_src_type(P), sofi_problem_src_azi(P), sofi_problem_src_amp(P),
                         sofi_problem_signals(P), sofi_problem_source_type(P)))
        return die_dev();
    if (sofi_set_snapshots(h, sofi_problem_snapcheck(P) ? D->nsnap : 0)) return die_dev();
    if (sofi_run(h, 1, p->nt)) return die_dev();
    double ms = sofi_last_run_ms(h);
    *loop_s += ms * 1e-3;
    printf("shot %d: time loop %.3f s (%.2f GLUPS)\n", ishot, ms * 1e-3,
           (double)p->nxg * p->nyg * p->nt / (ms * 1e-3) / 1e9);
    if (p->seismo) {
        int comps[5];
        int nc = sofi_problem_seismo_comps(P, comps);
        for (int c = 0; c < nc; ++c) {
            if (sofi_get_section(h, comps[c], D->sec)) return die_dev();
            if (sofi_problem_write_section(P, comps[c], D->sec)) return die_host();
        }
    }
    int ns_taken = sofi_snapshot_count(h);
    if (ns_taken > 0) {
        int comps[5];
        int nc = sofi_problem_snap_comps(P, comps);
        for (int k = 0; k < ns_taken; ++k) {
            for (int c = 0; c < nc; ++c) {
                if (sofi_get_snapshot(h, k, comps[c], D->snapb)) return die_dev();
                if (sofi_problem_write_snapshot(P, comps[c], D->snaps_written == 0, D->snapb, D->nxs * D->nys))
                    return die_host();
            }
            ++D->snaps_written;   /* the reference's snapshot counter is static across shots */
        }
    }
    return 0;
}

/* residuals (on the device, from the synthetic sections the forward pass
 * left there), adjoint pass and gradient of the shot just run forward */
static int adjoint_shot(driver *D, int ishot, float *g[3], double *loop_s)
{
    const sofi_params *p = D->p;
    const size_t n = (size_t)D->ntr * p->nt;
    float *data[2] = { NULL, NULL };
    int rc = 0;
    double l2 = 0.0;
    for (int c = SOFI_C_VX; c <= SOFI_C_VY && !rc; ++c) {
        if (!((c == SOFI_C_VX && (p->adjoint_type == 1 || p->adjoint_type == 3)) ||
              (c == SOFI_C_VY && (p->adjoint_type == 1 || p->adjoint_type == 2))))
            continue;
        data[c] = (float *)malloc(sizeof(float) * (n ? n : 1));
        if (sofi_read_data(p, ishot, c, D->ntr, data[c])) rc = die_host();
    }
    if (!rc) {
        int iv[3];
        float fv[3];
        sofi_problem_residual_opts(D->P, iv, fv);
        const sofi_residual_opts o = { iv[0], iv[1], fv[0], fv[1], iv[2], fv[2] };
        if (sofi_fwi_residual(D->h, &o, data[0], data[1], &l2) ||
            sofi_fwi_set_pass(D->h, 1) || sofi_reset(D->h) || sofi_run(D->h, 1, p->nt) ||
            sofi_fwi_gradient(D->h, g[0], g[1], g[2]) || sofi_fwi_set_pass(D->h, 0))
            rc = die_dev();
    }
    if (!rc) {
        double ms = sofi_last_run_ms(D->h);
        *loop_s += ms * 1e-3;
        printf("shot %d: adjoint time loop %.3f s, L2 %.6e\n", ishot, ms * 1e-3, l2);
        if (sofi_write_gradient(p, ishot, g[0], g[1], g[2])) rc = die_host();
    }
    free(data[0]);
    free(data[1]);
    return rc;
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "FATAL: Unexpected number of commandline arguments; single argument required: name of json parameter file.\n");
        return 1;
    }
    double t0 = now_s();
    driver D;
    memset(&D, 0, sizeof(D));
    sofi_problem *P = D.P = sofi_problem_load(argv[1]);
    if (!P) return die_host();
    const sofi_params *p = D.p = sofi_problem_params(P);
    const int fwi = p->mode == SOFI_MODE_FWI;
    int nshots = sofi_problem_nshots(P);
    if (sofi_problem_set_shot(P, 1)) return die_host();
    if (sofi_open(sofi_problem_desc(P), &D.h)) return die_dev();
    sofi_dev *h = D.h;
    if (sofi_set_medium(h, sofi_problem_arrays(P), sofi_problem_profiles(P))) return die_dev();
    D.ntr = sofi_problem_ntr(P);
    if (p->seismo && sofi_set_receivers(h, D.ntr, sofi_problem_rec_i(P), sofi_problem_rec_j(P))) return die_dev();
    D.nsnap = sofi_problem_nsnap(P);
    const size_t nsec = (size_t)(D.ntr > 0 ? D.ntr : 1) * (p->ns > 0 ? p->ns : 1);
    D.sec = (float *)malloc(sizeof(float) * nsec);
    sofi_snapshot_dims(p, &D.nxs, &D.nys);
    D.snapb = (float *)malloc(sizeof(float) * (size_t)D.nxs * D.nys);
    float *g[3] = { NULL, NULL, NULL };
    if (fwi) {
        for (int k = 0; k < 3; ++k) g[k] = (float *)calloc(sofi_plane(p->nxg, p->nyg), sizeof(float));
        if (sofi_fwi_set_model(h, sofi_problem_model(P, 0), sofi_problem_model(P, 1), sofi_problem_model(P, 2)))
            return die_dev();
    }
    printf("sofi2D (MI355X): %d x %d grid, NT=%d, %d shot(s), %d receiver(s)%s\n", p->nxg, p->nyg, p->nt, nshots,
           D.ntr, fwi ? ", MODE=FWI (gradients)" : "");
    double loop_s = 0.0;
    const int iters = fwi ? p->itermax : 1;
    for (int iter = 1; iter <= iters; ++iter) {
        if (fwi && sofi_fwi_gradient_sum(h, NULL, NULL, NULL, 1)) return die_dev();   /* init_inv */
        for (int ishot = 1; ishot <= nshots; ++ishot) {
            if (sofi_problem_set_shot(P, ishot)) return die_host();
            if (forward_shot(&D, ishot, &loop_s)) return 1;
            if (fwi && adjoint_shot(&D, ishot, g, &loop_s)) return 1;
        }
        if (fwi) {
            if (sofi_fwi_gradient_sum(h, g[0], g[1], g[2], 0)) return die_dev();
            if (sofi_write_gradient(p, 0, g[0], g[1], g[2])) return die_host();
        }
    }
    printf("Total time loop %.3f s, program %.3f s\n", loop_s, now_s() - t0);
    free(D.sec);
    free(D.snapb);
    for (int k = 0; k < 3; ++k) free(g[k]);
    sofi_close(h);
    sofi_problem_free(P);
    return 0;
}
