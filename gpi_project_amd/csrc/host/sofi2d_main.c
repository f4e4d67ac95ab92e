/*
 * sofi2d_main.c - the sofi2D command line on one MI355X.
 *
 *     sofi2D <file.json>
 *
 * Same argument check, input files and output files as the reference's
 * main (src/sofi2D.c:54-442) for MODE=FW: reads the parameter file and
 * the model/source/receiver files, runs every shot's time loop on the GPU
 * through libsofi_hip, writes SEIS_FILE_{vx,vy,p,div,curl}.su, the
 * SIGOUT_FILE source signal and the snapshots (already in the merged
 * layout that snapmerge would produce).  Errors print a message and exit
 * with status 1, as log_fatal does.
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

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "FATAL: Unexpected number of commandline arguments; single argument required: name of json parameter file.\n");
        return 1;
    }
    double t0 = now_s();
    sofi_problem *P = sofi_problem_load(argv[1]);
    if (!P) return die_host();
    const sofi_params *p = sofi_problem_params(P);
    if (p->mode != SOFI_MODE_FW) {
        fprintf(stderr, "FATAL: MODE=FWI is driven from the Python layer (gpi_project_amd.fwi); this driver runs MODE=FW.\n");
        return 1;
    }
    int nshots = sofi_problem_nshots(P);
    if (sofi_problem_set_shot(P, 1)) return die_host();
    sofi_dev *h = NULL;
    if (sofi_open(sofi_problem_desc(P), &h)) return die_dev();
    if (sofi_set_medium(h, sofi_problem_arrays(P), sofi_problem_profiles(P))) return die_dev();
    int ntr = sofi_problem_ntr(P);
    if (p->seismo && sofi_set_receivers(h, ntr, sofi_problem_rec_i(P), sofi_problem_rec_j(P))) return die_dev();
    int nsnap = sofi_problem_nsnap(P);
    float *sec = (float *)malloc(sizeof(float) * (size_t)(ntr > 0 ? ntr : 1) * (p->ns > 0 ? p->ns : 1));
    int nxs = 0, nys = 0;
    sofi_snapshot_dims(p, &nxs, &nys);
    float *snapb = (float *)malloc(sizeof(float) * (size_t)nxs * nys);
    int snaps_written = 0;   /* the reference's snapshot counter is static across shots */
    printf("sofi2D (MI355X): %d x %d grid, NT=%d, %d shot(s), %d receiver(s)\n", p->nxg, p->nyg, p->nt, nshots, ntr);
    double loop_s = 0.0;
    for (int ishot = 1; ishot <= nshots; ++ishot) {
        if (ishot > 1 && sofi_problem_set_shot(P, ishot)) return die_host();
        if (sofi_problem_write_signal(P)) return die_host();
        if (sofi_reset(h)) return die_dev();
        if (sofi_set_sources(h, sofi_problem_nsrc(P), sofi_problem_src_i(P), sofi_problem_src_j(P),
                             sofi_problem_src_type(P), sofi_problem_src_azi(P), sofi_problem_src_amp(P),
                             sofi_problem_signals(P), sofi_problem_source_type(P)))
            return die_dev();
        if (sofi_set_snapshots(h, sofi_problem_snapcheck(P) ? nsnap : 0)) return die_dev();
        if (sofi_run(h, 1, p->nt)) return die_dev();
        double ms = sofi_last_run_ms(h);
        loop_s += ms * 1e-3;
        printf("shot %d: time loop %.3f s (%.2f GLUPS)\n", ishot, ms * 1e-3,
               (double)p->nxg * p->nyg * p->nt / (ms * 1e-3) / 1e9);
        if (p->seismo) {
            int comps[5];
            int nc = sofi_problem_seismo_comps(P, comps);
            for (int c = 0; c < nc; ++c) {
                if (sofi_get_section(h, comps[c], sec)) return die_dev();
                if (sofi_problem_write_section(P, comps[c], sec)) return die_host();
            }
        }
        int ns_taken = sofi_snapshot_count(h);
        if (ns_taken > 0) {
            int comps[5];
            int nc = sofi_problem_snap_comps(P, comps);
            for (int k = 0; k < ns_taken; ++k) {
                for (int c = 0; c < nc; ++c) {
                    if (sofi_get_snapshot(h, k, comps[c], snapb)) return die_dev();
                    if (sofi_problem_write_snapshot(P, comps[c], snaps_written == 0, snapb, nxs * nys)) return die_host();
                }
                ++snaps_written;
            }
        }
    }
    printf("Total time loop %.3f s, program %.3f s\n", loop_s, now_s() - t0);
    free(sec);
    free(snapb);
    sofi_close(h);
    sofi_problem_free(P);
    return 0;
}
