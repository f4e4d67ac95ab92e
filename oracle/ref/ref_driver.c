/*
 * ref_driver.c - main() for the reference built from its own sources
 * (oracle/ref/Makefile).  TEST INFRASTRUCTURE ONLY.
 *
 * It calls the reference's own functions in the order of its main()
 * (src/sofi2D.c:113-404) for MODE=FW and MODE=FWI (one gradient per shot),
 * with one change: seismograms are gathered and written with catseis() +
 * saveseis_glob(), i.e. the MODE=FW branch of saveseis()
 * (src/saveseis.c:27-150), because saveseis.c also references the
 * FFTW-only calc_envelope() (needed for LNORM=8 only), which this image
 * cannot build.  It prints the wall time of each time_loop() as
 * "REF_TIME_LOOP <shot> <seconds> <ranks>".
 *
 * Built with -DTIME_LOOP_MI355X (sofi2D_ref_hip) it calls
 * time_loop_mi355x() instead of time_loop() for MODE=FW: the reference's
 * own setup and output around the GPU time loop, as a maintainer would
 * wire libsofi_hip.so into src/sofi2D.c:369.
 */
#include <unistd.h>

#include "fd.h"
#include "logging.h"

#ifdef TIME_LOOP_MI355X
/* oracle/ref/time_loop_mi355x.c: the time loop on the GPU through libsofi_hip.so */
void time_loop_mi355x(int ishot, int snapcheck, float *hc, float **srcpos_loc, float **signalx, int nsrc,
                      AcqVar *acq, MemModel *mpm, GlobVar *gv);
#endif

static void write_sections(int ishot, AcqVar *acq, GlobVar *gv)
{
    struct {
        int seismo_mask;
        float **sec;
        int type;
    } comps[5] = {
        { 1 | 8, NULL, 1 }, { 1 | 8, NULL, 2 }, { 2 | 8, NULL, 4 }, { 4 | 8, NULL, 5 }, { 4 | 8, NULL, 6 },
    };
    comps[0].sec = gv->SECTIONVX;
    comps[1].sec = gv->SECTIONVY;
    comps[2].sec = gv->SECTIONP;
    comps[3].sec = gv->SECTIONDIV;
    comps[4].sec = gv->SECTIONCURL;
    int mask = gv->SEISMO == 1 ? 1 : gv->SEISMO == 2 ? 2 : gv->SEISMO == 3 ? 4 : gv->SEISMO == 4 ? 8 : 0;
    for (int k = 0; k < 5; ++k) {
        if (!(comps[k].seismo_mask & mask)) continue;
        catseis(comps[k].sec, gv->SEISMO_FULLDATA, acq->recswitch, gv->NTRG, gv->NS);
        if (gv->MPID == 0) saveseis_glob(gv->SEISMO_FULLDATA, acq->recpos, acq->srcpos, ishot, gv->NS, comps[k].type, gv);
    }
}

int main(int argc, char **argv)
{
    int ishot, nshots, snapcheck, iter;
    float *hc = NULL;
    GlobVar gv = {.MPID = -1,.OUTNTIMESTEPINFO = 100,.NDT = 1,.IDX = 1,.IDY = 1 };
    GlobVarInv vinv = {.ITERMAX = 1,.DTINV = 1,.WORKFLOW_STAGE = 1,.LBFGS_ITER_START = 1,
        .FWI_RUN = 1,.GRADIENT_OPTIMIZATION = 1,.FREQ_NR = 1};
    AcqVar acq = { };
    MemWavefield mpw = { };
    MemModel mpm = { };
    MemInv minv = { };
    Perform perf = { };

    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &gv.MPID);
    log_init(NULL);
    if (gv.MPID == 0) {
        if (argc != 2) log_fatal("usage: sofi2D_ref <file.json>\n");
        if (access(argv[1], R_OK) != 0) log_fatal("Cannot open/read json parameter file %s.\n", argv[1]);
        read_par_json(argv[1], &gv, &vinv);
    }
    exchange_par(&gv, &vinv);
    if (gv.MODE == FWI) {   /* workflow and the low-pass start frequency (src/sofi2D.c:150-164) */
        if (vinv.USE_WORKFLOW) read_workflow(&gv, &vinv);
        if (vinv.TIME_FILT == 1) vinv.F_LOW_PASS = vinv.F_LOW_PASS_START;
        if (vinv.TIME_FILT == 2) {
            filter_frequencies(&vinv);
            vinv.F_LOW_PASS = vinv.F_LOW_PASS_EXT[vinv.FREQ_NR];
        }
    }
    check_fs(&gv, &vinv);
    log_set_output(NULL);
    MPI_Barrier(MPI_COMM_WORLD);
    initproc(&gv);
    nshots = acq_read(&acq, &gv);
    initmem(&mpm, &mpw, &minv, &gv, &vinv);
    initfd(&gv);
    hc = holbergcoeff(&gv);
    MPI_Barrier(MPI_COMM_WORLD);
    readmod(&mpm, &minv, &gv, &vinv);
    MPI_Barrier(MPI_COMM_WORLD);
    checkfd(hc, acq.srcpos, acq.nsrc, acq.recpos, &gv);
    if (gv.ABS_TYPE == 1) PML_pro(&mpm, &gv);
    if (gv.ABS_TYPE == 2) absorb(mpm.absorb_coeff, &gv);
    MPI_Barrier(MPI_COMM_WORLD);
    for (iter = 1; iter <= vinv.ITERMAX; iter++) {
        if (gv.MODE == FWI) {   /* src/sofi2D.c:293-299 */
            if (vinv.USE_WORKFLOW) apply_workflow(iter, &gv, &vinv);
            init_grad(iter, &minv, &gv, &vinv);
        }
        prepmod(&mpm, &gv);
        MPI_Barrier(MPI_COMM_WORLD);
        if (gv.MODE == FWI) init_inv(iter, &minv, &gv, &vinv);
        for (ishot = 1; ishot <= nshots; ishot++) {
            gv.SOURCE_SHAPE = gv.SOURCE_SHAPE_OLD;
            snapcheck = initsrc(ishot, nshots, &acq, &gv);
            zero_wavefield(iter, &mpw, &minv, &gv, &vinv);
            subgrid_bounds(1, gv.NX, 1, gv.NY, &gv);
            MPI_Barrier(MPI_COMM_WORLD);
            double t0 = MPI_Wtime();
#ifdef TIME_LOOP_MI355X
            if (gv.MODE == FW)
                time_loop_mi355x(ishot, snapcheck, hc, acq.srcpos_loc, acq.signals, acq.nsrc_loc, &acq, &mpm, &gv);
            else
#endif
                time_loop(iter, ishot, snapcheck, hc, acq.srcpos_loc, acq.signals, acq.signals, acq.nsrc_loc, 0,
                          &acq, &mpm, &mpw, &minv, &gv, &vinv, &perf);
            MPI_Barrier(MPI_COMM_WORLD);
            double t1 = MPI_Wtime();
            if (gv.MPID == 0) {
                int np;
                MPI_Comm_size(MPI_COMM_WORLD, &np);
                printf("REF_TIME_LOOP %d %.6f %d\n", ishot, t1 - t0, np);
                fflush(stdout);
            }
            if (gv.SEISMO) {
                if (gv.MODE == FWI) {
                    for (int i = 1; i <= gv.NTR; i++)
                        for (int j = 1; j <= gv.NS; j++) {
                            if (gv.SEISMO == 1 || gv.SEISMO == 4) {
                                minv.sectionvxcalc[i][j] = gv.SECTIONVX[i][j];
                                minv.sectionvycalc[i][j] = gv.SECTIONVY[i][j];
                            }
                            if (gv.SEISMO == 2 || gv.SEISMO == 4) minv.sectionpcalc[i][j] = gv.SECTIONP[i][j];
                        }
                }
                write_sections(ishot, &acq, &gv);
            }
            if (gv.MODE == FWI) {
                /* same order as src/sofi2D.c:379-391 */
                inversion(iter, ishot, snapcheck, hc, &acq, &mpm, &mpw, &minv, &gv, &vinv, &perf);
                for (int j = 1; j <= gv.NY; j = j + gv.IDY)
                    for (int i = 1; i <= gv.NX; i = i + gv.IDX) {
                        minv.gradRhos_shot[j][i] = 0;
                        minv.gradLam_shot[j][i] = 0;
                        minv.gradMu_shot[j][i] = 0;
                    }
                MPI_Barrier(MPI_COMM_WORLD);
                output_gradient(ishot, nshots, &minv, &gv, &vinv);
            }
        }
    }
    log_finalize();
    MPI_Finalize();
    return 0;
}
