/*
 * time_loop_mi355x.c - the reference-side binding of libsofi_hip.so: what a
 * maintainer adds to the reference's src/ to run its time loop on one
 * MI355X.  TEST INFRASTRUCTURE ONLY: oracle/ref/Makefile compiles it against
 * the reference's own headers (src/fd.h and what it includes) and links it
 * into sofi2D_ref_hip, whose main() (ref_driver.c built with
 * -DTIME_LOOP_MI355X) calls it in place of time_loop() at the point where
 * src/sofi2D.c:369 calls time_loop().
 *
 * It replaces
 *     void time_loop(int iter, int ishot, int snapcheck, float *hc,
 *                    float **srcpos_loc, float **signalx, float **signaly,
 *                    int nsrc, int sw, AcqVar *acq, MemModel *mpm,
 *                    MemWavefield *mpw, MemInv *minv, GlobVar *gv,
 *                    GlobVarInv *vinv, Perform *perf)
 * (src/fd.h:335-337, src/time_loop.c:27-326) for MODE=FW on one process
 * (NPROCX = NPROCY = 1), every WEQ but VAC_ISO, FDORDER_TIME 2 and 4, CPML,
 * damping frame or none, free surface, BOUNDARY 0 / 1.  On return the
 * reference's state is what its time_loop leaves for the caller:
 * gv->SECTION* hold the seismograms, the snapshot files are written (snap(),
 * src/snap_ssg.c:27-209, with its process-wide snapshot counter), and
 * gv->SOURCE_TYPE holds the last source's type (src/update_v_interior.c:57).
 * The wavefields in mpw are not copied back: nothing after time_loop reads
 * them in MODE=FW (src/sofi2D.c:369-395).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fd.h"
#include "logging.h"
#include "sofi_hip.h"

/* NR matrix a[lo..NY+g][lo..NX+g], lo = 1 - g (src/util.c:231-262), into a
 * fresh padded plane; cells outside the matrix's range stay zero */
static float *plane_of(float **a, int nx, int ny, int g)
{
    if (!a) return NULL;
    const int pitch = sofi_pitch(nx);
    float *p = calloc(sofi_plane(nx, ny), sizeof(float));
    if (!p) log_fatal("time_loop_mi355x: out of host memory\n");
    for (int j = 1 - g; j <= ny + g; j++)
        for (int i = 1 - g; i <= nx + g; i++)
            p[sofi_idx(pitch, i, j)] = a[j][i];
    return p;
}

/* f3tensor a[j][i][1..L] (src/util.c:363-408) over the same cells, into L
 * consecutive padded planes, plane l-1 for mechanism l */
static float *planes_of(float ***a, int nx, int ny, int g, int L)
{
    if (!a) return NULL;
    const int pitch = sofi_pitch(nx);
    const size_t pl = sofi_plane(nx, ny);
    float *p = calloc(pl * (size_t)L, sizeof(float));
    if (!p) log_fatal("time_loop_mi355x: out of host memory\n");
    for (int l = 1; l <= L; l++)
        for (int j = 1 - g; j <= ny + g; j++)
            for (int i = 1 - g; i <= nx + g; i++)
                p[(size_t)(l - 1) * pl + sofi_idx(pitch, i, j)] = a[j][i][l];
    return p;
}

/* vector(1, n) (src/util.c) into n + 1 floats indexed like it, [0] = 0 */
static float *prof_of(const float *v, int n)
{
    if (!v) return NULL;
    float *p = calloc((size_t)n + 1, sizeof(float));
    if (!p) log_fatal("time_loop_mi355x: out of host memory\n");
    for (int h = 1; h <= n; h++) p[h] = v[h];
    return p;
}

static void check(int rc)
{
    if (rc) log_fatal("%s\n", sofi_last_error());
}

/* snapshots this shot's loop takes: the schedule of src/time_loop.c:38-40,306-309 */
static int count_snapshots(const GlobVar *gv)
{
    int lsnap = iround(gv->TSNAP1 / gv->DT), isnap = iround(gv->TSNAPINC / gv->DT);
    int esnap = iround(gv->TSNAP2 / gv->DT), n = 0;
    for (int nt = 1; nt <= gv->NT; nt++)
        if (nt == lsnap && nt <= esnap) {
            ++n;
            lsnap += isnap;
        }
    return n;
}

void time_loop_mi355x(int ishot, int snapcheck, float *hc, float **srcpos_loc, float **signalx, int nsrc,
                      AcqVar *acq, MemModel *mpm, GlobVar *gv)
{
    static int nsnap_written = 0;   /* snap()'s counter, static across shots as in src/time_loop.c:36 */
    const int nx = gv->NX, ny = gv->NY, nd = gv->ND, L = gv->L;

    if (gv->NPROCX != 1 || gv->NPROCY != 1)
        log_fatal("time_loop_mi355x: one process holds the grid (NPROCX = NPROCY = 1)\n");
    if (gv->MODE != FW) log_fatal("time_loop_mi355x: MODE=FW only\n");

    /* ---- the GlobVar scalars the loop reads ------------------------------ */
    sofi_desc d;
    memset(&d, 0, sizeof d);
    d.nx = nx;
    d.ny = ny;
    d.fdorder = gv->FDORDER;
    d.weq = gv->WEQ;
    d.L = L;
    d.nt = gv->NT;
    d.dt = gv->DT;
    d.dh = gv->DH;
    for (int k = 0; k < 7; k++) d.hc[k] = hc[k];
    d.fw = gv->FW;
    d.abs_type = gv->FW > 0 ? gv->ABS_TYPE : 0;
    d.free_surf = gv->FREE_SURF;
    for (int k = 1; k <= 4; k++) {
        d.gx[k] = gv->GX[k];
        d.gy[k] = gv->GY[k];
    }
    /* periodic in x: subgrid_bounds sets GX[4] = NX - 1 (src/subgrid_bounds.c:265-270),
     * which bounds only the (empty) right-frame loops; the library's loops over
     * the whole grid read GX[4] as the last column */
    if (gv->BOUNDARY == 1) d.gx[4] = nx;
    d.boundary = gv->BOUNDARY == 1;
    d.source_type = nsrc > 0 ? (int)srcpos_loc[8][nsrc] : gv->SOURCE_TYPE;
    d.seismo = gv->SEISMO;
    d.ndt = gv->NDT;
    d.ns = gv->NS;
    d.snap = snapcheck ? gv->SNAP : 0;
    for (int k = 1; k <= 4; k++) d.snap_idx[k] = gv->SNAPIDX[k];
    d.idx = gv->IDX;
    d.idy = gv->IDY;
    d.lsnap = iround(gv->TSNAP1 / gv->DT);
    d.isnap = iround(gv->TSNAPINC / gv->DT);
    d.esnap = iround(gv->TSNAP2 / gv->DT);
    d.mode = SOFI_MODE_FW;
    d.dtinv = 1;
    d.fdorder_time = gv->FDORDER_TIME;

    sofi_dev *h = NULL;
    check(sofi_open(&d, &h));

    /* ---- MemModel after prepmod / PML_pro / absorb ---------------------- */
    const float *arrays[SOFI_A_NARRAYS];
    float *own[SOFI_A_NARRAYS];
    memset(arrays, 0, sizeof arrays);
    memset(own, 0, sizeof own);
#define PLANE(id, a) (arrays[id] = own[id] = plane_of((a), nx, ny, nd))
#define PLANES(id, a) (arrays[id] = own[id] = planes_of((a), nx, ny, nd, L))
    PLANE(SOFI_A_RIP, mpm->prip);
    PLANE(SOFI_A_RJP, mpm->prjp);
    PLANE(SOFI_A_PPI, mpm->ppi);
    PLANE(SOFI_A_PU, mpm->pu);
    PLANE(SOFI_A_PUIPJP, mpm->puipjp);
    switch (gv->WEQ) {
      case AC_ISO:
          PLANE(SOFI_A_G, mpm->g);
          break;
      case EL_ISO:
          PLANE(SOFI_A_F, mpm->f);
          PLANE(SOFI_A_G, mpm->g);
          PLANE(SOFI_A_FIPJP, mpm->fipjp);
          break;
      case VEL_ISO:
          PLANE(SOFI_A_F, mpm->f);
          PLANE(SOFI_A_G, mpm->g);
          PLANE(SOFI_A_FIPJP, mpm->fipjp);
          PLANE(SOFI_A_PTAUS, mpm->ptaus);
          PLANE(SOFI_A_PTAUP, mpm->ptaup);
          PLANE(SOFI_A_PTAUSIPJP, mpm->ptausipjp);
          PLANES(SOFI_A_DIP, mpm->dip);
          PLANES(SOFI_A_D, mpm->d);
          PLANES(SOFI_A_E, mpm->e);
          break;
      case EL_VTI:
      case EL_TTI:
          /* DT-scaled by prepmod (src/prepmod.c:119-137); pc11 / pc55 /
           * pc55ipjp are the ppi / pu / puipjp matrices (src/initmem_model.c) */
          PLANE(SOFI_A_C11, mpm->pc11);
          PLANE(SOFI_A_C13, mpm->pc13);
          PLANE(SOFI_A_C33, mpm->pc33);
          PLANE(SOFI_A_C55IPJP, mpm->pc55ipjp);
          if (gv->WEQ == EL_TTI) {
              PLANE(SOFI_A_C15, mpm->pc15);
              PLANE(SOFI_A_C35, mpm->pc35);
              PLANE(SOFI_A_C15IPJP, mpm->pc15ipjp);
              PLANE(SOFI_A_C35IPJP, mpm->pc35ipjp);
          }
          break;
      case VEL_VTI:
      case VEL_TTI:
          /* unrelaxed moduli and defects of prepare_update_s_vti / _tti */
          PLANE(SOFI_A_C11U, mpm->pc11u);
          PLANE(SOFI_A_C13U, mpm->pc13u);
          PLANE(SOFI_A_C33U, mpm->pc33u);
          PLANE(SOFI_A_C55IPJPU, mpm->pc55ipjpu);
          PLANES(SOFI_A_C11D, mpm->pc11d);
          PLANES(SOFI_A_C13D, mpm->pc13d);
          PLANES(SOFI_A_C33D, mpm->pc33d);
          PLANES(SOFI_A_C55IPJPD, mpm->pc55ipjpd);
          if (gv->WEQ == VEL_TTI) {
              PLANE(SOFI_A_C15U, mpm->pc15u);
              PLANE(SOFI_A_C35U, mpm->pc35u);
              PLANE(SOFI_A_C15IPJPU, mpm->pc15ipjpu);
              PLANE(SOFI_A_C35IPJPU, mpm->pc35ipjpu);
              PLANES(SOFI_A_C15D, mpm->pc15d);
              PLANES(SOFI_A_C35D, mpm->pc35d);
              PLANES(SOFI_A_C15IPJPD, mpm->pc15ipjpd);
              PLANES(SOFI_A_C35IPJPD, mpm->pc35ipjpd);
          }
          break;
      default:
          log_fatal("time_loop_mi355x: WEQ %d is not supported\n", (int)gv->WEQ);
    }
#undef PLANE
#undef PLANES
    /* absorb_coeff is matrix(1, NY, 1, NX): no ghost cells */
    if (d.abs_type == 2) arrays[SOFI_A_ABSORB] = own[SOFI_A_ABSORB] = plane_of(mpm->absorb_coeff, nx, ny, 0);

    const float *profiles[SOFI_P_NPROFILES];
    float *ownp[SOFI_P_NPROFILES];
    memset(profiles, 0, sizeof profiles);
    memset(ownp, 0, sizeof ownp);
    if (d.abs_type == 1) {
        /* CPML vectors, vector(1, 2 FW) (src/initmem_model.c, src/PML_pro.c) */
        const float *cpml[12] = { mpm->K_x, mpm->a_x, mpm->b_x, mpm->K_x_half, mpm->a_x_half, mpm->b_x_half,
                                  mpm->K_y, mpm->a_y, mpm->b_y, mpm->K_y_half, mpm->a_y_half, mpm->b_y_half };
        for (int k = 0; k < 12; k++) profiles[SOFI_P_K_X + k] = ownp[SOFI_P_K_X + k] = prof_of(cpml[k], 2 * gv->FW);
    }
    if (L > 0 && gv->WEQ != EL_ISO && gv->WEQ != EL_VTI && gv->WEQ != EL_TTI) {
        profiles[SOFI_P_ETA] = ownp[SOFI_P_ETA] = prof_of(mpm->peta, L);
        profiles[SOFI_P_BIP] = ownp[SOFI_P_BIP] = prof_of(mpm->bip, L);
        profiles[SOFI_P_CIP] = ownp[SOFI_P_CIP] = prof_of(mpm->cip, L);
    }
    check(sofi_set_medium(h, arrays, profiles));
    for (int k = 0; k < SOFI_A_NARRAYS; k++) free(own[k]);
    for (int k = 0; k < SOFI_P_NPROFILES; k++) free(ownp[k]);

    /* ---- sources: srcpos_loc rows 1, 2 (grid indices), 6 (amplitude),
     * 7 (azimuth / dip), 8 (type), signals[1..nsrc][1..NT] ----------------- */
    const int ns = nsrc > 0 ? nsrc : 1;
    int *si = malloc(ns * sizeof(int)), *sj = malloc(ns * sizeof(int)), *st = malloc(ns * sizeof(int));
    float *azi = malloc(ns * sizeof(float)), *amp = malloc(ns * sizeof(float));
    float *sig = malloc((size_t)ns * gv->NT * sizeof(float));
    if (!si || !sj || !st || !azi || !amp || !sig) log_fatal("time_loop_mi355x: out of host memory\n");
    for (int l = 1; l <= nsrc; l++) {
        si[l - 1] = (int)srcpos_loc[1][l];
        sj[l - 1] = (int)srcpos_loc[2][l];
        amp[l - 1] = srcpos_loc[6][l];
        azi[l - 1] = srcpos_loc[7][l];
        st[l - 1] = (int)srcpos_loc[8][l];
        memcpy(sig + (size_t)(l - 1) * gv->NT, &signalx[l][1], (size_t)gv->NT * sizeof(float));
    }
    check(sofi_set_sources(h, nsrc, si, sj, st, azi, amp, sig, d.source_type));
    free(si);
    free(sj);
    free(st);
    free(azi);
    free(amp);
    free(sig);

    /* ---- receivers: acq->recpos_loc[1..2][1..NTR] (src/seismo_ssg.c:28-67) */
    const int ntr = gv->SEISMO ? gv->NTR : 0;
    if (ntr > 0) {
        int *ri = malloc(ntr * sizeof(int)), *rj = malloc(ntr * sizeof(int));
        if (!ri || !rj) log_fatal("time_loop_mi355x: out of host memory\n");
        for (int r = 1; r <= ntr; r++) {
            ri[r - 1] = acq->recpos_loc[1][r];
            rj[r - 1] = acq->recpos_loc[2][r];
        }
        check(sofi_set_receivers(h, ntr, ri, rj));
        free(ri);
        free(rj);
    }

    const int nsnap = d.snap ? count_snapshots(gv) : 0;
    check(sofi_set_snapshots(h, nsnap));

    /* ---- the loop: zero_wavefield, then steps 1..NT ---------------------- */
    check(sofi_reset(h));
    check(sofi_run(h, 1, gv->NT));   /* "Simulation is unstable!" as src/time_loop.c:44-47 */
    gv->SOURCE_TYPE = d.source_type;

    /* ---- seismograms into gv->SECTION*[1..NTR][1..NS] ------------------- */
    if (ntr > 0) {
        float **sections[SOFI_NCOMP] = { gv->SECTIONVX, gv->SECTIONVY, gv->SECTIONP, gv->SECTIONDIV, gv->SECTIONCURL };
        float *sec = malloc((size_t)ntr * gv->NS * sizeof(float));
        if (!sec) log_fatal("time_loop_mi355x: out of host memory\n");
        for (int c = 0; c < SOFI_NCOMP; c++) {
            if (!sections[c]) continue;   /* allocated for the components SEISMO asks for */
            check(sofi_get_section(h, c, sec));
            for (int r = 0; r < ntr; r++)
                memcpy(&sections[c][r + 1][1], sec + (size_t)r * gv->NS, (size_t)gv->NS * sizeof(float));
        }
        free(sec);
    }

    /* ---- snapshot files, as snap() appends them ------------------------- */
    const int taken = sofi_snapshot_count(h);
    if (taken > 0) {
        static const char *names[SOFI_NCOMP] = { "vx", "vy", "p", "div", "curl" };
        const int comps[5][SOFI_NCOMP] = {   /* SNAP 1..4 -> components, -1 ends */
            { -1 },
            { SOFI_C_VX, SOFI_C_VY, -1 },
            { SOFI_C_P, -1 },
            { SOFI_C_DIV, SOFI_C_CURL, -1 },
            { SOFI_C_VX, SOFI_C_VY, SOFI_C_P, SOFI_C_DIV, SOFI_C_CURL },
        };
        const char *ext = gv->SNAP_FORMAT == 1 ? ".su" : gv->SNAP_FORMAT == 2 ? ".asc" : ".bin";
        const int nxs = (gv->SNAPIDX[2] - gv->SNAPIDX[1]) / gv->IDX + 1;
        const int nys = (gv->SNAPIDX[4] - gv->SNAPIDX[3]) / gv->IDY + 1;
        float *buf = malloc((size_t)nxs * nys * sizeof(float));
        if (!buf) log_fatal("time_loop_mi355x: out of host memory\n");
        for (int k = 0; k < taken; k++) {
            ++nsnap_written;
            for (int q = 0; q < SOFI_NCOMP && gv->SNAP >= 1 && gv->SNAP <= 4; q++) {
                const int c = comps[gv->SNAP][q];
                if (c < 0) break;
                char fname[STRING_SIZE * 2];
                snprintf(fname, sizeof fname, "%s%s.%s.%i.%i", gv->SNAP_FILE, ext, names[c], gv->POS[1], gv->POS[2]);
                check(sofi_get_snapshot(h, k, c, buf));   /* x-major [nxs][nys], snap()'s write order */
                FILE *fp = fopen(fname, nsnap_written == 1 ? "w" : "a");
                if (!fp) log_fatal("time_loop_mi355x: cannot open %s\n", fname);
                for (size_t n = 0; n < (size_t)nxs * nys; n++) writedsk(fp, buf[n], gv->SNAP_FORMAT);
                fclose(fp);
            }
        }
        free(buf);
    }
    check(sofi_close(h));
}
