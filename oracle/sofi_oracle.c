/*
 * sofi_oracle.c - CPU restatement of SOFI2D's time loop.  TEST
 * INFRASTRUCTURE ONLY: the checker for the HIP kernels; nothing in the
 * product links or calls it.
 *
 * Built twice (oracle/Makefile): REAL=float reproduces the reference's
 * arithmetic, including the expressions it promotes to double through
 * double literals ("2.0 * f * vyy"); REAL=double is the FP64 build every
 * FP32 result's error is also stated against.
 *
 * It operates on the padded layout of include/sofi_types.h and the arrays
 * the host library builds.  Each function names the reference code it
 * restates:
 *   time_loop             src/time_loop.c:27-326 (phase order of one step)
 *   force / adjoint injection  src/update_v_interior.c:43-90
 *   velocity update       src/update_v_interior.c:92-98, src/wavefield_update_v.c:27-48
 *   velocity CPML frame   src/update_v_PML.c:33-227, src/cpml_update.c:50-66
 *   velocity damping frame src/update_v_abs.c:32-192, src/abs_update.c:46-50
 *   FD operators          src/initfd.c:9-183 (op_s_fd*, op_v_fd*)
 *   v_derivatives         src/v_derivatives.c:27-50
 *   stress updates        src/wavefield_update_s_{el,visc,el_vti,visc_VTI,el_tti,visc_TTI}.c
 *   stress frames         src/update_s_*_{interior,PML,abs}.c, src/cpml_update.c:30-46
 *   psource / eqsource    src/psource.c:27-54, src/eqsource.c:25-52
 *   free surface          src/surface_elastic.c:27-90, src/surface.c:27-113
 *   seismo_ssg            src/seismo_ssg.c:28-67
 *   snap                  src/snap_ssg.c:28-209
 *   snap_store / calc_conv src/snap_store.c:27-47, src/calc_conv.c:30-73
 *   FDORDER_TIME = 4      src/time_loop.c:78-92,215-245, src/update_v_interior_4.c,
 *                         src/update_v_{PML,abs}_4.c, src/wavefield_update_v_4.c,
 *                         src/update_s_{elastic,visc}_{interior,PML,abs}_4.c,
 *                         src/wavefield_update_s_{el,visc}_4.c,
 *                         src/prepare_update_s_{el,visc}_4.c, src/util.c:89-109
 *
 * Every piece is checked in tests/test_oracle.py and tests/test_host.py
 * against the reference's golden files, the reference build in oracle/_ref
 * and closed forms.  This file is test infrastructure: only the tests, the
 * smoke check and the benchmark's CPU baseline use it, as the checker.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sofi_types.h"

#ifndef REAL
#define REAL float
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct oracle {
    sofi_desc d;
    int nx, ny, nd, pitch, L;
    size_t plane;
    REAL hc[7];
    REAL dhi;
    /* wavefield */
    REAL *vx, *vy, *sxx, *syy, *sxy;
    REAL *vxx, *vyy, *vyx, *vxy;       /* pvxx, pvyy, pvyx, pvxy */
    REAL *r, *p, *q;                   /* memory variables, L planes */
    /* CPML memory: x arrays [j 0..NY][h 0..2FW], y arrays [h 0..2FW][i 0..NX] */
    REAL *psi_sxx_x, *psi_sxy_x, *psi_syy_y, *psi_sxy_y;
    REAL *psi_vxx, *psi_vyx, *psi_vyy, *psi_vxy;
    int wx, wy;                        /* row lengths of the x and y psi arrays */
    /* materials (REAL copies) */
    REAL *a[SOFI_A_NARRAYS];
    REAL *prof[SOFI_P_NPROFILES];
    /* acquisition */
    int nsrc, *si, *sj, *stype;
    REAL *sazi, *samp;
    REAL *sig;                         /* [nsrc][nt], step nt at sig[k*nt + nt-1] */
    int ntr, *ri, *rj;
    REAL *sec[SOFI_NCOMP];             /* [ntr][ns] */
    /* snapshots */
    int nsnap_max, nsnap, nxs, nys;
    REAL *snapbuf;                     /* [nsnap][SOFI_NCOMP][nxs*nys] */
    int lsnap, lsamp;
    /* FWI */
    int sw;                            /* 0 forward, 1 adjoint */
    REAL *ux, *uy, *uxy, *pvxp1, *pvyp1;
    int nstore, hin, hin1;
    int conv_shift;                    /* test knob: the adjoint pass correlates with slot + conv_shift (0 = the schedule) */
    REAL *store;                       /* forward store [nstore][5][ny][nx]: pvxp1 pvyp1 sxx syy sxy (or ux uy uxy) */
    REAL *grad_lam, *grad_mu, *grad_rho;   /* gradLam_shot, gradMu_shot, gradRhos_shot */
    REAL *gsum[3];                     /* gradVp, gradVs, gradRho summed over shots */
    REAL *vp0, *vs0, *rho0, *muss, *lamss;
    int nadj, *ai, *aj;                /* adjoint sources (the receivers) */
    REAL *adj_x, *adj_y;               /* adjoint signals [nadj][nt] */
    /* FDORDER_TIME = 4: the Adams-Bashforth histories, [0] the newest;
     * rotated after each update as shift_var2 / shift_var3 rotate pointers */
    int fdt4;
    REAL *sv[2][4];                    /* svx_1..4, svy_1..4 */
    REAL *dvh[4][4];                   /* vxx_1..4, vyy_1..4, vxy_1..4, vyx_1..4 */
    REAL *mh[3][4];                    /* (pr, pr_2..4), (pp, ..), (pq, ..): L planes each */
} oracle;

#define IX(i, j) ((size_t)((j) - 1 + SOFI_PAD) * (size_t)o->pitch + (size_t)((i) - 1 + SOFI_PAD))
#define F(a, i, j) ((a)[IX(i, j)])
#define M(id, i, j) (o->a[id][IX(i, j)])
#define M3(id, i, j, l) (o->a[id][(size_t)((l) - 1) * o->plane + IX(i, j)])
#define MV(arr, i, j, l) ((arr)[(size_t)((l) - 1) * o->plane + IX(i, j)])
#define PX(arr, j, h) ((arr)[(size_t)(j) * (size_t)o->wx + (size_t)(h)])
#define PY(arr, h, i) ((arr)[(size_t)(h) * (size_t)o->wy + (size_t)(i)])
#define PR(id, k) (o->prof[id][k])

static REAL *zalloc(size_t n) { return (REAL *)calloc(n ? n : 1, sizeof(REAL)); }

/* ------------------------------------------------------------ lifecycle */

oracle *oracle_new(const sofi_desc *d, const float *const *arrays, const float *const *prof)
{
    oracle *o = (oracle *)calloc(1, sizeof(oracle));
    o->d = *d;
    o->nx = d->nx;
    o->ny = d->ny;
    o->nd = d->fdorder / 2;
    o->pitch = sofi_pitch(d->nx);
    o->plane = sofi_plane(d->nx, d->ny);
    o->L = d->L;
    for (int k = 0; k < 7; ++k) o->hc[k] = (REAL)d->hc[k];
    o->dhi = (REAL)1 / (REAL)d->dh; /* src/initfd.c:366: 1.f / DH */
    o->vx = zalloc(o->plane); o->vy = zalloc(o->plane);
    o->sxx = zalloc(o->plane); o->syy = zalloc(o->plane); o->sxy = zalloc(o->plane);
    if (d->weq == SOFI_AC_ISO) {   /* one pressure array: psyy is psxx (src/initmem_wavefield.c:31-33) */
        free(o->syy);
        o->syy = o->sxx;
    }
    o->vxx = zalloc(o->plane); o->vyy = zalloc(o->plane); o->vyx = zalloc(o->plane); o->vxy = zalloc(o->plane);
    int L = o->L > 0 ? o->L : 1;
    o->r = zalloc(o->plane * L); o->p = zalloc(o->plane * L); o->q = zalloc(o->plane * L);
    int fw = d->fw;
    o->wx = 2 * fw + 1;
    o->wy = o->nx + 1;
    size_t nxp = (size_t)(o->ny + 1) * o->wx, nyp = (size_t)(2 * fw + 1) * o->wy;
    o->psi_sxx_x = zalloc(nxp); o->psi_sxy_x = zalloc(nxp); o->psi_vxx = zalloc(nxp); o->psi_vyx = zalloc(nxp);
    o->psi_syy_y = zalloc(nyp); o->psi_sxy_y = zalloc(nyp); o->psi_vyy = zalloc(nyp); o->psi_vxy = zalloc(nyp);
    for (int k = 0; k < SOFI_A_NARRAYS; ++k) {
        if (!arrays[k]) continue;
        int planes = (k == SOFI_A_DIP || k == SOFI_A_D || k == SOFI_A_E || (k >= SOFI_A_C11D && k <= SOFI_A_C35IPJPD)) ? L : 1;
        size_t n = o->plane * planes;
        o->a[k] = zalloc(n);
        for (size_t q = 0; q < n; ++q) o->a[k][q] = (REAL)arrays[k][q];
    }
    for (int k = 0; k < SOFI_P_NPROFILES; ++k) {
        if (!prof || !prof[k]) continue;
        int n = (k >= SOFI_P_ETA) ? (o->L + 1) : (2 * fw + 1);
        o->prof[k] = zalloc(n);
        for (int q = 0; q < n; ++q) o->prof[k][q] = (REAL)prof[k][q];
    }
    o->lsnap = d->lsnap;
    o->lsamp = d->ndt;
    o->nxs = (d->snap_idx[2] - d->snap_idx[1]) / (d->idx > 0 ? d->idx : 1) + 1;
    o->nys = (d->snap_idx[4] - d->snap_idx[3]) / (d->idy > 0 ? d->idy : 1) + 1;
    o->hin = 1;
    o->hin1 = 1;
    o->fdt4 = d->fdorder_time == 4;
    if (o->fdt4) {
        for (int a = 0; a < 2; ++a)
            for (int k = 0; k < 4; ++k) o->sv[a][k] = zalloc(o->plane);
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 4; ++k) o->dvh[a][k] = zalloc(o->plane);
        REAL *cur[3] = { o->r, o->p, o->q };
        for (int a = 0; a < 3; ++a) {
            o->mh[a][0] = cur[a];
            for (int k = 1; k < 4; ++k) o->mh[a][k] = zalloc(o->plane * L);
        }
    }
    if (d->mode == SOFI_MODE_FWI) {
        o->ux = zalloc(o->plane); o->uy = zalloc(o->plane); o->uxy = zalloc(o->plane);
        o->pvxp1 = zalloc(o->plane); o->pvyp1 = zalloc(o->plane);
        o->grad_lam = zalloc(o->plane); o->grad_mu = zalloc(o->plane); o->grad_rho = zalloc(o->plane);
        for (int k = 0; k < 3; ++k) o->gsum[k] = zalloc(o->plane);
        o->vp0 = zalloc(o->plane); o->vs0 = zalloc(o->plane); o->rho0 = zalloc(o->plane);
        o->muss = zalloc(o->plane); o->lamss = zalloc(o->plane);
        int dtinv = d->dtinv > 0 ? d->dtinv : 1;
        o->nstore = (d->nt + dtinv - 1) / dtinv;
        o->store = zalloc((size_t)o->nstore * 5 * o->nx * o->ny);
    }
    return o;
}

void oracle_set_sources(oracle *o, int nsrc, const int *si, const int *sj, const int *stype, const float *sazi,
                        const float *samp, const float *signals)
{
    o->nsrc = nsrc;
    o->si = (int *)malloc(sizeof(int) * (nsrc + 1));
    o->sj = (int *)malloc(sizeof(int) * (nsrc + 1));
    o->stype = (int *)malloc(sizeof(int) * (nsrc + 1));
    o->sazi = zalloc(nsrc + 1);
    o->samp = zalloc(nsrc + 1);
    o->sig = zalloc((size_t)(nsrc > 0 ? nsrc : 1) * o->d.nt);
    for (int k = 0; k < nsrc; ++k) {
        o->si[k] = si[k]; o->sj[k] = sj[k]; o->stype[k] = stype[k];
        o->sazi[k] = sazi[k]; o->samp[k] = samp[k];
    }
    for (size_t q = 0; q < (size_t)nsrc * o->d.nt; ++q) o->sig[q] = signals[q];
}

void oracle_set_receivers(oracle *o, int ntr, const int *ri, const int *rj)
{
    o->ntr = ntr;
    o->ri = (int *)malloc(sizeof(int) * (ntr + 1));
    o->rj = (int *)malloc(sizeof(int) * (ntr + 1));
    for (int k = 0; k < ntr; ++k) { o->ri[k] = ri[k]; o->rj[k] = rj[k]; }
    for (int c = 0; c < SOFI_NCOMP; ++c) o->sec[c] = zalloc((size_t)(ntr > 0 ? ntr : 1) * (o->d.ns > 0 ? o->d.ns : 1));
}

void oracle_set_snapshots(oracle *o, int nsnap_max)
{
    o->nsnap_max = nsnap_max;
    o->snapbuf = zalloc((size_t)(nsnap_max > 0 ? nsnap_max : 1) * SOFI_NCOMP * o->nxs * o->nys);
}

/* ------------------------------------------------------- FD operators */

/* op_v_fd*: stress derivatives without 1/DH (src/initfd.c:97-183) */
static void op_v(const oracle *o, int i, int j, REAL *sxx_x, REAL *sxy_x, REAL *sxy_y, REAL *syy_y)
{
    const REAL *hc = o->hc;
    REAL a = 0, b = 0, c = 0, e = 0;
    for (int m = 1; m <= o->nd; ++m) {
        a += hc[m] * (F(o->sxx, i + m, j) - F(o->sxx, i - m + 1, j));
        b += hc[m] * (F(o->sxy, i + m - 1, j) - F(o->sxy, i - m, j));
        c += hc[m] * (F(o->sxy, i, j + m - 1) - F(o->sxy, i, j - m));
        e += hc[m] * (F(o->syy, i, j + m) - F(o->syy, i, j - m + 1));
    }
    *sxx_x = a; *sxy_x = b; *sxy_y = c; *syy_y = e;
}

/* op_s_fd*: velocity derivatives scaled by 1/DH (src/initfd.c:9-95) */
static void op_s(const oracle *o, int i, int j, REAL *vxx, REAL *vyx, REAL *vxy, REAL *vyy)
{
    const REAL *hc = o->hc;
    REAL a = 0, b = 0, c = 0, e = 0;
    for (int m = 1; m <= o->nd; ++m) {
        a += hc[m] * (F(o->vx, i + m - 1, j) - F(o->vx, i - m, j));
        b += hc[m] * (F(o->vy, i + m, j) - F(o->vy, i - m + 1, j));
        c += hc[m] * (F(o->vx, i, j + m) - F(o->vx, i, j - m + 1));
        e += hc[m] * (F(o->vy, i, j + m - 1) - F(o->vy, i, j - m));
    }
    *vxx = a * o->dhi; *vyx = b * o->dhi; *vxy = c * o->dhi; *vyy = e * o->dhi;
}

/* ----------------------------------------------------- velocity phase */

static void wavefield_update_v(oracle *o, int i, int j, REAL sxx_x, REAL sxy_x, REAL sxy_y, REAL syy_y)
{
    REAL dtdh = (REAL)o->d.dt / (REAL)o->d.dh; /* DTDH = DT / DH, src/prepmod.c:37 */
    if (o->d.mode == SOFI_MODE_FWI) {
        if (o->sw == 0) {
            F(o->pvxp1, i, j) = M(SOFI_A_RIP, i, j) * (sxx_x + sxy_y) / o->d.dh;
            F(o->pvyp1, i, j) = M(SOFI_A_RJP, i, j) * (sxy_x + syy_y) / o->d.dh;
        } else if (o->d.velocity == 0) {
            F(o->pvxp1, i, j) += F(o->vx, i, j) * o->d.dt;
            F(o->pvyp1, i, j) += F(o->vy, i, j) * o->d.dt;
        } else {
            F(o->pvxp1, i, j) = F(o->vx, i, j);
            F(o->pvyp1, i, j) = F(o->vy, i, j);
        }
    }
    F(o->vx, i, j) += (sxx_x + sxy_y) * dtdh * M(SOFI_A_RIP, i, j);
    F(o->vy, i, j) += (sxy_x + syy_y) * dtdh * M(SOFI_A_RJP, i, j);
}

static void cpml_v_x(oracle *o, int h, int j, REAL *sxx_x, REAL *sxy_x)
{
    PX(o->psi_sxx_x, j, h) = PR(SOFI_P_B_X_HALF, h) * PX(o->psi_sxx_x, j, h) + PR(SOFI_P_A_X_HALF, h) * (*sxx_x);
    *sxx_x = (*sxx_x) / PR(SOFI_P_K_X_HALF, h) + PX(o->psi_sxx_x, j, h);
    PX(o->psi_sxy_x, j, h) = PR(SOFI_P_B_X, h) * PX(o->psi_sxy_x, j, h) + PR(SOFI_P_A_X, h) * (*sxy_x);
    *sxy_x = (*sxy_x) / PR(SOFI_P_K_X, h) + PX(o->psi_sxy_x, j, h);
}

static void cpml_v_y(oracle *o, int i, int h, REAL *sxy_y, REAL *syy_y)
{
    PY(o->psi_syy_y, h, i) = PR(SOFI_P_B_Y_HALF, h) * PY(o->psi_syy_y, h, i) + PR(SOFI_P_A_Y_HALF, h) * (*syy_y);
    *syy_y = (*syy_y) / PR(SOFI_P_K_Y_HALF, h) + PY(o->psi_syy_y, h, i);
    PY(o->psi_sxy_y, h, i) = PR(SOFI_P_B_Y, h) * PY(o->psi_sxy_y, h, i) + PR(SOFI_P_A_Y, h) * (*sxy_y);
    *sxy_y = (*sxy_y) / PR(SOFI_P_K_Y, h) + PY(o->psi_sxy_y, h, i);
}

/* one frame cell of update_v_PML / update_v_abs */
static void v_cell(oracle *o, int i, int j)
{
    const sofi_desc *d = &o->d;
    REAL sxx_x, sxy_x, sxy_y, syy_y;
    op_v(o, i, j, &sxx_x, &sxy_x, &sxy_y, &syy_y);
    int xf = (i <= d->gx[2]) || (i > d->gx[3]);
    int yf = (j <= d->gy[2]) || (j > d->gy[3]);
    if (d->abs_type == 1 && (xf || yf)) {
        if (xf) cpml_v_x(o, i <= d->gx[2] ? i : i - o->nx + 2 * d->fw, j, &sxx_x, &sxy_x);
        if (yf) cpml_v_y(o, i, j <= d->gy[2] ? j : j - o->ny + 2 * d->fw, &sxy_y, &syy_y);
    }
    wavefield_update_v(o, i, j, sxx_x, sxy_x, sxy_y, syy_y);
    if (d->abs_type == 2 && (xf || yf)) {
        F(o->vx, i, j) *= M(SOFI_A_ABSORB, i, j);
        F(o->vy, i, j) *= M(SOFI_A_ABSORB, i, j);
    }
}

static void update_v(oracle *o, int nt)
{
    const sofi_desc *d = &o->d;
    if (o->sw == 0) {
        for (int l = 0; l < o->nsrc; l++) {
            int i = o->si[l], j = o->sj[l];
            float azi_rad = (float)(o->sazi[l] * M_PI / 180);
            REAL amp = (REAL)((d->dt * o->sig[(size_t)l * d->nt + nt - 1]) / (d->dh * d->dh));
            switch (o->stype[l]) {
              case 2: F(o->vx, i, j) += M(SOFI_A_RIP, i, j) * amp; break;
              case 3: F(o->vy, i, j) += M(SOFI_A_RJP, i, j) * amp; break;
              case 4:
                  F(o->vx, i, j) += sin(azi_rad) * (M(SOFI_A_RIP, i, j) * amp);
                  F(o->vy, i, j) += cos(azi_rad) * (M(SOFI_A_RJP, i, j) * amp);
                  break;
            }
        }
    } else {
        for (int l = 0; l < o->nadj; l++) {
            int i = o->ai[l], j = o->aj[l];
            REAL fx = o->adj_x[(size_t)l * d->nt + nt - 1], fy = o->adj_y[(size_t)l * d->nt + nt - 1];
            switch (d->adjoint_type) {
              case 1: F(o->vx, i, j) += fx; F(o->vy, i, j) += fy; break;
              case 2: F(o->vy, i, j) += fy; break;
              case 3: F(o->vx, i, j) += fx; break;
            }
        }
    }
    /* interior, then the frame (update_v_interior, update_v_PML/abs) */
    for (int j = d->gy[2] + 1; j <= d->gy[3]; j++)
        for (int i = d->gx[2] + 1; i <= d->gx[3]; i++) {
            REAL sxx_x, sxy_x, sxy_y, syy_y;
            op_v(o, i, j, &sxx_x, &sxy_x, &sxy_y, &syy_y);
            wavefield_update_v(o, i, j, sxx_x, sxy_x, sxy_y, syy_y);
        }
    /* the frame is updated only by update_v_PML (ABS_TYPE 1) or update_v_abs
     * (ABS_TYPE 2); with any other ABS_TYPE its cells are never updated
     * (src/time_loop.c:64-71) */
    if (d->fw && (d->abs_type == 1 || d->abs_type == 2)) {
        for (int j = d->gy[1]; j <= d->gy[4]; j++)
            for (int i = d->gx[1]; i <= d->gx[4]; i++) {
                int inner = (i > d->gx[2] && i <= d->gx[3] && j > d->gy[2] && j <= d->gy[3]);
                if (!inner) v_cell(o, i, j);
            }
    }
}

/* ------------------------------------------------------ stress phase */

static void v_derivatives(oracle *o)
{
    for (int j = 1; j <= o->ny; j++)
        for (int i = 1; i <= o->nx; i++) {
            REAL a, b, c, e;
            op_s(o, i, j, &a, &b, &c, &e);
            F(o->vxx, i, j) = a;
            F(o->vyy, i, j) = e;
            F(o->vyx, i, j) = b;
            F(o->vxy, i, j) = c;
        }
}

static void cpml_s_x(oracle *o, int i, int j, int h)
{
    PX(o->psi_vxx, j, h) = PR(SOFI_P_B_X, h) * PX(o->psi_vxx, j, h) + PR(SOFI_P_A_X, h) * F(o->vxx, i, j);
    F(o->vxx, i, j) = F(o->vxx, i, j) / PR(SOFI_P_K_X, h) + PX(o->psi_vxx, j, h);
    PX(o->psi_vyx, j, h) = PR(SOFI_P_B_X_HALF, h) * PX(o->psi_vyx, j, h) + PR(SOFI_P_A_X_HALF, h) * F(o->vyx, i, j);
    F(o->vyx, i, j) = F(o->vyx, i, j) / PR(SOFI_P_K_X_HALF, h) + PX(o->psi_vyx, j, h);
}

static void cpml_s_y(oracle *o, int i, int j, int h)
{
    PY(o->psi_vyy, h, i) = PR(SOFI_P_B_Y, h) * PY(o->psi_vyy, h, i) + PR(SOFI_P_A_Y, h) * F(o->vyy, i, j);
    F(o->vyy, i, j) = F(o->vyy, i, j) / PR(SOFI_P_K_Y, h) + PY(o->psi_vyy, h, i);
    PY(o->psi_vxy, h, i) = PR(SOFI_P_B_Y_HALF, h) * PY(o->psi_vxy, h, i) + PR(SOFI_P_A_Y_HALF, h) * F(o->vxy, i, j);
    F(o->vxy, i, j) = F(o->vxy, i, j) / PR(SOFI_P_K_Y_HALF, h) + PY(o->psi_vxy, h, i);
}

static void store_u(oracle *o, int i, int j, REAL u1, REAL u2, REAL u3)
{
    if (o->d.mode == SOFI_MODE_FWI) {
        F(o->uxy, i, j) = u1 / o->d.dt;
        F(o->ux, i, j) = u2 / o->d.dt;
        F(o->uy, i, j) = u3 / o->d.dt;
    }
}

/* wavefield_update_s_ac (src/wavefield_update_s_ac.c:27-33): sxx and syy
 * are one array, so the pressure takes g vxx, then g vyy */
static void s_ac(oracle *o, int i, int j)
{
    F(o->sxx, i, j) += (M(SOFI_A_G, i, j) * F(o->vxx, i, j));
    F(o->syy, i, j) += (M(SOFI_A_G, i, j) * F(o->vyy, i, j));
}

/* wavefield_update_s_el (src/wavefield_update_s_el.c:28-48) */
static void s_el(oracle *o, int i, int j)
{
    REAL vxx = F(o->vxx, i, j), vyy = F(o->vyy, i, j), vyx = F(o->vyx, i, j), vxy = F(o->vxy, i, j);
    REAL u1 = M(SOFI_A_FIPJP, i, j) * (vyx + vxy);
    REAL u2 = (M(SOFI_A_G, i, j) * (vxx + vyy)) - (2.0 * M(SOFI_A_F, i, j) * vyy);
    REAL u3 = (M(SOFI_A_G, i, j) * (vxx + vyy)) - (2.0 * M(SOFI_A_F, i, j) * vxx);
    F(o->sxy, i, j) += u1;
    F(o->sxx, i, j) += u2;
    F(o->syy, i, j) += u3;
    store_u(o, i, j, u1, u2, u3);
}

/* wavefield_update_s_visc (src/wavefield_update_s_visc.c:27-68) */
static void s_visc(oracle *o, int i, int j)
{
    REAL dthalbe = (REAL)(o->d.dt / 2.0);
    REAL sumr_old = 0, sump_old = 0, sumq_old = 0, sumr = 0, sump = 0, sumq = 0;
    REAL vxx = F(o->vxx, i, j), vyy = F(o->vyy, i, j), vyx = F(o->vyx, i, j), vxy = F(o->vxy, i, j);
    const REAL *bip = o->prof[SOFI_P_BIP], *cip = o->prof[SOFI_P_CIP];
    for (int l = 1; l <= o->L; l++) {
        sumr_old += MV(o->r, i, j, l);
        sump_old += MV(o->p, i, j, l);
        sumq_old += MV(o->q, i, j, l);
        MV(o->r, i, j, l) = bip[l] * (MV(o->r, i, j, l) * cip[l] - (M3(SOFI_A_DIP, i, j, l) * (vxy + vyx)));
        MV(o->p, i, j, l) = bip[l] * (MV(o->p, i, j, l) * cip[l] - (M3(SOFI_A_E, i, j, l) * (vxx + vyy)) +
                                     (2.0 * M3(SOFI_A_D, i, j, l) * vyy));
        MV(o->q, i, j, l) = bip[l] * (MV(o->q, i, j, l) * cip[l] - (M3(SOFI_A_E, i, j, l) * (vxx + vyy)) +
                                     (2.0 * M3(SOFI_A_D, i, j, l) * vxx));
        sumr += MV(o->r, i, j, l);
        sump += MV(o->p, i, j, l);
        sumq += MV(o->q, i, j, l);
    }
    REAL u1 = (M(SOFI_A_FIPJP, i, j) * (vxy + vyx)) + (dthalbe * (sumr_old + sumr));
    REAL u2 = (M(SOFI_A_G, i, j) * (vxx + vyy)) - (2.0 * M(SOFI_A_F, i, j) * vyy) + (dthalbe * (sump_old + sump));
    REAL u3 = (M(SOFI_A_G, i, j) * (vxx + vyy)) - (2.0 * M(SOFI_A_F, i, j) * vxx) + (dthalbe * (sumq_old + sumq));
    F(o->sxy, i, j) += u1;
    F(o->sxx, i, j) += u2;
    F(o->syy, i, j) += u3;
    store_u(o, i, j, u1, u2, u3);
}

/* wavefield_update_s_el_vti (src/wavefield_update_s_el_vti.c:27-47) */
static void s_el_vti(oracle *o, int i, int j)
{
    REAL vxx = F(o->vxx, i, j), vyy = F(o->vyy, i, j), vyx = F(o->vyx, i, j), vxy = F(o->vxy, i, j);
    REAL u1 = (M(SOFI_A_C55IPJP, i, j) * (vxy + vyx));
    REAL u2 = ((M(SOFI_A_C11, i, j) * vxx) + (M(SOFI_A_C13, i, j) * vyy));
    REAL u3 = ((M(SOFI_A_C13, i, j) * vxx) + (M(SOFI_A_C33, i, j) * vyy));
    F(o->sxy, i, j) += u1;
    F(o->sxx, i, j) += u2;
    F(o->syy, i, j) += u3;
    store_u(o, i, j, u1, u2, u3);
}

/* wavefield_update_s_visc_VTI (src/wavefield_update_s_visc_VTI.c:27-66) */
static void s_visc_vti(oracle *o, int i, int j)
{
    REAL dthalbe = (REAL)(o->d.dt / 2.0);
    REAL sumr_old = 0, sump_old = 0, sumq_old = 0, sumr = 0, sump = 0, sumq = 0;
    REAL vxx = F(o->vxx, i, j), vyy = F(o->vyy, i, j), vyx = F(o->vyx, i, j), vxy = F(o->vxy, i, j);
    const REAL *bip = o->prof[SOFI_P_BIP], *cip = o->prof[SOFI_P_CIP];
    for (int l = 1; l <= o->L; l++) {
        sumr_old += MV(o->r, i, j, l);
        sump_old += MV(o->p, i, j, l);
        sumq_old += MV(o->q, i, j, l);
        MV(o->r, i, j, l) = bip[l] * (MV(o->r, i, j, l) * cip[l] - (M3(SOFI_A_C55IPJPD, i, j, l) * (vxy + vyx)));
        MV(o->p, i, j, l) = bip[l] * (MV(o->p, i, j, l) * cip[l] - (M3(SOFI_A_C11D, i, j, l) * vxx) -
                                     (M3(SOFI_A_C13D, i, j, l) * vyy));
        MV(o->q, i, j, l) = bip[l] * (MV(o->q, i, j, l) * cip[l] - (M3(SOFI_A_C13D, i, j, l) * vxx) -
                                     (M3(SOFI_A_C33D, i, j, l) * vyy));
        sumr += MV(o->r, i, j, l);
        sump += MV(o->p, i, j, l);
        sumq += MV(o->q, i, j, l);
    }
    REAL u1 = (M(SOFI_A_C55IPJPU, i, j) * (vxy + vyx)) + (dthalbe * (sumr_old + sumr));
    REAL u2 = (M(SOFI_A_C11U, i, j) * vxx) + (M(SOFI_A_C13U, i, j) * vyy) + (dthalbe * (sump_old + sump));
    REAL u3 = (M(SOFI_A_C13U, i, j) * vxx) + (M(SOFI_A_C33U, i, j) * vyy) + (dthalbe * (sumq_old + sumq));
    F(o->sxy, i, j) += u1;
    F(o->sxx, i, j) += u2;
    F(o->syy, i, j) += u3;
    store_u(o, i, j, u1, u2, u3);
}

/* wavefield_update_s_el_tti (src/wavefield_update_s_el_tti.c:27-52) */
static void s_el_tti(oracle *o, int i, int j)
{
    REAL vxxipjp = 0.25 * (F(o->vxx, i, j) + F(o->vxx, i, j + 1) + F(o->vxx, i + 1, j) + F(o->vxx, i + 1, j + 1));
    REAL vyyipjp = 0.25 * (F(o->vyy, i, j) + F(o->vyy, i, j + 1) + F(o->vyy, i + 1, j) + F(o->vyy, i + 1, j + 1));
    REAL vij = (0.25 * (F(o->vyx, i, j) + F(o->vyx, i, j - 1) + F(o->vyx, i - 1, j) + F(o->vyx, i - 1, j - 1))) +
               (0.25 * (F(o->vxy, i, j) + F(o->vxy, i, j - 1) + F(o->vxy, i - 1, j) + F(o->vxy, i - 1, j - 1)));
    REAL v = F(o->vxy, i, j) + F(o->vyx, i, j);
    REAL u1 = ((M(SOFI_A_C55IPJP, i, j) * v) + (M(SOFI_A_C15IPJP, i, j) * vxxipjp) + (M(SOFI_A_C35IPJP, i, j) * vyyipjp));
    REAL u2 = ((M(SOFI_A_C11, i, j) * F(o->vxx, i, j)) + (M(SOFI_A_C13, i, j) * F(o->vyy, i, j)) + (M(SOFI_A_C15, i, j) * vij));
    REAL u3 = ((M(SOFI_A_C13, i, j) * F(o->vxx, i, j)) + (M(SOFI_A_C33, i, j) * F(o->vyy, i, j)) + (M(SOFI_A_C35, i, j) * vij));
    F(o->sxy, i, j) += u1;
    F(o->sxx, i, j) += u2;
    F(o->syy, i, j) += u3;
    store_u(o, i, j, u1, u2, u3);
}

/* wavefield_update_s_visc_TTI (src/wavefield_update_s_visc_TTI.c:27-84) */
static void s_visc_tti(oracle *o, int i, int j)
{
    REAL dthalbe = (REAL)(o->d.dt / 2.0);
    REAL sumr_old = 0, sump_old = 0, sumq_old = 0, sumr = 0, sump = 0, sumq = 0;
    REAL vxxipjp = 0.25 * (F(o->vxx, i, j) + F(o->vxx, i, j + 1) + F(o->vxx, i + 1, j) + F(o->vxx, i + 1, j + 1));
    REAL vyyipjp = 0.25 * (F(o->vyy, i, j) + F(o->vyy, i, j + 1) + F(o->vyy, i + 1, j) + F(o->vyy, i + 1, j + 1));
    REAL vyxij = 0.25 * (F(o->vyx, i, j) + F(o->vyx, i, j - 1) + F(o->vyx, i - 1, j) + F(o->vyx, i - 1, j - 1));
    REAL vxyij = 0.25 * (F(o->vxy, i, j) + F(o->vxy, i, j - 1) + F(o->vxy, i - 1, j) + F(o->vxy, i - 1, j - 1));
    REAL vij = vyxij + vxyij;
    REAL v = F(o->vxy, i, j) + F(o->vyx, i, j);
    REAL vxx = F(o->vxx, i, j), vyy = F(o->vyy, i, j);
    const REAL *bip = o->prof[SOFI_P_BIP], *cip = o->prof[SOFI_P_CIP];
    for (int l = 1; l <= o->L; l++) {
        sumr_old += MV(o->r, i, j, l);
        sump_old += MV(o->p, i, j, l);
        sumq_old += MV(o->q, i, j, l);
        MV(o->r, i, j, l) = bip[l] * (MV(o->r, i, j, l) * cip[l] - (M3(SOFI_A_C55IPJPD, i, j, l) * v) -
                                     (M3(SOFI_A_C15IPJPD, i, j, l) * vxxipjp) - (M3(SOFI_A_C35IPJPD, i, j, l) * vyyipjp));
        MV(o->p, i, j, l) = bip[l] * (MV(o->p, i, j, l) * cip[l] - (M3(SOFI_A_C11D, i, j, l) * vxx) -
                                     (M3(SOFI_A_C13D, i, j, l) * vyy) - (M3(SOFI_A_C15D, i, j, l) * vij));
        MV(o->q, i, j, l) = bip[l] * (MV(o->q, i, j, l) * cip[l] - (M3(SOFI_A_C13D, i, j, l) * vxx) -
                                     (M3(SOFI_A_C33D, i, j, l) * vyy) - (M3(SOFI_A_C35D, i, j, l) * vij));
        sumr += MV(o->r, i, j, l);
        sump += MV(o->p, i, j, l);
        sumq += MV(o->q, i, j, l);
    }
    REAL u1 = (M(SOFI_A_C55IPJPU, i, j) * v) + (M(SOFI_A_C15IPJPU, i, j) * vxxipjp) + (M(SOFI_A_C35IPJPU, i, j) * vyyipjp) +
              (dthalbe * (sumr_old + sumr));
    REAL u2 = (M(SOFI_A_C11U, i, j) * vxx) + (M(SOFI_A_C13U, i, j) * vyy) + (M(SOFI_A_C15U, i, j) * vij) + (dthalbe * (sump_old + sump));
    REAL u3 = (M(SOFI_A_C13U, i, j) * vxx) + (M(SOFI_A_C33U, i, j) * vyy) + (M(SOFI_A_C35U, i, j) * vij) + (dthalbe * (sumq_old + sumq));
    F(o->sxy, i, j) += u1;
    F(o->sxx, i, j) += u2;
    F(o->syy, i, j) += u3;
    store_u(o, i, j, u1, u2, u3);
}

static void s_cell(oracle *o, int i, int j)
{
    switch (o->d.weq) {
      case SOFI_AC_ISO: s_ac(o, i, j); break;
      case SOFI_EL_ISO: s_el(o, i, j); break;
      case SOFI_VEL_ISO: s_visc(o, i, j); break;
      case SOFI_EL_VTI: s_el_vti(o, i, j); break;
      case SOFI_VEL_VTI: s_visc_vti(o, i, j); break;
      case SOFI_EL_TTI: s_el_tti(o, i, j); break;
      case SOFI_VEL_TTI: s_visc_tti(o, i, j); break;
    }
}

static void abs_s(oracle *o, int i, int j)
{
    REAL c = M(SOFI_A_ABSORB, i, j);
    if (o->d.weq == SOFI_AC_ISO) {   /* abs_update_s_ac2 (src/abs_update.c:28-31): the pressure once */
        F(o->sxx, i, j) *= c;
        return;
    }
    F(o->sxy, i, j) *= c;
    F(o->syy, i, j) *= c;
    F(o->sxx, i, j) *= c;
}

/* CPML on the stress-side derivatives of a frame cell (cpml_update_s_x/_y) */
static void cpml_s_cell(oracle *o, int i, int j)
{
    const sofi_desc *d = &o->d;
    int xf = (i <= d->gx[2]) || (i > d->gx[3]);
    int yf = (j <= d->gy[2]) || (j > d->gy[3]);
    if (xf) cpml_s_x(o, i, j, i <= d->gx[2] ? i : i - o->nx + 2 * d->fw);
    if (yf) cpml_s_y(o, i, j, j <= d->gy[2] ? j : j - o->ny + 2 * d->fw);
}

static int in_frame(const oracle *o, int i, int j)
{
    const sofi_desc *d = &o->d;
    return !(i > d->gx[2] && i <= d->gx[3] && j > d->gy[2] && j <= d->gy[3]);
}

static void update_s(oracle *o)
{
    const sofi_desc *d = &o->d;
    int tti = (d->weq == SOFI_EL_TTI || d->weq == SOFI_VEL_TTI);
    for (int j = d->gy[2] + 1; j <= d->gy[3]; j++)
        for (int i = d->gx[2] + 1; i <= d->gx[3]; i++) s_cell(o, i, j);
    if (!d->fw) return;
    if (d->abs_type == 1 && !tti) {
        for (int j = d->gy[1]; j <= d->gy[4]; j++)
            for (int i = d->gx[1]; i <= d->gx[4]; i++)
                if (in_frame(o, i, j)) {
                    cpml_s_cell(o, i, j);
                    s_cell(o, i, j);
                }
    } else if (d->abs_type == 1 && tti) {
        /* update_s_*_tti_PML works region by region, in the order left,
         * right, top, bottom, left-top, left-bottom, right-top, right-bottom:
         * the CPML over the region, then the stress update on the region
         * shrunk by two cells at the outer edges.  Side cells therefore read
         * corner derivatives before the corner CPML has modified them.  The
         * elastic right-bottom CPML loop stops at GY[4]-1
         * (src/update_s_elastic_TTI_PML.c:32-145, src/update_s_visc_TTI_PML.c:132). */
        const int X[3][2] = { { d->gx[1], d->gx[2] }, { d->gx[2] + 1, d->gx[3] }, { d->gx[3] + 1, d->gx[4] } };
        const int Y[3][2] = { { d->gy[1], d->gy[2] }, { d->gy[2] + 1, d->gy[3] }, { d->gy[3] + 1, d->gy[4] } };
        /* (x band, y band): 0 = low frame, 1 = inner, 2 = high frame */
        const int order[8][2] = { { 0, 1 }, { 2, 1 }, { 1, 0 }, { 1, 2 }, { 0, 0 }, { 0, 2 }, { 2, 0 }, { 2, 2 } };
        for (int r = 0; r < 8; ++r) {
            int bx = order[r][0], by = order[r][1];
            int jc1 = Y[by][1];
            if (d->weq == SOFI_EL_TTI && bx == 2 && by == 2) jc1 = d->gy[4] - 1;
            for (int j = Y[by][0]; j <= jc1; j++)
                for (int i = X[bx][0]; i <= X[bx][1]; i++) cpml_s_cell(o, i, j);
            int i0 = X[bx][0] + (bx == 0 ? 2 : 0), i1 = X[bx][1] - (bx == 2 ? 2 : 0);
            int j0 = Y[by][0] + (by == 0 ? 2 : 0), j1 = Y[by][1] - (by == 2 ? 2 : 0);
            for (int j = j0; j <= j1; j++)
                for (int i = i0; i <= i1; i++) s_cell(o, i, j);
        }
    } else if (d->abs_type == 2) {
        for (int j = d->gy[1]; j <= d->gy[4]; j++)
            for (int i = d->gx[1]; i <= d->gx[4]; i++)
                if (in_frame(o, i, j)) {
                    s_cell(o, i, j);
                    abs_s(o, i, j);
                }
    }
}

/* psource (src/psource.c:27-54) */
static void psource(oracle *o, int nt, const REAL *sig)
{
    const sofi_desc *d = &o->d;
    for (int l = 0; l < o->nsrc; l++) {
        int i = o->si[l], j = o->sj[l];
        const REAL *s = sig + (size_t)l * d->nt - 1; /* s[nt] = signal at step nt */
        REAL amp = (s[nt]) / (d->dh * d->dh);
        if (nt == 1) amp = s[nt + 1] / (2.0 * d->dh * d->dh);
        if ((nt > 1) && (nt < d->nt)) amp = (s[nt + 1] - s[nt - 1]) / (2.0 * d->dh * d->dh);
        if (nt == d->nt) amp = -s[nt - 1] / (2.0 * d->dh * d->dh);
        F(o->sxx, i, j) += amp;
        F(o->syy, i, j) += amp;
    }
}

/* eqsource (src/eqsource.c:25-52) */
static void eqsource(oracle *o, int nt)
{
    const sofi_desc *d = &o->d;
    for (int l = 0; l < o->nsrc; l++) {
        int i = o->si[l], j = o->sj[l];
        REAL dip = (REAL)(o->sazi[l] * M_PI / 180.0);
        REAL scale_amp = d->dt / (d->dh * d->dh);
        REAL amp = o->samp[l] * scale_amp * o->sig[(size_t)l * d->nt + nt - 1];
        REAL m11 = -sin(2 * dip), m12 = -cos(2 * dip), m22 = sin(2 * dip);
        F(o->sxx, i, j) += amp * m11;
        F(o->sxy, i, j) += 0.3 * amp * m12;
        F(o->sxy, i, j - 1) += 0.3 * amp * m12;
        F(o->sxy, i - 1, j) += 0.3 * amp * m12;
        F(o->syy, i, j) += amp * m22;
    }
}

/* surface_elastic / surface (src/surface_elastic.c:27-90, src/surface.c:27-113) */
static void free_surface(oracle *o)
{
    const sofi_desc *d = &o->d;
    int fdoh = o->nd, j = 1;
    REAL dhi = (REAL)(1.0 / d->dh);
    REAL dthalbe = (REAL)(d->dt / 2.0);
    int visco = o->L > 0;
    for (int i = 1; i <= o->nx; i++) {
        F(o->syy, i, j) = 0.0;
        if (visco)
            for (int l = 1; l <= o->L; l++) MV(o->q, i, j, l) = 0.0;
        REAL vxx = 0.0, vyy = 0.0;
        for (int m = 1; m <= fdoh; m++) {
            F(o->syy, i, j - m) = -F(o->syy, i, j + m);
            F(o->sxy, i, j - m) = -F(o->sxy, i, j + m - 1);
            vxx += o->hc[m] * (F(o->vx, i + m - 1, j) - F(o->vx, i - m, j));
            vyy += o->hc[m] * (F(o->vy, i, j + m - 1) - F(o->vy, i, j - m));
        }
        vxx *= dhi;
        vyy *= dhi;
        if (d->abs_type == 1 && !d->boundary) {   /* no x-CPML at a periodic edge (src/surface_elastic.c:64-78) */
            if (i <= d->fw) {
                PX(o->psi_vxx, j, i) = PR(SOFI_P_B_X, i) * PX(o->psi_vxx, j, i) + PR(SOFI_P_A_X, i) * vxx;
                vxx = vxx / PR(SOFI_P_K_X, i) + PX(o->psi_vxx, j, i);
            }
            if (i >= o->nx - d->fw + 1) {
                int h1 = i - o->nx + 2 * d->fw;
                PX(o->psi_vxx, j, h1) = PR(SOFI_P_B_X, h1) * PX(o->psi_vxx, j, h1) + PR(SOFI_P_A_X, h1) * vxx;
                vxx = vxx / PR(SOFI_P_K_X, h1) + PX(o->psi_vxx, j, h1);
            }
        }
        if (!visco) {
            REAL fjm = M(SOFI_A_PU, i, j) * 2.0;
            REAL g = M(SOFI_A_PPI, i, j);
            F(o->sxx, i, j) += -(d->dt * (g - fjm) * (g - fjm) * vxx / g) - (d->dt * (g - fjm) * vyy);
        } else {
            REAL sump = 0.0;
            for (int l = 1; l <= o->L; l++) sump += MV(o->p, i, j, l);
            REAL fjm = M(SOFI_A_PU, i, j) * 2.0 * (1.0 + o->L * M(SOFI_A_PTAUS, i, j));
            REAL g = M(SOFI_A_PPI, i, j) * (1.0 + o->L * M(SOFI_A_PTAUP, i, j));
            F(o->sxx, i, j) += -(d->dt * (g - fjm) * (g - fjm) * vxx / g) - (d->dt * (g - fjm) * vyy) - (dthalbe * sump);
            sump = 0.0;
            for (int l = 1; l <= o->L; l++) {
                REAL eta = o->prof[SOFI_P_ETA][l];
                REAL bjm = eta / (1.0 + (eta * 0.5));
                REAL djm = 2.0 * M(SOFI_A_PU, i, j) * M(SOFI_A_PTAUS, i, j);
                REAL e = M(SOFI_A_PPI, i, j) * M(SOFI_A_PTAUP, i, j);
                MV(o->p, i, j, l) += bjm * (((djm - e) * ((fjm / g) - 1.0) * vxx) - ((djm - e) * vyy));
                sump += MV(o->p, i, j, l);
            }
            F(o->sxx, i, j) += (dthalbe * sump);
        }
    }
}

/* seismo_ssg (src/seismo_ssg.c:28-67) */
static void seismo(oracle *o, int lsamp)
{
    const sofi_desc *d = &o->d;
    int ins = lsamp / d->ndt;
    for (int k = 0; k < o->ntr; k++) {
        int i = o->ri[k], j = o->rj[k];
        size_t s = (size_t)k * d->ns + (ins - 1);
        /* (the acoustic path has no pu; it records SEISMO 1 and 2 only) */
        REAL ppi = M(SOFI_A_PPI, i, j), pu = o->a[SOFI_A_PU] ? M(SOFI_A_PU, i, j) : 0;
        switch (d->seismo) {
          case 1:
              o->sec[SOFI_C_VX][s] = F(o->vx, i, j);
              o->sec[SOFI_C_VY][s] = F(o->vy, i, j);
              break;
          case 2:
              o->sec[SOFI_C_P][s] = -F(o->sxx, i, j) - F(o->syy, i, j);
              break;
          case 3:
              o->sec[SOFI_C_DIV][s] = (F(o->vxx, i, j) + F(o->vyy, i, j)) * sqrt(ppi);
              o->sec[SOFI_C_CURL][s] = (F(o->vxy, i, j) - F(o->vyx, i, j)) * sqrt(pu);
              break;
          case 4:
              o->sec[SOFI_C_DIV][s] = (F(o->vxx, i, j) + F(o->vyy, i, j)) * sqrt(ppi);
              o->sec[SOFI_C_CURL][s] = (F(o->vxy, i, j) - F(o->vyx, i, j)) * sqrt(pu);
              o->sec[SOFI_C_VX][s] = F(o->vx, i, j);
              o->sec[SOFI_C_VY][s] = F(o->vy, i, j);
              o->sec[SOFI_C_P][s] = ((3.0 * ppi - 4.0 * pu) / (2.0 * ppi - 2.0 * pu)) * (-F(o->sxx, i, j) - F(o->syy, i, j)) / 3;
              break;
        }
    }
}

/* snap (src/snap_ssg.c:28-209): x-major over the decimated window */
static void snap(oracle *o)
{
    const sofi_desc *d = &o->d;
    if (o->nsnap >= o->nsnap_max) return;
    size_t per = (size_t)o->nxs * o->nys;
    REAL *base = o->snapbuf + (size_t)o->nsnap * SOFI_NCOMP * per;
    REAL dhi = (REAL)(1.0 / d->dh);
    size_t k = 0;
    for (int i = d->snap_idx[1]; i <= d->snap_idx[2]; i += d->idx)
        for (int j = d->snap_idx[3]; j <= d->snap_idx[4]; j += d->idy, ++k) {
            base[SOFI_C_VX * per + k] = F(o->vx, i, j);
            base[SOFI_C_VY * per + k] = F(o->vy, i, j);
            base[SOFI_C_P * per + k] = -F(o->sxx, i, j) - F(o->syy, i, j);
            REAL vyx = 0, vxy = 0, vxx = 0, vyy = 0;
            for (int m = 1; m <= o->nd; m++) {
                vyx += o->hc[m] * (F(o->vy, i + m, j) - F(o->vy, i - m + 1, j));
                vxy += o->hc[m] * (F(o->vx, i, j + m) - F(o->vx, i, j - m + 1));
            }
            vyx *= dhi;
            vxy *= dhi;
            base[SOFI_C_CURL * per + k] = o->a[SOFI_A_PU] ? (vxy - vyx) * sqrt(M(SOFI_A_PU, i, j)) : 0;
            for (int m = 1; m <= o->nd; m++) {
                vxx += o->hc[m] * (F(o->vx, i + m - 1, j) - F(o->vx, i - m, j));
                vyy += o->hc[m] * (F(o->vy, i, j + m - 1) - F(o->vy, i, j - m));
            }
            vxx *= dhi;
            vyy *= dhi;
            base[SOFI_C_DIV * per + k] = (vxx + vyy) * sqrt(M(SOFI_A_PPI, i, j));
        }
    o->nsnap++;
}

/* ------------------------------------------------------------- FWI */

#define ST(slot, k, i, j) (o->store[(((size_t)(slot) * 5 + (k)) * o->ny + ((j) - 1)) * o->nx + ((i) - 1)])

/* snap_store (src/snap_store.c:27-47): the forward fields of one step.  The
 * reference writes slot hin = nt (src/time_loop.c:288-292); with DTINV > 1
 * that index runs past its NT/DTINV slots, so the slot used here is the
 * store step's rank (nt - 1) / DTINV, the same as the reference's at
 * DTINV = 1. */
static void fwi_store(oracle *o, int slot)
{
    for (int j = 1; j <= o->ny; j++)
        for (int i = 1; i <= o->nx; i++) {
            ST(slot, 0, i, j) = F(o->pvxp1, i, j);
            ST(slot, 1, i, j) = F(o->pvyp1, i, j);
            if (o->d.velocity == 0) {
                ST(slot, 2, i, j) = F(o->sxx, i, j);
                ST(slot, 3, i, j) = F(o->syy, i, j);
                ST(slot, 4, i, j) = F(o->sxy, i, j);
            } else {
                ST(slot, 2, i, j) = F(o->ux, i, j);
                ST(slot, 3, i, j) = F(o->uy, i, j);
                ST(slot, 4, i, j) = F(o->uxy, i, j);
            }
        }
}

/* calc_conv (src/calc_conv.c:30-73): zero-lag correlation of the adjoint
 * fields with the forward fields of slot `slot` */
static void fwi_conv(oracle *o, int slot)
{
    const REAL DT = o->d.dt;
    for (int j = 1; j <= o->ny; j++)
        for (int i = 1; i <= o->nx; i++) {
            REAL muss = F(o->muss, i, j), lamss = F(o->lamss, i, j);
            REAL frx = ST(slot, 0, i, j), fry = ST(slot, 1, i, j);
            REAL fx = ST(slot, 2, i, j), fy = ST(slot, 3, i, j), fu = ST(slot, 4, i, j);
            REAL sxx = F(o->sxx, i, j), syy = F(o->syy, i, j), sxy = F(o->sxy, i, j);
            F(o->grad_rho, i, j) += -DT * ((F(o->pvxp1, i, j) * frx) + (F(o->pvyp1, i, j) * fry));
            F(o->grad_lam, i, j) += -DT * (1.0 / (4.0 * (lamss + muss) * (lamss + muss))) * ((fx + fy) * (sxx + syy));
            if (M(SOFI_A_PU, i, j) > 0.0)
                F(o->grad_mu, i, j) += -DT * (((1.0 / (muss * muss)) * (fu * sxy)) +
                                              ((1.0 / 4.0) * ((fx + fy) * (sxx + syy)) / ((lamss + muss) * (lamss + muss))) +
                                              ((1.0 / 4.0) * ((fx - fy) * (sxx - syy)) / (muss * muss)));
        }
}

/* the model of the iteration (minv->Vp0/Vs0/Rho0, src/readmod_elastic.c:126-128)
 * and the moduli calc_conv forms from it (src/calc_conv.c:45-46) */
void oracle_fwi_set_model(oracle *o, const float *vp0, const float *vs0, const float *rho0)
{
    for (size_t q = 0; q < o->plane; ++q) {
        o->vp0[q] = vp0[q];
        o->vs0[q] = vs0[q];
        o->rho0[q] = rho0[q];
        o->muss[q] = o->vs0[q] * o->vs0[q] * o->rho0[q];
        o->lamss[q] = o->vp0[q] * o->vp0[q] * o->rho0[q] - 2.0 * o->muss[q];
    }
}

/* adjoint sources at the receivers with signals [n][nt] (src/inversion.c:205-223) */
void oracle_fwi_set_adjoint(oracle *o, int n, const int *ai, const int *aj, const float *sx, const float *sy)
{
    free(o->ai); free(o->aj); free(o->adj_x); free(o->adj_y);
    o->nadj = n;
    o->ai = (int *)malloc(sizeof(int) * (n + 1));
    o->aj = (int *)malloc(sizeof(int) * (n + 1));
    o->adj_x = zalloc((size_t)(n > 0 ? n : 1) * o->d.nt);
    o->adj_y = zalloc((size_t)(n > 0 ? n : 1) * o->d.nt);
    for (int k = 0; k < n; ++k) { o->ai[k] = ai[k]; o->aj[k] = aj[k]; }
    for (size_t q = 0; q < (size_t)n * o->d.nt; ++q) { o->adj_x[q] = sx[q]; o->adj_y[q] = sy[q]; }
}

/* zero_wavefield (src/zero_wavefield.c:28-130), including the FWI fields, and
 * the per-call counters of time_loop; sw selects the pass that follows */
/* Test knob, for showing that a wrong correlation slot is detected: the
 * adjoint pass correlates its fields with store slot + shift (slots outside
 * the store are skipped).  0 restores the schedule. */
void oracle_fwi_set_conv_shift(oracle *o, int shift) { o->conv_shift = shift; }

void oracle_reset(oracle *o, int sw)
{
    REAL *planes[] = { o->vx, o->vy, o->sxx, o->syy, o->sxy, o->ux, o->uy, o->uxy, o->pvxp1, o->pvyp1 };
    for (size_t k = 0; k < sizeof(planes) / sizeof(planes[0]); ++k)
        if (planes[k]) memset(planes[k], 0, o->plane * sizeof(REAL));
    int L = o->L > 0 ? o->L : 1;
    memset(o->r, 0, o->plane * L * sizeof(REAL));
    memset(o->p, 0, o->plane * L * sizeof(REAL));
    memset(o->q, 0, o->plane * L * sizeof(REAL));
    size_t nxp = (size_t)(o->ny + 1) * o->wx, nyp = (size_t)(2 * o->d.fw + 1) * o->wy;
    REAL *px[] = { o->psi_sxx_x, o->psi_sxy_x, o->psi_vxx, o->psi_vyx };
    REAL *py[] = { o->psi_syy_y, o->psi_sxy_y, o->psi_vyy, o->psi_vxy };
    for (int k = 0; k < 4; ++k) {
        memset(px[k], 0, nxp * sizeof(REAL));
        memset(py[k], 0, nyp * sizeof(REAL));
    }
    if (o->fdt4) {   /* zero_wavefield clears every history (src/zero_wavefield.c:88-127) */
        for (int a = 0; a < 2; ++a)
            for (int k = 0; k < 4; ++k) memset(o->sv[a][k], 0, o->plane * sizeof(REAL));
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 4; ++k) memset(o->dvh[a][k], 0, o->plane * sizeof(REAL));
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 4; ++k) memset(o->mh[a][k], 0, o->plane * L * sizeof(REAL));
    }
    o->sw = sw;
    o->lsamp = o->d.ndt;
    o->lsnap = o->d.lsnap;
    o->hin = 1;
    o->hin1 = 1;
}

/* gradient of the shot (src/inversion.c:226-249): vp/vs/rho' from the
 * lambda/mu/rho accumulators, added to the sums; the accumulators are then
 * cleared (src/sofi2D.c:382-388).  Outputs the shot's gradients. */
int oracle_fwi_gradient(oracle *o, float *gvp, float *gvs, float *grho)
{
    size_t n = o->plane;
    memset(gvp, 0, n * sizeof(float));
    memset(gvs, 0, n * sizeof(float));
    memset(grho, 0, n * sizeof(float));
    for (int j = 1; j <= o->ny; j += o->d.idy > 0 ? o->d.idy : 1)
        for (int i = 1; i <= o->nx; i += o->d.idx > 0 ? o->d.idx : 1) {
            size_t q = IX(i, j);
            REAL Vp0 = o->vp0[q], Vs0 = o->vs0[q], Rho0 = o->rho0[q];
            REAL gl = o->grad_lam[q], gm = o->grad_mu[q], gr = o->grad_rho[q];
            REAL vp = 2.0 * Rho0 * Vp0 * gl;
            REAL vs = (-4.0 * Vs0 * Rho0 * gl) + 2.0 * Rho0 * Vp0 * gm;
            REAL rho = ((((Vp0 * Vp0) - (2.0 * Vs0 * Vs0)) * gl) + (Vs0 * Vs0 * gm) + gr);
            gvp[q] = (float)vp;
            gvs[q] = (float)vs;
            grho[q] = (float)rho;
            o->gsum[0][q] += vp;
            o->gsum[1][q] += vs;
            o->gsum[2][q] += rho;
        }
    memset(o->grad_lam, 0, n * sizeof(REAL));
    memset(o->grad_mu, 0, n * sizeof(REAL));
    memset(o->grad_rho, 0, n * sizeof(REAL));
    return 0;
}

/* gradients summed over the shots so far (minv->gradVp/gradVs/gradRho) */
int oracle_fwi_gradient_sum(const oracle *o, double *gvp, double *gvs, double *grho)
{
    for (size_t q = 0; q < o->plane; ++q) {
        gvp[q] = o->gsum[0][q];
        gvs[q] = o->gsum[1][q];
        grho[q] = o->gsum[2][q];
    }
    return 0;
}

/* ------------------------------------------------ 4th order in time */

/* Adams-Bashforth coefficients (src/wavefield_update_v_4.c:29-33: floats) */
#define AB1 ((REAL)(float)(13.0 / 12.0))
#define AB2 ((REAL)(float)(-5.0 / 24.0))
#define AB3 ((REAL)(float)(1.0 / 6.0))
#define AB4 ((REAL)(float)(-1.0 / 24.0))

/* shift_var2 / shift_var3 (src/util.c:89-109): the oldest buffer becomes
 * the newest, to be overwritten by the next update */
static void ab_shift(REAL **v)
{
    REAL *t = v[3];
    v[3] = v[2];
    v[2] = v[1];
    v[1] = v[0];
    v[0] = t;
}

/* wavefield_update_v_4 (src/wavefield_update_v_4.c:25-47) */
static void v4_update(oracle *o, int i, int j, REAL sxx_x, REAL sxy_x, REAL sxy_y, REAL syy_y)
{
    REAL dtdh = (REAL)o->d.dt / (REAL)o->d.dh;
    REAL **sx = o->sv[0], **sy = o->sv[1];
    F(sx[0], i, j) = sxx_x + sxy_y;
    F(sy[0], i, j) = sxy_x + syy_y;
    F(o->vx, i, j) += (AB1 * F(sx[0], i, j) + AB2 * F(sx[1], i, j) + AB3 * F(sx[2], i, j) + AB4 * F(sx[3], i, j)) *
                      dtdh * M(SOFI_A_RIP, i, j);
    F(o->vy, i, j) += (AB1 * F(sy[0], i, j) + AB2 * F(sy[1], i, j) + AB3 * F(sy[2], i, j) + AB4 * F(sy[3], i, j)) *
                      dtdh * M(SOFI_A_RJP, i, j);
}

/* update_v_interior_4 + update_v_PML_4 / update_v_abs_4: sources first
 * (no adjoint sources in this loop), the interior, then the frame.  In the
 * CPML frame the left side and the four corners only advance the CPML
 * memory; their velocities are never updated (src/update_v_PML_4.c:51-57,
 * 91-136).  Any ABS_TYPE other than 1 or 2 leaves the frame alone. */
static void update_v4(oracle *o, int nt)
{
    const sofi_desc *d = &o->d;
    if (o->sw == 0)
        for (int l = 0; l < o->nsrc; l++) {
            int i = o->si[l], j = o->sj[l];
            float azi_rad = (float)(o->sazi[l] * M_PI / 180);
            REAL amp = (REAL)((d->dt * o->sig[(size_t)l * d->nt + nt - 1]) / (d->dh * d->dh));
            switch (o->stype[l]) {
              case 2: F(o->vx, i, j) += M(SOFI_A_RIP, i, j) * amp; break;
              case 3: F(o->vy, i, j) += M(SOFI_A_RJP, i, j) * amp; break;
              case 4:
                  F(o->vx, i, j) += sin(azi_rad) * (M(SOFI_A_RIP, i, j) * amp);
                  F(o->vy, i, j) += cos(azi_rad) * (M(SOFI_A_RJP, i, j) * amp);
                  break;
            }
        }
    for (int j = d->gy[2] + 1; j <= d->gy[3]; j++)
        for (int i = d->gx[2] + 1; i <= d->gx[3]; i++) {
            REAL a, b, c, e;
            op_v(o, i, j, &a, &b, &c, &e);
            v4_update(o, i, j, a, b, c, e);
        }
    if (d->fw && (d->abs_type == 1 || d->abs_type == 2))
        for (int j = d->gy[1]; j <= d->gy[4]; j++)
            for (int i = d->gx[1]; i <= d->gx[4]; i++) {
                if (!in_frame(o, i, j)) continue;
                REAL a, b, c, e;
                op_v(o, i, j, &a, &b, &c, &e);
                int xl = i <= d->gx[2], xr = i > d->gx[3], yt = j <= d->gy[2], yb = j > d->gy[3];
                if (d->abs_type == 1) {
                    if (xl || xr) cpml_v_x(o, xl ? i : i - o->nx + 2 * d->fw, j, &a, &b);
                    if (yt || yb) cpml_v_y(o, i, yt ? j : j - o->ny + 2 * d->fw, &c, &e);
                    /* right side, top and bottom update; left side and corners do not */
                    if (!xl && !(xr && (yt || yb))) v4_update(o, i, j, a, b, c, e);
                } else {
                    v4_update(o, i, j, a, b, c, e);
                    F(o->vx, i, j) *= M(SOFI_A_ABSORB, i, j);
                    F(o->vy, i, j) *= M(SOFI_A_ABSORB, i, j);
                }
            }
    ab_shift(o->sv[0]);
    ab_shift(o->sv[1]);
}

/* wavefield_update_s_el_4 (src/wavefield_update_s_el_4.c:27-70) and
 * wavefield_update_s_visc_4 (src/wavefield_update_s_visc_4.c:27-98) on the
 * (CPML-modified) derivatives of the cell */
static void s4_cell(oracle *o, int i, int j)
{
    const REAL dh = (REAL)o->d.dh, dhi = (REAL)(1.0 / o->d.dh), dthalbe = (REAL)(o->d.dt / 2.0);
    REAL *src[4] = { o->vxx, o->vyy, o->vxy, o->vyx };
    REAL sum[4];
    for (int a = 0; a < 4; ++a) {
        REAL **h = o->dvh[a];
        F(h[0], i, j) = F(src[a], i, j) * dh;
        sum[a] = AB1 * F(h[0], i, j) + AB2 * F(h[1], i, j) + AB3 * F(h[2], i, j) + AB4 * F(h[3], i, j);
    }
    const REAL sumxx = sum[0], sumyy = sum[1], sumxy = sum[2], sumyx = sum[3];
    REAL u1, u2, u3;
    if (o->d.weq != SOFI_VEL_ISO) {
        u1 = M(SOFI_A_FIPJP, i, j) * (sumxy + sumyx) * dhi;
        u2 = ((M(SOFI_A_G, i, j) * (sumxx + sumyy)) - (2.0 * M(SOFI_A_F, i, j) * sumyy)) * dhi;
        u3 = ((M(SOFI_A_G, i, j) * (sumxx + sumyy)) - (2.0 * M(SOFI_A_F, i, j) * sumxx)) * dhi;
    } else {
        const REAL *bip = o->prof[SOFI_P_BIP], *cip = o->prof[SOFI_P_CIP];
        REAL **r = o->mh[0], **p = o->mh[1], **q = o->mh[2];
        REAL sumr_old = 0, sump_old = 0, sumq_old = 0, sumr = 0, sump = 0, sumq = 0;
        for (int l = 1; l <= o->L; l++) {
            sumr_old += AB1 * MV(r[0], i, j, l) + AB2 * MV(r[1], i, j, l) + AB3 * MV(r[2], i, j, l) + AB4 * MV(r[3], i, j, l);
            sump_old += AB1 * MV(p[0], i, j, l) + AB2 * MV(p[1], i, j, l) + AB3 * MV(p[2], i, j, l) + AB4 * MV(p[3], i, j, l);
            sumq_old += AB1 * MV(q[0], i, j, l) + AB2 * MV(q[1], i, j, l) + AB3 * MV(q[2], i, j, l) + AB4 * MV(q[3], i, j, l);
            REAL ctemp = 2 * (1 - cip[l]) / AB1;
            MV(r[3], i, j, l) = bip[l] * (MV(r[0], i, j, l) * cip[l] - ctemp * AB2 * (MV(r[0], i, j, l) + MV(r[1], i, j, l)) -
                                          ctemp * AB3 * (MV(r[1], i, j, l) + MV(r[2], i, j, l)) -
                                          ctemp * AB4 * (MV(r[2], i, j, l) + MV(r[3], i, j, l)) -
                                          (M3(SOFI_A_DIP, i, j, l) * (sumxy + sumyx) * dhi));
            MV(p[3], i, j, l) = bip[l] * (MV(p[0], i, j, l) * cip[l] - ctemp * AB2 * (MV(p[0], i, j, l) + MV(p[1], i, j, l)) -
                                          ctemp * AB3 * (MV(p[1], i, j, l) + MV(p[2], i, j, l)) -
                                          ctemp * AB4 * (MV(p[2], i, j, l) + MV(p[3], i, j, l)) -
                                          (M3(SOFI_A_E, i, j, l) * (sumxx + sumyy) * dhi) +
                                          (2.0 * M3(SOFI_A_D, i, j, l) * sumyy * dhi));
            MV(q[3], i, j, l) = bip[l] * (MV(q[0], i, j, l) * cip[l] - ctemp * AB2 * (MV(q[0], i, j, l) + MV(q[1], i, j, l)) -
                                          ctemp * AB3 * (MV(q[1], i, j, l) + MV(q[2], i, j, l)) -
                                          ctemp * AB4 * (MV(q[2], i, j, l) + MV(q[3], i, j, l)) -
                                          (M3(SOFI_A_E, i, j, l) * (sumxx + sumyy) * dhi) +
                                          (2.0 * M3(SOFI_A_D, i, j, l) * sumxx * dhi));
            sumr += AB1 * MV(r[3], i, j, l) + AB2 * MV(r[0], i, j, l) + AB3 * MV(r[1], i, j, l) + AB4 * MV(r[2], i, j, l);
            sump += AB1 * MV(p[3], i, j, l) + AB2 * MV(p[0], i, j, l) + AB3 * MV(p[1], i, j, l) + AB4 * MV(p[2], i, j, l);
            sumq += AB1 * MV(q[3], i, j, l) + AB2 * MV(q[0], i, j, l) + AB3 * MV(q[1], i, j, l) + AB4 * MV(q[2], i, j, l);
        }
        u1 = (M(SOFI_A_FIPJP, i, j) * (sumxy + sumyx)) * dhi + (dthalbe * (sumr_old + sumr));
        u2 = (M(SOFI_A_G, i, j) * (sumxx + sumyy) * dhi) - (2.0 * M(SOFI_A_F, i, j) * sumyy * dhi) +
             (dthalbe * (sump_old + sump));
        u3 = (M(SOFI_A_G, i, j) * (sumxx + sumyy) * dhi) - (2.0 * M(SOFI_A_F, i, j) * sumxx * dhi) +
             (dthalbe * (sumq_old + sumq));
    }
    F(o->sxy, i, j) += u1;
    F(o->sxx, i, j) += u2;
    F(o->syy, i, j) += u3;
}

/* update_s_{elastic,visc}_{interior,PML,abs}_4, then the shifts of the
 * histories (src/time_loop.c:215-245): the memory variables' newest buffer
 * is the one just written */
static void update_s4(oracle *o)
{
    const sofi_desc *d = &o->d;
    for (int j = d->gy[2] + 1; j <= d->gy[3]; j++)
        for (int i = d->gx[2] + 1; i <= d->gx[3]; i++) s4_cell(o, i, j);
    if (d->fw)
        for (int j = d->gy[1]; j <= d->gy[4]; j++)
            for (int i = d->gx[1]; i <= d->gx[4]; i++) {
                if (!in_frame(o, i, j)) continue;
                if (d->abs_type == 1) {
                    cpml_s_cell(o, i, j);
                    s4_cell(o, i, j);
                } else {
                    s4_cell(o, i, j);
                    abs_s(o, i, j);
                }
            }
    if (d->weq == SOFI_VEL_ISO) {
        for (int a = 0; a < 3; ++a) ab_shift(o->mh[a]);
        o->r = o->mh[0][0];
        o->p = o->mh[1][0];
        o->q = o->mh[2][0];
    }
    for (int a = 0; a < 4; ++a) ab_shift(o->dvh[a]);
}

/* ------------------------------------------- periodic boundary in x */

/* exchange_v with BOUNDARY = 1 on one process (src/exchange_v.c:91-148):
 * the process is its own left and right neighbour, so the ghost columns
 * wrap around; the array in the vx position gets ND - 1 columns on the
 * right and ND on the left, the one in the vy position ND on the right and
 * ND - 1 on the left (rows 1..NY) */
static void exch_v(oracle *o, REAL *a, REAL *b)
{
    const int nx = o->nx, nd = o->nd;
    for (int j = 1; j <= o->ny; j++) {
        for (int l = 1; l <= nd; l++) F(b, nx + l, j) = F(b, l, j);
        for (int l = 1; l <= nd - 1; l++) F(a, nx + l, j) = F(a, l, j);
        for (int l = 1; l <= nd - 1; l++) F(b, 1 - l, j) = F(b, nx - l + 1, j);
        for (int l = 1; l <= nd; l++) F(a, 1 - l, j) = F(a, nx - l + 1, j);
    }
}

/* exchange_s with BOUNDARY = 1 on one process (src/exchange_s.c:92-146,
 * 193-236): sxy gets ND - 1 columns on the right and ND on the left, sxx ND
 * on the right and ND - 1 on the left; the acoustic path has sxx only */
static void exch_s(oracle *o)
{
    const int nx = o->nx, nd = o->nd, ac = o->d.weq == SOFI_AC_ISO;
    for (int j = 1; j <= o->ny; j++) {
        if (!ac) {
            for (int l = 1; l <= nd - 1; l++) F(o->sxy, nx + l, j) = F(o->sxy, l, j);
            for (int l = 1; l <= nd; l++) F(o->sxy, 1 - l, j) = F(o->sxy, nx - l + 1, j);
        }
        for (int l = 1; l <= nd; l++) F(o->sxx, nx + l, j) = F(o->sxx, l, j);
        for (int l = 1; l <= nd - 1; l++) F(o->sxx, 1 - l, j) = F(o->sxx, nx - l + 1, j);
    }
}

/* --------------------------------------------------------- one step */

void oracle_step(oracle *o, int nt)
{
    const sofi_desc *d = &o->d;
    const int per = d->boundary;
    if (o->fdt4) update_v4(o, nt);
    else update_v(o, nt);
    if (per) exch_v(o, o->vx, o->vy);
    v_derivatives(o);
    if (per) {   /* the derivative arrays too (src/time_loop.c:107-110) */
        exch_v(o, o->vxx, o->vyy);
        if (d->weq >= SOFI_EL_ISO) exch_v(o, o->vyx, o->vxy);
    }
    if (o->fdt4) update_s4(o);
    else update_s(o);
    int stype = d->source_type;
    /* (ADJOINT_TYPE 4, a pressure adjoint source through psource, is refused
     * by the host library: the reference injects sectionvxdiff there while
     * calc_res fills sectionpdiff) */
    if (o->sw == 0 && stype == 1) psource(o, nt, o->sig);
    if (o->sw == 0 && stype == 5) eqsource(o, nt);
    if (d->free_surf) free_surface(o);
    if (per) exch_s(o);
    if (d->seismo && nt == o->lsamp && nt < d->nt) {
        seismo(o, o->lsamp);
        o->lsamp += d->ndt;
    }
    /* FWI: forward store every DTINV steps, correlation in the adjoint pass at
     * the steps whose time-reversed forward step was stored
     * (src/time_loop.c:287-303) */
    if (d->mode == SOFI_MODE_FWI) {
        int dtinv = d->dtinv > 0 ? d->dtinv : 1;
        if (o->sw == 0 && nt == o->hin1) {
            fwi_store(o, (nt - 1) / dtinv);
            o->hin1 += dtinv;
        }
        int f = d->nt - nt + 1;
        if (o->sw == 1 && (f - 1) % dtinv == 0) {
            const int slot = (f - 1) / dtinv + o->conv_shift;
            if (slot >= 0 && slot < o->nstore) fwi_conv(o, slot);
        }
        o->hin++;
    }
    if (d->snap && nt == o->lsnap && nt <= d->esnap) {
        snap(o);
        o->lsnap += d->isnap;
    }
}

void oracle_run(oracle *o, int nt0, int nt1)
{
    for (int nt = nt0; nt <= nt1; ++nt) oracle_step(o, nt);
}

/* ----------------------------------------------------------- results */

static const REAL *field_ptr(const oracle *o, int id)
{
    switch (id) {
      case 0: return o->vx;
      case 1: return o->vy;
      case 2: return o->sxx;
      case 3: return o->syy;
      case 4: return o->sxy;
      case 5: return o->vxx;
      case 6: return o->vyy;
      case 7: return o->vyx;
      case 8: return o->vxy;
      default: return NULL;
    }
}

/* copies field id (0 vx, 1 vy, 2 sxx, 3 syy, 4 sxy, 5..8 dv) in the padded layout */
int oracle_get_field(const oracle *o, int id, float *out)
{
    const REAL *f = field_ptr(o, id);
    if (!f) return -1;
    for (size_t q = 0; q < o->plane; ++q) out[q] = (float)f[q];
    return 0;
}

int oracle_get_field_f64(const oracle *o, int id, double *out)
{
    const REAL *f = field_ptr(o, id);
    if (!f) return -1;
    for (size_t q = 0; q < o->plane; ++q) out[q] = (double)f[q];
    return 0;
}

/* memory variable plane l (1-based) of r (0), p (1), q (2) */
int oracle_get_memvar(const oracle *o, int which, int l, float *out)
{
    const REAL *f = which == 0 ? o->r : which == 1 ? o->p : o->q;
    for (size_t q = 0; q < o->plane; ++q) out[q] = (float)f[(size_t)(l - 1) * o->plane + q];
    return 0;
}

int oracle_get_section(const oracle *o, int comp, double *out)
{
    size_t n = (size_t)o->ntr * o->d.ns;
    for (size_t q = 0; q < n; ++q) out[q] = (double)o->sec[comp][q];
    return 0;
}

int oracle_snapshot_count(const oracle *o) { return o->nsnap; }

int oracle_get_snapshot(const oracle *o, int k, int comp, double *out)
{
    if (k < 0 || k >= o->nsnap) return -1;
    size_t per = (size_t)o->nxs * o->nys;
    const REAL *b = o->snapbuf + (size_t)k * SOFI_NCOMP * per + (size_t)comp * per;
    for (size_t q = 0; q < per; ++q) out[q] = (double)b[q];
    return 0;
}

void oracle_free(oracle *o)
{
    if (!o) return;
    if (o->syy == o->sxx) o->syy = NULL;
    REAL *bufs[] = { o->vx, o->vy, o->sxx, o->syy, o->sxy, o->vxx, o->vyy, o->vyx, o->vxy, o->r, o->p, o->q,
                     o->psi_sxx_x, o->psi_sxy_x, o->psi_syy_y, o->psi_sxy_y, o->psi_vxx, o->psi_vyx, o->psi_vyy,
                     o->psi_vxy, o->sazi, o->samp, o->sig, o->snapbuf, o->ux, o->uy, o->uxy, o->pvxp1, o->pvyp1,
                     o->grad_lam, o->grad_mu, o->grad_rho, o->adj_x, o->adj_y, o->store, o->gsum[0], o->gsum[1],
                     o->gsum[2], o->vp0, o->vs0, o->rho0, o->muss, o->lamss };
    for (size_t k = 0; k < sizeof(bufs) / sizeof(bufs[0]); ++k) free(bufs[k]);
    if (o->fdt4) {   /* o->r, o->p, o->q (freed above) are one of each history */
        for (int a = 0; a < 2; ++a)
            for (int k = 0; k < 4; ++k) free(o->sv[a][k]);
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 4; ++k) free(o->dvh[a][k]);
        REAL *cur[3] = { o->r, o->p, o->q };
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 4; ++k)
                if (o->mh[a][k] != cur[a]) free(o->mh[a][k]);
    }
    for (int k = 0; k < SOFI_A_NARRAYS; ++k) free(o->a[k]);
    for (int k = 0; k < SOFI_P_NPROFILES; ++k) free(o->prof[k]);
    for (int c = 0; c < SOFI_NCOMP; ++c) free(o->sec[c]);
    free(o->si); free(o->sj); free(o->stype); free(o->ri); free(o->rj); free(o->ai); free(o->aj);
    free(o);
}

/* size of the oracle's step in bytes of REAL (lets tests check which build they loaded) */
/* timedomain_filt (src/timedomain_filt.c:28-113) with libseife's
 * Butterworth cascade (contrib/Seitosh/libseife/cseife.c:160-190,
 * cseife_rfk.c, cseife_rekfl.c), in double as the reference: each trace is
 * filtered as NS + 1 samples (a leading 0), forward, rounded to float,
 * reversed, filtered again and reversed back; low-pass at period tlp, then,
 * when thp > 0, high-pass at thp. */
static void rec_filter(double *x, int n, double f0, double f1, double f2, double g1, double g2)
{
    double xm1 = 0, xm2 = 0, ym1 = 0, ym2 = 0;
    for (int k = 0; k < n; ++k) {
        double in = x[k];
        x[k] = f0 * in + f1 * xm1 + f2 * xm2 + g1 * ym1 + g2 * ym2;
        xm2 = xm1;
        xm1 = in;
        ym2 = ym1;
        ym1 = x[k];
    }
}

static void bw_cascade(double *x, int n, double t0s, int order, int hp)
{
    /* t0s: the corner period in samples */
    const double w = 8.0 * atan(1.0) / t0s;
    if (hp) {
        double first = x[0];
        for (int k = 0; k < n; ++k) x[k] -= first;
    }
    if (order % 2 == 1) {
        double g1 = (2.0 - w) / (2.0 + w);
        if (hp) rec_filter(x, n, 2.0 / (2.0 + w), -(2.0 / (2.0 + w)), 0, g1, 0);
        else rec_filter(x, n, w / (2.0 + w), w / (2.0 + w), 0, g1, 0);
    }
    for (int s = 1; s <= order / 2; ++s) {
        const double h = sin(2.0 * atan(1.0) * (2 * s - 1) / order);
        const double w2 = w * w;
        const double A = 1.0 - w * h + w2 / 4.0, B = -2.0 + w2 / 2.0, Cc = 1.0 + w * h + w2 / 4.0;
        double f0 = hp ? 1.0 / Cc : w2 / 4.0 / Cc;
        double f1 = hp ? -(f0) - f0 : f0 + f0;
        rec_filter(x, n, f0, f1, f0, -B / Cc, -A / Cc);
    }
}

void oracle_time_filter(int ntr, int ns, float dt, float f_low, float f_high, int order, float *traces)
{
    double *u = (double *)calloc((size_t)ns + 1, sizeof(double)), *v = (double *)calloc((size_t)ns + 1, sizeof(double));
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1 && !f_high) break;
        const double period = 1.0 / (double)(pass == 0 ? f_low : f_high);
        for (int tr = 0; tr < ntr; ++tr) {
            float *t = traces + (size_t)tr * ns;
            u[0] = 0;
            for (int k = 1; k <= ns; ++k) u[k] = t[k - 1];
            bw_cascade(u, ns + 1, period / (double)dt, order, pass);
            v[0] = 0;
            for (int k = 1; k <= ns; ++k) v[ns + 1 - k] = (double)(float)u[k];
            bw_cascade(v, ns + 1, period / (double)dt, order, pass);
            for (int k = 1; k <= ns; ++k) t[k - 1] = (float)v[ns + 1 - k];
        }
    }
    free(u);
    free(v);
}

/* calc_res (src/calc_res.c:127-301): the adjoint source of one component
 * for LNORM 1-5, 7 and 8, traces [ntr][ns], time-reversed into diff, and
 * the misfit the reference accumulates for a test shot (src/calc_res.c:
 * 276-291).  RMS = 1 (src/calc_res.c:166); the envelope and Hilbert arrays
 * of LNORM 8 stay at the zeros matrix() gives them (their computation is
 * commented out, src/calc_res.c:175-187). */
static void integrate_trace(const float *x, REAL *out, int ns, int velocity, REAL dt)
{
    REAL acc = 0;
    for (int j = 0; j < ns; ++j) {
        if (velocity == 0) {
            acc += x[j];
            out[j] = acc * dt;
        } else {
            out[j] = x[j];
        }
    }
}

int oracle_residual(int lnorm, int velocity, float dt, float waterlevel, int ntr, int ns, const float *data,
                    const float *calc, float *diff, double *misfit)
{
    REAL *is = (REAL *)malloc(sizeof(REAL) * (size_t)ns), *id = (REAL *)malloc(sizeof(REAL) * (size_t)ns);
    if (!is || !id) {
        free(is);
        free(id);
        return -1;
    }
    REAL l2 = 0;
    for (int tr = 0; tr < ntr; ++tr) {
        const float *syn = calc + (size_t)tr * ns, *obs = data + (size_t)tr * ns;
        integrate_trace(syn, is, ns, velocity, (REAL)dt);
        integrate_trace(obs, id, ns, velocity, (REAL)dt);
        REAL nd = 0, nsyn = 0, dot = 0;
        if (lnorm == 5 || lnorm == 7) {
            for (int j = 0; j < ns; ++j) {
                nd += id[j] * id[j];
                nsyn += is[j] * is[j];
                dot += id[j] * is[j];
            }
            nd = nd == 0 ? (REAL)1 : (REAL)sqrt(nd);
            nsyn = nsyn == 0 ? (REAL)1 : (REAL)sqrt(nsyn);
        }
        float *out = diff + (size_t)tr * ns;
        for (int j = 0; j < ns; ++j) {
            const REAL r = (REAL)obs[j] - (REAL)syn[j];     /* (sectiondata - section) / RMS */
            REAL v = 0;
            if (lnorm == 1) v = r > 0 ? (REAL)1 : (r < 0 ? (REAL)-1 : (REAL)0);
            else if (lnorm == 2) v = is[j] - id[j];
            else if (lnorm == 3) v = r / (1 + r * r);
            else if (lnorm == 4) v = (REAL)tanh(r);
            else if (lnorm == 5 || lnorm == 7) v = is[j] * dot / (nsyn * nsyn * nsyn * nd) - id[j] / (nsyn * nd);
            else if (lnorm == 8) v = is[j] / ((REAL)0 + (REAL)waterlevel) * ((REAL)0 - (REAL)0) - (REAL)0;
            out[ns - 1 - j] = (float)v;
            if (lnorm == 2) l2 += v * v;
            else if (lnorm == 5 || lnorm == 7) l2 -= id[j] * is[j] / (nd * nsyn);
        }
    }
    free(is);
    free(id);
    *misfit = (double)l2;
    return 0;
}

int oracle_real_size(void) { return (int)sizeof(REAL); }
