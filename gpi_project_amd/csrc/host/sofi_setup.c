/*
 * sofi_setup.c - FD coefficients, stability check and the time-loop descriptor.
 *
 *   holbergcoeff  src/holbergcoeff.c:38-100
 *   checkfd       src/checkfd.c:32-291 (fatal checks; warnings go to stderr)
 *   subgrid_bounds src/subgrid_bounds.c:56-271 (one process: GX = {1, FW, NX-FW, NX})
 *   time_loop schedule constants  src/time_loop.c:38-40
 */
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "sofi_host_internal.h"

/* Staggered-grid Taylor coefficients of order 2N in closed form:
 *   c_m = (-1)^(m+1) [(2N-1)!!]^2 / (2^(2N-2) (N+m-1)! (N-m)! (2m-1)^2),
 * evaluated as one division of exact integers so that the result is the
 * correctly rounded value of the rational (the reference lists the same
 * rationals, e.g. 9/8, -1/24, src/holbergcoeff.c:42-48). */
static double taylor_coeff(int N, int m)
{
    long long df = 1;
    for (int k = 2 * N - 1; k > 1; k -= 2) df *= k;
    long long num = df * df;
    long long fa = 1, fb = 1;
    for (int k = 2; k <= N + m - 1; ++k) fa *= k;
    for (int k = 2; k <= N - m; ++k) fb *= k;
    long long den = (1LL << (2 * N - 2)) * fa * fb * (long long)(2 * m - 1) * (2 * m - 1);
    double c = (double)num / (double)den;
    return (m % 2 == 1) ? c : -c;
}

/* Holberg (1987) group-velocity optimised coefficients for maximum relative
 * errors of 0.1, 0.5, 1 and 3 %; column 0 is the minimum number of grid
 * points per shortest wavelength.  Published constants, as tabulated for
 * MAXRELERROR = 1..4 in src/holbergcoeff.c:50-81. */
static const double holberg_tab[4][6][7] = {
    { { 49.7, 1.0010 },
      { 8.32, 1.1382, -0.046414 },
      { 4.77, 1.1965, -0.078804, 0.0081781 },
      { 3.69, 1.2257, -0.099537, 0.018063, -0.0026274 },
      { 3.19, 1.2415, -0.11231, 0.026191, -0.0064682, 0.001191 },
      { 2.91, 1.2508, -0.12034, 0.032131, -0.010142, 0.0029857, -0.00066667 } },
    { { 22.2, 1.0050 },
      { 5.65, 1.1534, -0.052806 },
      { 3.74, 1.2111, -0.088313, 0.011768 },
      { 3.11, 1.2367, -0.10815, 0.023113, -0.0046905 },
      { 2.80, 1.2496, -0.11921, 0.031130, -0.0093272, 0.0025161 },
      { 2.62, 1.2568, -0.12573, 0.036423, -0.013132, 0.0047484, -0.0015979 } },
    { { 15.8, 1.0100 },
      { 4.80, 1.1640, -0.057991 },
      { 3.39, 1.2192, -0.094070, 0.014608 },
      { 2.90, 1.2422, -0.11269, 0.026140, -0.0064054 },
      { 2.65, 1.2534, -0.12257, 0.033755, -0.011081, 0.0036784 },
      { 2.51, 1.2596, -0.12825, 0.038550, -0.014763, 0.0058619, -0.0024538 } },
    { { 9.16, 1.0300 },
      { 3.47, 1.1876, -0.072518 },
      { 2.91, 1.2341, -0.10569, 0.022589 },
      { 2.61, 1.2516, -0.12085, 0.032236, -0.011459 },
      { 2.45, 1.2596, -0.12829, 0.038533, -0.014681, 0.0072580 },
      { 2.36, 1.2640, -0.13239, 0.042217, -0.017803, 0.0081959, -0.0051848 } },
};

/* minimum points per wavelength of the Taylor operators (src/holbergcoeff.c:42-48) */
static const double taylor_ppw[6] = { 23.0, 8.0, 6.0, 5.0, 5.0, 4.0 };

void sofi_holberg(int fdorder, int maxrelerror, float hc[7])
{
    int N = fdorder / 2;
    for (int k = 0; k < 7; ++k) hc[k] = 0.0f;
    if (maxrelerror == 0) {
        hc[0] = (float)taylor_ppw[N - 1];
        for (int m = 1; m <= N; ++m) hc[m] = (float)taylor_coeff(N, m);
    } else {
        for (int k = 0; k <= 6; ++k) hc[k] = (float)holberg_tab[maxrelerror - 1][N - 1][k];
    }
}

#define RECP(row, k) (a->recpos[(size_t)(row) * (size_t)(a->ntr + 1) + (size_t)(k)])
#define SRCP(row, l) (a->srcpos[(size_t)(row) * (size_t)(a->nsrc + 1) + (size_t)(l)])

int sofi_checkfd(sofi_params *p, const sofi_model *m, const sofi_acq *a)
{
    float hc[7];
    sofi_holberg(p->fdorder, p->maxrelerror, hc);
    float cmax = m->vsmax > m->vpmax ? m->vsmax : m->vpmax;
    float cmin = m->vsmin < m->vpmin ? m->vsmin : m->vpmin;
    float temporal = p->fdorder_time == 4 ? 1.5f : 1.0f;   /* 3.0 / 2.0 for Adams-Bashforth (src/checkfd.c:48-52) */
    float fmax = (float)(2.0 / p->ts);
    float dhstab = (cmin / (hc[0] * fmax));
    float gamma = (float)(fabs(hc[1]) + fabs(hc[2]) + fabs(hc[3]) + fabs(hc[4]) + fabs(hc[5]) + fabs(hc[6]));
    float dtstab = (float)(p->dh / (sqrt(2) * gamma * cmax * temporal));
    if (p->dh > dhstab)
        fprintf(stderr, "WARNING: Grid dispersion will influence wave propagation, choose a smaller grid spacing (DH).\n");
    if (p->dt > dtstab)
        return sofi_fail("The simulation will get unstable, choose a smaller DT (DT=%e > dtstab=%e).", p->dt, dtstab);
    if (p->snap && p->tsnap2 > p->time) p->tsnap2 = p->time;
    if (p->seismo) {
        if (p->seis_format == 1 && p->ns > USHRT_MAX)
            return sofi_fail("Maximum number of samples per trace in SU format: %d. Your value: %d", USHRT_MAX, p->ns);
        float srec_minx = p->nxg * p->dh, srec_miny = p->nyg * p->dh, srec_maxx = -1.0f, srec_maxy = -1.0f;
        if (p->readrec == 0) {
            srec_maxx = p->xrec1 > p->xrec2 ? p->xrec1 : p->xrec2;
            srec_minx = p->xrec1 > p->xrec2 ? p->xrec2 : p->xrec1;
            srec_maxy = p->yrec1 > p->yrec2 ? p->yrec1 : p->yrec2;
            srec_miny = p->yrec1 > p->yrec2 ? p->yrec2 : p->yrec1;
        } else {
            for (int k = 1; k <= a->ntr; k++) {
                float x = (RECP(1, k) - 1) * p->dh, y = (RECP(2, k) - 1) * p->dh;
                if (x > srec_maxx) srec_maxx = x;
                if (y > srec_maxy) srec_maxy = y;
                if (x < srec_minx) srec_minx = x;
                if (y < srec_miny) srec_miny = y;
            }
        }
        if (srec_maxx < 0.0 || srec_maxy < 0.0 || srec_minx < 0.0 || srec_miny < 0.0 ||
            srec_maxx > (p->nxg - 1) * p->dh || srec_maxy > (p->nyg - 1) * p->dh)
            return sofi_fail("Coordinate of at least one receiver location is outside the global grid.");
    }
    if (p->srcrec == 1) {
        float srec_minx = p->nxg * p->dh, srec_miny = p->nyg * p->dh, srec_maxx = -1.0f, srec_maxy = -1.0f;
        for (int k = 1; k <= a->nsrc; k++) {
            if (SRCP(1, k) > srec_maxx) srec_maxx = SRCP(1, k);
            if (SRCP(2, k) > srec_maxy) srec_maxy = SRCP(2, k);
            if (SRCP(1, k) < srec_minx) srec_minx = SRCP(1, k);
            if (SRCP(2, k) < srec_miny) srec_miny = SRCP(2, k);
        }
        if (srec_maxx < 0.0 || srec_maxy < 0.0 || srec_minx < 0.0 || srec_miny < 0.0 ||
            srec_maxx > (p->nxg - 1) * p->dh || srec_maxy > (p->nyg - 1) * p->dh)
            return sofi_fail("Coordinate of at least one source location is outside the global grid.");
    }
    if (p->nxg < p->fw || p->nyg < p->fw)
        return sofi_fail("Width of absorbing boundary (FW=%d grid points) is larger than at least one subdomain dimension.", p->fw);
    return 0;
}

void sofi_fill_desc(const sofi_params *p, const sofi_shot *s, sofi_desc *d)
{
    memset(d, 0, sizeof(*d));
    d->nx = p->nxg;
    d->ny = p->nyg;
    d->fdorder = p->fdorder;
    d->weq = p->weq;
    d->L = p->L;
    d->nt = p->nt;
    d->dt = p->dt;
    d->dh = p->dh;
    sofi_holberg(p->fdorder, p->maxrelerror, d->hc);
    d->fw = p->fw;
    d->abs_type = p->fw > 0 ? p->abs_type : 0;
    d->free_surf = p->free_surf;
    /* subgrid_bounds(1, NX, 1, NY) on a 1x1 process grid */
    d->gx[1] = 1; d->gx[2] = p->fw; d->gx[3] = p->nxg - p->fw; d->gx[4] = p->nxg;
    d->gy[1] = 1; d->gy[2] = p->fw; d->gy[3] = p->nyg - p->fw; d->gy[4] = p->nyg;
    if (p->free_surf == 1) d->gy[2] = 0;
    /* periodic in x: no left / right frame (src/subgrid_bounds.c:265-270 on
     * one process: GX[2] = 0, GX[3] = NX).  The reference also sets GX[4] =
     * NX - 1, which bounds only the loops over the right frame
     * (GX[3] + 1 .. GX[4]), empty either way; GX[4] stays NX here so that
     * loops over the whole grid keep the last column. */
    if (p->boundary == 1) {
        d->gx[2] = 0;
        d->gx[3] = p->nxg;
    }
    d->boundary = p->boundary == 1;
    d->source_type = s ? s->source_type : p->source_type;
    d->seismo = p->seismo;
    d->ndt = p->ndt;
    d->ns = p->ns;
    d->snap = (s && s->snapcheck) ? p->snap : 0;
    d->idx = p->idx;
    d->idy = p->idy;
    d->snap_idx[1] = 1;
    d->snap_idx[2] = 1 + ((p->nxg - 1) / p->idx) * p->idx;
    d->snap_idx[3] = 1;
    d->snap_idx[4] = 1 + ((p->nyg - 1) / p->idy) * p->idy;
    d->lsnap = sofi_iround(p->tsnap1 / p->dt);
    d->isnap = sofi_iround(p->tsnapinc / p->dt);
    d->esnap = sofi_iround(p->tsnap2 / p->dt);
    d->mode = p->mode;
    d->dtinv = p->dtinv;
    d->velocity = p->velocity;
    d->adjoint_type = p->adjoint_type;
    d->fdorder_time = p->fdorder_time;
}
