/*
 * sofi_medium.c - model input and material pre-processing.
 *
 * Reads MFILE.{vp,vs,rho[,qp,qs][,epsilon,delta][,theta]} (raw float32,
 * x-major: NY floats per x position; or the same with a .su suffix) and
 * computes every per-cell coefficient the time loop reads, with the
 * arithmetic (float vs. double promotions, evaluation order) of:
 *   readmod_elastic      src/readmod_elastic.c:103-135
 *   readmod_visco        src/readmod_visco.c:58-167
 *   readmod_elastic_vti  src/readmod_elastic_vti.c:123-160
 *   readmod_visco_vti    src/readmod_visco_vti.c:116-232
 *   readmod_elastic_tti  src/readmod_elastic_tti.c:126-185
 *   readmod_visco_tti    src/readmod_visco_tti.c:116-232
 *   prepmod              src/prepmod.c:28-194
 *   matcopy_elastic      src/matcopy_elastic.c:30-128 (zero material halo at the grid edge)
 *   matcopy              src/matcopy.c:28-140 (VEL_ISO: halo filled periodically, see below)
 *   av_mue / av_rho / av_tau  src/av_mue.c:27-43, src/av_rho.c:28-36, src/av_tau.c:27-34
 *   prepare_update_s_{el,visc,vti,tti}  src/prepare_update_s_*.c
 *   dt_mult              src/util.c:37-46
 * The halo rule matters for the last row and column: matcopy_elastic leaves
 * the halo at the global edge at zero (rip[j][NX] = 2/rho, mu_ipjp = 0 there),
 * whereas matcopy (used for VEL_ISO only) exchanges with neighbours whose
 * indices wrap around the process grid (src/initproc.c:58-65), so on the
 * global grid the halo row NY+1 is row 1 and column NX+1 is column 1.
 */
#define _GNU_SOURCE
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include "sofi_host_internal.h"

#define V_IGNORE 1.0

/* ------------------------------------------------------------ file input */

/* Reads npara model files into para[k][i*ny + j] (x-major, 0-based). */
static int read_model_files(const sofi_params *p, int npara, const char **suffix, float **para)
{
    char fn[2 * SOFI_STR];
    int nx = p->nxg, ny = p->nyg;
    snprintf(fn, sizeof(fn), "%s.%s.su", p->mfile, suffix[0]);
    int is_su = access(fn, R_OK) == 0;
    for (int k = 0; k < npara; ++k) {
        snprintf(fn, sizeof(fn), is_su ? "%s.%s.su" : "%s.%s", p->mfile, suffix[k]);
        FILE *fp = fopen(fn, "rb");
        if (!fp) return sofi_fail("Could not open model file %s.", fn);
        para[k] = (float *)malloc((size_t)nx * ny * sizeof(float));
        if (!para[k]) { fclose(fp); return sofi_fail("out of memory reading %s", fn); }
        if (is_su) {
            /* SU: 240-byte trace header (ns at byte 114) + ns samples per trace */
            for (int i = 0; i < nx; ++i) {
                unsigned char hdr[240];
                if (fread(hdr, 1, 240, fp) != 240) { fclose(fp); return sofi_fail("%s has fewer than NX=%d traces.", fn, nx); }
                unsigned short ns;
                memcpy(&ns, hdr + 114, 2);
                if (ns < ny) { fclose(fp); return sofi_fail("%s has fewer than NY=%d samples.", fn, ny); }
                if (fread(para[k] + (size_t)i * ny, sizeof(float), ny, fp) != (size_t)ny) { fclose(fp); return sofi_fail("short read in %s", fn); }
                if (ns > ny) fseek(fp, (long)(ns - ny) * 4, SEEK_CUR);
            }
        } else {
            size_t got = fread(para[k], sizeof(float), (size_t)nx * ny, fp);
            if (got != (size_t)nx * ny) {
                fclose(fp);
                return sofi_fail("Model file %s holds %zu floats, expected NX*NY=%zu.", fn, got, (size_t)nx * ny);
            }
        }
        fclose(fp);
    }
    return 0;
}

/* --------------------------------------------------------- averaging */

static void av_mue(const float *u, float *uipjp, int nx, int ny, int pitch_)
{
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            float a = AT(u, i, j), b = AT(u, i + 1, j), c = AT(u, i, j + 1), d = AT(u, i + 1, j + 1);
            if (a == 0.f || b == 0.f || c == 0.f || d == 0.f) {
                AT(uipjp, i, j) = 0.f;
            } else {
                double sum = (1.0 / a) + (1.0 / b) + (1.0 / c) + (1.0 / d);
                AT(uipjp, i, j) = (sum != 0.0) ? (float)(4.0 / sum) : 0.0f;
            }
        }
}

static void av_rho(const float *rho, float *rip, float *rjp, int nx, int ny, int pitch_)
{
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            AT(rjp, i, j) = (float)(1.0 / (0.5 * (AT(rho, i, j) + AT(rho, i, j + 1))));
            AT(rip, i, j) = (float)(1.0 / (0.5 * (AT(rho, i, j) + AT(rho, i + 1, j))));
        }
}

static void av_tau(const float *t, float *tipjp, int nx, int ny, int pitch_)
{
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++)
            AT(tipjp, i, j) = (float)(0.25 * (AT(t, i, j) + AT(t, i + 1, j) + AT(t, i, j + 1) + AT(t, i + 1, j + 1)));
}

/* periodic material halo of width one (matcopy on a wrapped process grid) */
static void halo_periodic(float *a, int nx, int ny, int pitch_)
{
    for (int i = 1; i <= nx; i++) {
        AT(a, i, ny + 1) = AT(a, i, 1);
        AT(a, i, 0) = AT(a, i, ny);
    }
    for (int j = 0; j <= ny + 1; j++) {
        AT(a, nx + 1, j) = AT(a, 1, j);
        AT(a, 0, j) = AT(a, nx, j);
    }
}

/* matcopy_elastic's left / right exchange with BOUNDARY = 1 on one
 * process: columns 0 and NX + 1 of rows 0..NY + 1 wrap around
 * (src/matcopy_elastic.c:84-119); the zero rows 0 and NY + 1 stay zero */
static void halo_periodic_x(float *a, int nx, int ny, int pitch_)
{
    for (int j = 0; j <= ny + 1; j++) {
        AT(a, nx + 1, j) = AT(a, 1, j);
        AT(a, 0, j) = AT(a, nx, j);
    }
}

static void dt_mult(float *a, float dt, int nx, int ny, int pitch_)
{
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++) AT(a, i, j) = AT(a, i, j) * dt;
}

/* ------------------------------------------------------------ readers */

static float *NEW(sofi_model *m, int id, int planes)
{
    m->a[id] = sofi_alloc_plane(m->nx, m->ny, planes);
    return m->a[id];
}

static void minmax_init(sofi_model *m)
{
    m->vpmin = FLT_MAX;
    m->vpmax = 0.0f;
    m->vsmin = FLT_MAX;
    m->vsmax = 0.0f;
}

static void minmax(sofi_model *m, float vp, float vs, float vpmin, float vpmax, float vsmin, float vsmax)
{
    if (vpmin < m->vpmin && vp > V_IGNORE) m->vpmin = vpmin;
    if (vpmax > m->vpmax && vp > V_IGNORE) m->vpmax = vpmax;
    if (vsmin < m->vsmin && vs > V_IGNORE) m->vsmin = vsmin;
    if (vsmax > m->vsmax && vs > V_IGNORE) m->vsmax = vsmax;
}

static void keep_model(sofi_model *m, int i, int j, float vp, float vs, float rho)
{
    int pitch_ = m->pitch;
    if (m->vp0) {
        AT(m->vp0, i, j) = vp;
        AT(m->vs0, i, j) = vs;
        AT(m->rho0, i, j) = rho;
    }
}

/* relaxation parameters shared by the visco readers */
static void visco_taus(const sofi_params *p, sofi_model *m, float *pts, float *ws)
{
    float *eta = m->prof[SOFI_P_ETA];
    for (int l = 1; l <= p->L; l++) {
        pts[l] = (float)(1.0 / (2.0 * M_PI * p->fl[l]));
        eta[l] = p->dt / pts[l];
    }
    *ws = (float)(2.0 * M_PI * p->f_ref);
}

/* readmod_acoustic (src/readmod_acoustic.c:31-129) and prepmod for AC_ISO
 * (src/prepmod.c:40-43, matcopy_elastic + av_rho; prepare_update_s_ac,
 * src/prepare_update_s_ac.c:26-33: g = pi * DT).  The S velocities are 0,
 * which checkfd takes as the minimum velocity. */
static int build_ac_iso(const sofi_params *p, sofi_model *m, float **para)
{
    int pitch_ = m->pitch;
    float *ppi = NEW(m, SOFI_A_PPI, 1), *rho = m->prho;
    minmax_init(m);
    for (int i = 1; i <= p->nxg; i++)
        for (int j = 1; j <= p->nyg; j++) {
            size_t k = (size_t)(i - 1) * p->nyg + (j - 1);
            float vp = para[0][k], r = para[1][k];
            minmax(m, vp, 0.0f, vp, vp, 0.0f, 0.0f);
            AT(rho, i, j) = r;
            AT(ppi, i, j) = vp * vp * r;
        }
    m->vsmin = 0.0f;
    m->vsmax = 0.0f;
    if (p->boundary == 1) {
        halo_periodic_x(rho, m->nx, m->ny, pitch_);
        halo_periodic_x(ppi, m->nx, m->ny, pitch_);
    }
    av_rho(rho, NEW(m, SOFI_A_RIP, 1), NEW(m, SOFI_A_RJP, 1), m->nx, m->ny, pitch_);
    float *g = NEW(m, SOFI_A_G, 1);
    for (int j = 1; j <= m->ny; j++)
        for (int i = 1; i <= m->nx; i++) AT(g, i, j) = AT(ppi, i, j) * p->dt;
    return 0;
}

static int build_el_iso(const sofi_params *p, sofi_model *m, float **para)
{
    int pitch_ = m->pitch;
    float *pu = NEW(m, SOFI_A_PU, 1), *ppi = NEW(m, SOFI_A_PPI, 1), *rho = m->prho;
    minmax_init(m);
    for (int i = 1; i <= p->nxg; i++)
        for (int j = 1; j <= p->nyg; j++) {
            size_t k = (size_t)(i - 1) * p->nyg + (j - 1);
            float vp = para[0][k], vs = para[1][k], r = para[2][k];
            minmax(m, vp, vs, vp, vp, vs, vs);
            keep_model(m, i, j, vp, vs, r);
            AT(pu, i, j) = vs * vs * r;
            AT(rho, i, j) = r;
            AT(ppi, i, j) = vp * vp * r;
        }
    /* prepmod EL_ISO (matcopy_elastic: the x halo wraps with BOUNDARY = 1) */
    if (p->boundary == 1) {
        halo_periodic_x(rho, m->nx, m->ny, pitch_);
        halo_periodic_x(ppi, m->nx, m->ny, pitch_);
        halo_periodic_x(pu, m->nx, m->ny, pitch_);
    }
    float *puipjp = NEW(m, SOFI_A_PUIPJP, 1);
    av_mue(pu, puipjp, m->nx, m->ny, pitch_);
    av_rho(rho, NEW(m, SOFI_A_RIP, 1), NEW(m, SOFI_A_RJP, 1), m->nx, m->ny, pitch_);
    float *fipjp = NEW(m, SOFI_A_FIPJP, 1), *f = NEW(m, SOFI_A_F, 1), *g = NEW(m, SOFI_A_G, 1);
    /* with FDORDER_TIME = 4, prepare_update_s_el_4 computes these into locals
     * and stores nothing (src/prepare_update_s_el_4.c:33-45): f, g and fipjp
     * keep the zeros of their allocation, and the stresses change only
     * through the explosive / moment sources */
    if (p->fdorder_time == 4) return 0;
    for (int j = 1; j <= m->ny; j++)
        for (int i = 1; i <= m->nx; i++) {
            AT(fipjp, i, j) = AT(puipjp, i, j) * p->dt;
            AT(f, i, j) = AT(pu, i, j) * p->dt;
            AT(g, i, j) = AT(ppi, i, j) * p->dt;
        }
    return 0;
}

static int build_vel_iso(const sofi_params *p, sofi_model *m, float **para)
{
    int pitch_ = m->pitch, L = p->L;
    float pts[SOFI_MAXL + 1], ws;
    visco_taus(p, m, pts, &ws);
    float *pu = NEW(m, SOFI_A_PU, 1), *ppi = NEW(m, SOFI_A_PPI, 1), *rho = m->prho;
    float *taus = NEW(m, SOFI_A_PTAUS, 1), *taup = NEW(m, SOFI_A_PTAUP, 1);
    minmax_init(m);
    for (int i = 1; i <= p->nxg; i++)
        for (int j = 1; j <= p->nyg; j++) {
            size_t k = (size_t)(i - 1) * p->nyg + (j - 1);
            float vp = para[0][k], vs = para[1][k], r = para[2][k];
            /* relaxation times from Q, or given (the internal model, para[5], para[6]) */
            float tp = para[5] ? para[5][k] : (float)(2.0 / (para[3][k] * L));
            float ts = para[6] ? para[6][k] : (float)(2.0 / (para[4][k] * L));
            float sumu = 0.0f, sumpi = 0.0f;
            for (int l = 1; l <= L; l++) {
                sumu += ((ws * ws * pts[l] * pts[l] * ts) / (1.0 + ws * ws * pts[l] * pts[l]));
                sumpi += ((ws * ws * pts[l] * pts[l] * tp) / (1.0 + ws * ws * pts[l] * pts[l]));
            }
            float vsmin = (float)(vs / sqrt(1.0 + sumu));
            float vsmax = (float)(vsmin * sqrt(1 + L * ts));
            float vpmin = (float)(vp / sqrt(1.0 + sumpi));
            float vpmax = (float)(vpmin * sqrt(1 + L * tp));
            minmax(m, vp, vs, vpmin, vpmax, vsmin, vsmax);
            keep_model(m, i, j, vp, vs, r);
            float muv = (float)(vs * vs * r / (1.0 + sumu));
            float piv = (float)(vp * vp * r / (1.0 + sumpi));
            AT(taus, i, j) = ts;
            AT(taup, i, j) = tp;
            AT(pu, i, j) = muv;
            AT(rho, i, j) = r;
            AT(ppi, i, j) = piv;
        }
    /* prepmod VEL_ISO: matcopy (periodic halo), averages, prepare_update_s_visc */
    halo_periodic(rho, m->nx, m->ny, pitch_);
    halo_periodic(ppi, m->nx, m->ny, pitch_);
    halo_periodic(pu, m->nx, m->ny, pitch_);
    halo_periodic(taus, m->nx, m->ny, pitch_);
    halo_periodic(taup, m->nx, m->ny, pitch_);
    float *puipjp = NEW(m, SOFI_A_PUIPJP, 1), *tausipjp = NEW(m, SOFI_A_PTAUSIPJP, 1);
    av_mue(pu, puipjp, m->nx, m->ny, pitch_);
    av_rho(rho, NEW(m, SOFI_A_RIP, 1), NEW(m, SOFI_A_RJP, 1), m->nx, m->ny, pitch_);
    av_tau(taus, tausipjp, m->nx, m->ny, pitch_);
    float *fipjp = NEW(m, SOFI_A_FIPJP, 1), *f = NEW(m, SOFI_A_F, 1), *g = NEW(m, SOFI_A_G, 1);
    float *dip = NEW(m, SOFI_A_DIP, L), *d = NEW(m, SOFI_A_D, L), *e = NEW(m, SOFI_A_E, L);
    float *eta = m->prof[SOFI_P_ETA], *bip = m->prof[SOFI_P_BIP], *cip = m->prof[SOFI_P_CIP];
    size_t pl = sofi_plane(m->nx, m->ny);
    for (int j = 1; j <= m->ny; j++)
        for (int i = 1; i <= m->nx; i++) {
            AT(fipjp, i, j) = (float)(AT(puipjp, i, j) * p->dt * (1.0 + L * AT(tausipjp, i, j)));
            AT(f, i, j) = (float)(AT(pu, i, j) * p->dt * (1.0 + L * AT(taus, i, j)));
            AT(g, i, j) = (float)(AT(ppi, i, j) * p->dt * (1.0 + L * AT(taup, i, j)));
            for (int l = 1; l <= L; l++) {
                if (p->fdorder_time == 4) {
                    /* prepare_update_s_visc_4 (src/prepare_update_s_visc_4.c:36-48):
                     * the relaxation factors carry the first Adams-Bashforth weight */
                    const float c1 = (float)(13.0 / 12.0);
                    bip[l] = (float)(1.0 / (1.0 + (c1 * eta[l] * 0.5)));
                    cip[l] = (float)(1.0 - (c1 * eta[l] * 0.5));
                } else {
                    bip[l] = (float)(1.0 / (1.0 + (eta[l] * 0.5)));
                    cip[l] = (float)(1.0 - (eta[l] * 0.5));
                }
                size_t o = (size_t)(l - 1) * pl + sofi_idx(pitch_, i, j);
                dip[o] = AT(puipjp, i, j) * eta[l] * AT(tausipjp, i, j);
                d[o] = AT(pu, i, j) * eta[l] * AT(taus, i, j);
                e[o] = AT(ppi, i, j) * eta[l] * AT(taup, i, j);
            }
        }
    return 0;
}

/* Thomsen -> VTI stiffness (src/readmod_elastic_vti.c:146-149) */
static void thomsen(float r, float vp, float vs, float eps, float del, float *c11, float *c13, float *c33, float *c55)
{
    float c33_ = r * vp * vp;
    float c55_ = r * vs * vs;
    float c11_ = (float)(c33_ * (2.0 * eps + 1.0));
    float c13_ = (float)(sqrt((2.0 * del * c33_ * (c33_ - c55_)) + ((c33_ - c55_) * (c33_ - c55_))) - c55_);
    *c11 = c11_;
    *c13 = c13_;
    *c33 = c33_;
    *c55 = c55_;
}

static void aniso_minmax(sofi_model *m, float vp, float vs, float eps, float del,
                         float *vpmin, float *vpmax, float *vsmin, float *vsmax)
{
    float vp_90 = (float)(vp * sqrt(2.0 * eps + 1.0));
    float vs_45 = vs > 0.0 ? (float)(vs * (1.0 + pow(vp / vs, 2) * 0.25 * (eps - del))) : 0.0f;
    (void)m;
    *vpmax = vp_90 > vp ? vp_90 : vp;
    *vpmin = vp_90 < vp ? vp_90 : vp;
    *vsmax = vs_45 > vs ? vs_45 : vs;
    *vsmin = vs_45 < vs ? vs_45 : vs;
}

/* Bond rotation by theta (src/readmod_elastic_tti.c:160-182) */
static void bond(float c11, float c13, float c33, float c55, float theta, float *o11, float *o13, float *o33,
                 float *o55, float *o15, float *o35)
{
    float t = (float)(theta * M_PI / 180.0);
    float l1 = (float)cos(t), l2 = (float)sin(t);
    float l12 = l1 * l1, l22 = l2 * l2, l14 = l12 * l12, l24 = l22 * l22, l13 = l1 * l12, l23 = l2 * l22;
    float a1 = (float)(2.0 * c13 + 4.0 * c55);
    float a3 = (float)(c11 + c33 - 4.0 * c55);
    float a4 = (float)(c11 + c33 - 2.0 * c13);
    float a5 = (float)(c13 - c11 + 2.0 * c55);
    float a6 = (float)(c13 - c33 + 2.0 * c55);
    *o11 = c11 * l14 + c33 * l24 + a1 * l12 * l22;
    *o33 = c11 * l24 + c33 * l14 + a1 * l12 * l22;
    *o13 = a3 * l12 * l22 + c13 * (l14 + l24);
    *o55 = a4 * l12 * l22 + c55 * (l12 - l22) * (l12 - l22);
    *o15 = a5 * l13 * l2 - a6 * l1 * l23;
    *o35 = a5 * l23 * l1 - a6 * l2 * l13;
}

/* The stiffness of an internal anisotropic model (READMOD = 0): the same in
 * every cell; tau (visco) from the parameter TAU, theta in degrees (TTI). */
typedef struct {
    float c11, c13, c33, c55, rho, tau, theta;
} sofi_builtin_aniso;

static int build_aniso(const sofi_params *p, sofi_model *m, float **para, const sofi_builtin_aniso *bi)
{
    int pitch_ = m->pitch, L = p->L, weq = p->weq;
    int visco = (weq == SOFI_VEL_VTI || weq == SOFI_VEL_TTI);
    int tti = (weq == SOFI_EL_TTI || weq == SOFI_VEL_TTI);
    /* para order: vp vs rho eps delta [qp qs] [theta] */
    int iq = 5, it = visco ? 7 : 5;
    float pts[SOFI_MAXL + 1], ws = 0.0f;
    if (visco) visco_taus(p, m, pts, &ws);
    float *c11 = NEW(m, SOFI_A_PPI, 1), *c55 = NEW(m, SOFI_A_PU, 1);
    float *c13 = NEW(m, SOFI_A_C13, 1), *c33 = NEW(m, SOFI_A_C33, 1), *rho = m->prho;
    float *c15 = tti ? NEW(m, SOFI_A_C15, 1) : NULL, *c35 = tti ? NEW(m, SOFI_A_C35, 1) : NULL;
    /* tau arrays (visco): 11, 33, 13, 55, 15, 35 */
    float *t11 = NULL, *t33 = NULL, *t13 = NULL, *t55 = NULL, *t15 = NULL, *t35 = NULL;
    if (visco) {
        t11 = sofi_alloc_plane(m->nx, m->ny, 1);
        t33 = sofi_alloc_plane(m->nx, m->ny, 1);
        t13 = sofi_alloc_plane(m->nx, m->ny, 1);
        t55 = sofi_alloc_plane(m->nx, m->ny, 1);
        if (tti) {
            t15 = sofi_alloc_plane(m->nx, m->ny, 1);
            t35 = sofi_alloc_plane(m->nx, m->ny, 1);
        }
    }
    minmax_init(m);
    for (int i = 1; i <= p->nxg; i++)
        for (int j = 1; j <= p->nyg; j++) {
            size_t k = (size_t)(i - 1) * p->nyg + (j - 1);
            float vp, vs, r, eps, del, a11, a13, a33, a55;
            if (bi) {
                /* velocities and Thomsen parameters of the given stiffness,
                 * for the stability checks only */
                r = bi->rho;
                a11 = bi->c11;
                a13 = bi->c13;
                a33 = bi->c33;
                a55 = bi->c55;
                vp = (float)sqrt(a33 / r);
                vs = (float)sqrt(a55 / r);
                eps = (float)((a11 - a33) / (2.0 * a33));
                del = (float)(((a13 + a55) * (double)(a13 + a55) - (a33 - a55) * (double)(a33 - a55)) /
                              (2.0 * a33 * (a33 - a55)));
            } else {
                vp = para[0][k];
                vs = para[1][k];
                r = para[2][k];
                eps = para[3][k];
                del = para[4][k];
                thomsen(r, vp, vs, eps, del, &a11, &a13, &a33, &a55);
            }
            float vpmin, vpmax, vsmin, vsmax;
            aniso_minmax(m, vp, vs, eps, del, &vpmin, &vpmax, &vsmin, &vsmax);
            float tau11 = 0, tau33 = 0, tau13 = 0, tau55 = 0;
            if (visco) {
                float tp = bi ? bi->tau : (float)(2.0 / (para[iq][k] * L));
                float ts = bi ? bi->tau : (float)(2.0 / (para[iq + 1][k] * L));
                float sumu = 0.0f, sumpi = 0.0f;
                for (int l = 1; l <= L; l++) {
                    sumu += ((ws * ws * pts[l] * pts[l] * ts) / (1.0 + ws * ws * pts[l] * pts[l]));
                    sumpi += ((ws * ws * pts[l] * pts[l] * tp) / (1.0 + ws * ws * pts[l] * pts[l]));
                }
                vsmin = (float)(vsmin / sqrt(1.0 + sumu));
                vsmax = (float)(vsmax * sqrt((1 + L * ts) / (1.0 + sumu)));
                vpmin = (float)(vpmin / sqrt(1.0 + sumpi));
                vpmax = (float)(vpmax * sqrt((1 + L * tp) / (1.0 + sumpi)));
                tau11 = tau33 = tp;
                tau55 = ts;
                a11 = (float)(a11 / (1.0 + sumpi));
                a33 = (float)(a33 / (1.0 + sumpi));
                a55 = (float)(a55 / (1.0 + sumu));
                tau13 = (float)((a11 * L * tau11 - 2.0 * a55 * L * tau55) / (a11 - 2.0 * a55));
                float sumc13 = 0.0f;
                for (int l = 1; l <= L; l++)
                    sumc13 += ((ws * ws * pts[l] * pts[l] * tau13) / (1.0 + ws * ws * pts[l] * pts[l]));
                a13 = (float)(a13 / (1.0 + sumc13));
            }
            minmax(m, vp, vs, vpmin, vpmax, vsmin, vsmax);
            keep_model(m, i, j, vp, vs, r);
            if (tti) {
                float o11, o13, o33, o55, o15, o35;
                bond(a11, a13, a33, a55, bi ? bi->theta : para[it][k], &o11, &o13, &o33, &o55, &o15, &o35);
                AT(c11, i, j) = o11;
                AT(c33, i, j) = o33;
                AT(c13, i, j) = o13;
                AT(c55, i, j) = o55;
                AT(c15, i, j) = o15;
                AT(c35, i, j) = o35;
            } else {
                AT(c11, i, j) = a11;
                AT(c13, i, j) = a13;
                AT(c33, i, j) = a33;
                AT(c55, i, j) = a55;
            }
            AT(rho, i, j) = r;
            if (visco) {
                AT(t11, i, j) = tau11;
                AT(t33, i, j) = tau33;
                AT(t13, i, j) = tau13;
                AT(t55, i, j) = tau55;
                if (tti) {
                    AT(t15, i, j) = tau13;
                    AT(t35, i, j) = tau13;
                }
            }
        }
    /* prepmod: averages with a zero halo (matcopy_elastic at the global edge;
     * with BOUNDARY = 1 its x halo wraps: rho and c55, c15 / c35 (TTI), the
     * tau arrays that are averaged (visco), src/prepmod.c:56-95) */
    int nx = m->nx, ny = m->ny;
    if (p->boundary == 1) {
        float *wrap[] = { rho, c55, visco ? NULL : c11, tti ? c15 : NULL, tti ? c35 : NULL, visco ? t55 : NULL,
                          visco && tti ? t15 : NULL, visco && tti ? t35 : NULL };
        for (size_t k = 0; k < sizeof(wrap) / sizeof(wrap[0]); ++k)
            if (wrap[k]) halo_periodic_x(wrap[k], nx, ny, pitch_);
    }
    float *c55ipjp = NEW(m, SOFI_A_PUIPJP, 1);
    av_mue(c55, c55ipjp, nx, ny, pitch_);
    float *c15ipjp = NULL, *c35ipjp = NULL;
    if (tti) {
        c15ipjp = NEW(m, SOFI_A_C15IPJP, 1);
        c35ipjp = NEW(m, SOFI_A_C35IPJP, 1);
        av_mue(c15, c15ipjp, nx, ny, pitch_);
        av_mue(c35, c35ipjp, nx, ny, pitch_);
    }
    av_rho(rho, NEW(m, SOFI_A_RIP, 1), NEW(m, SOFI_A_RJP, 1), nx, ny, pitch_);
    if (!visco) {
        /* dt_mult in place (src/prepmod.c:134-152); the aliases ppi = pc11 and
         * puipjp = pc55ipjp therefore carry the DT factor afterwards */
        dt_mult(c11, p->dt, nx, ny, pitch_);
        dt_mult(c33, p->dt, nx, ny, pitch_);
        dt_mult(c13, p->dt, nx, ny, pitch_);
        if (tti) {
            dt_mult(c15, p->dt, nx, ny, pitch_);
            dt_mult(c35, p->dt, nx, ny, pitch_);
        }
        dt_mult(c55ipjp, p->dt, nx, ny, pitch_);
        if (tti) {
            dt_mult(c35ipjp, p->dt, nx, ny, pitch_);
            dt_mult(c15ipjp, p->dt, nx, ny, pitch_);
        }
        /* kernels read the named arrays; share storage with the aliases */
        m->a[SOFI_A_C11] = c11;
        m->a[SOFI_A_C55IPJP] = c55ipjp;
        return 0;
    }
    /* visco VTI/TTI: tau averages and prepare_update_s_{vti,tti} */
    float *t55ipjp = sofi_alloc_plane(nx, ny, 1), *t15ipjp = NULL, *t35ipjp = NULL;
    av_tau(t55, t55ipjp, nx, ny, pitch_);
    if (tti) {
        t15ipjp = sofi_alloc_plane(nx, ny, 1);
        t35ipjp = sofi_alloc_plane(nx, ny, 1);
        av_tau(t15, t15ipjp, nx, ny, pitch_);
        av_tau(t35, t35ipjp, nx, ny, pitch_);
    }
    float *eta = m->prof[SOFI_P_ETA], *bip = m->prof[SOFI_P_BIP], *cip = m->prof[SOFI_P_CIP];
    float *u11 = NEW(m, SOFI_A_C11U, 1), *u13 = NEW(m, SOFI_A_C13U, 1), *u33 = NEW(m, SOFI_A_C33U, 1);
    float *u55 = NEW(m, SOFI_A_C55IPJPU, 1);
    float *d11 = NEW(m, SOFI_A_C11D, L), *d13 = NEW(m, SOFI_A_C13D, L), *d33 = NEW(m, SOFI_A_C33D, L);
    float *d55 = NEW(m, SOFI_A_C55IPJPD, L);
    float *u15 = NULL, *u35 = NULL, *u15i = NULL, *u35i = NULL, *d15 = NULL, *d35 = NULL, *d15i = NULL, *d35i = NULL;
    if (tti) {
        u15 = NEW(m, SOFI_A_C15U, 1);
        u35 = NEW(m, SOFI_A_C35U, 1);
        u15i = NEW(m, SOFI_A_C15IPJPU, 1);
        u35i = NEW(m, SOFI_A_C35IPJPU, 1);
        d15 = NEW(m, SOFI_A_C15D, L);
        d35 = NEW(m, SOFI_A_C35D, L);
        d15i = NEW(m, SOFI_A_C15IPJPD, L);
        d35i = NEW(m, SOFI_A_C35IPJPD, L);
    }
    size_t pl = sofi_plane(nx, ny);
    float dt = p->dt;
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            size_t c = sofi_idx(pitch_, i, j);
            u55[c] = (float)(c55ipjp[c] * dt * (1.0 + L * t55ipjp[c]));
            u13[c] = (float)(c13[c] * dt * (1.0 + L * t13[c]));
            u11[c] = (float)(c11[c] * dt * (1.0 + L * t11[c]));
            u33[c] = (float)(c33[c] * dt * (1.0 + L * t33[c]));
            if (tti) {
                u15[c] = (float)(c15[c] * dt * (1.0 + L * t15[c]));
                u35[c] = (float)(c35[c] * dt * (1.0 + L * t35[c]));
                u15i[c] = (float)(c15ipjp[c] * dt * (1.0 + L * t15ipjp[c]));
                u35i[c] = (float)(c35ipjp[c] * dt * (1.0 + L * t35ipjp[c]));
            }
            for (int l = 1; l <= L; l++) {
                bip[l] = (float)(1.0 / (1.0 + (eta[l] * 0.5)));
                cip[l] = (float)(1.0 - (eta[l] * 0.5));
                size_t o = (size_t)(l - 1) * pl + c;
                d55[o] = c55ipjp[c] * eta[l] * t55ipjp[c];
                d13[o] = c13[c] * eta[l] * t13[c];
                d33[o] = c33[c] * eta[l] * t33[c];
                d11[o] = c11[c] * eta[l] * t11[c];
                if (tti) {
                    d15i[o] = c15ipjp[c] * eta[l] * t15ipjp[c];
                    d35i[o] = c35ipjp[c] * eta[l] * t35ipjp[c];
                    d15[o] = c15[c] * eta[l] * t15[c];
                    d35[o] = c35[c] * eta[l] * t35[c];
                }
            }
        }
    m->a[SOFI_A_C11] = c11;
    m->a[SOFI_A_C55IPJP] = c55ipjp;
    free(t11); free(t33); free(t13); free(t55); free(t15); free(t35);
    free(t55ipjp); free(t15ipjp); free(t35ipjp);
    return 0;
}

/* ------------------------------------------------- CPML and damping frame */

/* PML_pro (src/PML_pro.c:86-334): d, K, alpha' on integer and half grids,
 * b = exp(-(d/K + alpha') DT), a = d (b-1) / (K (d + K alpha')). */
static void pml_side(const sofi_params *p, float *K, float *Kh, float *a, float *ah, float *b, float *bh,
                     int n_global, int is_x)
{
    int fw = p->fw;
    const float alpha_max_PML = (float)(2.0 * M_PI * (p->fpml / 2.0));
    float thickness = (float)fw * p->dh;
    float Rcoef = 0.001f;
    float d0 = (float)(-(p->npower + 1) * p->vppml * log(Rcoef) / (2.0 * thickness));
    float origin_lo = thickness;
    float origin_hi = (n_global - 1) * p->dh - thickness;
    float *d = (float *)calloc(2 * fw + 1, sizeof(float)), *dh = (float *)calloc(2 * fw + 1, sizeof(float));
    float *al = (float *)calloc(2 * fw + 1, sizeof(float)), *alh = (float *)calloc(2 * fw + 1, sizeof(float));
    for (int side = 0; side < 2; ++side) {
        int i0 = side == 0 ? 1 : n_global - fw + 1, i1 = side == 0 ? fw : n_global;
        for (int i = i0; i <= i1; i++) {
            int h = side == 0 ? i : i - n_global + 2 * fw;
            K[h] = 1.0f;
            Kh[h] = 1.0f;
            float val = p->dh * (i - 1);
            float absc = side == 0 ? origin_lo - val : val - origin_hi;
            if (absc >= 0.0) {
                float an = absc / thickness;
                d[h] = (float)(d0 * pow(an, p->npower));
                K[h] = (float)(1.0 + (p->k_max_cpml - 1.0) * pow(an, p->npower));
                al[h] = (float)(alpha_max_PML * (1.0 - an));
            }
            absc = side == 0 ? (float)(origin_lo - (val + p->dh / 2.0)) : (float)(val + p->dh / 2.0 - origin_hi);
            if (absc >= 0.0) {
                float an = absc / thickness;
                dh[h] = (float)(d0 * pow(an, p->npower));
                Kh[h] = (float)(1.0 + (p->k_max_cpml - 1.0) * pow(an, p->npower));
                alh[h] = (float)(alpha_max_PML * (1.0 - an));
            }
            if (is_x) { /* the clamp exists for the x profiles only */
                if (al[h] < 0.0) al[h] = 0.0f;
                if (alh[h] < 0.0) alh[h] = 0.0f;
            }
            b[h] = (float)exp(-(d[h] / K[h] + al[h]) * p->dt);
            bh[h] = (float)exp(-(dh[h] / Kh[h] + alh[h]) * p->dt);
            if (fabsf(d[h]) > 1.0e-6) a[h] = (float)(d[h] * (b[h] - 1.0) / (K[h] * (d[h] + K[h] * al[h])));
            if (fabsf(dh[h]) > 1.0e-6) ah[h] = (float)(dh[h] * (bh[h] - 1.0) / (Kh[h] * (dh[h] + Kh[h] * alh[h])));
        }
    }
    free(d); free(dh); free(al); free(alh);
}

static void build_pml(const sofi_params *p, sofi_model *m)
{
    int n = 2 * p->fw + 1;
    for (int k = SOFI_P_K_X; k <= SOFI_P_B_Y_HALF; ++k) m->prof[k] = (float *)calloc(n, sizeof(float));
    pml_side(p, m->prof[SOFI_P_K_X], m->prof[SOFI_P_K_X_HALF], m->prof[SOFI_P_A_X], m->prof[SOFI_P_A_X_HALF],
             m->prof[SOFI_P_B_X], m->prof[SOFI_P_B_X_HALF], p->nxg, 1);
    pml_side(p, m->prof[SOFI_P_K_Y], m->prof[SOFI_P_K_Y_HALF], m->prof[SOFI_P_A_Y], m->prof[SOFI_P_A_Y_HALF],
             m->prof[SOFI_P_B_Y], m->prof[SOFI_P_B_Y_HALF], p->nyg, 0);
}

/* absorb (src/absorb.c:34-121), one process: frames written left, right,
 * top (unless free surface), bottom, later writes winning at corners. */
static void build_absorb(const sofi_params *p, sofi_model *m)
{
    int pitch_ = m->pitch, fw = p->fw, nx = m->nx, ny = m->ny;
    float *ac = NEW(m, SOFI_A_ABSORB, 1);
    float amp = (float)(1.0 - p->damping / 100.0);
    float *coeff = (float *)calloc(fw + 1, sizeof(float));
    float a = (float)sqrt(-log(amp) / ((fw - 1) * (fw - 1)));
    for (int i = 1; i <= fw; i++) coeff[i] = (float)exp(-(a * a * (fw - i) * (fw - i)));
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) AT(ac, i, j) = 1.0f;
    /* periodic in x (BOUNDARY = 1): no left / right frame, and the top and
     * bottom frames run over the whole width (src/absorb.c:61-112) */
    const int per = p->boundary == 1;
    int yb = 1, ye = ny;
    for (int i = 1; i <= fw && !per; i++) {
        if (!p->free_surf) yb = i;
        ye = ny - i + 1;
        for (int j = yb; j <= ye; j++) AT(ac, i, j) = coeff[i];
    }
    yb = 1;
    ye = ny;
    for (int i = 1; i <= fw && !per; i++) {
        int ii = nx - i + 1;
        if (!p->free_surf) yb = i;
        ye = ny - i + 1;
        for (int j = yb; j <= ye; j++) AT(ac, ii, j) = coeff[i];
    }
    if (!p->free_surf) {
        for (int j = 1; j <= fw; j++) {
            int xb = per ? 1 : j, xe = per ? nx : nx - j + 1;
            for (int i = xb; i <= xe; i++) AT(ac, i, j) = coeff[j];
        }
    }
    for (int j = 1; j <= fw; j++) {
        int jj = ny - j + 1, xb = per ? 1 : j, xe = per ? nx : nx - j + 1;
        for (int i = xb; i <= xe; i++) AT(ac, i, jj) = coeff[j];
    }
    free(coeff);
}

/* ------------------------------------------------------------- driver */

/* ----------------------------------------- internal models (READMOD = 0)
 *
 * The reference's built-in test media, with their constants, handed to the
 * same prepmod as a model read from files:
 *   model_elastic      src/hh_elastic.c:37-88      a 100 m layer of vp 0,
 *                      vs 1, rho 1 over vp 3500, vs 2000, rho 2000
 *   model_visco        src/hh_visco.c:38-102       the same, tau_p = tau_s =
 *                      0.01, and 2.0 for 95 m <= y <= 110 m
 *   model_elastic_VTI  src/hh_elastic_VTI.c:39-72  c11 34.3, c33 22.7, c55
 *                      5.4, c13 10.7 GPa, rho 2000 (Jones and Wang 1981)
 *   model_visco_vti    src/hh_visco_VTI.c:41-123   the same, every tau = TAU
 *   model_elastic_TTI  src/hh_elastic_TTI.c:42-104 the VTI medium rotated by 0
 *   model_visco_tti    src/hh_visco_TTI.c:44-150   the visco VTI medium
 *                      rotated by 30 degrees
 * (their layer switches at j < -1 and y < -1 never fire). */
static int build_internal(const sofi_params *p, sofi_model *m, float **para)
{
    const size_t n = (size_t)p->nxg * p->nyg;
    const int vel = p->weq == SOFI_VEL_ISO;
    if (p->weq == SOFI_EL_ISO || vel) {
        for (int k = 0; k < 3; ++k) para[k] = (float *)malloc(n * sizeof(float));
        if (vel) {
            para[5] = (float *)malloc(n * sizeof(float));
            para[6] = (float *)malloc(n * sizeof(float));
        }
        for (int k = 0; k < (vel ? 7 : 3); ++k)
            if ((k < 3 || k >= 5) && !para[k]) return sofi_fail("out of memory building the internal model");
        for (int i = 1; i <= p->nxg; i++)
            for (int j = 1; j <= p->nyg; j++) {
                const size_t k = (size_t)(i - 1) * p->nyg + (j - 1);
                const float y = (float)j * p->dh;
                const int top = y <= 100.0f;
                para[0][k] = top ? 0.0f : 3500.0f;
                para[1][k] = top ? 1.0f : 2000.0f;
                para[2][k] = top ? 1.0f : 2000.0f;
                if (vel) {
                    const float t = (y >= 95.0f && y <= 110.0f) ? 2.0f : 0.01f;
                    para[5][k] = t;
                    para[6][k] = t;
                }
            }
        return vel ? build_vel_iso(p, m, para) : build_el_iso(p, m, para);
    }
    sofi_builtin_aniso bi = { 34.3e9f, 10.7e9f, 22.7e9f, 5.4e9f, 2000.0f, p->tau, 0.0f };
    switch (p->weq) {
      case SOFI_EL_VTI:
      case SOFI_VEL_VTI:
      case SOFI_EL_TTI: return build_aniso(p, m, para, &bi);
      case SOFI_VEL_TTI: bi.theta = 30.0f; return build_aniso(p, m, para, &bi);
    }
    return sofi_fail("Internal model for your chosen WEQ not implemented.");
}

int sofi_build_model(const sofi_params *p, sofi_model *m)
{
    memset(m, 0, sizeof(*m));
    m->nx = p->nxg;
    m->ny = p->nyg;
    m->pitch = sofi_pitch(m->nx);
    m->rows = sofi_rows(m->ny);
    m->L = p->L;
    m->weq = p->weq;
    int L = p->L > 0 ? p->L : 0;
    m->prof[SOFI_P_ETA] = (float *)calloc(L + 1, sizeof(float));
    m->prof[SOFI_P_BIP] = (float *)calloc(L + 1, sizeof(float));
    m->prof[SOFI_P_CIP] = (float *)calloc(L + 1, sizeof(float));
    m->prho = sofi_alloc_plane(m->nx, m->ny, 1);
    if (p->mode == SOFI_MODE_FWI) {
        m->vp0 = sofi_alloc_plane(m->nx, m->ny, 1);
        m->vs0 = sofi_alloc_plane(m->nx, m->ny, 1);
        m->rho0 = sofi_alloc_plane(m->nx, m->ny, 1);
    }
    float *para[8] = { 0 };
    int rc = 0;
    if (p->readmod == 0) {
        rc = build_internal(p, m, para);
        for (int k = 0; k < 8; ++k) free(para[k]);
        if (rc) return rc;
        if (p->fw > 0 && p->abs_type == 1) build_pml(p, m);
        if (p->fw > 0 && p->abs_type == 2) build_absorb(p, m);
        return 0;
    }
    switch (p->weq) {
      case SOFI_AC_ISO: {
          const char *sfx[] = { "vp", "rho" };
          rc = read_model_files(p, 2, sfx, para);
          if (!rc) rc = build_ac_iso(p, m, para);
          break;
      }
      case SOFI_EL_ISO: {
          const char *sfx[] = { "vp", "vs", "rho" };
          rc = read_model_files(p, 3, sfx, para);
          if (!rc) rc = build_el_iso(p, m, para);
          break;
      }
      case SOFI_VEL_ISO: {
          const char *sfx[] = { "vp", "vs", "rho", "qp", "qs" };
          rc = read_model_files(p, 5, sfx, para);
          if (!rc) rc = build_vel_iso(p, m, para);
          break;
      }
      case SOFI_EL_VTI: {
          const char *sfx[] = { "vp", "vs", "rho", "epsilon", "delta" };
          rc = read_model_files(p, 5, sfx, para);
          if (!rc) rc = build_aniso(p, m, para, NULL);
          break;
      }
      case SOFI_VEL_VTI: {
          const char *sfx[] = { "vp", "vs", "rho", "epsilon", "delta", "qp", "qs" };
          rc = read_model_files(p, 7, sfx, para);
          if (!rc) rc = build_aniso(p, m, para, NULL);
          break;
      }
      case SOFI_EL_TTI: {
          const char *sfx[] = { "vp", "vs", "rho", "epsilon", "delta", "theta" };
          rc = read_model_files(p, 6, sfx, para);
          if (!rc) rc = build_aniso(p, m, para, NULL);
          break;
      }
      case SOFI_VEL_TTI: {
          const char *sfx[] = { "vp", "vs", "rho", "epsilon", "delta", "qp", "qs", "theta" };
          rc = read_model_files(p, 8, sfx, para);
          if (!rc) rc = build_aniso(p, m, para, NULL);
          break;
      }
      default:
          rc = sofi_fail("Unknown WEQ.");
    }
    for (int k = 0; k < 8; ++k) free(para[k]);
    if (rc) return rc;
    if (p->fw > 0 && p->abs_type == 1) build_pml(p, m);
    if (p->fw > 0 && p->abs_type == 2) build_absorb(p, m);
    return 0;
}

void sofi_free_model(sofi_model *m)
{
    /* aliases share storage: free each distinct pointer once */
    for (int k = 0; k < SOFI_A_NARRAYS; ++k) {
        if (!m->a[k]) continue;
        for (int q = k + 1; q < SOFI_A_NARRAYS; ++q)
            if (m->a[q] == m->a[k]) m->a[q] = NULL;
        free(m->a[k]);
        m->a[k] = NULL;
    }
    for (int k = 0; k < SOFI_P_NPROFILES; ++k) { free(m->prof[k]); m->prof[k] = NULL; }
    free(m->prho); free(m->vp0); free(m->vs0); free(m->rho0);
    m->prho = m->vp0 = m->vs0 = m->rho0 = NULL;
}
