/*
 * sofi_fwi.c - host side of the FWI gradient path: the measured data, the
 * adjoint sources (residuals) and the gradient files.
 *
 *   sofi_read_data      inseis      src/inseis.c:28-80
 *   sofi_residual       calc_res    src/calc_res.c:26-324 (LNORM 1-5, 7, 8)
 *   sofi_time_filter    timedomain_filt src/timedomain_filt.c:28-113 (TIME_FILT)
 *   sofi_write_gradient output_gradient + writemod
 *                                   src/output_gradient.c:25-68, src/writemod.c:28-45
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sofi_host_internal.h"
#include "../common/sofi_iir.h"

/* Traces 1..ntr of DATA_DIR_<vx|vy|p>.su.shot<ishot>: each a 240-byte SU
 * header (skipped) and NS float32 samples, into section[ntr][ns]. */
int sofi_read_data(const sofi_params *p, int ishot, int comp, int ntr, float *section)
{
    const char *name = comp == SOFI_C_VX ? "vx" : comp == SOFI_C_VY ? "vy" : comp == SOFI_C_P ? "p" : NULL;
    if (!name) return sofi_fail("sofi_read_data: component %d has no measured data file.", comp);
    char path[2 * SOFI_STR + 32];
    snprintf(path, sizeof(path), "%s_%s.su.shot%d", p->data_dir, name, ishot);
    FILE *fp = fopen(path, "r");
    if (!fp) return sofi_fail("Was not able to read %s: Seismograms for inversion were not found.", path);
    for (int r = 0; r < ntr; r++) {
        if (fseek(fp, 240, SEEK_CUR) != 0) {
            fclose(fp);
            return sofi_fail("Could not skip over trace header in SU file %s.", path);
        }
        if (fread(section + (size_t)r * p->ns, sizeof(float), (size_t)p->ns, fp) != (size_t)p->ns) {
            fclose(fp);
            return sofi_fail("Short read in SU file %s (trace %d).", path, r + 1);
        }
    }
    fclose(fp);
    return 0;
}

/* ---------------------------------------------------- TIME_FILT filters */

/* Recursive (IIR) section y[k] = f0 x[k] + f1 x[k-1] + f2 x[k-2] + g1 y[k-1]
 * + g2 y[k-2] over x[0..n-1] in place, zero initial state: the recursion of
 * libseife's rekfl, which timedomain_filt uses (src/timedomain_filt.c:62-99,
 * contrib/Seitosh/libseife/cseife_rekfl.c). */
static void iir_run(double *x, int n, sofi_iir2 c)
{
    double x1 = 0., x2 = 0., y1 = 0., y2 = 0.;
    for (int k = 0; k < n; ++k) {
        const double xn = x[k];
        const double y = c.f0 * xn + c.f1 * x1 + c.f2 * x2 + c.g1 * y1 + c.g2 * y2;
        x[k] = y;
        x2 = x1;
        x1 = xn;
        y2 = y1;
        y1 = y;
    }
}

/* Butterworth of order m (sofi_iir.h): a high-pass first removes x[0]
 * from every sample, then the sections run one after the other. */
static void butterworth(double *x, int n, double dt, double period, int m, int highpass)
{
    if (highpass) {
        const double x0 = x[0];
        for (int k = 0; k < n; ++k) x[k] -= x0;
    }
    sofi_iir2 sec[SOFI_IIR_MAXSEC];
    const int ns = sofi_iir_butterworth(dt, period, m, highpass, sec);
    for (int k = 0; k < ns; ++k) iir_run(x, n, sec[k]);
}

/* One pass of timedomain_filt over a trace: samples 1..NS at x[1..NS] with
 * x[0] = 0 (the reference filters NS + 1 values, its buffer's element 0
 * included), filtered forward, rounded to float, reversed, filtered again,
 * reversed and rounded back. */
static void zero_phase(float *tr, int ns, double dt, double period, int m, int highpass, double *a, double *b)
{
    a[0] = 0.;
    for (int j = 1; j <= ns; ++j) a[j] = (double)tr[j - 1];
    butterworth(a, ns + 1, dt, period, m, highpass);
    b[0] = 0.;
    for (int j = 1; j <= ns; ++j) b[ns - j + 1] = (double)(float)a[j];
    butterworth(b, ns + 1, dt, period, m, highpass);
    for (int j = 1; j <= ns; ++j) tr[j - 1] = (float)b[ns - j + 1];
}

int sofi_time_filter(const sofi_params *p, int ntr, float *traces)
{
    if (!p->time_filt) return 0;
    const int ns = p->ns;
    double *a = (double *)malloc(sizeof(double) * (size_t)(ns + 1)), *b = (double *)malloc(sizeof(double) * (size_t)(ns + 1));
    if (!a || !b) {
        free(a);
        free(b);
        return sofi_fail("out of memory");
    }
    const double t_lp = 1.0 / (double)p->f_low_pass;
    for (int r = 0; r < ntr; ++r) zero_phase(traces + (size_t)r * ns, ns, p->dt, t_lp, p->filt_order, 0, a, b);
    if (p->f_high_pass) {
        const double t_hp = 1.0 / (double)p->f_high_pass;
        for (int r = 0; r < ntr; ++r) zero_phase(traces + (size_t)r * ns, ns, p->dt, t_hp, p->filt_order, 1, a, b);
    }
    free(a);
    free(b);
    return 0;
}

/* calc_res (src/calc_res.c:26-324), one component, traces 0..ntr-1 of
 * [ntr][NS], the adjoint source stored time-reversed (diff[i][NS - j + 1]):
 *   LNORM 1: sign(measured - synthetic) of the raw sections;
 *   LNORM 2: synthetic - measured, both integrated in time first (running
 *            sum times DT) when VELOCITY = 0;
 *   LNORM 3: r / (1 + r^2) and LNORM 4: tanh(r), r = measured - synthetic
 *            of the raw sections ("not up to date" in the reference);
 *   LNORM 5, 7: the derivative of the normalised correlation of the
 *            (integrated) traces, with the trace norms taken as 1 for a
 *            zero trace;
 *   LNORM 8: the envelope residual, whose envelope and Hilbert transforms
 *            the reference leaves at zero (their calls are commented out):
 *            s / (0 + WATERLEVEL_LNORM8) * (0 - 0) - 0.
 * Float arithmetic in the reference's order; RMS is 1.  *l2 receives the
 * misfit calc_res accumulates for a test shot (float): LNORM 2 the sum of
 * squares, LNORM 5 / 7 minus the sum of the normalised products, LNORM 8
 * the (zero) envelope differences, else 0. */
int sofi_residual(const sofi_params *p, int ntr, const float *data, const float *calc, float *diff, double *l2)
{
    const int ns = p->ns, ln = p->lnorm;
    if (ln < 1 || ln > 8 || ln == 6) return sofi_fail("LNORM=%d has no residual in the reference.", ln);
    if (p->time_filt) {   /* both sections filtered first (src/inversion.c:68-84) */
        const size_t n = (size_t)ntr * (size_t)ns;
        float *fd = (float *)malloc(sizeof(float) * (n ? n : 1)), *fc = (float *)malloc(sizeof(float) * (n ? n : 1));
        if (!fd || !fc) {
            free(fd);
            free(fc);
            return sofi_fail("out of memory");
        }
        memcpy(fd, data, sizeof(float) * n);
        memcpy(fc, calc, sizeof(float) * n);
        sofi_params q = *p;
        int rc = sofi_time_filter(&q, ntr, fd) || sofi_time_filter(&q, ntr, fc);
        q.time_filt = 0;
        if (!rc) rc = sofi_residual(&q, ntr, fd, fc, diff, l2);
        free(fd);
        free(fc);
        return rc;
    }
    float *is = (float *)malloc(sizeof(float) * (size_t)(ns > 0 ? ns : 1));
    float *isd = (float *)malloc(sizeof(float) * (size_t)(ns > 0 ? ns : 1));
    if (!is || !isd) {
        free(is);
        free(isd);
        return sofi_fail("out of memory");
    }
    const float RMS = 1.0f, wl = p->waterlevel_lnorm8;
    float misfit = 0.0f;
    for (int r = 0; r < ntr; r++) {
        const float *s = calc + (size_t)r * ns, *sd = data + (size_t)r * ns;
        float *df = diff + (size_t)r * ns;
        float acc_s = 0.0f, acc_sd = 0.0f;
        for (int j = 0; j < ns; j++) {
            if (p->velocity == 0) {
                acc_s += s[j];
                is[j] = acc_s * p->dt;
                acc_sd += sd[j];
                isd[j] = acc_sd * p->dt;
            } else {
                is[j] = s[j];
                isd[j] = sd[j];
            }
        }
        float abs_sd = 0.0f, abs_s = 0.0f, mult = 0.0f;
        if (ln == 5 || ln == 7) {
            for (int j = 0; j < ns; j++) {
                abs_sd += isd[j] * isd[j];
                abs_s += is[j] * is[j];
                mult += isd[j] * is[j];
            }
            abs_sd = abs_sd == 0 ? 1.0f : (float)sqrt(abs_sd);
            abs_s = abs_s == 0 ? 1.0f : (float)sqrt(abs_s);
        }
        for (int j = 1; j <= ns; j++) {
            float v = 0.0f;
            const float rr = (sd[j - 1] - s[j - 1]) / RMS;
            switch (ln) {
              case 1: v = (rr > 0 ? 1.0f : rr < 0 ? -1.0f : 0.0f) / RMS; break;
              case 2: v = is[j - 1] - isd[j - 1]; break;
              case 3: v = (rr / (1 + (rr * rr))) / RMS; break;
              case 4: v = (float)tanh(rr) / RMS; break;
              case 5:
              case 7:
                  v = ((is[j - 1] * mult) / (abs_s * abs_s * abs_s * abs_sd)) - (isd[j - 1] / (abs_s * abs_sd));
                  break;
              case 8: v = is[j - 1] / (0.0f + wl) * (0.0f - 0.0f) - 0.0f; break;
            }
            df[ns - j] = v;
            if (ln == 2) misfit += v * v;
            if (ln == 5 || ln == 7) misfit -= (isd[j - 1] * is[j - 1]) / (abs_sd * abs_s);
            if (ln == 8) misfit += (0.0f - 0.0f) * (0.0f - 0.0f);
        }
    }
    free(is);
    free(isd);
    if (l2) *l2 = misfit;
    return 0;
}

static int write_plane(const sofi_params *p, const char *path, const float *a, int nx, int ny)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return sofi_fail("Could not open model output file %s for writing.", path);
    const int pitch_ = sofi_pitch(nx);
    const int idx = p->idx > 0 ? p->idx : 1, idy = p->idy > 0 ? p->idy : 1;
    for (int i = 1; i <= nx; i += idx)
        for (int j = 1; j <= ny; j += idy) {
            float v = AT(a, i, j);
            fwrite(&v, sizeof(float), 1, fp);
        }
    fclose(fp);
    return 0;
}

/* GRADIENT_gradient.{vp,vs,rho}.shot<ishot> (ishot > 0) or the sums over all
 * shots GRADIENT_gradient.{vp,vs,rho} (ishot = 0); planes in the padded
 * layout, written x-major (i outer, j inner) as writemod/mergemod do. */
int sofi_write_gradient(const sofi_params *p, int ishot, const float *gvp, const float *gvs, const float *grho)
{
    static const char *nm[3] = { "vp", "vs", "rho" };
    const float *g[3] = { gvp, gvs, grho };
    for (int k = 0; k < 3; k++) {
        char path[2 * SOFI_STR + 32];
        if (ishot > 0) snprintf(path, sizeof(path), "%s_gradient.%s.shot%i", p->gradient_file, nm[k], ishot);
        else snprintf(path, sizeof(path), "%s_gradient.%s", p->gradient_file, nm[k]);
        if (write_plane(p, path, g[k], p->nxg, p->nyg)) return -1;
    }
    return 0;
}
