/*
 * sofi_fwi.c - host side of the FWI gradient path: the measured data, the
 * adjoint sources (residuals) and the gradient files.
 *
 *   sofi_read_data      inseis      src/inseis.c:28-80
 *   sofi_residual       calc_res    src/calc_res.c:26-324 (LNORM 2)
 *   sofi_write_gradient output_gradient + writemod
 *                                   src/output_gradient.c:25-68, src/writemod.c:28-45
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <string.h>

#include "sofi_host_internal.h"

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

/* L2 residual of calc_res: with VELOCITY = 0 both sections are integrated
 * in time (running sum times DT), then diff = synthetic - measured, stored
 * time-reversed (diff[i][NS - j + 1]).  *l2 receives the sum of diff^2 (the
 * test-shot misfit of the reference, accumulated in float). */
int sofi_residual(const sofi_params *p, int ntr, const float *data, const float *calc, float *diff, double *l2)
{
    if (p->lnorm != 2) return sofi_fail("LNORM=%d is not supported by this build (L2 only).", p->lnorm);
    const int ns = p->ns;
    float misfit = 0.0f;
    for (int r = 0; r < ntr; r++) {
        const float *s = calc + (size_t)r * ns, *sd = data + (size_t)r * ns;
        float *df = diff + (size_t)r * ns;
        float intseis_s = 0.0f, intseis_sd = 0.0f;
        for (int j = 1; j <= ns; j++) {
            float is, isd;
            if (p->velocity == 0) {
                intseis_s += s[j - 1];
                is = intseis_s * p->dt;
                intseis_sd += sd[j - 1];
                isd = intseis_sd * p->dt;
            } else {
                is = s[j - 1];
                isd = sd[j - 1];
            }
            const int invtime = ns - j + 1;
            df[invtime - 1] = is - isd;
            misfit += df[invtime - 1] * df[invtime - 1];
        }
    }
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
