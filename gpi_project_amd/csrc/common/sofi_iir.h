/*
 * sofi_iir.h - the Butterworth cascade of the TIME_FILT filter, shared by
 * the host library (sofi_fwi.c, which filters on the CPU) and the device
 * library (sofi_kernels.hip, which filters the resident sections).
 *
 * timedomain_filt (src/timedomain_filt.c:28-113) filters with libseife's
 * seife_butterworth (contrib/Seitosh/libseife/cseife.c:160-190): an odd
 * order starts with a first-order section, then m/2 second-order sections
 * with damping sin(pi/2 (2j - 1) / m); each section is a bilinear-transform
 * recursion y[k] = f0 x[k] + f1 x[k-1] + f2 x[k-2] + g1 y[k-1] + g2 y[k-2]
 * in double (libseife's rfk coefficients for its lp1 / hp1 / lp2 / hp2
 * types, the recursion of rekfl).  A cascade of causal sections applied one
 * after the other over a trace gives the same numbers as passing every
 * sample through all sections in turn.
 */
#ifndef SOFI_IIR_H
#define SOFI_IIR_H

#include <math.h>

#define SOFI_IIR_MAXSEC 9   /* ORDER <= 17 */

typedef struct { double f0, f1, f2, g1, g2; } sofi_iir2;

/* section of period t0 (in samples): first order (h < 0) or second order
 * with damping h; low- or high-pass */
static inline sofi_iir2 sofi_iir_section(double t0, double h, int highpass)
{
    sofi_iir2 c = { 0., 0., 0., 0., 0. };
    const double eps = (8. * atan(1.)) / t0;
    if (h < 0) {
        c.g1 = (2. - eps) / (2. + eps);
        if (highpass) {
            c.f0 = 2. / (2. + eps);
            c.f1 = -(c.f0);
        } else {
            c.f0 = eps / (2. + eps);
            c.f1 = c.f0;
        }
        return c;
    }
    const double e2 = eps * eps;
    const double a = 1. - eps * h + e2 / 4., b = -2. + e2 / 2., cc = 1. + eps * h + e2 / 4.;
    c.g1 = -b / cc;
    c.g2 = -a / cc;
    if (highpass) {
        c.f0 = 1. / cc;
        c.f1 = -(c.f0) - c.f0;
    } else {
        c.f0 = e2 / 4. / cc;
        c.f1 = c.f0 + c.f0;
    }
    c.f2 = c.f0;
    return c;
}

/* the sections of a Butterworth filter of order m, corner period `period`
 * [s], sampling interval dt [s], in cascade order; returns their number */
static inline int sofi_iir_butterworth(double dt, double period, int m, int highpass, sofi_iir2 *sec)
{
    int n = 0;
    if (m % 2) sec[n++] = sofi_iir_section(period / dt, -1., highpass);
    const double pih = 2. * atan(1.);
    for (int j = 1; j <= m / 2; ++j) sec[n++] = sofi_iir_section(period / dt, sin(pih * (2 * j - 1) / m), highpass);
    return n;
}

#endif /* SOFI_IIR_H */
