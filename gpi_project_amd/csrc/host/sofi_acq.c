/*
 * sofi_acq.c - receivers, sources, source wavelets and per-shot setup.
 *
 *   receiver()  src/receiver.c:28-190  (file, receiver array, or a straight line)
 *   sources()   src/sources.c:28-300   (SOURCE_FILE lines "x y td fc amp [azimuth] [type]"; plane wave)
 *   splitsrc()  src/splitsrc.c:27-81   (grid index = floor(x/DH)+1)
 *   wavelet()   src/wavelet.c:33-149   (Ricker, Fuchs-Mueller, file, sin^3, Berlage, Klauder)
 *   read_srcsig src/read_srcsig.c:38-185 (SIGNAL_FILE: text, one sample per line, or SU)
 *   initsrc()   src/initsrc.c:30-101
 *   acq_read()  src/acq_read.c:25-80   (scan_topo, src/scan_topo.c:32-85, for SOURCE_TOPO / REC_TOPO)
 * One device holds the whole grid, so the "local" sources and receivers are
 * all of them (GGRID = 1..NX x 1..NY).
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "sofi_host_internal.h"

#define SRC(row, l) (a->srcpos[(size_t)(row) * (size_t)(a->nsrc + 1) + (size_t)(l)])
#define REC(row, k) (a->recpos[(size_t)(row) * (size_t)(a->ntr + 1) + (size_t)(k)])

/* scan_topo (src/scan_topo.c:32-85): for every x position the first grid
 * row whose P velocity differs from the top cell's (1 when the trace is
 * uniform), from MFILE.vp.su (traces of NY samples after a 240-byte header)
 * or MFILE.vp.  topo[1..NX]; NULL on error. */
static int *scan_topo(const sofi_params *p)
{
    char fn[SOFI_STR + 16];
    snprintf(fn, sizeof(fn), "%s.vp.su", p->mfile);
    int su = access(fn, R_OK) == 0;
    if (!su) {
        snprintf(fn, sizeof(fn), "%s.vp", p->mfile);
        if (access(fn, R_OK) != 0) { sofi_fail("Could not open P-wave velocity model to scan topography."); return NULL; }
    }
    FILE *fp = fopen(fn, "rb");
    if (!fp) { sofi_fail("Could not open P-wave velocity model %s.", fn); return NULL; }
    int *topo = (int *)calloc((size_t)p->nxg + 1, sizeof(int));
    float *tr = (float *)malloc(sizeof(float) * (size_t)p->nyg);
    for (int i = 1; i <= p->nxg; i++) {
        if ((su && fseek(fp, 240, SEEK_CUR) != 0) || fread(tr, sizeof(float), (size_t)p->nyg, fp) != (size_t)p->nyg) {
            fclose(fp);
            free(tr);
            free(topo);
            sofi_fail("Short read in %s while scanning the topography.", fn);
            return NULL;
        }
        topo[i] = 1;
        for (int j = 2; j <= p->nyg; j++)
            if (tr[j - 1] != tr[0]) {
                topo[i] = j;
                break;
            }
    }
    free(tr);
    fclose(fp);
    return topo;
}

static int topo_at(const sofi_params *p, const int *topo, int ix, int *out)
{
    if (ix < 1 || ix > p->nxg) return sofi_fail("Topography: grid column %d is outside 1..%d.", ix, p->nxg);
    *out = topo[ix];
    return 0;
}

static int read_receivers(sofi_params *p, sofi_acq *a, const int *topo)
{
    if (p->readrec) {
        char buf[SOFI_STR], tok[SOFI_STR];
        FILE *fp = fopen(p->rec_file, "r");
        if (!fp) return sofi_fail("Receiver file %s could not be opened.", p->rec_file);
        int n = 0;
        while (fgets(buf, SOFI_STR, fp)) {
            int comment = strchr(buf, '#') || strchr(buf, '%');
            if (!comment && sscanf(buf, "%255s", tok) == 1) ++n;
        }
        rewind(fp);
        int *tmp = (int *)calloc((size_t)4 * (n + 1), sizeof(int));
        for (int itr = 1; itr <= n; itr++) {
            float xrec = 0, yrec = 0;
            if (fscanf(fp, "%f%f\n", &xrec, &yrec) != 2) { fclose(fp); free(tmp); return sofi_fail("Receiver file %s: bad line %d.", p->rec_file, itr); }
            tmp[1 * (n + 1) + itr] = (int)floor((xrec + p->refrec[1]) / p->dh) + 1;
            if (p->rec_topo) {   /* below the topography: no REFREC offset, no +1 (src/receiver.c:70-72) */
                int t = 1;
                if (topo_at(p, topo, tmp[1 * (n + 1) + itr], &t)) { fclose(fp); free(tmp); return -1; }
                tmp[2 * (n + 1) + itr] = t + (int)floor(yrec / p->dh);
            } else {
                tmp[2 * (n + 1) + itr] = (int)floor((yrec + p->refrec[2]) / p->dh) + 1;
            }
            tmp[3 * (n + 1) + itr] = itr;
        }
        fclose(fp);
        /* drop co-located receivers (src/receiver.c:78-92) */
        int flag = 0;
        for (int i = 1; i <= n - 1; i++)
            for (int k = i + 1; k <= n; k++)
                if (tmp[1 * (n + 1) + i] == tmp[1 * (n + 1) + k] && tmp[2 * (n + 1) + i] == tmp[2 * (n + 1) + k])
                    tmp[1 * (n + 1) + k] = -(++flag);
        a->ntr = n - flag;
        a->recpos = (int *)calloc((size_t)4 * (a->ntr + 1), sizeof(int));
        int k2 = 0;
        for (int i = 1; i <= n; i++)
            if (tmp[1 * (n + 1) + i] > 0) {
                ++k2;
                REC(1, k2) = tmp[1 * (n + 1) + i];
                REC(2, k2) = tmp[2 * (n + 1) + i];
                REC(3, k2) = tmp[3 * (n + 1) + i];
            }
        free(tmp);
    } else if (p->rec_array > 0) {
        a->ntr = (1 + (p->nxg - 2 * p->fw) / p->drx) * p->rec_array;
        a->recpos = (int *)calloc((size_t)4 * (a->ntr + 1), sizeof(int));
        int itr = 0;
        for (int n = 0; n <= p->rec_array - 1; n++) {
            int j = sofi_iround((p->rec_array_depth + p->rec_array_dist * (float)n) / p->dh);
            for (int i = p->fw; i <= p->nxg - p->fw; i += p->drx) {
                itr++;
                REC(1, itr) = i;
                REC(2, itr) = j;
                REC(3, itr) = itr;
            }
        }
    } else {
        /* straight line; the reference keeps these in float (src/receiver.c:33-35) */
        float nxrec1 = (int)floor(p->xrec1 / p->dh) + 1;
        float nyrec1 = (int)floor(p->yrec1 / p->dh) + 1;
        float nxrec2 = (int)floor(p->xrec2 / p->dh) + 1;
        float nyrec2 = (int)floor(p->yrec2 / p->dh) + 1;
        if ((nxrec2 - nxrec1) == 0 && (nyrec2 - nyrec1) == 0) {
            a->ntr = 1;
            a->recpos = (int *)calloc(4 * 2, sizeof(int));
            REC(1, 1) = (int)nxrec1;
            REC(2, 1) = (int)nyrec1;
            REC(3, 1) = 1;
        } else if ((nyrec2 - nyrec1) == 0) {
            a->ntr = (int)floor((nxrec2 - nxrec1) / p->ngeoph) + 1;
            a->recpos = (int *)calloc((size_t)4 * (a->ntr + 1), sizeof(int));
            float nxrec = nxrec1;
            for (int n = 1; n <= a->ntr; n++) {
                float nyrec = nyrec1 + ((nyrec2 - nyrec1) / (nxrec2 - nxrec1) * (nxrec - nxrec1));
                int itr = (int)floor((nxrec - nxrec1) / p->ngeoph) + 1;
                REC(1, itr) = (int)nxrec;
                REC(2, itr) = (int)nyrec;
                REC(3, itr) = itr;
                nxrec += p->ngeoph;
            }
        } else if ((nxrec2 - nxrec1) == 0) {
            a->ntr = (int)floor((nyrec2 - nyrec1) / p->ngeoph) + 1;
            a->recpos = (int *)calloc((size_t)4 * (a->ntr + 1), sizeof(int));
            float nyrec = nyrec1;
            for (int n = 1; n <= a->ntr; n++) {
                float nxrec = nxrec1 + ((nxrec2 - nxrec1) / (nyrec2 - nyrec1) * (nyrec - nyrec1));
                int itr = (int)floor((nyrec - nyrec1) / p->ngeoph) + 1;
                REC(1, itr) = (int)nxrec;
                REC(2, itr) = (int)nyrec;
                REC(3, itr) = itr;
                nyrec += p->ngeoph;
            }
        } else {
            return sofi_fail("No horizontal or vertical receiver line is specified in the json file (use READREC=1 for an arbitrary line).");
        }
    }
    for (int itr = 1; itr <= a->ntr; itr++) {
        if (REC(1, itr) == 0) REC(1, itr) = 1;
        if (REC(2, itr) == 0) REC(2, itr) = 1;
    }
    return 0;
}

static int read_sources(sofi_params *p, sofi_acq *a, const int *topo)
{
    if (p->srcrec == 1) {
        char buf[SOFI_STR], tok[SOFI_STR];
        FILE *fp = fopen(p->source_file, "r");
        if (!fp) return sofi_fail("Source file %s could not be opened.", p->source_file);
        a->nsrc = 0;
        while (fgets(buf, SOFI_STR, fp))
            if (!strchr(buf, '#') && sscanf(buf, "%255s", tok) == 1) ++a->nsrc;
        rewind(fp);
        a->srcpos = (float *)calloc((size_t)(SOFI_NSPAR + 1) * (a->nsrc + 1), sizeof(float));
        float fc = 0.0f;
        int l = 0;
        while (fgets(buf, 255, fp)) {
            if (strchr(buf, '#') || sscanf(buf, "%255s", tok) != 1) continue;
            ++l;
            float xsrc = 0, ysrc = 0, tshift = 0;
            float v[16] = { 0 };
            int nvarin;
            if (p->source_shape <= 4) {
                nvarin = sscanf(buf, "%f%f%f%f%f%f%f", &xsrc, &ysrc, &tshift, &v[5], &v[6], &v[7], &v[8]);
                if (nvarin < 5) { fclose(fp); return sofi_fail("Missing parameter in SOURCE_FILE (source %d needs x y td fc amp).", l); }
                if (nvarin == 5) v[7] = 0.0f;
                if (nvarin <= 6) v[8] = (float)p->source_type;
            } else if (p->source_shape == 5) {
                /* Berlage: x y td fc amp n alpha phi0 [azimuth] [type] (src/sources.c:142-200) */
                nvarin = sscanf(buf, "%f%f%f%f%f%f%f%f%f%f", &xsrc, &ysrc, &tshift, &v[5], &v[6], &v[9], &v[10], &v[11], &v[7], &v[8]);
                if (nvarin < 8) { fclose(fp); return sofi_fail("Missing parameter in SOURCE_FILE for the Berlage wavelet (source %d).", l); }
                if (nvarin == 8) v[7] = 0.0f;
                if (nvarin <= 9) v[8] = (float)p->source_type;
            } else {
                /* Klauder: x y td fmin fmax amp tsweep width [azimuth] [type]; fc = (fmin+fmax)/2 (src/sources.c:247-310) */
                nvarin = sscanf(buf, "%f%f%f%f%f%f%f%f%f%f", &xsrc, &ysrc, &tshift, &v[9], &v[10], &v[6], &v[11], &v[12], &v[7], &v[8]);
                if (nvarin < 8) { fclose(fp); return sofi_fail("Missing parameter in SOURCE_FILE for the Klauder wavelet (source %d).", l); }
                if (nvarin == 8) v[7] = 0.0f;
                if (nvarin <= 9) v[8] = (float)p->source_type;
                v[5] = (v[9] + v[10]) / 2;
            }
            SRC(1, l) = xsrc;
            SRC(2, l) = ysrc;
            if (p->source_topo) {   /* depth below the topography (src/sources.c:122-126) */
                int t = 1;
                if (topo_at(p, topo, (int)floor(xsrc / p->dh) + 1, &t)) { fclose(fp); return -1; }
                SRC(2, l) = (t - 1) * p->dh + ysrc;
            }
            SRC(3, l) = 0.0f;
            SRC(4, l) = tshift;
            for (int r = 5; r <= SOFI_NSPAR; ++r) SRC(r, l) = v[r];
            fc = SRC(5, l);
        }
        fclose(fp);
        /* TS = 1 / (largest centre frequency) (src/sources.c:314-318) */
        for (l = 1; l <= a->nsrc; l++)
            if (SRC(5, l) > fc) fc = SRC(5, l);
        p->ts = (float)(1.0 / fc);
    } else if (p->srcrec == 2) {
        if (p->plane_wave_depth <= 0) return sofi_fail("SRCREC parameter specifies PLANE_WAVE excitation, but PLANE_WAVE_DEPTH<=0!");
        if (p->source_shape > 4) return sofi_fail("Plane wave is only implemented for Ricker, Fuchs-Mueller, sin^3 or an external wavelet!");
        float tan_phi = (float)tan(p->plane_wave_angle * M_PI / 180.0);
        float dz = (float)(p->nxg - 1) * p->dh * tan_phi;
        if (!((p->plane_wave_depth + dz) <= (p->nyg - 1) * p->dh)) return sofi_fail("Maximum depth of plane wave exceeds model depth.");
        a->nsrc = p->nxg;
        a->srcpos = (float *)calloc((size_t)(SOFI_NSPAR + 1) * (a->nsrc + 1), sizeof(float));
        for (int i = 1; i <= p->nxg; i++) {
            float x = (i - 1) * p->dh;
            SRC(1, i) = x;
            SRC(2, i) = p->plane_wave_depth + (tan_phi * x);
            SRC(4, i) = 0.0f;
            SRC(5, i) = (float)(1.0 / p->ts);
            SRC(6, i) = 1.0f;
            SRC(7, i) = 0.0f;
            SRC(8, i) = (float)p->source_type;
        }
    } else {
        return sofi_fail("SRCREC parameter is invalid (SRCREC!=1 or SRCREC!=2)! No source parameters specified!");
    }
    return 0;
}

int sofi_read_acq(sofi_params *p, sofi_acq *a)
{
    memset(a, 0, sizeof(*a));
    int *topo = NULL;
    if (p->source_topo || p->rec_topo) {   /* acq_read (src/acq_read.c:33-36) */
        topo = scan_topo(p);
        if (!topo) return -1;
    }
    if (p->seismo) {
        int rc = read_receivers(p, a, topo);
        if (rc) { free(topo); return rc; }
    }
    int rc = read_sources(p, a, topo);
    free(topo);
    if (rc) return rc;
    a->nshots = p->run_multiple_shots ? a->nsrc : 1;
    if (p->snapshot_end == -9999 || p->snapshot_end > a->nshots) p->snapshot_end = a->nshots;
    return 0;
}

void sofi_free_acq(sofi_acq *a)
{
    free(a->srcpos);
    free(a->recpos);
    memset(a, 0, sizeof(*a));
}

/* SIGNAL_FILE (src/read_srcsig.c:38-185).  The type follows the name: a
 * name containing ".su" (any case) is an SU file whose first trace is the
 * signal, with its header's dt (microseconds) as the sampling interval and
 * delrt required to be 0; anything else is text, one sample per line (the
 * first number of the line; a line that holds none counts as a 0 sample),
 * blank lines skipped, lines starting with '#' comments, of which one of the
 * form "# DT = <value>" gives the sampling interval.  A sampling interval
 * that differs from DT is an error; none is accepted. */
static float *read_srcsig(const sofi_params *p, int *nts)
{
    *nts = 0;
    float dt = 0.0f;
    int dt_found = 0;
    float *sig = NULL;
    if (strcasestr(p->signal_file, ".su")) {
        FILE *fp = fopen(p->signal_file, "rb");
        if (!fp) { sofi_fail("Could not open SU file %s for reading.", p->signal_file); return NULL; }
        unsigned char h[240];
        if (fread(h, 1, 240, fp) != 240) {
            fclose(fp);
            sofi_fail("Short read in SU file %s.", p->signal_file);
            return NULL;
        }
        uint16_t ns, sudt;
        int16_t delrt;
        memcpy(&delrt, h + 108, 2);
        memcpy(&ns, h + 114, 2);
        memcpy(&sudt, h + 116, 2);
        if (ns < 1) {
            fclose(fp);
            sofi_fail("SU file has no number of samples (ns) in first SU header.");
            return NULL;
        }
        sig = (float *)calloc(ns, sizeof(float));
        if (fread(sig, sizeof(float), ns, fp) != ns) {
            fclose(fp);
            free(sig);
            sofi_fail("Short read in SU file while reading data.");
            return NULL;
        }
        fclose(fp);
        if (delrt != 0) {
            free(sig);
            sofi_fail("Delay recording time != 0 not supported for source signature.");
            return NULL;
        }
        if (sudt > 0) {
            dt = (float)(sudt * 1e-6);
            dt_found = 1;
        }
        *nts = ns;
    } else {
        FILE *fp = fopen(p->signal_file, "r");
        if (!fp) { sofi_fail("Could not open file %s for reading.", p->signal_file); return NULL; }
        char line[256];
        int n = 0;
        while (fgets(line, sizeof(line), fp)) {
            char *c = line;
            while (isspace((unsigned char)*c)) ++c;
            if (*c == '\0') continue;
            if (*c == '#') {
                ++c;
                const char *q = strcasestr(c, "DT");
                if (!q || !(q = strcasestr(c, "="))) continue;
                ++q;
                while (isspace((unsigned char)*q)) ++q;
                float v;
                if (sscanf(q, "%e", &v) == 1) {
                    dt = v;
                    dt_found = 1;
                }
                continue;
            }
            ++n;
        }
        if (n == 0) {
            fclose(fp);
            sofi_fail("Source signature file did not contain valid samples.");
            return NULL;
        }
        sig = (float *)calloc(n, sizeof(float));
        rewind(fp);
        int k = 0;
        while (k < n && fgets(line, sizeof(line), fp)) {
            char *c = line;
            while (isspace((unsigned char)*c)) ++c;
            if (*c == '\0' || *c == '#') continue;
            if (sscanf(c, "%e", &sig[k]) != 1) sig[k] = 0.0f;
            ++k;
        }
        fclose(fp);
        *nts = n;
    }
    if (dt_found && dt != p->dt) {
        free(sig);
        sofi_fail("Sampling interval of source signature does not match FD step size! (dt source signature: %fs, "
                  "FD step size: %fs)", dt, p->dt);
        return NULL;
    }
    return sig;
}

int sofi_init_shot(const sofi_params *p, const sofi_acq *a, int ishot, int prev_source_type, sofi_shot *s)
{
    memset(s, 0, sizeof(*s));
    int first = p->run_multiple_shots ? ishot : 1;
    int last = p->run_multiple_shots ? ishot : a->nsrc;
    int n = 0;
    for (int l = first; l <= last; ++l) {
        int ia = (int)floor(SRC(1, l) / p->dh) + 1, ja = (int)floor(SRC(2, l) / p->dh) + 1;
        if (ia >= 1 && ia <= p->nxg && ja >= 1 && ja <= p->nyg) ++n;
    }
    s->nsrc = n;
    s->src_i = (int *)calloc(n + 1, sizeof(int));
    s->src_j = (int *)calloc(n + 1, sizeof(int));
    s->src_type = (int *)calloc(n + 1, sizeof(int));
    s->src_azi = (float *)calloc(n + 1, sizeof(float));
    s->src_amp = (float *)calloc(n + 1, sizeof(float));
    s->signals = (float *)calloc((size_t)(n > 0 ? n : 1) * p->nt, sizeof(float));
    float *sig_file = NULL;
    int nts = 0;
    if (p->source_shape == 3 && n > 0) {
        sig_file = read_srcsig(p, &nts);
        if (!sig_file) return -1;
    }
    int k = 0;
    for (int l = first; l <= last; ++l) {
        int ia = (int)floor(SRC(1, l) / p->dh) + 1, ja = (int)floor(SRC(2, l) / p->dh) + 1;
        if (!(ia >= 1 && ia <= p->nxg && ja >= 1 && ja <= p->nyg)) continue;
        s->src_i[k] = ia;
        s->src_j[k] = ja;
        s->src_type[k] = (int)SRC(8, l);
        s->src_azi[k] = SRC(7, l);
        s->src_amp[k] = SRC(6, l);
        /* wavelet() for this source */
        float tshift = SRC(4, l), fc = SRC(5, l), amp_a = SRC(6, l);
        float ts = (float)(1.0 / fc);
        float sigmax = 0.0f;
        float *sig = s->signals + (size_t)k * p->nt;
        for (int nt = 1; nt <= p->nt; nt++) {
            float t = (float)nt * p->dt;
            float amp = 0.0f, tau;
            switch (p->source_shape) {
              case 1:
                  tau = (float)(M_PI * (t - 1.5 * ts - tshift) / (ts));
                  amp = (float)(((1.0 - 2.0 * tau * tau) * exp(-tau * tau)));
                  break;
              case 2:
                  if ((t < tshift) || (t > (tshift + ts))) amp = 0.0f;
                  else amp = (float)(sin(2.0 * M_PI * (t - tshift) * fc) - 0.5 * sin(4.0 * M_PI * (t - tshift) * fc));
                  break;
              case 3:
                  amp = (nt <= nts) ? sig_file[nt - 1] : 0.0f;
                  break;
              case 4:
                  if ((t < tshift) || (t > (tshift + ts))) amp = 0.0f;
                  else amp = (float)pow(sin(M_PI * (t - tshift) / ts), 3.0);
                  break;
              case 5: {
                  float nexp = SRC(9, l), alpha = SRC(10, l), phi0 = (float)(SRC(11, l) * M_PI / 180);
                  if (nexp <= 0) { free(sig_file); return sofi_fail("No valid time exponent for Berlage wavelet (N>0) specified!"); }
                  if (t < tshift) amp = 0.0f;
                  else amp = (float)(amp_a * pow((t - tshift), nexp) * exp(-alpha * (t - tshift)) * cos(2 * M_PI * fc * (t - tshift) + phi0));
                  break;
              }
              case 6: {
                  float fmin = SRC(9, l), fmax = SRC(10, l), T = SRC(11, l), width = SRC(12, l);
                  float f2 = (fmax - fmin) / T;
                  tau = (width / fc);
                  if ((t < tshift) || (t > (tau * 2) + tshift)) {
                      amp = 0.0f;
                  } else {
                      /* The reference writes the carrier as exp(2 PI I fc x) with
                       * <complex.h> but without <tgmath.h>, so the real exp()
                       * receives the argument's real part, 0, and the carrier is
                       * 1: its Klauder wavelet is the sweep's sinc envelope alone
                       * (src/wavelet.c:109-123), kept here as it computes it. */
                      float x = t - tau - tshift;   /* T - x in float, as there */
                      amp = (float)(sin(M_PI * f2 * x * (T - x)) / (M_PI * f2 * x));
                      if (t < (tau * 0.2) + tshift) {
                          float taper = (float)((1 - cos(((t - tshift) / tau / 0.2) * M_PI)) / 2);
                          amp = amp * taper;
                      } else if (t > (tau * 1.8) + tshift) {
                          float taper = (float)((1 - cos(((2 * tau + tshift - t) / 0.2 / tau) * M_PI)) / 2);
                          amp = amp * taper;
                      }
                  }
                  break;
              }
              default:
                  free(sig_file);
                  return sofi_fail("No valid source wavelet (SOURCE_SHAPE) specified!");
            }
            if (fabs(amp) > sigmax) sigmax = (float)fabs(amp);
            sig[nt - 1] = amp * amp_a;
        }
        if (0.0 == sigmax) { free(sig_file); return sofi_fail("Source signal contains only zeros!"); }
        if (p->source_shape == 5 || p->source_shape == 6)
            for (int nt = 1; nt <= p->nt; nt++) sig[nt - 1] = sig[nt - 1] / sigmax;
        ++k;
    }
    free(sig_file);
    /* update_v_interior() overwrites SOURCE_TYPE with each local source's
     * type on every forward step (src/update_v_interior.c:53), so from the
     * first step on psource/eqsource are selected by the last source's type. */
    s->source_type = n > 0 ? s->src_type[n - 1] : prev_source_type;
    s->snapcheck = (p->snap > 0 && ((ishot - p->snapshot_start) % p->snapshot_incr) == 0 && ishot <= p->snapshot_start) ? 1 : 0;
    return 0;
}

void sofi_free_shot(sofi_shot *s)
{
    free(s->src_i); free(s->src_j); free(s->src_type); free(s->src_azi); free(s->src_amp); free(s->signals);
    memset(s, 0, sizeof(*s));
}
