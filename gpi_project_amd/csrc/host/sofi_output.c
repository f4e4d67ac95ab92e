/*
 * sofi_output.c - seismogram, source-signal and snapshot files.
 *
 * SU traces follow outseis_glob (src/outseis_glob.c:30-104): a 240-byte SU
 * trace header in native byte order followed by ns float32 samples; the
 * header fields set there are reproduced field by field at their standard
 * SU byte offsets.  ASCII (one "%e" per line) and raw binary are the other
 * two formats.  File names follow saveseis_glob (src/saveseis_glob.c:60-73).
 * Snapshots are written directly in the merged layout snapmerge produces
 * (src/merge.c:85-111): per snapshot, x-major (i outer, j inner) over the
 * decimated window, snapshots appended one after the other.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sofi_host_internal.h"

static void put_i32(unsigned char *h, int off, int32_t v) { memcpy(h + off, &v, 4); }
static void put_i16(unsigned char *h, int off, int16_t v) { memcpy(h + off, &v, 2); }
static void put_u16(unsigned char *h, int off, uint16_t v) { memcpy(h + off, &v, 2); }
static void put_f32(unsigned char *h, int off, float v) { memcpy(h + off, &v, 4); }

int sofi_outseis(const char *path, const float *section, const int *recpos, int ntr,
                 const float *srcpos, int nsrc, int ns, int seis_form, int ishot, float dh,
                 float dt, int ndt, const float *refrec)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return sofi_fail("Could not open %s for writing.", path);
#define RP(row, k) recpos[(size_t)(row) * (size_t)(ntr + 1) + (size_t)(k)]
    switch (seis_form) {
      case 1: {
          const float xshift = 800.0f, yshift = 800.0f;
          float XS = srcpos[(size_t)1 * (nsrc + 1) + ishot];
          float YS = srcpos[(size_t)2 * (nsrc + 1) + ishot];
          for (int k = 1; k <= ntr; k++) {
              unsigned char h[240];
              memset(h, 0, sizeof(h));
              float xr = (RP(1, k) - 1) * dh;
              float yr = (RP(2, k) - 1) * dh;
              float x = xr - refrec[1];
              float y = yr - refrec[2];
              int tracl = RP(3, k);
              put_i32(h, 0, tracl);                                   /* tracl */
              put_i32(h, 4, tracl);                                   /* tracr */
              put_i32(h, 8, ishot);                                   /* fldr */
              put_i32(h, 12, k);                                      /* tracf */
              put_i32(h, 16, ishot);                                  /* ep */
              put_i32(h, 20, tracl);                                  /* cdp */
              put_i16(h, 28, 1);                                      /* trid */
              put_i32(h, 36, sofi_iround(sqrt((XS - xr) * (XS - xr) + (YS - yr) * (YS - yr)))); /* offset */
              put_i32(h, 40, sofi_iround(yr * 1000.0));               /* gelev */
              put_i32(h, 48, sofi_iround(YS * 1000.0));               /* sdepth */
              put_i32(h, 60, sofi_iround(((360.0 / (2.0 * M_PI)) * atan2(x - xshift, y - yshift)) * 1000.0)); /* swdep */
              put_i16(h, 68, -1000);                                  /* scalel */
              put_i16(h, 70, -1000);                                  /* scalco */
              put_i32(h, 72, sofi_iround(XS * 1000.0));               /* sx */
              put_i32(h, 80, sofi_iround(xr * 1000.0));               /* gx */
              put_u16(h, 114, (uint16_t)ns);                          /* ns */
              uint16_t dtu = (uint16_t)sofi_iround((float)ndt * dt * 1.0e6);
              put_u16(h, 116, dtu);                                   /* dt [us] */
              put_f32(h, 180, (float)((float)dtu * 1.0e-6));          /* d1 */
              if (fwrite(h, 1, 240, fp) != 240 || fwrite(section + (size_t)(k - 1) * ns, sizeof(float), ns, fp) != (size_t)ns) {
                  fclose(fp);
                  return sofi_fail("Short write to %s.", path);
              }
          }
          break;
      }
      case 2:
          for (int k = 0; k < ntr; k++)
              for (int s = 0; s < ns; s++) fprintf(fp, "%e\n", section[(size_t)k * ns + s]);
          break;
      case 3:
          for (int k = 0; k < ntr; k++)
              if (fwrite(section + (size_t)k * ns, sizeof(float), ns, fp) != (size_t)ns) {
                  fclose(fp);
                  return sofi_fail("Short write to %s.", path);
              }
          break;
      default:
          fprintf(stderr, "WARNING: Unknown data format for seismograms. No output written!\n");
    }
#undef RP
    fclose(fp);
    return 0;
}

static const char *seis_ext(int f)
{
    switch (f) {
      case 1: return "su";
      case 2: return "txt";
      case 3: return "bin";
      default: return "sgy";
    }
}

static const char *comp_name[SOFI_NCOMP] = { "vx", "vy", "p", "div", "curl" };

int sofi_write_section(const sofi_params *p, const sofi_acq *a, int ishot, int comp, const float *section)
{
    char fn[3 * SOFI_STR];
    if (p->run_multiple_shots)
        snprintf(fn, sizeof(fn), "%s_%s.%s.shot%d", p->seis_file, comp_name[comp], seis_ext(p->seis_format), ishot);
    else
        snprintf(fn, sizeof(fn), "%s_%s.%s", p->seis_file, comp_name[comp], seis_ext(p->seis_format));
    return sofi_outseis(fn, section, a->recpos, a->ntr, a->srcpos, a->nsrc, p->ns, p->seis_format, ishot, p->dh,
                        p->dt, p->ndt, p->refrec);
}

/* Source signal of the first local source (src/initsrc.c:62-91). */
int sofi_write_signal(const sofi_params *p, const sofi_acq *a, const sofi_shot *s, int ishot)
{
    if (!(s->nsrc > 0 && p->sigout == 1)) return 0;
    const char *ext;
    switch (p->sigout_format) {
      case 1: ext = "su"; break;
      case 2: ext = "txt"; break;
      case 3: ext = "bin"; break;
      default: return sofi_fail("Unknown SIGOUT_FORMAT encountered.");
    }
    char fn[3 * SOFI_STR];
    snprintf(fn, sizeof(fn), "%s.shot%d.%s", p->sigout_file, ishot, ext);
    int rp[4 * 2] = { 0, 0, 0, 0, 0, 0, 0, 0 };
    rp[1 * 2 + 1] = s->src_i[0];
    rp[2 * 2 + 1] = s->src_j[0];
    rp[3 * 2 + 1] = 0;
    return sofi_outseis(fn, s->signals, rp, 1, a->srcpos, a->nsrc, p->nt, p->sigout_format, ishot, p->dh, p->dt,
                        p->ndt, p->refrec);
}

int sofi_count_snapshots(const sofi_params *p)
{
    int lsnap = sofi_iround(p->tsnap1 / p->dt), isnap = sofi_iround(p->tsnapinc / p->dt),
        esnap = sofi_iround(p->tsnap2 / p->dt), n = 0;
    for (int nt = 1; nt <= p->nt; nt++)
        if (nt == lsnap && nt <= esnap) {
            ++n;
            lsnap += isnap;
        }
    return n;
}

/* decimated window: i = 1, 1+IDX, ... <= NX; j likewise (src/initproc.c:84-111) */
int sofi_snapshot_dims(const sofi_params *p, int *nxs, int *nys)
{
    *nxs = (p->nxg - 1) / p->idx + 1;
    *nys = (p->nyg - 1) / p->idy + 1;
    return 0;
}

static int append_floats(const char *fn, int first, int format, const float *data, size_t n)
{
    FILE *fp = fopen(fn, first ? "w" : "a");
    if (!fp) return sofi_fail("Could not open snapshot file %s.", fn);
    if (format == 2) {
        for (size_t k = 0; k < n; ++k) fprintf(fp, "%e\n", data[k]);
    } else if (n > 0 && fwrite(data, sizeof(float), n, fp) != n) {
        fclose(fp);
        return sofi_fail("Short write to %s.", fn);
    }
    fclose(fp);
    return 0;
}

/* The decimated snapshot points of one process's subgrid along one axis, as
 * zero-based indices into the global list 1, 1+id, 1+2id, ...: the points
 * of [g0, g1] (src/initproc.c:86-114).  Returns the count (0: none). */
static int rank_points(int ng, int id, int g0, int g1, int *first)
{
    int lo = -1, hi = -2;
    for (int i = 1, k = 0; i <= ng; i += id, ++k)
        if (i >= g0 && i <= g1) {
            if (lo < 0) lo = k;
            hi = k;
        }
    *first = lo;
    return hi - lo + 1 > 0 ? hi - lo + 1 : 0;
}

/* Append one snapshot of component comp (n floats, x-major). first != 0
 * truncates the files (the reference opens with "w" for its first snapshot).
 * Written twice: the merged file snapmerge produces (src/merge.c:85-111),
 * and the per-process files the reference's NPROCX x NPROCY processes write
 * themselves, SNAP_FILE<ext>.<comp>.<px>.<py> (src/snap_ssg.c:62-67), each
 * holding the points of its subgrid x-major, so that the reference's own
 * snapmerge applies to them.  The per-process files are skipped when the
 * subgrids do not divide the grid (the reference refuses such runs,
 * src/initproc.c:41-47). */
int sofi_write_snapshot(const sofi_params *p, int comp, int first, const float *data, int n)
{
    char fn[3 * SOFI_STR + 32];
    const char *ext = p->snap_format == 2 ? ".asc" : ".bin";
    snprintf(fn, sizeof(fn), "%s%s.%s", p->snap_file, ext, comp_name[comp]);
    if (append_floats(fn, first, p->snap_format, data, (size_t)n)) return -1;
    const int npx = p->nprocx, npy = p->nprocy;
    if (npx < 1 || npy < 1 || p->nxg % npx || p->nyg % npy) return 0;
    int nxs, nys;
    sofi_snapshot_dims(p, &nxs, &nys);
    if (n != nxs * nys) return sofi_fail("snapshot of %d values, expected %d x %d", n, nxs, nys);
    const int lx = p->nxg / npx, ly = p->nyg / npy;
    float *buf = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    if (!buf) return sofi_fail("out of memory");
    for (int px = 0; px < npx; ++px)
        for (int py = 0; py < npy; ++py) {
            int i0, j0;
            const int ni = rank_points(p->nxg, p->idx, 1 + px * lx, (px + 1) * lx, &i0);
            const int nj = rank_points(p->nyg, p->idy, 1 + py * ly, (py + 1) * ly, &j0);
            size_t m = 0;
            for (int a = 0; a < ni; ++a)
                for (int b = 0; b < nj; ++b) buf[m++] = data[(size_t)(i0 + a) * nys + (j0 + b)];
            snprintf(fn, sizeof(fn), "%s%s.%s.%d.%d", p->snap_file, ext, comp_name[comp], px, py);
            if (append_floats(fn, first, p->snap_format, buf, m)) {
                free(buf);
                return -1;
            }
        }
    free(buf);
    return 0;
}
