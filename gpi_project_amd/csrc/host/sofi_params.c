/*
 * sofi_params.c - reader for the reference's line-based "JSON" parameter file.
 *
 * Format (src/json_parser.c:30-190): one or more "name" : "value" pairs per
 * line; lines containing "comment"/"Comment" are skipped; a line with k
 * names separated by commas carries k comma-separated values (1..5 pairs);
 * names and values are cut at the first blank.  Keys are looked up exactly
 * as src/json_parser.c:221-330 does (first match wins, integers must have no
 * fractional part, trailing garbage is an error).  Which keys are required,
 * their defaults and the derived values (NT, NS, ND) follow
 * src/read_par_json.c:108-620.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <unistd.h>

#include "sofi_host.h"
#include "sofi_host_internal.h"

static _Thread_local char g_err[1024];   /* per thread: the library is used from several threads */

const char *sofi_host_last_error(void) { return g_err; }

int sofi_fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return -1;
}

/* ---------------------------------------------------------------- objects */

#define MAXOBJ 1024

typedef struct {
    int n;
    char name[MAXOBJ][SOFI_STR];
    char value[MAXOBJ][SOFI_STR];
    int used[MAXOBJ];
} objlist;

static int count_char(const char *s, char c)
{
    int n = 0;
    for (; *s; ++s) n += (*s == c);
    return n;
}

/* keep the first blank-delimited token (remove_blankspaces_around_string) */
static void first_token(const char *in, char *out)
{
    out[0] = '\0';
    sscanf(in, "%255s", out);
}

static int add_obj(objlist *o, const char *name, const char *value)
{
    if (o->n >= MAXOBJ) return sofi_fail("too many objects in parameter file");
    first_token(name, o->name[o->n]);
    first_token(value, o->value[o->n]);
    o->used[o->n] = 0;
    o->n++;
    return 0;
}

/* split "a,b,c" into k fields */
static int split_commas(const char *s, int k, char out[][SOFI_STR])
{
    int f = 0, c = 0;
    for (; *s && f < k; ++s) {
        if (*s == ',' && f < k - 1) {
            out[f][c] = '\0';
            f++;
            c = 0;
        } else if (c < SOFI_STR - 1) {
            out[f][c++] = *s;
        }
    }
    out[f][c] = '\0';
    return f + 1;
}

static int read_objects(const char *path, objlist *o)
{
    char line[SOFI_STR];
    int lineno = 0;
    FILE *fp = fopen(path, "r");
    if (!fp) return sofi_fail("Could not open file %s for reading.", path);
    o->n = 0;
    while (fgets(line, SOFI_STR, fp)) {
        lineno++;
        if (!(strchr(line, ':') && strchr(line, '"'))) continue;
        if (strstr(line, "comment") || strstr(line, "Comment")) continue;
        int nq = count_char(line, '"');
        int nc = count_char(line, ',');
        if (nq != 4) {
            fclose(fp);
            return sofi_fail("Error in input file, line %i, only 4 (two pairs) of double quotes are allowed per line, but found %i!", lineno, nq);
        }
        /* isolate the two quoted strings */
        char q[4][SOFI_STR];
        const char *p = line;
        for (int k = 0; k < 2; ++k) {
            const char *a = strchr(p, '"');
            const char *b = a ? strchr(a + 1, '"') : NULL;
            if (!a || !b) {
                fclose(fp);
                return sofi_fail("Error in input file, line %i, cannot read object name and object value!", lineno);
            }
            size_t len = (size_t)(b - a - 1);
            if (len >= SOFI_STR) len = SOFI_STR - 1;
            memcpy(q[k], a + 1, len);
            q[k][len] = '\0';
            p = b + 1;
        }
        int npairs;
        switch (nc) {
          case 0: case 1: npairs = 1; break;
          case 3: npairs = 2; break;
          case 5: npairs = 3; break;
          case 7: npairs = 4; break;
          case 9: npairs = 5; break;
          default:
              fclose(fp);
              return sofi_fail("Error in input file, line %i, only 0, 1, 3, 5, 7 or 9 commas are allowed per line, but found %i!", lineno, nc);
        }
        if (npairs == 1) {
            if (add_obj(o, q[0], q[1])) { fclose(fp); return -1; }
        } else {
            char nm[5][SOFI_STR], vl[5][SOFI_STR];
            if (split_commas(q[0], npairs, nm) != npairs || split_commas(q[1], npairs, vl) != npairs) {
                fclose(fp);
                return sofi_fail("Error in input file, line %i, cannot read %d object names and values!", lineno, npairs);
            }
            for (int k = 0; k < npairs; ++k)
                if (add_obj(o, nm[k], vl[k])) { fclose(fp); return -1; }
        }
    }
    fclose(fp);
    return 0;
}

/* returns 0 when found (as the reference's get_*_from_objectlist), 1 when absent, -1 on error */
static int find(objlist *o, const char *key)
{
    for (int i = 0; i < o->n; ++i)
        if (strcmp(o->name[i], key) == 0) {
            o->used[i] = 1;
            return i;
        }
    return -1;
}

static int get_int(objlist *o, const char *key, int *v)
{
    int i = find(o, key);
    if (i < 0) return 1;
    if (strlen(o->value[i]) == 0) return sofi_fail("Error in input file, value of object %s is empty!", key), -1;
    char *end = NULL;
    double d = strtod(o->value[i], &end);
    if (end && strlen(end) > 0)
        return sofi_fail("Error in input file, value of object %s contains more than one float: '%s'!", key, end), -1;
    double ip;
    if (modf(d, &ip) != 0)
        return sofi_fail("Error in input file, value of object %s is not an int : %f!", key, d), -1;
    *v = atoi(o->value[i]);
    return 0;
}

static int get_float(objlist *o, const char *key, float *v)
{
    int i = find(o, key);
    if (i < 0) return 1;
    if (strlen(o->value[i]) == 0) return sofi_fail("Error in input file, value of object %s is empty!", key), -1;
    char *end = NULL;
    double d = strtod(o->value[i], &end);
    if (end && strlen(end) > 0) {
        for (const char *c = end; *c; ++c)
            if (*c != ' ')
                return sofi_fail("Error in input file, value of object %s contains more than one float: '%s'!", key, end), -1;
    }
    *v = (float)d;
    return 0;
}

static int get_str(objlist *o, const char *key, char *v)
{
    int i = find(o, key);
    if (i < 0) return 1;
    if (strlen(o->value[i]) == 0) return sofi_fail("Error in input file, value of object %s is empty!", key), -1;
    strcpy(v, o->value[i]);
    return 0;
}

#define REQ_INT(K, F) do { int r_ = get_int(o, K, &(F)); if (r_ < 0) goto fail; if (r_) { sofi_fail("Variable %s could not be retrieved from the json input file!", K); goto fail; } } while (0)
#define REQ_FLT(K, F) do { int r_ = get_float(o, K, &(F)); if (r_ < 0) goto fail; if (r_) { sofi_fail("Variable %s could not be retrieved from the json input file!", K); goto fail; } } while (0)
#define REQ_STR(K, F) do { int r_ = get_str(o, K, (F)); if (r_ < 0) goto fail; if (r_) { sofi_fail("Variable %s could not be retrieved from the json input file!", K); goto fail; } } while (0)
#define OPT_INT(K, F, D) do { int r_ = get_int(o, K, &(F)); if (r_ < 0) goto fail; if (r_) (F) = (D); } while (0)
#define OPT_FLT(K, F, D) do { int r_ = get_float(o, K, &(F)); if (r_ < 0) goto fail; if (r_) (F) = (D); } while (0)
#define OPT_STR(K, F, D) do { int r_ = get_str(o, K, (F)); if (r_ < 0) goto fail; if (r_) strcpy((F), (D)); } while (0)

static const char *weq_names[8] = { "AC_ISO", "VAC_ISO", "EL_ISO", "VEL_ISO", "EL_VTI", "VEL_VTI", "EL_TTI", "VEL_TTI" };
static const char *mode_names[3] = { "FWI", "FW", "RTM" };

static int apply_workflow_stage1(sofi_params *p);

/* a relative file name of the parameter file taken relative to base_dir */
static int rebase(const char *base, char *path)
{
    if (!base || !base[0] || !path[0] || path[0] == '/') return 0;
    char tmp[SOFI_STR];
    const int n = snprintf(tmp, sizeof(tmp), "%s/%s", base, path);
    if (n < 0 || n >= (int)sizeof(tmp)) return sofi_fail("path %s/%s is too long", base, path);
    strcpy(path, tmp);
    return 0;
}

int sofi_read_params(const char *json_file, sofi_params *p) { return sofi_read_params_in(json_file, NULL, p); }

int sofi_read_params_in(const char *json_arg, const char *base, sofi_params *p)
{
    objlist *o = (objlist *)calloc(1, sizeof(objlist));
    if (!o) return sofi_fail("out of memory");
    memset(p, 0, sizeof(*p));
    if (base && strlen(base) >= SOFI_STR) {
        free(o);
        return sofi_fail("base directory %s is too long", base);
    }
    if (base) strcpy(p->base_dir, base);
    char json_file[SOFI_STR];
    if (strlen(json_arg) >= SOFI_STR) {
        free(o);
        return sofi_fail("parameter file name %s is too long", json_arg);
    }
    strcpy(json_file, json_arg);
    if (rebase(p->base_dir, json_file)) {
        free(o);
        return -1;
    }
    /* defaults of the GlobVar / GlobVarInv initialisers in src/sofi2D.c:86-89 */
    p->outntimestepinfo = 100;
    p->ndt = 1;
    p->idx = 1;
    p->idy = 1;
    p->itermax = 1;
    p->dtinv = 1;

    const char *suffix = strcasestr(json_file, ".json");
    if (!suffix) { sofi_fail("Parameter file %s has no .json suffix.", json_file); goto fail; }
    if (access(json_file, R_OK) != 0) { sofi_fail("Cannot open/read json parameter file %s.", json_file); goto fail; }
    if (read_objects(json_file, o)) goto fail;

    OPT_STR("LOG_VERBOSITY", p->log_verbosity, "INFO");
    REQ_INT("NPROCX", p->nprocx);
    REQ_INT("NPROCY", p->nprocy);
    REQ_INT("FDORDER", p->fdorder);
    {
        int r = get_int(o, "FDORDER_TIME", &p->fdorder_time);
        if (r < 0) goto fail;
        if (r) p->fdorder_time = 2;
        else if (p->fdorder_time != 2 && p->fdorder_time != 4) { sofi_fail("Only FDORDER_TIME 2 or 4 are supported!"); goto fail; }
    }
    REQ_INT("MAXRELERROR", p->maxrelerror);
    REQ_INT("NX", p->nxg);
    REQ_INT("NY", p->nyg);
    REQ_FLT("DH", p->dh);
    REQ_FLT("TIME", p->time);
    REQ_FLT("DT", p->dt);
    {
        char s[SOFI_STR];
        int r = get_str(o, "MODE", s);
        if (r < 0) goto fail;
        if (r) p->mode = SOFI_MODE_FW;
        else {
            int found = -1;
            for (int i = 0; i < 3; ++i) if (!strncasecmp(s, mode_names[i], SOFI_STR - 1)) { found = i; break; }
            if (found < 0) { sofi_fail("Unknown run mode (MODE) '%s' specified in parameter file.", s); goto fail; }
            if (found == 2) { sofi_fail("MODE '%s' is not yet implemented.", s); goto fail; }
            p->mode = found;
        }
    }
    if (p->mode == SOFI_MODE_FWI && p->fdorder_time == 4) p->fdorder_time = 2;
    {
        char s[SOFI_STR];
        int r = get_str(o, "WEQ", s);
        if (r < 0) goto fail;
        if (r) p->weq = SOFI_EL_ISO;
        else {
            int found = -1;
            for (int i = 0; i < 8; ++i) if (!strncasecmp(s, weq_names[i], SOFI_STR - 1)) { found = i; break; }
            if (found < 0) { sofi_fail("Unknown wave equation (WEQ) '%s' specified in parameter file.", s); goto fail; }
            p->weq = found;
        }
    }
    /* sources */
    p->source_topo = 0;
    REQ_INT("SOURCE_TYPE", p->source_type);
    {
        int r = get_int(o, "SIGOUT", &p->sigout);
        if (r < 0) goto fail;
        if (!r && p->sigout == 1) {
            REQ_STR("SIGOUT_FILE", p->sigout_file);
            REQ_INT("SIGOUT_FORMAT", p->sigout_format);
        }
    }
    REQ_INT("SOURCE_SHAPE", p->source_shape);
    if (p->source_shape == 3) REQ_STR("SIGNAL_FILE", p->signal_file);
    REQ_INT("SRCREC", p->srcrec);
    if (p->srcrec == 1) {
        REQ_STR("SOURCE_FILE", p->source_file);
        REQ_INT("RUN_MULTIPLE_SHOTS", p->run_multiple_shots);
        OPT_INT("SOURCE_TOPO", p->source_topo, 0);
        if (p->source_topo < 0 || p->source_topo > 1) { sofi_fail("Variable SOURCE_TOPO needs to be 0 or 1!"); goto fail; }
    }
    if (p->srcrec == 2) {
        REQ_FLT("PLANE_WAVE_DEPTH", p->plane_wave_depth);
        if (p->plane_wave_depth > 0.0f) {
            REQ_FLT("PLANE_WAVE_ANGLE", p->plane_wave_angle);
            REQ_FLT("TS", p->ts);
        }
    }
    /* boundaries */
    REQ_INT("FREE_SURF", p->free_surf);
    REQ_INT("BOUNDARY", p->boundary);
    REQ_INT("FW", p->fw);
    REQ_INT("ABS_TYPE", p->abs_type);
    if (p->abs_type == 1) {
        REQ_FLT("NPOWER", p->npower);
        REQ_FLT("K_MAX_CPML", p->k_max_cpml);
        REQ_FLT("FPML", p->fpml);
        REQ_FLT("VPPML", p->vppml);
    }
    if (p->abs_type == 2) REQ_FLT("DAMPING", p->damping);
    /* snapshots */
    REQ_INT("SNAP", p->snap);
    if (p->snap > 0) {
        REQ_INT("SNAP_FORMAT", p->snap_format);
        REQ_FLT("TSNAP1", p->tsnap1);
        REQ_FLT("TSNAP2", p->tsnap2);
        REQ_FLT("TSNAPINC", p->tsnapinc);
        REQ_STR("SNAP_FILE", p->snap_file);
    }
    OPT_INT("SNAPSHOT_START", p->snapshot_start, 1);
    OPT_INT("SNAPSHOT_END", p->snapshot_end, -9999);
    OPT_INT("SNAPSHOT_INCR", p->snapshot_incr, 1);
    REQ_INT("IDX", p->idx);
    REQ_INT("IDY", p->idy);
    if (p->snap && p->weq < SOFI_EL_ISO && p->snap > 2) { sofi_fail("Output of curl impossible in case of acoustic modelling. Choose SNAP=1,2"); goto fail; }
    /* seismograms */
    p->rec_topo = 0;
    REQ_INT("SEISMO", p->seismo);
    if (p->seismo > 0) {
        REQ_STR("REC_FILE", p->rec_file);
        REQ_STR("SEIS_FILE", p->seis_file);
        REQ_INT("READREC", p->readrec);
        OPT_INT("REC_TOPO", p->rec_topo, 0);
        if (p->rec_topo < 0 || p->rec_topo > 1) { sofi_fail("Variable REC_TOPO needs to be 0 or 1!"); goto fail; }
        if (p->readrec == 0) {
            REQ_FLT("XREC1", p->xrec1);
            REQ_FLT("XREC2", p->xrec2);
            REQ_FLT("YREC1", p->yrec1);
            REQ_FLT("YREC2", p->yrec2);
        }
        REQ_INT("NDT", p->ndt);
        REQ_INT("SEIS_FORMAT", p->seis_format);
        REQ_INT("REC_ARRAY", p->rec_array);
        if (p->rec_array > 0) {
            REQ_INT("DRX", p->drx);
            REQ_FLT("REC_ARRAY_DEPTH", p->rec_array_depth);
            REQ_FLT("REC_ARRAY_DIST", p->rec_array_dist);
        }
        REQ_FLT("REFRECX", p->refrec[1]);
        REQ_FLT("REFRECY", p->refrec[2]);
        REQ_FLT("NGEOPH", p->ngeoph);
    }
    if (p->seismo && p->weq < SOFI_EL_ISO && p->seismo > 2) { sofi_fail("Output of curl impossible in case of acoustic modelling. Choose SEISMO=1,2"); goto fail; }
    if ((p->source_topo == 1 || p->rec_topo == 1) && p->free_surf == 1) {
        sofi_fail("Source or receiver topography option cannot be combined with free-surface option.");
        goto fail;
    }
    /* model and log */
    REQ_STR("MFILE", p->mfile);
    OPT_INT("WRITE_MODELFILES", p->write_modelfiles, 0);
    REQ_INT("LOG", p->log);
    REQ_STR("LOG_FILE", p->log_file);
    REQ_INT("READMOD", p->readmod);
    OPT_INT("OUT_TIMESTEP_INFO", p->outntimestepinfo, 100);
    OPT_FLT("TAU", p->tau, 0.0001f);
    REQ_INT("L", p->L);
    if (p->L == 0 && (p->weq == SOFI_VEL_ISO || p->weq == SOFI_VEL_VTI || p->weq == SOFI_VEL_TTI || p->weq == SOFI_VAC_ISO)) {
        sofi_fail("L>0 required for viscoacoustic/elastic simulation!");
        goto fail;
    }
    if (p->L > 0 && (p->weq == SOFI_AC_ISO || p->weq == SOFI_EL_ISO || p->weq == SOFI_EL_VTI || p->weq == SOFI_EL_TTI)) p->L = 0;
    if (p->L > 0) {
        REQ_FLT("F_REF", p->f_ref);
        if (p->L > SOFI_MAXL) { sofi_fail("More than five relaxation parameters (L>5) not implemented!"); goto fail; }
        static const char *fln[6] = { "", "FL1", "FL2", "FL3", "FL4", "FL5" };
        for (int l = p->L; l >= 1; --l) REQ_FLT(fln[l], p->fl[l]);
    }
    p->nt = (int)floor(p->time / p->dt + 0.5);
    if (p->mode == SOFI_MODE_FWI) {
        if (p->idx != 1 || p->idy != 1) { p->idx = 1; p->idy = 1; }
        OPT_INT("ITERMAX", p->itermax, 1);
        OPT_STR("DATA_DIR", p->data_dir, "");
        OPT_INT("ADJOINT_TYPE", p->adjoint_type, 1);
        OPT_INT("VELOCITY", p->velocity, 0);
        OPT_INT("LNORM", p->lnorm, 2);
        if (p->lnorm == 8) REQ_FLT("WATERLEVEL_LNORM8", p->waterlevel_lnorm8);
        OPT_INT("DTINV", p->dtinv, 1);
        OPT_STR("GRADIENT", p->gradient_file, "gradient");
        OPT_INT("TIME_FILT", p->time_filt, 0);
        if (p->time_filt) {   /* src/read_par_json_fwi.c:497-545, src/sofi2D.c:156-164 */
            OPT_FLT("F_HIGH_PASS", p->f_high_pass, 0.0f);
            REQ_INT("ORDER", p->filt_order);
            if (p->time_filt == 1) {
                float f_end, f_incr;   /* required by the reference, unused by its loop */
                REQ_FLT("F_LOW_PASS_START", p->f_low_pass);
                REQ_FLT("F_LOW_PASS_END", f_end);
                REQ_FLT("F_LOW_PASS_INCR", f_incr);
                (void)f_end;
                (void)f_incr;
            } else if (p->time_filt == 2) {
                char ff[SOFI_STR];
                REQ_STR("FREQ_FILE", ff);
                if (rebase(p->base_dir, ff)) goto fail;
                /* filter_frequencies (src/filter_frequencies.c:25-60): the
                 * first nonzero number opening a line; FREQ_NR = 1 */
                FILE *fq = fopen(ff, "r");
                if (!fq) { sofi_fail("Freqency file could not be opened!"); goto fail; }
                char line[SOFI_STR];
                float v = 0.0f;
                int found = 0;
                while (!found && fgets(line, sizeof(line), fq))
                    if (sscanf(line, "%f", &v) == 1 && v != 0) found = 1;
                fclose(fq);
                if (!found) { sofi_fail("FREQ_FILE %s holds no frequency.", ff); goto fail; }
                p->f_low_pass = v;
            }
        }
        OPT_INT("EPRECOND", p->eprecond, 0);
        OPT_INT("TRKILL", p->trkill, 0);
        OPT_INT("TIMEWIN", p->timewin, 0);
        OPT_INT("USE_WORKFLOW", p->use_workflow, 0);
        if (p->use_workflow) REQ_STR("FILE_WORKFLOW", p->file_workflow);
        if (p->use_workflow && rebase(p->base_dir, p->file_workflow)) goto fail;
        if (p->use_workflow && apply_workflow_stage1(p)) goto fail;
        if (p->dtinv <= 0) { sofi_fail("Variable DTINV must be a positive integer!"); goto fail; }
    }
    /* the other file names of the parameter file */
    {
        char *names[] = { p->sigout_file, p->signal_file, p->source_file, p->snap_file, p->rec_file, p->seis_file,
                          p->mfile, p->log_file, p->data_dir, p->gradient_file };
        for (size_t k = 0; k < sizeof(names) / sizeof(names[0]); ++k)
            if (rebase(p->base_dir, names[k])) goto fail;
    }
    /* NS = iround(NT / NDT) with integer division (src/read_par_json.c:516) */
    p->ns = p->nt / p->ndt;
    p->nd = p->fdorder / 2;
    p->dtdh = p->dt / p->dh;
    /* file checks (src/read_par_json.c:532-574) */
    if (p->source_shape == 3 && access(p->signal_file, R_OK) != 0) { sofi_fail("Signal file %s does not exist or cannot be read.", p->signal_file); goto fail; }
    if (p->srcrec == 1 && access(p->source_file, R_OK) != 0) { sofi_fail("Source file %s does not exist or cannot be read.", p->source_file); goto fail; }
    if (p->readrec && access(p->rec_file, R_OK) != 0) { sofi_fail("Receiver file %s does not exist or cannot be read.", p->rec_file); goto fail; }
    free(o);
    return 0;
fail:
    free(o);
    return -1;
}

/* read_workflow + apply_workflow (src/read_workflow.c:32-113,
 * src/apply_workflow.c:28-110): the first line of FILE_WORKFLOW is a header;
 * every later line holding no '#' and something to read is a stage of 14
 * numbers.  The reference's loop never leaves stage 1 (nothing advances
 * WORKFLOW_STAGE), so every iteration takes TIME_FILT (column 5), when it
 * is positive F_HIGH_PASS and F_LOW_PASS (6, 7), and EPRECOND (8) from the
 * first stage; the other columns steer the model update, which the
 * reference does not run.  EPRECOND_MAX, the largest EPRECOND of the file
 * and the parameters, decides whether the energy arrays exist. */
static int apply_workflow_stage1(sofi_params *p)
{
    FILE *fp = fopen(p->file_workflow, "r");
    if (!fp) return sofi_fail("Workflow file  %s could not be opened.", p->file_workflow);
    char line[SOFI_STR], tok[SOFI_STR];
    float w[15] = { 0 }, emax = (float)p->eprecond;
    int stages = 0, lines = 0;
    /* the reference sizes its table by the readable lines less one, the
     * header it expects among them (src/read_workflow.c:48-63) */
    while (fgets(line, sizeof(line), fp))
        if (!strchr(line, '#') && sscanf(line, "%255s", tok) == 1) ++lines;
    rewind(fp);
    if (!fgets(line, sizeof(line), fp)) line[0] = '\0';        /* header */
    while (fgets(line, sizeof(line), fp)) {
        if (strchr(line, '#') || sscanf(line, "%255s", tok) != 1) continue;
        float v[15] = { 0 };
        sscanf(line, "%f%f%f%f%f%f%f%f%f%f%f%f%f%f", &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9],
               &v[10], &v[11], &v[12], &v[13], &v[14]);
        if (++stages > lines - 1) {
            fclose(fp);
            return sofi_fail("read_workflow.c: buffer not large enough to store all workflow parameters - "
                             "programming error (the first line of %s is its header).", p->file_workflow);
        }
        if (stages == 1) memcpy(w, v, sizeof(w));
        if (v[8] > emax) emax = v[8];
    }
    fclose(fp);
    if (!stages) return sofi_fail("Workflow file %s holds no stage.", p->file_workflow);
    p->time_filt = (int)w[5];
    if (p->time_filt > 0) {
        p->f_high_pass = w[6];
        p->f_low_pass = w[7];
    }
    p->eprecond = (int)w[8];
    if (p->eprecond > 0 && emax != 1.0f && emax != 3.0f)
        return sofi_fail("MODE=FWI: the workflow's EPRECOND writes an array the reference does not allocate "
                         "(src/initmem_fwi.c:57-62, src/init_inv.c:28-33); refused.");
    return 0;
}

/* What this build runs; everything else is refused with a message rather
 * than run wrongly. */
int sofi_check_supported(const sofi_params *p)
{
    if ((p->weq < SOFI_EL_ISO || p->weq > SOFI_VEL_TTI) && p->weq != SOFI_AC_ISO)
        return sofi_fail("WEQ %s is not supported by the reference's time loop (src/time_loop.c:123-212).", weq_names[p->weq]);
    if (p->weq == SOFI_AC_ISO) {
        /* the acoustic path keeps one pressure array (psyy aliases psxx,
         * src/initmem_wavefield.c:31-33) and allocates no sxy */
        if (p->fdorder_time == 4) return sofi_fail("WEQ not yet implemented for 4th order in time.");
        if (p->free_surf)
            return sofi_fail("FREE_SURF with AC_ISO: the reference's surface_elastic reads the unallocated pu "
                             "(src/surface_elastic.c:80, src/initmem_model.c:32-39); refused.");
        if (p->source_type == 5)
            return sofi_fail("SOURCE_TYPE=5 with AC_ISO: eqsource writes the unallocated sxy (src/eqsource.c); refused.");
        if (p->mode == SOFI_MODE_FWI)
            return sofi_fail("MODE=FWI with AC_ISO: calc_conv reads the unallocated psxy (src/calc_conv.c:58-60, "
                             "src/initmem_wavefield.c:31-33); refused.");
    }
    if (p->fdorder_time == 4) {
        /* the reference's 4th-order-in-time loop exists for the isotropic
         * rheologies only (src/prepmod.c:162-187, src/time_loop.c:215-245) */
        if (p->weq != SOFI_EL_ISO && p->weq != SOFI_VEL_ISO)
            return sofi_fail("WEQ not yet implemented for 4th order in time.");
        if (p->fw > 0 && p->abs_type != 1 && p->abs_type != 2)
            return sofi_fail("FDORDER_TIME=4 needs ABS_TYPE 1 or 2 with FW > 0 (the reference damps the stresses with "
                             "an absorbing frame it does not allocate otherwise).");
    }
    if (p->fdorder != 2 && p->fdorder != 4 && p->fdorder != 6 && p->fdorder != 8 && p->fdorder != 10 && p->fdorder != 12)
        return sofi_fail("Error in selection of FD coefficients: wrong FDORDER!");
    if (p->maxrelerror < 0 || p->maxrelerror > 4)
        return sofi_fail("Error in selection of FD coefficients: wrong choice of maximum relative error!");
    /* READMOD = 0: the internal models of src/hh_*.c (no acoustic one,
     * src/readmod.c:60-84; under MODE=FWI the reference averages starting
     * models that the internal models never set, src/readmod.c:86-97) */
    if (p->readmod != 0 && p->readmod != 1) return sofi_fail("READMOD=%d: 0 (internal model) or 1 (model files).", p->readmod);
    if (p->readmod == 0 && p->weq == SOFI_AC_ISO) return sofi_fail("Internal model for your chosen WEQ not implemented.");
    if (p->readmod == 0 && p->mode == SOFI_MODE_FWI)
        return sofi_fail("READMOD=0 with MODE=FWI: the internal models set no starting model for the gradient.");
    if (p->free_surf && (p->weq == SOFI_VEL_VTI || p->weq == SOFI_VEL_TTI))
        return sofi_fail("FREE_SURF with VEL_VTI/VEL_TTI dereferences unallocated tau arrays in the reference (src/surface.c:94-95); refused.");
    if (p->snap && p->snap_format != 2 && p->snap_format != 3)
        return sofi_fail("SNAP_FORMAT=%d is not supported (the reference supports 2 and 3, src/writedsk.c).", p->snap_format);
    if (p->source_shape == 7) return sofi_fail("SOURCE_SHAPE=7 (STF inversion) is not supported.");
    if (p->mode == SOFI_MODE_FWI) {
        /* the gradient path: every elastic and viscoelastic rheology (the
         * stress updates write ux, uy, uxy for the store, calc_conv is the
         * same for all), the residuals of every LNORM calc_res has */
        if (p->ndt != 1)
            return sofi_fail("MODE=FWI needs NDT=1: the adjoint sources are read at every time step (src/update_v_interior.c:79-86).");
        if (p->adjoint_type < 1 || p->adjoint_type > 3)
            return sofi_fail("MODE=FWI: ADJOINT_TYPE=%d is not supported (1: vx and vy, 2: vy, 3: vx).", p->adjoint_type);
        if (p->seismo != 1 && p->seismo != 4) return sofi_fail("MODE=FWI needs SEISMO=1 or 4 (vx and vy sections).");
        if (p->lnorm < 1 || p->lnorm > 8 || p->lnorm == 6)
            return sofi_fail("MODE=FWI: LNORM=%d has no residual in the reference (src/calc_res.c: 1-5, 7, 8).", p->lnorm);
        if (p->time_filt < 0 || p->time_filt > 2) return sofi_fail("MODE=FWI: TIME_FILT=%d (0, 1 or 2).", p->time_filt);
        if (p->time_filt && (p->filt_order < 1 || p->filt_order > 17))
            return sofi_fail("MODE=FWI: ORDER=%d of the time filter is not supported (1..17).", p->filt_order);
        /* trace killing from a file: calc_res skips killed traces and leaves
         * their residual buffer as the previous shot left it
         * (src/calc_res.c:62-125,196-197); not built */
        if (p->trkill == 1) return sofi_fail("MODE=FWI: TRKILL=1 (trace killing) is not supported by this build.");
        /* EPRECOND 1 / 3 accumulate the source and receiver wavefield
         * energies Ws, Wr (src/eprecond.c:9-18, src/time_loop.c:297-301),
         * which nothing in the reference reads back: the outputs do not
         * depend on them, so the key is accepted and nothing is accumulated.
         * EPRECOND 2 zeroes We_sum (src/init_inv.c:28-33), which only 1 and
         * 3 allocate (src/initmem_fwi.c:57-62). */
        if (p->eprecond == 2 || p->eprecond < 0 || p->eprecond > 3)
            return sofi_fail("MODE=FWI: EPRECOND=%d writes an unallocated array in the reference (src/init_inv.c:28-33); refused.", p->eprecond);
    }
    return 0;
}
