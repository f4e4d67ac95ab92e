// sofi_kernels.hip - SOFI2D velocity-stress time loop on CDNA4 (gfx950).
//
// One handle = one shot's state resident in HBM, one HIP stream.  A time
// step is (src/time_loop.c:42-322):
//   force injection -> velocity update (interior + CPML/damping frame fused)
//   -> stress update (v-derivatives computed in-kernel, CPML/damping fused)
//   -> [div/curl sampling] -> explosive/moment source -> free surface
//   -> [vx/vy/p sampling] -> [snapshot]
// The reference's exchange_v/exchange_s and the full-grid derivative arrays
// of v_derivatives disappear for ISO/VTI: every stress cell recomputes its
// four velocity derivatives from vx/vy.  TTI needs neighbouring derivatives
// after the frame CPML has modified them in place, so it keeps the four
// derivative arrays and the reference's pass order (interior, frame CPML,
// frame update).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <map>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "sofi_hip.h"
#include "../common/sofi_iir.h"

namespace {

thread_local std::string g_err;

int fail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
int fail(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return -1;
}

#define HIPCHK(x)                                                                                     \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return fail("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

// ------------------------------------------------------------------ params

struct Grid {
    int nx, ny, pitch, fw, abs_type, free_surf, weq, L, nd;
    int gx2, gx3, gy2, gy3;
    int plane;      // floats per padded plane
    int wx, wy;     // row lengths of x-frame and y-frame CPML memory
    int ntx;        // tiles along x of the one-cell-per-thread update kernels
    int nty;        // tiles along y of the same
    // tiles (BX x BY) that hold a cell of the absorbing frame or one next to
    // it: tile rows at the top / bottom, tile columns at the left / right
    // (the TTI frame passes run over these tiles only)
    int ftop, fbot, fl, fr;
    int periodic;   // BOUNDARY = 1: the grid wraps around in x
    int nb, spt;    // shot slots of a batch; slots per thread of the update kernels
    int k_lo, k_hi; // the slots this launch updates (a group of the batch)
    int nxp, nyp;   // floats per slot of the x-frame and y-frame CPML memory
    float dt, dh, dtdh, dhi, dthalbe;
    float dhi4;     // 1.0 / DH rounded from double (the 4th-order stress updates' dhi)
    float hc[7];
};

// Wavefield of a handle.  A batch of nb shots sharing one medium keeps its
// slots one after the other in every array (slot k of a plane array at
// k * plane, of the memory variables at k * L * plane, of the CPML memory at
// k * nxp / k * nyp); at(k) gives slot k's view.
struct Fields {
    float *vx, *vy, *sxx, *syy, *sxy;
    float *r, *p, *q;                          // L planes each
    float *dvxx, *dvyy, *dvyx, *dvxy;           // TTI only
    float *psi_sxx_x, *psi_sxy_x, *psi_syy_y, *psi_sxy_y;
    float *psi_vxx, *psi_vyx, *psi_vyy, *psi_vxy;
    // MODE=FWI, one plane each per slot: the rho-weighted accelerations of
    // the forward pass (the adjoint velocities or their time integral in the
    // adjoint pass), the stress increments over DT, and the shot's lambda /
    // mu / rho gradient accumulators
    float *pvxp1, *pvyp1, *ux, *uy, *uxy, *glam, *gmu, *grho;
    int *unstable;                             // per slot: first step whose start finds vy(NX/2, NY/2) = NaN

    __host__ __device__ __forceinline__ Fields at(int k, const Grid &g) const
    {
        Fields s = *this;
        if (k == 0) return s;
        const size_t pl = (size_t)k * g.plane, ml = pl * (size_t)(g.L > 0 ? g.L : 1);
        const size_t xp = (size_t)k * g.nxp, yp = (size_t)k * g.nyp;
        s.vx += pl;
        s.vy += pl;
        s.sxx += pl;
        s.syy += pl;
        s.sxy += pl;
        if (r) {
            s.r += ml;
            s.p += ml;
            s.q += ml;
        }
        if (dvxx) {
            s.dvxx += pl;
            s.dvyy += pl;
            s.dvyx += pl;
            s.dvxy += pl;
        }
        if (pvxp1) {
            s.pvxp1 += pl;
            s.pvyp1 += pl;
            s.ux += pl;
            s.uy += pl;
            s.uxy += pl;
            s.glam += pl;
            s.gmu += pl;
            s.grho += pl;
        }
        if (psi_sxx_x) {
            s.psi_sxx_x += xp;
            s.psi_sxy_x += xp;
            s.psi_vxx += xp;
            s.psi_vyx += xp;
            s.psi_syy_y += yp;
            s.psi_sxy_y += yp;
            s.psi_vyy += yp;
            s.psi_vxy += yp;
        }
        s.unstable += k;
        return s;
    }
};

struct Mat {
    const float *rip, *rjp;
    const float *f, *g, *fipjp;                 // EL_ISO / VEL_ISO
    const float *bd, *be, *br;                  // VEL_ISO defect bases: pu*taus, ppi*taup, puipjp*tausipjp
    const float *c11, *c13, *c33, *c15, *c35, *c55i, *c15i, *c35i;   // elastic VTI/TTI (DT-scaled)
    const float *u11, *u13, *u33, *u15, *u35, *u55i, *u15i, *u35i;   // visco VTI/TTI unrelaxed
    const float *d11, *d13, *d33, *d15, *d35, *d55i, *d15i, *d35i;   // visco VTI/TTI defects (L planes)
    const float *absorb;
    const float *pu, *ppi, *ptaus, *ptaup;      // snapshots
    const float *surf;                          // free-surface coefficients of row 1, SURF_N x NX
    const float *prof;                          // CPML profiles, 12 x (2FW+1)
    float eta[SOFI_MAXL + 1], bip[SOFI_MAXL + 1], cip[SOFI_MAXL + 1];
    float bjm[SOFI_MAXL + 1];                   // free surface: eta_l / (1 + eta_l / 2)
};

__device__ __forceinline__ int IDX(const Grid &g, int i, int j) { return (j - 1 + SOFI_PAD) * g.pitch + (i - 1 + SOFI_PAD); }

// CPML profile k (enum sofi_profile) at index h
__device__ __forceinline__ float PRF(const Grid &g, const Mat &m, int k, int h) { return m.prof[k * (2 * g.fw + 1) + h]; }

__device__ __forceinline__ bool in_xframe(const Grid &g, int i) { return i <= g.gx2 || i > g.gx3; }
__device__ __forceinline__ bool in_yframe(const Grid &g, int j) { return j <= g.gy2 || j > g.gy3; }
__device__ __forceinline__ int hx(const Grid &g, int i) { return i <= g.gx2 ? i : i - g.nx + 2 * g.fw; }
__device__ __forceinline__ int hy(const Grid &g, int j) { return j <= g.gy2 ? j : j - g.ny + 2 * g.fw; }

// ------------------------------------------------------------ FD stencils

// op_v_fd* (src/initfd.c:97-183): stress derivatives, no 1/DH
template <int ND>
__device__ __forceinline__ void op_v(const Grid &g, const float *__restrict__ sxx, const float *__restrict__ syy,
                                     const float *__restrict__ sxy, int c, float &sxx_x, float &sxy_x,
                                     float &sxy_y, float &syy_y)
{
    const int P = g.pitch;
    float a = 0.f, b = 0.f, d = 0.f, e = 0.f;
#pragma unroll
    for (int m = 1; m <= ND; ++m) {
        const float h = g.hc[m];
        a += h * (sxx[c + m] - sxx[c - m + 1]);
        b += h * (sxy[c + m - 1] - sxy[c - m]);
        d += h * (sxy[c + (m - 1) * P] - sxy[c - m * P]);
        e += h * (syy[c + m * P] - syy[c - (m - 1) * P]);
    }
    sxx_x = a;
    sxy_x = b;
    sxy_y = d;
    syy_y = e;
}

// op_s_fd* (src/initfd.c:9-95): velocity derivatives times 1/DH
template <int ND>
__device__ __forceinline__ void op_s(const Grid &g, const float *__restrict__ vx, const float *__restrict__ vy, int c,
                                     float &vxx, float &vyx, float &vxy, float &vyy)
{
    const int P = g.pitch;
    float a = 0.f, b = 0.f, d = 0.f, e = 0.f;
#pragma unroll
    for (int m = 1; m <= ND; ++m) {
        const float h = g.hc[m];
        a += h * (vx[c + m - 1] - vx[c - m]);
        b += h * (vy[c + m] - vy[c - m + 1]);
        d += h * (vx[c + m * P] - vx[c - (m - 1) * P]);
        e += h * (vy[c + (m - 1) * P] - vy[c - m * P]);
    }
    vxx = a * g.dhi;
    vyx = b * g.dhi;
    vxy = d * g.dhi;
    vyy = e * g.dhi;
}

// CPML on the stress-side derivatives of one frame cell (cpml_update_s_x/_y,
// src/cpml_update.c:30-46); updates the psi memory and returns modified values.
__device__ __forceinline__ void cpml_dv(const Grid &g, const Mat &m, Fields &f, int i, int j, bool xf, bool yf,
                                        float &vxx, float &vyx, float &vxy, float &vyy)
{
    if (xf) {
        const int h = hx(g, i);
        const int o = j * g.wx + h;
        float px = PRF(g, m, SOFI_P_B_X, h) * f.psi_vxx[o] + PRF(g, m, SOFI_P_A_X, h) * vxx;
        f.psi_vxx[o] = px;
        vxx = vxx / PRF(g, m, SOFI_P_K_X, h) + px;
        float py = PRF(g, m, SOFI_P_B_X_HALF, h) * f.psi_vyx[o] + PRF(g, m, SOFI_P_A_X_HALF, h) * vyx;
        f.psi_vyx[o] = py;
        vyx = vyx / PRF(g, m, SOFI_P_K_X_HALF, h) + py;
    }
    if (yf) {
        const int h = hy(g, j);
        const int o = h * g.wy + i;
        float py = PRF(g, m, SOFI_P_B_Y, h) * f.psi_vyy[o] + PRF(g, m, SOFI_P_A_Y, h) * vyy;
        f.psi_vyy[o] = py;
        vyy = vyy / PRF(g, m, SOFI_P_K_Y, h) + py;
        float px = PRF(g, m, SOFI_P_B_Y_HALF, h) * f.psi_vxy[o] + PRF(g, m, SOFI_P_A_Y_HALF, h) * vxy;
        f.psi_vxy[o] = px;
        vxy = vxy / PRF(g, m, SOFI_P_K_Y_HALF, h) + px;
    }
}

// read-only version: the modified derivative from the psi value already
// written this step (used by the div/curl sampler)
__device__ __forceinline__ void cpml_dv_read(const Grid &g, const Mat &m, const Fields &f, int i, int j, bool xf,
                                             bool yf, float &vxx, float &vyx, float &vxy, float &vyy)
{
    if (xf) {
        const int h = hx(g, i);
        const int o = j * g.wx + h;
        vxx = vxx / PRF(g, m, SOFI_P_K_X, h) + f.psi_vxx[o];
        vyx = vyx / PRF(g, m, SOFI_P_K_X_HALF, h) + f.psi_vyx[o];
    }
    if (yf) {
        const int h = hy(g, j);
        const int o = h * g.wy + i;
        vyy = vyy / PRF(g, m, SOFI_P_K_Y, h) + f.psi_vyy[o];
        vxy = vxy / PRF(g, m, SOFI_P_K_Y_HALF, h) + f.psi_vxy[o];
    }
}

// ------------------------------------------------------------- sources

// Sources of every slot of the handle, one list (slot[l] = the shot slot of
// source l); signals [n][nt].
struct Src {
    const int *i, *j, *type, *slot;
    const float *azi, *amp;
    const float *sig;   // [nsrc][nt]
    const double *sc;   // [nsrc][2]: sin and cos of the azimuth (degrees -> float radians), host libm
    int n, nt;
};

// Injection fused into the update kernels for shots with few sources: the
// thread owning a source cell adds the sources in source order, so that
// coincident sources sum as in the reference.
// Receivers of every slot of the handle, one list: receiver k samples slot
// shot[k] at (i[k], j[k]); its trace of component c starts at
// sec + off[k] + c * cst[k] (each slot's sections are [5][ntr][ns]).
struct Rec {
    const int *i, *j, *shot;
    const long long *off, *cst;
    const double *sq_ppi, *sq_pu, *pfac;   // sqrt(ppi), sqrt(pu), (3ppi-4pu)/(2ppi-2pu)
    float *sec;
    int n, ns;
};

constexpr int FUSE_MAX = 8;
struct Fuse {
    int nforce;         // sources whose forces the velocity kernel adds (0: none)
    int npsrc;          // sources whose explosive term the stress kernel adds (0: none)
    int surf;           // the stress kernel applies the free surface
    int c[FUSE_MAX];    // cell offsets of the sources, in source order
    int soff;           // index of the first of these sources in the handle's source list
    // FWI (src/wavefield_update_v.c:31-43, src/wavefield_update_s_el.c:43-47):
    // velocity kernel: 0 nothing, 1 forward rho-weighted accelerations,
    // 2 adjoint time integral of v, 3 adjoint copy of v (into pvxp1/pvyp1);
    // stress kernel: u / DT of the EL_ISO update into ux/uy/uxy
    int fwi_acc, fwi_u;
    const int *ntb;     // inside a captured block of steps: the block's base step (nt is then local)
    // vx / vy sampling carried by the stress launch of a one-shot handle
    // (SEISMO = 1): blocks ntile.. of the launch sample the receivers, whose
    // velocities the stress update does not change
    int sblk, ntile, ins, ndt;
    Rec rec;
};

// The explosive sources of the slots a slot-block stress launch updates,
// added by the thread that updates the source's cell in the source's slot
// after the update (psource, src/psource.c), in source order: slot k[w],
// cell offset c[w] within the slot, source l[w] of the handle's list
struct PFuse {
    int n;
    int k[FUSE_MAX], c[FUSE_MAX], l[FUSE_MAX];
};

// the time step of a launch: nt, or inside a captured block of steps the
// block's local step nt plus its base step, which the device holds so that
// one captured block serves every block of the run (the update kernels read
// it only in the threads that inject a source or flag an instability)
__device__ __forceinline__ int step_of(const int *ntb, int nt) { return ntb ? nt + *ntb : nt; }

// Sources grouped by cell, for injection one thread per distinct cell; a
// cell's sources are added in source order, as the reference's loop does.
struct Inj {
    int n;              // distinct cells
    const int *cell;    // [n] cell offsets
    const int *first;   // [n + 1] start of each cell's run in idx
    const int *idx;     // source indices
};

// force of source l at step nt on the velocities of its cell, added before
// the velocity update (src/update_v_interior.c:49-71)
__device__ __forceinline__ void force_add(const Grid &g, const Src &s, int l, int nt, float rip, float rjp, float &vx,
                                          float &vy)
{
    const float amp = (g.dt * s.sig[l * s.nt + nt - 1]) / (g.dh * g.dh);
    const int t = s.type[l];
    if (t == 2) vx += rip * amp;
    if (t == 3) vy += rjp * amp;
    if (t == 4) {
        vx = (float)((double)vx + s.sc[2 * l] * (double)(rip * amp));
        vy = (float)((double)vy + s.sc[2 * l + 1] * (double)(rjp * amp));
    }
}

// psource (src/psource.c:27-54): centred time derivative of the signal
__device__ __forceinline__ float psource_amp(const Grid &g, const Src &s, int l, int nt)
{
    const float *sg = s.sig + l * s.nt - 1;   // sg[k] = signal at step k
    const double h2 = 2.0 * g.dh * g.dh;
    if (nt == 1) return s.nt >= 2 ? (float)(sg[2] / h2) : 0.0f;
    if (nt < s.nt) return (float)((sg[nt + 1] - sg[nt - 1]) / h2);
    return (float)(-sg[nt - 1] / h2);
}

// force sources when they are not fused: one thread per source cell (cell
// offsets count the slots of a batch; the medium is shared by the slots)
__global__ void k_force(Grid g, Fields f, Mat m, Src s, Inj inj, int nt, const int *ntb)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= inj.n) return;
    nt = step_of(ntb, nt);
    const int c = inj.cell[k], cm = c % g.plane;
    float vx = f.vx[c], vy = f.vy[c];
    const float rip = m.rip[cm], rjp = m.rjp[cm];
    for (int q = inj.first[k]; q < inj.first[k + 1]; ++q) force_add(g, s, inj.idx[q], nt, rip, rjp, vx, vy);
    f.vx[c] = vx;
    f.vy[c] = vy;
}

// explosive sources when they are not fused: one thread per source cell
__global__ void k_psource(Grid g, Fields f, Src s, Inj inj, int nt, const int *ntb)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= inj.n) return;
    nt = step_of(ntb, nt);
    const int c = inj.cell[k];
    if (f.syy == f.sxx) {     // acoustic: sxx += amp, then syy (the same array) += amp
        float p = f.sxx[c];
        for (int q = inj.first[k]; q < inj.first[k + 1]; ++q) {
            const float amp = psource_amp(g, s, inj.idx[q], nt);
            p += amp;
            p += amp;
        }
        f.sxx[c] = p;
        return;
    }
    float sxx = f.sxx[c], syy = f.syy[c];
    for (int q = inj.first[k]; q < inj.first[k + 1]; ++q) {
        const float amp = psource_amp(g, s, inj.idx[q], nt);
        sxx += amp;
        syy += amp;
    }
    f.sxx[c] = sxx;
    f.syy[c] = syy;
}

// adjoint sources at the receivers (src/update_v_interior.c:72-90):
// ADJOINT_TYPE 1 adds the x and y residuals, 2 only y, 3 only x
__global__ void k_adjoint(Grid g, Fields f, Inj inj, const float *sx, const float *sy, int nt_total, int type, int nt,
                          const int *ntb)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= inj.n) return;
    nt = step_of(ntb, nt);
    const int c = inj.cell[k];
    float vx = f.vx[c], vy = f.vy[c];
    for (int q = inj.first[k]; q < inj.first[k + 1]; ++q) {
        const size_t o = (size_t)inj.idx[q] * nt_total + (nt - 1);
        if (type == 1 || type == 3) vx += sx[o];
        if (type == 1 || type == 2) vy += sy[o];
    }
    f.vx[c] = vx;
    f.vy[c] = vy;
}

// eqsource (src/eqsource.c:25-52): double-couple moment tensor; the
// sources listed (their updates touch neighbouring cells), one thread in order
__global__ void k_eqsource(Grid g, Fields f, Src s, const int *list, int n, int nt, const int *ntb, int lo, int hi)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    nt = step_of(ntb, nt);
    for (int q = 0; q < n; ++q) {
        const int l = list[q];
        if (s.slot[l] < lo || s.slot[l] >= hi) continue;     // slots outside the group being advanced
        const int i = s.i[l], j = s.j[l];
        const float dip = (float)(s.azi[l] * M_PI / 180.0);
        const float scale = g.dt / (g.dh * g.dh);
        const float amp = s.amp[l] * scale * s.sig[l * s.nt + nt - 1];
        const float m11 = (float)(-sin(2.0 * dip)), m12 = (float)(-cos(2.0 * dip)), m22 = (float)sin(2.0 * dip);
        const int c = s.slot[l] * g.plane + IDX(g, i, j);
        f.sxx[c] += amp * m11;
        f.sxy[c] = (float)((double)f.sxy[c] + 0.3 * amp * m12);
        f.sxy[c - g.pitch] = (float)((double)f.sxy[c - g.pitch] + 0.3 * amp * m12);
        f.sxy[c - 1] = (float)((double)f.sxy[c - 1] + 0.3 * amp * m12);
        f.syy[c] += amp * m22;
    }
}

// --------------------------------------------------------- free surface

// Per-column coefficients of the row-1 correction, formed on the host from
// the material at row 1 with the reference's rounding (src/surface.c:53-58,
// src/surface_elastic.c:52-54): with fjm = 2 mu (visco: 2 mu (1 + L tau_s))
// and gg = pi (visco: pi (1 + L tau_p)), t1 = DT (gg - fjm)^2, t2 = DT (gg -
// fjm); visco only: de = 2 mu tau_s - pi tau_p, fg = fjm / gg - 1.
enum { SURF_T1 = 0, SURF_T2, SURF_G, SURF_DE, SURF_FG, SURF_N };

// surface_elastic / surface at j = 1 (src/surface_elastic.c:27-90,
// src/surface.c:27-113) for column i, except the imaging of syy and sxy into
// the ghost rows: zero the q memory variables, correct sxx (and the p memory
// variables) with the horizontal derivative only; in the x-CPML the second
// update of psi_vxx[1][h] in the same step is reproduced.  sxx is the cell's
// value after the stress update and source injection; returns the new one.
template <int ND, bool VISCO>
__device__ __forceinline__ float surface_sxx(const Grid &g, Fields &f, const Mat &m, int i, int c, float sxx)
{
    const int j = 1, P = g.pitch;
    const float dhi = (float)(1.0 / g.dh);
    if (VISCO)
        for (int l = 1; l <= g.L; ++l) f.q[(l - 1) * g.plane + c] = 0.0f;
    float vxx = 0.f, vyy = 0.f;
#pragma unroll
    for (int m1 = 1; m1 <= ND; ++m1) {
        vxx += g.hc[m1] * (f.vx[c + m1 - 1] - f.vx[c - m1]);
        vyy += g.hc[m1] * (f.vy[c + (m1 - 1) * P] - f.vy[c - m1 * P]);
    }
    vxx *= dhi;
    vyy *= dhi;
    if (g.abs_type == 1 && !g.periodic) {   // no x-CPML at a periodic edge (src/surface_elastic.c:64-78)
        if (i <= g.fw) {
            const int o = j * g.wx + i;
            const float ps = PRF(g, m, SOFI_P_B_X, i) * f.psi_vxx[o] + PRF(g, m, SOFI_P_A_X, i) * vxx;
            f.psi_vxx[o] = ps;
            vxx = vxx / PRF(g, m, SOFI_P_K_X, i) + ps;
        }
        if (i >= g.nx - g.fw + 1) {
            const int h = i - g.nx + 2 * g.fw, o = j * g.wx + h;
            const float ps = PRF(g, m, SOFI_P_B_X, h) * f.psi_vxx[o] + PRF(g, m, SOFI_P_A_X, h) * vxx;
            f.psi_vxx[o] = ps;
            vxx = vxx / PRF(g, m, SOFI_P_K_X, h) + ps;
        }
    }
    const float *sc = m.surf + (i - 1);
    const float t1 = sc[SURF_T1 * g.nx], t2 = sc[SURF_T2 * g.nx], gg = sc[SURF_G * g.nx];
    if (!VISCO) {
        sxx += -(t1 * vxx / gg) - (t2 * vyy);
    } else {
        float sump = 0.f;
        for (int l = 1; l <= g.L; ++l) sump += f.p[(l - 1) * g.plane + c];
        sxx += -(t1 * vxx / gg) - (t2 * vyy) - (g.dthalbe * sump);
        const float de = sc[SURF_DE * g.nx], fg = sc[SURF_FG * g.nx];
        sump = 0.f;
        for (int l = 1; l <= g.L; ++l) {
            float &pv = f.p[(l - 1) * g.plane + c];
            pv += m.bjm[l] * ((de * fg * vxx) - (de * vyy));
            sump += pv;
        }
        sxx += (g.dthalbe * sump);
    }
    return sxx;
}

// the free surface as a pass of its own, after the stress update and sources
template <int ND, bool VISCO>
__global__ void k_surface(Grid g, Fields f0, Mat m)
{
    Fields f = f0.at(g.k_lo + blockIdx.y, g);   // shot slot of a batch
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i > g.nx) return;
    const int P = g.pitch, c = IDX(g, i, 1);
    f.syy[c] = 0.0f;
#pragma unroll
    for (int m1 = 1; m1 <= ND; ++m1) {
        f.syy[c - m1 * P] = -f.syy[c + m1 * P];
        f.sxy[c - m1 * P] = -f.sxy[c + (m1 - 1) * P];
    }
    f.sxx[c] = surface_sxx<ND, VISCO>(g, f, m, i, c, f.sxx[c]);
}

// ------------------------------------------------------------ tile mapping

// Tiles of the ISO/VTI update kernels are numbered row-major over the grid
// in workgroup order.  The hardware hands workgroups round-robin to the
// eight XCDs in that order, so at any moment the whole GPU works on one
// narrow band of rows and a tile's vertical stencil halo was fetched moments
// before by the tiles above it (from the memory-side cache when another XCD
// fetched it).  Measured on the 4000 x 2000 viscoelastic workload, this
// order beat giving each XCD a contiguous band of its own by 4 % (stress)
// and 3 % (velocity).
__device__ __forceinline__ int tile_id() { return blockIdx.x; }

// Tile shapes: CELL, one thread per cell in BX x BY tiles (also the TTI
// passes); VEC2 / VEC4, one thread per VW = 2 / 4 consecutive cells of a row
// in (VW BX) x BY tiles, every stencil row and coefficient fetched as one
// aligned VW-float vector (interior rows start on 256-B boundaries and the
// first of a thread's cells sits at a column offset divisible by VW).
constexpr int BX = 64, BY = 4;

// seismo_ssg vx / vy / p (src/seismo_ssg.c:40-64) of receiver k
__device__ __forceinline__ void seis_sample(const Grid &g, const Fields &f0, const Rec &r, int k, int ins, int seismo)
{
    const Fields f = f0.at(r.shot[k], g);
    const int c = IDX(g, r.i[k], r.j[k]);
    float *t = r.sec + r.off[k] + (ins - 1);
    const long long cs = r.cst[k];
    if (seismo == 1 || seismo == 4) {
        t[SOFI_C_VX * cs] = f.vx[c];
        t[SOFI_C_VY * cs] = f.vy[c];
    }
    if (seismo == 2) t[SOFI_C_P * cs] = -f.sxx[c] - f.syy[c];
    if (seismo == 4) t[SOFI_C_P * cs] = (float)(r.pfac[k] * (double)(-f.sxx[c] - f.syy[c]) / 3);
}

// the sampling blocks of a stress launch (Fuse::sblk); true when this block is one
__device__ __forceinline__ bool fused_seis(const Grid &g, const Fields &f0, const Fuse &fz)
{
    if (fz.sblk == 0 || (int)blockIdx.x < fz.ntile) return false;
    const int k = ((int)blockIdx.x - fz.ntile) * (BX * BY) + threadIdx.y * BX + threadIdx.x;
    if (k < fz.rec.n) seis_sample(g, f0, fz.rec, k, fz.ntb ? (fz.ins + *fz.ntb) / fz.ndt : fz.ins, 1);
    return true;
}

// loads and stores of the one-shot cell kernels' arrays that are touched only
// at the thread's own cell: with nontemporal hints when NT (see k_velocity)
template <bool NT>
__device__ __forceinline__ float ld1(const float *p, int c)
{
    if (NT) return __builtin_nontemporal_load(p + c);
    return p[c];
}
template <bool NT>
__device__ __forceinline__ void st1(float *p, int c, float v)
{
    if (NT) __builtin_nontemporal_store(v, p + c);
    else p[c] = v;
}

template <int VW>
using vecf = float __attribute__((ext_vector_type(VW)));

template <int VW>
__device__ __forceinline__ vecf<VW> ldv(const float *p, int c)
{
    return *reinterpret_cast<const vecf<VW> *>(p + c);
}
template <int VW>
__device__ __forceinline__ void stv(float *p, int c, vecf<VW> v)
{
    *reinterpret_cast<vecf<VW> *>(p + c) = v;
}

// ---------------------------------------------------------- velocity update

// update_v_interior + update_v_PML/update_v_abs + wavefield_update_v
// (src/update_v_interior.c:49-98, src/update_v_PML.c, src/update_v_abs.c,
// src/wavefield_update_v.c:46-47) of cell (i, j) given the four stress
// derivatives of op_v (no 1/DH), on the cell's velocities in registers.
// cpml_update_v_x / _y (src/cpml_update.c:50-66) on the stress derivatives of
// a frame cell
__device__ __forceinline__ void cpml_v(const Grid &g, Fields &f, const Mat &m, int i, int j, bool xf, bool yf,
                                       float &sxx_x, float &sxy_x, float &sxy_y, float &syy_y)
{
    if (xf) {
        const int h = hx(g, i), o = j * g.wx + h;
        float a = PRF(g, m, SOFI_P_B_X_HALF, h) * f.psi_sxx_x[o] + PRF(g, m, SOFI_P_A_X_HALF, h) * sxx_x;
        f.psi_sxx_x[o] = a;
        sxx_x = sxx_x / PRF(g, m, SOFI_P_K_X_HALF, h) + a;
        float b = PRF(g, m, SOFI_P_B_X, h) * f.psi_sxy_x[o] + PRF(g, m, SOFI_P_A_X, h) * sxy_x;
        f.psi_sxy_x[o] = b;
        sxy_x = sxy_x / PRF(g, m, SOFI_P_K_X, h) + b;
    }
    if (yf) {
        const int h = hy(g, j), o = h * g.wy + i;
        float a = PRF(g, m, SOFI_P_B_Y_HALF, h) * f.psi_syy_y[o] + PRF(g, m, SOFI_P_A_Y_HALF, h) * syy_y;
        f.psi_syy_y[o] = a;
        syy_y = syy_y / PRF(g, m, SOFI_P_K_Y_HALF, h) + a;
        float b = PRF(g, m, SOFI_P_B_Y, h) * f.psi_sxy_y[o] + PRF(g, m, SOFI_P_A_Y, h) * sxy_y;
        f.psi_sxy_y[o] = b;
        sxy_y = sxy_y / PRF(g, m, SOFI_P_K_Y, h) + b;
    }
}

__device__ __forceinline__ void velocity_core(const Grid &g, Fields &f, const Mat &m, const Src &s, const Fuse &fz,
                                              int nt, int i, int j, int c, float sxx_x, float sxy_x, float sxy_y,
                                              float syy_y, float rip, float rjp, float &vx, float &vy)
{
    // instability guard of time_loop (src/time_loop.c:44-47), before this step's update
    if (i == g.nx / 2 && j == g.ny / 2 && isnan(vy) && *f.unstable == 0) *f.unstable = step_of(fz.ntb, nt);
    const bool xf = in_xframe(g, i), yf = in_yframe(g, j), frame = xf || yf;
    if (frame && g.abs_type != 1 && g.abs_type != 2) {
        // a frame that is never updated still receives its forces (the
        // reference adds them to every source cell first)
        for (int k = 0; k < fz.nforce; ++k)
            if (c == fz.c[k]) force_add(g, s, fz.soff + k, step_of(fz.ntb, nt), rip, rjp, vx, vy);
        return;
    }
    if (g.abs_type == 1 && frame) cpml_v(g, f, m, i, j, xf, yf, sxx_x, sxy_x, sxy_y, syy_y);
    for (int k = 0; k < fz.nforce; ++k)
        if (c == fz.c[k]) force_add(g, s, fz.soff + k, step_of(fz.ntb, nt), rip, rjp, vx, vy);
    if (fz.fwi_acc == 1) {
        f.pvxp1[c] = rip * (sxx_x + sxy_y) / g.dh;
        f.pvyp1[c] = rjp * (sxy_x + syy_y) / g.dh;
    } else if (fz.fwi_acc == 2) {
        f.pvxp1[c] += vx * g.dt;
        f.pvyp1[c] += vy * g.dt;
    } else if (fz.fwi_acc == 3) {
        f.pvxp1[c] = vx;
        f.pvyp1[c] = vy;
    }
    vx = vx + (sxx_x + sxy_y) * g.dtdh * rip;
    vy = vy + (sxy_x + syy_y) * g.dtdh * rjp;
    if (g.abs_type == 2 && frame) {
        const float a = m.absorb[c];
        vx *= a;
        vy *= a;
    }
}

// In the one-shot kernels (ONE), arrays that a kernel touches only at the cell
// itself (vx, vy, rip, rjp here; the stresses, memory variables and
// coefficients in k_stress) are read and written with nontemporal hints, so
// that they stream past the L2 and leave it to the stencil arrays' rows, which
// neighbouring tiles read again (viscoelastic FD8 on 4000 x 2000: 9 % less time
// per step, stress kernel 72 -> 77 % of 8 TB/s).  The vector-shape kernels of
// the elastic rheologies ran slower with the same hints (-17 % for a batch of
// EL_VTI shots on 2000 x 1000) and keep plain accesses, and so does the
// velocity kernel of the TTI rheologies (NT false: the TTI stress kernel
// reads vx, vy right after for its derivatives; with the hints one EL_TTI
// shot on 2000 x 1000 ran 11 % slower).
//
// one thread per cell; blockIdx.y picks g.spt shot slots of a batch, which
// the thread updates in turn with the cell's coefficients read once (ONE: the
// single slot of a one-shot handle, without the slot loop, whose run-time
// trip count and per-slot offsets cost the one-shot kernels 10-45 % on
// MI355X)
// tile t of BX x TBY cells (tiles row by row, g.ntx to a row)
template <int ND, bool ONE, bool NT, int TBY = BY>
__device__ __forceinline__ void velocity_tile(const Grid &g, const Fields &f0, const Mat &m, const Src &s,
                                              const Fuse &fz, int nt, int t)
{
    const int ty = t / g.ntx, tx = t - ty * g.ntx;
    const int i = 1 + tx * BX + threadIdx.x;
    const int j = 1 + ty * TBY + threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    const int c = IDX(g, i, j);
    const float rip = ld1<NT>(m.rip, c), rjp = ld1<NT>(m.rjp, c);
    const int k0 = ONE ? 0 : g.k_lo + blockIdx.y * g.spt, k1 = ONE ? 1 : min(k0 + g.spt, g.k_hi);
    for (int k = k0; k < k1; ++k) {
        Fields f = ONE ? f0 : f0.at(k, g);
        float sxx_x, sxy_x, sxy_y, syy_y;
        op_v<ND>(g, f.sxx, f.syy, f.sxy, c, sxx_x, sxy_x, sxy_y, syy_y);
        float vx = ld1<NT>(f.vx, c), vy = ld1<NT>(f.vy, c);
        velocity_core(g, f, m, s, fz, nt, i, j, c, sxx_x, sxy_x, sxy_y, syy_y, rip, rjp, vx, vy);
        st1<NT>(f.vx, c, vx);
        st1<NT>(f.vy, c, vy);
    }
}
template <int ND, bool ONE, bool NT = ONE>
__global__ __launch_bounds__(BX *BY) void k_velocity(Grid g, Fields f0, Mat m, Src s, Fuse fz, int nt)
{
    velocity_tile<ND, ONE, NT>(g, f0, m, s, fz, nt, tile_id());
}

// The one-shot velocity update (one slot, one thread per cell) with R
// vertically adjacent cells per thread, in tiles of BX x R BY cells: the y
// stencils of the R cells share their sxy and syy rows, 2 ND + R - 1 loads of
// each instead of 2 ND R.  Each cell's sums are op_v's, in its order.  With R
// = 2 the velocity kernel of the viscoelastic FD8 headline workload moves 6 %
// more bytes per second than with one cell per thread (R = 3 and 4 gain less;
// the same scheme made the stress kernel, 13 VGPRs larger, 12 % slower, and
// one EL_TTI shot on 2000 x 1000 FD6 5 % slower, whose velocity update keeps
// one cell per thread).
constexpr int VROWS = 2;
template <int ND, bool NT, int R>
__global__ __launch_bounds__(BX *BY) void k_velocity_r(Grid g, Fields f, Mat m, Src s, Fuse fz, int nt)
{
    const int t = tile_id(), ty = t / g.ntx, tx = t - ty * g.ntx;
    const int i = 1 + tx * BX + threadIdx.x;
    const int j = 1 + ty * R * BY + R * threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    const int nr = min(R, g.ny - j + 1);
    const int c = IDX(g, i, j), P = g.pitch;
    // rows j - ND .. j + ND + R - 2 of sxy and j - ND + 1 .. j + ND + R - 1 of syy
    float qy[2 * ND + R - 1], qs[2 * ND + R - 1];
#pragma unroll
    for (int r = 0; r < 2 * ND + R - 1; ++r) {
        const bool need = r < 2 * ND + nr - 1;
        qy[r] = need ? f.sxy[c + (r - ND) * P] : 0.f;
        qs[r] = need ? f.syy[c + (r - ND + 1) * P] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
        if (q >= nr) break;
        const int cq = c + q * P;
        float a = 0.f, b = 0.f, d = 0.f, e = 0.f;
#pragma unroll
        for (int mm = 1; mm <= ND; ++mm) {
            const float h = g.hc[mm];
            a += h * (f.sxx[cq + mm] - f.sxx[cq - mm + 1]);
            b += h * (f.sxy[cq + mm - 1] - f.sxy[cq - mm]);
            d += h * (qy[ND + q + mm - 1] - qy[ND + q - mm]);
            e += h * (qs[ND - 1 + q + mm] - qs[ND + q - mm]);
        }
        const float rip = ld1<NT>(m.rip, cq), rjp = ld1<NT>(m.rjp, cq);
        float vx = ld1<NT>(f.vx, cq), vy = ld1<NT>(f.vy, cq);
        velocity_core(g, f, m, s, fz, nt, i, j + q, cq, a, b, d, e, rip, rjp, vx, vy);
        st1<NT>(f.vx, cq, vx);
        st1<NT>(f.vy, cq, vy);
    }
}

// VW cells per thread: the x stencils come from 2G+1 vectors of the row,
// the y stencils from one vector per row; cells past NX keep their (zero)
// padding values
template <int ND, int VW, bool PRE>
__device__ __forceinline__ void velocity_v_slot(const Grid &g, Fields f, const Mat &m, const Src &s, const Fuse &fz,
                                                int nt, int i0, int j, const vecf<VW> *pre)
{
    constexpr int G = (ND + VW - 1) / VW;
    using V = vecf<VW>;
    const int c0 = IDX(g, i0, j), P = g.pitch;
    float xs[(2 * G + 1) * VW], xy[(2 * G + 1) * VW];   // sxx, sxy of row j: columns i0 - G VW .. i0 + (G + 1) VW - 1
#pragma unroll
    for (int q = -G; q <= G; ++q) {
        const V a = ldv<VW>(f.sxx, c0 + VW * q), b = ldv<VW>(f.sxy, c0 + VW * q);
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            xs[(q + G) * VW + k] = a[k];
            xy[(q + G) * VW + k] = b[k];
        }
    }
    float d[VW] = {}, e[VW] = {};
#pragma unroll
    for (int mm = 1; mm <= ND; ++mm) {
        const float h = g.hc[mm];
        // sxy rows j + m - 1 and j - m, syy rows j + m and j - m + 1
        V a1;
        if (mm == 1) {
#pragma unroll
            for (int k = 0; k < VW; ++k) a1[k] = xy[G * VW + k];
        } else {
            a1 = ldv<VW>(f.sxy, c0 + (mm - 1) * P);
        }
        const V a0 = ldv<VW>(f.sxy, c0 - mm * P);
        const V b1 = ldv<VW>(f.syy, c0 + mm * P), b0 = ldv<VW>(f.syy, c0 - (mm - 1) * P);
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            d[k] += h * (a1[k] - a0[k]);
            e[k] += h * (b1[k] - b0[k]);
        }
    }
    V vx = ldv<VW>(f.vx, c0), vy = ldv<VW>(f.vy, c0);
    const V rip = PRE ? pre[0] : ldv<VW>(m.rip, c0), rjp = PRE ? pre[1] : ldv<VW>(m.rjp, c0);
#pragma unroll
    for (int k = 0; k < VW; ++k) {
        if (i0 + k > g.nx) break;
        float a = 0.f, b = 0.f;
#pragma unroll
        for (int mm = 1; mm <= ND; ++mm) {
            const float h = g.hc[mm];
            a += h * (xs[G * VW + k + mm] - xs[G * VW + k - mm + 1]);
            b += h * (xy[G * VW + k + mm - 1] - xy[G * VW + k - mm]);
        }
        float x = vx[k], y = vy[k];
        velocity_core(g, f, m, s, fz, nt, i0 + k, j, c0 + k, a, b, d[k], e[k], rip[k], rjp[k], x, y);
        vx[k] = x;
        vy[k] = y;
    }
    stv<VW>(f.vx, c0, vx);
    stv<VW>(f.vy, c0, vy);
}

// one slot (ONE: a one-shot handle) or blockIdx.y's g.spt slots of a batch
template <int ND, int VW, bool ONE>
__global__ __launch_bounds__(BX *BY) void k_velocity_v(Grid g, Fields f0, Mat m, Src s, Fuse fz, int nt)
{
    const int ntxv = (g.nx + BX * VW - 1) / (BX * VW);
    const int t = tile_id(), ty = t / ntxv, tx = t - ty * ntxv;
    const int i0 = 1 + (tx * BX + threadIdx.x) * VW;
    const int j = 1 + ty * BY + threadIdx.y;
    if (i0 > g.nx || j > g.ny) return;
    if (ONE) {
        velocity_v_slot<ND, VW, false>(g, f0, m, s, fz, nt, i0, j, nullptr);
        return;
    }
    // a batch: the cells' coefficients once for the thread's slots
    const int c0 = IDX(g, i0, j);
    const vecf<VW> pre[2] = { ldv<VW>(m.rip, c0), ldv<VW>(m.rjp, c0) };
    const int k0 = g.k_lo + blockIdx.y * g.spt, k1 = min(k0 + g.spt, g.k_hi);
    for (int k = k0; k < k1; ++k) velocity_v_slot<ND, VW, true>(g, f0.at(k, g), m, s, fz, nt, i0, j, pre);
}

// --------------------------------------------------------- stress updates

// Per-cell rheology updates on registers: (sxx, syy, sxy) hold the old
// stresses on entry and the updated ones on exit; dv are the (possibly
// CPML-modified) derivatives.

// MODE=FWI: the stress increments over DT (ux, uy, uxy) that every
// rheology's update writes (src/wavefield_update_s_el.c:43-47,
// src/wavefield_update_s_visc.c:61-66, src/wavefield_update_s_el_vti.c:42-46,
// src/wavefield_update_s_visc_VTI.c:61-65, src/wavefield_update_s_el_tti.c:47-51,
// src/wavefield_update_s_visc_TTI.c:79-83); only the forward store steps
// with VELOCITY = 1 read them (src/snap_store.c:35-43), so they are written
// in those steps only
__device__ __forceinline__ void fwi_u_store(const Fuse &fz, const Fields &f, float dt, int c, float u1, float u2,
                                            float u3)
{
    if (fz.fwi_u) {
        f.uxy[c] = u1 / dt;
        f.ux[c] = u2 / dt;
        f.uy[c] = u3 / dt;
    }
}

// wavefield_update_s_el (src/wavefield_update_s_el.c:28-48)
__device__ __forceinline__ void upd_el(float f_, float g_, float fipjp, float vxx, float vyx, float vxy, float vyy,
                                       float &sxx, float &syy, float &sxy, const Fuse &fz, const Fields &f, float dt, int c)
{
    const float u1 = fipjp * (vyx + vxy);
    const float ff = 2.0f * f_;
    const float u2 = (g_ * (vxx + vyy)) - (ff * vyy);
    const float u3 = (g_ * (vxx + vyy)) - (ff * vxx);
    sxy += u1;
    sxx += u2;
    syy += u3;
    fwi_u_store(fz, f, dt, c, u1, u2, u3);
}

// wavefield_update_s_el_vti (src/wavefield_update_s_el_vti.c:27-47)
__device__ __forceinline__ void upd_el_vti(float c11, float c13, float c33, float c55i, float vxx, float vyx,
                                           float vxy, float vyy, float &sxx, float &syy, float &sxy, const Fuse &fz,
                                           const Fields &f, float dt, int c)
{
    const float u1 = c55i * (vxy + vyx);
    const float u2 = (c11 * vxx) + (c13 * vyy);
    const float u3 = (c13 * vxx) + (c33 * vyy);
    sxy += u1;
    sxx += u2;
    syy += u3;
    fwi_u_store(fz, f, dt, c, u1, u2, u3);
}

// Memory-variable sums of the viscoelastic updates: old (r0, p0, q0) and new
// (r, p, q) summed over the relaxation mechanisms in their order.
struct VSum {
    float r0 = 0.f, p0 = 0.f, q0 = 0.f, r = 0.f, p = 0.f, q = 0.f;
};

// mechanism l of wavefield_update_s_visc (src/wavefield_update_s_visc.c:27-68);
// the defect coefficients dip/d/e[l] = eta_l * (per-cell base) are formed in registers
__device__ __forceinline__ void visc_level(float eta, float bip, float cip, float br, float bd, float be, float sxyv,
                                           float sdiag, float vxx, float vyy, float &r, float &p, float &q, VSum &u)
{
    u.r0 += r;
    u.p0 += p;
    u.q0 += q;
    const float dip = eta * br, d2 = 2.0f * (eta * bd), e = eta * be;
    r = bip * (r * cip - dip * sxyv);
    p = bip * (p * cip - e * sdiag + d2 * vyy);
    q = bip * (q * cip - e * sdiag + d2 * vxx);
    u.r += r;
    u.p += p;
    u.q += q;
}

__device__ __forceinline__ void visc_final(float dthalbe, float f_, float g_, float fipjp, float sxyv, float sdiag,
                                           float vxx, float vyy, const VSum &u, float &sxx, float &syy, float &sxy,
                                           const Fuse &fz, const Fields &f, float dt, int c)
{
    const float ff = 2.0f * f_;
    const float u1 = (fipjp * sxyv) + (dthalbe * (u.r0 + u.r));
    const float u2 = (g_ * sdiag) - (ff * vyy) + (dthalbe * (u.p0 + u.p));
    const float u3 = (g_ * sdiag) - (ff * vxx) + (dthalbe * (u.q0 + u.q));
    sxy += u1;
    sxx += u2;
    syy += u3;
    fwi_u_store(fz, f, dt, c, u1, u2, u3);
}

// mechanism l of wavefield_update_s_visc_VTI (src/wavefield_update_s_visc_VTI.c:27-66)
__device__ __forceinline__ void visc_vti_level(float bip, float cip, float d11, float d13, float d33, float d55i,
                                               float sxyv, float vxx, float vyy, float &r, float &p, float &q,
                                               VSum &u)
{
    u.r0 += r;
    u.p0 += p;
    u.q0 += q;
    r = bip * (r * cip - (d55i * sxyv));
    p = bip * (p * cip - (d11 * vxx) - (d13 * vyy));
    q = bip * (q * cip - (d13 * vxx) - (d33 * vyy));
    u.r += r;
    u.p += p;
    u.q += q;
}

__device__ __forceinline__ void visc_vti_final(float dthalbe, float u11, float u13, float u33, float u55i, float sxyv,
                                               float vxx, float vyy, const VSum &u, float &sxx, float &syy,
                                               float &sxy, const Fuse &fz, const Fields &f, float dt, int c)
{
    const float u1 = (u55i * sxyv) + (dthalbe * (u.r0 + u.r));
    const float u2 = (u11 * vxx) + (u13 * vyy) + (dthalbe * (u.p0 + u.p));
    const float u3 = (u13 * vxx) + (u33 * vyy) + (dthalbe * (u.q0 + u.q));
    sxy += u1;
    sxx += u2;
    syy += u3;
    fwi_u_store(fz, f, dt, c, u1, u2, u3);
}

// Derivatives of a cell from the raw op_s sums (no 1/DH): scaled, and CPML in
// the frame (cpml_update_s_x/_y).  Returns false for a frame cell that the
// reference never updates (ABS_TYPE other than 1 or 2).
__device__ __forceinline__ bool stress_dv(const Grid &g, Fields &f, const Mat &m, int i, int j, float a, float b,
                                          float d, float e, float &vxx, float &vyx, float &vxy, float &vyy)
{
    const bool xf = in_xframe(g, i), yf = in_yframe(g, j), frame = xf || yf;
    if (frame && g.abs_type != 1 && g.abs_type != 2) return false;
    vxx = a * g.dhi;
    vyx = b * g.dhi;
    vxy = d * g.dhi;
    vyy = e * g.dhi;
    if (g.abs_type == 1 && frame) cpml_dv(g, m, f, i, j, xf, yf, vxx, vyx, vxy, vyy);
    return true;
}

// What follows the rheology update for one cell: damping in the frame
// (abs_update_s, src/abs_update.c:39-44), then what the reference does next
// to the same cell when it is fused here: the explosive sources (psource) and
// the free surface (imaging into the ghost rows by the cells that are imaged,
// correction of row 1).
template <int ND, bool VISCO>
__device__ __forceinline__ void stress_post(const Grid &g, Fields &f, const Mat &m, const Src &s, const Fuse &fz,
                                            int nt, int i, int j, int c, float &sxx, float &syy, float &sxy)
{
    if (g.abs_type == 2 && (in_xframe(g, i) || in_yframe(g, j))) {
        const float q = m.absorb[c];
        sxy *= q;
        syy *= q;
        sxx *= q;
    }
    for (int k = 0; k < fz.npsrc; ++k)
        if (c == fz.c[k]) {
            const float amp = psource_amp(g, s, fz.soff + k, step_of(fz.ntb, nt));
            sxx += amp;
            syy += amp;
        }
    if (fz.surf && j <= ND + 1) {
        const int P = g.pitch;
        if (j >= 2) f.syy[c + (2 - 2 * j) * P] = -syy;   // syy(1 - m) = -syy(1 + m)
        if (j <= ND) f.sxy[c + (1 - 2 * j) * P] = -sxy;  // sxy(1 - m) = -sxy(m)
        if (j == 1) {
            syy = 0.0f;
            sxx = surface_sxx<ND, VISCO>(g, f, m, i, c, sxx);
        }
    }
}

// one thread per cell; the velocity derivatives are recomputed from vx, vy
// (replaces v_derivatives + exchange_v) (src/update_s_*_{interior,PML,abs}.c).
// blockIdx.y picks g.spt shot slots of a batch, updated in turn with the
// cell's material coefficients read once (ONE: a one-shot handle's single
// slot, as in k_velocity).
template <int ND, int WEQ, bool ONE, int TBY = BY>
__device__ __forceinline__ void stress_tile(const Grid &g, const Fields &f0, const Mat &m, const Src &s,
                                            const Fuse &fz, int nt, int t)
{
    constexpr bool VISCO = WEQ == SOFI_VEL_ISO || WEQ == SOFI_VEL_VTI;
    const int ty = t / g.ntx, tx = t - ty * g.ntx;
    const int i = 1 + tx * BX + threadIdx.x;
    const int j = 1 + ty * TBY + threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    const int c = IDX(g, i, j), P = g.pitch;
    // the cell's coefficients: EL_ISO f g fipjp; VEL_ISO + defect bases;
    // EL_VTI c11 c13 c33 c55ipjp; VEL_VTI the unrelaxed ones (defects per level below)
    float c1 = 0.f, c2 = 0.f, c3 = 0.f, c4 = 0.f, br = 0.f, bd = 0.f, be = 0.f;
    if (WEQ == SOFI_EL_ISO || WEQ == SOFI_VEL_ISO) {
        c1 = ld1<ONE>(m.f, c);
        c2 = ld1<ONE>(m.g, c);
        c3 = ld1<ONE>(m.fipjp, c);
    }
    if (WEQ == SOFI_VEL_ISO) {
        br = ld1<ONE>(m.br, c);
        bd = ld1<ONE>(m.bd, c);
        be = ld1<ONE>(m.be, c);
    }
    if (WEQ == SOFI_EL_VTI) {
        c1 = m.c11[c];
        c2 = m.c13[c];
        c3 = m.c33[c];
        c4 = m.c55i[c];
    }
    if (WEQ == SOFI_VEL_VTI) {
        c1 = m.u11[c];
        c2 = m.u13[c];
        c3 = m.u33[c];
        c4 = m.u55i[c];
    }
    const int k0 = ONE ? 0 : g.k_lo + blockIdx.y * g.spt, k1 = ONE ? 1 : min(k0 + g.spt, g.k_hi);
    for (int k = k0; k < k1; ++k) {
        Fields f = ONE ? f0 : f0.at(k, g);
        float a = 0.f, b = 0.f, d = 0.f, e = 0.f;
#pragma unroll
        for (int mm = 1; mm <= ND; ++mm) {
            const float h = g.hc[mm];
            a += h * (f.vx[c + mm - 1] - f.vx[c - mm]);
            b += h * (f.vy[c + mm] - f.vy[c - mm + 1]);
            d += h * (f.vx[c + mm * P] - f.vx[c - (mm - 1) * P]);
            e += h * (f.vy[c + (mm - 1) * P] - f.vy[c - mm * P]);
        }
        float vxx, vyx, vxy, vyy;
        if (!stress_dv(g, f, m, i, j, a, b, d, e, vxx, vyx, vxy, vyy)) continue;
        float sxx = ld1<ONE>(f.sxx, c), syy = ld1<ONE>(f.syy, c),
              sxy = ld1<ONE>(f.sxy, c);
        if (WEQ == SOFI_EL_ISO) upd_el(c1, c2, c3, vxx, vyx, vxy, vyy, sxx, syy, sxy, fz, f, g.dt, c);
        if (WEQ == SOFI_EL_VTI) upd_el_vti(c1, c2, c3, c4, vxx, vyx, vxy, vyy, sxx, syy, sxy, fz, f, g.dt, c);
        if (WEQ == SOFI_VEL_ISO) {
            const float sxyv = vxy + vyx, sdiag = vxx + vyy;
            VSum u;
#pragma unroll
            for (int l = 1; l <= SOFI_MAXL; ++l) {
                if (l > g.L) break;
                const int o = (l - 1) * g.plane + c;
                float r = ld1<ONE>(f.r, o), p = ld1<ONE>(f.p, o),
                      q = ld1<ONE>(f.q, o);
                visc_level(m.eta[l], m.bip[l], m.cip[l], br, bd, be, sxyv, sdiag, vxx, vyy, r, p, q, u);
                st1<ONE>(f.r, o, r);
                st1<ONE>(f.p, o, p);
                st1<ONE>(f.q, o, q);
            }
            visc_final(g.dthalbe, c1, c2, c3, sxyv, sdiag, vxx, vyy, u, sxx, syy, sxy, fz, f, g.dt, c);
        }
        if (WEQ == SOFI_VEL_VTI) {
            const float sxyv = vxy + vyx;
            VSum u;
#pragma unroll
            for (int l = 1; l <= SOFI_MAXL; ++l) {
                if (l > g.L) break;
                const int o = (l - 1) * g.plane + c;
                float r = f.r[o], p = f.p[o], q = f.q[o];
                visc_vti_level(m.bip[l], m.cip[l], m.d11[o], m.d13[o], m.d33[o], m.d55i[o], sxyv, vxx, vyy, r, p, q, u);
                f.r[o] = r;
                f.p[o] = p;
                f.q[o] = q;
            }
            visc_vti_final(g.dthalbe, c1, c2, c3, c4, sxyv, vxx, vyy, u, sxx, syy, sxy, fz, f, g.dt, c);
        }
        stress_post<ND, VISCO>(g, f, m, s, fz, nt, i, j, c, sxx, syy, sxy);
        st1<ONE>(f.sxx, c, sxx);
        st1<ONE>(f.syy, c, syy);
        st1<ONE>(f.sxy, c, sxy);
    }
}
template <int ND, int WEQ, bool ONE>
__global__ __launch_bounds__(BX *BY) void k_stress(Grid g, Fields f0, Mat m, Src s, Fuse fz, int nt)
{
    if (ONE && fused_seis(g, f0, fz)) return;
    stress_tile<ND, WEQ, ONE>(g, f0, m, s, fz, nt, tile_id());
}

// VW cells per thread (see k_velocity_v); inactive cells (past NX, or
// frame cells that are never updated) keep their stresses and memory
// variables
template <int ND, int WEQ, int VW, bool PRE>
__device__ __forceinline__ void stress_v_slot(const Grid &g, Fields f, const Mat &m, const Src &s, const Fuse &fz,
                                              int nt, int i0, int j, const vecf<VW> *pre)
{
    constexpr bool VISCO = WEQ == SOFI_VEL_ISO || WEQ == SOFI_VEL_VTI;
    constexpr int G = (ND + VW - 1) / VW;
    using V = vecf<VW>;
    const int c0 = IDX(g, i0, j), P = g.pitch;
    float xv[(2 * G + 1) * VW], yv[(2 * G + 1) * VW];   // vx, vy of row j: columns i0 - G VW .. i0 + (G + 1) VW - 1
#pragma unroll
    for (int q = -G; q <= G; ++q) {
        const V a = ldv<VW>(f.vx, c0 + VW * q), b = ldv<VW>(f.vy, c0 + VW * q);
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            xv[(q + G) * VW + k] = a[k];
            yv[(q + G) * VW + k] = b[k];
        }
    }
    float d[VW] = {}, e[VW] = {};
#pragma unroll
    for (int mm = 1; mm <= ND; ++mm) {
        const float h = g.hc[mm];
        // vx rows j + m and j - m + 1, vy rows j + m - 1 and j - m
        const V a1 = ldv<VW>(f.vx, c0 + mm * P);
        V a0, b1;
        if (mm == 1) {
#pragma unroll
            for (int k = 0; k < VW; ++k) {
                a0[k] = xv[G * VW + k];
                b1[k] = yv[G * VW + k];
            }
        } else {
            a0 = ldv<VW>(f.vx, c0 - (mm - 1) * P);
            b1 = ldv<VW>(f.vy, c0 + (mm - 1) * P);
        }
        const V b0 = ldv<VW>(f.vy, c0 - mm * P);
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            d[k] += h * (a1[k] - a0[k]);
            e[k] += h * (b1[k] - b0[k]);
        }
    }
    float vxx[VW], vyx[VW], vxy[VW], vyy[VW];
    bool act[VW];
#pragma unroll
    for (int k = 0; k < VW; ++k) {
        act[k] = false;
        vxx[k] = vyx[k] = vxy[k] = vyy[k] = 0.f;
        if (i0 + k > g.nx) continue;
        float a = 0.f, b = 0.f;
#pragma unroll
        for (int mm = 1; mm <= ND; ++mm) {
            const float h = g.hc[mm];
            a += h * (xv[G * VW + k + mm - 1] - xv[G * VW + k - mm]);
            b += h * (yv[G * VW + k + mm] - yv[G * VW + k - mm + 1]);
        }
        act[k] = stress_dv(g, f, m, i0 + k, j, a, b, d[k], e[k], vxx[k], vyx[k], vxy[k], vyy[k]);
    }
    const V sxx0 = ldv<VW>(f.sxx, c0), syy0 = ldv<VW>(f.syy, c0), sxy0 = ldv<VW>(f.sxy, c0);
    float sxx[VW], syy[VW], sxy[VW];
#pragma unroll
    for (int k = 0; k < VW; ++k) {
        sxx[k] = sxx0[k];
        syy[k] = syy0[k];
        sxy[k] = sxy0[k];
    }
    if (WEQ == SOFI_EL_ISO) {
        const V c1 = PRE ? pre[0] : ldv<VW>(m.f, c0), c2 = PRE ? pre[1] : ldv<VW>(m.g, c0),
                c3 = PRE ? pre[2] : ldv<VW>(m.fipjp, c0);
#pragma unroll
        for (int k = 0; k < VW; ++k)
            if (act[k]) upd_el(c1[k], c2[k], c3[k], vxx[k], vyx[k], vxy[k], vyy[k], sxx[k], syy[k], sxy[k], fz, f, g.dt, c0 + k);
    }
    if (WEQ == SOFI_EL_VTI) {
        const V c1 = PRE ? pre[0] : ldv<VW>(m.c11, c0), c2 = PRE ? pre[1] : ldv<VW>(m.c13, c0),
                c3 = PRE ? pre[2] : ldv<VW>(m.c33, c0), c4 = PRE ? pre[3] : ldv<VW>(m.c55i, c0);
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            Fuse fk = fz;
            fk.fwi_u = fz.fwi_u && act[k];
            upd_el_vti(c1[k], c2[k], c3[k], c4[k], vxx[k], vyx[k], vxy[k], vyy[k], sxx[k], syy[k], sxy[k], fk, f, g.dt,
                       c0 + k);
        }
    }
    if (VISCO) {
        VSum u[VW];
        float sxyv[VW], sdiag[VW];
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            sxyv[k] = vxy[k] + vyx[k];
            sdiag[k] = vxx[k] + vyy[k];
        }
        V br, bd, be;
        if (WEQ == SOFI_VEL_ISO) {
            br = ldv<VW>(m.br, c0);
            bd = ldv<VW>(m.bd, c0);
            be = ldv<VW>(m.be, c0);
        }
#pragma unroll
        for (int l = 1; l <= SOFI_MAXL; ++l) {
            if (l > g.L) break;
            const int o = (l - 1) * g.plane + c0;
            V r = ldv<VW>(f.r, o), p = ldv<VW>(f.p, o), q = ldv<VW>(f.q, o);
            V d11, d13, d33, d55;
            if (WEQ == SOFI_VEL_VTI) {
                d11 = ldv<VW>(m.d11, o);
                d13 = ldv<VW>(m.d13, o);
                d33 = ldv<VW>(m.d33, o);
                d55 = ldv<VW>(m.d55i, o);
            }
#pragma unroll
            for (int k = 0; k < VW; ++k) {
                float rr = r[k], pp = p[k], qq = q[k];
                if (WEQ == SOFI_VEL_ISO)
                    visc_level(m.eta[l], m.bip[l], m.cip[l], br[k], bd[k], be[k], sxyv[k], sdiag[k], vxx[k], vyy[k],
                               rr, pp, qq, u[k]);
                else
                    visc_vti_level(m.bip[l], m.cip[l], d11[k], d13[k], d33[k], d55[k], sxyv[k], vxx[k], vyy[k], rr,
                                   pp, qq, u[k]);
                if (act[k]) {
                    r[k] = rr;
                    p[k] = pp;
                    q[k] = qq;
                }
            }
            stv<VW>(f.r, o, r);
            stv<VW>(f.p, o, p);
            stv<VW>(f.q, o, q);
        }
        if (WEQ == SOFI_VEL_ISO) {
            const V c1 = ldv<VW>(m.f, c0), c2 = ldv<VW>(m.g, c0), c3 = ldv<VW>(m.fipjp, c0);
#pragma unroll
            for (int k = 0; k < VW; ++k) {
                Fuse fk = fz;
                fk.fwi_u = fz.fwi_u && act[k];
                visc_final(g.dthalbe, c1[k], c2[k], c3[k], sxyv[k], sdiag[k], vxx[k], vyy[k], u[k], sxx[k], syy[k],
                           sxy[k], fk, f, g.dt, c0 + k);
            }
        } else {
            const V c1 = ldv<VW>(m.u11, c0), c2 = ldv<VW>(m.u13, c0), c3 = ldv<VW>(m.u33, c0), c4 = ldv<VW>(m.u55i, c0);
#pragma unroll
            for (int k = 0; k < VW; ++k) {
                Fuse fk = fz;
                fk.fwi_u = fz.fwi_u && act[k];
                visc_vti_final(g.dthalbe, c1[k], c2[k], c3[k], c4[k], sxyv[k], vxx[k], vyy[k], u[k], sxx[k], syy[k],
                               sxy[k], fk, f, g.dt, c0 + k);
            }
        }
    }
    V sxxn = sxx0, syyn = syy0, sxyn = sxy0;
#pragma unroll
    for (int k = 0; k < VW; ++k) {
        if (!act[k]) continue;
        stress_post<ND, VISCO>(g, f, m, s, fz, nt, i0 + k, j, c0 + k, sxx[k], syy[k], sxy[k]);
        sxxn[k] = sxx[k];
        syyn[k] = syy[k];
        sxyn[k] = sxy[k];
    }
    stv<VW>(f.sxx, c0, sxxn);
    stv<VW>(f.syy, c0, syyn);
    stv<VW>(f.sxy, c0, sxyn);
}

// one slot (ONE: a one-shot handle) or blockIdx.y's g.spt slots of a batch
template <int ND, int WEQ, int VW, bool ONE>
__global__ __launch_bounds__(BX *BY) void k_stress_v(Grid g, Fields f0, Mat m, Src s, Fuse fz, int nt)
{
    if (ONE && fused_seis(g, f0, fz)) return;
    const int ntxv = (g.nx + BX * VW - 1) / (BX * VW);
    const int t = tile_id(), ty = t / ntxv, tx = t - ty * ntxv;
    const int i0 = 1 + (tx * BX + threadIdx.x) * VW;
    const int j = 1 + ty * BY + threadIdx.y;
    if (i0 > g.nx || j > g.ny) return;
    if (ONE) {
        stress_v_slot<ND, WEQ, VW, false>(g, f0, m, s, fz, nt, i0, j, nullptr);
        return;
    }
    // a batch: the elastic coefficients once for the thread's slots
    const int c0 = IDX(g, i0, j);
    vecf<VW> pre[4];
    if (WEQ == SOFI_EL_ISO) {
        pre[0] = ldv<VW>(m.f, c0);
        pre[1] = ldv<VW>(m.g, c0);
        pre[2] = ldv<VW>(m.fipjp, c0);
    }
    if (WEQ == SOFI_EL_VTI) {
        pre[0] = ldv<VW>(m.c11, c0);
        pre[1] = ldv<VW>(m.c13, c0);
        pre[2] = ldv<VW>(m.c33, c0);
        pre[3] = ldv<VW>(m.c55i, c0);
    }
    constexpr bool P = WEQ == SOFI_EL_ISO || WEQ == SOFI_EL_VTI;
    const int k0 = g.k_lo + blockIdx.y * g.spt, k1 = min(k0 + g.spt, g.k_hi);
    for (int k = k0; k < k1; ++k) stress_v_slot<ND, WEQ, VW, P>(g, f0.at(k, g), m, s, fz, nt, i0, j, pre);
}

// ------------------------------------ 4th order in time (Adams-Bashforth)

// the reference's weights, float constants (src/wavefield_update_v_4.c:29-33)
constexpr float AB1 = (float)(13.0 / 12.0), AB2 = (float)(-5.0 / 24.0), AB3 = (float)(1.0 / 6.0),
                AB4 = (float)(-1.0 / 24.0);

// The histories of the 4th-order loop, [0] the newest: the velocity
// kernel writes sx/sy[0], the stress kernel dv[.][0] and the memory
// variables' [3]; the host then rotates the pointers as shift_var2 /
// shift_var3 do (src/util.c:89-109), so that a buffer is reused four steps
// later.
struct Hist4 {
    float *sx[4], *sy[4];          // svx_1..4, svy_1..4: sxx_x + sxy_y, sxy_x + syy_y
    float *dv[4][4];               // vxx, vyy, vxy, vyx times DH
    float *r[4], *p[4], *q[4];     // memory variables, L planes each
};

// update_v_interior_4 + update_v_PML_4 / update_v_abs_4 with
// wavefield_update_v_4 (src/update_v_interior_4.c:32-368,
// src/update_v_PML_4.c:33-136, src/update_v_abs_4.c:32-192,
// src/wavefield_update_v_4.c:25-47), one thread per cell.  Forces first; in
// the CPML frame the left side and the four corners only advance the CPML
// memory and keep their velocities, as the reference's loops do.
template <int ND>
__global__ __launch_bounds__(BX *BY) void k_velocity4(Grid g, Fields f, Mat m, Src s, Fuse fz, int nt, Hist4 H)
{
    const int t = tile_id(), ty = t / g.ntx, tx = t - ty * g.ntx;
    const int i = 1 + tx * BX + threadIdx.x;
    const int j = 1 + ty * BY + threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    const int c = IDX(g, i, j);
    const float rip = m.rip[c], rjp = m.rjp[c];
    float sxx_x, sxy_x, sxy_y, syy_y;
    op_v<ND>(g, f.sxx, f.syy, f.sxy, c, sxx_x, sxy_x, sxy_y, syy_y);
    float vx = f.vx[c], vy = f.vy[c];
    if (i == g.nx / 2 && j == g.ny / 2 && isnan(vy) && *f.unstable == 0) *f.unstable = step_of(fz.ntb, nt);
    for (int k = 0; k < fz.nforce; ++k)
        if (c == fz.c[k]) force_add(g, s, fz.soff + k, step_of(fz.ntb, nt), rip, rjp, vx, vy);
    const bool xf = in_xframe(g, i), yf = in_yframe(g, j), frame = xf || yf;
    bool upd = !frame || g.abs_type == 2;
    if (frame && g.abs_type == 1) {
        cpml_v(g, f, m, i, j, xf, yf, sxx_x, sxy_x, sxy_y, syy_y);
        upd = i > g.gx2 && !(i > g.gx3 && yf);     // right side, top and bottom only
    }
    if (upd) {
        const float sx = sxx_x + sxy_y, sy = sxy_x + syy_y;
        H.sx[0][c] = sx;
        H.sy[0][c] = sy;
        vx += (AB1 * sx + AB2 * H.sx[1][c] + AB3 * H.sx[2][c] + AB4 * H.sx[3][c]) * g.dtdh * rip;
        vy += (AB1 * sy + AB2 * H.sy[1][c] + AB3 * H.sy[2][c] + AB4 * H.sy[3][c]) * g.dtdh * rjp;
        if (frame) {     // ABS_TYPE 2: damping (abs_update_v, src/abs_update.c:46-50)
            if (g.abs_type == 2) {
                const float a = m.absorb[c];
                vx *= a;
                vy *= a;
            }
        }
    }
    f.vx[c] = vx;
    f.vy[c] = vy;
}

// update_s_{elastic,visc}_{interior,PML,abs}_4 with
// wavefield_update_s_el_4 / wavefield_update_s_visc_4
// (src/wavefield_update_s_el_4.c:27-70, src/wavefield_update_s_visc_4.c:27-98),
// one thread per cell: the derivatives (CPML-modified in the frame) times DH
// into the histories, their Adams-Bashforth sums, the memory variables with
// their own histories (written into [3]), the stress update, then damping,
// the fused sources and the free surface as in the 2nd-order kernel, which
// see the memory variables just written.
template <int ND, bool VISCO>
__global__ __launch_bounds__(BX *BY) void k_stress4(Grid g, Fields f, Mat m, Src s, Fuse fz, int nt, Hist4 H)
{
    if (fused_seis(g, f, fz)) return;
    const int t = tile_id(), ty = t / g.ntx, tx = t - ty * g.ntx;
    const int i = 1 + tx * BX + threadIdx.x;
    const int j = 1 + ty * BY + threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    const int c = IDX(g, i, j), P = g.pitch;
    float a = 0.f, b = 0.f, d = 0.f, e = 0.f;
#pragma unroll
    for (int mm = 1; mm <= ND; ++mm) {
        const float h = g.hc[mm];
        a += h * (f.vx[c + mm - 1] - f.vx[c - mm]);
        b += h * (f.vy[c + mm] - f.vy[c - mm + 1]);
        d += h * (f.vx[c + mm * P] - f.vx[c - (mm - 1) * P]);
        e += h * (f.vy[c + (mm - 1) * P] - f.vy[c - mm * P]);
    }
    float vxx, vyx, vxy, vyy;
    if (!stress_dv(g, f, m, i, j, a, b, d, e, vxx, vyx, vxy, vyy)) return;
    const float dhi = g.dhi4;
    const float cur[4] = { vxx * g.dh, vyy * g.dh, vxy * g.dh, vyx * g.dh };
    float sum[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        H.dv[q][0][c] = cur[q];
        sum[q] = AB1 * cur[q] + AB2 * H.dv[q][1][c] + AB3 * H.dv[q][2][c] + AB4 * H.dv[q][3][c];
    }
    const float sumxx = sum[0], sumyy = sum[1], sumxy = sum[2], sumyx = sum[3];
    const float f_ = m.f[c], g_ = m.g[c], fipjp = m.fipjp[c];
    float sxx = f.sxx[c], syy = f.syy[c], sxy = f.sxy[c];
    Fields fo = f;
    float u1, u2, u3;
    if (!VISCO) {
        u1 = fipjp * (sumxy + sumyx) * dhi;
        u2 = ((g_ * (sumxx + sumyy)) - (2.0f * f_ * sumyy)) * dhi;
        u3 = ((g_ * (sumxx + sumyy)) - (2.0f * f_ * sumxx)) * dhi;
    } else {
        fo.r = H.r[3];      // the memory variables after this step
        fo.p = H.p[3];
        fo.q = H.q[3];
        const float br = m.br[c], bd = m.bd[c], be = m.be[c];
        float r_old = 0.f, p_old = 0.f, q_old = 0.f, r_new = 0.f, p_new = 0.f, q_new = 0.f;
#pragma unroll
        for (int l = 1; l <= SOFI_MAXL; ++l) {
            if (l > g.L) break;
            const int o = (l - 1) * g.plane + c;
            const float eta = m.eta[l], bip = m.bip[l], cip = m.cip[l];
            const float dip = eta * br, dd = eta * bd, ee = eta * be;
            const float ctemp = 2.0f * (1.0f - cip) / AB1;
            const float r0 = H.r[0][o], r1 = H.r[1][o], r2 = H.r[2][o], r3 = H.r[3][o];
            const float p0 = H.p[0][o], p1 = H.p[1][o], p2 = H.p[2][o], p3 = H.p[3][o];
            const float q0 = H.q[0][o], q1 = H.q[1][o], q2 = H.q[2][o], q3 = H.q[3][o];
            r_old += AB1 * r0 + AB2 * r1 + AB3 * r2 + AB4 * r3;
            p_old += AB1 * p0 + AB2 * p1 + AB3 * p2 + AB4 * p3;
            q_old += AB1 * q0 + AB2 * q1 + AB3 * q2 + AB4 * q3;
            const float rn = bip * (r0 * cip - ctemp * AB2 * (r0 + r1) - ctemp * AB3 * (r1 + r2) -
                                    ctemp * AB4 * (r2 + r3) - (dip * (sumxy + sumyx) * dhi));
            const float pn = bip * (p0 * cip - ctemp * AB2 * (p0 + p1) - ctemp * AB3 * (p1 + p2) -
                                    ctemp * AB4 * (p2 + p3) - (ee * (sumxx + sumyy) * dhi) + (2.0f * dd * sumyy * dhi));
            const float qn = bip * (q0 * cip - ctemp * AB2 * (q0 + q1) - ctemp * AB3 * (q1 + q2) -
                                    ctemp * AB4 * (q2 + q3) - (ee * (sumxx + sumyy) * dhi) + (2.0f * dd * sumxx * dhi));
            H.r[3][o] = rn;
            H.p[3][o] = pn;
            H.q[3][o] = qn;
            r_new += AB1 * rn + AB2 * r0 + AB3 * r1 + AB4 * r2;
            p_new += AB1 * pn + AB2 * p0 + AB3 * p1 + AB4 * p2;
            q_new += AB1 * qn + AB2 * q0 + AB3 * q1 + AB4 * q2;
        }
        u1 = (fipjp * (sumxy + sumyx)) * dhi + (g.dthalbe * (r_old + r_new));
        u2 = (g_ * (sumxx + sumyy) * dhi) - (2.0f * f_ * sumyy * dhi) + (g.dthalbe * (p_old + p_new));
        u3 = (g_ * (sumxx + sumyy) * dhi) - (2.0f * f_ * sumxx * dhi) + (g.dthalbe * (q_old + q_new));
    }
    sxy += u1;
    sxx += u2;
    syy += u3;
    stress_post<ND, VISCO>(g, fo, m, s, fz, nt, i, j, c, sxx, syy, sxy);
    f.sxx[c] = sxx;
    f.syy[c] = syy;
    f.sxy[c] = sxy;
}

// ------------------------------------------------------------ acoustic

// update_s_acoustic_{interior,PML,abs} with wavefield_update_s_ac
// (src/update_s_acoustic_interior.c:33-53, src/update_s_acoustic_PML.c:33-124,
// src/update_s_acoustic_abs.c:107-171, src/wavefield_update_s_ac.c:27-33),
// one thread per cell: vxx, vyy with the x / y CPML of the reference's
// cpml_update_s_x_ac / _y_ac (src/cpml_update.c:70-80) in the frame, then
// the pressure (sxx and syy are one array) takes g vxx and then g vyy; the
// damping frame scales it once (abs_update_s_ac2, src/abs_update.c:28-31);
// fused explosive sources add each amplitude twice (psource adds it to sxx
// and to syy).
template <int ND, bool ONE, int TBY = BY>
__device__ __forceinline__ void stress_ac_tile(const Grid &g, const Fields &f0, const Mat &m, const Src &s,
                                               const Fuse &fz, int nt, int t)
{
    const int ty = t / g.ntx, tx = t - ty * g.ntx;
    const int i = 1 + tx * BX + threadIdx.x;
    const int j = 1 + ty * TBY + threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    const bool xf = in_xframe(g, i), yf = in_yframe(g, j), frame = xf || yf;
    if (frame && g.abs_type != 1 && g.abs_type != 2) return;
    const int c = IDX(g, i, j), P = g.pitch;
    const float gc = m.g[c];
    const int k0 = ONE ? 0 : g.k_lo + blockIdx.y * g.spt, k1 = ONE ? 1 : min(k0 + g.spt, g.k_hi);
    for (int k = k0; k < k1; ++k) {
        Fields f = ONE ? f0 : f0.at(k, g);
        float a = 0.f, e = 0.f;
#pragma unroll
        for (int mm = 1; mm <= ND; ++mm) {
            const float h = g.hc[mm];
            a += h * (f.vx[c + mm - 1] - f.vx[c - mm]);
            e += h * (f.vy[c + (mm - 1) * P] - f.vy[c - mm * P]);
        }
        float vxx = a * g.dhi, vyy = e * g.dhi;
        if (frame && g.abs_type == 1) {
            if (xf) {
                const int hh = hx(g, i), o = j * g.wx + hh;
                const float px = PRF(g, m, SOFI_P_B_X, hh) * f.psi_vxx[o] + PRF(g, m, SOFI_P_A_X, hh) * vxx;
                f.psi_vxx[o] = px;
                vxx = vxx / PRF(g, m, SOFI_P_K_X, hh) + px;
            }
            if (yf) {
                const int hh = hy(g, j), o = hh * g.wy + i;
                const float py = PRF(g, m, SOFI_P_B_Y, hh) * f.psi_vyy[o] + PRF(g, m, SOFI_P_A_Y, hh) * vyy;
                f.psi_vyy[o] = py;
                vyy = vyy / PRF(g, m, SOFI_P_K_Y, hh) + py;
            }
        }
        float p = f.sxx[c];
        p += gc * vxx;
        p += gc * vyy;
        if (frame && g.abs_type == 2) p *= m.absorb[c];
        for (int q = 0; q < fz.npsrc; ++q)
            if (c == fz.c[q]) {
                const float amp = psource_amp(g, s, fz.soff + q, step_of(fz.ntb, nt));
                p += amp;
                p += amp;
            }
        f.sxx[c] = p;
    }
}
template <int ND, bool ONE>
__global__ __launch_bounds__(BX *BY) void k_stress_ac(Grid g, Fields f0, Mat m, Src s, Fuse fz, int nt)
{
    if (ONE && fused_seis(g, f0, fz)) return;
    stress_ac_tile<ND, ONE>(g, f0, m, s, fz, nt, tile_id());
}

// ------------------------------------------------------ work-list launches

// One handle's part of a work-list launch: its grid, fields, medium and
// sources, and what its velocity / stress updates fuse (forces; explosive
// sources and the free surface).  Fuse::ntb points at the handle's step
// before the first step of the launch sequence, so that local step `it` of a
// launch is the handle's step *ntb + it.
struct Shot {
    Grid g;
    Fields f;
    Mat m;
    Src s;
    Fuse fv, fs;
    Rec r;          // its receivers (sampled by k_seis_w)
    int seismo;
};

// The one-thread-per-cell update kernels over several one-shot handles on
// grids of any size, one launch for all of them, in tiles of BX x TBY cells
// (TBY waves per block): item w = (handle, tile) covers `rows` tiles down the
// tile column from that tile; block b takes items b, b + gridDim.x, ... (a
// persistent grid when gridDim.x is below the item count).  Items 0..n-1
// belong to the handles that still advance (the handles with the most steps
// come first).  Every cell runs the code of the one-shot kernel, so each
// handle's results are those of its own launches, bit for bit, whatever the
// tile height, rows and grid size.
template <int ND, bool NT, int TBY>
__global__ __launch_bounds__(BX *TBY) void k_velocity_w(const Shot *__restrict__ sh, const int2 *__restrict__ item,
                                                        int n, int rows, int it)
{
    for (int w = blockIdx.x; w < n; w += gridDim.x) {
        const int2 e = item[w];
        const Shot &d = sh[e.x];
        const int last = d.g.ntx * ((d.g.ny + TBY - 1) / TBY);
        for (int r = 0, t = e.y; r < rows && t < last; ++r, t += d.g.ntx)
            velocity_tile<ND, true, NT, TBY>(d.g, d.f, d.m, d.s, d.fv, it, t);
    }
}

// The samplers of a work-list step, one launch for every handle that
// samples in it: job q (blockIdx.y) samples the receivers of handle shot[q]
// at sample index ins[q], div / curl (DV) or vx / vy / p (seismo_ssg,
// src/seismo_ssg.c:40-64), as k_seis_dv / k_seis do for one handle.
constexpr int SEIS_JOBS = 64;
struct SeisJobs {
    int shot[SEIS_JOBS], ins[SEIS_JOBS];
};
template <int ND, bool DV>
__global__ void k_seis_w(const Shot *__restrict__ sh, SeisJobs jobs)
{
    const Shot &d = sh[jobs.shot[blockIdx.y]];
    const int k = blockIdx.x * blockDim.x + threadIdx.x, ins = jobs.ins[blockIdx.y];
    if (k >= d.r.n) return;
    if (!DV) {
        seis_sample(d.g, d.f, d.r, k, ins, d.seismo);
        return;
    }
    const Grid &g = d.g;
    const Fields f = d.f.at(d.r.shot[k], g);
    const int i = d.r.i[k], j = d.r.j[k], c = IDX(g, i, j);
    float vxx, vyx, vxy, vyy;
    op_s<ND>(g, f.vx, f.vy, c, vxx, vyx, vxy, vyy);
    const bool xf = in_xframe(g, i), yf = in_yframe(g, j);
    if (g.abs_type == 1 && (xf || yf)) cpml_dv_read(g, d.m, f, i, j, xf, yf, vxx, vyx, vxy, vyy);
    float *t = d.r.sec + d.r.off[k] + (ins - 1);
    t[SOFI_C_DIV * d.r.cst[k]] = (float)((double)(vxx + vyy) * d.r.sq_ppi[k]);
    t[SOFI_C_CURL * d.r.cst[k]] = (float)((double)(vxy - vyx) * d.r.sq_pu[k]);
}

template <int ND, int WEQ, int TBY>
__global__ __launch_bounds__(BX *TBY) void k_stress_w(const Shot *__restrict__ sh, const int2 *__restrict__ item,
                                                      int n, int rows, int it)
{
    for (int w = blockIdx.x; w < n; w += gridDim.x) {
        const int2 e = item[w];
        const Shot &d = sh[e.x];
        const int last = d.g.ntx * ((d.g.ny + TBY - 1) / TBY);
        for (int r = 0, t = e.y; r < rows && t < last; ++r, t += d.g.ntx) {
            if constexpr (WEQ == SOFI_AC_ISO) stress_ac_tile<ND, true, TBY>(d.g, d.f, d.m, d.s, d.fs, it, t);
            else stress_tile<ND, WEQ, true, TBY>(d.g, d.f, d.m, d.s, d.fs, it, t);
        }
    }
}

// ------------------------------------------------------------ TTI passes

// The TTI frame tiles: tile q of the list (frame_tile_count of them) ->
// tile column tx and row ty.  First the tile rows at the top and at the
// bottom, whole; then, row by row in between, the tile columns at the left
// and at the right.
__host__ __device__ __forceinline__ int frame_tile_count(const Grid &g)
{
    return (g.ftop + g.fbot) * g.ntx + (g.nty - g.ftop - g.fbot) * (g.fl + g.fr);
}
__device__ __forceinline__ void frame_tile(const Grid &g, int q, int &tx, int &ty)
{
    const int top = g.ftop * g.ntx, bot = g.fbot * g.ntx;
    if (q < top) {
        ty = q / g.ntx;
        tx = q - ty * g.ntx;
        return;
    }
    q -= top;
    if (q < bot) {
        const int r = q / g.ntx;
        ty = g.nty - g.fbot + r;
        tx = q - r * g.ntx;
        return;
    }
    q -= bot;
    const int w = g.fl + g.fr, r = q / w, col = q - r * w;
    ty = g.ftop + r;
    tx = col < g.fl ? col : g.ntx - g.fr + (col - g.fl);
}
__device__ __forceinline__ bool is_frame_tile(const Grid &g, int tx, int ty)
{
    return ty < g.ftop || ty >= g.nty - g.fbot || tx < g.fl || tx >= g.ntx - g.fr;
}


// The TTI CPML frame (src/update_s_elastic_TTI_PML.c:32-145,
// src/update_s_visc_TTI_PML.c).  The reference treats the frame region by
// region (left, right, top, bottom, then the four corners), CPML then update
// per region, so a side cell reads its neighbours' derivatives after the
// sides' CPML and before the corners' CPML, and a corner cell reads them
// after both.  Here k_stress_tti_f applies the CPML of every frame cell as it
// stores the frame derivatives (the elastic right-bottom corner loop stops
// at GY[4]-1, src/update_s_elastic_TTI_PML.c:132 vs
// src/update_s_visc_TTI_PML.c:132, so those cells keep their raw ones), and
// k_tti_frame then updates sides and corners in one launch over the frame
// tiles (blockIdx.x) and slots (blockIdx.y): a side cell next to a corner
// recomputes that corner cell's raw derivatives from vx, vy instead of
// reading the stored ones.  Both updates only read the derivative arrays, so
// they do not depend on each other.  The updates cover the ranges shrunk by
// two cells at the outer edges.

// the TTI stress update of one frame cell from its derivatives and their
// 4-point averages (src/wavefield_update_s_el_tti.c:27-52,
// src/wavefield_update_s_visc_TTI.c:27-84)
template <int WEQ>
__device__ __forceinline__ void tti_frame_update(const Grid &g, const Fields &f, const Mat &m, const Fuse &fz, int c,
                                                 float vxx, float vyy, float vxxipjp, float vyyipjp, float vij, float v)
{
    float u1, u2, u3;
    if (WEQ == SOFI_EL_TTI) {
        u1 = (m.c55i[c] * v) + (m.c15i[c] * vxxipjp) + (m.c35i[c] * vyyipjp);
        u2 = (m.c11[c] * vxx) + (m.c13[c] * vyy) + (m.c15[c] * vij);
        u3 = (m.c13[c] * vxx) + (m.c33[c] * vyy) + (m.c35[c] * vij);
    } else {
        float sr0 = 0.f, sp0 = 0.f, sq0 = 0.f, sr = 0.f, sp = 0.f, sq = 0.f;
#pragma unroll
        for (int l = 1; l <= SOFI_MAXL; ++l) {
            if (l > g.L) break;
            const int o = (l - 1) * g.plane + c;
            const float r0 = f.r[o], p0 = f.p[o], q0 = f.q[o];
            sr0 += r0;
            sp0 += p0;
            sq0 += q0;
            const float bip = m.bip[l], cip = m.cip[l];
            const float r1 = bip * (r0 * cip - (m.d55i[o] * v) - (m.d15i[o] * vxxipjp) - (m.d35i[o] * vyyipjp));
            const float p1 = bip * (p0 * cip - (m.d11[o] * vxx) - (m.d13[o] * vyy) - (m.d15[o] * vij));
            const float q1 = bip * (q0 * cip - (m.d13[o] * vxx) - (m.d33[o] * vyy) - (m.d35[o] * vij));
            f.r[o] = r1;
            f.p[o] = p1;
            f.q[o] = q1;
            sr += r1;
            sp += p1;
            sq += q1;
        }
        u1 = (m.u55i[c] * v) + (m.u15i[c] * vxxipjp) + (m.u35i[c] * vyyipjp) + (g.dthalbe * (sr0 + sr));
        u2 = (m.u11[c] * vxx) + (m.u13[c] * vyy) + (m.u15[c] * vij) + (g.dthalbe * (sp0 + sp));
        u3 = (m.u13[c] * vxx) + (m.u33[c] * vyy) + (m.u35[c] * vij) + (g.dthalbe * (sq0 + sq));
    }
    fwi_u_store(fz, f, g.dt, c, u1, u2, u3);
    float sxy = f.sxy[c] + u1, sxx = f.sxx[c] + u2, syy = f.syy[c] + u3;
    f.sxy[c] = sxy;
    f.sxx[c] = sxx;
    f.syy[c] = syy;
}

// a frame cell the frame update updates: not within two cells of the outer
// edge of its frame side
__device__ __forceinline__ bool tti_frame_updated(const Grid &g, int i, int j)
{
    return (in_xframe(g, i) || in_yframe(g, j)) &&
           !((i <= g.gx2 && i < 1 + 2) || (i > g.gx3 && i > g.nx - 2) || (j <= g.gy2 && j < 1 + 2) ||
             (j > g.gy3 && j > g.ny - 2));
}

// the frame cell whose derivatives keep their raw values (no CPML): the
// elastic right-bottom corner's last row
__device__ __forceinline__ bool tti_cpml_skipped(const Grid &g, int i, int j)
{
    return g.weq == SOFI_EL_TTI && i > g.gx3 && j > g.gy3 && j == g.ny;
}

// explosive sources of cell c at step nt, in source order (psource)
__device__ __forceinline__ void psource_cell(const Grid &g, const Fields &f, const Src &src, const Fuse &fz, int c,
                                             int nt)
{
    for (int k = 0; k < fz.npsrc; ++k)
        if (c == fz.c[k]) {
            const float amp = psource_amp(g, src, fz.soff + k, step_of(fz.ntb, nt));
            f.sxx[c] += amp;
            f.syy[c] += amp;
        }
}

// The frame update, the last stress kernel of a TTI step with a CPML frame.
// With fz.npsrc (one-shot handles) it also adds the explosive sources after
// every update (psource, src/time_loop.c:248-250): the thread of an updated
// frame cell for that cell, the first thread of the launch for all other
// cells, whose stresses no kernel touches any more.
template <int ND, int WEQ>
__global__ __launch_bounds__(BX *BY) void k_tti_frame(Grid g, Fields f0, Mat m, Fuse fz, Src src, int nt)
{
    const Fields f = f0.at(g.k_lo + blockIdx.y, g);   // shot slot of a batch
    if (fz.npsrc > 0 && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0)
        for (int k = 0; k < fz.npsrc; ++k) {
            const int c = fz.c[k], ci = c % g.pitch - SOFI_PAD + 1, cj = c / g.pitch - SOFI_PAD + 1;
            if (tti_frame_updated(g, ci, cj)) continue;
            bool first = true;      // a cell's sources are added once, at its first source
            for (int q = 0; q < k; ++q) first = first && fz.c[q] != c;
            if (first) psource_cell(g, f, src, fz, c, nt);
        }
    int tx, ty;
    frame_tile(g, blockIdx.x, tx, ty);
    const int i = 1 + tx * BX + threadIdx.x;
    const int j = 1 + ty * BY + threadIdx.y;
    if (i > g.nx || j > g.ny || !tti_frame_updated(g, i, j)) return;
    const int c = IDX(g, i, j), P = g.pitch;
    const bool corner = in_xframe(g, i) && in_yframe(g, j);
    // the derivatives of cell (i + di, j + dj) as this update reads them: for a
    // side cell, a corner cell's raw ones (its CPML comes after the sides'
    // update in the reference), else the stored ones
    auto rd = [&](int di, int dj, float &xx, float &yy, float &yx, float &xy) {
        const int ii = i + di, jj = j + dj, cc = c + di + dj * P;
        if (!corner && in_xframe(g, ii) && in_yframe(g, jj) && ii >= 1 && ii <= g.nx && jj >= 1 && jj <= g.ny) {
            op_s<ND>(g, f.vx, f.vy, cc, xx, yx, xy, yy);
        } else {
            xx = f.dvxx[cc];
            yy = f.dvyy[cc];
            yx = f.dvyx[cc];
            xy = f.dvxy[cc];
        }
    };
    float x00, y00, yx00, xy00, x01, y01, d0, d1, x10, y10, x11, y11;
    float yxm0, xym0, yx0m, xy0m, yxmm, xymm;
    rd(0, 0, x00, y00, yx00, xy00);
    rd(0, 1, x01, y01, d0, d1);
    rd(1, 0, x10, y10, d0, d1);
    rd(1, 1, x11, y11, d0, d1);
    rd(0, -1, d0, d1, yx0m, xy0m);
    rd(-1, 0, d0, d1, yxm0, xym0);
    rd(-1, -1, d0, d1, yxmm, xymm);
    const float vxxipjp = 0.25f * (x00 + x01 + x10 + x11);
    const float vyyipjp = 0.25f * (y00 + y01 + y10 + y11);
    const float vyxij = 0.25f * (yx00 + yx0m + yxm0 + yxmm);
    const float vxyij = 0.25f * (xy00 + xy0m + xym0 + xymm);
    tti_frame_update<WEQ>(g, f, m, fz, c, x00, y00, vxxipjp, vyyipjp, vyxij + vxyij, xy00 + yx00);
    psource_cell(g, f, src, fz, c, nt);
}


// ------------------------------------------------- periodic boundary in x
//
// exchange_v / exchange_s with BOUNDARY = 1 on one process: the process is
// its own left and right neighbour, so the ghost columns of rows 1..NY wrap
// around the grid (src/exchange_v.c:91-148, src/exchange_s.c:92-146,193-236).
// The reference fills the array in the vx / sxy position with ND - 1
// columns on the right and ND on the left, the one in the vy / sxx position
// with ND on the right and ND - 1 on the left; the same here.  a may be null
// (the acoustic path has no sxy).  One thread per row, blockIdx.y the slot.
__global__ void k_wrap_x(Grid g, float *a, float *b)
{
    const int j = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (j > g.ny) return;
    const size_t off = (size_t)(g.k_lo + blockIdx.y) * g.plane;
    const int r0 = IDX(g, 1, j), nx = g.nx, nd = g.nd;
    if (a) {
        float *p = a + off + r0;                 // p[i - 1] = a[j][i]
        for (int l = 1; l <= nd - 1; ++l) p[nx + l - 1] = p[l - 1];
        for (int l = 1; l <= nd; ++l) p[-l] = p[nx - l];
    }
    float *q = b + off + r0;
    for (int l = 1; l <= nd; ++l) q[nx + l - 1] = q[l - 1];
    for (int l = 1; l <= nd - 1; ++l) q[-l] = q[nx - l];
}

// ------------------------------------------ batches: S slots per thread
//
// The update kernels of a batch in its "slot-block" shape: one thread per
// cell for S consecutive shot slots (S fixed at compile time), the cell's
// material coefficients read once for them, every slot's stencil and state
// loaded before any slot is updated (the slots' planes are disjoint parts of
// the same allocations, so the compiler cannot move one slot's loads above
// another's stores by itself), then the slots updated and stored in turn.
// blockIdx.x is the tile, blockIdx.y the group of S slots, so consecutive
// workgroups read one tile's coefficients.  Sources, the free surface and
// sampling run as separate launches; the per-cell arithmetic of a slot is
// that of the one-thread-per-cell kernels.

template <int ND, int S>
__global__ __launch_bounds__(BX *BY) void k_velocity_b(Grid g, Fields f0, Mat m, int nt, const int *ntb, int fwi_acc)
{
    const int t = blockIdx.x, ty = t / g.ntx, tx = t - ty * g.ntx;
    const int i = 1 + tx * BX + threadIdx.x;
    const int j = 1 + ty * BY + threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    const int c = IDX(g, i, j);
    const float rip = m.rip[c], rjp = m.rjp[c];
    const int k0 = g.k_lo + blockIdx.y * S;
    float a[S], b[S], d[S], e[S], vx[S], vy[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const Fields f = f0.at(k0 + q, g);
        op_v<ND>(g, f.sxx, f.syy, f.sxy, c, a[q], b[q], d[q], e[q]);
        vx[q] = f.vx[c];
        vy[q] = f.vy[c];
    }
    const Src s0{};
    Fuse fz{};
    fz.ntb = ntb;
    fz.fwi_acc = fwi_acc;     // MODE=FWI: the accelerations / adjoint velocities of every slot
#pragma unroll
    for (int q = 0; q < S; ++q) {
        Fields f = f0.at(k0 + q, g);
        velocity_core(g, f, m, s0, fz, nt, i, j, c, a[q], b[q], d[q], e[q], rip, rjp, vx[q], vy[q]);
        f.vx[c] = vx[q];
        f.vy[c] = vy[q];
    }
}

template <int ND, int WEQ, int S>
__global__ __launch_bounds__(BX *BY) void k_stress_b(Grid g, Fields f0, Mat m, int nt, const int *ntb, int fwi_u, Src s,
                                                     PFuse pf)
{
    constexpr bool VISCO = WEQ == SOFI_VEL_ISO || WEQ == SOFI_VEL_VTI;
    const int t = blockIdx.x, ty = t / g.ntx, tx = t - ty * g.ntx;
    const int i = 1 + tx * BX + threadIdx.x;
    const int j = 1 + ty * BY + threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    const int c = IDX(g, i, j), P = g.pitch;
    float c1 = 0.f, c2 = 0.f, c3 = 0.f, c4 = 0.f, br = 0.f, bd = 0.f, be = 0.f;
    if (WEQ == SOFI_EL_ISO || WEQ == SOFI_VEL_ISO) {
        c1 = m.f[c];
        c2 = m.g[c];
        c3 = m.fipjp[c];
    }
    if (WEQ == SOFI_VEL_ISO) {
        br = m.br[c];
        bd = m.bd[c];
        be = m.be[c];
    }
    if (WEQ == SOFI_EL_VTI) {
        c1 = m.c11[c];
        c2 = m.c13[c];
        c3 = m.c33[c];
        c4 = m.c55i[c];
    }
    if (WEQ == SOFI_VEL_VTI) {
        c1 = m.u11[c];
        c2 = m.u13[c];
        c3 = m.u33[c];
        c4 = m.u55i[c];
    }
    const int k0 = g.k_lo + blockIdx.y * S;
    float a[S], b[S], d[S], e[S], sxx[S], syy[S], sxy[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const Fields f = f0.at(k0 + q, g);
        float a_ = 0.f, b_ = 0.f, d_ = 0.f, e_ = 0.f;
#pragma unroll
        for (int mm = 1; mm <= ND; ++mm) {
            const float h = g.hc[mm];
            a_ += h * (f.vx[c + mm - 1] - f.vx[c - mm]);
            b_ += h * (f.vy[c + mm] - f.vy[c - mm + 1]);
            d_ += h * (f.vx[c + mm * P] - f.vx[c - (mm - 1) * P]);
            e_ += h * (f.vy[c + (mm - 1) * P] - f.vy[c - mm * P]);
        }
        a[q] = a_;
        b[q] = b_;
        d[q] = d_;
        e[q] = e_;
        sxx[q] = f.sxx[c];
        syy[q] = f.syy[c];
        sxy[q] = f.sxy[c];
    }
    const Src s0{};
    Fuse fz{};
    fz.ntb = ntb;
    fz.fwi_u = fwi_u;         // MODE=FWI: the stress increments of every slot
#pragma unroll
    for (int q = 0; q < S; ++q) {
        Fields f = f0.at(k0 + q, g);
        float vxx, vyx, vxy, vyy;
        if (!stress_dv(g, f, m, i, j, a[q], b[q], d[q], e[q], vxx, vyx, vxy, vyy)) continue;
        float sx = sxx[q], sy = syy[q], sz = sxy[q];
        if (WEQ == SOFI_EL_ISO) upd_el(c1, c2, c3, vxx, vyx, vxy, vyy, sx, sy, sz, fz, f, g.dt, c);
        if (WEQ == SOFI_EL_VTI) upd_el_vti(c1, c2, c3, c4, vxx, vyx, vxy, vyy, sx, sy, sz, fz, f, g.dt, c);
        if (WEQ == SOFI_VEL_ISO) {
            const float sxyv = vxy + vyx, sdiag = vxx + vyy;
            VSum u;
#pragma unroll
            for (int l = 1; l <= SOFI_MAXL; ++l) {
                if (l > g.L) break;
                const int o = (l - 1) * g.plane + c;
                float r = f.r[o], p = f.p[o], qq = f.q[o];
                visc_level(m.eta[l], m.bip[l], m.cip[l], br, bd, be, sxyv, sdiag, vxx, vyy, r, p, qq, u);
                f.r[o] = r;
                f.p[o] = p;
                f.q[o] = qq;
            }
            visc_final(g.dthalbe, c1, c2, c3, sxyv, sdiag, vxx, vyy, u, sx, sy, sz, fz, f, g.dt, c);
        }
        if (WEQ == SOFI_VEL_VTI) {
            const float sxyv = vxy + vyx;
            VSum u;
#pragma unroll
            for (int l = 1; l <= SOFI_MAXL; ++l) {
                if (l > g.L) break;
                const int o = (l - 1) * g.plane + c;
                float r = f.r[o], p = f.p[o], qq = f.q[o];
                visc_vti_level(m.bip[l], m.cip[l], m.d11[o], m.d13[o], m.d33[o], m.d55i[o], sxyv, vxx, vyy, r, p, qq, u);
                f.r[o] = r;
                f.p[o] = p;
                f.q[o] = qq;
            }
            visc_vti_final(g.dthalbe, c1, c2, c3, c4, sxyv, vxx, vyy, u, sx, sy, sz, fz, f, g.dt, c);
        }
        stress_post<ND, VISCO>(g, f, m, s0, fz, nt, i, j, c, sx, sy, sz);
        for (int w = 0; w < pf.n; ++w)
            if (pf.k[w] == k0 + q && pf.c[w] == c) {
                const float amp = psource_amp(g, s, pf.l[w], step_of(ntb, nt));
                sx += amp;
                sy += amp;
            }
        f.sxx[c] = sx;
        f.syy[c] = sy;
        f.sxy[c] = sz;
    }
}

// TTI stress update with the velocity derivatives computed in the kernel:
// each workgroup first computes the four derivatives (op_s, raw: no CPML) of
// its BX x BY tile and a one-cell ring into LDS (zero outside 1..NX x 1..NY,
// the reference's never-written ghosts of the derivative arrays), then
// updates its cells from them (src/wavefield_update_s_el_tti.c:27-52,
// src/wavefield_update_s_visc_TTI.c:27-84).
//   * interior cells: updated here; they read the raw derivatives of their
//     neighbours, frame cells included, as the reference's interior update
//     runs before the frame CPML (src/time_loop.c:180-186);
//   * with the damping frame (ABS_TYPE 2) every frame cell is updated here
//     too, then damped (src/update_s_elastic_TTI_abs.c:35-99);
//   * with the CPML frame, tiles that hold a frame cell or one next to it
//     store the derivatives of their cells into the derivative arrays: the
//     frame cells' after their CPML, the others' raw; k_tti_frame (over
//     these tiles only) then updates the frame;
//   * explosive sources (fz.npsrc: one-shot handles with no CPML frame
//     pass to follow) are added to their cells after the update, as psource
//     does after the stress update (src/time_loop.c:248-250).
// S slots per thread (blockIdx.y: group of S slots of a batch; 1 for a
// one-shot handle).
template <int ND, int WEQ, int S>
__global__ __launch_bounds__(BX *BY) void k_stress_tti_f(Grid g, Fields f0, Mat m, Fuse fz, Src src, int nt)
{
    constexpr int LX = BX + 2, LY = BY + 2;
    __shared__ float sd[S][4][LY][LX];   // dvxx, dvyy, dvyx, dvxy of rows j0-1..j0+BY, columns i0-1..i0+BX
    const int t = blockIdx.x, ty = t / g.ntx, tx = t - ty * g.ntx;
    const int i0 = 1 + tx * BX, j0 = 1 + ty * BY;
    const int k0 = g.k_lo + blockIdx.y * S;
    const bool cpml = g.abs_type == 1 && g.fw > 0;
    const bool store_dv = cpml && is_frame_tile(g, tx, ty);
    const int tid = threadIdx.y * BX + threadIdx.x;
    for (int q = tid; q < LX * LY; q += BX * BY) {
        const int lj = q / LX, li = q - lj * LX;
        const int ii = i0 - 1 + li, jj = j0 - 1 + lj;
        const bool in = ii >= 1 && ii <= g.nx && jj >= 1 && jj <= g.ny;
        // periodic in x: the ghost columns 0 and NX + 1 of rows 1..NY hold
        // the wrapped derivatives as exchange_v leaves them in the derivative
        // arrays (src/time_loop.c:107-110): vxx and vyx (the vx position)
        // on the left, vyy and vxy (the vy position) on the right, the other
        // pair only when ND >= 2
        const bool wl = g.periodic && ii == 0 && jj >= 1 && jj <= g.ny;
        const bool wr = g.periodic && ii == g.nx + 1 && jj >= 1 && jj <= g.ny;
        const int cc = IDX(g, ii, jj), cw = wl ? IDX(g, g.nx, jj) : wr ? IDX(g, 1, jj) : cc;
        const bool own = store_dv && in && li >= 1 && li <= BX && lj >= 1 && lj <= BY;
        const bool ghost = store_dv && (wl || wr) && lj >= 1 && lj <= BY;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            float vxx = 0.f, vyx = 0.f, vxy = 0.f, vyy = 0.f;
            const Fields f = f0.at(k0 + s, g);
            if (in || wl || wr) op_s<ND>(g, f.vx, f.vy, cw, vxx, vyx, vxy, vyy);
            if (wl && ND < 2) vyy = vxy = 0.f;
            if (wr && ND < 2) vxx = vyx = 0.f;
            sd[s][0][lj][li] = vxx;
            sd[s][1][lj][li] = vyy;
            sd[s][2][lj][li] = vyx;
            sd[s][3][lj][li] = vxy;
            // raw derivatives of the tile's cells outside the frame (the
            // frame cells' follow below, after their CPML) and of periodic
            // ghost columns
            if ((own && !in_xframe(g, ii) && !in_yframe(g, jj)) || ghost) {
                f.dvxx[cc] = vxx;
                f.dvyy[cc] = vyy;
                f.dvyx[cc] = vyx;
                f.dvxy[cc] = vxy;
            }
        }
    }
    __syncthreads();
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    const int c = IDX(g, i, j);
    const int li = threadIdx.x + 1, lj = threadIdx.y + 1;
    const bool xf = in_xframe(g, i), yf = in_yframe(g, j), frame = xf || yf;
    const bool damp = frame && g.abs_type == 2 && g.fw > 0;
    if (frame && !damp) {
        // CPML frame: updated by k_tti_frame (no absorbing frame: never
        // updated); a frame cell's derivatives get their CPML here and are
        // stored for it (the tile's updates read the raw ones from LDS)
        if (store_dv) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                Fields f = f0.at(k0 + s, g);
                float vxx = sd[s][0][lj][li], vyy = sd[s][1][lj][li], vyx = sd[s][2][lj][li], vxy = sd[s][3][lj][li];
                if (!tti_cpml_skipped(g, i, j)) cpml_dv(g, m, f, i, j, xf, yf, vxx, vyx, vxy, vyy);
                f.dvxx[c] = vxx;
                f.dvyy[c] = vyy;
                f.dvyx[c] = vyx;
                f.dvxy[c] = vxy;
            }
        }
        return;
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const Fields f = f0.at(k0 + s, g);
        const float(*X)[LX] = sd[s][0], (*Y)[LX] = sd[s][1], (*YX)[LX] = sd[s][2], (*XY)[LX] = sd[s][3];
        const float vxx = X[lj][li], vyy = Y[lj][li];
        const float vxxipjp = 0.25f * (X[lj][li] + X[lj + 1][li] + X[lj][li + 1] + X[lj + 1][li + 1]);
        const float vyyipjp = 0.25f * (Y[lj][li] + Y[lj + 1][li] + Y[lj][li + 1] + Y[lj + 1][li + 1]);
        const float vyxij = 0.25f * (YX[lj][li] + YX[lj - 1][li] + YX[lj][li - 1] + YX[lj - 1][li - 1]);
        const float vxyij = 0.25f * (XY[lj][li] + XY[lj - 1][li] + XY[lj][li - 1] + XY[lj - 1][li - 1]);
        const float vij = vyxij + vxyij;
        const float v = XY[lj][li] + YX[lj][li];
        float u1, u2, u3;
        if (WEQ == SOFI_EL_TTI) {
            u1 = (m.c55i[c] * v) + (m.c15i[c] * vxxipjp) + (m.c35i[c] * vyyipjp);
            u2 = (m.c11[c] * vxx) + (m.c13[c] * vyy) + (m.c15[c] * vij);
            u3 = (m.c13[c] * vxx) + (m.c33[c] * vyy) + (m.c35[c] * vij);
        } else {
            float sr0 = 0.f, sp0 = 0.f, sq0 = 0.f, sr = 0.f, sp = 0.f, sq = 0.f;
#pragma unroll
            for (int l = 1; l <= SOFI_MAXL; ++l) {
                if (l > g.L) break;
                const int o = (l - 1) * g.plane + c;
                const float r0 = f.r[o], p0 = f.p[o], q0 = f.q[o];
                sr0 += r0;
                sp0 += p0;
                sq0 += q0;
                const float bip = m.bip[l], cip = m.cip[l];
                const float r1 = bip * (r0 * cip - (m.d55i[o] * v) - (m.d15i[o] * vxxipjp) - (m.d35i[o] * vyyipjp));
                const float p1 = bip * (p0 * cip - (m.d11[o] * vxx) - (m.d13[o] * vyy) - (m.d15[o] * vij));
                const float q1 = bip * (q0 * cip - (m.d13[o] * vxx) - (m.d33[o] * vyy) - (m.d35[o] * vij));
                f.r[o] = r1;
                f.p[o] = p1;
                f.q[o] = q1;
                sr += r1;
                sp += p1;
                sq += q1;
            }
            u1 = (m.u55i[c] * v) + (m.u15i[c] * vxxipjp) + (m.u35i[c] * vyyipjp) + (g.dthalbe * (sr0 + sr));
            u2 = (m.u11[c] * vxx) + (m.u13[c] * vyy) + (m.u15[c] * vij) + (g.dthalbe * (sp0 + sp));
            u3 = (m.u13[c] * vxx) + (m.u33[c] * vyy) + (m.u35[c] * vij) + (g.dthalbe * (sq0 + sq));
        }
        fwi_u_store(fz, f, g.dt, c, u1, u2, u3);
        float sxy = f.sxy[c] + u1, sxx = f.sxx[c] + u2, syy = f.syy[c] + u3;
        if (damp) {
            const float a = m.absorb[c];
            sxy *= a;
            syy *= a;
            sxx *= a;
        }
        if (S == 1)
            for (int k = 0; k < fz.npsrc; ++k)
                if (c == fz.c[k]) {
                    const float amp = psource_amp(g, src, fz.soff + k, step_of(fz.ntb, nt));
                    sxx += amp;
                    syy += amp;
                }
        f.sxy[c] = sxy;
        f.sxx[c] = sxx;
        f.syy[c] = syy;
    }
}

// ------------------------------------------------------------- samplers

// seismo_ssg div/curl (src/seismo_ssg.c:51-58): the post-CPML derivatives;
// runs right after the stress update, before the free surface touches psi.
template <int ND, bool TTI>
__global__ void k_seis_dv(Grid g, Fields f0, Mat m, Rec r, int ins, const int *ntb, int ndt)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= r.n) return;
    if (ntb) ins = (ins + *ntb) / ndt;      // ins: the block's local step
    const Fields f = f0.at(r.shot[k], g);
    const int i = r.i[k], j = r.j[k], c = IDX(g, i, j);
    float vxx, vyx, vxy, vyy;
    if (TTI) {
        // the derivative arrays hold the post-CPML values of the CPML frame;
        // elsewhere the derivatives are the raw ones, recomputed here
        const bool xf = in_xframe(g, i), yf = in_yframe(g, j);
        if (g.abs_type == 1 && g.fw > 0 && (xf || yf)) {
            vxx = f.dvxx[c];
            vyy = f.dvyy[c];
            vyx = f.dvyx[c];
            vxy = f.dvxy[c];
        } else {
            op_s<ND>(g, f.vx, f.vy, c, vxx, vyx, vxy, vyy);
        }
    } else {
        op_s<ND>(g, f.vx, f.vy, c, vxx, vyx, vxy, vyy);
        const bool xf = in_xframe(g, i), yf = in_yframe(g, j);
        if (g.abs_type == 1 && (xf || yf)) cpml_dv_read(g, m, f, i, j, xf, yf, vxx, vyx, vxy, vyy);
    }
    float *t = r.sec + r.off[k] + (ins - 1);
    t[SOFI_C_DIV * r.cst[k]] = (float)((double)(vxx + vyy) * r.sq_ppi[k]);
    t[SOFI_C_CURL * r.cst[k]] = (float)((double)(vxy - vyx) * r.sq_pu[k]);
}

// seismo_ssg vx / vy / p (src/seismo_ssg.c:40-64)
__global__ void k_seis(Grid g, Fields f0, Rec r, int ins, int seismo, const int *ntb, int ndt)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= r.n) return;
    if (ntb) ins = (ins + *ntb) / ndt;      // ins: the block's local step
    seis_sample(g, f0, r, k, ins, seismo);
}


// snap (src/snap_ssg.c:68-205): x-major decimated window; curl and div
// recomputed from vx, vy without CPML
template <int ND>
__global__ void k_snap(Grid g, Fields f, Mat m, float *out, int i0, int di, int nxs, int j0, int dj, int nys)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nxs * nys) return;
    const int ii = k / nys, jj = k % nys;
    const int i = i0 + ii * di, j = j0 + jj * dj, c = IDX(g, i, j), P = g.pitch;
    const size_t per = (size_t)nxs * nys;
    out[SOFI_C_VX * per + k] = f.vx[c];
    out[SOFI_C_VY * per + k] = f.vy[c];
    out[SOFI_C_P * per + k] = -f.sxx[c] - f.syy[c];
    const float dhi = (float)(1.0 / g.dh);
    float vyx = 0.f, vxy = 0.f, vxx = 0.f, vyy = 0.f;
#pragma unroll
    for (int m1 = 1; m1 <= ND; ++m1) {
        vyx += g.hc[m1] * (f.vy[c + m1] - f.vy[c - m1 + 1]);
        vxy += g.hc[m1] * (f.vx[c + m1 * P] - f.vx[c - (m1 - 1) * P]);
        vxx += g.hc[m1] * (f.vx[c + m1 - 1] - f.vx[c - m1]);
        vyy += g.hc[m1] * (f.vy[c + (m1 - 1) * P] - f.vy[c - m1 * P]);
    }
    vyx *= dhi;
    vxy *= dhi;
    vxx *= dhi;
    vyy *= dhi;
    out[SOFI_C_CURL * per + k] = m.pu ? (float)((double)(vxy - vyx) * sqrt((double)m.pu[c])) : 0.0f;   // (none: acoustic)
    out[SOFI_C_DIV * per + k] = (float)((double)(vxx + vyy) * sqrt((double)m.ppi[c]));
}

// ------------------------------------------------------------------ FWI

struct Fwi {
    float *gsum;                               // gradVp, gradVs, gradRho summed over shots (3 planes)
    float *gshot;                              // a shot's gradVp, gradVs, gradRho (3 planes)
    const float *vp0, *vs0, *rho0, *muss, *lamss;
    float *store;                              // per slot [nstore][5][ny][nx]
    size_t sstride;                            // floats per slot of the store
    int velocity;
};

// snap_store (src/snap_store.c:27-47): the forward fields of one store step
// into store slot `slot` (compact [5][ny][nx] per store slot) of every shot
// slot g.k_lo + blockIdx.y
__global__ __launch_bounds__(BX *BY) void k_fwi_store(Grid g, Fields f0, Fwi w, int slot, const int *ntb, int dtinv)
{
    const int i = 1 + blockIdx.x * BX + threadIdx.x;
    const int j = 1 + blockIdx.z * BY + threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    if (ntb) slot += *ntb / dtinv;          // in a captured block (its base step is a multiple of DTINV)
    const int k = g.k_lo + blockIdx.y;
    const Fields f = f0.at(k, g);
    const int c = IDX(g, i, j);
    const size_t n = (size_t)g.nx * g.ny;
    float *st = w.store + (size_t)k * w.sstride;
    const size_t o = (size_t)slot * 5 * n + (size_t)(j - 1) * g.nx + (i - 1);
    st[o] = f.pvxp1[c];
    st[o + n] = f.pvyp1[c];
    if (w.velocity == 0) {
        st[o + 2 * n] = f.sxx[c];
        st[o + 3 * n] = f.syy[c];
        st[o + 4 * n] = f.sxy[c];
    } else {
        st[o + 2 * n] = f.ux[c];
        st[o + 3 * n] = f.uy[c];
        st[o + 4 * n] = f.uxy[c];
    }
}

// calc_conv (src/calc_conv.c:30-73): zero-lag correlation of the adjoint
// fields with the forward fields of store slot `slot`, with the reference's
// double promotions (1.0 / (4.0 ...), 1.0 / 4.0), for every shot slot
// g.k_lo + blockIdx.y
__global__ __launch_bounds__(BX *BY) void k_fwi_conv(Grid g, Fields f0, Mat m, Fwi w, int slot, const int *ntb, int dtinv)
{
    const int i = 1 + blockIdx.x * BX + threadIdx.x;
    const int j = 1 + blockIdx.z * BY + threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    if (ntb) slot -= *ntb / dtinv;          // in a captured block: the forward step runs backwards
    const int k = g.k_lo + blockIdx.y;
    const Fields f = f0.at(k, g);
    const int c = IDX(g, i, j);
    const size_t n = (size_t)g.nx * g.ny;
    const float *st = w.store + (size_t)k * w.sstride;
    const size_t o = (size_t)slot * 5 * n + (size_t)(j - 1) * g.nx + (i - 1);
    const float frx = st[o], fry = st[o + n];
    const float fx = st[o + 2 * n], fy = st[o + 3 * n], fu = st[o + 4 * n];
    const float sxx = f.sxx[c], syy = f.syy[c], sxy = f.sxy[c];
    const float muss = w.muss[c], lamss = w.lamss[c], DT = g.dt;
    f.grho[c] += -DT * ((f.pvxp1[c] * frx) + (f.pvyp1[c] * fry));
    f.glam[c] = (float)((double)f.glam[c] +
                        (double)(-DT) * (1.0 / (4.0 * (double)(lamss + muss) * (double)(lamss + muss))) *
                            (double)((fx + fy) * (sxx + syy)));
    if (m.pu[c] > 0.0f) {
        const double a = (1.0 / (double)(muss * muss)) * (double)(fu * sxy);
        const double b = (1.0 / 4.0) * (double)((fx + fy) * (sxx + syy)) / (double)((lamss + muss) * (lamss + muss));
        const double e = (1.0 / 4.0) * (double)((fx - fy) * (sxx - syy)) / (double)(muss * muss);
        f.gmu[c] = (float)((double)f.gmu[c] + (double)(-DT) * (a + b + e));
    }
}

// gradient of shot slot k (src/inversion.c:226-249): vp / vs / rho' from the
// lambda / mu / rho accumulators, added to the sums; accumulators cleared
__global__ __launch_bounds__(BX *BY) void k_fwi_grad(Grid g, Fields f0, Fwi w, int idx, int idy, int k)
{
    const int i = 1 + blockIdx.x * BX + threadIdx.x;
    const int j = 1 + blockIdx.y * BY + threadIdx.y;
    if (i > g.nx || j > g.ny) return;
    const Fields f = f0.at(k, g);
    const int c = IDX(g, i, j);
    float vp = 0.f, vs = 0.f, rho = 0.f;
    if ((i - 1) % idx == 0 && (j - 1) % idy == 0) {
        const double Vp0 = w.vp0[c], Vs0 = w.vs0[c], Rho0 = w.rho0[c];
        const float gl = f.glam[c], gm = f.gmu[c], gr = f.grho[c];
        vp = (float)(2.0 * Rho0 * Vp0 * gl);
        vs = (float)((-4.0 * Vs0 * Rho0 * gl) + 2.0 * Rho0 * Vp0 * gm);
        rho = (float)((((double)(w.vp0[c] * w.vp0[c]) - (2.0 * Vs0 * Vs0)) * gl) + (double)(w.vs0[c] * w.vs0[c] * gm) + gr);
        w.gsum[c] += vp;
        w.gsum[g.plane + c] += vs;
        w.gsum[2 * g.plane + c] += rho;
    }
    w.gshot[c] = vp;
    w.gshot[g.plane + c] = vs;
    w.gshot[2 * g.plane + c] = rho;
    f.glam[c] = 0.f;
    f.gmu[c] = 0.f;
    f.grho[c] = 0.f;
}

} // namespace


// ------------------------------------------------ FWI residuals on the device
//
// calc_res (src/calc_res.c:127-301) on the shot's resident sections, one
// thread per trace walking its samples in order, so that the running sums
// (the time integral when VELOCITY = 0, the trace norms of LNORM 5 / 7) and
// every sample are the host restatement's float operations in the same order
// (no contraction into fused multiply-adds, which the host build does not
// form either).

// one TIME_FILT pass (src/timedomain_filt.c:28-113) over each trace of in:
// the Butterworth sections in cascade, in double, from a zero state, the
// result rounded to float and written time-reversed into out (the reference
// filters its buffer, rounds to float and reverses it, twice)
struct TFilt {
    int n;
    double f0[SOFI_IIR_MAXSEC], f1[SOFI_IIR_MAXSEC], f2[SOFI_IIR_MAXSEC], g1[SOFI_IIR_MAXSEC], g2[SOFI_IIR_MAXSEC];
};

__global__ void k_tfilt(const float *in, float *out, int ntr, int ns, TFilt f)
{
#pragma clang fp contract(off)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ntr) return;
    double x1[SOFI_IIR_MAXSEC], x2[SOFI_IIR_MAXSEC], y1[SOFI_IIR_MAXSEC], y2[SOFI_IIR_MAXSEC];
#pragma unroll
    for (int q = 0; q < SOFI_IIR_MAXSEC; ++q) x1[q] = x2[q] = y1[q] = y2[q] = 0.0;
    const float *a = in + (size_t)r * ns;
    float *b = out + (size_t)r * ns;
    for (int k = 0; k < ns; ++k) {
        double x = (double)a[k];
#pragma unroll
        for (int q = 0; q < SOFI_IIR_MAXSEC; ++q) {
            if (q >= f.n) break;
            const double y = f.f0[q] * x + f.f1[q] * x1[q] + f.f2[q] * x2[q] + f.g1[q] * y1[q] + f.g2[q] * y2[q];
            x2[q] = x1[q];
            x1[q] = x;
            y2[q] = y1[q];
            y1[q] = y;
            x = y;
        }
        b[ns - 1 - k] = (float)x;
    }
}

// the adjoint source of trace r, time-reversed, into adj[r][nt] (the first
// ns samples), and the trace's share of the misfit calc_res accumulates:
// LNORM 1 sign(measured - synthetic); 2 synthetic - measured, integrated
// (running sum times DT) when VELOCITY = 0; 3 Cauchy r / (1 + r^2); 4
// tanh(r); 5 / 7 the normalised correlation's derivative of the
// (integrated) traces, a zero trace norm taken as 1; 8 the envelope residual
// with the envelopes the reference never fills, s / (0 + WATERLEVEL) * (0 - 0) - 0
__global__ void k_residual(const float *syn, const float *obs, float *adj, float *partial, int ntr, int ns, int nt,
                           int lnorm, int velocity, float dt, float wl)
{
#pragma clang fp contract(off)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ntr) return;
    const float *s = syn + (size_t)r * ns, *sd = obs + (size_t)r * ns;
    float *df = adj + (size_t)r * nt;
    const float RMS = 1.0f;
    float abs_sd = 0.0f, abs_s = 0.0f, mult = 0.0f;
    if (lnorm == 5 || lnorm == 7) {
        float acc_s = 0.0f, acc_sd = 0.0f;
        for (int j = 0; j < ns; ++j) {
            float is = s[j], isd = sd[j];
            if (velocity == 0) {
                acc_s += s[j];
                is = acc_s * dt;
                acc_sd += sd[j];
                isd = acc_sd * dt;
            }
            abs_sd += isd * isd;
            abs_s += is * is;
            mult += isd * is;
        }
        abs_sd = abs_sd == 0 ? 1.0f : (float)sqrt((double)abs_sd);
        abs_s = abs_s == 0 ? 1.0f : (float)sqrt((double)abs_s);
    }
    float misfit = 0.0f, acc_s = 0.0f, acc_sd = 0.0f;
    for (int j = 1; j <= ns; ++j) {
        float is = s[j - 1], isd = sd[j - 1];
        if (velocity == 0) {
            acc_s += s[j - 1];
            is = acc_s * dt;
            acc_sd += sd[j - 1];
            isd = acc_sd * dt;
        }
        float v = 0.0f;
        const float rr = (sd[j - 1] - s[j - 1]) / RMS;
        switch (lnorm) {
          case 1: v = (rr > 0 ? 1.0f : rr < 0 ? -1.0f : 0.0f) / RMS; break;
          case 2: v = is - isd; break;
          case 3: v = (rr / (1 + (rr * rr))) / RMS; break;
          case 4: v = (float)tanh((double)rr) / RMS; break;
          case 5:
          case 7: v = ((is * mult) / (abs_s * abs_s * abs_s * abs_sd)) - (isd / (abs_s * abs_sd)); break;
          case 8: v = is / (0.0f + wl) * (0.0f - 0.0f) - 0.0f; break;
        }
        df[ns - j] = v;
        if (lnorm == 2) misfit += v * v;
        if (lnorm == 5 || lnorm == 7) misfit -= (isd * is) / (abs_sd * abs_s);
        if (lnorm == 8) misfit += (0.0f - 0.0f) * (0.0f - 0.0f);
    }
    partial[r] = misfit;
}

// the traces' misfit shares summed in trace order into *out
__global__ void k_misfit_sum(const float *partial, int ntr, float *out)
{
#pragma clang fp contract(off)
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float m = 0.0f;
    for (int r = 0; r < ntr; ++r) m += partial[r];
    *out = m;
}

// n floats set to v (a fill kernel rather than hipMemsetAsync, which must
// not be recorded where a caller captures the stream into a graph)
__global__ void k_fill(float *p, size_t n, float v)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// ====================================================================== handle

struct sofi_dev {
    sofi_desc d;
    Grid g;
    Fields f;
    Mat m;
    Src s;
    Rec r;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    struct Alloc {
        void *user, *raw;
        size_t bytes;
    };
    std::vector<Alloc> allocs;
    size_t bytes = 0;
    // device-side copies owned by the handle
    float *mat[SOFI_A_NARRAYS] = {};
    float *base_d = nullptr, *base_e = nullptr, *base_r = nullptr;
    float *prof = nullptr;
    float *surf = nullptr;
    // acquisition of every shot slot (host copies); combined into the device
    // lists Src / Rec by sync_acq before the next time step
    struct Slot {
        std::vector<int> si, sj, st, ri, rj;
        std::vector<float> azi, amp, sig;
        int source_type = 0;
    };
    std::vector<Slot> slots;
    int nb = 1, sel = 0;                   // slots of the handle; the slot the per-shot calls address
    bool src_dirty = true, rec_dirty = true;
    int *src_ibuf = nullptr;               // device: i, j, type, slot of every source
    float *src_fbuf = nullptr, *src_sig = nullptr;
    double *src_sc = nullptr;
    int source_type = 0;                   // slot 0's SOURCE_TYPE (one-shot handles)
    Inj src_inj{}, psrc_inj{};             // all sources / explosive ones, grouped by cell
    int *src_inj_buf = nullptr, *psrc_inj_buf = nullptr;
    int *eq_list = nullptr;                // moment-tensor sources, in order
    int n_eq = 0;
    bool any_force = false;                // a source of type 2, 3 or 4 (force)
    Fuse fz{};                             // cells of the sources when nsrc <= FUSE_MAX (fz.nforce, npsrc unset)
    bool fusable = false;
    // a batch whose slots each have at most FUSE_MAX sources and no moment
    // sources: launches of one slot carry that slot's sources (slot_fz[k])
    bool slot_fusable = false;
    std::vector<Fuse> slot_fz;
    int shape_v = 0, shape_s = 0;          // velocity / stress kernel shapes: SOFI_LAUNCH_*
    int bs = 0;                            // a batch's slot blocks: slots per thread of k_*_b (0: slot kernels)
    // a batch's launch configuration as the autotuner chooses it
    struct BatchCfg {
        int bs, res, streams, grp, spt;    // slot blocks, resident group, streams, slots per launch / thread
    };
    bool tune_on = true, tuned = false;    // measured on the first run unless fixed through the API
    BatchCfg tune_cfg{};
    int spt_mode = 0, grp_mode = 0;        // a batch's slots per thread and per launch (0: tuned in the first run)
    int grp = 0;                           // slots per launch of the update kernels (0: all)
    bool spt_tuned = false;
    hipEvent_t tune0 = nullptr, tune1 = nullptr;
    int *rec_ibuf = nullptr;               // device: i, j, slot of every receiver
    long long *rec_lbuf = nullptr;         // device: trace offset and component stride of every receiver
    double *rec_d = nullptr;
    float *sec = nullptr;                  // sections of all slots
    std::vector<long long> sec_base;       // start of each slot's [5][ntr][ns] sections
    size_t sec_len = 0;
    float *snapbuf = nullptr;
    int nsnap_max = 0, nsnap = 0, nxs = 0, nys = 0;
    int lsamp = 0, lsnap = 0;
    bool medium_set = false;
    // sampled per-kernel timing (HIP events around the velocity and stress kernels)
    int time_every = 0;                    // 0 = off; else every time_every-th step is timed
    std::vector<hipEvent_t> tev;           // 4 events per timed step
    int tev_used = 0;
    double t_vel_ms = 0.0, t_str_ms = 0.0;
    int t_samples = 0;
    // host copies needed for receiver factors
    std::vector<float> h_ppi, h_pu;
    // FWI (desc.mode == SOFI_MODE_FWI)
    Fwi w{};
    int nstore = 0;
    int sw = 0;                            // 0 forward pass (store), 1 adjoint pass (correlate)
    bool model_set = false;
    Inj adj_inj{};
    int *adj_inj_buf = nullptr;
    float *adj_sx = nullptr, *adj_sy = nullptr;
    int nadj = 0;
    std::vector<int> adj_start;            // first adjoint row of each slot (receivers slot after slot)
    bool adj_at_rec = false;               // the adjoint arrays and plan are those of the slots' receivers
    bool adj_custom = false;               // the adjoint sources were set by sofi_fwi_set_adjoint
    float *rbuf = nullptr;                 // workspace of sofi_fwi_residual (measured data, filtered copies, shares)
    size_t rbuf_len = 0;
    int dev = 0;                           // the device the handle was opened on
    // captured blocks of time steps: one graph of graph_k steps whose kernels
    // read the block's base step from ntb, replayed for every aligned block
    int *ntb = nullptr;
    int graph_mode = SOFI_GRAPH_AUTO, graph_steps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t gbuild = nullptr;           // while a block is recorded: the graph its launches go to
    hipGraphNode_t glast = nullptr;        // the last node recorded (each node follows it)
    // argument values and pointer arrays of the recorded kernel nodes, held
    // until the graph is destroyed (nothing that a node's parameters point
    // to may die before the node: a runtime or tool that reads them after
    // hipGraphAddKernelNode returns finds them alive)
    struct NodeArgs {
        std::shared_ptr<void> vals;
        std::vector<void *> ptrs;
    };
    std::vector<NodeArgs> gargs;
    hipError_t gerr = hipSuccess;
    int graph_k = 0;
    unsigned gen = 1, graph_gen = 0;       // launch configuration: current / of the captured graph
    long long graph_blocks = 0;            // blocks replayed since the handle opened
    // a batch's resident groups: res_mode slots (0: chosen from the cache
    // budget) advanced through a run's whole step range before the next
    // group; the slots of the group being advanced, and where each slot's
    // entries start in the injection and receiver lists
    // 4th order in time: the rotating histories
    bool fdt4 = false;
    Hist4 h4{};
    int res_mode = 0;
    int cur_lo = 0, cur_hi = 1;
    // resident groups on several streams at once: the extra streams (the
    // handle's own is the first), their join events, the fork event
    int str_mode = 0;
    hipStream_t cur = nullptr;             // the stream launches go to (null: the handle's)
    std::vector<hipStream_t> xs;
    std::vector<hipEvent_t> xev;
    hipEvent_t evfork = nullptr;
    std::vector<int> src_inj_start, psrc_inj_start, rec_start;
    // sofi_run_many: while a step of the handle is recorded for the work-list
    // launches (wl), its update kernels' arguments go to wl_shot and its other
    // launches to *defer, tagged with the stage they belong to (0 before the
    // velocity update, 1 between the updates, 2 after the stress update)
    struct Deferred {
        int stage;
        std::function<hipError_t(hipStream_t)> run;
    };
    bool wl = false;
    Shot wl_shot{};
    int wl_dv = -1, wl_seis = -1;          // this step's div/curl and vx/vy/p sample index (-1: none)
    std::vector<Deferred> *defer = nullptr;
    int dstage = 0;
};

namespace {

// Every device array starts 224 B past a 256-B boundary: a padded plane's
// cell (1, 1) lies 8 rows and 8 columns (32 B) past the array start and the
// row pitch is a multiple of 64 floats, so every interior row then begins on
// a 256-B boundary and a wavefront's 64 cells of a row are two whole 128-B
// cache lines.
constexpr size_t ALIGN_SHIFT = 224;

// the tiles of the TTI frame passes: tiles holding a cell of the absorbing
// frame or one next to it (the frame passes read the derivatives of the
// frame cells' neighbours); rows 1..GY[2]+1 and GY[3]..NY, columns
// 1..GX[2]+1 and GX[3]..NX
void frame_tiles(Grid &g)
{
    g.ftop = g.gy2 > 0 ? g.gy2 / BY + 1 : 0;
    g.fbot = g.gy3 < g.ny ? g.nty - (g.gy3 - 1) / BY : 0;
    g.fl = g.gx2 > 0 ? g.gx2 / BX + 1 : 0;
    g.fr = g.gx3 < g.nx ? g.ntx - (g.gx3 - 1) / BX : 0;
    if (g.ftop > g.nty) g.ftop = g.nty;
    if (g.ftop + g.fbot > g.nty) g.fbot = g.nty - g.ftop;
    if (g.fl > g.ntx) g.fl = g.ntx;
    if (g.fl + g.fr > g.ntx) g.fr = g.ntx - g.fl;
}

int dalloc(sofi_dev *h, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 4;
    void *raw = nullptr;
    hipError_t e = hipMalloc(&raw, bytes + 256);
    if (e != hipSuccess) return fail("hipMalloc(%zu) failed: %s", bytes + 256, hipGetErrorString(e));
    if ((uintptr_t)raw % 256) {
        (void)hipFree(raw);
        return fail("hipMalloc returned a pointer that is not 256-B aligned");
    }
    // zeroed on the handle's stream and waited for: the handle's streams are
    // non-blocking, so a fill on the null stream would not be ordered before
    // the kernels and copies that follow on them
    e = hipMemsetAsync(raw, 0, bytes + 256, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        (void)hipFree(raw);
        return fail("zeroing a new allocation failed: %s", hipGetErrorString(e));
    }
    *p = (char *)raw + ALIGN_SHIFT;
    h->allocs.push_back({ *p, raw, bytes + 256 });
    h->bytes += bytes + 256;
    return 0;
}

// host -> device copy ordered on the handle's stream and waited for (a copy
// on the null stream is not ordered with the handle's non-blocking stream)
int h2d(sofi_dev *h, void *dst, const void *src, size_t bytes)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// release one allocation owned by the handle (per-shot buffers replaced on re-upload)
template <typename T>
void dfree(sofi_dev *h, T **p)
{
    if (!*p) return;
    for (size_t k = 0; k < h->allocs.size(); ++k)
        if (h->allocs[k].user == (void *)*p) {
            h->bytes -= h->allocs[k].bytes;
            (void)hipFree(h->allocs[k].raw);
            h->allocs.erase(h->allocs.begin() + k);
            break;
        }
    *p = nullptr;
}

template <typename T>
int dalloc_t(sofi_dev *h, T **p, size_t n)
{
    return dalloc(h, (void **)p, n * sizeof(T));
}

int upload(sofi_dev *h, float **dst, const float *src, size_t n)
{
    if (!*dst && dalloc_t(h, dst, n)) return -1;
    if (h2d(h, *dst, src, n * sizeof(float))) return -1;
    return 0;
}

// Groups sources by cell (first occurrence order of the cells, source order
// within a cell) into one device buffer [cell | first | idx].
// Entry k is cell (si[k], sj[k]) of shot slot slots[k] (slot 0 if slots is
// null) and records source index ids[k] (k if ids is null).
int build_inj(sofi_dev *h, int n, const int *si, const int *sj, const int *slots, const int *ids, Inj *inj,
              int **buf, std::vector<int> *starts = nullptr)
{
    dfree(h, buf);
    *inj = Inj{};
    std::vector<int> cells, order;
    std::vector<std::vector<int>> members;
    std::unordered_map<int, size_t> slot;
    for (int k = 0; k < n; ++k) {
        const int c = (slots ? slots[k] : 0) * h->g.plane + (int)sofi_idx(h->g.pitch, si[k], sj[k]);
        auto it = slot.find(c);
        if (it == slot.end()) {
            it = slot.emplace(c, cells.size()).first;
            cells.push_back(c);
            members.emplace_back();
        }
        members[it->second].push_back(ids ? ids[k] : k);
    }
    std::vector<int> first(1, 0);
    for (const auto &mbr : members) {
        order.insert(order.end(), mbr.begin(), mbr.end());
        first.push_back((int)order.size());
    }
    const int nc = (int)cells.size();
    std::vector<int> all;
    all.insert(all.end(), cells.begin(), cells.end());
    all.insert(all.end(), first.begin(), first.end());
    all.insert(all.end(), order.begin(), order.end());
    if (dalloc_t(h, buf, all.size())) return -1;
    if (h2d(h, *buf, all.data(), all.size() * sizeof(int))) return -1;
    inj->n = nc;
    inj->cell = *buf;
    inj->first = *buf + nc;
    inj->idx = *buf + 2 * nc + 1;
    if (starts) {
        // cells appear slot by slot (the slots' sources are listed in slot
        // order): where each slot's cells start
        starts->assign(h->nb + 1, nc);
        for (int q = nc - 1; q >= 0; --q) (*starts)[cells[q] / h->g.plane] = q;
        for (int k = h->nb - 1; k >= 0; --k) (*starts)[k] = std::min((*starts)[k], (*starts)[k + 1]);
    }
    return 0;
}

bool planes_of(int id) { return id == SOFI_A_DIP || id == SOFI_A_D || id == SOFI_A_E || (id >= SOFI_A_C11D && id <= SOFI_A_C35IPJPD); }

} // namespace

extern "C" {

const char *sofi_last_error(void) { return g_err.c_str(); }

int sofi_open(const sofi_desc *desc, sofi_dev **out) { return sofi_open_batch(desc, 1, out); }

int sofi_open_batch(const sofi_desc *desc, int nshots, sofi_dev **out)
{
    if (!desc || !out) return fail("sofi_open: null argument");
    if (nshots < 1 || nshots > 4096) return fail("sofi_open_batch: %d shot slots (1..4096)", nshots);
    const sofi_desc &d = *desc;
    if ((double)nshots * (double)sofi_plane(d.nx, d.ny) * (d.L > 0 ? d.L : 1) > 2147483647.0)
        return fail("sofi_open_batch: %d slots of a %d x %d grid exceed the 2^31-float index range", nshots, d.nx, d.ny);
    if (d.nx < 1 || d.ny < 1) return fail("sofi_open: bad grid %d x %d", d.nx, d.ny);
    if (d.fdorder < 2 || d.fdorder > 12 || (d.fdorder & 1)) return fail("sofi_open: unsupported FDORDER %d", d.fdorder);
    if ((d.weq < SOFI_EL_ISO || d.weq > SOFI_VEL_TTI) && d.weq != SOFI_AC_ISO)
        return fail("sofi_open: unsupported WEQ %d", d.weq);
    if (d.weq == SOFI_AC_ISO && (d.free_surf || d.mode == SOFI_MODE_FWI || d.fdorder_time == 4))
        return fail("sofi_open: AC_ISO runs without free surface, FWI or 4th order in time (as the reference)");
    if (d.L < 0 || d.L > SOFI_MAXL) return fail("sofi_open: L=%d out of range", d.L);
    bool visco = d.weq == SOFI_VEL_ISO || d.weq == SOFI_VEL_VTI || d.weq == SOFI_VEL_TTI;
    if (visco && d.L < 1) return fail("sofi_open: L>0 required for viscoelastic rheologies");
    if (d.fw < 0 || 2 * d.fw > d.nx || 2 * d.fw > d.ny) return fail("sofi_open: FW=%d does not fit the grid", d.fw);
    if (d.fdorder_time != 0 && d.fdorder_time != 2 && d.fdorder_time != 4)
        return fail("sofi_open: FDORDER_TIME=%d (2 or 4)", d.fdorder_time);
    if (d.fdorder_time == 4) {
        if (d.weq != SOFI_EL_ISO && d.weq != SOFI_VEL_ISO) return fail("WEQ not yet implemented for 4th order in time.");
        if (nshots != 1) return fail("sofi_open_batch: 4th order in time runs one shot per handle");
        if (d.mode == SOFI_MODE_FWI) return fail("sofi_open: MODE=FWI runs 2nd order in time");
        if (d.fw > 0 && d.abs_type != 1 && d.abs_type != 2)
            return fail("sofi_open: 4th order in time needs ABS_TYPE 1 or 2 with FW > 0");
    }
    if (d.free_surf && (d.weq == SOFI_VEL_VTI || d.weq == SOFI_VEL_TTI))
        return fail("sofi_open: free surface with VEL_VTI/VEL_TTI is not supported (the reference reads unallocated tau arrays there)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("sofi_open: no HIP device available");
    sofi_dev *h = new (std::nothrow) sofi_dev();
    if (!h) return fail("out of memory");
    h->d = d;
    h->nb = nshots;
    h->cur_hi = nshots;
    h->slots.resize(nshots);
    Grid &g = h->g;
    g.nb = nshots;
    g.spt = 1;
    g.k_lo = 0;
    g.k_hi = nshots;
    g.nx = d.nx;
    g.ny = d.ny;
    g.pitch = sofi_pitch(d.nx);
    g.plane = (int)sofi_plane(d.nx, d.ny);
    g.fw = d.fw;
    g.abs_type = d.abs_type;
    g.free_surf = d.free_surf;
    g.weq = d.weq;
    g.L = visco ? d.L : 0;
    g.nd = d.fdorder / 2;
    g.gx2 = d.gx[2];
    g.gx3 = d.gx[3];
    g.gy2 = d.gy[2];
    g.gy3 = d.gy[3];
    g.wx = 2 * d.fw + 1;
    g.wy = d.nx + 1;
    g.ntx = (d.nx + BX - 1) / BX;
    g.nty = (d.ny + BY - 1) / BY;
    g.periodic = d.boundary == 1;
    frame_tiles(g);
    g.nxp = (d.ny + 1) * g.wx;
    g.nyp = g.wx * g.wy;
    g.dt = d.dt;
    g.dh = d.dh;
    g.dtdh = d.dt / d.dh;
    g.dhi = 1.0f / d.dh;
    g.dhi4 = (float)(1.0 / (double)d.dh);
    g.dthalbe = (float)(d.dt / 2.0);
    for (int k = 0; k < 7; ++k) g.hc[k] = d.hc[k];
    if (hipGetDevice(&h->dev) != hipSuccess) {
        sofi_close(h);
        return fail("sofi_open: hipGetDevice failed");
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
        hipEventCreate(&h->tune0) != hipSuccess || hipEventCreate(&h->tune1) != hipSuccess) {
        sofi_close(h);
        return fail("sofi_open: stream/event creation failed");
    }
    const size_t pl = (size_t)g.plane, nb = (size_t)nshots;
    Fields &f = h->f;
    int rc = 0;
    rc |= dalloc_t(h, &f.vx, pl * nb);
    rc |= dalloc_t(h, &f.vy, pl * nb);
    rc |= dalloc_t(h, &f.sxx, pl * nb);
    // the acoustic path keeps one pressure array: syy is sxx, as psyy is
    // psxx in the reference (src/initmem_wavefield.c:31-33); sxy stays zero
    if (d.weq == SOFI_AC_ISO) f.syy = f.sxx;
    else rc |= dalloc_t(h, &f.syy, pl * nb);
    rc |= dalloc_t(h, &f.sxy, pl * nb);
    rc |= dalloc_t(h, &f.unstable, nb);
    rc |= dalloc_t(h, &h->ntb, 1);
    if (visco) {
        rc |= dalloc_t(h, &f.r, pl * g.L * nb);
        rc |= dalloc_t(h, &f.p, pl * g.L * nb);
        rc |= dalloc_t(h, &f.q, pl * g.L * nb);
    }
    if (d.fdorder_time == 4) {
        h->fdt4 = true;
        Hist4 &H = h->h4;
        for (int k = 0; k < 4; ++k) {
            rc |= dalloc_t(h, &H.sx[k], pl);
            rc |= dalloc_t(h, &H.sy[k], pl);
            for (int q = 0; q < 4; ++q) rc |= dalloc_t(h, &H.dv[q][k], pl);
        }
        if (visco) {
            H.r[0] = f.r;
            H.p[0] = f.p;
            H.q[0] = f.q;
            for (int k = 1; k < 4; ++k) {
                rc |= dalloc_t(h, &H.r[k], pl * g.L);
                rc |= dalloc_t(h, &H.p[k], pl * g.L);
                rc |= dalloc_t(h, &H.q[k], pl * g.L);
            }
        }
    }
    if (d.weq == SOFI_EL_TTI || d.weq == SOFI_VEL_TTI) {
        rc |= dalloc_t(h, &f.dvxx, pl * nb);
        rc |= dalloc_t(h, &f.dvyy, pl * nb);
        rc |= dalloc_t(h, &f.dvyx, pl * nb);
        rc |= dalloc_t(h, &f.dvxy, pl * nb);
    }
    if (d.abs_type == 1 && d.fw > 0) {
        const size_t nxp = (size_t)g.nxp * nb, nyp = (size_t)g.nyp * nb;
        rc |= dalloc_t(h, &f.psi_sxx_x, nxp);
        rc |= dalloc_t(h, &f.psi_sxy_x, nxp);
        rc |= dalloc_t(h, &f.psi_vxx, nxp);
        rc |= dalloc_t(h, &f.psi_vyx, nxp);
        rc |= dalloc_t(h, &f.psi_syy_y, nyp);
        rc |= dalloc_t(h, &f.psi_sxy_y, nyp);
        rc |= dalloc_t(h, &f.psi_vyy, nyp);
        rc |= dalloc_t(h, &f.psi_vxy, nyp);
    }
    if (d.mode == SOFI_MODE_FWI) {
        if (d.weq == SOFI_AC_ISO) {
            sofi_close(h);
            return fail("sofi_open: MODE=FWI needs an elastic or viscoelastic WEQ (calc_conv reads sxy)");
        }
        if (d.adjoint_type < 1 || d.adjoint_type > 3) {
            sofi_close(h);
            return fail("sofi_open: ADJOINT_TYPE=%d is not supported (1, 2 or 3)", d.adjoint_type);
        }
        const int dtinv = d.dtinv > 0 ? d.dtinv : 1;
        h->nstore = (d.nt + dtinv - 1) / dtinv;
        Fwi &w = h->w;
        w.velocity = d.velocity;
        // per shot slot: the FWI planes of Fields and a store of its own
        for (float **q : { &f.pvxp1, &f.pvyp1, &f.ux, &f.uy, &f.uxy, &f.glam, &f.gmu, &f.grho })
            rc |= dalloc_t(h, q, pl * nb);
        rc |= dalloc_t(h, &w.gsum, 3 * pl) | dalloc_t(h, &w.gshot, 3 * pl);
        w.sstride = (size_t)h->nstore * 5 * d.nx * d.ny;
        rc |= dalloc_t(h, &w.store, w.sstride * nb);
    }
    if (rc) {
        std::string e = g_err;
        sofi_close(h);
        g_err = e;
        return -1;
    }
    h->lsamp = d.ndt;
    h->lsnap = d.lsnap;
    *out = h;
    return 0;
}

int sofi_set_medium(sofi_dev *h, const float *const *arrays, const float *const *profiles)
{
    if (!h || !arrays) return fail("sofi_set_medium: null argument");
    const sofi_desc &d = h->d;
    const Grid &g = h->g;
    size_t pl = (size_t)g.plane;
    Mat &m = h->m;
    auto need = [&](int id) -> int {
        if (!arrays[id]) return fail("sofi_set_medium: material array %d required for WEQ %d is missing", id, d.weq);
        return upload(h, &h->mat[id], arrays[id], planes_of(id) ? pl * g.L : pl);
    };
    int rc = 0;
    rc |= need(SOFI_A_RIP);
    rc |= need(SOFI_A_RJP);
    if (d.weq != SOFI_AC_ISO) rc |= need(SOFI_A_PU);
    rc |= need(SOFI_A_PPI);
    if (rc) return -1;
    switch (d.weq) {
      case SOFI_AC_ISO:
          rc |= need(SOFI_A_G);
          break;
      case SOFI_EL_ISO:
          rc |= need(SOFI_A_F) | need(SOFI_A_G) | need(SOFI_A_FIPJP);
          break;
      case SOFI_VEL_ISO: {
          rc |= need(SOFI_A_F) | need(SOFI_A_G) | need(SOFI_A_FIPJP) | need(SOFI_A_PTAUS) | need(SOFI_A_PTAUP);
          if (rc) return -1;
          for (int id : { SOFI_A_PUIPJP, SOFI_A_PTAUSIPJP })
              if (!arrays[id]) return fail("sofi_set_medium: array %d required for VEL_ISO is missing", id);
          // per-cell defect bases: d_l = eta_l*pu*taus, e_l = eta_l*ppi*taup, dip_l = eta_l*puipjp*tausipjp
          std::vector<float> bd(pl), be(pl), br(pl);
          const float *pu = arrays[SOFI_A_PU], *ppi = arrays[SOFI_A_PPI], *ts = arrays[SOFI_A_PTAUS],
                      *tp = arrays[SOFI_A_PTAUP], *pui = arrays[SOFI_A_PUIPJP], *tsi = arrays[SOFI_A_PTAUSIPJP];
          for (size_t k = 0; k < pl; ++k) {
              bd[k] = pu[k] * ts[k];
              be[k] = ppi[k] * tp[k];
              br[k] = pui[k] * tsi[k];
          }
          rc |= upload(h, &h->base_d, bd.data(), pl) | upload(h, &h->base_e, be.data(), pl) | upload(h, &h->base_r, br.data(), pl);
          break;
      }
      case SOFI_EL_VTI:
          rc |= need(SOFI_A_C11) | need(SOFI_A_C13) | need(SOFI_A_C33) | need(SOFI_A_C55IPJP);
          break;
      case SOFI_VEL_VTI:
          rc |= need(SOFI_A_C11U) | need(SOFI_A_C13U) | need(SOFI_A_C33U) | need(SOFI_A_C55IPJPU);
          rc |= need(SOFI_A_C11D) | need(SOFI_A_C13D) | need(SOFI_A_C33D) | need(SOFI_A_C55IPJPD);
          break;
      case SOFI_EL_TTI:
          rc |= need(SOFI_A_C11) | need(SOFI_A_C13) | need(SOFI_A_C33) | need(SOFI_A_C15) | need(SOFI_A_C35);
          rc |= need(SOFI_A_C55IPJP) | need(SOFI_A_C15IPJP) | need(SOFI_A_C35IPJP);
          break;
      case SOFI_VEL_TTI:
          rc |= need(SOFI_A_C11U) | need(SOFI_A_C13U) | need(SOFI_A_C33U) | need(SOFI_A_C15U) | need(SOFI_A_C35U);
          rc |= need(SOFI_A_C55IPJPU) | need(SOFI_A_C15IPJPU) | need(SOFI_A_C35IPJPU);
          rc |= need(SOFI_A_C11D) | need(SOFI_A_C13D) | need(SOFI_A_C33D) | need(SOFI_A_C15D) | need(SOFI_A_C35D);
          rc |= need(SOFI_A_C55IPJPD) | need(SOFI_A_C15IPJPD) | need(SOFI_A_C35IPJPD);
          break;
    }
    if (rc) return -1;
    if (d.abs_type == 2 && d.fw > 0) {
        if (need(SOFI_A_ABSORB)) return -1;
    }
    if (d.abs_type == 1 && d.fw > 0) {
        if (!profiles) return fail("sofi_set_medium: CPML profiles missing");
        int n = 2 * d.fw + 1;
        std::vector<float> pr((size_t)12 * n);
        for (int k = 0; k < 12; ++k) {
            if (!profiles[k]) return fail("sofi_set_medium: CPML profile %d missing", k);
            memcpy(pr.data() + (size_t)k * n, profiles[k], n * sizeof(float));
        }
        if (upload(h, &h->prof, pr.data(), pr.size())) return -1;
    }
    m = Mat();
    m.rip = h->mat[SOFI_A_RIP];
    m.rjp = h->mat[SOFI_A_RJP];
    m.f = h->mat[SOFI_A_F];
    m.g = h->mat[SOFI_A_G];
    m.fipjp = h->mat[SOFI_A_FIPJP];
    m.bd = h->base_d;
    m.be = h->base_e;
    m.br = h->base_r;
    m.c11 = h->mat[SOFI_A_C11];
    m.c13 = h->mat[SOFI_A_C13];
    m.c33 = h->mat[SOFI_A_C33];
    m.c15 = h->mat[SOFI_A_C15];
    m.c35 = h->mat[SOFI_A_C35];
    m.c55i = h->mat[SOFI_A_C55IPJP];
    m.c15i = h->mat[SOFI_A_C15IPJP];
    m.c35i = h->mat[SOFI_A_C35IPJP];
    m.u11 = h->mat[SOFI_A_C11U];
    m.u13 = h->mat[SOFI_A_C13U];
    m.u33 = h->mat[SOFI_A_C33U];
    m.u15 = h->mat[SOFI_A_C15U];
    m.u35 = h->mat[SOFI_A_C35U];
    m.u55i = h->mat[SOFI_A_C55IPJPU];
    m.u15i = h->mat[SOFI_A_C15IPJPU];
    m.u35i = h->mat[SOFI_A_C35IPJPU];
    m.d11 = h->mat[SOFI_A_C11D];
    m.d13 = h->mat[SOFI_A_C13D];
    m.d33 = h->mat[SOFI_A_C33D];
    m.d15 = h->mat[SOFI_A_C15D];
    m.d35 = h->mat[SOFI_A_C35D];
    m.d55i = h->mat[SOFI_A_C55IPJPD];
    m.d15i = h->mat[SOFI_A_C15IPJPD];
    m.d35i = h->mat[SOFI_A_C35IPJPD];
    m.absorb = h->mat[SOFI_A_ABSORB];
    m.pu = h->mat[SOFI_A_PU];
    m.ppi = h->mat[SOFI_A_PPI];
    m.ptaus = h->mat[SOFI_A_PTAUS];
    m.ptaup = h->mat[SOFI_A_PTAUP];
    m.prof = h->prof;
    if (d.free_surf) {
        // row-1 coefficients of the free surface (see SURF_*)
        const float *pu = arrays[SOFI_A_PU], *ppi = arrays[SOFI_A_PPI];
        const float *ts = arrays[SOFI_A_PTAUS], *tp = arrays[SOFI_A_PTAUP];
        const bool visco = g.L > 0;
        if (visco && (!ts || !tp)) return fail("sofi_set_medium: the free surface needs the tau arrays");
        std::vector<float> sc((size_t)SURF_N * d.nx, 0.0f);
        for (int i = 1; i <= d.nx; ++i) {
            const size_t c = sofi_idx(g.pitch, i, 1);
            float fjm, gg;
            if (!visco) {
                fjm = pu[c] * 2.0f;
                gg = ppi[c];
            } else {
                fjm = (float)(pu[c] * 2.0 * (1.0 + g.L * ts[c]));
                gg = (float)(ppi[c] * (1.0 + g.L * tp[c]));
                const float djm = (float)(2.0 * pu[c] * ts[c]);
                const float e = ppi[c] * tp[c];
                sc[(size_t)SURF_DE * d.nx + i - 1] = djm - e;
                sc[(size_t)SURF_FG * d.nx + i - 1] = (fjm / gg) - 1.0f;
            }
            sc[(size_t)SURF_T1 * d.nx + i - 1] = g.dt * (gg - fjm) * (gg - fjm);
            sc[(size_t)SURF_T2 * d.nx + i - 1] = g.dt * (gg - fjm);
            sc[(size_t)SURF_G * d.nx + i - 1] = gg;
        }
        if (upload(h, &h->surf, sc.data(), sc.size())) return -1;
        m.surf = h->surf;
    }
    if (g.L > 0) {
        if (!profiles || !profiles[SOFI_P_ETA] || !profiles[SOFI_P_BIP] || !profiles[SOFI_P_CIP])
            return fail("sofi_set_medium: relaxation profiles (eta, bip, cip) missing");
        for (int l = 1; l <= g.L; ++l) {
            m.eta[l] = profiles[SOFI_P_ETA][l];
            m.bip[l] = profiles[SOFI_P_BIP][l];
            m.cip[l] = profiles[SOFI_P_CIP][l];
            const float eta = m.eta[l];
            m.bjm[l] = (float)(eta / (1.0 + (eta * 0.5)));
        }
    }
    h->h_ppi.assign(arrays[SOFI_A_PPI], arrays[SOFI_A_PPI] + pl);
    if (arrays[SOFI_A_PU]) h->h_pu.assign(arrays[SOFI_A_PU], arrays[SOFI_A_PU] + pl);
    else h->h_pu.assign(pl, 0.0f);     // the acoustic path has no shear modulus
    h->medium_set = true;
    h->gen++;
    return 0;
}

int sofi_set_sources(sofi_dev *h, int nsrc, const int *src_i, const int *src_j, const int *src_type,
                     const float *src_azi, const float *src_amp, const float *signals, int source_type)
{
    if (!h) return fail("sofi_set_sources: null handle");
    if (nsrc < 0) return fail("sofi_set_sources: nsrc < 0");
    const sofi_desc &d = h->d;
    for (int k = 0; k < nsrc; ++k) {
        if (src_i[k] < 1 || src_i[k] > d.nx || src_j[k] < 1 || src_j[k] > d.ny)
            return fail("sofi_set_sources: source %d at (%d, %d) is outside the %d x %d grid", k, src_i[k], src_j[k], d.nx, d.ny);
    }
    sofi_dev::Slot &sl = h->slots[h->sel];
    sl.si.assign(src_i, src_i + nsrc);
    sl.sj.assign(src_j, src_j + nsrc);
    sl.st.assign(src_type, src_type + nsrc);
    sl.azi.assign(src_azi, src_azi + nsrc);
    sl.amp.assign(src_amp, src_amp + nsrc);
    sl.sig.assign(signals, signals + (size_t)nsrc * d.nt);
    sl.source_type = source_type;
    h->src_dirty = true;
    return 0;
}

int sofi_set_receivers(sofi_dev *h, int ntr, const int *rec_i, const int *rec_j)
{
    if (!h) return fail("sofi_set_receivers: null handle");
    if (!h->medium_set) return fail("sofi_set_receivers: call sofi_set_medium first");
    const sofi_desc &d = h->d;
    if (ntr < 0) return fail("sofi_set_receivers: ntr < 0");
    for (int k = 0; k < ntr; ++k)
        if (rec_i[k] < 1 || rec_i[k] > d.nx || rec_j[k] < 1 || rec_j[k] > d.ny)
            return fail("sofi_set_receivers: receiver %d at (%d, %d) is outside the grid", k, rec_i[k], rec_j[k]);
    sofi_dev::Slot &sl = h->slots[h->sel];
    sl.ri.assign(rec_i, rec_i + ntr);
    sl.rj.assign(rec_j, rec_j + ntr);
    h->rec_dirty = true;
    h->adj_at_rec = false;
    return 0;
}

int sofi_select_shot(sofi_dev *h, int k)
{
    if (!h) return fail("sofi_select_shot: null handle");
    if (k < 0 || k >= h->nb) return fail("sofi_select_shot: slot %d of %d", k, h->nb);
    h->sel = k;
    return 0;
}

} // extern "C"

namespace {

// The device lists of sources and receivers of all slots (rebuilt when a
// slot's acquisition changed).
int adj_at_receivers(sofi_dev *h);

int sync_acq(sofi_dev *h)
{
    if (!h->src_dirty && !h->rec_dirty) return 0;
    const sofi_desc &d = h->d;
    const Grid &g = h->g;
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->src_dirty) {
        std::vector<int> ci, cj, ct, cslot, gi, gj, pi, pj, eq;
        std::vector<float> cazi, camp, csig;
        std::vector<double> sc;
        for (int k = 0; k < h->nb; ++k) {
            const sofi_dev::Slot &sl = h->slots[k];
            for (size_t l = 0; l < sl.si.size(); ++l) {
                const int id = (int)ci.size();
                ci.push_back(sl.si[l]);
                cj.push_back(sl.sj[l]);
                ct.push_back(sl.st[l]);
                cslot.push_back(k);
                cazi.push_back(sl.azi[l]);
                camp.push_back(sl.amp[l]);
                const float azi = (float)(sl.azi[l] * M_PI / 180);
                sc.push_back(sin((double)azi));
                sc.push_back(cos((double)azi));
                if (sl.source_type == 5) eq.push_back(id);
            }
            csig.insert(csig.end(), sl.sig.begin(), sl.sig.end());
        }
        const int n = (int)ci.size();
        dfree(h, &h->src_ibuf);
        dfree(h, &h->src_fbuf);
        dfree(h, &h->src_sig);
        dfree(h, &h->src_sc);
        dfree(h, &h->eq_list);
        const size_t nn = (size_t)(n > 0 ? n : 1);
        int rc = dalloc_t(h, &h->src_ibuf, 4 * nn) | dalloc_t(h, &h->src_fbuf, 2 * nn) |
                 dalloc_t(h, &h->src_sig, nn * d.nt) | dalloc_t(h, &h->src_sc, 2 * nn) |
                 dalloc_t(h, &h->eq_list, eq.empty() ? 1 : eq.size());
        if (rc) return -1;
        if (n > 0) {
            std::vector<int> ib(ci);
            ib.insert(ib.end(), cj.begin(), cj.end());
            ib.insert(ib.end(), ct.begin(), ct.end());
            ib.insert(ib.end(), cslot.begin(), cslot.end());
            std::vector<float> fb(cazi);
            fb.insert(fb.end(), camp.begin(), camp.end());
            if (h2d(h, h->src_ibuf, ib.data(), ib.size() * sizeof(int))) return -1;
            if (h2d(h, h->src_fbuf, fb.data(), fb.size() * sizeof(float))) return -1;
            if (h2d(h, h->src_sig, csig.data(), csig.size() * sizeof(float))) return -1;
            if (h2d(h, h->src_sc, sc.data(), sc.size() * sizeof(double))) return -1;
        }
        if (!eq.empty()) if (h2d(h, h->eq_list, eq.data(), eq.size() * sizeof(int))) return -1;
        h->n_eq = (int)eq.size();
        Src &sv = h->s;
        sv.i = h->src_ibuf;
        sv.j = h->src_ibuf + nn;
        sv.type = h->src_ibuf + 2 * nn;
        sv.slot = h->src_ibuf + 3 * nn;
        sv.azi = h->src_fbuf;
        sv.amp = h->src_fbuf + nn;
        sv.sig = h->src_sig;
        sv.sc = h->src_sc;
        sv.n = n;
        sv.nt = d.nt;
        // injection plans over the slots' cells (cell offsets count the slots)
        std::vector<int> all_i, all_j, ps_i, ps_j, all_id, ps_id;
        h->any_force = false;
        for (int l = 0; l < n; ++l) {
            if (ct[l] >= 2 && ct[l] <= 4) h->any_force = true;
        }
        auto plan = [&](bool explosive_only, Inj *inj, int **buf, std::vector<int> *starts) -> int {
            std::vector<int> ii, jj, kk, idx;
            for (int l = 0; l < n; ++l) {
                if (explosive_only && h->slots[cslot[l]].source_type != 1) continue;
                ii.push_back(ci[l]);
                jj.push_back(cj[l]);
                kk.push_back(cslot[l]);
                idx.push_back(l);
            }
            return build_inj(h, (int)ii.size(), ii.data(), jj.data(), kk.data(), idx.data(), inj, buf, starts);
        };
        if (plan(false, &h->src_inj, &h->src_inj_buf, &h->src_inj_start) ||
            plan(true, &h->psrc_inj, &h->psrc_inj_buf, &h->psrc_inj_start))
            return -1;
        h->source_type = h->slots[0].source_type;
        h->fusable = h->nb == 1 && n <= FUSE_MAX;
        h->fz = Fuse{};
        if (h->fusable)
            for (int l = 0; l < n; ++l) h->fz.c[l] = (int)sofi_idx(g.pitch, ci[l], cj[l]);
        // per slot: the list holds the slots' sources slot after slot
        h->slot_fusable = h->nb > 1;
        h->slot_fz.assign(h->nb, Fuse{});
        for (int k = 0, l0 = 0; k < h->nb; ++k) {
            const int nk = (int)h->slots[k].si.size();
            Fuse &fk = h->slot_fz[k];
            fk.soff = l0;
            if (nk > FUSE_MAX || h->slots[k].source_type == 5) h->slot_fusable = false;
            bool force = false;
            for (int l = l0; l < l0 + nk && l - l0 < FUSE_MAX; ++l) {
                fk.c[l - l0] = (int)sofi_idx(g.pitch, ci[l], cj[l]);
                if (ct[l] >= 2 && ct[l] <= 4) force = true;
            }
            fk.nforce = force ? nk : 0;
            fk.npsrc = h->slots[k].source_type == 1 ? nk : 0;
            l0 += nk;
        }
        h->src_dirty = false;
        h->gen++;
    }
    if (h->rec_dirty) {
        std::vector<int> ri, rj, rs;
        std::vector<long long> off, cst;
        std::vector<double> sq_ppi, sq_pu, pfac;
        h->sec_base.assign(h->nb, 0);
        h->rec_start.assign(h->nb + 1, 0);
        long long base = 0;
        for (int k = 0; k < h->nb; ++k) {
            const sofi_dev::Slot &sl = h->slots[k];
            h->rec_start[k] = (int)ri.size();
            const long long ntr = (long long)sl.ri.size(), ns = d.ns > 0 ? d.ns : 1;
            h->sec_base[k] = base;
            for (long long r = 0; r < ntr; ++r) {
                ri.push_back(sl.ri[r]);
                rj.push_back(sl.rj[r]);
                rs.push_back(k);
                off.push_back(base + r * ns);
                cst.push_back(ntr * ns);
                const size_t c = sofi_idx(g.pitch, sl.ri[r], sl.rj[r]);
                const double ppi = h->h_ppi[c], pu = h->h_pu[c];
                sq_ppi.push_back(sqrt(ppi));
                sq_pu.push_back(sqrt(pu));
                pfac.push_back((3.0 * (float)ppi - 4.0 * (float)pu) / (2.0 * (float)ppi - 2.0 * (float)pu));
            }
            base += (long long)SOFI_NCOMP * ntr * ns;
        }
        h->rec_start[h->nb] = (int)ri.size();
        const int n = (int)ri.size();
        const size_t nn = (size_t)(n > 0 ? n : 1);
        dfree(h, &h->rec_ibuf);
        dfree(h, &h->rec_lbuf);
        dfree(h, &h->rec_d);
        dfree(h, &h->sec);
        h->sec_len = (size_t)(base > 0 ? base : 1);
        int rc = dalloc_t(h, &h->rec_ibuf, 3 * nn) | dalloc_t(h, &h->rec_lbuf, 2 * nn) |
                 dalloc_t(h, &h->rec_d, 3 * nn) | dalloc_t(h, &h->sec, h->sec_len);
        if (rc) return -1;
        if (n > 0) {
            std::vector<int> ib(ri);
            ib.insert(ib.end(), rj.begin(), rj.end());
            ib.insert(ib.end(), rs.begin(), rs.end());
            std::vector<long long> lb(off);
            lb.insert(lb.end(), cst.begin(), cst.end());
            std::vector<double> db(sq_ppi);
            db.insert(db.end(), sq_pu.begin(), sq_pu.end());
            db.insert(db.end(), pfac.begin(), pfac.end());
            if (h2d(h, h->rec_ibuf, ib.data(), ib.size() * sizeof(int))) return -1;
            if (h2d(h, h->rec_lbuf, lb.data(), lb.size() * sizeof(long long))) return -1;
            if (h2d(h, h->rec_d, db.data(), db.size() * sizeof(double))) return -1;
        }
        Rec &r = h->r;
        r.i = h->rec_ibuf;
        r.j = h->rec_ibuf + nn;
        r.shot = h->rec_ibuf + 2 * nn;
        r.off = h->rec_lbuf;
        r.cst = h->rec_lbuf + nn;
        r.sq_ppi = h->rec_d;
        r.sq_pu = h->rec_d + nn;
        r.pfac = h->rec_d + 2 * nn;
        r.sec = h->sec;
        r.n = n;
        r.ns = d.ns;
        h->rec_dirty = false;
        h->gen++;
        // an FWI handle's adjoint sources sit at its receivers unless set
        // elsewhere: their arrays exist from the first forward pass on
        if (d.mode == SOFI_MODE_FWI && !h->adj_custom && adj_at_receivers(h)) return -1;
    }
    return 0;
}

} // namespace

extern "C" {

int sofi_set_snapshots(sofi_dev *h, int nsnap)
{
    if (!h) return fail("sofi_set_snapshots: null handle");
    if (nsnap > 0 && h->nb > 1) return fail("sofi_set_snapshots: snapshots are taken by one-shot handles only");
    const sofi_desc &d = h->d;
    h->nxs = (d.snap_idx[2] - d.snap_idx[1]) / (d.idx > 0 ? d.idx : 1) + 1;
    h->nys = (d.snap_idx[4] - d.snap_idx[3]) / (d.idy > 0 ? d.idy : 1) + 1;
    if (d.snap_idx[1] < 1 || d.snap_idx[2] > d.nx || d.snap_idx[3] < 1 || d.snap_idx[4] > d.ny)
        return fail("sofi_set_snapshots: snapshot window outside the grid");
    h->nsnap_max = nsnap > 0 ? nsnap : 0;
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(h, &h->snapbuf);
    if (h->nsnap_max > 0 &&
        dalloc_t(h, &h->snapbuf, (size_t)h->nsnap_max * SOFI_NCOMP * h->nxs * h->nys))
        return -1;
    h->nsnap = 0;
    return 0;
}

int sofi_reset(sofi_dev *h)
{
    if (!h) return fail("sofi_reset: null handle");
    const size_t nb = (size_t)h->nb, pl = (size_t)h->g.plane * sizeof(float);
    Fields &f = h->f;
    for (float *p : { f.vx, f.vy, f.sxx, f.syy, f.sxy, f.dvxx, f.dvyy, f.dvyx, f.dvxy })
        if (p) HIPCHK(hipMemsetAsync(p, 0, pl * nb, h->stream));
    for (float *p : { f.r, f.p, f.q })
        if (p) HIPCHK(hipMemsetAsync(p, 0, pl * h->g.L * nb, h->stream));
    const size_t nxp = (size_t)h->g.nxp * sizeof(float) * nb, nyp = (size_t)h->g.nyp * sizeof(float) * nb;
    for (float *p : { f.psi_sxx_x, f.psi_sxy_x, f.psi_vxx, f.psi_vyx })
        if (p) HIPCHK(hipMemsetAsync(p, 0, nxp, h->stream));
    for (float *p : { f.psi_syy_y, f.psi_sxy_y, f.psi_vyy, f.psi_vxy })
        if (p) HIPCHK(hipMemsetAsync(p, 0, nyp, h->stream));
    if (h->sec) HIPCHK(hipMemsetAsync(h->sec, 0, h->sec_len * sizeof(float), h->stream));
    for (float *p : { f.pvxp1, f.pvyp1, f.ux, f.uy, f.uxy })
        if (p) HIPCHK(hipMemsetAsync(p, 0, pl * nb, h->stream));
    HIPCHK(hipMemsetAsync(f.unstable, 0, sizeof(int) * nb, h->stream));
    if (h->fdt4) {   // zero_wavefield clears every history (src/zero_wavefield.c:88-127)
        Hist4 &H = h->h4;
        for (int k = 0; k < 4; ++k) {
            HIPCHK(hipMemsetAsync(H.sx[k], 0, pl, h->stream));
            HIPCHK(hipMemsetAsync(H.sy[k], 0, pl, h->stream));
            for (int q = 0; q < 4; ++q) HIPCHK(hipMemsetAsync(H.dv[q][k], 0, pl, h->stream));
            for (float *m : { H.r[k], H.p[k], H.q[k] })
                if (m) HIPCHK(hipMemsetAsync(m, 0, pl * h->g.L, h->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    h->nsnap = 0;
    h->lsamp = h->d.ndt;
    h->lsnap = h->d.lsnap;
    return 0;
}

} // extern "C"

namespace {

// cells per thread of an ISO/VTI update kernel of the given shape (1: CELL,
// 2: VEC2, 4: VEC4)
int vec_width(const Grid &g, int shape)
{
    if (shape == SOFI_LAUNCH_CELL) return 1;
    if (shape == SOFI_LAUNCH_VEC2) return 2;
    if (shape == SOFI_LAUNCH_VEC4) return 4;
    // automatic, as measured on MI355X: the elastic stress updates stream
    // only 11-12 arrays, too few bytes in flight at one float per thread, and
    // VEC2 for both kernels won by 26 % (EL_ISO and EL_VTI, 2000 x 1000 FD6:
    // 48.4 vs 38.5 GLUPS); the viscoelastic ones stream 20-40 arrays and one
    // cell per thread won (VEL_ISO L=3, 4000 x 2000 FD8: the vector shapes
    // need 80-150 VGPRs and lose more to occupancy than they gain)
    return (g.weq == SOFI_EL_ISO || g.weq == SOFI_EL_VTI) ? 2 : 1;
}

// One time step.  With ntb set the step is captured into a block: nt is the
// block's local step (1..K), the kernels add the base step held at ntb, and
// the host-side sampling state is advanced for the whole block by the
// caller; such a block starts on a sampling boundary, so the local step has
// the sampling phase of the global one.
// the entries of slots lo..hi-1 of a batch's injection / receiver lists
// (each slot's entries are contiguous, starting at st[slot])
Inj inj_part(const Inj &a, const std::vector<int> &st, int lo, int hi)
{
    Inj b = a;
    if (st.empty()) return b;
    b.n = st[hi] - st[lo];
    b.cell = a.cell + st[lo];
    b.first = a.first + st[lo];
    return b;
}

Rec rec_part(const Rec &a, const std::vector<int> &st, int lo, int hi)
{
    Rec b = a;
    if (st.empty()) return b;
    const int o = st[lo];
    b.n = st[hi] - o;
    b.i = a.i + o;
    b.j = a.j + o;
    b.shot = a.shot + o;
    b.off = a.off + o;
    b.cst = a.cst + o;
    b.sq_ppi = a.sq_ppi + o;
    b.sq_pu = a.sq_pu + o;
    b.pfac = a.pfac + o;
    return b;
}

// the oldest of four history buffers becomes the newest (shift_var2 /
// shift_var3, src/util.c:89-109)
void rot4(float **v)
{
    float *t = v[3];
    v[3] = v[2];
    v[2] = v[1];
    v[1] = v[0];
    v[0] = t;
}

// A launch of the time loop: on the handle's stream, or, while a block of
// steps is recorded, a kernel node of that block's graph after the previous
// one (graphs are built node by node, not by stream capture, so other
// threads' synchronous HIP calls cannot invalidate a recording)
template <typename... P, typename... A>
void launch(sofi_dev *h, void (*k)(P...), dim3 grid, dim3 block, A &&...a)
{
    using Tup = std::tuple<std::decay_t<P>...>;
    if (h->defer) {
        // a step recorded for the work-list launches: issued later, in its stage
        auto v = std::make_shared<Tup>(std::forward<A>(a)...);
        h->defer->push_back({h->dstage, [k, grid, block, v](hipStream_t st) {
            void *args[sizeof...(P) > 0 ? sizeof...(P) : 1];
            std::apply([&](auto &...x) {
                int q = 0;
                ((args[q++] = (void *)&x), ...);
            }, *v);
            return hipLaunchKernel(reinterpret_cast<const void *>(k), grid, block, args, 0, st);
        }});
        return;
    }
    if (!h->gbuild) {
        Tup v(std::forward<A>(a)...);
        void *args[sizeof...(P) > 0 ? sizeof...(P) : 1];
        std::apply([&](auto &...x) {
            int q = 0;
            ((args[q++] = (void *)&x), ...);
        }, v);
        const hipError_t e =
            hipLaunchKernel(reinterpret_cast<const void *>(k), grid, block, args, 0, h->cur ? h->cur : h->stream);
        if (e != hipSuccess && h->gerr == hipSuccess) h->gerr = e;
        return;
    }
    // a node of the block being recorded: its argument values live in the
    // handle until the graph is dropped
    sofi_dev::NodeArgs na;
    auto v = std::make_shared<Tup>(std::forward<A>(a)...);
    na.ptrs.resize(sizeof...(P) > 0 ? sizeof...(P) : 1);
    std::apply([&](auto &...x) {
        int q = 0;
        ((na.ptrs[q++] = (void *)&x), ...);
    }, *v);
    na.vals = v;
    h->gargs.push_back(std::move(na));
    hipKernelNodeParams p{};
    p.func = reinterpret_cast<void *>(k);
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = 0;
    p.kernelParams = h->gargs.back().ptrs.data();
    p.extra = nullptr;
    hipGraphNode_t n = nullptr;
    const hipError_t e = hipGraphAddKernelNode(&n, h->gbuild, h->glast ? &h->glast : nullptr, h->glast ? 1 : 0, &p);
    if (e != hipSuccess) {
        if (h->gerr == hipSuccess) h->gerr = e;
    } else {
        h->glast = n;
    }
}

// the periodic wrap of the ghost columns of a (vx / sxy position) and b (vy
// / sxx position) for slots lo..hi-1 (BOUNDARY = 1)
void wrap_x(sofi_dev *h, Grid gg, int lo, int hi, float *a, float *b)
{
    if (!gg.periodic || hi <= lo) return;
    gg.k_lo = lo;
    launch(h, k_wrap_x, dim3((unsigned)((gg.ny + 255) / 256), (unsigned)(hi - lo)), dim3(256), gg, a, b);
}

// the TTI CPML frame of slots lo..hi-1 over the frame tiles, after
// k_stress_tti_f has applied the frame's CPML: the sides' and corners'
// update (src/update_s_elastic_TTI_PML.c:32-145)
template <int ND>
void tti_frame(sofi_dev *h, Grid gg, int lo, int hi, const Fields &f, const Mat &m, const Fuse &fz, int nt)
{
    const int nft = frame_tile_count(gg);
    if (gg.fw <= 0 || gg.abs_type != 1 || nft <= 0 || hi <= lo) return;
    gg.k_lo = lo;
    gg.k_hi = hi;
    const dim3 tb(BX, BY), tg((unsigned)nft, (unsigned)(hi - lo));
    if (gg.weq == SOFI_EL_TTI) launch(h, k_tti_frame<ND, SOFI_EL_TTI>, tg, tb, gg, f, m, fz, h->s, nt);
    else launch(h, k_tti_frame<ND, SOFI_VEL_TTI>, tg, tb, gg, f, m, fz, h->s, nt);
}

// Launches of a batch's update kernels in slot blocks of S slots per thread
// (k_velocity_b, k_stress_b, k_stress_tti_f) over slots lo..hi-1; the slots
// that do not fill a block of S go in blocks of one.
template <int ND, template <int> class K, typename... A>
void launch_blocks(sofi_dev *h, Grid gg, int lo, int hi, int S, A &&...a)
{
    const dim3 blk(BX, BY);
    const unsigned ntile = (unsigned)(gg.ntx * gg.nty);
    const int n = (hi - lo) / S;
    if (n > 0) {
        gg.k_lo = lo;
        gg.k_hi = lo + n * S;
        if (S == 4) launch(h, K<4>::fn(), dim3(ntile, n), blk, gg, a...);
        else if (S == 2) launch(h, K<2>::fn(), dim3(ntile, n), blk, gg, a...);
        else launch(h, K<1>::fn(), dim3(ntile, n), blk, gg, a...);
    }
    const int r0 = lo + n * S;
    if (r0 < hi && S != 1) {
        gg.k_lo = r0;
        gg.k_hi = hi;
        launch(h, K<1>::fn(), dim3(ntile, hi - r0), blk, gg, a...);
    }
}
template <int ND>
struct VelB {
    template <int S>
    struct K {
        static auto fn() { return k_velocity_b<ND, S>; }
    };
};
template <int ND, int WEQ>
struct StrB {
    template <int S>
    struct K {
        static auto fn() { return k_stress_b<ND, WEQ, S>; }
    };
};
template <int ND, int WEQ>
struct TtiF {
    template <int S>
    struct K {
        static auto fn() { return k_stress_tti_f<ND, WEQ, S>; }
    };
};

template <int ND>
int step_nd(sofi_dev *h, int nt, const int *ntb)
{
    const Grid &g = h->g;
    Fields &f = h->f;
    const Mat &m = h->m;
    const sofi_desc &d = h->d;
    hipStream_t st = h->cur ? h->cur : h->stream;
    const bool tti = d.weq == SOFI_EL_TTI || d.weq == SOFI_VEL_TTI;
    // FWI: the forward pass stores every DTINV-th step; the adjoint pass
    // injects the residuals at the receivers and correlates at the steps whose
    // time-reversed forward step was stored (src/time_loop.c:287-303); it
    // records no seismograms and takes no snapshots
    const bool fwi = d.mode == SOFI_MODE_FWI, adj = fwi && h->sw == 1;
    // in a captured block nt is local and the block's base step a multiple
    // of DTINV, so the store / correlation pattern of the local step is the
    // global one
    const bool inblk = ntb != nullptr;
    const int dtinv = d.dtinv > 0 ? d.dtinv : 1, fstep = d.nt - nt + 1, ndt = d.ndt > 0 ? d.ndt : 1;
    const bool store = fwi && !adj && (nt - 1) % dtinv == 0;
    const bool conv = adj && (fstep - 1) % dtinv == 0;
    // the slots this step advances: all, or the resident group of a batch
    const int lo0 = h->cur_lo, hi0 = h->cur_hi;
    const bool part = lo0 > 0 || hi0 < h->nb;
    const Rec rr = part ? rec_part(h->r, h->rec_start, lo0, hi0) : h->r;
    const Inj src_inj = part ? inj_part(h->src_inj, h->src_inj_start, lo0, hi0) : h->src_inj;
    const Inj psrc_inj = part ? inj_part(h->psrc_inj, h->psrc_inj_start, lo0, hi0) : h->psrc_inj;
    const bool sample = !adj && d.seismo && rr.n > 0 && (inblk ? nt % ndt == 0 : (nt == h->lsamp && nt < d.nt));
    const bool sample_dv = sample && (d.seismo == 3 || d.seismo == 4);
    const int ins = inblk ? nt : h->lsamp / ndt;       // in a block: the local step (k_seis adds the base)
    const int nsrc = adj ? 0 : h->s.n;     // the shot's sources act in the forward pass only
    // what the update kernels do besides the update: forces before the
    // velocity update; explosive sources and the free surface after the
    // stress update (not in steps that sample div/curl, which read the CPML
    // memory before the free surface updates it again)
    Fuse fv = h->fz, fs = h->fz;
    fv.npsrc = fs.nforce = 0;
    fv.surf = 0;
    fv.nforce = h->fusable && h->any_force ? nsrc : 0;
    // sources fused into the update kernels (one-shot handles with few
    // sources; batches launched one slot at a time, each launch with its
    // slot's sources)
    // (TTI: not with a frame that is never updated, ABS_TYPE other than 1 or
    // 2, where a source cell would be skipped by every update kernel)
    const bool src_fused = h->fusable && h->source_type != 5 && (!tti || g.fw == 0 || g.abs_type == 1 || g.abs_type == 2);
    // a batch in slot blocks (h->bs slots per thread, all slots of the group
    // in one launch per kernel); forces, the free surface and sampling then
    // run as launches of their own, explosive sources too unless the stress
    // launch carries them (below)
    const bool bm = h->nb > 1 && h->bs > 0 && !h->fdt4 && d.weq != SOFI_AC_ISO;
    const int span = hi0 - lo0, grp = !bm && h->grp > 0 && h->grp < span ? h->grp : span;
    // slot blocks of ISO / VTI rheologies: the explosive sources of the
    // slots (at most FUSE_MAX of them) ride in the stress launch
    PFuse pf{};
    bool pf_fused = false;
    if (bm && !tti && nsrc > 0 && h->slot_fz.size() == (size_t)h->nb) {
        pf_fused = true;
        for (int k = lo0; k < hi0 && pf_fused; ++k) {
            const Fuse &sf = h->slot_fz[k];
            for (int q = 0; q < sf.npsrc; ++q) {
                if (pf.n == FUSE_MAX || q >= FUSE_MAX) {
                    pf_fused = false;
                    break;
                }
                pf.k[pf.n] = k;
                pf.c[pf.n] = sf.c[q];
                pf.l[pf.n] = sf.soff + q;
                ++pf.n;
            }
        }
        if (!pf_fused) pf = PFuse{};
    }
    const bool tti_pf = bm && tti && span == 1 && nsrc > 0 && h->slot_fz.size() == (size_t)h->nb &&
                        h->slots[lo0].source_type == 1 && h->slot_fz[lo0].npsrc > 0 &&
                        h->slot_fz[lo0].npsrc <= FUSE_MAX && (g.fw == 0 || g.abs_type == 1 || g.abs_type == 2);
    pf_fused = pf_fused || tti_pf;
    const bool slot_fused = !bm && !tti && !h->fdt4 && h->slot_fusable && grp == 1 && nsrc > 0;
    fs.npsrc = src_fused && h->source_type == 1 ? nsrc : 0;
    // the free surface follows the explosive / moment sources, so it is fused
    // when they are, or when there are none to inject this step
    const bool no_stress_src = nsrc == 0 || (h->psrc_inj.n == 0 && h->n_eq == 0);
    // (in a work-list run, not in any step of a handle that samples div/curl,
    // so that its update kernels' arguments are the same in every step)
    const bool dv_any = h->wl ? d.seismo == 3 || d.seismo == 4 : sample_dv;
    fs.surf = !bm && !tti && (src_fused || slot_fused || no_stress_src) && d.free_surf && !dv_any &&
              (g.fw == 0 || g.abs_type == 1 || g.abs_type == 2);
    fv.fwi_acc = store ? 1 : adj ? (d.velocity == 0 ? 2 : (conv ? 3 : 0)) : 0;
    fs.fwi_u = store && d.velocity == 1;
    fv.ntb = fs.ntb = ntb;
    fv.sblk = fs.sblk = 0;
    // vx / vy sampling rides on the stress launch of a one-shot handle
    // (SEISMO = 1, no TTI; the sampler addresses slot 0 of the handle's fields)
    const bool seis_fused = sample && d.seismo == 1 && !tti && h->nb == 1 && !h->wl;
    if (nsrc > 0 && h->any_force && !fv.nforce && !slot_fused && src_inj.n > 0)
        launch(h, k_force, dim3((src_inj.n + 63) / 64), dim3(64), g, f, m, h->s, src_inj, nt, ntb);
    if (adj && h->nadj > 0)
        launch(h, k_adjoint, dim3((h->adj_inj.n + 63) / 64), dim3(64), g, f, h->adj_inj, h->adj_sx,
                           h->adj_sy, d.nt, d.adjoint_type, nt, ntb);
    hipEvent_t *te = nullptr;
    if (!inblk && !h->wl && h->time_every > 0 && nt % h->time_every == 0 && h->tev_used + 4 <= (int)h->tev.size()) {
        te = &h->tev[h->tev_used];
        h->tev_used += 4;
    }
    const int nty = (g.ny + BY - 1) / BY;
    const dim3 blk(BX, BY);
    // the update kernels, launched per group of h->grp slots (all slots of a
    // one-shot handle or of an untuned batch: one group)
    if (te) (void)hipEventRecord(te[0], st);
    for (int lo = lo0; lo < hi0; lo += grp) {
        Grid gg = g;
        gg.k_lo = lo;
        gg.k_hi = lo + grp < hi0 ? lo + grp : hi0;
        const int nchunk = (gg.k_hi - lo + g.spt - 1) / g.spt;   // groups of g.spt slots (blockIdx.y)
        if (h->fdt4) {
            // 4th order in time (one slot, one thread per cell): the
            // histories rotate after each update, as shift_var2 after the
            // velocity update and shift_var3 / shift_var2 after the stress
            // update (src/time_loop.c:89-91,226-244); the memory variables
            // the rest of the step sees are the ones just written
            const dim3 t4((g.nx + BX - 1) / BX * nty);
            launch(h, k_velocity4<ND>, t4, blk, gg, f, m, h->s, fv, nt, h->h4);
            wrap_x(h, gg, lo, gg.k_hi, f.vx, f.vy);
            rot4(h->h4.sx);
            rot4(h->h4.sy);
            if (te) {
                (void)hipEventRecord(te[1], st);
                (void)hipEventRecord(te[2], st);
            }
            Fuse fss = fs;
            dim3 ts = t4;
            if (seis_fused) {
                fss.sblk = (rr.n + BX * BY - 1) / (BX * BY);
                fss.ntile = (int)t4.x;
                fss.ins = ins;
                fss.ndt = ndt;
                fss.rec = rr;
                ts.x += fss.sblk;
            }
            if (g.L > 0) launch(h, k_stress4<ND, true>, ts, blk, gg, f, m, h->s, fss, nt, h->h4);
            else launch(h, k_stress4<ND, false>, ts, blk, gg, f, m, h->s, fss, nt, h->h4);
            for (int q = 0; q < 4; ++q) rot4(h->h4.dv[q]);
            if (g.L > 0) {
                rot4(h->h4.r);
                rot4(h->h4.p);
                rot4(h->h4.q);
                f.r = h->h4.r[0];
                f.p = h->h4.p[0];
                f.q = h->h4.q[0];
            }
            continue;
        }
        if (bm) {
            const int S = h->bs;
            launch_blocks<ND, VelB<ND>::template K>(h, gg, lo, gg.k_hi, S, f, m, nt, ntb, fv.fwi_acc);
            wrap_x(h, gg, lo, gg.k_hi, f.vx, f.vy);
            if (te && lo == lo0 && grp == span) {
                (void)hipEventRecord(te[1], st);
                (void)hipEventRecord(te[2], st);
            }
            switch (d.weq) {
              case SOFI_EL_ISO: launch_blocks<ND, StrB<ND, SOFI_EL_ISO>::template K>(h, gg, lo, gg.k_hi, S, f, m, nt, ntb, fs.fwi_u, h->s, pf); break;
              case SOFI_VEL_ISO: launch_blocks<ND, StrB<ND, SOFI_VEL_ISO>::template K>(h, gg, lo, gg.k_hi, S, f, m, nt, ntb, fs.fwi_u, h->s, pf); break;
              case SOFI_EL_VTI: launch_blocks<ND, StrB<ND, SOFI_EL_VTI>::template K>(h, gg, lo, gg.k_hi, S, f, m, nt, ntb, fs.fwi_u, h->s, pf); break;
              case SOFI_VEL_VTI: launch_blocks<ND, StrB<ND, SOFI_VEL_VTI>::template K>(h, gg, lo, gg.k_hi, S, f, m, nt, ntb, fs.fwi_u, h->s, pf); break;
              default: {
                  // a group of one slot: its explosive sources ride in the
                  // last stress kernel, as for a one-shot handle
                  Fuse f0{};
                  f0.fwi_u = fs.fwi_u;
                  f0.ntb = ntb;
                  if (tti_pf) {
                      const Fuse &sf = h->slot_fz[lo];
                      memcpy(f0.c, sf.c, sizeof(sf.c));
                      f0.soff = sf.soff;
                      f0.npsrc = sf.npsrc;
                  }
                  Fuse fin = f0;
                  if (g.abs_type == 1 && g.fw > 0 && frame_tile_count(g) > 0) fin.npsrc = 0;
                  if (d.weq == SOFI_EL_TTI) launch_blocks<ND, TtiF<ND, SOFI_EL_TTI>::template K>(h, gg, lo, gg.k_hi, S, f, m, fin, h->s, nt);
                  else launch_blocks<ND, TtiF<ND, SOFI_VEL_TTI>::template K>(h, gg, lo, gg.k_hi, S, f, m, fin, h->s, nt);
                  tti_frame<ND>(h, gg, lo, gg.k_hi, f, m, f0, nt);
              }
            }
            continue;
        }
        // a launch of one slot runs the one-shot kernels on that slot's fields
        const bool one = gg.k_hi - gg.k_lo == 1;
        const Fields fk = one ? f.at(lo, g) : f;
        Fuse fv_l = fv;
        if (slot_fused) {
            const Fuse &sf = h->slot_fz[lo];
            memcpy(fv_l.c, sf.c, sizeof(sf.c));
            fv_l.soff = sf.soff;
            fv_l.nforce = sf.nforce;
        }
        int vw = vec_width(g, h->shape_v);
        dim3 tiles((g.nx + BX * vw - 1) / (BX * vw) * nty, nchunk);
        const dim3 vtiles((unsigned)(g.ntx * ((g.ny + VROWS * BY - 1) / (VROWS * BY))));   // one slot, VROWS rows per thread
        if (h->wl) {
            // the work-list launch of sofi_run_many covers this update
            h->wl_shot.g = gg;
            h->wl_shot.f = fk;
            h->wl_shot.m = m;
            h->wl_shot.s = h->s;
            h->wl_shot.fv = fv_l;
            h->wl_shot.fv.ntb = h->ntb;
            h->wl_shot.r = h->r;
            h->wl_shot.seismo = d.seismo;
            h->dstage = 1;
        } else if (vw == 4 && one) launch(h, k_velocity_v<ND, 4, true>, tiles, blk, gg, fk, m, h->s, fv_l, nt);
        else if (vw == 4) launch(h, k_velocity_v<ND, 4, false>, tiles, blk, gg, fk, m, h->s, fv_l, nt);
        else if (vw == 2 && one) launch(h, k_velocity_v<ND, 2, true>, tiles, blk, gg, fk, m, h->s, fv_l, nt);
        else if (vw == 2) launch(h, k_velocity_v<ND, 2, false>, tiles, blk, gg, fk, m, h->s, fv_l, nt);
        else if (one && tti) launch(h, k_velocity<ND, true, false>, tiles, blk, gg, fk, m, h->s, fv_l, nt);
        else if (one) launch(h, k_velocity_r<ND, true, VROWS>, vtiles, blk, gg, fk, m, h->s, fv_l, nt);
        else launch(h, k_velocity<ND, false>, tiles, blk, gg, fk, m, h->s, fv_l, nt);
        wrap_x(h, gg, lo, gg.k_hi, f.vx, f.vy);
        vw = vec_width(g, h->shape_s);
        tiles = dim3((g.nx + BX * vw - 1) / (BX * vw) * nty, nchunk);
        // vx / vy sampling in extra blocks of a one-slot stress launch
        Fuse fss = fs;
        if (slot_fused) {
            const Fuse &sf = h->slot_fz[lo];
            memcpy(fss.c, sf.c, sizeof(sf.c));
            fss.soff = sf.soff;
            fss.npsrc = sf.npsrc;
        }
        if (seis_fused && one) {
            fss.sblk = (rr.n + BX * BY - 1) / (BX * BY);
            fss.ntile = (int)tiles.x;
            fss.ins = ins;
            fss.ndt = ndt;
            fss.rec = rr;
            tiles.x += fss.sblk;
        }
        if (te && lo == lo0 && grp == span) {
            (void)hipEventRecord(te[1], st);
            (void)hipEventRecord(te[2], st);
        }
#define SOFI_STRESS(W)                                                                                      \
        do {                                                                                                \
            if (vw == 4 && one) launch(h, k_stress_v<ND, W, 4, true>, tiles, blk, gg, fk, m, h->s, fss, nt);   \
            else if (vw == 4) launch(h, k_stress_v<ND, W, 4, false>, tiles, blk, gg, fk, m, h->s, fss, nt);    \
            else if (vw == 2 && one) launch(h, k_stress_v<ND, W, 2, true>, tiles, blk, gg, fk, m, h->s, fss, nt);   \
            else if (vw == 2) launch(h, k_stress_v<ND, W, 2, false>, tiles, blk, gg, fk, m, h->s, fss, nt);    \
            else if (one) launch(h, k_stress<ND, W, true>, tiles, blk, gg, fk, m, h->s, fss, nt);              \
            else launch(h, k_stress<ND, W, false>, tiles, blk, gg, fk, m, h->s, fss, nt);            \
        } while (0)
        if (h->wl) {
            h->wl_shot.fs = fss;
            h->wl_shot.fs.ntb = h->ntb;
            h->dstage = 2;
            continue;
        }
        switch (d.weq) {
          case SOFI_AC_ISO: {     // one thread per cell whatever the stress shape
              dim3 ta((g.nx + BX - 1) / BX * nty, nchunk);
              if (seis_fused && one) ta.x += fss.sblk;
              if (seis_fused && one) fss.ntile = (int)((g.nx + BX - 1) / BX * nty);
              if (one) launch(h, k_stress_ac<ND, true>, ta, blk, gg, fk, m, h->s, fss, nt);
              else launch(h, k_stress_ac<ND, false>, ta, blk, gg, fk, m, h->s, fss, nt);
              break;
          }
          case SOFI_EL_ISO: SOFI_STRESS(SOFI_EL_ISO); break;
          case SOFI_VEL_ISO: SOFI_STRESS(SOFI_VEL_ISO); break;
          case SOFI_EL_VTI: SOFI_STRESS(SOFI_EL_VTI); break;
          case SOFI_VEL_VTI: SOFI_STRESS(SOFI_VEL_VTI); break;
          default: {
              // the interior (and a damping frame) with the derivatives
              // computed in the kernel, one slot per thread; then the CPML
              // frame region by region over the frame tiles
              // (src/time_loop.c:180-186)
              // fused explosive sources: added by the last stress kernel,
              // k_tti_frame when the CPML frame passes run
              const bool frame_passes = g.abs_type == 1 && g.fw > 0 && frame_tile_count(g) > 0;
              Fuse fin = fss;
              if (frame_passes) fin.npsrc = 0;
              if (d.weq == SOFI_EL_TTI) launch_blocks<ND, TtiF<ND, SOFI_EL_TTI>::template K>(h, gg, lo, gg.k_hi, 1, f, m, fin, h->s, nt);
              else launch_blocks<ND, TtiF<ND, SOFI_VEL_TTI>::template K>(h, gg, lo, gg.k_hi, 1, f, m, fin, h->s, nt);
              tti_frame<ND>(h, gg, lo, gg.k_hi, f, m, fss, nt);
          }
        }
    }
    if (te && grp != span) {
        (void)hipEventRecord(te[1], st);
        (void)hipEventRecord(te[2], st);
    }
#undef SOFI_STRESS
    if (te) (void)hipEventRecord(te[3], st);
    if (sample_dv && h->wl) {
        h->wl_dv = ins;                    // the work-list step's sampler launch
    } else if (sample_dv) {
        dim3 rb(64), rg((rr.n + 63) / 64);
        if (tti) launch(h, k_seis_dv<ND, true>, rg, rb, g, f, m, rr, ins, ntb, ndt);
        else launch(h, k_seis_dv<ND, false>, rg, rb, g, f, m, rr, ins, ntb, ndt);
    }
    if (nsrc > 0 && !src_fused && !slot_fused) {
        if (psrc_inj.n > 0 && !pf_fused)
            launch(h, k_psource, dim3((psrc_inj.n + 63) / 64), dim3(64), g, f, h->s, psrc_inj, nt, ntb);
        if (h->n_eq > 0)
            launch(h, k_eqsource, dim3(1), dim3(64), g, f, h->s, h->eq_list, h->n_eq, nt, ntb, lo0, hi0);
    }
    if (d.free_surf && !fs.surf) {
        Grid gs = g;
        gs.k_lo = lo0;
        dim3 sb(256), sg((g.nx + 255) / 256, span);
        if (g.L > 0) launch(h, k_surface<ND, true>, sg, sb, gs, f, m);
        else launch(h, k_surface<ND, false>, sg, sb, gs, f, m);
    }
    // the stresses' ghost columns wrap once the step's stresses are final
    // (exchange_s after the sources and the free surface, src/time_loop.c:273)
    wrap_x(h, g, lo0, hi0, d.weq == SOFI_AC_ISO ? nullptr : f.sxy, f.sxx);
    if (sample && d.seismo != 3 && h->wl) {
        h->wl_seis = ins;
    } else if (sample && d.seismo != 3 && !seis_fused) {
        dim3 rb(64), rg((rr.n + 63) / 64);
        launch(h, k_seis, rg, rb, g, f, rr, ins, d.seismo, ntb, ndt);
    }
    if (!inblk && d.seismo && nt == h->lsamp && nt < d.nt) h->lsamp += d.ndt;
    if (store || conv) {
        // every shot slot of the step (blockIdx.y), rows of tiles in blockIdx.z
        Grid gs = g;
        gs.k_lo = lo0;
        dim3 tb(BX, BY), tg((g.nx + BX - 1) / BX, span, (g.ny + BY - 1) / BY);
        if (store) launch(h, k_fwi_store, tg, tb, gs, f, h->w, (nt - 1) / dtinv, ntb, dtinv);
        else launch(h, k_fwi_conv, tg, tb, gs, f, m, h->w, (fstep - 1) / dtinv, ntb, dtinv);
    }
    if (!inblk && !adj && d.snap && nt == h->lsnap && nt <= d.esnap) {
        if (h->nsnap < h->nsnap_max) {
            int n = h->nxs * h->nys;
            float *out = h->snapbuf + (size_t)h->nsnap * SOFI_NCOMP * n;
            launch(h, k_snap<ND>, dim3((n + 255) / 256), dim3(256), g, f, m, out, d.snap_idx[1],
                               d.idx, h->nxs, d.snap_idx[3], d.idy, h->nys);
            h->nsnap++;
        }
        h->lsnap += d.isnap;
    }
    return 0;
}

int step(sofi_dev *h, int nt, const int *ntb = nullptr)
{
    switch (h->g.nd) {
      case 1: return step_nd<1>(h, nt, ntb);
      case 2: return step_nd<2>(h, nt, ntb);
      case 3: return step_nd<3>(h, nt, ntb);
      case 4: return step_nd<4>(h, nt, ntb);
      case 5: return step_nd<5>(h, nt, ntb);
      case 6: return step_nd<6>(h, nt, ntb);
    }
    return fail("unsupported ND %d", h->g.nd);
}

constexpr size_t GRAPH_AUTO_CELLS = size_t(1) << 20;
constexpr int GRAPH_STEPS = 64;

// steps of a captured block for this handle (0: run step by step)
int graph_len(const sofi_dev *h)
{
    const sofi_desc &d = h->d;
    if (h->graph_mode == SOFI_GRAPH_OFF || h->time_every > 0) return 0;
    if (h->cur_lo > 0 || h->cur_hi < h->nb) return 0;       // a batch advanced by resident groups
    if (h->fdt4) return 0;                                  // its histories rotate every step
    if (h->graph_mode == SOFI_GRAPH_AUTO && (size_t)d.nx * d.ny * h->nb > GRAPH_AUTO_CELLS) return 0;
    int unit = d.seismo && d.ndt > 0 ? d.ndt : 1;
    if (d.mode == SOFI_MODE_FWI) {
        const int dtinv = d.dtinv > 0 ? d.dtinv : 1;
        int a = unit, b = dtinv;                          // unit = lcm(NDT, DTINV)
        while (b) { const int t = a % b; a = b; b = t; }
        unit = unit / a * dtinv;
    }
    const int want = h->graph_steps > 0 ? h->graph_steps : GRAPH_STEPS;
    return (want + unit - 1) / unit * unit;
}

// a block of k steps may start at nt when it ends before nt_last and the
// loop's last step, starts on a sampling boundary with the sampling state
// there, and holds no snapshot
bool block_fits(const sofi_dev *h, int k, int nt, int nt_last)
{
    const sofi_desc &d = h->d;
    const int last = nt + k - 1, ndt = d.ndt > 0 ? d.ndt : 1;
    if (k <= 0 || last > nt_last || last >= d.nt) return false;
    if (d.seismo && ((nt - 1) % ndt != 0 || h->lsamp != nt - 1 + ndt)) return false;
    if (d.snap && h->lsnap >= nt && h->lsnap <= last && h->lsnap <= d.esnap) return false;
    if (d.mode == SOFI_MODE_FWI && (nt - 1) % (d.dtinv > 0 ? d.dtinv : 1) != 0) return false;
    return true;
}

__global__ void k_set_base(int *ntb, int v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *ntb = v;
}

void graph_drop(sofi_dev *h)
{
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->gexec = nullptr;
    h->graph = nullptr;
    h->graph_k = 0;
    h->gargs.clear();
}

// steps nt .. nt + k - 1 from the captured block (captured first when the
// launch configuration changed since)
int run_block(sofi_dev *h, int nt, int k)
{
    if (!h->gexec || h->graph_gen != h->gen || h->graph_k != k) {
        graph_drop(h);
        hipGraph_t gr = nullptr;
        HIPCHK(hipGraphCreate(&gr, 0));
        h->gbuild = gr;
        h->glast = nullptr;
        h->gerr = hipSuccess;
        int rc = 0;
        for (int l = 1; l <= k && !rc; ++l) rc = step(h, l, h->ntb);
        h->gbuild = nullptr;
        const hipError_t e = h->gerr;
        h->gerr = hipSuccess;
        if (rc || e != hipSuccess) {
            (void)hipGraphDestroy(gr);
            h->gargs.clear();
            return rc ? rc : fail("recording a block of steps failed: %s", hipGetErrorString(e));
        }
        h->graph = gr;
        const hipError_t ei = hipGraphInstantiate(&h->gexec, gr, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            h->gexec = nullptr;
            graph_drop(h);
            return fail("hipGraphInstantiate failed: %s", hipGetErrorString(ei));
        }
        h->graph_k = k;
        h->graph_gen = h->gen;
    }
    hipLaunchKernelGGL(k_set_base, dim3(1), dim3(64), 0, h->stream, h->ntb, nt - 1);
    HIPCHK(hipGraphLaunch(h->gexec, h->stream));
    if (h->d.seismo) h->lsamp += k;
    h->graph_blocks++;
    return 0;
}

// slots of a batch advanced together through a run's whole step range: the
// largest group such that the shared medium and the wavefields of the groups
// that run at once (one per stream) fit the memory-side cache (256 MB
// Infinity Cache; 240 MB used), so that every step of a group reads what the
// previous kernels left there; at least one slot; all slots when the medium
// and one slot do not fit (the medium is then shared by all of them instead).
// Measured on 16 EL_VTI shots of 2000 x 1000 FD6 on two streams: groups of 2
// (233 MB at once) 70.5 GLUPS, of 3 (316 MB) 59.2, of 1 64.2; 16 EL_TTI
// shots: groups of 1 (250 MB) 42.3, of 2 41.7, of 3 37.9.
int resident_slots(const sofi_dev *h)
{
    if (h->nb == 1) return 1;
    if (h->d.mode == SOFI_MODE_FWI) return h->nb;     // the adjoint sources and the store span every slot
    if (h->res_mode > 0) return h->res_mode < h->nb ? h->res_mode : h->nb;
    const sofi_desc &d = h->d;
    const size_t pl = (size_t)h->g.plane * sizeof(float);
    const bool visco = d.weq == SOFI_VEL_ISO || d.weq == SOFI_VEL_VTI || d.weq == SOFI_VEL_TTI;
    const bool tti = d.weq == SOFI_EL_TTI || d.weq == SOFI_VEL_TTI;
    const size_t L = visco ? (size_t)d.L : 0;
    const size_t slot = pl * (5 + 3 * L + (tti ? 4 : 0));
    size_t arrays = 4;
    switch (d.weq) {
      case SOFI_EL_ISO: arrays += 3; break;
      case SOFI_VEL_ISO: arrays += 6; break;
      case SOFI_EL_VTI: arrays += 4; break;
      case SOFI_VEL_VTI: arrays += 4 + 4 * L; break;
      case SOFI_EL_TTI: arrays += 8; break;
      default: arrays += 8 + 8 * L; break;
    }
    const size_t budget = (size_t)240 << 20, medium = arrays * pl;
    const size_t nst = h->str_mode > 0 ? (size_t)h->str_mode : 2;     // groups at once
    if (medium + slot > budget) return h->nb;
    const size_t n = (budget - medium) / (slot * nst);
    return n < 1 ? 1 : n < (size_t)h->nb ? (int)n : h->nb;
}

// streams a batch's resident groups run on at once (1 when the batch is one
// group)
int batch_streams(const sofi_dev *h, int ngroups)
{
    if (ngroups <= 1) return 1;
    const int s = h->str_mode > 0 ? h->str_mode : 2;
    return s < ngroups ? s : ngroups;
}

// steps nt_first..nt_last of the slots cur_lo..cur_hi-1: a batch's launch
// groups and slots per thread, when not fixed, measured on the first steps
// (three per candidate, within the range; the results do not depend on
// them), then captured blocks where they apply, else step by step
int run_range(sofi_dev *h, int nt_first, int nt_last)
{
    int nt = nt_first;
    const int span = h->cur_hi - h->cur_lo;
    if (h->nb > 1 && h->bs == 0 && (h->spt_mode == 0 || h->grp_mode == 0) && !h->spt_tuned) {
        std::vector<std::pair<int, int>> cand;   // (slots per launch, slots per thread)
        for (int gr : { span, 16, 8, 4, 1 }) {
            if (gr > span || (gr != span && gr >= span)) continue;
            if (h->grp_mode > 0 && gr != h->grp_mode) continue;
            for (int sp : { 1, 2, 4, 8 }) {
                if (sp > gr || (h->spt_mode > 0 && sp != h->spt_mode)) continue;
                if (std::find(cand.begin(), cand.end(), std::make_pair(gr, sp)) == cand.end()) cand.emplace_back(gr, sp);
            }
        }
        if (!cand.empty() && nt_last - nt_first + 1 >= 3 * (int)cand.size()) {
            float best = 1e30f;
            std::pair<int, int> best_c = cand[0];
            for (const auto &c : cand) {
                h->grp = c.first;
                h->g.spt = c.second;
                if (step(h, nt++)) return -1;
                HIPCHK(hipEventRecord(h->tune0, h->cur ? h->cur : h->stream));
                if (step(h, nt++) || step(h, nt++)) return -1;
                HIPCHK(hipEventRecord(h->tune1, h->cur ? h->cur : h->stream));
                HIPCHK(hipEventSynchronize(h->tune1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, h->tune0, h->tune1));
                if (ms < best) {
                    best = ms;
                    best_c = c;
                }
            }
            h->grp = best_c.first;
            h->g.spt = best_c.second;
            h->spt_tuned = true;
            h->gen++;
        } else {
            h->grp = h->grp_mode > 0 ? h->grp_mode : (span < 4 ? span : 4);
            h->g.spt = h->spt_mode > 0 ? h->spt_mode : (h->grp < 4 ? h->grp : 4);
            h->gen++;
        }
    }
    const int k = graph_len(h);
    while (nt <= nt_last) {
        if (block_fits(h, k, nt, nt_last)) {
            if (run_block(h, nt, k)) return -1;
            nt += k;
        } else {
            if (step(h, nt)) return -1;
            ++nt;
        }
    }
    return 0;
}

} // namespace

namespace {

// the adjoint-source arrays [ntr][NT] and their injection plan at the
// receivers of the handle's shot, zero (what the device residuals fill)
int adj_at_receivers(sofi_dev *h)
{
    // the receivers of every slot, slot after slot; slot k's rows of the
    // adjoint arrays start at adj_start[k]
    std::vector<int> ri, rj, rs;
    h->adj_start.assign(h->nb + 1, 0);
    for (int k = 0; k < h->nb; ++k) {
        const sofi_dev::Slot &sl = h->slots[k];
        h->adj_start[k] = (int)ri.size();
        ri.insert(ri.end(), sl.ri.begin(), sl.ri.end());
        rj.insert(rj.end(), sl.rj.begin(), sl.rj.end());
        rs.insert(rs.end(), sl.ri.size(), k);
    }
    const int n = (int)ri.size();
    h->adj_start[h->nb] = n;
    if (h->adj_at_rec && h->nadj == n) return 0;
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(h, &h->adj_sx);
    dfree(h, &h->adj_sy);
    const size_t len = (size_t)(n > 0 ? n : 1) * h->d.nt;
    if (dalloc_t(h, &h->adj_sx, len) || dalloc_t(h, &h->adj_sy, len)) return -1;
    HIPCHK(hipMemsetAsync(h->adj_sx, 0, len * sizeof(float), h->stream));
    HIPCHK(hipMemsetAsync(h->adj_sy, 0, len * sizeof(float), h->stream));
    if (build_inj(h, n, ri.data(), rj.data(), rs.data(), nullptr, &h->adj_inj, &h->adj_inj_buf)) return -1;
    h->nadj = n;
    h->adj_at_rec = true;
    h->adj_custom = false;
    h->gen++;
    return 0;
}

// the Butterworth sections of one TIME_FILT pass (sofi_iir.h)
TFilt tfilt(double dt, double period, int order, int highpass)
{
    sofi_iir2 sec[SOFI_IIR_MAXSEC];
    TFilt f{};
    f.n = sofi_iir_butterworth(dt, period, order, highpass, sec);
    for (int q = 0; q < f.n; ++q) {
        f.f0[q] = sec[q].f0;
        f.f1[q] = sec[q].f1;
        f.f2[q] = sec[q].f2;
        f.g1[q] = sec[q].g1;
        f.g2[q] = sec[q].g2;
    }
    return f;
}

// floats of the residual workspace for ntr traces of ns samples: a filtered
// synthetic section, a filter buffer, a filtered measured section, the
// traces' misfit shares and the two components' misfits
size_t residual_ws_floats(int ntr, int ns) { return 3 * (size_t)ntr * ns + (size_t)ntr + 2; }

const char *residual_check(const sofi_dev *h, const sofi_residual_opts *o)
{
    if (!h) return "null handle";
    if (!o) return "null options";
    if (h->d.mode != SOFI_MODE_FWI) return "the handle was not opened with MODE=FWI";
    if (o->lnorm < 1 || o->lnorm > 8 || o->lnorm == 6) return "LNORM has no residual in the reference (1-5, 7, 8)";
    if (o->time_filt && (o->order < 1 || o->order > 2 * SOFI_IIR_MAXSEC - 1 || !(o->f_low_pass > 0.0f) ||
                         o->f_high_pass < 0.0f))
        return "time filter: ORDER must be 1..17 and the corner frequencies positive";
    if (h->rec_dirty || !h->sec || h->slots[h->sel].ri.empty()) return "no synthetic sections: run the forward pass first";
    if (h->d.ndt != 1) return "MODE=FWI needs NDT = 1";
    return nullptr;
}

// calc_res of the components ADJOINT_TYPE uses, on the handle's stream:
// obs[c] the measured section of component c (device, [ntr][ns]), the
// synthetic one the handle's section; TIME_FILT filters both first
// (src/inversion.c:68-84).  The residuals go time-reversed into the
// adjoint arrays, the components' misfits into the last two floats of ws.
// rows of the selected slot in the adjoint arrays: its receivers' (first,
// count), or all of a one-slot handle's custom adjoint sources
void adj_rows(const sofi_dev *h, int *first, int *count)
{
    if (h->adj_at_rec && (int)h->adj_start.size() == h->nb + 1) {
        *first = h->adj_start[h->sel];
        *count = h->adj_start[h->sel + 1] - *first;
    } else {
        *first = 0;
        *count = h->nadj;
    }
}

int residual_core(sofi_dev *h, const sofi_residual_opts &o, const float *const obs[2], float *ws)
{
    const sofi_desc &d = h->d;
    int first, ntr;
    adj_rows(h, &first, &ntr);
    const int ns = d.ns;
    const size_t nn = (size_t)ntr * ns;
    float *syncopy = ws, *tmp = ws + nn, *obsf = ws + 2 * nn, *partial = ws + 3 * nn, *mis = partial + ntr;
    hipStream_t st = h->stream;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, mis, (size_t)2, 0.0f);
    const dim3 tb(64), tg((unsigned)((ntr + 63) / 64));
    const double dt = (double)d.dt;
    for (int c = 0; c < 2 && ntr > 0; ++c) {
        const bool used = c == 0 ? (d.adjoint_type == 1 || d.adjoint_type == 3)
                                 : (d.adjoint_type == 1 || d.adjoint_type == 2);
        if (!used) continue;
        float *adj = (c == 0 ? h->adj_sx : h->adj_sy) + (size_t)first * d.nt;
        const float *sy = h->sec + h->sec_base[h->sel] + (size_t)c * nn, *ob = obs[c];
        if (o.time_filt) {
            // timedomain_filt: the zero-phase low-pass, then the high-pass (src/timedomain_filt.c:57-111)
            const TFilt lp = tfilt(dt, 1.0 / (double)o.f_low_pass, o.order, 0);
            hipLaunchKernelGGL(k_tfilt, tg, tb, 0, st, sy, tmp, ntr, ns, lp);
            hipLaunchKernelGGL(k_tfilt, tg, tb, 0, st, (const float *)tmp, syncopy, ntr, ns, lp);
            hipLaunchKernelGGL(k_tfilt, tg, tb, 0, st, ob, tmp, ntr, ns, lp);
            hipLaunchKernelGGL(k_tfilt, tg, tb, 0, st, (const float *)tmp, obsf, ntr, ns, lp);
            if (o.f_high_pass > 0.0f) {
                const TFilt hp = tfilt(dt, 1.0 / (double)o.f_high_pass, o.order, 1);
                hipLaunchKernelGGL(k_tfilt, tg, tb, 0, st, (const float *)syncopy, tmp, ntr, ns, hp);
                hipLaunchKernelGGL(k_tfilt, tg, tb, 0, st, (const float *)tmp, syncopy, ntr, ns, hp);
                hipLaunchKernelGGL(k_tfilt, tg, tb, 0, st, (const float *)obsf, tmp, ntr, ns, hp);
                hipLaunchKernelGGL(k_tfilt, tg, tb, 0, st, (const float *)tmp, obsf, ntr, ns, hp);
            }
            sy = syncopy;
            ob = obsf;
        }
        hipLaunchKernelGGL(k_residual, tg, tb, 0, st, sy, ob, adj, partial, ntr, ns, d.nt, o.lnorm, d.velocity, d.dt,
                           o.waterlevel_lnorm8);
        hipLaunchKernelGGL(k_misfit_sum, dim3(1), dim3(64), 0, st, (const float *)partial, ntr, mis + c);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

} // namespace

extern "C" {

int sofi_fwi_residual(sofi_dev *h, const sofi_residual_opts *o, const float *obs_x, const float *obs_y, double *misfit)
{
    if (const char *e = residual_check(h, o)) return fail("sofi_fwi_residual: %s", e);
    const sofi_desc &d = h->d;
    const bool ux = d.adjoint_type == 1 || d.adjoint_type == 3, uy = d.adjoint_type == 1 || d.adjoint_type == 2;
    if ((ux && !obs_x) || (uy && !obs_y)) return fail("sofi_fwi_residual: measured section missing for ADJOINT_TYPE=%d", d.adjoint_type);
    if (adj_at_receivers(h)) return -1;
    int first, ntr;
    adj_rows(h, &first, &ntr);
    const size_t nn = (size_t)ntr * d.ns, need = 2 * nn + residual_ws_floats(ntr, d.ns);
    if (h->rbuf_len < need) {
        HIPCHK(hipStreamSynchronize(h->stream));
        dfree(h, &h->rbuf);
        h->rbuf_len = 0;
        if (dalloc_t(h, &h->rbuf, need)) return -1;
        h->rbuf_len = need;
    }
    float *ox = h->rbuf, *oy = h->rbuf + nn, *ws = h->rbuf + 2 * nn;
    if (ux && nn) HIPCHK(hipMemcpyAsync(ox, obs_x, nn * sizeof(float), hipMemcpyHostToDevice, h->stream));
    if (uy && nn) HIPCHK(hipMemcpyAsync(oy, obs_y, nn * sizeof(float), hipMemcpyHostToDevice, h->stream));
    const float *obs[2] = { ox, oy };
    if (residual_core(h, *o, obs, ws)) return -1;
    float mis[2] = { 0.f, 0.f };
    HIPCHK(hipMemcpyAsync(mis, ws + 3 * nn + ntr, sizeof(mis), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (misfit) *misfit = (double)mis[0] + (double)mis[1];
    return 0;
}

int sofi_fwi_get_adjoint(sofi_dev *h, int comp, float *out)
{
    if (!h) return fail("sofi_fwi_get_adjoint: null handle");
    if (h->d.mode != SOFI_MODE_FWI) return fail("sofi_fwi_get_adjoint: the handle was not opened with MODE=FWI");
    if (comp != SOFI_C_VX && comp != SOFI_C_VY) return fail("sofi_fwi_get_adjoint: component %d (0 vx, 1 vy)", comp);
    if (!out) return fail("sofi_fwi_get_adjoint: null output");
    const float *a = comp == SOFI_C_VX ? h->adj_sx : h->adj_sy;
    int first, n;
    adj_rows(h, &first, &n);
    HIPCHK(hipStreamSynchronize(h->stream));
    if (a && n > 0)
        HIPCHK(hipMemcpy(out, a + (size_t)first * h->d.nt, (size_t)n * h->d.nt * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int sofi_fwi_adjoint_count(const sofi_dev *h)
{
    if (!h) return 0;
    int first, n;
    adj_rows(h, &first, &n);
    return n;
}

} // extern "C"

// ------------------------------------------- residuals that allocate nothing
//
// sofi_fwi_residual_noalloc runs sofi_fwi_residual's kernels on a workspace
// in this library's static device array `scratch` instead of one it
// allocates, for callers that cannot allocate device memory.  Each GPU holds
// its own copy of the array (SCRATCH_FLOATS floats, 16 MiB) from the first
// use until the process ends; its device address is looked up once per
// device with hipGetSymbolAddress and used from then on (host code cannot
// take the array's address by name).  Calls on one GPU share the array and
// are serialised by a per-device lock held until the call's work has
// finished on the device.

constexpr size_t SCRATCH_FLOATS = size_t(4) << 20;
constexpr int SCRATCH_MAXDEV = 64;
static __device__ __align__(256) float scratch[SCRATCH_FLOATS];

namespace {

std::mutex g_scratch_mu[SCRATCH_MAXDEV];
std::mutex g_scratch_lookup;
float *g_scratch_ptr[SCRATCH_MAXDEV] = {};

// the device address of `scratch` on device dev (current), looked up once
int scratch_of(int dev, float **out)
{
    if (dev < 0 || dev >= SCRATCH_MAXDEV) return fail("device %d: no scratch slot (at most %d devices)", dev, SCRATCH_MAXDEV);
    std::lock_guard<std::mutex> lk(g_scratch_lookup);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<float *>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// true when p is device memory (of any device) or memory the device can address
bool device_pointer(const void *p)
{
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

} // namespace

extern "C" {

int sofi_fwi_residual_noalloc(sofi_dev *h, const sofi_residual_opts *o, const float *obs_x, const float *obs_y,
                              float *misfit)
{
    if (const char *e = residual_check(h, o)) return fail("sofi_fwi_residual_noalloc: %s", e);
    const sofi_desc &d = h->d;
    const bool ux = d.adjoint_type == 1 || d.adjoint_type == 3, uy = d.adjoint_type == 1 || d.adjoint_type == 2;
    if ((ux && !obs_x) || (uy && !obs_y))
        return fail("sofi_fwi_residual_noalloc: measured section missing for ADJOINT_TYPE=%d", d.adjoint_type);
    if ((ux && !device_pointer(obs_x)) || (uy && !device_pointer(obs_y)) || (misfit && !device_pointer(misfit)))
        return fail("sofi_fwi_residual_noalloc: the measured sections and the misfit must be device memory");
    if (!h->adj_at_rec || (int)h->adj_start.size() != h->nb + 1 ||
        h->adj_start[h->sel + 1] - h->adj_start[h->sel] != (int)h->slots[h->sel].ri.size())
        return fail("sofi_fwi_residual_noalloc: the adjoint sources are not those of the receivers (set by "
                    "sofi_fwi_set_adjoint or receivers changed): call sofi_fwi_residual once first");
    int first, ntr;
    adj_rows(h, &first, &ntr);
    const size_t need = residual_ws_floats(ntr, d.ns);
    if (need > SCRATCH_FLOATS)
        return fail("sofi_fwi_residual_noalloc: %d traces of %d samples need %zu floats of scratch, more than its "
                    "%zu; use sofi_fwi_residual", ntr, d.ns, need, SCRATCH_FLOATS);
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != h->dev) return fail("sofi_fwi_residual_noalloc: device %d is current, the handle's is %d", cur, h->dev);
    float *ws = nullptr;
    if (scratch_of(h->dev, &ws)) return -1;
    std::lock_guard<std::mutex> lk(g_scratch_mu[h->dev]);
    const float *obs[2] = { obs_x, obs_y };
    if (residual_core(h, *o, obs, ws)) return -1;
    if (misfit) HIPCHK(hipMemcpyAsync(misfit, ws + 3 * (size_t)ntr * d.ns + ntr, 2 * sizeof(float),
                                      hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

} // extern "C"


namespace {

// steps nt_first..nt_last of every slot: a batch advances its slots in
// resident groups (all slots together when one group holds them), each group
// through the whole step range with the sampling state of the range's
// start; groups go to S streams at once (the extra streams wait for the work
// enqueued so far on the handle's stream, which waits for them at the end)
constexpr int RUN_TURN = 32;

int run_groups(sofi_dev *h, int nt_first, int nt_last)
{
    const int G = resident_slots(h), lsamp0 = h->lsamp, ngroups = (h->nb + G - 1) / G;
    const int S = batch_streams(h, ngroups);
    if (S > 1) {
        while ((int)h->xs.size() < S - 1) {
            hipStream_t x = nullptr;
            hipEvent_t e = nullptr;
            HIPCHK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
            h->xs.push_back(x);
            HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            h->xev.push_back(e);
        }
        if (!h->evfork) HIPCHK(hipEventCreateWithFlags(&h->evfork, hipEventDisableTiming));
        HIPCHK(hipEventRecord(h->evfork, h->stream));
        for (int q = 0; q < S - 1; ++q) HIPCHK(hipStreamWaitEvent(h->xs[q], h->evfork, 0));
    }
    int rc = 0;
    if (S == 1) {
        for (int gi = 0; gi < ngroups && !rc; ++gi) {
            h->cur_lo = gi * G;
            h->cur_hi = h->cur_lo + G < h->nb ? h->cur_lo + G : h->nb;
            h->lsamp = lsamp0;
            rc = run_range(h, nt_first, nt_last);
        }
    } else {
        // the S groups that run at once (one per stream) are enqueued in turns
        // of RUN_TURN steps each, so that no stream waits while the host
        // enqueues another group's whole step range; each group keeps its own
        // sampling state
        std::vector<int> ls(ngroups, lsamp0);
        for (int g0 = 0; g0 < ngroups && !rc; g0 += S) {
            const int g1 = g0 + S < ngroups ? g0 + S : ngroups;
            for (int a = nt_first; a <= nt_last && !rc; a += RUN_TURN) {
                const int b = a + RUN_TURN - 1 < nt_last ? a + RUN_TURN - 1 : nt_last;
                for (int gi = g0; gi < g1 && !rc; ++gi) {
                    h->cur_lo = gi * G;
                    h->cur_hi = h->cur_lo + G < h->nb ? h->cur_lo + G : h->nb;
                    h->cur = gi % S == 0 ? nullptr : h->xs[gi % S - 1];
                    h->lsamp = ls[gi];
                    rc = run_range(h, a, b);
                    ls[gi] = h->lsamp;
                }
            }
        }
    }
    h->cur = nullptr;
    h->cur_lo = 0;
    h->cur_hi = h->nb;
    for (int q = 0; q < S - 1; ++q) {
        HIPCHK(hipEventRecord(h->xev[q], h->xs[q]));
        HIPCHK(hipStreamWaitEvent(h->stream, h->xev[q], 0));
    }
    return rc ? -1 : 0;
}

// ------------------------------------------------------------ autotuner
//
// A batch's launch configuration (slot blocks and their slots per thread,
// or the slot-loop kernels one slot or all slots per launch; resident
// groups; streams) is measured on the first steps of its first run and
// kept for every later run of the handle and for every later batch of the
// same rheology, FD order, grid, slot count and boundaries in the process.
// Each candidate advances all slots by one step untimed (code loading, the
// caches filled) and TUNE_STEPS timed steps.  The candidates are of one
// family of kernels (slot blocks; the slot-loop kernels for AC_ISO), whose
// results are the same bits for every launch, so the choice changes no bit.

struct TuneKey {
    int weq, nd, nx, ny, nb, abs_type, free_surf, L, periodic, mode;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(weq, nd, nx, ny, nb, abs_type, free_surf, L, periodic, mode) <
               std::tie(o.weq, o.nd, o.nx, o.ny, o.nb, o.abs_type, o.free_surf, o.L, o.periodic, o.mode);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, sofi_dev::BatchCfg> g_tune_cache;
constexpr int TUNE_STEPS = 4;

void apply_cfg(sofi_dev *h, const sofi_dev::BatchCfg &c)
{
    h->bs = c.bs;
    h->res_mode = c.res;
    h->str_mode = c.streams;
    h->grp_mode = c.grp;
    h->spt_mode = c.spt;
    h->grp = c.grp;
    h->g.spt = c.spt > 0 ? c.spt : 1;
    h->spt_tuned = true;
    h->gen++;
}

int batch_tune(sofi_dev *h, int *nt, int nt_last)
{
    const sofi_desc &d = h->d;
    if (h->nb == 1 || !h->tune_on || h->tuned || h->fdt4) return 0;
    const TuneKey key{ d.weq, h->g.nd, d.nx, d.ny, h->nb, d.abs_type, d.free_surf, h->g.L, h->g.periodic, d.mode };
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            apply_cfg(h, it->second);
            h->tuned = true;
            h->tune_cfg = it->second;
            return 0;
        }
    }
    const int nb = h->nb;
    std::vector<sofi_dev::BatchCfg> cand;
    // the candidates of one family of kernels, whose results are the same
    // bits whatever the launch: slot blocks (1, 2 or 4 slots per thread) over
    // all slots, or over cache-sized groups on two streams; for AC_ISO, which
    // has no slot blocks, the slot-loop kernels (one slot per launch in
    // cache-sized groups on two streams; all slots per launch, up to four
    // per thread)
    if (d.weq != SOFI_AC_ISO) {
        for (int bs : { 1, 2, 4 }) {
            if (bs > nb) continue;
            cand.push_back({ bs, nb, 1, 0, 0 });
            cand.push_back({ bs, 0, 2, 0, 0 });
        }
    } else {
        cand.push_back({ 0, 0, 2, 1, 1 });
        cand.push_back({ 0, nb, 1, nb, nb >= 4 ? 4 : 1 });
    }
    const int need = (int)cand.size() * (TUNE_STEPS + 1);
    if (nt_last - *nt + 1 < 2 * need) {
        // too short a run to time: the first candidate, so that the steps
        // run the kernels of the family a later tuned run chooses from
        if (h->bs != cand[0].bs) apply_cfg(h, cand[0]);
        return 0;
    }
    float best = 1e30f;
    sofi_dev::BatchCfg bestc = cand[0];
    for (const auto &c : cand) {
        apply_cfg(h, c);
        if (run_groups(h, *nt, *nt)) return -1;
        ++*nt;
        HIPCHK(hipEventRecord(h->tune0, h->stream));
        if (run_groups(h, *nt, *nt + TUNE_STEPS - 1)) return -1;
        *nt += TUNE_STEPS;
        HIPCHK(hipEventRecord(h->tune1, h->stream));
        HIPCHK(hipEventSynchronize(h->tune1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, h->tune0, h->tune1));
        if (ms < best) {
            best = ms;
            bestc = c;
        }
    }
    apply_cfg(h, bestc);
    h->tuned = true;
    h->tune_cfg = bestc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = bestc;
    return 0;
}

// ------------------------------------------- work-list runs (sofi_run_many)

// what keeps a handle out of a work-list run (null: nothing)
const char *many_refusal(const sofi_dev *h)
{
    if (!h->medium_set) return "its medium is not set";
    if (h->nb != 1) return "it is a batch handle";
    if (h->d.mode == SOFI_MODE_FWI) return "it runs MODE=FWI";
    if (h->fdt4) return "it runs 4th order in time";
    switch (h->d.weq) {
      case SOFI_EL_ISO:
      case SOFI_VEL_ISO:
      case SOFI_EL_VTI:
      case SOFI_VEL_VTI:
      case SOFI_AC_ISO: return nullptr;
    }
    return "the TTI rheologies have no work-list kernels";
}

// a work-list launch: tiles per item (down a tile column), blocks (0: one
// block per item), tile height in rows = waves per block (2, 4 or 8)
struct ManyCfg {
    int rows, blocks, waves;
};
std::mutex g_many_mu;
std::map<std::tuple<int, int, int, int, int>, ManyCfg> g_many_cache;
thread_local ManyCfg g_many_last{1, 0, BY};
thread_local int g_many_tuned = 0;

struct ManyCtx {
    hipStream_t st = nullptr;
    std::vector<sofi_dev *> hs;     // in work-list order: the handles with the most steps first
    std::vector<int> first, len;    // their first step and step count
    std::vector<sofi_dev::Deferred> dl;
    std::vector<Shot> tab;          // this step's descriptors
    std::vector<Shot> sent;         // the descriptors on the device
    Shot *dtab = nullptr;
    int2 *ditem = nullptr;
    int rows = 0, waves = 0;        // tiles per item and tile height of the uploaded items
    std::vector<int> cum;           // cum[a]: items of the first a handles
    hipError_t err = hipSuccess;
};

// the items of `rows` tiles of BX x waves cells each, handle after handle,
// bands of tile rows in order and the tile columns of a band side by side
int many_items(ManyCtx &c, int rows, int waves)
{
    if (c.rows == rows && c.waves == waves) return 0;
    std::vector<int2> it;
    c.cum.assign(1, 0);
    for (size_t p = 0; p < c.hs.size(); ++p) {
        const Grid &g = c.hs[p]->g;
        const int nty = (g.ny + waves - 1) / waves;
        for (int ty = 0; ty < nty; ty += rows)
            for (int tx = 0; tx < g.ntx; ++tx) it.push_back(make_int2((int)p, ty * g.ntx + tx));
        c.cum.push_back((int)it.size());
    }
    // the launches that read the old items end first
    HIPCHK(hipStreamSynchronize(c.st));
    HIPCHK(hipMemcpy(c.ditem, it.data(), it.size() * sizeof(int2), hipMemcpyHostToDevice));
    c.rows = rows;
    c.waves = waves;
    return 0;
}

// one update launch of a work-list step in tiles of BX x TBY cells
template <int ND, int TBY>
void many_launch(hipStream_t st, int weq, bool stress, unsigned grid, const Shot *tab, const int2 *item, int n,
                 int rows, int it)
{
    const dim3 blk(BX, TBY);
    if (!stress) {
        hipLaunchKernelGGL((k_velocity_w<ND, true, TBY>), dim3(grid), blk, 0, st, tab, item, n, rows, it);
        return;
    }
    switch (weq) {
      case SOFI_EL_ISO: hipLaunchKernelGGL((k_stress_w<ND, SOFI_EL_ISO, TBY>), dim3(grid), blk, 0, st, tab, item, n, rows, it); break;
      case SOFI_VEL_ISO: hipLaunchKernelGGL((k_stress_w<ND, SOFI_VEL_ISO, TBY>), dim3(grid), blk, 0, st, tab, item, n, rows, it); break;
      case SOFI_EL_VTI: hipLaunchKernelGGL((k_stress_w<ND, SOFI_EL_VTI, TBY>), dim3(grid), blk, 0, st, tab, item, n, rows, it); break;
      case SOFI_AC_ISO: hipLaunchKernelGGL((k_stress_w<ND, SOFI_AC_ISO, TBY>), dim3(grid), blk, 0, st, tab, item, n, rows, it); break;
      default: hipLaunchKernelGGL((k_stress_w<ND, SOFI_VEL_VTI, TBY>), dim3(grid), blk, 0, st, tab, item, n, rows, it);
    }
}

// local step k (0-based) of every handle that still advances: each handle's
// step recorded (its update kernels' arguments and its other launches), then
// its launches before the velocity update, the velocity update of all of
// them, the launches between, the stress update of all of them, the rest
template <int ND>
int many_step(ManyCtx &c, int k, ManyCfg cfg)
{
    int a = 0;
    while (a < (int)c.hs.size() && c.len[a] > k) ++a;
    c.dl.clear();
    for (int p = 0; p < a; ++p) {
        sofi_dev *h = c.hs[p];
        memset(static_cast<void *>(&h->wl_shot), 0, sizeof(Shot));
        h->wl = true;
        h->defer = &c.dl;
        h->dstage = 0;
        h->wl_dv = h->wl_seis = -1;
        const int rc = step(h, c.first[p] + k);
        h->wl = false;
        h->defer = nullptr;
        if (rc) return rc;
        if (h->dstage != 2) return fail("sofi_run_many: step %d of a handle recorded no update kernels", c.first[p] + k);
        memcpy(static_cast<void *>(&c.tab[p]), &h->wl_shot, sizeof(Shot));
    }
    // the descriptors are uploaded when they change (in the first step: the
    // arguments of a handle's update kernels are the same in every step of a
    // work-list run), after the launches that read the old ones
    if (c.sent.empty() || memcmp(c.sent.data(), c.tab.data(), (size_t)a * sizeof(Shot)) != 0) {
        HIPCHK(hipStreamSynchronize(c.st));
        HIPCHK(hipMemcpy(c.dtab, c.tab.data(), c.tab.size() * sizeof(Shot), hipMemcpyHostToDevice));
        c.sent = c.tab;
    }
    if (many_items(c, cfg.rows, cfg.waves)) return -1;
    const int nit = c.cum[a];
    const unsigned grid = (unsigned)(cfg.blocks > 0 && cfg.blocks < nit ? cfg.blocks : nit);
    auto issue = [&](int stage) {
        for (auto &e : c.dl)
            if (e.stage == stage) {
                const hipError_t r = e.run(c.st);
                if (r != hipSuccess && c.err == hipSuccess) c.err = r;
            }
    };
    auto update = [&](bool stress) {
        const int weq = c.hs[0]->d.weq;
        switch (cfg.waves) {
          case 2: many_launch<ND, 2>(c.st, weq, stress, grid, c.dtab, c.ditem, nit, cfg.rows, k + 1); break;
          case 8: many_launch<ND, 8>(c.st, weq, stress, grid, c.dtab, c.ditem, nit, cfg.rows, k + 1); break;
          default: many_launch<ND, 4>(c.st, weq, stress, grid, c.dtab, c.ditem, nit, cfg.rows, k + 1);
        }
    };
    // the samplers of every handle that samples in this step, one launch each
    // for div / curl (before the rest of the step, as k_seis_dv runs) and for
    // vx / vy / p (after it, as k_seis runs)
    auto sample = [&](bool dv) {
        SeisJobs jobs{};
        int nj = 0, nrec = 0;
        for (int p = 0; p <= a; ++p) {
            if (nj == SEIS_JOBS || (p == a && nj > 0)) {
                const dim3 sg((unsigned)((nrec + 63) / 64), (unsigned)nj);
                if (dv) hipLaunchKernelGGL((k_seis_w<ND, true>), sg, dim3(64), 0, c.st, c.dtab, jobs);
                else hipLaunchKernelGGL((k_seis_w<ND, false>), sg, dim3(64), 0, c.st, c.dtab, jobs);
                nj = nrec = 0;
            }
            if (p == a) break;
            const sofi_dev *h = c.hs[p];
            const int ins = dv ? h->wl_dv : h->wl_seis;
            if (ins < 0 || h->r.n == 0) continue;
            jobs.shot[nj] = p;
            jobs.ins[nj] = ins;
            nrec = std::max(nrec, h->r.n);
            ++nj;
        }
    };
    issue(0);
    update(false);
    issue(1);
    update(true);
    sample(true);
    issue(2);
    sample(false);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess && c.err == hipSuccess) c.err = e;
    return 0;
}

constexpr int MANY_TUNE_STEPS = 2;

// all steps of a work-list run: the launch fixed by the caller, cached for
// this rheology, FD order, handle count and size, or measured first (each
// candidate on MANY_TUNE_STEPS steps of the run after one untimed step;
// results do not depend on it, every candidate runs the same per-cell code)
template <int ND>
int run_many_nd(ManyCtx &c, const int *launch)
{
    const sofi_dev *h0 = c.hs[0];
    long long tiles = 0;
    for (const sofi_dev *h : c.hs) tiles += (long long)h->g.ntx * h->g.nty;
    int lg = 0;
    while ((1ll << (lg + 1)) <= tiles) ++lg;
    const auto key = std::make_tuple(ND, h0->d.weq, h0->g.L, (int)c.hs.size(), lg);
    ManyCfg cfg{1, 0, BY};
    int k = 0;
    g_many_tuned = 0;
    bool cached = false;
    if (launch) {
        cfg = ManyCfg{launch[0], launch[1], launch[2]};
    } else {
        std::lock_guard<std::mutex> lk(g_many_mu);
        const auto f = g_many_cache.find(key);
        if (f != g_many_cache.end()) {
            cfg = f->second;
            cached = true;
            g_many_tuned = 1;
        }
    }
    std::vector<ManyCfg> cand;
    for (int waves : { 2, 4, 8 })
        for (int rows : { 1, 2, 4 })
            for (int blocks : { 0, 2048 }) cand.push_back(ManyCfg{rows, blocks, waves});
    if (!launch && !cached && c.len[0] >= 1 + (int)cand.size() * MANY_TUNE_STEPS) {
        if (many_step<ND>(c, k++, cfg)) return -1;
        std::vector<hipEvent_t> ev(2 * cand.size(), nullptr);
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
        int rc = 0;
        for (size_t q = 0; q < cand.size() && !rc; ++q) {
            rc = many_items(c, cand[q].rows, cand[q].waves);     // outside the timed steps
            if (!rc && hipEventRecord(ev[2 * q], c.st) != hipSuccess) rc = fail("sofi_run_many: hipEventRecord failed");
            for (int s = 0; s < MANY_TUNE_STEPS && !rc; ++s) rc = many_step<ND>(c, k++, cand[q]);
            if (!rc && hipEventRecord(ev[2 * q + 1], c.st) != hipSuccess) rc = fail("sofi_run_many: hipEventRecord failed");
        }
        if (!rc && hipEventSynchronize(ev.back()) != hipSuccess) rc = fail("sofi_run_many: hipEventSynchronize failed");
        float best = 0.f;
        for (size_t q = 0; q < cand.size() && !rc; ++q) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[2 * q], ev[2 * q + 1]) != hipSuccess) {
                rc = fail("sofi_run_many: hipEventElapsedTime failed");
                break;
            }
            if (q == 0 || ms < best) {
                best = ms;
                cfg = cand[q];
            }
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        if (rc) return rc;
        std::lock_guard<std::mutex> lk(g_many_mu);
        g_many_cache[key] = cfg;
        g_many_tuned = 1;
    }
    for (; k < c.len[0]; ++k)
        if (many_step<ND>(c, k, cfg)) return -1;
    g_many_last = cfg;
    return 0;
}

} // namespace

extern "C" {

int sofi_run(sofi_dev *h, int nt_first, int nt_last)
{
    if (!h) return fail("sofi_run: null handle");
    if (!h->medium_set) return fail("sofi_run: medium not set");
    if (nt_first < 1 || nt_last > h->d.nt || nt_first > nt_last) return fail("sofi_run: bad step range %d..%d (NT=%d)", nt_first, nt_last, h->d.nt);
    if (sync_acq(h)) return -1;
    h->gerr = hipSuccess;
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    int nt = nt_first;
    if (batch_tune(h, &nt, nt_last)) return -1;
    const int rc = nt <= nt_last ? run_groups(h, nt, nt_last) : 0;
    if (rc) return -1;
    if (h->gerr != hipSuccess) {
        const hipError_t e = h->gerr;
        h->gerr = hipSuccess;
        return fail("kernel launch failed: %s", hipGetErrorString(e));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    for (int k = 0; k + 4 <= h->tev_used; k += 4) {
        float a = 0.f, b = 0.f;
        HIPCHK(hipEventElapsedTime(&a, h->tev[k], h->tev[k + 1]));
        HIPCHK(hipEventElapsedTime(&b, h->tev[k + 2], h->tev[k + 3]));
        h->t_vel_ms += a;
        h->t_str_ms += b;
        h->t_samples++;
    }
    h->tev_used = 0;
    std::vector<int> bad(h->nb, 0);
    HIPCHK(hipMemcpy(bad.data(), h->f.unstable, sizeof(int) * h->nb, hipMemcpyDeviceToHost));
    for (int k = 0; k < h->nb; ++k)
        if (bad[k]) {
            if (h->nb == 1) return fail("Simulation is unstable: vy(NX/2, NY/2) is NaN at time step %d.", bad[k]);
            return fail("Simulation is unstable: vy(NX/2, NY/2) of shot slot %d is NaN at time step %d.", k, bad[k]);
        }
    return 0;
}

int sofi_run_many(sofi_dev *const *hs, int n, const int *nt_first, const int *nt_last, const int *launch)
{
    if (!hs || n < 1 || !nt_first || !nt_last) return fail("sofi_run_many: null argument or no handles");
    if (launch && (launch[0] < 1 || launch[0] > 64 || launch[1] < 0 || (launch[2] != 2 && launch[2] != 4 && launch[2] != 8)))
        return fail("sofi_run_many: launch {%d rows, %d blocks, %d waves}: rows 1..64, blocks >= 0, waves 2, 4 or 8",
                    launch[0], launch[1], launch[2]);
    for (int i = 0; i < n; ++i) {
        const sofi_dev *h = hs[i];
        if (!h) return fail("sofi_run_many: handle %d is null", i);
        for (int q = 0; q < i; ++q)
            if (hs[q] == h) return fail("sofi_run_many: handle %d is handle %d again", i, q);
        if (const char *why = many_refusal(h)) return fail("sofi_run_many: handle %d: %s", i, why);
        if (h->dev != hs[0]->dev || h->g.nd != hs[0]->g.nd || h->d.weq != hs[0]->d.weq)
            return fail("sofi_run_many: handle %d differs from handle 0 in device, FD order or rheology", i);
        if (nt_first[i] < 1 || nt_last[i] > h->d.nt || nt_first[i] > nt_last[i])
            return fail("sofi_run_many: handle %d: bad step range %d..%d (NT=%d)", i, nt_first[i], nt_last[i], h->d.nt);
    }
    ManyCtx c;
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return nt_last[a] - nt_first[a] > nt_last[b] - nt_first[b];
    });
    long long tiles = 0;
    for (int i : order) {
        sofi_dev *h = hs[i];
        if (sync_acq(h)) return -1;
        h->gerr = hipSuccess;
        c.hs.push_back(h);
        c.first.push_back(nt_first[i]);
        c.len.push_back(nt_last[i] - nt_first[i] + 1);
        tiles += (long long)h->g.ntx * ((h->g.ny + 1) / 2);     // the most items: 1 tile of 2 rows each
    }
    c.tab.assign(n, Shot{});
    c.st = hs[0]->stream;
    for (int i = 0; i < n; ++i) {
        HIPCHK(hipEventRecord(hs[i]->ev0, hs[i]->stream));
        HIPCHK(hipStreamWaitEvent(c.st, hs[i]->ev0, 0));
    }
    HIPCHK(hipMalloc(&c.dtab, sizeof(Shot) * (size_t)n));
    if (hipMalloc(&c.ditem, sizeof(int2) * (size_t)tiles) != hipSuccess) {
        (void)hipFree(c.dtab);
        return fail("sofi_run_many: hipMalloc of %lld work items failed", tiles);
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int rc = hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess
                 ? fail("sofi_run_many: hipEventCreate failed") : 0;
    for (size_t p = 0; p < c.hs.size() && !rc; ++p)
        hipLaunchKernelGGL(k_set_base, dim3(1), dim3(64), 0, c.st, c.hs[p]->ntb, c.first[p] - 1);
    if (!rc && hipEventRecord(t0, c.st) != hipSuccess) rc = fail("sofi_run_many: hipEventRecord failed");
    if (!rc) {
        switch (hs[0]->g.nd) {
          case 1: rc = run_many_nd<1>(c, launch); break;
          case 2: rc = run_many_nd<2>(c, launch); break;
          case 3: rc = run_many_nd<3>(c, launch); break;
          case 4: rc = run_many_nd<4>(c, launch); break;
          case 5: rc = run_many_nd<5>(c, launch); break;
          case 6: rc = run_many_nd<6>(c, launch); break;
          default: rc = fail("unsupported ND %d", hs[0]->g.nd);
        }
    }
    float ms = 0.f;
    if (hipEventRecord(t1, c.st) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
        hipEventElapsedTime(&ms, t0, t1) != hipSuccess)
        rc = rc ? rc : fail("sofi_run_many: timing events failed");
    (void)hipStreamSynchronize(c.st);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(c.ditem);
    (void)hipFree(c.dtab);
    if (rc) return -1;
    if (c.err != hipSuccess) return fail("sofi_run_many: kernel launch failed: %s", hipGetErrorString(c.err));
    for (int i = 0; i < n; ++i) {
        sofi_dev *h = hs[i];
        h->last_ms = ms;
        int bad = 0;
        HIPCHK(hipMemcpy(&bad, h->f.unstable, sizeof(int), hipMemcpyDeviceToHost));
        if (bad) return fail("Simulation is unstable: vy(NX/2, NY/2) of handle %d is NaN at time step %d.", i, bad);
    }
    return 0;
}

int sofi_get_many_config(int *rows, int *blocks, int *waves, int *tuned)
{
    if (rows) *rows = g_many_last.rows;
    if (blocks) *blocks = g_many_last.blocks;
    if (waves) *waves = g_many_last.waves;
    if (tuned) *tuned = g_many_tuned;
    return 0;
}

int sofi_set_batch_launch(sofi_dev *h, int group, int spt)
{
    if (!h) return fail("sofi_set_batch_launch: null handle");
    if (group < 0 || group > h->nb) return fail("sofi_set_batch_launch: group %d (0..%d)", group, h->nb);
    if (spt < 0 || spt > (group > 0 ? group : h->nb)) return fail("sofi_set_batch_launch: %d slots per thread", spt);
    h->grp_mode = group;
    h->spt_mode = spt;
    h->spt_tuned = false;
    h->grp = group;
    h->g.spt = spt > 0 ? spt : 1;
    if (group > 0 || spt > 0) h->tune_on = false;
    h->gen++;
    return 0;
}

int sofi_set_batch_block(sofi_dev *h, int slots)
{
    if (!h) return fail("sofi_set_batch_block: null handle");
    if (slots != 0 && slots != 1 && slots != 2 && slots != 4)
        return fail("sofi_set_batch_block: slots per thread must be 0, 1, 2 or 4 (got %d)", slots);
    h->bs = slots;
    h->tune_on = false;
    h->gen++;
    return 0;
}

int sofi_get_batch_block(const sofi_dev *h) { return h ? h->bs : 0; }

int sofi_get_batch_launch(const sofi_dev *h, int *group, int *spt)
{
    if (!h) return fail("sofi_get_batch_launch: null handle");
    if (group) *group = h->grp > 0 ? h->grp : h->nb;
    if (spt) *spt = h->g.spt;
    return 0;
}

double sofi_last_run_ms(const sofi_dev *h) { return h ? h->last_ms : 0.0; }

int sofi_get_section(sofi_dev *h, int comp, float *out)
{
    if (!h || !out) return fail("sofi_get_section: null argument");
    if (comp < 0 || comp >= SOFI_NCOMP) return fail("sofi_get_section: bad component %d", comp);
    if (sync_acq(h)) return -1;
    const size_t ntr = h->slots[h->sel].ri.size(), n = ntr * h->d.ns;
    if (n == 0) return 0;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(out, h->sec + h->sec_base[h->sel] + (size_t)comp * n, n * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int sofi_snapshot_count(const sofi_dev *h) { return h ? h->nsnap : 0; }

int sofi_get_snapshot(sofi_dev *h, int k, int comp, float *out)
{
    if (!h || !out) return fail("sofi_get_snapshot: null argument");
    if (k < 0 || k >= h->nsnap) return fail("sofi_get_snapshot: snapshot %d not taken (have %d)", k, h->nsnap);
    if (comp < 0 || comp >= SOFI_NCOMP) return fail("sofi_get_snapshot: bad component %d", comp);
    size_t n = (size_t)h->nxs * h->nys;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(out, h->snapbuf + ((size_t)k * SOFI_NCOMP + comp) * n, n * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int sofi_get_field(sofi_dev *h, int id, float *out)
{
    if (!h || !out) return fail("sofi_get_field: null argument");
    float *p = nullptr;
    switch (id) {
      case 0: p = h->f.vx; break;
      case 1: p = h->f.vy; break;
      case 2: p = h->f.sxx; break;
      case 3: p = h->f.syy; break;
      case 4: p = h->f.sxy; break;
      default: return fail("sofi_get_field: bad field id %d", id);
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(out, p + (size_t)h->sel * h->g.plane, (size_t)h->g.plane * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int sofi_close(sofi_dev *h)
{
    if (!h) return 0;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (hipStream_t x : h->xs) (void)hipStreamSynchronize(x);
    graph_drop(h);
    for (hipStream_t x : h->xs) (void)hipStreamDestroy(x);
    for (hipEvent_t e : h->xev) (void)hipEventDestroy(e);
    if (h->evfork) (void)hipEventDestroy(h->evfork);
    for (hipEvent_t e : h->tev) (void)hipEventDestroy(e);
    for (const auto &a : h->allocs) (void)hipFree(a.raw);
    for (hipEvent_t e : { h->ev0, h->ev1, h->tune0, h->tune1 })
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

size_t sofi_footprint(const sofi_dev *h) { return h ? h->bytes : 0; }

size_t sofi_estimate_footprint(const sofi_desc *d, int nshots, int ntr, int nsnap)
{
    if (!d || d->nx < 1 || d->ny < 1 || nshots < 1) return 0;
    // the allocations of sofi_open_batch, sofi_set_medium, the acquisition and
    // sofi_set_snapshots, each rounded up by the 256 B of its alignment
    const size_t pl = sofi_plane(d->nx, d->ny) * sizeof(float), nb = (size_t)nshots, A = 256;
    const bool visco = d->weq == SOFI_VEL_ISO || d->weq == SOFI_VEL_VTI || d->weq == SOFI_VEL_TTI;
    const bool tti = d->weq == SOFI_EL_TTI || d->weq == SOFI_VEL_TTI;
    const size_t L = visco ? (size_t)d->L : 0, wx = 2 * (size_t)d->fw + 1;
    size_t b = 5 * (pl * nb + A) + (nb * sizeof(int) + A);                  // v, s, instability flags
    if (visco) b += 3 * (pl * L * nb + A);                                    // memory variables
    if (tti) b += 4 * (pl * nb + A);                                          // derivative arrays
    if (d->fdorder_time == 4) {                                               // Adams-Bashforth histories
        b += 24 * (pl + A);
        if (visco) b += 9 * (pl * L + A);
    }
    if (d->abs_type == 1 && d->fw > 0)
        b += 4 * ((size_t)(d->ny + 1) * wx * 4 * nb + A) + 4 * (wx * (size_t)(d->nx + 1) * 4 * nb + A);
    size_t arrays = 4;                                                        // rip rjp pu ppi
    switch (d->weq) {
      case SOFI_EL_ISO: arrays += 3; break;
      case SOFI_VEL_ISO: arrays += 5 + 3; break;                              // f g fipjp taus taup + bases
      case SOFI_EL_VTI: arrays += 4; break;
      case SOFI_VEL_VTI: arrays += 4 + 4 * L; break;
      case SOFI_EL_TTI: arrays += 8; break;
      case SOFI_VEL_TTI: arrays += 8 + 8 * L; break;
    }
    if (d->abs_type == 2 && d->fw > 0) arrays += 1;
    b += arrays * (pl + A);
    if (d->abs_type == 1 && d->fw > 0) b += 12 * wx * 4 + A;
    if (d->free_surf) b += 5 * (size_t)d->nx * 4 + A;
    const size_t ns = (size_t)(d->ns > 0 ? d->ns : 1), nr = (size_t)(ntr > 0 ? ntr : 0) * nb;
    b += (SOFI_NCOMP * nr * ns * 4 + A) + (nr * 3 * 4 + A) + (nr * 2 * 8 + A) + (nr * 3 * 8 + A);
    if (nsnap > 0) {
        const size_t nxs = (size_t)((d->snap_idx[2] - d->snap_idx[1]) / (d->idx > 0 ? d->idx : 1) + 1);
        const size_t nys = (size_t)((d->snap_idx[4] - d->snap_idx[3]) / (d->idy > 0 ? d->idy : 1) + 1);
        b += (size_t)nsnap * SOFI_NCOMP * nxs * nys * 4 + A;
    }
    if (d->mode == SOFI_MODE_FWI) {
        const size_t dtinv = d->dtinv > 0 ? (size_t)d->dtinv : 1, nstore = ((size_t)d->nt + dtinv - 1) / dtinv;
        b += 8 * (pl * nb + A) + 2 * (3 * pl + A) + 5 * (pl + A);
        b += nstore * 5 * (size_t)d->nx * d->ny * 4 * nb + A;
        b += 2 * ((size_t)(ntr > 0 ? ntr : 1) * nb * d->nt * 4 + A);
    }
    b += (size_t)1 << 20;                                                     // sources and small buffers
    return b;
}

int sofi_fwi_set_model(sofi_dev *h, const float *vp0, const float *vs0, const float *rho0)
{
    if (!h || !vp0 || !vs0 || !rho0) return fail("sofi_fwi_set_model: null argument");
    if (h->d.mode != SOFI_MODE_FWI) return fail("sofi_fwi_set_model: the handle was not opened with MODE=FWI");
    const size_t pl = (size_t)h->g.plane;
    // the moduli calc_conv forms per cell (src/calc_conv.c:45-46), in the
    // reference's float / double rounding
    std::vector<float> muss(pl), lamss(pl);
    for (size_t q = 0; q < pl; ++q) {
        muss[q] = vs0[q] * vs0[q] * rho0[q];
        lamss[q] = (float)(vp0[q] * vp0[q] * rho0[q] - 2.0 * muss[q]);
    }
    float *b[5] = { const_cast<float *>(h->w.vp0), const_cast<float *>(h->w.vs0), const_cast<float *>(h->w.rho0),
                    const_cast<float *>(h->w.muss), const_cast<float *>(h->w.lamss) };
    const float *src[5] = { vp0, vs0, rho0, muss.data(), lamss.data() };
    for (int k = 0; k < 5; ++k)
        if (upload(h, &b[k], src[k], pl)) return -1;
    h->w.vp0 = b[0];
    h->w.vs0 = b[1];
    h->w.rho0 = b[2];
    h->w.muss = b[3];
    h->w.lamss = b[4];
    h->model_set = true;
    h->gen++;
    return 0;
}

int sofi_fwi_set_pass(sofi_dev *h, int sw)
{
    if (!h) return fail("sofi_fwi_set_pass: null handle");
    if (h->d.mode != SOFI_MODE_FWI) return fail("sofi_fwi_set_pass: the handle was not opened with MODE=FWI");
    if (sw != 0 && sw != 1) return fail("sofi_fwi_set_pass: sw must be 0 (forward) or 1 (adjoint)");
    if (sw == 1 && !h->model_set) return fail("sofi_fwi_set_pass: sofi_fwi_set_model first");
    if (h->sw != sw) h->gen++;            // the passes launch different kernels
    h->sw = sw;
    return 0;
}

int sofi_fwi_set_adjoint(sofi_dev *h, int n, const int *ai, const int *aj, const float *sig_x, const float *sig_y)
{
    if (!h) return fail("sofi_fwi_set_adjoint: null handle");
    if (h->d.mode != SOFI_MODE_FWI) return fail("sofi_fwi_set_adjoint: the handle was not opened with MODE=FWI");
    if (h->nb != 1) return fail("sofi_fwi_set_adjoint: one-slot handles only (batches: sofi_fwi_residual per slot)");
    if (n < 0 || (n > 0 && (!ai || !aj || !sig_x || !sig_y))) return fail("sofi_fwi_set_adjoint: bad arguments");
    for (int k = 0; k < n; ++k)
        if (ai[k] < 1 || ai[k] > h->d.nx || aj[k] < 1 || aj[k] > h->d.ny)
            return fail("sofi_fwi_set_adjoint: source %d at (%d, %d) is outside the grid", k, ai[k], aj[k]);
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(h, &h->adj_sx);
    dfree(h, &h->adj_sy);
    const size_t ns = (size_t)(n > 0 ? n : 1) * h->d.nt;
    if (dalloc_t(h, &h->adj_sx, ns) || dalloc_t(h, &h->adj_sy, ns)) return -1;
    if (n > 0) {
        if (h2d(h, h->adj_sx, sig_x, (size_t)n * h->d.nt * sizeof(float))) return -1;
        if (h2d(h, h->adj_sy, sig_y, (size_t)n * h->d.nt * sizeof(float))) return -1;
    }
    if (build_inj(h, n, ai, aj, nullptr, nullptr, &h->adj_inj, &h->adj_inj_buf)) return -1;
    h->nadj = n;
    h->adj_at_rec = false;
    h->adj_custom = true;
    h->gen++;
    return 0;
}

int sofi_fwi_gradient(sofi_dev *h, float *grad_vp, float *grad_vs, float *grad_rho)
{
    if (!h) return fail("sofi_fwi_gradient: null handle");
    if (h->d.mode != SOFI_MODE_FWI || !h->model_set) return fail("sofi_fwi_gradient: no FWI model set");
    const Grid &g = h->g;
    dim3 tb(BX, BY), tg((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY);
    hipLaunchKernelGGL(k_fwi_grad, tg, tb, 0, h->stream, g, h->f, h->w, h->d.idx > 0 ? h->d.idx : 1,
                       h->d.idy > 0 ? h->d.idy : 1, h->sel);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t pl = (size_t)g.plane;
    float *out[3] = { grad_vp, grad_vs, grad_rho };
    for (int k = 0; k < 3; ++k)
        if (out[k]) HIPCHK(hipMemcpy(out[k], h->w.gshot + k * pl, pl * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int sofi_fwi_gradient_sum(sofi_dev *h, float *grad_vp, float *grad_vs, float *grad_rho, int reset)
{
    if (!h) return fail("sofi_fwi_gradient_sum: null handle");
    if (h->d.mode != SOFI_MODE_FWI) return fail("sofi_fwi_gradient_sum: the handle was not opened with MODE=FWI");
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t pl = (size_t)h->g.plane;
    float *out[3] = { grad_vp, grad_vs, grad_rho };
    for (int k = 0; k < 3; ++k)
        if (out[k]) HIPCHK(hipMemcpy(out[k], h->w.gsum + k * pl, pl * sizeof(float), hipMemcpyDeviceToHost));
    if (reset) {
        HIPCHK(hipMemsetAsync(h->w.gsum, 0, 3 * pl * sizeof(float), h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return 0;
}

int sofi_set_launch(sofi_dev *h, int velocity_shape, int stress_shape)
{
    if (!h) return fail("sofi_set_launch: null handle");
    for (int sh : { velocity_shape, stress_shape })
        if (sh < SOFI_LAUNCH_AUTO || sh > SOFI_LAUNCH_VEC4) return fail("sofi_set_launch: bad shape %d", sh);
    h->shape_v = velocity_shape;
    h->shape_s = stress_shape;
    h->gen++;
    return 0;
}

int sofi_set_graph(sofi_dev *h, int mode, int steps)
{
    if (!h) return fail("sofi_set_graph: null handle");
    if (mode < SOFI_GRAPH_AUTO || mode > SOFI_GRAPH_ON) return fail("sofi_set_graph: bad mode %d", mode);
    if (steps < 0 || steps > 4096) return fail("sofi_set_graph: %d steps per block (0..4096)", steps);
    h->graph_mode = mode;
    h->graph_steps = steps;
    return 0;
}

long long sofi_graph_blocks(const sofi_dev *h) { return h ? h->graph_blocks : 0; }

int sofi_set_batch_resident(sofi_dev *h, int slots)
{
    if (!h) return fail("sofi_set_batch_resident: null handle");
    if (slots < 0 || slots > h->nb) return fail("sofi_set_batch_resident: %d slots (0..%d)", slots, h->nb);
    h->res_mode = slots;
    h->spt_tuned = false;
    if (slots > 0) h->tune_on = false;
    h->gen++;
    return 0;
}

int sofi_get_batch_resident(const sofi_dev *h) { return h ? resident_slots(h) : 0; }

int sofi_set_batch_streams(sofi_dev *h, int streams)
{
    if (!h) return fail("sofi_set_batch_streams: null handle");
    if (streams < 0 || streams > 8) return fail("sofi_set_batch_streams: %d streams (0..8)", streams);
    h->str_mode = streams;
    if (streams > 0) h->tune_on = false;
    return 0;
}

int sofi_set_batch_tuning(sofi_dev *h, int on)
{
    if (!h) return fail("sofi_set_batch_tuning: null handle");
    h->tune_on = on != 0;
    h->tuned = false;
    return 0;
}

int sofi_get_batch_config(const sofi_dev *h, int *cfg)
{
    if (!h || !cfg) return fail("sofi_get_batch_config: null argument");
    cfg[0] = h->bs;
    cfg[1] = h->nb > 1 ? resident_slots(h) : 1;
    cfg[2] = h->nb > 1 ? batch_streams(h, (h->nb + cfg[1] - 1) / cfg[1]) : 1;
    cfg[3] = h->grp > 0 ? h->grp : h->nb;
    cfg[4] = h->bs > 0 ? h->bs : h->g.spt;
    cfg[5] = h->tuned ? 1 : 0;
    return 0;
}

int sofi_set_kernel_timing(sofi_dev *h, int every, int max_samples)
{
    if (!h) return fail("sofi_set_kernel_timing: null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    for (hipEvent_t e : h->tev) (void)hipEventDestroy(e);
    h->tev.clear();
    h->tev_used = 0;
    h->t_vel_ms = h->t_str_ms = 0.0;
    h->t_samples = 0;
    h->time_every = every > 0 ? every : 0;
    if (h->time_every) {
        int n = 4 * (max_samples > 0 ? max_samples : 1);
        h->tev.resize(n);
        for (int k = 0; k < n; ++k) HIPCHK(hipEventCreate(&h->tev[k]));
    }
    return 0;
}

int sofi_get_kernel_timing(const sofi_dev *h, double *velocity_ms, double *stress_ms, int *samples)
{
    if (!h) return fail("sofi_get_kernel_timing: null handle");
    int n = h->t_samples;
    if (velocity_ms) *velocity_ms = n ? h->t_vel_ms / n : 0.0;
    if (stress_ms) *stress_ms = n ? h->t_str_ms / n : 0.0;
    if (samples) *samples = n;
    return 0;
}

} // extern "C"
