/*
 * sofi_types.h - plain C types shared by the host driver, the HIP library
 * (libsofi_hip.so) and the CPU oracle.
 *
 * Everything here is plain C: ints, floats, doubles and pointers.  No HIP or
 * torch type crosses the boundary.
 *
 * Grid layout ("padded layout")
 * -----------------------------
 * The reference stores every 2-D field as a float matrix indexed
 * [j][i] with j in [1-ND, NY+ND] and i in [1-ND, NX+ND]; the interior is
 * 1..NX x 1..NY and the ghost cells stay zero at the global edges
 * (src/initmem_wavefield.c:42-51, src/util.c:231-262).
 * Here every 2-D field is one row-major float array of
 *     rows  = NY + 2*SOFI_PAD          (SOFI_PAD = 8 >= ND_max = 6)
 *     pitch = roundup(NX + 2*SOFI_PAD, 64)  floats (256-byte rows)
 * and cell (i, j) (1-based, reference indexing) lives at
 *     (j - 1 + SOFI_PAD) * pitch + (i - 1 + SOFI_PAD).
 * i = 1 therefore sits at a 32-byte aligned column, so groups of four
 * interior cells starting at i = 1, 5, 9, ... are float4-aligned.
 * Memory-variable and defect "tensors" (the reference's f3tensor [j][i][l],
 * src/util.c:363-408) are stored as L consecutive planes, plane l-1 for
 * mechanism l (structure-of-arrays).
 */
#ifndef SOFI_TYPES_H
#define SOFI_TYPES_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SOFI_PAD 8
#define SOFI_MAXL 5
#define SOFI_MAXND 6

/* Wave-equation codes, same numbering as the reference (src/enums.h). */
enum sofi_weq {
    SOFI_AC_ISO = 0,
    SOFI_VAC_ISO = 1,
    SOFI_EL_ISO = 2,
    SOFI_VEL_ISO = 3,
    SOFI_EL_VTI = 4,
    SOFI_VEL_VTI = 5,
    SOFI_EL_TTI = 6,
    SOFI_VEL_TTI = 7
};

/* Run modes, same numbering as the reference (src/enums.h). */
enum sofi_mode { SOFI_MODE_FWI = 0, SOFI_MODE_FW = 1 };

static inline int sofi_pitch(int nx) { return ((nx + 2 * SOFI_PAD + 63) / 64) * 64; }
static inline int sofi_rows(int ny) { return ny + 2 * SOFI_PAD; }
static inline size_t sofi_plane(int nx, int ny) { return (size_t)sofi_pitch(nx) * (size_t)sofi_rows(ny); }
/* linear index of reference cell (i, j), 1-based */
static inline size_t sofi_idx(int pitch, int i, int j)
{
    return (size_t)(j - 1 + SOFI_PAD) * (size_t)pitch + (size_t)(i - 1 + SOFI_PAD);
}

/*
 * Static description of one shot's time loop: grid, scheme, boundaries,
 * acquisition sizes and output schedule.  It carries the scalars the
 * reference keeps in GlobVar (src/globvar_struct.h:40-203) that the time
 * loop reads (src/time_loop.c:27-326).
 */
typedef struct sofi_desc {
    int nx, ny;             /* NX, NY: grid points (one device holds the whole grid) */
    int fdorder;            /* FDORDER in {2,4,6,8,10,12} */
    int weq;                /* enum sofi_weq */
    int L;                  /* relaxation mechanisms (0 for elastic rheologies) */
    int nt;                 /* NT: time steps */
    float dt, dh;           /* DT [s], DH [m] */
    float hc[7];            /* Holberg coefficients hc[1..ND]; hc[0] = points per wavelength */
    int fw;                 /* FW: width of the absorbing frame */
    int abs_type;           /* 1 = CPML, 2 = exponential damping, other = none */
    int free_surf;          /* 1 = free surface at j = 1 */
    int gx[5], gy[5];       /* subgrid bounds GX[1..4], GY[1..4] (src/subgrid_bounds.c) */
    int source_type;        /* SOURCE_TYPE that selects psource/eqsource after the first step */
    int seismo;             /* SEISMO: 0 none, 1 vx/vy, 2 p, 3 div/curl, 4 all */
    int ndt;                /* NDT: seismogram sampling interval in steps */
    int ns;                 /* NS: samples per trace */
    int snap;               /* SNAP: 0 none, 1 vx/vy, 2 p, 3 div/curl, 4 all */
    int snap_idx[5];        /* SNAPIDX[1..4]: decimated snapshot window */
    int idx, idy;           /* snapshot decimation */
    int lsnap, isnap, esnap;/* first snapshot step, increment, last step */
    int mode;               /* enum sofi_mode */
    /* FWI (forward store / adjoint) */
    int dtinv;              /* DTINV: store interval of the forward wavefield */
    int velocity;           /* VELOCITY switch of the adjoint (src/wavefield_update_v.c:37-44) */
    int adjoint_type;       /* ADJOINT_TYPE (src/update_v_interior.c:72-90) */
    int fdorder_time;       /* FDORDER_TIME: 2 (leapfrog) or 4 (Adams-Bashforth, src/time_loop.c:78-92,215-245); 0 = 2 */
    int boundary;           /* BOUNDARY: 1 = periodic in x (exchange_v / exchange_s wrap the grid, src/exchange_v.c:94-148) */
} sofi_desc;

/*
 * Identifiers of the per-cell material arrays and 1-D profiles a caller
 * hands to the time loop.  2-D arrays are in the padded layout above;
 * "xL" arrays hold L planes.  Which ones a rheology needs is listed in
 * sofi_medium_required().
 */
enum sofi_array {
    SOFI_A_RIP = 0,    /* 1/rho at i+1/2 (src/av_rho.c) */
    SOFI_A_RJP,        /* 1/rho at j+1/2 */
    SOFI_A_PPI,        /* P modulus (relaxed for visco); pc11 for VTI/TTI */
    SOFI_A_PU,         /* S modulus (relaxed for visco); pc55 for VTI/TTI */
    SOFI_A_PUIPJP,     /* harmonic average of PU at (i+1/2, j+1/2) */
    SOFI_A_PTAUS,      /* tau_s (visco iso) */
    SOFI_A_PTAUP,      /* tau_p (visco iso) */
    SOFI_A_PTAUSIPJP,  /* tau_s averaged at (i+1/2, j+1/2) */
    SOFI_A_F,          /* mu * DT (*(1+L tau_s)) */
    SOFI_A_G,          /* pi * DT (*(1+L tau_p)) */
    SOFI_A_FIPJP,      /* mu_ipjp * DT (*(1+L tau_s_ipjp)) */
    SOFI_A_DIP,        /* xL: puipjp*eta_l*tausipjp */
    SOFI_A_D,          /* xL: pu*eta_l*taus */
    SOFI_A_E,          /* xL: ppi*eta_l*taup */
    SOFI_A_C11,        /* VTI/TTI stiffness (DT-scaled for elastic) */
    SOFI_A_C13,
    SOFI_A_C33,
    SOFI_A_C15,
    SOFI_A_C35,
    SOFI_A_C55IPJP,
    SOFI_A_C15IPJP,
    SOFI_A_C35IPJP,
    SOFI_A_C11U,       /* visco VTI/TTI unrelaxed moduli * DT */
    SOFI_A_C13U,
    SOFI_A_C33U,
    SOFI_A_C15U,
    SOFI_A_C35U,
    SOFI_A_C55IPJPU,
    SOFI_A_C15IPJPU,
    SOFI_A_C35IPJPU,
    SOFI_A_C11D,       /* xL: visco VTI/TTI modulus defects */
    SOFI_A_C13D,
    SOFI_A_C33D,
    SOFI_A_C15D,
    SOFI_A_C35D,
    SOFI_A_C55IPJPD,
    SOFI_A_C15IPJPD,
    SOFI_A_C35IPJPD,
    SOFI_A_ABSORB,     /* absorb_coeff (ABS_TYPE = 2) */
    SOFI_A_NARRAYS
};

/* 1-D profiles: CPML (index 1..2FW, stored 0..2FW) and visco scalars (1..L, stored 0..L). */
enum sofi_profile {
    SOFI_P_K_X = 0, SOFI_P_A_X, SOFI_P_B_X,
    SOFI_P_K_X_HALF, SOFI_P_A_X_HALF, SOFI_P_B_X_HALF,
    SOFI_P_K_Y, SOFI_P_A_Y, SOFI_P_B_Y,
    SOFI_P_K_Y_HALF, SOFI_P_A_Y_HALF, SOFI_P_B_Y_HALF,
    SOFI_P_ETA, SOFI_P_BIP, SOFI_P_CIP,
    SOFI_P_NPROFILES
};

/* Output components, numbering of the reference's section types. */
enum sofi_comp { SOFI_C_VX = 0, SOFI_C_VY, SOFI_C_P, SOFI_C_DIV, SOFI_C_CURL, SOFI_NCOMP };

#ifdef __cplusplus
}
#endif
#endif /* SOFI_TYPES_H */
