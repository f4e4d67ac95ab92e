/* Internal helpers of libsofi_host (not part of the public interface). */
#ifndef SOFI_HOST_INTERNAL_H
#define SOFI_HOST_INTERNAL_H

#include <math.h>
#include <stdlib.h>
#include "sofi_host.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int sofi_fail(const char *fmt, ...);

/* cell (i, j) of a padded-layout array, reference indexing */
#define AT(arr, i, j) ((arr)[sofi_idx(pitch_, (i), (j))])

static inline float *sofi_alloc_plane(int nx, int ny, int nplanes)
{
    return (float *)calloc(sofi_plane(nx, ny) * (size_t)(nplanes > 0 ? nplanes : 1), sizeof(float));
}

static inline int sofi_iround(double x) { return (int)floor(x + 0.5); }

#endif
