/*
 * bbfmm.h -- core value types of the C++ drop-in API.
 *
 * Same names and members as PBBFMM3D's include/bbfmm.h:29-39 (vector3) and
 * :104-107 (doft), so that caller code compiles unchanged.  The octree node
 * type and the BLAS/LAPACK prototypes of that header are not part of this
 * API: the tree lives on the GPU behind the C ABI.
 */
#ifndef PBBFMM3D_AMD_BBFMM_H
#define PBBFMM3D_AMD_BBFMM_H

#include <cstdlib>

/* 3-D point, FP64, array-of-structs (bbfmm.h:29-39). */
struct vector3 {
    vector3() : x(0), y(0), z(0) {}
    vector3(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
    double x;
    double y;
    double z;
};

/* Degrees of freedom per source / target point (bbfmm.h:104-107).  Only
 * scalar kernels (s = f = 1) are supported, as in every reference caller. */
struct doft {
    int s;
    int f;
};

/* Uniform random number in [xmin, xmax] from rand() (bbfmm.h:52-53), as the
 * reference examples use it to generate inputs. */
#define frand(xmin, xmax) \
    ((double)(xmin) + (double)((xmax) - (xmin)) * rand() / (double)RAND_MAX)

#endif
