/* grim_io.h — grim's primitive dumps and restart files (HDF5), host side.
 *
 * grim writes its grids with PETSc's HDF5 viewer (grid::dump, grid::load,
 * src/grid/grid.cpp:367-381, 482-499): one dataset holding the bulk of the
 * DMDA global vector in natural order — dims (N3, N2, N1, dof) for dim 3,
 * (N2, N1, dof) for dim 2, (N1, dof) for dim 1, the variable index fastest,
 * IEEE FP64.  The torus writes "primitives" datasets to primVarsT%06d.h5
 * (torus/problem.cpp:778-840) and, for a restart, to restartVarsT%08d.h5
 * named in restartFileName.txt with the time in restartFileTime.txt
 * (:880-906); the timeStepper constructor reads them back
 * (src/timestepper/timestepper.cpp:387-441).
 *
 * libgrim_io.so loads the HDF5 C library at run time (GRIM_HDF5_LIB, else
 * libhdf5.so on the loader path, else /opt/conda/lib/libhdf5.so): it has no
 * link-time dependency and no device code.  Arrays are host memory in the
 * layout of include/grim_hip.h (dof planes of grim_grid_zones(c) doubles,
 * ghost zones included, X1 fastest).  Status: 0 = OK, negative = error
 * (grim_io_last_error() has the message).
 */
#ifndef GRIM_IO_H_
#define GRIM_IO_H_

#include "grim_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* grid::dump (grid.cpp:367-381): the bulk zones of dof planes as dataset
 * `dataset` of a new file `file` (an existing file is replaced). */
int grim_io_dump(const grim_grid_config* c, int dof, const double* prim, const char* dataset,
                 const char* file);

/* grid::load (grid.cpp:482-499): the dataset back into the bulk zones of dof
 * planes; ghost zones are left as they are.  The dataset's shape must be the
 * grid's. */
int grim_io_load(const grim_grid_config* c, int dof, double* prim, const char* dataset,
                 const char* file);

/* grid::dumpVTS (grid.cpp:383-480) of the dump variables: a VTK XML
 * StructuredGrid over the N1 x N2 x N3 bulk zone centres, written by this
 * library (no PETSc, no HDF5): vars holds nvars planes of N1*N2*N3 doubles
 * (X1 fastest) followed by the three planes x1, x2, x3 of the points — the
 * layout grim_dump_vars produces, with nvars = dof + 2.  Each plane becomes
 * a Float64 point-data array named names[v]; raw appended data with UInt64
 * size headers, little-endian. */
int grim_io_dump_vts(const grim_grid_config* c, int nvars, const char* const* names, const double* vars,
                     const char* file);

const char* grim_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* GRIM_IO_H_ */
