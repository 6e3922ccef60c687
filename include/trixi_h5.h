/* trixi_h5.h -- host-side writer of Trixi's HDF5 solution files (libtrixih5.so).
 *
 * Replaces the file output of Trixi's SaveSolutionCallback for a DG solver
 * (Trixi.jl src/callbacks_step/save_solution_dg.jl, `save_solution_file`),
 * which the reference reaches through its `wrap_array_native` hook
 * (src/solvers/dg.jl:24-32) and uses in its examples with
 * `solution_variables = cons2prim` (examples/euler_ec_1d.jl:37-39,
 * examples/eulermulti_ec_2d.jl:40-43, examples/hypdiff_nonperiodic_2d.jl:50-53).
 *
 * The state arrives on the host already unpacked to Trixi's layout and
 * converted to the solution variables: libtrixihip.so's trixi_hip_unpack
 * gives u[v, i, j, k, e] (column-major); variable v's slice, in that order,
 * is dataset variables_<v+1>.  Separate from libtrixihip.so so that the
 * compute library has no HDF5 dependency.  Every function returns a status
 * (0 = OK) and sets a thread-local message read by trixi_h5_last_error.
 */
#ifndef TRIXI_H5_H
#define TRIXI_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { TRIXI_H5_OK = 0, TRIXI_H5_EINVAL = 1, TRIXI_H5_EIO = 2 };

/* Root attributes of a solution file (save_solution_dg.jl: ndims, equations,
 * polydeg, n_vars, n_elements, mesh_type, mesh_file, time, dt, timestep). */
typedef struct {
    int32_t ndims;          /* 1, 2 or 3 */
    int32_t polydeg;        /* nodes per direction - 1 */
    int32_t n_vars;         /* datasets variables_1 .. variables_<n_vars> */
    int32_t n_element_vars; /* datasets element_variables_1 .. */
    int64_t n_elements;
    int64_t timestep;
    double time;
    double dt;
    const char* equations;  /* Trixi's equation type name, e.g. "CompressibleEulerEquations2D" */
    const char* mesh_type;  /* "TreeMesh" */
    const char* mesh_file;  /* file name of the mesh file, "" when none is written */
} trixi_h5_solution_info;

/* Writes (truncating) `filename`.  data: n_vars slices of
 * (polydeg+1)^ndims * n_elements float64 each, node index fastest;
 * varnames: n_vars UTF-8 names (attribute "name" of each dataset);
 * element_data: n_element_vars slices of n_elements float64 (may be NULL when
 * n_element_vars is 0), element_varnames their names. */
int trixi_h5_save_solution(const char* filename, const trixi_h5_solution_info* info, const double* data,
                           const char* const* varnames, const double* element_data,
                           const char* const* element_varnames);

const char* trixi_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TRIXI_H5_H */
