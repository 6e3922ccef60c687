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

/* Cell tree of a TreeMesh, as Trixi's save_mesh_file writes it to mesh.h5
 * (Trixi.jl src/meshes/mesh_io.jl, TreeMesh method; read back by Trixi's
 * load_mesh and by Trixi2Vtk to place the solution files' elements).  The
 * reference's TreeMesh is Trixi's own (src/solvers/cache.jl:130-158), so the
 * mesh file of a HIP run is the one Trixi would write for the same mesh. */
typedef struct {
    int32_t ndims;             /* 1, 2 or 3 */
    int32_t periodicity[3];    /* per dimension, 0 or 1 */
    int64_t n_cells;           /* cells of the tree, leaves and parents */
    int64_t capacity;          /* n_cells_max of the TreeMesh */
    int64_t n_leaf_cells;
    int64_t minimum_level;     /* over the leaf cells */
    int64_t maximum_level;
    double center_level_0[3];  /* centre of the root cell */
    double length_level_0;     /* edge length of the root cell */
} trixi_h5_mesh_info;

/* Writes (truncating) `filename`: root attributes mesh_type ("TreeMesh"),
 * ndims, n_cells, capacity, n_leaf_cells, minimum_level, maximum_level,
 * center_level_0, length_level_0, periodicity and the int64 datasets
 * parent_ids [n_cells], child_ids [n_cells][2^ndims], neighbor_ids
 * [n_cells][2 ndims], levels [n_cells] and the float64 dataset coordinates
 * [n_cells][ndims] (row-major, the transposes of Trixi's column-major
 * arrays, as HDF5.jl stores a Julia matrix).  Cell ids are 1-based, 0 for
 * none. */
int trixi_h5_save_mesh(const char* filename, const trixi_h5_mesh_info* info, const int64_t* parent_ids,
                       const int64_t* child_ids, const int64_t* neighbor_ids, const int64_t* levels,
                       const double* coordinates);

const char* trixi_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TRIXI_H5_H */
