// Host-side HDF5 writer of Trixi's solution and TreeMesh files (libtrixih5.so).
//
// Writes the file that Trixi's SaveSolutionCallback writes for a DG solver
// (Trixi.jl src/callbacks_step/save_solution_dg.jl, `save_solution_file`;
// used by the reference's examples, e.g. examples/euler_ec_1d.jl:37-39):
//   root attributes  ndims, equations, polydeg, n_vars, n_elements,
//                    mesh_type, mesh_file, time, dt, timestep
//   datasets         variables_<v> (v = 1..n_vars): 1-D float64 of
//                    nnodes^ndims * n_elements values, node index i fastest,
//                    then j, k, then the element; attribute "name"
//                    element_variables_<v>: 1-D float64 of n_elements;
//                    attribute "name"
// Integers are written as 64-bit, reals as float64 and strings as scalar
// fixed-length UTF-8 strings, as HDF5.jl writes Julia Int, Float64 and String.
//
// This library is separate from libtrixihip.so so that the compute library
// carries no HDF5 dependency; it links the HDF5 C library (1.10 API).

#include "../../include/trixi_h5.h"

#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <string>

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
    g_err = msg;
    return TRIXI_H5_EIO;
}

struct Handle {  // closes an HDF5 id with the given routine on scope exit
    hid_t id;
    herr_t (*close)(hid_t);
    ~Handle() {
        if (id >= 0) close(id);
    }
};

bool attr_write(hid_t obj, const char* name, hid_t mem_type, hid_t file_type, const void* value) {
    Handle space{H5Screate(H5S_SCALAR), H5Sclose};
    if (space.id < 0) return false;
    Handle attr{H5Acreate2(obj, name, file_type, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose};
    if (attr.id < 0) return false;
    return H5Awrite(attr.id, mem_type, value) >= 0;
}

bool attr_i64(hid_t obj, const char* name, int64_t v) {
    return attr_write(obj, name, H5T_NATIVE_INT64, H5T_STD_I64LE, &v);
}

bool attr_f64(hid_t obj, const char* name, double v) {
    return attr_write(obj, name, H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, &v);
}

bool attr_str(hid_t obj, const char* name, const char* s) {
    size_t n = std::strlen(s);
    Handle type{H5Tcopy(H5T_C_S1), H5Tclose};
    if (type.id < 0) return false;
    if (H5Tset_size(type.id, n > 0 ? n : 1) < 0 || H5Tset_cset(type.id, H5T_CSET_UTF8) < 0 ||
        H5Tset_strpad(type.id, H5T_STR_NULLTERM) < 0)
        return false;
    char empty = 0;
    return attr_write(obj, name, type.id, type.id, n > 0 ? s : &empty);
}

bool dataset_f64(hid_t file, const char* name, const double* data, int64_t n, const char* varname) {
    hsize_t dims[1] = {static_cast<hsize_t>(n)};
    Handle space{H5Screate_simple(1, dims, nullptr), H5Sclose};
    if (space.id < 0) return false;
    Handle ds{H5Dcreate2(file, name, H5T_IEEE_F64LE, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose};
    if (ds.id < 0) return false;
    if (n > 0 && H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) return false;
    return attr_str(ds.id, "name", varname);
}

bool attr_f64_vec(hid_t obj, const char* name, const double* v, hsize_t n) {
    Handle space{H5Screate_simple(1, &n, nullptr), H5Sclose};
    if (space.id < 0) return false;
    Handle attr{H5Acreate2(obj, name, H5T_IEEE_F64LE, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose};
    if (attr.id < 0) return false;
    return H5Awrite(attr.id, H5T_NATIVE_DOUBLE, v) >= 0;
}

// Julia Bool vectors: one byte per value (HDF5.jl's 8-bit Bool)
bool attr_bool_vec(hid_t obj, const char* name, const int32_t* v, hsize_t n) {
    unsigned char b[3] = {0, 0, 0};
    for (hsize_t i = 0; i < n; ++i) b[i] = v[i] ? 1 : 0;
    Handle space{H5Screate_simple(1, &n, nullptr), H5Sclose};
    if (space.id < 0) return false;
    Handle attr{H5Acreate2(obj, name, H5T_STD_B8LE, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose};
    if (attr.id < 0) return false;
    return H5Awrite(attr.id, H5T_NATIVE_B8, b) >= 0;
}

bool dataset_nd(hid_t file, const char* name, hid_t mem_type, hid_t file_type, const void* data, int rank,
                const hsize_t* dims) {
    Handle space{H5Screate_simple(rank, dims, nullptr), H5Sclose};
    if (space.id < 0) return false;
    Handle ds{H5Dcreate2(file, name, file_type, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose};
    if (ds.id < 0) return false;
    hsize_t total = 1;
    for (int r = 0; r < rank; ++r) total *= dims[r];
    return total == 0 || H5Dwrite(ds.id, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
}

}  // namespace

extern "C" {

const char* trixi_h5_last_error(void) { return g_err.c_str(); }

int trixi_h5_save_solution(const char* filename, const trixi_h5_solution_info* info, const double* data,
                           const char* const* varnames, const double* element_data,
                           const char* const* element_varnames) {
    if (!filename || !info || !info->equations || !info->mesh_type || !info->mesh_file) {
        g_err = "null filename or info string";
        return TRIXI_H5_EINVAL;
    }
    if (info->ndims < 1 || info->ndims > 3 || info->polydeg < 0 || info->n_vars < 1 || info->n_elements < 0 ||
        info->n_element_vars < 0) {
        g_err = "invalid ndims, polydeg, n_vars, n_elements or n_element_vars";
        return TRIXI_H5_EINVAL;
    }
    if ((!data && info->n_elements > 0) || !varnames || (info->n_element_vars > 0 && !element_varnames) ||
        (info->n_element_vars > 0 && info->n_elements > 0 && !element_data)) {
        g_err = "null data or name array";
        return TRIXI_H5_EINVAL;
    }
    for (int v = 0; v < info->n_vars; ++v)
        if (!varnames[v]) {
            g_err = "null variable name";
            return TRIXI_H5_EINVAL;
        }
    for (int v = 0; v < info->n_element_vars; ++v)
        if (!element_varnames[v]) {
            g_err = "null element variable name";
            return TRIXI_H5_EINVAL;
        }
    int64_t per_var = info->n_elements;
    for (int d = 0; d < info->ndims; ++d) per_var *= (info->polydeg + 1);

    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors come back as status codes, not printed
    Handle file{H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose};
    if (file.id < 0) return fail(std::string("cannot create ") + filename);
    bool ok = attr_i64(file.id, "ndims", info->ndims) && attr_str(file.id, "equations", info->equations) &&
              attr_i64(file.id, "polydeg", info->polydeg) && attr_i64(file.id, "n_vars", info->n_vars) &&
              attr_i64(file.id, "n_elements", info->n_elements) &&
              attr_str(file.id, "mesh_type", info->mesh_type) && attr_str(file.id, "mesh_file", info->mesh_file) &&
              attr_f64(file.id, "time", info->time) && attr_f64(file.id, "dt", info->dt) &&
              attr_i64(file.id, "timestep", info->timestep);
    if (!ok) return fail("writing the file attributes failed");
    char name[64];
    for (int v = 0; v < info->n_vars; ++v) {
        std::snprintf(name, sizeof name, "variables_%d", v + 1);
        if (!dataset_f64(file.id, name, data + v * per_var, per_var, varnames[v]))
            return fail(std::string("writing ") + name + " failed");
    }
    for (int v = 0; v < info->n_element_vars; ++v) {
        std::snprintf(name, sizeof name, "element_variables_%d", v + 1);
        if (!dataset_f64(file.id, name, element_data + v * info->n_elements, info->n_elements,
                         element_varnames[v]))
            return fail(std::string("writing ") + name + " failed");
    }
    if (H5Fflush(file.id, H5F_SCOPE_LOCAL) < 0) return fail("flushing the file failed");
    return TRIXI_H5_OK;
}

int trixi_h5_save_mesh(const char* filename, const trixi_h5_mesh_info* info, const int64_t* parent_ids,
                       const int64_t* child_ids, const int64_t* neighbor_ids, const int64_t* levels,
                       const double* coordinates) {
    if (!filename || !info) {
        g_err = "null filename or info";
        return TRIXI_H5_EINVAL;
    }
    const int nd = info->ndims;
    if (nd < 1 || nd > 3 || info->n_cells < 1 || info->capacity < info->n_cells || info->n_leaf_cells < 1 ||
        info->n_leaf_cells > info->n_cells || info->minimum_level < 0 || info->maximum_level < info->minimum_level) {
        g_err = "invalid ndims, cell counts or levels";
        return TRIXI_H5_EINVAL;
    }
    if (!parent_ids || !child_ids || !neighbor_ids || !levels || !coordinates) {
        g_err = "null tree array";
        return TRIXI_H5_EINVAL;
    }
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Handle file{H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose};
    if (file.id < 0) return fail(std::string("cannot create ") + filename);
    const hsize_t n = static_cast<hsize_t>(info->n_cells);
    bool ok = attr_str(file.id, "mesh_type", "TreeMesh") && attr_i64(file.id, "ndims", nd) &&
              attr_i64(file.id, "n_cells", info->n_cells) && attr_i64(file.id, "capacity", info->capacity) &&
              attr_i64(file.id, "n_leaf_cells", info->n_leaf_cells) &&
              attr_i64(file.id, "minimum_level", info->minimum_level) &&
              attr_i64(file.id, "maximum_level", info->maximum_level) &&
              attr_f64_vec(file.id, "center_level_0", info->center_level_0, nd) &&
              attr_f64(file.id, "length_level_0", info->length_level_0) &&
              attr_bool_vec(file.id, "periodicity", info->periodicity, nd);
    if (!ok) return fail("writing the mesh attributes failed");
    const hsize_t d1[1] = {n};
    const hsize_t dch[2] = {n, static_cast<hsize_t>(1) << nd};
    const hsize_t dnb[2] = {n, static_cast<hsize_t>(2 * nd)};
    const hsize_t dxy[2] = {n, static_cast<hsize_t>(nd)};
    ok = dataset_nd(file.id, "parent_ids", H5T_NATIVE_INT64, H5T_STD_I64LE, parent_ids, 1, d1) &&
         dataset_nd(file.id, "child_ids", H5T_NATIVE_INT64, H5T_STD_I64LE, child_ids, 2, dch) &&
         dataset_nd(file.id, "neighbor_ids", H5T_NATIVE_INT64, H5T_STD_I64LE, neighbor_ids, 2, dnb) &&
         dataset_nd(file.id, "levels", H5T_NATIVE_INT64, H5T_STD_I64LE, levels, 1, d1) &&
         dataset_nd(file.id, "coordinates", H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, coordinates, 2, dxy);
    if (!ok) return fail("writing the tree datasets failed");
    if (H5Fflush(file.id, H5F_SCOPE_LOCAL) < 0) return fail("flushing the file failed");
    return TRIXI_H5_OK;
}

}  // extern "C"
