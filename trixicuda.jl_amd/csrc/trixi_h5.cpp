// Host-side HDF5 writer of Trixi's solution files (libtrixih5.so).
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

}  // extern "C"
