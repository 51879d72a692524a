// pk_gdf.cpp -- GDF-v2 data files and bin/gdf_distance (host code,
// libpiernik_gdf.so, linked with the HDF5 C library).
//
// Writer: the layout of write_to_hdf5_v2 / h5_write_to_single_file_v2
// (src/IO/hdf5/common_hdf5.F90:936-1232, data_hdf5.F90:696-763) with the
// attribute and dataset types of helpers_hdf5.F90 and set_get_attributes.F90,
// so that PIERNIK's own readers (bin/gdf_distance, the GDF frontend of yt)
// read it as a PIERNIK data file.  Field values come from pk_datafield
// (libpiernik_hip, computed on the device) or from host arrays.
//
// Comparator: bin/gdf_distance's measures (is_comparable, base_compare,
// compare_dlists with vector detection, compare_data per grid level and the
// total 1 - prod(1 - norm)) restated over the HDF5 C API.
#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "../../include/piernik_gdf.h"
#include "../../include/piernik_hip.h"

namespace {

std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// An HDF5 handle closed when it leaves scope.
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid() = default;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
    ~Hid()
    {
        if (id >= 0 && close) close(id);
    }
    bool ok() const { return id >= 0; }
};

#define H5TRY(expr, what)                                                        \
    do {                                                                         \
        if ((expr) < 0) return fail(PK_EIO, "HDF5: %s failed", what);           \
    } while (0)

// ---------------------------------------------------------------- writing --

// create_real_attribute_dim1 / set_attr_R: 1-D FP64 attribute
int attr_f64(hid_t loc, const char *name, const double *v, hsize_t n)
{
    Hid sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
    if (!sp.ok()) return fail(PK_EIO, "HDF5: dataspace of %s", name);
    Hid a(H5Acreate2(loc, name, H5T_NATIVE_DOUBLE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return fail(PK_EIO, "HDF5: attribute %s", name);
    H5TRY(H5Awrite(a.id, H5T_NATIVE_DOUBLE, v), name);
    return PK_OK;
}

// create_int_attribute / set_attr_I: 1-D native int attribute
int attr_i32(hid_t loc, const char *name, const int32_t *v, hsize_t n)
{
    Hid sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
    if (!sp.ok()) return fail(PK_EIO, "HDF5: dataspace of %s", name);
    Hid a(H5Acreate2(loc, name, H5T_NATIVE_INT32, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return fail(PK_EIO, "HDF5: attribute %s", name);
    H5TRY(H5Awrite(a.id, H5T_NATIVE_INT32, v), name);
    return PK_OK;
}

// create_int8_attribute: 1-D little-endian int64 attribute
int attr_i64(hid_t loc, const char *name, const int64_t *v, hsize_t n)
{
    Hid sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
    if (!sp.ok()) return fail(PK_EIO, "HDF5: dataspace of %s", name);
    Hid a(H5Acreate2(loc, name, H5T_STD_I64LE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return fail(PK_EIO, "HDF5: attribute %s", name);
    H5TRY(H5Awrite(a.id, H5T_NATIVE_INT64, v), name);
    return PK_OK;
}

// create_str_attribute: a scalar C string of a Fortran character(len=len)
// value, blank-padded to len, stored with its terminating NUL
int attr_str_padded(hid_t loc, const char *name, const std::string &s, size_t len)
{
    std::string v = s.substr(0, len);
    v.resize(len, ' ');
    Hid t(H5Tcopy(H5T_C_S1), H5Tclose);
    H5TRY(H5Tset_size(t.id, len + 1), name);
    Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
    Hid a(H5Acreate2(loc, name, t.id, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return fail(PK_EIO, "HDF5: attribute %s", name);
    H5TRY(H5Awrite(a.id, t.id, v.c_str()), name);
    return PK_OK;
}

// set_attr_C: a one-element array of a blank-padded Fortran string of the
// value's length
int attr_str_fortran(hid_t loc, const char *name, const std::string &s)
{
    const size_t len = std::max<size_t>(s.size(), 1);
    std::string v = s;
    v.resize(len, ' ');
    Hid t(H5Tcopy(H5T_FORTRAN_S1), H5Tclose);
    H5TRY(H5Tset_size(t.id, len), name);
    const hsize_t one = 1;
    Hid sp(H5Screate_simple(1, &one, nullptr), H5Sclose);
    Hid a(H5Acreate2(loc, name, t.id, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return fail(PK_EIO, "HDF5: attribute %s", name);
    H5TRY(H5Awrite(a.id, t.id, v.data()), name);
    return PK_OK;
}

// h5ltset_attribute_string_f: a scalar NUL-terminated C string
int attr_cstr(hid_t loc, const char *name, const char *s)
{
    Hid t(H5Tcopy(H5T_C_S1), H5Tclose);
    H5TRY(H5Tset_size(t.id, strlen(s) + 1), name);
    Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
    Hid a(H5Acreate2(loc, name, t.id, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return fail(PK_EIO, "HDF5: attribute %s", name);
    H5TRY(H5Awrite(a.id, t.id, s), name);
    return PK_OK;
}

// create_dataset_*: an integer dataset of C dims (the Fortran shape reversed)
int dset_int(hid_t loc, const char *name, hid_t ftype, hid_t mtype, const void *v, int rank, const hsize_t *dims)
{
    Hid sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    Hid d(H5Dcreate2(loc, name, ftype, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return fail(PK_EIO, "HDF5: dataset %s", name);
    H5TRY(H5Dwrite(d.id, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, v), name);
    return PK_OK;
}

// create_dataset_real_scalar plus its "unit" attribute (create_units_description)
int dset_unit(hid_t loc, const char *name, double val, const std::string &unit, size_t unit_len)
{
    Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
    Hid d(H5Dcreate2(loc, name, H5T_NATIVE_DOUBLE, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return fail(PK_EIO, "HDF5: dataset %s", name);
    H5TRY(H5Dwrite(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &val), name);
    return attr_str_padded(d.id, "unit", unit, unit_len);
}

// gdf_translate with gdf_strict (data_hdf5.F90:162-262), for the fields the
// library produces and the names it passes through.
std::string translate(const std::string &v)
{
    const size_t n = v.size();
    auto fl = [&](size_t i) { return n > i && (v[i] == 'd' || v[i] == 'n' || v[i] == 'i'); };
    if (n == 4 && v.compare(0, 3, "den") == 0 && fl(3)) return "density";
    if (n == 4 && v.compare(0, 2, "vl") == 0 && v[2] >= 'x' && v[2] <= 'z' && fl(3)) return std::string("velocity_") + v[2];
    if (n == 5 && v.compare(0, 3, "mom") == 0 && v[3] >= 'x' && v[3] <= 'z' && fl(4))
        return std::string("momentum_") + v[3];
    if (n == 4 && v.compare(0, 3, "ene") == 0 && (v[3] == 'n' || v[3] == 'i')) return "energy_density";
    if (n == 4 && v.compare(0, 3, "eth") == 0 && (v[3] == 'n' || v[3] == 'i')) return "specific_energy";
    if (n == 4 && v.compare(0, 3, "pre") == 0 && (v[3] == 'n' || v[3] == 'i')) return "pressure";
    if (n == 4 && v.compare(0, 3, "tem") == 0 && (v[3] == 'n' || v[3] == 'i')) return "temperature";
    if (n == 4 && v.compare(0, 3, "mag") == 0 && v[3] >= 'x' && v[3] <= 'z') return std::string("mag_field_") + v[3];
    if (v == "magB") return "magnetic_field_magnitude";
    if (v == "v") return "total_velocity";
    return v;
}

// datafields_descr (data_hdf5.F90:74-160): cgs factor, staggering, unit label
struct Descr {
    double f2cgs = 1.0;
    long long stag = 0;
    std::string fu = "fixme";
};

Descr descr(const std::string &v, const pk_gdf_meta &m)
{
    const double cm = m.cm, gram = m.gram, sek = m.sek;
    const double erg = gram * (cm * cm) / (sek * sek);  // units.F90:370
    const double fpi = 4.0 * 3.14159265358979323846;
    Descr d;
    const size_t n = v.size();
    auto fl = [&](size_t i) { return n > i && (v[i] == 'd' || v[i] == 'n' || v[i] == 'i'); };
    if (n == 4 && v.compare(0, 3, "den") == 0 && fl(3)) {
        d.fu = "\\rm{g}/\\rm{cm}^3";
        d.f2cgs = 1.0 / (gram / (cm * cm * cm));
    } else if ((n == 4 && v.compare(0, 2, "vl") == 0 && v[2] >= 'x' && v[2] <= 'z' && fl(3)) || v == "v" ||
               v == "c_s" || v == "cs") {
        d.fu = "\\rm{cm}/\\rm{s}";
        d.f2cgs = 1.0 / (cm / sek);
    } else if (v == "enen" || v == "enei") {
        d.fu = "\\rm{erg}/\\rm{cm}^3";
        d.f2cgs = 1.0 / (erg / (cm * cm * cm));
    } else if (v == "ethn" || v == "ethi") {
        d.fu = "\\rm{erg}/\\rm{g}";
        d.f2cgs = 1.0 / (erg / gram);
    } else if (v == "pren" || v == "prei") {
        d.fu = "\\rm{g}/\\rm{cm}/\\rm{s}^2";
        d.f2cgs = 1.0 / (gram / cm / (sek * sek));
    } else if (v == "temn" || v == "temi") {
        d.fu = "\\rm{K}";
        d.f2cgs = 1.0;
    } else if (v == "magx" || v == "magy" || v == "magz" || v == "magB") {
        d.fu = "\\rm{Gs}";
        d.f2cgs = 1.0 / (fpi * std::sqrt(cm / (m.miu0 * gram)) * sek);
        d.stag = 1;
    }
    return d;
}

std::string trimmed(const char *s, size_t cap)
{
    std::string v(s, strnlen(s, cap));
    while (!v.empty() && v.back() == ' ') v.pop_back();
    return v;
}

// get_unit (units.F90:492-559) of a dataset name
void unit_of(const std::string &f, const pk_gdf_meta &m, double *val, std::string *s)
{
    enum { L = 0, M, T, V, B, TEMP, E };
    std::string nm[7];
    for (int i = 0; i < 7; i++) nm[i] = trimmed(m.s_lmtvB[i], sizeof m.s_lmtvB[i]);
    const double *u = m.lmtvB;
    const bool complex_e = nm[E] == "complex";
    if (f == "dend" || f == "deni" || f == "denn" || f == "density") {
        *val = u[M] / (u[L] * u[L] * u[L]);
        *s = nm[M] + "/" + nm[L] + "**3";
    } else if ((f.size() == 4 && f.compare(0, 2, "vl") == 0) || f == "velocity_x" || f == "velocity_y" ||
               f == "velocity_z") {
        *val = u[V];
        *s = nm[V];
    } else if (f == "enen" || f == "enei" || f == "energy_density") {
        *val = u[M] / u[L] / (u[T] * u[T]);
        *s = complex_e ? nm[M] + "/" + nm[L] + "/" + nm[T] + "**2" : nm[E] + "/" + nm[L] + "**3";
    } else if (f == "ethn" || f == "ethi" || f == "specific_energy") {
        *val = (u[L] * u[L]) / (u[T] * u[T]);
        *s = complex_e ? nm[L] + "**2 /" + nm[T] + "**2" : nm[E] + "/" + nm[M];
    } else if (f == "pren" || f == "prei" || f == "pressure") {
        *val = u[M] / u[L] / (u[T] * u[T]);
        *s = nm[M] + "/" + nm[L] + "/" + nm[T] + "**2";
    } else if (f == "temn" || f == "temi" || f == "temperature") {
        *val = 1.0;
        *s = nm[TEMP];
    } else if (f == "magx" || f == "magy" || f == "magz" || f == "mag_field_x" || f == "mag_field_y" ||
               f == "mag_field_z") {
        *val = u[B];
        *s = nm[B];
    } else {
        *val = 1.0;
        *s = "dimensionless";
    }
}

const size_t kCbuff = 32, kUnits = 5 * 32, kFmax = 60;  // constants.F90:138-140, grid_data_format.F90

// Field values of block c, field f into buf (n_b cells, x fastest): FP64,
// or FP32 when fp32 is set and the source can round on its side (*rounded
// tells which it wrote).
typedef int (*FieldSource)(const void *ctx, int c, int f, const pk_gdf_cg &cg, const char *name, void *buf,
                           bool fp32, bool *rounded);

int src_device(const void *, int, int, const pk_gdf_cg &cg, const char *name, void *buf, bool fp32, bool *rounded)
{
    // FP32 datasets: rounded on the device, half the bytes over PCIe and no
    // conversion in HDF5 (the same round-to-nearest)
    const int rc = fp32 ? pk_datafield_f32(cg.cg_id, name, (float *)buf) : pk_datafield(cg.cg_id, name, (double *)buf);
    if (rc != PK_OK) return fail(rc, "pk_datafield(%lld, %s): %s", (long long)cg.cg_id, name, pk_last_error());
    *rounded = fp32;
    return PK_OK;
}

struct HostData {
    const double *const *data;
    int nfields;
};

int src_host(const void *ctx, int c, int f, const pk_gdf_cg &cg, const char *, void *buf, bool, bool *rounded)
{
    const HostData *h = (const HostData *)ctx;
    const double *p = h->data[(size_t)c * h->nfields + f];
    if (!p) return fail(PK_EINVAL, "no data for block %d field %d", c, f);
    memcpy(buf, p, sizeof(double) * (size_t)cg.n_b[0] * cg.n_b[1] * cg.n_b[2]);
    *rounded = false;
    return PK_OK;
}

int write_file(const char *path, const pk_gdf_meta *mp, int32_t ncg, const pk_gdf_cg *cgs, int32_t nfields,
               const char *const *fields, FieldSource src, const void *ctx)
{
    if (!path || !mp || ncg < 1 || !cgs || nfields < 0 || (nfields > 0 && !fields))
        return fail(PK_EINVAL, "bad arguments");
    const pk_gdf_meta &m = *mp;
    if (m.nfluids < 1 || m.nfluids > 2) return fail(PK_EINVAL, "nfluids must be 1 or 2");
    std::vector<std::string> names(nfields), dnames(nfields);
    std::set<std::string> seen;
    for (int f = 0; f < nfields; f++) {
        if (!fields[f] || !*fields[f]) return fail(PK_EINVAL, "empty field name");
        names[f] = fields[f];
        dnames[f] = m.gdf_strict ? translate(names[f]) : names[f];
        if (!seen.insert(dnames[f]).second)
            return fail(PK_EINVAL, "two fields are written as dataset '%s' (the reference's gdf_strict names "
                                   "do not tell the fluids apart; set gdf_strict = 0)", dnames[f].c_str());
    }
    for (int c = 0; c < ncg; c++)
        for (int d = 0; d < 3; d++)
            if (cgs[c].n_b[d] < 1) return fail(PK_EINVAL, "block %d has no cells along %d", c, d);

    H5open();
    Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return fail(PK_EIO, "cannot create %s", path);
    int rc;
    const hid_t fid = file.id;

    // set_common_attributes_v2 (common_hdf5.F90:771-813)
    const double piernik_v2 = 2.03;  // piernik_hdf5_version2, dataio_pub.F90:47
    const double plt = -99999.99999;
    const int32_t minus1 = -1;
    if ((rc = attr_f64(fid, "time", &m.time, 1)) || (rc = attr_f64(fid, "timestep", &m.timestep, 1)) ||
        (rc = attr_f64(fid, "piernik", &piernik_v2, 1)) || (rc = attr_f64(fid, "last_log_time", &m.last_log_time, 1)) ||
        (rc = attr_f64(fid, "last_tsl_time", &m.last_tsl_time, 1)) ||
        (rc = attr_f64(fid, "last_hdf_time", &m.last_hdf_time, 1)) ||
        (rc = attr_f64(fid, "last_res_time", &m.last_res_time, 1)) || (rc = attr_f64(fid, "last_plt_time", &plt, 1)) ||
        (rc = attr_f64(fid, "c_all_old", &m.c_all_old, 1)) || (rc = attr_f64(fid, "stepcfl", &m.stepcfl, 1)) ||
        (rc = attr_f64(fid, "cfl_c", &m.cfl_c, 1)) || (rc = attr_f64(fid, "magic_mass", m.magic_mass, m.nfluids)) ||
        (rc = attr_i32(fid, "nstep", &m.nstep, 1)) || (rc = attr_i32(fid, "nres", &m.nres, 1)) ||
        (rc = attr_i32(fid, "nhdf", &m.nhdf, 1)) || (rc = attr_i32(fid, "nimg", &minus1, 1)) ||
        (rc = attr_i32(fid, "require_problem_IC", &m.require_problem_IC, 1)) ||
        (rc = attr_str_fortran(fid, "problem_name", trimmed(m.problem_name, sizeof m.problem_name))) ||
        (rc = attr_str_fortran(fid, "domain", "phys_domain")) ||
        (rc = attr_str_fortran(fid, "run_id", trimmed(m.run_id, sizeof m.run_id))))
        return rc;

    // gdf_create_format_stamp
    {
        Hid g(H5Gcreate2(fid, "gridded_data_format", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (!g.ok()) return fail(PK_EIO, "HDF5: group gridded_data_format");
        if ((rc = attr_cstr(g.id, "data_software", "piernik")) ||
            (rc = attr_cstr(g.id, "data_software_version", "1.1")))
            return rc;
    }
    // gdf_create_simulation_parameters (cartesian, domain_dump = 'phys_domain')
    {
        Hid g(H5Gcreate2(fid, "simulation_parameters", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (!g.ok()) return fail(PK_EIO, "HDF5: group simulation_parameters");
        const int64_t refine_by = 2, dim = m.eff_dim, cosmo = 0, ghosts = 0;
        const int64_t dd[3] = {m.n_d[0], m.n_d[1], m.n_d[2]};
        const double le[3] = {m.edge[0][0], m.edge[1][0], m.edge[2][0]};
        const double re[3] = {m.edge[0][1], m.edge[1][1], m.edge[2][1]};
        const int32_t ordering = 1, geometry = 0, bc[6] = {0, 0, 0, 0, 0, 0};
        if ((rc = attr_i64(g.id, "refine_by", &refine_by, 1)) || (rc = attr_i64(g.id, "dimensionality", &dim, 1)) ||
            (rc = attr_i64(g.id, "cosmological_simulation", &cosmo, 1)) ||
            (rc = attr_i64(g.id, "num_ghost_zones", &ghosts, 1)) ||
            (rc = attr_i64(g.id, "domain_dimensions", dd, 3)) || (rc = attr_f64(g.id, "domain_left_edge", le, 3)) ||
            (rc = attr_f64(g.id, "domain_right_edge", re, 3)) || (rc = attr_f64(g.id, "current_time", &m.time, 1)) ||
            (rc = attr_i32(g.id, "field_ordering", &ordering, 1)) || (rc = attr_i32(g.id, "geometry", &geometry, 1)) ||
            (rc = attr_str_padded(g.id, "unique_identifier", "Piernik", 12)) ||
            (rc = attr_i32(g.id, "boundary_conditions", bc, 6)))
            return rc;
    }
    // /data with one group per block (write_to_hdf5_v2, :1084-1196)
    std::vector<double> buf;
    {
        Hid gl(H5Gcreate2(fid, "data", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (!gl.ok()) return fail(PK_EIO, "HDF5: group data");
        if ((rc = attr_i32(gl.id, "cg_count", &ncg, 1))) return rc;
        const hid_t ftype = m.h5_64bit ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT;
        for (int c = 0; c < ncg; c++) {
            const pk_gdf_cg &cg = cgs[c];
            char gname[32];
            snprintf(gname, sizeof gname, "grid_%010d", c);
            Hid g(H5Gcreate2(gl.id, gname, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
            if (!g.ok()) return fail(PK_EIO, "HDF5: group %s", gname);
            const double nb[3] = {(double)cg.n_b[0], (double)cg.n_b[1], (double)cg.n_b[2]};
            int32_t off[3];
            for (int d = 0; d < 3; d++) {
                if (cg.off[d] > 2147483647ll || cg.off[d] < -2147483647ll)
                    return fail(PK_EINVAL, "block %d: offset too large for the file's int32", c);
                off[d] = (int32_t)cg.off[d];
            }
            if ((rc = attr_i32(g.id, "level", &cg.level, 1)) || (rc = attr_f64(g.id, "n_b", nb, 3)) ||
                (rc = attr_i32(g.id, "off", off, 3)) || (rc = attr_f64(g.id, "left_edge", cg.left_edge, 3)) ||
                (rc = attr_f64(g.id, "right_edge", cg.right_edge, 3)) || (rc = attr_f64(g.id, "dl", cg.dl, 3)))
                return rc;
            const hsize_t dims[3] = {(hsize_t)cg.n_b[2], (hsize_t)cg.n_b[1], (hsize_t)cg.n_b[0]};
            buf.resize((size_t)cg.n_b[0] * cg.n_b[1] * cg.n_b[2]);
            for (int f = 0; f < nfields; f++) {
                bool rounded = false;
                if ((rc = src(ctx, c, f, cg, names[f].c_str(), buf.data(), !m.h5_64bit, &rounded))) return rc;
                Hid sp(H5Screate_simple(3, dims, nullptr), H5Sclose);
                Hid d(H5Dcreate2(g.id, dnames[f].c_str(), ftype, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                      H5Dclose);
                if (!d.ok()) return fail(PK_EIO, "HDF5: dataset %s of %s", dnames[f].c_str(), gname);
                H5TRY(H5Dwrite(d.id, rounded ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                               buf.data()),
                      "field write");
            }
        }
    }
    // gdf_create_root_datasets: Fortran (3, n) arrays are C (n, 3)
    {
        std::vector<int32_t> dims(3 * (size_t)ncg), level(ncg), part(ncg, 0);
        std::vector<int64_t> left(3 * (size_t)ncg), parent(ncg, -1);
        for (int c = 0; c < ncg; c++) {
            for (int d = 0; d < 3; d++) {
                dims[3 * c + d] = cgs[c].n_b[d];
                left[3 * c + d] = cgs[c].off[d];
            }
            level[c] = cgs[c].level;
        }
        const hsize_t d2[2] = {(hsize_t)ncg, 3}, d1[1] = {(hsize_t)ncg}, dp[2] = {(hsize_t)ncg, 1};
        if ((rc = dset_int(fid, "grid_dimensions", H5T_STD_I32LE, H5T_NATIVE_INT32, dims.data(), 2, d2)) ||
            (rc = dset_int(fid, "grid_left_index", H5T_STD_I64LE, H5T_NATIVE_INT64, left.data(), 2, d2)) ||
            (rc = dset_int(fid, "grid_level", H5T_STD_I32LE, H5T_NATIVE_INT32, level.data(), 1, d1)) ||
            (rc = dset_int(fid, "grid_parent_id", H5T_STD_I64LE, H5T_NATIVE_INT64, parent.data(), 1, d1)) ||
            (rc = dset_int(fid, "grid_particle_count", H5T_STD_I32LE, H5T_NATIVE_INT32, part.data(), 2, dp)))
            return rc;
    }
    // describe_domains (:1201-1219)
    {
        Hid gl(H5Gcreate2(fid, "domains", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (!gl.ok()) return fail(PK_EIO, "HDF5: group domains");
        {
            Hid g(H5Gcreate2(gl.id, "base", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
            if (!g.ok()) return fail(PK_EIO, "HDF5: group domains/base");
            if ((rc = attr_i32(g.id, "n_d", m.n_d, 3))) return rc;
            const char pref[3] = {'x', 'y', 'z'};
            for (int d = 0; d < 3; d++) {
                char nm[32];
                snprintf(nm, sizeof nm, "%c-edge_position", pref[d]);
                if ((rc = attr_f64(g.id, nm, m.edge[d], 2))) return rc;
                snprintf(nm, sizeof nm, "%c-boundary_type", pref[d]);
                if ((rc = attr_i32(g.id, nm, m.bnd[d], 2))) return rc;
            }
        }
        const int32_t zero = 0;
        if ((rc = attr_i32(gl.id, "fine_count", &zero, 1))) return rc;
    }
    // field_types (create_datafields_descrs) and dataset_units (create_units_description)
    {
        Hid g(H5Gcreate2(fid, "field_types", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (!g.ok()) return fail(PK_EIO, "HDF5: group field_types");
        for (int f = 0; f < nfields; f++) {
            const Descr d = descr(names[f], m);
            Hid h(H5Gcreate2(g.id, dnames[f].c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
            if (!h.ok()) return fail(PK_EIO, "HDF5: group field_types/%s", dnames[f].c_str());
            const int64_t stag = d.stag;
            if ((rc = attr_f64(h.id, "field_to_cgs", &d.f2cgs, 1)) || (rc = attr_i64(h.id, "staggering", &stag, 1)) ||
                (rc = attr_str_padded(h.id, "field_units", d.fu, kFmax)) ||
                (rc = attr_str_padded(h.id, "field_name", names[f], kFmax)))
                return rc;
        }
    }
    {
        Hid g(H5Gcreate2(fid, "dataset_units", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (!g.ok()) return fail(PK_EIO, "HDF5: group dataset_units");
        const char *base[5] = {"length_unit", "mass_unit", "time_unit", "velocity_unit", "magnetic_unit"};
        for (int i = 0; i < 5; i++)
            if ((rc = dset_unit(g.id, base[i], m.lmtvB[i], trimmed(m.s_lmtvB[i], sizeof m.s_lmtvB[i]), kCbuff)))
                return rc;
        for (int f = 0; f < nfields; f++) {
            double val;
            std::string su;
            unit_of(dnames[f], m, &val, &su);
            if ((rc = dset_unit(g.id, dnames[f].c_str(), val, su, kUnits))) return rc;
        }
    }
    H5TRY(H5Fflush(fid, H5F_SCOPE_LOCAL), "flush");
    return PK_OK;
}

// ---------------------------------------------------------------- reading --

int read_attr_f64(hid_t loc, const char *obj, const char *name, std::vector<double> *v)
{
    if (H5Aexists_by_name(loc, obj, name, H5P_DEFAULT) <= 0) return fail(PK_EIO, "no attribute %s of %s", name, obj);
    Hid a(H5Aopen_by_name(loc, obj, name, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return fail(PK_EIO, "cannot open attribute %s of %s", name, obj);
    Hid sp(H5Aget_space(a.id), H5Sclose);
    const hssize_t n = H5Sget_simple_extent_npoints(sp.id);
    if (n < 1) return fail(PK_EIO, "empty attribute %s of %s", name, obj);
    v->resize((size_t)n);
    H5TRY(H5Aread(a.id, H5T_NATIVE_DOUBLE, v->data()), name);
    return PK_OK;
}

int read_dset_f64(hid_t loc, const char *name, std::vector<double> *v)
{
    Hid d(H5Dopen2(loc, name, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return fail(PK_EIO, "no dataset %s", name);
    Hid sp(H5Dget_space(d.id), H5Sclose);
    const hssize_t n = H5Sget_simple_extent_npoints(sp.id);
    v->resize((size_t)std::max<hssize_t>(n, 0));
    if (n > 0) H5TRY(H5Dread(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v->data()), name);
    return PK_OK;
}

int member_names(hid_t loc, const char *group, std::vector<std::string> *out)
{
    Hid g(H5Gopen2(loc, group, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return fail(PK_EIO, "no group %s", group);
    H5G_info_t info;
    H5TRY(H5Gget_info(g.id, &info), "group info");
    out->clear();
    for (hsize_t i = 0; i < info.nlinks; i++) {
        const ssize_t len = H5Lget_name_by_idx(g.id, ".", H5_INDEX_NAME, H5_ITER_INC, i, nullptr, 0, H5P_DEFAULT);
        if (len < 0) return fail(PK_EIO, "member of %s", group);
        std::string s((size_t)len + 1, '\0');
        H5Lget_name_by_idx(g.id, ".", H5_INDEX_NAME, H5_ITER_INC, i, &s[0], (size_t)len + 1, H5P_DEFAULT);
        s.resize((size_t)len);
        out->push_back(s);
    }
    return PK_OK;
}

struct File {
    Hid f;
    std::string path;
};

int open_file(const char *path, File *F)
{
    H5open();
    F->path = path;
    F->f.id = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    F->f.close = H5Fclose;
    if (!F->f.ok()) return fail(PK_EIO, "cannot open %s", path);
    return PK_OK;
}

// is_comparable (bin/gdf_distance:23-82)
int comparable(File &a, File &b)
{
    std::vector<double> v;
    double ver[2], dim[2], geo[2];
    File *fs[2] = {&a, &b};
    for (int i = 0; i < 2; i++) {
        const hid_t id = fs[i]->f.id;
        int rc;
        if ((rc = read_attr_f64(id, "/", "piernik", &v))) return rc;
        ver[i] = v[0];
        if ((int)ver[i] != 2)
            return fail(PK_EINVAL, "Only Piernik v2 HDF (GDF) files are supported, not %g.", ver[i]);
        if ((rc = read_attr_f64(id, "simulation_parameters", "dimensionality", &v))) return rc;
        dim[i] = v[0];
        if ((rc = read_attr_f64(id, "simulation_parameters", "geometry", &v))) return rc;
        geo[i] = v[0];
        std::vector<double> lev;
        if ((rc = read_dset_f64(id, "grid_level", &lev))) return rc;
        std::set<int> levl(lev.begin(), lev.end());
        if (!levl.count(0)) return fail(PK_EINVAL, "Cannot find base level in '%s'.", fs[i]->path.c_str());
        if ((int)levl.size() != *levl.rbegin() + 1) return fail(PK_EINVAL, "Missing levels in '%s'.", fs[i]->path.c_str());
        if (levl.size() > 1) {
            if ((rc = read_attr_f64(id, "simulation_parameters", "refine_by", &v))) return rc;
            if (v[0] != 2.0)
                return fail(PK_EINVAL, "Only refine factor of 2 is supported. Cannot interpret '%s'.", fs[i]->path.c_str());
        }
    }
    if ((int)ver[0] != (int)ver[1]) return fail(PK_EINVAL, "Cannot compare different major file format revisions");
    if (dim[0] != dim[1]) return fail(PK_EINVAL, "Cannot compare files with different dimensionality");
    if (geo[0] != geo[1]) return fail(PK_EINVAL, "Cannot compare files with different geometries yet");
    return PK_OK;
}

// compare_data (:227-244): |a - b| / (|a| + |b|), Frobenius norms
double compare_data(const std::vector<double> &x, const std::vector<double> &y)
{
    long double s1 = 0, s2 = 0, sd = 0;
    for (size_t i = 0; i < x.size(); i++) {
        s1 += (long double)x[i] * x[i];
        s2 += (long double)y[i] * y[i];
        const long double d = (long double)x[i] - y[i];
        sd += d * d;
    }
    const double n1 = (double)sqrtl(s1), n2 = (double)sqrtl(s2), nd = (double)sqrtl(sd);
    return (n1 + n2 != 0.0) ? nd / (n1 + n2) : 0.0;
}

// collect_dataset (:247-272): the fields of one vector on a whole grid level
// (components stacked, x fastest), from every block of that level
int collect(File &F, const std::vector<std::string> &fields, int level, std::vector<double> *out, int nd[3],
            bool *met)
{
    std::vector<double> v;
    int rc;
    if ((rc = read_attr_f64(F.f.id, "domains/base", "n_d", &v))) return rc;
    for (int d = 0; d < 3; d++) nd[d] = (int)v[d] * (1 << level);
    const size_t ncell = (size_t)nd[0] * nd[1] * nd[2];
    out->assign(fields.size() * ncell, 0.0);
    if ((rc = read_attr_f64(F.f.id, "data", "cg_count", &v))) return rc;
    const int ncg = (int)v[0];
    *met = false;
    for (int g = 0; g < ncg; g++) {
        char gname[48];
        snprintf(gname, sizeof gname, "data/grid_%010d", g);
        if ((rc = read_attr_f64(F.f.id, gname, "level", &v))) return rc;
        if ((int)v[0] != level) continue;
        *met = true;
        std::vector<double> off, nb;
        if ((rc = read_attr_f64(F.f.id, gname, "off", &off)) || (rc = read_attr_f64(F.f.id, gname, "n_b", &nb)))
            return rc;
        int o[3], n[3];
        for (int d = 0; d < 3; d++) {
            o[d] = (int)off[d];
            n[d] = (int)nb[d];
            if (o[d] < 0 || o[d] + n[d] > nd[d]) return fail(PK_EINVAL, "%s lies outside its level", gname);
        }
        Hid gid(H5Gopen2(F.f.id, gname, H5P_DEFAULT), H5Gclose);
        for (size_t i = 0; i < fields.size(); i++) {
            std::vector<double> ds;
            if ((rc = read_dset_f64(gid.id, fields[i].c_str(), &ds))) return rc;
            if (ds.size() != (size_t)n[0] * n[1] * n[2]) return fail(PK_EINVAL, "%s/%s has the wrong size", gname, fields[i].c_str());
            double *dst = out->data() + i * ncell;
            for (int z = 0; z < n[2]; z++)
                for (int y = 0; y < n[1]; y++)
                    for (int x = 0; x < n[0]; x++)
                        dst[(size_t)(o[0] + x) + (size_t)nd[0] * ((o[1] + y) + (size_t)nd[1] * (o[2] + z))] =
                            ds[(size_t)x + (size_t)n[0] * (y + (size_t)n[1] * z)];
        }
    }
    return PK_OK;
}

bool has_two_digits_after(const std::string &f, const std::string &prefix)
{
    for (size_t p = f.find(prefix); p != std::string::npos; p = f.find(prefix, p + 1)) {
        const size_t q = p + prefix.size();
        if (q + 2 <= f.size() && isdigit((unsigned char)f[q]) && isdigit((unsigned char)f[q + 1])) return true;
    }
    return false;
}

// compare_dlists (:124-212): common fields, grouped into vectors
double field_lists(File &a, File &b, std::vector<std::vector<std::string>> *vec, int *rc)
{
    std::vector<std::string> f1, f2;
    if ((*rc = member_names(a.f.id, "data/grid_0000000000", &f1)) ||
        (*rc = member_names(b.f.id, "data/grid_0000000000", &f2)))
        return 0.0;
    auto drop = [](std::vector<std::string> &v) { v.erase(std::remove(v.begin(), v.end(), "particles"), v.end()); };
    drop(f1);
    drop(f2);
    std::set<std::string> s1(f1.begin(), f1.end()), s2(f2.begin(), f2.end()), fc;
    for (const auto &x : s1)
        if (s2.count(x)) fc.insert(x);
    std::vector<std::string> v0;
    for (const auto &f : fc) {
        const size_t n = f.size();
        if ((n >= 1 && f[n - 1] == 'x') || (n >= 2 && f[n - 2] == 'x') || (n >= 2 && f.compare(n - 2, 2, "01") == 0))
            v0.push_back(f);
    }
    std::set<std::string> used;
    for (const auto &v : v0) {
        const size_t n = v.size();
        std::vector<std::string> vv;
        if (v[n - 1] == 'x') {
            vv.push_back(v);
            for (const auto &f : fc)
                for (char i : {'y', 'z'})
                    if (f == v.substr(0, n - 1) + i) vv.push_back(f);
        } else if (n >= 2 && v[n - 2] == 'x') {
            vv.push_back(v);
            for (const auto &f : fc)
                for (char i : {'y', 'z'})
                    if (f == v.substr(0, n - 2) + i + v[n - 1]) vv.push_back(f);
        } else {
            for (const auto &f : fc)
                if (has_two_digits_after(f, v.substr(0, n - 2))) vv.push_back(f);
        }
        if (vv.size() > 1) {
            vec->push_back(vv);
            for (const auto &x : vv) used.insert(x);
        }
    }
    for (const auto &f : fc)
        if (!used.count(f)) vec->push_back({f});
    const size_t n1 = s1.size(), n2 = s2.size(), nc = fc.size();
    *rc = PK_OK;
    if (n1 == nc && n2 == nc) return 0.0;
    return 1.0 - (double)nc / (double)(n1 + n2 - nc);
}

std::string vector_name(std::vector<std::string> f)
{
    if (f.size() == 1) return f[0];
    std::sort(f.begin(), f.end());
    const std::string &a = f.front(), &z = f.back();
    const size_t n = a.size();
    if (a[n - 1] == 'x') return a.substr(0, n - 1) + "[x.." + z.back() + "]";
    if (n >= 2 && a[n - 2] == 'x') return a.substr(0, n - 2) + "[x.." + z[z.size() - 2] + "]" + a[n - 1];
    if (n >= 2 && a.compare(n - 2, 2, "01") == 0) return a.substr(0, n - 2) + "[01.." + z.substr(z.size() - 2) + "]";
    return "[" + a + ".." + z + "]";
}

}  // namespace

extern "C" {

void pk_gdf_meta_default(pk_gdf_meta *m)
{
    if (!m) return;
    memset(m, 0, sizeof *m);
    m->nfluids = 1;
    m->eff_dim = 3;
    m->gdf_strict = 1;
    for (int d = 0; d < 3; d++) {
        m->n_d[d] = 1;
        m->edge[d][0] = 0.0;
        m->edge[d][1] = 1.0;
    }
    snprintf(m->problem_name, sizeof m->problem_name, "%s", "piernik_amd");
    snprintf(m->run_id, sizeof m->run_id, "%s", "aaa");
    m->cm = m->gram = m->sek = 1.0;
    m->miu0 = 4.0 * 3.14159265358979323846;
    for (int i = 0; i < 5; i++) m->lmtvB[i] = 1.0;
    const char *nm[7] = {"cm", "g", "s", "cm / s", "gauss", "K", "erg"};  // units_set = 'cgs'
    for (int i = 0; i < 7; i++) snprintf(m->s_lmtvB[i], sizeof m->s_lmtvB[i], "%s", nm[i]);
}

int pk_gdf_write(const char *path, const pk_gdf_meta *meta, int32_t ncg, const pk_gdf_cg *cgs, int32_t nfields,
                 const char *const *fields)
{
    return write_file(path, meta, ncg, cgs, nfields, fields, src_device, nullptr);
}

int pk_gdf_write_host(const char *path, const pk_gdf_meta *meta, int32_t ncg, const pk_gdf_cg *cgs, int32_t nfields,
                      const char *const *fields, const double *const *data)
{
    if (!data && nfields > 0) return fail(PK_EINVAL, "no data");
    HostData h{data, nfields};
    return write_file(path, meta, ncg, cgs, nfields, fields, src_host, &h);
}

int pk_gdf_translate(const char *field, char *out, int32_t out_len)
{
    if (!field || !out || out_len < 1) return fail(PK_EINVAL, "bad arguments");
    const std::string t = translate(field);
    if ((int32_t)t.size() >= out_len) return fail(PK_EINVAL, "output too short");
    memcpy(out, t.c_str(), t.size() + 1);
    return PK_OK;
}

int pk_gdf_read(const char *path, int32_t grid, const char *name, double *out, int64_t n)
{
    if (!path || !name || !out || grid < 0) return fail(PK_EINVAL, "bad arguments");
    File F;
    int rc;
    if ((rc = open_file(path, &F))) return rc;
    char gname[48];
    snprintf(gname, sizeof gname, "data/grid_%010d", grid);
    Hid g(H5Gopen2(F.f.id, gname, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return fail(PK_EIO, "no group %s in %s", gname, path);
    std::vector<double> v;
    if ((rc = read_dset_f64(g.id, name, &v))) return rc;
    if ((int64_t)v.size() != n) return fail(PK_EINVAL, "%s/%s holds %zu values, not %lld", gname, name, v.size(), (long long)n);
    memcpy(out, v.data(), sizeof(double) * v.size());
    return PK_OK;
}

int pk_gdf_distance(const char *path1, const char *path2, int32_t lev1, int32_t lev2, int32_t cap, double *norms,
                    int32_t *levels, char *names, int64_t names_len, double *total)
{
    if (!path1 || !path2 || cap < 0 || (cap > 0 && (!norms || !levels || !names)) || !total)
        return -fail(PK_EINVAL, "bad arguments");
    File a, b;
    int rc;
    if ((rc = open_file(path1, &a)) || (rc = open_file(path2, &b))) return -rc;
    if ((rc = comparable(a, b))) return -rc;

    // base_compare (:405-436)
    std::map<std::string, double> base;
    std::vector<double> v1, v2;
    if ((rc = read_attr_f64(a.f.id, "/", "time", &v1)) || (rc = read_attr_f64(b.f.id, "/", "time", &v2))) return -rc;
    {
        const double t1 = v1[0], t2 = v2[0];
        base["time"] = (t1 > 0.0 && t2 > 0.0) ? std::fabs(t1 - t2) / (t1 + t2) : (t1 == 0.0 ? 0.0 : 1.0);
    }
    std::vector<double> l1, r1, l2, r2;
    if ((rc = read_attr_f64(a.f.id, "simulation_parameters", "domain_left_edge", &l1)) ||
        (rc = read_attr_f64(a.f.id, "simulation_parameters", "domain_right_edge", &r1)) ||
        (rc = read_attr_f64(b.f.id, "simulation_parameters", "domain_left_edge", &l2)) ||
        (rc = read_attr_f64(b.f.id, "simulation_parameters", "domain_right_edge", &r2)))
        return -rc;
    {
        double vol1 = 1, vol2 = 1, volo = 1;
        bool disjoint = false;
        for (int d = 0; d < 3; d++) {
            const double lo = std::max(l1[d], l2[d]), hi = std::min(r1[d], r2[d]);
            if (hi - lo < 0.0) disjoint = true;
            vol1 *= r1[d] - l1[d];
            vol2 *= r2[d] - l2[d];
            volo *= hi - lo;
        }
        if (vol1 == 0.0 && vol2 == 0.0) return -fail(PK_EINVAL, "Both domains have no volume!");
        base["domains"] = disjoint ? 1.0 : 1.0 - volo / (vol1 + vol2 - volo);
    }
    if ((rc = read_attr_f64(a.f.id, "simulation_parameters", "domain_dimensions", &v1)) ||
        (rc = read_attr_f64(b.f.id, "simulation_parameters", "domain_dimensions", &v2)))
        return -rc;
    {
        double r = 0.0;
        for (size_t i = 0; i < v1.size(); i++) r = 1.0 - (1.0 - r) * (1.0 - std::fabs(v1[i] - v2[i]) / (v1[i] + v2[i]));
        base["resolution"] = r;
    }
    base["boundary types (not checked)"] = 0.0;
    std::vector<std::vector<std::string>> vec;
    base["datafield lists"] = field_lists(a, b, &vec, &rc);
    if (rc) return -rc;
    if (base["domains"] < 1.0) base["AMR (not checked)"] = 0.0;
    if (!(base["domains"] == 0.0 && base["resolution"] == 0.0))
        return -fail(PK_EINVAL, "Comparing domains of different size or resolution is not implemented yet");

    // compare_datafields_by_level over the requested levels
    std::vector<std::pair<int, int>> lv;
    if (lev1 < 0) {
        std::vector<double> g1, g2;
        if ((rc = read_dset_f64(a.f.id, "grid_level", &g1)) || (rc = read_dset_f64(b.f.id, "grid_level", &g2)))
            return -rc;
        const int m = (int)std::min(*std::max_element(g1.begin(), g1.end()), *std::max_element(g2.begin(), g2.end()));
        for (int l = 0; l <= m; l++) lv.push_back({l, l});
    } else {
        lv.push_back({lev1, lev2});
    }
    std::map<std::string, std::vector<double>> perlev;  // field name -> norm per level
    for (const auto &L : lv) {
        for (const auto &f : vec) {
            std::vector<double> d1, d2;
            int nd1[3], nd2[3];
            bool m1, m2;
            if ((rc = collect(a, f, L.first, &d1, nd1, &m1)) || (rc = collect(b, f, L.second, &d2, nd2, &m2))) return -rc;
            if (!m1) return -fail(PK_EINVAL, "Level %d not met in %s.", L.first, path1);
            if (!m2) return -fail(PK_EINVAL, "Level %d not met in %s.", L.second, path2);
            if (nd1[0] != nd2[0] || nd1[1] != nd2[1] || nd1[2] != nd2[2])
                return -fail(PK_EINVAL, "Arrays cannot be compared: level shapes differ");
            perlev["datafield `" + vector_name(f) + "'"].push_back(compare_data(d1, d2));
        }
    }
    // the total as the two-file mode prints it (:491-520), keys in sorted order
    double tot = 0.0;
    int n = 0;
    size_t used = 0;
    auto emit = [&](const std::string &name, int level, double val) -> int {
        tot += (1.0 - tot) * val;
        if (n < cap) {
            if (used + name.size() + 1 > (size_t)names_len) return fail(PK_EINVAL, "names buffer too short");
            memcpy(names + used, name.c_str(), name.size() + 1);
            used += name.size() + 1;
            norms[n] = val;
            levels[n] = level;
        }
        n++;
        return PK_OK;
    };
    for (const auto &kv : base)
        if ((rc = emit(kv.first, -1, kv.second))) return -rc;
    for (const auto &kv : perlev)
        for (size_t i = 0; i < kv.second.size(); i++)
            if ((rc = emit(kv.first, lv[i].first, kv.second[i]))) return -rc;
    *total = tot;
    return n;
}

const char *pk_gdf_last_error(void) { return g_err.c_str(); }

}  // extern "C"
