/*
 * piernik_gdf.h -- GDF-v2 data files (PIERNIK's "v2" HDF5 output) and the
 * file comparison of bin/gdf_distance, in libpiernik_gdf.so.
 *
 * The library is separate from libpiernik_hip.so because it links the HDF5 C
 * library; the field values come from libpiernik_hip's pk_datafield (the
 * device state) or from host arrays.  The calls return PK_OK (0) or a PK_E*
 * code (piernik_hip.h); pk_gdf_last_error() gives the message.
 *
 * The file layout follows write_to_hdf5_v2 and h5_write_to_single_file_v2
 * (src/IO/hdf5/common_hdf5.F90:936-1232, data_hdf5.F90:696-763):
 *   /                      attributes of set_common_attributes_v2 (:771-813),
 *                          piernik = 2.03 (dataio_pub.F90:47)
 *   /gridded_data_format   data_software, data_software_version (grid_data_format.F90)
 *   /simulation_parameters gdf_create_simulation_parameters (cartesian, phys_domain)
 *   /grid_dimensions, /grid_left_index, /grid_level, /grid_parent_id,
 *   /grid_particle_count   gdf_create_root_datasets
 *   /data                  cg_count; /data/grid_%010d: level, n_b, off,
 *                          left_edge, right_edge, dl and one dataset per
 *                          field, named by gdf_translate (gdf_strict), of the
 *                          block's n_b cells, FP32 (or FP64 with h5_64bit)
 *   /domains               fine_count; /domains/base: n_d, [xyz]-edge_position,
 *                          [xyz]-boundary_type
 *   /field_types/<field>   field_to_cgs, staggering, field_units, field_name
 *                          (create_datafields_descrs, datafields_descr)
 *   /dataset_units         the five base units and one per field (create_units_description)
 * Not written: the problem.par and env text copies, particles.
 */
#ifndef PIERNIK_GDF_H
#define PIERNIK_GDF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One block of the file (a leaf grid container). */
typedef struct pk_gdf_cg {
    int64_t cg_id;          /* pk_gdf_write: a block registered with libpiernik_hip */
    int32_t level;          /* refinement level, 0 = base (cg%l%id) */
    int32_t n_b[3];         /* interior cells (cg%n_b) */
    int64_t off[3];         /* first interior cell on its level (cg%my_se(:, LO)) */
    double left_edge[3];    /* cg%fbnd(:, LO) */
    double right_edge[3];   /* cg%fbnd(:, HI) */
    double dl[3];           /* cg%dl */
} pk_gdf_cg;

/* The run: domain, time and bookkeeping (global, domain, dataio_pub, units). */
typedef struct pk_gdf_meta {
    double time, timestep;            /* t, dt */
    int32_t nstep, nres, nhdf;
    int32_t require_problem_IC;
    double last_log_time, last_tsl_time, last_hdf_time, last_res_time;
    double c_all_old, stepcfl, cfl_c;
    int32_t nfluids;
    double magic_mass[2];
    int32_t n_d[3];                   /* base-level cells (dom%n_d) */
    double edge[3][2];                /* dom%edge(:, LO:HI) */
    int32_t bnd[3][2];                /* dom%bnd(:, LO:HI) boundary types */
    int32_t eff_dim;                  /* dom%eff_dim */
    int32_t h5_64bit;                 /* 1: FP64 datasets; 0: FP32 (the reference's default) */
    int32_t gdf_strict;               /* 1: GDF dataset names (gdf_translate; the default), 0: short names */
    char problem_name[64];
    char run_id[16];
    /* units (units.F90): code units of one centimetre, gram, second; miu0;
     * the GDF base units lmtvB and their names (length, mass, time,
     * velocity, magnetic field, temperature, energy -- "complex" when energy
     * is derived) */
    double cm, gram, sek, miu0;
    double lmtvB[5];
    char s_lmtvB[7][16];
} pk_gdf_meta;

/* Defaults: CGS units (cm = gram = sek = 1, miu0 = 4 pi), time 0, one fluid. */
void pk_gdf_meta_default(pk_gdf_meta *meta);

/* Write a data file of ncg blocks registered with libpiernik_hip, their
 * fields computed on the device (pk_datafield); fields are the reference's
 * short names (hdf_vars, e.g. "deni", "vlxi", "prei", "magx", "psi"). */
int pk_gdf_write(const char *path, const pk_gdf_meta *meta, int32_t ncg, const pk_gdf_cg *cgs, int32_t nfields,
                 const char *const *fields);

/* The same from host arrays: data[c * nfields + f] holds field f of block c
 * in the reference's (nx, ny, nz) order (cg_id is not used). */
int pk_gdf_write_host(const char *path, const pk_gdf_meta *meta, int32_t ncg, const pk_gdf_cg *cgs, int32_t nfields,
                      const char *const *fields, const double *const *data);

/* gdf_translate with gdf_strict (data_hdf5.F90:162-262): the dataset name of
 * a short field name, into out (at least 64 bytes). */
int pk_gdf_translate(const char *field, char *out, int32_t out_len);

/* Read dataset `name` of block (group grid_%010d) `grid` into out (n values
 * as FP64, file order: x fastest). */
int pk_gdf_read(const char *path, int32_t grid, const char *name, double *out, int64_t n);

/* bin/gdf_distance for two files (piernik_gdf_compare with levels lev1,
 * lev2, or every common level when lev1 < 0).  Up to cap difference
 * measures, each in [0, 1]: their names NUL-separated into names
 * (names_len bytes), values into norms, grid levels into levels (-1 for the
 * measures of the whole file); *total is the tool's "Total difference"
 * (1 - prod(1 - norm)).  Returns the number of measures, or minus a PK_E*
 * code; files the tool would refuse (is_comparable, differing domains or
 * resolution) give -PK_EINVAL with the tool's reason. */
int pk_gdf_distance(const char *path1, const char *path2, int32_t lev1, int32_t lev2, int32_t cap, double *norms,
                    int32_t *levels, char *names, int64_t names_len, double *total);

const char *pk_gdf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
