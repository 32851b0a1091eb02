// Observable files in ReaDDy's HDF5 layout (libreaddy_h5.so), for writers that run off the
// simulation's critical path (the serving loop calls it from a writer thread).
//
// Layout read by readdy.Trajectory (wrappers/python/src/python/readdy/api/trajectory.py:420-480):
//   /readdy/observables/<name>/time      [T]          uint64 step index
//   /readdy/observables/<name>/data      [T][n_types] uint64 particles per type        (NParticles.cpp:83)
//   /readdy/observables/<name>/counts/<reaction id>   [T] uint64 events per interval  (ReactionCounts)
//   /readdy/observables/<name>/bin_centers [B], distribution [T][B] double, time [T]
//                                                   (RadialDistribution.cpp:158-170)
//   /readdy/observables/<name>/data [T][B] double, time [T]     (HistogramAlongAxis.cpp:83-96)
//   /readdy/observables/<name>/data [T] vlen of {x, y, z} double, time [T]    (Positions.cpp:71-86,
//                                                   Forces.cpp:80-91)
//   /readdy/observables/<name>/{types vlen u16, ids vlen u64, positions vlen {x,y,z}}, time [T]
//                                                   (Particles.cpp:62-92)
//   /readdy/trajectory/<name>/records [R] {id u64, typeId u16, flavor u8, pos f64[3]},
//                             limits [T][2] uint64 (a frame's [begin, end) in records), time [T]
//                                                   (FlatTrajectory, Trajectory.cpp:112-158; Types.cpp:79-88)
//   /readdy/observables/<name>/records [T] vlen of {reaction_id u64, reaction_type i32, educts u64[2],
//                             products u64[2], position f64[3], types_from u16[2]}, time [T]
//                                                   (Reactions.cpp:72-86; Types.cpp:66-77)
//   /readdy/config/particle_types        [n] {name str, flavor str, type_id u64, diffusion_constant f64}
//   /readdy/config/registered_reactions  [n] {name str, id u64, n_educts u64, n_products u64, rate,
//                                         educt_distance, product_distance f64, educt_types u16[2],
//                                         product_types u16[2]}
//   /readdy/config/general               JSON string {kbt, box_volume, box_size, pbc}
//                                                   (IOUtils.cpp:58-176, as readdy.Trajectory reads them)
// plus string attributes on /readdy/config describing the request.
//
// Every time-indexed dataset is chunked and unlimited along time, as the reference's observables create
// them with their flushStride as the chunk size (Observable.h:157-199, TimeSeriesWriter.h:64-66): a
// writer called again for the same observable appends its frames (a flat trajectory's limits continue
// from the records already written), so a recorder can flush every flushStride frames instead of
// holding the whole series in memory.
//
// C ABI (all return 0 on success, -1 on an HDF5 error):
//   int64_t rdy_h5_create(const char *path);                    file handle (> 0) or -1
//   int rdy_h5_set_chunk(int64_t f, uint64_t rows);   chunk rows (time points) of datasets created from
//                                                      now on in this file (the flushStride; default 64)
//   int rdy_h5_n_particles(int64_t f, const char *name, uint64_t T, uint32_t n_types,
//                          const uint64_t *time, const uint64_t *counts);
//   int rdy_h5_reaction_counts(int64_t f, const char *name, uint64_t T, uint32_t n_reactions,
//                              const uint64_t *time, const uint64_t *counts);   counts [T][n_reactions]
//   int rdy_h5_radial_distribution(int64_t f, const char *name, uint64_t T, uint32_t n_bins,
//                                  const uint64_t *time, const double *centers, const double *dist);
//   int rdy_h5_series_f64(int64_t f, const char *name, uint64_t T, uint32_t width, const uint64_t *time,
//                         const double *data);                    data [T][width]
//   int rdy_h5_series_f64_nd(int64_t f, const char *name, uint64_t T, uint32_t ndims, const uint32_t *dims,
//                            const uint64_t *time, const double *data);
//       data [T][dims[0]]..[dims[ndims-1]], ndims <= 3: Energy's [T] (ndims 0, Energy.cpp:69-71) and
//       Virial's [T][3][3] (Virial.cpp:76-78)
//   int rdy_h5_vlen_series(int64_t f, const char *name, const char *dataset, int elem, uint64_t T,
//                          const uint64_t *frame_sizes, const void *data, const uint64_t *time);
//       elem 0: {x, y, z} doubles, 1: uint16, 2: uint64; the frames' elements back to back in data;
//       time [T] is written beside the dataset when not null
//   int rdy_h5_u64_table(int64_t f, const char *name, const char *dataset, uint64_t rows, uint32_t cols,
//                        const uint64_t *data);        /readdy/observables/<name>/<dataset>, [rows] or [rows][cols]
//   int rdy_h5_flat_trajectory(int64_t f, const char *name, uint64_t T, const uint64_t *time,
//                              const uint64_t *frame_sizes, const uint64_t *ids, const uint16_t *types,
//                              const uint8_t *flavors, const double *pos);          pos [R][3]
//   int rdy_h5_reactions(int64_t f, const char *name, uint64_t T, const uint64_t *time,
//                        const uint64_t *frame_sizes, const uint64_t *reaction_id, const int32_t *type,
//                        const uint64_t *educts, const uint64_t *products, const double *position,
//                        const uint16_t *types_from);          educts, products, types_from [R][2]; position [R][3]
//   int rdy_h5_particle_types(int64_t f, uint32_t n, const char *const *names, const double *diffusion,
//                             const char *const *flavors);             type ids 0 .. n-1
//   int rdy_h5_registered_reactions(int64_t f, uint32_t n, const char *const *names, const uint64_t *ids,
//                                   const uint64_t *n_educts, const uint64_t *n_products, const double *rate,
//                                   const double *educt_distance, const double *product_distance,
//                                   const uint16_t *educt_types, const uint16_t *product_types);   [n][2]
//   int rdy_h5_general(int64_t f, const char *json);
//   int rdy_h5_attribute(int64_t f, const char *key, const char *value);
//   int rdy_h5_close(int64_t f);
//   int rdy_h5_read_u64(const char *path, const char *dataset, uint64_t cap, uint64_t *out, uint64_t *n);
//       reads a uint64 dataset by absolute path (checks and tools); *n = its element count
//   int rdy_h5_read_f64(const char *path, const char *dataset, uint64_t cap, double *out, uint64_t *n);
//   int rdy_h5_read_records(const char *path, const char *dataset, uint64_t cap, uint64_t *ids,
//                           uint16_t *types, uint8_t *flavors, double *pos, uint64_t *n);
//   int rdy_h5_read_reactions(const char *path, const char *dataset, uint64_t cap_frames, uint64_t *frame_sizes,
//                             uint64_t cap, uint64_t *reaction_id, int32_t *type, uint64_t *educts,
//                             uint64_t *products, double *position, uint16_t *types_from, uint64_t *n_frames,
//                             uint64_t *n);   *n_frames, *n = frames and records in the file
//   int rdy_h5_read_vlen(const char *path, const char *dataset, int elem, uint64_t cap_frames,
//                        uint64_t *frame_sizes, uint64_t cap, void *out, uint64_t *n_frames, uint64_t *n);
//   int rdy_h5_read_string(const char *path, const char *dataset, uint64_t cap, char *out);
//       a scalar string dataset into out (NUL-terminated; -1 when it needs more than cap bytes)
#include <hdf5.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace {

std::mutex g_chunk_mu;
std::map<hid_t, hsize_t> g_chunk;  // chunk rows per open file

hsize_t chunk_rows(hid_t g) {
    const hid_t f = H5Iget_file_id(g);
    std::lock_guard<std::mutex> lk(g_chunk_mu);
    auto it = g_chunk.find(f);
    const hsize_t r = it == g_chunk.end() ? 64 : it->second;
    H5Fclose(f);  // H5Iget_file_id took a reference
    return r;
}

// Appends dims[0] rows (each of shape dims[1 .. rank-1]) to the dataset `name` of group g, creating it
// chunked and unlimited along axis 0 when absent; *prev (if not null) = its row count before.
int append(hid_t g, const char *name, hid_t ftype, hid_t mtype, int rank, const hsize_t *dims, const void *data,
           hsize_t *prev = nullptr) {
    hsize_t cur[4] = {0, 0, 0, 0}, nd[4], start[4] = {0, 0, 0, 0};
    hid_t ds;
    if (H5Lexists(g, name, H5P_DEFAULT) > 0) {
        ds = H5Dopen2(g, name, H5P_DEFAULT);
        if (ds < 0) return -1;
        hid_t sp = H5Dget_space(ds);
        const int r = H5Sget_simple_extent_ndims(sp);
        if (r != rank) {
            H5Sclose(sp);
            H5Dclose(ds);
            return -1;
        }
        H5Sget_simple_extent_dims(sp, cur, nullptr);
        H5Sclose(sp);
        for (int a = 1; a < rank; ++a)
            if (cur[a] != dims[a]) {
                H5Dclose(ds);
                return -1;
            }
    } else {
        hsize_t zero[4], maxd[4], chunk[4];
        for (int a = 0; a < rank; ++a) {
            zero[a] = a == 0 ? 0 : dims[a];
            maxd[a] = a == 0 ? H5S_UNLIMITED : dims[a];
            chunk[a] = a == 0 ? std::max<hsize_t>(chunk_rows(g), 1) : std::max<hsize_t>(dims[a], 1);
        }
        hid_t sp = H5Screate_simple(rank, zero, maxd);
        hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
        H5Pset_chunk(dcpl, rank, chunk);
        ds = H5Dcreate2(g, name, ftype, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT);
        H5Pclose(dcpl);
        H5Sclose(sp);
        if (ds < 0) return -1;
        for (int a = 1; a < rank; ++a) cur[a] = dims[a];
    }
    if (prev) *prev = cur[0];
    int rc = 0;
    if (dims[0] > 0) {
        for (int a = 0; a < rank; ++a) nd[a] = a == 0 ? cur[0] + dims[0] : dims[a];
        rc = H5Dset_extent(ds, nd) < 0 ? -1 : 0;
        if (rc == 0) {
            hid_t fs = H5Dget_space(ds);
            start[0] = cur[0];
            H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, dims, nullptr);
            hid_t ms = H5Screate_simple(rank, dims, nullptr);
            rc = H5Dwrite(ds, mtype, ms, fs, H5P_DEFAULT, data) < 0 ? -1 : 0;
            H5Sclose(ms);
            H5Sclose(fs);
        }
    }
    H5Dclose(ds);
    return rc;
}

int append_u64(hid_t g, const char *name, int rank, const hsize_t *dims, const uint64_t *data, hsize_t *prev = nullptr) {
    return append(g, name, H5T_STD_U64LE, H5T_NATIVE_UINT64, rank, dims, data, prev);
}

int append_f64(hid_t g, const char *name, int rank, const hsize_t *dims, const double *data) {
    return append(g, name, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, rank, dims, data);
}

// failures are reported through the return codes: HDF5's own error-stack printing stays off
struct QuietErrors {
    QuietErrors() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
} quiet_errors;

hid_t ensure_group(hid_t parent, const char *name) {
    if (H5Lexists(parent, name, H5P_DEFAULT) > 0) return H5Gopen2(parent, name, H5P_DEFAULT);
    return H5Gcreate2(parent, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
}

// /readdy/observables/<name>, created as needed
hid_t observable_group(hid_t f, const char *name) {
    hid_t g0 = ensure_group(f, "readdy");
    if (g0 < 0) return -1;
    hid_t g1 = ensure_group(g0, "observables");
    H5Gclose(g0);
    if (g1 < 0) return -1;
    hid_t g2 = ensure_group(g1, name);
    H5Gclose(g1);
    return g2;
}

// one FlatTrajectory record in memory (the reference's TrajectoryEntry fields; TrajectoryEntry.h)
struct Record {
    uint64_t id;
    uint16_t typeId;
    uint8_t flavor;
    double pos[3];
};

// memory and file compound types of Record (the file type packed, little endian)
void record_types(hid_t *mem, hid_t *file) {
    const hsize_t three[1] = {3};
    hid_t am = H5Tarray_create2(H5T_NATIVE_DOUBLE, 1, three);
    *mem = H5Tcreate(H5T_COMPOUND, sizeof(Record));
    H5Tinsert(*mem, "id", HOFFSET(Record, id), H5T_NATIVE_UINT64);
    H5Tinsert(*mem, "typeId", HOFFSET(Record, typeId), H5T_NATIVE_UINT16);
    H5Tinsert(*mem, "flavor", HOFFSET(Record, flavor), H5T_NATIVE_UINT8);
    H5Tinsert(*mem, "pos", HOFFSET(Record, pos), am);
    H5Tclose(am);
    hid_t af = H5Tarray_create2(H5T_IEEE_F64LE, 1, three);
    *file = H5Tcreate(H5T_COMPOUND, 8 + 2 + 1 + 24);
    H5Tinsert(*file, "id", 0, H5T_STD_U64LE);
    H5Tinsert(*file, "typeId", 8, H5T_STD_U16LE);
    H5Tinsert(*file, "flavor", 10, H5T_STD_U8LE);
    H5Tinsert(*file, "pos", 11, af);
    H5Tclose(af);
}

// /readdy/config, created as needed
hid_t config_group(hid_t f) {
    hid_t g0 = ensure_group(f, "readdy");
    if (g0 < 0) return -1;
    hid_t g = ensure_group(g0, "config");
    H5Gclose(g0);
    return g;
}

// a variable-length C string type
hid_t vlen_string() {
    hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, H5T_VARIABLE);
    return t;
}

// writes n records of a compound type (used as memory and file type) as a 1-d dataset
int write_compound(hid_t g, const char *name, hid_t type, hsize_t n, const void *data) {
    hid_t space = H5Screate_simple(1, &n, nullptr);
    if (space < 0) return -1;
    hid_t ds = H5Dcreate2(g, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    int rc = ds < 0 ? -1 : (H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0 ? -1 : 0);
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(space);
    return rc;
}

// element type of a variable-length series: {x, y, z} doubles (Types.cpp:57-64), uint16, uint64
hid_t vlen_elem_type(int elem, size_t *size) {
    if (elem == 0) {
        hid_t t = H5Tcreate(H5T_COMPOUND, 3 * sizeof(double));
        H5Tinsert(t, "x", 0, H5T_NATIVE_DOUBLE);
        H5Tinsert(t, "y", sizeof(double), H5T_NATIVE_DOUBLE);
        H5Tinsert(t, "z", 2 * sizeof(double), H5T_NATIVE_DOUBLE);
        *size = 3 * sizeof(double);
        return t;
    }
    *size = elem == 1 ? 2 : 8;
    return H5Tcopy(elem == 1 ? H5T_NATIVE_UINT16 : H5T_NATIVE_UINT64);
}

// one reaction record in memory (the fields the reference's ReactionRecord type writes, Types.cpp:66-77)
struct RecordH5 {
    uint64_t reaction_id;
    int32_t reaction_type;
    uint64_t educts[2];
    uint64_t products[2];
    double position[3];
    uint16_t types_from[2];
};

hid_t reaction_record_type() {
    const hsize_t two[1] = {2}, three[1] = {3};
    hid_t u64x2 = H5Tarray_create2(H5T_NATIVE_UINT64, 1, two);
    hid_t f64x3 = H5Tarray_create2(H5T_NATIVE_DOUBLE, 1, three);
    hid_t u16x2 = H5Tarray_create2(H5T_NATIVE_UINT16, 1, two);
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(RecordH5));
    H5Tinsert(t, "reaction_id", HOFFSET(RecordH5, reaction_id), H5T_NATIVE_UINT64);
    H5Tinsert(t, "reaction_type", HOFFSET(RecordH5, reaction_type), H5T_NATIVE_INT32);
    H5Tinsert(t, "educts", HOFFSET(RecordH5, educts), u64x2);
    H5Tinsert(t, "products", HOFFSET(RecordH5, products), u64x2);
    H5Tinsert(t, "position", HOFFSET(RecordH5, position), f64x3);
    H5Tinsert(t, "types_from", HOFFSET(RecordH5, types_from), u16x2);
    H5Tclose(u64x2);
    H5Tclose(f64x3);
    H5Tclose(u16x2);
    return t;
}

struct TypeInfo {  // ParticleTypeInfo (IOUtils.h:71-76)
    const char *name;
    const char *flavor;
    uint64_t type_id;
    double diffusion_constant;
};

struct ReactionInfo {  // ReactionInfo (IOUtils.h:59-70)
    const char *name;
    uint64_t id, n_educts, n_products;
    double rate, educt_distance, product_distance;
    uint16_t educt_types[2], product_types[2];
};

int write_f64(hid_t g, const char *name, int rank, const hsize_t *dims, const double *data) {
    hid_t space = H5Screate_simple(rank, dims, nullptr);
    if (space < 0) return -1;
    hid_t ds = H5Dcreate2(g, name, H5T_IEEE_F64LE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    int rc = -1;
    if (ds >= 0) {
        rc = H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0 ? -1 : 0;
        H5Dclose(ds);
    }
    H5Sclose(space);
    return rc;
}

}  // namespace

extern "C" {

int64_t rdy_h5_create(const char *path) {
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    return f < 0 ? -1 : (int64_t)f;
}

int rdy_h5_set_chunk(int64_t f, uint64_t rows) {
    if (f < 0 || rows == 0) return -1;
    std::lock_guard<std::mutex> lk(g_chunk_mu);
    g_chunk[(hid_t)f] = (hsize_t)rows;
    return 0;
}

int rdy_h5_n_particles(int64_t f, const char *name, uint64_t T, uint32_t n_types, const uint64_t *time,
                       const uint64_t *counts) {
    hid_t g = observable_group((hid_t)f, name);
    if (g < 0) return -1;
    hsize_t dt[1] = {T}, dd[2] = {T, n_types};
    int rc = append_u64(g, "time", 1, dt, time);
    if (rc == 0) rc = append_u64(g, "data", 2, dd, counts);
    H5Gclose(g);
    return rc;
}

int rdy_h5_reaction_counts(int64_t f, const char *name, uint64_t T, uint32_t n_reactions, const uint64_t *time,
                           const uint64_t *counts) {
    hid_t g = observable_group((hid_t)f, name);
    if (g < 0) return -1;
    hsize_t dt[1] = {T};
    int rc = append_u64(g, "time", 1, dt, time);
    hid_t c = rc == 0 ? ensure_group(g, "counts") : -1;
    if (c < 0) rc = -1;
    std::vector<uint64_t> col(std::max<uint64_t>(T, 1));
    for (uint32_t r = 0; rc == 0 && r < n_reactions; ++r) {
        for (uint64_t t = 0; t < T; ++t) col[t] = counts[t * n_reactions + r];
        rc = append_u64(c, std::to_string(r).c_str(), 1, dt, col.data());
    }
    if (c >= 0) H5Gclose(c);
    H5Gclose(g);
    return rc;
}

int rdy_h5_radial_distribution(int64_t f, const char *name, uint64_t T, uint32_t n_bins, const uint64_t *time,
                               const double *centers, const double *dist) {
    hid_t g = observable_group((hid_t)f, name);
    if (g < 0) return -1;
    hsize_t dt[1] = {T}, dc[1] = {n_bins}, dd[2] = {T, n_bins};
    int rc = H5Lexists(g, "bin_centers", H5P_DEFAULT) > 0 ? 0 : write_f64(g, "bin_centers", 1, dc, centers);
    if (rc == 0) rc = append_f64(g, "distribution", 2, dd, dist);
    if (rc == 0) rc = append_u64(g, "time", 1, dt, time);
    H5Gclose(g);
    return rc;
}

int rdy_h5_u64_table(int64_t f, const char *name, const char *dataset, uint64_t rows, uint32_t cols,
                     const uint64_t *data) {
    hid_t g = observable_group((hid_t)f, name);
    if (g < 0) return -1;
    hsize_t d[2] = {rows, cols};
    const int rc = append_u64(g, dataset, cols > 1 ? 2 : 1, d, data);
    H5Gclose(g);
    return rc;
}

int rdy_h5_series_f64(int64_t f, const char *name, uint64_t T, uint32_t width, const uint64_t *time,
                      const double *data) {
    hid_t g = observable_group((hid_t)f, name);
    if (g < 0) return -1;
    hsize_t dt[1] = {T}, dd[2] = {T, width};
    int rc = append_f64(g, "data", 2, dd, data);
    if (rc == 0) rc = append_u64(g, "time", 1, dt, time);
    H5Gclose(g);
    return rc;
}

int rdy_h5_series_f64_nd(int64_t f, const char *name, uint64_t T, uint32_t ndims, const uint32_t *dims,
                         const uint64_t *time, const double *data) {
    if (ndims > 3 || (ndims && !dims)) return -1;
    hid_t g = observable_group((hid_t)f, name);
    if (g < 0) return -1;
    hsize_t dt[1] = {T}, dd[4] = {T, 0, 0, 0};
    for (uint32_t a = 0; a < ndims; ++a) dd[a + 1] = dims[a];
    int rc = append_f64(g, "data", 1 + (int)ndims, dd, data);
    if (rc == 0) rc = append_u64(g, "time", 1, dt, time);
    H5Gclose(g);
    return rc;
}

int rdy_h5_vlen_series(int64_t f, const char *name, const char *dataset, int elem, uint64_t T,
                       const uint64_t *frame_sizes, const void *data, const uint64_t *time) {
    if (elem < 0 || elem > 2) return -1;
    hid_t g = observable_group((hid_t)f, name);
    if (g < 0) return -1;
    size_t esz = 0;
    hid_t et = vlen_elem_type(elem, &esz);
    hid_t vt = H5Tvlen_create(et);
    std::vector<hvl_t> frames(T);
    const char *p = (const char *)data;
    for (uint64_t t = 0; t < T; ++t) {
        frames[t].len = frame_sizes[t];
        frames[t].p = frame_sizes[t] ? (void *)p : nullptr;
        p += frame_sizes[t] * esz;
    }
    hsize_t dt[1] = {T};
    int rc = append(g, dataset, vt, vt, 1, dt, frames.data());
    H5Tclose(vt);
    H5Tclose(et);
    if (rc == 0 && time) rc = append_u64(g, "time", 1, dt, time);
    H5Gclose(g);
    return rc;
}

int rdy_h5_flat_trajectory(int64_t f, const char *name, uint64_t T, const uint64_t *time, const uint64_t *frame_sizes,
                           const uint64_t *ids, const uint16_t *types, const uint8_t *flavors, const double *pos) {
    hid_t g0 = ensure_group((hid_t)f, "readdy");
    if (g0 < 0) return -1;
    hid_t g1 = ensure_group(g0, "trajectory");
    H5Gclose(g0);
    if (g1 < 0) return -1;
    // an empty name puts the datasets into /readdy/trajectory itself (Trajectory.cpp:115-116)
    hid_t g = g1;
    if (name[0]) {
        g = ensure_group(g1, name);
        H5Gclose(g1);
    }
    if (g < 0) return -1;
    uint64_t total = 0;
    std::vector<uint64_t> limits(2 * T);
    for (uint64_t t = 0; t < T; ++t) {
        limits[2 * t] = total;
        total += frame_sizes[t];
        limits[2 * t + 1] = total;
    }
    std::vector<Record> rec(total);
    for (uint64_t k = 0; k < total; ++k) {
        rec[k].id = ids[k];
        rec[k].typeId = types[k];
        rec[k].flavor = flavors[k];
        for (int a = 0; a < 3; ++a) rec[k].pos[a] = pos[3 * k + a];
    }
    hid_t mem, file;
    record_types(&mem, &file);
    hsize_t dr[1] = {total}, before = 0;
    int rc = append(g, "records", file, mem, 1, dr, rec.data(), &before);
    H5Tclose(mem);
    H5Tclose(file);
    for (auto &l : limits) l += before;  // frames appended after earlier ones continue their records
    hsize_t dl[2] = {T, 2}, dt[1] = {T};
    if (rc == 0) rc = append_u64(g, "limits", 2, dl, limits.data());
    if (rc == 0) rc = append_u64(g, "time", 1, dt, time);
    H5Gclose(g);
    return rc;
}

int rdy_h5_reactions(int64_t f, const char *name, uint64_t T, const uint64_t *time, const uint64_t *frame_sizes,
                     const uint64_t *reaction_id, const int32_t *type, const uint64_t *educts,
                     const uint64_t *products, const double *position, const uint16_t *types_from) {
    hid_t g = observable_group((hid_t)f, name);
    if (g < 0) return -1;
    uint64_t total = 0;
    for (uint64_t t = 0; t < T; ++t) total += frame_sizes[t];
    std::vector<RecordH5> rec(total);
    for (uint64_t k = 0; k < total; ++k) {
        RecordH5 &x = rec[k];
        x.reaction_id = reaction_id[k];
        x.reaction_type = type[k];
        for (int j = 0; j < 2; ++j) {
            x.educts[j] = educts[2 * k + j];
            x.products[j] = products[2 * k + j];
            x.types_from[j] = types_from[2 * k + j];
        }
        for (int d = 0; d < 3; ++d) x.position[d] = position[3 * k + d];
    }
    std::vector<hvl_t> frames(T);
    uint64_t at = 0;
    for (uint64_t t = 0; t < T; ++t) {
        frames[t].len = frame_sizes[t];
        frames[t].p = frame_sizes[t] ? (void *)(rec.data() + at) : nullptr;
        at += frame_sizes[t];
    }
    hid_t ct = reaction_record_type();
    hid_t vt = H5Tvlen_create(ct);
    hsize_t dt[1] = {T};
    int rc = append(g, "records", vt, vt, 1, dt, frames.data());
    H5Tclose(vt);
    H5Tclose(ct);
    if (rc == 0) rc = append_u64(g, "time", 1, dt, time);
    H5Gclose(g);
    return rc;
}

int rdy_h5_particle_types(int64_t f, uint32_t n, const char *const *names, const double *diffusion,
                          const char *const *flavors) {
    hid_t g = config_group((hid_t)f);
    if (g < 0) return -1;
    std::vector<TypeInfo> info(n);
    for (uint32_t k = 0; k < n; ++k) info[k] = {names[k], flavors[k], k, diffusion[k]};
    hid_t s = vlen_string();
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(TypeInfo));
    H5Tinsert(t, "name", HOFFSET(TypeInfo, name), s);
    H5Tinsert(t, "flavor", HOFFSET(TypeInfo, flavor), s);
    H5Tinsert(t, "type_id", HOFFSET(TypeInfo, type_id), H5T_NATIVE_UINT64);
    H5Tinsert(t, "diffusion_constant", HOFFSET(TypeInfo, diffusion_constant), H5T_NATIVE_DOUBLE);
    const int rc = write_compound(g, "particle_types", t, n, info.data());
    H5Tclose(t);
    H5Tclose(s);
    H5Gclose(g);
    return rc;
}

int rdy_h5_registered_reactions(int64_t f, uint32_t n, const char *const *names, const uint64_t *ids,
                                const uint64_t *n_educts, const uint64_t *n_products, const double *rate,
                                const double *educt_distance, const double *product_distance,
                                const uint16_t *educt_types, const uint16_t *product_types) {
    hid_t g = config_group((hid_t)f);
    if (g < 0) return -1;
    std::vector<ReactionInfo> info(n);
    for (uint32_t k = 0; k < n; ++k) {
        info[k] = {names[k], ids[k], n_educts[k], n_products[k], rate[k], educt_distance[k], product_distance[k],
                   {educt_types[2 * k], educt_types[2 * k + 1]}, {product_types[2 * k], product_types[2 * k + 1]}};
    }
    const hsize_t two[1] = {2};
    hid_t s = vlen_string();
    hid_t pair = H5Tarray_create2(H5T_NATIVE_UINT16, 1, two);
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(ReactionInfo));
    H5Tinsert(t, "name", HOFFSET(ReactionInfo, name), s);
    H5Tinsert(t, "id", HOFFSET(ReactionInfo, id), H5T_NATIVE_UINT64);
    H5Tinsert(t, "n_educts", HOFFSET(ReactionInfo, n_educts), H5T_NATIVE_UINT64);
    H5Tinsert(t, "n_products", HOFFSET(ReactionInfo, n_products), H5T_NATIVE_UINT64);
    H5Tinsert(t, "rate", HOFFSET(ReactionInfo, rate), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "educt_distance", HOFFSET(ReactionInfo, educt_distance), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "product_distance", HOFFSET(ReactionInfo, product_distance), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "educt_types", HOFFSET(ReactionInfo, educt_types), pair);
    H5Tinsert(t, "product_types", HOFFSET(ReactionInfo, product_types), pair);
    const int rc = write_compound(g, "registered_reactions", t, n, info.data());
    H5Tclose(t);
    H5Tclose(pair);
    H5Tclose(s);
    H5Gclose(g);
    return rc;
}

// a scalar NUL-terminated fixed-length string, the form H5LTmake_dataset_string writes
int rdy_h5_general(int64_t f, const char *json) {
    hid_t g = config_group((hid_t)f);
    if (g < 0) return -1;
    hid_t type = H5Tcopy(H5T_C_S1);
    H5Tset_size(type, std::string(json).size() + 1);
    hid_t space = H5Screate(H5S_SCALAR);
    hid_t ds = H5Dcreate2(g, "general", type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const int rc = ds < 0 ? -1 : (H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, json) < 0 ? -1 : 0);
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(space);
    H5Tclose(type);
    H5Gclose(g);
    return rc;
}

int rdy_h5_attribute(int64_t f, const char *key, const char *value) {
    hid_t g = config_group((hid_t)f);
    if (g < 0) return -1;
    hid_t type = H5Tcopy(H5T_C_S1);
    H5Tset_size(type, std::string(value).size() + 1);
    hid_t space = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(g, key, type, space, H5P_DEFAULT, H5P_DEFAULT);
    int rc = a < 0 ? -1 : (H5Awrite(a, type, value) < 0 ? -1 : 0);
    if (a >= 0) H5Aclose(a);
    H5Sclose(space);
    H5Tclose(type);
    H5Gclose(g);
    return rc;
}

int rdy_h5_close(int64_t f) {
    {
        std::lock_guard<std::mutex> lk(g_chunk_mu);
        g_chunk.erase((hid_t)f);
    }
    return H5Fclose((hid_t)f) < 0 ? -1 : 0;
}

int rdy_h5_read_u64(const char *path, const char *dataset, uint64_t cap, uint64_t *out, uint64_t *n) {
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    int rc = -1;
    hid_t ds = H5Dopen2(f, dataset, H5P_DEFAULT);
    if (ds >= 0) {
        hid_t space = H5Dget_space(ds);
        const hssize_t count = H5Sget_simple_extent_npoints(space);
        *n = count < 0 ? 0 : (uint64_t)count;
        if (count >= 0 && (uint64_t)count <= cap)
            rc = H5Dread(ds, H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0 ? -1 : 0;
        H5Sclose(space);
        H5Dclose(ds);
    }
    H5Fclose(f);
    return rc;
}

int rdy_h5_read_f64(const char *path, const char *dataset, uint64_t cap, double *out, uint64_t *n) {
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    int rc = -1;
    hid_t ds = H5Dopen2(f, dataset, H5P_DEFAULT);
    if (ds >= 0) {
        hid_t space = H5Dget_space(ds);
        const hssize_t count = H5Sget_simple_extent_npoints(space);
        *n = count < 0 ? 0 : (uint64_t)count;
        if (count >= 0 && (uint64_t)count <= cap)
            rc = H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0 ? -1 : 0;
        H5Sclose(space);
        H5Dclose(ds);
    }
    H5Fclose(f);
    return rc;
}

int rdy_h5_read_records(const char *path, const char *dataset, uint64_t cap, uint64_t *ids, uint16_t *types,
                        uint8_t *flavors, double *pos, uint64_t *n) {
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    int rc = -1;
    hid_t ds = H5Dopen2(f, dataset, H5P_DEFAULT);
    if (ds >= 0) {
        hid_t space = H5Dget_space(ds);
        const hssize_t count = H5Sget_simple_extent_npoints(space);
        *n = count < 0 ? 0 : (uint64_t)count;
        if (count >= 0 && (uint64_t)count <= cap) {
            std::vector<Record> rec((size_t)count);
            hid_t mem, file;
            record_types(&mem, &file);
            rc = H5Dread(ds, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, rec.data()) < 0 ? -1 : 0;
            H5Tclose(mem);
            H5Tclose(file);
            for (size_t k = 0; rc == 0 && k < rec.size(); ++k) {
                ids[k] = rec[k].id;
                types[k] = rec[k].typeId;
                flavors[k] = rec[k].flavor;
                for (int a = 0; a < 3; ++a) pos[3 * k + a] = rec[k].pos[a];
            }
        }
        H5Sclose(space);
        H5Dclose(ds);
    }
    H5Fclose(f);
    return rc;
}

int rdy_h5_read_reactions(const char *path, const char *dataset, uint64_t cap_frames, uint64_t *frame_sizes,
                          uint64_t cap, uint64_t *reaction_id, int32_t *type, uint64_t *educts, uint64_t *products,
                          double *position, uint16_t *types_from, uint64_t *n_frames, uint64_t *n) {
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    int rc = -1;
    hid_t ds = H5Dopen2(f, dataset, H5P_DEFAULT);
    if (ds >= 0) {
        hid_t space = H5Dget_space(ds);
        const hssize_t T = H5Sget_simple_extent_npoints(space);
        hid_t ct = reaction_record_type();
        hid_t vt = H5Tvlen_create(ct);
        std::vector<hvl_t> frames(T > 0 ? (size_t)T : 0);
        if (T >= 0 && H5Dread(ds, vt, H5S_ALL, H5S_ALL, H5P_DEFAULT, frames.data()) >= 0) {
            uint64_t total = 0;
            for (const hvl_t &v : frames) total += v.len;
            *n_frames = (uint64_t)T;
            *n = total;
            if ((uint64_t)T <= cap_frames && total <= cap) {
                uint64_t k = 0;
                for (size_t t = 0; t < frames.size(); ++t) {
                    frame_sizes[t] = frames[t].len;
                    const RecordH5 *x = (const RecordH5 *)frames[t].p;
                    for (size_t i = 0; i < frames[t].len; ++i, ++k) {
                        reaction_id[k] = x[i].reaction_id;
                        type[k] = x[i].reaction_type;
                        for (int j = 0; j < 2; ++j) {
                            educts[2 * k + j] = x[i].educts[j];
                            products[2 * k + j] = x[i].products[j];
                            types_from[2 * k + j] = x[i].types_from[j];
                        }
                        for (int d = 0; d < 3; ++d) position[3 * k + d] = x[i].position[d];
                    }
                }
                rc = 0;
            }
            H5Dvlen_reclaim(vt, space, H5P_DEFAULT, frames.data());
        }
        H5Tclose(vt);
        H5Tclose(ct);
        H5Sclose(space);
        H5Dclose(ds);
    }
    H5Fclose(f);
    return rc;
}

int rdy_h5_read_vlen(const char *path, const char *dataset, int elem, uint64_t cap_frames, uint64_t *frame_sizes,
                     uint64_t cap, void *out, uint64_t *n_frames, uint64_t *n) {
    if (elem < 0 || elem > 2) return -1;
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    int rc = -1;
    hid_t ds = H5Dopen2(f, dataset, H5P_DEFAULT);
    if (ds >= 0) {
        hid_t space = H5Dget_space(ds);
        const hssize_t T = H5Sget_simple_extent_npoints(space);
        size_t esz = 0;
        hid_t et = vlen_elem_type(elem, &esz);
        hid_t vt = H5Tvlen_create(et);
        std::vector<hvl_t> frames(T > 0 ? (size_t)T : 0);
        if (T >= 0 && H5Dread(ds, vt, H5S_ALL, H5S_ALL, H5P_DEFAULT, frames.data()) >= 0) {
            uint64_t total = 0;
            for (const hvl_t &v : frames) total += v.len;
            *n_frames = (uint64_t)T;
            *n = total;
            if ((uint64_t)T <= cap_frames && total <= cap) {
                char *o = (char *)out;
                for (size_t t = 0; t < frames.size(); ++t) {
                    frame_sizes[t] = frames[t].len;
                    if (frames[t].len) std::memcpy(o, frames[t].p, frames[t].len * esz);
                    o += frames[t].len * esz;
                }
                rc = 0;
            }
            H5Dvlen_reclaim(vt, space, H5P_DEFAULT, frames.data());
        }
        H5Tclose(vt);
        H5Tclose(et);
        H5Sclose(space);
        H5Dclose(ds);
    }
    H5Fclose(f);
    return rc;
}

int rdy_h5_read_string(const char *path, const char *dataset, uint64_t cap, char *out) {
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    int rc = -1;
    hid_t ds = H5Dopen2(f, dataset, H5P_DEFAULT);
    if (ds >= 0) {
        hid_t ft = H5Dget_type(ds);
        if (H5Tget_class(ft) == H5T_STRING && !H5Tis_variable_str(ft)) {
            const size_t size = H5Tget_size(ft);
            if (size + 1 <= cap) {
                hid_t mt = H5Tcopy(H5T_C_S1);
                H5Tset_size(mt, size + 1);
                rc = H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0 ? -1 : 0;
                H5Tclose(mt);
            }
        }
        H5Tclose(ft);
        H5Dclose(ds);
    }
    H5Fclose(f);
    return rc;
}

}  // extern "C"
