// Observable files in ReaDDy's HDF5 layout (libreaddy_h5.so), for writers that run off the
// simulation's critical path (the serving loop calls it from a writer thread).
//
// Layout read by readdy.Trajectory (wrappers/python/src/python/readdy/api/trajectory.py:420-480):
//   /readdy/observables/<name>/time      [T]          uint64 step index
//   /readdy/observables/<name>/data      [T][n_types] uint64 particles per type        (NParticles.cpp:83)
//   /readdy/observables/<name>/counts/<reaction id>   [T] uint64 events per interval  (ReactionCounts)
// plus string attributes on /readdy/config describing the request.
//
// C ABI (all return 0 on success, -1 on an HDF5 error):
//   int64_t rdy_h5_create(const char *path);                    file handle (> 0) or -1
//   int rdy_h5_n_particles(int64_t f, const char *name, uint64_t T, uint32_t n_types,
//                          const uint64_t *time, const uint64_t *counts);
//   int rdy_h5_reaction_counts(int64_t f, const char *name, uint64_t T, uint32_t n_reactions,
//                              const uint64_t *time, const uint64_t *counts);   counts [T][n_reactions]
//   int rdy_h5_attribute(int64_t f, const char *key, const char *value);
//   int rdy_h5_close(int64_t f);
//   int rdy_h5_read_u64(const char *path, const char *dataset, uint64_t cap, uint64_t *out, uint64_t *n);
//       reads a uint64 dataset by absolute path (checks and tools); *n = its element count
#include <hdf5.h>

#include <cstdint>
#include <string>
#include <vector>

namespace {

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

int write_u64(hid_t g, const char *name, int rank, const hsize_t *dims, const uint64_t *data) {
    hid_t space = H5Screate_simple(rank, dims, nullptr);
    if (space < 0) return -1;
    hid_t ds = H5Dcreate2(g, name, H5T_STD_U64LE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    int rc = -1;
    if (ds >= 0) {
        rc = H5Dwrite(ds, H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0 ? -1 : 0;
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

int rdy_h5_n_particles(int64_t f, const char *name, uint64_t T, uint32_t n_types, const uint64_t *time,
                       const uint64_t *counts) {
    hid_t g = observable_group((hid_t)f, name);
    if (g < 0) return -1;
    hsize_t dt[1] = {T}, dd[2] = {T, n_types};
    int rc = write_u64(g, "time", 1, dt, time);
    if (rc == 0) rc = write_u64(g, "data", 2, dd, counts);
    H5Gclose(g);
    return rc;
}

int rdy_h5_reaction_counts(int64_t f, const char *name, uint64_t T, uint32_t n_reactions, const uint64_t *time,
                           const uint64_t *counts) {
    hid_t g = observable_group((hid_t)f, name);
    if (g < 0) return -1;
    hsize_t dt[1] = {T};
    int rc = write_u64(g, "time", 1, dt, time);
    hid_t c = rc == 0 ? ensure_group(g, "counts") : -1;
    if (c < 0) rc = -1;
    std::vector<uint64_t> col(T);
    for (uint32_t r = 0; rc == 0 && r < n_reactions; ++r) {
        for (uint64_t t = 0; t < T; ++t) col[t] = counts[t * n_reactions + r];
        rc = write_u64(c, std::to_string(r).c_str(), 1, dt, col.data());
    }
    if (c >= 0) H5Gclose(c);
    H5Gclose(g);
    return rc;
}

int rdy_h5_attribute(int64_t f, const char *key, const char *value) {
    hid_t g0 = ensure_group((hid_t)f, "readdy");
    if (g0 < 0) return -1;
    hid_t g = ensure_group(g0, "config");
    H5Gclose(g0);
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

int rdy_h5_close(int64_t f) { return H5Fclose((hid_t)f) < 0 ? -1 : 0; }

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

}  // extern "C"
