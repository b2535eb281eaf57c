// h5store.hpp -- MARQOV's HDF5 run files, one per simulation, written and read
// with the HDF5 C library.
//
// The layout is the reference's (core.h:650-768 setupHDF5Container / createstep,
// core.h:820-917 dumprng / dumpstatespace, cachecontainer.h:166-205 scalars,
// cachecontainer.h:238-458 time series, doc/io.rst:33-47):
//
//   /step<N>                          one per run of the configuration: 0 for a fresh
//                                     start, then one more per resumed run
//     environment/                    Code, Version, hostname, startingdate (strings)
//     config/marqovconfig/            beta (double) and id, repid, seed, nsteps,
//                                     warmupsteps, gameloopsteps, ncluster, nmetro
//     config/lattice/                 name, dim, npoints (RegularHypercubic);
//                                     name, npoints (GraphFromCSV, "unknown")
//     config/hamiltonian/             Model (string) and the Hamiltonian's
//                                     constructor arguments (param<i> or their names)
//     observables/<name>              one time series per observable: 1-D,
//                                     extendible, 4 MiB chunks, deflate level 9
//     state/RNG                       name of the generator
//     state/rngstate                  the words that restart it (uint64)
//     state/hamiltonianstatespace     N state vectors (fixed-size arrays of int32 or
//                                     double) in the lattice's site order; 4 MiB
//                                     chunks, shuffle + deflate level 9
//
// Strings are fixed-length C strings of the value's length, scalars scalar
// datasets, as the reference's dumpscalartoH5 writes them.  A step is written as a
// whole at the end of a run, so a file never holds a step without its state.
#pragma once

#include <hdf5.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace h5store {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// An HDF5 identifier closed by the given function when the holder goes away.
class Id {
public:
    Id() = default;
    Id(hid_t id, herr_t (*close)(hid_t), const std::string& what) : id_(id), close_(close)
    {
        if (id_ < 0) throw Error("HDF5: " + what);
    }
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    Id(Id&& o) noexcept : id_(o.id_), close_(o.close_) { o.id_ = -1; }
    Id& operator=(Id&& o) noexcept
    {
        if (this != &o) {
            reset();
            id_ = o.id_;
            close_ = o.close_;
            o.id_ = -1;
        }
        return *this;
    }
    ~Id() { reset(); }
    hid_t get() const { return id_; }
    void reset()
    {
        if (id_ >= 0 && close_) close_(id_);
        id_ = -1;
    }

private:
    hid_t id_ = -1;
    herr_t (*close_)(hid_t) = nullptr;
};

inline void check(herr_t e, const std::string& what)
{
    if (e < 0) throw Error("HDF5: " + what);
}

// The library prints its own error stack unless told not to; errors here become
// exceptions with the failing step named.
inline void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

inline bool is_hdf5(const std::string& path)
{
    quiet();
    return H5Fis_hdf5(path.c_str()) > 0;
}

inline Id create_file(const std::string& path)
{
    quiet();
    return Id(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, "cannot create " + path);
}

inline Id open_file(const std::string& path, bool write)
{
    quiet();
    return Id(H5Fopen(path.c_str(), write ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose,
              "cannot open " + path);
}

inline Id group(hid_t loc, const std::string& name, const char* comment = nullptr)
{
    Id g(H5Gcreate2(loc, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "cannot create group " + name);
    if (comment) check(H5Oset_comment(g.get(), comment), "comment on " + name);
    return g;
}

inline Id open_group(hid_t loc, const std::string& name)
{
    return Id(H5Gopen2(loc, name.c_str(), H5P_DEFAULT), H5Gclose, "no group " + name);
}

// Highest N of the root's step<N> groups, -1 without any (findstep, core.h:618-627).
inline int last_step(hid_t file)
{
    int best = -1;
    auto visit = [](hid_t, const char* name, const H5L_info_t*, void* data) -> herr_t {
        const std::size_t n = std::strlen(name);
        if (n > 4 && std::strncmp(name, "step", 4) == 0) {
            long v = 0;
            bool digits = true;
            for (std::size_t i = 4; i < n && digits; ++i) {
                digits = name[i] >= '0' && name[i] <= '9';
                v = v * 10 + (name[i] - '0');
                if (v > 1000000000L) digits = false;
            }
            if (digits && v > *static_cast<int*>(data)) *static_cast<int*>(data) = (int)v;
        }
        return 0;
    };
    hsize_t idx = 0;
    check(H5Literate(file, H5_INDEX_NAME, H5_ITER_NATIVE, &idx, visit, &best), "cannot list the file's steps");
    return best;
}

// ---- scalars (dumpscalartoH5)
template <class T>
hid_t native();
template <>
inline hid_t native<double>() { return H5T_NATIVE_DOUBLE; }
template <>
inline hid_t native<int32_t>() { return H5T_NATIVE_INT32; }
template <>
inline hid_t native<int64_t>() { return H5T_NATIVE_INT64; }
template <>
inline hid_t native<uint32_t>() { return H5T_NATIVE_UINT32; }
template <>
inline hid_t native<uint64_t>() { return H5T_NATIVE_UINT64; }

template <class T>
void scalar(hid_t loc, const std::string& key, const T& v)
{
    Id sp(H5Screate(H5S_SCALAR), H5Sclose, "scalar space");
    Id ds(H5Dcreate2(loc, key.c_str(), native<T>(), sp.get(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
          "cannot create " + key);
    check(H5Dwrite(ds.get(), native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), "cannot write " + key);
}

inline void string(hid_t loc, const std::string& key, const std::string& v)
{
    Id ty(H5Tcopy(H5T_C_S1), H5Tclose, "string type");
    check(H5Tset_size(ty.get(), v.empty() ? 1 : v.size()), "string size");
    Id sp(H5Screate(H5S_SCALAR), H5Sclose, "scalar space");
    Id ds(H5Dcreate2(loc, key.c_str(), ty.get(), sp.get(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
          "cannot create " + key);
    const std::string buf = v.empty() ? std::string(1, '\0') : v;
    check(H5Dwrite(ds.get(), ty.get(), H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()), "cannot write " + key);
}

template <class T>
T read_scalar(hid_t loc, const std::string& key)
{
    Id ds(H5Dopen2(loc, key.c_str(), H5P_DEFAULT), H5Dclose, "no dataset " + key);
    T v{};
    check(H5Dread(ds.get(), native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), "cannot read " + key);
    return v;
}

inline std::string read_string(hid_t loc, const std::string& key)
{
    Id ds(H5Dopen2(loc, key.c_str(), H5P_DEFAULT), H5Dclose, "no dataset " + key);
    Id ft(H5Dget_type(ds.get()), H5Tclose, "type of " + key);
    if (H5Tget_class(ft.get()) != H5T_STRING || H5Tis_variable_str(ft.get()) > 0)
        throw Error("HDF5: " + key + " is not a fixed-length string");
    const std::size_t n = H5Tget_size(ft.get());
    Id mt(H5Tcopy(H5T_C_S1), H5Tclose, "string type");
    check(H5Tset_size(mt.get(), n + 1), "string size");
    std::vector<char> buf(n + 1, '\0');
    check(H5Dread(ds.get(), mt.get(), H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()), "cannot read " + key);
    return std::string(buf.data());
}

// ---- series: n records, each a scalar (array = false, width 1) or a fixed-size
// array of `width` values, extendible along the record axis
template <class T>
void series(hid_t loc, const std::string& key, const T* data, std::size_t n, std::size_t width, bool array,
            std::size_t chunk, bool shuffle, const char* comment)
{
    const hsize_t dims[1] = {(hsize_t)n}, maxdims[1] = {H5S_UNLIMITED}, cdims[1] = {(hsize_t)chunk};
    Id sp(H5Screate_simple(1, dims, maxdims), H5Sclose, "space of " + key);
    Id ty;
    if (!array) {
        ty = Id(H5Tcopy(native<T>()), H5Tclose, "type of " + key);
    } else {
        const hsize_t adims[1] = {(hsize_t)width};
        ty = Id(H5Tarray_create2(native<T>(), 1, adims), H5Tclose, "array type of " + key);
    }
    Id pl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose, "creation list of " + key);
    check(H5Pset_chunk(pl.get(), 1, cdims), "chunks of " + key);
    if (shuffle) check(H5Pset_shuffle(pl.get()), "shuffle of " + key);
    check(H5Pset_deflate(pl.get(), 9), "deflate of " + key);
    // zero fill value, except for records of 8,192 values or more, whose fill value would
    // not fit the dataset's object header (the reference sets none there either,
    // cachecontainer.h:406-410)
    if (width < 8192) {
        const std::vector<T> fill(width, T(0));
        check(H5Pset_fill_value(pl.get(), ty.get(), fill.data()), "fill value of " + key);
    }
    Id ds(H5Dcreate2(loc, key.c_str(), ty.get(), sp.get(), H5P_DEFAULT, pl.get(), H5P_DEFAULT), H5Dclose,
          "cannot create " + key);
    if (comment) check(H5Oset_comment(ds.get(), comment), "comment on " + key);
    if (n > 0) check(H5Dwrite(ds.get(), ty.get(), H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "cannot write " + key);
}

// Elements per 4 MiB chunk for records of `bytes` bytes (cachecontainer.h:262,
// core.h:887), at least one.
inline std::size_t chunk_of(std::size_t bytes) { return std::max<std::size_t>(1, (4096u * 1024u) / bytes); }

// Records per chunk of a vector-valued series of `width` values of `bytes` each:
// 4 MiB / value / width + 1, whatever the vector's length (cachecontainer.h:403).
inline std::size_t vector_chunk(std::size_t bytes, std::size_t width) { return (4096u * 1024u) / bytes / width + 1; }

// A fixed-length array (rngstate, cachecontainer.h dumprng: dims = maxdims = n, no chunks).
inline void words(hid_t loc, const std::string& key, const std::vector<uint64_t>& v)
{
    const hsize_t dims[1] = {(hsize_t)v.size()};
    Id sp(H5Screate_simple(1, dims, dims), H5Sclose, "space of " + key);
    Id pl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose, "creation list of " + key);
    const uint64_t zero = 0;
    check(H5Pset_fill_value(pl.get(), H5T_NATIVE_UINT64, &zero), "fill value of " + key);
    Id ds(H5Dcreate2(loc, key.c_str(), H5T_NATIVE_UINT64, sp.get(), H5P_DEFAULT, pl.get(), H5P_DEFAULT), H5Dclose,
          "cannot create " + key);
    if (!v.empty())
        check(H5Dwrite(ds.get(), H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), "cannot write " + key);
}

inline std::vector<uint64_t> read_words(hid_t loc, const std::string& key)
{
    Id ds(H5Dopen2(loc, key.c_str(), H5P_DEFAULT), H5Dclose, "no dataset " + key);
    Id sp(H5Dget_space(ds.get()), H5Sclose, "space of " + key);
    if (H5Sget_simple_extent_ndims(sp.get()) != 1) throw Error("HDF5: " + key + " is not one-dimensional");
    hsize_t n = 0;
    H5Sget_simple_extent_dims(sp.get(), &n, nullptr);
    std::vector<uint64_t> v((std::size_t)n);
    if (n) check(H5Dread(ds.get(), H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), "cannot read " + key);
    return v;
}

// n records of `width` values from a series written by series(); a stored record
// must hold `width` values (a scalar counts as one), converted to T by the library.
template <class T>
std::vector<T> read_series(hid_t loc, const std::string& key, std::size_t expect_n, std::size_t width)
{
    Id ds(H5Dopen2(loc, key.c_str(), H5P_DEFAULT), H5Dclose, "no dataset " + key);
    Id sp(H5Dget_space(ds.get()), H5Sclose, "space of " + key);
    if (H5Sget_simple_extent_ndims(sp.get()) != 1) throw Error("HDF5: " + key + " is not one-dimensional");
    hsize_t n = 0;
    H5Sget_simple_extent_dims(sp.get(), &n, nullptr);
    if ((std::size_t)n != expect_n)
        throw Error("HDF5: " + key + " holds " + std::to_string(n) + " records, expected " + std::to_string(expect_n));
    Id ft(H5Dget_type(ds.get()), H5Tclose, "type of " + key);
    std::size_t fw = 1;
    if (H5Tget_class(ft.get()) == H5T_ARRAY) {
        if (H5Tget_array_ndims(ft.get()) != 1) throw Error("HDF5: " + key + " records are not vectors");
        hsize_t ad = 0;
        H5Tget_array_dims2(ft.get(), &ad);
        fw = (std::size_t)ad;
    }
    if (fw != width)
        throw Error("HDF5: " + key + " records have " + std::to_string(fw) + " values, expected " + std::to_string(width));
    Id mt;
    if (H5Tget_class(ft.get()) != H5T_ARRAY) {
        mt = Id(H5Tcopy(native<T>()), H5Tclose, "memory type");
    } else {
        const hsize_t adims[1] = {(hsize_t)width};
        mt = Id(H5Tarray_create2(native<T>(), 1, adims), H5Tclose, "memory array type");
    }
    std::vector<T> v((std::size_t)n * width);
    if (n) check(H5Dread(ds.get(), mt.get(), H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), "cannot read " + key);
    return v;
}

}  // namespace h5store
