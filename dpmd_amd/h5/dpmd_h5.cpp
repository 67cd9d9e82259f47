// libdpmd_h5: HDF5 I/O in the reference's layout, and its meta.h5 particle loader.
//
// Host-only C++ over the HDF5 C library.  The layout (dataset names, ranks, native types,
// scalar datasets, UTF-8 variable-length strings, (T, N[, 2]) trajectories chunked along T,
// /ragged/<field>/step_<k>) is the one written by the reference's io::OutputManager and read
// by BaseParticle::load_from_hdf5; see include/dpmd/dpmd_h5.h.
#include "dpmd/dpmd_h5.h"

#include <hdf5.h>

#include <algorithm>
#include <cstring>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

struct ArgError : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

template <class F>
int guard(F&& f) {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);   // errors are reported through the status
    try {
        f();
        return DPMD_OK;
    } catch (const ArgError& e) {
        g_error = e.what();
        return DPMD_ERR_ARGUMENT;
    } catch (const std::exception& e) {
        g_error = e.what();
        return DPMD_ERR_STATE;
    }
}

void need(bool ok, const std::string& what) {
    if (!ok) throw std::runtime_error(what);
}

// Owns one HDF5 identifier.
struct Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    ~Id() {
        if (id >= 0 && close) close(id);
    }
    operator hid_t() const { return id; }
};

hid_t file_id(int64_t f) {
    if (f <= 0 || H5Iget_type((hid_t)f) != H5I_FILE) throw ArgError("dpmd_h5: invalid file handle");
    return (hid_t)f;
}

std::string str(const char* s, const char* what) {
    if (!s) throw ArgError(std::string("dpmd_h5: null ") + what);
    return s;
}

hid_t native(int type) {
    if (type == DPMD_H5_F64) return H5T_NATIVE_DOUBLE;
    if (type == DPMD_H5_I32) return H5T_NATIVE_INT;
    throw ArgError("dpmd_h5: unknown element type " + std::to_string(type));
}

size_t elem_size(int type) { return type == DPMD_H5_F64 ? 8 : 4; }

// A link exists at every prefix of the path (H5Lexists needs the parents to exist).
bool exists(hid_t loc, const std::string& path) {
    if (path.empty() || path == "/") return true;
    size_t pos = path[0] == '/' ? 1 : 0;
    while (true) {
        const size_t slash = path.find('/', pos);
        const std::string prefix = path.substr(0, slash);
        if (!prefix.empty() && prefix != "/") {
            const htri_t e = H5Lexists(loc, prefix.c_str(), H5P_DEFAULT);
            if (e <= 0) return false;
        }
        if (slash == std::string::npos) return true;
        pos = slash + 1;
    }
}

void remove_link(hid_t loc, const std::string& path) {
    if (exists(loc, path)) need(H5Ldelete(loc, path.c_str(), H5P_DEFAULT) >= 0, "dpmd_h5: cannot delete " + path);
}

// Link-creation properties that create missing parent groups.
hid_t new_link_props() {
    const hid_t p = H5Pcreate(H5P_LINK_CREATE);
    need(p >= 0, "dpmd_h5: link properties");
    if (H5Pset_create_intermediate_group(p, 1) < 0) {
        H5Pclose(p);
        throw std::runtime_error("dpmd_h5: link properties");
    }
    return p;
}

hid_t open_dataset_id(hid_t f, const std::string& path) {
    if (!exists(f, path)) throw ArgError("dpmd_h5: missing dataset " + path);
    const hid_t d = H5Dopen2(f, path.c_str(), H5P_DEFAULT);
    need(d >= 0, "dpmd_h5: cannot open dataset " + path);
    return d;
}

std::vector<hsize_t> dims_of(hid_t d) {
    Id sp(H5Dget_space(d), H5Sclose);
    const int nd = H5Sget_simple_extent_ndims(sp);
    need(nd >= 0, "dpmd_h5: dataspace");
    std::vector<hsize_t> dims((size_t)nd);
    if (nd) H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
    return dims;
}

void write_dataset(hid_t f, const std::string& path, hid_t mtype, hid_t ftype, int rank, const hsize_t* dims,
                   const void* data) {
    remove_link(f, path);
    Id sp(rank ? H5Screate_simple(rank, dims, nullptr) : H5Screate(H5S_SCALAR), H5Sclose);
    need(sp >= 0, "dpmd_h5: dataspace for " + path);
    Id lcpl(new_link_props(), H5Pclose);
    Id d(H5Dcreate2(f, path.c_str(), ftype, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    need(d >= 0, "dpmd_h5: cannot create dataset " + path);
    hsize_t n = 1;
    for (int k = 0; k < rank; ++k) n *= dims[k];
    if (n) need(H5Dwrite(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0, "dpmd_h5: cannot write " + path);
}

std::string read_string_at(hid_t f, const std::string& path) {
    Id d(open_dataset_id(f, path), H5Dclose);
    Id ft(H5Dget_type(d), H5Tclose);
    if (H5Tget_class(ft) != H5T_STRING) throw ArgError("dpmd_h5: " + path + " is not a string");
    if (H5Tis_variable_str(ft) > 0) {
        Id mt(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(mt, H5T_VARIABLE);
        H5Tset_cset(mt, H5Tget_cset(ft));
        char* p = nullptr;
        need(H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, &p) >= 0, "dpmd_h5: cannot read " + path);
        std::string s = p ? p : "";
        if (p) H5free_memory(p);
        return s;
    }
    const size_t n = H5Tget_size(ft);
    std::vector<char> buf(n + 1, '\0');
    Id mt(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mt, n);
    H5Tset_cset(mt, H5Tget_cset(ft));
    H5Tset_strpad(mt, H5Tget_strpad(ft));
    need(H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) >= 0, "dpmd_h5: cannot read " + path);
    return std::string(buf.data());
}

template <class T>
std::vector<T> read_all(hid_t f, const std::string& path, int type, std::vector<hsize_t>* dims_out = nullptr) {
    Id d(open_dataset_id(f, path), H5Dclose);
    auto dims = dims_of(d);
    size_t n = 1;
    for (hsize_t x : dims) n *= (size_t)x;
    std::vector<T> v(n);
    if (n) need(H5Dread(d, native(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) >= 0, "dpmd_h5: cannot read " + path);
    if (dims_out) *dims_out = dims;
    return v;
}

}  // namespace

extern "C" {

const char* dpmd_h5_last_error(void) { return g_error.c_str(); }

int dpmd_h5_open(const char* path, int mode, int64_t* file) {
    return guard([&] {
        const std::string p = str(path, "path");
        if (!file) throw ArgError("dpmd_h5_open: null handle");
        hid_t f = -1;
        if (mode == DPMD_H5_CREATE) {
            f = H5Fcreate(p.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        } else if (mode == DPMD_H5_READ_WRITE) {
            f = std::filesystem::exists(p) ? H5Fopen(p.c_str(), H5F_ACC_RDWR, H5P_DEFAULT)
                                           : H5Fcreate(p.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        } else if (mode == DPMD_H5_READ_ONLY) {
            if (!std::filesystem::exists(p)) throw ArgError("dpmd_h5_open: no such file " + p);
            f = H5Fopen(p.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
        } else {
            throw ArgError("dpmd_h5_open: unknown mode " + std::to_string(mode));
        }
        need(f >= 0, "dpmd_h5_open: cannot open " + p);
        *file = (int64_t)f;
    });
}

int dpmd_h5_close(int64_t file) {
    return guard([&] { need(H5Fclose(file_id(file)) >= 0, "dpmd_h5_close failed"); });
}

int dpmd_h5_flush(int64_t file) {
    return guard([&] { need(H5Fflush(file_id(file), H5F_SCOPE_GLOBAL) >= 0, "dpmd_h5_flush failed"); });
}

int dpmd_h5_exists(int64_t file, const char* path, int* out) {
    return guard([&] {
        if (!out) throw ArgError("dpmd_h5_exists: null output");
        *out = exists(file_id(file), str(path, "path")) ? 1 : 0;
    });
}

int dpmd_h5_delete(int64_t file, const char* path) {
    return guard([&] { remove_link(file_id(file), str(path, "path")); });
}

int dpmd_h5_move(int64_t file, const char* from, const char* to) {
    return guard([&] {
        const hid_t f = file_id(file);
        const std::string a = str(from, "source"), b = str(to, "destination");
        if (!exists(f, a)) throw ArgError("dpmd_h5_move: missing " + a);
        remove_link(f, b);
        Id lcpl(new_link_props(), H5Pclose);
        need(H5Lmove(f, a.c_str(), f, b.c_str(), lcpl, H5P_DEFAULT) >= 0, "dpmd_h5_move: " + a + " -> " + b);
    });
}

int dpmd_h5_write_array(int64_t file, const char* path, int type, int64_t rows, int cols, const void* data) {
    return guard([&] {
        const hid_t f = file_id(file);
        if (rows < 0 || (cols != 0 && cols != 2)) throw ArgError("dpmd_h5_write_array: bad shape");
        if (rows && !data) throw ArgError("dpmd_h5_write_array: null data");
        const hsize_t dims[2] = {(hsize_t)rows, 2};
        write_dataset(f, str(path, "path"), native(type), native(type), cols ? 2 : 1, dims, data);
    });
}

int dpmd_h5_write_int(int64_t file, const char* path, int value) {
    return guard([&] {
        write_dataset(file_id(file), str(path, "path"), H5T_NATIVE_INT, H5T_NATIVE_INT, 0, nullptr, &value);
    });
}

int dpmd_h5_write_string(int64_t file, const char* path, const char* value) {
    return guard([&] {
        const std::string v = str(value, "value");
        Id t(H5Tcopy(H5T_C_S1), H5Tclose);
        need(H5Tset_size(t, H5T_VARIABLE) >= 0 && H5Tset_cset(t, H5T_CSET_UTF8) >= 0, "dpmd_h5: string type");
        const char* p = v.c_str();
        write_dataset(file_id(file), str(path, "path"), t, t, 0, nullptr, &p);
    });
}

int dpmd_h5_shape(int64_t file, const char* path, int* ndims, int64_t* dims3, int* type) {
    return guard([&] {
        Id d(open_dataset_id(file_id(file), str(path, "path")), H5Dclose);
        auto dims = dims_of(d);
        if (dims.size() > 3) throw ArgError("dpmd_h5_shape: rank above 3");
        if (ndims) *ndims = (int)dims.size();
        if (dims3)
            for (int k = 0; k < 3; ++k) dims3[k] = k < (int)dims.size() ? (int64_t)dims[k] : 0;
        if (type) {
            Id t(H5Dget_type(d), H5Tclose);
            const H5T_class_t c = H5Tget_class(t);
            *type = c == H5T_FLOAT ? DPMD_H5_F64 : c == H5T_INTEGER ? DPMD_H5_I32 : -1;
        }
    });
}

int dpmd_h5_read_array(int64_t file, const char* path, int type, void* out, int64_t count) {
    return guard([&] {
        Id d(open_dataset_id(file_id(file), str(path, "path")), H5Dclose);
        auto dims = dims_of(d);
        int64_t n = 1;
        for (hsize_t x : dims) n *= (int64_t)x;
        if (n != count) throw ArgError("dpmd_h5_read_array: " + std::string(path) + " holds " + std::to_string(n) +
                                       " elements, not " + std::to_string(count));
        if (n && !out) throw ArgError("dpmd_h5_read_array: null output");
        if (n) need(H5Dread(d, native(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, out) >= 0, "dpmd_h5: cannot read " + std::string(path));
    });
}

int dpmd_h5_read_int(int64_t file, const char* path, int* value) {
    return guard([&] {
        if (!value) throw ArgError("dpmd_h5_read_int: null output");
        auto v = read_all<int>(file_id(file), str(path, "path"), DPMD_H5_I32);
        if (v.size() != 1) throw ArgError("dpmd_h5_read_int: " + std::string(path) + " is not a scalar");
        *value = v[0];
    });
}

int dpmd_h5_read_string(int64_t file, const char* path, char* buf, int64_t capacity) {
    return guard([&] {
        const std::string s = read_string_at(file_id(file), str(path, "path"));
        if (!buf || capacity < (int64_t)s.size() + 1) throw ArgError("dpmd_h5_read_string: buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

int dpmd_h5_list(int64_t file, const char* group, char* buf, int64_t capacity) {
    int status = DPMD_OK;
    const int st = guard([&] {
        const hid_t f = file_id(file);
        const std::string g = str(group, "group");
        if (!exists(f, g)) throw ArgError("dpmd_h5_list: missing group " + g);
        Id gid(H5Gopen2(f, g.c_str(), H5P_DEFAULT), H5Gclose);
        need(gid >= 0, "dpmd_h5_list: cannot open group " + g);
        H5G_info_t info;
        need(H5Gget_info(gid, &info) >= 0, "dpmd_h5_list: group info");
        std::string out;
        for (hsize_t k = 0; k < info.nlinks; ++k) {
            const ssize_t len = H5Lget_name_by_idx(gid, ".", H5_INDEX_NAME, H5_ITER_INC, k, nullptr, 0, H5P_DEFAULT);
            need(len >= 0, "dpmd_h5_list: link name");
            std::string name((size_t)len + 1, '\0');
            H5Lget_name_by_idx(gid, ".", H5_INDEX_NAME, H5_ITER_INC, k, name.data(), name.size(), H5P_DEFAULT);
            name.resize((size_t)len);
            H5O_info_t oi;
            const char* kind = "other";
            if (H5Oget_info_by_name(gid, name.c_str(), &oi, H5P_DEFAULT) >= 0)
                kind = oi.type == H5O_TYPE_GROUP ? "group" : oi.type == H5O_TYPE_DATASET ? "dataset" : "other";
            out += std::string(kind) + " " + name + "\n";
        }
        if (!buf || capacity < (int64_t)out.size() + 1) {
            status = DPMD_ERR_OVERFLOW;
            throw std::runtime_error("dpmd_h5_list: buffer of " + std::to_string(capacity) + " bytes is too small");
        }
        std::memcpy(buf, out.c_str(), out.size() + 1);
    });
    return status != DPMD_OK ? status : st;
}

int dpmd_h5_write_frame(int64_t file, const char* path, int type, int64_t n, int cols, const void* data, int64_t row) {
    return guard([&] {
        const hid_t f = file_id(file);
        const std::string p = str(path, "path");
        if (n <= 0 || row < 0 || (cols != 0 && cols != 2) || !data) throw ArgError("dpmd_h5_write_frame: bad frame");
        const int rank = cols ? 3 : 2;
        hsize_t dims[3] = {(hsize_t)row + 1, (hsize_t)n, 2};
        if (!exists(f, p)) {
            // Chunks of whole frames, about 1 MiB or one frame when frames are larger.
            const hsize_t frame = (hsize_t)n * (cols ? 2 : 1) * elem_size(type);
            const hsize_t per = std::max<hsize_t>(1, std::min<hsize_t>(64, ((hsize_t)1 << 20) / frame));
            hsize_t maxd[3] = {H5S_UNLIMITED, (hsize_t)n, 2}, chunk[3] = {per, (hsize_t)n, 2};
            Id sp(H5Screate_simple(rank, dims, maxd), H5Sclose);
            Id dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
            need(sp >= 0 && dcpl >= 0 && H5Pset_chunk(dcpl, rank, chunk) >= 0, "dpmd_h5: frame dataset " + p);
            Id lcpl(new_link_props(), H5Pclose);
            Id d(H5Dcreate2(f, p.c_str(), native(type), sp, lcpl, dcpl, H5P_DEFAULT), H5Dclose);
            need(d >= 0, "dpmd_h5: cannot create " + p);
        }
        Id d(open_dataset_id(f, p), H5Dclose);
        auto cur = dims_of(d);
        if ((int)cur.size() != rank || cur[1] != (hsize_t)n || (cols && cur[2] != 2))
            throw ArgError("dpmd_h5_write_frame: frame shape differs from " + p);
        if (cur[0] < (hsize_t)row + 1) {
            cur[0] = (hsize_t)row + 1;
            need(H5Dset_extent(d, cur.data()) >= 0, "dpmd_h5: cannot extend " + p);
        }
        Id fs(H5Dget_space(d), H5Sclose);
        hsize_t start[3] = {(hsize_t)row, 0, 0}, count[3] = {1, (hsize_t)n, 2};
        need(H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr) >= 0, "dpmd_h5: hyperslab");
        Id ms(H5Screate_simple(rank, count, nullptr), H5Sclose);
        need(H5Dwrite(d, native(type), ms, fs, H5P_DEFAULT, data) >= 0, "dpmd_h5: cannot write frame of " + p);
    });
}

int dpmd_h5_trim_frames(int64_t file, int64_t rows) {
    return guard([&] {
        const hid_t f = file_id(file);
        if (rows < 0) throw ArgError("dpmd_h5_trim_frames: negative rows");
        H5G_info_t info;
        need(H5Gget_info(f, &info) >= 0, "dpmd_h5: root group");
        for (hsize_t k = 0; k < info.nlinks; ++k) {
            char name[512];
            if (H5Lget_name_by_idx(f, ".", H5_INDEX_NAME, H5_ITER_INC, k, name, sizeof name, H5P_DEFAULT) < 0) continue;
            H5O_info_t oi;
            if (H5Oget_info_by_name(f, name, &oi, H5P_DEFAULT) < 0 || oi.type != H5O_TYPE_DATASET) continue;
            Id d(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
            Id sp(H5Dget_space(d), H5Sclose);
            const int nd = H5Sget_simple_extent_ndims(sp);
            if (nd < 1 || nd > 3) continue;
            hsize_t cur[3], mx[3];
            H5Sget_simple_extent_dims(sp, cur, mx);
            if (mx[0] != H5S_UNLIMITED || cur[0] <= (hsize_t)rows) continue;
            cur[0] = (hsize_t)rows;
            need(H5Dset_extent(d, cur) >= 0, std::string("dpmd_h5: cannot trim ") + name);
        }
    });
}

}  // extern "C"

// ---- BaseParticle::load_from_hdf5 (base_particle.hpp:120-257, base_point_particle.hpp:221-242,
// base_poly_particle.hpp:379-409, rigid_bumpy.cu:1856-1881) ------------------------------

namespace {

struct Loader {
    dpmd_ctx* c;
    hid_t f;
    int kind;
    int S = 0, N = 0, Nv = 0;

    void ok(int st, const std::string& what) {
        if (st != DPMD_OK) {
            const char* m = dpmd_last_error(c);
            throw std::runtime_error("dpmd_h5_load: " + what + ": " + (m ? m : "error"));
        }
    }
    int64_t field_len(int fid) {
        int64_t n = 0;
        ok(dpmd_field(c, fid, nullptr, &n), "field length");
        return n;
    }
    // (n,) dataset into one field
    template <class T>
    void set1(const std::string& path, int fid, int type) {
        std::vector<hsize_t> dims;
        auto v = read_all<T>(f, path, type, &dims);
        const int64_t n = field_len(fid);
        if (dims.size() != 1 || (int64_t)dims[0] != n)
            throw ArgError("dpmd_h5_load: " + path + " must be (" + std::to_string(n) + ",)");
        ok(dpmd_set_field(c, fid, v.data(), n), path);
    }
    // (n, 2) dataset into two fields
    template <class T>
    void set2(const std::string& path, int fx, int fy, int type) {
        std::vector<hsize_t> dims;
        auto v = read_all<T>(f, path, type, &dims);
        const int64_t n = field_len(fx);
        if (dims.size() != 2 || (int64_t)dims[0] != n || dims[1] != 2)
            throw ArgError("dpmd_h5_load: " + path + " must be (" + std::to_string(n) + ", 2)");
        std::vector<T> x((size_t)n), y((size_t)n);
        for (int64_t k = 0; k < n; ++k) { x[(size_t)k] = v[(size_t)(2 * k)]; y[(size_t)k] = v[(size_t)(2 * k + 1)]; }
        ok(dpmd_set_field(c, fx, x.data(), n), path);
        ok(dpmd_set_field(c, fy, y.data(), n), path);
    }
    bool has(const std::string& g, const char* name) { return exists(f, "/" + g + "/" + name); }
    void opt1(const std::string& g, const char* name, int fid) {
        if (has(g, name)) set1<double>("/" + g + "/" + name, fid, DPMD_H5_F64);
    }
    void opt2(const std::string& g, const char* name, int fx, int fy) {
        if (has(g, name)) set2<double>("/" + g + "/" + name, fx, fy, DPMD_H5_F64);
    }
    // A field the context derives: the file's values must agree with the derived ones.
    void check_int(const std::string& path, int fid) {
        if (!exists(f, path)) return;
        auto v = read_all<int>(f, path, DPMD_H5_I32);
        const int64_t n = field_len(fid);
        std::vector<int> d((size_t)n);
        ok(dpmd_get_field(c, fid, d.data(), n, DPMD_ORDER_CANONICAL), path);
        if (v != d) throw ArgError("dpmd_h5_load: " + path + " disagrees with the system and particle offsets");
    }

    // load_from_hdf5_group and its point / poly / rigid extras: whatever the group holds
    void load_group(const std::string& g) {
        opt2(g, "box_size", DPMD_F_BOX_SIZE_X, DPMD_F_BOX_SIZE_Y);
        if (has(g, "cell_dim")) set2<int>("/" + g + "/cell_dim", DPMD_F_CELL_DIM_X, DPMD_F_CELL_DIM_Y, DPMD_H5_I32);
        opt1(g, "packing_fraction", DPMD_F_PACKING_FRACTION);
        opt1(g, "pressure", DPMD_F_PRESSURE);
        opt1(g, "temperature", DPMD_F_TEMPERATURE);
        opt1(g, "area", DPMD_F_AREA);
        opt1(g, "pe", DPMD_F_PE);
        opt1(g, "ke", DPMD_F_KE);
        opt1(g, "pe_total", DPMD_F_PE_TOTAL);
        opt1(g, "ke_total", DPMD_F_KE_TOTAL);
        opt1(g, "fpower_total", DPMD_F_FPOWER_TOTAL);
        opt2(g, "pos", DPMD_F_POS_X, DPMD_F_POS_Y);
        opt2(g, "vel", DPMD_F_VEL_X, DPMD_F_VEL_Y);
        opt2(g, "force", DPMD_F_FORCE_X, DPMD_F_FORCE_Y);
        if (kind == DPMD_RIGID_BUMPY) {
            opt2(g, "vertex_pos", DPMD_F_VERTEX_POS_X, DPMD_F_VERTEX_POS_Y);
            opt2(g, "vertex_force", DPMD_F_VERTEX_FORCE_X, DPMD_F_VERTEX_FORCE_Y);
            opt1(g, "vertex_pe", DPMD_F_VERTEX_PE);
            opt1(g, "angle", DPMD_F_ANGLE);
            opt1(g, "torque", DPMD_F_TORQUE);
            opt1(g, "angular_vel", DPMD_F_ANGULAR_VEL);
        }
    }

    void run(const std::string& location) {
        int method = 0;
        ok(dpmd_ctx_info(c, &kind, &method), "context");
        if (exists(f, "/static/class_name")) {
            const std::string cls = read_string_at(f, "/static/class_name");
            const char* want = kind == DPMD_RIGID_BUMPY ? "RigidBumpy" : "Disk";
            if (cls != want) throw ArgError("dpmd_h5_load: meta.h5 holds a " + cls + ", not a " + want);
        }
        auto scalar = [&](const char* p) {
            auto a = read_all<int>(f, p, DPMD_H5_I32);
            if (a.size() != 1) throw ArgError(std::string("dpmd_h5_load: ") + p + " is not a scalar");
            return a[0];
        };
        S = scalar("/static/n_systems");
        N = scalar("/static/n_particles");
        Nv = scalar("/static/n_vertices");
        const std::string nm = read_string_at(f, "/static/neighbor_method");
        if (nm == "Naive") method = DPMD_NEIGHBOR_NAIVE;
        else if (nm == "Cell") method = DPMD_NEIGHBOR_CELL;
        else throw ArgError("dpmd_h5_load: invalid neighbor method " + nm);
        ok(dpmd_set_neighbor_method(c, method), "set_neighbor_method");
        ok(dpmd_allocate_systems(c, S), "allocate_systems");
        ok(dpmd_allocate_particles(c, N), "allocate_particles");
        if (kind == DPMD_RIGID_BUMPY) ok(dpmd_allocate_vertices(c, Nv), "allocate_vertices");

        // load_static_from_hdf5_group (+ point / poly / rigid extras)
        set1<int>("/static/system_offset", DPMD_F_SYSTEM_OFFSET, DPMD_H5_I32);
        if (method == DPMD_NEIGHBOR_CELL) {
            set2<int>("/static/cell_dim", DPMD_F_CELL_DIM_X, DPMD_F_CELL_DIM_Y, DPMD_H5_I32);
            set1<double>("/static/neighbor_cutoff", DPMD_F_NEIGHBOR_CUTOFF, DPMD_H5_F64);
            set1<double>("/static/thresh2", DPMD_F_THRESH2, DPMD_H5_F64);
        }
        set1<double>("/static/e_interaction", DPMD_F_E_INTERACTION, DPMD_H5_F64);
        if (kind == DPMD_RIGID_BUMPY) {
            set1<double>("/static/vertex_mass", DPMD_F_VERTEX_MASS, DPMD_H5_F64);
            set1<double>("/static/vertex_rad", DPMD_F_VERTEX_RAD, DPMD_H5_F64);
            set1<int>("/static/n_vertices_per_particle", DPMD_F_N_VERTICES_PER_PARTICLE, DPMD_H5_I32);
            set1<double>("/static/mass", DPMD_F_MASS, DPMD_H5_F64);
            set1<double>("/static/moment_inertia", DPMD_F_MOMENT_INERTIA, DPMD_H5_F64);
            set1<double>("/static/area", DPMD_F_AREA, DPMD_H5_F64);
        } else {
            set1<double>("/static/mass", DPMD_F_MASS, DPMD_H5_F64);
            set1<double>("/static/rad", DPMD_F_RAD, DPMD_H5_F64);
        }
        load_group("static");
        if (location != "init" && location != "final" && location != "restart")
            throw ArgError("dpmd_h5_load: invalid location " + location);
        if (!exists(f, "/" + location)) throw ArgError("dpmd_h5_load: meta.h5 has no /" + location + " group");
        load_group(location);

        ok(dpmd_sync(c), "sync");
        check_int("/static/system_id", DPMD_F_SYSTEM_ID);
        check_int("/static/system_size", DPMD_F_SYSTEM_SIZE);
        if (method == DPMD_NEIGHBOR_CELL) check_int("/static/cell_system_start", DPMD_F_CELL_SYSTEM_START);
        if (kind == DPMD_RIGID_BUMPY) {
            check_int("/static/particle_offset", DPMD_F_PARTICLE_OFFSET);
            check_int("/static/vertex_particle_id", DPMD_F_VERTEX_PARTICLE_ID);
            check_int("/static/vertex_system_id", DPMD_F_VERTEX_SYSTEM_ID);
            check_int("/static/vertex_system_offset", DPMD_F_VERTEX_SYSTEM_OFFSET);
        }
        ok(dpmd_init_neighbors(c), "init_neighbors");
    }
};

}  // namespace

extern "C" {

int dpmd_h5_load(dpmd_ctx* ctx, const char* run_dir, const char* location) {
    return guard([&] {
        if (!ctx) throw ArgError("dpmd_h5_load: null context");
        const std::string meta = str(run_dir, "run directory") + "/meta.h5";
        if (!std::filesystem::exists(meta)) throw ArgError("dpmd_h5_load: meta file does not exist: " + meta);
        Id f(H5Fopen(meta.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
        need(f >= 0, "dpmd_h5_load: cannot open " + meta);
        Loader L{ctx, f, DPMD_DISK};
        L.run(str(location, "location"));
    });
}

}  // extern "C"
