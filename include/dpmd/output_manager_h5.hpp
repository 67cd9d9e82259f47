// io::H5Backend -- the reference's HDF5 container for io::OutputManager (link libdpmd_h5).
//
//   io::OutputManager<md::disk::Disk, io::H5Backend> om(P, "run_dir", 2);
//
// writes run_dir/meta.h5 and run_dir/trajectory.h5 as the reference's OutputManager does
// (output_manager.hpp:169-311 for the files and groups, :704-932 for trajectories, :1000-1037
// for restart and final groups): /static holds class_name, neighbor_method, n_particles,
// n_systems, n_vertices and the static fields; /init and /final the state fields; /restart is
// rebuilt as /restart_tmp and moved into place with its step; trajectory.h5 holds /timestep
// (T, 1) and one (T, N) or (T, N, 2) dataset per field, frames of a changed length going to
// /ragged/<field>/step_<k>.  BaseParticle::load_from_hdf5 is dpmd_h5_load.
#pragma once

#include <filesystem>
#include <stdexcept>
#include <string>

#include "dpmd_h5.h"
#include "output_manager.hpp"

namespace io {

class H5Backend {
public:
    H5Backend() = default;
    H5Backend(const H5Backend&) = delete;
    H5Backend& operator=(const H5Backend&) = delete;
    ~H5Backend() {
        try {
            close();
        } catch (...) {
        }
    }

    bool open(const std::string& path, bool append, bool meta, bool traj) {
        namespace fs = std::filesystem;
        fs::create_directories(path);
        const std::string mp = path + "/meta.h5", tp = path + "/trajectory.h5";
        const bool existed = fs::exists(mp);
        if (meta) ok(dpmd_h5_open(mp.c_str(), DPMD_H5_READ_WRITE, &meta_));
        if (traj) {
            if (!append) fs::remove(tp);
            ok(dpmd_h5_open(tp.c_str(), DPMD_H5_READ_WRITE, &traj_));
            rows_ = 0;
            int e = 0;
            ok(dpmd_h5_exists(traj_, "/timestep", &e));
            if (e) {
                int nd = 0, t = 0;
                int64_t d[3];
                ok(dpmd_h5_shape(traj_, "/timestep", &nd, d, &t));
                rows_ = d[0];
            }
        }
        return existed;
    }
    void write_info(const RunInfo& r) {
        ok(dpmd_h5_write_string(meta_, "/static/class_name", r.class_name.c_str()));
        ok(dpmd_h5_write_string(meta_, "/static/neighbor_method", r.neighbor_method.c_str()));
        ok(dpmd_h5_write_int(meta_, "/static/n_particles", r.n_particles));
        ok(dpmd_h5_write_int(meta_, "/static/n_systems", r.n_systems));
        ok(dpmd_h5_write_int(meta_, "/static/n_vertices", r.n_vertices));
        ok(dpmd_h5_flush(meta_));
    }
    void write_group(const std::string& g, const Fields& f) {
        for (auto& [n, a] : f) write_array(meta_, "/" + g + "/" + n, a);
        ok(dpmd_h5_flush(meta_));
    }
    void write_restart(int step, const Fields& f) {
        ok(dpmd_h5_delete(meta_, "/restart_tmp"));
        ok(dpmd_h5_write_int(meta_, "/restart_tmp/step", step));
        for (auto& [n, a] : f) write_array(meta_, "/restart_tmp/" + n, a);
        ok(dpmd_h5_move(meta_, "/restart_tmp", "/restart"));
        ok(dpmd_h5_flush(meta_));
    }
    void write_final_step(int) {}   // the reference's /final holds fields only
    void write_frame(int step, const Fields& f) {
        const int s = step;
        ok(dpmd_h5_write_frame(traj_, "/timestep", DPMD_H5_I32, 1, 0, &s, rows_));
        for (auto& [n, a] : f) {
            const int st = a.rows > 0 ? dpmd_h5_write_frame(traj_, ("/" + n).c_str(), type(a), a.rows,
                                                            a.cols == 2 ? 2 : 0, a.bytes.data(), rows_)
                                      : DPMD_ERR_ARGUMENT;
            if (st == DPMD_ERR_ARGUMENT)   // ragged: a frame of another length
                write_array(traj_, "/ragged/" + n + "/step_" + std::to_string(step), a);
            else
                ok(st);
        }
        ++rows_;
    }
    void close() {
        if (meta_ > 0) ok(dpmd_h5_close(meta_));
        if (traj_ > 0) ok(dpmd_h5_close(traj_));
        meta_ = traj_ = 0;
    }

private:
    static void ok(int st) {
        if (st != DPMD_OK) {
            const std::string m = dpmd_h5_last_error();
            if (st == DPMD_ERR_ARGUMENT) throw std::invalid_argument(m);
            throw std::runtime_error(m);
        }
    }
    static int type(const HostField& a) {
        if (a.dtype == "<f8") return DPMD_H5_F64;
        if (a.dtype == "<i4") return DPMD_H5_I32;
        throw std::invalid_argument("H5Backend: unsupported element type " + a.dtype);
    }
    static void write_array(int64_t f, const std::string& path, const HostField& a) {
        ok(dpmd_h5_write_array(f, path.c_str(), type(a), a.rows, a.cols == 2 ? 2 : 0, a.bytes.data()));
    }
    int64_t meta_ = 0, traj_ = 0, rows_ = 0;
};

}  // namespace io
