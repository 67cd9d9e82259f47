// io::OutputManager -- asynchronous trajectory / snapshot writer over the C ABI.
//
// Reference interface: include/utils/output_manager.hpp:124-1085 (io::OutputManager<P>): named
// fields are captured to the host at chosen steps, put back in canonical particle order when
// the particles are cell-sorted, and written by background workers with a bounded queue; the
// run's initial state, static fields, periodic restart snapshots and final state go to their
// own groups.  Field names are the reference's (pos, vel, force, pe, ke, rad, mass, area,
// angle, angular_vel, torque, box_size, packing_fraction, pe_total, ke_total, temperature,
// pressure, n_contacts_total, ...); per-system totals are computed before they are captured,
// as the reference's registry preprocess hooks do.
//
// The container is a policy.  io::H5Backend (dpmd/output_manager_h5.hpp, libdpmd_h5) writes
// the reference's meta.h5 / trajectory.h5.  io::NpyBackend (here, no dependencies) writes the
// same groups as a directory of NumPy files:
//
//   <path>/meta/static/run.json                              class name, neighbor method, sizes
//   <path>/meta/{static,init,restart,final}/<field>.npy     one array per field (+ step.npy)
//   <path>/trajectory/<field>.bin, <field>.json              frames appended; dtype and frame shape
//   <path>/trajectory/timestep.bin                           int64 step of every frame
//   <path>/trajectory/ragged/<field>/step_<k>.npy            frames whose shape changed
//
// dpmd_amd/output.py reads both containers back.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iterator>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "dpmd.h"
#include "particles.hpp"

namespace io {

enum class LogMode { Disabled, Linear, Logarithmic };

// output_manager.hpp:24-31, :621-637: Linear logs every `interval` steps; Logarithmic logs at
// interval, 2 interval, ... advancing its next step past the current one.
struct LogPolicy {
    LogMode mode = LogMode::Disabled;
    int interval = 0;
    int next_log_step = 0;
    bool due(int step) {
        switch (mode) {
            case LogMode::Disabled: return false;
            case LogMode::Linear: return interval > 0 && step % interval == 0;
            case LogMode::Logarithmic:
                if (interval <= 0 || step < next_log_step) return false;
                do next_log_step += interval; while (next_log_step <= step);
                return true;
        }
        return false;
    }
};

// One captured field: raw bytes, element type and shape (rows x cols, cols = 1 or 2).
struct HostField {
    std::string dtype;        // "<f8" or "<i4"
    int64_t rows = 0, cols = 1;
    std::vector<char> bytes;
};

namespace detail {

// NumPy .npy format version 1.0: magic, header length, a Python-literal dict, padded with
// spaces to a multiple of 64 bytes and ended by a newline; then the C-order data.
inline void write_npy(const std::string& path, const HostField& f) {
    std::ostringstream h;
    h << "{'descr': '" << f.dtype << "', 'fortran_order': False, 'shape': (" << f.rows;
    if (f.cols > 1) h << ", " << f.cols;
    h << (f.cols > 1 ? "), }" : ",), }");
    std::string hdr = h.str();
    const size_t base = 10;   // magic (6) + version (2) + header length (2)
    size_t total = base + hdr.size() + 1;
    hdr.append((64 - total % 64) % 64, ' ');
    hdr.push_back('\n');
    const std::string tmp = path + ".tmp";
    {
        std::ofstream o(tmp, std::ios::binary);
        if (!o) throw std::runtime_error("OutputManager: cannot write " + tmp);
        const char magic[8] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0};
        o.write(magic, 8);
        const uint16_t len = (uint16_t)hdr.size();
        const char l2[2] = {(char)(len & 0xff), (char)(len >> 8)};
        o.write(l2, 2);
        o.write(hdr.data(), (std::streamsize)hdr.size());
        o.write(f.bytes.data(), (std::streamsize)f.bytes.size());
    }
    std::filesystem::rename(tmp, path);   // atomic replacement (restart snapshots)
}

inline std::string json_escape(const std::string& s) {
    std::string o;
    for (char ch : s) { if (ch == '"' || ch == '\\') o.push_back('\\'); o.push_back(ch); }
    return o;
}

// A frame header file: {"dtype": "<f8", "shape": [rows, cols]} as written below.
inline bool read_frame_header(const std::string& path, std::string& dtype, int64_t& rows, int64_t& cols) {
    std::ifstream in(path);
    if (!in) return false;
    std::string t((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    const size_t d = t.find("\"dtype\": \""), sh = t.find("\"shape\": [");
    if (d == std::string::npos || sh == std::string::npos) return false;
    dtype = t.substr(d + 10, 3);
    rows = std::strtoll(t.c_str() + sh + 10, nullptr, 10);
    const size_t comma = t.find(',', sh), close = t.find(']', sh);
    cols = (comma != std::string::npos && comma < close) ? std::strtoll(t.c_str() + comma + 1, nullptr, 10) : 1;
    return true;
}

}  // namespace detail

using Fields = std::vector<std::pair<std::string, HostField>>;

// What the reference records beside the static fields (output_manager.hpp:248-256).
struct RunInfo {
    std::string class_name, neighbor_method;
    int n_particles = 0, n_systems = 0, n_vertices = 0;
};

// Directory of NumPy files.  Calls come from one writer at a time.
class NpyBackend {
public:
    // Returns whether the run's metadata already existed (append mode then keeps it).
    bool open(const std::string& path, bool append, bool meta, bool traj) {
        namespace fs = std::filesystem;
        path_ = path;
        const bool existed = fs::exists(path_ + "/meta/static/run.json");
        if (meta) for (const char* g : {"static", "init", "final"}) fs::create_directories(path_ + "/meta/" + g);
        if (traj) {
            if (!append) fs::remove_all(path_ + "/trajectory");
            fs::create_directories(path_ + "/trajectory");
            for (auto& e : fs::directory_iterator(path_ + "/trajectory")) {
                if (e.path().extension() != ".json") continue;
                Shape s;
                if (detail::read_frame_header(e.path().string(), s.dtype, s.rows, s.cols))
                    shapes_[e.path().stem().string()] = s;
            }
        }
        return existed;
    }
    void write_info(const RunInfo& r) {
        std::ofstream o(path_ + "/meta/static/run.json");
        o << "{\"class_name\": \"" << detail::json_escape(r.class_name) << "\", \"neighbor_method\": \""
          << detail::json_escape(r.neighbor_method) << "\", \"n_particles\": " << r.n_particles
          << ", \"n_systems\": " << r.n_systems << ", \"n_vertices\": " << r.n_vertices << "}\n";
    }
    void write_group(const std::string& g, const Fields& f) {
        for (auto& [n, a] : f) detail::write_npy(path_ + "/meta/" + g + "/" + n + ".npy", a);
    }
    // Written beside and then swapped in, as output_manager.hpp:1000-1011 does with /restart_tmp.
    void write_restart(int step, const Fields& f) {
        namespace fs = std::filesystem;
        const std::string tmp = path_ + "/meta/restart_tmp", dst = path_ + "/meta/restart";
        fs::remove_all(tmp);
        fs::create_directories(tmp);
        for (auto& [n, a] : f) detail::write_npy(tmp + "/" + n + ".npy", a);
        detail::write_npy(tmp + "/step.npy", step_field(step));
        fs::remove_all(dst);
        fs::rename(tmp, dst);
    }
    void write_final_step(int step) { detail::write_npy(path_ + "/meta/final/step.npy", step_field(step)); }
    void write_frame(int step, const Fields& f) {
        namespace fs = std::filesystem;
        const std::string d = path_ + "/trajectory/";
        for (auto& [n, a] : f) {
            auto it = shapes_.find(n);
            if (it == shapes_.end()) {
                Shape s{a.dtype, a.rows, a.cols};
                std::ofstream h(d + n + ".json");
                h << "{\"dtype\": \"" << a.dtype << "\", \"shape\": [" << a.rows;
                if (a.cols > 1) h << ", " << a.cols;
                h << "]}\n";
                it = shapes_.emplace(n, s).first;
            }
            if (it->second.dtype != a.dtype || it->second.rows != a.rows || it->second.cols != a.cols) {
                fs::create_directories(d + "ragged/" + n);
                detail::write_npy(d + "ragged/" + n + "/step_" + std::to_string(step) + ".npy", a);
                continue;
            }
            std::ofstream o(d + n + ".bin", std::ios::binary | std::ios::app);
            o.write(a.bytes.data(), (std::streamsize)a.bytes.size());
        }
        std::ofstream ts(d + "timestep.bin", std::ios::binary | std::ios::app);
        const int64_t s = step;
        ts.write((const char*)&s, 8);
    }
    void close() {}

private:
    struct Shape {
        std::string dtype;
        int64_t rows = 0, cols = 1;
    };
    static HostField step_field(int step) {
        HostField st{"<i8", 1, 1, std::vector<char>(8)};
        const int64_t s = step;
        std::copy_n((const char*)&s, 8, st.bytes.data());
        return st;
    }
    std::string path_;
    std::map<std::string, Shape> shapes_;
};

template <class ParticleT, class Backend = NpyBackend>
class OutputManager {
public:
    OutputManager(ParticleT& particles, const std::string& path, int max_workers = 1, bool append_mode = false)
        : p_(particles), path_(path), max_workers_(std::max(1, max_workers)), append_mode_(append_mode) {}
    ~OutputManager() {
        try {
            stop_workers();
            backend_.close();
        } catch (...) {
        }
    }
    Backend& backend() { return backend_; }
    OutputManager(const OutputManager&) = delete;
    OutputManager& operator=(const OutputManager&) = delete;

    // ---- policy (output_manager.hpp:134-157) -----------------------------------------------
    void set_trajectory_fields(const std::vector<std::string>& n) { traj_names_ = n; }
    void set_extra_init_fields(const std::vector<std::string>& n) { extra_init_ = n; }
    void set_extra_final_fields(const std::vector<std::string>& n) { extra_final_ = n; }
    void set_extra_static_fields(const std::vector<std::string>& n) { extra_static_ = n; }
    void set_extra_restart_fields(const std::vector<std::string>& n) { extra_restart_ = n; }
    void set_trajectory_log_policy(LogMode m, int interval) { set_policy(traj_, m, interval); }
    void set_restart_log_policy(LogMode m, int interval) { set_policy(restart_, m, interval); }
    void set_console_log_policy(LogMode m, int interval) { set_policy(console_, m, interval); }
    void set_trajectory_interval(int v) { set_trajectory_log_policy(LogMode::Linear, v); }
    void set_restart_interval(int v) { set_restart_log_policy(LogMode::Linear, v); }
    void set_console_interval(int v) { set_console_log_policy(LogMode::Linear, v); }
    void set_console_fields(const std::vector<std::string>& n) { console_names_ = n; }
    void set_queue_limit(size_t tasks) { queue_limit_ = std::max<size_t>(1, tasks); queue_limit_set_ = true; }
    void enable_meta(bool on) { enable_meta_ = on; }
    void enable_trajectory(bool on) { enable_traj_ = on; }
    void enable_console(bool on) { enable_console_ = on; }

    // ---- run (output_manager.hpp:169-375) -------------------------------------------------
    void initialize() {
        const bool existed = backend_.open(path_, append_mode_, enable_meta_, enable_traj_);
        if (enable_meta_) {
            backend_.write_info(run_info());
            if (!(append_mode_ && existed)) {
                backend_.write_group("static", capture_all(with(static_names(), extra_static_)));
                backend_.write_group("init", capture_all(with(state_names(), extra_init_)));
            }
        }
        start_workers();
    }

    void step(int step_num) {
        current_step_ = step_num;
        if (enable_console_ && !console_names_.empty() && console_.due(step_num)) {
            std::ostringstream os;
            os << "step " << step_num;
            for (const auto& n : console_names_) {
                HostField f = capture(n);
                os << "  " << n << "=" << first_values(f);
            }
            std::printf("%s\n", os.str().c_str());
        }
        if (enable_traj_ && !traj_names_.empty() && traj_.due(step_num)) {
            Task t{Kind::Trajectory, step_num, {}};
            for (const auto& n : traj_names_) t.fields.emplace_back(n, capture(n));
            enqueue(std::move(t));
        }
        if (enable_meta_ && restart_.due(step_num)) enqueue_group(Kind::Restart, all_restart_names());
    }

    void finalize() {
        if (enable_meta_) {
            enqueue_group(Kind::Restart, all_restart_names());
            enqueue_group(Kind::Final, with(state_names(), extra_final_));
        }
        stop_workers();
        backend_.close();
    }

private:
    enum class Kind { Trajectory, Restart, Final };
    struct Task {
        Kind kind;
        int step;
        std::vector<std::pair<std::string, HostField>> fields;
        uint64_t seq = 0;    // capture order; files are written in this order
    };

    // ---- field registry -----------------------------------------------------------------
    struct Spec {
        int fx, fy;           // C ABI field ids (fy < 0: one column)
        bool is_int;
        std::function<void()> pre;
    };
    static constexpr bool kRigid = ParticleT::kKind == DPMD_RIGID_BUMPY;

    // Default groups (base_particle.hpp:746-760, base_point_particle.hpp:246-257,
    // base_poly_particle.hpp:414-425, rigid_bumpy.cu:1887-1893).
    std::vector<std::string> static_names() const {
        std::vector<std::string> n{"system_id", "system_size", "system_offset", "box_size"};
        if (p_.get_neighbor_method() == md::NeighborMethod::Cell)
            n = with(n, {"cell_dim", "cell_system_start", "neighbor_cutoff", "thresh2"});
        if (kRigid)
            return with(n, {"e_interaction", "vertex_mass", "vertex_rad", "n_vertices_per_particle", "particle_offset",
                            "vertex_particle_id", "vertex_system_id", "vertex_system_offset", "mass", "moment_inertia",
                            "area"});
        return with(n, {"e_interaction", "mass", "rad", "area"});
    }
    std::vector<std::string> state_names() const {
        if (kRigid) return {"vertex_pos", "pos", "vel", "force", "torque", "angular_vel", "angle"};
        return {"pos", "vel", "force"};
    }
    // A restart snapshot holds what restores the run: static and state fields, the box (which
    // compression changes) and the caller's extras.
    std::vector<std::string> all_restart_names() const {
        return with(with(static_names(), state_names()), with(extra_static_, extra_restart_));
    }
    static std::vector<std::string> with(std::vector<std::string> a, const std::vector<std::string>& b) {
        for (const auto& x : b)
            if (std::find(a.begin(), a.end(), x) == a.end()) a.push_back(x);
        return a;
    }

    // Field name -> C ABI field ids and the computation that must run before it is read
    // (the registry and its preprocess hooks, base_particle.hpp:764-972,
    // rigid_bumpy.cu:1895-1991).
    Spec spec(const std::string& n) {
        dpmd_ctx* c = p_.ctx();
        auto ok = [c](int st) { df::detail::check(st, c); };
        struct Ids { int fx, fy; bool is_int; };
        static const std::map<std::string, Ids> ids = {
            {"pos", {DPMD_F_POS_X, DPMD_F_POS_Y, false}}, {"vel", {DPMD_F_VEL_X, DPMD_F_VEL_Y, false}},
            {"force", {DPMD_F_FORCE_X, DPMD_F_FORCE_Y, false}}, {"pe", {DPMD_F_PE, -1, false}},
            {"ke", {DPMD_F_KE, -1, false}}, {"rad", {DPMD_F_RAD, -1, false}}, {"mass", {DPMD_F_MASS, -1, false}},
            {"area", {DPMD_F_AREA, -1, false}}, {"angle", {DPMD_F_ANGLE, -1, false}},
            {"angular_vel", {DPMD_F_ANGULAR_VEL, -1, false}}, {"torque", {DPMD_F_TORQUE, -1, false}},
            {"moment_inertia", {DPMD_F_MOMENT_INERTIA, -1, false}},
            {"stress_tensor_x", {DPMD_F_STRESS_XX, DPMD_F_STRESS_XY, false}},
            {"stress_tensor_y", {DPMD_F_STRESS_YX, DPMD_F_STRESS_YY, false}},
            {"overlaps", {DPMD_F_OVERLAP, -1, false}}, {"contact_count", {DPMD_F_CONTACT_COUNT, -1, false}},
            {"box_size", {DPMD_F_BOX_SIZE_X, DPMD_F_BOX_SIZE_Y, false}},
            {"e_interaction", {DPMD_F_E_INTERACTION, -1, false}}, {"neighbor_cutoff", {DPMD_F_NEIGHBOR_CUTOFF, -1, false}},
            {"thresh2", {DPMD_F_THRESH2, -1, false}}, {"packing_fraction", {DPMD_F_PACKING_FRACTION, -1, false}},
            {"pe_total", {DPMD_F_PE_TOTAL, -1, false}}, {"ke_total", {DPMD_F_KE_TOTAL, -1, false}},
            {"fpower_total", {DPMD_F_FPOWER_TOTAL, -1, false}}, {"pressure", {DPMD_F_PRESSURE, -1, false}},
            {"temperature", {DPMD_F_TEMPERATURE, -1, false}},
            {"stress_tensor_total_x", {DPMD_F_STRESS_TOTAL_XX, DPMD_F_STRESS_TOTAL_XY, false}},
            {"stress_tensor_total_y", {DPMD_F_STRESS_TOTAL_YX, DPMD_F_STRESS_TOTAL_YY, false}},
            {"vertex_pos", {DPMD_F_VERTEX_POS_X, DPMD_F_VERTEX_POS_Y, false}},
            {"vertex_force", {DPMD_F_VERTEX_FORCE_X, DPMD_F_VERTEX_FORCE_Y, false}},
            {"vertex_pe", {DPMD_F_VERTEX_PE, -1, false}}, {"vertex_rad", {DPMD_F_VERTEX_RAD, -1, false}},
            {"vertex_mass", {DPMD_F_VERTEX_MASS, -1, false}},
            {"system_offset", {DPMD_F_SYSTEM_OFFSET, -1, true}}, {"system_size", {DPMD_F_SYSTEM_SIZE, -1, true}},
            {"cell_dim", {DPMD_F_CELL_DIM_X, DPMD_F_CELL_DIM_Y, true}},
            {"cell_system_start", {DPMD_F_CELL_SYSTEM_START, -1, true}},
            {"n_contacts_total", {DPMD_F_N_CONTACTS_TOTAL, -1, true}}, {"system_id", {DPMD_F_SYSTEM_ID, -1, true}},
            {"static_index", {DPMD_F_STATIC_INDEX, -1, true}}, {"contacts", {DPMD_F_CONTACTS, -1, true}},
            {"particle_offset", {DPMD_F_PARTICLE_OFFSET, -1, true}},
            {"n_vertices_per_particle", {DPMD_F_N_VERTICES_PER_PARTICLE, -1, true}},
            {"vertex_particle_id", {DPMD_F_VERTEX_PARTICLE_ID, -1, true}},
            {"vertex_system_id", {DPMD_F_VERTEX_SYSTEM_ID, -1, true}},
            {"vertex_system_offset", {DPMD_F_VERTEX_SYSTEM_OFFSET, -1, true}}};
        std::function<void()> pre;
        if (n == "ke") pre = [=] { ok(dpmd_compute_ke(c)); };
        else if (n == "pe_total") pre = [=] { ok(dpmd_compute_pe_total(c)); };
        else if (n == "ke_total") pre = [=] { ok(dpmd_compute_ke_total(c)); };
        else if (n == "temperature") pre = [=] { ok(dpmd_compute_temperature(c)); };
        else if (n == "pressure") pre = [=] { ok(dpmd_compute_pressure(c)); };
        else if (n == "packing_fraction") pre = [=] { ok(dpmd_compute_packing_fraction(c)); };
        else if (n == "fpower_total") pre = [=] { ok(dpmd_compute_fpower_total(c)); };
        else if (n == "contacts") pre = [=] { ok(dpmd_compute_contacts(c)); };
        else if (n == "n_contacts_total") pre = [=] { ok(dpmd_compute_n_contacts_total(c)); };
        else if (n == "overlaps" || n == "contact_count") pre = [=] { ok(dpmd_compute_overlaps(c)); };
        else if (n.rfind("stress_tensor", 0) == 0) pre = [=] { ok(dpmd_compute_stress_tensor(c)); };
        auto d = ids.find(n);
        if (d == ids.end()) throw std::invalid_argument("OutputManager: unknown field '" + n + "'");
        return Spec{d->second.fx, d->second.fy, d->second.is_int, pre};
    }

    // Device -> host copy of one field, canonical order (the reindexing of
    // output_manager.hpp:532-574 happens in dpmd_get_field).
    HostField capture(const std::string& n) {
        dpmd_ctx* c = p_.ctx();
        Spec sp = spec(n);
        if (sp.pre) sp.pre();
        HostField f;
        f.dtype = sp.is_int ? "<i4" : "<f8";
        const int es = sp.is_int ? 4 : 8;
        int64_t len = 0;
        df::detail::check(dpmd_field(c, sp.fx, nullptr, &len), c);
        f.rows = len;
        f.cols = sp.fy >= 0 ? 2 : 1;
        f.bytes.resize((size_t)len * es * f.cols);
        std::vector<char> col((size_t)len * es);
        for (int k = 0; k < f.cols; ++k) {
            const int fid = k == 0 ? sp.fx : sp.fy;
            df::detail::check(dpmd_get_field(c, fid, col.data(), len, DPMD_ORDER_CANONICAL), c);
            for (int64_t r = 0; r < len; ++r)
                std::copy(col.data() + r * es, col.data() + (r + 1) * es, f.bytes.data() + (r * f.cols + k) * es);
        }
        return f;
    }
    static std::string first_values(const HostField& f) {
        std::ostringstream os;
        const int64_t n = std::min<int64_t>(f.rows * f.cols, 4);
        for (int64_t k = 0; k < n; ++k) {
            if (k) os << ",";
            if (f.dtype == "<f8") { double v; std::copy_n(f.bytes.data() + k * 8, 8, (char*)&v); os << v; }
            else { int32_t v; std::copy_n(f.bytes.data() + k * 4, 4, (char*)&v); os << v; }
        }
        return os.str();
    }

    Fields capture_all(const std::vector<std::string>& names) {
        Fields f;
        for (const auto& n : names) f.emplace_back(n, capture(n));
        return f;
    }
    RunInfo run_info() {
        RunInfo r;
        r.class_name = kRigid ? "RigidBumpy" : "Disk";
        r.neighbor_method = p_.get_neighbor_method() == md::NeighborMethod::Cell ? "Cell" : "Naive";
        df::detail::check(dpmd_sizes(p_.ctx(), &r.n_systems, &r.n_particles, &r.n_vertices, nullptr), p_.ctx());
        return r;
    }
    void enqueue_group(Kind k, const std::vector<std::string>& names) {
        Task t{k, current_step_, {}};
        for (const auto& n : names) t.fields.emplace_back(n, capture(n));
        enqueue(std::move(t));
    }

    // ---- workers with blocking back-pressure (output_manager.hpp:640-702): one worker means
    // the caller writes synchronously. ------------------------------------------------------
    void start_workers() {
        stop_ = false;
        if (max_workers_ <= 1) return;
        if (!queue_limit_set_) queue_limit_ = (size_t)max_workers_;
        for (int w = 0; w < max_workers_; ++w) workers_.emplace_back([this] { worker(); });
    }
    void stop_workers() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) if (t.joinable()) t.join();
        workers_.clear();
        if (error_) std::rethrow_exception(error_);
    }
    void enqueue(Task&& t) {
        t.seq = next_seq_++;
        std::unique_lock<std::mutex> lk(mu_);
        space_.wait(lk, [&] { return queue_.size() < queue_limit_ || workers_.empty(); });
        if (workers_.empty()) { lk.unlock(); process(t); return; }
        queue_.push_back(std::move(t));
        cv_.notify_one();
    }
    void worker() {
        for (;;) {
            Task t;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
                if (queue_.empty()) return;
                t = std::move(queue_.front());
                queue_.pop_front();
            }
            space_.notify_all();
            try { process(t); } catch (...) { std::lock_guard<std::mutex> lk(mu_); error_ = std::current_exception(); }
        }
    }
    // Workers take tasks in capture order but may finish out of it, so each waits for its turn:
    // frames land in step order and a newer restart snapshot is never overwritten by an older one.
    void process(Task& t) {
        {
            std::unique_lock<std::mutex> lk(io_);
            turn_.wait(lk, [&] { return written_ == t.seq; });
        }
        try {
            write(t);
        } catch (...) {
            finish_turn();
            throw;
        }
        finish_turn();
    }
    void finish_turn() {
        {
            std::lock_guard<std::mutex> lk(io_);
            ++written_;
        }
        turn_.notify_all();
    }
    void write(Task& t) {
        if (t.kind == Kind::Trajectory) {
            backend_.write_frame(t.step, t.fields);
        } else if (t.kind == Kind::Restart) {
            backend_.write_restart(t.step, t.fields);
        } else {
            backend_.write_group("final", t.fields);
            backend_.write_final_step(t.step);
        }
    }
    void set_policy(LogPolicy& p, LogMode m, int interval) {
        p.mode = interval > 0 ? m : LogMode::Disabled;
        p.interval = std::max(0, interval);
        p.next_log_step = m == LogMode::Logarithmic ? p.interval : 0;
    }

    ParticleT& p_;
    std::string path_;
    int max_workers_;
    bool append_mode_;
    bool enable_meta_ = true, enable_traj_ = true, enable_console_ = false;
    LogPolicy traj_, restart_, console_;
    std::vector<std::string> traj_names_, extra_init_, extra_final_, extra_static_, extra_restart_, console_names_;
    int current_step_ = 0;
    size_t queue_limit_ = 1;
    bool queue_limit_set_ = false;
    std::vector<std::thread> workers_;
    std::deque<Task> queue_;
    std::mutex mu_, io_;
    std::condition_variable cv_, space_, turn_;
    uint64_t next_seq_ = 0, written_ = 0;
    bool stop_ = false;
    std::exception_ptr error_;
    Backend backend_;
};

}  // namespace io
