// Host-only check of include/dpmd/output_manager.hpp.  Links nothing from the library:
//   output_format_test <dir>                 writes .npy files with io::detail::write_npy and
//                                            prints the steps each log policy logs;
//   output_format_test <dir> run <workers> [h5]
//                                            runs io::OutputManager for 30 steps over a
//                                            stand-in of the C ABI whose fields are closed-form
//                                            functions of the field id and the step, with the
//                                            NumPy container or (built with DPMD_WITH_H5 and
//                                            linked to libdpmd_h5) the HDF5 one.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "dpmd/output_manager.hpp"
#ifdef DPMD_WITH_H5
#include "dpmd/output_manager_h5.hpp"
#endif

// ---- stand-in C ABI: S = 2 systems, N = 12 particles, cell order = a fixed permutation -----
namespace {
int g_step = 0;
std::vector<std::string> g_calls;
const int kS = 2, kN = 12;
bool is_int(int f) { return (f >= 64 && f < 96) || f >= 112; }
int64_t length(int f) {
    if (f == DPMD_F_SYSTEM_OFFSET || f == DPMD_F_CELL_SYSTEM_START) return kS + 1;
    if ((f >= 32 && f < 80)) return kS;
    return kN;
}
}  // namespace

extern "C" {
const char* dpmd_last_error(const dpmd_ctx*) { return "stand-in error"; }
int dpmd_sizes(const dpmd_ctx*, int* S, int* N, int* Nv, int* C) {
    if (S) *S = kS;
    if (N) *N = kN;
    if (Nv) *Nv = 0;
    if (C) *C = 0;
    return DPMD_OK;
}
int dpmd_field(dpmd_ctx*, int f, void** p, int64_t* n) {
    if (p) *p = nullptr;
    if (n) *n = length(f);
    return DPMD_OK;
}
// Field f at step t, canonical entry k: ints f + k + t, doubles f + k / 4 + t / 1000.
int dpmd_get_field(dpmd_ctx*, int f, void* host, int64_t n, int order) {
    if (order != DPMD_ORDER_CANONICAL || n != length(f)) return DPMD_ERR_ARGUMENT;
    for (int64_t k = 0; k < n; ++k) {
        if (is_int(f)) ((int32_t*)host)[k] = (int32_t)(f + k + g_step);
        else ((double*)host)[k] = f + 0.25 * (double)k + 0.001 * g_step;
    }
    return DPMD_OK;
}
#define STAND_IN(name)                                           \
    int dpmd_compute_##name(dpmd_ctx*) {                         \
        g_calls.push_back(#name ":" + std::to_string(g_step));   \
        return DPMD_OK;                                          \
    }
STAND_IN(ke) STAND_IN(pe_total) STAND_IN(ke_total) STAND_IN(temperature) STAND_IN(pressure)
STAND_IN(packing_fraction) STAND_IN(fpower_total) STAND_IN(contacts) STAND_IN(n_contacts_total)
STAND_IN(overlaps) STAND_IN(stress_tensor)
#undef STAND_IN
}

struct StandInDisk {
    static constexpr int kKind = DPMD_DISK;
    dpmd_ctx* ctx() const { return nullptr; }
    md::NeighborMethod get_neighbor_method() const { return md::NeighborMethod::Cell; }
};

template <class Backend>
static int run(const std::string& d, int workers) {
    StandInDisk P;
    io::OutputManager<StandInDisk, Backend> om(P, d, workers);
    om.set_trajectory_fields({"pos", "pe_total", "system_offset"});
    om.set_trajectory_interval(5);
    om.set_restart_log_policy(io::LogMode::Logarithmic, 8);
    om.set_extra_final_fields({"packing_fraction"});
    om.set_queue_limit(2);
    g_step = 0;
    om.initialize();
    for (int k = 1; k <= 30; ++k) {
        g_step = k;
        om.step(k);
    }
    om.finalize();
    for (const auto& c : g_calls) std::cout << c << "\n";
    try {
        io::OutputManager<StandInDisk, Backend> bad(P, d + "_bad", 1);
        bad.set_trajectory_fields({"no_such_field"});
        bad.set_trajectory_interval(1);
        bad.enable_meta(false);
        bad.initialize();
        bad.step(1);
        return 1;
    } catch (const std::invalid_argument& e) {
        std::cout << "error: " << e.what() << "\n";
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 4 && std::string(argv[2]) == "run") return run<io::NpyBackend>(argv[1], std::atoi(argv[3]));
#ifdef DPMD_WITH_H5
    if (argc == 5 && std::string(argv[2]) == "run" && std::string(argv[4]) == "h5")
        return run<io::H5Backend>(argv[1], std::atoi(argv[3]));
#endif
    if (argc != 2) { std::cerr << "usage: output_format_test <dir> [run <workers>]\n"; return 2; }
    const std::string d = argv[1];
    io::HostField a{"<f8", 5, 2, std::vector<char>(5 * 2 * 8)};
    for (int k = 0; k < 10; ++k) { const double v = 0.5 * k - 1.0; std::memcpy(a.bytes.data() + 8 * k, &v, 8); }
    io::detail::write_npy(d + "/f8_5x2.npy", a);
    io::HostField b{"<i4", 7, 1, std::vector<char>(7 * 4)};
    for (int k = 0; k < 7; ++k) { const int32_t v = k * k - 3; std::memcpy(b.bytes.data() + 4 * k, &v, 4); }
    io::detail::write_npy(d + "/i4_7.npy", b);
    io::HostField e{"<f8", 0, 1, {}};
    io::detail::write_npy(d + "/empty.npy", e);
    io::LogPolicy lin{io::LogMode::Linear, 4, 0}, lg{io::LogMode::Logarithmic, 3, 3}, off{};
    std::cout << "linear:";
    for (int s = 0; s <= 20; ++s) if (lin.due(s)) std::cout << " " << s;
    std::cout << "\nlogarithmic:";
    for (int s = 0; s <= 20; s += 2) if (lg.due(s)) std::cout << " " << s;
    std::cout << "\ndisabled:";
    for (int s = 0; s <= 20; ++s) if (off.due(s)) std::cout << " " << s;
    std::cout << "\n";
    return 0;
}
