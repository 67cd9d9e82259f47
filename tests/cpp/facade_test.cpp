// Drop-in check of the C++ facade: a driver written against the reference's API
// (md::disk::Disk, df::DeviceField1D/2D, md::integrators::*, md::routines::minimize_fire,
// nvt_rescale_run, jam_binary_search_pbc), in the call order of the reference's tests (tests/test_disk.cu:120-207,
// tests/fire_disk_pbc.cu).  It writes the final canonical state to a binary file so the
// test harness can compare it with the same sequence issued through the Python binding.
//
// With a second argument it then runs 60 more velocity-Verlet steps under io::OutputManager
// (the reference's include/utils/output_manager.hpp), writing its output directory there; with
// a third (built with DPMD_FACADE_H5 and linked to libdpmd_h5) a second manager writes the same
// run as meta.h5 / trajectory.h5 into that directory.
//
// usage: facade_test <out.bin> [<output dir> [<h5 output dir>]]
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <vector>

#include "dpmd/output_manager.hpp"
#ifdef DPMD_FACADE_H5
#include "dpmd/output_manager_h5.hpp"
#endif
#include "dpmd/routines.hpp"

int main(int argc, char** argv) {
    if (argc < 2 || argc > 4) { std::cerr << "usage: facade_test <out.bin> [<output dir> [<h5 output dir>]]\n"; return 2; }
    const int S = 2, n_per = 128, N = S * n_per;
    const double phi = 0.80;
    std::vector<double> rad(N), mass(N, 1.0), area(N);
    for (int s = 0; s < S; ++s)
        for (int k = 0; k < n_per; ++k) rad[s * n_per + k] = k < n_per / 2 ? 0.5 : 0.7;
    double a_sys = 0.0;
    for (int k = 0; k < n_per; ++k) a_sys += M_PI * rad[k] * rad[k];
    for (int i = 0; i < N; ++i) area[i] = M_PI * rad[i] * rad[i];
    const double L = std::sqrt(a_sys / phi);
    std::vector<int> off = {0, n_per, 2 * n_per}, sizes = {n_per, n_per}, sid(N);
    for (int i = 0; i < N; ++i) sid[i] = i / n_per;

    md::disk::Disk P;
    P.set_neighbor_method(md::NeighborMethod::Cell);
    P.allocate_systems(S);
    P.allocate_particles(N);
    P.neighbor_cutoff.from_host(std::vector<double>(S, 0.3));
    P.thresh2.from_host(std::vector<double>(S, 0.15 * 0.15));
    P.system_id.from_host(sid);
    P.system_size.from_host(sizes);
    P.system_offset.from_host(off);
    P.cell_dim.from_host(std::vector<int>(S, 0), std::vector<int>(S, 0));   // derived
    P.box_size.from_host(std::vector<double>(S, L), std::vector<double>(S, L));
    P.e_interaction.from_host(std::vector<double>(S, 1.0));
    P.rad.from_host(rad);
    P.mass.from_host(mass);
    P.area.from_host(area);
    P.set_rng_seed(11);
    P.stateless_rand_uniform_pos(0.0, L, 0.0, L, 0);
    P.sync_box();
    P.sync_system();
    P.sync_neighbors();
    P.sync_cells();
    P.sync_class_constants();
    P.init_neighbors();

    P.compute_forces();
    P.compute_pe_total();
    std::vector<double> pe0;
    P.pe_total.to_host(pe0);

    df::DeviceField1D<double> dt(S);
    dt.fill(1e-2);
    const int steps = md::routines::minimize_fire(P, dt, 300, 1e-16, 1e-16);

    P.set_rng_seed(3);
    P.stateless_rand_normal_vel(0.0, 1e-2, 0.0, 1e-2, 1);
    P.set_average_velocity(0.0, 0.0);
    P.set_temperature(1e-4);
    md::integrators::VelocityVerlet<md::disk::Disk> vv(P, 1e-2);
    vv.init();
    for (int k = 0; k < 50; ++k) vv.step();
    vv.run(50);
    md::integrators::DampedVelocityVerlet<md::disk::Disk> dvv(P, 1e-2, 0.2);
    dvv.init();
    for (int k = 0; k < 20; ++k) dvv.step();
    dvv.run(20);
    df::DeviceField1D<double> inc = P.per_system(2e-3), T = P.per_system(2e-4);
    md::routines::nvt_rescale_run(P, 60, &inc, 25, &T, 10);
    df::DeviceField1D<double> jdt(S);
    jdt.fill(1e-2);
    const int rounds = md::routines::jam_binary_search_pbc(P, jdt, 3, 500, 1e-10, 1e-10, 5e-3, 1e-4);
    P.compute_n_contacts_total();
    P.compute_pair_dist();
    P.compute_pair_forces();
    P.compute_hessian();
    std::vector<int> nct;
    P.n_contacts_total.to_host(nct);
    std::vector<double> hxx;
    P.hessian_xx.x.to_host(hxx);
    double trace = 0.0;
    for (double v : hxx) trace += v;
    P.compute_pe_total();
    P.compute_ke_total();
    P.compute_pressure();

    std::vector<double> x, y, vx, vy, pe, ke, pr;
    P.pos.to_host(x, y);
    P.vel.to_host(vx, vy);
    P.pe_total.to_host(pe);
    P.ke_total.to_host(ke);
    P.pressure.to_host(pr);
    std::vector<int> si;
    P.static_index.to_host(si);
    std::ofstream f(argv[1], std::ios::binary);
    auto put = [&](const std::vector<double>& v, bool canon) {
        for (int c = 0; c < (int)v.size(); ++c) {
            const double val = canon ? v[si[c]] : v[c];
            f.write((const char*)&val, sizeof(double));
        }
    };
    put(x, true); put(y, true); put(vx, true); put(vy, true);
    put(pe0, false); put(pe, false); put(ke, false); put(pr, false);
    std::cout << "facade_test: fire_steps=" << steps << " jam_rounds=" << rounds << " contacts=" << nct[0] << ","
              << nct[1] << " hessian_xx_sum=" << trace << " n_neighbors=" << P.n_neighbors()
              << " n_cells=" << P.n_cells() << " pe=" << pe[0] << "," << pe[1] << std::endl;

    if (argc >= 3) {   // the reference's OutputManager call pattern around an integrator loop
        auto configure = [](auto& om) {
            om.set_trajectory_fields({"pos", "vel", "pe_total", "ke_total", "temperature", "contacts",
                                      "n_contacts_total"});
            om.set_trajectory_interval(7);
            om.set_restart_interval(25);
            om.set_extra_init_fields({"packing_fraction"});
            om.set_extra_final_fields({"pe", "pressure", "stress_tensor_total_x"});
            om.set_queue_limit(3);
            om.initialize();
        };
        io::OutputManager<md::disk::Disk> om(P, argv[2], 2);
        configure(om);
#ifdef DPMD_FACADE_H5
        std::unique_ptr<io::OutputManager<md::disk::Disk, io::H5Backend>> oh;
        if (argc == 4) {
            oh = std::make_unique<io::OutputManager<md::disk::Disk, io::H5Backend>>(P, argv[3], 2);
            configure(*oh);
        }
#endif
        md::integrators::VelocityVerlet<md::disk::Disk> vv2(P, 1e-2);
        vv2.init();
        for (int k = 1; k <= 60; ++k) {
            vv2.step();
            om.step(k);
#ifdef DPMD_FACADE_H5
            if (oh) oh->step(k);
#endif
        }
        om.finalize();
#ifdef DPMD_FACADE_H5
        if (oh) oh->finalize();
#endif
        std::cout << "facade_test: output written to " << argv[2] << (argc == 4 ? " and " : "")
                  << (argc == 4 ? argv[3] : "") << std::endl;
    }
    return 0;
}
