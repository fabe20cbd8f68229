// sophia_run -- C++ host driver over the C ABI: the counterpart of the
// reference's per-GPU time loop ISPH_Calc (ISPH_Calc.cuh:238-577) for one
// GPU.  It reads the solver settings and a particle deck with the library's
// readers, uploads the deck, runs the step loop, prints the reference's
// periodic statistics line (ISPH_Calc.cuh:500-521) and, with --vtk, writes
// the reference's VTK frames every plot-output steps (ISPH_Calc.cuh:525-540).
//
// Settings: a solv.txt in the reference's keyword format (--solv), then any
// number of "key value" files naming sph_params members, e.g.
//   kernel_type 3
//   ppe_solver 1
//   ppe_tol 1e-6
// plus the run controls dt, time, time_end, freq_output; later files win.
//
// usage: sophia_run deck.txt [--solv solv.txt] [params.txt ...] [--steps N]
//                    [--vtk DIR] [--dump out.bin]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/urbansph.h"

static bool read_params(const char* path, sph_params& prm, sph_run_config& run) {
    std::ifstream f(path);
    if (!f) return false;
    std::map<std::string, int*> ints = {
        {"kernel_type", &prm.kernel_type}, {"time_type", &prm.time_type}, {"freq_filt", &prm.freq_filt},
        {"fp_solve", &prm.fp_solve}, {"fv_solve", &prm.fv_solve}, {"fg_solve", &prm.fg_solve},
        {"con_solve", &prm.con_solve}, {"boussinesq_solve", &prm.boussinesq_solve},
        {"turbulence_model", &prm.turbulence_model}, {"concn_solve", &prm.concn_solve},
        {"noslip_bc", &prm.noslip_bc}, {"artificial_viscosity", &prm.artificial_viscosity},
        {"most_mode", &prm.most_mode}, {"ppe_solver", &prm.ppe_solver}, {"precision", &prm.precision},
        {"ppe_maxit", &prm.ppe_maxit}, {"ppe_max_refine", &prm.ppe_max_refine}, {"freq_output", &run.freq_output}};
    std::map<std::string, double*> dbl = {
        {"kappa", &prm.kappa}, {"u_limit", &prm.u_limit}, {"mu_fluid", &prm.mu_fluid}, {"mu_wall", &prm.mu_wall},
        {"Cs_sgs", &prm.Cs_sgs}, {"L2", &prm.L2}, {"L3", &prm.L3}, {"L_left", &prm.L_left},
        {"L_right", &prm.L_right}, {"outlet_copy_dx", &prm.outlet_copy_dx}, {"most_z0", &prm.most_z0},
        {"most_z0h", &prm.most_z0h}, {"most_L", &prm.most_L}, {"ppe_tol", &prm.ppe_tol},
        {"ppe_inner_tol", &prm.ppe_inner_tol}, {"rho_air", &prm.rho_air}, {"dt", &run.dt}, {"time", &run.time},
        {"time_end", &run.time_end}};
    std::string k, v;
    while (f >> k >> v) {
        if (ints.count(k)) *ints[k] = std::atoi(v.c_str());
        else if (dbl.count(k)) *dbl[k] = std::atof(v.c_str());
        else std::fprintf(stderr, "unknown parameter %s (ignored)\n", k.c_str());
    }
    return true;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s deck.txt [--solv solv.txt] [params.txt ...] [--steps N] [--vtk DIR] "
                             "[--dump out.bin]\n", argv[0]);
        return 2;
    }
    sph_params prm;
    sph_params_default(&prm);
    sph_run_config run;
    std::memset(&run, 0, sizeof(run));
    run.dt = 1e-4;
    run.time_end = 1e-3;
    run.freq_output = 10;
    int max_steps = -1;
    const char* dump = nullptr;
    const char* vtk_dir = nullptr;
    for (int a = 2; a < argc; a++) {
        if (!std::strcmp(argv[a], "--steps") && a + 1 < argc) max_steps = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--dump") && a + 1 < argc) dump = argv[++a];
        else if (!std::strcmp(argv[a], "--vtk") && a + 1 < argc) vtk_dir = argv[++a];
        else if (!std::strcmp(argv[a], "--solv") && a + 1 < argc) {
            if (sph_read_solv(argv[++a], &prm, &run)) { std::fprintf(stderr, "cannot read %s\n", argv[a]); return 2; }
            if (run.unknown[0]) std::fprintf(stderr, "solv: unrecognised values for: %s\n", run.unknown);
        } else if (!read_params(argv[a], prm, run)) { std::fprintf(stderr, "cannot read %s\n", argv[a]); return 2; }
    }
    if (run.freq_output <= 0) run.freq_output = 10;
    int n = 0;
    if (sph_read_deck(argv[1], nullptr, 0, &n) || n == 0) {
        std::fprintf(stderr, "cannot read deck %s\n", argv[1]);
        return 2;
    }
    std::vector<sph_part1> P(n);
    if (sph_read_deck(argv[1], P.data(), n, &n)) { std::fprintf(stderr, "cannot read deck %s\n", argv[1]); return 2; }
    std::vector<const char*> labels;
    for (int q = 0; q < run.num_plot; q++) labels.push_back(run.plot[q]);
    if (vtk_dir && labels.empty())
        for (const char* l : {"p_type", "velocity", "pressure", "vis_t"}) labels.push_back(l);
    std::vector<sph_part1> O1;
    std::vector<sph_part2> O2;
    std::vector<sph_part3> O3;
    auto save_vtk = [&](sph_ctx* c, int count) {
        O1.resize(n);
        O2.resize(n);
        O3.resize(n);
        if (sph_download(c, O1.data(), O2.data(), O3.data(), n)) return false;
        char path[1024];
        std::snprintf(path, sizeof(path), "%s/fluid_%dstp.vtk", vtk_dir, count);
        return sph_write_vtk(path, O1.data(), O2.data(), O3.data(), n, labels.data(), (int)labels.size()) == 0;
    };
    sph_ctx* c = nullptr;
    int rc = sph_create(&prm, n, &c);
    if (rc) { std::fprintf(stderr, "sph_create: %d\n", rc); return 1; }
    if ((rc = sph_upload(c, P.data(), nullptr, nullptr, n))) {
        std::fprintf(stderr, "sph_upload: %s\n", sph_last_error(c));
        return 1;
    }
    sph_domain d;
    sph_domain_info(c, &d);
    std::printf("Total number of particles=%d\nNI=%d, NJ=%d, NK=%d\nCell Size(dcell) %f\n", n, d.NI, d.NJ, d.NK, d.dcell);
    int count = (int)std::floor(run.time / run.dt + 0.5);   // ISPH_Calc.cuh:256
    double t = run.time;
    auto t0 = std::chrono::steady_clock::now();
    int done = 0;
    while (true) {
        if ((rc = sph_step(c, count, run.dt, t))) {
            std::fprintf(stderr, "sph_step(count=%d): %s\n", count, sph_last_error(c));
            return 1;
        }
        t += run.dt;
        count++;
        done++;
        if (count % run.freq_output == 0) {
            float umax, drho, fmax;
            if ((rc = sph_stats(c, &umax, &drho, &fmax))) { std::fprintf(stderr, "sph_stats: %s\n", sph_last_error(c)); return 1; }
            double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::printf("Simulation time: %5.5f sec\nComputing time: %5.2f sec\n", t, wall);
            std::printf("%d\t tu_max=%5.2f\tftotal_max=%5.5f\n\n", count, umax, fmax);
            if (vtk_dir && !save_vtk(c, count)) { std::fprintf(stderr, "cannot write VTK in %s\n", vtk_dir); return 1; }
        }
        if (t >= run.time_end || (max_steps > 0 && done >= max_steps)) break;
    }
    if ((rc = sph_synchronize(c))) { std::fprintf(stderr, "%s\n", sph_last_error(c)); return 1; }
    double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("steps=%d particle-steps/s=%.4e\n", done, (double)n * done / wall);
    if (dump) {
        std::vector<sph_part1> out(n);
        sph_download(c, out.data(), nullptr, nullptr, n);
        FILE* fo = std::fopen(dump, "wb");
        if (fo) { std::fwrite(out.data(), sizeof(sph_part1), n, fo); std::fclose(fo); }
    }
    sph_destroy(c);
    return 0;
}
