// sophia_run -- C++ host driver over the C ABI: the counterpart of the
// reference's per-GPU time loop ISPH_Calc (ISPH_Calc.cuh:238-577) for one
// GPU.  It reads a particle deck, uploads it, runs the step loop and prints
// the reference's periodic statistics line (ISPH_Calc.cuh:500-521).
//
// Deck format: the reference's tab-header particle file
// (function_init.cuh:550-676): the first line lists integer column codes
// separated by tabs (1 x, 2 y, 3 z, 4 ux, 5 uy, 6 uz, 7 m, 8 p_type, 9 h,
// 10 temp, 11 pres, 12 rho, 15 concn, 26 h_ref, 27 m_ref; other codes are
// read and ignored), then one value per column for each particle.
//
// Parameters: "key value" lines naming sph_params members, e.g.
//   kernel_type 3
//   ppe_solver 1
//   ppe_tol 1e-6
// plus the run controls dt, time, time_end, freq_output.
//
// usage: sophia_run deck.txt [params.txt] [--steps N] [--dump out.bin]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/urbansph.h"

static bool read_deck(const char* path, std::vector<sph_part1>& P) {
    std::ifstream f(path);
    if (!f) return false;
    std::string header;
    if (!std::getline(f, header)) return false;
    std::vector<int> codes;
    {
        std::stringstream ss(header);
        std::string tok;
        while (std::getline(ss, tok, '\t'))
            if (!tok.empty()) codes.push_back(std::atoi(tok.c_str()));
    }
    if (codes.empty()) return false;
    std::string v;
    while (true) {
        sph_part1 p;
        std::memset(&p, 0, sizeof(p));
        size_t k = 0;
        for (; k < codes.size(); k++) {
            if (!(f >> v)) break;
            double d = std::atof(v.c_str());
            switch (codes[k]) {
            case 1: p.x = (float)d; break;
            case 2: p.y = (float)d; break;
            case 3: p.z = (float)d; break;
            case 4: p.ux = (float)d; break;
            case 5: p.uy = (float)d; break;
            case 6: p.uz = (float)d; break;
            case 7: p.m = (float)d; break;
            case 8: p.p_type = std::atoi(v.c_str()); break;
            case 9: p.h = (float)d; break;
            case 10: p.temp = (float)d; break;
            case 11: p.pres = (float)d; break;
            case 12: p.rho = (float)d; break;
            case 15: p.concn = (float)d; break;
            case 26: p.h_ref = (float)d; break;
            case 27: p.m_ref = (float)d; break;
            default: break;
            }
        }
        if (k < codes.size()) break;
        p.i_type = 1;
        P.push_back(p);
    }
    return !P.empty();
}

struct Run {
    double dt = 1e-4, time = 0.0, time_end = 1e-3;
    int freq_output = 10;
};

static bool read_params(const char* path, sph_params& prm, Run& run) {
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
        std::fprintf(stderr, "usage: %s deck.txt [params.txt] [--steps N] [--dump out.bin]\n", argv[0]);
        return 2;
    }
    sph_params prm;
    sph_params_default(&prm);
    Run run;
    int max_steps = -1;
    const char* dump = nullptr;
    for (int a = 2; a < argc; a++) {
        if (!std::strcmp(argv[a], "--steps") && a + 1 < argc) max_steps = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--dump") && a + 1 < argc) dump = argv[++a];
        else if (!read_params(argv[a], prm, run)) { std::fprintf(stderr, "cannot read %s\n", argv[a]); return 2; }
    }
    std::vector<sph_part1> P;
    if (!read_deck(argv[1], P)) { std::fprintf(stderr, "cannot read deck %s\n", argv[1]); return 2; }
    const int n = (int)P.size();
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
