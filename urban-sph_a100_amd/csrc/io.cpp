// Input and output formats of SOPHIA runs, host C++ on the public records:
//
//   read_solv_input            function_init.cuh:166-463    -> sph_read_solv
//   read_input (+ count)       function_init.cuh:483-676    -> sph_read_deck
//   save_vtk_bin_single_flag   functions_OUTPUT.cuh:121-616 -> sph_write_vtk
//   read_table                 function_init.cuh:724-970    -> sph_read_tables
//   interp2                    functions_PROP.cuh:8-40      -> sph_table_interp
//
// solv.txt is a whitespace-separated token stream: a keyword, a fixed number
// of filler tokens (":", "type", ...), then the value.  The keyword table
// gives each keyword's filler count as the reference reads it.
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <string>
#include <utility>
#include <vector>

#include "../../include/urbansph.h"

namespace {

using Opts = std::initializer_list<std::pair<const char*, int>>;

struct Solv {
    std::vector<std::string> tok;
    size_t k = 0;
    sph_run_config* run;

    // the value after `fillers` tokens; false at the end of the file
    bool value(int fillers, std::string& v) {
        k += fillers;
        if (k >= tok.size()) return false;
        v = tok[k++];
        return true;
    }
    void unknown(const std::string& key) {
        size_t used = std::strlen(run->unknown);
        if (used + key.size() + 2 <= sizeof(run->unknown)) {
            if (used) std::strcat(run->unknown, " ");
            std::strcat(run->unknown, key.c_str());
        }
    }
    // option keyword: the value names one of `opts`, else the target keeps
    // its value and the keyword is listed as unknown
    bool option(const std::string& key, int fillers, int* target, Opts opts) {
        std::string v;
        if (!value(fillers, v)) return false;
        for (auto& o : opts)
            if (v == o.first) { *target = o.second; return true; }
        unknown(key);
        return true;
    }
    bool integer(int fillers, int* target) {
        std::string v;
        if (!value(fillers, v)) return false;
        *target = std::atoi(v.c_str());
        return true;
    }
    bool real(int fillers, double* target) {
        std::string v;
        if (!value(fillers, v)) return false;
        *target = std::atof(v.c_str());
        return true;
    }
};

const Opts YES_NO = {{"YES", 1}, {"NO", 0}};

}  // namespace

extern "C" int sph_read_solv(const char* path, sph_params* prm, sph_run_config* run) {
    if (!path || !prm || !run) return SPH_E_ARG;
    std::ifstream f(path);
    if (!f) return SPH_E_ARG;
    std::memset(run, 0, sizeof(*run));
    run->dt = 1e-4;
    run->time_end = 1e-3;
    run->freq_output = 10;
    // cleared before reading, as function_init.cuh:168-181 does
    prm->fp_solve = prm->fv_solve = prm->fg_solve = prm->con_solve = 0;
    prm->boussinesq_solve = prm->kgc_solve = prm->turbulence_model = prm->concn_solve = 0;
    Solv S{{}, 0, run};
    for (std::string w; f >> w;) S.tok.push_back(w);
    bool more = true;
    while (more && S.k < S.tok.size()) {
        const std::string key = S.tok[S.k++];
        if (key == "solver_type") more = S.option(key, 1, &run->solver_type, {{"WCSPH", 0}, {"ISPH", 1}});
        else if (key == "dimension(1/2/3):") more = S.integer(0, &prm->dim);
        else if (key == "property_table(YES/NO):") more = S.option(key, 0, &run->prop_table, YES_NO);
        else if (key == "atmospheric-stability")
            more = S.option(key, 1, &run->atmos_stab, {{"Unstable", 0}, {"Neutral", 1}, {"Stable", 2}});
        else if (key == "terrain-type")
            more = S.option(key, 1, &run->terrain_type,
                            {{"Flat", 0}, {"Grass", 1}, {"Forest", 2}, {"City", 3}, {"Ocean", 4}});
        else if (key == "kernel")
            more = S.option(key, 2, &prm->kernel_type,
                            {{"Quartic", SPH_QUARTIC}, {"Gaussian", SPH_GAUSSIAN}, {"Quintic", SPH_QUINTIC},
                             {"Wendland2", SPH_WENDLAND2}, {"Wendland4", SPH_WENDLAND4}, {"Wendland6", SPH_WENDLAND6}});
        else if (key == "time")
            more = S.option(key, 2, &prm->time_type, {{"Euler", SPH_EULER}, {"Predictor_Corrector", SPH_PRE_COR}});
        else if (key == "structure") more = S.option(key, 1, &run->structure, YES_NO);
        else if (key == "open")
            more = S.option(key, 2, &run->open_boundary, {{"OPEN", 1}, {"PERIODIC", 2}, {"NO", 0}});
        else if (key == "pressure-force") more = S.option(key, 1, &prm->fp_solve, YES_NO);
        else if (key == "viscous-force")
            more = S.option(key, 1, &prm->fv_solve,
                            {{"NO", SPH_FV_NONE}, {"Cleary", SPH_FV_CLEARY}, {"Monaghan", SPH_FV_MONAGHAN}, {"Violeau", 3}});
        else if (key == "turbulence-model")
            more = S.option(key, 1, &prm->turbulence_model,
                            {{"Laminar", SPH_LAMINAR}, {"k-lm", 1}, {"k-e", 2}, {"HB", 3}, {"SSM", SPH_SSM}, {"DDF", SPH_DDF}});
        else if (key == "gravitational-force") more = S.option(key, 1, &prm->fg_solve, YES_NO);
        else if (key == "Conduction(YES/NO):") more = S.option(key, 0, &prm->con_solve, YES_NO);
        else if (key == "Boussinesq-natural-convection") more = S.option(key, 1, &prm->boussinesq_solve, YES_NO);
        else if (key == "Concentration-diffusion(YES/NO):") more = S.option(key, 0, &prm->concn_solve, YES_NO);
        else if (key == "kernel-gradient-correction")
            more = S.option(key, 1, &prm->kgc_solve, {{"NO", 0}, {"KGC", 1}, {"FPM", 2}, {"DFPM", 3}, {"KGF", 4}});
        else if (key == "switch_ptype") more = S.option(key, 1, &run->switch_ptype, YES_NO);
        else if (key == "noslip_boundary") more = S.option(key, 1, &prm->noslip_bc, YES_NO);
        else if (key == "APR_solve")
            more = S.option(key, 1, &run->apr_solve, {{"Cell", 1}, {"Switch", 2}, {"Block", 3}, {"NO", 0}});
        else if (key == "Block_configuration")
            more = S.integer(1, &run->block_config[0]) && S.integer(1, &run->block_config[1]) &&
                   S.integer(1, &run->block_config[2]);
        else if (key == "Variable")
            // h_change: 1 Abrupt (children get h/2), 0 NO (children keep h);
            // the reference's 2 (Smooth) takes neither branch of the split
            more = S.option(key, 1, &prm->apr_h_change, {{"Abrupt", 1}, {"Smooth", 2}, {"NO", 0}});
        else if (key == "Buffersize:") more = S.real(0, &run->buffersize);
        else if (key == "reference-pressure") more = S.real(1, &run->p_ref);
        else if (key == "kappa:") more = S.real(0, &prm->kappa);
        else if (key == "minimum-iteration:") more = S.integer(0, &run->min_iteration);
        else if (key == "maximum-iteration:") more = S.integer(0, &run->max_iteration);
        else if (key == "pressure-convergence") more = S.real(1, &run->dp_th);
        else if (key == "density-convergence") more = S.real(1, &run->drho_th);
        else if (key == "pressure-relaxation") more = S.real(1, &run->p_relaxation);
        else if (key == "time-step") more = S.real(1, &run->dt);
        else if (key == "start-time") more = S.real(1, &run->time);
        else if (key == "end-time") more = S.real(1, &run->time_end);
        else if (key == "neighbor") more = S.option(key, 3, &run->nb_cell_type, {{"3X3", 0}, {"5X5", 1}});
        else if (key == "timestep") more = S.option(key, 2, &run->timestep_update, YES_NO);
        else if (key == "structure-timestep") more = S.real(1, &run->dt_structure);
        else if (key == "filtering") more = S.integer(1, &prm->freq_filt);
        else if (key == "temperature-filtering") more = S.integer(1, &run->freq_temp);
        else if (key == "plot-output") more = S.integer(1, &run->freq_output);
        else if (key == "APR_frequency:") more = S.integer(0, &run->freq_apr);
        else if (key == "h-reset") more = S.integer(1, &run->freq_h_reset);
        else if (key == "BAPR_frequency:") more = S.integer(0, &run->freq_bapr);
        else if (key == "plot-variables:") {
            int np = 0;
            more = S.integer(0, &np);
            run->num_plot = 0;
            for (int q = 0; more && q < np; q++) {
                std::string v;
                more = S.value(0, v);
                if (more && run->num_plot < SPH_MAX_PLOT) {
                    std::snprintf(run->plot[run->num_plot], sizeof(run->plot[0]), "%s", v.c_str());
                    run->num_plot++;
                }
            }
        } else if (key == "velocity-limit") more = S.real(1, &prm->u_limit);
    }
    return 0;
}

extern "C" int sph_read_deck(const char* path, sph_part1* P1, int cap, int* n_out) {
    if (!path || !n_out || (P1 && cap < 0)) return SPH_E_ARG;
    *n_out = 0;
    std::ifstream f(path);
    if (!f) return SPH_E_ARG;
    std::string header;
    if (!std::getline(f, header)) return SPH_E_ARG;
    // the header splits at tabs with empty fields skipped and the line end
    // kept in the last field, so a trailing tab before the newline adds a
    // column whose code is 0 (read and skipped), as strtok on the line does
    header += '\n';
    std::vector<int> codes;
    for (size_t a = 0; a < header.size();) {
        size_t b = header.find('\t', a);
        if (b == std::string::npos) b = header.size();
        if (b > a) codes.push_back(std::atoi(header.substr(a, b - a).c_str()));
        a = b + 1;
    }
    if (codes.empty()) return SPH_E_ARG;
    int n = 0;
    for (std::string v;;) {
        sph_part1 p;
        std::memset(&p, 0, sizeof(p));
        size_t k = 0;
        for (; k < codes.size() && (f >> v); k++) {
            const float d = (float)std::atof(v.c_str());
            switch (codes[k]) {
            case 1: p.x = d; break;
            case 2: p.y = d; break;
            case 3: p.z = d; break;
            case 4: p.ux = d; break;
            case 5: p.uy = d; break;
            case 6: p.uz = d; break;
            case 7: p.m = d; break;
            case 8: p.p_type = std::atoi(v.c_str()); break;
            case 9: p.h = d; break;
            case 10: p.temp = d; break;
            case 11: p.pres = d; break;
            case 12: p.rho = d; break;
            case 15: p.concn = d; break;
            case 26: p.h_ref = d; break;
            case 27: p.m_ref = d; break;
            default: break;
            }
        }
        if (k < codes.size()) break;
        p.i_type = 1;
        if (P1) {
            if (n >= cap) return SPH_E_NOMEM;
            P1[n] = p;
        }
        n++;
    }
    *n_out = n;
    return 0;
}

namespace {

void put_be(FILE* f, const void* v) {
    unsigned char b[4];
    std::memcpy(b, v, 4);
    std::swap(b[0], b[3]);
    std::swap(b[1], b[2]);
    std::fwrite(b, 1, 4, f);
}

enum Src { S1, S2, S3 };
enum Kind { INT, FLT };

// One FIELD array: a label selects it when its first `len` characters equal
// those of `match` (the strncmp table of functions_OUTPUT.cuh:172-603, in
// its order); `name` is the array name written, `off` the member offsets of
// its components in the record.
struct Entry {
    const char* match;
    int len;
    const char* name;
    Src src;
    Kind kind;
    int ncomp;
    size_t off[3];
};

#define P1O(m) offsetof(sph_part1, m)
#define P3O(m) offsetof(sph_part3, m)
const Entry TABLE[] = {
    {"buffer_type", 3, "buffer_type", S1, INT, 1, {P1O(buffer_type)}},
    {"i_type", 3, "i_type", S1, INT, 1, {P1O(i_type)}},
    {"p_type", 3, "p_type", S1, INT, 1, {P1O(p_type)}},
    {"source", 5, "source", S1, INT, 1, {P1O(source)}},
    {"rho", 3, "rho", S1, FLT, 1, {P1O(rho)}},
    {"vol", 3, "vol", S1, FLT, 1, {P1O(vol)}},
    {"velocity", 4, "velocity", S1, FLT, 3, {P1O(ux), P1O(uy), P1O(uz)}},
    {"normal", 5, "normal", S3, FLT, 3, {P3O(nx), P3O(ny), P3O(nz)}},
    {"b_force", 5, "fb", S1, FLT, 3, {P1O(fbx), P1O(fby), P1O(fbz)}},
    {"MOST_force", 7, "fm", S1, FLT, 2, {P1O(fmx), P1O(fmy)}},
    {"MOST_heat", 7, "hm", S1, FLT, 1, {P1O(hmz)}},
    {"pressure", 3, "pressure", S1, FLT, 1, {P1O(pres)}},
    {"vis_t", 5, "vis_t", S1, FLT, 1, {P1O(vis_t)}},
    {"temp", 3, "temp", S1, FLT, 1, {P1O(temp)}},
    {"flt_s", 3, "flt_s", S1, FLT, 1, {P1O(flt_s)}},
    {"concn", 5, "concn", S1, FLT, 1, {P1O(concn)}},
    {"concentration", 6, "concentration", S1, FLT, 1, {P1O(concentration)}},
    {"enthalpy", 4, "enthalpy", S1, FLT, 1, {P1O(enthalpy)}},
    {"mass", 4, "mass", S1, FLT, 1, {P1O(m)}},
    {"eli", 3, "eli", S1, FLT, 1, {P1O(eli)}},
    {"smoothing", 5, "h", S1, FLT, 1, {P1O(h)}},
    {"monitor", 5, "monitor", S1, FLT, 3, {P1O(float1), P1O(float2), P1O(float3)}},
    {"ncell", 5, "ncell", S1, INT, 1, {P1O(ncell)}},
    {"rho_ref", 5, "rho_ref", S2, FLT, 1, {offsetof(sph_part2, rho_ref)}},
    {"k_turb", 4, "k_turb", S1, FLT, 1, {P1O(k_turb)}},
    {"e_turb", 4, "e_turb", S1, FLT, 1, {P1O(e_turb)}},
    {"dtemp", 5, "dtemp", S3, FLT, 1, {P3O(dtemp)}},
};
#undef P1O
#undef P3O

}  // namespace

extern "C" int sph_write_vtk(const char* path, const sph_part1* P1, const sph_part2* P2, const sph_part3* P3, int n,
                             const char* const* labels, int nlabels) {
    if (!path || n < 0 || (n > 0 && !P1) || nlabels < 0 || (nlabels > 0 && !labels)) return SPH_E_ARG;
    std::vector<const Entry*> arrays;
    for (int q = 0; q < nlabels; q++) {
        if (!labels[q]) return SPH_E_ARG;
        for (const Entry& e : TABLE)
            if (!std::strncmp(labels[q], e.match, e.len)) {
                if ((e.src == S2 && !P2) || (e.src == S3 && !P3)) return SPH_E_ARG;
                arrays.push_back(&e);
            }
    }
    std::vector<int> idx;
    for (int i = 0; i < n; i++)
        if (P1[i].i_type == 1 || P1[i].i_type == 2) idx.push_back(i);
    const int np = (int)idx.size();
    FILE* f = std::fopen(path, "wb");
    if (!f) return SPH_E_ARG;
    std::fprintf(f, "# vtk DataFile Version 3.0\n");
    std::fprintf(f, "Print out results in vtk format\n");
    std::fprintf(f, "BINARY\n");
    std::fprintf(f, "DATASET POLYDATA\n");
    std::fprintf(f, "POINTS\t%d\tfloat\n", np);
    for (int i : idx) {
        put_be(f, &P1[i].x);
        put_be(f, &P1[i].y);
        put_be(f, &P1[i].z);
    }
    std::fprintf(f, "POINT_DATA\t%d\n", np);
    std::fprintf(f, "FIELD FieldData\t%d\n", (int)arrays.size());
    for (const Entry* e : arrays) {
        std::fprintf(f, "%s\t%d\t%d\t%s\n", e->name, e->ncomp, np, e->kind == INT ? "int" : "float");
        const char* base = e->src == S1 ? (const char*)P1 : e->src == S2 ? (const char*)P2 : (const char*)P3;
        const size_t stride = e->src == S1 ? sizeof(sph_part1) : e->src == S2 ? sizeof(sph_part2) : sizeof(sph_part3);
        for (int i : idx)
            for (int c = 0; c < e->ncomp; c++) put_be(f, base + stride * (size_t)i + e->off[c]);
    }
    const bool ok = std::ferror(f) == 0;
    return (std::fclose(f) == 0 && ok) ? 0 : SPH_E_ARG;
}

// ------------------------------------------------------------------------
// Property tables: read_table (function_init.cuh:724-970).  A block opens
// with "#p_type N" and closes with "#end"; inside it a line starting with
// //T, //h, //k, //c(p) or //m(u) announces a comma-separated line of
// values.  Blocks are stored back to back; a block's length is that of the
// last array read in it, and block i starts where block i-1 ends.
namespace {

bool getline_c(std::ifstream& f, std::string& s) {
    if (!std::getline(f, s)) return false;
    s += '\n';
    return true;
}

// the line's fields between spaces, empty ones skipped (strtok(line, " "))
std::vector<std::string> fields(const std::string& s) {
    std::vector<std::string> out;
    for (size_t a = 0; a < s.size();) {
        size_t b = s.find(' ', a);
        if (b == std::string::npos) b = s.size();
        if (b > a) out.push_back(s.substr(a, b - a));
        a = b + 1;
    }
    return out;
}

}  // namespace

extern "C" int sph_read_tables(const char* path, sph_prop_tables* t) {
    if (!path || !t) return SPH_E_ARG;
    std::ifstream f(path);
    if (!f) return SPH_E_ARG;
    std::memset(t, 0, sizeof(*t));
    int N_data[SPH_MAX_TABLES] = {0};
    int data_numbers = 0;
    std::string line;
    while (getline_c(f, line)) {
        const std::vector<std::string> tok = fields(line);
        if (tok.empty() || tok[0].compare(0, 7, "#p_type") != 0) continue;
        if (tok.size() < 2) return SPH_E_ARG;
        const int number = std::atoi(tok[1].c_str());
        if (number < 0 || number >= SPH_MAX_TABLES || t->ntables >= SPH_MAX_TABLES) return SPH_E_NOMEM;
        t->ntables++;
        int count2 = 0;
        while (true) {
            if (!getline_c(f, line)) return SPH_E_ARG;   // a block without #end
            if (line.compare(0, 4, "#end") == 0) break;
            const std::vector<std::string> hf = fields(line);
            const std::string head = hf.empty() ? std::string() : hf[0];
            float* dst = nullptr;
            if (head.compare(0, 3, "//T") == 0) dst = t->T;
            else if (head.compare(0, 3, "//h") == 0) dst = t->h;
            else if (head.compare(0, 3, "//k") == 0) dst = t->k;
            else if (head.compare(0, 3, "//c") == 0) dst = t->cp;
            else if (head.compare(0, 3, "//m") == 0) dst = t->mu;
            if (!dst) continue;
            if (!getline_c(f, line)) return SPH_E_ARG;
            int count = 0;
            for (size_t a = 0; a < line.size();) {
                size_t b = line.find(',', a);
                if (b == std::string::npos) b = line.size();
                if (b > a) {
                    if (data_numbers + count >= SPH_TABLE_CAP) return SPH_E_NOMEM;
                    dst[data_numbers + count] = (float)std::atof(line.substr(a, b - a).c_str());
                    count++;
                }
                a = b + 1;
            }
            count2 = count;
        }
        data_numbers += count2;
        N_data[number] = data_numbers;
    }
    t->ndata = data_numbers;
    if (t->ntables > 0) {
        t->index[0] = 0;
        t->size[0] = N_data[0];
        for (int i = 1; i < t->ntables; i++) {
            t->index[i] = N_data[i - 1];
            t->size[i] = N_data[i] - N_data[i - 1];
        }
    }
    return 0;
}

// interp2 (functions_PROP.cuh:8-40): piecewise-linear in float, extended
// linearly past both ends, with the reference's order of operations.
extern "C" float sph_table_interp(const float* xd, const float* yd, int size, float x) {
    if (!xd || !yd || size < 2) return std::nanf("");
    const int e = size - 1;
    float y;
    if (xd[e] < x) {
        y = yd[e] - yd[e - 1];
        y /= (xd[e] - xd[e - 1]);
        y *= (x - xd[e]);
        y += yd[e];
        return y;
    }
    if (x <= xd[0]) {
        y = yd[1] - yd[0];
        y /= (xd[1] - xd[0]);
        y *= (x - xd[0]);
        y += yd[0];
        return y;
    }
    for (int i = 0; i < e; i++)
        if (xd[i] < x && x <= xd[i + 1]) {
            y = yd[i + 1] - yd[i];
            y /= (xd[i + 1] - xd[i]);
            y *= (x - xd[i]);
            y += yd[i];
            return y;
        }
    return std::nanf("");   // x is NaN
}
