// scft_drive.cpp -- C++ host driver of the MI355X SCFT string method.
//
// Reproduces the order of the reference's main() (scft.cu:869-1191) on one GPU, with every
// replica on that GPU: parameters from input.dat, initial fields (win files or the analytic
// profile), the W+ pre-relaxation, solve_field (string iterations with periodic check / FEs /
// vtk / win output), the final free energy and outputs. File formats follow scft.cu:486-512,
// 771-775, 786-788, 843-850.
//
//   scft_drive [input.dat]
#include <signal.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "scft_host.hpp"

namespace {

volatile sig_atomic_t g_stop = 0;
void on_sigint(int) { g_stop = 7; }  // scft.cu:299-302

using Params = std::unordered_map<std::string, std::vector<double>>;

// input.dat parser (scft.cu:80-114): '#' comments, trimmed lines, "key = v1 v2 ...",
// whitespace removed from the key, values read as doubles.
Params read_parameters(const std::string& fname)
{
    Params out;
    std::ifstream in(fname);
    std::string line;
    while (std::getline(in, line)) {
        const size_t c = line.find('#');
        if (c != std::string::npos) line.erase(c);
        auto notsp = [](unsigned char ch) { return !std::isspace(ch); };
        line.erase(line.begin(), std::find_if(line.begin(), line.end(), notsp));
        line.erase(std::find_if(line.rbegin(), line.rend(), notsp).base(), line.end());
        if (line.empty()) continue;
        std::istringstream iss(line);
        std::string key;
        if (!std::getline(iss, key, '=')) continue;
        key.erase(std::remove_if(key.begin(), key.end(), ::isspace), key.end());
        std::vector<double> v;
        double x;
        while (iss >> x) v.push_back(x);
        out[key] = v;
    }
    return out;
}

bool has(const Params& p, const char* k) { return p.count(k) && !p.at(k).empty(); }

// Defaults of scft.cu:35-38, 871-872, 916-918; keys of scft.cu:120-251 (plus "chiN" and
// "flag", the README's names, and runtime strn / precision / maxiter / springsteps).
struct Run {
    scft_params p{};
    int readwin = 1, maxiter = 4000, prerelax = 5000;
    int relaxbox = 0;         // bit i: relax D_i to zero stress after the field solve (dostring = 0)
    double stresstol = 1e-6;  // |dF/dD_i| bar of that relaxation
};

Run set_parameters(const Params& P)
{
    Run r;
    scft_params& p = r.p;
    p.m[0] = p.m[1] = p.m[2] = 50;
    p.D[0] = p.D[1] = p.D[2] = 4;
    p.N = 60;
    p.f = 0.8;
    p.alpha = 0.1;
    p.ens = 2;
    p.phidb = 0;
    p.chi = 30;
    p.nrep = 48;
    p.precision = SCFT_FP64;
    p.fix[0] = p.fix[1] = 0;
    p.dostring = 1;
    p.justFE = 0;
    p.springsteps = 0;
    p.device = 0;
    p.fe_fp64 = 1;
    auto get = [&](const char* k, double& dst) { if (has(P, k)) dst = P.at(k).front(); };
    auto geti = [&](const char* k, int& dst) { if (has(P, k)) dst = (int)P.at(k).front(); };
    get("chi", p.chi);
    get("chiN", p.chi);
    get("f", p.f);
    geti("N", p.N);
    get("alpha", p.alpha);
    geti("readwin", r.readwin);
    geti("flag", r.readwin);
    geti("ens", p.ens);
    get("phidb", p.phidb);
    geti("justFE", p.justFE);
    geti("dostring", p.dostring);
    geti("strn", p.nrep);
    geti("maxiter", r.maxiter);
    geti("prerelax", r.prerelax);
    geti("device", p.device);
    geti("fe_fp64", p.fe_fp64);
    get("phihA", p.phihA);
    geti("anderson", p.anderson);
    geti("autotune", p.autotune);
    geti("ckpt", p.ckpt);
    get("anderson_lambda", p.anderson_lambda);
    get("anderson_start", p.anderson_start);
    geti("relaxbox", r.relaxbox);
    get("stresstol", r.stresstol);
    if (has(P, "precision")) p.precision = P.at("precision").front() == 32 ? SCFT_FP32 : SCFT_FP64;
    if (has(P, "m")) for (size_t i = 0; i < P.at("m").size() && i < 3; ++i) p.m[i] = (int)P.at("m")[i];
    if (has(P, "D")) for (size_t i = 0; i < P.at("D").size() && i < 3; ++i) p.D[i] = P.at("D")[i];
    if (has(P, "fix")) for (size_t i = 0; i < P.at("fix").size() && i < 2; ++i) p.fix[i] = (int)P.at("fix")[i];
    return r;
}

bool fexist(const std::string& f)
{
    if (FILE* fp = fopen(f.c_str(), "r")) {
        fclose(fp);
        return true;
    }
    return false;
}

long count_lines(const std::string& f)  // scft.cu:259-280
{
    FILE* fp = fopen(f.c_str(), "r");
    if (!fp) return -1;
    long n = 0;
    int ch;
    while ((ch = fgetc(fp)) != EOF)
        if (ch == '\n') ++n;
    fclose(fp);
    return n;
}

void tovtk(const std::string& fname, const int* m, const double* D, const double* data)  // scft.cu:486-512
{
    FILE* out = fopen(fname.c_str(), "w");
    if (!out) return;
    fprintf(out, "# vtk DataFile Version 2.0\n");
    fprintf(out, "CT Density\n");
    fprintf(out, "ASCII\n\n");
    fprintf(out, "DATASET STRUCTURED_POINTS\n");
    fprintf(out, "DIMENSIONS %d %d %d\n", m[2], m[1], m[0]);
    fprintf(out, "ORIGIN 0.000000 0.000000 0.000000\n");
    fprintf(out, "SPACING %lf %lf %lf\n\n", D[2] / m[2], D[1] / m[1], D[0] / m[0]);
    fprintf(out, "POINT_DATA %ld\n", (long)m[0] * m[1] * m[2]);
    fprintf(out, "SCALARS scalars float\n");
    fprintf(out, "LOOKUP_TABLE default\n\n");
    for (int x = 0; x < m[0]; x++)
        for (int y = 0; y < m[1]; y++) {
            for (int z = 0; z < m[2]; z++) fprintf(out, "%.4lf\t", data[z + y * m[2] + x * m[2] * m[1]]);
            fprintf(out, "\n");
        }
    fclose(out);
}

void write_win(const std::string& fname, const double* W, long long M)  // scft.cu:786-788
{
    FILE* out = fopen(fname.c_str(), "w");
    if (!out) return;
    for (long long r = 0; r < M; r++) fprintf(out, "%.6lf %.6lf\n", W[r], W[r + M]);
    fclose(out);
}

std::string elapsed(std::chrono::steady_clock::time_point t0)  // scft.cu:322-345
{
    long s = (long)std::chrono::duration_cast<std::chrono::seconds>(std::chrono::steady_clock::now() - t0).count();
    long d = s / 86400, h = (s % 86400) / 3600, mi = (s % 3600) / 60, se = s % 60;
    char b[64];
    if (d > 0) snprintf(b, sizeof b, "%ld, %02ld:%02ld:%02ld", d, h, mi, se);
    else if (h > 0) snprintf(b, sizeof b, "%ld:%02ld:%02ld", h, mi, se);
    else if (mi > 0) snprintf(b, sizeof b, "%ld:%02ld", mi, se);
    else snprintf(b, sizeof b, "%lds", se);
    return b;
}

}  // namespace

int main(int argc, char** argv)
{
    const auto t0 = std::chrono::steady_clock::now();
    signal(SIGINT, on_sigint);
    const std::string finc = argc > 1 ? argv[1] : "input.dat";
    Params P;
    if (fexist(finc)) {
        printf("Reading input from %s\n\n", finc.c_str());
        P = read_parameters(finc);
    } else {
        printf("%s not found. Using default values for inputs.\n", finc.c_str());
    }
    Run run = set_parameters(P);
    scft_params& p = run.p;
    const int R = p.nrep;
    const long long M = (long long)p.m[0] * p.m[1] * p.m[2];
    printf("chiN=%g f=%g N=%d alpha=%g ens=%d phidb=%g m=%d %d %d D=%g %g %g strn=%d fix=%d %d dostring=%d "
           "justFE=%d readwin=%d precision=%s\n",
           p.chi, p.f, p.N, p.alpha, p.ens, p.phidb, p.m[0], p.m[1], p.m[2], p.D[0], p.D[1], p.D[2], R, p.fix[0],
           p.fix[1], p.dostring, p.justFE, run.readwin, p.precision == SCFT_FP32 ? "fp32" : "fp64");
    if (R < 2 && p.dostring == 1) {
        fprintf(stderr, "the string method needs strn >= 2 (use dostring = 0 for one replica)\n");
        return 1;
    }

    // initial fields (scft.cu:1057-1098); rows [R][2M] = (W-, W+)
    std::vector<double> W((size_t)R * 2 * M, 0.0);
    if (run.readwin == 1) {
        bool cpto0 = false;
        for (int ii = 0; ii < R; ii++) {
            const std::string winf = "win" + std::to_string(ii);
            if (fexist(winf) && count_lines(winf) == M) {
                FILE* in = fopen(winf.c_str(), "r");
                for (long long r = 0; r < M; r++)
                    if (fscanf(in, "%lf %lf", &W[(size_t)ii * 2 * M + r], &W[(size_t)ii * 2 * M + r + M]) != 2) break;
                fclose(in);
                printf("win%d read.\n", ii);
            } else {
                printf("win%d is too small: %ld.\n", ii, count_lines(winf));
                p.springsteps += 1;
                if (ii != 0) std::copy(W.begin() + (size_t)(ii - 1) * 2 * M, W.begin() + (size_t)ii * 2 * M,
                                       W.begin() + (size_t)ii * 2 * M);
                else cpto0 = true;
            }
        }
        if (cpto0 && R > 1) std::copy(W.begin() + 2 * M, W.begin() + 4 * M, W.begin());
    } else {
        for (int x = 0; x < p.m[0]; x++)
            for (int y = 0; y < p.m[1]; y++)
                for (int z = 0; z < p.m[2]; z++) {
                    const long long r = ((long long)x * p.m[1] + y) * p.m[2] + z;
                    const double rr = std::sqrt((z - p.m[2] / 2.0) * (z - p.m[2] / 2.0) +
                                                (y - p.m[1] / 2.0) * (y - p.m[1] / 2.0));
                    W[r] += -0.5 * p.chi * std::exp(-1.0 * std::pow((rr - p.m[1] / 4.0) / 8.0, 2.0));
                    W[r] += -0.5 * p.chi * std::exp(-1.0 * std::pow((rr - p.m[1] / 8.0) / 8.0, 2.0));
                }
    }

    try {
        scft_host::density D2(p);
        scft_ctx* ctx = D2.ctx();
        const double fr = std::round(p.N * p.f) / p.N;  // scft.cu:1008
        scft_host::check(ctx, scft_set_fields(ctx, W.data()), "set_fields");
        printf("Setup\n");
        std::vector<double> phi((size_t)R * 4 * M);
        auto pull = [&]() {
            scft_host::check(ctx, scft_get_fields(ctx, W.data()), "get_fields");
            scft_host::check(ctx, scft_get_densities(ctx, phi.data()), "get_densities");
        };
        // W+ before the string (scft.cu:1105), every replica
        scft_host::check(ctx, scft_solve_w_plus(ctx, run.prerelax, 1e-4, nullptr, nullptr), "solve_field0");
        pull();
        for (int ii = 0; ii < R; ii++) {
            tovtk("rhoA_" + std::to_string(ii) + ".vtk", p.m, p.D, &phi[(size_t)ii * 4 * M]);
            tovtk("rhocB_" + std::to_string(ii) + ".vtk", p.m, p.D, &phi[(size_t)ii * 4 * M + 3 * M]);
            write_win("win" + std::to_string(ii), &W[(size_t)ii * 2 * M], M);
        }
        std::vector<int> doiis(R, 1);
        if (p.fix[0] == 1) doiis[0] = 0;
        if (p.fix[1] == 1) doiis[R - 1] = 0;
        std::vector<double> FEs(R), FEs0(R), FEs1(R), phct(R), alf(R);
        double FE = 0;
        auto free_energy = [&]() {
            scft_host::check(ctx, scft_free_energy(ctx, &FE, FEs.data(), FEs0.data(), FEs1.data(), phct.data(),
                                                   alf.data()),
                             "FreeE");
            FILE* out = fopen("FEs", "w");  // scft.cu:843-850
            for (int ii = 0; ii < R; ii++)
                fprintf(out, "%lf\t%.10lf\t%.10lf\n", alf[ii] / alf[R - 1], FEs[ii], phct[ii]);
            fclose(out);
        };
        (void)fr;
        if (p.justFE != 1) {
            printf("Solving fields with %d springsteps. (procid=0)\n", p.springsteps);
            std::vector<double> errs(R), errs3(R), x(R), xref(R), dotWs(R), dotWs2(R), nder(R), ndev(R);
            std::vector<int> steps(R);
            scft_iter_report rep{errs.data(), errs3.data(), x.data(), xref.data(), dotWs.data(), dotWs2.data(),
                                 nder.data(), ndev.data(), steps.data(), 0.0, 0.0};
            const int skip = 50;
            for (int k = 0; k < run.maxiter; k++) {
                scft_host::check(ctx, scft_string_iterate(ctx, k, 1e-3, alf.data(), &rep), "solve_field");
                const bool out_now = (k % skip == 0 || k == run.maxiter - 1 || g_stop == 7);
                if (out_now) {
                    double errtot = 0;
                    for (int ii = 0; ii < R; ii++) errtot += errs3[ii] * errs3[ii] / R;
                    errtot = std::sqrt(errtot);
                    printf("error  %5d/%d %.4lE/%.1lE %.2lg.\tEs: %.2lE %.2lE ... %.2lE ... %.2lE %.2lE... %.2lg.. "
                           "%.1lf (Time: %s)\n",
                           k, run.maxiter, errtot, 1e-3, 0.05, errs3[0], errs3[std::min(1, R - 1)], errs3[R / 2],
                           errs3[std::max(R - 2, 0)], errs3[R - 1], rep.fitup, rep.stepav, elapsed(t0).c_str());
                    fflush(stdout);
                    if (!(errtot == errtot)) throw scft_host::error(SCFT_ERR_NAN, "string error is NaN");
                    FILE* out = fopen("check", "w");  // scft.cu:771-775
                    fprintf(out, "%d %lf %lf %lf\n", k, errtot, 0.05, FE);
                    for (int ii = 0; ii < R; ii++) fprintf(out, "%d %.12lf %.12lf\n", ii, errs[ii], errs3[ii]);
                    fprintf(out, "\n");
                    fclose(out);
                    free_energy();
                    pull();
                    for (int ii = 0; ii < R; ii++) {
                        if (!doiis[ii]) continue;
                        tovtk("rhoA_" + std::to_string(ii) + ".vtk", p.m, p.D, &phi[(size_t)ii * 4 * M]);
                        write_win("win" + std::to_string(ii), &W[(size_t)ii * 2 * M], M);
                    }
                }
                if (g_stop == 7) {
                    printf("Time's up (procid=0).\n");
                    break;
                }
            }
            printf("Fields solved (0) (Time: %s)\n", elapsed(t0).c_str());
        }
        if (run.relaxbox & 7) {
            if (R != 1 || p.dostring != 0)
                throw scft_host::error(SCFT_ERR_INVALID, "relaxbox needs strn = 1 and dostring = 0");
            double g[3] = {0, 0, 0};
            int nbox = 0;
            scft_host::check(ctx, scft_relax_box(ctx, run.relaxbox, run.stresstol, 1e-9, 5000, 50, p.D, g, &nbox),
                             "relax_box");
            printf("Box relaxed after %d boxes: D = %.10g %.10g %.10g, dF/dD = %.3e %.3e %.3e (Time: %s)\n", nbox,
                   p.D[0], p.D[1], p.D[2], g[0], g[1], g[2], elapsed(t0).c_str());
        }
        free_energy();
        printf("Free energy caculated (procid = 0) (Time: %s)\n", elapsed(t0).c_str());
        if (FE == FE) {  // scft.cu:1138-1155
            pull();
            for (int ii = 0; ii < R; ii++) {
                if (!doiis[ii]) continue;
                write_win("win" + std::to_string(ii), &W[(size_t)ii * 2 * M], M);
                tovtk("rhoA_" + std::to_string(ii) + ".vtk", p.m, p.D, &phi[(size_t)ii * 4 * M]);
                tovtk("rhocB_" + std::to_string(ii) + ".vtk", p.m, p.D, &phi[(size_t)ii * 4 * M + 3 * M]);
            }
        }
        printf("vtks and wins printed\n");
    } catch (const scft_host::error& e) {
        fprintf(stderr, "scft_drive: %s\n", e.what());
        return 2;
    }
    printf("Done (0) (Time: %s)\n", elapsed(t0).c_str());
    return 0;
}
