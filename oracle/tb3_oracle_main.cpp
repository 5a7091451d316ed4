/*
 * tb3_oracle_main.cpp -- command-line driver over the CPU oracle (test
 * infrastructure only).  Restates the reference's `main` at one process
 * (main.cpp:516-583): the flags of Utility::cliParse (utility/cli.hpp:47-187),
 * the CSV reader (utility/utility.cpp:14-24), Simulation::Start
 * (simulation/Simulation.cpp:28-81) and the per-step CSV writer
 * (Simulation.cpp:74-79, 99-133; utility.cpp:54-68).
 *
 * Extra flags (not in the reference):
 *   --oracle-mode brute|owner   sweep implementation (identical results)
 *   --threads N                 OpenMP threads for the owner sweep
 *   --timing                    print {"allTimes": [ns], "max": ns} per step
 *                               (the MEASURESIMSTEP_3BMDA line, Simulation.cpp:50-72)
 *   --counts                    print the accepted-triplet count of each step
 *   --full-precision            write CSV values with %.17g instead of %g
 *   --ranks P                   run as `mpiexec -n P` does (orc_sweep_ranks for
 *                               NATA and AUTA, orc_simulate_ranks_p3bca for
 *                               P3BCA on the decomposition {P}); --counts then
 *                               prints each rank's two counts per step
 *   --cart X,Y[,Z]              P3BCA's Cartesian decomposition (the ranks are
 *                               its product), as the reference's table gives it
 */
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "tb3_oracle.h"

namespace {

struct Args {
    int iterations = 0;
    double cutoff = 0.0;
    double deltaT = 0.0;
    std::string inputCSV, outputCSV;
    int algorithm = 0; /* NATAType */
    double g[3] = {0.0, 0.0, 0.0};
    int mode = 1;
    int threads = 0;
    bool timing = false, counts = false, fullPrecision = false;
    int ranks = 1;
    std::vector<int> cart; /* P3BCA's Cartesian decomposition; {ranks} by default */
};

void printHelp()
{
    std::cout << "Help: \n"
              << "Options:\n"
              << "\t-h,--help\t\tShow this help message\n"
              << "\t-a,--algorithm\t\talgorithm to use (\"nata\", \"p3bca\", \"auta\")\n"
              << "\t-i,--iterations\t\tnum of iterations to simulate\n"
              << "\t-d,--delta\t\tduration of one simulation step\n"
              << "\t-gx,--gravityX\t\tgravitational force in x-direction\n"
              << "\t-gy,--gravityY\t\tgravitational force in y-direction\n"
              << "\t-gz,--gravityZ\t\tgravitational force in z-direction\n"
              << "\t-csv,--csv\t\tcsv file with particles\n"
              << "\t-o,--out\t\tcsv base file name\n"
              << "\t-c,--cutoff\t\tcutoff distance" << std::endl;
}

Args parse(const std::vector<std::string> &args)
{
    Args a;
    for (size_t i = 0; i < args.size(); i++) {
        try {
            if (args[i].rfind("-", 0) == 0) {
                const std::string flag = args[i].substr(1);
                auto need = [&]() -> std::string {
                    if (args.size() <= i + 1) {
                        printHelp();
                        std::exit(1);
                    }
                    return args[++i];
                };
                if (flag == "c" || flag == "-cutoff") a.cutoff = std::stod(need());
                else if (flag == "i" || flag == "-iterations") a.iterations = std::stoi(need());
                else if (flag == "a" || flag == "-algorithm") {
                    const std::string v = need();
                    if (v == "nata") a.algorithm = 0;
                    else if (v == "p3bca") a.algorithm = 1;
                    else if (v == "auta") a.algorithm = 2;
                    else {
                        printHelp();
                        std::exit(1);
                    }
                } else if (flag == "d" || flag == "-delta") a.deltaT = std::stod(need());
                else if (flag == "csv" || flag == "-csv") a.inputCSV = need();
                else if (flag == "o" || flag == "-out") a.outputCSV = need();
                else if (flag == "gx" || flag == "-gravityX") a.g[0] = std::stod(need());
                else if (flag == "gy" || flag == "-gravityY") a.g[1] = std::stod(need());
                else if (flag == "gz" || flag == "-gravityZ") a.g[2] = std::stod(need());
                else if (flag == "decomp" || flag == "-decomp" || flag == "op" || flag == "-outProfile" ||
                         flag == "benchyaml" || flag == "-benchyaml")
                    (void)need();
                else if (flag == "-oracle-mode") a.mode = (need() == "brute") ? 0 : 1;
                else if (flag == "-threads") a.threads = std::stoi(need());
                else if (flag == "-timing") a.timing = true;
                else if (flag == "-counts") a.counts = true;
                else if (flag == "-full-precision") a.fullPrecision = true;
                else if (flag == "-ranks") a.ranks = std::stoi(need());
                else if (flag == "-cart") {
                    std::stringstream ss(need());
                    std::string v;
                    a.cart.clear();
                    while (std::getline(ss, v, ',')) a.cart.push_back(std::stoi(v));
                }
                else std::cout << "CLI parser: unrecognized option: " << flag << std::endl;
            } else {
                std::cout << "CLI parser: unrecognized option: " << args[i] << std::endl;
            }
        } catch (std::exception &) {
            printHelp();
            std::exit(1);
        }
    }
    return a;
}

struct Table {
    std::vector<int> id;
    std::vector<double> p, v, acc, m, f;
};

/* utility.cpp:14-24: header row skipped, ID = column 0 as int, columns 1..10 as double. */
bool readCSV(const std::string &file, Table &t)
{
    std::ifstream in(file);
    if (!in) return false;
    std::string line;
    if (!std::getline(in, line)) return true;
    while (std::getline(in, line)) {
        if (line.empty() || line == "\r") continue;
        std::vector<std::string> cells;
        std::stringstream ss(line);
        std::string cell;
        while (std::getline(ss, cell, ',')) cells.push_back(cell);
        if (cells.size() < 11) return false;
        t.id.push_back(std::stoi(cells[0]));
        for (int c = 0; c < 3; ++c) t.p.push_back(std::stod(cells[1 + c]));
        for (int c = 0; c < 3; ++c) t.v.push_back(std::stod(cells[4 + c]));
        for (int c = 0; c < 3; ++c) t.acc.push_back(std::stod(cells[7 + c]));
        t.m.push_back(std::stod(cells[10]));
    }
    t.f.assign(t.p.size(), 0.0);
    return true;
}

/* utility.cpp:54-68 with the ID sort of Simulation.cpp:127-128. */
void writeCSV(const std::string &file, const Table &t, bool full)
{
    const size_t n = t.id.size();
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return t.id[a] < t.id[b]; });
    FILE *fp = std::fopen(file.c_str(), "w");
    if (!fp) return;
    std::fputs("ID, pX, pY, pZ, vX, vY, vZ, aX, aY, aZ, m, fX, fY, fZ\n", fp);
    const char *fmt = full ? "%.17g" : "%g";
    char buf[64];
    for (size_t r = 0; r < n; ++r) {
        const size_t i = order[r];
        std::fprintf(fp, "%d", t.id[i]);
        const double vals[14] = {t.p[3 * i],   t.p[3 * i + 1],   t.p[3 * i + 2],   t.v[3 * i],   t.v[3 * i + 1],
                                 t.v[3 * i + 2], t.acc[3 * i],     t.acc[3 * i + 1], t.acc[3 * i + 2], t.m[i],
                                 t.f[3 * i],   t.f[3 * i + 1],   t.f[3 * i + 2],   0};
        for (int c = 0; c < 13; ++c) {
            std::snprintf(buf, sizeof(buf), fmt, vals[c]);
            std::fputs(", ", fp);
            std::fputs(buf, fp);
        }
        std::fputs("\n", fp);
    }
    std::fclose(fp);
}

struct CbState {
    Args *a;
    Table *t;
    std::string stem;
    std::chrono::steady_clock::time_point start;
    std::vector<uint64_t> *counts;
};

void stepCallback(int step, void *user)
{
    CbState *s = static_cast<CbState *>(user);
    if (s->a->timing) {
        const auto end = std::chrono::steady_clock::now();
        const long long ns = std::chrono::duration_cast<std::chrono::nanoseconds>(end - s->start).count();
        std::cout << "{\"allTimes\": [" << ns << "], \"max\": " << ns << "}" << std::endl;
    }
    if (!s->a->outputCSV.empty()) writeCSV(s->stem + "_" + std::to_string(step) + ".csv", *s->t, s->a->fullPrecision);
    s->start = std::chrono::steady_clock::now();
}

}  // namespace

int main(int argc, char **argv)
{
    std::vector<std::string> args(argv + 1, argv + argc);
    Args a = parse(args);
    Table t;
    if (!readCSV(a.inputCSV, t)) {
        std::cerr << "cannot read particles from '" << a.inputCSV << "'" << std::endl;
        return 1;
    }
    const int64_t n = (int64_t)t.id.size();
    std::vector<uint64_t> counts((size_t)std::max(a.iterations, 1), 0);
    CbState st{&a, &t, a.outputCSV.substr(0, a.outputCSV.find_last_of('.')), std::chrono::steady_clock::now(),
               &counts};
    if (!a.cart.empty()) {
        a.ranks = 1;
        for (int d : a.cart) a.ranks *= d;
    }
    if (a.ranks > 1) {
        /* Simulation::Start with the multi-rank sweep: the integrator and the
         * CSV writer act per particle, so only the sweep and its counts differ */
        if (a.algorithm == 1) {
            const int it = std::max(a.iterations, 0);
            std::vector<uint64_t> bcs((size_t)std::max(it, 1) * a.ranks), pcs(bcs.size());
            const std::vector<int> dims = a.cart.empty() ? std::vector<int>{a.ranks} : a.cart;
            const int rc = orc_simulate_ranks_p3bca((int)dims.size(), dims.data(), n, t.p.data(), t.v.data(),
                                                    t.acc.data(), t.m.data(), t.f.data(), it, a.deltaT, a.g, a.cutoff,
                                                    1.0, bcs.data(), pcs.data(), stepCallback, &st);
            if (rc == -2) {
                std::cerr << "invalid cutoff: " << a.cutoff << " for " << a.ranks << " ranks" << std::endl;
                return 1;
            }
            if (rc != 0) {
                std::cerr << "--ranks: P3BCA needs a positive cutoff and 1 to 3 decomposition dimensions" << std::endl;
                return 1;
            }
            if (a.counts)
                for (int s = 0; s < it; ++s)
                    for (int r = 0; r < a.ranks; ++r)
                        std::cout << "step " << s << ": proc " << r << ": " << pcs[(size_t)s * a.ranks + r]
                                  << " particle interactions, " << bcs[(size_t)s * a.ranks + r]
                                  << " buffer interactions" << std::endl;
            return 0;
        }
        std::vector<uint64_t> bc((size_t)a.ranks), pc((size_t)a.ranks);
        for (int s = 0; s < a.iterations; ++s) {
            orc_predictor(n, t.p.data(), t.v.data(), t.acc.data(), a.deltaT);
            if (orc_sweep_ranks(a.algorithm, a.ranks, n, t.p.data(), 1.0, t.f.data(), bc.data(), pc.data(), nullptr,
                                0, nullptr) != 0) {
                std::cerr << "--ranks: unsupported rank count" << std::endl;
                return 1;
            }
            orc_update(n, t.f.data(), t.m.data(), t.v.data(), t.acc.data(), a.deltaT, a.g);
            stepCallback(s, &st);
            if (a.counts)
                for (int r = 0; r < a.ranks; ++r)
                    std::cout << "step " << s << ": proc " << r << ": " << pc[(size_t)r]
                              << " particle interactions, " << bc[(size_t)r] << " buffer interactions" << std::endl;
        }
        return 0;
    }
    orc_simulate(n, t.p.data(), t.v.data(), t.acc.data(), t.m.data(), t.f.data(), a.algorithm, a.iterations,
                 a.deltaT, a.g, a.cutoff, 1.0, counts.data(), a.mode, a.threads, stepCallback, &st);
    if (a.counts)
        for (int s = 0; s < a.iterations; ++s)
            std::cout << "step " << s << ": " << counts[(size_t)s] << " particle interactions, 1 buffer interactions"
                      << std::endl;
    return 0;
}
