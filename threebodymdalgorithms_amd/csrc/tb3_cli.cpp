// tb3_cli.cpp -- `tb3`, the drop-in for the reference's `main` at one GPU.
//
// Same flags as Utility::cliParse (utility/cli.hpp:47-187): -a/--algorithm
// (nata|p3bca|auta), -i/--iterations, -d/--delta, -gx/-gy/-gz
// (--gravityX/Y/Z), -c/--cutoff, -csv/--csv, -o/--out, -op/--outProfile;
// -decomp and -benchyaml are accepted and ignored (they only steer
// multi-process runs).  Unknown flags are reported and skipped (cli.hpp:171-178);
// a flag without its value or an unparsable value prints the help and exits 1
// (cli.hpp:57-60, 181-184).  Defaults are the reference's zero-initialised
// arguments (main.cpp:28): NATA, 0 iterations, dt 0, cutoff 0, g = 0.
//
// Input: the particle CSV of getParticlesFromCSV (utility/utility.cpp:14-24),
// header row skipped, columns ID, pX, pY, pZ, vX, vY, vZ, aX, aY, aZ, m (a
// 14-column output file is also accepted, the force columns are ignored).
// Output: after every step `<stem>_<i>.csv` with stem = the -o value up to its
// last '.', rows sorted by ID, header
// "ID, pX, pY, pZ, vX, vY, vZ, aX, aY, aZ, m, fX, fY, fZ", ", " separators and
// %g (6 significant digits) values, as Simulation.cpp:74-79, 99-133 and
// utility.cpp:54-68.
//
// Extra flags (not in the reference):
//   --precision exact|fp64|fp32   arithmetic of the sweep (default exact:
//                                 bit-identical to the reference)
//   --device N                    GPU ordinal
//   --full-precision              write %.17g instead of %g
//   --timing                      print {"allTimes": [ns], "max": ns} per step
//                                 (the MEASURESIMSTEP_3BMDA line, Simulation.cpp:50-72)
//   --counts                      print the interaction counts of step 0 as the
//                                 reference's VLEVEL line (main.cpp:561-569)
//
// -op <file> writes the profile JSON of a PROFILE_3BMDA build of the reference
// (main.cpp:127-451) at one process, for the keys this path has, in the
// reference's key order (its timer map is ordered by name): per-step
// "SumUpParticles" times in ns (Algorithm.cpp:237-270: the device time of the
// kernel that ends the pass, which forms b0 + (b1 + b2) on the exact path and
// integrates; the fast paths accumulate one force and do no separate sum, so
// there it is the integration alone), "calculateForces" and
// "calculateInteractions" times (the sweep kernel's device time: on the GPU
// the potential is evaluated inside the sweep), the
// hitrate (accepted / candidate triplets, Algorithm.cpp:92-138 counters, with
// the reference's averaging: step 0's value over the number of recorded
// hitrates, main.cpp:403-416, and one "0" member per step in
// "hitrate per proc") and "stvr" (0 at one process).  Doubles are written as
// rapidjson's Writer does: shortest round-trip digits, ".0" on integral values.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/tb3.h"

namespace {

struct Args {
    int iterations = 0;
    double cutoff = 0.0;
    double deltaT = 0.0;
    std::string inputCSV, outputCSV;
    int algorithm = TB3_NATA;
    double g[3] = {0.0, 0.0, 0.0};
    int precision = TB3_FP64_EXACT;
    int device = 0;
    bool fullPrecision = false, timing = false, counts = false;
    std::string outputProfile;
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
              << "\t-o,--out\t\tcsv base file name that will be used to create all csv outputs from each "
                 "simulation step\n"
              << "\t-c,--cutoff\t\tcutoff distance\n"
              << "\t--precision\t\texact | fp64 | fp32 (sweep arithmetic, default exact)\n"
              << "\t--device\t\tGPU ordinal\n"
              << "\t--full-precision\twrite %.17g values\n"
              << "\t--timing\t\tprint per-step wall time JSON\n"
              << "\t--counts\t\tprint the interaction counts" << std::endl;
}

Args parse(const std::vector<std::string> &args)
{
    Args a;
    for (size_t i = 0; i < args.size(); i++) {
        try {
            if (args[i].rfind("-", 0) == 0) {
                const std::string flag = args[i].substr(1);
                auto value = [&]() -> std::string {
                    if (args.size() <= i + 1) {
                        printHelp();
                        std::exit(1);
                    }
                    return args[++i];
                };
                if (flag == "c" || flag == "-cutoff") {
                    a.cutoff = std::stod(value());
                } else if (flag == "i" || flag == "-iterations") {
                    a.iterations = std::stoi(value());
                } else if (flag == "a" || flag == "-algorithm") {
                    const std::string v = value();
                    if (v == "nata") a.algorithm = TB3_NATA;
                    else if (v == "p3bca") a.algorithm = TB3_P3BCA;
                    else if (v == "auta") a.algorithm = TB3_AUTA;
                    else {
                        printHelp();
                        std::exit(1);
                    }
                } else if (flag == "decomp" || flag == "-decomp") {
                    const std::string v = value();
                    if (v != "optimal" && v != "naive") {
                        printHelp();
                        std::exit(1);
                    }
                } else if (flag == "d" || flag == "-delta") {
                    a.deltaT = std::stod(value());
                } else if (flag == "csv" || flag == "-csv") {
                    a.inputCSV = value();
                } else if (flag == "o" || flag == "-out") {
                    a.outputCSV = value();
                } else if (flag == "op" || flag == "-outProfile") {
                    a.outputProfile = value();
                } else if (flag == "benchyaml" || flag == "-benchyaml") {
                    (void)value();
                } else if (flag == "gx" || flag == "-gravityX") {
                    a.g[0] = std::stod(value());
                } else if (flag == "gy" || flag == "-gravityY") {
                    a.g[1] = std::stod(value());
                } else if (flag == "gz" || flag == "-gravityZ") {
                    a.g[2] = std::stod(value());
                } else if (flag == "-precision") {
                    const std::string v = value();
                    if (v == "exact") a.precision = TB3_FP64_EXACT;
                    else if (v == "fp64") a.precision = TB3_FP64;
                    else if (v == "fp32") a.precision = TB3_FP32_MIXED;
                    else {
                        printHelp();
                        std::exit(1);
                    }
                } else if (flag == "-device") {
                    a.device = std::stoi(value());
                } else if (flag == "-full-precision") {
                    a.fullPrecision = true;
                } else if (flag == "-timing") {
                    a.timing = true;
                } else if (flag == "-counts") {
                    a.counts = true;
                } else if (flag == "h" || flag == "-help") {
                    printHelp();
                    std::exit(0);
                } else {
                    std::cout << "CLI parser: unrecognized option: " << flag << std::endl;
                }
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

struct Particles {
    std::vector<int32_t> id;
    std::vector<double> p, v, a, m, f;
    size_t size() const { return id.size(); }
};

// getParticlesFromCSV (utility.cpp:14-24): rapidcsv with the first row as
// column labels; each cell parsed with std::stod / std::stoi.
bool readCSV(const std::string &file, Particles &t, std::string &err)
{
    std::ifstream in(file);
    if (!in) {
        err = "cannot open '" + file + "'";
        return false;
    }
    std::string line;
    if (!std::getline(in, line)) return true;
    size_t row = 1;
    while (std::getline(in, line)) {
        ++row;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) continue;
        std::vector<std::string> cells;
        std::stringstream ss(line);
        std::string cell;
        while (std::getline(ss, cell, ',')) cells.push_back(cell);
        if (cells.size() < 11) {
            err = file + ":" + std::to_string(row) + ": expected at least 11 columns";
            return false;
        }
        try {
            t.id.push_back((int32_t)std::stoi(cells[0]));
            for (int c = 0; c < 3; ++c) t.p.push_back(std::stod(cells[1 + c]));
            for (int c = 0; c < 3; ++c) t.v.push_back(std::stod(cells[4 + c]));
            for (int c = 0; c < 3; ++c) t.a.push_back(std::stod(cells[7 + c]));
            t.m.push_back(std::stod(cells[10]));
        } catch (std::exception &) {
            err = file + ":" + std::to_string(row) + ": unparsable value";
            return false;
        }
    }
    t.f.assign(t.p.size(), 0.0);
    return true;
}

// writeStepToCSVWithForces (utility.cpp:54-68) after the ID sort of
// Simulation.cpp:127-128.
bool writeCSV(const std::string &file, const Particles &t, bool full)
{
    const size_t n = t.size();
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return t.id[a] < t.id[b]; });
    FILE *fp = std::fopen(file.c_str(), "w");
    if (!fp) return false;
    std::fputs("ID, pX, pY, pZ, vX, vY, vZ, aX, aY, aZ, m, fX, fY, fZ\n", fp);
    const char *fmt = full ? ", %.17g" : ", %g";
    for (size_t r = 0; r < n; ++r) {
        const size_t i = order[r];
        std::fprintf(fp, "%d", t.id[i]);
        const double vals[13] = {t.p[3 * i], t.p[3 * i + 1], t.p[3 * i + 2], t.v[3 * i], t.v[3 * i + 1],
                                 t.v[3 * i + 2], t.a[3 * i], t.a[3 * i + 1], t.a[3 * i + 2], t.m[i],
                                 t.f[3 * i], t.f[3 * i + 1], t.f[3 * i + 2]};
        for (double v : vals) std::fprintf(fp, fmt, v);
        std::fputs("\n", fp);
    }
    std::fclose(fp);
    return true;
}

// A double as rapidjson's Writer prints it (Grisu2 shortest digits, then
// Prettify with maxDecimalPlaces 324): integral values keep ".0", decimal
// notation while the decimal point falls within 21 digits and down to 1e-6,
// exponents without padding otherwise.
std::string jsonDouble(double v)
{
    if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
    std::string sci(buf, r.ptr);  // [-]d[.ddd]e(+|-)XX
    std::string sign;
    if (sci[0] == '-') {
        sign = "-";
        sci = sci.substr(1);
    }
    const size_t e = sci.find('e');
    std::string digits = sci.substr(0, e);
    digits.erase(std::remove(digits.begin(), digits.end(), '.'), digits.end());
    const int exp10 = std::stoi(sci.substr(e + 1));  // value = d.ddd x 10^exp10
    const int length = (int)digits.size();
    const int k = exp10 - (length - 1);  // value = digits x 10^k
    const int kk = length + k;           // position of the decimal point
    std::string out;
    if (k >= 0 && kk <= 21) {
        out = digits + std::string(k, '0') + ".0";
    } else if (0 < kk && kk <= 21) {
        out = digits.substr(0, kk) + "." + digits.substr(kk);
    } else if (-6 < kk && kk <= 0) {
        out = "0." + std::string(-kk, '0') + digits;
    } else {
        const int x = kk - 1;
        out = digits.substr(0, 1);
        if (length > 1) out += "." + digits.substr(1);
        out += "e" + std::to_string(x);
    }
    return sign + out;
}

// The profile JSON (see the header comment) for one process.
bool writeProfile(const std::string &path, const std::vector<long long> &sweepNs,
                  const std::vector<long long> &sumNs, const std::vector<double> &hitrates)
{
    std::ostringstream o;
    auto times = [&](const char *key, const std::vector<long long> &ns) {
        long long total = 0;
        for (long long t : ns) total += t;
        const double steps = (double)ns.size();
        o << "\"" << key << "\":{\"time unit\":\"nanosecond\",\"avg time all proc\":"
          << jsonDouble(steps > 0 ? (double)total / steps : 0.0) << ",\"total time all proc\":"
          << jsonDouble((double)total) << ",\"total avg time union proc\":" << jsonDouble((double)total)
          << ",\"total max time union proc\":" << total << ",\"times per proc\":{\"0\":[";
        for (size_t i = 0; i < ns.size(); ++i) o << (i ? "," : "") << ns[i];
        o << "]}}";
    };
    o << "{\"times\":{";
    if (!sweepNs.empty()) {
        times("SumUpParticles", sumNs);
        o << ",";
        times("calculateForces", sweepNs);
        o << ",";
        times("calculateInteractions", sweepNs);
    }
    o << "},\"hitrate\":{";
    if (!hitrates.empty()) o << "\"avg hitrate per sim step\":[" << jsonDouble(hitrates[0] / (double)hitrates.size()) << "]";
    o << "},\"hitrate per proc\":{";
    for (size_t j = 0; j < hitrates.size(); ++j) o << (j ? "," : "") << "\"0\":" << jsonDouble(hitrates[j]);
    o << "},\"stvr\":[";
    if (!sweepNs.empty()) o << "{\"rank\":[0],\"stvr\":0.0}";
    o << "]}";
    std::ofstream f(path);
    if (!f) return false;
    f << o.str();
    return (bool)f;
}

int die(tb3_ctx *ctx, const char *what, tb3_status st)
{
    std::cerr << "tb3: " << what << " failed (" << (int)st << "): " << tb3_last_error(ctx) << std::endl;
    if (ctx) tb3_destroy(ctx);
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    const Args a = parse(std::vector<std::string>(argv + 1, argv + argc));
    Particles t;
    std::string err;
    if (!readCSV(a.inputCSV, t, err)) {
        std::cerr << "tb3: " << err << std::endl;
        return 1;
    }
    const int64_t n = (int64_t)t.size();

    tb3_ctx *ctx = nullptr;
    tb3_status st = tb3_create(a.device, &ctx);
    if (st != TB3_OK) return die(nullptr, "tb3_create", st);
    st = tb3_set_particles(ctx, n, t.id.data(), t.p.data(), t.v.data(), t.a.data(), t.m.data());
    if (st != TB3_OK) return die(ctx, "tb3_set_particles", st);
    // the AxilrodTeller(1.0) of main.cpp:41,66,91
    st = tb3_configure(ctx, a.algorithm, 1.0, a.cutoff, nullptr, a.precision);
    if (st != TB3_OK) return die(ctx, "tb3_configure", st);

    const std::string stem = a.outputCSV.substr(0, a.outputCSV.find_last_of('.'));
    std::vector<uint64_t> counts;
    const bool profile = !a.outputProfile.empty();
    std::vector<long long> sweepNs, sumNs;
    std::vector<double> hitrates;
    // candidate triplets of the one-process i<j<k sweep (Algorithm.cpp:72-92)
    const double candidates = n >= 3 ? (double)n * (double)(n - 1) * (double)(n - 2) / 6.0 : 0.0;
    double sweepMs = 0.0, sumMs = 0.0;
    if (profile) {
        st = tb3_set_profiling(ctx, 1);
        if (st != TB3_OK) return die(ctx, "tb3_set_profiling", st);
    }
    for (int i = 0; i < a.iterations; ++i) {
        const auto start = std::chrono::steady_clock::now();
        uint64_t nb = 0, nt = 0;
        st = tb3_step(ctx, a.deltaT, a.g, &nb, &nt);
        if (st != TB3_OK) return die(ctx, "tb3_step", st);
        counts.push_back(nt);
        if (profile) {
            double prof[6] = {0, 0, 0, 0, 0, 0};
            st = tb3_get_profile(ctx, prof, 6);
            if (st != TB3_OK) return die(ctx, "tb3_get_profile", st);
            sweepNs.push_back(std::llround((prof[3] - sweepMs) * 1e6));
            sweepMs = prof[3];
            sumNs.push_back(std::llround((prof[4] - sumMs) * 1e6));
            sumMs = prof[4];
            // the reference records a hitrate only when there were candidates
            if (candidates > 0) hitrates.push_back((double)nt / candidates);
        }
        if (a.timing) {
            const long long ns = std::chrono::duration_cast<std::chrono::nanoseconds>(
                                     std::chrono::steady_clock::now() - start)
                                     .count();
            std::cout << "{\"allTimes\": [" << ns << "], \"max\": " << ns << "}" << std::endl;
        }
        if (!a.outputCSV.empty()) {
            st = tb3_get_particles(ctx, n, nullptr, t.p.data(), t.v.data(), t.a.data(), nullptr, t.f.data());
            if (st != TB3_OK) return die(ctx, "tb3_get_particles", st);
            const std::string file = stem + "_" + std::to_string(i) + ".csv";
            if (!writeCSV(file, t, a.fullPrecision)) {
                std::cerr << "tb3: cannot write '" << file << "'" << std::endl;
                tb3_destroy(ctx);
                return 1;
            }
        }
    }
    if (profile && !writeProfile(a.outputProfile, sweepNs, sumNs, hitrates)) {
        std::cerr << "tb3: cannot write '" << a.outputProfile << "'" << std::endl;
        tb3_destroy(ctx);
        return 1;
    }
    if (a.counts)
        std::cout << "I'm proc 0 and have done " << (counts.empty() ? 0 : counts[0])
                  << " particle interactions actually, and " << (counts.empty() ? 0 : 1) << " buffer interactions"
                  << std::endl;
    tb3_destroy(ctx);
    return 0;
}
