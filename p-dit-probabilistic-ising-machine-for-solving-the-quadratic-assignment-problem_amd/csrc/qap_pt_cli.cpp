// qap_pt_cli.cpp -- command-line front end with the reference solver's flags and log
// lines (CPU-PIM/CPU_PIM.cpp:478-685), on top of the C-ABI in include/qap_pt.h.
//
//   qap_pt_cli [flags] FILE
//     -b beta      largest beta in natural-log units (CPU_PIM.cpp:526-533); without -b the
//                  ladder is chosen automatically (CPU_PIM.cpp:366-476, see below)
//     -m ratio     multiplicative beta step between rungs (:555-561)
//     -r R         replicas per ladder (:569-580), default 10
//     -t T         independent ladders (the reference's threads, :588-599), default 20
//     -w megas     exchange intervals (:600-606), default 10000
//     -i k         sweeps per exchange interval (NUMITERA, :548-554), default 4
//     -g goal      stop when the best cost is <= goal (:534-540)
//     -G goal      print the best permutation if the best cost is < goal (:541-547)
//     -s seconds   timeout (:581-587)
//     -o dir       output folder for log.txt / sweep<t>.txt (:562-568)
//     -al / -ah    automatic-mode proportions (:508-524)
//     --seed S     Philox seed (the reference seeds from std::random_device)
//     --tmin T --tmax T   geometric temperature ladder instead of -b/-m
//     --cpu        run on the host's cores (qap_pt_solve_cpu: the same chain, no device
//                  call), as CPU_PIM does; needs an explicit ladder (-b or --tmin/--tmax)
//                  and writes the best energy once, at the end (no sweep<t>.txt)
//
// FILE holds n, then F (n*n), then D (n*n), whitespace separated (readProblem,
// CPU_PIM.cpp:79-107).  Progress lines are written after every chunk of exchange
// intervals, so the "Time (s)" of an improvement is resolved to one chunk.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/qap_pt.h"

namespace {

struct Options {
    std::string file, outdir;
    bool have_beta = false;
    double beta0_ln = 0.0, ratio = 1.2;
    int R = 10, T = 20, megas = 10000, numitera = 4;
    bool have_goal = false;
    long long goal = 0, goal_print = 0;
    double timeout = 0.0;
    double lower_prop = 0.1, higher_prop = 0.9;
    unsigned long long seed = 0x5eedULL;
    double tmin = 0.0, tmax = 0.0;
    bool one_shot = false;
    bool cpu = false;
};

bool read_problem(const std::string &path, int &n, std::vector<int32_t> &F, std::vector<int32_t> &D) {
    std::ifstream in(path);
    if (!in.is_open()) {
        std::cerr << "Error: Could not open the file (" << path << ")." << std::endl;
        return false;
    }
    in >> n;
    if (!in || n < 1 || n > QAP_PT_MAX_N) {
        std::cerr << "Error: N must be in [1, " << QAP_PT_MAX_N << "]." << std::endl;
        return false;
    }
    F.resize((size_t)n * n);
    D.resize((size_t)n * n);
    for (auto &x : F) in >> x;
    for (auto &x : D) in >> x;
    if (!in) {
        std::cerr << "Error: the file ends before 2 n^2 matrix entries." << std::endl;
        return false;
    }
    return true;
}

int check(int rc) {
    if (rc != QAP_PT_OK) std::cerr << "qap_pt error " << rc << ": " << qap_pt_last_error() << std::endl;
    return rc;
}

// Automatic ladder (CPU_PIM.cpp:366-476) through qap_pt_auto_beta, with the
// reference's log lines.
bool auto_select(const Options &o, int n, const std::vector<int32_t> &F, const std::vector<int32_t> &D,
                 double &beta0_ln, double &ratio, std::ostream &log) {
    qap_problem prob{n, F.data(), D.data()};
    qap_pt_params base{};
    base.R = o.R;
    base.T = o.T;
    base.sweeps_per_exchange = o.numitera;
    base.seed = o.seed;
    base.kick_after = 100;
    base.flags = o.one_shot ? QAP_PT_ONE_SHOT : 0u;  // trials run the chosen dynamics
    const auto t0 = std::chrono::steady_clock::now();
    double avg[QAP_PT_AUTO_STEPS];
    if (check(qap_pt_auto_beta(&prob, &base, o.lower_prop, o.higher_prop, avg, &beta0_ln, &ratio)))
        return false;
    for (int i = 0; i < QAP_PT_AUTO_STEPS; ++i) {
        const double b = std::exp2(QAP_PT_AUTO_MIN_EXP + i);
        log << "Beta: " << b << ", Average Energy: " << avg[i] << std::endl;
        std::cout << "Beta: " << b << ", Average Energy: " << avg[i] << std::endl;
    }
    const double maxBeta = beta0_ln * 1.44269504089, minBeta = maxBeta / std::pow(ratio, o.R);
    const long long secs = (long long)std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    log << "Selecting minimum beta of " << minBeta << std::endl
        << "Selecting maximum beta of " << maxBeta << std::endl
        << "Selecting Beta Parameters took " << secs << " seconds." << std::endl;
    std::cout << "Selecting minimum beta of " << minBeta << std::endl
              << "Selecting maximum beta of " << maxBeta << std::endl
              << "Selecting Beta Parameters took " << secs << " seconds." << std::endl
              << std::endl;
    return true;
}

int finish(const Options &o, int n, const qap_pt_result &res, const std::vector<int32_t> &best_perm,
           std::ostream &log, std::chrono::steady_clock::time_point t0);

}  // namespace

int main(int argc, char **argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char *flag) -> const char * {
            if (i + 1 >= argc) {
                std::cerr << "Error: " << flag << " option requires a value." << std::endl;
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "-h" || a == "--help") {
            std::cout << "Usage: " << argv[0] << " [OPTIONS] FILE\n"
                      << "  -b -m -r -t -w -i -g -G -s -o -al -ah as the reference CPU_PIM;\n"
                      << "  --seed S, --tmin T --tmax T (geometric temperature ladder)\n"
                      << "  --one-shot: one-shot parallel-proposal steps (the GPU-CUPIM driver's\n"
                      << "              dynamics) instead of p-dit sweeps; -i counts steps\n"
                      << "  --cpu: the same run on the host's cores (needs -b or --tmin/--tmax)\n";
            return 0;
        } else if (a == "-b") {
            o.beta0_ln = std::atof(next("-b"));
            o.have_beta = true;
        } else if (a == "-m") o.ratio = std::atof(next("-m"));
        else if (a == "-r") o.R = std::atoi(next("-r"));
        else if (a == "-t") o.T = std::atoi(next("-t"));
        else if (a == "-w") o.megas = std::atoi(next("-w"));
        else if (a == "-i") o.numitera = std::atoi(next("-i"));
        else if (a == "-g") {
            o.goal = std::atoll(next("-g"));
            o.have_goal = true;
        } else if (a == "-G") o.goal_print = std::atoll(next("-G"));
        else if (a == "-s") o.timeout = std::atof(next("-s"));
        else if (a == "-o") o.outdir = next("-o");
        else if (a == "-al") o.lower_prop = std::atof(next("-al"));
        else if (a == "-ah") o.higher_prop = std::atof(next("-ah"));
        else if (a == "--seed") o.seed = std::strtoull(next("--seed"), nullptr, 0);
        else if (a == "--tmin") o.tmin = std::atof(next("--tmin"));
        else if (a == "--tmax") o.tmax = std::atof(next("--tmax"));
        else if (a == "--one-shot") o.one_shot = true;
        else if (a == "--cpu") o.cpu = true;
        else o.file = a;
    }
    if (o.file.empty()) {
        std::cerr << "Error: No input file specified." << std::endl;
        return 1;
    }
    if (o.R < 1 || o.T < 1 || o.megas < 0 || o.numitera < 1) {
        std::cerr << "Error: -r, -t, -i must be >= 1 and -w >= 0." << std::endl;
        return 1;
    }
    if (o.cpu && (o.one_shot || (!o.have_beta && !(o.tmin > 0 && o.tmax > 0)))) {
        std::cerr << "Error: --cpu runs p-dit sweeps with an explicit ladder (-b, or --tmin and "
                     "--tmax); the automatic scan and --one-shot run on the device."
                  << std::endl;
        return 1;
    }
    if (o.outdir.empty()) o.outdir = o.file + ".out";
    if (mkdir(o.outdir.c_str(), 0755) != 0)
        std::cerr << "Error: Folder '" << o.outdir << "' already exists or could not be created." << std::endl;
    int n = 0;
    std::vector<int32_t> F, D;
    if (!read_problem(o.file, n, F, D)) return 1;

    std::ofstream log(o.outdir + "/log.txt");
    const auto t0 = std::chrono::steady_clock::now();
    log << "Iterations between swaps " << o.numitera << std::endl
        << "Number of Swaps " << o.megas << std::endl
        << "Number of Replicas " << o.R << std::endl
        << "Number of Threads " << o.T << std::endl;
    std::vector<double> betas(o.R);
    if (o.tmin > 0 && o.tmax > 0) {
        for (int l = 0; l < o.R; ++l)
            betas[l] = 1.0 / (o.tmax * std::pow(o.tmin / o.tmax, o.R > 1 ? (double)l / (o.R - 1) : 1.0));
    } else {
        double b0 = o.beta0_ln, m = o.ratio;
        if (!o.have_beta) {
            log << "Automatic mode enabled, Beta parameters will be determined automatically." << std::endl;
            if (!auto_select(o, n, F, D, b0, m, log)) return 1;
        }
        double run = b0;  // CPU_PIM.cpp:113-118: rung R-1 at beta0, each colder one divided by m
        for (int l = o.R - 1; l >= 0; --l, run /= m) betas[l] = run;
        log << "Maximum Beta " << b0 * 1.44269504089 << std::endl
            << "Multiplicative Beta between iterations " << m << std::endl;
        std::cout << "Maximum Beta " << b0 * 1.44269504089 << std::endl
                  << "Multiplicative Beta between iterations " << m << std::endl;
    }
    log << "Input file: " << o.file << std::endl;

    qap_problem prob{n, F.data(), D.data()};
    qap_pt_params par{};
    par.R = o.R;
    par.T = o.T;
    par.beta_ln = betas.data();
    par.sweeps = (int64_t)o.megas * o.numitera;
    par.sweeps_per_exchange = o.numitera;
    par.seed = o.seed;
    par.kick_after = 100;
    par.has_goal = o.have_goal ? 1 : 0;
    par.goal = o.goal;
    par.timeout_s = o.timeout;
    par.flags = o.one_shot ? QAP_PT_ONE_SHOT : 0u;
    std::vector<int32_t> best_perm(n);
    qap_pt_result res{};
    res.best_perm = best_perm.data();
    if (o.cpu) {  // one call on the host's cores; the best is reported at the end
        if (check(qap_pt_solve_cpu(&prob, &par, &res))) return 1;
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        log << "Best Energy Found: " << res.best_cost << " Time (s): " << el << std::endl;
        std::cout << "Best Energy Found: " << res.best_cost << " Time (s): " << el << std::endl;
        return finish(o, n, res, best_perm, log, t0);
    }
    qap_pt_pool *pool = nullptr;
    if (check(qap_pt_pool_create(1, &prob, &par, &pool))) return 1;
    int64_t shown = INT64_MAX;
    // sweep<t>.txt for every ladder t (CPU_PIM.cpp:320,660-664): "megas elapsed_ms best"
    // whenever the best over the ladder's replicas improved (replicas are ladder-major)
    std::vector<std::ofstream> sweep_log(o.T);
    for (int t = 0; t < o.T; ++t)
        sweep_log[t].open(o.outdir + "/sweep" + std::to_string(t) + ".txt");
    std::vector<int64_t> rep_best((size_t)o.R * o.T), ladder_shown(o.T, INT64_MAX);
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(o.megas, 100));
    for (int64_t m = 0; m < o.megas; m += chunk) {
        const int64_t step = std::min<int64_t>(chunk, o.megas - m);
        if (check(qap_pt_pool_run(pool, step * o.numitera, nullptr, nullptr))) return 1;
        if (check(qap_pt_pool_result(pool, 0, &res))) return 1;
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (res.best_cost < shown) {
            shown = res.best_cost;
            log << "Best Energy Found: " << shown << " Time (s): " << el << std::endl;
            std::cout << "Best Energy Found: " << shown << " Time (s): " << el << std::endl;
        }
        if (check(qap_pt_pool_state(pool, 0, nullptr, nullptr, rep_best.data(), nullptr, nullptr,
                                    nullptr)))
            return 1;
        for (int t = 0; t < o.T; ++t) {
            const auto first = rep_best.begin() + (size_t)t * o.R;
            const int64_t lb = *std::min_element(first, first + o.R);
            if (lb < ladder_shown[t]) {
                ladder_shown[t] = lb;
                sweep_log[t] << m + step << " " << (long long)(el * 1000) << " " << lb << std::endl;
            }
        }
        if (res.sweeps_done < (m + step) * o.numitera) break;  // goal or timeout
        if (o.timeout > 0 && el >= o.timeout) break;
    }
    qap_pt_pool_destroy(pool);
    return finish(o, n, res, best_perm, log, t0);
}

namespace {

// The closing lines (CPU_PIM.cpp:667-684): the permutation with -G, the single best energy,
// the gap to the goal and the elapsed time.
int finish(const Options &o, int n, const qap_pt_result &res, const std::vector<int32_t> &best_perm,
           std::ostream &log, std::chrono::steady_clock::time_point t0) {
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (o.goal_print > 0 && res.best_cost < o.goal_print) {
        for (int i = 0; i < n; ++i) {
            log << best_perm[i] << " ";
            std::cout << best_perm[i] << " ";
        }
        log << std::endl;
        std::cout << std::endl;
    }
    std::cout << std::endl << "Halting program." << std::endl;
    log << "Single best energy found: " << res.best_cost << std::endl;
    std::cout << "Single best energy found: " << res.best_cost << std::endl;
    if (o.have_goal && o.goal != 0)
        std::cout << "AFD " << (1.0 * res.best_cost - o.goal) / o.goal * 100.0 << std::endl;
    log << "Elapsed time: " << el << " seconds." << std::endl;
    std::cout << "Elapsed time: " << el << " seconds." << std::endl;
    return 0;
}

}  // namespace
