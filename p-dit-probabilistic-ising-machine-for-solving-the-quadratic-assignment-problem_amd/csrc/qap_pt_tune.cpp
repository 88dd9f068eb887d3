// qap_pt_tune.cpp -- the tuned sweep-shape tables (written by the package's autotune.py
// on the device it describes; read here by the shape choice at pool creation): one for
// full-chip loads (shapes_gfx950.txt, rates of kernel time with many replicas per CU) and
// one for latency-bound pools of at most one replica per CU (shapes_latency_gfx950.txt,
// rates of wall time through the launches such a pool runs).
//
// Format, one row per line, '#' starts a comment:
//   n_max  dense|sparse  lds32|lds16|lazy|hbm  nt  cpl  [swap-attempts/s]
// Rows of one flow class and placement are taken in increasing n_max; an instance uses
// the first row of its class and placement with n <= n_max.  The optional rate lets the
// placement choice compare placements that all fit.
#include <dlfcn.h>

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "qap_pt_internal.h"
#include "qap_pt_launch.h"

namespace qap {
namespace {

std::vector<TunedShape> g_table[2];  // [0] full-chip, [1] latency-bound
std::once_flag g_table_once;

std::string default_table_path(const char *file) {
    Dl_info info{};
    if (!dladdr(reinterpret_cast<void *>(&tuned_shape), &info) || !info.dli_fname) return {};
    std::string lib = info.dli_fname;
    const size_t slash = lib.rfind('/');
    const std::string dir = slash == std::string::npos ? std::string(".") : lib.substr(0, slash);
    return dir + "/../tune/" + file;
}

void load_table(std::vector<TunedShape> &table, const char *env_name, const char *file) {
    const char *env = std::getenv(env_name);
    const std::string path = env ? std::string(env) : default_table_path(file);
    if (path.empty()) return;
    std::ifstream in(path);
    if (!in.is_open()) return;
    std::string line;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ss(line);
        TunedShape t{};
        std::string flow, place;
        if (!(ss >> t.n_max >> flow >> place >> t.nt >> t.cpl)) continue;
        if (!(ss >> t.rate)) t.rate = 0.0;
        if (flow != "dense" && flow != "sparse") continue;
        t.sparse = flow == "sparse";
        if (place == "lds32") t.im = IM_LDS32;
        else if (place == "lds16") t.im = IM_LDS16;
        else if (place == "hbm") t.im = IM_HBM;
        else if (place == "lazy") t.im = IM_LAZY;
        else continue;
        if (t.n_max < 1 || !sweep_shape_supported(t.nt, t.cpl, t.im)) continue;
        table.push_back(t);
    }
    std::stable_sort(table.begin(), table.end(),
                     [](const TunedShape &a, const TunedShape &b) { return a.n_max < b.n_max; });
}

void load_tables() {
    load_table(g_table[0], "QAP_PT_TUNE_TABLE", "shapes_gfx950.txt");
    load_table(g_table[1], "QAP_PT_TUNE_LATENCY_TABLE", "shapes_latency_gfx950.txt");
}

}  // namespace

const TunedShape *tuned_shape(int n, bool sparse, int im, bool latency) {
    std::call_once(g_table_once, load_tables);
    for (const TunedShape &t : g_table[latency ? 1 : 0])
        if (t.sparse == (sparse ? 1 : 0) && t.im == im && n <= t.n_max) return &t;
    return nullptr;
}

}  // namespace qap
