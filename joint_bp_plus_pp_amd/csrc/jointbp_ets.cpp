// jointbp_ets — command line of the MI355X joint-BP decoder.
//
// Keeps the reference program's flags, input formats and result lines
// (jointbp_ets.cpp:4733-4999 usage and parsing, 5003-5211 code summary,
// 5426-5860 single decodes, 5862-6449 --simulate loop and Run Summary) for
// post-processing on (its default, ETS / flip / OSD) or off (--no-pp), and
// runs every decode through libjbp (include/jbp.h): trials are sampled,
// decoded and judged on the device in batches; with post-processing the
// trials still unconverged after 99 iterations finish on the device one
// iteration per launch, with the repairs on host threads.
//
// Differences, by design:
//   * trapping-set tables are looked up where the reference looks (the
//     params file's stem as a directory under the working directory, then
//     the bare file name) and, failing both, in that directory next to the
//     params file;
//   * there is no CPU decoder: if the device cannot be used the program
//     reports the error and exits (the reference falls back to its CPU
//     decoder, jointbp_ets.cpp:5421-5423);
//   * --precision 32|64 selects FP32 messages (default) or the FP64 decoder
//     that reproduces the reference's CPU decoder trial for trial.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sys/stat.h>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "jbp.h"
#include "jbp_code.h"

namespace {

const size_t kBoxWidth = 72;

std::string fixed(double v, int prec) {
    std::ostringstream o;
    o << std::fixed << std::setprecision(prec) << v;
    return o.str();
}

// Value scaled to a unit with 2/1/0 decimals (jointbp_ets.cpp:203-237).
std::string scaled(double v, double scale, const char *suffix) {
    double s = v / scale;
    int prec = std::fabs(s) < 10.0 ? 2 : (std::fabs(s) < 100.0 ? 1 : 0);
    return fixed(s, prec) + suffix;
}
std::string fmt_latency(double sec) {
    return std::fabs(sec) < 1e-3 ? scaled(sec, 1e-6, "us") : scaled(sec, 1e-3, "ms");
}
std::string fmt_qbps(double q) {
    double a = std::fabs(q);
    if (a >= 1e9) return scaled(q, 1e9, "G");
    if (a >= 1e6) return scaled(q, 1e6, "M");
    if (a >= 1e3) return scaled(q, 1e3, "k");
    return scaled(q, 1.0, "");
}
const char *tf(bool b) { return b ? "true" : "false"; }

void section(const std::string &title) {
    std::string bar = "+" + std::string(kBoxWidth, '-') + "+";
    std::cout << "\n" << bar << "\n| " << title << std::string(kBoxWidth - 1 - title.size(), ' ')
              << "|\n" << bar << "\n";
}

// key=value tokens joined by two spaces, wrapped at the box width.
void kv_lines(const std::vector<std::pair<std::string, std::string>> &items,
              const std::string &prefix = "") {
    std::string line;
    for (const auto &kv : items) {
        std::string tok = kv.first + "=" + kv.second;
        if (line.empty()) {
            line = tok;
        } else if (prefix.size() + line.size() + 2 + tok.size() <= kBoxWidth) {
            line += "  " + tok;
        } else {
            std::cout << prefix << line << "\n";
            line = tok;
        }
    }
    if (!line.empty()) std::cout << prefix << line << "\n";
}

// Wilson score interval (jointbp_ets.cpp:4388-4399).
bool wilson(long long n, long long k, double z, double &lo, double &hi) {
    if (n <= 0) return false;
    double nn = static_cast<double>(n), ph = static_cast<double>(k) / nn, z2 = z * z;
    double den = 1.0 + z2 / nn;
    double c = (ph + z2 / (2.0 * nn)) / den;
    double h = z * std::sqrt((ph * (1.0 - ph) + z2 / (4.0 * nn)) / nn) / den;
    lo = std::max(0.0, c - h);
    hi = std::min(1.0, c + h);
    return true;
}

uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

bool read_ints(const std::string &path, std::vector<int> &out, bool as_bits) {
    std::ifstream in(path);
    if (!in) return false;
    out.clear();
    int v = 0;
    while (in >> v) out.push_back(as_bits ? (v ? 1 : 0) : v);
    return true;
}

bool write_ints(const std::string &path, const std::vector<int> &v) {
    std::ofstream o(path);
    if (!o) return false;
    for (int x : v) o << x << "\n";
    return true;
}

// --iter-hist-out (jointbp_ets.cpp:4623-4643)
bool write_hist(const std::string &path, const std::vector<long long> &hist, long long total) {
    if (path.empty() || total <= 0 || hist.size() <= 1) return true;
    std::ofstream o(path);
    if (!o) return false;
    o << "# costs " << fixed(0.0, 8) << " " << fixed(0.0, 8) << "\n# iter fraction\n";
    for (size_t i = 1; i < hist.size(); ++i) {
        if (!hist[i]) continue;
        o << i << " " << fixed(static_cast<double>(hist[i]) / static_cast<double>(total), 10) << "\n";
    }
    return true;
}

void usage(const char *prog) {
    std::cerr << "Usage: " << prog << " --params FILE [options]\n\n"
              << "Input modes:\n"
              << "  --simulate            random trials (--p, --trials, --min-fails, --seed, --trial-index)\n"
              << "  --sx FILE --sz FILE   decode one syndrome pair (one 0/1 per line)\n"
              << "  --err FILE            decode the syndromes of one error (0=I,1=X,2=Z,3=Y per line)\n"
              << "Decoder:\n"
              << "  --p N (0.05)  --max-iter N (50)  --damping R (0.0)\n"
              << "  --freeze-syn          freeze a side once its syndrome matches (FP64, --no-pp)\n"
              << "  --precision 32|64     FP32 messages (default) or the FP64 decoder that\n"
              << "                        reproduces the reference CPU decoder trial for trial\n"
              << "  --no-pp               disable post-processing (ETS / flip / OSD)\n"
              << "  --flip-hist N (5)     flip history window of post-processing\n"
              << "  --ets6-x/-z FILE  --ets12-x/-z FILE (--ets-x/-z)  --ets8-path4 FILE\n"
              << "Randomness / trials:\n"
              << "  --seed N  --trial-index N  --trials N (1)  --min-fails N  --report-every N (100)\n"
              << "Output:\n"
              << "  --report-fail  --est FILE  --iter-hist-out FILE\n"
              << "Device:\n"
              << "  --cuda  --cuda-device N (0)  --cuda-graph  --cuda-check-warmup N  --cuda-check-interval N\n"
              << "Accepted for compatibility and ignored: --progress-tsv, --progress-every,\n"
              << "  --save-fail-prefix\n";
}

struct Progress {
    long long done = 0, failures = 0, bp_fail = 0, stab = 0, iters = 0;
    long long pp = 0, flip = 0, ets_used = 0, ets_saved = 0, ets6_used = 0, ets6_saved = 0;
    long long ets12_used = 0, ets12_saved = 0;
};

bool file_exists(const std::string &path) {
    struct stat st;
    return !path.empty() && stat(path.c_str(), &st) == 0;
}

std::string strip_dir(const std::string &path) {
    size_t slash = path.find_last_of("/\\");
    return slash == std::string::npos ? path : path.substr(slash + 1);
}

std::string dir_of(const std::string &path) {
    size_t slash = path.find_last_of("/\\");
    return slash == std::string::npos ? std::string(".") : path.substr(0, slash);
}

// resolve_existing_path (jointbp_ets.cpp:4439-4446), then next to the
// params file.
std::string resolve_table(const std::string &path, const std::string &params_dir) {
    if (file_exists(path)) return path;
    const std::string base = strip_dir(path);
    if (base != path && file_exists(base)) return base;
    if (!path.empty() && path[0] != '/' && file_exists(params_dir + "/" + path))
        return params_dir + "/" + path;
    return "";
}

void report_progress(const Progress &p, double elapsed, long long k, double latency) {
    double fer = static_cast<double>(p.failures) / static_cast<double>(p.done);
    double fps = elapsed > 0 ? p.done / elapsed : 0.0;
    section("PROGRESS");
    kv_lines({{"trials", std::to_string(p.done)}, {"failures", std::to_string(p.failures)},
              {"bp_fail", std::to_string(p.bp_fail)}, {"FER", fixed(fer, 6)},
              {"elapsed_s", fixed(elapsed, 2) + "s"}, {"latency", fmt_latency(latency)},
              {"fps", fixed(fps, 2)}, {"qbps", fmt_qbps(fps * static_cast<double>(k))},
              {"avg_iter", fixed(static_cast<double>(p.iters) / p.done, 2)},
              {"iters", std::to_string(p.iters)}, {"stab_success", std::to_string(p.stab)},
              {"pp_success", std::to_string(p.pp)}});
}

}  // namespace

int main(int argc, char **argv) {
    if (argc == 1) {
        usage(argv[0]);
        return 0;
    }
    std::string params_path, sx_path, sz_path, err_path, est_path, hist_path;
    double p_err = 0.05, damping = 0.0;
    int max_iter = 50, cuda_device = 0, warmup = 0, interval = 1, precision = JBP_FP32;
    int flip_hist = 5;
    std::string ets6_x = "ets_6_2.txt", ets6_z = "ets_6_2.txt", ets12_x = "ets_12_2.txt",
                ets12_z = "ets_12_2.txt", ets8_path4 = "ets_8_2_path4.txt";
    bool ets6_x_set = false, ets6_z_set = false, ets12_x_set = false, ets12_z_set = false,
         ets8_path4_set = false;
    bool simulate = false, verbose = false, verbose_all = false, report_fail = false;
    bool report_ets = false, freeze_syn = false, enable_pp = true, use_cuda = false, use_graph = false;
    unsigned long long seed = 0;
    bool seed_set = false;
    long long trials = 1, min_fails = 0, report_every = 100, trial_index = -1;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto need = [&](int n) {
            if (i + n >= argc) {
                usage(argv[0]);
                std::exit(1);
            }
        };
        try {
            if (a == "--params") { need(1); params_path = argv[++i]; }
            else if (a == "--sx") { need(1); sx_path = argv[++i]; }
            else if (a == "--sz") { need(1); sz_path = argv[++i]; }
            else if (a == "--err") { need(1); err_path = argv[++i]; }
            else if (a == "--est") { need(1); est_path = argv[++i]; }
            else if (a == "--p") { need(1); p_err = std::stod(argv[++i]); }
            else if (a == "--max-iter") { need(1); max_iter = std::stoi(argv[++i]); }
            else if (a == "--flip-hist") { need(1); flip_hist = std::max(0, std::stoi(argv[++i])); }
            else if (a == "--ets6-x") { need(1); ets6_x = argv[++i]; ets6_x_set = true; }
            else if (a == "--ets6-z") { need(1); ets6_z = argv[++i]; ets6_z_set = true; }
            else if (a == "--ets12-x" || a == "--ets-x") { need(1); ets12_x = argv[++i]; ets12_x_set = true; }
            else if (a == "--ets12-z" || a == "--ets-z") { need(1); ets12_z = argv[++i]; ets12_z_set = true; }
            else if (a == "--ets8-path4") { need(1); ets8_path4 = argv[++i]; ets8_path4_set = true; }
            else if (a == "--freeze-syn") { freeze_syn = true; }
            else if (a == "--damping") { need(1); damping = std::stod(argv[++i]); }
            else if (a == "--simulate") { simulate = true; }
            else if (a == "--verbose-all") { verbose = verbose_all = true; }
            else if (a == "--verbose") { need(1); verbose = std::stoi(argv[++i]) != 0; }
            else if (a == "--seed") { need(1); seed = std::stoull(argv[++i]); seed_set = true; }
            else if (a == "--save-fail-prefix" || a == "--progress-tsv" || a == "--progress-every") {
                need(1);
                ++i;
            }
            else if (a == "--trials") { need(1); trials = std::stoll(argv[++i]); }
            else if (a == "--min-fails") { need(1); min_fails = std::stoll(argv[++i]); }
            else if (a == "--report-every") { need(1); report_every = std::stoll(argv[++i]); }
            else if (a == "--report-fail") { report_fail = true; }
            else if (a == "--report-ets") { report_ets = true; }
            else if (a == "--iter-hist-out") { need(1); hist_path = argv[++i]; }
            else if (a == "--no-pp") { enable_pp = false; }
            else if (a == "--cuda") { use_cuda = true; }
            else if (a == "--cuda-device") { need(1); cuda_device = std::stoi(argv[++i]); }
            else if (a == "--cuda-graph") { use_graph = true; }
            else if (a == "--cuda-check-warmup") { need(1); warmup = std::max(0, std::stoi(argv[++i])); }
            else if (a == "--cuda-check-interval") {
                need(1);
                interval = std::stoi(argv[++i]);
                if (interval <= 0) interval = 1;
            }
            else if (a == "--trial-index") { need(1); trial_index = std::stoll(argv[++i]); }
            else if (a == "--precision") {
                need(1);
                precision = std::stoi(argv[++i]);
                if (precision != JBP_FP32 && precision != JBP_FP64) {
                    std::cerr << "--precision must be 32 or 64\n";
                    return 1;
                }
            }
            else if (a == "--help" || a == "-h") { usage(argv[0]); return 0; }
            else { usage(argv[0]); return 1; }
        } catch (const std::exception &) {
            std::cerr << "invalid value for " << a << "\n";
            return 1;
        }
    }
    (void)use_cuda;
    if (!seed_set) {  // jointbp_ets.cpp:5003-5008
        seed = static_cast<unsigned long long>(std::chrono::duration_cast<std::chrono::microseconds>(
            std::chrono::system_clock::now().time_since_epoch()).count());
    }
    if (params_path.empty()) {
        std::cerr << "--params is required\n";
        return 1;
    }
    char err[512] = {0};
    jbp_code *code = nullptr;
    if (jbp_code_from_params(params_path.c_str(), &code, err, sizeof(err)) != JBP_OK) {
        std::cerr << "Failed to read params from " << params_path << "\n";
        return 1;
    }
    int64_t dims[8];
    jbp_code_dims(code, dims);
    const int n = static_cast<int>(dims[0]), mx = static_cast<int>(dims[1]), mz = static_cast<int>(dims[2]);
    const int L = static_cast<int>(dims[7]), L2 = L / 2;
    std::vector<int64_t> abcd(4 * static_cast<size_t>(L2));
    jbp_code_params(code, abcd.data());

    std::cout << "AI Report: JointBP ETS (MI355X)\n";
    section("Input");
    std::cout << "params_path=" << params_path << "\n";
    auto pairs = [&](int o1, int o2) {
        std::ostringstream o;
        o << "[";
        for (int i = 0; i < L2; ++i) o << (i ? " " : "") << "(" << abcd[o1 + i] << "," << abcd[o2 + i] << ")";
        return o.str() + "]";
    };
    std::cout << "fi=" << pairs(0, L2) << "\ngi=" << pairs(2 * L2, 3 * L2) << "\n";

    int64_t d[15];
    jbp_code_diagnostics(code, 1, d);
    const long long k = n - d[0] - d[1];
    section("Code Summary");
    kv_lines({{"P", std::to_string(dims[5])}, {"J", std::to_string(dims[6])}, {"L", std::to_string(L)},
              {"L2", std::to_string(L2)}, {"p", fixed(p_err, 6)}, {"n", std::to_string(n)},
              {"mx", std::to_string(mx)}, {"mz", std::to_string(mz)}, {"rank_x", std::to_string(d[0])},
              {"rank_z", std::to_string(d[1])}, {"stab_rank", std::to_string(d[0] + d[1])},
              {"k", std::to_string(k)}, {"rate", fixed(n ? static_cast<double>(k) / n : 0.0, 6)}});
    for (int s = 0; s < 2; ++s) {
        const double edges = static_cast<double>(dims[3 + s]);
        const int m = s == 0 ? mx : mz;
        kv_lines({{"edges", std::to_string(dims[3 + s])},
                  {"var_deg", std::to_string(d[2 + 4 * s]) + "/" + fixed(edges / n, 2) + "/" + std::to_string(d[3 + 4 * s])},
                  {"check_deg", std::to_string(d[4 + 4 * s]) + "/" + fixed(m ? edges / m : 0.0, 2) + "/" + std::to_string(d[5 + 4 * s])}},
                 s == 0 ? "X: " : "Z: ");
    }
    section("Graph Checks");
    auto girth = [](int64_t g) { return g <= 0 ? std::string("inf") : std::to_string(g); };
    kv_lines({{"hx_hz_orthogonal", tf(d[10] == 0)}, {"hx_hz_odd_pairs", std::to_string(d[10])},
              {"girth_x", girth(d[11])}, {"girth_z", girth(d[12])}});

    // trapping-set tables (jointbp_ets.cpp:5031-5053 paths, 5211-5400 load
    // and the ETS Load table)
    jbp_graph_desc gd;
    jbp_code_graph(code, &gd);
    jbp_pp_tables *tables = nullptr;
    jbp_pp_tables_create(&tables);
    {
        std::string stem = strip_dir(params_path);
        const size_t dot = stem.find_last_of('.');
        if (dot != std::string::npos) stem = stem.substr(0, dot);
        const std::string pdir = dir_of(params_path);
        if (!ets6_x_set) ets6_x = stem + "/ets_6_2.txt";
        if (!ets6_z_set) ets6_z = stem + "/ets_6_2.txt";
        if (!ets12_x_set) ets12_x = stem + "/ets_12_2.txt";
        if (!ets12_z_set) ets12_z = stem + "/ets_12_2.txt";
        if (!ets8_path4_set) ets8_path4 = stem + "/ets_8_2_path4.txt";
        struct Row {
            std::string label;
            int cnt[2];
        };
        std::vector<Row> rows;
        auto add = [&](int size, const std::string &label, const std::string &px, const std::string &pz) {
            const std::string rx = resolve_table(px, pdir), rz = resolve_table(pz, pdir);
            Row r{label, {-1, -1}};
            if (jbp_pp_tables_add(tables, &gd, dims[5], size, label.c_str(), rx.empty() ? nullptr : rx.c_str(),
                                  rz.empty() ? nullptr : rz.c_str(), r.cnt, err, sizeof(err)) != JBP_OK)
                r.cnt[0] = r.cnt[1] = -1;
            rows.push_back(r);
        };
        for (int size : {6, 8, 10, 12, 14, 16, 18, 20, 22}) {
            const std::string gen = stem + "/ets_" + std::to_string(size) + "_2.txt";
            if (size == 6) add(6, "6", ets6_x, ets6_z);
            else if (size == 12) add(12, "12", ets12_x, ets12_z);
            else add(size, std::to_string(size), gen, gen);
            if (size == 8) add(8, "8p4", ets8_path4, ets8_path4);
        }
        section("ETS Load");
        auto val = [](int c) { return c < 0 ? std::string("-") : std::to_string(c); };
        auto status = [](const Row &r) {
            if (r.cnt[0] > 0 || r.cnt[1] > 0) return std::string("ok");
            if (r.cnt[0] == 0 || r.cnt[1] == 0) return std::string("empty");
            return std::string("missing");
        };
        size_t lw = 0, xw = 1, zw = 1, sw = 0;
        for (const Row &r : rows) {
            lw = std::max(lw, r.label.size() + 4);
            xw = std::max(xw, val(r.cnt[0]).size());
            zw = std::max(zw, val(r.cnt[1]).size());
            sw = std::max(sw, status(r).size());
        }
        for (const Row &r : rows) {
            std::ostringstream o;
            o << std::left << std::setw(static_cast<int>(lw)) << ("ETS" + r.label + ":") << " X="
              << std::right << std::setw(static_cast<int>(xw)) << val(r.cnt[0]) << " Z="
              << std::setw(static_cast<int>(zw)) << val(r.cnt[1]) << " status=" << std::left
              << std::setw(static_cast<int>(sw)) << status(r);
            std::cout << o.str() << "\n";
        }
    }
    if (verbose || verbose_all || report_ets) {
        std::cerr << "Per-iteration diagnostics (--verbose, --verbose-all, --report-ets) are not "
                     "available on the device decoder.\n";
        return 1;
    }
    if (simulate && !err_path.empty()) {
        std::cerr << "--err cannot be used with --simulate\n";
        return 1;
    }
    if (simulate && trial_index >= 0 && (trials != 1 || min_fails > 0)) {
        std::cerr << "--trial-index requires --trials 1 and no --min-fails\n";
        return 1;
    }
    std::vector<int> sx, sz, errv;
    if (!simulate) {
        if (trials != 1 || min_fails > 0) {
            std::cerr << "--trials/--min-fails are only valid with --simulate\n";
            return 1;
        }
        if (!err_path.empty() && (!sx_path.empty() || !sz_path.empty())) {
            std::cerr << "--err cannot be combined with --sx/--sz\n";
            return 1;
        }
        if (!err_path.empty()) {
            if (!read_ints(err_path, errv, false)) {
                std::cerr << "Failed to read error vector from " << err_path << "\n";
                return 1;
            }
            if (static_cast<int>(errv.size()) != n) {
                std::cerr << "Error vector size mismatch\n";
                return 1;
            }
            for (int v : errv)
                if (v < 0 || v > 3) {
                    std::cerr << "Error vector has invalid value\n";
                    return 1;
                }
        } else {
            if (sx_path.empty() || sz_path.empty()) {
                std::cerr << "--sx and --sz are required unless --simulate is set\n";
                return 1;
            }
            if (!read_ints(sx_path, sx, true) || !read_ints(sz_path, sz, true)) {
                std::cerr << "Failed to read syndrome files\n";
                return 1;
            }
            if (static_cast<int>(sx.size()) != mx || static_cast<int>(sz.size()) != mz) {
                std::cerr << "Syndrome size mismatch\n";
                return 1;
            }
        }
    }

    if (freeze_syn && (enable_pp || precision != JBP_FP64)) {
        std::cerr << "--freeze-syn runs on the FP64 decoder with post-processing off; add "
                     "--precision 64 --no-pp.\n";
        return 1;
    }
    jbp_ctx *ctx = nullptr;
    if (jbp_create(&gd, cuda_device, &ctx, err, sizeof(err)) != JBP_OK) {
        std::cerr << "CUDA init failed: " << err << " (no CPU decoder in this build)\n";
        return 1;
    }
    if (enable_pp && jbp_set_pp_tables(ctx, tables, err, sizeof(err)) != JBP_OK) {
        std::cerr << "post-processing tables: " << err << "\n";
        jbp_destroy(ctx);
        return 1;
    }
    jbp_pp_tables_free(tables);
    // prior = normalize({1-p, p/3, p/3, p/3})  (jointbp_ets.cpp:5484-5485)
    double prior[4] = {1.0 - p_err, p_err / 3.0, p_err / 3.0, p_err / 3.0};
    {
        double s = prior[0] + prior[1] + prior[2] + prior[3];
        for (double &v : prior) v = s <= 0.0 ? 0.25 : v / s;
    }

    if (!simulate) {
        if (!errv.empty()) {
            sx.assign(mx, 0);
            sz.assign(mz, 0);
            for (int c = 0; c < mx; ++c)
                for (int e = gd.x_check_offsets[c]; e < gd.x_check_offsets[c + 1]; ++e)
                    sx[c] ^= errv[gd.x_edge_var[e]] & 1;
            for (int c = 0; c < mz; ++c)
                for (int e = gd.z_check_offsets[c]; e < gd.z_check_offsets[c + 1]; ++e)
                    sz[c] ^= errv[gd.z_edge_var[e]] >> 1;
        }
        std::vector<int> est(n, 0);
        jbp_decode_result res;
        uint8_t ppf = JBP_PPF_BP_OK;
        auto t0 = std::chrono::steady_clock::now();
        if (enable_pp) {
            std::vector<uint32_t> bx((mx + 31) / 32 + 1, 0u), bz((mz + 31) / 32 + 1, 0u);
            for (int c = 0; c < mx; ++c) bx[c >> 5] |= static_cast<uint32_t>(sx[c] != 0) << (c & 31);
            for (int c = 0; c < mz; ++c) bz[c >> 5] |= static_cast<uint32_t>(sz[c] != 0) << (c & 31);
            uint8_t match = 0;
            std::memset(&res, 0, sizeof(res));
            if (jbp_decode_batch_pp(ctx, 1, bx.data(), bz.data(), prior, max_iter, damping, precision,
                                    flip_hist, &res.iterations, &match, &ppf, est.data(), err,
                                    sizeof(err)) != JBP_OK) {
                std::cerr << "CUDA decode failed: " << err << " (no CPU decoder in this build)\n";
                jbp_destroy(ctx);
                return 1;
            }
            res.syndrome_match = match;
        } else if (jbp_decode(ctx, sx.data(), sz.data(), prior, max_iter, warmup, interval, 0, use_graph ? 1 : 0,
                              freeze_syn ? 1 : 0, damping, precision, &res, est.data(), err, sizeof(err)) != JBP_OK) {
            std::cerr << "CUDA decode failed: " << err << " (no CPU decoder in this build)\n";
            jbp_destroy(ctx);
            return 1;
        }
        double latency = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        bool ok = res.syndrome_match != 0, exact = false, stab = false, success = ok;
        if (!errv.empty()) {
            exact = (est == errv);
            if (ok) {
                std::vector<int> dxv(n), dzv(n);
                for (int v = 0; v < n; ++v) {
                    dxv[v] = (errv[v] ^ est[v]) & 1;
                    dzv[v] = ((errv[v] ^ est[v]) >> 1) & 1;
                }
                uint8_t in_z = 0, in_x = 0;
                if (jbp_in_row_space(ctx, 1, 1, dxv.data(), &in_z, err, sizeof(err)) != JBP_OK ||
                    jbp_in_row_space(ctx, 0, 1, dzv.data(), &in_x, err, sizeof(err)) != JBP_OK) {
                    std::cerr << "residual test failed: " << err << "\n";
                    jbp_destroy(ctx);
                    return 1;
                }
                stab = in_z && in_x;
                success = ok && stab;
            }
        }
        // jointbp_ets.cpp:5740-5745, 5794-5801
        const bool pp_success = success && (ppf & JBP_PPF_PP_SUCCESS);
        section("Decode Result");
        std::cout << "iterations=" << res.iterations << " syndrome_match=" << tf(ok) << "\n"
                  << "decode_success=" << tf(success) << "\npp_enabled=" << tf(enable_pp) << "\n"
                  << "pp_success=" << tf(pp_success) << "\npp_success_ets="
                  << tf(success && (ppf & JBP_PPF_ETS_SAVED)) << "\npp_success_flip="
                  << tf(success && (ppf & JBP_PPF_FLIP_SAVED)) << "\navg_latency=" << fmt_latency(latency) << "\n";
        if (!errv.empty())
            std::cout << "exact_match=" << tf(exact) << "\nstabilizer_equiv=" << tf(stab) << "\n";
        if (!est_path.empty()) {
            if (!write_ints(est_path, est)) {
                std::cerr << "Failed to write estimate to " << est_path << "\n";
                jbp_destroy(ctx);
                return 1;
            }
            std::cout << "Saved estimate: " << est_path << "\n";
        }
        if (!hist_path.empty()) {
            std::vector<long long> hist(static_cast<size_t>(max_iter) + 1, 0);
            if (res.iterations >= 1 && res.iterations <= max_iter) hist[res.iterations] = 1;
            if (!write_hist(hist_path, hist, 1)) std::cerr << "Failed to write histogram output: " << hist_path << "\n";
        }
        jbp_destroy(ctx);
        return success ? 0 : 2;
    }

    if (trials <= 0) {
        std::cerr << "--trials must be positive in --simulate mode\n";
        jbp_destroy(ctx);
        return 1;
    }
    if (trials != 1 && !est_path.empty()) {
        std::cerr << "--est is only supported with --trials 1\n";
        jbp_destroy(ctx);
        return 1;
    }
    const long long first_idx = trial_index >= 0 ? trial_index : 1;
    const long long block = report_every > 0
                                ? std::max<long long>(1, (65536 / report_every)) * report_every
                                : 65536;
    std::vector<long long> hist(static_cast<size_t>(max_iter) + 1, 0);
    Progress pr;
    long long exact_total = 0;
    auto start = std::chrono::steady_clock::now();
    std::vector<int> it;
    std::vector<uint8_t> okv, succ, exv, ppv;
    while (pr.done < trials) {
        if (min_fails > 0 && pr.failures >= min_fails) break;
        const long long cnt = std::min(block, trials - pr.done);
        it.assign(cnt, 0);
        okv.assign(cnt, 0);
        succ.assign(cnt, 0);
        exv.assign(cnt, 0);
        ppv.assign(cnt, 0);
        jbp_trial_out out = {it.data(), okv.data(), succ.data(), exv.data(), nullptr, nullptr, ppv.data()};
        jbp_sim_cfg cfg = {seed, first_idx + pr.done, cnt, p_err, max_iter, damping, precision,
                           warmup, interval, 0, 0, min_fails > 0 ? min_fails - pr.failures : 0};
        cfg.pp = enable_pp ? 1 : 0;
        cfg.flip_hist = flip_hist;
        cfg.freeze_syn = freeze_syn ? 1 : 0;
        jbp_run_stats st;
        if (jbp_simulate(ctx, &cfg, &out, &st, err, sizeof(err)) != JBP_OK) {
            std::cerr << "CUDA decode failed: " << err << " (no CPU decoder in this build)\n";
            jbp_destroy(ctx);
            return 1;
        }
        double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        for (long long i = 0; i < st.done; ++i) {
            const long long idx = first_idx + pr.done;
            const bool ok = okv[i] != 0, s = succ[i] != 0, ex = exv[i] != 0;
            const uint8_t f = ppv[i];
            const bool bp_ok = (f & JBP_PPF_BP_OK) != 0;
            pr.done++;
            pr.iters += it[i];
            pr.failures += s ? 0 : 1;
            pr.bp_fail += bp_ok ? 0 : 1;
            pr.stab += (s && !ex) ? 1 : 0;
            // the Run Summary counters (jointbp_ets.cpp:6113-6228) from the
            // per-trial flags
            pr.pp += (f & JBP_PPF_PP_SUCCESS) ? 1 : 0;
            pr.flip += (f & JBP_PPF_FLIP_SAVED) ? 1 : 0;
            pr.ets_used += (f & JBP_PPF_ETS_USED) ? 1 : 0;
            pr.ets6_used += (f & JBP_PPF_ETS6) ? 1 : 0;
            pr.ets12_used += (f & JBP_PPF_ETS12) ? 1 : 0;
            if (f & JBP_PPF_ETS_SAVED) {
                pr.ets_saved++;
                pr.ets6_saved += (f & JBP_PPF_ETS6) ? 1 : 0;
                pr.ets12_saved += (f & JBP_PPF_ETS12) ? 1 : 0;
            }
            exact_total += ex ? 1 : 0;
            if (it[i] >= 1 && it[i] <= max_iter) hist[it[i]]++;
            if (report_fail && !s) {
                std::cout << "[fail] trial=" << idx << " iter=" << it[i] << " bp_ok=" << tf(bp_ok)
                          << " syndrome_match=" << tf(ok) << " ets6=" << tf(f & JBP_PPF_ETS6)
                          << " ets12=" << tf(f & JBP_PPF_ETS12) << " trial_seed="
                          << splitmix64(seed ^ (static_cast<uint64_t>(idx) * 0x9e3779b97f4a7c15ull)) << "\n";
            }
            if (report_every > 0 && pr.done % report_every == 0)
                report_progress(pr, elapsed, k, elapsed / pr.done);
        }
        if (st.done < cnt) break;  // --min-fails reached
    }
    double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    double fer = pr.done ? static_cast<double>(pr.failures) / pr.done : 0.0;
    double lo = 0, hi = 0;
    bool has_ci = wilson(pr.done, pr.failures, 1.96, lo, hi);
    section("Run Summary");
    std::cout << "trials=" << pr.done << " failures=" << pr.failures << " bp_fail=" << pr.bp_fail
              << " pp_success=" << pr.pp << " flip_pp_success=" << pr.flip << " stab_success=" << pr.stab
              << " ets_used=" << pr.ets_used << " ets_saved=" << pr.ets_saved << " ets6_used=" << pr.ets6_used
              << " ets6_saved=" << pr.ets6_saved << " ets12_used=" << pr.ets12_used
              << " ets12_saved=" << pr.ets12_saved << " fer=" << fixed(fer, 8);
    if (has_ci) std::cout << " ci95=[" << fixed(lo, 8) << "," << fixed(hi, 8) << "]";
    std::cout << " iters=" << pr.iters << " avg_iter=" << fixed(pr.done ? static_cast<double>(pr.iters) / pr.done : 0.0, 2)
              << " avg_latency=" << fmt_latency(pr.done ? elapsed / pr.done : 0.0)
              << " exact_rate=" << fixed(pr.done ? static_cast<double>(exact_total) / pr.done : 0.0, 6)
              << " elapsed_s=" << fixed(elapsed, 2) << "s\n";
    if (!hist_path.empty() && !write_hist(hist_path, hist, pr.done))
        std::cerr << "Failed to write histogram output: " << hist_path << "\n";
    jbp_destroy(ctx);
    jbp_code_free(code);
    return 0;
}
