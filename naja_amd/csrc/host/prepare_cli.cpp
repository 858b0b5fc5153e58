// `naja prepare`: produces the conditioned model directories the sampler
// reads (cmd_prepare.cpp:213-457), with the start-point LP run by the
// library's own interior-point solver instead of Gurobi.
//
// Per model of --model-list under --models-root:
//   * with --conditioner-cmd: run it once per model (unless gem/ is present
//     with a matching conditioning.signature), check its reaction ids against
//     the base model's, validate both bound sets, and write the tightened
//     bounds as extra rows over the base model's rounded variables
//     (naja_extra_from_bounds; extra_A/extra_b CSV, or no files when nothing
//     tightened), then the signature;
//   * drop an empty extra pair (inherit.cpp:14-29); inherit the base model's
//     rounding files by symlink or copy when its own are missing
//     (inherit.cpp:31-71, with INHERITED_FROM.txt);
//   * after all models: where extra rows exist and the current start violates
//     them (1e-9), replace the start by the cube centre of the augmented
//     polytope (feasible_start_files.cpp:41-90), models in parallel.
// Bulk-mode config generation (--out-bulk-config) belongs to the reference's
// multi-model job runner, which this build does not have; the flag is
// rejected with that reason.
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <exception>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/naja_amd.h"
#include "host_io.h"

using namespace naja::host;

namespace {

struct Contract {
    std::string model_dir, rounding_dir, gem_dir, model_name;
    std::string prefix() const { return rounding_dir + "/" + model_name + "_rounding"; }
};

[[noreturn]] void usage(const std::string& msg)
{
    std::cerr << "error: " << msg << "\n\nusage:\n  naja prepare --models-root <dir> --model-list <file> "
                 "[--base-model-dir <dir>] [--mode symlink|copy] [--out-model-list <file>] [--dry-run] "
                 "[--conditioner-cmd <cmd>] [--conditioner-arg <arg>]... [--skip-conditioning-if-gem-present] "
                 "[--force-conditioning] [--verbose]\n";
    std::exit(2);
}

std::string arg_value(int& i, int argc, char** argv, const std::string& flag)
{
    if (i + 1 >= argc) usage("missing value for " + flag);
    return std::string(argv[++i]);
}

// model_contract.cpp:53-67
Contract parse_model_dir(const std::string& dir)
{
    std::string md = dir;
    if (!md.empty() && md.back() == '/') md.pop_back();
    if (!is_directory(md)) throw std::runtime_error("model-dir is not a directory: " + md);
    const size_t slash = md.find_last_of('/');
    Contract c;
    c.model_dir = md;
    c.rounding_dir = md + "/rounding";
    c.gem_dir = md + "/gem";
    c.model_name = slash == std::string::npos ? md : md.substr(slash + 1);
    return c;
}

// model_contract.cpp:69-84 (backmap = true: T and shift required)
void validate_contract(const Contract& c)
{
    if (!is_directory(c.rounding_dir)) throw std::runtime_error("missing rounding/ dir: " + c.rounding_dir);
    const std::string p = c.prefix() + "_";
    require_nonempty_file(p + "A.csv", "rounding_A");
    require_nonempty_file(p + "b.csv", "rounding_b");
    require_nonempty_file(p + "start.csv", "rounding_start");
    require_nonempty_file(p + "T.csv", "rounding_T");
    require_nonempty_file(p + "shift.csv", "rounding_shift");
    const bool ha = path_exists(p + "extra_A.csv"), hb = path_exists(p + "extra_b.csv");
    if (ha != hb)
        throw std::runtime_error("extra constraints must be both-present or both-absent: " + p + "extra_A.csv / " + p +
                                 "extra_b.csv");
    if (ha) {
        require_nonempty_file(p + "extra_A.csv", "extra_A");
        require_nonempty_file(p + "extra_b.csv", "extra_b");
    }
}

std::vector<std::string> load_model_list(const std::string& path)
{
    require_nonempty_file(path, "model list");
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("cannot open model list: " + path);
    std::vector<std::string> out;
    std::string line;
    auto ws = [](unsigned char ch) { return ch == ' ' || ch == '\t' || ch == '\r' || ch == '\n'; };
    while (std::getline(f, line)) {
        while (!line.empty() && ws((unsigned char)line.front())) line.erase(line.begin());
        while (!line.empty() && ws((unsigned char)line.back())) line.pop_back();
        if (!line.empty() && line[0] != '#') out.push_back(line);
    }
    if (out.empty()) throw std::runtime_error("no models in list: " + path);
    return out;
}

std::vector<std::string> trimmed_lines(const std::string& path)
{
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("cannot read: " + path);
    std::vector<std::string> out;
    std::string line;
    auto ws = [](unsigned char ch) { return ch == ' ' || ch == '\t' || ch == '\r' || ch == '\n'; };
    while (std::getline(f, line)) {
        while (!line.empty() && ws((unsigned char)line.front())) line.erase(line.begin());
        while (!line.empty() && ws((unsigned char)line.back())) line.pop_back();
        if (!line.empty()) out.push_back(line);
    }
    return out;
}

std::string read_all(const std::string& path)
{
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("cannot read: " + path);
    std::ostringstream s;
    s << f.rdbuf();
    return s.str();
}

bool file_is_empty(const std::string& path)
{
    struct stat st;
    if (stat(path.c_str(), &st) != 0) throw std::runtime_error("cannot stat: " + path);
    return st.st_size == 0;
}

void remove_if_exists(const std::string& path)
{
    if (!path_exists(path)) {
        struct stat st;  // a dangling symlink still has to go
        if (lstat(path.c_str(), &st) != 0) return;
    }
    if (std::remove(path.c_str()) != 0) throw std::runtime_error("cannot remove: " + path);
}

void write_matrix_csv(const std::string& path, const std::vector<double>& M, int rows, int cols, int prec)
{
    std::ofstream f(path);
    if (!f.is_open()) throw std::runtime_error("cannot write: " + path);
    f << std::setprecision(prec);
    for (int i = 0; i < rows; ++i) {
        for (int j = 0; j < cols; ++j) {
            if (j) f << ",";
            f << M[(size_t)i * cols + j];
        }
        f << "\n";
    }
}

void write_vector_csv(const std::string& path, const std::vector<double>& v, int prec)
{
    std::ofstream f(path);
    if (!f.is_open()) throw std::runtime_error("cannot write: " + path);
    f << std::setprecision(prec);
    for (double x : v) f << x << "\n";
}

void napi(int rc, const char* what)
{
    if (rc != NAJA_OK) throw std::runtime_error(std::string(what) + ": " + naja_last_error());
}

// cmd_prepare.cpp:43-57 (std::hash of the base reaction-id text)
std::string conditioning_signature(const std::string& cmd, const std::vector<std::string>& args,
                                   const std::string& base_rxn_path, const std::string& row_id)
{
    std::string sig = cmd;
    for (const auto& a : args) sig.append("\nARG ").append(a);
    sig.append("\nROW_ID ").append(row_id);
    sig.append("\nBASE_RXN_SHA ").append(std::to_string(std::hash<std::string>{}(read_all(base_rxn_path))));
    return sig;
}

// cmd_prepare.cpp:80-93
void validate_bounds(const std::vector<double>& lb, const std::vector<double>& ub, size_t expected)
{
    if (lb.size() != ub.size()) throw std::runtime_error("l_bounds size != u_bounds size");
    if (lb.size() != expected)
        throw std::runtime_error("bounds length mismatch: got " + std::to_string(lb.size()) + " expected " +
                                 std::to_string(expected));
    for (size_t i = 0; i < lb.size(); ++i) {
        if (!std::isfinite(lb[i]) || !std::isfinite(ub[i]))
            throw std::runtime_error("non-finite bound at index " + std::to_string(i));
        if (lb[i] > ub[i]) throw std::runtime_error("l_bounds > u_bounds at index " + std::to_string(i));
    }
}

// cmd_prepare.cpp:111-156 over the C ABI
void build_extra_from_bounds(const Contract& c, const std::string& base_prefix, const std::vector<double>& lb_base,
                             const std::vector<double>& ub_base, const std::vector<double>& lb_new,
                             const std::vector<double>& ub_new)
{
    const Matrix T = load_csv_matrix(base_prefix + "_T.csv");
    const std::vector<double> shift = load_csv_vector(base_prefix + "_shift.csv");
    if ((size_t)T.rows != lb_base.size()) throw std::runtime_error("T rows != base bounds length");
    if (shift.size() != (size_t)T.rows) throw std::runtime_error("shift length != T rows");
    std::vector<double> EA((size_t)2 * T.rows * T.cols), Eb((size_t)2 * T.rows);
    int k = 0;
    napi(naja_extra_from_bounds(T.v.data(), shift.data(), T.rows, T.cols, lb_base.data(), ub_base.data(),
                                lb_new.data(), ub_new.data(), 1e-9, EA.data(), Eb.data(), &k),
         "build_extra_from_bounds");
    const std::string ea = c.prefix() + "_extra_A.csv", eb = c.prefix() + "_extra_b.csv";
    if (k == 0) {
        remove_if_exists(ea);
        remove_if_exists(eb);
        return;
    }
    Eb.resize((size_t)k);
    write_matrix_csv(ea, EA, k, T.cols, 12);
    write_vector_csv(eb, Eb, 12);
}

// inherit.cpp:14-29
void normalize_extra_constraints(const Contract& c)
{
    const std::string ea = c.prefix() + "_extra_A.csv", eb = c.prefix() + "_extra_b.csv";
    const bool ha = path_exists(ea), hb = path_exists(eb);
    if (ha != hb) throw std::runtime_error("extra constraints must be both-present or both-absent: " + ea + " / " + eb);
    if (!ha) return;
    if (file_is_empty(ea) || file_is_empty(eb)) {
        remove_if_exists(ea);
        remove_if_exists(eb);
    }
}

bool rounding_present_enough(const Contract& c)
{
    if (!is_directory(c.rounding_dir)) return false;
    for (const char* s : {"_A.csv", "_b.csv", "_start.csv", "_T.csv", "_shift.csv"})
        if (!path_exists(c.prefix() + s)) return false;
    return true;
}

// inherit.cpp:31-71
void inherit_rounding(const Contract& base, const Contract& target, const std::string& mode)
{
    if (mode != "symlink" && mode != "copy") throw std::runtime_error("invalid inherit mode: " + mode);
    ensure_dir(target.rounding_dir);
    for (const char* suf : {"A.csv", "b.csv", "T.csv", "shift.csv", "start.csv"}) {
        const std::string src = base.model_dir + "/rounding/" + base.model_name + "_rounding_" + suf;
        const std::string dst = target.rounding_dir + "/" + target.model_name + "_rounding_" + suf;
        require_nonempty_file(src, std::string("base rounding ") + suf);
        remove_if_exists(dst);
        if (mode == "symlink") {
            if (symlink(src.c_str(), dst.c_str()) != 0) throw std::runtime_error("symlink failed: " + dst + " <- " + src);
        } else {
            std::ifstream in(src, std::ios::binary);
            std::ofstream out(dst, std::ios::binary);
            if (!in.is_open() || !out.is_open()) throw std::runtime_error("copy failed: " + dst + " <- " + src);
            out << in.rdbuf();
        }
    }
    std::ofstream f(target.rounding_dir + "/INHERITED_FROM.txt");
    if (!f.is_open()) throw std::runtime_error("cannot write provenance file: " + target.rounding_dir + "/INHERITED_FROM.txt");
    f << "base_model_dir=" << base.model_dir << "\ntarget_model_dir=" << target.model_dir << "\nmode=" << mode
      << "\ncreated_at=" << current_timestamp() << "\n";
}

bool feasible(const Matrix& A, const std::vector<double>& b, const std::vector<double>& x, double eps)
{
    for (int i = 0; i < A.rows; ++i) {
        double acc = 0.0;
        for (int j = 0; j < A.cols; ++j) acc += A.at(i, j) * x[(size_t)j];
        if (acc > b[(size_t)i] + eps) return false;
    }
    return true;
}

// feasible_start_files.cpp:41-90
void ensure_feasible_start(const Contract& c)
{
    const std::string p = c.prefix();
    const bool ha = path_exists(p + "_extra_A.csv"), hb = path_exists(p + "_extra_b.csv");
    if (!ha && !hb) return;
    if (ha != hb)
        throw std::runtime_error("extra constraints must be both-present or both-absent: " + p + "_extra_A.csv / " + p +
                                 "_extra_b.csv");
    Matrix A = load_csv_matrix(p + "_A.csv");
    std::vector<double> b = load_csv_vector(p + "_b.csv");
    const Matrix E = load_csv_matrix(p + "_extra_A.csv");
    const std::vector<double> f = load_csv_vector(p + "_extra_b.csv");
    if ((size_t)E.rows != f.size()) throw std::runtime_error("extra constraint matrices have mismatched dimensions");
    if (E.cols != A.cols) throw std::runtime_error("extra constraint matrix has wrong number of columns");
    A.v.insert(A.v.end(), E.v.begin(), E.v.end());
    A.rows += E.rows;
    b.insert(b.end(), f.begin(), f.end());
    const std::string start = p + "_start.csv";
    if (path_exists(start)) {
        try {
            const std::vector<double> x = load_csv_vector(start);
            if ((int)x.size() == A.cols && feasible(A, b, x, 1e-9)) return;
        } catch (const std::exception&) {
            // recompute below
        }
    }
    std::vector<double> x((size_t)A.cols);
    double r = 0.0;
    napi(naja_cube_center(A.v.data(), b.data(), A.rows, A.cols, nullptr, x.data(), &r, 1), "axis_aligned_cube_center");
    if (!feasible(A, b, x, 1e-9)) throw std::runtime_error("feasible_start_files: cube centre violates A x <= b");
    remove_if_exists(start);  // may be a symlink into the base model
    write_vector_csv(start, x, 17);
}

void ensure_feasible_starts_parallel(const std::vector<Contract>& cs)
{
    if (cs.empty()) return;
    size_t workers = 1;
    if (cs.size() > 1) {
        const unsigned hw = std::thread::hardware_concurrency();
        workers = std::max<size_t>(2, std::min<size_t>(std::min<size_t>(hw ? hw : 8, 16), cs.size()));
    }
    std::atomic<size_t> next{0};
    std::exception_ptr first;
    std::mutex mu;
    auto work = [&]() {
        for (size_t i; (i = next.fetch_add(1)) < cs.size();) {
            try {
                ensure_feasible_start(cs[i]);
                validate_contract(cs[i]);
            } catch (...) {
                std::lock_guard<std::mutex> g(mu);
                if (!first) first = std::current_exception();
            }
        }
    };
    std::vector<std::thread> pool;
    for (size_t t = 0; t < workers; ++t) pool.emplace_back(work);
    for (auto& t : pool) t.join();
    if (first) std::rethrow_exception(first);
}

}  // namespace

int cmd_prepare(int argc, char** argv)
{
    std::string models_root, model_list, base_model_dir, mode = "symlink", out_model_list, conditioner_cmd;
    std::vector<std::string> conditioner_args;
    bool dry_run = false, skip_if_gem = true, force = false, verbose = false;
    for (int i = 0; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--models-root") models_root = arg_value(i, argc, argv, a);
        else if (a == "--model-list") model_list = arg_value(i, argc, argv, a);
        else if (a == "--base-model-dir") base_model_dir = arg_value(i, argc, argv, a);
        else if (a == "--mode") mode = arg_value(i, argc, argv, a);
        else if (a == "--out-model-list") out_model_list = arg_value(i, argc, argv, a);
        else if (a == "--dry-run") dry_run = true;
        else if (a == "--conditioner-cmd") conditioner_cmd = arg_value(i, argc, argv, a);
        else if (a == "--conditioner-arg") conditioner_args.push_back(arg_value(i, argc, argv, a));
        else if (a == "--skip-conditioning-if-gem-present") skip_if_gem = true;
        else if (a == "--force-conditioning") force = true;
        else if (a == "--verbose") verbose = true;
        else if (a == "--out-bulk-config")
            usage("--out-bulk-config configures the reference's bulk job runner, which this build does not have");
        else usage("unknown flag: " + a);
    }
    (void)verbose;
    if (models_root.empty()) usage("missing --models-root");
    if (model_list.empty()) usage("missing --model-list");
    if (!base_model_dir.empty() && mode != "symlink" && mode != "copy") usage("invalid --mode: " + mode);
    if (!is_directory(models_root)) throw std::runtime_error("models-root is not a directory: " + models_root);
    const std::vector<std::string> models = load_model_list(model_list);
    const bool do_inherit = !base_model_dir.empty();
    Contract base;
    if (do_inherit) {
        base = parse_model_dir(base_model_dir);
        validate_contract(base);
    }
    if (!conditioner_cmd.empty() && base_model_dir.empty()) usage("--conditioner-cmd requires --base-model-dir");

    std::vector<std::string> prepared;
    std::vector<Contract> for_start;
    for (const auto& name : models) {
        const Contract c = parse_model_dir(models_root + "/" + name);
        const std::string rxn = c.gem_dir + "/reaction_ids.txt", lb = c.gem_dir + "/l_bounds.csv",
                          ub = c.gem_dir + "/u_bounds.csv";
        const bool gem_present = is_directory(c.gem_dir) && path_exists(rxn) && path_exists(lb) && path_exists(ub);
        if (!conditioner_cmd.empty()) {
            const std::string sig =
                conditioning_signature(conditioner_cmd, conditioner_args, base_model_dir + "/gem/reaction_ids.txt", name);
            const std::string sig_path = c.gem_dir + "/conditioning.signature";
            bool need = force || !gem_present || !skip_if_gem;
            if (gem_present && skip_if_gem && !force) {
                if (path_exists(sig_path)) {
                    if (read_all(sig_path) != sig)
                        throw std::runtime_error("conditioning signature mismatch; use --force-conditioning to overwrite");
                    need = false;
                } else {
                    need = true;  // provenance unknown: condition again
                }
            }
            if (dry_run) {
                prepared.push_back(name);
                continue;
            }
            if (need) {
                ensure_dir(c.gem_dir);
                std::string cmd = conditioner_cmd + " --base-model-dir " + base_model_dir + " --out-model-dir " +
                                  c.model_dir + " --row-id " + name;
                for (const auto& a : conditioner_args) cmd += " " + a;
                const int rc = std::system(cmd.c_str());
                if (rc != 0) throw std::runtime_error("conditioner failed with exit code " + std::to_string(rc));
            }
            if (!path_exists(rxn) || !path_exists(lb) || !path_exists(ub))
                throw std::runtime_error("conditioner did not produce required gem files for " + c.model_dir);
            const auto base_rxn = trimmed_lines(base_model_dir + "/gem/reaction_ids.txt");
            if (base_rxn != trimmed_lines(rxn)) throw std::runtime_error("reaction_ids mismatch vs base for " + c.model_dir);
            const std::vector<double> lbv = load_csv_vector(lb), ubv = load_csv_vector(ub);
            validate_bounds(lbv, ubv, base_rxn.size());
            const std::vector<double> lbb = load_csv_vector(base_model_dir + "/gem/l_bounds.csv"),
                                      ubb = load_csv_vector(base_model_dir + "/gem/u_bounds.csv");
            validate_bounds(lbb, ubb, lbv.size());
            ensure_dir(c.rounding_dir);
            build_extra_from_bounds(c, base_model_dir + "/rounding/" + base.model_name + "_rounding", lbb, ubb, lbv, ubv);
            std::ofstream sf(sig_path);
            if (!sf.is_open()) throw std::runtime_error("cannot write signature: " + sig_path);
            sf << sig;
        } else if (!gem_present) {
            throw std::runtime_error("gem/ missing and no conditioner provided for " + c.model_dir);
        }
        if (!dry_run) {
            ensure_dir(c.rounding_dir);
            normalize_extra_constraints(c);
        }
        if (!rounding_present_enough(c)) {
            if (!do_inherit)
                throw std::runtime_error("rounding contract missing/invalid and no --base-model-dir provided: " +
                                         c.model_dir);
            if (!dry_run) {
                inherit_rounding(base, c, mode);
                normalize_extra_constraints(c);
            }
        }
        validate_contract(c);
        if (!dry_run && do_inherit && !path_exists(c.prefix() + "_extra_A.csv") &&
            path_exists(c.gem_dir + "/l_bounds.csv") && path_exists(base_model_dir + "/gem/l_bounds.csv") &&
            read_all(c.gem_dir + "/l_bounds.csv") != read_all(base_model_dir + "/gem/l_bounds.csv"))
            std::cerr << "*** WARNING: " << c.model_name
                      << " has different gem bounds from base but NO extra constraints. ***\n"
                      << "  Sampling will use the unconditioned base polytope.\n"
                      << "  To fix: rerun prepare with --conditioner-cmd, or condition + prepare with "
                         "--force-conditioning.\n";
        if (!dry_run) for_start.push_back(c);
        prepared.push_back(name);
    }
    if (!dry_run) ensure_feasible_starts_parallel(for_start);
    if (!out_model_list.empty() && !dry_run) {
        std::ofstream f(out_model_list);
        if (!f.is_open()) throw std::runtime_error("cannot write out-model-list: " + out_model_list);
        for (const auto& m : prepared) f << m << "\n";
    }
    std::cout << "OK\nmodels_root     :: " << make_absolute_path(models_root)
              << "\nmodel_list      :: " << make_absolute_path(model_list) << "\nprepared        :: " << prepared.size()
              << " / " << models.size() << "\n";
    if (!out_model_list.empty()) std::cout << "out_model_list  :: " << make_absolute_path(out_model_list) << "\n";
    if (dry_run) std::cout << "dry_run         :: true\n";
    return 0;
}
