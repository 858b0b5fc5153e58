// `naja prepare`: turns a list of model directories into directories the
// sampler can read, with conditioning expressed as extra rows over a shared
// base model's rounded variables.
//
// The observable contract is the reference's (cmd_prepare.cpp:213-457,
// inherit.cpp:14-71, feasible_start_files.cpp:41-90): its flags, the
// conditioner's command line, the gem/ and rounding/ file names, the
// conditioning.signature text and the INHERITED_FROM.txt keys, the summary
// lines. The driver itself is organised differently:
//
//   * every model is a PrepareJob (its paths, whether gem/ is complete, what
//     happened to it); the base model is loaded once, lazily, as a BaseModel
//     (reaction ids, bounds, T and shift);
//   * stage 1 runs per job in list order: decide whether the conditioner
//     must run (signature check), run it, and re-derive the extra rows from
//     the conditioned bounds through the library (naja_extra_from_bounds);
//   * stage 2 completes each job's rounding file set: an empty extra pair is
//     dropped, missing core files are materialised from the base (symlink or
//     copy), and the set is checked;
//   * stage 3 repairs, on a pool of host threads, the start of every job
//     whose extra rows cut off its current start (cube centre of the
//     augmented polytope by the library's own interior-point LP, in place of
//     the reference's Gurobi model).
// `--out-bulk-config` configures the reference's multi-model job runner,
// which this build does not have, and is rejected with that reason.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <exception>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/naja_amd.h"
#include "host_io.h"

namespace fs = std::filesystem;
using naja::host::Matrix;

namespace {

// ---------------------------------------------------------------------------
// Files of one model directory.

enum class Role { A, b, start, T, shift, extra_A, extra_b };
constexpr Role kCoreRoles[] = {Role::A, Role::b, Role::T, Role::shift, Role::start};

const char* role_suffix(Role r)
{
    switch (r) {
    case Role::A: return "A.csv";
    case Role::b: return "b.csv";
    case Role::start: return "start.csv";
    case Role::T: return "T.csv";
    case Role::shift: return "shift.csv";
    case Role::extra_A: return "extra_A.csv";
    case Role::extra_b: return "extra_b.csv";
    }
    return "";
}

struct ModelPaths {
    std::string dir;   // as given, without a trailing slash
    std::string name;  // last path component (model_contract.cpp:53-67)

    static ModelPaths of(std::string d)
    {
        while (d.size() > 1 && d.back() == '/') d.pop_back();
        if (!naja::host::is_directory(d)) throw std::runtime_error("model-dir is not a directory: " + d);
        const size_t cut = d.find_last_of('/');
        return ModelPaths{d, cut == std::string::npos ? d : d.substr(cut + 1)};
    }
    std::string rounding_dir() const { return dir + "/rounding"; }
    std::string gem_dir() const { return dir + "/gem"; }
    std::string file(Role r) const { return rounding_dir() + "/" + name + "_rounding_" + role_suffix(r); }
    std::string gem(const char* leaf) const { return gem_dir() + "/" + leaf; }
};

bool exists_or_dangling(const std::string& p)
{
    std::error_code ec;
    return fs::symlink_status(p, ec).type() != fs::file_type::not_found;
}

bool has_file(const std::string& p)
{
    std::error_code ec;
    return fs::exists(p, ec);
}

bool zero_bytes(const std::string& p)
{
    std::error_code ec;
    const auto n = fs::file_size(p, ec);
    if (ec) throw std::runtime_error("cannot stat: " + p);
    return n == 0;
}

void unlink_path(const std::string& p)
{
    std::error_code ec;
    fs::remove(p, ec);  // removes a symlink itself, never its target
    if (ec) throw std::runtime_error("cannot remove: " + p);
}

std::string slurp(const std::string& p)
{
    std::ifstream in(p, std::ios::binary);
    if (!in) throw std::runtime_error("cannot read: " + p);
    return std::string(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
}

void spit(const std::string& p, const std::string& text, const char* what)
{
    std::ofstream out(p, std::ios::binary | std::ios::trunc);
    if (!out) throw std::runtime_error(std::string("cannot write ") + what + ": " + p);
    out << text;
}

// Non-empty lines with surrounding blanks stripped; '#' lines dropped when
// `comments` (the model list format).
std::vector<std::string> text_lines(const std::string& p, bool comments)
{
    std::vector<std::string> out;
    std::istringstream in(slurp(p));
    for (std::string line; std::getline(in, line);) {
        const size_t a = line.find_first_not_of(" \t\r\n");
        if (a == std::string::npos) continue;
        const size_t z = line.find_last_not_of(" \t\r\n");
        std::string t = line.substr(a, z - a + 1);
        if (comments && t[0] == '#') continue;
        out.push_back(std::move(t));
    }
    return out;
}

// The CSV layout the reference writes for extra rows: 12 significant digits.
std::string csv_text(const double* v, int rows, int cols, int digits)
{
    std::ostringstream o;
    o << std::setprecision(digits);
    for (int r = 0; r < rows; ++r) {
        for (int c = 0; c < cols; ++c) o << (c ? "," : "") << v[(size_t)r * cols + c];
        o << "\n";
    }
    return o.str();
}

// The extra pair must be both present or both absent, each non-empty
// (model_contract.cpp:30-40).
void check_extra_pair(const ModelPaths& m, bool require_nonempty)
{
    const bool ea = has_file(m.file(Role::extra_A)), eb = has_file(m.file(Role::extra_b));
    if (ea != eb)
        throw std::runtime_error("extra constraints must be both-present or both-absent: " + m.file(Role::extra_A) +
                                 " / " + m.file(Role::extra_b));
    if (ea && require_nonempty) {
        naja::host::require_nonempty_file(m.file(Role::extra_A), "extra_A");
        naja::host::require_nonempty_file(m.file(Role::extra_b), "extra_b");
    }
}

// The sampler's contract with T and shift required (model_contract.cpp:69-84).
void check_contract(const ModelPaths& m)
{
    if (!naja::host::is_directory(m.rounding_dir()))
        throw std::runtime_error("missing rounding/ dir: " + m.rounding_dir());
    for (Role r : {Role::A, Role::b, Role::start, Role::T, Role::shift})
        naja::host::require_nonempty_file(m.file(r), std::string("rounding_") + role_suffix(r));
    check_extra_pair(m, true);
}

// ---------------------------------------------------------------------------
// Options.

struct Options {
    std::string models_root, model_list, base_dir, mode = "symlink", out_list, cond_cmd;
    std::vector<std::string> cond_args;
    bool dry_run = false, force = false;
};

[[noreturn]] void die_usage(const std::string& why)
{
    std::cerr << "error: " << why
              << "\n\nusage:\n  naja prepare --models-root <dir> --model-list <file> [--base-model-dir <dir>] "
                 "[--mode symlink|copy] [--out-model-list <file>] [--dry-run] [--conditioner-cmd <cmd>] "
                 "[--conditioner-arg <arg>]... [--skip-conditioning-if-gem-present] [--force-conditioning] "
                 "[--verbose]\n";
    std::exit(2);
}

Options parse_options(int argc, char** argv)
{
    Options o;
    const std::map<std::string, std::function<void(const std::string&)>> valued = {
        {"--models-root", [&](const std::string& v) { o.models_root = v; }},
        {"--model-list", [&](const std::string& v) { o.model_list = v; }},
        {"--base-model-dir", [&](const std::string& v) { o.base_dir = v; }},
        {"--mode", [&](const std::string& v) { o.mode = v; }},
        {"--out-model-list", [&](const std::string& v) { o.out_list = v; }},
        {"--conditioner-cmd", [&](const std::string& v) { o.cond_cmd = v; }},
        {"--conditioner-arg", [&](const std::string& v) { o.cond_args.push_back(v); }},
    };
    const std::map<std::string, std::function<void()>> switches = {
        {"--dry-run", [&] { o.dry_run = true; }},
        {"--force-conditioning", [&] { o.force = true; }},
        {"--skip-conditioning-if-gem-present", [] {}},  // the default; kept for the flag set
        {"--verbose", [] {}},
    };
    for (int i = 0; i < argc; ++i) {
        const std::string flag = argv[i];
        if (auto s = switches.find(flag); s != switches.end()) {
            s->second();
        } else if (auto v = valued.find(flag); v != valued.end()) {
            if (i + 1 >= argc) die_usage("missing value for " + flag);
            v->second(argv[++i]);
        } else if (flag == "--out-bulk-config") {
            die_usage("--out-bulk-config configures the reference's bulk job runner, which this build does not have");
        } else {
            die_usage("unknown flag: " + flag);
        }
    }
    if (o.models_root.empty()) die_usage("missing --models-root");
    if (o.model_list.empty()) die_usage("missing --model-list");
    if (!o.base_dir.empty() && o.mode != "symlink" && o.mode != "copy") die_usage("invalid --mode: " + o.mode);
    return o;
}

// ---------------------------------------------------------------------------
// The base model and the conditioner.

struct BaseModel {
    ModelPaths paths;
    // Loaded on first use (only conditioning needs them).
    std::optional<std::vector<std::string>> rxn;
    std::optional<std::vector<double>> lb, ub;
    std::optional<Matrix> T;
    std::vector<double> shift;

    const std::vector<std::string>& reaction_ids()
    {
        if (!rxn) rxn = text_lines(paths.gem("reaction_ids.txt"), false);
        return *rxn;
    }
    void load_bounds()
    {
        if (!lb) {
            lb = naja::host::load_csv_vector(paths.gem("l_bounds.csv"));
            ub = naja::host::load_csv_vector(paths.gem("u_bounds.csv"));
        }
    }
    const Matrix& rounding_T()
    {
        if (!T) {
            T = naja::host::load_csv_matrix(paths.file(Role::T));
            shift = naja::host::load_csv_vector(paths.file(Role::shift));
        }
        return *T;
    }
};

// Bounds of one model: equal lengths, finite, lb <= ub (cmd_prepare.cpp:80-93).
void check_bounds(const std::vector<double>& lo, const std::vector<double>& hi, size_t n)
{
    if (lo.size() != hi.size()) throw std::runtime_error("l_bounds size != u_bounds size");
    if (lo.size() != n)
        throw std::runtime_error("bounds length mismatch: got " + std::to_string(lo.size()) + " expected " +
                                 std::to_string(n));
    for (size_t i = 0; i < n; ++i) {
        if (!(std::isfinite(lo[i]) && std::isfinite(hi[i])))
            throw std::runtime_error("non-finite bound at index " + std::to_string(i));
        if (lo[i] > hi[i]) throw std::runtime_error("l_bounds > u_bounds at index " + std::to_string(i));
    }
}

struct Conditioner {
    std::string cmd;
    std::vector<std::string> args;

    // The text stored in gem/conditioning.signature: the command, its extra
    // arguments, the row id and the std::hash of the base reaction-id file
    // (the format cmd_prepare.cpp:43-57 writes, so directories prepared by
    // either tool are recognised by the other).
    std::string signature(const std::string& row, const std::string& base_rxn_file) const
    {
        std::ostringstream s;
        s << cmd;
        for (const auto& a : args) s << "\nARG " << a;
        s << "\nROW_ID " << row << "\nBASE_RXN_SHA " << std::hash<std::string>{}(slurp(base_rxn_file));
        return s.str();
    }

    // `<cmd> --base-model-dir <base> --out-model-dir <dir> --row-id <row> <args...>`
    void run(const std::string& base_dir, const ModelPaths& m) const
    {
        std::string line = cmd + " --base-model-dir " + base_dir + " --out-model-dir " + m.dir + " --row-id " + m.name;
        for (const auto& a : args) line += " " + a;
        const int rc = std::system(line.c_str());
        if (rc != 0) throw std::runtime_error("conditioner failed with exit code " + std::to_string(rc));
    }
};

// ---------------------------------------------------------------------------
// Jobs.

struct PrepareJob {
    std::string row;  // the model-list entry
    ModelPaths paths;
    bool gem_complete = false;

    bool gem_files_present() const
    {
        return has_file(paths.gem("reaction_ids.txt")) && has_file(paths.gem("l_bounds.csv")) &&
               has_file(paths.gem("u_bounds.csv"));
    }
};

// Whether the conditioner must (re)run: always when forced or when gem/ is
// incomplete; otherwise only when no signature proves which run produced the
// gem files. A signature from other settings is an error unless forced.
bool conditioner_needed(const PrepareJob& job, const std::string& expected, bool force)
{
    if (force || !job.gem_complete) return true;
    const std::string sig = job.paths.gem("conditioning.signature");
    if (!has_file(sig)) return true;
    if (slurp(sig) != expected)
        throw std::runtime_error("conditioning signature mismatch; use --force-conditioning to overwrite");
    return false;
}

// Extra rows of the conditioned bounds over the base's rounded variables,
// written as the model's extra pair (or removed when nothing tightened).
void derive_extra_rows(const PrepareJob& job, BaseModel& base)
{
    const ModelPaths& m = job.paths;
    if (text_lines(m.gem("reaction_ids.txt"), false) != base.reaction_ids())
        throw std::runtime_error("reaction_ids mismatch vs base for " + m.dir);
    const std::vector<double> lo = naja::host::load_csv_vector(m.gem("l_bounds.csv"));
    const std::vector<double> hi = naja::host::load_csv_vector(m.gem("u_bounds.csv"));
    check_bounds(lo, hi, base.reaction_ids().size());
    base.load_bounds();
    check_bounds(*base.lb, *base.ub, lo.size());
    const Matrix& T = base.rounding_T();
    if ((size_t)T.rows != base.lb->size()) throw std::runtime_error("T rows != base bounds length");
    if (base.shift.size() != (size_t)T.rows) throw std::runtime_error("shift length != T rows");

    std::vector<double> EA((size_t)2 * T.rows * T.cols), Eb((size_t)2 * T.rows);
    int k = 0;
    if (naja_extra_from_bounds(T.v.data(), base.shift.data(), T.rows, T.cols, base.lb->data(), base.ub->data(),
                               lo.data(), hi.data(), 1e-9, EA.data(), Eb.data(), &k) != NAJA_OK)
        throw std::runtime_error(std::string("build_extra_from_bounds: ") + naja_last_error());
    naja::host::ensure_dir(m.rounding_dir());
    if (k == 0) {
        unlink_path(m.file(Role::extra_A));
        unlink_path(m.file(Role::extra_b));
        return;
    }
    spit(m.file(Role::extra_A), csv_text(EA.data(), k, T.cols, 12), "extra_A");
    spit(m.file(Role::extra_b), csv_text(Eb.data(), k, 1, 12), "extra_b");
}

// An extra pair where either file is empty means "no extra rows".
void drop_empty_extra_pair(const ModelPaths& m)
{
    check_extra_pair(m, false);
    if (!has_file(m.file(Role::extra_A))) return;
    if (zero_bytes(m.file(Role::extra_A)) || zero_bytes(m.file(Role::extra_b))) {
        unlink_path(m.file(Role::extra_A));
        unlink_path(m.file(Role::extra_b));
    }
}

bool core_set_present(const ModelPaths& m)
{
    if (!naja::host::is_directory(m.rounding_dir())) return false;
    return std::all_of(std::begin(kCoreRoles), std::end(kCoreRoles), [&](Role r) { return has_file(m.file(r)); });
}

// The base's core rounding files appear in the target under the target's
// names, as symlinks or copies, with a provenance note beside them.
void materialise_from_base(const ModelPaths& base, const ModelPaths& target, const std::string& mode)
{
    naja::host::ensure_dir(target.rounding_dir());
    for (Role r : kCoreRoles) {
        const std::string from = base.file(r), to = target.file(r);
        naja::host::require_nonempty_file(from, std::string("base rounding ") + role_suffix(r));
        if (exists_or_dangling(to)) unlink_path(to);
        std::error_code ec;
        if (mode == "symlink") {
            fs::create_symlink(from, to, ec);
            if (ec) throw std::runtime_error("symlink failed: " + to + " <- " + from);
        } else {
            fs::copy_file(from, to, fs::copy_options::overwrite_existing, ec);
            if (ec) throw std::runtime_error("copy failed: " + to + " <- " + from);
        }
    }
    std::ostringstream note;
    note << "base_model_dir=" << base.dir << "\ntarget_model_dir=" << target.dir << "\nmode=" << mode
         << "\ncreated_at=" << naja::host::current_timestamp() << "\n";
    spit(target.rounding_dir() + "/INHERITED_FROM.txt", note.str(), "provenance file");
}

// ---------------------------------------------------------------------------
// Start repair.

bool inside(const Matrix& A, const std::vector<double>& b, const std::vector<double>& x, double tol)
{
    for (int i = 0; i < A.rows; ++i) {
        double ax = 0.0;
        for (int j = 0; j < A.cols; ++j) ax += A.at(i, j) * x[(size_t)j];
        if (ax > b[(size_t)i] + tol) return false;
    }
    return true;
}

// With extra rows present, a start outside [A; E] x <= [b; f] (1e-9) is
// replaced by the cube centre of that polytope. A symlinked start is
// unlinked first, so the base model's file is never written through.
void repair_start(const ModelPaths& m)
{
    check_extra_pair(m, false);
    if (!has_file(m.file(Role::extra_A))) return;
    Matrix A = naja::host::load_csv_matrix(m.file(Role::A));
    std::vector<double> b = naja::host::load_csv_vector(m.file(Role::b));
    const Matrix E = naja::host::load_csv_matrix(m.file(Role::extra_A));
    const std::vector<double> f = naja::host::load_csv_vector(m.file(Role::extra_b));
    if ((size_t)E.rows != f.size()) throw std::runtime_error("extra constraint matrices have mismatched dimensions");
    if (E.cols != A.cols) throw std::runtime_error("extra constraint matrix has wrong number of columns");
    A.v.insert(A.v.end(), E.v.begin(), E.v.end());
    A.rows += E.rows;
    b.insert(b.end(), f.begin(), f.end());

    const std::string start = m.file(Role::start);
    if (has_file(start)) {
        try {
            const std::vector<double> x = naja::host::load_csv_vector(start);
            if ((int)x.size() == A.cols && inside(A, b, x, 1e-9)) return;
        } catch (const std::exception&) {
            // an unreadable start is replaced like an infeasible one
        }
    }
    std::vector<double> x((size_t)A.cols);
    double r = 0.0;
    if (naja_cube_center(A.v.data(), b.data(), A.rows, A.cols, nullptr, x.data(), &r, 1) != NAJA_OK)
        throw std::runtime_error(std::string("axis_aligned_cube_center: ") + naja_last_error());
    if (!inside(A, b, x, 1e-9)) throw std::runtime_error("feasible_start_files: cube centre violates A x <= b");
    if (exists_or_dangling(start)) unlink_path(start);
    spit(start, csv_text(x.data(), A.cols, 1, 17), "start");
}

// Runs fn on every job on up to 16 host threads; the first failure is rethrown
// after all threads have finished.
void for_each_parallel(const std::vector<const PrepareJob*>& jobs, const std::function<void(const PrepareJob&)>& fn)
{
    if (jobs.empty()) return;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t n_threads = jobs.size() == 1 ? 1 : std::clamp<size_t>(std::min<size_t>(hw, 16), 2, jobs.size());
    std::atomic<size_t> cursor{0};
    std::exception_ptr failure;
    std::mutex mu;
    auto worker = [&] {
        for (size_t i = cursor++; i < jobs.size(); i = cursor++) {
            try {
                fn(*jobs[i]);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu);
                if (!failure) failure = std::current_exception();
            }
        }
    };
    std::vector<std::thread> threads;
    for (size_t t = 1; t < n_threads; ++t) threads.emplace_back(worker);
    worker();
    for (auto& t : threads) t.join();
    if (failure) std::rethrow_exception(failure);
}

}  // namespace

int cmd_prepare(int argc, char** argv)
{
    const Options opt = parse_options(argc, argv);
    if (!naja::host::is_directory(opt.models_root))
        throw std::runtime_error("models-root is not a directory: " + opt.models_root);
    naja::host::require_nonempty_file(opt.model_list, "model list");
    const std::vector<std::string> rows = text_lines(opt.model_list, true);
    if (rows.empty()) throw std::runtime_error("no models in list: " + opt.model_list);

    std::unique_ptr<BaseModel> base;
    if (!opt.base_dir.empty()) {
        base = std::make_unique<BaseModel>();
        base->paths = ModelPaths::of(opt.base_dir);
        check_contract(base->paths);
    }
    std::optional<Conditioner> cond;
    if (!opt.cond_cmd.empty()) {
        if (!base) die_usage("--conditioner-cmd requires --base-model-dir");
        cond = Conditioner{opt.cond_cmd, opt.cond_args};
    }

    std::vector<PrepareJob> jobs;
    jobs.reserve(rows.size());
    for (const auto& row : rows) {
        PrepareJob job{row, ModelPaths::of(opt.models_root + "/" + row)};
        job.gem_complete = naja::host::is_directory(job.paths.gem_dir()) && job.gem_files_present();
        jobs.push_back(std::move(job));
    }

    std::vector<const PrepareJob*> prepared, to_repair;
    for (PrepareJob& job : jobs) {
        const ModelPaths& m = job.paths;
        // Stage 1: conditioning and extra rows.
        if (cond) {
            const std::string expected = cond->signature(m.name, base->paths.gem("reaction_ids.txt"));
            const bool run = conditioner_needed(job, expected, opt.force);
            if (opt.dry_run) {  // a dry run reports conditioned models without touching them
                prepared.push_back(&job);
                continue;
            }
            if (run) {
                naja::host::ensure_dir(m.gem_dir());
                cond->run(opt.base_dir, m);
            }
            if (!job.gem_files_present())
                throw std::runtime_error("conditioner did not produce required gem files for " + m.dir);
            derive_extra_rows(job, *base);
            spit(m.gem("conditioning.signature"), expected, "signature");
        } else if (!job.gem_complete) {
            throw std::runtime_error("gem/ missing and no conditioner provided for " + m.dir);
        }

        // Stage 2: a complete rounding file set.
        if (!opt.dry_run) {
            naja::host::ensure_dir(m.rounding_dir());
            drop_empty_extra_pair(m);
        }
        if (!core_set_present(m)) {
            if (!base)
                throw std::runtime_error("rounding contract missing/invalid and no --base-model-dir provided: " + m.dir);
            if (!opt.dry_run) {
                materialise_from_base(base->paths, m, opt.mode);
                drop_empty_extra_pair(m);
            }
        }
        check_contract(m);
        if (!opt.dry_run && base && !has_file(m.file(Role::extra_A)) && has_file(m.gem("l_bounds.csv")) &&
            has_file(base->paths.gem("l_bounds.csv")) &&
            slurp(m.gem("l_bounds.csv")) != slurp(base->paths.gem("l_bounds.csv")))
            std::cerr << "*** WARNING: " << m.name << " has different gem bounds from base but NO extra constraints. ***\n"
                      << "  Sampling will use the unconditioned base polytope.\n"
                      << "  To fix: rerun prepare with --conditioner-cmd, or condition + prepare with "
                         "--force-conditioning.\n";
        if (!opt.dry_run) to_repair.push_back(&job);
        prepared.push_back(&job);
    }

    // Stage 3: starts that the extra rows cut off, models in parallel.
    for_each_parallel(to_repair, [](const PrepareJob& job) {
        repair_start(job.paths);
        check_contract(job.paths);
    });

    if (!opt.out_list.empty() && !opt.dry_run) {
        std::string text;
        for (const PrepareJob* j : prepared) text += j->row + "\n";
        spit(opt.out_list, text, "out-model-list");
    }
    std::cout << "OK\nmodels_root     :: " << naja::host::make_absolute_path(opt.models_root)
              << "\nmodel_list      :: " << naja::host::make_absolute_path(opt.model_list)
              << "\nprepared        :: " << prepared.size() << " / " << rows.size() << "\n";
    if (!opt.out_list.empty())
        std::cout << "out_model_list  :: " << naja::host::make_absolute_path(opt.out_list) << "\n";
    if (opt.dry_run) std::cout << "dry_run         :: true\n";
    return 0;
}
