// `naja` host binary: the drop-in command-line surface over libnaja_amd.so.
//
//   naja sample run   (cmd_run.cpp:44-229 + run_sampling_job, job_sampling.cu:82-443)
//   naja validate     (validate_cli.cpp:49-296)
//   naja prepare      (cmd_prepare.cpp:213-457; prepare_cli.cpp)
//   naja sample calibrate-rounding (cmd_calibrate_rounding.cpp:24-195)
//
// Flags, the model-directory contract, the run-directory layout and the
// output files (samples.npy, profile.json, config_generated.txt,
// config_used.txt, run_manifest.json, bounds outputs) follow the reference;
// the sampling, backmap and diagnostics run through the C ABI on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <memory>
#include <vector>

#include "../../../include/naja_amd.h"
#include "host_io.h"

using namespace naja::host;

namespace {

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

void napi(int rc, const char* what)
{
    if (rc == NAJA_OK) return;
    const std::string msg = std::string(what) + ": " + naja_last_error();
    if (rc == NAJA_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    explicit DevBuf(size_t count) : n(count)
    {
        if (count) HIPCHK(hipMalloc((void**)&p, sizeof(T) * count));
    }
    DevBuf(const T* host, size_t count) : DevBuf(count)
    {
        if (count) HIPCHK(hipMemcpy(p, host, sizeof(T) * count, hipMemcpyHostToDevice));
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// ------------------------------------------------------------------ config

struct RunConfig {
    std::string DATA_DIR, MODEL_NAME, OUT_DIR, MODEL_DIR, ROUND_PREFIX;
    int N_CHAINS = -1, N_SAMPLES = -1, TPB_SS = 128, GPU_DEVICE = 0, THINNING = 0;
    bool BACK_TRANSFORM = false, WRITE_DATA = false, VERBOSE = false, STATUS = true;
    bool BOUNDS_FILTER = false, WRITE_SAMPLES_VALID = false;
    int STREAM_CHUNK = 0;  // > 0: stream samples.npy in chunks of this many samples per chain
    double BOUNDS_EPS = 1e-6, PAIR_PROB = 0.0, EXTRA_CONSTRAINT_EPS = 0.0, CONSTRAINT_EPS = 0.0;
    double AFFINE_HULL_TOL = 0.0, KSPARSE_PROB = 0.0;
    int RESYNC_INTERVAL = 0, ITER_ROUNDING_PASSES = 0, ITER_ROUNDING_WARMUP = 0, KSPARSE_K = 8;
    std::string PAIR_SCHEDULE, EXTRA_CONSTRAINTS = "auto", START_POLICY = "file";
    bool HAS_SEED = false;
    uint64_t SEED = 0;
    int CHAINS_PER_GROUP = 1;
    std::string source_file;

    std::string file(const std::string& suffix) const { return ROUND_PREFIX + "_" + suffix + ".csv"; }
    void derive_paths()
    {
        if (!DATA_DIR.empty() && DATA_DIR.back() == '/') DATA_DIR.pop_back();
        if (!OUT_DIR.empty() && OUT_DIR.back() == '/') OUT_DIR.pop_back();
        MODEL_DIR = DATA_DIR + "/" + MODEL_NAME;
        ROUND_PREFIX = MODEL_DIR + "/rounding/" + MODEL_NAME + "_rounding";
    }
};

[[noreturn]] void die_usage(const std::string& msg)
{
    std::cerr << "error: " << msg << "\n\n";
    std::cerr << "usage:\n";
    std::cerr << "  naja sample run --model-dir <dir> --out-root <dir> --gpu <id> --n-chains <n> --n-samples <n> "
                 "[--tpb <n>] [--thinning <n>] [--start-policy file|cube_center] [--extra-constraints "
                 "auto|ignore|require] [--extra-constraint-eps <eps>] [--constraint-eps <eps>] [--pair-prob <p>] "
                 "[--pair-schedule <csv>] [--ksparse-prob <p>] [--ksparse-k <k>] [--resync-interval <n>] "
                 "[--iter-rounding-passes <n>] [--iter-rounding-warmup <n>] [--affine-hull-tol <tol>] [--backmap] "
                 "[--write-npy] [--stream-chunk <n>] [--bounds-policy ignore|filter] [--bounds-eps <eps>] [--write-samples-valid] "
                 "[--seed <s>] [--chains-per-group 1|4] [--verbose] [--quiet] [--dry-run] [--print-config]\n";
    std::cerr << "  naja validate --samples <file.npy> [--model-dir <dir>] [--n-chains 4] "
                 "[--chain-layout interleaved|naja-blocks]\n";
    std::exit(2);
}

std::string next_arg(int& i, int argc, char** argv, const std::string& flag)
{
    if (i + 1 >= argc) die_usage("missing value for " + flag);
    return std::string(argv[++i]);
}

void write_generated_config(const std::string& path, const RunConfig& c)
{
    std::ofstream f(path);
    if (!f.is_open()) throw std::runtime_error("cannot write generated config: " + path);
    f << "DATA_DIR=" << c.DATA_DIR << "\nMODEL_NAME=" << c.MODEL_NAME << "\nOUT_DIR=" << c.OUT_DIR << "\n";
    f << "N_CHAINS=" << c.N_CHAINS << "\nN_SAMPLES=" << c.N_SAMPLES << "\nTPB_SS=" << c.TPB_SS << "\n";
    f << "GPU_DEVICE=" << c.GPU_DEVICE << "\nBACK_TRANSFORM=" << (c.BACK_TRANSFORM ? "true" : "false") << "\n";
    f << "WRITE_DATA=" << (c.WRITE_DATA ? "true" : "false") << "\nVERBOSE=" << (c.VERBOSE ? "true" : "false") << "\n";
    f << "STATUS=" << (c.STATUS ? "true" : "false") << "\n";
    if (c.THINNING > 0) f << "THINNING=" << c.THINNING << "\n";
    if (c.PAIR_PROB > 0.0) f << "PAIR_PROB=" << std::setprecision(12) << c.PAIR_PROB << "\n";
    if (c.RESYNC_INTERVAL > 0) f << "RESYNC_INTERVAL=" << c.RESYNC_INTERVAL << "\n";
    if (c.ITER_ROUNDING_PASSES > 0) f << "ITER_ROUNDING_PASSES=" << c.ITER_ROUNDING_PASSES << "\n";
    if (c.ITER_ROUNDING_WARMUP > 0) f << "ITER_ROUNDING_WARMUP=" << c.ITER_ROUNDING_WARMUP << "\n";
    if (c.EXTRA_CONSTRAINT_EPS > 0.0) f << "EXTRA_CONSTRAINT_EPS=" << std::setprecision(12) << c.EXTRA_CONSTRAINT_EPS << "\n";
    if (c.CONSTRAINT_EPS > 0.0) f << "CONSTRAINT_EPS=" << std::setprecision(12) << c.CONSTRAINT_EPS << "\n";
    if (!c.PAIR_SCHEDULE.empty()) f << "PAIR_SCHEDULE=" << c.PAIR_SCHEDULE << "\n";
    if (!c.EXTRA_CONSTRAINTS.empty()) f << "EXTRA_CONSTRAINTS=" << c.EXTRA_CONSTRAINTS << "\n";
    if (!c.START_POLICY.empty()) f << "START_POLICY=" << c.START_POLICY << "\n";
    if (c.AFFINE_HULL_TOL > 0.0) f << "AFFINE_HULL_TOL=" << std::setprecision(12) << c.AFFINE_HULL_TOL << "\n";
    if (c.KSPARSE_PROB > 0.0) f << "KSPARSE_PROB=" << std::setprecision(12) << c.KSPARSE_PROB << "\n";
    if (c.KSPARSE_K != 8) f << "KSPARSE_K=" << c.KSPARSE_K << "\n";
    if (c.BOUNDS_FILTER) {
        f << "BOUNDS_POLICY=filter\nBOUNDS_EPS=" << std::setprecision(12) << c.BOUNDS_EPS << "\n";
        f << "WRITE_SAMPLES_VALID=" << (c.WRITE_SAMPLES_VALID ? "true" : "false") << "\n";
    } else {
        f << "BOUNDS_POLICY=ignore\n";
    }
    if (c.HAS_SEED) f << "SEED=" << c.SEED << "\n";
    if (c.CHAINS_PER_GROUP != 1) f << "CHAINS_PER_GROUP=" << c.CHAINS_PER_GROUP << "\n";
}

std::string allocate_run_dir(const std::string& out_root, const std::string& name)
{
    ensure_dir(out_root);
    const std::string date = current_datestamp_compact();
    for (int idx = 1; idx < 1000000; ++idx) {
        std::ostringstream o;
        o << out_root << "/" << name << "_" << date << "_" << std::setw(3) << std::setfill('0') << idx;
        if (!path_exists(o.str())) {
            ensure_dir(o.str());
            return o.str();
        }
    }
    throw std::runtime_error("could not allocate run directory under " + out_root);
}

std::string json_escape(const std::string& s)
{
    std::string o;
    for (char ch : s) {
        if (ch == '"' || ch == '\\') o.push_back('\\');
        o.push_back(ch);
    }
    return o;
}

void write_run_manifest(const RunConfig& c, const std::vector<std::string>& argv_full)
{
    std::ofstream f(c.OUT_DIR + "/run_manifest.json");
    if (!f.is_open()) throw std::runtime_error("cannot write run manifest");
    f << "{\n  \"tool\": \"naja_amd\",\n  \"version\": " << naja_version() << ",\n  \"created\": \""
      << current_timestamp() << "\",\n  \"argv\": [";
    for (size_t i = 0; i < argv_full.size(); ++i) f << (i ? ", " : "") << "\"" << json_escape(argv_full[i]) << "\"";
    f << "],\n  \"model\": \"" << json_escape(c.MODEL_NAME) << "\",\n  \"model_dir\": \""
      << json_escape(make_absolute_path(c.MODEL_DIR)) << "\",\n  \"run_dir\": \""
      << json_escape(make_absolute_path(c.OUT_DIR)) << "\",\n  \"config_generated\": \""
      << json_escape(make_absolute_path(c.source_file)) << "\"\n}\n";
}

struct Profile {
    double load = 0, upload = 0, sampling = 0, download = 0, backtransform = 0, write = 0, total = 0;
    int n_chains = 0, n_samples = 0, reduced_dim = 0, original_dim = 0, constraints = 0, thinning = 0;
    double throughput = 0;
};

// profile.h:28-66
void write_profile_json(const std::string& path, const Profile& p, const RunConfig& c)
{
    std::ofstream f(path);
    if (!f.is_open()) throw std::runtime_error("cannot write profile: " + path);
    f << "{\n  \"timestamp\": \"" << current_timestamp() << "\",\n  \"config\": {\n";
    f << "    \"n_chains\": " << p.n_chains << ",\n    \"n_samples\": " << p.n_samples << ",\n";
    f << "    \"reduced_dim\": " << p.reduced_dim << ",\n    \"original_dim\": " << p.original_dim << ",\n";
    f << "    \"constraints\": " << p.constraints << ",\n    \"thinning\": " << p.thinning << ",\n";
    f << "    \"tpb_ss\": " << c.TPB_SS << ",\n    \"gpu_device\": " << c.GPU_DEVICE << ",\n";
    f << "    \"back_transform\": " << (c.BACK_TRANSFORM ? "true" : "false") << "\n  },\n  \"timing\": {\n";
    f << std::fixed << std::setprecision(4);
    f << "    \"load_polytope_s\": " << p.load << ",\n    \"upload_to_gpu_s\": " << p.upload << ",\n";
    f << "    \"gpu_sampling_s\": " << p.sampling << ",\n    \"download_from_gpu_s\": " << p.download << ",\n";
    f << "    \"backtransform_s\": " << p.backtransform << ",\n    \"write_output_s\": " << p.write << ",\n";
    f << "    \"total_s\": " << p.total << "\n  },\n  \"performance\": {\n";
    f << "    \"total_samples\": " << (p.n_chains * p.n_samples) << ",\n";
    f << "    \"throughput_samples_per_sec\": " << std::setprecision(1) << p.throughput << ",\n";
    f << "    \"sampling_fraction\": " << std::setprecision(3) << (p.total > 0 ? p.sampling / p.total : 0.0)
      << "\n  }\n}\n";
}

void write_config_snapshot(const RunConfig& c, const Profile& p)
{
    std::ofstream dst(c.OUT_DIR + "/config_used.txt");
    if (!dst.is_open()) throw std::runtime_error("cannot write config snapshot");
    dst << "# Source config: " << c.source_file << "\n# Captured at: " << current_timestamp() << "\n";
    dst << "# Output directory: " << make_absolute_path(c.OUT_DIR) << "\n\n";
    std::ifstream src(c.source_file);
    if (!src.is_open()) throw std::runtime_error("cannot read source config: " + c.source_file);
    dst << src.rdbuf();
    dst << "\n# Derived metadata\nMODEL_NAME=" << c.MODEL_NAME << "\n";
    dst << "DATA_DIR_ABS=" << make_absolute_path(c.DATA_DIR) << "\nMODEL_DIR_ABS=" << make_absolute_path(c.MODEL_DIR)
        << "\n";
    dst << "ROUND_PREFIX_ABS=" << make_absolute_path(c.ROUND_PREFIX) << "\n";
    for (const char* s : {"A", "b", "start", "T", "shift"}) dst << s << "_FILE=" << make_absolute_path(c.file(s)) << "\n";
    dst << "NPY_FILE=" << make_absolute_path(c.OUT_DIR + "/samples.npy") << "\n";
    dst << "PROFILE_FILE=" << make_absolute_path(c.OUT_DIR + "/profile.json") << "\n";
    dst << "N_CHAINS=" << p.n_chains << "\nN_SAMPLES=" << p.n_samples << "\nTHINNING=" << p.thinning << "\n";
    dst << "GPU_DEVICE=" << c.GPU_DEVICE << "\nBACK_TRANSFORM=" << (c.BACK_TRANSFORM ? "true" : "false") << "\n";
    dst << "WRITE_DATA=" << (c.WRITE_DATA ? "true" : "false") << "\n";
    dst << "BOUNDS_POLICY=" << (c.BOUNDS_FILTER ? "filter" : "ignore") << "\n";
    if (c.BOUNDS_FILTER) {
        dst << "BOUNDS_EPS=" << std::setprecision(12) << c.BOUNDS_EPS << "\n";
        dst << "WRITE_SAMPLES_VALID=" << (c.WRITE_SAMPLES_VALID ? "true" : "false") << "\n";
    }
}

// ------------------------------------------------------------- polytope prep

struct Polytope {
    Matrix A;               // m x d row-major FP64
    std::vector<double> b;  // m
    std::vector<double> x0; // d
};

// extra_constraints.cpp:8-58
bool augment_extra(Polytope& P, const std::string& mode_in, const RunConfig& c)
{
    std::string mode = mode_in;
    std::transform(mode.begin(), mode.end(), mode.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
    int m = 0;  // 0 auto, 1 ignore, 2 require
    if (mode.empty() || mode == "auto") m = 0;
    else if (mode == "ignore" || mode == "off" || mode == "false" || mode == "0") m = 1;
    else if (mode == "require" || mode == "on" || mode == "true" || mode == "1") m = 2;
    else throw std::invalid_argument("invalid extra constraints mode: " + mode_in);
    const std::string ea = c.file("extra_A"), eb = c.file("extra_b");
    const bool ha = path_exists(ea), hb = path_exists(eb);
    if ((ha || hb) && !(ha && hb))
        throw std::runtime_error("incomplete extra constraint files (need both A_extra and b_extra)");
    if (m == 1) return false;
    if (!ha) {
        if (m == 2) throw std::runtime_error("extra constraints required but not provided");
        return false;
    }
    Matrix E = load_csv_matrix(ea);
    std::vector<double> f = load_csv_vector(eb);
    if ((size_t)E.rows != f.size()) throw std::runtime_error("extra constraint matrices have mismatched dimensions");
    if (E.cols != P.A.cols) throw std::runtime_error("extra constraint matrix has wrong number of columns");
    if (c.EXTRA_CONSTRAINT_EPS > 0.0)
        for (double& v : f) v += c.EXTRA_CONSTRAINT_EPS;
    P.A.v.insert(P.A.v.end(), E.v.begin(), E.v.end());
    P.A.rows += E.rows;
    P.b.insert(P.b.end(), f.begin(), f.end());
    return true;
}

std::vector<double> residual(const Polytope& P)  // A x0 - b
{
    std::vector<double> r((size_t)P.A.rows);
    for (int i = 0; i < P.A.rows; ++i) {
        double acc = 0.0;
        for (int j = 0; j < P.A.cols; ++j) acc += P.A.at(i, j) * P.x0[(size_t)j];
        r[(size_t)i] = acc - P.b[(size_t)i];
    }
    return r;
}

// feasibility.h:12-35
void require_feasible_start(const Polytope& P, double eps, const std::string& tag)
{
    const std::vector<double> r = residual(P);
    double mx = -INFINITY;
    int viol = 0;
    for (double v : r) {
        mx = std::max(mx, v);
        if (v > eps) ++viol;
    }
    if (!(mx <= eps)) {
        std::ostringstream o;
        o << "infeasible start point (" << tag << "): max(A*x0 - b)=" << std::setprecision(12) << mx
          << " > eps=" << std::setprecision(12) << eps << ", violated=" << viol << "/" << r.size();
        throw std::runtime_error(o.str());
    }
}

struct PairSchedule {
    std::vector<int> i, j;
    std::vector<float> c, s;
};

// schedule_io.cpp:12-75
PairSchedule load_pair_schedule(const std::string& path)
{
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("cannot open pair schedule: " + path);
    PairSchedule S;
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        while (!line.empty() && (line.front() == ' ' || line.front() == '\t')) line.erase(line.begin());
        if (line.empty() || line[0] == '#') continue;
        std::stringstream ss(line);
        std::string cell;
        std::vector<std::string> cells;
        while (std::getline(ss, cell, ',')) cells.push_back(cell);
        if (cells.size() != 4)
            throw std::runtime_error("pair schedule parse error at line " + std::to_string(lineno) +
                                     ": expected 4 columns");
        const int ii = std::stoi(cells[0]), jj = std::stoi(cells[1]);
        const double c = std::stod(cells[2]), s = std::stod(cells[3]);
        if (ii < 0 || jj < 0)
            throw std::runtime_error("pair schedule parse error at line " + std::to_string(lineno) + ": negative index");
        if (!std::isfinite(c) || !std::isfinite(s))
            throw std::runtime_error("pair schedule parse error at line " + std::to_string(lineno) + ": non-finite c/s");
        if (!(c * c + s * s > 0.0))
            throw std::runtime_error("pair schedule parse error at line " + std::to_string(lineno) + ": zero norm");
        S.i.push_back(ii);
        S.j.push_back(jj);
        S.c.push_back((float)c);
        S.s.push_back((float)s);
    }
    if (S.i.empty()) throw std::runtime_error("pair schedule is empty: " + path);
    return S;
}

// jacobi.cpp:8-21
std::pair<double, double> jacobi_rotation_cs(double var_i, double var_j, double cov_ij)
{
    if (!(var_i >= 0.0) || !(var_j >= 0.0))
        throw std::invalid_argument("jacobi_rotation_cs: variances must be nonnegative");
    if (!std::isfinite(var_i) || !std::isfinite(var_j) || !std::isfinite(cov_ij))
        throw std::invalid_argument("jacobi_rotation_cs: inputs must be finite");
    if (cov_ij == 0.0) return {1.0, 0.0};
    const double theta = 0.5 * std::atan2(2.0 * cov_ij, var_i - var_j);
    return {std::cos(theta), std::sin(theta)};
}

// Column-major FP32 copy of A for the device.
std::vector<float> colmajor_f32(const Matrix& A)
{
    std::vector<float> out((size_t)A.rows * A.cols);
    for (int j = 0; j < A.cols; ++j)
        for (int i = 0; i < A.rows; ++i) out[(size_t)j * A.rows + i] = (float)A.at(i, j);
    return out;
}

naja_chr_params base_params(const RunConfig& c, int nspc, int thin, int nchains, uint64_t seed)
{
    naja_chr_params p{};
    p.nspc = nspc;
    p.thinning = thin;
    p.nchains = nchains;
    p.tpb = c.TPB_SS;
    p.seed = seed;
    p.pair_prob = c.PAIR_PROB;
    p.resync_interval = c.RESYNC_INTERVAL;
    p.ksparse_prob = c.KSPARSE_PROB;
    p.ksparse_k = c.KSPARSE_K;
    p.pair_mode = 1;
    p.chains_per_group = c.CHAINS_PER_GROUP;
    return p;
}

// runtime_warmup.cpp:16-108: seeded shuffle into d/2 disjoint pairs; each pass
// runs one warmup chain (thin 1), measures per-pair population var/cov and
// sets the pair's Jacobi rotation.
PairSchedule runtime_pair_schedule(const RunConfig& c, const float* A_d, const double* b_d, int m, int d,
                                   const std::vector<double>& x0, uint64_t seed, int passes, int warmup)
{
    std::vector<int> perm((size_t)d);
    for (int i = 0; i < d; ++i) perm[(size_t)i] = i;
    std::mt19937_64 rng(seed ^ 0x9e3779b97f4a7c15ULL);
    std::shuffle(perm.begin(), perm.end(), rng);
    const int np = d / 2;
    PairSchedule S;
    for (int k = 0; k < np; ++k) {
        S.i.push_back(perm[(size_t)(2 * k)]);
        S.j.push_back(perm[(size_t)(2 * k + 1)]);
        S.c.push_back(1.0f);
        S.s.push_back(0.0f);
    }
    std::vector<float> x0f(x0.begin(), x0.end());
    DevBuf<float> X(x0f.data(), x0f.size());
    DevBuf<float> W((size_t)d * warmup);
    std::vector<float> Wh((size_t)d * warmup);
    for (int pass = 0; pass < passes; ++pass) {
        const int mode = pass == 0 ? 1 : 2;
        DevBuf<int> pi(S.i.data(), S.i.size()), pj(S.j.data(), S.j.size());
        DevBuf<float> pc(S.c.data(), S.c.size()), ps(S.s.data(), S.s.size());
        naja_chr_params p = base_params(c, warmup, 1, 1, seed ^ (uint64_t)(pass + 1));
        p.pair_prob = std::min(std::max(c.PAIR_PROB, 0.05), 0.5);
        p.ksparse_prob = 0.0;
        p.pair_mode = mode;
        p.chains_per_group = 1;
        if (mode == 2) {
            p.pair_i = pi.p;
            p.pair_j = pj.p;
            p.pair_c = pc.p;
            p.pair_s = ps.p;
            p.n_pairs = np;
        }
        napi(naja_chr_sample(A_d, b_d, m, d, X.p, &p, W.p, nullptr, nullptr), "warmup CHR");
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(Wh.data(), W.p, sizeof(float) * Wh.size(), hipMemcpyDeviceToHost));
        for (int k = 0; k < np; ++k) {
            const int ii = S.i[(size_t)k], jj = S.j[(size_t)k];
            double mi = 0, mj = 0;
            for (int t = 0; t < warmup; ++t) {
                mi += Wh[(size_t)t * d + ii];
                mj += Wh[(size_t)t * d + jj];
            }
            mi /= warmup;
            mj /= warmup;
            double vi = 0, vj = 0, cij = 0;
            for (int t = 0; t < warmup; ++t) {
                const double di = Wh[(size_t)t * d + ii] - mi, dj = Wh[(size_t)t * d + jj] - mj;
                vi += di * di;
                vj += dj * dj;
                cij += di * dj;
            }
            const auto cs = jacobi_rotation_cs(vi / warmup, vj / warmup, cij / warmup);
            S.c[(size_t)k] = (float)cs.first;
            S.s[(size_t)k] = (float)cs.second;
        }
    }
    return S;
}

std::vector<std::string> read_nonempty_trimmed_lines(const std::string& path)
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
    if (out.empty()) throw std::runtime_error("empty file: " + path);
    return out;
}

// bounds_filter.cpp:67-179 on backmapped rows (ns x dim, row-major).
void bounds_filter_and_write(const RunConfig& c, const std::vector<float>& V, int64_t ns, int dim)
{
    if (!c.BACK_TRANSFORM)
        throw std::runtime_error(
            "bounds filtering requires BACK_TRANSFORM=true (samples must be in GEM/original space)");
    const std::string gem = c.MODEL_DIR + "/gem";
    require_nonempty_file(gem + "/reaction_ids.txt", "gem/reaction_ids.txt");
    require_nonempty_file(gem + "/l_bounds.csv", "gem/l_bounds.csv");
    require_nonempty_file(gem + "/u_bounds.csv", "gem/u_bounds.csv");
    const std::vector<std::string> ids = read_nonempty_trimmed_lines(gem + "/reaction_ids.txt");
    const std::vector<double> lb = load_csv_vector(gem + "/l_bounds.csv");
    const std::vector<double> ub = load_csv_vector(gem + "/u_bounds.csv");
    if ((int)ids.size() != dim)
        throw std::runtime_error("reaction_ids length mismatch: expected " + std::to_string(dim) + " got " +
                                 std::to_string(ids.size()));
    if ((int)lb.size() != dim || (int)ub.size() != dim)
        throw std::runtime_error("bounds length mismatch: expected " + std::to_string(dim) + " got lb=" +
                                 std::to_string(lb.size()) + " ub=" + std::to_string(ub.size()));
    std::vector<uint8_t> mask((size_t)ns, 1);
    std::vector<double> per((size_t)ns, 0.0), worst_rxn((size_t)dim, 0.0);
    double max_v = 0.0;
    const double eps = c.BOUNDS_EPS;
    for (int64_t j = 0; j < ns; ++j) {
        bool ok = true;
        double worst = 0.0;
        for (int i = 0; i < dim; ++i) {
            const double x = V[(size_t)j * dim + i];
            const double vlo = (lb[(size_t)i] - x) - eps, vhi = (x - ub[(size_t)i]) - eps;
            const double v = vlo > vhi ? vlo : vhi;
            if (v > 0.0) {
                ok = false;
                max_v = std::max(max_v, v);
                worst_rxn[(size_t)i] = std::max(worst_rxn[(size_t)i], v);
                worst = std::max(worst, v);
            }
        }
        mask[(size_t)j] = ok ? 1 : 0;
        per[(size_t)j] = worst;
    }
    int64_t n_valid = 0;
    for (uint8_t bv : mask) n_valid += bv != 0;
    save_npy_u8(c.OUT_DIR + "/valid_mask.npy", mask);
    {
        std::ofstream vf(c.OUT_DIR + "/valid_fraction.txt");
        if (!vf.is_open()) throw std::runtime_error("cannot write valid_fraction.txt");
        vf << std::setprecision(12) << (ns ? (double)n_valid / (double)ns : 0.0) << "\n";
    }
    if (c.WRITE_SAMPLES_VALID) {
        // npy.h:15-53 writes (dim, n_valid) float64, C order
        std::vector<double> sv((size_t)dim * n_valid);
        int64_t k = 0;
        for (int64_t j = 0; j < ns; ++j) {
            if (!mask[(size_t)j]) continue;
            for (int i = 0; i < dim; ++i) sv[(size_t)i * n_valid + k] = V[(size_t)j * dim + i];
            ++k;
        }
        save_npy_f64(c.OUT_DIR + "/samples_valid.npy", sv.data(), dim, n_valid);
    }
    double mean_v = 0.0;
    for (double v : per) mean_v += v;
    mean_v = ns ? mean_v / (double)ns : 0.0;
    double p99 = 0.0;
    if (ns > 0) {
        const size_t k = (size_t)((99 * (ns - 1)) / 100);
        std::nth_element(per.begin(), per.begin() + (ptrdiff_t)k, per.end());
        p99 = per[k];
    }
    const int kmax = std::min(dim, 10);
    std::vector<std::pair<int, double>> top;
    for (int i = 0; i < dim; ++i) top.emplace_back(i, worst_rxn[(size_t)i]);
    std::partial_sort(top.begin(), top.begin() + kmax, top.end(),
                      [](const auto& a, const auto& b) { return a.second > b.second; });
    top.resize((size_t)kmax);
    while (!top.empty() && top.back().second <= 0.0) top.pop_back();
    std::ofstream f(c.OUT_DIR + "/bounds_report.json");
    if (!f.is_open()) throw std::runtime_error("Cannot write bounds report");
    f << "{\n  \"model\": \"" << c.MODEL_NAME << "\",\n  \"model_dir\": \"" << make_absolute_path(c.MODEL_DIR)
      << "\",\n  \"run_dir\": \"" << make_absolute_path(c.OUT_DIR) << "\",\n";
    f << "  \"eps\": " << std::setprecision(12) << eps << ",\n  \"n_total\": " << ns << ",\n  \"n_valid\": " << n_valid
      << ",\n  \"n_invalid\": " << (ns - n_valid) << ",\n";
    f << "  \"invalid_fraction\": " << std::setprecision(12) << (ns ? (double)(ns - n_valid) / (double)ns : 0.0)
      << ",\n  \"mean_violation\": " << mean_v << ",\n  \"p99_violation\": " << p99 << ",\n  \"max_violation\": "
      << max_v << ",\n  \"outputs\": {\n    \"valid_mask\": \"" << make_absolute_path(c.OUT_DIR + "/valid_mask.npy")
      << "\",\n    \"samples_valid\": "
      << (c.WRITE_SAMPLES_VALID ? ("\"" + make_absolute_path(c.OUT_DIR + "/samples_valid.npy") + "\"") : "null")
      << "\n  },\n  \"top_violations\": [\n";
    for (size_t i = 0; i < top.size(); ++i) {
        const int idx = top[i].first;
        f << "    {\"index\": " << idx << ", \"reaction\": \"" << ids[(size_t)idx] << "\", \"worst_violation\": "
          << std::setprecision(12) << top[i].second << "}" << (i + 1 != top.size() ? "," : "") << "\n";
    }
    f << "  ]\n}\n";
}

// Affine hull of the start point: n_file x k basis (row-major) and x0.
struct Hull {
    bool enabled = false;
    std::vector<double> basis, shift;
};

// The map applied to the sampled variables before output: the model
// transform (T, shift) when --backmap is set, composed with the hull as
// (T B, T x0 + shift) (job_sampling.cu:288-302); (B, x0) for a hull without
// --backmap (job_sampling.cu:343-348); nothing otherwise.
struct OutputTransform {
    bool on = false;
    int rows = 0;
    std::vector<float> T, shift;
};

OutputTransform output_transform(const RunConfig& c, int d, int d_file, const Hull& hull)
{
    OutputTransform o;
    if (c.BACK_TRANSFORM) {
        status_phase(c.STATUS, "loading transform");
        const Matrix T = load_csv_matrix(c.file("T"));
        const std::vector<double> shift = load_csv_vector(c.file("shift"));
        const int expected = hull.enabled ? d_file : d;
        if (T.cols != expected)
            throw std::runtime_error("dimension mismatch: transform matrix cols (" + std::to_string(T.cols) +
                                     ") != reduced dim (" + std::to_string(expected) + ")");
        if ((int)shift.size() != T.rows)
            throw std::invalid_argument("backmap_into: shift length mismatch with transformation rows");
        o.on = true;
        o.rows = T.rows;
        o.T.resize((size_t)T.rows * d);
        o.shift.resize((size_t)T.rows);
        for (int i = 0; i < T.rows; ++i) {
            if (hull.enabled) {
                double sh = shift[(size_t)i];
                for (int j = 0; j < d_file; ++j) sh += T.at(i, j) * hull.shift[(size_t)j];
                o.shift[(size_t)i] = (float)sh;
                for (int l = 0; l < d; ++l) {
                    double acc = 0.0;
                    for (int j = 0; j < d_file; ++j) acc += T.at(i, j) * hull.basis[(size_t)j * d + l];
                    o.T[(size_t)i * d + l] = (float)acc;
                }
            } else {
                o.shift[(size_t)i] = (float)shift[(size_t)i];
                for (int l = 0; l < d; ++l) o.T[(size_t)i * d + l] = (float)T.at(i, l);
            }
        }
    } else if (hull.enabled) {
        o.on = true;
        o.rows = d_file;
        o.T.assign(hull.basis.begin(), hull.basis.end());
        o.shift.assign(hull.shift.begin(), hull.shift.end());
    }
    return o;
}

// ------------------------------------------------------------ sampling job

int run_sampling_job(RunConfig c)
{
    ensure_dir(c.OUT_DIR);
    const double t_total = now_seconds();
    Profile prof;
    HIPCHK(hipSetDevice(c.GPU_DEVICE));
    if (c.VERBOSE) {
        std::cout << "config\n  model           " << c.MODEL_NAME << "\n  gpu_device      " << c.GPU_DEVICE
                  << "\n  n_chains        " << c.N_CHAINS << "\n  n_samples       " << c.N_SAMPLES
                  << "\n  tpb_ss          " << c.TPB_SS << "\n";
        if (c.PAIR_PROB > 0.0) std::cout << "  pair_prob       " << c.PAIR_PROB << "\n";
        if (c.RESYNC_INTERVAL > 0) std::cout << "  resync_interval " << c.RESYNC_INTERVAL << "\n";
        std::cout << "  back_transform  " << (c.BACK_TRANSFORM ? "yes" : "no") << "\n\n";
    }

    status_phase(c.STATUS, "loading polytope");
    double t0 = now_seconds();
    Polytope P;
    P.A = load_csv_matrix(c.file("A"));
    P.b = load_csv_vector(c.file("b"));
    P.x0 = load_csv_vector(c.file("start"));
    const bool extra_loaded = augment_extra(P, c.EXTRA_CONSTRAINTS, c);
    if (c.CONSTRAINT_EPS > 0.0)
        for (double& v : P.b) v += c.CONSTRAINT_EPS;
    prof.load = now_seconds() - t0;
    int m = P.A.rows, d = P.A.cols;
    const int d_file = d;
    if ((int)P.x0.size() != d)
        throw std::runtime_error("dimension mismatch: start point size (" + std::to_string(P.x0.size()) +
                                 ") != cols (" + std::to_string(d) + ")");
    if ((int)P.b.size() != m)
        throw std::runtime_error("dimension mismatch: rhs vector size (" + std::to_string(P.b.size()) +
                                 ") != rows (" + std::to_string(m) + ")");
    if (c.START_POLICY == "cube_center") {
        // job_sampling.cu:162-167: the centre of the largest axis-aligned cube
        // (own interior-point LP in place of Gurobi).
        status_phase(c.STATUS, "start policy: cube_center");
        std::vector<double> xc((size_t)d);
        double r = 0.0;
        napi(naja_cube_center(P.A.v.data(), P.b.data(), m, d, nullptr, xc.data(), &r, 0), "axis_aligned_cube_center");
        P.x0 = xc;
        if (c.VERBOSE) std::cout << "  start_cube_r     " << std::setprecision(6) << r << "\n" << std::defaultfloat;
    }
    require_feasible_start(P, 1e-9, extra_loaded ? "A+extra" : "A");
    {
        const std::vector<double> r = residual(P);
        int n_tight = 0;
        for (double v : r)
            if (-v < 1e-6) ++n_tight;
        if (n_tight >= d)
            std::cerr << "\n*** WARNING: polytope appears degenerate for model " << c.MODEL_NAME << " ***\n"
                      << "  tight constraints: " << n_tight << " / " << m << "\n  reduced dimension: " << d
                      << "\n  The feasible set may be a single point. Samples will likely be identical.\n"
                      << "  Check rounding quality (tight constraint rank vs dimension).\n\n";
    }
    // Affine-hull reduction (job_sampling.cu:196-235): rows tight at x0 become
    // equalities and sampling runs in their kernel, x = x0 + B z.
    Hull hull;
    if (c.AFFINE_HULL_TOL > 0.0 && d > 0) {
        std::vector<double> B((size_t)d * d);
        int k = 0, n_tight = 0;
        napi(naja_affine_hull(P.A.v.data(), P.b.data(), P.x0.data(), m, d, c.AFFINE_HULL_TOL, 1e-10, B.data(), &k,
                              &n_tight),
             "affine hull");
        if (n_tight > 0 && k > 0 && k < d) {
            hull.enabled = true;
            hull.basis.assign(B.begin(), B.begin() + (size_t)d * k);
            hull.shift = P.x0;
            const std::vector<double> r = residual(P);
            Matrix Ar;
            Ar.rows = m;
            Ar.cols = k;
            Ar.v.assign((size_t)m * k, 0.0);
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < d; ++j) {
                    const double a = P.A.at(i, j);
                    for (int l = 0; l < k; ++l) Ar.at(i, l) += a * hull.basis[(size_t)j * k + l];
                }
            for (int i = 0; i < m; ++i) P.b[(size_t)i] = -r[(size_t)i];
            P.A = std::move(Ar);
            P.x0.assign((size_t)k, 0.0);
            d = k;
            if (c.VERBOSE)
                std::cout << "  affine_hull_tol  " << std::setprecision(12) << c.AFFINE_HULL_TOL
                          << "\n  affine_hull_rows " << n_tight << "\n  affine_hull_dim  " << k << " (from " << d_file
                          << ")\n" << std::defaultfloat;
            require_feasible_start(P, 1e-9, "affine-hull reduced");
        } else if (c.VERBOSE) {
            std::cout << "  affine_hull_tol  " << std::setprecision(12) << c.AFFINE_HULL_TOL
                      << (n_tight == 0 ? " (no tight rows)" : " (no reduction; kernel dim=" + std::to_string(k) + ")")
                      << "\n" << std::defaultfloat;
        }
    }

    const int thin = c.THINNING > 0 ? c.THINNING : std::max(d / 6, 1);
    if (c.VERBOSE)
        std::cout << "  constraints     " << m << "\n  reduced_dim     " << d << "\n  thinning        " << thin
                  << "\n  time            " << std::fixed << std::setprecision(3) << prof.load << "s\n\n"
                  << std::defaultfloat;
    prof.n_chains = c.N_CHAINS;
    prof.n_samples = c.N_SAMPLES;
    prof.reduced_dim = d;
    prof.constraints = m;
    prof.thinning = thin;

    // job_sampling.cu:252-254 default seed; --seed overrides it.
    uint64_t seed;
    if (c.HAS_SEED) {
        seed = c.SEED;
    } else {
        std::hash<std::string> h;
        int s = (int)(h(c.MODEL_NAME) ^ (h(c.MODEL_DIR) << 1) ^ (uint64_t)(c.GPU_DEVICE * 0x9e3779b9));
        if (s == 0) s = 1;
        seed = (uint64_t)(int64_t)s;
    }

    status_phase(c.STATUS, "uploading to gpu");
    t0 = now_seconds();
    const std::vector<float> Acm = colmajor_f32(P.A);
    DevBuf<float> A_d(Acm.data(), Acm.size());
    DevBuf<double> b_d(P.b.data(), P.b.size());
    std::vector<float> X0((size_t)d * c.N_CHAINS);
    for (int ch = 0; ch < c.N_CHAINS; ++ch)
        for (int j = 0; j < d; ++j) X0[(size_t)ch * d + j] = (float)P.x0[(size_t)j];
    DevBuf<float> X_d(X0.data(), X0.size());
    prof.upload = now_seconds() - t0;

    // rounding plan (plan.cpp:12-62): schedule file -> mode 2; warmup -> mode 2
    PairSchedule sched;
    int pair_mode = 1;
    if (!c.PAIR_SCHEDULE.empty()) {
        status_phase(c.STATUS, "loading pair schedule");
        sched = load_pair_schedule(c.PAIR_SCHEDULE);
        pair_mode = 2;
    } else if (c.ITER_ROUNDING_PASSES > 0 && c.ITER_ROUNDING_WARMUP > 0 && d >= 2) {
        status_phase(c.STATUS, "iterative rounding-lite (warmup)");
        sched = runtime_pair_schedule(c, A_d.p, b_d.p, m, d, P.x0, seed, c.ITER_ROUNDING_PASSES,
                                      c.ITER_ROUNDING_WARMUP);
        pair_mode = 2;
        if (c.VERBOSE)
            std::cout << "  iter_rounding_passes " << c.ITER_ROUNDING_PASSES << "\n  iter_rounding_warmup "
                      << c.ITER_ROUNDING_WARMUP << "\n  iter_rounding_pairs  " << (d / 2) << "\n";
    }
    DevBuf<int> pi(sched.i.data(), sched.i.size()), pj(sched.j.data(), sched.j.size());
    DevBuf<float> pc(sched.c.data(), sched.c.size()), ps(sched.s.data(), sched.s.size());

    if (c.STATUS)
        std::cout << "> sampling " << c.N_CHAINS << " chains x " << c.N_SAMPLES << " samples (thin=" << thin << ", "
                  << (long long)c.N_CHAINS * c.N_SAMPLES * thin << " MCMC steps, " << m << " constraints x " << d
                  << " dims)" << std::endl;

    naja_chr_params p = base_params(c, c.N_SAMPLES, thin, c.N_CHAINS, seed);
    p.pair_mode = pair_mode;
    if (pair_mode == 2) {
        p.pair_i = pi.p;
        p.pair_j = pj.p;
        p.pair_c = pc.p;
        p.pair_s = ps.p;
        p.n_pairs = (int)sched.i.size();
    }
    const int64_t ns = (int64_t)c.N_CHAINS * c.N_SAMPLES;
    if (c.STREAM_CHUNK > 0) {
        // Chunked sampling straight into samples.npy (chr.cu:101-230 streaming;
        // the file is written as chunks arrive, never held whole in memory).
        std::unique_ptr<DevBuf<float>> T_d, s_d;
        int out_cols = d;
        const OutputTransform tr = output_transform(c, d, d_file, hull);
        if (tr.on) {
            T_d.reset(new DevBuf<float>(tr.T.data(), tr.T.size()));
            s_d.reset(new DevBuf<float>(tr.shift.data(), tr.shift.size()));
            out_cols = tr.rows;
        }
        prof.original_dim = out_cols;
        status_phase(c.STATUS, "gpu sampling (streamed)");
        t0 = now_seconds();
        std::unique_ptr<NpyStreamWriter> w;
        if (c.WRITE_DATA) w.reset(new NpyStreamWriter(c.OUT_DIR + "/samples.npy", ns, out_cols));
        struct Ctx {
            NpyStreamWriter* w;
        } ctx{w.get()};
        auto sink = [](const float* rows, int64_t n_rows, int, int64_t, void* user) -> int {
            Ctx* cx = static_cast<Ctx*>(user);
            try {
                if (cx->w) cx->w->append(rows, n_rows);
            } catch (const std::exception& e) {
                std::cerr << e.what() << std::endl;
                return 1;
            }
            return 0;
        };
        napi(naja_chr_sample_streamed(A_d.p, b_d.p, m, d, X_d.p, &p, c.STREAM_CHUNK, T_d ? T_d->p : nullptr,
                                      s_d ? s_d->p : nullptr, out_cols, sink, &ctx, nullptr),
             "CoordinateHitAndRunStreamed");
        if (w) w->close();
        prof.sampling = now_seconds() - t0;
        prof.throughput = (double)c.N_CHAINS * c.N_SAMPLES / prof.sampling;
        prof.total = now_seconds() - t_total;
        write_profile_json(c.OUT_DIR + "/profile.json", prof, c);
        write_config_snapshot(c, prof);
        if (c.STATUS)
            std::cout << "> done (streamed): " << std::fixed << std::setprecision(1) << prof.sampling << "s"
                      << std::defaultfloat << std::endl;
        return 0;
    }
    DevBuf<float> Y_d((size_t)ns * d);
    std::vector<float> out;
    int out_dim = d;
    const OutputTransform tr = output_transform(c, d, d_file, hull);
    if (tr.on) {
        prof.original_dim = tr.rows;
        DevBuf<float> T_d(tr.T.data(), tr.T.size()), s_d(tr.shift.data(), tr.shift.size());
        DevBuf<float> V_d((size_t)ns * tr.rows);
        status_phase(c.STATUS, hull.enabled && !c.BACK_TRANSFORM ? "gpu sampling+backmap (affine hull)"
                                                                   : "gpu sampling+backmap");
        t0 = now_seconds();
        napi(naja_chr_sample(A_d.p, b_d.p, m, d, X_d.p, &p, Y_d.p, nullptr, nullptr), "CoordinateHitAndRun");
        napi(naja_backmap(T_d.p, s_d.p, Y_d.p, V_d.p, tr.rows, d, ns, nullptr), "backmap_into");
        HIPCHK(hipDeviceSynchronize());
        prof.sampling = now_seconds() - t0;
        status_phase(c.STATUS, "downloading");
        t0 = now_seconds();
        out.resize((size_t)ns * tr.rows);
        HIPCHK(hipMemcpy(out.data(), V_d.p, sizeof(float) * out.size(), hipMemcpyDeviceToHost));
        prof.download = now_seconds() - t0;
        out_dim = tr.rows;
    } else {
        status_phase(c.STATUS, "gpu sampling");
        t0 = now_seconds();
        napi(naja_chr_sample(A_d.p, b_d.p, m, d, X_d.p, &p, Y_d.p, nullptr, nullptr), "CoordinateHitAndRun");
        HIPCHK(hipDeviceSynchronize());
        prof.sampling = now_seconds() - t0;
        prof.original_dim = d;
        status_phase(c.STATUS, "downloading");
        t0 = now_seconds();
        out.resize((size_t)ns * d);
        HIPCHK(hipMemcpy(out.data(), Y_d.p, sizeof(float) * out.size(), hipMemcpyDeviceToHost));
        prof.download = now_seconds() - t0;
    }
    prof.throughput = (double)c.N_CHAINS * c.N_SAMPLES / prof.sampling;
    if (c.STATUS)
        std::cout << "> done: " << std::fixed << std::setprecision(1) << prof.sampling << "s, " << std::setprecision(0)
                  << prof.throughput << " samples/s" << std::defaultfloat << std::endl;

    if (c.WRITE_DATA) {
        status_phase(c.STATUS, "writing npy");
        t0 = now_seconds();
        save_npy_f32(c.OUT_DIR + "/samples.npy", out.data(), ns, out_dim);
        prof.write = now_seconds() - t0;
    }
    if (c.BOUNDS_FILTER) {
        status_phase(c.STATUS, "bounds filter (GEM validity)");
        bounds_filter_and_write(c, out, ns, out_dim);
    }
    prof.total = now_seconds() - t_total;
    write_profile_json(c.OUT_DIR + "/profile.json", prof, c);
    write_config_snapshot(c, prof);
    if (c.VERBOSE) std::cout << "output      :: " << make_absolute_path(c.OUT_DIR) << std::endl;
    return 0;
}

// cmd_run.cpp:44-229
int cmd_run(int argc, char** argv)
{
    RunConfig c;
    std::string model_dir, out_root, bounds_policy = "ignore";
    bool quiet = false, dry_run = false, print_config = false;
    for (int i = 0; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--model-dir") model_dir = next_arg(i, argc, argv, a);
        else if (a == "--out-root") out_root = next_arg(i, argc, argv, a);
        else if (a == "--gpu") c.GPU_DEVICE = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--n-chains") c.N_CHAINS = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--n-samples") c.N_SAMPLES = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--tpb") c.TPB_SS = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--thinning") c.THINNING = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--pair-prob") c.PAIR_PROB = std::stod(next_arg(i, argc, argv, a));
        else if (a == "--resync-interval") c.RESYNC_INTERVAL = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--iter-rounding-passes") c.ITER_ROUNDING_PASSES = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--iter-rounding-warmup") c.ITER_ROUNDING_WARMUP = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--extra-constraint-eps") c.EXTRA_CONSTRAINT_EPS = std::stod(next_arg(i, argc, argv, a));
        else if (a == "--constraint-eps") c.CONSTRAINT_EPS = std::stod(next_arg(i, argc, argv, a));
        else if (a == "--pair-schedule" || a == "--pair-schedule-file") c.PAIR_SCHEDULE = next_arg(i, argc, argv, a);
        else if (a == "--start-policy") c.START_POLICY = next_arg(i, argc, argv, a);
        else if (a == "--ksparse-prob") c.KSPARSE_PROB = std::stod(next_arg(i, argc, argv, a));
        else if (a == "--ksparse-k") c.KSPARSE_K = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--extra-constraints") c.EXTRA_CONSTRAINTS = next_arg(i, argc, argv, a);
        else if (a == "--affine-hull-tol") c.AFFINE_HULL_TOL = std::stod(next_arg(i, argc, argv, a));
        else if (a == "--backmap") c.BACK_TRANSFORM = true;
        else if (a == "--write-npy") c.WRITE_DATA = true;
        else if (a == "--stream-chunk") c.STREAM_CHUNK = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--verbose") c.VERBOSE = true;
        else if (a == "--bounds-policy") bounds_policy = next_arg(i, argc, argv, a);
        else if (a == "--bounds-eps") c.BOUNDS_EPS = std::stod(next_arg(i, argc, argv, a));
        else if (a == "--write-samples-valid") c.WRITE_SAMPLES_VALID = true;
        else if (a == "--quiet") quiet = true;
        else if (a == "--dry-run") dry_run = true;
        else if (a == "--print-config") print_config = true;
        else if (a == "--seed") {
            c.SEED = std::stoull(next_arg(i, argc, argv, a));
            c.HAS_SEED = true;
        } else if (a == "--chains-per-group") c.CHAINS_PER_GROUP = std::stoi(next_arg(i, argc, argv, a));
        else die_usage("unknown flag: " + a);
    }
    if (model_dir.empty()) die_usage("missing --model-dir");
    if (out_root.empty()) die_usage("missing --out-root");
    if (c.N_CHAINS <= 0) die_usage("invalid --n-chains");
    if (c.N_SAMPLES <= 0) die_usage("invalid --n-samples");
    if (c.PAIR_PROB < 0.0 || c.PAIR_PROB > 1.0) die_usage("invalid --pair-prob");
    if (c.RESYNC_INTERVAL < 0) die_usage("invalid --resync-interval");
    if (c.ITER_ROUNDING_PASSES < 0) die_usage("invalid --iter-rounding-passes");
    if (c.ITER_ROUNDING_WARMUP < 0) die_usage("invalid --iter-rounding-warmup");
    if (c.EXTRA_CONSTRAINT_EPS < 0.0) die_usage("invalid --extra-constraint-eps (must be >=0)");
    if (c.CONSTRAINT_EPS < 0.0) die_usage("invalid --constraint-eps (must be >=0)");
    if (c.START_POLICY != "file" && c.START_POLICY != "cube_center") die_usage("invalid --start-policy (file|cube_center)");
    if (c.KSPARSE_PROB < 0.0 || c.KSPARSE_PROB > 1.0) die_usage("invalid --ksparse-prob (must be in [0,1])");
    if (c.KSPARSE_K <= 0) die_usage("invalid --ksparse-k (must be >0)");
    if (c.AFFINE_HULL_TOL < 0.0) die_usage("invalid --affine-hull-tol (must be >=0)");
    if (c.EXTRA_CONSTRAINTS != "auto" && c.EXTRA_CONSTRAINTS != "ignore" && c.EXTRA_CONSTRAINTS != "require")
        die_usage("invalid --extra-constraints (auto|ignore|require)");
    if (bounds_policy != "ignore" && bounds_policy != "filter") die_usage("invalid --bounds-policy: " + bounds_policy);
    if (c.STREAM_CHUNK < 0) die_usage("--stream-chunk must be >= 0");
    if (c.STREAM_CHUNK > 0 && bounds_policy == "filter")
        die_usage("--stream-chunk writes samples.npy in pieces; --bounds-policy filter needs them all in memory");
    if (c.CHAINS_PER_GROUP != 1 && c.CHAINS_PER_GROUP != 4) die_usage("invalid --chains-per-group (1|4)");
    if (bounds_policy == "filter" && !c.BACK_TRANSFORM) throw std::runtime_error("bounds-policy=filter requires --backmap");
    c.STATUS = !quiet;
    c.BOUNDS_FILTER = bounds_policy == "filter";
    if (!c.PAIR_SCHEDULE.empty()) (void)load_pair_schedule(c.PAIR_SCHEDULE);

    // model_contract.cpp:53-84
    std::string md = model_dir;
    if (!md.empty() && md.back() == '/') md.pop_back();
    if (!is_directory(md)) throw std::runtime_error("model-dir is not a directory: " + md);
    const size_t slash = md.find_last_of('/');
    c.MODEL_NAME = slash == std::string::npos ? md : md.substr(slash + 1);
    c.DATA_DIR = slash == std::string::npos ? "." : md.substr(0, slash);
    const std::string rdir = md + "/rounding";
    if (!is_directory(rdir)) throw std::runtime_error("missing rounding/ dir: " + rdir);
    const std::string pre = rdir + "/" + c.MODEL_NAME + "_rounding_";
    require_nonempty_file(pre + "A.csv", "rounding_A");
    require_nonempty_file(pre + "b.csv", "rounding_b");
    require_nonempty_file(pre + "start.csv", "rounding_start");
    if (c.BACK_TRANSFORM) {
        require_nonempty_file(pre + "T.csv", "rounding_T");
        require_nonempty_file(pre + "shift.csv", "rounding_shift");
    }
    {
        const bool ha = path_exists(pre + "extra_A.csv"), hb = path_exists(pre + "extra_b.csv");
        if (ha != hb)
            throw std::runtime_error("extra constraints must be both-present or both-absent: " + pre + "extra_A.csv / " +
                                     pre + "extra_b.csv");
        if (ha) {
            require_nonempty_file(pre + "extra_A.csv", "extra_A");
            require_nonempty_file(pre + "extra_b.csv", "extra_b");
        }
    }
    const std::string inherited = rdir + "/INHERITED_FROM.txt";
    if (path_exists(inherited)) {
        std::cout << "rounding_provenance :: " << inherited << "\n";
        std::ifstream f(inherited);
        std::string line;
        while (std::getline(f, line))
            if (!line.empty()) std::cout << "  " << line << "\n";
    }
    c.OUT_DIR = allocate_run_dir(out_root, c.MODEL_NAME);
    c.derive_paths();

    status_phase(c.STATUS, "validate model contract");
    status_phase(c.STATUS, "write generated config");
    const std::string gen = c.OUT_DIR + "/config_generated.txt";
    write_generated_config(gen, c);
    c.source_file = make_absolute_path(gen);
    if (print_config) {
        std::cout << "config_generated    :: " << c.source_file << "\n---\n";
        std::ifstream f(gen);
        std::string line;
        while (std::getline(f, line)) std::cout << line << "\n";
        std::cout << "---\n";
    }
    std::vector<std::string> full = {"naja", "sample", "run"};
    for (int i = 0; i < argc; ++i) full.emplace_back(argv[i]);
    write_run_manifest(c, full);
    status_kv(c.STATUS, "model", c.MODEL_NAME);
    status_kv(c.STATUS, "output", make_absolute_path(c.OUT_DIR));
    status_kv(c.STATUS, "bounds", c.BOUNDS_FILTER ? "filter" : "ignore");
    if (c.BOUNDS_FILTER) status_kv(c.STATUS, "bounds_eps", std::to_string(c.BOUNDS_EPS));
    if (dry_run) {
        status_phase(c.STATUS, "dry-run: not executing sampling");
        return 0;
    }
    status_phase(c.STATUS, "sampling");
    const int rc = run_sampling_job(c);
    if (rc != 0) throw std::runtime_error("sampling failed with return code " + std::to_string(rc));
    return 0;
}

// ----------------------------------------------------------------- validate

float percentile(std::vector<float> v, float q)
{
    if (v.empty()) return 0.0f;
    std::sort(v.begin(), v.end());
    const float pos = q * (float)(v.size() - 1);
    const int lo = (int)std::floor(pos), hi = (int)std::ceil(pos);
    if (lo == hi) return v[(size_t)lo];
    const float t = pos - (float)lo;
    return v[(size_t)lo] * (1.0f - t) + v[(size_t)hi] * t;
}

void json_array(std::ostream& o, const std::vector<float>& v)
{
    o << "[";
    for (size_t i = 0; i < v.size(); ++i) {
        if (i) o << ",";
        o << std::setprecision(4) << v[i];
    }
    o << "]";
}

// validate_cli.cpp:49-296
int cmd_validate(int argc, char** argv)
{
    std::string samples_path, model_dir, layout_s = "interleaved";
    int n_chains = 4;
    for (int i = 0; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--samples") samples_path = next_arg(i, argc, argv, a);
        else if (a == "--model-dir") model_dir = next_arg(i, argc, argv, a);
        else if (a == "--n-chains") n_chains = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--chain-layout") layout_s = next_arg(i, argc, argv, a);
        else die_usage("unknown flag: " + a);
    }
    if (samples_path.empty()) die_usage("missing --samples");
    if (n_chains < 1) die_usage("--n-chains must be >= 1");
    int layout;
    if (layout_s == "interleaved") layout = NAJA_LAYOUT_INTERLEAVED;
    else if (layout_s == "naja-blocks" || layout_s == "blocks") layout = NAJA_LAYOUT_BLOCKS;
    else die_usage("invalid --chain-layout (interleaved|naja-blocks)");

    std::cerr << "loading " << samples_path << "..." << std::endl;
    NpyArray a = load_npy_2d(samples_path);
    // smaller axis = dim (validate_cli.cpp:93-95); build (n_total, dim) rows
    const int64_t dim = std::min(a.rows, a.cols), n_total = std::max(a.rows, a.cols);
    std::vector<float> rows((size_t)dim * n_total);
    if (a.rows <= a.cols) {  // file (dim, nsamp)
        for (int64_t d = 0; d < dim; ++d)
            for (int64_t s = 0; s < n_total; ++s) rows[(size_t)(s * dim + d)] = a.f32[(size_t)(d * n_total + s)];
    } else {
        rows.swap(a.f32);
    }
    const int64_t spc = n_total / n_chains;
    std::cerr << "  shape: (" << dim << ", " << n_total << ") = " << n_chains << " chains x " << spc << " samples"
              << std::endl;
    if (n_total % n_chains != 0)
        throw std::runtime_error("n_total (" + std::to_string(n_total) + ") not divisible by n_chains (" +
                                 std::to_string(n_chains) + ")");
    int active = 0;
    for (int64_t d = 0; d < dim; ++d) {
        double mean = 0.0;
        for (int64_t s = 0; s < n_total; ++s) mean += rows[(size_t)(s * dim + d)];
        mean /= (double)n_total;
        double var = 0.0;
        for (int64_t s = 0; s < n_total; ++s) {
            const double df = rows[(size_t)(s * dim + d)] - mean;
            var += df * df;
        }
        if (var / (double)n_total > 1e-20) ++active;
    }

    std::cerr << "  computing ESS..." << std::endl;
    std::cerr << "  computing split-R-hat..." << std::endl;
    std::vector<float> ess((size_t)dim), rhat((size_t)dim);
    {
        DevBuf<float> S(rows.data(), rows.size()), E((size_t)dim), R((size_t)dim);
        napi(naja_diag(S.p, (int)dim, (int)n_total, n_chains, layout, E.p, R.p, nullptr, nullptr), "diagnostics");
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(ess.data(), E.p, sizeof(float) * ess.size(), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(rhat.data(), R.p, sizeof(float) * rhat.size(), hipMemcpyDeviceToHost));
    }
    const float ess_min = *std::min_element(ess.begin(), ess.end());
    const float ess_max = *std::max_element(ess.begin(), ess.end());
    const float ess_p10 = percentile(ess, 0.10f), ess_med = percentile(ess, 0.50f);
    const float r_med = percentile(rhat, 0.50f), r_max = *std::max_element(rhat.begin(), rhat.end());
    int a11 = 0, a12 = 0;
    for (float r : rhat) {
        if ((double)r > 1.1) ++a11;
        if ((double)r > 1.2) ++a12;
    }

    bool has_bounds = false, has_chords = false;
    int64_t n_viol = 0, n_checked = 0;
    double worst = 0.0;
    std::vector<double> chord;
    if (!model_dir.empty()) {
        std::string md = model_dir;
        if (!md.empty() && md.back() == '/') md.pop_back();
        const size_t sl = md.find_last_of('/');
        const std::string name = sl == std::string::npos ? md : md.substr(sl + 1);
        const std::string pre = md + "/rounding/" + name + "_rounding_";
        const std::string lb_path = md + "/gem/l_bounds.csv", ub_path = md + "/gem/u_bounds.csv";
        if (path_exists(pre + "T.csv") && path_exists(lb_path)) {
            // bounds_check.cpp:8-39: <= 1000 evenly strided samples, eps 1e-6
            std::cerr << "  checking bounds compliance..." << std::endl;
            const Matrix T = load_csv_matrix(pre + "T.csv");
            const std::vector<double> shift = load_csv_vector(pre + "shift.csv");
            const std::vector<double> lb = load_csv_vector(lb_path), ub = load_csv_vector(ub_path);
            const int64_t step = std::max<int64_t>(1, n_total / 1000);
            for (int64_t j = 0; j < n_total; j += step) {
                for (int r = 0; r < T.rows; ++r) {
                    double v = shift[(size_t)r];
                    for (int k = 0; k < T.cols; ++k) v += T.at(r, k) * (double)rows[(size_t)(j * dim + k)];
                    ++n_checked;
                    const double viol = std::max(lb[(size_t)r] - v - 1e-6, v - ub[(size_t)r] - 1e-6);
                    if (viol > 0.0) {
                        ++n_viol;
                        worst = std::max(worst, viol);
                    }
                }
            }
            has_bounds = true;
        }
        if (path_exists(pre + "A.csv") && path_exists(pre + "b.csv") && path_exists(pre + "start.csv")) {
            Polytope P;
            P.A = load_csv_matrix(pre + "A.csv");
            P.b = load_csv_vector(pre + "b.csv");
            P.x0 = load_csv_vector(pre + "start.csv");
            if (path_exists(pre + "extra_A.csv") && path_exists(pre + "extra_b.csv")) {
                const Matrix E = load_csv_matrix(pre + "extra_A.csv");
                const std::vector<double> f = load_csv_vector(pre + "extra_b.csv");
                P.A.v.insert(P.A.v.end(), E.v.begin(), E.v.end());
                P.A.rows += E.rows;
                P.b.insert(P.b.end(), f.begin(), f.end());
            }
            // chord_lengths.cpp:9-42
            const std::vector<double> r = residual(P);
            chord.resize((size_t)P.A.cols);
            for (int j = 0; j < P.A.cols; ++j) {
                double imin = INFINITY, imax = -INFINITY;
                for (int i = 0; i < P.A.rows; ++i) {
                    const double s = -r[(size_t)i], ae = P.A.at(i, j);
                    double inv;
                    if (s == 0.0) inv = ae > 0.0 ? INFINITY : (ae < 0.0 ? -INFINITY : 0.0);
                    else {
                        inv = ae / s;
                        if (std::isnan(inv)) inv = 0.0;
                    }
                    imin = std::min(imin, inv);
                    imax = std::max(imax, inv);
                }
                chord[(size_t)j] = 1.0 / imax - 1.0 / imin;
            }
            has_chords = true;
        }
    }

    std::string status = "OK";
    if (a12 > dim / 4) status = "FAIL";
    else if (a11 > dim / 4 || ess_min < 10) status = "WARN";

    std::cerr << "\n[" << status << "] ESS min=" << std::setprecision(0) << std::fixed << ess_min << " p10=" << ess_p10
              << " med=" << ess_med << "  R-hat med=" << std::setprecision(3) << r_med << " max=" << r_max
              << " >1.1=" << a11 << " >1.2=" << a12;
    if (has_bounds) std::cerr << "  bounds=" << n_viol << "/" << n_checked;
    std::cerr << std::endl;

    std::cout << "{\n  \"samples_file\": \"" << samples_path << "\",\n  \"shape\": [" << n_total << ", " << dim
              << "],\n  \"n_chains\": " << n_chains << ",\n  \"n_samples_per_chain\": " << spc
              << ",\n  \"chain_layout\": \"" << (layout == NAJA_LAYOUT_BLOCKS ? "naja-blocks" : "interleaved")
              << "\",\n  \"active_dims\": " << active << ",\n  \"status\": \"" << status << "\",\n";
    std::cout << "  \"ess\": {\n    \"min\": " << std::setprecision(1) << std::fixed << ess_min << ",\n    \"p10\": "
              << ess_p10 << ",\n    \"median\": " << ess_med << ",\n    \"max\": " << ess_max << "\n  },\n";
    std::cout << "  \"rhat\": {\n    \"median\": " << std::setprecision(4) << r_med << ",\n    \"max\": " << r_max
              << ",\n    \"above_1_1\": " << a11 << ",\n    \"above_1_2\": " << a12 << "\n  },\n";
    if (has_bounds)
        std::cout << "  \"bounds\": {\n    \"violations\": " << n_viol << ",\n    \"checked\": " << n_checked
                  << ",\n    \"worst\": " << std::setprecision(6) << worst << "\n  },\n";
    if (has_chords) {
        int pos = 0;
        double cmin = 1e30, cmax = -INFINITY;
        for (double v : chord) {
            if (v > 0) {
                ++pos;
                cmin = std::min(cmin, v);
            }
            cmax = std::max(cmax, v);
        }
        if (pos == 0) cmin = 0;
        std::cout << "  \"chords\": {\n    \"positive_dims\": " << pos << ",\n    \"min_positive\": "
                  << std::setprecision(6) << cmin << ",\n    \"max\": " << cmax << "\n  },\n";
    }
    std::cout << "  \"per_dim_ess\": ";
    json_array(std::cout, ess);
    std::cout << ",\n  \"per_dim_rhat\": ";
    json_array(std::cout, rhat);
    std::cout << "\n}\n";
    return 0;
}

}  // namespace

// Per-pair population variance/covariance of one warmup chain (rows of d),
// accumulated into the sums over models (cmd_calibrate_rounding.cpp:136-156).
void accumulate_pair_moments(const float* W, int d, int warmup, const PairSchedule& S, std::vector<double>& vi_sum,
                             std::vector<double>& vj_sum, std::vector<double>& cov_sum)
{
    for (size_t k = 0; k < S.i.size(); ++k) {
        const int ii = S.i[k], jj = S.j[k];
        double mi = 0, mj = 0;
        for (int t = 0; t < warmup; ++t) {
            mi += W[(size_t)t * d + ii];
            mj += W[(size_t)t * d + jj];
        }
        mi /= warmup;
        mj /= warmup;
        double vi = 0, vj = 0, cij = 0;
        for (int t = 0; t < warmup; ++t) {
            const double di = W[(size_t)t * d + ii] - mi, dj = W[(size_t)t * d + jj] - mj;
            vi += di * di;
            vj += dj * dj;
            cij += di * dj;
        }
        vi_sum[k] += vi / warmup;
        vj_sum[k] += vj / warmup;
        cov_sum[k] += cij / warmup;
    }
}

// cmd_calibrate_rounding.cpp:24-195: one Jacobi pair schedule for a set of
// models. Disjoint pairs from a seeded shuffle; each pass runs one warmup
// chain per model on rotated pairs (pair mode 2), averages the per-pair
// moments over models and sets each pair's rotation. Models that share A and
// b (the conditioned variants of one base model) run as one batched launch
// per pass (variant = model index, seed ^ (pass + 1)); otherwise one launch
// per model with the reference's seed ^ model ^ (pass + 1).
int cmd_calibrate_rounding(int argc, char** argv)
{
    std::string models_root, model_list, out_csv;
    int warmup = 2000, passes = 2, seed = 1, max_models = 0, gpu = 0;
    double pair_prob = 0.3, extra_eps = 0.0;
    for (int i = 0; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--models-root") models_root = next_arg(i, argc, argv, a);
        else if (a == "--model-list") model_list = next_arg(i, argc, argv, a);
        else if (a == "--out") out_csv = next_arg(i, argc, argv, a);
        else if (a == "--warmup") warmup = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--passes") passes = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--pair-prob") pair_prob = std::stod(next_arg(i, argc, argv, a));
        else if (a == "--extra-constraint-eps") extra_eps = std::stod(next_arg(i, argc, argv, a));
        else if (a == "--seed") seed = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--max-models") max_models = std::stoi(next_arg(i, argc, argv, a));
        else if (a == "--gpu") gpu = std::stoi(next_arg(i, argc, argv, a));
        else die_usage("unknown flag: " + a);
    }
    if (models_root.empty()) die_usage("missing --models-root");
    if (model_list.empty()) die_usage("missing --model-list");
    if (out_csv.empty()) die_usage("missing --out");
    if (warmup <= 0) die_usage("invalid --warmup");
    if (passes <= 0) die_usage("invalid --passes");
    if (!(pair_prob > 0.0 && pair_prob <= 1.0)) die_usage("invalid --pair-prob (0,1]");
    if (extra_eps < 0.0) die_usage("invalid --extra-constraint-eps (must be >=0)");
    if (max_models < 0) die_usage("invalid --max-models (must be >=0)");
    if (seed == 0) seed = 1;
    std::vector<std::string> names;
    for (const auto& l : read_nonempty_trimmed_lines(model_list))
        if (l[0] != '#') names.push_back(l);
    if (names.empty()) throw std::runtime_error("no models in list: " + model_list);
    if (max_models > 0 && (int)names.size() > max_models) names.resize((size_t)max_models);

    // Load every model: base rows, extra rows (+ eps), start.
    struct Model {
        Matrix A;
        std::vector<double> b, x0;
        Matrix E;
        std::vector<double> f;
    };
    std::vector<Model> ms;
    for (const auto& name : names) {
        std::string md = make_absolute_path(models_root + "/" + name);
        if (!md.empty() && md.back() == '/') md.pop_back();
        const std::string mname = md.substr(md.find_last_of('/') + 1);
        const std::string pre = md + "/rounding/" + mname + "_rounding_";
        Model m;
        m.A = load_csv_matrix(pre + "A.csv");
        m.b = load_csv_vector(pre + "b.csv");
        m.x0 = load_csv_vector(pre + "start.csv");
        if (path_exists(pre + "extra_A.csv") && path_exists(pre + "extra_b.csv")) {
            m.E = load_csv_matrix(pre + "extra_A.csv");
            m.f = load_csv_vector(pre + "extra_b.csv");
            for (double& v : m.f) v += extra_eps;
            if ((size_t)m.E.rows != m.f.size() || m.E.cols != m.A.cols)
                throw std::runtime_error("extra constraint matrices have mismatched dimensions: " + pre);
        } else {
            m.E.cols = m.A.cols;
        }
        if ((int)m.x0.size() != m.A.cols || (int)m.b.size() != m.A.rows)
            throw std::runtime_error("dimension mismatch in " + pre);
        ms.push_back(std::move(m));
    }
    const int d = ms.front().A.cols;
    if (d < 2) throw std::runtime_error("calibrate-rounding requires reduced_dim >= 2");
    for (const auto& m : ms)
        if (m.A.cols != d) throw std::runtime_error("calibrate-rounding: models differ in reduced dimension");
    bool shared = true;
    for (const auto& m : ms) shared = shared && m.A.rows == ms[0].A.rows && m.A.v == ms[0].A.v && m.b == ms[0].b;

    std::vector<int> perm((size_t)d);
    for (int i = 0; i < d; ++i) perm[(size_t)i] = i;
    std::mt19937_64 rng((uint64_t)seed);
    std::shuffle(perm.begin(), perm.end(), rng);
    const int np = d / 2;
    PairSchedule S;
    std::vector<double> sc((size_t)np, 1.0), ss((size_t)np, 0.0);  // FP64 angles as the reference keeps them
    for (int k = 0; k < np; ++k) {
        S.i.push_back(perm[(size_t)(2 * k)]);
        S.j.push_back(perm[(size_t)(2 * k + 1)]);
        S.c.push_back(1.0f);
        S.s.push_back(0.0f);
    }
    HIPCHK(hipSetDevice(gpu));
    const int nm = (int)ms.size();
    const std::vector<float> A0cm = colmajor_f32(ms[0].A);
    DevBuf<float> A0_d(A0cm.data(), A0cm.size());
    DevBuf<double> b0_d(ms[0].b.data(), ms[0].b.size());
    // batched layout of every model's extra rows and starts
    std::vector<int> ks;
    std::vector<float> Ecm, X0;
    std::vector<double> fe;
    for (const auto& m : ms) {
        ks.push_back(m.E.rows);
        const std::vector<float> e = colmajor_f32(m.E);
        Ecm.insert(Ecm.end(), e.begin(), e.end());
        fe.insert(fe.end(), m.f.begin(), m.f.end());
        for (double v : m.x0) X0.push_back((float)v);
    }
    DevBuf<float> E_d(Ecm.data(), Ecm.size()), X_d(X0.data(), X0.size());
    DevBuf<double> f_d(fe.data(), fe.size());
    DevBuf<float> W_d((size_t)(shared ? nm : 1) * warmup * d);
    std::vector<float> W((size_t)nm * warmup * d);
    for (int pass = 0; pass < passes; ++pass) {
        DevBuf<int> pi(S.i.data(), S.i.size()), pj(S.j.data(), S.j.size());
        DevBuf<float> pc(S.c.data(), S.c.size()), ps(S.s.data(), S.s.size());
        naja_chr_params p{};
        p.nspc = warmup;
        p.thinning = 1;
        p.nchains = 1;
        p.tpb = 128;
        p.pair_prob = pair_prob;
        p.ksparse_k = 8;
        p.pair_mode = 2;
        p.pair_i = pi.p;
        p.pair_j = pj.p;
        p.pair_c = pc.p;
        p.pair_s = ps.p;
        p.n_pairs = np;
        p.chains_per_group = 1;
        if (shared) {
            p.seed = (uint64_t)(int64_t)(seed ^ (pass + 1));
            napi(naja_chr_sample_batched(A0_d.p, b0_d.p, ms[0].A.rows, d, nm, ks.data(), Ecm.empty() ? nullptr : E_d.p,
                                         fe.empty() ? nullptr : f_d.p, X_d.p, &p, W_d.p, nullptr, nullptr),
                 "warmup CHR (batched)");
            HIPCHK(hipDeviceSynchronize());
            HIPCHK(hipMemcpy(W.data(), W_d.p, sizeof(float) * W.size(), hipMemcpyDeviceToHost));
        } else {
            for (int mi = 0; mi < nm; ++mi) {
                Matrix A = ms[(size_t)mi].A;
                std::vector<double> b = ms[(size_t)mi].b;
                A.v.insert(A.v.end(), ms[(size_t)mi].E.v.begin(), ms[(size_t)mi].E.v.end());
                A.rows += ms[(size_t)mi].E.rows;
                b.insert(b.end(), ms[(size_t)mi].f.begin(), ms[(size_t)mi].f.end());
                const std::vector<float> Acm = colmajor_f32(A);
                DevBuf<float> Ad(Acm.data(), Acm.size());
                DevBuf<double> bd(b.data(), b.size());
                p.seed = (uint64_t)(int64_t)(seed ^ mi ^ (pass + 1));
                napi(naja_chr_sample(Ad.p, bd.p, A.rows, d, X_d.p + (size_t)mi * d, &p, W_d.p, nullptr, nullptr),
                     "warmup CHR");
                HIPCHK(hipDeviceSynchronize());
                HIPCHK(hipMemcpy(W.data() + (size_t)mi * warmup * d, W_d.p, sizeof(float) * (size_t)warmup * d,
                                 hipMemcpyDeviceToHost));
            }
        }
        std::vector<double> vi((size_t)np, 0.0), vj((size_t)np, 0.0), cv((size_t)np, 0.0);
        for (int mi = 0; mi < nm; ++mi)
            accumulate_pair_moments(W.data() + (size_t)mi * warmup * d, d, warmup, S, vi, vj, cv);
        for (int k = 0; k < np; ++k) {
            const auto cs = jacobi_rotation_cs(vi[(size_t)k] / nm, vj[(size_t)k] / nm, cv[(size_t)k] / nm);
            sc[(size_t)k] = cs.first;
            ss[(size_t)k] = cs.second;
            S.c[(size_t)k] = (float)cs.first;
            S.s[(size_t)k] = (float)cs.second;
        }
    }
    // schedule_io.cpp:77-89: i,j,c,s per line at the stream's default precision
    std::ofstream f(out_csv);
    if (!f.is_open()) throw std::runtime_error("cannot write pair schedule: " + out_csv);
    for (int k = 0; k < np; ++k)
        f << S.i[(size_t)k] << "," << S.j[(size_t)k] << "," << sc[(size_t)k] << "," << ss[(size_t)k] << "\n";
    f.close();
    std::cout << "OK\nout              :: " << make_absolute_path(out_csv) << "\nmodels           :: " << nm
              << "\nreduced_dim      :: " << d << "\npairs            :: " << np << "\npasses           :: " << passes
              << "\nwarmup           :: " << warmup << "\npair_prob        :: " << pair_prob
              << "\nextra_eps        :: " << extra_eps << "\nbatched          :: " << (shared ? "yes" : "no") << "\n";
    return 0;
}

int cmd_prepare(int argc, char** argv);

int main(int argc, char** argv)
{
    try {
        if (argc >= 3 && std::string(argv[1]) == "sample" && std::string(argv[2]) == "run")
            return cmd_run(argc - 3, argv + 3);
        if (argc >= 2 && std::string(argv[1]) == "validate") return cmd_validate(argc - 2, argv + 2);
        if (argc >= 2 && std::string(argv[1]) == "prepare") return cmd_prepare(argc - 2, argv + 2);
        if (argc >= 3 && std::string(argv[1]) == "sample" && std::string(argv[2]) == "calibrate-rounding")
            return cmd_calibrate_rounding(argc - 3, argv + 3);
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    std::cerr << "usage:\n  naja sample run <args...>\n  naja validate --samples <file.npy> [--model-dir <dir>] "
                 "[--n-chains 4] [--chain-layout interleaved|naja-blocks]\n  naja prepare --models-root <dir> "
                 "--model-list <file> [--base-model-dir <dir>] ...\n  naja sample calibrate-rounding --models-root <dir> "
                 "--model-list <file> --out <csv> [--warmup 2000] [--passes 2] [--pair-prob 0.3] [--seed 1] ...\n";
    return 2;
}
