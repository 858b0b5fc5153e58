// Host-side file formats and filesystem helpers of the `naja` binary.
//
// Formats are byte-compatible with the reference's:
//   * CSV matrices: one row per line, comma separated, parsed with
//     std::stod, empty lines skipped (csv_loader.h:12-49); vectors may be a
//     column, a row or a flattened matrix (csv_loader.h:51-60);
//   * NPY v1.0 writers with the reference's header padding rule
//     (npy.h:15-53 float64, npy.h:56-84 uint8, npy.h:96-141 float32 samples);
//   * status lines "> phase" and "key :: value" (status.h:8-16).
#pragma once
#include <cstdio>

#include <cstdint>
#include <string>
#include <vector>

namespace naja::host {

// Row-major dense matrix in FP64 (what the CSV files hold).
struct Matrix {
    int rows = 0;
    int cols = 0;
    std::vector<double> v;
    double& at(int r, int c) { return v[(size_t)r * cols + c]; }
    double at(int r, int c) const { return v[(size_t)r * cols + c]; }
};

Matrix load_csv_matrix(const std::string& path);
std::vector<double> load_csv_vector(const std::string& path);

// NPY header text exactly as npy.h builds it: dict, spaces so that
// (10 + len) is a multiple of 64, then '\n'.
std::string npy_header(const std::string& descr, const std::string& shape);
// (rows, cols) float32 C-order file from rows of `cols` floats.
void save_npy_f32(const std::string& path, const float* data, int64_t rows, int64_t cols);
void save_npy_f64(const std::string& path, const double* data, int64_t rows, int64_t cols);
void save_npy_u8(const std::string& path, const std::vector<uint8_t>& v);

// (rows, cols) float32 NPY written in pieces: the header (final shape) at
// open, rows appended as they arrive, close checks that all rows came.
class NpyStreamWriter {
public:
    NpyStreamWriter(const std::string& path, int64_t rows, int64_t cols);
    ~NpyStreamWriter();
    void append(const float* data, int64_t n_rows);
    void close();
    int64_t rows_written() const { return written_; }

private:
    std::string path_;
    std::FILE* f_ = nullptr;
    int64_t rows_, cols_, written_ = 0;
};

struct NpyArray {
    int64_t rows = 0, cols = 0;
    bool is_f64 = false;
    std::vector<float> f32;  // row-major rows x cols (f64 files are cast)
};
NpyArray load_npy_2d(const std::string& path);

bool path_exists(const std::string& p);
bool is_directory(const std::string& p);
void ensure_dir(const std::string& p);
std::string make_absolute_path(const std::string& p);
std::string current_timestamp();
std::string current_datestamp_compact();
void require_nonempty_file(const std::string& path, const std::string& what);

void status_phase(bool enabled, const std::string& msg);
void status_kv(bool enabled, const std::string& k, const std::string& v);

double now_seconds();

}  // namespace naja::host
