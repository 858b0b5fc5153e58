#include "host_io.h"

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace naja::host {

Matrix load_csv_matrix(const std::string& path)
{
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Cannot open file: " + path);
    Matrix m;
    std::string line, cell;
    std::vector<double> row;
    while (std::getline(f, line)) {
        row.clear();
        std::stringstream ss(line);
        while (std::getline(ss, cell, ',')) row.push_back(std::stod(cell));
        if (row.empty()) continue;
        if (m.rows == 0) m.cols = (int)row.size();
        if ((int)row.size() != m.cols)
            throw std::runtime_error("ragged CSV row " + std::to_string(m.rows + 1) + " in " + path);
        m.v.insert(m.v.end(), row.begin(), row.end());
        ++m.rows;
    }
    if (m.rows == 0) throw std::runtime_error("Empty CSV file: " + path);
    return m;
}

std::vector<double> load_csv_vector(const std::string& path)
{
    Matrix m = load_csv_matrix(path);
    if (m.cols == 1 || m.rows == 1) return m.v;
    // csv_loader.h:57: an r x c matrix flattens column-major
    std::vector<double> out((size_t)m.rows * m.cols);
    for (int c = 0; c < m.cols; ++c)
        for (int r = 0; r < m.rows; ++r) out[(size_t)c * m.rows + r] = m.at(r, c);
    return out;
}

std::string npy_header(const std::string& descr, const std::string& shape)
{
    std::string h = "{'descr': '" + descr + "', 'fortran_order': False, 'shape': (" + shape + "), }";
    const size_t total = 6 + 2 + 2 + h.size();
    const size_t pad = (64 - (total % 64)) % 64;
    h.append(pad, ' ');
    h.push_back('\n');
    return h;
}

static void write_npy(const std::string& path, const std::string& descr, const std::string& shape, const void* data,
                      size_t bytes)
{
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Cannot write file: " + path);
    const std::string h = npy_header(descr, shape);
    f.write("\x93NUMPY", 6);
    f.put(0x01);
    f.put(0x00);
    const uint16_t hl = (uint16_t)h.size();
    f.write(reinterpret_cast<const char*>(&hl), 2);
    f.write(h.data(), (std::streamsize)h.size());
    if (bytes) f.write(reinterpret_cast<const char*>(data), (std::streamsize)bytes);
    if (!f) throw std::runtime_error("write failed: " + path);
}

void save_npy_f32(const std::string& path, const float* data, int64_t rows, int64_t cols)
{
    write_npy(path, "<f4", std::to_string(rows) + ", " + std::to_string(cols), data,
              sizeof(float) * (size_t)rows * (size_t)cols);
}

void save_npy_f64(const std::string& path, const double* data, int64_t rows, int64_t cols)
{
    write_npy(path, "<f8", std::to_string(rows) + ", " + std::to_string(cols), data,
              sizeof(double) * (size_t)rows * (size_t)cols);
}

NpyStreamWriter::NpyStreamWriter(const std::string& path, int64_t rows, int64_t cols)
    : path_(path), rows_(rows), cols_(cols)
{
    f_ = std::fopen(path.c_str(), "wb");
    if (!f_) throw std::runtime_error("Cannot write file: " + path);
    const std::string h = npy_header("<f4", std::to_string(rows) + ", " + std::to_string(cols));
    const uint16_t hl = (uint16_t)h.size();
    const char magic[8] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 0x01, 0x00};
    if (std::fwrite(magic, 1, 8, f_) != 8 || std::fwrite(&hl, 2, 1, f_) != 1 ||
        std::fwrite(h.data(), 1, h.size(), f_) != h.size())
        throw std::runtime_error("write failed: " + path);
}

NpyStreamWriter::~NpyStreamWriter()
{
    if (f_) std::fclose(f_);
}

void NpyStreamWriter::append(const float* data, int64_t n_rows)
{
    if (!f_) throw std::runtime_error("append to a closed NPY stream: " + path_);
    if (written_ + n_rows > rows_) throw std::runtime_error("NPY stream overflow: " + path_);
    const size_t n = (size_t)n_rows * (size_t)cols_;
    if (std::fwrite(data, sizeof(float), n, f_) != n) throw std::runtime_error("write failed: " + path_);
    written_ += n_rows;
}

void NpyStreamWriter::close()
{
    if (!f_) return;
    const bool full = written_ == rows_;
    const bool ok = std::fclose(f_) == 0;
    f_ = nullptr;
    if (!full) throw std::runtime_error("NPY stream closed after " + std::to_string(written_) + " of " +
                                        std::to_string(rows_) + " rows: " + path_);
    if (!ok) throw std::runtime_error("write failed: " + path_);
}

void save_npy_u8(const std::string& path, const std::vector<uint8_t>& v)
{
    write_npy(path, "|u1", std::to_string(v.size()) + ",", v.data(), v.size());
}

NpyArray load_npy_2d(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("cannot open: " + path);
    char magic[6];
    f.read(magic, 6);
    if (std::strncmp(magic, "\x93NUMPY", 6) != 0) throw std::runtime_error("Invalid NPY file: " + path);
    char ver[2];
    f.read(ver, 2);
    uint32_t hl = 0;
    if (ver[0] == 1) {
        uint16_t h16;
        f.read(reinterpret_cast<char*>(&h16), 2);
        hl = h16;
    } else {
        f.read(reinterpret_cast<char*>(&hl), 4);
    }
    std::string header(hl, ' ');
    f.read(&header[0], hl);
    NpyArray a;
    const bool f32 = header.find("'<f4'") != std::string::npos;
    a.is_f64 = header.find("'<f8'") != std::string::npos;
    if (!f32 && !a.is_f64) throw std::runtime_error("unsupported dtype in " + path);
    if (header.find("'fortran_order': True") != std::string::npos)
        throw std::runtime_error("fortran_order arrays are not supported: " + path);
    const size_t sp = header.find("'shape':");
    const size_t p1 = header.find('(', sp);
    const size_t p2 = header.find(')', p1);
    std::string shape = header.substr(p1 + 1, p2 - p1 - 1);
    std::replace(shape.begin(), shape.end(), ',', ' ');
    std::stringstream ss(shape);
    ss >> a.rows >> a.cols;
    if (a.rows <= 0 || a.cols <= 0) throw std::runtime_error("expected a 2-D array in " + path);
    const size_t n = (size_t)a.rows * (size_t)a.cols;
    a.f32.resize(n);
    if (f32) {
        f.read(reinterpret_cast<char*>(a.f32.data()), (std::streamsize)(n * sizeof(float)));
    } else {
        std::vector<double> buf(n);
        f.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(n * sizeof(double)));
        for (size_t i = 0; i < n; ++i) a.f32[i] = (float)buf[i];
    }
    if (!f) throw std::runtime_error("truncated NPY data in " + path);
    return a;
}

bool path_exists(const std::string& p)
{
    struct stat st;
    return !p.empty() && stat(p.c_str(), &st) == 0;
}

bool is_directory(const std::string& p)
{
    struct stat st;
    return !p.empty() && stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

void ensure_dir(const std::string& p)
{
    if (p.empty() || is_directory(p)) return;
    const size_t slash = p.find_last_of('/');
    if (slash != std::string::npos && slash > 0) ensure_dir(p.substr(0, slash));
    if (mkdir(p.c_str(), 0755) != 0 && !is_directory(p)) throw std::runtime_error("cannot create directory: " + p);
}

std::string make_absolute_path(const std::string& p)
{
    char buf[PATH_MAX];
    if (realpath(p.c_str(), buf)) return std::string(buf);
    return p;
}

static std::string fmt_now(const char* fmt)
{
    std::time_t now = std::time(nullptr);
    char buf[64];
    if (std::strftime(buf, sizeof(buf), fmt, std::localtime(&now))) return buf;
    return {};
}

std::string current_timestamp() { return fmt_now("%Y-%m-%d %H:%M:%S"); }
std::string current_datestamp_compact() { return fmt_now("%Y%m%d"); }

void require_nonempty_file(const std::string& path, const std::string& what)
{
    if (!path_exists(path)) throw std::runtime_error("missing " + what + ": " + path);
    struct stat st;
    if (stat(path.c_str(), &st) != 0) throw std::runtime_error("cannot stat " + what + ": " + path);
    if (st.st_size == 0) throw std::runtime_error("empty " + what + ": " + path);
}

void status_phase(bool enabled, const std::string& msg)
{
    if (enabled) std::cout << "> " << msg << std::endl;
}

void status_kv(bool enabled, const std::string& k, const std::string& v)
{
    if (enabled) std::cout << k << " :: " << v << std::endl;
}

double now_seconds()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace naja::host
