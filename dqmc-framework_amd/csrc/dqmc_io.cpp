// On-disk formats of dqmc-framework (QuantumMonteCarlo::DqmcIO, include/dqmc_io.h:267-629):
// the Hubbard-Stratonovich field file (warm restarts) and the mean / bin files of the
// observables, the k-star list and the imaginary-time grid.  Host-only: no device work, so
// these run without a GPU.  Every number is written with the conversion of the reference's
// boost::format directives ("%| 20d|" is printf's "% 20d", "%| 20.10f|" is "% 20.10f").
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dqmc_amd.h"

namespace {

struct File {
    FILE* f;
    File(const char* path, const char* mode) : f(path ? std::fopen(path, mode) : nullptr) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

// k-stars of the square lattice, (kx, ky) per star (src/lattice/square.cpp:42-66)
void kstars(int side, std::vector<double>& kx, std::vector<double>& ky) {
    kx.clear();
    ky.clear();
    for (int i = (int)std::ceil(side / 2.0); i <= side; ++i)
        for (int j = (int)std::ceil(side / 2.0); j <= i; ++j) {
            kx.push_back((double)i / side * 2 * M_PI - M_PI);
            ky.push_back((double)j / side * 2 * M_PI - M_PI);
        }
}

// whitespace-separated fields of one line
std::vector<std::string> split(const char* line) {
    std::vector<std::string> out;
    const char* p = line;
    while (*p) {
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
        if (!*p) break;
        const char* q = p;
        while (*q && *q != ' ' && *q != '\t' && *q != '\r' && *q != '\n') ++q;
        out.emplace_back(p, q - p);
        p = q;
    }
    return out;
}

bool to_int(const std::string& s, long& v) {
    char* end = nullptr;
    v = std::strtol(s.c_str(), &end, 10);
    return end && *end == '\0' && !s.empty();
}
bool to_double(const std::string& s, double& v) {
    char* end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end && *end == '\0' && !s.empty();
}

}  // namespace

extern "C" {

// DqmcIO::output_bosonic_fields (include/dqmc_io.h:478-534): "L N", then "t i s" per site
int dqmc_write_fields_file(const char* path, int time_size, int space_size, const int8_t* fields) {
    if (!path || !fields || time_size < 1 || space_size < 1) return DQMC_ERR_ARG;
    File out(path, "w");
    if (!out.f) return DQMC_ERR_IO;
    std::fprintf(out.f, "% 20d% 20d\n", time_size, space_size);
    for (int t = 0; t < time_size; ++t)
        for (int i = 0; i < space_size; ++i)
            std::fprintf(out.f, "% 20d% 20d% 20.1f\n", t, i, (double)fields[(size_t)t * space_size + i]);
    return std::ferror(out.f) ? DQMC_ERR_IO : DQMC_OK;
}

// DqmcIO::read_bosonic_fields_from_file (include/dqmc_io.h:537-623): the header must match the
// chain's sizes; every "t i s" line sets one field (entries a file leaves out keep their value)
int dqmc_read_fields_file(const char* path, int time_size, int space_size, int8_t* fields) {
    if (!path || !fields || time_size < 1 || space_size < 1) return DQMC_ERR_ARG;
    File in(path, "r");
    if (!in.f) return DQMC_ERR_IO;
    std::vector<char> buf(4096);
    if (!std::fgets(buf.data(), (int)buf.size(), in.f)) return DQMC_ERR_IO;
    auto head = split(buf.data());
    long L = 0, N = 0;
    if (head.size() < 2 || !to_int(head[0], L) || !to_int(head[1], N)) return DQMC_ERR_IO;
    if (L != time_size || N != space_size) return DQMC_ERR_ARG;   // inconsistent time or space size
    while (std::fgets(buf.data(), (int)buf.size(), in.f)) {
        auto w = split(buf.data());
        if (w.empty()) continue;
        long t = 0, i = 0;
        double s = 0.0;
        if (w.size() < 3 || !to_int(w[0], t) || !to_int(w[1], i) || !to_double(w[2], s)) return DQMC_ERR_IO;
        if (t < 0 || t >= L || i < 0 || i >= N || (s != 1.0 && s != -1.0)) return DQMC_ERR_IO;
        fields[(size_t)t * N + i] = s > 0 ? 1 : -1;
    }
    return DQMC_OK;
}

// DqmcIO::output_k_stars (include/dqmc_io.h:432-455)
int dqmc_write_kstars_file(const char* path, int side) {
    if (!path || side < 2) return DQMC_ERR_ARG;
    std::vector<double> kx, ky;
    kstars(side, kx, ky);
    File out(path, "w");
    if (!out.f) return DQMC_ERR_IO;
    std::fprintf(out.f, "% 20d\n", (int)kx.size());
    for (size_t k = 0; k < kx.size(); ++k) std::fprintf(out.f, "% 20d% 20.10f% 20.10f\n", (int)k, kx[k], ky[k]);
    return std::ferror(out.f) ? DQMC_ERR_IO : DQMC_OK;
}

// DqmcIO::output_imaginary_time_grids (include/dqmc_io.h:458-475)
int dqmc_write_tgrids_file(const char* path, int time_size, double beta) {
    if (!path || time_size < 1 || !(beta > 0.0)) return DQMC_ERR_ARG;
    const double dtau = beta / time_size;
    File out(path, "w");
    if (!out.f) return DQMC_ERR_IO;
    std::fprintf(out.f, "% 20d% 20.5f% 20.5f\n", time_size, beta, dtau);
    for (int t = 0; t < time_size; ++t) std::fprintf(out.f, "% 20d% 20.10f\n", t, t * dtau);
    return std::ferror(out.f) ? DQMC_ERR_IO : DQMC_OK;
}

// DqmcIO::output_observable, file branch (include/dqmc_io.h:305-355): mean, error and
// relative error; rank 0 scalar, 1 vector of `rows`, 2 matrix rows x cols (element (i,j) at
// i + rows j, Eigen's order)
int dqmc_write_observable_file(const char* path, int rank, int rows, int cols, const double* mean, const double* err) {
    if (!path || !mean || !err || rank < 0 || rank > 2) return DQMC_ERR_ARG;
    if ((rank >= 1 && rows < 1) || (rank == 2 && cols < 1)) return DQMC_ERR_ARG;
    File out(path, "w");
    if (!out.f) return DQMC_ERR_IO;
    if (rank == 0) {
        std::fprintf(out.f, "% 20.10f% 20.10f% 20.10f\n", mean[0], err[0], err[0] / mean[0]);
    } else if (rank == 1) {
        std::fprintf(out.f, "% 20d\n", rows);
        for (int i = 0; i < rows; ++i) std::fprintf(out.f, "% 20d% 20.10f% 20.10f% 20.10f\n", i, mean[i], err[i], err[i] / mean[i]);
    } else {
        std::fprintf(out.f, "% 20d% 20d\n", rows, cols);
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < cols; ++j) {
                const size_t e = (size_t)j * rows + i;
                std::fprintf(out.f, "% 20d% 20d% 20.10f% 20.10f% 20.10f\n", i, j, mean[e], err[e], err[e] / mean[e]);
            }
    }
    return std::ferror(out.f) ? DQMC_ERR_IO : DQMC_OK;
}

// DqmcIO::output_observable_in_bins (include/dqmc_io.h:365-429): bins[b * size + e] with the
// element order of dqmc_write_observable_file
int dqmc_write_observable_bins_file(const char* path, int rank, int n_bins, int rows, int cols, const double* bins) {
    if (!path || !bins || rank < 0 || rank > 2 || n_bins < 0) return DQMC_ERR_ARG;
    if ((rank >= 1 && rows < 1) || (rank == 2 && cols < 1)) return DQMC_ERR_ARG;
    File out(path, "w");
    if (!out.f) return DQMC_ERR_IO;
    if (rank == 0) {
        std::fprintf(out.f, "% 20d\n", n_bins);
        for (int b = 0; b < n_bins; ++b) std::fprintf(out.f, "% 20d% 20.10f\n", b, bins[b]);
    } else if (rank == 1) {
        std::fprintf(out.f, "% 20d% 20d\n", n_bins, rows);
        for (int b = 0; b < n_bins; ++b)
            for (int i = 0; i < rows; ++i) std::fprintf(out.f, "% 20d% 20d% 20.10f\n", b, i, bins[(size_t)b * rows + i]);
    } else {
        std::fprintf(out.f, "% 20d% 20d% 20d\n", n_bins, rows, cols);
        const size_t size = (size_t)rows * cols;
        for (int b = 0; b < n_bins; ++b)
            for (int i = 0; i < rows; ++i)
                for (int j = 0; j < cols; ++j)
                    std::fprintf(out.f, "% 20d% 20d% 20d% 20.10f\n", b, i, j, bins[b * size + (size_t)j * rows + i]);
    }
    return std::ferror(out.f) ? DQMC_ERR_IO : DQMC_OK;
}

}  // extern "C"
