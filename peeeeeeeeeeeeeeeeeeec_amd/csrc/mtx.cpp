// Matrix Market interchange (include/matrix_export.h; reference
// src/io/matrix_export.c). Host I/O: dense matrices are formatted in
// parallel slices (one "%.17e\n" line per value, column-major) and written
// in order, so the file is the reference's byte for byte.
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/matrix_export.h"
#include "common.h"

namespace {

int ensure_directory(const char *dir) {
    if (!dir || !dir[0]) return -1;
    std::string b(dir);
    for (size_t i = 1; i <= b.size(); ++i) {
        if (i == b.size() || b[i] == '/') {
            const std::string part = b.substr(0, i);
            if (!part.empty() && mkdir(part.c_str(), 0755) != 0 && errno != EEXIST) {
                PEEC_ERR("cannot create directory '%s': %s", part.c_str(), std::strerror(errno));
                return -1;
            }
        }
    }
    return 0;
}

// column-major "%.17e\n" lines of a row-major matrix
int write_dense_values(FILE *f, const double *data, size_t rows, size_t cols) {
    const size_t total = rows * cols;
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t chunk = 1u << 18;  // values per slice
    std::vector<std::string> bufs(hw);
    for (size_t base = 0; base < total; base += chunk * hw) {
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < hw; ++t)
            pool.emplace_back([&, t]() {
                std::string &s = bufs[t];
                s.clear();
                const size_t lo = base + (size_t)t * chunk, hi = std::min(total, lo + chunk);
                char line[64];
                for (size_t idx = lo; idx < hi; ++idx) {
                    const size_t col = idx / rows, row = idx % rows;
                    const int len = std::snprintf(line, sizeof line, "%.17e\n", data[row * cols + col]);
                    s.append(line, (size_t)len);
                }
            });
        for (std::thread &th : pool) th.join();
        for (unsigned t = 0; t < hw; ++t)
            if (!bufs[t].empty() && std::fwrite(bufs[t].data(), 1, bufs[t].size(), f) != bufs[t].size()) return -1;
    }
    return 0;
}

}  // namespace

extern "C" {

int matrix_export_dense_mtx(const char *filename, const double *data, size_t rows, size_t cols) {
    if (!filename || !data || rows == 0 || cols == 0) return -1;
    FILE *f = std::fopen(filename, "w");
    if (!f) {
        PEEC_ERR("cannot open matrix file '%s'.", filename);
        return -1;
    }
    std::fprintf(f, "%%%%MatrixMarket matrix array real general\n");
    std::fprintf(f, "%% PEEC matrix export\n");
    std::fprintf(f, "%zu %zu\n", rows, cols);
    const int st = write_dense_values(f, data, rows, cols);
    std::fclose(f);
    return st;
}

int matrix_export_incidence_mtx(const char *filename, const Mesh *mesh) {
    if (!filename || !mesh || !mesh->edges || mesh->n_edges == 0 || mesh->n_nodes == 0) return -1;
    FILE *f = std::fopen(filename, "w");
    if (!f) return -1;
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(f, "%% Incidence matrix A: Ne x Nv\n");
    std::fprintf(f, "%zu %zu %zu\n", mesh->n_edges, mesh->n_nodes, 2 * mesh->n_edges);
    for (size_t e = 0; e < mesh->n_edges; ++e) {
        std::fprintf(f, "%zu %d -1.0\n", e + 1, mesh->edges[2 * e] + 1);
        std::fprintf(f, "%zu %d 1.0\n", e + 1, mesh->edges[2 * e + 1] + 1);
    }
    std::fclose(f);
    return 0;
}

int matrix_export_resistance_mtx(const char *filename, const ResistanceMatrix *R) {
    if (!filename || !R || !R->diagonal || R->n == 0) return -1;
    FILE *f = std::fopen(filename, "w");
    if (!f) return -1;
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(f, "%% Diagonal PEEC resistance matrix R\n");
    std::fprintf(f, "%zu %zu %zu\n", R->n, R->n, R->n);
    for (size_t i = 0; i < R->n; ++i) std::fprintf(f, "%zu %zu %.17e\n", i + 1, i + 1, R->diagonal[i]);
    std::fclose(f);
    return 0;
}

int matrix_export_peec_set(const char *dir, const Mesh *mesh, const InductanceMatrix *L, const PotentialMatrix *P,
                           const Matrix *PAT, const ResistanceMatrix *R) {
    if (!dir || !mesh || !L || !P || !PAT || !R) return -1;
    if (ensure_directory(dir) != 0) return -1;
    const std::string d(dir);
    if (matrix_export_dense_mtx((d + "/L.mtx").c_str(), L->data, L->n, L->n) != 0) return -1;
    if (matrix_export_incidence_mtx((d + "/A.mtx").c_str(), mesh) != 0) return -1;
    if (matrix_export_dense_mtx((d + "/P.mtx").c_str(), P->data, P->n, P->n) != 0) return -1;
    if (matrix_export_dense_mtx((d + "/PAT.mtx").c_str(), PAT->data, PAT->rows, PAT->cols) != 0) return -1;
    if (matrix_export_resistance_mtx((d + "/R.mtx").c_str(), R) != 0) return -1;
    return 0;
}

int matrix_import_dense_mtx(const char *filename, double **data, size_t *rows, size_t *cols) {
    if (!filename || !data || !rows || !cols) return -1;
    *data = nullptr;
    FILE *f = std::fopen(filename, "r");
    if (!f) {
        PEEC_ERR("cannot open matrix file '%s'.", filename);
        return -1;
    }
    char line[512];
    bool header = false;
    size_t r = 0, c = 0;
    while (std::fgets(line, sizeof line, f)) {
        if (!header) {
            if (std::strncmp(line, "%%MatrixMarket matrix array real general", 40) != 0) break;
            header = true;
            continue;
        }
        if (line[0] == '%') continue;
        if (std::sscanf(line, "%zu %zu", &r, &c) != 2) r = c = 0;
        break;
    }
    if (!header || r == 0 || c == 0) {
        PEEC_ERR("'%s' is not a dense real Matrix Market file.", filename);
        std::fclose(f);
        return -1;
    }
    double *out = (double *)std::malloc(r * c * sizeof(double));
    if (!out) {
        std::fclose(f);
        return -1;
    }
    for (size_t idx = 0; idx < r * c; ++idx) {
        double v;
        if (std::fscanf(f, "%lf", &v) != 1) {
            PEEC_ERR("'%s': expected %zu values, got %zu.", filename, r * c, idx);
            std::free(out);
            std::fclose(f);
            return -1;
        }
        out[(idx % r) * c + idx / r] = v;
    }
    std::fclose(f);
    *data = out;
    *rows = r;
    *cols = c;
    return 0;
}

}  // extern "C"
