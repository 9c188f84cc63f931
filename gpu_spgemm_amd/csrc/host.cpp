// host.cpp -- host-side matrices for libtsg: synthetic generators, MTX I/O,
// transpose and product counts.  Plain C++ (compiled without the GPU
// toolchain); all output arrays are malloc'd and released by tsg_csr_free_host.
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "tsg.h"

namespace tsg {
int fail(int code, const char *fmt, ...);
}
using tsg::fail;

namespace {

// Counter-based SplitMix64: value i of stream s under seed is
// mix(mix(seed ^ s*K) + i), so generation order and thread count do not
// change any value.
inline uint64_t mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix(mix(seed ^ (stream * 0xD1B54A32D192ED03ull)) + i);
}
inline double u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }  // [0,1)

int alloc_csr(tsg_csr *o, int32_t m, int32_t n, int64_t nnz, int32_t dtype) {
    memset(o, 0, sizeof *o);
    o->m = m; o->n = n; o->nnz = nnz; o->dtype = dtype; o->location = TSG_HOST;
    o->rowptr = (int64_t *)calloc((size_t)m + 1, sizeof(int64_t));
    o->col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    o->val = malloc((dtype == TSG_F32 ? 4 : 8) * (size_t)(nnz ? nnz : 1));
    if (!o->rowptr || !o->col || !o->val) {
        free(o->rowptr); free(o->col); free(o->val);
        memset(o, 0, sizeof *o);
        return fail(TSG_ERR_OOM, "host allocation failed");
    }
    return TSG_OK;
}

inline void setv(tsg_csr *o, int64_t i, double v) {
    if (o->dtype == TSG_F32) ((float *)o->val)[i] = (float)v;
    else ((double *)o->val)[i] = v;
}
inline double getv(const tsg_csr *o, int64_t i) {
    return o->dtype == TSG_F32 ? (double)((const float *)o->val)[i] : ((const double *)o->val)[i];
}

template <typename F>
void parallel_for(int64_t n, F f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 100000) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([=] { f(n * t / nt, n * (t + 1) / nt); });
    for (auto &x : th) x.join();
}

// Stencil generator over an nx*ny*nz grid (row r = x + nx*(y + ny*z)); the
// offsets list is in ascending column order.
int stencil(int32_t nx, int32_t ny, int32_t nz, const std::vector<std::array<int, 3>> &offs, double diag,
            double off, tsg_csr *o) {
    int64_t n = (int64_t)nx * ny * nz;
    if (nx <= 0 || ny <= 0 || nz <= 0 || n > INT32_MAX) return fail(TSG_ERR_ARG, "bad grid %dx%dx%d", nx, ny, nz);
    std::vector<int64_t> cnt((size_t)n + 1, 0);
    auto inside = [&](int64_t x, int64_t y, int64_t z) { return x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz; };
    parallel_for(n, [&](int64_t a, int64_t b) {
        for (int64_t r = a; r < b; r++) {
            int64_t x = r % nx, y = (r / nx) % ny, z = r / ((int64_t)nx * ny);
            int64_t c = 0;
            for (auto &d : offs) c += inside(x + d[0], y + d[1], z + d[2]);
            cnt[r + 1] = c;
        }
    });
    for (int64_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    if (int rc = alloc_csr(o, (int32_t)n, (int32_t)n, cnt[n], TSG_F64)) return rc;
    memcpy(o->rowptr, cnt.data(), sizeof(int64_t) * ((size_t)n + 1));
    parallel_for(n, [&](int64_t a, int64_t b) {
        for (int64_t r = a; r < b; r++) {
            int64_t x = r % nx, y = (r / nx) % ny, z = r / ((int64_t)nx * ny), p = o->rowptr[r];
            for (auto &d : offs)
                if (inside(x + d[0], y + d[1], z + d[2])) {
                    int64_t c = (x + d[0]) + nx * ((y + d[1]) + (int64_t)ny * (z + d[2]));
                    o->col[p] = (int32_t)c;
                    ((double *)o->val)[p] = (c == r) ? diag : off;
                    p++;
                }
        }
    });
    return TSG_OK;
}

}  // namespace

extern "C" {

int tsg_csr_free_host(tsg_csr *h) {
    if (!h) return TSG_OK;
    if (h->location != TSG_HOST) return fail(TSG_ERR_ARG, "CSR is not on the host");
    free(h->rowptr); free(h->col); free(h->val);
    memset(h, 0, sizeof *h);
    return TSG_OK;
}

int tsg_csr_convert_dtype_host(tsg_csr *h, int32_t dtype) {
    if (!h || h->location != TSG_HOST) return fail(TSG_ERR_ARG, "host CSR required");
    if (dtype == h->dtype) return TSG_OK;
    void *nv = malloc((dtype == TSG_F32 ? 4 : 8) * (size_t)(h->nnz ? h->nnz : 1));
    if (!nv) return fail(TSG_ERR_OOM, "host allocation failed");
    for (int64_t i = 0; i < h->nnz; i++) {
        if (dtype == TSG_F32) ((float *)nv)[i] = (float)((double *)h->val)[i];
        else ((double *)nv)[i] = (double)((float *)h->val)[i];
    }
    free(h->val);
    h->val = nv;
    h->dtype = dtype;
    return TSG_OK;
}

int tsg_gen_er(int32_t n, int32_t per_row, uint64_t seed, tsg_csr *o) {
    if (n <= 0 || per_row < 0 || per_row > n) return fail(TSG_ERR_ARG, "bad ER parameters n=%d per_row=%d", n, per_row);
    if (int rc = alloc_csr(o, n, n, (int64_t)n * per_row, TSG_F64)) return rc;
    for (int64_t i = 0; i <= n; i++) o->rowptr[i] = i * per_row;
    parallel_for(n, [&](int64_t a, int64_t b) {
        std::vector<int32_t> cols;
        for (int64_t r = a; r < b; r++) {
            cols.clear();
            uint64_t k = 0;
            while ((int32_t)cols.size() < per_row) {   // distinct uniform columns by rejection
                int32_t c = (int32_t)(u01(rnd(seed, 1, (uint64_t)r * 4096 + k++)) * n);
                if (std::find(cols.begin(), cols.end(), c) == cols.end()) cols.push_back(c);
            }
            std::sort(cols.begin(), cols.end());
            for (int q = 0; q < per_row; q++) {
                o->col[r * per_row + q] = cols[q];
                ((double *)o->val)[r * per_row + q] = 2.0 * u01(rnd(seed, 2, (uint64_t)r * per_row + q)) - 1.0;
            }
        }
    });
    return TSG_OK;
}

int tsg_gen_er_var(int32_t n, int32_t lo, int32_t hi, uint64_t seed, tsg_csr *o) {
    if (n <= 0 || lo < 0 || hi < lo || hi > n) return fail(TSG_ERR_ARG, "bad ER parameters");
    std::vector<int64_t> rp((size_t)n + 1, 0);
    for (int64_t r = 0; r < n; r++) rp[r + 1] = rp[r] + lo + (int64_t)(u01(rnd(seed, 3, r)) * (hi - lo + 1));
    if (int rc = alloc_csr(o, n, n, rp[n], TSG_F64)) return rc;
    memcpy(o->rowptr, rp.data(), sizeof(int64_t) * ((size_t)n + 1));
    parallel_for(n, [&](int64_t a, int64_t b) {
        std::vector<int32_t> cols;
        for (int64_t r = a; r < b; r++) {
            int32_t k = (int32_t)(rp[r + 1] - rp[r]);
            cols.clear();
            uint64_t t = 0;
            while ((int32_t)cols.size() < k) {
                int32_t c = (int32_t)(u01(rnd(seed, 1, (uint64_t)r * 4096 + t++)) * n);
                if (std::find(cols.begin(), cols.end(), c) == cols.end()) cols.push_back(c);
            }
            std::sort(cols.begin(), cols.end());
            for (int q = 0; q < k; q++) {
                o->col[rp[r] + q] = cols[q];
                ((double *)o->val)[rp[r] + q] = 2.0 * u01(rnd(seed, 2, (uint64_t)rp[r] + q)) - 1.0;
            }
        }
    });
    return TSG_OK;
}

int tsg_gen_laplace2d(int32_t nx, int32_t ny, tsg_csr *o) {
    return stencil(nx, ny, 1, {{0, -1, 0}, {-1, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}}, 4.0, -1.0, o);
}

int tsg_gen_laplace3d(int32_t nx, int32_t ny, int32_t nz, tsg_csr *o) {
    return stencil(nx, ny, nz, {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
                   6.0, -1.0, o);
}

int tsg_gen_stencil27(int32_t nx, int32_t ny, int32_t nz, tsg_csr *o) {
    std::vector<std::array<int, 3>> offs;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) offs.push_back({dx, dy, dz});
    return stencil(nx, ny, nz, offs, 26.0, -1.0, o);
}

int tsg_gen_aggregation(int32_t nx, int32_t ny, int32_t nz, tsg_csr *o) {
    int64_t n = (int64_t)nx * ny * nz;
    if (nx <= 0 || ny <= 0 || nz <= 0 || n > INT32_MAX) return fail(TSG_ERR_ARG, "bad grid");
    int32_t cx = (nx + 1) / 2, cy = (ny + 1) / 2, cz = (nz + 1) / 2;
    if (int rc = alloc_csr(o, (int32_t)n, cx * cy * cz, n, TSG_F64)) return rc;
    for (int64_t r = 0; r < n; r++) {
        int64_t x = r % nx, y = (r / nx) % ny, z = r / ((int64_t)nx * ny);
        o->rowptr[r + 1] = r + 1;
        o->col[r] = (int32_t)((x / 2) + cx * ((y / 2) + (int64_t)cy * (z / 2)));
        ((double *)o->val)[r] = 1.0;
    }
    return TSG_OK;
}

int tsg_gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed, int32_t dtype,
                 tsg_csr *o) {
    if (scale <= 0 || scale > 30 || ef <= 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return fail(TSG_ERR_ARG, "bad R-MAT parameters");
    const int64_t n = (int64_t)1 << scale, ne = n * ef;
    std::vector<uint32_t> er((size_t)ne), ec((size_t)ne);
    parallel_for(ne, [&](int64_t s, int64_t e) {
        for (int64_t i = s; i < e; i++) {
            uint32_t r = 0, cc = 0;
            for (int l = 0; l < scale; l++) {     // one quadrant choice per level, top bit first
                double u = u01(rnd(seed, 5, (uint64_t)i * 64 + l));
                uint32_t br = 0, bc = 0;
                if (u < a) {} else if (u < a + b) bc = 1; else if (u < a + b + c) br = 1; else { br = 1; bc = 1; }
                r = (r << 1) | br;
                cc = (cc << 1) | bc;
            }
            er[i] = r; ec[i] = cc;
        }
    });
    // counting sort by row, then sort + merge duplicates inside each row
    std::vector<int64_t> rp((size_t)n + 1, 0);
    for (int64_t i = 0; i < ne; i++) rp[er[i] + 1]++;
    for (int64_t i = 0; i < n; i++) rp[i + 1] += rp[i];
    std::vector<uint32_t> cols((size_t)ne);
    {
        std::vector<int64_t> cur(rp.begin(), rp.end() - 1);
        for (int64_t i = 0; i < ne; i++) cols[cur[er[i]]++] = ec[i];
    }
    std::vector<uint32_t>().swap(er);
    std::vector<uint32_t>().swap(ec);
    std::vector<int64_t> uniq((size_t)n + 1, 0);
    parallel_for(n, [&](int64_t s, int64_t e) {
        for (int64_t r = s; r < e; r++) {
            auto b0 = cols.begin() + rp[r], b1 = cols.begin() + rp[r + 1];
            std::sort(b0, b1);
            uniq[r + 1] = std::unique(b0, b1) - b0;
        }
    });
    for (int64_t i = 0; i < n; i++) uniq[i + 1] += uniq[i];
    if (int rc = alloc_csr(o, (int32_t)n, (int32_t)n, uniq[n], dtype)) return rc;
    memcpy(o->rowptr, uniq.data(), sizeof(int64_t) * ((size_t)n + 1));
    parallel_for(n, [&](int64_t s, int64_t e) {
        for (int64_t r = s; r < e; r++) {
            int64_t k = uniq[r + 1] - uniq[r];
            for (int64_t q = 0; q < k; q++) {
                int64_t p = uniq[r] + q;
                o->col[p] = (int32_t)cols[rp[r] + q];
                // U(0,1]: 1 - [0,1)
                double v = 1.0 - u01(rnd(seed, 6, (uint64_t)p));
                if (dtype == TSG_F32) {
                    float f = (float)v;
                    ((float *)o->val)[p] = f > 0.0f ? f : 0x1.0p-24f;
                } else {
                    ((double *)o->val)[p] = v;
                }
            }
        }
    });
    return TSG_OK;
}

// ---------------------------------------------------------------- MTX
// Semantics of load_mtx (src/tile_format.h:51-90): flags from the banner line,
// '%' lines skipped, 1-based indices, symmetric off-diagonals mirrored,
// pattern -> 1.0, complex -> real part, sorted by (row,col); duplicates kept
// (ties in file order).
int tsg_load_mtx(const char *path, tsg_csr *o) {
    FILE *f = fopen(path, "r");
    if (!f) return fail(TSG_ERR_IO, "cannot open %s", path);
    char line[1024];
    bool sym = false, pat = false, cpx = false;
    if (fgets(line, sizeof line, f)) {
        sym = strstr(line, "symmetric") != nullptr;
        pat = strstr(line, "pattern") != nullptr;
        cpx = strstr(line, "complex") != nullptr;
    }
    bool got_size = false;
    while (fgets(line, sizeof line, f)) if (line[0] != '%') { got_size = true; break; }
    long long m = 0, n = 0, nf = 0;
    if (!got_size || sscanf(line, "%lld %lld %lld", &m, &n, &nf) != 3 || m < 0 || n < 0 || nf < 0 ||
        m > INT32_MAX || n > INT32_MAX) {
        fclose(f);
        return fail(TSG_ERR_IO, "%s: bad size line", path);
    }
    struct E { int32_t r, c; int64_t k; double v; };
    std::vector<E> e;
    e.reserve((size_t)nf * (sym ? 2 : 1));
    for (long long i = 0; i < nf; i++) {
        long long r, c;
        double v = 1.0, vi = 0.0;
        int ok;
        if (pat) ok = fscanf(f, "%lld %lld", &r, &c) == 2;
        else if (cpx) ok = fscanf(f, "%lld %lld %lf %lf", &r, &c, &v, &vi) == 4;
        else ok = fscanf(f, "%lld %lld %lf", &r, &c, &v) == 3;
        if (!ok || r < 1 || c < 1 || r > m || c > n) {
            fclose(f);
            return fail(TSG_ERR_IO, "%s: bad entry %lld", path, i + 1);
        }
        e.push_back({(int32_t)(r - 1), (int32_t)(c - 1), (int64_t)e.size(), v});
        if (sym && r != c) e.push_back({(int32_t)(c - 1), (int32_t)(r - 1), (int64_t)e.size(), v});
    }
    fclose(f);
    std::sort(e.begin(), e.end(), [](const E &x, const E &y) {
        return x.r != y.r ? x.r < y.r : x.c != y.c ? x.c < y.c : x.k < y.k;
    });
    if (int rc = alloc_csr(o, (int32_t)m, (int32_t)n, (int64_t)e.size(), TSG_F64)) return rc;
    for (size_t i = 0; i < e.size(); i++) {
        o->col[i] = e[i].c;
        ((double *)o->val)[i] = e[i].v;
        o->rowptr[e[i].r + 1]++;
    }
    for (int64_t i = 0; i < m; i++) o->rowptr[i + 1] += o->rowptr[i];
    return TSG_OK;
}

int tsg_write_mtx(const char *path, const tsg_csr *h) {
    if (!h || h->location != TSG_HOST) return fail(TSG_ERR_ARG, "host CSR required");
    FILE *f = fopen(path, "w");
    if (!f) return fail(TSG_ERR_IO, "cannot open %s for writing", path);
    fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%d %d %lld\n", h->m, h->n, (long long)h->nnz);
    for (int32_t r = 0; r < h->m; r++)
        for (int64_t j = h->rowptr[r]; j < h->rowptr[r + 1]; j++)
            fprintf(f, "%d %d %.17g\n", r + 1, h->col[j] + 1, getv(h, j));
    int rc = fclose(f);
    return rc ? fail(TSG_ERR_IO, "write failed") : TSG_OK;
}

// B = A^T by counting sort (src/spgemm_raw_v2.cu:78-89): rows of A^T keep
// ascending column order.
int tsg_csr_transpose_host(const tsg_csr *in, tsg_csr *o) {
    if (!in || in->location != TSG_HOST) return fail(TSG_ERR_ARG, "host CSR required");
    if (int rc = alloc_csr(o, in->n, in->m, in->nnz, in->dtype)) return rc;
    for (int64_t i = 0; i < in->nnz; i++) o->rowptr[in->col[i] + 1]++;
    for (int32_t i = 0; i < in->n; i++) o->rowptr[i + 1] += o->rowptr[i];
    std::vector<int64_t> off(o->rowptr, o->rowptr + in->n + 1);
    for (int32_t r = 0; r < in->m; r++)
        for (int64_t j = in->rowptr[r]; j < in->rowptr[r + 1]; j++) {
            int64_t p = off[in->col[j]]++;
            o->col[p] = r;
            setv(o, p, getv(in, j));
        }
    return TSG_OK;
}

int tsg_csr_products(const tsg_csr *A, const tsg_csr *B, uint64_t *products) {
    if (!A || !B || !products || A->location != TSG_HOST || B->location != TSG_HOST)
        return fail(TSG_ERR_ARG, "host CSRs required");
    if (A->n != B->m) return fail(TSG_ERR_ARG, "inner dimensions differ");
    uint64_t f = 0;
    for (int64_t j = 0; j < A->nnz; j++) f += (uint64_t)(B->rowptr[A->col[j] + 1] - B->rowptr[A->col[j]]);
    *products = f;
    return TSG_OK;
}

}  // extern "C"
