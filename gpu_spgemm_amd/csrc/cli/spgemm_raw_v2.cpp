// spgemm_raw_v2 -- command-line drop-in for the reference's host program
// (src/spgemm_raw_v2.cu:61-372): same flags (--matrix/-m, --aat/-a,
// --device/-d), same stdout lines, same exit codes, with every step run by
// libtsg's HIP kernels (CSR -> tiles and the B re-layout included, which the
// reference does on the host).
//
// Extra flags, all optional: --gen KIND:ARGS (a synthetic A instead of a file:
// er:N:PER_ROW:SEED, lap2d:NX:NY, lap3d:NX:NY:NZ, st27:NX:NY:NZ,
// rmat:SCALE:EF:SEED), --f32 1 (FP32 values), --repeat N (time the product N
// times, report the median), --scheduler plain|ctile.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/tsg.h"

namespace {

struct Args {
    std::string matrix, gen, scheduler = "plain";
    int aat = 0, device = 0, f32 = 0, repeat = 1;
};

void usage(const char *prog) {
    fprintf(stderr,
            "Usage: %s --matrix FILE [--aat 0|1] [--device N]\n"
            "       [--gen KIND:ARGS] [--f32 0|1] [--repeat N] [--scheduler plain|ctile]\n",
            prog);
}

bool parse_int(const char *s, int *out) {
    char *end = nullptr;
    long v = strtol(s, &end, 10);
    if (!s[0] || *end) return false;
    *out = (int)v;
    return true;
}

// argparse semantics of the reference: --matrix is required, integer flags
// must parse; a failure prints the message and the usage and returns 1.
bool parse(int argc, char **argv, Args &a) {
    for (int i = 1; i < argc; i++) {
        const std::string k = argv[i];
        auto val = [&](const char *name) -> const char * {
            if (i + 1 >= argc) {
                fprintf(stderr, "%s: expected 1 argument\n", name);
                return nullptr;
            }
            return argv[++i];
        };
        const char *v = nullptr;
        if (k == "--matrix" || k == "-m") {
            if (!(v = val("--matrix"))) return false;
            a.matrix = v;
        } else if (k == "--aat" || k == "-a" || k == "--device" || k == "-d" || k == "--f32" || k == "--repeat") {
            if (!(v = val(k.c_str()))) return false;
            int *dst = (k == "--aat" || k == "-a") ? &a.aat : (k == "--device" || k == "-d") ? &a.device
                       : k == "--f32" ? &a.f32 : &a.repeat;
            if (!parse_int(v, dst)) {
                fprintf(stderr, "%s: invalid integer '%s'\n", k.c_str(), v);
                return false;
            }
        } else if (k == "--gen") {
            if (!(v = val("--gen"))) return false;
            a.gen = v;
        } else if (k == "--scheduler") {
            if (!(v = val("--scheduler"))) return false;
            a.scheduler = v;
        } else {
            fprintf(stderr, "Unknown argument: %s\n", k.c_str());
            return false;
        }
    }
    if (a.matrix.empty() && a.gen.empty()) {
        fprintf(stderr, "--matrix: required.\n");
        return false;
    }
    if (a.repeat < 1 || (a.scheduler != "plain" && a.scheduler != "ctile")) return false;
    return true;
}

int generate(const std::string &spec, tsg_csr *out) {
    std::vector<long long> v;
    std::string kind = spec.substr(0, spec.find(':'));
    for (size_t p = spec.find(':'); p != std::string::npos; p = spec.find(':', p + 1))
        v.push_back(atoll(spec.c_str() + p + 1));
    auto need = [&](size_t n) { return v.size() == n; };
    if (kind == "er" && need(3)) return tsg_gen_er((int)v[0], (int)v[1], (uint64_t)v[2], out);
    if (kind == "lap2d" && need(2)) return tsg_gen_laplace2d((int)v[0], (int)v[1], out);
    if (kind == "lap3d" && need(3)) return tsg_gen_laplace3d((int)v[0], (int)v[1], (int)v[2], out);
    if (kind == "st27" && need(3)) return tsg_gen_stencil27((int)v[0], (int)v[1], (int)v[2], out);
    if (kind == "rmat" && need(3))
        return tsg_gen_rmat((int)v[0], (int)v[1], 0.57, 0.19, 0.19, (uint64_t)v[2], TSG_F64, out);
    fprintf(stderr, "bad --gen spec '%s'\n", spec.c_str());
    return TSG_ERR_ARG;
}

int die(const char *what) {
    fprintf(stderr, "%s failed: %s\n", what, tsg_last_error_string());
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    Args a;
    if (!parse(argc, argv, a)) {
        usage(argv[0]);
        return 1;
    }
    tsg_context *ctx = nullptr;
    if (tsg_create(a.device, &ctx) != TSG_OK) return die("tsg_create");
    char name[256] = {0};
    tsg_device_name(ctx, name, sizeof name);
    printf("Device [%d] %s\n", a.device, name);

    tsg_csr hA{}, hB{};
    if (!a.gen.empty()) {
        printf("Generating %s\n", a.gen.c_str());
        if (generate(a.gen, &hA) != TSG_OK) return die("generator");
    } else {
        printf("Loading %s\n", a.matrix.c_str());
        if (tsg_load_mtx(a.matrix.c_str(), &hA) != TSG_OK) {
            // load_mtx prints and exits with 1 when the file cannot be opened
            printf("Cannot open %s\n", a.matrix.c_str());
            return 1;
        }
    }
    printf("A: %d x %d, nnz = %lld\n", hA.m, hA.n, (long long)hA.nnz);
    if (a.aat) {
        if (tsg_csr_transpose_host(&hA, &hB) != TSG_OK) return die("transpose");
    } else if (hA.m != hA.n) {
        printf("A^2 requires square\n");
        return 1;
    }
    if (a.f32) {
        if (tsg_csr_convert_dtype_host(&hA, TSG_F32) != TSG_OK) return die("dtype");
        if (a.aat && tsg_csr_convert_dtype_host(&hB, TSG_F32) != TSG_OK) return die("dtype");
    }
    const tsg_csr &hBr = a.aat ? hB : hA;
    uint64_t flops = 0;
    if (tsg_csr_products(&hA, &hBr, &flops) != TSG_OK) return die("products");
    printf("SpGEMM flops = %llu\n", (unsigned long long)flops);

    // CSR -> tiles on the GPU (the reference converts on the host, untimed)
    tsg_csr dA{}, dB{};
    tsg_tile_matrix tA{}, tB{};
    if (tsg_csr_upload(ctx, &hA, &dA) != TSG_OK || tsg_csr2tile(ctx, &dA, &tA) != TSG_OK) return die("A conversion");
    if (tsg_csr_upload(ctx, &hBr, &dB) != TSG_OK || tsg_csr2tile(ctx, &dB, &tB) != TSG_OK ||
        tsg_build_csc_tiles(ctx, &tB) != TSG_OK)
        return die("B conversion");
    tsg_csr_free(ctx, &dA);
    tsg_csr_free(ctx, &dB);

    tsg_options o{};
    o.scheduler = TSG_SCHED_PLAIN;
    o.numeric_variant = a.scheduler == "ctile" ? 1 : 0;
    o.timing = 1;
    std::vector<tsg_stats> runs;
    tsg_tile_matrix C{};
    for (int r = 0; r < a.repeat; r++) {
        if (r) tsg_tile_free(ctx, &C);
        tsg_stats st{};
        if (tsg_spgemm(ctx, &tA, &tB, &C, &o, &st) != TSG_OK) {
            printf("HIP error: %s\n", tsg_last_error_string());
            return 1;
        }
        runs.push_back(st);
    }
    std::sort(runs.begin(), runs.end(), [](const tsg_stats &x, const tsg_stats &y) { return x.ms_total < y.ms_total; });
    const tsg_stats &st = runs[runs.size() / 2];
    const double ms = st.ms_step1 + st.ms_step2 + st.ms_step3;
    printf("\nRaw Tile SpGEMM v2: tiles_C = %lld, nnzC = %lld\n", (long long)st.numtileC, (long long)st.nnzC);
    printf("  Step1 (tile structure) = %.2f ms\n", st.ms_step1);
    printf("  Step2 (symbolic+scan)  = %.2f ms\n", st.ms_step2);
    printf("  Step3 (numeric v2)     = %.2f ms\n", st.ms_step3);
    printf("GPU time = %.2f ms, GFlops = %.2f\n", ms, 2.0 * (double)flops / (ms * 1e6));

    tsg_tile_free(ctx, &C);
    tsg_tile_free(ctx, &tA);
    tsg_tile_free(ctx, &tB);
    tsg_csr_free_host(&hA);
    if (a.aat) tsg_csr_free_host(&hB);
    tsg_destroy(ctx);
    printf("Done.\n");
    return 0;
}
