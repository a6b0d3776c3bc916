// drot_cli -- command-line experiments over libdrot.so with the file formats,
// argument lists and outputs of the reference's experiment programs:
//
//   drot_cli multiexp-qr  <cmat_dir> <n_rows> <n_cols> <n_tests> <alpha> <q_weight>
//                         <use_warmup_init> <eps> <max_iters> <is_benchmark>
//                                                  (core/multiexp_qr.cu:100-136)
//   drot_cli multiexp-glr <cmat_dir> <out_dir> <n_rows> <n_cols> <n_groups> <n_tests>
//                         <alpha> <r_weight> <use_warmup_init> <eps> <max_iters>
//                         <is_benchmark>           (core/multiexp_glr.cu:120-158)
//   drot_cli singleexp-qr  <cmat_file> <n_rows> <n_cols> <alpha> <q_weight>
//                          <use_warmup_init>       (core/singleexp_qr.cu:60-86)
//   drot_cli singleexp-glr <cmat_file> <n_rows> <n_cols> <n_groups> <rho> <r_weight>
//                          <use_warmup_init>       (core/singleexp_glr.cu:55-82)
//
// Formats (kept identical so existing data directories and result parsers work):
//   cost files   raw fp32, row-major n_rows x n_cols, transposed on load into the
//                solver's column-major layout (core/utils.hpp:8-25); named
//                cmatrix_<n_rows>_<n_cols>_<idx> (QR) or
//                cmatrix_<n_rows>_<n_cols>_<n_groups>_<idx> (GL)
//                (core/multiexp_qr.cu:36-37, core/multiexp_glr.cu:39-40);
//   summaries    mean and std over the tests after dropping the first one
//                (core/utils.hpp:27-47), printed as in core/multiexp_qr.cu:57-80;
//   TSV          the benchmark schema of core/multiexp_qr.cu:85-90 and
//                core/multiexp_glr.cu:93-98;
//   plan dumps   "%.10f " per entry, one line per row (core/utils.hpp:49-62), for
//                GL to <out_dir>/drot-gl_<n_rows>_<n_cols>_<n_groups>_<idx>; the
//                single-experiment programs print X column by column, then the
//                objective and the iteration count (core/singleexp_qr.cu:44-52).
// As in the reference, a wrong argument count prints the expected arguments and
// runs with the default values.  The solves run on the GPU through the host
// wrappers drot_qr_wrapper_f32 / drot_glr_wrapper_f32.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <string>
#include <vector>

#include "drot.h"

namespace {

std::vector<float> load_cost(const std::string &filename, int n_rows, int n_cols) {
    const size_t total = (size_t)n_rows * (size_t)n_cols;
    std::vector<float> values(total, 0.0f), out(total);
    std::ifstream ifs(filename, std::ifstream::binary);
    if (!ifs) fprintf(stderr, "drot_cli: cannot open %s\n", filename.c_str());
    ifs.read(reinterpret_cast<char *>(values.data()), (std::streamsize)(total * sizeof(float)));
    for (int col = 0; col < n_cols; col++)
        for (int row = 0; row < n_rows; row++)
            out[(size_t)col * n_rows + row] = values[(size_t)row * n_cols + col];
    return out;
}

// Mean and standard deviation after dropping the first entry (warm-up).
void mean_std(const std::vector<float> &d, float *mean_out, float *std_out) {
    if (d.empty()) {
        *mean_out = 0.0f;
        *std_out = 0.0f;
        return;
    }
    const int off = std::min((int)d.size() - 1, 1);
    float mean = 0.0f, sd = 0.0f;
    for (size_t i = off; i < d.size(); i++) mean += d[i];
    mean = mean / (float)(d.size() - off);
    for (size_t i = off; i < d.size(); i++) sd += (d[i] - mean) * (d[i] - mean);
    *mean_out = mean;
    *std_out = sqrtf(sd / (float)(d.size() - off));
}

void save_trans(const float *x, const std::string &filename, int n_rows, int n_cols) {
    FILE *f = fopen(filename.c_str(), "w");
    if (!f) {
        fprintf(stderr, "drot_cli: cannot write %s\n", filename.c_str());
        return;
    }
    for (int row = 0; row < n_rows; row++) {
        for (int col = 0; col < n_cols; col++) fprintf(f, "%.10f ", x[(size_t)col * n_rows + row]);
        fprintf(f, "\n");
    }
    fclose(f);
}

void report_failure(const char *what) {
    fprintf(stderr, "drot_cli: %s failed: %s\n", what, drot_last_error());
}

struct Runs {
    std::vector<float> prep, run, iters, objs;
};

void print_summary(const char *weight_label, float r_weight, int n_tests, const Runs &R) {
    printf("%s = %.4f, # Tests = %d\n", weight_label, r_weight, n_tests - 1);
    float m, s;
    mean_std(R.run, &m, &s);
    const float mean_dur = m;
    printf("\tTime (ms)               = %12.8f (std = %.8f)\n", m, s);
    mean_std(R.iters, &m, &s);
    const float mean_it = m;
    printf("\t# Iteration             = %12.8f (std = %.8f)\n", m, s);
    mean_std(R.objs, &m, &s);
    printf("\tObjective               = %12.8f (std = %.8f)\n", m, s);
    printf("\tTime Per Iteration (ms) = %12.8f\n", mean_dur / mean_it);
    mean_std(R.prep, &m, &s);
    printf("\tPreperation Time (ms)   = %12.8f (std = %.8f)\n", m, s);
}

int multiexp_qr(int argc, char **argv) {
    int n_rows = 512, n_cols = 512, n_tests = 1, benchmark = 0, use_warm = 1, max_iters = 20000;
    float step = 2.0f / (float)(n_rows + n_cols), r_weight = 0.0f, eps = 1e-4f;
    std::string dir = "./data/";
    if (argc == 11) {
        dir = argv[1];
        n_rows = atoi(argv[2]);
        n_cols = atoi(argv[3]);
        n_tests = atoi(argv[4]);
        step = (float)atof(argv[5]) / (float)(n_rows + n_cols);
        r_weight = (float)atof(argv[6]);
        use_warm = atoi(argv[7]) > 0;
        eps = (float)atof(argv[8]);
        max_iters = atoi(argv[9]);
        benchmark = atoi(argv[10]);
    } else {
        printf("Expected arguments: <cmat_dir> <n_rows> <n_cols> <n_tests> <alpha> <q_weight> "
               "<use_warmup_init> <eps> <max_iters> <is_benchmark>\n");
        printf("Using default values...\n");
    }
    if (benchmark == 0) {
        printf("n_rows: %d, n_cols: %d, n_tests: %d\n", n_rows, n_cols, n_tests);
        printf("step_size: %.8f, q_weight: %.8f\n", step, r_weight);
    }
    std::vector<float> p(n_rows, 1 / (float)n_rows), q(n_cols, 1 / (float)n_cols);
    Runs R;
    for (int idx = 0; idx < n_tests; idx++) {
        const std::string fn = dir + "/cmatrix_" + std::to_string(n_rows) + "_" +
                               std::to_string(n_cols) + "_" + std::to_string(idx);
        std::vector<float> cost = load_cost(fn, n_rows, n_cols);
        float dur = 0, prep = 0, obj = 0;
        int it = -1;  // stays negative if the wrapper fails (it returns NULL either way)
        drot_qr_wrapper_f32(cost.data(), p.data(), q.data(), n_rows, n_cols, step, r_weight,
                            max_iters, eps, &dur, &prep, &it, &obj, use_warm, 0);
        if (it < 0) {
            report_failure("drot_qr_wrapper_f32");
            it = 0;
        }
        R.prep.push_back(prep);
        R.run.push_back(dur);
        R.iters.push_back((float)it);
        R.objs.push_back(obj);
    }
    if (benchmark == 0) {
        print_summary("R Weight", r_weight, n_tests, R);
    } else {
        printf("test_index\tmethod\tn\tm\teps\tuse_warmup_init\trho\tr_weight\tmax_iters\t"
               "runtime_ms\tobjective\tn_iteration\truntime_ms_per_iteration\n");
        for (int idx = 1; idx < n_tests; idx++)
            printf("%d\tquad_drot\t%d\t%d\t%.8f\t%d\t%.8f\t%.8f\t%d\t%.8f\t%.8f\t%d\t%.8f\n", idx,
                   n_rows, n_cols, eps, use_warm, step * (n_rows + n_cols), r_weight, max_iters,
                   R.run[idx], R.objs[idx], (int)R.iters[idx], R.run[idx] / R.iters[idx]);
    }
    return 0;
}

int multiexp_glr(int argc, char **argv) {
    int n_rows = 2000, n_cols = 2000, n_groups = 2, n_tests = 10, benchmark = 0, use_warm = 1;
    int max_iters = 20000;
    float step = 2.0f / (float)(n_rows + n_cols), r_weight = 0.0f, eps = 1e-4f;
    std::string dir = "./class/", outdir = "./tmp/";
    if (argc == 13) {
        dir = argv[1];
        outdir = argv[2];
        n_rows = atoi(argv[3]);
        n_cols = atoi(argv[4]);
        n_groups = atoi(argv[5]);
        n_tests = atoi(argv[6]);
        step = (float)atof(argv[7]) / (float)(n_rows + n_cols);
        r_weight = (float)atof(argv[8]);
        use_warm = atoi(argv[9]) > 0;
        eps = (float)atof(argv[10]);
        max_iters = atoi(argv[11]);
        benchmark = atoi(argv[12]);
    } else {
        printf("Expected arguments: <cmat_dir> <out_dir> <n_rows> <n_cols> <n_groups> <n_tests> "
               "<alpha> <r_weight> <use_warmup_init> <eps> <max_iters> <is_benchmark>\n");
        printf("Using default values...\n");
    }
    if (benchmark == 0) {
        printf("n_rows: %d, n_cols: %d, n_tests: %d\n", n_rows, n_cols, n_tests);
        printf("step_size: %.8f, r_weight: %.8f\n", step, r_weight);
    }
    std::vector<float> p(n_rows, 1 / (float)n_rows), q(n_cols, 1 / (float)n_cols);
    Runs R;
    const std::string tag = std::to_string(n_rows) + "_" + std::to_string(n_cols) + "_" +
                            std::to_string(n_groups) + "_";
    for (int idx = 0; idx < n_tests; idx++) {
        std::vector<float> cost = load_cost(dir + "/cmatrix_" + tag + std::to_string(idx), n_rows,
                                            n_cols);
        float dur = 0, prep = 0, obj = 0;
        int it = 0;
        float *x = drot_glr_wrapper_f32(cost.data(), p.data(), q.data(), n_rows, n_cols, n_groups,
                                        step, r_weight, max_iters, eps, &dur, &prep, &it, &obj,
                                        use_warm, 1);
        if (x) {
            save_trans(x, outdir + "/drot-gl_" + tag + std::to_string(idx), n_rows, n_cols);
            drot_free(x);
        } else {
            report_failure("drot_glr_wrapper_f32");
        }
        R.prep.push_back(prep);
        R.run.push_back(dur);
        R.iters.push_back((float)it);
        R.objs.push_back(obj);
    }
    if (benchmark == 0) {
        print_summary("R Weight", r_weight, n_tests, R);
    } else {
        printf("test_index\tmethod\tn\tm\tn_class\teps\tuse_warmup_init\trho\tr_weight\tmax_iters\t"
               "runtime_ms\tobjective\tn_iteration\truntime_ms_per_iteration\n");
        for (int idx = 1; idx < n_tests; idx++)
            printf("%d\tgl_drot\t%d\t%d\t%d\t%.8f\t%d\t%.8f\t%.8f\t%d\t%.8f\t%.8f\t%d\t%.8f\n", idx,
                   n_rows, n_cols, n_groups, eps, use_warm, step * (n_rows + n_cols), r_weight,
                   max_iters, R.run[idx], R.objs[idx], (int)R.iters[idx],
                   R.run[idx] / R.iters[idx]);
    }
    return 0;
}

void print_single(const float *x, int n_rows, int n_cols, float obj, int it) {
    for (int j = 0; j < n_cols; j++) {
        for (int i = 0; i < n_rows; i++) printf("%.10f ", x[i + (size_t)j * n_rows]);
        printf("\n");
    }
    printf("%.10f\n", obj);
    printf("%d\n", it);
}

int singleexp_qr(int argc, char **argv) {
    int n_rows = 512, n_cols = 512, use_warm = 1;
    float step = 2.0f / (float)(n_rows + n_cols), r_weight = 0.0f;
    std::string file = "./data/";
    if (argc == 7) {
        file = argv[1];
        n_rows = atoi(argv[2]);
        n_cols = atoi(argv[3]);
        step = (float)atof(argv[4]) / (float)(n_rows + n_cols);
        r_weight = (float)atof(argv[5]);
        use_warm = atoi(argv[6]) > 0;
    } else {
        printf("Expected arguments: <cmat_dir> <n_rows> <n_cols> <alpha> <q_weight>\n");
        printf("Using default values...\n");
    }
    std::vector<float> p(n_rows, 1 / (float)n_rows), q(n_cols, 1 / (float)n_cols);
    std::vector<float> cost = load_cost(file, n_rows, n_cols);
    float dur = 0, prep = 0, obj = 0;
    int it = 0;
    float *x = drot_qr_wrapper_f32(cost.data(), p.data(), q.data(), n_rows, n_cols, step, r_weight,
                                   20000, 1e-4f, &dur, &prep, &it, &obj, use_warm, 1);
    if (!x) {
        report_failure("drot_qr_wrapper_f32");
        return 1;
    }
    print_single(x, n_rows, n_cols, obj, it);
    drot_free(x);
    return 0;
}

int singleexp_glr(int argc, char **argv) {
    int n_rows = 2000, n_cols = 2000, n_groups = 2, use_warm = 1;
    float step = 2.0f / (float)(n_rows + n_cols), r_weight = 0.0f;
    std::string file = "./tmp/cmatrix.tmp";
    if (argc == 8) {
        file = argv[1];
        n_rows = atoi(argv[2]);
        n_cols = atoi(argv[3]);
        n_groups = atoi(argv[4]);
        step = (float)atof(argv[5]) / (float)(n_rows + n_cols);
        r_weight = (float)atof(argv[6]);
        use_warm = atoi(argv[7]) > 0;
    } else {
        printf("Expected arguments: <cmat_dir> <n_rows> <n_cols> <n_groups> <rho> <r_weight> "
               "<use_warmup_init> <is_benchmark>\n");
        printf("Using default values...\n");
    }
    std::vector<float> p(n_rows, 1 / (float)n_rows), q(n_cols, 1 / (float)n_cols);
    std::vector<float> cost = load_cost(file, n_rows, n_cols);
    float dur = 0, prep = 0, obj = 0;
    int it = 0;
    float *x = drot_glr_wrapper_f32(cost.data(), p.data(), q.data(), n_rows, n_cols, n_groups, step,
                                    r_weight, 20000, 1e-4f, &dur, &prep, &it, &obj, use_warm, 1);
    if (!x) {
        report_failure("drot_glr_wrapper_f32");
        return 1;
    }
    print_single(x, n_rows, n_cols, obj, it);
    drot_free(x);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    const char *usage =
        "usage: drot_cli {multiexp-qr|multiexp-glr|singleexp-qr|singleexp-glr} <arguments of the "
        "reference program>\n";
    if (argc < 2) {
        fputs(usage, stderr);
        return 2;
    }
    const std::string cmd = argv[1];
    // the subcommand's own argument vector, with argv[0] = the subcommand
    if (cmd == "multiexp-qr") return multiexp_qr(argc - 1, argv + 1);
    if (cmd == "multiexp-glr") return multiexp_glr(argc - 1, argv + 1);
    if (cmd == "singleexp-qr") return singleexp_qr(argc - 1, argv + 1);
    if (cmd == "singleexp-glr") return singleexp_glr(argc - 1, argv + 1);
    fputs(usage, stderr);
    return 2;
}
