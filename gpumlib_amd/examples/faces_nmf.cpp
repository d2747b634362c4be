// Headless NMF training driver with the semantics of GPUMLib's FacesNMF demo
// (reference src/examples/FacesNMF/FacesNMFDlg.cpp), written against the
// drop-in C++ API (HostMatrix + NMF_Multiplicative*).
//
//   * images: 19 x 19 binary PGM (P5, maxval 255) files face00001.pgm ... as the
//     columns of V, scaled by 1/255 or histogram-equalised, raster read as the reference
//     reads it (LoadImage, :409-472; see load_pgm), --dump-v writes V as loaded;
//     or a seeded synthetic V (--synthetic n m r0 seed);
//   * training: DoIteration(true) until the quality improvement changes and drops
//     below tol, or the iteration cap (Train, :658-689);
//   * test projection: a second NMF on the test columns with W fixed, DoIteration(false)
//     (OnBnClickedTrainTest, :731-756);
//   * groups: m/group independent problems of rank r/(m/group) on consecutive
//     column blocks, assembled into W and a block-diagonal H (TrainUsingGroups, :582-656);
//   * output: W.csv, Htrain.csv, Htest.csv, one line per column, 20 fixed digits (Save, :913-935).
// Initial W and H come from the seeded generator (the demo used rand()).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <fstream>
#include <iomanip>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gpumlib_nmf.h"
#include "../src/algorithms/nmf/nmf_multiplicative_divergence.h"
#include "../src/algorithms/nmf/nmf_multiplicative_euclidian.h"

using namespace GPUMLib;

namespace {

struct Options {
    std::string pgm_dir, out_dir = ".";
    int images = 2429, test_images = 0;
    int syn_n = 0, syn_m = 0, syn_r0 = 0;
    unsigned long long seed = 0x5EED;
    int rank = 49;
    int rule = GPUMLIB_NMF_EUCLIDEAN;
    int max_iterations = 1000;
    double tol = 0.01;
    bool equalize = false;
    bool pgm_standard = false;
    bool load_only = false;
    std::string dump_v;
    int group = 0;
};

// The reference reads the raster right after the maxval token (`imageFile >> maxGrayValue`
// then `read`, FacesNMFDlg.cpp:433-441), so the whitespace byte that ends the header becomes
// pixel 0 and the last raster byte is never read: V is shifted by one byte against the image.
// That is the default here, so V is the reference's V; standard = true skips that one byte as
// the Netpbm format defines (no reference counterpart).
bool load_pgm(const std::string& path, int column, bool equalize, bool standard, HostMatrix<cudafloat>& V) {
    std::ifstream f(path, std::ios::binary);
    std::string magic;
    int w = 0, h = 0, maxv = 0;
    f >> magic >> w >> h >> maxv;
    if (!f || magic != "P5" || w != 19 || h != 19 || maxv != 255) return false;
    if (standard) f.get();  // the single whitespace byte before the raster
    const int pixels = w * h;
    std::vector<unsigned char> buf(pixels);
    f.read(reinterpret_cast<char*>(buf.data()), pixels);
    if (!f) return false;
    if (equalize) {
        float hist[256] = {0};
        for (int b = 0; b < pixels; b++) hist[buf[b]]++;
        for (int i = 1; i < 256; i++) hist[i] += hist[i - 1];
        int i = 0;
        while (hist[i] == 0) i++;
        const float cdfmin = hist[i];
        for (; i < 256; i++) hist[i] = (hist[i] - cdfmin) / (pixels - cdfmin);
        for (int b = 0; b < pixels; b++) V(b, column) = hist[buf[b]];
    } else {
        for (int b = 0; b < pixels; b++) V(b, column) = buf[b] / 255.0f;
    }
    return true;
}

void save_csv(const HostMatrix<cudafloat>& x, const std::string& file) {
    std::ofstream f(file);
    f << std::fixed << std::setprecision(20);
    for (size_t c = 0; c < x.Columns(); c++) {
        for (size_t r = 0; r < x.Rows();) {
            f << x(r, c);
            if (++r < x.Rows()) f << ",";
        }
        f << "\n";
    }
}

HostMatrix<cudafloat> random_matrix(size_t rows, size_t cols, unsigned long long seed, int stream) {
    HostMatrix<cudafloat> x(rows, cols, ColumnMajor);
    gpumlib_nmf_gen_uniform_f32(seed, stream, rows * cols, x.Pointer());
    return x;
}

std::unique_ptr<NMF> make_nmf(int rule, HostMatrix<cudafloat>& v, HostMatrix<cudafloat>& w, HostMatrix<cudafloat>& h) {
    if (rule == GPUMLIB_NMF_KL) return std::unique_ptr<NMF>(new NMF_MultiplicativeDivergence(v, w, h));
    return std::unique_ptr<NMF>(new NMF_MultiplicativeEuclidianDistance(v, w, h));
}

// Train until the quality improvement changes and falls below tol, or the cap.
int train(NMF& nmf, double tol, int cap, bool updateW) {
    cudafloat previous = nmf.QualityImprovement();
    int it = 0;
    while (it < cap) {
        nmf.DoIteration(updateW);
        it++;
        const cudafloat q = nmf.QualityImprovement();
        if (q != previous) {
            if (q < tol) break;
            previous = q;
        }
    }
    return it;
}

HostMatrix<cudafloat> columns(const HostMatrix<cudafloat>& V, size_t c0, size_t c1) {
    HostMatrix<cudafloat> x(V.Rows(), c1 - c0, ColumnMajor);
    for (size_t c = c0; c < c1; c++)
        for (size_t r = 0; r < V.Rows(); r++) x(r, c - c0) = V(r, c);
    return x;
}

int usage() {
    fprintf(stderr,
            "usage: faces_nmf (--pgm DIR [--images N] [--equalize] [--pgm-standard] | --synthetic n m r0 seed)\n"
            "                 [--rank R] [--method euclidean|kl] [--max-iterations N] [--tol T]\n"
            "                 [--groups G] [--test-images T] [--seed S] [--out DIR] [--dump-v FILE] [--load-only]\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* { return (i + 1 < argc) ? argv[++i] : nullptr; };
        const char* v = nullptr;
        if (a == "--pgm" && (v = next())) o.pgm_dir = v;
        else if (a == "--images" && (v = next())) o.images = atoi(v);
        else if (a == "--equalize") o.equalize = true;
        else if (a == "--pgm-standard") o.pgm_standard = true;
        else if (a == "--dump-v" && (v = next())) o.dump_v = v;
        else if (a == "--load-only") o.load_only = true;
        else if (a == "--synthetic" && i + 4 < argc) {
            o.syn_n = atoi(argv[++i]);
            o.syn_m = atoi(argv[++i]);
            o.syn_r0 = atoi(argv[++i]);
            o.seed = strtoull(argv[++i], nullptr, 0);
        } else if (a == "--rank" && (v = next())) o.rank = atoi(v);
        else if (a == "--method" && (v = next())) o.rule = strcmp(v, "kl") == 0 ? GPUMLIB_NMF_KL : GPUMLIB_NMF_EUCLIDEAN;
        else if (a == "--max-iterations" && (v = next())) o.max_iterations = atoi(v);
        else if (a == "--tol" && (v = next())) o.tol = atof(v);
        else if (a == "--groups" && (v = next())) o.group = atoi(v);
        else if (a == "--test-images" && (v = next())) o.test_images = atoi(v);
        else if (a == "--seed" && (v = next())) o.seed = strtoull(v, nullptr, 0);
        else if (a == "--out" && (v = next())) o.out_dir = v;
        else return usage();
    }
    // V: all columns (training first, then test)
    HostMatrix<cudafloat> Vall;
    if (!o.pgm_dir.empty()) {
        Vall.ResizeWithoutPreservingData(19 * 19, o.images);
        Vall = HostMatrix<cudafloat>(19 * 19, o.images, ColumnMajor);
        for (int k = 0; k < o.images; k++) {
            std::ostringstream p;
            p << o.pgm_dir << "/face" << std::setfill('0') << std::setw(5) << (k + 1) << ".pgm";
            if (!load_pgm(p.str(), k, o.equalize, o.pgm_standard, Vall)) {
                fprintf(stderr, "cannot load %s\n", p.str().c_str());
                return 1;
            }
        }
    } else if (o.syn_n > 0) {
        Vall = HostMatrix<cudafloat>(o.syn_n, o.syn_m, ColumnMajor);
        if (o.rule == GPUMLIB_NMF_KL)
            gpumlib_nmf_gen_poisson_v(o.seed, o.syn_n, o.syn_m, o.syn_r0, Vall.Pointer());
        else
            gpumlib_nmf_gen_euclidean_v(o.seed, o.syn_n, o.syn_m, o.syn_r0, 0.01f, Vall.Pointer());
    } else {
        return usage();
    }
    const size_t n = Vall.Rows(), mall = Vall.Columns();
    if (!o.dump_v.empty()) {  // V as loaded: n x columns float32, column-major
        FILE* fv = fopen(o.dump_v.c_str(), "wb");
        if (!fv || fwrite(Vall.Pointer(), sizeof(cudafloat), Vall.Elements(), fv) != Vall.Elements()) return 1;
        fclose(fv);
    }
    if (o.load_only) {  // loading and --dump-v only: nothing touches the GPU
        printf("{\"n\": %zu, \"columns\": %zu}\n", n, mall);
        return 0;
    }
    const size_t mtest = (size_t)o.test_images, mtrain = mall - mtest;
    const int r = o.rank;
    HostMatrix<cudafloat> V = columns(Vall, 0, mtrain);
    auto t0 = std::chrono::steady_clock::now();

    HostMatrix<cudafloat> W, H;
    int iterations = 0;
    if (o.group > 0) {
        const size_t subs = mtrain / o.group;
        const int rr = (int)(r / subs);
        W = HostMatrix<cudafloat>(n, r, ColumnMajor);
        H = HostMatrix<cudafloat>(r, mtrain, ColumnMajor);
        memset(W.Pointer(), 0, sizeof(cudafloat) * W.Elements());
        memset(H.Pointer(), 0, sizeof(cudafloat) * H.Elements());
        for (size_t sm = 0; sm < subs; sm++) {
            HostMatrix<cudafloat> v = columns(V, sm * o.group, (sm + 1) * o.group);
            HostMatrix<cudafloat> w = random_matrix(n, rr, o.seed + sm, 4);
            HostMatrix<cudafloat> h = random_matrix(rr, o.group, o.seed + sm, 5);
            std::unique_ptr<NMF> nmf = make_nmf(o.rule, v, w, h);
            iterations += train(*nmf, o.tol, o.max_iterations, true);
            HostMatrix<cudafloat> ws = nmf->GetW(), hs = nmf->GetH();
            for (size_t i = 0; i < n; i++)
                for (int a = 0; a < rr; a++) W(i, sm * rr + a) = ws(i, a);
            for (int a = 0; a < rr; a++)
                for (int u = 0; u < o.group; u++) H(sm * rr + a, sm * o.group + u) = hs(a, u);
        }
    } else {
        HostMatrix<cudafloat> w = random_matrix(n, r, o.seed, 4);
        HostMatrix<cudafloat> h = random_matrix(r, mtrain, o.seed, 5);
        std::unique_ptr<NMF> nmf = make_nmf(o.rule, V, w, h);
        iterations = train(*nmf, o.tol, o.max_iterations, true);
        W = nmf->GetW();
        H = nmf->GetH();
        if (nmf->LastError() != hipSuccess) {
            fprintf(stderr, "library error %d\n", (int)nmf->LastError());
            return 1;
        }
    }
    int test_iterations = 0;
    HostMatrix<cudafloat> Htest;
    if (mtest > 0) {
        HostMatrix<cudafloat> vt = columns(Vall, mtrain, mall);
        HostMatrix<cudafloat> wt(W);
        HostMatrix<cudafloat> ht = random_matrix(r, mtest, o.seed, 6);
        std::unique_ptr<NMF> nmf = make_nmf(o.rule, vt, wt, ht);
        test_iterations = train(*nmf, o.tol, o.max_iterations, false);
        Htest = nmf->GetH();
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    save_csv(W, o.out_dir + "/W.csv");
    save_csv(H, o.out_dir + "/Htrain.csv");
    if (mtest > 0) save_csv(Htest, o.out_dir + "/Htest.csv");
    printf("{\"n\": %zu, \"m_train\": %zu, \"m_test\": %zu, \"rank\": %d, \"method\": \"%s\", \"groups\": %d, "
           "\"iterations\": %d, \"test_iterations\": %d, \"seconds\": %.6f}\n",
           n, mtrain, mtest, r, o.rule == GPUMLIB_NMF_KL ? "kl" : "euclidean", o.group, iterations, test_iterations,
           secs);
    return 0;
}
