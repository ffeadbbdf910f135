// C++ drop-in check: a PBBFMM3D-style caller, unchanged except for the
// device functor its user kernel names.  Mirrors the reference's
// examples/get_input_through_routine_{standard_kernel,mykernel}.cpp flow:
// build the tree, run H2_3D_Compute, compare with DirectCalc3D.
// Flags as the reference's examples: -n order -N points -L length -m nCols
// -level tree_level -eps eps -cheb use_chebyshev; -cache DIR keeps the
// operators in DIR as the reference's cache files.
// Exit: 0 ok, 1 accuracy failure, 3 no GPU (after a successful precompute).
#include <cstdio>
#include <cstring>
#include <iostream>

#include "bbfmm3d.hpp"

// a user kernel in the style of the reference README (exp(-r), non-homogeneous)
class myKernel : public H2_3D_Tree {
  public:
    myKernel(double L, int tree_level, int order, double eps, int cheb)
        : H2_3D_Tree(L, tree_level, order, eps, cheb) {}
    void SetKernelProperty() override
    {
        homogen = 0;
        symmetry = 1;
        kernelType = "myKernel";
        device_kernel = PBB_KERNEL_EXPONENTIAL;  // c * exp(-r / a)
        device_param[0] = 1.0;
        device_param[1] = 1.0;
    }
    double EvaluateKernel(const vector3 &t, const vector3 &s) override
    {
        double dx = s.x - t.x, dy = s.y - t.y, dz = s.z - t.z;
        return std::exp(-std::sqrt(dx * dx + dy * dy + dz * dz));
    }
};

template <typename T>
static int run(const char *name, T &tree, std::vector<vector3> &pts, std::vector<double> &w, int nCols,
               double tol, double mixed_tol, const char *cache)
{
    if (cache) tree.cache_dir = cache;
    tree.buildFMMTree();
    std::printf("%s: ranks %d/%d\n", name, tree.cutoff.f, tree.cutoff.s);
    if (cache) std::printf("%s: operator cache %s\n", name, tree.cache_hit ? "read" : "written");
    const int N = (int)pts.size();
    std::vector<double> out((size_t)N * nCols, 0.0);
    try {
        H2_3D_Compute<T> compute(tree, pts, pts, w, nCols, out);
        H2_3D_Compute<T> again(tree, pts, pts, w, nCols, out);  // a tree serves many matvecs
    } catch (const std::runtime_error &e) {
        std::printf("%s: %s\n", name, e.what());
        return std::strstr(e.what(), "no HIP device") ? 3 : 2;
    }
    for (double &v : out) v *= 0.5;  // two accumulated matvecs
    int rows = std::min(N, 400);
    std::vector<double> exact((size_t)rows * nCols);
    DirectCalc3D(&tree, pts, pts, w, nCols, exact, rows);
    double err = ComputeError(out, exact, N, nCols, rows);
    std::printf("%s: relative error for first %d rows: %.3e\n", name, rows, err);
    if (err >= tol) return 1;

    // the same matvec in the f16x2 mixed mode, against the FP64 result
    {
        std::vector<double> out_m((size_t)N * nCols, 0.0);
        tree.precision = PBB_PREC_MIXED_F16X2;
        try {
            H2_3D_Compute<T> mixed(tree, pts, pts, w, nCols, out_m);
        } catch (const std::runtime_error &e) {
            std::printf("%s: %s\n", name, e.what());
            return 2;
        }
        tree.precision = PBB_PREC_FP64;
        double d2 = 0, s2 = 0;
        for (size_t i = 0; i < out.size(); i++) {
            d2 += (out_m[i] - out[i]) * (out_m[i] - out[i]);
            s2 += out[i] * out[i];
        }
        const double rel = std::sqrt(d2 / s2);
        std::printf("%s: f16x2 mode against FP64: relative difference %.3e\n", name, rel);
        if (!(rel < mixed_tol)) return 1;
    }

    // separate targets: Nf = N / 3 other points of the same cube, the N
    // points as sources (weights column-major Ns x nCols, outputs Nf x nCols)
    const int Nf = N / 3;
    std::vector<vector3> tgt(Nf);
    const double L = tree.L;
    for (auto &p : tgt) p = vector3(frand(-0.5, 0.5) * L, frand(-0.5, 0.5) * L, frand(-0.5, 0.5) * L);
    std::vector<double> out_t((size_t)Nf * nCols, 0.0);
    try {
        H2_3D_Compute<T> separate(tree, tgt, pts, w, nCols, out_t);
    } catch (const std::runtime_error &e) {
        std::printf("%s: %s\n", name, e.what());
        return 2;
    }
    rows = std::min(Nf, 400);
    double diff = 0, sum = 0;
    for (int k = 0; k < nCols; k++)
        for (int i = 0; i < rows; i++) {
            double acc = 0;
            for (int j = 0; j < N; j++) {
                double v = tree.EvaluateKernel(tgt[i], pts[j]);
                if (std::isinf(v)) v = 0;
                acc += v * w[(size_t)k * N + j];
            }
            const double d = out_t[(size_t)k * Nf + i] - acc;
            diff += d * d;
            sum += acc * acc;
        }
    err = std::sqrt(diff / sum);
    std::printf("%s: separate targets (%d) and sources (%d): relative error for first %d rows: %.3e\n", name, Nf,
                N, rows, err);
    return err < tol ? 0 : 1;
}

int main(int argc, char **argv)
{
    double L = 1.0, eps = 1e-7;
    int order = 5, N = 20000, nCols = 2, level = 4, cheb = 1;
    const char *cache = nullptr;
    for (int i = 1; i + 1 < argc; i++) {
        if (!std::strcmp(argv[i], "-n")) order = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "-N")) N = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "-L")) L = std::atof(argv[++i]);
        else if (!std::strcmp(argv[i], "-m")) nCols = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "-level")) level = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "-eps")) eps = std::atof(argv[++i]);
        else if (!std::strcmp(argv[i], "-cheb")) cheb = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "-cache")) cache = argv[++i];
    }
    std::vector<vector3> pts(N);
    std::vector<double> w((size_t)N * nCols);
    for (auto &v : w) v = frand(0, 1);
    for (auto &p : pts) p = vector3(frand(-0.5, 0.5) * L, frand(-0.5, 0.5) * L, frand(-0.5, 0.5) * L);

    kernel_Laplacian lap(L, level, order, eps, cheb);
    int rc = run("kernel_Laplacian", lap, pts, w, nCols, 1e-5, 1e-6, cache);
    if (rc) return rc;
    myKernel mk(L, level, order, eps, cheb);
    return run("myKernel", mk, pts, w, nCols, 1e-6, 1e-6, cache);
}
