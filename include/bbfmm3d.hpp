/*
 * bbfmm3d.hpp -- umbrella header of the C++ drop-in API (as PBBFMM3D's
 * include/bbfmm3d.hpp), plus the accuracy helpers DirectCalc3D and
 * ComputeError with the signatures and semantics of include/test.hpp:18-47
 * (host FP64, isinf -> 0, relative 2-norm over num_rows x nCols).
 */
#ifndef PBBFMM3D_AMD_BBFMM3D_HPP
#define PBBFMM3D_AMD_BBFMM3D_HPP

#include <cmath>
#include <vector>

#include "bbfmm.h"
#include "compute.hpp"
#include "H2_3D_Tree.hpp"
#include "kernel_Types.hpp"

/* output_exact[k*num_rows + i] = sum_j K(target_i, source_j) w[k*N + j] for
 * the first num_rows targets (test.hpp:18-34). */
template <typename T>
void DirectCalc3D(T *FMMtree, const std::vector<vector3> &target, const std::vector<vector3> &source,
                  std::vector<double> &weight, int nCols, std::vector<double> &output_exact, int num_rows)
{
    const int Ns = (int)source.size();
    const int Nf = (int)target.size();
    for (int k = 0; k < nCols; k++)
        for (int i = 0; i < num_rows; i++) {
            double acc = 0;
            for (int j = 0; j < Ns; j++) {
                double v = FMMtree->EvaluateKernel(target[i], source[j]);
                if (std::isinf(v)) v = 0;
                acc += v * weight[(size_t)k * Nf + j];
            }
            output_exact[(size_t)k * num_rows + i] = acc;
        }
}

/* ||output - outputdir|| / ||outputdir|| over the first num_rows rows of each
 * column (test.hpp:36-47). */
inline double ComputeError(std::vector<double> &output, std::vector<double> &outputdir, int Nf, int nCols,
                           int num_rows)
{
    double diff = 0, sum = 0;
    for (int k = 0; k < nCols; k++)
        for (int i = 0; i < num_rows; i++) {
            double d = output[(size_t)k * Nf + i] - outputdir[(size_t)k * num_rows + i];
            diff += d * d;
            sum += outputdir[(size_t)k * num_rows + i] * outputdir[(size_t)k * num_rows + i];
        }
    return std::sqrt(diff) / std::sqrt(sum);
}

#endif
