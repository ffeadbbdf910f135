/*
 * compute.hpp -- the matvec of the C++ drop-in API.
 *
 * template<class T> H2_3D_Compute(T& tree, target, source, weight, nCols,
 * output) with the constructor signature of PBBFMM3D's include/compute.hpp:
 * 20 and its contract (compute.hpp:67-149): all work happens in the
 * constructor; weights are column-major weight[c*Ns + j]; the result is
 * ADDED to output[c*Nf + i] (compute.hpp:747).  Sources equal to the
 * targets run as one point set; other sources, of any number, as separate
 * sets in one tree centred on the targets' bbox (pbb_matvec_prec).  The
 * arithmetic is the tree's `precision` (FP64 unless the caller sets a mixed
 * mode).  The work runs on the current GPU.  Errors throw
 * std::runtime_error instead of exiting.
 */
#ifndef PBBFMM3D_AMD_COMPUTE_HPP
#define PBBFMM3D_AMD_COMPUTE_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "H2_3D_Tree.hpp"
#include "kernel_Types.hpp"

template <typename T>
class H2_3D_Compute {
  public:
    H2_3D_Compute(T &FMMTree, const std::vector<vector3> &target, const std::vector<vector3> &source,
                  std::vector<double> &weight, int nCols, std::vector<double> &output)
        : tree(&FMMTree), Ns((int)source.size()), Nf((int)target.size()), nCols(nCols)
    {
        if (nCols < 1) throw std::runtime_error("H2_3D_Compute: nCols must be >= 1");
        if (Ns < 1 || Nf < 1) throw std::runtime_error("H2_3D_Compute: no targets or no sources");
        if (weight.size() < (size_t)Ns * nCols || output.size() < (size_t)Nf * nCols)
            throw std::runtime_error("H2_3D_Compute: weight must hold Ns * nCols values, output Nf * nCols");
        static_assert(sizeof(vector3) == 3 * sizeof(double), "vector3 must be three packed doubles");
        pbb_status st = pbb_matvec_prec(FMMTree.ops(), &target[0].x, (int64_t)Nf, &source[0].x, (int64_t)Ns,
                                        weight.data(), nCols, output.data(), FMMTree.precision);
        if (st != PBB_OK) throw std::runtime_error(std::string("H2_3D_Compute: ") + pbb_last_error());
    }
    T *tree;
    int Ns;
    int Nf;
    int nCols;
};

#endif
