/*
 * kernel_Types.hpp -- built-in kernels of the C++ drop-in API.
 *
 * The scalar kernels of PBBFMM3D's include/kernel_Types.hpp:10-90 with the
 * properties and formulas of src/kernel_Types.cpp, each bound to its device
 * functor.  kernel_OneOverR4 (dof 9/6 in the reference, kernel_Types.cpp:
 * 34-40, which the reference's engine cannot run) and the declared-only
 * kernel_LaplacianForce / kernel_Stokes are not provided.  Parameterised
 * variants (kernel_GaussianScaled, kernel_Exponential, kernel_SoftLaplacian)
 * are the user kernels that the reference's README describes writing.
 */
#ifndef PBBFMM3D_AMD_KERNEL_TYPES_HPP
#define PBBFMM3D_AMD_KERNEL_TYPES_HPP

#include <cmath>

#include "H2_3D_Tree.hpp"

namespace pbb_detail {
inline double dist2(const vector3 &t, const vector3 &s)
{
    double dx = s.x - t.x, dy = s.y - t.y, dz = s.z - t.z;
    return dx * dx + dy * dy + dz * dz;
}
}  // namespace pbb_detail

#define PBB_KERNEL_CLASS(NAME)                                                             \
    NAME(double L, int tree_level, int interpolation_order, double epsilon, int use_chebyshev) \
        : H2_3D_Tree(L, tree_level, interpolation_order, epsilon, use_chebyshev) {}

/* 1/r (kernel_Types.cpp:7-28) */
class kernel_Laplacian : public H2_3D_Tree {
  public:
    PBB_KERNEL_CLASS(kernel_Laplacian)
    void SetKernelProperty() override
    {
        homogen = -1; symmetry = 1; kernelType = "Laplacian";
        device_kernel = PBB_KERNEL_LAPLACIAN;
    }
    double EvaluateKernel(const vector3 &t, const vector3 &s) override
    {
        return 1. / std::sqrt(pbb_detail::dist2(t, s));
    }
};

/* exp(-r^2) (kernel_Types.cpp:56-74) */
class kernel_Gaussian : public H2_3D_Tree {
  public:
    PBB_KERNEL_CLASS(kernel_Gaussian)
    void SetKernelProperty() override
    {
        homogen = 0; symmetry = 1; kernelType = "Gaussian";
        device_kernel = PBB_KERNEL_GAUSSIAN; device_param[0] = 1.0;
    }
    double EvaluateKernel(const vector3 &t, const vector3 &s) override
    {
        return std::exp(-pbb_detail::dist2(t, s));
    }
};

/* 0.5 log r^2, 0 at r = 0 (kernel_Types.cpp:77-99) */
class kernel_Logarithm : public H2_3D_Tree {
  public:
    PBB_KERNEL_CLASS(kernel_Logarithm)
    void SetKernelProperty() override
    {
        homogen = 0; symmetry = 1; kernelType = "Logarithm";
        device_kernel = PBB_KERNEL_LOGARITHM;
    }
    double EvaluateKernel(const vector3 &t, const vector3 &s) override
    {
        double r2 = pbb_detail::dist2(t, s);
        return r2 == 0 ? 0 : 0.5 * std::log(r2);
    }
};

/* 1/r^2 (kernel_Types.cpp:101-121) */
class kernel_OneOverR2 : public H2_3D_Tree {
  public:
    PBB_KERNEL_CLASS(kernel_OneOverR2)
    void SetKernelProperty() override
    {
        homogen = -2; symmetry = 1; kernelType = "OneOverR2";
        device_kernel = PBB_KERNEL_ONEOVERR2;
    }
    double EvaluateKernel(const vector3 &t, const vector3 &s) override
    {
        double ri = 1. / std::sqrt(pbb_detail::dist2(t, s));
        return ri * ri;
    }
};

/* 1 + r^2 (kernel_Types.cpp:123-138) */
class kernel_Quadric : public H2_3D_Tree {
  public:
    PBB_KERNEL_CLASS(kernel_Quadric)
    void SetKernelProperty() override
    {
        homogen = 0; symmetry = 1; kernelType = "Quadric";
        device_kernel = PBB_KERNEL_QUADRIC;
    }
    double EvaluateKernel(const vector3 &t, const vector3 &s) override
    {
        return 1.0 + pbb_detail::dist2(t, s);
    }
};

/* 1 / (1 + r^2) (kernel_Types.cpp:140-155) */
class kernel_InverseQuadric : public H2_3D_Tree {
  public:
    PBB_KERNEL_CLASS(kernel_InverseQuadric)
    void SetKernelProperty() override
    {
        homogen = 0; symmetry = 1; kernelType = "InverseQuadric";
        device_kernel = PBB_KERNEL_INV_QUADRIC;
    }
    double EvaluateKernel(const vector3 &t, const vector3 &s) override
    {
        return 1.0 / (1.0 + pbb_detail::dist2(t, s));
    }
};

/* 0.5 r^2 log r^2, 0 at r = 0 (kernel_Types.cpp:157-178) */
class kernel_ThinPlateSpline : public H2_3D_Tree {
  public:
    PBB_KERNEL_CLASS(kernel_ThinPlateSpline)
    void SetKernelProperty() override
    {
        homogen = 0; symmetry = 1; kernelType = "ThinPlateSpline";
        device_kernel = PBB_KERNEL_THIN_PLATE;
    }
    double EvaluateKernel(const vector3 &t, const vector3 &s) override
    {
        double r2 = pbb_detail::dist2(t, s);
        return r2 == 0 ? 0 : 0.5 * r2 * std::log(r2);
    }
};

/* exp(-r^2 / a^2) */
class kernel_GaussianScaled : public H2_3D_Tree {
  public:
    kernel_GaussianScaled(double L, int tree_level, int interpolation_order, double epsilon,
                          int use_chebyshev, double a_)
        : H2_3D_Tree(L, tree_level, interpolation_order, epsilon, use_chebyshev), a(a_) {}
    void SetKernelProperty() override
    {
        homogen = 0; symmetry = 1; kernelType = "GaussianScaled";
        device_kernel = PBB_KERNEL_GAUSSIAN; device_param[0] = a;
    }
    double EvaluateKernel(const vector3 &t, const vector3 &s) override
    {
        return std::exp(-pbb_detail::dist2(t, s) / (a * a));
    }
    double a;
};

/* c * exp(-r / a) (exponential covariance) */
class kernel_Exponential : public H2_3D_Tree {
  public:
    kernel_Exponential(double L, int tree_level, int interpolation_order, double epsilon,
                       int use_chebyshev, double a_, double c_ = 1.0)
        : H2_3D_Tree(L, tree_level, interpolation_order, epsilon, use_chebyshev), a(a_), c(c_) {}
    void SetKernelProperty() override
    {
        homogen = 0; symmetry = 1; kernelType = "Exponential";
        device_kernel = PBB_KERNEL_EXPONENTIAL; device_param[0] = a; device_param[1] = c;
    }
    double EvaluateKernel(const vector3 &t, const vector3 &s) override
    {
        return c * std::exp(-std::sqrt(pbb_detail::dist2(t, s)) / a);
    }
    double a, c;
};

/* 1 / sqrt(r^2 + delta) (softened 1/r) */
class kernel_SoftLaplacian : public H2_3D_Tree {
  public:
    kernel_SoftLaplacian(double L, int tree_level, int interpolation_order, double epsilon,
                         int use_chebyshev, double delta_)
        : H2_3D_Tree(L, tree_level, interpolation_order, epsilon, use_chebyshev), delta(delta_) {}
    void SetKernelProperty() override
    {
        homogen = 0; symmetry = 1; kernelType = "SoftLaplacian";
        device_kernel = PBB_KERNEL_SOFT_LAPLACIAN; device_param[0] = delta;
    }
    double EvaluateKernel(const vector3 &t, const vector3 &s) override
    {
        return 1. / std::sqrt(pbb_detail::dist2(t, s) + delta);
    }
    double delta;
};

#undef PBB_KERNEL_CLASS

#endif
