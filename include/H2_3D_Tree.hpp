/*
 * H2_3D_Tree.hpp -- the kernel/tree base class of the C++ drop-in API.
 *
 * Mirrors PBBFMM3D's H2_3D_Tree (include/H2_3D_Tree.hpp:16-116): the same
 * constructor H2_3D_Tree(L, tree_level, interpolation_order, epsilon,
 * use_chebyshev), buildFMMTree(), and the virtuals EvaluateKernel (pure)
 * and SetKernelProperty that a kernel class overrides.  buildFMMTree()
 * runs the FP64 operator precompute through pbb_ops_create, evaluating the
 * kernel through the virtual EvaluateKernel on the host.
 *
 * What differs: the near field runs on the GPU, so a kernel class also names
 * a device functor in SetKernelProperty (device_kernel / device_param, see
 * pbb_kernel_id in pbbfmm3d_amd.h).  A class that sets none can still build
 * its operators, but H2_3D_Compute rejects it.  Trees are reusable: any
 * number of H2_3D_Compute calls may use one built tree.
 *
 * Operator cache: with cache_dir set (empty by default: no files),
 * buildFMMTree() works as the reference's PrecomputeM2L
 * (H2_3D_Tree.cpp:74-99): it reads {cache_dir}/{kernelType}Cheb{K,U,V}{n}.bin
 * when they fit (PrecomputeAvailable, :102-131), and otherwise precomputes
 * and writes them, in the reference's file format.
 */
#ifndef PBBFMM3D_AMD_H2_3D_TREE_HPP
#define PBBFMM3D_AMD_H2_3D_TREE_HPP

#include <stdexcept>
#include <string>

#include "bbfmm.h"
#include "pbbfmm3d_amd.h"

class H2_3D_Tree {
  public:
    H2_3D_Tree(double L_, int tree_level_, int interpolation_order_, double epsilon_,
               int use_chebyshev_)
        : L(L_), tree_level(tree_level_), interpolation_order(interpolation_order_),
          epsilon(epsilon_), use_chebyshev(use_chebyshev_)
    {
        dof = &dof_storage;
        dof->s = dof->f = 1;
        cutoff.s = cutoff.f = 0;
        for (double &p : device_param) p = 0;
    }
    virtual ~H2_3D_Tree()
    {
        if (ops_) pbb_ops_destroy(ops_);
    }
    H2_3D_Tree(const H2_3D_Tree &) = delete;
    H2_3D_Tree &operator=(const H2_3D_Tree &) = delete;

    /* Precompute (H2_3D_Tree.cpp:40-66).  Throws std::runtime_error. */
    void buildFMMTree()
    {
        SetKernelProperty();
        if (dof->s != 1 || dof->f != 1)
            throw std::runtime_error("only scalar kernels (dof 1/1) are supported");
        pbb_kernel_desc d{};
        d.id = device_kernel;
        d.symmetry = symmetry;
        d.homogen = homogen;
        for (int i = 0; i < 4; i++) d.param[i] = device_param[i];
        if (ops_) {
            pbb_ops_destroy(ops_);
            ops_ = nullptr;
        }
        pbb_status st = PBB_ERR_NOT_FOUND;
        std::string fk, fu, fv;
        cache_hit = false;
        if (!cache_dir.empty() && use_chebyshev == 1) {
            const std::string base = cache_dir + "/" + kernelType + "Cheb";
            const std::string n = std::to_string(interpolation_order) + ".bin";
            fk = base + "K" + n;
            fu = base + "U" + n;
            fv = base + "V" + n;
            st = pbb_ops_read_cache(&d, L, tree_level, interpolation_order, epsilon, fk.c_str(), fu.c_str(),
                                    fv.c_str(), &ops_);
        }
        cache_hit = st == PBB_OK;
        if (st != PBB_OK) {  // no usable cache: precompute (and write it)
            st = pbb_ops_create(&d, &H2_3D_Tree::eval_trampoline, this, L, tree_level, interpolation_order,
                                epsilon, use_chebyshev, &ops_);
            if (st != PBB_OK) throw std::runtime_error(std::string("buildFMMTree: ") + pbb_last_error());
            if (!fk.empty() && pbb_ops_write_cache(ops_, fk.c_str(), fu.c_str(), fv.c_str()) != PBB_OK)
                throw std::runtime_error(std::string("buildFMMTree: ") + pbb_last_error());
        }
        pbb_ops_info info{};
        pbb_ops_query(ops_, &info);
        cutoff.f = info.rank_f;
        cutoff.s = info.rank_s;
        homogen = -homogen; /* stored flipped, as H2_3D_Tree.cpp:45 does */
        computed = true;
    }

    virtual double EvaluateKernel(const vector3 &targetpos, const vector3 &sourcepos) = 0;
    virtual void SetKernelProperty() {}

    pbb_ops *ops() const
    {
        if (!ops_) throw std::runtime_error("buildFMMTree() has not been called");
        return ops_;
    }
    /* the device functor of this kernel, with homogen as declared */
    pbb_kernel_desc device_desc() const
    {
        pbb_kernel_desc d{};
        d.id = device_kernel;
        d.symmetry = symmetry;
        d.homogen = computed ? -homogen : homogen;
        for (int i = 0; i < 4; i++) d.param[i] = device_param[i];
        return d;
    }

    /* public state, as the reference exposes it (H2_3D_Tree.hpp:22-56) */
    doft *dof;
    double L;
    int tree_level;
    int interpolation_order;
    double epsilon;
    double homogen = 0;
    int symmetry = 1;
    int use_chebyshev;
    doft cutoff;
    std::string kernelType;
    std::string cache_dir;  /* operator cache directory; empty: no cache files */
    bool cache_hit = false; /* the last buildFMMTree() read its operators from cache_dir */
    bool computed = false;
    int device_kernel = PBB_KERNEL_NONE;
    double device_param[4];
    /* precision mode of H2_3D_Compute on this tree (pbb_precision); FP64,
       the reference's arithmetic, unless the caller chooses a mixed mode */
    int precision = PBB_PREC_FP64;

  private:
    static double eval_trampoline(void *ctx, const double *t, const double *s)
    {
        H2_3D_Tree *self = static_cast<H2_3D_Tree *>(ctx);
        return self->EvaluateKernel(vector3(t[0], t[1], t[2]), vector3(s[0], s[1], s[2]));
    }
    doft dof_storage;
    pbb_ops *ops_ = nullptr;
};

#endif
