// The binding a 3bem maintainer adds to 3bem's cpp/ (beside fmm.h) to run the FMM
// matvec on the MI355X library: an OperatorI (cpp/operator.h:14-19) with
// FMMOperator's constructor signature (cpp/fmm.h:146-147) and apply / n_rows /
// n_cols / clone (cpp/fmm.h:149, 233-236), whose work is done through the C ABI
// of include/tbem_fmm.h (libtbem_fmm.so). It includes 3bem's own headers, so it
// is compiled inside a 3bem build; in this repository it is test infrastructure
// (tests/cpp/ref_adapter/Makefile compiles it against the reference's headers).
#ifndef TBEM_FMM_MI355X_H
#define TBEM_FMM_MI355X_H

#include <memory>
#include <vector>

#include "fmm.h"
#include "nbody_operator.h"

struct tbem_fmm_s;

namespace tbem {

// The kernel's id in the C ABI (TBEM_K_*), reading mu and nu from the elastic
// kernels' public members; -1 for kernels the GPU library does not evaluate
// (gravity and user subclasses).
template <size_t R, size_t C>
int mi355x_kernel_id(const Kernel<3, R, C>& K, double* mu, double* nu);

template <size_t R, size_t C>
struct GPUFMMOperator : public OperatorI {
    GPUFMMOperator(const Kernel<3, R, C>& K, const NBodyData<3>& data, const FMMConfig& config);
    ~GPUFMMOperator();
    GPUFMMOperator(const GPUFMMOperator&) = delete;
    GPUFMMOperator& operator=(const GPUFMMOperator&) = delete;

    size_t n_rows() const override;
    size_t n_cols() const override;
    std::vector<double> apply(const std::vector<double>& x) const override;
    // shares the device plan (octrees, C2E, task lists) with this operator
    std::unique_ptr<OperatorI> clone() const override;

  private:
    GPUFMMOperator() = default;
    tbem_fmm_s* h = nullptr;
};

// FMMOperator<3,R,C> on the GPU where the kernel is one of the library's, the
// reference's own CPU FMMOperator otherwise (and for every 2D operator).
template <size_t dim, size_t R, size_t C>
std::unique_ptr<OperatorI> make_fmm_operator(const Kernel<dim, R, C>& K, const NBodyData<dim>& data,
                                             const FMMConfig& config);

}  // namespace tbem

#endif
