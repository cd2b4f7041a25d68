// ORACLE — test infrastructure only. Nothing in the product imports, links or calls this.
//
// A second Python module around the reference's own CPU SpMM, pbr_batched_matmul_cpu
// (Backend/graph_link/pbr_matrix/pbr_matrix.cpp:17-68), built by `make -C oracle ref`
// together with the reference's module graph_link_core (pbr_matrix.cpp + bindings.cpp).
// The reference's bindings.cpp binds that function without including torch's type caster
// for at::Tensor (it includes only ATen headers), so the binding it defines rejects every
// tensor argument. This file binds the same compiled function again from a translation unit
// that includes <torch/extension.h>; the PBR classes it takes are the ones graph_link_core
// registers (import graph_link_core first).
#include <torch/extension.h>

#include "pbr_matrix/pbr_matrix.hpp"

template <typename I, typename S>
static at::Tensor spmm(const pbr_matrix_t<I, S>& p, at::Tensor x) {
  return pbr_batched_matmul_cpu<I, S>(p, x);
}

PYBIND11_MODULE(graph_link_ref_spmm, m) {
  m.def("pbr_batched_matmul_cpu", &spmm<int32_t, float>);
  m.def("pbr_batched_matmul_cpu", &spmm<int64_t, float>);
  m.def("pbr_batched_matmul_cpu", &spmm<int32_t, double>);
  m.def("pbr_batched_matmul_cpu", &spmm<int64_t, double>);
}
