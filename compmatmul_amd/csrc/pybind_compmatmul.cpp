// Python module "compmatmul": the drop-in boundary.
//
// Same functions, argument names, defaults, validation order, messages and
// exception types as the reference's pybind11 module (compmatmul.cpp:13-209):
//   gemm(C, AT, B)
//   matmul_fft(C, AT, B, d, b, seed=None, hash=None)
//   matmul_fwht(C, AT, B, d, b, seed=None, hash=None)
// The computation goes through the C-ABI of libcompmatmul_amd.so.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdint>
#include <new>
#include <optional>
#include <stdexcept>
#include <string>

#include "../../include/compmatmul_amd.h"

namespace py = pybind11;

namespace {

// compmatmul.cpp:13-74
void check_array_arguments(py::array C, py::array AT, py::array B) {
  const auto f64 = py::dtype::of<double>();
  if (!C.dtype().is(f64) || !AT.dtype().is(f64) || !B.dtype().is(f64))
    throw std::invalid_argument("compmatmul only supports double precision floats for now");
  if (C.ndim() != 2)
    throw std::invalid_argument("C must have ndim=2 (got " + std::to_string(C.ndim()) + ")");
  if (AT.ndim() != 2)
    throw std::invalid_argument("AT must have ndim=2 (got " + std::to_string(AT.ndim()) + ")");
  if (B.ndim() != 2)
    throw std::invalid_argument("B must have ndim=2 (got " + std::to_string(B.ndim()) + ")");
  const auto contig = py::detail::npy_api::constants::NPY_ARRAY_C_CONTIGUOUS_;
  for (const py::array* a : {&C, &AT, &B})
    if ((a->flags() & contig) != contig)
      throw std::invalid_argument("The arrays must be C-contiguous");
  const int q = int(AT.shape(0));
  const int p = int(AT.shape(1));
  const int r = int(B.shape(1));
  if (B.shape(0) != q || C.shape(0) != p || C.shape(1) != r)
    throw std::invalid_argument("Shape mismatch! Note that AT is assumed to be transposed");
  if (reinterpret_cast<uintptr_t>(AT.data()) % 64 != 0)
    throw std::invalid_argument("AT is not aligned at a 64-byte boundary!");
  if (reinterpret_cast<uintptr_t>(B.data()) % 64 != 0)
    throw std::invalid_argument("B is not aligned at a 64-byte boundary!");
  if (reinterpret_cast<uintptr_t>(C.data()) % 64 != 0)
    throw std::invalid_argument("C is not aligned at a 64-byte boundary!");
}

// compmatmul.cpp:87-105
void check_other_arguments(int d, int b, const std::optional<std::string>& hash) {
  if (b < 2 || (b & (b - 1)) != 0)
    throw std::invalid_argument("b must be power of 2 greater than 1 (got " + std::to_string(b) + ")");
  if (d < 1) throw std::invalid_argument("d must be positive (got " + std::to_string(d) + ")");
  if (hash && *hash != "multiply shift" && *hash != "multiply add shift" &&
      *hash != "tabulation8" && *hash != "tabulation16")
    throw std::invalid_argument("Invalid hash `" + *hash +
                                "' provided; hash must be one of `multiply shift', "
                                "`multiply add shift', `tabulation8', or `tabulation16'");
}

int family_of(const std::optional<std::string>& hash) {
  if (!hash || *hash == "multiply add shift") return CM_MULTIPLY_ADD_SHIFT;
  if (*hash == "multiply shift") return CM_MULTIPLY_SHIFT;
  if (*hash == "tabulation8") return CM_TABULATION8;
  if (*hash == "tabulation16") return CM_TABULATION16;
  throw std::runtime_error("Tried to use an invalid hash");
}

void raise(int rc) {
  if (rc == CM_OK) return;
  const std::string msg = cm_last_error();
  if (rc == CM_EINVAL) throw std::invalid_argument(msg);
  if (rc == CM_ENOMEM) throw std::bad_alloc();
  throw std::runtime_error(msg);
}

void gemm(py::array C, py::array AT, py::array B) {
  check_array_arguments(C, AT, B);
  const int q = int(AT.shape(0)), p = int(AT.shape(1)), r = int(B.shape(1));
  raise(cm_gemm(p, q, r, static_cast<double*>(C.mutable_data()),
                static_cast<const double*>(AT.data()), static_cast<const double*>(B.data())));
}

template <int Variant>
void matmul(py::array C, py::array AT, py::array B, int d, int b, std::optional<uint64_t> seed,
            std::optional<std::string> hash) {
  check_array_arguments(C, AT, B);
  check_other_arguments(d, b, hash);
  const int q = int(AT.shape(0)), p = int(AT.shape(1)), r = int(B.shape(1));
  const int family = family_of(hash);
  auto fn = Variant == CM_FFT ? cm_matmul_fft : cm_matmul_fwht;
  raise(fn(p, q, r, static_cast<double*>(C.mutable_data()), static_cast<const double*>(AT.data()),
           static_cast<const double*>(B.data()), d, b, seed.value_or(0), seed ? 1 : 0, family));
}

}  // namespace

PYBIND11_MODULE(compmatmul, m) {
  m.doc() = "Compressed Matrix Multiplication Python bindings (MI355X)";
  m.def("gemm", &gemm,
        "Computes C = A^T B using an FP64 GEMM on the GPU",
        py::arg("C"), py::arg("AT"), py::arg("B"));
  m.def("matmul_fft", &matmul<CM_FFT>,
        "Computes C = A^T B using compressed matrix multiplication with FFT",
        py::arg("C"), py::arg("AT"), py::arg("B"), py::arg("d"), py::arg("b"),
        py::arg("seed") = py::none(), py::arg("hash") = py::none());
  m.def("matmul_fwht", &matmul<CM_FWHT>,
        "Computes C = A^T B using compressed matrix multiplication with FWHT",
        py::arg("C"), py::arg("AT"), py::arg("B"), py::arg("d"), py::arg("b"),
        py::arg("seed") = py::none(), py::arg("hash") = py::none());
}
