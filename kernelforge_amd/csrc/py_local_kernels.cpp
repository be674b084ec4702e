// py_local_kernels.cpp -- pybind11 module `kernelforge_amd.local_kernels`.
//
// Function names, keyword names (X, dX, Q, N, sigma, X1, ... N2, alpha, alpha_desc,
// alpha_desc_F, alpha_E), argument order, shapes and error types follow the
// reference module (src/local_kernels_bindings.cpp:759-1536,
// python/kernelforge/local_kernels.pyi).  Inputs are copied to the device, the
// kernels run through the C ABI of libkf_hip.so, and float64 NumPy arrays return.
#include "py_common.hpp"

using namespace kfpy;

namespace {

struct Set {
    int nm, ma, rep;
};

int parse_precision(const std::string &p) {
    if (p == "fp64" || p == "float64") return KF_PREC_F64;
    if (p == "bf16x3") return KF_PREC_BF16X3;
    throw py::value_error("precision must be 'fp64' or 'bf16x3'");
}

// Download a float32 or float64 device result into a NumPy array of that dtype.
py::array download_result(void *dev, size_t n, int prec, std::vector<py::ssize_t> shape) {
    if (prec == KF_PREC_F64) {
        py::array_t<double> out(shape);
        hip_check(hipMemcpy(out.mutable_data(), dev, n * sizeof(double), hipMemcpyDeviceToHost), "D2H");
        return out;
    }
    py::array_t<float> out(shape);
    hip_check(hipMemcpy(out.mutable_data(), dev, n * sizeof(float), hipMemcpyDeviceToHost), "D2H");
    return out;
}

Set check_set(const carray<double> &X, const carray<int> &Q, const carray<int> &N,
              const char *name) {
    if (X.ndim() != 3 || Q.ndim() != 2 || N.ndim() != 1)
        throw py::value_error(std::string("Expect shapes: ") + name +
                              "(nm,max_atoms,rep), Q(nm,max_atoms), N(nm).");
    Set s{(int)X.shape(0), (int)X.shape(1), (int)X.shape(2)};
    if (Q.shape(0) != s.nm || Q.shape(1) != s.ma) throw py::value_error("Q shape mismatch.");
    if (N.shape(0) != s.nm) throw py::value_error("N length mismatch.");
    return s;
}

py::array kernel_gaussian_symm_rfp(carray<double> X, carray<int> Q, carray<int> N, double sigma,
                                   const std::string &precision) {
    const int prec = parse_precision(precision);
    const Set s = check_set(X, Q, N, "X");
    const size_t nt = (size_t)s.nm * (s.nm + 1) / 2;
    if (s.nm == 0) return download_result(nullptr, 0, prec, {(py::ssize_t)0});
    DevBuf<double> dX = to_device(X);
    DevBuf<int> dQ = to_device(Q), dN = to_device(N);
    DevBuf<double> dK(nt);  // large enough for either precision
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = kf_local_kernel_symm_rfp(prec, dX.p, dQ.p, dN.p, s.nm, s.ma, s.rep, sigma, dK.p,
                                      nullptr);
    }
    check(rc, "kernel_gaussian_symm_rfp");
    return download_result(dK.p, nt, prec, {(py::ssize_t)nt});
}

py::array kernel_gaussian_symm(carray<double> X, carray<int> Q, carray<int> N, double sigma,
                               const std::string &precision) {
    const int prec = parse_precision(precision);
    const Set s = check_set(X, Q, N, "X");
    const size_t nn = (size_t)s.nm * s.nm;
    if (s.nm == 0) return download_result(nullptr, 0, prec, {0, 0});
    DevBuf<double> dX = to_device(X);
    DevBuf<int> dQ = to_device(Q), dN = to_device(N);
    DevBuf<double> dK(nn);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = kf_local_kernel_symm(prec, dX.p, dQ.p, dN.p, s.nm, s.ma, s.rep, sigma, dK.p, nullptr);
    }
    check(rc, "kernel_gaussian_symm");
    return download_result(dK.p, nn, prec, {(py::ssize_t)s.nm, (py::ssize_t)s.nm});
}

py::array kernel_gaussian(carray<double> X1, carray<double> X2, carray<int> Q1, carray<int> Q2,
                          carray<int> N1, carray<int> N2, double sigma,
                          const std::string &precision) {
    const int prec = parse_precision(precision);
    if (X1.ndim() != 3 || X2.ndim() != 3 || Q1.ndim() != 2 || Q2.ndim() != 2 || N1.ndim() != 1 ||
        N2.ndim() != 1)
        throw py::value_error(
            "Expect shapes: X1(nm1,max_atoms1,rep), X2(nm2,max_atoms2,rep), Q1(nm1,max_atoms1), "
            "Q2(nm2,max_atoms2), N1(nm1), N2(nm2).");
    const Set a = check_set(X1, Q1, N1, "X1"), b = check_set(X2, Q2, N2, "X2");
    if (a.rep != b.rep) throw py::value_error("X2 rep_size mismatch.");
    const size_t nn = (size_t)a.nm * b.nm;
    if (a.nm == 0 || b.nm == 0)
        return download_result(nullptr, 0, prec, {(py::ssize_t)a.nm, (py::ssize_t)b.nm});
    DevBuf<double> dX1 = to_device(X1), dX2 = to_device(X2);
    DevBuf<int> dQ1 = to_device(Q1), dQ2 = to_device(Q2), dN1 = to_device(N1), dN2 = to_device(N2);
    DevBuf<double> dK(nn);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = kf_local_kernel(prec, dX1.p, dX2.p, dQ1.p, dQ2.p, dN1.p, dN2.p, a.nm, b.nm, a.ma,
                             b.ma, a.rep, sigma, dK.p, nullptr);
    }
    check(rc, "kernel_gaussian");
    return download_result(dK.p, nn, prec, {(py::ssize_t)a.nm, (py::ssize_t)b.nm});
}

}  // namespace

PYBIND11_MODULE(local_kernels, m) {
    m.doc() = "Local (element-matched) Gaussian kernels for FCHL19 (MI355X / gfx950)";
    m.def("kernel_gaussian", &kernel_gaussian, py::arg("X1"), py::arg("X2"), py::arg("Q1"),
          py::arg("Q2"), py::arg("N1"), py::arg("N2"), py::arg("sigma"), py::kw_only(),
          py::arg("precision") = "fp64",
          "K[a,b] = sum over label-matched atom pairs of exp(-|x1-x2|^2/2sigma^2), (nm1, nm2).\n"
          "precision='bf16x3' (keyword only, an extension): BF16x3 MFMA contraction, float32 out.");
    m.def("kernel_gaussian_symm", &kernel_gaussian_symm, py::arg("X"), py::arg("Q"), py::arg("N"),
          py::arg("sigma"), py::kw_only(), py::arg("precision") = "fp64",
          "Symmetric local kernel, dense (nm, nm).");
    m.def("kernel_gaussian_symm_rfp", &kernel_gaussian_symm_rfp, py::arg("X"), py::arg("Q"),
          py::arg("N"), py::arg("sigma"), py::kw_only(), py::arg("precision") = "fp64",
          "Symmetric local kernel in RFP (TRANSR='N', UPLO='U'), length nm(nm+1)/2.");
}
