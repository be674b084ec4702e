// py_kernelmath.cpp -- pybind11 module `kernelforge_amd.kernelmath`.
//
// Same functions, argument names, defaults and error types as the reference
// (src/math_bindings.cpp:248-322, python/kernelforge/kernelmath.pyi) for the
// Cholesky/RFP subset on the hot path.  All arithmetic runs on the GPU through
// kf_rfp_potrf / kf_rfp_potrs / kf_trttf / kf_tfttr.
#include "py_common.hpp"

using namespace kfpy;

namespace {

char norm_uplo(char u) { return (u == 'l' || u == 'L') ? 'L' : 'U'; }
char norm_tr(char t) { return (t == 't' || t == 'T') ? 'T' : 'N'; }
char swap_uplo(char u) { return u == 'U' ? 'L' : 'U'; }

// cho_solve_rfp: (K + l2 I) alpha = y with K in RFP (TRANSR='N', UPLO='U').
// K_rfp is overwritten with the Cholesky factor (reaches the caller's array when it
// was already C-contiguous float64, as in the reference, src/math_bindings.cpp:49-76).
py::array_t<double> cho_solve_rfp(carray<double> K_rfp, carray<double> y, double l2) {
    if (K_rfp.ndim() != 1 || y.ndim() != 1)
        throw std::runtime_error("K_rfp must be 1D RFP and y must be 1D");
    const long n = (long)y.shape(0);
    if ((size_t)K_rfp.shape(0) != (size_t)n * (n + 1) / 2)
        throw std::runtime_error("K_rfp length must be n*(n+1)/2");
    if (n <= 0) throw std::runtime_error("n must be > 0");
    DevBuf<double> dK = to_device(K_rfp);
    DevBuf<double> db = to_device(y);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = kf_rfp_potrf(KF_PREC_F64, dK.p, (int)n, l2, nullptr);
    }
    if (rc > 0) {
        dK.download(K_rfp.mutable_data(), (size_t)K_rfp.size());
        throw std::runtime_error("DPFTRF failed, info=" + std::to_string(rc));
    }
    check(rc, "cho_solve_rfp");
    {
        py::gil_scoped_release nogil;
        rc = kf_rfp_potrs(KF_PREC_F64, dK.p, (int)n, db.p, 1, nullptr);
    }
    check(rc, "cho_solve_rfp");
    dK.download(K_rfp.mutable_data(), (size_t)K_rfp.size());
    py::array_t<double> alpha(n);
    db.download(alpha.mutable_data(), (size_t)n);
    return alpha;
}

// solve_cholesky: dense K (n, n) C-order, uses its lower triangle (the reference's
// dpotrf('U') on the C-order buffer, src/math.cpp:69-115); K is left unchanged for a
// symmetric input, as the reference restores it.
py::array_t<double> solve_cholesky(carray<double> K, carray<double> y, double regularize) {
    if (K.ndim() != 2 || K.shape(0) != K.shape(1)) throw std::runtime_error("K must be n x n.");
    if (y.ndim() != 1 || y.shape(0) != K.shape(0)) throw std::runtime_error("y must have length n.");
    const int n = (int)K.shape(0);
    if (n <= 0) throw std::runtime_error("n must be > 0");
    DevBuf<double> dA = to_device(K);
    DevBuf<double> dR((size_t)n * (n + 1) / 2);
    DevBuf<double> db = to_device(y);
    int rc;
    {
        py::gil_scoped_release nogil;
        // C-order lower triangle == column-major upper of A^T: RFP(N, U) of the matrix
        rc = kf_trttf('N', 'U', n, dA.p, n, dR.p, nullptr);
        if (rc == KF_OK) rc = kf_rfp_potrf(KF_PREC_F64, dR.p, n, regularize, nullptr);
    }
    if (rc > 0) throw std::runtime_error("Cholesky decomposition failed, info=" + std::to_string(rc));
    check(rc, "solve_cholesky");
    {
        py::gil_scoped_release nogil;
        rc = kf_rfp_potrs(KF_PREC_F64, dR.p, n, db.p, 1, nullptr);
    }
    check(rc, "solve_cholesky");
    py::array_t<double> alpha(n);
    db.download(alpha.mutable_data(), (size_t)n);
    return alpha;
}

py::array_t<double> full_to_rfp(carray<double> A, char uplo, char transr) {
    if (A.ndim() != 2 || A.shape(0) != A.shape(1)) throw std::runtime_error("A must be square (n x n).");
    uplo = norm_uplo(uplo);
    transr = norm_tr(transr);
    const int n = (int)A.shape(0);
    const size_t nt = (size_t)n * (n + 1) / 2;
    DevBuf<double> dA = to_device(A);
    DevBuf<double> dR(nt);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = kf_trttf(transr, swap_uplo(uplo), n, dA.p, n > 0 ? n : 1, dR.p, nullptr);
    }
    check(rc, "dtrttf_");
    py::array_t<double> out((py::ssize_t)nt);
    dR.download(out.mutable_data(), nt);
    return out;
}

py::array_t<double> rfp_to_full(carray<double> ARF, int n, char uplo, char transr) {
    if (n <= 0) throw std::runtime_error("n must be > 0");
    if (ARF.ndim() != 1) throw std::runtime_error("ARF must be a 1-D array.");
    const size_t nt = (size_t)n * (n + 1) / 2;
    if ((size_t)ARF.shape(0) != nt) throw std::runtime_error("ARF length must be n*(n+1)/2.");
    uplo = norm_uplo(uplo);
    transr = norm_tr(transr);
    DevBuf<double> dR = to_device(ARF);
    DevBuf<double> dA((size_t)n * n);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = kf_tfttr(transr, swap_uplo(uplo), n, dR.p, dA.p, n, 1, nullptr);
    }
    check(rc, "dtfttr_");
    py::array_t<double> out({(py::ssize_t)n, (py::ssize_t)n});
    dA.download(out.mutable_data(), (size_t)n * n);
    return out;
}

}  // namespace

PYBIND11_MODULE(kernelmath, m) {
    m.doc() = "Cholesky / RFP solvers (MI355X / gfx950)";
    m.def("solve_cholesky", &solve_cholesky, py::arg("K"), py::arg("y"),
          py::arg("regularize") = 0.0,
          "Solve (K + regularize I) alpha = y by Cholesky (reference src/math.cpp:69).");
    m.def("cho_solve_rfp", &cho_solve_rfp, py::arg("K_rfp"), py::arg("y"), py::arg("l2") = 0.0,
          "Solve (K + l2*I) @ alpha = y where K is in RFP packed format (TRANSR='N', UPLO='U').\n"
          "K_rfp is overwritten with the Cholesky factor -- pass K_rfp.copy() to preserve it.");
    m.def("full_to_rfp", &full_to_rfp, py::arg("A"), py::arg("uplo") = 'U',
          py::arg("transr") = 'N', "Full (n x n, C-order) -> RFP (length n(n+1)/2).");
    m.def("rfp_to_full", &rfp_to_full, py::arg("ARF"), py::arg("n"), py::arg("uplo") = 'U',
          py::arg("transr") = 'N', "RFP -> full symmetric (n x n, C-order).");
    m.def(
        "get_blas_info", []() { return std::string("kernelforge_amd HIP (gfx950), no host BLAS"); },
        "Backend identification string.");
}
