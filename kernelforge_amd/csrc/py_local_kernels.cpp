// py_local_kernels.cpp -- pybind11 module `kernelforge_amd.local_kernels`.
//
// Function names, keyword names (X, dX, Q, N, sigma, X1, ... N2, alpha, alpha_desc,
// alpha_desc_F, alpha_E), argument order, shapes and error types follow the
// reference module (src/local_kernels_bindings.cpp:759-1536,
// python/kernelforge/local_kernels.pyi).  Inputs are copied to the device, the
// kernels run through the C ABI of libkf_hip.so, and float64 NumPy arrays return.
#include <cmath>

#include "py_common.hpp"

using namespace kfpy;

namespace {

struct Set {
    int nm, ma, rep;
};

int parse_precision(const std::string &p) {
    if (p == "fp64" || p == "float64") return KF_PREC_F64;
    if (p == "bf16x3") return KF_PREC_BF16X3;
    if (p == "i8x6") return KF_PREC_I8X6;
    throw py::value_error("precision must be 'fp64', 'i8x6' or 'bf16x3'");
}

// Download a float32 or float64 device result into a NumPy array of that dtype.
py::array download_result(void *dev, size_t n, int prec, std::vector<py::ssize_t> shape) {
    if (prec != KF_PREC_BF16X3) {
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


// dX must be (nm, max_atoms, rep, 3*max_atoms) for the set it belongs to.
void check_dx(const carray<double> &dX, const Set &s) {
    if (dX.ndim() != 4 || dX.shape(0) != s.nm || dX.shape(1) != s.ma || dX.shape(2) != s.rep ||
        dX.shape(3) != 3 * (py::ssize_t)s.ma)
        throw py::value_error("dX shape mismatch.");
}

py::array_t<double> download_f64(const DevBuf<double> &d, std::vector<py::ssize_t> shape) {
    py::array_t<double> out(shape);
    d.download(out.mutable_data(), (size_t)out.size());
    return out;
}

struct Two {
    Set a, b;
};
Two check_two(const carray<double> &X1, const carray<double> &X2, const carray<int> &Q1,
              const carray<int> &Q2, const carray<int> &N1, const carray<int> &N2) {
    if (X1.ndim() != 3 || X2.ndim() != 3 || Q1.ndim() != 2 || Q2.ndim() != 2 || N1.ndim() != 1 ||
        N2.ndim() != 1)
        throw py::value_error(
            "Expect shapes: X1(nm1,max_atoms1,rep), X2(nm2,max_atoms2,rep), Q1(nm1,max_atoms1), "
            "Q2(nm2,max_atoms2), N1(nm1), N2(nm2).");
    Two t{check_set(X1, Q1, N1, "X1"), check_set(X2, Q2, N2, "X2")};
    if (t.a.rep != t.b.rep) throw py::value_error("X2 rep_size mismatch.");
    return t;
}

using RectFn = int (*)(const double *, const double *, const double *, const double *,
                       const int *, const int *, const int *, const int *, int, int, int, int,
                       int, double, double *, void *);

// Rectangular derivative forms.  rows/cols: 0 = energies only, 1 = forces only,
// 2 = energies then forces.
py::array_t<double> rect_form(const char *name, RectFn fn, const carray<double> &X1,
                              const carray<double> *dX1, const carray<double> &X2,
                              const carray<double> *dX2, const carray<int> &Q1,
                              const carray<int> &Q2, const carray<int> &N1, const carray<int> &N2,
                              double sigma, int rows, int cols) {
    const Two t = check_two(X1, X2, Q1, Q2, N1, N2);
    if (dX1) check_dx(*dX1, t.a);
    if (dX2) check_dx(*dX2, t.b);
    const long naq1 = naq_of(N1, t.a.ma), naq2 = naq_of(N2, t.b.ma);
    const py::ssize_t R = rows == 0 ? t.a.nm : rows == 1 ? naq1 : t.a.nm + naq1;
    const py::ssize_t C = cols == 0 ? t.b.nm : cols == 1 ? naq2 : t.b.nm + naq2;
    if (!(std::isfinite(sigma) && sigma > 0.0)) throw py::value_error("sigma must be > 0");
    if (R == 0 || C == 0) return py::array_t<double>(std::vector<py::ssize_t>{R, C});
    DevBuf<double> d1 = to_device(X1), d2 = to_device(X2);
    DevBuf<double> j1 = dX1 ? to_device(*dX1) : DevBuf<double>();
    DevBuf<double> j2 = dX2 ? to_device(*dX2) : DevBuf<double>();
    DevBuf<int> q1 = to_device(Q1), q2 = to_device(Q2), n1 = to_device(N1), n2 = to_device(N2);
    DevBuf<double> out((size_t)R * C);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = fn(d1.p, d2.p, j1.p, j2.p, q1.p, q2.p, n1.p, n2.p, t.a.nm, t.b.nm, t.a.ma, t.b.ma,
                t.a.rep, sigma, out.p, nullptr);
        if (rc == KF_OK) rc = kf_stream_synchronize(nullptr);
    }
    check(rc, name);
    return download_f64(out, {R, C});
}

int jac_adapter(const double *x1, const double *x2, const double *, const double *dx2,
                const int *q1, const int *q2, const int *n1, const int *n2, int nm1, int nm2,
                int ma1, int ma2, int rep, double sigma, double *K, void *st) {
    return kf_local_kernel_jacobian(x1, x2, dx2, q1, q2, n1, n2, nm1, nm2, ma1, ma2, rep, sigma,
                                    K, st);
}
int jac_t_adapter(const double *x1, const double *x2, const double *dx1, const double *,
                  const int *q1, const int *q2, const int *n1, const int *n2, int nm1, int nm2,
                  int ma1, int ma2, int rep, double sigma, double *K, void *st) {
    return kf_local_kernel_jacobian_t(x1, x2, dx1, q1, q2, n1, n2, nm1, nm2, ma1, ma2, rep,
                                      sigma, K, st);
}

py::array_t<double> kernel_gaussian_jacobian(carray<double> X1, carray<double> X2,
                                             carray<double> dX2, carray<int> Q1, carray<int> Q2,
                                             carray<int> N1, carray<int> N2, double sigma) {
    return rect_form("kernel_gaussian_jacobian", jac_adapter, X1, nullptr, X2, &dX2, Q1, Q2, N1,
                     N2, sigma, 0, 1);
}
py::array_t<double> kernel_gaussian_jacobian_t(carray<double> X1, carray<double> dX1,
                                               carray<double> X2, carray<int> Q1,
                                               carray<int> Q2, carray<int> N1, carray<int> N2,
                                               double sigma) {
    return rect_form("kernel_gaussian_jacobian_t", jac_t_adapter, X1, &dX1, X2, nullptr, Q1, Q2,
                     N1, N2, sigma, 1, 0);
}
py::array_t<double> kernel_gaussian_hessian(carray<double> X1, carray<double> dX1,
                                            carray<double> X2, carray<double> dX2,
                                            carray<int> Q1, carray<int> Q2, carray<int> N1,
                                            carray<int> N2, double sigma) {
    return rect_form("kernel_gaussian_hessian", kf_local_kernel_hessian, X1, &dX1, X2, &dX2, Q1,
                     Q2, N1, N2, sigma, 1, 1);
}
py::array_t<double> kernel_gaussian_full(carray<double> X1, carray<double> dX1, carray<double> X2,
                                         carray<double> dX2, carray<int> Q1, carray<int> Q2,
                                         carray<int> N1, carray<int> N2, double sigma) {
    return rect_form("kernel_gaussian_full", kf_local_kernel_full, X1, &dX1, X2, &dX2, Q1, Q2,
                     N1, N2, sigma, 2, 2);
}

using SymmFn = int (*)(const double *, const double *, const int *, const int *, int, int, int,
                       double, double *, void *);

// Symmetric derivative forms: with_e adds the nm energy rows; rfp packs the upper
// triangle (TRANSR='N', UPLO='U').
py::array_t<double> symm_form(const char *name, SymmFn fn, const carray<double> &X,
                              const carray<double> &dX, const carray<int> &Q, const carray<int> &N,
                              double sigma, bool with_e, bool rfp) {
    const Set s = check_set(X, Q, N, "X");
    check_dx(dX, s);
    if (!(std::isfinite(sigma) && sigma > 0.0)) throw py::value_error("sigma must be > 0");
    const py::ssize_t n = (with_e ? s.nm : 0) + naq_of(N, s.ma);
    const size_t len = rfp ? (size_t)n * (n + 1) / 2 : (size_t)n * n;
    std::vector<py::ssize_t> shape = rfp ? std::vector<py::ssize_t>{(py::ssize_t)len}
                                         : std::vector<py::ssize_t>{n, n};
    if (n == 0) return py::array_t<double>(shape);
    DevBuf<double> dx = to_device(X), dj = to_device(dX);
    DevBuf<int> q = to_device(Q), nn = to_device(N);
    DevBuf<double> out(len);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = fn(dx.p, dj.p, q.p, nn.p, s.nm, s.ma, s.rep, sigma, out.p, nullptr);
        if (rc == KF_OK) rc = kf_stream_synchronize(nullptr);
    }
    check(rc, name);
    return download_f64(out, shape);
}


// ---- matrix-free products (alpha_desc and the three matvecs) ----
py::array_t<double> alpha_desc_py(carray<double> dX2, carray<int> Q2, carray<int> N2,
                                  carray<double> alpha) {
    if (dX2.ndim() != 4 || Q2.ndim() != 2 || N2.ndim() != 1 || alpha.ndim() != 1)
        throw py::value_error("shape error");
    const int nm = (int)dX2.shape(0), ma = (int)dX2.shape(1), rep = (int)dX2.shape(2);
    if (dX2.shape(3) != 3 * (py::ssize_t)ma) throw py::value_error("dX2 shape mismatch");
    if (Q2.shape(0) != nm || Q2.shape(1) != ma) throw py::value_error("Q2 shape mismatch");
    if (N2.shape(0) != nm) throw py::value_error("N2 length mismatch");
    if (alpha.shape(0) < naq_of(N2, ma)) throw py::value_error("alpha length mismatch");
    std::vector<py::ssize_t> shape{nm, ma, rep};
    if ((size_t)nm * ma * rep == 0) return py::array_t<double>(shape);
    DevBuf<double> dj = to_device(dX2), da = to_device(alpha), out((size_t)nm * ma * rep);
    DevBuf<int> dq = to_device(Q2), dn = to_device(N2);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = kf_local_alpha_desc(dj.p, dq.p, dn.p, nm, ma, rep, da.p, out.p, nullptr);
        if (rc == KF_OK) rc = kf_stream_synchronize(nullptr);
    }
    check(rc, "kernel_gaussian_local_compute_alpha_desc");
    return download_f64(out, shape);
}

// mode bit 1: E, bit 2: F (dX1 given), bit 4: alpha_E given
py::object matvec_py(const char *name, int mode, const carray<double> &X1,
                     const carray<double> *dX1, const carray<double> &X2,
                     const carray<double> &AD, const carray<double> *aE, const carray<int> &Q1,
                     const carray<int> &Q2, const carray<int> &N1, const carray<int> &N2,
                     double sigma) {
    const Two t = check_two(X1, X2, Q1, Q2, N1, N2);
    if (dX1) check_dx(*dX1, t.a);
    if (AD.ndim() != 3 || AD.shape(0) != t.b.nm || AD.shape(1) != t.b.ma ||
        AD.shape(2) != t.b.rep)
        throw py::value_error("alpha_desc shape mismatch");
    if (aE && (aE->ndim() != 1 || aE->shape(0) != t.b.nm))
        throw py::value_error("alpha_E length mismatch");
    if (!(std::isfinite(sigma) && sigma > 0.0)) throw py::value_error("sigma must be > 0");
    const long naq1 = naq_of(N1, t.a.ma);
    DevBuf<double> E((mode & 1) ? t.a.nm : 0), F((mode & 2) ? naq1 : 0);
    if (t.a.nm > 0) {
        DevBuf<double> d1 = to_device(X1), d2 = to_device(X2), dad = to_device(AD);
        DevBuf<double> j1 = dX1 ? to_device(*dX1) : DevBuf<double>();
        DevBuf<double> de = aE ? to_device(*aE) : DevBuf<double>();
        DevBuf<int> q1 = to_device(Q1), q2 = to_device(Q2), n1 = to_device(N1), n2 = to_device(N2);
        int rc;
        {
            py::gil_scoped_release nogil;
            if (mode == 7)
                rc = kf_local_full_matvec(d1.p, j1.p, d2.p, dad.p, de.p, q1.p, q2.p, n1.p, n2.p,
                                          t.a.nm, t.b.nm, t.a.ma, t.b.ma, t.a.rep, sigma, E.p,
                                          F.p, nullptr);
            else if (mode == 2)
                rc = kf_local_hessian_matvec(d1.p, j1.p, d2.p, dad.p, q1.p, q2.p, n1.p, n2.p,
                                             t.a.nm, t.b.nm, t.a.ma, t.b.ma, t.a.rep, sigma, F.p,
                                             nullptr);
            else
                rc = kf_local_jacobian_t_matvec(d1.p, d2.p, dad.p, q1.p, q2.p, n1.p, n2.p,
                                                t.a.nm, t.b.nm, t.a.ma, t.b.ma, t.a.rep, sigma,
                                                E.p, nullptr);
            if (rc == KF_OK) rc = kf_stream_synchronize(nullptr);
        }
        check(rc, name);
    }
    py::array_t<double> e = download_f64(E, {(py::ssize_t)E.n});
    py::array_t<double> f = download_f64(F, {(py::ssize_t)F.n});
    if (mode == 7) return py::make_tuple(e, f);
    return mode == 2 ? py::object(f) : py::object(e);
}

}  // namespace

PYBIND11_MODULE(local_kernels, m) {
    m.doc() = "Local (element-matched) Gaussian kernels for FCHL19 (MI355X / gfx950)";
    m.def("kernel_gaussian", &kernel_gaussian, py::arg("X1"), py::arg("X2"), py::arg("Q1"),
          py::arg("Q2"), py::arg("N1"), py::arg("N2"), py::arg("sigma"), py::kw_only(),
          py::arg("precision") = "fp64",
          "K[a,b] = sum over label-matched atom pairs of exp(-|x1-x2|^2/2sigma^2), (nm1, nm2).\n"
          "precision (keyword only, an extension): 'i8x6' FP64-accurate int8-split contraction\n"
          "(six 7-bit slices per row, exact int32 MFMA products; entries ~1e-11 relative, float64\n"
          "out) or 'bf16x3' BF16x3 MFMA contraction (float32 out).");
    m.def("kernel_gaussian_symm", &kernel_gaussian_symm, py::arg("X"), py::arg("Q"), py::arg("N"),
          py::arg("sigma"), py::kw_only(), py::arg("precision") = "fp64",
          "Symmetric local kernel, dense (nm, nm).");
    m.def("kernel_gaussian_symm_rfp", &kernel_gaussian_symm_rfp, py::arg("X"), py::arg("Q"),
          py::arg("N"), py::arg("sigma"), py::kw_only(), py::arg("precision") = "fp64",
          "Symmetric local kernel in RFP (TRANSR='N', UPLO='U'), length nm(nm+1)/2.");
    m.def("kernel_gaussian_jacobian", &kernel_gaussian_jacobian, py::arg("X1"), py::arg("X2"),
          py::arg("dX2"), py::arg("Q1"), py::arg("Q2"), py::arg("N1"), py::arg("N2"),
          py::arg("sigma"), "Jacobian kernel dK/dR2, (nm1, naq2).");
    m.def("kernel_gaussian_jacobian_t", &kernel_gaussian_jacobian_t, py::arg("X1"),
          py::arg("dX1"), py::arg("X2"), py::arg("Q1"), py::arg("Q2"), py::arg("N1"),
          py::arg("N2"), py::arg("sigma"),
          "Transposed Jacobian kernel -dK/dR1, (naq1, nm2); forces = K_t @ alpha.");
    m.def("kernel_gaussian_hessian", &kernel_gaussian_hessian, py::arg("X1"), py::arg("dX1"),
          py::arg("X2"), py::arg("dX2"), py::arg("Q1"), py::arg("Q2"), py::arg("N1"),
          py::arg("N2"), py::arg("sigma"), "Hessian kernel, (naq1, naq2).");
    m.def("kernel_gaussian_full", &kernel_gaussian_full, py::arg("X1"), py::arg("dX1"),
          py::arg("X2"), py::arg("dX2"), py::arg("Q1"), py::arg("Q2"), py::arg("N1"),
          py::arg("N2"), py::arg("sigma"),
          "Full energy+force kernel [[scalar, jacobian], [jacobian_t, hessian]], "
          "(nm1+naq1, nm2+naq2).");
    m.def(
        "kernel_gaussian_hessian_symm",
        [](carray<double> X, carray<double> dX, carray<int> Q, carray<int> N, double sigma) {
            return symm_form("kernel_gaussian_hessian_symm", kf_local_kernel_hessian_symm, X, dX,
                             Q, N, sigma, false, false);
        },
        py::arg("X"), py::arg("dX"), py::arg("Q"), py::arg("N"), py::arg("sigma"),
        "Symmetric Hessian kernel, (naq, naq).");
    m.def(
        "kernel_gaussian_hessian_symm_rfp",
        [](carray<double> X, carray<double> dX, carray<int> Q, carray<int> N, double sigma) {
            return symm_form("kernel_gaussian_hessian_symm_rfp",
                             kf_local_kernel_hessian_symm_rfp, X, dX, Q, N, sigma, false, true);
        },
        py::arg("X"), py::arg("dX"), py::arg("Q"), py::arg("N"), py::arg("sigma"),
        "Symmetric Hessian kernel in RFP (TRANSR='N', UPLO='U'), length naq(naq+1)/2.");
    m.def(
        "kernel_gaussian_full_symm",
        [](carray<double> X, carray<double> dX, carray<int> Q, carray<int> N, double sigma) {
            return symm_form("kernel_gaussian_full_symm", kf_local_kernel_full_symm, X, dX, Q, N,
                             sigma, true, false);
        },
        py::arg("X"), py::arg("dX"), py::arg("Q"), py::arg("N"), py::arg("sigma"),
        "Symmetric full energy+force kernel, (nm+naq, nm+naq).");
    m.def(
        "kernel_gaussian_full_symm_rfp",
        [](carray<double> X, carray<double> dX, carray<int> Q, carray<int> N, double sigma) {
            return symm_form("kernel_gaussian_full_symm_rfp", kf_local_kernel_full_symm_rfp, X,
                             dX, Q, N, sigma, true, true);
        },
        py::arg("X"), py::arg("dX"), py::arg("Q"), py::arg("N"), py::arg("sigma"),
        "Symmetric full kernel in RFP (TRANSR='N', UPLO='U'), length BIG(BIG+1)/2, "
        "BIG = nm + naq.");
    m.def("kernel_gaussian_local_compute_alpha_desc", &alpha_desc_py, py::arg("dX2"),
          py::arg("Q2"), py::arg("N2"), py::arg("alpha"),
          "alpha_desc[b,i,:] = dX2[b,i,:,:3N_b] @ alpha[offs_b:offs_b+3N_b], (nm2, max_atoms2, rep).");
    m.def(
        "kernel_gaussian_local_hessian_matvec",
        [](carray<double> X1, carray<double> dX1, carray<double> X2, carray<double> alpha_desc,
           carray<int> Q1, carray<int> Q2, carray<int> N1, carray<int> N2, double sigma) {
            return matvec_py("kernel_gaussian_local_hessian_matvec", 2, X1, &dX1, X2, alpha_desc,
                             nullptr, Q1, Q2, N1, N2, sigma);
        },
        py::arg("X1"), py::arg("dX1"), py::arg("X2"), py::arg("alpha_desc"), py::arg("Q1"),
        py::arg("Q2"), py::arg("N1"), py::arg("N2"), py::arg("sigma"),
        "Forces F (naq1,) = Hessian kernel @ alpha via alpha_desc.");
    m.def(
        "kernel_gaussian_local_jacobian_t_matvec",
        [](carray<double> X1, carray<double> X2, carray<double> alpha_desc, carray<int> Q1,
           carray<int> Q2, carray<int> N1, carray<int> N2, double sigma) {
            return matvec_py("kernel_gaussian_local_jacobian_t_matvec", 1, X1, nullptr, X2,
                             alpha_desc, nullptr, Q1, Q2, N1, N2, sigma);
        },
        py::arg("X1"), py::arg("X2"), py::arg("alpha_desc"), py::arg("Q1"), py::arg("Q2"),
        py::arg("N1"), py::arg("N2"), py::arg("sigma"),
        "Energies E (nm1,) = Jacobian kernel @ alpha via alpha_desc.");
    m.def(
        "kernel_gaussian_local_full_matvec",
        [](carray<double> X1, carray<double> dX1, carray<double> X2,
           carray<double> alpha_desc_F, carray<double> alpha_E, carray<int> Q1, carray<int> Q2,
           carray<int> N1, carray<int> N2, double sigma) {
            return matvec_py("kernel_gaussian_local_full_matvec", 7, X1, &dX1, X2, alpha_desc_F,
                             &alpha_E, Q1, Q2, N1, N2, sigma);
        },
        py::arg("X1"), py::arg("dX1"), py::arg("X2"), py::arg("alpha_desc_F"),
        py::arg("alpha_E"), py::arg("Q1"), py::arg("Q2"), py::arg("N1"), py::arg("N2"),
        py::arg("sigma"), "(E, F) = full energy+force kernel @ [alpha_E; alpha_F].");
}
