// py_fchl19_repr.cpp -- pybind11 module `kernelforge_amd.fchl19_repr`.
//
// Same functions, argument names, defaults and error types as the reference module
// (src/fchl19_repr_bindings.cpp:219-315, python/kernelforge/fchl19_repr.pyi), plus
// batched entry points that take the padded (nm, max_atoms, ...) layout the local
// kernels consume.  Every call runs kf_fchl19_f64 on the GPU.
#include <cmath>

#include "py_common.hpp"

using namespace kfpy;

namespace {

struct Basis {
    std::vector<int> elements;
    std::vector<double> Rs2, Rs3;
    int nabasis = 0;
    double w3 = 0.0;
};

// Grids as the reference bindings build them: Rs2 = rcut*i/nRs2, Rs3 = acut*i/nRs3
// (i = 1..n), nabasis = 2*nFourier (forced >= 2 on the gradient path), and
// w3 = sqrt(eta3/pi) * three_body_weight (src/fchl19_repr_bindings.cpp:72-87,125-164).
Basis make_basis(std::vector<int> elements, int nRs2, int nRs3, int nFourier, double rcut,
                 double acut, double eta3, double weight, bool grad) {
    Basis b;
    if (elements.empty()) elements = {1, 6, 7, 8, 16};
    b.elements = elements;
    for (int i = 1; i <= nRs2; ++i) b.Rs2.push_back(rcut * double(i) / double(nRs2));
    for (int i = 1; i <= nRs3; ++i) b.Rs3.push_back(acut * double(i) / double(nRs3));
    b.nabasis = 2 * nFourier;
    if (grad && b.nabasis < 2) b.nabasis = 2;
    b.w3 = std::sqrt(eta3 / M_PI) * weight;
    return b;
}

long rep_size_of(const Basis &b) {
    const long ne = (long)b.elements.size();
    return ne * (long)b.Rs2.size() + ne * (ne + 1) / 2 * (long)b.Rs3.size() * b.nabasis;
}

std::vector<double> coords_vector(const py::array &arr, size_t natoms) {
    auto buf = arr.cast<carray<double>>();
    if (buf.ndim() == 2) {
        if ((size_t)buf.shape(0) != natoms || buf.shape(1) != 3)
            throw py::value_error("coords must have shape (n_atoms, 3)");
    } else if (buf.ndim() == 1) {
        if ((size_t)buf.shape(0) != natoms * 3)
            throw py::value_error("coords 1D length must be n_atoms*3");
    } else {
        throw py::value_error("coords must be a 1D or 2D NumPy array");
    }
    return std::vector<double>(buf.data(), buf.data() + natoms * 3);
}

// Run the batched device kernel on host arrays.
void run_batch(const std::vector<double> &C, const std::vector<int> &Q, const std::vector<int> &N,
               int nm, int ma, const Basis &b, double eta2, double eta3, double zeta,
               double rcut, double acut, double d2, double d3, bool grad, double *X,
               double *dX) {
    const long rs = rep_size_of(b);
    DevBuf<double> dC = to_device(C);
    DevBuf<int> dQ = to_device(Q), dN = to_device(N);
    DevBuf<double> dXo((size_t)nm * ma * rs);
    DevBuf<double> ddX(grad ? (size_t)nm * ma * rs * 3 * ma : 0);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = kf_fchl19_f64(dC.p, dQ.p, dN.p, nm, ma, b.elements.data(), (int)b.elements.size(),
                           b.Rs2.data(), (int)b.Rs2.size(), b.Rs3.data(), (int)b.Rs3.size(),
                           b.nabasis, eta2, eta3, zeta, rcut, acut, d2, d3, b.w3, dXo.p,
                           grad ? ddX.p : nullptr, nullptr);
    }
    if (rc == KF_ERR_ELEMENT)
        throw std::runtime_error("nuclear_z contains an element not present in elements");
    check(rc, "generate_fchl_acsf");
    dXo.download(X, (size_t)nm * ma * rs);
    if (grad) ddX.download(dX, (size_t)nm * ma * rs * 3 * ma);
}

py::array_t<double> generate_fchl_acsf(const py::array &coords, const py::array &nuclear_z,
                                       std::vector<int> elements, int nRs2, int nRs3,
                                       int nFourier, double eta2, double eta3, double zeta,
                                       double rcut, double acut, double two_body_decay,
                                       double three_body_decay, double three_body_weight) {
    auto z = nuclear_z.cast<carray<int>>();
    if (z.ndim() != 1) throw py::value_error("array must be 1D");
    const size_t n = (size_t)z.shape(0);
    if (n == 0) throw py::value_error("n_atoms must be > 0");
    std::vector<double> C = coords_vector(coords, n);
    std::vector<int> Q(z.data(), z.data() + n), N{(int)n};
    Basis b = make_basis(elements, nRs2, nRs3, nFourier, rcut, acut, eta3, three_body_weight, false);
    py::array_t<double> out({(py::ssize_t)n, (py::ssize_t)rep_size_of(b)});
    run_batch(C, Q, N, 1, (int)n, b, eta2, eta3, zeta, rcut, acut, two_body_decay,
              three_body_decay, false, out.mutable_data(), nullptr);
    return out;
}

py::tuple generate_fchl_acsf_and_gradients(const py::array &coords, const py::array &nuclear_z,
                                           std::vector<int> elements, int nRs2, int nRs3,
                                           int nFourier, double eta2, double eta3, double zeta,
                                           double rcut, double acut, double two_body_decay,
                                           double three_body_decay, double three_body_weight) {
    auto z = nuclear_z.cast<carray<int>>();
    if (z.ndim() != 1) throw py::value_error("array must be 1D");
    const size_t n = (size_t)z.shape(0);
    if (n == 0) throw py::value_error("n_atoms must be > 0");
    std::vector<double> C = coords_vector(coords, n);
    std::vector<int> Q(z.data(), z.data() + n), N{(int)n};
    Basis b = make_basis(elements, nRs2, nRs3, nFourier, rcut, acut, eta3, three_body_weight, true);
    const long rs = rep_size_of(b);
    py::array_t<double> rep({(py::ssize_t)n, (py::ssize_t)rs});
    py::array_t<double> grad({(py::ssize_t)n, (py::ssize_t)rs, (py::ssize_t)(3 * n)});
    run_batch(C, Q, N, 1, (int)n, b, eta2, eta3, zeta, rcut, acut, two_body_decay,
              three_body_decay, true, rep.mutable_data(), grad.mutable_data());
    return py::make_tuple(rep, grad);
}

// Batched: coords (nm, max_atoms, 3), Q (nm, max_atoms), N (nm,).
py::object generate_batch(const carray<double> &coords, const carray<int> &Q, const carray<int> &N,
                          bool with_gradients, std::vector<int> elements, int nRs2, int nRs3,
                          int nFourier, double eta2, double eta3, double zeta, double rcut,
                          double acut, double two_body_decay, double three_body_decay,
                          double three_body_weight) {
    if (coords.ndim() != 3 || coords.shape(2) != 3 || Q.ndim() != 2 || N.ndim() != 1)
        throw py::value_error("expected coords (nm, max_atoms, 3), Q (nm, max_atoms), N (nm,)");
    const int nm = (int)coords.shape(0), ma = (int)coords.shape(1);
    if (Q.shape(0) != nm || Q.shape(1) != ma || N.shape(0) != nm)
        throw py::value_error("Q/N shape mismatch");
    Basis b = make_basis(elements, nRs2, nRs3, nFourier, rcut, acut, eta3, three_body_weight,
                         with_gradients);
    const long rs = rep_size_of(b);
    std::vector<double> C(coords.data(), coords.data() + coords.size());
    std::vector<int> q(Q.data(), Q.data() + Q.size()), nn(N.data(), N.data() + N.size());
    py::array_t<double> X({(py::ssize_t)nm, (py::ssize_t)ma, (py::ssize_t)rs});
    if (!with_gradients) {
        run_batch(C, q, nn, nm, ma, b, eta2, eta3, zeta, rcut, acut, two_body_decay,
                  three_body_decay, false, X.mutable_data(), nullptr);
        return X;
    }
    py::array_t<double> dX(
        {(py::ssize_t)nm, (py::ssize_t)ma, (py::ssize_t)rs, (py::ssize_t)(3 * ma)});
    run_batch(C, q, nn, nm, ma, b, eta2, eta3, zeta, rcut, acut, two_body_decay,
              three_body_decay, true, X.mutable_data(), dX.mutable_data());
    return py::make_tuple(X, dX);
}

}  // namespace

PYBIND11_MODULE(fchl19_repr, m) {
    m.doc() = "FCHL19 representation and gradients (MI355X / gfx950)";
    m.def(
        "compute_rep_size",
        [](size_t nelements, size_t nbasis2, size_t nbasis3, size_t nabasis) {
            return nelements * nbasis2 + nelements * (nelements + 1) / 2 * nbasis3 * nabasis;
        },
        py::arg("nelements"), py::arg("nbasis2"), py::arg("nbasis3"), py::arg("nabasis"),
        "Compute the total representation size per-atom.");
    const std::vector<int> def_el{1, 6, 7, 8, 16};
#define KF_FCHL_ARGS                                                                          \
    py::arg("elements") = def_el, py::arg("nRs2") = 24, py::arg("nRs3") = 20,                 \
    py::arg("nFourier") = 1, py::arg("eta2") = 0.32, py::arg("eta3") = 2.7,                   \
    py::arg("zeta") = M_PI, py::arg("rcut") = 8.0, py::arg("acut") = 8.0,                     \
    py::arg("two_body_decay") = 1.8, py::arg("three_body_decay") = 0.57,                      \
    py::arg("three_body_weight") = 13.4
    m.def("generate_fchl_acsf", &generate_fchl_acsf, py::arg("coords"), py::arg("nuclear_z"),
          KF_FCHL_ARGS,
          "FCHL19 representation (n_atoms, rep_size) of one molecule (reference: "
          "src/fchl19_repr.cpp:89).");
    m.def("generate_fchl_acsf_and_gradients", &generate_fchl_acsf_and_gradients,
          py::arg("coords"), py::arg("nuclear_z"), KF_FCHL_ARGS,
          "Representation and its Jacobian (rep (n, rs), grad (n, rs, 3n)) of one molecule "
          "(reference: src/fchl19_repr.cpp:308).");
    m.def("generate_fchl_acsf_batch", &generate_batch, py::arg("coords"), py::arg("Q"),
          py::arg("N"), py::arg("with_gradients") = false, KF_FCHL_ARGS,
          "Padded batch: coords (nm, max_atoms, 3), Q (nm, max_atoms), N (nm,) -> X "
          "(nm, max_atoms, rs) or (X, dX (nm, max_atoms, rs, 3*max_atoms)).");
#undef KF_FCHL_ARGS
}
