// py_fchl19v2_repr.cpp -- pybind11 module `kernelforge_amd.fchl19v2_repr`.
//
// Same functions, argument names, defaults and error types as the reference module
// (src/fchl19v2_repr_bindings.cpp:235-308, python/kernelforge/fchl19v2_repr.pyi), plus a
// batched entry point over the padded (nm, max_atoms, ...) layout.  Every call runs
// kf_fchl19v2_f64 on the GPU.
#include <cmath>
#include <string>

#include "py_common.hpp"

using namespace kfpy;

namespace {

const char *const kTwoBody[] = {"log_normal", "gaussian_r", "gaussian_log_r",
                                "gaussian_r_no_pow", "bessel"};
const char *const kThreeBody[] = {"odd_fourier_rbar", "cosine_rbar", "odd_fourier_split_r",
                                  "cosine_split_r", "cosine_split_r_no_atm"};

int type_index(const std::string &s, const char *const *names, const char *what) {
    for (int i = 0; i < 5; ++i)
        if (s == names[i]) return i;
    throw py::value_error(std::string("Unknown ") + what + ": " + s);
}

bool split_type(int ab) { return ab >= 2; }
bool odd_type(int ab) { return ab == 0 || ab == 2; }

struct Basis {
    std::vector<int> elements;
    std::vector<double> Rs2, Rs3, Rs3m;
    int tb = 0, ab = 0, nabasis = 0;
    double w3 = 0.0;
};

// Grids, nabasis and w3 as the reference bindings build them: Rs = cut*i/n (i = 1..n),
// nabasis = 2 nFourier for the odd-Fourier bases and nCosine (if > 0, else 2 nFourier)
// for the cosine series, w3 = sqrt(eta3/pi) * three_body_weight
// (src/fchl19v2_repr_bindings.cpp:56-88,120).
Basis make_basis(std::vector<int> elements, int nRs2, int nRs3, int nRs3_minus, int nFourier,
                 int nCosine, double rcut, double acut, double eta3, double weight,
                 const std::string &tb, const std::string &ab) {
    Basis b;
    if (elements.empty()) elements = {1, 6, 7, 8, 16};
    b.elements = elements;
    b.tb = type_index(tb, kTwoBody, "two_body_type");
    b.ab = type_index(ab, kThreeBody, "three_body_type");
    for (int i = 1; i <= nRs2; ++i) b.Rs2.push_back(rcut * double(i) / double(nRs2));
    for (int i = 1; i <= nRs3; ++i) b.Rs3.push_back(acut * double(i) / double(nRs3));
    for (int i = 1; i <= nRs3_minus; ++i) b.Rs3m.push_back(acut * double(i) / double(nRs3_minus));
    b.nabasis = odd_type(b.ab) ? 2 * nFourier : (nCosine > 0 ? nCosine : 2 * nFourier);
    b.w3 = std::sqrt(eta3 / M_PI) * weight;
    return b;
}

// src/fchl19v2_repr.cpp:118-141
long rep_size(long nelem, long nb2, long nb3, long nab, int ab, long nb3m) {
    const long npair = nelem * (nelem + 1) / 2;
    if (split_type(ab)) {
        if (nb3m <= 0) throw py::value_error("nbasis3_minus must be > 0 for SplitR variants");
        return nelem * nb2 + npair * nb3 * nb3m * nab;
    }
    return nelem * nb2 + npair * nb3 * nab;
}

long rep_size_of(const Basis &b) {
    return rep_size((long)b.elements.size(), (long)b.Rs2.size(), (long)b.Rs3.size(), b.nabasis,
                    b.ab, (long)b.Rs3m.size());
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

void run_batch(const std::vector<double> &C, const std::vector<int> &Q, const std::vector<int> &N,
               int nm, int ma, const Basis &b, double eta2, double eta3, double eta3_minus,
               double zeta, double rcut, double acut, double d2, double d3, bool grad, double *X,
               double *dX) {
    const long rs = rep_size_of(b);
    DevBuf<double> dC = to_device(C);
    DevBuf<int> dQ = to_device(Q), dN = to_device(N);
    DevBuf<double> dXo((size_t)nm * ma * rs);
    DevBuf<double> ddX(grad ? (size_t)nm * ma * rs * 3 * ma : 0);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = kf_fchl19v2_f64(dC.p, dQ.p, dN.p, nm, ma, b.elements.data(), (int)b.elements.size(),
                             b.Rs2.data(), (int)b.Rs2.size(), b.Rs3.data(), (int)b.Rs3.size(),
                             b.Rs3m.data(), (int)b.Rs3m.size(), b.nabasis, b.tb, b.ab, eta2, eta3,
                             eta3_minus, zeta, rcut, acut, d2, d3, b.w3, dXo.p,
                             grad ? ddX.p : nullptr, nullptr);
    }
    if (rc == KF_ERR_ELEMENT)
        throw std::runtime_error("nuclear_z contains an element not present in elements");
    check(rc, "fchl19v2_repr.generate");
    dXo.download(X, (size_t)nm * ma * rs);
    if (grad) ddX.download(dX, (size_t)nm * ma * rs * 3 * ma);
}

struct Args {
    std::vector<int> elements;
    int nRs2, nRs3, nRs3_minus, nFourier, nCosine;
    double eta2, eta3, eta3_minus, zeta, rcut, acut, d2, d3, weight;
    std::string tb, ab;
    Basis basis() const {
        return make_basis(elements, nRs2, nRs3, nRs3_minus, nFourier, nCosine, rcut, acut, eta3,
                          weight, tb, ab);
    }
};

py::object generate_one(const py::array &coords, const py::array &nuclear_z, const Args &a,
                        bool grad) {
    const size_t n = (size_t)nuclear_z.cast<py::array>().shape(0);
    if (n == 0) throw py::value_error("n_atoms must be > 0");
    std::vector<double> C = coords_vector(coords, n);
    auto z = nuclear_z.cast<carray<int>>();
    if (z.ndim() != 1) throw py::value_error("array must be 1D");
    std::vector<int> Q(z.data(), z.data() + n), N{(int)n};
    const Basis b = a.basis();
    const long rs = rep_size_of(b);
    py::array_t<double> rep({(py::ssize_t)n, (py::ssize_t)rs});
    if (!grad) {
        run_batch(C, Q, N, 1, (int)n, b, a.eta2, a.eta3, a.eta3_minus, a.zeta, a.rcut, a.acut, a.d2,
                  a.d3, false, rep.mutable_data(), nullptr);
        return rep;
    }
    py::array_t<double> g({(py::ssize_t)n, (py::ssize_t)rs, (py::ssize_t)(3 * n)});
    run_batch(C, Q, N, 1, (int)n, b, a.eta2, a.eta3, a.eta3_minus, a.zeta, a.rcut, a.acut, a.d2,
              a.d3, true, rep.mutable_data(), g.mutable_data());
    return py::make_tuple(rep, g);
}

}  // namespace

PYBIND11_MODULE(fchl19v2_repr, m) {
    m.doc() = "FCHL19v2 representation with selectable two-body and three-body basis functions "
              "(MI355X / gfx950)";
    m.def(
        "compute_rep_size",
        [](size_t nelements, size_t nbasis2, size_t nbasis3, size_t nabasis,
           const std::string &three_body_type, size_t nbasis3_minus) {
            const int ab = type_index(three_body_type, kThreeBody, "three_body_type");
            return rep_size((long)nelements, (long)nbasis2, (long)nbasis3, (long)nabasis, ab,
                            (long)nbasis3_minus);
        },
        py::arg("nelements"), py::arg("nbasis2"), py::arg("nbasis3"), py::arg("nabasis"),
        py::arg("three_body_type") = "odd_fourier_rbar", py::arg("nbasis3_minus") = 0,
        "Compute the total representation size per atom.");
    const std::vector<int> def_el{1, 6, 7, 8, 16};
    auto bind = [&](const char *name, bool grad, const char *doc) {
        m.def(
            name,
            [grad](const py::array &coords, const py::array &nuclear_z, std::vector<int> elements,
                   int nRs2, int nRs3, int nRs3_minus, int nFourier, int nCosine, double eta2,
                   double eta3, double eta3_minus, double zeta, double rcut, double acut,
                   double two_body_decay, double three_body_decay, double three_body_weight,
                   const std::string &two_body_type, const std::string &three_body_type) {
                Args a{elements, nRs2, nRs3, nRs3_minus, nFourier, nCosine, eta2, eta3,
                       eta3_minus, zeta, rcut, acut, two_body_decay, three_body_decay,
                       three_body_weight, two_body_type, three_body_type};
                return generate_one(coords, nuclear_z, a, grad);
            },
            py::arg("coords"), py::arg("nuclear_z"), py::arg("elements") = def_el,
            py::arg("nRs2") = 24, py::arg("nRs3") = 20, py::arg("nRs3_minus") = 0,
            py::arg("nFourier") = 1, py::arg("nCosine") = 0, py::arg("eta2") = 0.32,
            py::arg("eta3") = 2.7, py::arg("eta3_minus") = 2.7, py::arg("zeta") = M_PI,
            py::arg("rcut") = 8.0, py::arg("acut") = 8.0, py::arg("two_body_decay") = 1.8,
            py::arg("three_body_decay") = 0.57, py::arg("three_body_weight") = 13.4,
            py::arg("two_body_type") = "log_normal",
            py::arg("three_body_type") = "odd_fourier_rbar", doc);
    };
    bind("generate", false,
         "FCHL19v2 representation (n_atoms, rep_size) of one molecule (reference: "
         "src/fchl19v2_repr.cpp:2874).");
    bind("generate_and_gradients", true,
         "Representation and Jacobian (rep (n, rs), grad (n, rs, 3n)) of one molecule "
         "(reference: src/fchl19v2_repr.cpp:2919).");
    m.def(
        "generate_batch",
        [](const carray<double> &coords, const carray<int> &Q, const carray<int> &N,
           bool with_gradients, std::vector<int> elements, int nRs2, int nRs3, int nRs3_minus,
           int nFourier, int nCosine, double eta2, double eta3, double eta3_minus, double zeta,
           double rcut, double acut, double two_body_decay, double three_body_decay,
           double three_body_weight, const std::string &two_body_type,
           const std::string &three_body_type) -> py::object {
            if (coords.ndim() != 3 || coords.shape(2) != 3 || Q.ndim() != 2 || N.ndim() != 1)
                throw py::value_error("expected coords (nm, max_atoms, 3), Q (nm, max_atoms), N (nm,)");
            const int nm = (int)coords.shape(0), ma = (int)coords.shape(1);
            if (Q.shape(0) != nm || Q.shape(1) != ma || N.shape(0) != nm)
                throw py::value_error("Q/N shape mismatch");
            Args a{elements, nRs2, nRs3, nRs3_minus, nFourier, nCosine, eta2, eta3, eta3_minus,
                   zeta, rcut, acut, two_body_decay, three_body_decay, three_body_weight,
                   two_body_type, three_body_type};
            const Basis b = a.basis();
            const long rs = rep_size_of(b);
            std::vector<double> C(coords.data(), coords.data() + coords.size());
            std::vector<int> q(Q.data(), Q.data() + Q.size()), nn(N.data(), N.data() + N.size());
            py::array_t<double> X({(py::ssize_t)nm, (py::ssize_t)ma, (py::ssize_t)rs});
            if (!with_gradients) {
                run_batch(C, q, nn, nm, ma, b, eta2, eta3, eta3_minus, zeta, rcut, acut,
                          two_body_decay, three_body_decay, false, X.mutable_data(), nullptr);
                return X;
            }
            py::array_t<double> dX(
                {(py::ssize_t)nm, (py::ssize_t)ma, (py::ssize_t)rs, (py::ssize_t)(3 * ma)});
            run_batch(C, q, nn, nm, ma, b, eta2, eta3, eta3_minus, zeta, rcut, acut,
                      two_body_decay, three_body_decay, true, X.mutable_data(), dX.mutable_data());
            return py::make_tuple(X, dX);
        },
        py::arg("coords"), py::arg("Q"), py::arg("N"), py::arg("with_gradients") = false,
        py::arg("elements") = def_el, py::arg("nRs2") = 24, py::arg("nRs3") = 20,
        py::arg("nRs3_minus") = 0, py::arg("nFourier") = 1, py::arg("nCosine") = 0,
        py::arg("eta2") = 0.32, py::arg("eta3") = 2.7, py::arg("eta3_minus") = 2.7,
        py::arg("zeta") = M_PI, py::arg("rcut") = 8.0, py::arg("acut") = 8.0,
        py::arg("two_body_decay") = 1.8, py::arg("three_body_decay") = 0.57,
        py::arg("three_body_weight") = 13.4, py::arg("two_body_type") = "log_normal",
        py::arg("three_body_type") = "odd_fourier_rbar",
        "Padded batch: coords (nm, max_atoms, 3), Q (nm, max_atoms), N (nm,) -> X "
        "(nm, max_atoms, rs) or (X, dX (nm, max_atoms, rs, 3*max_atoms)).");
}
