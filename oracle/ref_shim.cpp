// ref_shim.cpp -- C entry points over the reference's own FP64 CPU sources.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile (target `ref`) into
// oracle/_ref/ from the reference's source files where they lie
// (src/fchl19_repr.cpp, src/fchl19v2_repr.cpp, src/math.cpp), to validate the oracle
// restatement.
// Never shipped as product code; tests skip when oracle/_ref is absent.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "fchl19_repr.hpp"
#include "fchl19v2_repr.hpp"
#include "math.hpp"

// Grids exactly as the reference bindings build them
// (src/fchl19_repr_bindings.cpp:72-84 for the rep path, :125-144 for the grad path).
static void grids(int nRs2, int nRs3, int nFourier, double rcut, double acut, bool grad_path,
                  std::vector<double> &Rs2, std::vector<double> &Rs3, std::vector<double> &Ts) {
    for (int i = 1; i <= nRs2; ++i) Rs2.push_back(rcut * double(i) / double(nRs2));
    for (int i = 1; i <= nRs3; ++i) Rs3.push_back(acut * double(i) / double(nRs3));
    int nT = 2 * nFourier;
    if (grad_path && nT < 2) nT = 2;
    for (int i = 0; i < nT; ++i) Ts.push_back(nT == 1 ? 0.0 : M_PI * double(i) / double(nT - 1));
}

extern "C" int kfr_fchl19(const double *coords, const int *z, int n, const int *elements,
                          int nelem, int nRs2, int nRs3, int nFourier, double eta2, double eta3,
                          double zeta, double rcut, double acut, double d2, double d3,
                          double weight, double *rep_out, double *grad_out) {
    try {
        std::vector<double> c(coords, coords + 3 * n);
        std::vector<int> zz(z, z + n), el(elements, elements + nelem);
        std::vector<double> Rs2, Rs3, Ts, rep, grad;
        grids(nRs2, nRs3, nFourier, rcut, acut, grad_out != nullptr, Rs2, Rs3, Ts);
        const double w3 = std::sqrt(eta3 / M_PI) * weight;
        if (grad_out) {
            kf::fchl19::generate_fchl_acsf_and_gradients(c, zz, el, Rs2, Rs3, Ts, eta2, eta3,
                                                         zeta, rcut, acut, d2, d3, w3, rep, grad);
            std::memcpy(grad_out, grad.data(), grad.size() * sizeof(double));
        } else {
            kf::fchl19::generate_fchl_acsf(c, zz, el, Rs2, Rs3, Ts, eta2, eta3, zeta, rcut, acut,
                                           d2, d3, w3, rep);
        }
        std::memcpy(rep_out, rep.data(), rep.size() * sizeof(double));
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

// solve_cholesky_rfp (src/math.cpp:134-171): +l2 on the RFP diagonal, dpftrf, dpftrs.
extern "C" int kfr_cho_solve_rfp(double *arf, const double *y, int n, double l2, double *alpha) {
    try {
        kf::math::solve_cholesky_rfp(arf, y, n, alpha, l2, 'U', 'N');
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

// FCHL19v2 (src/fchl19v2_repr.cpp:2874-2986) with grids, nabasis and w3 built as its
// bindings build them (src/fchl19v2_repr_bindings.cpp:56-88,120-140).  tb, ab index the
// TwoBodyType / ThreeBodyType enumerators (src/fchl19v2_repr.hpp:11-27).
extern "C" int kfr_fchl19v2(const double *coords, const int *z, int n, const int *elements,
                            int nelem, int nRs2, int nRs3, int nRs3_minus, int nFourier,
                            int nCosine, double eta2, double eta3, double eta3_minus, double zeta,
                            double rcut, double acut, double d2, double d3, double weight, int tb,
                            int ab, double *rep_out, double *grad_out) {
    try {
        namespace v2 = kf::fchl19v2;
        std::vector<double> c(coords, coords + 3 * n);
        std::vector<int> zz(z, z + n), el(elements, elements + nelem);
        std::vector<double> Rs2, Rs3, Rs3m, rep, grad;
        for (int i = 1; i <= nRs2; ++i) Rs2.push_back(rcut * double(i) / double(nRs2));
        for (int i = 1; i <= nRs3; ++i) Rs3.push_back(acut * double(i) / double(nRs3));
        for (int i = 1; i <= nRs3_minus; ++i) Rs3m.push_back(acut * double(i) / double(nRs3_minus));
        const auto tbt = static_cast<v2::TwoBodyType>(tb);
        const auto abt = static_cast<v2::ThreeBodyType>(ab);
        const bool odd = abt == v2::ThreeBodyType::OddFourier_Rbar ||
                         abt == v2::ThreeBodyType::OddFourier_SplitR;
        const int nab = odd ? 2 * nFourier : (nCosine > 0 ? nCosine : 2 * nFourier);
        const double w3 = std::sqrt(eta3 / M_PI) * weight;
        if (grad_out) {
            v2::generate_and_gradients(c, zz, el, Rs2, Rs3, Rs3m, eta2, eta3, eta3_minus, zeta,
                                       rcut, acut, d2, d3, w3, tbt, abt, nab, rep, grad);
            std::memcpy(grad_out, grad.data(), grad.size() * sizeof(double));
        } else {
            v2::generate(c, zz, el, Rs2, Rs3, Rs3m, eta2, eta3, eta3_minus, zeta, rcut, acut, d2,
                         d3, w3, tbt, abt, nab, rep);
        }
        std::memcpy(rep_out, rep.data(), rep.size() * sizeof(double));
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}
