// feature_compact.hip -- gather descriptor sets onto their non-zero columns.
//
// A descriptor column that is zero for every atom of the sets a call works on, in X and in
// the Jacobian, adds nothing to any kernel entry or matvec product built from them (the
// FCHL19 channels of neighbour elements a data set does not contain: N and S for C9H8O4
// conformers use 312 of the 720 columns).  The derivative kernels and the matrix-free
// solver gather their sets once onto the used columns and run on those; the results
// differ only in the order of the FP64 additions.
#include <algorithm>
#include <vector>

#include "feature_compact.hpp"

namespace kf {
namespace {

__device__ inline int clamp_n(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

__global__ __launch_bounds__(256) void fill_int_kernel(int *p, int n, int v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// flag[f] = 1 when column f is non-zero in X or dX for some atom of the set.  One
// work-group per molecule; dynamic LDS of rep ints.
__global__ __launch_bounds__(256) void active_cols_kernel(const double *__restrict__ X,
                                                          const double *__restrict__ dX,
                                                          const int *__restrict__ N, int ma,
                                                          int rep, int *__restrict__ flag) {
    extern __shared__ int lf[];
    const int a = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int f = tid; f < rep; f += 256) lf[f] = 0;
    __syncthreads();
    const int na = clamp_n(N[a], ma), nr = na * rep;
    const double *x = X + (size_t)a * ma * rep;
    for (int e = tid; e < nr; e += 256)
        if (x[e] != 0.0) lf[e % rep] = 1;
    if (dX) {
        const int L = 3 * ma;  // a column of one atom's Jacobian is a run of L coordinates
        const double *j = dX + (size_t)a * ma * rep * L;
        for (int r = wave; r < nr; r += 4) {
            bool nz = false;
            for (int c = lane; c < L; c += 64) nz |= j[(size_t)r * L + c] != 0.0;
            if (__ballot(nz) && lane == 0) lf[r % rep] = 1;
        }
    }
    __syncthreads();
    for (int f = tid; f < rep; f += 256)
        if (lf[f]) flag[f] = 1;
}

// Y[r][g][c] = X[r][cols[g]][c]: row r of rep runs of L doubles -> nf runs.  One
// work-group per row.
__global__ __launch_bounds__(256) void gather_cols_kernel(const double *__restrict__ X, int rep,
                                                          int L, const int *__restrict__ cols,
                                                          int nf, double *__restrict__ Y) {
    const size_t r = blockIdx.x;
    const double *src = X + r * rep * L;
    double *dst = Y + r * nf * L;
    const int n = nf * L;
    for (int e = threadIdx.x; e < n; e += 256) {
        const int g = e / L, c = e - g * L;
        dst[e] = src[(size_t)cols[g] * L + c];
    }
}

// Gather X (and dX, if given) of one set onto the columns cols (device) into ws.
int gather_set(const double *X, const double *dX, int nm, int ma, int rep, const int *cols,
               int nf, int ws, hipStream_t stream, const double **Xc, const double **dXc) {
    const size_t rows = (size_t)nm * ma, L = 3 * (size_t)ma;
    const size_t nx = rows * nf, nj = dX ? rows * nf * L : 0;
    double *buf = static_cast<double *>(workspace(ws, sizeof(double) * (nx + nj + 1), stream));
    KF_REQUIRE(buf, KF_ERR_NOMEM, "workspace allocation failed (%zu B)", sizeof(double) * (nx + nj));
    *Xc = buf;
    *dXc = dX ? buf + nx : nullptr;
    if (rows == 0) return KF_OK;
    hipLaunchKernelGGL(gather_cols_kernel, dim3((unsigned)rows), dim3(256), 0, stream, X, rep, 1,
                       cols, nf, buf);
    KF_HIP_TRY(hipGetLastError());
    if (dX) {
        hipLaunchKernelGGL(gather_cols_kernel, dim3((unsigned)rows), dim3(256), 0, stream, dX, rep,
                           (int)L, cols, nf, buf + nx);
        KF_HIP_TRY(hipGetLastError());
    }
    return KF_OK;
}

}  // namespace

int compact_columns(CompactSets &c, int rep, int ws1, int ws2, hipStream_t stream, int *nf_out) {
    *nf_out = rep;
    if (rep < 1 || (c.nm1 == 0 && c.nm2 == 0)) return KF_OK;
    const size_t lds = sizeof(int) * rep;
    if (lds > 65536) return KF_OK;  // wider than one work-group's flags: run on every column
    const bool same = c.x1 == c.x2 && c.dx1 == c.dx2 && c.n1 == c.n2 && c.nm1 == c.nm2 &&
                      c.ma1 == c.ma2;
    int *flag = static_cast<int *>(workspace(WS_AUX1, sizeof(int) * (2 * (size_t)rep + 1), stream));
    int *hflag = static_cast<int *>(host_staging(WS_AUX1, sizeof(int) * rep));
    KF_REQUIRE(flag && hflag, KF_ERR_NOMEM, "workspace allocation failed");
    hipLaunchKernelGGL(fill_int_kernel, dim3((rep + 255) / 256), dim3(256), 0, stream, flag, rep, 0);
    KF_HIP_TRY(hipGetLastError());
    if (c.nm1 > 0) {
        hipLaunchKernelGGL(active_cols_kernel, dim3(c.nm1), dim3(256), lds, stream, c.x1, c.dx1,
                           c.n1, c.ma1, rep, flag);
        KF_HIP_TRY(hipGetLastError());
    }
    if (!same && c.nm2 > 0) {
        hipLaunchKernelGGL(active_cols_kernel, dim3(c.nm2), dim3(256), lds, stream, c.x2, c.dx2,
                           c.n2, c.ma2, rep, flag);
        KF_HIP_TRY(hipGetLastError());
    }
    KF_HIP_TRY(hipMemcpyAsync(hflag, flag, sizeof(int) * rep, hipMemcpyDeviceToHost, stream));
    KF_HIP_TRY(hipStreamSynchronize(stream));
    std::vector<int> cols;
    for (int f = 0; f < rep; ++f)
        if (hflag[f]) cols.push_back(f);
    if ((int)cols.size() == rep) return KF_OK;
    if (cols.empty()) cols.push_back(0);  // all-zero sets: one (zero) column keeps rep >= 1
    const int nf = (int)cols.size();
    std::copy(cols.begin(), cols.end(), hflag);
    int *dcols = flag + rep;
    KF_HIP_TRY(hipMemcpyAsync(dcols, hflag, sizeof(int) * nf, hipMemcpyHostToDevice, stream));
    int rc;
    const double *x1 = nullptr, *dx1 = nullptr, *x2 = nullptr, *dx2 = nullptr;
    if ((rc = gather_set(c.x1, c.dx1, c.nm1, c.ma1, rep, dcols, nf, ws1, stream, &x1, &dx1)))
        return rc;
    if (same) {
        x2 = x1;
        dx2 = dx1;
    } else if ((rc = gather_set(c.x2, c.dx2, c.nm2, c.ma2, rep, dcols, nf, ws2, stream, &x2,
                                &dx2))) {
        return rc;
    }
    c.x1 = x1; c.dx1 = dx1; c.x2 = x2; c.dx2 = dx2;
    *nf_out = nf;
    // the staging copy above must land before hflag is reused by a later call
    KF_HIP_TRY(hipStreamSynchronize(stream));
    return KF_OK;
}

}  // namespace kf
