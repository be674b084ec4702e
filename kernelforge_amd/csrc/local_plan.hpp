// local_plan.hpp -- host-side packing plan for the element-matched kernels.
//
// Atoms of one set are sorted by (label, molecule, atom) -- a stable counting
// sort over the distinct labels -- so that, for every label, the rows of each
// molecule are one contiguous range of the packed matrix.  mol_row[l*(nm+1)+a]
// is the first packed row of molecule a for label l (mol_row[..+nm] = end), and
// src[row] = a*max_atoms + i names the source row of X.  This is the layout the
// reference builds per label (src/local_kernels.cpp:59-149,325-378), kept for all
// labels at once so one launch covers the ragged batch.
#pragma once

#include <algorithm>
#include <vector>

namespace kf {

struct LocalPlan {
    int nm = 0, max_atoms = 0, nlabels = 0, rows = 0;
    std::vector<int> labels;   // sorted distinct labels present in this set
    std::vector<int> mol_row;  // nlabels * (nm + 1)
    std::vector<int> src;      // rows
};

inline int clamp_atoms(int n, int ma) { return n < 0 ? 0 : (n > ma ? ma : n); }

// Build the plan of one set against a given (sorted) label list.
inline void build_plan(const int *Q, const int *N, int nm, int ma, const std::vector<int> &labels,
                       LocalPlan &p) {
    p.nm = nm;
    p.max_atoms = ma;
    p.labels = labels;
    p.nlabels = (int)labels.size();
    const int L = p.nlabels;
    p.mol_row.assign((size_t)L * (nm + 1), 0);
    // per (label, molecule) counts
    std::vector<int> cnt((size_t)L * nm, 0);
    // label -> index: a direct table for the small non-negative labels atomic numbers are,
    // a binary search for anything else
    constexpr int kDirect = 256;
    int direct[kDirect];
    std::fill(direct, direct + kDirect, -1);
    for (int l = 0; l < L; ++l)
        if (labels[l] >= 0 && labels[l] < kDirect) direct[labels[l]] = l;
    auto lab_idx = [&](int q) {
        if (q >= 0 && q < kDirect) return direct[q];
        auto it = std::lower_bound(labels.begin(), labels.end(), q);
        return (it != labels.end() && *it == q) ? int(it - labels.begin()) : -1;
    };
    // labels are few; cache the index of every atom
    std::vector<int> li((size_t)nm * ma, -1);
    for (int a = 0; a < nm; ++a) {
        const int na = clamp_atoms(N[a], ma);
        for (int i = 0; i < na; ++i) {
            const int l = lab_idx(Q[(size_t)a * ma + i]);
            li[(size_t)a * ma + i] = l;
            if (l >= 0) cnt[(size_t)l * nm + a]++;
        }
    }
    int row = 0;
    for (int l = 0; l < L; ++l) {
        for (int a = 0; a < nm; ++a) {
            p.mol_row[(size_t)l * (nm + 1) + a] = row;
            row += cnt[(size_t)l * nm + a];
        }
        p.mol_row[(size_t)l * (nm + 1) + nm] = row;
    }
    p.rows = row;
    p.src.assign(row > 0 ? row : 1, 0);
    std::vector<int> fill(p.mol_row);
    for (int a = 0; a < nm; ++a) {
        const int na = clamp_atoms(N[a], ma);
        for (int i = 0; i < na; ++i) {
            const int l = li[(size_t)a * ma + i];
            if (l < 0) continue;
            p.src[fill[(size_t)l * (nm + 1) + a]++] = a * ma + i;
        }
    }
}

// Sorted distinct labels of the atoms (i < N[a]).  Labels in [0, 256) (atomic numbers)
// are collected in a presence table; any others are sorted and merged in.
inline std::vector<int> distinct_labels(const int *Q, const int *N, int nm, int ma) {
    constexpr int kDirect = 256;
    bool seen[kDirect] = {};
    std::vector<int> other;
    for (int a = 0; a < nm; ++a) {
        const int na = clamp_atoms(N[a], ma);
        const int *q = Q + (size_t)a * ma;
        for (int i = 0; i < na; ++i) {
            if (q[i] >= 0 && q[i] < kDirect) seen[q[i]] = true;
            else other.push_back(q[i]);
        }
    }
    std::vector<int> out;
    for (int v = 0; v < kDirect; ++v)
        if (seen[v]) out.push_back(v);
    if (!other.empty()) {
        std::sort(other.begin(), other.end());
        other.erase(std::unique(other.begin(), other.end()), other.end());
        std::vector<int> merged(out.size() + other.size());
        merged.erase(std::merge(out.begin(), out.end(), other.begin(), other.end(), merged.begin()),
                     merged.end());
        out.swap(merged);
    }
    return out;
}

}  // namespace kf
