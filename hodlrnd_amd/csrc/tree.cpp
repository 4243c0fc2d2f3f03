// tree.cpp -- 2^d tree and neighbour / interaction lists on the host.
//
// Bit-exact with the reference's construction:
//   HDD_tree.hpp:43-69     uniform depth: split every node while the largest
//                          node of the newest level exceeds Nmax;
//   HDD_clusters.hpp:89-133 split at mid = 0.5*(a+b) per dimension in order
//                          0..d-1, points with x[d] < mid go low; children
//                          in Morton order with dimension 0 most significant;
//                          a node's points keep ascending point ids;
//   HDD_clusters.hpp:22-41,70-79 box side L = largest extent, centre = mid;
//   HDD_node.hpp:24-31     admissible unless the max-norm centre distance is
//                          below 1.5 L_self and interaction_type > d';
//   HDD_clusters.hpp:142-162 interaction_type from the squared centre
//                          distance over max(L);
//   HDD_node.hpp:150-180   list order: siblings, then children of each of
//                          the parent's neighbours in the parent's order.
// The implementation is a level-synchronous stable counting sort over one
// permutation array, so every node is a contiguous range of tree positions.
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace hodlr {

namespace {

// box side and centre of node g from its bounds
void finish_box(hodlr_tree &t, int64_t g) {
    const int d = t.ndim;
    const double *a = &t.lo[g * d], *b = &t.hi[g * d];
    double L = std::fabs(a[0] - b[0]);
    for (int i = 0; i < d; i++) {
        double e = std::fabs(a[i] - b[i]);
        if (e > L) L = e;
    }
    t.len[g] = L;
    for (int i = 0; i < d; i++) t.centre[g * d + i] = 0.5 * (a[i] + b[i]);
}

int sharing_type(const hodlr_tree &t, int64_t ga, int64_t gb) {
    const int d = t.ndim;
    const double *ca = &t.centre[ga * d], *cb = &t.centre[gb * d];
    double L = t.len[ga] > t.len[gb] ? t.len[ga] : t.len[gb];
    double s = 0.0;
    for (int i = 0; i < d; i++) s += (ca[i] - cb[i]) * (ca[i] - cb[i]);
    double q = std::sqrt(s) / L;
    q *= q;
    if (static_cast<double>(d) + 1 > q) return static_cast<int>(static_cast<double>(d) - std::round(q));
    return -1;
}

bool is_far(const hodlr_tree &t, int64_t gself, int64_t go) {
    const int d = t.ndim;
    const double *ca = &t.centre[gself * d], *cb = &t.centre[go * d];
    double dist = std::fabs(ca[0] - cb[0]);
    for (int i = 0; i < d; i++) {
        double e = std::fabs(ca[i] - cb[i]);
        if (e > dist) dist = e;
    }
    if (dist < 1.5 * t.len[gself] && sharing_type(t, gself, go) > t.dprime) return false;
    return true;
}

}  // namespace

void build_tree(hodlr_tree &t) {
    const int d = t.ndim;
    const int64_t nch = int64_t(1) << d;
    t.perm.resize(t.n);
    for (int64_t i = 0; i < t.n; i++) t.perm[i] = i;
    t.level_off = {0, 1};
    t.start = {0};
    t.count = {t.n};
    t.len.assign(1, 0.0);
    t.centre.assign(d, 0.0);
    finish_box(t, 0);
    t.nbr_ptr = {0, 0};
    t.il_ptr = {0, 0};
    t.nbr.clear();
    t.il.clear();

    std::vector<int64_t> scratch(t.n);
    std::vector<int> side(t.n);
    int64_t biggest = t.n;
    int l = 0;
    while (biggest > t.nmax) {
        const int64_t p0 = t.level_off[l], np = t.level_off[l + 1] - p0;
        const int64_t c0 = t.level_off[l + 1], ncnew = np * nch;
        // more than nmax coincident points never split (HDD_tree.hpp:43-69
        // would deepen without end): stop once a level would hold more than
        // max(2^20, 64 n) nodes
        if (ncnew > std::max<int64_t>(int64_t(1) << 20, 64 * t.n))
            throw ArgError{HODLR_E_ARG, "tree level " + std::to_string(l + 1) + " would hold " +
                                            std::to_string(ncnew) + " nodes: more than nmax points share every box "
                                                                    "(coincident points?)"};
        t.level_off.push_back(c0 + ncnew);
        t.start.resize(c0 + ncnew);
        t.count.resize(c0 + ncnew);
        t.lo.resize((c0 + ncnew) * d);
        t.hi.resize((c0 + ncnew) * d);
        t.len.resize(c0 + ncnew);
        t.centre.resize((c0 + ncnew) * d);
        for (int64_t p = 0; p < np; p++) {
            const int64_t gp = p0 + p;
            double mid[kMaxDim];
            for (int k = 0; k < d; k++) mid[k] = 0.5 * (t.lo[gp * d + k] + t.hi[gp * d + k]);
            // child boxes
            for (int64_t q = 0; q < nch; q++) {
                const int64_t gc = c0 + p * nch + q;
                for (int k = 0; k < d; k++) {
                    const bool upper = (q >> (d - 1 - k)) & 1;
                    t.lo[gc * d + k] = upper ? mid[k] : t.lo[gp * d + k];
                    t.hi[gc * d + k] = upper ? t.hi[gp * d + k] : mid[k];
                }
                finish_box(t, gc);
            }
            // stable counting sort of the parent's range into its children
            const int64_t s = t.start[gp], c = t.count[gp];
            int64_t cnt[1 << kMaxDim] = {0};
            for (int64_t a = 0; a < c; a++) {
                const double *x = &t.pts[t.perm[s + a] * d];
                int q = 0;
                for (int k = 0; k < d; k++) q = (q << 1) | (x[k] < mid[k] ? 0 : 1);
                side[s + a] = q;
                cnt[q]++;
            }
            int64_t off = s;
            for (int64_t q = 0; q < nch; q++) {
                const int64_t gc = c0 + p * nch + q;
                t.start[gc] = off;
                t.count[gc] = cnt[q];
                cnt[q] = off;
                off += t.count[gc];
            }
            for (int64_t a = 0; a < c; a++) scratch[cnt[side[s + a]]++] = t.perm[s + a];
            for (int64_t a = 0; a < c; a++) t.perm[s + a] = scratch[s + a];
        }
        // lists of the new level
        for (int64_t j = 0; j < ncnew; j++) {
            const int64_t g = c0 + j, par = j / nch, gpar = p0 + par;
            auto classify = [&](int64_t other) {
                if (is_far(t, g, c0 + other)) t.il.push_back(other);
                else t.nbr.push_back(other);
            };
            for (int64_t q = 0; q < nch; q++) {
                const int64_t sib = par * nch + q;
                if (sib != j) classify(sib);
            }
            for (int64_t a = t.nbr_ptr[gpar]; a < t.nbr_ptr[gpar + 1]; a++)
                for (int64_t q = 0; q < nch; q++) classify(t.nbr[a] * nch + q);
            t.nbr_ptr.push_back(static_cast<int64_t>(t.nbr.size()));
            t.il_ptr.push_back(static_cast<int64_t>(t.il.size()));
        }
        biggest = 0;
        for (int64_t j = 0; j < ncnew; j++) biggest = t.count[c0 + j] > biggest ? t.count[c0 + j] : biggest;
        l++;
    }
    t.nlevels = l + 1;
}

}  // namespace hodlr
