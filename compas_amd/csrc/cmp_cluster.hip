// Host clustering of refinement tags (compas_amd/regrid.py's step 2): Berger-Rigoutsos on a level's
// tagged blocks. No counterpart in the reference's sources, which hand their TagBoxArray to AMReX's
// clustering (AmrCore::MakeNewGrids with amr.grid_eff); this is the library's own clustering, the
// same cuts in the same order as regrid._split, without its per-box numpy overhead.
#include <array>
#include <cmath>
#include <vector>

#include "../../include/compas_amd.h"

namespace {

struct Blocks {
    const unsigned char* b;
    int dim;
    int n[3];
    long long st[3];
};

// The boxes of the tagged blocks inside [lo, hi] (block indices, numpy axis order), appended to out
// as lo[3], hi[3]: a box at least eff tagged (after shrinking to its tags' bounding box) is kept,
// any other is cut at the tag-free plane nearest its middle, else where the second difference of
// the tag signature changes sign most strongly, else in the middle of its longest side.
void split(const Blocks& B, std::array<int, 3> lo, std::array<int, 3> hi, double eff, std::vector<int>& out)
{
    const int dim = B.dim;
    std::vector<long long> sig[3];
    for (int ax = 0; ax < dim; ++ax) sig[ax].assign((size_t)(hi[ax] - lo[ax] + 1), 0);
    long long count = 0;
    const int h1 = (dim > 1) ? hi[1] : 0, l1 = (dim > 1) ? lo[1] : 0;
    const int h2 = (dim > 2) ? hi[2] : 0, l2 = (dim > 2) ? lo[2] : 0;
    for (int i0 = lo[0]; i0 <= hi[0]; ++i0)
        for (int i1 = l1; i1 <= h1; ++i1)
            for (int i2 = l2; i2 <= h2; ++i2) {
                if (!B.b[i0 * B.st[0] + i1 * B.st[1] + i2 * B.st[2]]) continue;
                ++count;
                ++sig[0][i0 - lo[0]];
                if (dim > 1) ++sig[1][i1 - l1];
                if (dim > 2) ++sig[2][i2 - l2];
            }
    if (count == 0) return;
    // shrink to the tags' bounding box (the planes cut away hold no tags: the signatures of what is
    // left are the same sums, sliced)
    long long size = 1;
    bool single = true;
    for (int ax = 0; ax < dim; ++ax) {
        int a = 0, b = (int)sig[ax].size() - 1;
        while (sig[ax][a] == 0) ++a;
        while (sig[ax][b] == 0) --b;
        sig[ax] = std::vector<long long>(sig[ax].begin() + a, sig[ax].begin() + b + 1);
        hi[ax] = lo[ax] + b;
        lo[ax] = lo[ax] + a;
        size *= hi[ax] - lo[ax] + 1;
        single = single && hi[ax] == lo[ax];
    }
    if ((double)count / (double)size >= eff || single) {
        for (int ax = 0; ax < 3; ++ax) out.push_back(lo[ax]);
        for (int ax = 0; ax < 3; ++ax) out.push_back(hi[ax]);
        return;
    }
    bool have = false, hole = false;
    double best = 0.0;
    int bax = 0, bc = 0;
    for (int ax = 0; ax < dim; ++ax) {
        const int len = (int)sig[ax].size();
        for (int c = 1; c < len - 1; ++c)
            if (sig[ax][c] == 0) {
                const double score = std::fabs((double)c - (double)(len - 1) / 2.0);
                if (!have || score < best) {
                    have = true;
                    best = score;
                    bax = ax;
                    bc = c;
                    hole = true;
                }
            }
    }
    if (!have)
        for (int ax = 0; ax < dim; ++ax) {
            const int len = (int)sig[ax].size();
            if (len < 4) continue;
            std::vector<double> lap((size_t)len - 2);
            for (int c = 0; c < len - 2; ++c)
                lap[c] = (double)sig[ax][c] - 2.0 * (double)sig[ax][c + 1] + (double)sig[ax][c + 2];
            for (int c = 0; c + 1 < len - 2; ++c)
                if (lap[c] * lap[c + 1] < 0.0) {
                    const double jump = std::fabs(lap[c] - lap[c + 1]);
                    if (!have || -jump < best) {
                        have = true;
                        best = -jump;
                        bax = ax;
                        bc = c + 2;
                    }
                }
        }
    if (!have) {
        for (int ax = 1; ax < dim; ++ax)
            if (hi[ax] - lo[ax] > hi[bax] - lo[bax]) bax = ax;
        bc = (hi[bax] - lo[bax] + 1) / 2;
    }
    std::array<int, 3> a_hi = hi, b_lo = lo;
    a_hi[bax] = lo[bax] + bc - 1;
    b_lo[bax] = lo[bax] + bc + (hole ? 1 : 0);
    split(B, lo, a_hi, eff, out);
    split(B, b_lo, hi, eff, out);
}

}  // namespace

extern "C" int cmp_cluster_blocks(const unsigned char* blocks, int dim, const int* shape, double grid_eff,
                                  int max_boxes, int* boxes, int* nboxes)
{
    if (!blocks || !shape || !boxes || !nboxes || dim < 1 || dim > 3 || max_boxes < 0) return CMP_ERR_ARG;
    Blocks B{blocks, dim, {1, 1, 1}, {0, 0, 0}};
    for (int d = 0; d < dim; ++d) {
        if (shape[d] < 1) return CMP_ERR_ARG;
        B.n[d] = shape[d];
    }
    long long st = 1;
    for (int d = dim - 1; d >= 0; --d) {
        B.st[d] = st;
        st *= B.n[d];
    }
    std::vector<int> out;
    std::array<int, 3> lo{0, 0, 0}, hi{0, 0, 0};
    for (int d = 0; d < dim; ++d) hi[d] = B.n[d] - 1;
    split(B, lo, hi, grid_eff, out);
    const int n = (int)(out.size() / 6);
    if (n > max_boxes) return CMP_ERR_ARG;
    for (size_t q = 0; q < out.size(); ++q) boxes[q] = out[q];
    *nboxes = n;
    return CMP_OK;
}
