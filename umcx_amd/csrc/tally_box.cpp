// tally_box.cpp -- greedy growth of the LDS tally tile (tally_box.h).
#include "tally_box.h"

namespace umcx {

namespace {

struct View {
    const float* cnt;
    uint64_t nyz, nxyz;
    int nz;
    double at(int g, int x, int y, int z) const {
        return cnt[(uint64_t)g * nxyz + (uint64_t)x * nyz + (uint64_t)y * nz + z];
    }
};

// sum over the box [l, h) (dims g, x, y, z)
double box_sum(const View& v, const int* l, const int* h) {
    double s = 0;
    for (int g = l[0]; g < h[0]; g++)
        for (int x = l[1]; x < h[1]; x++)
            for (int y = l[2]; y < h[2]; y++)
                for (int z = l[3]; z < h[3]; z++) s += v.at(g, x, y, z);
    return s;
}

uint64_t cells(const int* l, const int* h) {
    uint64_t c = 1;
    for (int d = 0; d < 4; d++) c *= (uint64_t)(h[d] - l[d]);
    return c;
}

}  // namespace

// Start at the densest cell; at each step add the face slab (one of 8: two sides of 4 dims)
// with the most counts per added cell, while the box holds at most `budget` cells.
TallyBox choose_box(const float* cnt, int nx, int ny, int nz, int nt, uint64_t budget) {
    const uint64_t nyz = (uint64_t)ny * nz, nxyz = nyz * nx, n = nxyz * (uint64_t)nt;
    TallyBox none{0, 0, 0, 0, 0, 0, 0, 0};
    if (n == 0 || budget == 0) return none;
    uint64_t best = 0;
    for (uint64_t i = 1; i < n; i++)
        if (cnt[i] > cnt[best]) best = i;
    if (!(cnt[best] > 0.f)) return none;
    const View v{cnt, nyz, nxyz, nz};
    int lo[4], hi[4];
    lo[0] = (int)(best / nxyz);
    uint64_t r = best % nxyz;
    lo[1] = (int)(r / nyz);
    r %= nyz;
    lo[2] = (int)(r / nz);
    lo[3] = (int)(r % nz);
    for (int d = 0; d < 4; d++) hi[d] = lo[d] + 1;
    const int ext[4] = {nt, nx, ny, nz};
    while (true) {
        int bd = -1, bside = 0;
        double bscore = 0;
        const uint64_t cur = cells(lo, hi);
        for (int d = 0; d < 4; d++)
            for (int side = 0; side < 2; side++) {
                const int idx = side ? hi[d] : lo[d] - 1;
                if (idx < 0 || idx >= ext[d]) continue;
                int l[4], h[4];
                for (int k = 0; k < 4; k++) { l[k] = lo[k]; h[k] = hi[k]; }
                if (side) h[d]++; else l[d]--;
                const uint64_t nc = cells(l, h);
                if (nc > budget) continue;
                l[d] = idx;
                h[d] = idx + 1;
                const double add = box_sum(v, l, h);
                const double score = add / (double)(nc - cur);
                if (add > 0 && score > bscore) {
                    bscore = score;
                    bd = d;
                    bside = side;
                }
            }
        if (bd < 0) break;
        if (bside) hi[bd]++; else lo[bd]--;
    }
    return {lo[1], lo[2], lo[3], lo[0], hi[1] - lo[1], hi[2] - lo[2], hi[3] - lo[3], hi[0] - lo[0]};
}

std::vector<TallyBox> choose_boxes(const std::vector<CountGrid>& grids, uint64_t budget) {
    std::vector<TallyBox> out(grids.size());
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < (long)grids.size(); i++) {
        const CountGrid& g = grids[i];
        out[i] = choose_box(g.cnt, g.nx, g.ny, g.nz, g.nt, budget);
    }
    return out;
}

std::vector<double> grid_totals(const std::vector<CountGrid>& grids) {
    std::vector<double> out(grids.size(), 0.0);
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < (long)grids.size(); i++) {
        const CountGrid& g = grids[i];
        if (g.cnt) out[i] = box_count(g, TallyBox{0, 0, 0, 0, g.nx, g.ny, g.nz, g.nt});
    }
    return out;
}

std::vector<double> box_counts(const std::vector<CountGrid>& grids, const std::vector<TallyBox>& boxes) {
    std::vector<double> out(grids.size(), 0.0);
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < (long)grids.size(); i++)
        if (grids[i].cnt) out[i] = box_count(grids[i], boxes[i]);
    return out;
}

double box_count(const CountGrid& g, const TallyBox& b) {
    if (b[4] == 0) return 0.0;
    const uint64_t nyz = (uint64_t)g.ny * g.nz, nxyz = nyz * g.nx;
    const View v{g.cnt, nyz, nxyz, g.nz};
    const int l[4] = {b[3], b[0], b[1], b[2]}, h[4] = {b[3] + b[7], b[0] + b[4], b[1] + b[5], b[2] + b[6]};
    return box_sum(v, l, h);
}

double box_share(const CountGrid& g, const TallyBox& b) {
    const double tot = box_count(g, TallyBox{0, 0, 0, 0, g.nx, g.ny, g.nz, g.nt});
    if (!(tot > 0) || b[4] == 0) return 0.0;
    return box_count(g, b) / tot;
}

}  // namespace umcx
