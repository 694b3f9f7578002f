// tally_box.h -- choice of the LDS tally tile from a pilot count of deposits.
//
// A pilot run tallies how often photons deposit into each (gate, voxel) cell. The tile is the
// axis-aligned (gate, x, y, z) box, within a cell budget, that greedy growth from the densest
// cell finds to hold most of those deposits; deposits inside it accumulate in LDS and reach HBM
// once per workgroup.
#pragma once
#include <array>
#include <cstdint>
#include <vector>

namespace umcx {

// {x0, y0, z0, g0, sx, sy, sz, sg}; sx == 0 when the grid holds no deposit
using TallyBox = std::array<int, 8>;

// cnt: nt x nx x ny x nz counts, memory order [t][x][y][z]
TallyBox choose_box(const float* cnt, int nx, int ny, int nz, int nt, uint64_t budget);

struct CountGrid {
    const float* cnt;
    int nx, ny, nz, nt;
};
// choose_box for many grids at once (OpenMP over grids)
std::vector<TallyBox> choose_boxes(const std::vector<CountGrid>& grids, uint64_t budget);

// counts of each grid, and counts of each grid inside its box (OpenMP over grids; null grids: 0)
std::vector<double> grid_totals(const std::vector<CountGrid>& grids);
std::vector<double> box_counts(const std::vector<CountGrid>& grids, const std::vector<TallyBox>& boxes);

// counts inside a box, and the fraction of a grid's counts that fall inside it
double box_count(const CountGrid& g, const TallyBox& b);
double box_share(const CountGrid& g, const TallyBox& b);

}  // namespace umcx
