// ogs::ogs_t on one rank: T maps, row blocks, GlobalToLocal.
//
// Reference behaviour (libs/ogs/ogsSetup.cpp:87-290, 492-579, 833-849;
// libs/ogs/ogsOperator.cpp:474-617):
//  - ids equal to 0 are dropped (masked);
//  - nodes with the same |id| form one gather row; rows are numbered by the
//    first appearance of the group in local order (with unique=true every
//    group has a positive member, so all groups are "positive" groups);
//  - each row lists its local nodes in ascending order (colIdsT), rows
//    delimited by rowStartsT;
//  - rows are cut into blocks of at most 512 entries by recursive bisection
//    at the first row start past the block's midpoint (blockRowStartsT);
//    a row longer than 512 entries is an error;
//  - GlobalToLocal[n] = row of local node n, or -1 when masked.
// Construction here is a single sort of (id, local index) pairs and prefix
// sums, parallel with OpenMP.
#include <omp.h>

#include <algorithm>
#include <numeric>
#include <parallel/algorithm>

#include "bpx.hpp"

namespace bpx {
namespace ogs {

static constexpr dlong kBlockNnz = 512;  // include/ogs/ogsOperator.hpp:38

// Recursive bisection of [0, Nrows) into blocks with <= kBlockNnz entries.
// Each pass halves every oversized block at the first row whose start lies
// past start + ceil(size/2) (keeping at least one row on the right), which
// reproduces the reference's level-by-level bisection exactly.
static int block_rows(dlong Nrows, const std::vector<dlong>& rowStarts, std::vector<dlong>& blocks) {
  blocks.clear();
  if (Nrows == 0) return BPX_OK;
  for (dlong r = 0; r < Nrows; ++r)
    if (rowStarts[r + 1] - rowStarts[r] > kBlockNnz)
      return set_error(BPX_ERR_OGS, "ogs: a global node has multiplicity above 512");
  std::vector<dlong> cur = {0, Nrows}, next;
  bool split = true;
  while (split) {
    split = false;
    next.clear();
    next.push_back(0);
    for (size_t b = 0; b + 1 < cur.size(); ++b) {
      const dlong s = cur[b], e = cur[b + 1];
      const dlong size = rowStarts[e] - rowStarts[s];
      if (size > kBlockNnz) {
        const dlong target = rowStarts[s] + (size + 1) / 2;
        // first row index in (s, e] whose start exceeds target
        dlong mid = (dlong)(std::upper_bound(rowStarts.begin() + s, rowStarts.begin() + e, target) -
                            rowStarts.begin());
        if (mid == e) --mid;
        if (mid > s && mid != next.back()) next.push_back(mid);
        split = true;
      }
      if (e != next.back()) next.push_back(e);
    }
    cur.swap(next);
  }
  blocks = cur;
  return BPX_OK;
}

int ogs_t::Setup(dlong N_, const hlong* ids) {
  if (N_ < 0) return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: N < 0");
  N = N_;
  // (|id|, local index) of the unmasked nodes
  std::vector<std::pair<hlong, dlong>> key;
  key.reserve(N);
  for (dlong n = 0; n < N; ++n)
    if (ids[n] != 0) key.emplace_back(ids[n] < 0 ? -ids[n] : ids[n], n);
  __gnu_parallel::sort(key.begin(), key.end());
  const dlong nnz = (dlong)key.size();
  nnzT = nnz;

  // group boundaries in sorted order; the first member is the smallest index
  std::vector<dlong> gstart;
  gstart.reserve(nnz / 2 + 1);
  for (dlong i = 0; i < nnz; ++i)
    if (i == 0 || key[i].first != key[i - 1].first) gstart.push_back(i);
  const dlong Ngroups = (dlong)gstart.size();
  gstart.push_back(nnz);

  // row number = rank of the group's first appearance in local order
  std::vector<dlong> first_flag(N, 0);
#pragma omp parallel for
  for (dlong g = 0; g < Ngroups; ++g) first_flag[key[gstart[g]].second] = 1;
  std::vector<dlong> row_at(N, 0);
  {
    dlong acc = 0;
    for (dlong n = 0; n < N; ++n) {
      row_at[n] = acc;
      acc += first_flag[n];
    }
  }
  Ngather = Ngroups;
  rowStartsT.assign(Ngroups + 1, 0);
  std::vector<dlong> group_row(Ngroups);
#pragma omp parallel for
  for (dlong g = 0; g < Ngroups; ++g) {
    const dlong row = row_at[key[gstart[g]].second];
    group_row[g] = row;
    rowStartsT[row + 1] = gstart[g + 1] - gstart[g];
  }
  for (dlong r = 0; r < Ngroups; ++r) rowStartsT[r + 1] += rowStartsT[r];
  colIdsT.assign(nnz, 0);
#pragma omp parallel for
  for (dlong g = 0; g < Ngroups; ++g) {
    dlong o = rowStartsT[group_row[g]];
    for (dlong i = gstart[g]; i < gstart[g + 1]; ++i) colIdsT[o++] = key[i].second;
  }
  int st = block_rows(Ngroups, rowStartsT, blockRowStartsT);
  if (st) return st;
  NrowBlocksT = blockRowStartsT.empty() ? 0 : (dlong)blockRowStartsT.size() - 1;
  return BPX_OK;
}

// Maps with caller-numbered rows: local node n joins row rowOf[n] (or none
// when rowOf[n] < 0); each row lists its nodes in ascending order. Used for
// sums over the node lattice, whose rows are the lattice points in order.
int ogs_t::SetupRows(dlong N_, dlong Nrows, const dlong* rowOf) {
  if (N_ < 0 || Nrows < 0) return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: N or Nrows < 0");
  N = N_;
  Ngather = Nrows;
  rowStartsT.assign((size_t)Nrows + 1, 0);
  for (dlong n = 0; n < N; ++n) {
    if (rowOf[n] >= Nrows) return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: row index out of range");
    if (rowOf[n] >= 0) ++rowStartsT[rowOf[n] + 1];
  }
  for (dlong r = 0; r < Nrows; ++r) rowStartsT[r + 1] += rowStartsT[r];
  nnzT = rowStartsT[Nrows];
  colIdsT.assign(nnzT, 0);
  std::vector<dlong> fill(rowStartsT.begin(), rowStartsT.end() - 1);
  for (dlong n = 0; n < N; ++n)
    if (rowOf[n] >= 0) colIdsT[fill[rowOf[n]]++] = n;
  int st = block_rows(Nrows, rowStartsT, blockRowStartsT);
  if (st) return st;
  NrowBlocksT = blockRowStartsT.empty() ? 0 : (dlong)blockRowStartsT.size() - 1;
  return BPX_OK;
}

void ogs_t::SetupGlobalToLocalMapping(dlong* G2L) const {
  std::fill(G2L, G2L + N, -1);
#pragma omp parallel for
  for (dlong r = 0; r < Ngather; ++r)
    for (dlong g = rowStartsT[r]; g < rowStartsT[r + 1]; ++g) G2L[colIdsT[g]] = r;
}

int ogs_t::Upload(platform_t& p) {
  if (int st = o_rowStartsT.upload(rowStartsT, p.stream)) return st;
  if (int st = o_colIdsT.upload(colIdsT, p.stream)) return st;
  if (int st = o_blockRowStartsT.upload(blockRowStartsT, p.stream)) return st;
  return BPX_OK;
}

template <typename T>
int ogs_t::Gather(platform_t& p, T* o_gv, const T* o_v, int K) const {
  return launch_ogs_gather<T>(p.stream, NrowBlocksT, K, o_blockRowStartsT.ptr(), o_rowStartsT.ptr(),
                              o_colIdsT.ptr(), o_v, o_gv, gatherMode);
}

template <typename T>
int ogs_t::Scatter(platform_t& p, T* o_v, const T* o_gv, int K) const {
  return launch_ogs_scatter<T>(p.stream, NrowBlocksT, K, o_blockRowStartsT.ptr(), o_rowStartsT.ptr(),
                               o_colIdsT.ptr(), o_gv, o_v);
}

int ogs_t::GatherScatter(platform_t& p, dfloat* o_v, int K) const {
  return launch_ogs_gather_scatter<double>(p.stream, NrowBlocksT, K, o_blockRowStartsT.ptr(),
                                           o_rowStartsT.ptr(), o_colIdsT.ptr(), o_v);
}

template int ogs_t::Gather<double>(platform_t&, double*, const double*, int) const;
template int ogs_t::Gather<float>(platform_t&, float*, const float*, int) const;
template int ogs_t::Scatter<double>(platform_t&, double*, const double*, int) const;
template int ogs_t::Scatter<float>(platform_t&, float*, const float*, int) const;

}  // namespace ogs
}  // namespace bpx
