// Host symbolic analysis for the supernodal Cholesky behind the ISpdSolver slot.
//
// Input is the full symmetric CSC pattern the reference hands to ISpdSolver.Analyze
// (src/ImpressionFEA.Core/Solver/Linear/ISpdSolver.cs:28-32). Pipeline:
//   1. compress columns with identical structure into blocks (the 3 DOFs of a node),
//   2. order the block graph (nested dissection, or a caller permutation),
//   3. elimination tree, postorder, column counts on the block graph, expanded to scalars,
//   4. fundamental supernodes + relaxed amalgamation, supernode row structures,
//   5. level schedule, panel/update storage offsets, the A->panel scatter map and extend-add maps.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace ifea {

struct SymbolicOptions {
  int leaf_blocks = 32;        // nested dissection stops at this many blocks
  int relax_small = 16;        // amalgamate freely up to this many columns
  double relax_zeros = 0.10;   // ... and up to 128 columns while explicit zeros stay below this fraction
  int relax_max = 192;
  // optional coordinates (3 per original column): nested dissection then cuts planes through the blocks' positions
  const double* coords = nullptr;
};

struct Symbolic {
  int n = 0;
  // ordering (pivot k eliminates original column perm[k])
  std::vector<int> perm, iperm;
  // scalar elimination tree and column counts of L (diagonal included) in pivot order
  std::vector<int> parent, colcount;
  long long nnzL = 0;                 // exact structural non-zeros of L
  double flops = 0;                   // sum of colcount^2 (the factor flop count)
  // block compression (blocks in pivot order)
  int nblocks = 0;
  std::vector<int> block_size;        // per original block
  // supernodes (pivot order, contiguous columns)
  int ns = 0;
  std::vector<int> sn_first;          // [ns+1] first pivot column
  std::vector<int> sn_parent;         // [ns] parent supernode or -1
  std::vector<long long> sn_rowptr;   // [ns+1] offsets into sn_rows
  std::vector<int> sn_rows;           // pivot-order row indices; the first k_s are the supernode's own columns
  std::vector<int> sn_level;          // height above the leaves
  std::vector<int> col_to_sn;         // [n]
  // storage: L panel (m x k column-major) and update block ((m-k)^2 column-major)
  std::vector<long long> Loff, Uoff, Voff;  // Voff: forward-solve update vector offset (m - k)
  long long Lsize = 0, Usize = 0, Vsize = 0;
  long long Usum = 0;                 // sum of all update blocks (storage without reuse)
  // extend-add: for child c, rel index of each of its update rows in the parent's row list
  std::vector<long long> rel_off;     // [ns+1]
  std::vector<int> rel;               // positions in parent's row list
  // A -> L scatter: full-CSC position -> L storage offset (lower part of P A P^T only)
  std::vector<int> amap_src;
  std::vector<long long> amap_dst;
  // diagonal positions (full CSC index of A(j,j)) of each original column
  int max_level = 0;
  std::string error;
};

// Returns 0, or a negative code (-1 bad input, -2 bad permutation).
int symbolic_analyze(int n, const int* colPtr, const int* rowIdx, const int* userPerm, const SymbolicOptions& opt,
                     Symbolic& S);

}  // namespace ifea
