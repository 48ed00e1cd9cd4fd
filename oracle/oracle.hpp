// CPU FP64 restatement of the TwoPhaseAMR.jl pressure-projection path.
//
// TEST INFRASTRUCTURE ONLY. This code is the checker: tests/, the smoke
// check and bench.py's cpu_baseline leg may load it; the product library
// (twophaseamr.jl_amd/) never links, calls or reads anything here.
//
// Each routine cites the reference file:line it restates. Parity pins (the
// reference's published numbers, known answers and invariants) are exercised
// by tests/test_oracle_*.py; functions without such a pin are marked
// "parity unpinned" in their comment.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <unordered_set>
#include <vector>

namespace orc {

struct Key {
  int level, i, j;
};

inline uint64_t kpack(int level, int i, int j) {
  return (uint64_t(uint32_t(level)) << 56) | (uint64_t(uint32_t(i)) << 28) |
         uint64_t(uint32_t(j));
}

// LeafBlock, src/mesh/topology/mesh.jl:13-19
struct Block {
  Key key;
  double origin[2], extent[2], spacing[2], volume;
};

// InteriorFace, mesh.jl:32-38 (cells as 0-based global indices,
// mesh.jl:28-30 `_cell_index`)
struct Face {
  int minus, plus, axis;  // axis 0 = x, 1 = y
  double center[2], area;
};

// BoundaryFace, mesh.jl:40-46
struct BFace {
  int cell, axis, side;
  double center[2], area;
};

// BlockTreeMesh, mesh.jl:226-246 (only what the pressure path reads)
struct Mesh {
  int base[2], bs[2], layout[2];
  double origin[2], extent[2], root_extent[2];
  bool periodic[2];
  int max_level;  // refinement ceiling
  std::unordered_set<uint64_t> internal;
  std::vector<Block> blocks;
  std::vector<Face> faces;
  std::vector<BFace> bfaces;
  int ncells() const { return int(blocks.size()) * bs[0] * bs[1]; }
  int max_leaf_level() const;
};

// SparseMatrixCSC with 0-based indices.
struct CSC {
  int m = 0, n = 0;
  std::vector<int> colptr, rowval;
  std::vector<double> nzval;
  int nnz() const { return int(rowval.size()); }
};

struct Stats {
  int cycles, converged, compatible, pad;
  double initial_l2, initial_linf, final_l2, final_linf, target_linf,
      residual_scale, removed_rhs_mean;
};

std::unique_ptr<Mesh> build_mesh(const int base[2], const int bs[2],
                                 const double origin[2], const double extent[2],
                                 const bool periodic[2], int max_level,
                                 const std::vector<Key>& refined);

CSC inverse_metric(const Mesh& mesh, bool* diagonal);

// fill_halo! (src/mesh/variables.jl:205-229) with the halo schedule of
// mesh.jl:944-1037; returns the number of physical halo entries
int fill_halo(const Mesh& M, int halo, const double* cells, const double* physical_values, double* out);

struct Components {
  std::vector<int> labels;  // root cell index, -1 for inactive
  std::vector<int> anchors;  // ungauged component roots (increasing)
  std::vector<int> slots;    // per cell: slot+1 of ungauged component, 0 otherwise
  std::vector<double> volumes;  // per slot
};

struct Operator {
  const Mesh* mesh;
  std::unique_ptr<Mesh> owned_mesh;
  CSC K;
  bool K_diagonal;
  std::vector<double> kappa, sqrt_kappa, volumes;
  CSC A;  // assembled symmetric stiffness (pressure: matrix == stiffness)
  // cell-face incidence in increasing face order (a face's minus entry
  // before its plus entry): the order a face loop scatters to a cell
  std::vector<int> inc_ptr, inc_face;
  std::vector<signed char> inc_sign;
  // open (homogeneous Dirichlet) pressure sides left, right, bottom, top;
  // per boundary face its coefficient and gradient factor 2 k_b / h, per
  // cell the boundary diagonal sum |f| 2 k_b / h over its open faces
  // (elliptic.jl:76-152)
  bool open[4] = {false, false, false, false};
  std::vector<double> bcoef, bgrad, bdiag;
  Components comp;
  bool singular;
};

struct Level {
  Operator* op;
  std::vector<double> invdiag;
  double omega;
  CSC P;  // prolongation to this level from the next coarser one (fine x coarse)
  std::vector<int> comp_index;  // per cell: 1-based component slot or 0
  std::vector<double> comp_volumes;
  std::vector<double> x, b, r, t;
  // row form of P (built on first use): per fine row, its entries in
  // increasing column order
  std::vector<int> prow_ptr, prow_col, prow_ent;
};

struct Solver {
  std::vector<std::unique_ptr<Operator>> ops;  // ops[0] is the fine operator
  std::vector<Level> levels;
  int pre, post, coarse_size, max_levels;
  // coarse dense LU (row-major, partial pivoting)
  int coarse_n = 0;
  bool coarse_augmented = false;
  std::vector<double> lu;
  std::vector<int> piv;
  // fine-level factorized action data
  std::vector<int> cf_offsets, cf_faces, cf_signs;
  bool use_mg = true;
};

}  // namespace orc
