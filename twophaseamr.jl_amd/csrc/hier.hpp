// Host-side topology and hierarchy construction for the GPU pressure solver.
//
// The library receives a block quadtree (the refined node set of a
// BlockTreeMesh) and per-face coefficients in the reference's face order,
// and builds, per multigrid level, the block-structured data the HIP
// kernels consume: regular 5-point conductances, compact CSR rows for
// the cells next to coarse-fine interfaces, block neighbour and transfer
// descriptors, and the dense inverse of the coarsest augmented operator.
#pragma once
#include <cstdint>
#include <memory>
#include <utility>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace tpamr {

// Allocator whose value-less construct leaves elements default-initialised:
// resizing a large array does not zero it on one thread, and the threads that
// fill it touch its pages first.
template <typename T>
struct NoInit : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInit<U>;
  };
  NoInit() = default;
  template <typename U>
  NoInit(const NoInit<U>&) noexcept {}
  template <typename U, typename... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
  template <typename U>
  void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
    ::new (static_cast<void*>(p)) U;
  }
};
template <typename T>
using hvec = std::vector<T, NoInit<T>>;

struct BlockKey {
  int level, i, j;
};

inline uint64_t key64(int level, int i, int j) {
  return (uint64_t(uint32_t(level)) << 56) | (uint64_t(uint32_t(i)) << 28) | uint64_t(uint32_t(j));
}

// One mesh of the hierarchy (a BlockTreeMesh, src/mesh/topology/mesh.jl).
struct TreeMesh {
  int base[2], bs[2], layout[2];
  double origin[2], extent[2], root_extent[2];
  bool periodic[2];
  int max_level;  // refinement ceiling
  int max_leaf = 0;
  std::unordered_set<uint64_t> internal;  // refined nodes
  std::vector<BlockKey> blocks;           // leaves sorted by (level, j, i)
  std::unordered_map<uint64_t, int> block_of;
  std::vector<double> bspacing;  // 2 per block
  std::vector<double> borigin;   // 2 per block: lower-left corner
  // interior faces, reference order
  hvec<int> fminus, fplus;
  hvec<int8_t> faxis;
  hvec<double> farea, fcx, fcy;
  // boundary faces, reference order
  hvec<int> bcell;
  hvec<int8_t> baxis, bside;
  hvec<double> barea;
  int nb() const { return bs[0] * bs[1]; }
  int ncells() const { return int(blocks.size()) * nb(); }
  int nfaces() const { return int(fminus.size()); }
  int block_level(int cell) const { return blocks[cell / nb()].level; }
};

std::unique_ptr<TreeMesh> make_tree_mesh(const int base[2], const int bs[2], const double origin[2],
                                         const double extent[2], const bool periodic[2], int max_level,
                                         const std::vector<BlockKey>& refined);

// Inverse face metric K: diagonal plus symmetric off-diagonal pairs.
struct FaceMetric {
  hvec<double> diag;                      // per face
  std::vector<int> off_ptr, off_face;     // CSR of off-diagonal entries by face
  std::vector<double> off_val;
  bool diagonal() const { return off_face.empty(); }
};

FaceMetric face_metric(const TreeMesh& m);

// One multigrid level in device layout.
struct LevelData {
  std::unique_ptr<TreeMesh> owned;
  const TreeMesh* mesh = nullptr;
  FaceMetric K;
  int nblocks = 0, ncells = 0;
  std::vector<double> kappa;  // per face
  hvec<uint8_t> face_regular;
  // face -> conductance slots: x faces in cx[(bx+1)*by] per block, y faces in cy[bx*(by+1)]
  hvec<int64_t> slot_a, slot_b;  // flat indices (or -1)
  std::vector<double> cx, cy;           // conductances (0 where unused)
  // aligned device layout of the same conductances: west face of every cell
  // (nb per block), east face of the last column (by per block), south face
  // of every cell (nb per block), north face of the top row (bx per block)
  std::vector<double> cw, ce, cs, cn;
  std::vector<int> nbr;                 // 4 per block: W, E, S, N same-level neighbour or -1
  std::vector<double> volume;           // per block
  // irregular rows, grouped by block
  std::vector<int> blk_irr;             // nblocks + 1 row offsets
  std::vector<int> irr_local;           // local cell index per row
  std::vector<int> irr_ptr;             // row -> entries
  std::vector<int> irr_col;
  std::vector<double> irr_val, irr_bound;  // A values; factorized floor-bound values
  std::vector<double> irr_diag;
  std::vector<int> cell_irr;  // per cell: irregular row index or -1
  std::vector<double> dinv;   // 1 / diag(A)
  // Device form of the irregular rows. A row of A splits into the
  // conductances of the cell's regular (and open boundary) faces, which its
  // stencil slots hold like any cell's, and the couplings of its non-regular
  // faces (coarse-fine faces and faces with off-diagonal metric terms). The
  // ELL keeps the off-diagonal entries of that second part only: per block,
  // the rows ordered by decreasing length with their entries in groups of
  // four (entry 4g+t of the block's row x at (blk_irr[b] * ng + g * nr + x) *
  // 4 + t, ng = ceil(ell_w / 4)); ell_loc[blk_irr[b] + x] holds local cell |
  // length << 16. Columns are LDS slots of the block's tile: interior cells,
  // the W, E, S, N halo strips, then the block's external cells (build_ell).
  // ell_diag is the full diagonal of A (the Jacobi diagonal), ell_bdiag the
  // non-regular faces' part of the floor-bound diagonal.
  int ell_w = 0;
  int ext_max = 0;  // most external cells of one block
  std::vector<uint32_t> ell_loc;
  std::vector<double> ell_diag, ell_bdiag;  // per row, same order as ell_loc
  std::vector<uint16_t> ell_ref;
  std::vector<double> ell_val, ell_bound;
  // per CSR entry: whether a term comes from a non-regular face (its
  // off-diagonal entries are the ELL's); ell_dpos: for a row's diagonal entry
  // its row's position in ell_loc order (-1 for the others)
  std::vector<uint8_t> ell_part;
  std::vector<int> ell_dpos;
  std::vector<int> blk_ext, ext_col;
  // structure of the value refresh (level_pattern, built once per topology):
  // per irregular entry its face-metric terms [term_ptr[e], term_ptr[e+1])
  // (faces f, g, K entry, sign, diagonal flag) and its ELL value position;
  // per row its diagonal entry and its position in ell_loc order
  bool planned = false;
  hvec<int> term_ptr, term_f, term_g;
  hvec<double> term_k, term_s;
  hvec<uint8_t> term_diag;
  std::vector<int> ell_pos, ell_row, irr_diag_e;
  double omega = 1, omega_p = 0;
  // transfer to the next coarser level (absent on the coarsest)
  std::vector<int> xfer;  // 2 per block: coarse block, mode (-1 identity, 0..3 quadrant)
  // face-coefficient restriction from the previous (finer) level
  std::vector<int> rptr;
  hvec<int> ridx;
  hvec<double> rw;
  // Open (homogeneous Dirichlet) pressure sides, left / right / bottom / top
  // (PoissonOperator's velocity_boundary, elliptic.jl:76-88). Per boundary
  // face its coefficient kb (the finest level: the caller's; coarser ones:
  // the overlap-weighted restriction brptr/bridx/brw of the finer level's);
  // the open faces (bopen, boundary-face order) with their conductance
  // |f| 2 kb / h (bw, the boundary diagonal term of elliptic.jl:133-152) in
  // the cell's conductance slot beyond that side (bslot: flat index into cx
  // (x faces) or cy (y faces); bdst: the device code index * 4 + array of the
  // regular faces): the halo value there is zero. irr_bptr / irr_bface: each
  // irregular row's open faces; irr_bd: their conductances summed in face
  // order, which the row's full diagonal includes. A level with an open face
  // of positive conductance is nonsingular.
  bool open[4] = {false, false, false, false};
  std::vector<double> kb;
  std::vector<int> brptr, bridx;
  std::vector<double> brw;
  std::vector<int> bopen, bdst, irr_bptr, irr_bface;
  std::vector<int64_t> bslot;
  std::vector<double> bw, irr_bd;
  // Pressure components of the level (`_refresh_elliptic_components!`,
  // elliptic_solver.jl:39-109): connected over faces with a nonzero
  // coefficient; a component with an open face of positive conductance is
  // gauged, the others (ncomp, slots 1..ncomp in increasing root order) are
  // projected and gauged each on its own. connected: one component holding
  // every cell. comp_slot per cell (0: gauged), comp_root (0-based smallest
  // cell) and comp_vol (cell volumes summed in cell order) per slot. The
  // component sums of a level with several components run in chunks of
  // GAUGE_CHUNK cells (`_chunked_slot_sums!`, threading.jl:202-283): one
  // partial per (chunk, slot) pair present, chunk-major (pair_chunk,
  // pair_slot), and per slot its pairs in chunk order (slot_pair_ptr /
  // slot_pairs). total_volume: the level's cell volumes (per block).
  int ncomp = 1;
  bool connected = true;
  double total_volume = 0;
  std::vector<int> comp_slot, comp_root, pair_chunk, pair_slot, slot_pair_ptr, slot_pairs;
  std::vector<double> comp_vol;
  bool singular = true;
  bool any_open() const { return open[0] || open[1] || open[2] || open[3]; }
  // one ungauged component holding every cell, or none: no per-level projection
  bool simple_gauge() const { return ncomp == 0 || (ncomp == 1 && connected); }
};

// Cells per chunk of the component sums.
constexpr int GAUGE_CHUNK = 1024;

struct Hierarchy {
  std::vector<std::unique_ptr<LevelData>> levels;
  // coarsest level: the dense operator, augmented [A V; V^T 0] ((n + 1) x
  // (n + 1)) when the levels are singular, A itself (n x n) when open sides
  // anchor them; row major, coarse_N rows; the device inverts it
  int coarse_n = 0, coarse_N = 0;
  std::vector<double> coarse_aug;
  // the finest level singular (an ungauged component): the solve projects
  // the rhs and removes volume means; otherwise no mean is removed.
  // general_gauge: some level has several components or gauged cells beside
  // an ungauged component, so the V-cycle projects and gauges every such
  // level's components (cycles.jl:8-77)
  bool singular = true, general_gauge = false;
  double total_volume = 0;
  // levels taken over from the previous hierarchy (_tree_reuse_suffix!):
  // this hierarchy's levels from reused_new on are the previous one's from
  // reused_old on (-1: none)
  int reused_new = -1, reused_old = -1;
};

// Halo refresh schedule of a haloed cell field (fill_halo!, halo width h):
// per block the (bx + 2h) x (by + 2h) array with the x index fastest; every
// halo position is an entry with donor cells and weights (ptr / donor /
// weight), or a physical-boundary entry (one donor: the nearest interior
// cell; phys = its rank among the physical entries, -1 otherwise). Entries
// are in the reference's schedule order.
struct HaloSchedule {
  int halo = 0;
  std::vector<int> dest;  // offset in the haloed output
  std::vector<int> ptr, donor;
  std::vector<double> weight;
  std::vector<int> phys;
  int nphysical = 0;
};

HaloSchedule halo_schedule(const TreeMesh& m, int halo);

// Largest coarsest level (cells) of the dense coarse solve: its augmented
// inverse is 8.6 GB of FP64 at this size (N^2 stays below 2^31), built by a
// blocked Gauss-Jordan on the device and applied as one matrix-vector
// product per V-cycle.
constexpr int COARSE_MAX = 32768;

struct MultigridConfig {
  int coarse_size = 64, pre = 2, post = 2, max_levels = 32;
};

// Multigrid hierarchy of a tree mesh (TreeGeometricMultigridPreconditioner,
// geometric_preconditioner.jl:171-442). With `previous` (a hierarchy of an
// earlier mesh built with the same options), the first coarse mesh whose
// tree topology equals one of its levels ends the coarsening: that level and
// the coarser ones are taken over from it (_tree_reuse_suffix!,
// tree_coarsening.jl:214-226), and `previous` loses them.
// `open` (NULL: none) are the open pressure sides (left, right, bottom, top);
// they shape the irregular rows, so a change of open sides needs a new
// hierarchy, and `previous` must have been built with the same sides.
std::unique_ptr<Hierarchy> build_hierarchy(std::unique_ptr<TreeMesh> fine, const MultigridConfig& cfg,
                                           Hierarchy* previous = nullptr, const bool* open = nullptr);
// Boundary coefficients of the finest level's boundary faces (one per
// boundary face, reference order): restricted to every level, with each
// level's open-face conductances, row boundary diagonals and singular flags;
// the next refresh_coefficients / level value pass uses them.
void set_boundary_coefficients(Hierarchy& H, const double* kb);
// `_boundary_material` (elliptic.jl:90-131) for interior face coefficients:
// every open boundary face of `m` takes the area-weighted mean of the
// interior coefficients on the opposite side of its cell along its axis;
// closed faces 0.
std::vector<double> boundary_material(const TreeMesh& m, const bool open[4], const double* kappa);
// _same_tree_topology (tree_coarsening.jl:84-121): domain, block size and
// leaf keys (faces and geometry follow from them)
bool same_tree_topology(const TreeMesh& a, const TreeMesh& b);
// Refresh all coefficient-dependent values for new fine face coefficients
// (nonnegative; zeros split the domain into components).
void refresh_coefficients(Hierarchy& H, const double* kappa);
// The components of every level for positive face coefficients on every
// level (one connected component; gauged by open faces), and the
// hierarchy's gauge flags: the state after a device refresh.
void connected_components(Hierarchy& H);
// Values of the coarsest level alone from its own face coefficients, and its
// dense augmented matrix (the device refreshes the other levels).
void refresh_coarsest(Hierarchy& H, const double* kappa_coarsest);

}  // namespace tpamr
