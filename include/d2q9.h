/*
 * d2q9.h -- C ABI of the MI355X (gfx950) D2Q9 lattice Boltzmann engine.
 *
 * One shared library, libd2q9.so, built from pylabolt_amd/csrc/.  Plain C
 * types only: device pointers are void*, the HIP stream is passed as void*
 * (a hipStream_t).  Every entry point returns 0 on success, a negative
 * D2Q9_E* code for a rejected argument, or a positive hipError_t; the text
 * of the last error is available from d2q9_last_error().  Apart from the
 * HIP graph wrappers, no entry point allocates memory: callers pass every
 * buffer, including reduction workspaces.
 *
 * Reference interfaces replaced (file:line in Malyadeep/pylabolt):
 *   d2q9_init_eq      <- parallel/{cpu,gpu}/equilibrium_kernels.py:38-78
 *                        initialize_pop_density_based_second_order, bound by
 *                        base/collision_operator.py:337-374 initialize_pop
 *   d2q9_collide      <- parallel/{cpu,gpu}/compute_fields_kernels.py:5-65,
 *                        force_field_kernels.py:5-25 and
 *                        collision_kernels.py:5-153 (BGK_*), bound by
 *                        base/compute_fields_operator.py, force_operator.py
 *                        and collision_operator.py:289-335 collide
 *   d2q9_stream_bc    <- parallel/{cpu,gpu}/streaming_kernels.py:5-47
 *                        scalar_based_kernel, MPI_kernels.py:34-99 periodic
 *                        wrap, fluid_boundary_kernels.py:5-152, bound by
 *                        base/streaming_operator.py stream,
 *                        parallel/MPI_operator.py:155-589 halo_exchange and
 *                        base/boundary_operator.py:138-202 set_boundary
 *   d2q9_fused_step   <- one whole solvers/fluidLB.py:206-253
 *                        single_time_step (stream + boundaries of step n,
 *                        then moments + collision of step n+1)
 *   d2q9_moments      <- compute_fields_kernels.py:5-65 alone
 *   d2q9_mass_fp64    <- no reference kernel; FP64 mass check
 *   d2q9_momentum_exchange <- parallel/cpu/force_torque_kernels.py:11-141
 *                        (obstacle force/torque and wall force), bound by
 *                        base/obstacle_operator.py:75-113 and
 *                        base/boundary_operator.py:204-294
 *   d2q9_residues     <- parallel/cpu/compute_residues_kernels.py:6-73,
 *                        utils/residues.py:171-317
 *   d2q9_graph_*      <- solvers/fluidLB.py:270-287 (CUDA graph of a step)
 *   d2q9_batched_step <- no reference counterpart (many lattices per launch;
 *                        each lattice's step is d2q9_fused_step)
 *   d2q9_adjoint_step,
 *   d2q9_adjoint_init_grad <- no reference counterpart (reverse-mode step)
 */
#ifndef D2Q9_H
#define D2Q9_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* storage of the populations in HBM */
enum {
  D2Q9_STORE_F64 = 0,     /* double, compute in double (the parity path) */
  D2Q9_STORE_F32 = 1,     /* float of (f_i - w_i) (dev_scale 1), compute float */
  D2Q9_STORE_F16DEV = 2,  /* half of (f_i - w_i) * dev_scale_i, compute float */
  D2Q9_STORE_BF16DEV = 3  /* bfloat16 of (f_i - w_i) (dev_scale 1), compute float */
};
enum { D2Q9_BGK = 0, D2Q9_MRT = 1 };
enum { D2Q9_FORCING_NONE = 0, D2Q9_GUO_LINEAR = 1, D2Q9_GUO_SECOND_ORDER = 2 };
enum { D2Q9_FSRC_GRAVITY = 0, D2Q9_FSRC_FIELD = 1 };

/* per-cell flags (uint8, one per padded cell) */
enum {
  D2Q9_CELL_FLUID = 1,      /* not solid, not ghost: updated every step */
  D2Q9_CELL_SOLID = 2,
  D2Q9_CELL_NEAR_SOLID = 4, /* some pull neighbour is solid */
  D2Q9_CELL_BC = 8,         /* boundary node: bc_slot/bc_nodes valid */
  D2Q9_CELL_STORE_MOM = 16  /* rho/u read by a neighbour next step */
};

/* boundary rule of one direction of one node: kind | (element << 4) */
enum {
  D2Q9_RULE_STREAM = 0,   /* keep the streamed value */
  D2Q9_RULE_BB = 1,       /* f[inv] = f*[out] */
  D2Q9_RULE_VEL = 2,      /* moving wall: f*[out] - 2 w rho cs^-2 c.U */
  D2Q9_RULE_PRES = 3,     /* anti-bounce-back at fixed density */
  D2Q9_RULE_ZG = 4        /* copy from the inward neighbour, streamed */
};

/* error codes (negative) */
enum {
  D2Q9_E_ARG = -1,        /* null pointer, bad enum or size */
  D2Q9_E_TOO_LARGE = -2,  /* workspace too small */
  D2Q9_E_UNSUPPORTED = -3 /* combination not implemented */
};

typedef struct d2q9_bc_elem {
  double vec[2];    /* fixed_velocity wall velocity */
  double scalar;    /* fixed_pressure density */
  int32_t normal[2];/* inward normal */
  int32_t type;     /* D2Q9_RULE_* of this element */
  int32_t pad;
} d2q9_bc_elem;

typedef struct d2q9_bc_node {
  uint16_t rule[9];
  uint16_t pad;
} d2q9_bc_node;

/*
 * One lattice, padded cells (i, j), 0 <= i <= nx+1, 0 <= j <= ny+1.
 *
 * Populations (AoSoA): direction k of cell (i, j) is element
 *   ((i * nbj + jj / B) * 9 + k) * B + jj % B,  jj = j + B - 1,
 * with B = 64 * cpl cells per block; cpl = 1 (FP64), 2 (FP32 and 16-bit),
 * so that one lane of a wave moves 8 bytes (4 bytes in 16-bit storage) per
 * direction.  A buffer holds
 * pop_len = (nx + 2) * nbj * 9 * B elements, nbj >= (ny + B) / B + 1.
 *
 * Per-cell side arrays (flags, bc_slot, moments [rho|ux|uy], force field
 * [fx|fy]) are pitched planes: cell (i, j) is element
 * origin + i * pitch + j of each plane of `plane` elements.
 */
/* Single-precision copies of the scalar constants, filled by the caller
 * from the double fields (the FP32-compute kernels read these). */
typedef struct d2q9_fconst {
  float w[9], rates[9], scale[9], inv_scale[9];
  float omega, inv_cs_2, inv_cs_4, float_min;
  float gravity[2];
} d2q9_fconst;

typedef struct d2q9_lattice {
  int32_t nx, ny, pitch, origin;
  int64_t plane;
  int32_t cpl, nbj;
  int64_t pop_len;
  int32_t storage, collision, forcing, force_source;
  int32_t x_periodic, y_periodic;
  int32_t tile_j, tile_i;       /* launch tile: tile_j == 64 * cpl, tile_i == 4 */
  double omega, inv_cs_2, inv_cs_4, float_min;
  double w[9];
  double gravity[2];
  double rates[9];              /* MRT relaxation rates */
  double dev_scale[9];          /* 16-bit storage scale per direction; a
                                 * power of two keeps decoding exact
                                 * (cf.inv_scale = 1 / scale), else every
                                 * decoded value carries the same relative
                                 * rounding error of inv_scale */
  d2q9_fconst cf;
  const uint8_t* flags;         /* [plane] */
  /* Tile lists: tiles[0 .. n_bulk) are bulk tiles (no solid cell; cells
   * past the lattice are masked), tiles[n_bulk .. n_bulk + n_general)
   * general tiles (holding a solid cell).  Tile t covers rows
   * 1 + (t / ntj) * tile_i .. and cells 1 + (t % ntj) * tile_j ..,
   * ntj = ceil(ny / tile_j).  NULL: every tile is handled as general.  The
   * bulk sweep treats every cell as plain interior fluid; fix_cells lists
   * (as origin + i * pitch + j) the cells of bulk tiles that are not --
   * solid neighbour, boundary rule, stored moment, periodic edge -- and a
   * second pass recomputes them.  A tile in neither list (n_bulk +
   * n_general may be less than the tile count) has its fluid cells in
   * fix_cells; its solid cells keep the encoded zero d2q9_init_eq wrote. */
  const int32_t* tiles;
  int32_t n_bulk, n_general;
  const int32_t* fix_cells;
  int32_t n_fix, pad_fix;
  /* Masked-cell bits of the bulk tiles: for the tile at position p of the
   * bulk list and its row r, words [(p * tile_i + r) * cpl, + cpl) hold one
   * bit per cell of the block (bit b of word w: cell 1 + tj*tile_j + 64w + b).
   * The bulk sweep does not store those cells: the special ones are
   * recomputed by the fix-up pass, the ones past the lattice's last column
   * (partial tiles) are never written. */
  const uint64_t* bulk_masks;
  const int32_t* bc_slot;       /* [plane], read where D2Q9_CELL_BC */
  const d2q9_bc_node* bc_nodes;
  const d2q9_bc_elem* bc_elems;
  const void* force_field;      /* [2][plane] when force_source == FIELD
                                 * (needs a Guo forcing: D2Q9_E_UNSUPPORTED
                                 * with forcing NONE) */
  /* Wall cells swept inside bulk tiles (null: none).  wall_index[p] is the
   * slot of bulk tile p in wall_masks, or -1 when the tile has no such
   * cell.  For slot t and row r, words [((t * tile_i + r) * 6 + s) * cpl,
   * + cpl) mark (one bit per cell, as bulk_masks) the cells whose only
   * boundary rules are those of one straight wall:
   *   s = 0: zero-gradient, directions 2,5,6 from the cell at j + 1,
   *   s = 1: zero-gradient, directions 4,7,8 from the cell at j - 1,
   *          (both as that cell pulled them, before any boundary rule);
   *   s = 2..5: bounce-back on the wall of normal (0,1), (0,-1), (1,0),
   *          (-1,0): directions 2,5,6 / 4,7,8 / 1,5,8 / 3,6,7 take the
   *          cell's own post-collision value of the opposite direction.
   * Those cells are not masked in bulk_masks and not in the fix-up list. */
  const int32_t* wall_index;
  const uint64_t* wall_masks;
} d2q9_lattice;

/* Populations f = feq(rho0, u0) on fluid cells of both buffers, the encoded
 * zero elsewhere.  mom = [rho | ux | uy] planes of the initial fields. */
int d2q9_init_eq(const d2q9_lattice* L, void* f_a, void* f_b, const void* mom,
                 void* stream);

/* Moments + collision only (the first step's pre-collision state is the
 * initial equilibrium): f_dst = C(f_src).  Writes rho/u into mom_dst where
 * D2Q9_CELL_STORE_MOM, or on every fluid cell if write_moments. */
int d2q9_collide(const d2q9_lattice* L, const void* f_src, void* f_dst,
                 const void* mom_src, void* mom_dst, int write_moments,
                 void* stream);

/* Stream + periodic wrap + boundary rules only: f_dst = pop_new. */
int d2q9_stream_bc(const d2q9_lattice* L, const void* f_src, void* f_dst,
                   const void* mom_src, void* stream);

/* The sweeps (d2q9_collide, d2q9_stream_bc, d2q9_fused_step) launch the
 * bulk tiles on `stream` and the general tiles plus fix-up cells on an
 * auxiliary stream, forked from and joined back to `stream` with events
 * (both passes write disjoint cells), so the two overlap; the library
 * creates one auxiliary stream and event pair per caller stream (and
 * device) on first use and keeps them for the process (at most 4096
 * caller streams).  Work is ordered on `stream` exactly as if it ran there
 * alone, including under HIP graph capture, and host threads that drive
 * different streams do not interact; calls on one stream from several
 * threads must be ordered by the caller, as for any work on one stream.
 * Exception: a fused step in FP32 storage, or in FP16/BF16 storage with
 * n_general == 0, is one kernel on `stream`. */

/* Fused: stream + boundaries of step n, moments + collision of step n+1. */
int d2q9_fused_step(const d2q9_lattice* L, const void* f_src, void* f_dst,
                    const void* mom_src, void* mom_dst, int write_moments,
                    void* stream);

/* Two fused steps: f_dst and mom_dst as after d2q9_fused_step(f_src ->
 * f_tmp) and d2q9_fused_step(f_tmp -> f_dst), bitwise in FP64 and FP32
 * storage; in 16-bit storage the deep tiles' intermediate step stays in
 * FP32 (only the stored step is rounded).  It has no counterpart in the reference: it replaces two of its
 * time steps (solvers/fluidLB.py:206-253) by one sweep over the interior.
 * deep_tiles lists tiles of d2q9_deep_rows() rows x tile_j cells (tile t:
 * rows 1 + (t / ntj) * rows .., cells 1 + (t % ntj) * tile_j .., ntj =
 * ceil(ny / tile_j)) whose cells and whose neighbours one cell deep are plain
 * fluid (flag D2Q9_CELL_FLUID only, no periodic edge); the caller
 * guarantees it (a tile whose grown region leaves the lattice is skipped,
 * so a bad list cannot make an access out of bounds).  In a lattice
 * periodic in x with at least 2 d2q9_deep_rows() rows, rows past the first
 * and last are read from their periodic images, so tiles may reach them,
 * and a tile that would pass row nx is moved up to end at row nx (cells of
 * two tiles are written twice with the same bits).  deep_walls (null:
 * none) gives per deep tile the straight wall its grown region may hold at
 * j = 1 (bits 0-1) and at j = ny (bits 2-3): 1 zero-gradient, 2
 * bounce-back, every cell of that column in the grown rows a wall cell of
 * wall_masks sides 0/2 resp. 1/3; the ghost column beside it is then the
 * only cell of the grown region outside the lattice.  Each is swept once for both steps, step n + 1 of the tile
 * grown by one cell held in LDS.  L_rim_a and L_rim_b describe the same
 * lattice as L with tile and fix-up lists restricted to the rest (the rim):
 * L_rim_b covers every cell outside the deep tiles, L_rim_a those cells and
 * also every cell they pull from.  The rim runs as two single steps through
 * f_tmp / mom_tmp (scratch of the sizes of f_src / mom_src) on the
 * auxiliary stream, forked and joined as for d2q9_fused_step; when both rim
 * descriptors list no tile and no fix-up cell, the deep tiles alone run on
 * `stream`, with no fork and join.  All three population buffers and all
 * three moment buffers must differ. */
int d2q9_fused_step2(const d2q9_lattice* L, const d2q9_lattice* L_rim_a,
                     const d2q9_lattice* L_rim_b, const int32_t* deep_tiles,
                     const uint8_t* deep_walls, int32_t n_deep, const void* f_src,
                     void* f_dst, void* f_tmp,
                     const void* mom_src, void* mom_dst, void* mom_tmp,
                     int write_moments, void* stream);

/* Rows of a deep tile of d2q9_fused_step2. */
int d2q9_deep_rows(void);

/* rho/u of the populations f (e.g. pop_new) on every fluid cell. */
int d2q9_moments(const d2q9_lattice* L, const void* f, void* mom,
                 void* stream);

/* FP64 sum of all populations over fluid cells into *out (device double).
 * work: device doubles, at least d2q9_reduce_work_len(L) of them. */
int64_t d2q9_reduce_work_len(const d2q9_lattice* L);
int d2q9_mass_fp64(const d2q9_lattice* L, const void* f, double* work,
                   double* out, void* stream);

/* d2q9_mass_fp64 with no workspace argument and no allocation.  It has no
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  The
 * partial sums live in a static device array of 2048 doubles (16 KiB per
 * GPU, held until the process ends), which bounds nothing about the
 * lattice (a reduction launches at most 2048 workgroups); a launch needing
 * more returns D2Q9_E_TOO_LARGE.  Calls that overlap on one GPU share that
 * array and race: order them on one stream.  Same kernels, same result,
 * bit for bit, as d2q9_mass_fp64. */
int d2q9_mass_fp64_static(const d2q9_lattice* L, const void* f, double* out,
                          void* stream);

/* Residue sums over fluid cells of mom vs mom_old (FP64):
 * out[0..5] = sum drho^2, sum rho_old^2, sum dux^2, sum ux_old^2,
 *             sum duy^2, sum uy_old^2; then mom_old = mom. */
int d2q9_residues(const d2q9_lattice* L, const void* mom, void* mom_old,
                  double* work, double* out, void* stream);

/* Plain copy kernel (bandwidth reference). */
int d2q9_copy(void* dst, const void* src, int64_t bytes, void* stream);

/* HIP graph capture of the work issued on `stream` between begin and end
 * (end instantiates the graph). */
int d2q9_graph_begin(void* stream);
int d2q9_graph_end(void* stream, void** graph_exec);
int d2q9_graph_launch(void* graph_exec, void* stream);
int d2q9_graph_destroy(void* graph_exec);

/*
 * Momentum exchange on obstacle and wall links.  A force node is a fluid
 * cell and a 9-bit mask of directions k whose link it sums:
 *   F += f*[k] c_k - f_new[inv k] c_{inv k}     (FP64 sums)
 * with f* the post-collision populations of the current step (f_pop) and
 * f_new their streamed values after the boundary rules (read from
 * f_pop_new when it is given, else recomputed for the node from f_pop and
 * mom).  A group with torque != 0 also sums r_x F_y - r_y F_x over its
 * links, r = (i - 1, j - 1) - ref with the minimum image across periodic
 * sides.  Nodes are sorted by group; each group owns n_blocks whole blocks
 * of 256 nodes starting at block first_block (padding nodes have mask 0),
 * n_nodes is a multiple of 256, work holds 3 doubles per block and out 3
 * per group (fx, fy, torque).  The sums are in a fixed order: a run is
 * reproducible bit for bit.
 */
typedef struct d2q9_force_node {
  int64_t cell;   /* origin + i * pitch + j */
  int32_t group;
  uint32_t mask;  /* bit k: sum the link of direction k */
} d2q9_force_node;

typedef struct d2q9_force_group {
  double ref[2];
  int32_t torque;
  int32_t first_block;
  int32_t n_blocks;
  int32_t pad;
} d2q9_force_group;

int d2q9_momentum_exchange(const d2q9_lattice* L, const void* f_pop,
                           const void* f_pop_new, const void* mom,
                           const d2q9_force_node* nodes, int64_t n_nodes,
                           const d2q9_force_group* groups, int n_groups,
                           double* work, double* out, void* stream);

/*
 * Batched sweep: many lattices advance one fused step in one persistent
 * launch.  Every entry carries its lattice descriptor and its two
 * population and moment buffers; at parity p a lattice streams from f[p]
 * into f[1-p] (moments mom[p] -> mom[1-p]).  All entries share one storage
 * type (FP64 or FP32; others return D2Q9_E_UNSUPPORTED) and the workgroup
 * shape 64 x tile_i; each may have its own size,
 * boundaries, collision model and forcing.
 *
 * Work items (lattice, code): code >= 0 is position `code` of that
 * lattice's tile list (a bulk tile if code < n_bulk, else a general tile);
 * code < 0 is the chunk of fix-up cells starting at 64 * tile_i * (-code-1).
 * The items of one step write disjoint cells, so they run in any order.
 * queue_bounds[0..8]: XCD x (workgroup b lives on XCD b % 8) owns items
 * [queue_bounds[x], queue_bounds[x+1]).  policy 0: static, workgroup r of
 * an XCD takes that XCD's items r, r + grid/8, ...; policy 1: dynamic, each
 * workgroup claims `chunk` items at a time from its own XCD queue with an
 * atomic counter, then helps the other queues.  waves: 0 lets the compiler
 * allocate registers freely; 6 or 8 selects a build of the kernel whose
 * registers allow that many waves per SIMD (more latency hiding, some
 * register spills).  has_general = 0 promises that no item is a general
 * tile (no lattice holds solid cells) and selects a build without that
 * path.  counters: 16 int32, zero before the first
 * step; a step of parity p uses counters[8p .. 8p+8) and clears the other
 * half for the next step.  grid must be a multiple of 8.  Entries, items,
 * bounds and counters are device memory.
 */
typedef struct d2q9_batch_entry {
  d2q9_lattice L;
  void* f[2];
  void* mom[2];
} d2q9_batch_entry;

typedef struct d2q9_batch_item {
  int32_t lattice;
  int32_t code;
} d2q9_batch_item;

int d2q9_batched_step(const d2q9_batch_entry* entries,
                      const d2q9_batch_item* items,
                      const int32_t* queue_bounds, int32_t* counters,
                      int parity, int storage, int tile_i, int grid,
                      int policy, int chunk, int waves, int has_general,
                      int write_moments, void* stream);

/*
 * Adjoint of one FP64 step (BGK, Guo forcing, per-cell force field;
 * bounce-back, periodic and static-solid boundaries only, so that streaming
 * permutes the populations).  lam_src = dL/d(post-collision populations of
 * this step), lam_dst receives dL/d(post-collision populations of the
 * previous step) when `streamed` (the transpose of pull streaming), or
 * dL/d(pre-collision populations) at each cell when not (the first step,
 * which collides the initial equilibrium without streaming).  mom holds
 * this step's rho/u (stored by the forward pass), seed_u (nullable) the
 * loss gradient with respect to this step's velocity [2][plane], and
 * grad_force [2][plane] accumulates dL/dF.  Layouts as the populations /
 * side arrays of d2q9_lattice.
 */
int d2q9_adjoint_step(const d2q9_lattice* L, int streamed, const void* lam_src,
                      void* lam_dst, const void* mom, const void* seed_u,
                      void* grad_force, void* stream);

/* dL/d(rho0, u0) [3][plane] from mu = dL/d(initial equilibrium populations)
 * and mom0 = the initial rho/u. */
int d2q9_adjoint_init_grad(const d2q9_lattice* L, const void* mu,
                           const void* mom0, void* grad_init, void* stream);

/* Text of the last error on this thread. */
const char* d2q9_last_error(void);

/* Digest of the sources and compiler flags the library was built from ("h"
 * + 32 hex digits); a loader compares it with its own tree's to refuse a
 * stale library.  No reference counterpart.  No GPU needed. */
const char* d2q9_source_hash(void);

/* Struct layout for binding checks: sizeof(d2q9_lattice), sizeof of the two
 * boundary structs, then offsetof(d2q9_lattice, f) for f in plane, cpl,
 * pop_len, storage, tile_j, omega, w, gravity, rates, dev_scale, cf,
 * flags, tiles, n_bulk, fix_cells, n_fix, bulk_masks, bc_slot, bc_nodes,
 * bc_elems, force_field.  Returns the count written (24) or D2Q9_E_ARG if n
 * is too small.  No GPU needed. */
int d2q9_abi_layout(int64_t* out, int n);

#ifdef __cplusplus
}
#endif

#endif /* D2Q9_H */
