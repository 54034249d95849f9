/*
 * h2g -- MI355X-native hierarchical (H2) matrix-vector product behind a C ABI.
 *
 * One shared library (libh2g.so) replaces the per-node traversal that H2Factory.jl runs for
 * LinearAlgebra.mul! on an H2 map.  The method (tree, leaf collections, transfers, couplings,
 * near operator) is described once in an h2g_core_desc and uploaded into a read-only core;
 * each h2g_ws workspace owns the mutable device buffers of one matvec at a time.
 *
 * All arrays are plain host pointers in the core description; x and y of h2g_matvec are
 * device pointers unless H2G_HOST_PTRS is set.  Matrices are column-major (Julia layout).
 * Complex values are interleaved (re, im) pairs.
 */
#ifndef H2G_H
#define H2G_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Each maps to the Julia exception the reference raises for the same
 * condition: H2G_EBUSY  -> ConcurrentWorkspaceError (src/maps/workspace.jl:284-343),
 * H2G_EINVAL -> ArgumentError (src/maps/workspace.jl:124-135,
 *               src/farinteractions/i2otranslations/i2oprotocol.jl:180-203),
 * H2G_EDIM   -> DimensionMismatch (src/farinteractions/i2otranslations/symmetryreducedcollection.jl:19-26). */
enum {
    H2G_OK = 0,
    H2G_EBUSY = 1,
    H2G_EINVAL = 2,
    H2G_EDIM = 3,
    H2G_ENOMEM = 4,
    H2G_EHIP = 5,
    H2G_EUNSUPPORTED = 6,
    H2G_EAGAIN = 7     /* h2g_ws_query: work of the workspace is still in flight */
};

/* Element types. */
enum { H2G_F64 = 0, H2G_C64 = 1, H2G_F32 = 2, H2G_C32 = 3 };

/* Orientation of the product: A, transpose(A), adjoint(A)
 * (src/maps/traits.jl:4-53, LinearMaps.TransposeMap / AdjointMap). */
enum { H2G_OP_N = 0, H2G_OP_T = 1, H2G_OP_C = 2 };

/* Which product h2g_matvec computes.
 *   H2G_MUL     mul!(y, A, x[, a, b])        src/maps/abstracth2map.jl:42-86
 *   H2G_FARMUL  farmul!(y, A, x[, a, b])     src/farinteractions/farmul.jl:8-70
 *   H2G_NEARMUL nearmul!(y, A, x[, a, b])    src/nearinteractions/nearmul.jl:8-14 */
enum { H2G_MUL = 3, H2G_FARMUL = 2, H2G_NEARMUL = 1 };

/* Storage of the couplings (I2O / M2L).
 *   DENSE        one K x K matrix per translation id
 *                (src/farinteractions/i2otranslations/i2otranslationcollection.jl:157-221,351-409)
 *   TOEPLITZ     one L^3 spectrum D = fft(kappa) per id, L = 2p-1, kappa in wrap order
 *                (src/farinteractions/i2otranslations/toeplitzi2ocollection.jl:9-193,
 *                 src/moments/toeplitz/toeplitzmul.jl:86-305)
 *   SYMDENSE     canonical K x K matrices + grid permutation per pair
 *                (src/farinteractions/i2otranslations/symmetryreduceddense.jl:16-293)
 *   SYMTOEPLITZ  canonical spectra + spectrum permutation per pair
 *                (src/farinteractions/i2otranslations/symmetryreducedcollection.jl:11-35,356-387)
 *   DIAGONAL     one vector of K coefficients per translation id, applied elementwise (the
 *                MLFMA-style coupling): out (+)= coef .* in for forward and transpose,
 *                conj(coef) .* in for the adjoint
 *                (src/farinteractions/i2otranslations/i2otranslationcollection.jl:415-495,
 *                 src/moments/abstractmoment.jl:420-510) */
enum { H2G_M2L_DENSE = 0, H2G_M2L_TOEPLITZ = 1, H2G_M2L_SYMDENSE = 2, H2G_M2L_SYMTOEPLITZ = 3, H2G_M2L_DIAGONAL = 4 };

/* Arithmetic of the far field in a workspace.
 *   PLAIN  everything in the core's type (Float64 / ComplexF64): the drop-in.
 *   MIXED  P2M, M2M, M2L, L2L, L2P in FP32 (ComplexF32); the near field of real cores reads an
 *          FP32 copy of the near values (made at core creation, or by h2g_method_fill) and
 *          multiplies and accumulates in FP64; the near+far sum and all precomputed spectra stay
 *          FP64.
 *   BF16X2 / BF16X3  MIXED, with the dense (and symmetry-reduced dense) couplings on BF16 MFMA:
 *          each FP32 operand is split into BF16 planes, hi + lo (X2: products hi.hi + hi.lo +
 *          lo.hi) or hi + mid + lo (X3: the six products of total order <= 2), accumulated in
 *          FP32 by the MFMA and added into FP64 every 64 reductions.  Toeplitz couplings run as in
 *          MIXED. */
enum { H2G_PREC_PLAIN = 0, H2G_PREC_MIXED = 1, H2G_PREC_BF16X2 = 2, H2G_PREC_BF16X3 = 3 };

/* Flags of h2g_matvec. */
enum {
    H2G_HOST_PTRS = 1, /* x, y point to host memory (staged through the workspace)          */
    H2G_ASYNC = 2      /* return after enqueueing on the workspace stream; with
                          H2G_HOST_PTRS the host buffers must be pinned and stay alive until
                          h2g_ws_query (or h2g_ws_sync) reports the workspace idle             */
};

/* One side (test = rows/targets, trial = columns/sources) of the map: its octree and its leaf
 * data.  Boxes are level-contiguous; the children of a box are contiguous on the next level.
 * Replaces the H2Trees tree plus testmomentcollections / trialmomentcollections and the I2I
 * level translator that the core accessors return (docs/src/implementing/required.md:23-104,
 * src/maps/leveltranslator.jl:100-158). */
typedef struct h2g_side_desc {
    int64_t n_dofs;               /* degrees of freedom of this side                          */
    int32_t n_levels;             /* levels 0 .. n_levels-1                                    */
    int32_t n_boxes;
    const int64_t *perm;          /* [n_dofs] tree position -> original dof (values(tree, leaf)) */
    const int32_t *level_offsets; /* [n_levels+1] boxes of level l: [off[l], off[l+1])         */
    const int32_t *parent;        /* [n_boxes] parent box, -1 for the root                     */
    const int32_t *sector;        /* [n_boxes] child index 0..7 in its parent, -1 for the root  */
    const int64_t *dof_begin;     /* [n_boxes] tree-order dof range of the box's subtree        */
    const int64_t *dof_end;
    const uint8_t *is_leaf;       /* [n_boxes]                                                  */
    const void *collection;       /* K x n_dofs (NULL: filled later by h2g_method_fill),
                                     column j = leaf basis at tree-ordered dof j:
                                     P2M m = C x (linkmomenttocollection.jl:56-63), L2P y = C^T m
                                     (linkmomenttocollection.jl:386-411).  Core type.           */
    const void *transfers;        /* [n_levels][8] K x K: transfers[l][s] maps a child of sector s
                                     at level l to its parent (leveltranslator.jl:111-116).
                                     Entries of level 0 are ignored.  Core type.               */
    const int32_t *coord;         /* optional [n_boxes][3] integer lattice coordinate of each box
                                     at its level (child = 2*parent + sector bits).  When given,
                                     levels whose couplings follow the parent-neighbour lattice run
                                     the lattice-blocked M2L kernels; NULL keeps the generic ones. */
} h2g_side_desc;

/* Everything a method supplies to a GPU core.  Replaces the core accessors of an H2Factory
 * GalerkinH2Core / PetrovGalerkinH2Core (src/maps/core.jl:17-54). */
typedef struct h2g_core_desc {
    int32_t dtype;                /* H2G_F64 or H2G_C64                                        */
    int32_t K;                    /* moment length per box                                     */
    int32_t p;                    /* nodes per axis for the Toeplitz kinds (K == p^3), else 0  */
    int32_t galerkin;             /* 1: one tree for both sides, `trial` is ignored            */
    h2g_side_desc test;
    h2g_side_desc trial;

    /* couplings: forward pairs (receiving test box, translating trial box) */
    int32_t m2l_kind;
    int64_t n_pairs;
    const int32_t *pair_test;
    const int32_t *pair_trial;
    const int32_t *pair_op;       /* translation id (canonical id for the SYM kinds)           */
    const int32_t *pair_sym;      /* permutation id (SYM kinds), else NULL                    */
    int64_t n_ops;
    const void *ops;              /* DENSE/SYMDENSE: n_ops K x K, core type.
                                     TOEPLITZ/SYMTOEPLITZ: n_ops L^3 ComplexF64 spectra, x fastest
                                     DIAGONAL: n_ops K-vectors, core type                      */
    int32_t n_perms;
    const int32_t *perms;         /* SYMDENSE: [n_perms][K] grid permutation (0-based):
                                       t[perm[i], perm[j]] = canonical[i, j]
                                     SYMTOEPLITZ: [n_perms][L^3]: D_pair[j] = D_canon[perm[j]]  */

    /* near field: block-sparse over (test leaf, trial leaf) pairs; block b is the
     * n_rows x n_cols column-major matrix at near_values + near_offset[b]
     * (the stored operator nearmul! multiplies, src/nearinteractions/nearmul.jl:8-14) */
    int64_t n_near_blocks;
    const int32_t *near_test;
    const int32_t *near_trial;
    const int64_t *near_offset;
    int64_t n_near_values;
    const void *near_values;      /* core type (NULL: filled later by h2g_method_fill)        */

    double alpha_re, alpha_im;    /* the core's far-field prefactor alpha(A)                   */
    int32_t precisions;           /* bitmask (1 << H2G_PREC_*) of workspace precisions to support */
    int32_t device;               /* HIP device ordinal                                        */
} h2g_core_desc;

typedef struct h2g_core h2g_core;
typedef struct h2g_ws h2g_ws;

/* Workspace options (a workspace is the mutable half of an H2Factory map: H2Workspace,
 * src/maps/workspace.jl:36-100). */
typedef struct h2g_ws_opts {
    int32_t precision;            /* H2G_PREC_PLAIN or H2G_PREC_MIXED                         */
    int32_t max_rhs;              /* columns per pass; wider x is processed in column blocks  */
    int32_t use_graph;            /* capture each (op, parts, nrhs) pass into a hipGraph      */
    int32_t timing;               /* record per-phase / per-level device times (HIP events on the
                                     workspace stream; passes are then not captured in graphs) */
    void *stream;                 /* a hipStream_t the workspace runs on (shared by several workspaces,
                                     e.g. a serving loop's pool of as many streams as the device has
                                     hardware queues; the near field then runs in order with the far
                                     field), or NULL for two streams of its own */
    /* Tile shapes (0: the library's default; the autotuner in h2g.autotune measures them per map): */
    int32_t m2l_groups;           /* lattice Toeplitz M2L of wide FP32 batches: workgroups per
                                     frequency (groups of Morton-consecutive parent blocks, at least
                                     4 blocks each); default 512                                   */
    int32_t near_cols;            /* near field of real maps: columns per workgroup, 16, 32 or 64
                                     (one to four 16-column MFMA tiles per wave); default
                                     min(64, 16 ceil(nrhs / 16))                                   */
    int32_t persistent;           /* 1: the lattice Toeplitz coupling stage of every level in one
                                     persistent launch (FP32 batches of >= 4 columns; each XCD
                                     works through its own task list, largest level first); the
                                     workspace then keeps the spectra of all lattice levels at
                                     once (about 1.45x the spectrum memory of the per-level
                                     launches).  0: one launch per level.                        */
    int32_t dense_pairs;          /* dense couplings of real maps: 1 runs the receiving boxes two
                                     by two (a lattice neighbour where one is free), each operator
                                     the two apply read once for both; -1 one receiving box per
                                     workgroup; 0 the library's choice: pairs for batches of at
                                     most 16 columns when the map's operators in the product's
                                     precision exceed 192 MB (then the coupling stage streams
                                     them from HBM)                                              */
} h2g_ws_opts;

/* One per-phase timing record, shaped like the reference's TimingRecord
 * (src/diagnostics/timing.jl:11-20: operation, level, node, mode, seconds). */
typedef struct h2g_timing_record {
    int32_t phase;                /* the operation: 0 gather, 1 near, 2 p2m, 3 m2m, 4 m2l
                                     modulate-in (FFT), 5 m2l apply (Hadamard / dense GEMM),
                                     6 m2l modulate-out (iFFT), 7 l2l, 8 l2p, 9 scatter          */
    int32_t level;                /* -1 for whole-vector phases (missing)                     */
    float ms;
    int32_t mode;                 /* orientation of the product: H2G_OP_N / T / C             */
    int32_t node;                 /* -1: records are per level, not per node (missing)        */
} h2g_timing_record;

/* Structure of a core: the reference's h2summary and translation_counts
 * (src/diagnostics/structure.jl:40-118); the tree fields describe the test side. */
typedef struct h2g_summary {
    int64_t rows, cols;
    int32_t dtype, m2l_kind, galerkin, n_levels;
    int32_t n_nodes, n_leaves;
    int64_t min_leaf_dofs, max_leaf_dofs;
    double mean_leaf_dofs;
    int64_t far_pairs;            /* M2L pairs                                                */
    int64_t unique_translations;  /* distinct coupling operators (translation id, symmetry)  */
    double reuse;                 /* far_pairs / unique_translations (NaN without pairs)     */
    int64_t near_blocks, near_values;
    int32_t lattice_levels;       /* levels whose Toeplitz couplings run the lattice kernels */
    int32_t dense_mfma;           /* dense couplings on the MFMA kernels (1) or the box GEMM */
    int64_t device_bytes;
} h2g_summary;
int h2g_core_summary(const h2g_core *core, h2g_summary *out);

/* Read-only checks of a core, like the reference's checkplans / checkmoments / sanitycheck
 * (src/diagnostics/structure.jl:150-310): one record per issue found. */
enum {
    H2G_ISSUE_NO_FAR_PAIRS = 1,           /* warning: the far field is empty                  */
    H2G_ISSUE_SELF_FAR_PAIR = 2,          /* error: a node translated with itself             */
    H2G_ISSUE_UNRESOLVED_TRANSLATION = 3, /* error: a pair's operator id is out of range      */
    H2G_ISSUE_MISSING_MOMENT = 4,         /* error: a far-pair node without a moment           */
    H2G_ISSUE_NONUNIFORM_MOMENT_SIZE = 5, /* error: moments of one level differ in size        */
    H2G_ISSUE_GENERIC_COUPLING_LEVEL = 6  /* info: a Toeplitz level on the per-pair kernels   */
};
typedef struct h2g_issue {
    int32_t severity;             /* 0 info, 1 warning, 2 error                               */
    int32_t code;                 /* H2G_ISSUE_*                                              */
    int64_t count;                /* offending pairs / nodes / levels                         */
    int32_t level;                /* the level concerned, or -1                               */
} h2g_issue;
int h2g_core_sanitycheck(const h2g_core *core, h2g_issue *out, int32_t max, int32_t *n);

/* --- core ------------------------------------------------------------------------------ */
/* Builds the read-only device core (GalerkinH2Core / PetrovGalerkinH2Core). */
int h2g_core_create(const h2g_core_desc *desc, h2g_core **out);
/* Reference counting: a core is freed when its creator and every workspace released it. */
int h2g_core_retain(h2g_core *core);
int h2g_core_release(h2g_core *core);
/* size(A) (src/maps/abstracth2map.jl:88-98). */
int h2g_core_size(const h2g_core *core, int64_t *rows, int64_t *cols);
/* Device bytes held by the core (admission against a memory budget). */
int h2g_core_device_bytes(const h2g_core *core, int64_t *bytes);

/* --- workspace ------------------------------------------------------------------------- */
/* H2Workspace(core; ...) (src/maps/workspace.jl:81-100); similarworkspace is a second call
 * with the same core (src/maps/workspace.jl:109-113). */
int h2g_ws_create(h2g_core *core, const h2g_ws_opts *opts, h2g_ws **out);
int h2g_ws_destroy(h2g_ws *ws);
int h2g_ws_device_bytes(const h2g_ws *ws, int64_t *bytes);
/* The workspace's HIP stream (hipStream_t). */
void *h2g_ws_stream(h2g_ws *ws);
/* Moves a workspace created with a caller stream (h2g_ws_opts.stream) to another caller stream;
 * its next product starts after its last one (the new stream waits for the workspace's
 * completion event).  A serving loop uses it to send each batch to the least-loaded stream of
 * its pool.  H2G_EINVAL for a workspace with streams of its own. */
int h2g_ws_set_stream(h2g_ws *ws, void *stream);

/* y = op(A) x (alpha == beta == NULL) or y = alpha op(A) x + beta y, for `which` in
 * H2G_MUL / H2G_FARMUL / H2G_NEARMUL, exactly as the reference's 3- and 5-argument methods.
 * x is n_in x nrhs (leading dimension ldx), y is n_out x nrhs (ldy), column-major, in the
 * original dof order, of element type `dtype` (F64/C64 for a plain map; F32/C32 are
 * accepted and widened on entry / rounded on exit).  alpha and beta point to one scalar of
 * the complex-or-real type matching dtype.  Returns H2G_EBUSY when the workspace is in use
 * (acquireworkspace!, src/maps/workspace.jl:304-314). */
int h2g_matvec(h2g_ws *ws, int op, int which, int dtype, int64_t nrhs,
               const void *x, int64_t ldx, void *y, int64_t ldy,
               const void *alpha, const void *beta, int flags);

/* The same products as h2g_matvec without allocating anything: no workspace, no hipMalloc or
 * hipFree, no allocating library call.  It has no counterpart in the reference; it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The buffers of the product (x and y in tree order, moments,
 * spectra) are views into one static device array of 128 MiB per GPU, which the library holds
 * until the process ends; wider x is processed in column blocks that fit, and a map of which even
 * one column does not fit is refused with H2G_ENOMEM.  `precision` is H2G_PREC_PLAIN or
 * H2G_PREC_MIXED (the core must support it); x and y are device pointers; the product runs on
 * `stream` (a hipStream_t, NULL for the null stream), synchronously unless flags is H2G_ASYNC.
 * Every call on one GPU shares that array: a call made while another is inside the library
 * returns H2G_EBUSY, and calls whose device work overlaps (H2G_ASYNC on different streams) race,
 * so order them on one stream. */
int h2g_matvec_noalloc(h2g_core *core, int precision, int op, int which, int dtype, int64_t nrhs,
                       const void *x, int64_t ldx, void *y, int64_t ldy, const void *alpha, const void *beta,
                       void *stream, int flags);

/* Exclusive ownership of a workspace across several calls: acquireworkspace! /
 * releaseworkspace! (src/maps/workspace.jl:304-317).  h2g_matvec on a held workspace
 * returns H2G_EBUSY, like a nested withworkspace in the reference. */
int h2g_ws_acquire(h2g_ws *ws);
int h2g_ws_release(h2g_ws *ws);

/* Waits for the last product enqueued on the workspace (after H2G_ASYNC calls). */
int h2g_ws_sync(h2g_ws *ws);
/* H2G_OK when the last product enqueued on the workspace has finished, H2G_EAGAIN otherwise
 * (a completion event per workspace, so workspaces sharing a stream are told apart). */
int h2g_ws_query(h2g_ws *ws);

/* Timing records of the last matvec (opts.timing != 0). */
int h2g_ws_timing(h2g_ws *ws, h2g_timing_record *out, int32_t max, int32_t *n);

/* --- method assembly on the device ---------------------------------------------------- */
/* Fills, on the device, the near-field values and/or the leaf collections of a core that was
 * created with near_values == NULL and/or collection == NULL, for the point-kernel
 * interpolation method the benchmarks use: kernel 0 = Laplace 1/(4 pi r), 1 = Helmholtz
 * exp(i k r)/(4 pi r) (0 on the diagonal); collections are tensor Lagrange bases on the
 * `nodes` (reference coordinates in [-1, 1]) of each leaf box (center, half width).
 * The reference leaves this to the downstream method package (near assemblers,
 * src/nearinteractions/strategies.jl:67-81; collection builders behind
 * testmomentcollections / trialmomentcollections). */
typedef struct h2g_method_side_desc {
    const double *points;         /* [n_dofs][3] original order                               */
    const double *box_center;     /* [n_boxes][3]                                             */
    const double *box_halfwidth;  /* [n_boxes]                                                */
} h2g_method_side_desc;

typedef struct h2g_method_desc {
    int32_t kernel;
    double wavenumber;
    int32_t p;
    const double *nodes;          /* [p]                                                      */
    h2g_method_side_desc test;
    h2g_method_side_desc trial;   /* ignored for a Galerkin core                              */
    int32_t fill_near;
    int32_t fill_collections;
} h2g_method_desc;

int h2g_method_fill(h2g_core *core, const h2g_method_desc *desc);

/* ---- Operators of the point-kernel interpolation method, built on the device --------------
 * The couplings of the given translation ids, the level transfers and the symmetry
 * permutation tables, in the layouts h2g_core_desc takes.  The reference builds one operator per
 * translation id from a representative pair (src/farinteractions/i2otranslations/
 * i2otranslationcollection.jl:52-70), the Toeplitz spectra D = fft(kappa)
 * (ext/H2Toeplitz/toeplitzi2ocollection.jl:14-31) and the permutations toeplitzpermutation /
 * latticegridpermutation (src/farinteractions/i2otranslations/symmetryreducedcollection.jl:162-232).
 *   DENSE kinds: ops[o] = M (K x K column-major), M[k, j] = G(x_k - y_j - width_l * offset_o),
 *     x, y the p^3 grids (nodes scaled by width_l / 2, x fastest) of the receiving box at the
 *     origin and of the translating box; Float64 for Laplace, ComplexF64 for Helmholtz.
 *   TOEPLITZ kinds: ops[o] = D = fftn(kappa) (L^3 ComplexF64, x fastest, L = 2p-1),
 *     kappa[d] = G(delta (d - (p-1) offset_o)) over wrap-order d, delta = width_l / (p-1).
 *   transfers: [8] K x K column-major Float64, T_s[k, i] = parent basis k at child node i of
 *     sector s (the core's transfers of every level but 0).
 *   perms (SYM kinds): [48][K] grid permutations (SYMDENSE) or [48][L^3] spectrum permutations
 *     (SYMTOEPLITZ), group enumerated as for h2g_plan_build's pair_sym.
 * width_l = root_size / 2^level.  Outputs are host arrays; NULL fields are skipped. */
typedef struct h2g_ops_desc {
    int32_t kernel;               /* 0 = Laplace 1/(4 pi r), 1 = Helmholtz exp(i k r)/(4 pi r)  */
    double wavenumber;
    int32_t p;
    const double *nodes;          /* [p] reference nodes in [-1, 1]                           */
    double root_size;             /* edge of the level-0 box                                  */
    int32_t m2l_kind;             /* H2G_M2L_DENSE, TOEPLITZ, SYMDENSE or SYMTOEPLITZ          */
    int64_t n_keys;
    const int32_t *keys;          /* [n_keys][4] (level, ox, oy, oz) of each translation id   */
    int32_t device;
} h2g_ops_desc;

typedef struct h2g_ops_out {
    void *ops;
    double *transfers;
    int32_t *perms;
} h2g_ops_out;

int h2g_method_operators(const h2g_ops_desc *desc, const h2g_ops_out *out);

/* ---- Tree and traversal plans, built on the device ----------------------------------------
 * The octree of each side and the pair lists a core description needs (h2g_core_desc: the side
 * arrays, pair_test / pair_trial / pair_op / pair_sym, near_test / near_trial / near_offset).
 * In the reference they come from H2Trees, which the method hands to H2Factory: the tree, the
 * plans of assembleplans (src/maps/interactionstrategy.jl:50-60: isnear and the translating
 * nodes) and the translation ids of translationmetadata
 * (src/farinteractions/i2otranslations/i2otranslationcollection.jl:19-84).  Definitions:
 *   - points get 21-bit integer coordinates floor((x - origin) / size * 2^21) (clipped) and are
 *     put in tree order by a stable sort of their Morton codes (x in the lowest bit);
 *   - a box is split while it holds more than leaf_size points and its level is below
 *     max_level; only non-empty children exist; boxes are stored level by level, Morton order
 *     within a level; a child's sector is bx + 2 by + 4 bz;
 *   - M2L pairs: same-level boxes whose parents are adjacent and that are not adjacent, grouped
 *     by test box, trial boxes ascending; translation id = rank of the pair's (level, offset)
 *     among the present ones, or of (level, canonical offset) for symmetric plans, where the
 *     canonical offset sorts |offset| ascending and pair_sym is the index of the signed
 *     permutation S (S canonical = offset) in the 48-element group enumerated as permutations of
 *     (0, 1, 2) in lexicographic order times sign patterns (+++), (++-), ..., (---);
 *   - near pairs: a test leaf and a trial leaf whose ancestors at the coarser of the two leaf
 *     levels are adjacent, grouped by test leaf, trial ascending; near_offset = exclusive prefix
 *     sum of the block sizes n_rows * n_cols.
 * Points are host arrays [n][3].  h2g_plan_info gives the sizes; h2g_plan_copy copies into
 * host arrays of those sizes (NULL fields are skipped). */
typedef struct h2g_plan_opts {
    double origin[3];             /* lower corner of the root cube                            */
    double size;                  /* edge of the root cube                                    */
    int32_t leaf_size;            /* split while a box holds more points than this            */
    int32_t max_level;            /* deepest level (1..21); 0 means 21                        */
    int32_t symmetric;            /* 1: translation ids over canonical offsets + pair_sym     */
    int32_t device;               /* HIP device ordinal                                       */
} h2g_plan_opts;

typedef struct h2g_plan h2g_plan;

typedef struct h2g_plan_counts {
    int64_t n_dofs[2];            /* [test, trial] (trial == test for a Galerkin plan)         */
    int32_t n_levels[2];
    int32_t n_boxes[2];
    int64_t n_pairs;              /* M2L pairs                                                */
    int64_t n_keys;               /* distinct translation ids                                 */
    int64_t n_near_pairs;
    int64_t n_near_values;        /* sum of the near block sizes                              */
    int32_t galerkin;
} h2g_plan_counts;

typedef struct h2g_plan_tree_out {
    int64_t *perm;                /* [n_dofs] tree position -> original point                 */
    int64_t *level_offsets;       /* [n_levels + 1]                                           */
    int32_t *level;               /* [n_boxes]                                                */
    int32_t *coord;               /* [n_boxes][3] integer coordinate at the box's level       */
    uint64_t *key;                /* [n_boxes] Morton key at the box's level                  */
    int32_t *parent;              /* [n_boxes], -1 at the root                                */
    int32_t *sector;              /* [n_boxes], -1 at the root                                */
    int64_t *begin, *end;         /* [n_boxes] tree-order point range                         */
    uint8_t *is_leaf;             /* [n_boxes]                                                */
} h2g_plan_tree_out;

typedef struct h2g_plan_out {
    h2g_plan_tree_out test;
    h2g_plan_tree_out trial;      /* ignored for a Galerkin plan                              */
    int32_t *pair_test, *pair_trial, *pair_op;  /* [n_pairs]                                  */
    int32_t *pair_sym;            /* [n_pairs], symmetric plans only                          */
    int32_t *keys;                /* [n_keys][4]: (level, ox, oy, oz) of each translation id  */
    int32_t *near_test, *near_trial;            /* [n_near_pairs]                             */
    int64_t *near_offset;         /* [n_near_pairs]                                           */
} h2g_plan_out;

/* test_points [n_test][3]; trial_points NULL for a Galerkin plan (one tree for both sides). */
int h2g_plan_build(const double *test_points, int64_t n_test, const double *trial_points, int64_t n_trial,
                   const h2g_plan_opts *opts, h2g_plan **out);
int h2g_plan_info(const h2g_plan *plan, h2g_plan_counts *out);
int h2g_plan_copy(const h2g_plan *plan, const h2g_plan_out *out);
int h2g_plan_destroy(h2g_plan *plan);

/* Last error message of the calling thread. */
int h2g_last_error(char *buf, size_t len);
/* Library version (major * 10000 + minor * 100 + patch). */
int h2g_version(void);

#ifdef __cplusplus
}
#endif
#endif /* H2G_H */
