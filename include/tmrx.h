/*
  tmrx -- C-ABI of the MI355X hex8 topology-optimisation solve path.

  Plain C types only: opaque handles, int status codes (0 = success,
  negative = error, message from tmrx_last_error()), caller-owned buffers or
  borrowed pointers whose lifetime is stated per call.

  Each entry point names the reference interface it replaces (file:line in
  smdogroup/tmr).  The reference has no status codes: it prints to stderr and
  returns NULL/0; here every failure is reported through the return value.
*/
#ifndef TMRX_H
#define TMRX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tmrx_forest tmrx_forest;
typedef struct tmrx_problem tmrx_problem;

/* Same layout as TMROctant (src/TMROctant.h:48-53). */
typedef struct {
  int32_t block;
  int32_t x, y, z;
  int32_t tag;
  int16_t level;
  int16_t info;
} tmrx_octant;

/* Last error message of the calling thread ("" when none). */
const char *tmrx_last_error(void);
/* Library version string. */
const char *tmrx_version(void);

/* ---------------------------------------------------------------------------
   Forest: TMROctForest (src/TMROctForest.h:51-181), one process.
   ------------------------------------------------------------------------- */

/* TMROctForest::TMROctForest(comm, mesh_order, interp_type)
   (src/TMROctForest.h:51-52).  Only mesh_order 2 builds nodes. */
int tmrx_forest_create(int mesh_order, int interp_type, tmrx_forest **out);
/* TMREntity::decref (src/TMRBase.h:151-188): frees the forest. */
void tmrx_forest_free(tmrx_forest *f);

/* TMROctForest::setConnectivity (src/TMROctForest.cpp:558-590). */
int tmrx_forest_set_connectivity(tmrx_forest *f, int num_nodes, const int *block_conn,
                                 int num_blocks);
/* TMROctForest::getConnectivity (src/TMROctForest.cpp:1633-1692).
   Borrowed pointers valid until the next set_connectivity. */
int tmrx_forest_get_connectivity(tmrx_forest *f, int *nblocks, int *nfaces, int *nedges,
                                 int *nnodes, const int **block_conn,
                                 const int **block_face_conn, const int **block_edge_conn,
                                 const int **block_face_ids);
/* Replaces TMROctForest::setTopology (src/TMROctForest.cpp:508-534): block
   vertex coordinates (num_nodes x 3); every block is mapped trilinearly. */
int tmrx_forest_set_node_points(tmrx_forest *f, const double *Xpts);
/* Replaces the entity names of TMRTopology used by getNodesWithName /
   getOctsWithName: kind 0 vertex, 1 edge, 2 face, 3 volume (block). */
int tmrx_forest_set_name(tmrx_forest *f, int kind, int index, const char *name);

/* TMROctForest::createTrees (src/TMROctForest.cpp:1744-1833). */
int tmrx_forest_create_trees(tmrx_forest *f, int level);
/* No counterpart in the reference (it restarts from design vectors only):
   the octants of a forest saved earlier (tmrx_forest_get_octants' layout),
   so that a checkpointed forest is rebuilt exactly.  Every octant must lie
   inside its block, aligned to its level, and the octants of each block must
   tile it without overlap or gap; they are sorted into Morton order and
   renumbered.  Frees the nodes (createNodes builds them again). */
int tmrx_forest_set_octants(tmrx_forest *f, const tmrx_octant *octs, int n);
/* No counterpart in the reference (createTrees takes one depth for every
   block): uniform trees of depth block_levels[b] in block b, so that blocks
   of different sizes carry elements of one size (nblocks must equal the
   number of blocks). */
int tmrx_forest_create_trees_per_block(tmrx_forest *f, const int *block_levels, int nblocks);
/* TMROctForest::refine (src/TMROctForest.cpp:2169-2329); refinement may be
   NULL (refine everything by one level). */
int tmrx_forest_refine(tmrx_forest *f, const int *refinement, int min_level, int max_level);
/* TMROctForest::balance (src/TMROctForest.cpp:2917-3089). */
int tmrx_forest_balance(tmrx_forest *f, int balance_corner);
/* The same balance on GPU `device` (balanceOctant, addFace/Edge/
   CornerNeighbors, src/TMROctForest.cpp:2525-3089, one thread per octant per
   level, radix sorts of the 128-bit position keys): the octants, their order
   and tags equal tmrx_forest_balance's bit for bit.  ms (may be null)
   receives the wall time of the pass. */
int tmrx_forest_balance_device(tmrx_forest *f, int balance_corner, int device, double *ms);
/* TMROctForest::coarsen (src/TMROctForest.cpp:2119-2164). New handle. */
int tmrx_forest_coarsen(tmrx_forest *f, tmrx_forest **out);
/* TMROctForest::duplicate (src/TMROctForest.cpp:2097-2109). New handle. */
int tmrx_forest_duplicate(tmrx_forest *f, tmrx_forest **out);
/* TMROctForest::repartition (src/TMROctForest.cpp:1922-2095): one process
   owns everything, so this only validates the handle. */
int tmrx_forest_repartition(tmrx_forest *f);
/* TMROctForest::createNodes (src/TMROctForest.cpp:4064-4268). */
int tmrx_forest_create_nodes(tmrx_forest *f);
/* The same createNodes on GPU `device` (computeDepFacesAndEdges,
   labelDependentNodes, createDependentConn and the node locations of
   src/TMROctForest.cpp:3619-3832, 5157-5680 as device passes with radix
   sorts of the node keys): octant info bits, connectivity, hanging
   constraints, node numbers and points equal tmrx_forest_create_nodes's bit
   for bit.  ms (may be null) receives the wall time of the pass. */
int tmrx_forest_create_nodes_device(tmrx_forest *f, int device, double *ms);

/* TMROctForest::getOctants (src/TMROctForest.cpp:1452-1456).  Borrowed,
   valid until the forest is refined/balanced/freed. */
int tmrx_forest_get_octants(tmrx_forest *f, const tmrx_octant **octs, int *n);
/* TMROctForest::getNodeConn (src/TMROctForest.cpp:5686-5704).  Borrowed.
   num_local_nodes receives the number of local nodes (the length of the node
   numbers, tmrx_forest_get_node_numbers); the reference declares it but never
   writes it (it assigns num_owned_nodes twice). */
int tmrx_forest_get_node_conn(tmrx_forest *f, const int **conn, int *num_elements,
                              int *num_owned_nodes, int *num_local_nodes);
/* TMROctForest::setMeshOrder (src/TMROctForest.cpp:1389-1427): frees the
   nodes and sets the element order; orders below 2 become 2 (the reference
   clamps to [2, MAX_ORDER]).  Only order 2 (trilinear hex8) is built here:
   a higher order returns an error.  interp_type as TMRInterpolationType
   (0 uniform, 1 Gauss-Lobatto, 2 Bernstein; order 2 knots are {-1, 1} for
   all three). */
int tmrx_forest_set_mesh_order(tmrx_forest *f, int mesh_order, int interp_type);
/* TMROctForest::getDepNodeConn (src/TMROctForest.cpp:5717-5729).
   Returns the number of dependent nodes (>= 0) or an error (< 0). */
int tmrx_forest_get_dep_node_conn(tmrx_forest *f, const int **ptr, const int **conn,
                                  const double **weights);
/* TMROctForest::getOwnedNodeRange (src/TMROctForest.cpp:1442-1447):
   range[0] = 0, range[1] = number of independent nodes. */
int tmrx_forest_get_owned_node_range(tmrx_forest *f, int range[2]);
/* TMROctForest::getNodeNumbers (src/TMROctForest.cpp:1461-1466). Borrowed. */
int tmrx_forest_get_node_numbers(tmrx_forest *f, const int **nums, int *n);
/* TMROctForest::getPoints (src/TMROctForest.cpp:1476-1481): num_local x 3,
   in the order of the sorted node numbers.  Borrowed. */
int tmrx_forest_get_points(tmrx_forest *f, const double **X, int *n);
/* TMROctForest::getNodesWithName (src/TMROctForest.cpp:5882-6203).
   Writes up to cap entries, returns the full count (call with cap 0 to size). */
int tmrx_forest_get_nodes_with_name(tmrx_forest *f, const char *name, int *out, int cap);
/* TMROctForest::getOctsWithName (src/TMROctForest.cpp:5747-5862).  info is
   the local face index, tag the element index. Same sizing rule. */
int tmrx_forest_get_octs_with_name(tmrx_forest *f, const char *name, tmrx_octant *out,
                                   int cap);
/* TMROctForest::createInterpolation (src/TMROctForest.cpp:6611-6793) into
   CSR instead of a TACSBVecInterp: rows are the fine independent nodes,
   columns the coarse ones, sorted and merged as TMRIndexWeight::uniqueSort.
   Borrowed arrays stored in the fine forest, valid until its next call. */
int tmrx_forest_create_interpolation(tmrx_forest *fine, tmrx_forest *coarse, int *nrows,
                                     int *ncols, const int **rowp, const int **cols,
                                     const double **weights);
/* The same createInterpolation on GPU `device` (findEnclosing and
   computeElemInterp, src/TMROctForest.cpp:6228-6594, one thread per fine
   independent node; the nodes of either forest are built there first if
   missing): equal to tmrx_forest_create_interpolation's CSR bit for bit.
   ms (may be null) receives the wall time of the pass. */
int tmrx_forest_create_interpolation_device(tmrx_forest *fine, tmrx_forest *coarse, int device,
                                            int *nrows, int *ncols, const int **rowp,
                                            const int **cols, const double **weights, double *ms);
/* No counterpart in the reference: interpolation between forests whose
   blocks differ (coarse spaces that are not nested, for meshes with no
   coarsen() level, e.g. cubes of 13 elements per edge under one of 7).  Each
   fine independent node is located physically in the coarse forest (the
   lowest-index block containing it, then its octant) and the row holds the
   coarse element's trilinear shape functions there, hanging nodes expanded
   as createInterpolation does (src/TMROctForest.cpp:6396-6594), merged and
   sorted by column.  Box blocks only.  On a coarsen() pair it equals
   tmrx_forest_create_interpolation's CSR up to rounding (1e-13).  Same
   borrowed-array rule. */
int tmrx_forest_create_interpolation_geometric(tmrx_forest *fine, tmrx_forest *coarse, int *nrows,
                                               int *ncols, const int **rowp, const int **cols,
                                               const double **weights);

/* TopOptUtils.interpolateDesignVec (tmr/TopOptUtils.py:270-307): the nodal
   vector x_orig (vars_per_node values per independent node of orig_forest)
   interpolated onto the independent nodes of new_forest,
   x_new = P x_orig with P = new_forest.createInterpolation(orig_forest), the
   product on the current HIP device.  on_device: both arrays are device
   pointers (else host arrays, staged). */
int tmrx_forest_interpolate_vec(tmrx_forest *new_forest, tmrx_forest *orig_forest,
                                int vars_per_node, const double *x_orig, double *x_new,
                                int on_device);

/* TMR_GenerateBinFile (src/TMR_STLTools.h:55-68; Python writeSTLToBin,
   tmr/TMR.pyx:4610-4631): the surface of the solid region {x >= cutoff} of
   the nodal field x (vars_per_node values per independent node, component
   x_offset) on the forest, as the level set inside the elements plus the
   solid parts of the outer block faces, normals out of the solid.  The
   level set is traced by marching tetrahedra (six per element around its
   0-7 diagonal) where the reference uses marching cubes, so the triangles
   differ while the surface is the same to second order.  File: int32
   triangle count, then 9 doubles per triangle.  ntris may be NULL. */
int tmrx_forest_generate_bin_file(tmrx_forest *f, const char *filename, const double *x,
                                  int vars_per_node, int x_offset, double cutoff, int *ntris);
/* TMR_ConvertBinToSTL (src/TMR_STLTools.h:70-77): the binary triangle file
   as an ASCII STL ("solid topology", one facet with its unit normal per
   triangle). */
int tmrx_convert_bin_to_stl(const char *binfile, const char *stlfile);
/* Volume output in place of TACS's ToFH5 .f5 file (the OutputCallback of
   examples/topology_optimization/cantilever/cantilever.py:154-182: mesh
   connectivity, nodes and element extras; TACS and its format are not in the
   reference): a legacy-VTK unstructured grid of the hex8 elements, with an
   optional nodal field (node_vpn values per independent node; dependent
   nodes interpolated; 3 components are written as VECTORS) and an optional
   element field (elem_vpe values per element).  NULL skips a field. */
int tmrx_forest_write_vtk(tmrx_forest *f, const char *filename, const double *node_vals,
                          int node_vpn, const char *node_name, const double *elem_vals,
                          int elem_vpe, const char *elem_name);

/* TMROctTACSTopoCreator's element densities from a filter forest
   (src/topology/TMR_TACSTopoCreator.cpp:55-186): rows = elements of f, columns
   = independent nodes of the filter; rho_e = sum_j w_ej x_j.  Borrowed arrays
   stored in f (shared with tmrx_forest_create_interpolation's). */
int tmrx_forest_element_center_weights(tmrx_forest *f, tmrx_forest *filter, int *nrows,
                                       int *ncols, const int **rowp, const int **cols,
                                       const double **weights);

/* ---------------------------------------------------------------------------
   Design filters on a filter forest's nodal field (TMRTopoFilter's
   applyFilter / applyTranspose, src/topology/TMRTopoFilter.h:41-66).
   ------------------------------------------------------------------------- */
typedef struct tmrx_filter tmrx_filter;
/* kind 0: TMRHelmholtzFilter (src/topology/TMRHelmholtzFilter.cpp:31-131):
     x_f = (M + r^2 K)^-1 M x with M the consistent mass and K the Laplacian
     of the trilinear element (TMRHexaHelmholtzModel,
     src/topology/TMRHelmholtzModel.cpp:85-99); the solve is a
     Jacobi-preconditioned CG to 1e-12 on the GPU (the reference: GMRES with
     multigrid to 1e-12).
   kind 1: TMRMatrixFilter (src/topology/TMRMatrixFilter.cpp:50-168):
     nterms terms of out = a o x + b o (M out) (Horner), D = M 1,
     a = 1 / (1 + r^2 / D^(2/3)), b = a r^2 / (D^(2/3) D), normalised by
     1 / filter(1).
   Hanging nodes enter through their constraints; the operators are over the
   forest's independent nodes.  device = HIP device index. */
int tmrx_filter_create(tmrx_forest *filter_forest, int kind, double r, int nterms, int device,
                       tmrx_filter **out);
/* TMRHelmholtzPUFilter (src/topology/TMRHelmholtzPUFilter.cpp:407-720; the
   base of TopOptUtils.Mfilter, tmr/TopOptUtils.py:785-890): a
   partition-of-unity filter over n nodes from per-node stencils computed by
   the caller: B (CSR rowp[n + 1], cols, values; off-diagonal stencil
   weights, column order free) and the stencil diagonal D (> 0).  Applied as
   F x = T^-1 sum_{k=0}^{nterms} (D^-1 B)^k D^-1 x (Horner), T^-1 = 1 / F(1);
   tmrx_filter_apply_transpose applies F^T (B^T inside).  Freed with
   tmrx_filter_free. */
int tmrx_filter_create_custom(int n, int nterms, const int *rowp, const int *cols, const double *B,
                              const double *D, int device, tmrx_filter **out);
void tmrx_filter_free(tmrx_filter *f);
/* applyFilter (TMRHelmholtzFilter::applyFilter, :152-255;
   TMRMatrixFilter::applyFilter, :191-211): in and out hold vars_per_node
   values per independent node, interleaved; on_device selects device
   pointers. */
int tmrx_filter_apply(tmrx_filter *f, int vars_per_node, const double *in, double *out,
                      int on_device);
/* applyTranspose (TMRHelmholtzFilter::applyTranspose, :257-361;
   TMRMatrixFilter::applyTranspose, :216-238): the adjoint of the filter, for
   gradients. */
int tmrx_filter_apply_transpose(tmrx_filter *f, int vars_per_node, const double *in, double *out,
                                int on_device);
/* Independent nodes of the filter forest and the CG iterations of the last
   Helmholtz solve. */
int tmrx_filter_info(tmrx_filter *f, int *num_nodes, int *last_iters);

/* Race check of owner-tile accumulation schedules.  No counterpart in the
   reference (its element loops run on the host); for callers that build or
   inspect schedules: the checker every tile build runs before upload.
   Records (24 bytes each: 8 uint16 slot positions, int32 element, uint8
   corner mask, colour, class, pad) of tile t are rec_ptr[t]..rec_ptr[t+1];
   nacc[t] is the tile's count of accumulated slots.  Returns 0 if the
   records are sorted by colour, no colour writes a slot twice and every
   written slot is an accumulator; else -1 with the reason in msg. */
int tmrx_check_tile_schedule(int ntiles, const int *rec_ptr, const void *recs, const int *nacc,
                             char *msg, int msglen);

/* ---------------------------------------------------------------------------
   Problem: the finite-element hierarchy and solver that TMR hands to TACS.
   One handle owns every multigrid level on one GPU.
   ------------------------------------------------------------------------- */

/* TMRStiffnessProperties (src/topology/TMROctConstitutive.cpp:26-71) with a
   single isotropic material (TACSMaterialProperties E, nu, ys).
   penalty_type 0 = SIMP (rho^q + k0), 1 = RAMP. */
typedef struct {
  double E, nu, ys;
  double q, k0, eps;
  int penalty_type;
} tmrx_material;

typedef struct {
  double rtol;      /* relative residual tolerance of the FP64 PCG */
  double atol;      /* absolute residual tolerance */
  int max_iters;
  int precision;    /* 0: FP64 V-cycle, 1: FP32 V-cycle inside FP64 PCG */
  int cheb_degree;  /* Chebyshev smoother degree */
  double cheb_lower, cheb_upper; /* interval as fractions of lambda_max */
  int eig_iters;    /* power iterations for lambda_max(D^-1 A) */
  int monitor;      /* print the residual history to stderr */
  int zero_guess;   /* 1: start from u = 0; 0: u holds the initial guess u0 on
                       entry, the PCG solves for the correction and rtol is
                       relative to the initial residual |f - K u0| (the Krylov
                       convention of TACS's GMRES) */
  int fp32_valu;    /* FP32 operator on the vector ALUs instead of the matrix cores */
  int no_graph;     /* enqueue each PCG iteration instead of replaying a captured HIP graph */
  int cheb_pre_degree; /* pre-smoothing degree (0: cheb_degree); post-smoothing uses cheb_degree */
} tmrx_solve_opts;

typedef struct {
  int iters;
  int converged;
  double rnorm0, rnorm;   /* |b| and the final recursive residual norm */
  double setup_ms, solve_ms;
} tmrx_solve_info;

/* Defaults: rtol 1e-8, FP32 V-cycle, Chebyshev degree 2 on
   [1/6, 1.1] lambda_max (TMR's smoother: degree 3 on [1/30, 1.1],
   src/TMR_RefinementTools.cpp:46-47), 200 iterations. */
void tmrx_solve_opts_default(tmrx_solve_opts *o);

/* Replaces TopOptUtils.createTopoProblem's assembler/multigrid construction
   (tmr/TopOptUtils.py:14-128): TMROctTACSCreator::createTACS per level
   (src/TMR_TACSCreator.cpp:332-461) and TMR_CreateTACSMg
   (src/TMR_RefinementTools.cpp:37-107).  forests[0] is the finest level;
   each coarser forest must nest in the finer one.  Boundary conditions are
   TMRBoundaryConditions entries (src/TMRBoundaryConditions.cpp:27-49): a
   name and a dof bit mask (bit d clamps displacement component d; 7 = all),
   applied on every level by getNodesWithName.  device = HIP device index. */
int tmrx_problem_create(int nlevels, tmrx_forest **forests, const tmrx_material *mat,
                        int nbcs, const char **bc_names, const int *bc_dof_masks, int device,
                        tmrx_problem **out);
void tmrx_problem_free(tmrx_problem *p);
/* TMR_CreateTACSMg (src/TMR_RefinementTools.h:32-35, .cpp:37-107) on the
   problem's level hierarchy.
   use_galerkin: coarse operators P^T A P (TACSMg's Galerkin mode), formed on
     the GPU at every setup from the finer level's element matrices; 0 keeps
     rediscretised coarse elements (volume-averaged material factors).
   use_chebyshev_smoother: 1 smooths with Chebyshev on D^-1 A
     (TACSChebyshevSmoother, :64-71); 0 with symmetric point SOR, one
     forward and one backward sweep before and after the coarse-grid
     correction, relaxation omega (TACSGaussSeidel, :93-94), on every
     smoothed level's assembled matrix.  The dofs are swept in colour-major
     node order (a greedy colouring of the level's node graph, so that a
     colour updates in parallel), not in TACS's row order: the smoother is
     the same SOR, its iterates follow that order.
   use_coarse_direct_solve: 1 solves the coarsest level directly (FP64
     Cholesky, TACSMg's direct solve, :83-86); 0 smooths it like the other
     levels, one Chebyshev sweep of cheb_degree steps from zero (the
     smoother TMR hands TACSMg there, :87-102).
   omega: the SOR relaxation factor, in (0, 2) (unused by Chebyshev).
   Invalidates the setup; the next solve redoes it. */
int tmrx_mg_create(tmrx_problem *p, double omega, int use_galerkin, int use_coarse_direct_solve,
                   int use_chebyshev_smoother);
/* No counterpart in the reference: the SOR sweep order of level `level`
   after a setup with use_chebyshev_smoother = 0, for checking the smoother
   against a sequential SOR in the same order.  nnodes receives the level's
   independent nodes, ncolors the colours; nodes[nnodes] (colour-major) and
   color_ptr[ncolors + 1] (colour c = nodes[color_ptr[c] .. color_ptr[c + 1]))
   are filled when not NULL.  Within a node the dofs go 0, 1, 2 forward and
   2, 1, 0 backward. */
int tmrx_problem_smoother_order(tmrx_problem *p, int level, int *nnodes, int *ncolors, int *nodes,
                                int *color_ptr);
/* Sizes of the finest level. */
int tmrx_problem_get_sizes(tmrx_problem *p, int *ndof, int *nelem, int *nlevels);
/* HIP stream (hipStream_t) used by every later call; NULL = default stream. */
int tmrx_problem_set_stream(tmrx_problem *p, void *stream);

/* Element densities on the finest level (one per element, element order of
   the forest); on_device != 0 means rho is a device pointer.  Replaces
   TMRTopoProblem::setDesignVars (src/topology/TMRTopoProblem.cpp:974) for
   element-constant densities. */
int tmrx_problem_set_design(tmrx_problem *p, const double *rho, int n, int on_device);
/* Nodal design variables: x[n] with one value per independent node of the
   finest forest (n = its node count); hanging nodes take the interpolated
   value.  Replaces the design path of TMROctConformTACSTopoCreator
   (src/topology/TMR_TACSTopoCreator.cpp:714-777: element i's design numbers
   are its nodes) with TMROctConstitutive::setDesignVars
   (src/topology/TMROctConstitutive.cpp:142-153): the density at each of the
   8 Gauss points is the trilinear interpolation of the corner values and the
   penalty (SIMP rho^q + k0, RAMP rho/(1+q(1-rho)) + k0) applies per point
   (TMROctConstitutive::evalTangentStiffness, :498-551).  Coarse levels get
   x_c = (P^T x_f) / (P^T 1), TMRConformFilter::setDesignVars
   (src/topology/TMRConformFilter.cpp:154-167) with TACSBVecInterp's
   multWeightTranspose.  After this call gradients are per node. */
int tmrx_problem_set_design_nodal(tmrx_problem *p, const double *x, int n, int on_device);
/* Design mode of the last set_design* call (0 element densities, 1 nodal)
   and the length of a design gradient (elements or finest-level nodes). */
int tmrx_problem_design_info(tmrx_problem *p, int *mode, int *n);
/* Nodal design values on multigrid level `level` (0 = finest), x[nodes of
   that level]: TMRConformFilter's per-level design vectors x[k]. */
int tmrx_problem_level_design(tmrx_problem *p, int level, double *x);
/* Prescribed values of the clamped dofs: ubc[n] (n = finest-level dofs,
   entries of free dofs ignored, all zero = homogeneous conditions).  The
   values of TMRBoundaryConditions::addBoundaryCondition
   (src/TMRBoundaryConditions.cpp:68-116) that TMROctTACSCreator::
   setBoundaryConditions hands to TACSAssembler::addBCs
   (src/TMR_TACSCreator.cpp:468-500).  Solutions take them on the clamped
   dofs; the free dofs solve K_ff u_f = f_f - K_fc u_c.  Compliance and its
   gradient then use the full u (-u^T dK/drho u, as TMR's compliance branch). */
int tmrx_problem_set_bc_values(tmrx_problem *p, const double *ubc, int n, int on_device);
/* Density-dependent multigrid setup (diagonals, eigenvalue bounds, coarse
   Cholesky): mg->assembleJacobian + mg->factor
   (src/topology/TMRTopoProblem.cpp:1000-1001). */
int tmrx_problem_setup(tmrx_problem *p, const tmrx_solve_opts *opts);
/* K(rho) u = f by FP64 flexible PCG with one multigrid V-cycle as the
   preconditioner; replaces ksm->solve (src/topology/TMRTopoProblem.cpp:1003-1006).
   f and u have ndof entries (3 per node, interleaved); on_device selects
   device pointers.  Returns 0 when converged, 1 when not, < 0 on error. */
int tmrx_problem_solve(tmrx_problem *p, const double *f, double *u, int on_device,
                       const tmrx_solve_opts *opts, tmrx_solve_info *info);
/* nrhs load cases with one operator: the loop of TMRTopoProblem::evalObjCon
   over load cases (src/topology/TMRTopoProblem.cpp:1003-1009, one ksm->solve
   per case) as one batched PCG.  f and u are [nrhs][ndof]; infos[nrhs].
   Every stage is one launch over all cases; each case stops at its own
   tolerance.  Load case 0 becomes the state read by compliance and the
   gradients.  Returns 1 if any case did not converge. */
int tmrx_problem_solve_multi(tmrx_problem *p, int nrhs, const double *f, double *u,
                             int on_device, const tmrx_solve_opts *opts,
                             tmrx_solve_info *infos);
/* Independent components: forests whose blocks fall into groups sharing no
   vertex hold several problems that are solved side by side in every
   launch, each with its own PCG scalars and stopping test (uneven problem
   batches; TMR would run one TMRTopoProblem per problem).  Per component of
   the last solve: iterations, converged flag and the true relative residual
   |f - K u| / |f| over its dofs. */
int tmrx_problem_num_components(tmrx_problem *p, int *ncomp);
/* The tuned choices: finest-level tile size and the FP32 operator form:
   0 vector ALUs, 1 matrix cores with bf16 three-way split operands, 2
   matrix cores with FP32 operands (non-zero: on the matrix cores). */
int tmrx_problem_tuning(tmrx_problem *p, int *tile_nodes, int *fp32_matrix_cores);
/* No counterpart in the reference: pin the FP32 operator form (0, 1, 2 as
   above) until the next autotune. */
int tmrx_problem_set_fp32_form(tmrx_problem *p, int form);
/* No counterpart in the reference: owner tiles of tile_nodes independent
   nodes at the finest level (64..4096; coarse levels shrink theirs as the
   autotuner does).  Overrides the autotuner until its next call. */
int tmrx_problem_set_tiling(tmrx_problem *p, int tile_nodes);
/* No counterpart in the reference: the owner tiles of one level (count and
   nominal size in independent nodes). */
int tmrx_problem_tile_info(tmrx_problem *p, int level, int *ntiles, int *tile_nodes);
/* Operator autotuner (no counterpart in the reference): rebuilds the owner
   tiles for each candidate size (independent nodes per finest-level tile,
   64..4096; coarse levels shrink theirs to keep >= 2048 tiles), times the
   finest-level FP32 operator in its three forms (vector ALUs, matrix cores
   with bf16-split or FP32 operands) and the FP64 operator, and keeps the
   size and FP32 form with the least
   4 x FP32 + 1 x FP64 time (one V-cycle's and one PCG iteration's share).
   ms[n] receives that time per candidate (-1: does not fit in LDS; may be
   NULL).  Needs a design; invalidates the current solution. */
int tmrx_problem_autotune(tmrx_problem *p, const int *tile_sizes, int n, int reps, int *chosen,
                          double *ms);
int tmrx_problem_component_info(tmrx_problem *p, int rhs, int comp, int *iters, int *converged,
                                double *rel_residual);
/* The Galerkin operator P^T A P of multigrid level `level` (1 <= level <
   nlevels) after a setup with use_galerkin: the matrix TACSMg would form
   there (src/TMR_RefinementTools.cpp:76-77), as 3 x 3 blocks over the
   independent nodes of that level, rows and columns of clamped dofs
   identity.  The coarsest level (level = nlevels - 1) is available in
   either mode when it is solved directly: the matrix that solve factors
   (:83-86).  n receives the block rows, nnzb the blocks; ptr[n + 1],
   col[nnzb] and val[9 nnzb] (row-major blocks) are filled when not NULL
   (call with NULL arrays to size). */
int tmrx_problem_level_matrix(tmrx_problem *p, int level, int *n, int *nnzb, int *ptr, int *col,
                              double *val);
/* Relative true residual |f - K u| / |f| of the last solve (FP64 operator). */
int tmrx_problem_true_residual(tmrx_problem *p, double *rel);
/* Make a stored load case (load f, displacement u; ndof each) the state read by
   the compliance / stress / gradient calls, as TMRTopoProblem does with
   assembler->setVariables(vars[i]) before each load case's gradient
   (src/topology/TMRTopoProblem.cpp:1275-1277). */
int tmrx_problem_set_state(tmrx_problem *p, const double *f, const double *u, int on_device);
/* Compliance f^T u of the last solve (src/topology/TMRTopoProblem.cpp:1034). */
int tmrx_problem_compliance(tmrx_problem *p, double *c);
/* d compliance / d rho_e for every element: the compliance branch of
   TMRTopoProblem::evalObjConGradient (src/topology/TMRTopoProblem.cpp:1275-1280),
   -psi^T dK/drho u with psi = u.  With a nodal design, d compliance / d x_j
   per finest-level node (TMROctConstitutive::addStressDVSens,
   src/topology/TMROctConstitutive.cpp:557-654). */
int tmrx_problem_compliance_grad(tmrx_problem *p, double *g, int on_device);
/* p-norm (exponent P >= 2) of the relaxed von Mises failure of the last
   solution, (sum_e sum_g (V_e/8) fail_eg^P)^(1/P) with
   fail = rho/(eps(1-rho)+rho) * vonMises(C0 B_g u_e)/ys
   (TMROctConstitutive::evalFailure, src/topology/TMROctConstitutive.cpp:1056-1104,
   aggregated like TACS's "pnorm-continuous" KSFailure used in
   examples/cylinder/cylinder.py:1040-1042). */
int tmrx_problem_stress_pnorm(tmrx_problem *p, double P, double *value);
/* No counterpart in the reference: tmrx_problem_stress_pnorm for callers
   that cannot allocate device memory (code under HIP graph capture, serving
   loops, harnesses that forbid hipMalloc).  Allocates nothing: the partial
   sums live in a fixed device array of the library (65,552 doubles, about
   512 KiB per GPU, held until the process ends), so problems of at most
   16,777,216 elements are accepted (larger ones return an error).  Every call
   on a GPU shares that array: calls that overlap on one GPU race. */
int tmrx_problem_stress_pnorm_noalloc(tmrx_problem *p, double P, double *value);
/* Its gradient w.r.t. the element densities: explicit relaxation term
   (TMROctConstitutive::addFailureDVSens, :1159-1215) minus psi^T dK/drho u,
   psi from the adjoint solve K psi = df/du -- the general-functional branch
   of TMRTopoProblem::evalObjConGradient (src/topology/TMRTopoProblem.cpp:1253-1271).
   Returns 1 if the adjoint solve did not converge.  With a nodal design the
   relaxation uses the Gauss-point density and g is per finest-level node. */
int tmrx_problem_stress_pnorm_grad(tmrx_problem *p, double P, double *g, int on_device,
                                   const tmrx_solve_opts *opts, tmrx_solve_info *info);
/* One component of the multigrid preconditioner on its own, in the
   V-cycle's precision (opts->precision), from a zero start -- the pieces
   TACSMg composes from TMR_CreateTACSMg's objects
   (src/TMR_RefinementTools.cpp:37-107): op 0 y = A_l x (the level operator,
   rediscretised or Galerkin); 1 y = R_l x (level l to l + 1, the transpose
   of createInterpolation's P, src/TMROctForest.cpp:6611-6793, clamped rows
   zero); 2 y = P_l x (level l + 1 to l); 3 y = the pre-smoothing sweep of
   level l with right-hand side x (Chebyshev, TACSChebyshevSmoother's role,
   :64-71, or one symmetric SOR sweep, TACSGaussSeidel's, :93-94);
   4 y = the coarsest level's solve (the direct solve, :83-86, or with
   use_coarse_direct_solve = 0 its Chebyshev sweep); 5 y = the V-cycle from
   level l.  x and y have the sizes of the input and output levels.
   lam_max (may be null) receives the level's estimate of the largest
   eigenvalue of Dinv A (0 where no Chebyshev smoother runs: SOR-smoothed
   levels and a coarsest level solved directly).  Sets the preconditioner
   up first if the design changed. */
int tmrx_problem_mg_op(tmrx_problem *p, int op, int level, const double *x, double *y,
                       int on_device, const tmrx_solve_opts *opts, double *lam_max);
/* y = K(rho) x on the finest level (precision 0 FP64, 1 FP32 on the matrix
   cores, 2 FP32 on the vector ALUs). */
int tmrx_problem_apply(tmrx_problem *p, const double *x, double *y, int on_device, int precision);
/* Consistent nodal load of a constant traction tr[3] on the faces named
   `name`: TopOptUtils.compute3DTractionLoad (tmr/TopOptUtils.py:241-267).
   f has ndof entries (host). */
int tmrx_problem_traction_load(tmrx_problem *p, const char *name, const double tr[3], double *f);

/* Measurement hooks (no counterpart in the reference, which records no
   timings): HIP events around the finest-level operator launches on the
   problem's stream.  enable = 0 stops; otherwise it is a mask of the launch
   kinds to time: 1 plain product, 2 residual b - A x, 4 Chebyshev update,
   16 product with dot, -1 all.  result: summed milliseconds, launch count and
   a mask of the precisions seen (1 FP32, 2 FP64). */
int tmrx_problem_kernel_timing(tmrx_problem *p, int enable);
int tmrx_problem_kernel_timing_result(tmrx_problem *p, double *total_ms, int *count,
                                      int *precision_mask);

#ifdef __cplusplus
}
#endif

#endif /* TMRX_H */
