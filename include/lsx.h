/*
 * lsx.h — C ABI of the MI355X pressure-projection solver (cut-cell Poisson operator,
 * geometric multigrid, FGMRES / BiCGStab) for LS-STAG.
 *
 * The entry points replace the reference's abstract solver interface
 * IterativeSolverInterface (src/Base/Solvers/iterative_solver_interface.h:31-73), whose
 * concrete classes MGFGMRES (src/Base/Solvers/FGMRES/mg_fgmres.h:33-52) and MGBiCGStab
 * (src/Base/Solvers/BiCGStab/mg_bicgstab.h:33-52) the Poisson equation selects at compile
 * time (src/DiscreteEquations/poisson_discrete_equation.h:19-61).  All arguments are plain
 * pointers and sizes; matrices use the reference CSR layout (sparse_matrix.h:220-228:
 * Number = row pointers int32[n+1], Portrait = column indices int32[nnz], Cell = values
 * double[nnz]) with rows numbered x fastest: num = x + N*y + N*M*z.
 *
 * Every function returns LSX_OK (0) or a negative status; lsx_last_error() gives the
 * message of the last failure on the calling thread.  A handle is used by one host thread
 * at a time; distinct handles are independent (no global mutable state per handle).
 * Work is ordered on the handle's HIP stream.
 */
#ifndef LSX_H_
#define LSX_H_

#ifdef __cplusplus
extern "C" {
#endif

#define LSX_OK 0
#define LSX_ERR_ARG (-1)      /* invalid argument (sizes, pointers, unsupported layout) */
#define LSX_ERR_HIP (-2)      /* HIP runtime failure */
#define LSX_ERR_STATE (-3)    /* call out of order (e.g. solve before set) */
#define LSX_ERR_NOTCONV (-4)  /* reserved: the solvers report non-convergence via it/norm_r */

#define LSX_KRYLOV_FGMRES 0   /* FGMRESInterface, fgmres_interface.h:34-177 */
#define LSX_KRYLOV_BICGSTAB 1 /* BiCGStabInterface, bicgstab_interface.h:37-140 */

#define LSX_MG_CASCADE 0 /* reference cascade: restrict RHS, coarsest LU, prolong + ADLJ
                            post-smoothing (multigrid_solver.h:48-75) */
#define LSX_MG_VCYCLE 1  /* V-cycle: red-black Gauss-Seidel, rediscretised face levels, pinned
                            cells as Dirichlet sinks on coarse levels; an operator without pinned
                            rows (a momentum operator) needs a diagonal beyond its couplings,
                            summed over the children on coarse levels (no reference counterpart) */

#define LSX_PREC_FP64 0  /* everything in FP64 (cascade or V-cycle) */
#define LSX_PREC_MIXED 1 /* V-cycle only: FP32 multigrid, FP16 coarse face coefficients, FP64
                            Krylov loop, operator and coarsest LU */
#define LSX_PREC_MIXED16 2 /* as LSX_PREC_MIXED with FP16 faces on the finest level too */

#define LSX_MEM_HOST 0
#define LSX_MEM_DEVICE 1

typedef struct lsx_solver lsx_solver;

typedef struct lsx_options {
  int krylov;        /* LSX_KRYLOV_*                                          */
  int restart;       /* FGMRES restart length (FGMRES_RESTART_NUM)            */
  int max_it;        /* iteration cap (MAX_ITER)                              */
  double eps;        /* absolute residual 2-norm tolerance (eps)              */
  int mg;            /* LSX_MG_*                                              */
  int n_post;        /* post-smoothing sweeps per level (n_post)             */
  int n_pre;         /* pre-smoothing sweeps (V-cycle only; the cascade has none) */
  double omega;      /* ADLJ damping (DEFAULT_OMEGA_3D = 0.805)               */
  int precision;     /* LSX_PREC_*                                            */
  int levels;        /* 0 = the reference's automatic level count; >= 2 = at most this many
                        levels, the finest included (V-cycle only: the cascade rejects a cap);
                        a V-cycle coarsest level of 65..4096 unknowns (capped, or automatic on
                        a grid with a short axis) is solved by its dense FP64 inverse */
  int z_head_fix;    /* 0 = reproduce the reference z-line head read (default) */
  int ortho;         /* LSX_ORTHO_*: FGMRES orthogonalisation                    */
  int coarse_fuse;   /* V-cycle: 1 = levels of <= 4096 cells in one launch (default), 0 = per level */
  int graphs;        /* V-cycle: 1 = replay each cycle from a captured HIP graph (one per basis slot) */
  int smoother;      /* V-cycle: LSX_SMOOTH_RBGS (default) or LSX_SMOOTH_LINE                  */
  int fused_sweep;   /* V-cycle red-black: 1 = both colours in one pass on a finest level of
                        >= 2^23 cells (default), 2 = on every level, 0 = never (half-sweeps)    */
  int line_levels;   /* V-cycle with LSX_SMOOTH_LINE: the finest line_levels levels are smoothed
                        with lines, coarser ones with red-black sweeps; 0 = every level (default) */
  int diag_faces;    /* V-cycle: 1 = on levels whose rows sum to zero (checked at set-up: every
                        pressure operator; pinned cells and their coarse sinks excepted), the line
                        passes and the residual restriction rebuild a_p from the faces instead of
                        reading 1/a_p (default); 0 = always read 1/a_p                            */
  int dot32;         /* mixed-precision FGMRES with classical Gram-Schmidt: 1 = the projections
                        h = V^T w are summed in FP64 over FP32 copies of the basis vectors (the
                        copies the V-cycle takes as its right-hand side), the update w - V h and
                        everything else stays on the FP64 basis (default); 0 = projections over
                        the FP64 basis                                                          */
} lsx_options;

#define LSX_SMOOTH_RBGS 0 /* red-black Gauss-Seidel                                             */
#define LSX_SMOOTH_LINE 1 /* line relaxation: x, y and z lines solved exactly in turn (line     \
                             Jacobi), for grids whose cells are long in some direction          */
#define LSX_SMOOTH_JACOBI 2 /* damped point Jacobi x + omega D^-1 (b - A x) on every level, with \
                               the omega of the options (in (0, 1])                           */

#define LSX_ORTHO_AUTO 0 /* MGS with the cascade (reference), CGS with the V-cycle           */
#define LSX_ORTHO_MGS 1  /* modified Gram-Schmidt (fgmres_interface.h:97-104)               */
#define LSX_ORTHO_CGS 2  /* classical Gram-Schmidt, two passes over the basis per step     */
#define LSX_ORTHO_CGS2 3 /* classical Gram-Schmidt with one reorthogonalisation            */

/* Fill defaults: FGMRES(13), max_it 200, eps 1e-5, cascade MG, n_post 3, omega 0.805, FP64
 * (base_task_description.h:41-46, ls_stag_real_constants.h:47, 62). */
void lsx_options_default(lsx_options* opt);

/* Create / destroy a solver handle.  stream: a hipStream_t, or NULL for a stream owned by
 * the handle.  Replaces construction of MGFGMRES / MGBiCGStab. */
int lsx_create(lsx_solver** out, const lsx_options* opt, void* stream);
int lsx_destroy(lsx_solver* s);

/* Set the matrix and build the preconditioner: replaces IterativeSolverInterface::set
 * (iterative_solver_interface.h:46-55) with its MultiGridSolver::set
 * (multigrid_solver.h:76-101).  The matrix must have the reference pressure-operator
 * pattern: each row's columns are a subset of {num-N*M, num-N, num-1, num, num+1, num+N,
 * num+N*M}, each at most once.  mem says whether the three arrays are host or device pointers.
 * Every level of the reference cascade that is coarsened must have all three dimensions even
 * (even-even-even transfers, coarse size n / 2) or all three odd (odd-odd-odd transfers, coarse
 * size (n + 1) / 2).  The V-cycle preconditioners coarsen any dimensions (n / 2 or (n + 1) / 2 cells,
 * the last coarse cell of an odd axis with one child). */
int lsx_set_csr(lsx_solver* s, int N, int M, int K, const int* rowptr, const int* col, const double* val, long nnz,
                int mem);

/* New values with the same pattern, then rebuild the hierarchy: replaces
 * rebuildPreconditioner (mg_fgmres.h:35-39, multigrid_solver.h:102-106).  Between the two calls
 * the operator is stale: solve, precondition and operator calls return LSX_ERR_STATE. */
int lsx_update_values(lsx_solver* s, const double* val, int mem);
int lsx_rebuild(lsx_solver* s);

/* Solve A x = b in place (x holds the initial guess, the reference passes p = 0):
 * replaces IterativeSolverInterface::solve (fgmres_interface.h:78-142,
 * bicgstab_interface.h:76-111).  *it and *norm_r mirror the public members it / norm_r. */
int lsx_solve(lsx_solver* s, const double* b, double* x, int mem, int* it, double* norm_r);

/* Tolerance and counters (public members eps / it / norm_r of the reference interface). */
int lsx_set_eps(lsx_solver* s, double eps);
int lsx_get_stats(const lsx_solver* s, int* it, double* norm_r, int* levels);

/* One preconditioner application z = M^{-1} v: replaces MGFGMRES::preconditioner
 * (mg_fgmres.h:46-50) / MultiGridSolver::solve; exposed for parity testing. */
int lsx_precondition(lsx_solver* s, const double* v, double* z, int mem);

/* y = A_level x on multigrid level `level` (0 = fine): replaces
 * RealValuedMeshFunction::setProduct (real_valued_mesh_function.h:81-88, 262-272). */
int lsx_apply_operator(lsx_solver* s, int level, const double* x, double* y, int mem);

/* ---- adjoint solves ------------------------------------------------------------------------
 * The reference has no adjoint; these serve gradient computations through the pressure solve.
 * lsx_solve_adjoint: lambda = A^{-T} g in place (lambda holds the initial guess), by the
 * configured Krylov method (FGMRES or BiCGStab) on A^T, right-preconditioned by the transpose of
 * the V-cycle, to the same absolute tolerance eps.  Needs LSX_MG_VCYCLE and a fine operator
 * symmetric off its pinned rows (lsx_fine_storage == 4); LSX_ERR_STATE otherwise.
 * lsx_precondition_transpose: z = (M^{-1})^T v for the V-cycle M^{-1} of lsx_precondition.
 * lsx_apply_operator_transpose: y = A^T x on the fine level (a pinned row of A, an identity
 * row, becomes a full row of A^T; the other rows drop their couplings to pinned cells). */
int lsx_solve_adjoint(lsx_solver* s, const double* g, double* lambda, int mem, int* it, double* norm_r);
int lsx_precondition_transpose(lsx_solver* s, const double* v, double* z, int mem);
int lsx_apply_operator_transpose(lsx_solver* s, const double* x, double* y, int mem);

/* `sweeps` sweeps of the V-cycle's smoother (opt.smoother) on the finest level, x in place;
 * zero_guess: x is taken as 0.  Exposed for parity tests of the smoothers. */
int lsx_smooth(lsx_solver* s, const double* b, double* x, int sweeps, int zero_guess, int mem);

/* Dimensions of level `level`; returns its cell count or a negative status. */
long lsx_level_dims(const lsx_solver* s, int level, int* N, int* M, int* K);

/* Copy level `level`'s operator as dense stencil coefficients to host memory:
 * npts = 7 ([b,s,w,p,e,n,t]) on the fine level, 27 on coarse levels (index
 * (dx+1)+3(dy+1)+9(dz+1)); layout coef[q*n + num].  Replaces reading CoarseGrid::AH
 * (coarse_grid.h:103-104).  Returns npts or a negative status. */
int lsx_level_stencil(const lsx_solver* s, int level, double* coef_host, long capacity);

/* Copy the diagonal of the fine-level operator (N*M*K values) to host memory.  Rows whose
 * diagonal is 0 are cut cells with every face closed; a consistent right-hand side is 0 there. */
int lsx_operator_diagonal(const lsx_solver* s, double* diag_host);

/* Storage of the fine operator used by the solve: 4 = symmetric planes [a_p, a_e, a_n, a_t]
 * (operator symmetric off its pinned identity rows), 7 = general [b,s,w,p,e,n,t].  No
 * reference counterpart (the reference keeps CSR, sparse_matrix.h:33-229).  Returns the plane
 * count or a negative status. */
int lsx_fine_storage(const lsx_solver* s);

/* Synchronise the handle's stream. */
int lsx_synchronize(lsx_solver* s);

/* Kernel timing (no reference counterpart; the reference's profiler times host sections,
 * profiler.h:53-274).  While enabled, the fine-level launches of each kernel class below are
 * bracketed by HIP events on the handle's stream.  lsx_kernel_stats reports the launch count,
 * the summed device time and the kernel's algorithmic HBM bytes per launch, averaged over the
 * launches (compulsory traffic: each array read or written once). */
#define LSX_K_RESIDUAL 0   /* r -= A c inside the smoother                 */
#define LSX_K_LINE_X 1     /* x-line solve: cascade c = w T^-1 r, ch += c;  */
#define LSX_K_LINE_Y 2     /*   V-cycle line relaxation, one pass of x, y  */
#define LSX_K_LINE_Z 3     /*   or z lines                                  */
#define LSX_K_OPERATOR 4   /* w = A z in the Krylov loop                    */
#define LSX_K_ORTHO 5      /* MGS step: fused w -= h v_i ; h' = w . v_{i+1}  */
#define LSX_K_SMOOTH 6     /* red-black launch of the fine level: one fused */
                           /*   sweep, or one colour (half a sweep)         */
#define LSX_K_CGS_DOT 7    /* block Gram-Schmidt h = V^T w (V-cycle loops)  */
#define LSX_K_CGS_UPDATE 8 /* w -= V h fused with V^T w or ||w|| (V-cycle)  */
#define LSX_K_COUNT 16
int lsx_kernel_timing(lsx_solver* s, int enable);
int lsx_kernel_stats(lsx_solver* s, int kernel, long* launches, double* total_ms, double* bytes_per_launch);

/* ---- cut-cell geometry of the pressure grid (built on the GPU) ------------------------------
 * Replaces the reference's mesh-function build for a static body: LSSTAGUniformMesh nodes and
 * widths (ls_stag_uniform_mesh.h:174-185), the level set (flow_sphere_task_description.h:57-74),
 * edge fractions, face/cell types and face fractions (FlowTaskDescription::calcAlpha /
 * calcTypeOfCells / calcTheta, flow_task_description.h:102-241), cut volumes and staggered
 * masses (LSSTAGUniformMesh::buildMeshFunctions, ls_stag_uniform_mesh.h:4047-4062, 1838-2826). */
typedef struct lsx_geometry lsx_geometry;

#define LSX_BODY_SPHERE 0    /* phi = r - |x - c|                                  */
#define LSX_BODY_ELLIPSOID 1 /* phi = min(a) (1 - |(x - c) / a|), axis-aligned axes a */
#define LSX_BODY_BEZIER6 2   /* phi = signed distance to a six-patch bicubic Bezier surface */

typedef struct lsx_body {
  int kind;
  double center[3];
  double axes[3];     /* sphere: axes[0] = radius; Bezier: the ellipsoid the search starts from */
  const double* ctrl; /* Bezier: 288 host values, patch f, point (i, j), coordinate c at
                         ((f * 16 + i * 4 + j) * 3 + c); patches on the cube faces +x, -x, +y,
                         -y, +z, -z with (u, v) along (y, z), (z, y), (z, x), (x, z), (x, y),
                         (y, x); copies of shared edge and corner points must be equal */
} lsx_body;

/* Six-patch bicubic Bezier fit of the ellipsoid with centre c and semi-axes a into ctrl[288]:
 * each cube edge curve is the least-squares cubic fit of the cube-sphere image of that edge
 * with the corners fixed, each patch interior the least-squares fit of the face image with the
 * boundary fixed.  No reference counterpart (the reference's Bezier bodies are 2D curves,
 * src/Bezier/bezier_approximation.h:39-100).  Host only. */
int lsx_bezier_ellipsoid(const double center[3], const double axes[3], double ctrl[288]);

/* Tensor grid with per-cell widths dx[N], dy[M], dz[K] (host arrays) and lower corner
 * (x0, y0, z0).  A uniform axis places centres at x0 + (k + 1/2) h as the reference does. */
int lsx_geometry_create(lsx_geometry** out, int N, int M, int K, const double* dx, const double* dy,
                        const double* dz, double x0, double y0, double z0, void* stream);
int lsx_geometry_destroy(lsx_geometry* g);
/* Build every field for `body`; cut cells may appear for x in [hs,hf), y in [vs,vf) (all z),
 * the reference's geometry box (1 <= hs < hf <= N-1, 1 <= vs < vf <= M-1). */
int lsx_geometry_build(lsx_geometry* g, const lsx_body* body, int hs, int hf, int vs, int vf);
/* Copy fields to host arrays of N*M*K entries (any pointer may be NULL): cell type codes,
 * east/north/top face fractions, cut volumes, upper-corner level set. */
int lsx_geometry_fields(const lsx_geometry* g, int* type, double* thx, double* thy, double* thz, double* vof,
                        double* phi);

/* Boundary data of the pressure equation's flux (FlowTaskDescription U_E, V_S, V_N, W_B, W_T,
 * flow_task_description.h:79-96). */
typedef struct lsx_flow_bc {
  double U_E, V_S, V_N, W_B, W_T;
  int outflow; /* 1: the east boundary flux is the last interior face of u_out (setBaseFlux) */
  int pin;     /* 1: b = 0 at the pinned cells N-1 and N*M-1 (fixPoint)                   */
} lsx_flow_bc;

/* Pressure right-hand side b = Dx u + Dy v + Dz w + flux for a static body, computed on the
 * device: replaces PoissonDiscreteEquation::recalcRightMember (poisson_discrete_equation.h:
 * 145-159) with DivergenceDiscreteOperatorX/Y/Z::calcD* (divergence_discrete_operator_x.h:
 * 119-159 and its y, z siblings) and the flux of LaplaceDiscreteOperator::initFlux, addFluxNSp,
 * addFluxTBp, setBaseFlux (laplace_discrete_operator.h:48-83, 744-750, 799-808).  Face arrays:
 * u has (N-1)*M*K values (index (y + M z)(N-1) + x), v has N(M-1)K (z N(M-1) + y N + x), w has
 * N*M*(K-1) (z N M + y N + x); u_west has M*K inflow velocities (pU_W, index y + M z); u_out is
 * the x-face field of the outflow flux (the reference's current u), NULL to use u.  Every array
 * is in `mem` space; b receives N*M*K values. */
int lsx_geometry_rhs(lsx_geometry* g, const double* u, const double* v, const double* w, const double* u_west,
                     const double* u_out, const lsx_flow_bc* bc, double* b, int mem);

/* Projection correction after the pressure solve, on the device: u = (G_x dp) inv_Mx + us and
 * likewise v, w, and p += dp * coef when p is not NULL.  Replaces SetSolutions::correction
 * (set_solutions.h:215-237, without its optional outflow clamps) with the gradient rows of
 * GradientDiscreteOperatorX/Y/Z::calc (gradient_discrete_operator_x.h:58-78 and its y, z
 * siblings) and the staggered cell types of LSSTAGUniformMesh::cvu/cvv/cvw
 * (ls_stag_uniform_mesh.h:108-130).  Face arrays as in lsx_geometry_rhs; every array in `mem`
 * space. */
int lsx_geometry_correct(lsx_geometry* g, const double* dp, const double* us, const double* vs, const double* ws,
                         double* u, double* v, double* w, double* p, double coef, int mem);

/* Set the solver's matrix to the cut-cell pressure operator of `g` (built on the device, no
 * CSR round trip) and build the preconditioner; pin != 0 turns rows (N-1,0,0) and
 * (N-1,M-1,0) into identity rows.  Replaces LaplaceDiscreteOperator::build
 * (laplace_discrete_operator.h:752-768, calcLap :107-686, fixPoint :809-821) followed by
 * IterativeSolverInterface::set. */
int lsx_set_geometry(lsx_solver* s, const lsx_geometry* g, int pin);

/* Set the solver's matrix to the cut-cell pressure operator assembled on the device from the
 * caller's own mesh fields and build the preconditioner: the stencil-native form of
 * IterativeSolverInterface::set for a caller that holds the reference's mesh functions
 * (LSSTAGUniformMesh::buildMeshFunctions, ls_stag_uniform_mesh.h:3968-4062) instead of the CSR.
 * dx[N], dy[M], dz[K]: cell widths of the tensor grid (m_delta); type: cell type codes
 * (CellTypes, cut_cells.h:26-168); thx, thy, thz: east, north and top face fractions (m_theta);
 * vof: cut-cell volumes (baseMesh.VOF; the staggered masses and their inverses are formed from
 * them, ls_stag_uniform_mesh.h:2804-2826, ls_stag_base_mesh_params.h:77-85); N*M*K values each,
 * all in `mem` space.  pin as in lsx_set_geometry.  The operator is the one LaplaceDiscreteOperator
 * ::calcLap (laplace_discrete_operator.h:107-686) builds; given the fields a geometry holds it is
 * bit-identical to lsx_set_geometry's. */
int lsx_set_fields(lsx_solver* s, int N, int M, int K, const double* dx, const double* dy, const double* dz,
                   const int* type, const double* thx, const double* thy, const double* thz, const double* vof, int pin,
                   int mem);

/* Set the solver's matrix to the momentum (Helmholtz) operator of velocity component `component`
 * (0 u, 1 v, 2 w) of `g` on its staggered mesh ((N-1) x M x K, N x (M-1) x K or N x M x (K-1)),
 * assembled on the device, and build the preconditioner: the stencil-native form of the set
 * call of HelmholtzDiscreteEquation (helmholtz_discrete_equation.h:99-129) with the operators of
 * HelmholtzDiscreteOperatorX / Y / Z::calcHelm* (helmholtz_discrete_operator_x.h:98-689 and the _y,
 * _z files).  nu = 1/Re (the task's nu), coef_time = COEF_INTEGR_TIME (1.5/dt,
 * base_task_description.h:337-339), flow != 0 for a flow task (outflow east boundary, no bottom
 * and top wall terms).  The geometry must be a box without cut cells (LSX_ERR_ARG otherwise: the
 * immersed-boundary terms of cut cells are not assembled); the staggered mesh has an odd
 * dimension on an even grid, so the solver needs options that accept it (the one-level cascade,
 * levels = 1).  Solve with lsx_solve as usual. */
int lsx_set_helmholtz(lsx_solver* s, const lsx_geometry* g, int component, double nu, double coef_time, int flow);

/* ---- objective and shape gradient ----------------------------------------------------------
 * The reference computes loads but has no adjoint; these entries have no counterpart there. */
#define LSX_OBJ_NONE 0   /* J depends on the geometry only through p                           */
#define LSX_OBJ_DRAG_X 1 /* pressure drag below: J also reads theta_x of the cut cells         */

/* Pressure drag J = sum over cut cells of (theta_x[i-1] - theta_x[i]) dy dz p[i]: the pressure
 * term of LoadsCalculator::calc (loads_calculator.h:97-116).  dJdp (may be NULL) gets dJ/dp. */
int lsx_geometry_drag(lsx_geometry* g, const double* p, int mem, double* J, double* dJdp);

/* Allocation-free variant of lsx_geometry_drag.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, time-stepping loops
 * that must not call hipMalloc).  p, J and dJdp (may be NULL) are device pointers; J receives
 * the same value lsx_geometry_drag returns, in device memory, with no host synchronisation
 * (ordered on the geometry's stream).  The reduction's partial sums live in a fixed device
 * array of the library (256 doubles, 2 KiB per GPU, held until the process ends), so the call
 * allocates nothing; grids above 2^31 - 1 cells are rejected.  Calls that overlap on one GPU
 * share that array and race. */
int lsx_geometry_drag_noalloc(lsx_geometry* g, const double* p, double* J, double* dJdp);

/* Shape gradient of an objective at the solution p of A p = b, given lambda = A^{-T} dJ/dp:
 * dJ/dphi per level-set corner (dJdphi: N*M*K values in `mem` space, may be NULL) and, for a
 * Bezier body, dJ/dctrl (288 host doubles, may be NULL).  u == NULL: b does not depend on the
 * geometry; otherwise b = lsx_geometry_rhs(u, v, w, u_west, u_out, bc) and its change with the
 * face fractions enters.  pin: as in lsx_set_geometry.  h_rel: level-set step of the central
 * differences of the cut-cell geometry, as a fraction of the smallest cell width (1e-7 is a
 * good default). */
int lsx_geometry_shape_gradient(lsx_geometry* g, const double* p, const double* lambda, int objective,
                                const double* u, const double* v, const double* w, const double* u_west,
                                const double* u_out, const lsx_flow_bc* bc, int pin, double h_rel, int mem,
                                double* dJdphi, double* dJdctrl);

/* Forward/backward pair for automatic differentiation through the pressure solve.
 * lsx_adjoint_forward: set the solver's operator from `g` (unless it already is), solve
 * A p = b in place like lsx_solve, and record p, g and its build on a new tape.
 * lsx_adjoint_backward: dJdb = A^{-T} dJdp (the adjoint solve, initial guess 0) and, when
 * dJdctrl is not NULL, dJ/dctrl through the geometry (objective, flow fields and h_rel as in
 * lsx_geometry_shape_gradient).  When the solver's operator was set from another geometry in
 * between, the backward call rebuilds the operator and the preconditioner hierarchy from the
 * taped geometry instead of having kept them.  LSX_ERR_STATE when the geometry was rebuilt
 * after the forward call. */
typedef struct lsx_tape lsx_tape;
int lsx_adjoint_forward(lsx_solver* s, lsx_geometry* g, int pin, const double* b, double* p, int mem, lsx_tape** tape,
                        int* it, double* norm_r);
int lsx_adjoint_backward(lsx_tape* t, const double* dJdp, int objective, const double* u, const double* v,
                         const double* w, const double* u_west, const double* u_out, const lsx_flow_bc* bc,
                         double h_rel, int mem, double* dJdb, double* dJdctrl, int* it, double* norm_r);
int lsx_tape_destroy(lsx_tape* t);

/* ---- state formats ---------------------------------------------------------------------------
 * Checkpoint of a time-stepped run in the reference's binary layout, replacing
 * SetSolutions::saveState / load (set_solutions.h:317-339 / 247-269): int32 it, double time, per
 * immersed body its centre and centre velocity (body_pos / body_vel: 3 doubles per body), per
 * body the loads F_aer_x, F_aer_y, F_aer_z, moment_aer (loads: 4 doubles per body,
 * BaseTaskDescription::saveState, base_task_description.h:151-160), then u, v, w, p of the
 * current solution and of the previous one as raw doubles (Solution::saveState, solution.h:
 * 76-81): u (N-1)*M*K, v N*(M-1)*K, w N*M*(K-1), p N*M*K values, N, M, K the pressure grid.
 * Solution arrays in `mem` space; body arrays on the host.  lsx_state_load checks the file size
 * against the grid and the body count before reading. */
int lsx_state_save(const char* path, int it, double time, int n_bodies, const double* body_pos, const double* body_vel,
                   const double* loads, int N, int M, int K, const double* u, const double* v, const double* w,
                   const double* p, const double* u_prev, const double* v_prev, const double* w_prev,
                   const double* p_prev, int mem);
int lsx_state_load(const char* path, int* it, double* time, int n_bodies, double* body_pos, double* body_vel,
                   double* loads, int N, int M, int K, double* u, double* v, double* w, double* p, double* u_prev,
                   double* v_prev, double* w_prev, double* p_prev, int mem);
/* Bytes of such a checkpoint, or a negative status. */
long lsx_state_bytes(int N, int M, int K, int n_bodies);

/* Wall velocities of the task (BaseTaskDescription U_*, V_*, W_*; the west wall's u is the
 * inflow profile u_west). */
typedef struct lsx_walls {
  double V_W, W_W;
  double U_E, V_E, W_E;
  double U_S, V_S, W_S;
  double U_N, V_N, W_N;
  double U_B, V_B, W_B;
  double U_T, V_T, W_T;
} lsx_walls;

/* Tecplot point file of the solution on the geometry's grid, replacing
 * SolTecplotPrinter::printField (sol_tecplot_printer.h:254-347): "x y z P SPD_x SPD_y SPD_z SPD"
 * on the (N+2) x (M+2) x (K+2) points of the cell centres and the walls, face velocities
 * averaged to the centres (cut cells weighted by their face fractions), in the reference's
 * point order and number format.  u, v, w, p (and u_west, M*K values, may be NULL) in `mem`
 * space; outflow: the east boundary is an outflow (flow tasks: the first and last cell planes
 * are repeated at the bottom and top walls). */
int lsx_geometry_write_tecplot(const lsx_geometry* g, const char* path, const double* u, const double* v,
                               const double* w, const double* p, const double* u_west, const lsx_walls* walls,
                               int outflow, int mem);
/* The same file from host arrays of a tensor grid: cell centres xc[N], yc[M], zc[K] and widths
 * dx[N], dy[M], dz[K] (the reference's m_nodes / m_delta), cell type codes (CellTypes,
 * cut_cells.h:26-168) and east / north / top face fractions (m_theta) of the N*M*K cells. */
int lsx_write_tecplot(const char* path, int N, int M, int K, const double* xc, const double* yc, const double* zc,
                      const double* dx, const double* dy, const double* dz, const int* type, const double* thx,
                      const double* thy, const double* thz, const double* u, const double* v, const double* w,
                      const double* p, const double* u_west, const lsx_walls* walls, int outflow);

/* Thread-local message of the last failure ("" if none). */
const char* lsx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* LSX_H_ */
