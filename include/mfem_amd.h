/*
 * mfem_amd.h -- C ABI of the MI355X (gfx950) per-timestep soft-body solve.
 *
 * Drop-in boundary for warp-mfem's RefinementSolver.step
 * (src/mfem/refinement/solver.py:429-579 in the reference).  The reference is
 * pure Python/Warp, so no C FFI exists there; these are the entry points a
 * ctypes binding of that path binds (see INTEGRATION.md for the stub).  The
 * Python mirror of the reference's solver API lives in
 * warp-mfem_amd/mfem_amd/solver.py and calls exactly these functions.
 *
 * Conventions
 *   - Every pointer inside mfem_solver is a device pointer the caller owns
 *     (the Python host allocates them as torch tensors on the GPU).  The
 *     library allocates nothing.
 *   - Every entry point enqueues work on `stream` (a hipStream_t, passed as
 *     void*) and returns without synchronising.
 *   - Return value: 0 on success, <0 on error (MFEM_E*); the message is
 *     available from mfem_last_error() on the calling thread.
 *   - Device-side anomalies (grid-barrier timeout, CG not converged, line
 *     search rejected) do not fail the call: they are recorded in
 *     status[] / stats[] (read them after synchronising), as the reference
 *     ignores CG non-convergence (solver.py:607) and line-search failure
 *     (line_search.py:143-161).
 */
#ifndef MFEM_AMD_H
#define MFEM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFEM_OK 0
#define MFEM_EINVAL -1   /* bad argument / inconsistent sizes */
#define MFEM_EHIP -2     /* a HIP runtime call failed */
#define MFEM_ECAPACITY -3 /* problem larger than the persistent kernels support */

#define MFEM_SHAPE_PLANE 0
#define MFEM_SHAPE_CAPSULE 1

/* stats[iter][k]: 0 cg_iters, 1 cg_relres (FP64, true residual), 2 ls_trials,
   3 ls_accepted, 4 alpha, 5 merit f0, 6 merit f1, 7 cg_restarts,
   8..13 PCG phase device-clock ticks (SpMV, reduce 1, update, reduce 2,
   residual checks, prologue), 14 start used (0 x~ - x; 3 the learned warm
   start's Galerkin start, the lowest quadratic-model energy over x~ - x,
   the learned guess, the last three Newton steps and the same Newton
   iteration's solution one time step earlier), 15 |g - H x0| /
   |g - H (x~ - x)| of the Galerkin start */
#define MFEM_STATS_PER_ITER 16
#define MFEM_MAX_STATS_ITERS 64
#define MFEM_RED_MAX_BLOCKS 2048
#define MFEM_MLP_NPARAMS 6288 /* 16x64 + 64 + 64x64 + 64 + 64x16 + 16 */

/* pcg_flags bits */
#define MFEM_PCG_PIPELINED 1   /* one grid reduction per CG iteration (pipelined PCG) */
#define MFEM_PCG_STREAM_ROWS 2 /* stream matrix rows from HBM instead of keeping them in VGPRs */
#define MFEM_PCG_LEARNED_GUESS 4 /* start from the lowest-energy combination of x~ - x, the learned guess, the last three Newton steps and the previous step's same iteration (needs dx_hist; with guess null the learned guess is left out and no network runs) */
#define MFEM_PCG_COARSE 8 /* block Jacobi plus a workgroup-local rigid-mode correction (pipelined, resident rows;
                             see coarse_w): M^-1 = D^-1 + P (P^T H_aa P)^+ P^T per 256-row workgroup */
#define MFEM_PCG_COOPERATIVE 16 /* launch the persistent kernels (PCG, line search) with
                                   hipLaunchCooperativeKernel.  Off: ordinary launches of at most
                                   the resident workgroup count (the kernels synchronise through
                                   their own bounded grid barriers; a workgroup that never became
                                   resident would end the launch with MFEM_ERR_BARRIER after the
                                   barrier time limit).  A cooperative launch costs 11-20 us of
                                   idle device before it starts on MI355X. */

/* status[] words */
#define MFEM_STATUS_ERRBITS 0 /* bit 0: grid barrier timed out; every kernel of a step (the persistent
                                 PCG and line search included) then returns at once, leaving the state
                                 untouched, until status and flags are cleared */
#define MFEM_STATUS_EPOCH 1   /* grid-barrier epoch counter (internal) */
/* error bits in status[MFEM_STATUS_ERRBITS] (any bit voids the step the same way) */
#define MFEM_ERR_BARRIER 1 /* a grid barrier timed out */
#define MFEM_ERR_CONTACT 2 /* self-contact candidates overflowed or a swept primitive spanned too many cells
                              (sc_count[2], sc_count[3] hold the MFEM_BP_* bits) */

/*
 * Solver description: sizes, parameters and device buffers.
 * Field meanings follow RefinementSolver's constructor kwargs
 * (src/mfem/refinement/solver.py:56-128, 241-245) and its buffers
 * (solver.py:170-304, additional_state.py:17-110, contact.py:778-862).
 */
typedef struct mfem_solver {
  /* sizes */
  int32_t n_verts;
  int32_t n_tets;
  int32_t n_tris;
  int32_t n_shapes;
  int32_t n_chunks; /* ceil(n_verts / 64): SELL-64 row chunks of the system matrix */
  int32_t n_slots;  /* total block slots over all chunks (each slot = 64 lanes) */
  /* integer parameters */
  int32_t energy;     /* 0 = ARAP, 1 = Neo-Hookean (solver.py:108-122) */
  int32_t iterations; /* Newton iterations per step */
  int32_t cg_maxiter;
  int32_t precond;     /* block-Jacobi on/off (solver.py:128) */
  int32_t line_search; /* (solver.py:185) */
  int32_t ls_max;      /* line_search_max_iterations */
  int32_t tri_contact; /* triangle-vs-capsule barrier (contact.py:417-775) */
  int32_t friction_on; /* any shape with mu > 0 (contact.py:937-947) */
  /* float parameters */
  float cg_tol;
  float ls_alpha0;
  float ls_tau;
  float ls_c;
  float ls_thr;
  float sing_thr; /* preconditioner_singular_threshold */
  float attach_k; /* attachment_stiffness */
  float d0, d1, kappa, eps_v;
  float qb[3];      /* barrier(d0), barrier'(d0), barrier''(d0) */
  float gravity[3]; /* model.gravity */
  int32_t max_width; /* widest chunk, in slots (max row length of the pattern) */
  int32_t pcg_flags; /* MFEM_PCG_* bits */
  int32_t pcg_variant; /* -1: chosen from pcg_flags; k >= 0: variant k (mfem_pcg_variant_info), set by an autotuner */
  /* model / topology (device) */
  const int32_t* tets;      /* [n_tets][4] */
  const float* dminv;       /* [n_tets][9] tet_poses, row-major */
  const float* materials;   /* [n_tets][3] mu, lambda, damping */
  const float* vol;         /* [n_tets] signed rest volume 1/(6 det DmInv) */
  const float* mass;        /* [n_verts] particle_mass */
  const float* inv_mass;    /* [n_verts] particle_inv_mass */
  const float* pin_pos;     /* [n_verts][3] attachment target (only where inv_mass == 0) */
  const int32_t* tris;      /* [n_tris][3] surface triangles */
  const int32_t* shape_type; /* [n_shapes] */
  const float* shape_xf;     /* [n_shapes][7] world pose: p(3), q(4) xyzw */
  const float* shape_scale;  /* [n_shapes][3] capsule radius, half height */
  const float* shape_mu;     /* [n_shapes] friction coefficient */
  /* system-matrix pattern: SELL-64 (64 rows per chunk, slot-major) */
  const int32_t* chunk_ptr;   /* [n_chunks + 1] first slot of each chunk */
  const int32_t* sell_col;    /* [n_slots][64] column vertex of each lane (padding: 0) */
  float* sell_val;            /* [n_slots][9][64] 3x3 block values, lane-minor */
  const int32_t* diag_lane;   /* [n_verts] lane-slot index (slot*64 + lane) of the diagonal block */
  const int32_t* contrib_ptr; /* [n_slots*64 + 1] per lane-slot range into contrib_src */
  const int32_t* contrib_src; /* (pair*n_tets + tet) << 1 | transpose; pair indexes the 10 tet blocks */
  const int32_t* vinc_ptr;    /* [n_verts + 1] vertex -> incident (tet, corner) range */
  const int32_t* vinc_src;    /* tet*4 + corner */
  /* state (device) */
  float* x;       /* [n_verts][3] particle_q being solved (state_out) */
  float* v;       /* [n_verts][3] particle_qd (state_out) */
  float* s;       /* [n_tets][6] tet_stretch */
  float* lam;     /* [n_tets][6] tet_lambda */
  float* x_tilde; /* [n_verts][3] predictor */
  float* x_start; /* [n_verts][3] start-of-step positions (friction anchor, velocity) */
  float* lag_n;   /* [n_verts][3] lagged contact normal */
  float* lag_fn;  /* [n_verts] lagged normal force */
  int32_t* lag_sid; /* [n_verts] lagged contact shape */
  /* per-tet work */
  float* G;     /* [n_tets][72] constraint Jacobian, 6x12 row-major */
  float* Hl;    /* [n_tets][21] H_lambda, packed upper triangle */
  float* gl;    /* [n_tets][6] g_lambda */
  float* Hsinv; /* [n_tets][21] (H_s)^-1, packed upper triangle */
  float* gs;    /* [n_tets][6] dPsi/ds */
  float* psi;   /* [n_tets] elastic energy (solver._elastic_energy) */
  float* Kb;    /* [10][n_tets][9] unique 3x3 blocks of G^T H_lambda G */
  float* fe;    /* [n_tets][12] G^T H_lambda g_lambda */
  /* per-vertex work */
  float* gx;    /* [n_verts][3] kinetic + attachment gradient */
  float* gB;    /* [n_verts][3] barrier gradient */
  float* gD;    /* [n_verts][3] friction gradient */
  float* Hv;    /* [n_verts][9] barrier + friction Hessian block */
  float* dist;  /* [n_verts] contact distance */
  float* eB;    /* [n_verts] barrier energy */
  float* eD;    /* [n_verts] friction energy */
  float* g;     /* [n_verts][3] Newton right-hand side */
  double* dx;   /* [n_verts][3] PCG solution */
  float* ds;    /* [n_tets][6] stretch update */
  /* persistent-kernel scratch */
  float* pz;    /* [2][n_chunks*64][3] published PCG vector (double-buffered) */
  double* px;   /* [2][n_chunks*64][3] published FP64 PCG vector (residual checks, pipelined variant;
                   scratch owned by the solve: its values carry publication tags in the two lowest
                   mantissa bits) */
  double* red;  /* [2 kernels][2][MFEM_RED_MAX_BLOCKS][8] 64-bit words: epoch-tagged halves of the grid
                   reduction partials (PCG region, line-search region) */
  uint32_t* flags; /* [MFEM_RED_MAX_BLOCKS] grid barrier arrival epochs */
  int32_t* status; /* [8] see MFEM_STATUS_* */
  double* stats;   /* [MFEM_MAX_STATS_ITERS][MFEM_STATS_PER_ITER] per Newton iteration */
  /* learned warm start (used when pcg_flags has MFEM_PCG_LEARNED_GUESS) */
  const float* mlp_w;      /* [MFEM_MLP_NPARAMS] FP32 master weights (biases are read from here) */
  const uint16_t* mlp_wbf; /* [MFEM_MLP_NPARAMS] BF16 copy (matrix operands) */
  uint16_t* ws_x;          /* [n_verts][16] BF16 features */
  float* ws_aux;           /* [n_verts][10] normalised candidates c1, c2, c3 and their scale */
  float* ws_y;             /* [n_verts][3] network outputs (blend weights) */
  float* guess;            /* [n_verts][3] learned initial guess */
  /* triangle-vs-capsule contact (used when tri_contact != 0) */
  const int32_t* tcon_ptr; /* [n_slots*64 + 1] per lane-slot list of triangle vertex pairs */
  const int32_t* tcon_src; /* tri * 9 + 3 * a + b: block (tris[tri][a], tris[tri][b]) */
  const int32_t* vtri_ptr; /* [n_verts + 1] vertex -> incident triangle corners */
  const int32_t* vtri_src; /* tri * 3 + corner */
  float* tri_dist;         /* [n_tris] distance to the nearest capsule minus its radius */
  float* tri_bary;         /* [n_tris][3] barycentric weights of the closest point */
  float* tri_nrm;          /* [n_tris][3] contact normal (capsule -> triangle) */
  int32_t* tri_sid;        /* [n_tris] shape id, -1 when no capsule */
  /* batches of independent bodies (n_bodies > 1): body b owns vertices
     [body_vert[b], body_vert[b+1]) starting at a multiple of
     256 * body_wg_align (whole PCG workgroups of 4 waves x chunks per wave
     for every variant whose chunks per wave divide body_wg_align), tets
     [body_tet[b], ...) and triangles
     [body_tri[b], ...); each body has its own CG scalars, convergence,
     line search and stats ([n_bodies][MFEM_MAX_STATS_ITERS][MFEM_STATS_PER_ITER]).
     No counterpart in the reference (one body per solver there). */
  int32_t n_bodies;          /* 0 or 1: one body */
  int32_t ls_blocks;         /* line-search workgroups when n_bodies > 1 */
  const int32_t* body_wg;    /* [n_bodies + 1] chunk ranges / 4 (256-row blocks) */
  const int32_t* body_vert;  /* [n_bodies + 1] */
  const int32_t* body_tet;   /* [n_bodies + 1] */
  const int32_t* body_tri;   /* [n_bodies + 1] */
  const int32_t* body_ls_wg; /* [n_bodies + 1] line-search workgroup ranges */
  int32_t* body_epoch;       /* [n_bodies] grid-barrier epochs */
  const int32_t* tet_energy; /* [n_tets] per-tet energy (0 ARAP, 1 Neo-Hookean) or null: `energy` for all */
  int32_t body_wg_align;     /* a batch: 1, 2, 4 or 8, dividing every body_wg[b] */
  /* workgroup-local preconditioner (no counterpart in the reference, whose
     CG uses block Jacobi): 0 = block Jacobi z = D^-1 r; an even count s > 0
     = s damped block-Jacobi sweeps on the rows of the PCG workgroup,
     z <- z + omega D^-1 (r - A_loc z) from z = D^-1 r, with A_loc the matrix
     restricted to the workgroup's rows: a polynomial in D^-1 A_loc times
     D^-1, symmetric positive definite for every even s and omega > 0.
     Standard (non-pipelined) PCG variants only. */
  int32_t local_sweeps;
  float local_omega;
  /* Newton-step history for the learned warm start's Galerkin start (no
     counterpart in the reference): [2][n_chunks*64][3], the two steps
     before dx, rotated by the PCG when it writes dx; null = no history
     (and the plain start).  Buffers dx and dx_hist hold at least
     n_chunks*64 rows (the PCG reads them whole, in 64-byte lines). */
  double* dx_hist;
  /* each Newton iteration's solution in the previous time step,
     [dx_same_iters][n_chunks*64][3], written by the PCG of that iteration
     and used by the next step's (the Galerkin start's sixth candidate; no
     counterpart in the reference); null = none */
  double* dx_same;
  int32_t dx_same_iters;
  /* Workgroup-local coarse correction (pcg_flags & MFEM_PCG_COARSE; no
     counterpart in the reference, whose CG uses block Jacobi): each PCG
     workgroup's 256 rows form one aggregate a, with the 6 rigid modes of its
     vertices about their centroid as basis P_a (3 translations, 3 rotations
     scaled by the aggregate's radius), smoothed row by row as
     P'_i = P_i - w_i D_i^-1 (H_aa P)_i (H_aa: the matrix restricted to the
     aggregate).  The preconditioner is
     z = D^-1 r + P' (P'^T H_aa P')^+ P'^T r on each aggregate: symmetric
     positive definite for any weights, no communication beyond the
     workgroup.  w_i = coarse_w[i] (e.g. a network's output) or
     coarse_omega for every vertex when coarse_w is null. */
  float* coarse_w; /* [n_verts] or null (written by mfem_coarse_weights) */
  float coarse_omega;
  /* the network that gives coarse_w (no counterpart in the reference): at
     the first Newton iteration of a step, when cw_wbf is set and
     MFEM_PCG_COARSE is on, mfem_step evaluates mfem_coarse_features into
     cw_x, the per-vertex MLP (same architecture as the warm start's, its own
     weights) into cw_y, and coarse_w = clamp(cw_y[i][0], 0, 2) */
  uint16_t* cw_x;          /* [n_verts][16] BF16 features */
  float* cw_y;             /* [n_verts][3] network outputs */
  const float* cw_w;       /* [MFEM_MLP_NPARAMS] FP32 master weights */
  const uint16_t* cw_wbf;  /* [MFEM_MLP_NPARAMS] BF16 copy */
  /* self-contact: the IPC barrier b(d) = -(d - sc_dhat)^2 ln(d / sc_dhat)
     (src/mfem/collision.py:529-546) on the distance of every point-triangle
     and edge-edge pair of surface primitives (closest points of
     collision.py:220-322, evaluated in FP64) below sc_dhat, inside the Newton
     solve: energy in the merit, gradient in g, the pair's Gauss-Newton
     Hessian kappa max(b'', 0) grad d grad d^T on its vertices' diagonal
     blocks, candidates from the spatial-hash broad phase on the boxes swept
     by the Newton step, and the line search's first step cut by a
     conservative time of impact (additive CCD) so that no pair reaches
     zero distance.  The reference's solve leaves this hook commented out
     (src/mfem/solver.py:443-447).  Single body only (n_bodies <= 1). */
  int32_t self_contact;      /* 1: on */
  int32_t n_sc_points;       /* surface vertices (query points of the point-triangle pairs) */
  int32_t n_sc_edges;        /* surface edges */
  int32_t sc_max_pairs;      /* capacity of each candidate list */
  int32_t sc_max_span;       /* broad-phase cells per axis a swept primitive may span (1..8) */
  float sc_dhat;             /* barrier activation distance */
  float sc_kappa;            /* barrier stiffness */
  float sc_cell;             /* broad-phase cell size */
  float sc_ccd_s;            /* CCD: a pair keeps at least this fraction of its distance (0.1) */
  const int32_t* sc_points;  /* [n_sc_points] vertex ids; the triangles are tris[] */
  const int32_t* sc_edges;   /* [n_sc_edges][2] */
  int32_t* sc_pt;            /* [sc_max_pairs][2] (vertex, triangle) candidates */
  int32_t* sc_ee;            /* [sc_max_pairs][2] (edge, edge) candidates, e1 < e2 */
  int32_t* sc_count;         /* [4] point-triangle and edge-edge counts, their broad-phase status (MFEM_BP_*) */
  double* sc_g;              /* [n_verts][3] barrier gradient (FP64 sums) */
  double* sc_h;              /* [n_verts][6] barrier Hessian diagonal blocks, sym6 order 00 11 22 01 02 12 */
  float* sc_alpha;           /* [1] largest line-search step the CCD allows (<= 1) */
  void* sc_ws;               /* broad-phase workspace */
  int64_t sc_ws_bytes;
  /* membrane shell over the surface triangles (src/mfem/refinement/
     shell_energy.py): per triangle E = mu_s A |F - R(F)|^2 / 2, F = [x1 - x0,
     x2 - x0] W the 3x2 in-plane deformation gradient, R its polar factor, A
     and W = Dm^-1 from the rest positions; its gradient joins g and the
     slope, its energy the merit, and its constant Gauss-Newton Hessian
     mu_s A (w_a . w_b) I3 on the triangle's vertex pairs joins H
     (solver.py:297-302, 638-639, 789-797, 864-867, 910-911).  tri_W,
     tri_area and shell_h are written by mfem_shell_refresh. */
  int32_t shell_on;          /* 1: on (needs n_tris > 0 and the tcon / vtri lists) */
  float shell_mu;            /* surface shear modulus mu_s in Pa m (shell_mu kwarg) */
  const float* rest;         /* [n_verts][3] rest positions (AdditionalState.rest_particle_q) */
  float* tri_W;              /* [n_tris][4] W row-major (0 for a degenerate rest triangle) */
  float* tri_area;           /* [n_tris] rest area (0 for a degenerate rest triangle) */
  float* shell_h;            /* [n_slots*64] per lane-slot sum of mu_s A w_a . w_b (block = value I3) */
  float* gS;                 /* [n_verts][3] shell gradient */
  /* 1: the element kernel stores only what assembly needs (K blocks,
     forces, energies) and the line search's first pass runs the local solve
     (lambda, ds) from the element recomputed at the same positions,
     forming its merit and slope terms there; mfem_local_solve is then a
     no-op.  Needs line_search.  0: G, H_l, g_l, H_s^-1, g_s are stored by
     mfem_update_system and read by mfem_local_solve and the line search
     (the reference's stage order, solver.py:533-557). */
  int32_t fused_local;
} mfem_solver;

/* ABI self-description (checked against the Python mirror by the tests) */
int64_t mfem_solver_sizeof(void);
int64_t mfem_solver_offsetof(const char* field);
int mfem_version(void);
int mfem_wall_clock_khz(void); /* rate of the device clock used in stats[8..13], -1 on error */
const char* mfem_last_error(void);

/* AdditionalState.from_model stretch init: s = sym6(S(F(x)))
   (src/mfem/refinement/kernels.py:55-75, additional_state.py:88-97) */
int mfem_init_stretch(const mfem_solver* S, void* stream);

/* Start of step: predictor x~ = x + v dt + g dt^2/2, friction anchor and
   lagged normal force (kernels.py:77-91; solver.py:499-520;
   contact.py:937-967). */
int mfem_begin_step(const mfem_solver* S, float dt, void* stream);

/* _update_system (solver.py:633-639): kinetic + attachment gradient
   (kernels.py:172-185, solver.py:946-947), vertex contact + friction
   (contact.py:969-1048), elastic energy / gradient / PSD Hessian
   (elastic_energy.py), constraints c, G (kernels.py:131-169) and the Schur
   blocks H_lambda, g_lambda (kernels.py:189-249). */
int mfem_update_system(const mfem_solver* S, float dt, void* stream);

/* Shell.refresh (src/mfem/refinement/shell_energy.py:137-186, 262-293):
   tri_W and tri_area of every triangle from the rest positions, then
   shell_h of every lane-slot from its tcon list.  Call once per topology
   (construction and after every refine pass that changed the mesh), with
   shell_on set. */
int mfem_shell_refresh(const mfem_solver* S, void* stream);

/* _assemble_system (solver.py:641-799): H into the SELL pattern, g. */
int mfem_assemble(const mfem_solver* S, float dt, void* stream);

/* _global_solve (solver.py:581-609): block-Jacobi PCG on H dx = g from the
   guess x~ - x (solver.py:530-531), FP64 reductions, one persistent launch.
   Writes stats[iter][0..1] and stats[iter][7]. */
int mfem_global_solve(const mfem_solver* S, int newton_iter, void* stream);

/* _local_solve (solver.py:611-631, kernels.py:252-283); lam <- new lambda. */
int mfem_local_solve(const mfem_solver* S, void* stream);

/* _line_search (solver.py:871-928, line_search.py:76-163), FP64 merit sums,
   one persistent launch; or x += dx, s += ds when line_search == 0
   (solver.py:555-557).  Writes stats[iter][2..6]. */
int mfem_line_search(const mfem_solver* S, float dt, int newton_iter, void* stream);

/* Velocity v = (x - x_start) / dt (kernels.py:351-363, solver.py:565-577). */
int mfem_end_step(const mfem_solver* S, float dt, void* stream);

/* RefinementSolver.step without refinement (solver.py:429-579): all of the
   above in order, no host synchronisation. */
int mfem_step(const mfem_solver* S, float dt, void* stream);

/* Learned warm start, no counterpart in the reference (its guess is x~ - x,
   solver.py:530-531).  Per vertex: candidates c1 = x~ - x, c2 = D^-1 (g - H c1)
   (one block-Jacobi step), c3 = the previous Newton step, normalised by
   s = |c1| + |c2| + |c3|, and 16 rotation-invariant features (BF16) into
   ws_x / ws_aux; feature 15 is the line-search step of Newton iteration
   newton_iter - 1 (read from stats). */
int mfem_ws_features(const mfem_solver* S, float dt, int newton_iter, void* stream);
/* guess = c3 + s (y0 c1 + y1 c2 + y2 c3) from the network outputs ws_y:
   the previous Newton step plus a learned correction. */
int mfem_ws_apply(const mfem_solver* S, void* stream);

/* Features of the coarse-weight network (no counterpart in the reference),
   from the assembled system, per vertex i of aggregate a (the 256-row block
   of PCG workgroup a), BF16 into S->cw_x[i][0..15]: 1, the share of the
   row's off-diagonal block norms that couple outside a, the mass term's
   share of tr D_i, the contact Hessian's share of tr D_i, |x_i - centroid|
   over a's RMS radius, off-diagonal over diagonal block norms, pinned,
   log10(tr D_i / mass term) / 4, blocks in the row / 16, blocks outside a /
   16, contact active (distance < d1), zeros. */
int mfem_coarse_features(const mfem_solver* S, float dt, void* stream);
/* coarse_w[i] = clamp(cw_y[i][0], 0, 2) for the active vertices. */
int mfem_coarse_weights(const mfem_solver* S, void* stream);

/* Per-vertex MLP (16 -> 64 -> 64 -> 3, GELU, BF16 MFMA with FP32 accumulate):
   y[n][0..2] for the BF16 inputs x[n][16]. */
int mfem_mlp_forward(const uint16_t* x, int n, const uint16_t* w_bf16, const float* w_f32, float* y, void* stream);
/* One Adam step on the loss (1/n) sum_n |y0 c1 + y1 c2 + y2 c3 - t|^2 with
   aux[n] = (c1, c2, c3, t): per-workgroup gradients into part (at most
   max_parts workgroups), fixed-order sum, Adam (beta 0.9/0.999, eps 1e-8,
   step >= 1) on w_f32, BF16 refresh; lr = 0 only computes the gradient into
   grad_out.  loss_out / grad_out may be null. */
int mfem_mlp_train_step(const uint16_t* x, const float* aux, int n, float* w_f32, uint16_t* w_bf16, float* adam_state,
                        float* part, int max_parts, float* part_loss, float* loss_out, float* grad_out, float lr,
                        int step, void* stream);
/* One Adam step from a given output gradient dy[n][3] = dL/dy (any loss on
   the network outputs, e.g. the relative initial CG residual): backward pass,
   fixed-order sum of the per-workgroup gradients, Adam, BF16 refresh.
   lr = 0 only computes the gradient into grad_out (may be null). */
int mfem_mlp_backward_step(const uint16_t* x, const float* dy, int n, float* w_f32, uint16_t* w_bf16,
                           float* adam_state, float* part, int max_parts, float* grad_out, float lr, int step,
                           void* stream);
const char* mfem_mlp_last_error(void);

/* ---------------------------------------------------------------------
   Self-contact building blocks (mfem_collide.hip)
   --------------------------------------------------------------------- */
#define MFEM_BP_SPAN 1    /* a primitive spans more than max_span cells per axis (not inserted / queried) */
#define MFEM_BP_ENTRIES 2 /* hash-table entry capacity exceeded */
#define MFEM_BP_PAIRS 4   /* more candidate pairs than max_pairs (n_pairs holds the full count) */

/* Exclusive (inclusive != 0: inclusive) prefix sum of n int32 values;
   total (device, may be null) receives the sum.  Replaces TiledScan /
   wp.utils.array_scan for int32 (src/mfem/scan.py:21-193,
   refinement/refinement.py:1185-1186). */
int mfem_scan_i32(const int32_t* in, int32_t* out, int n, int inclusive, int32_t* total, void* workspace,
                  size_t workspace_bytes, void* stream);
size_t mfem_scan_workspace_bytes(int n);

/* mfem_scan_i32 without a workspace argument and without any allocation (no
   counterpart in the reference): for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc).  Its tile sums live in a fixed library-held scratch
   array (one per GPU, ~256 KiB, held until the process ends), so n is at
   most MFEM_SCAN_NOALLOC_MAX (larger n: MFEM_ECAPACITY, nothing launched).
   Calls that overlap on one GPU (two streams at once) race on that scratch;
   calls in one stream order are safe. */
#define MFEM_SCAN_NOALLOC_MAX (1 << 26)
int mfem_scan_i32_noalloc(const int32_t* in, int32_t* out, int n, int inclusive, int32_t* total, void* stream);

/* Spatial-hash broad phase: all point-triangle pairs (points[i], tri) whose
   boxes inflated by d_hat / 2 overlap, the point not a vertex of the
   triangle, as (vertex id, tri index) sorted by query order then tri index;
   edge-edge: all (e1, e2), e1 < e2, no shared vertex.  cell = grid cell
   size (>= the largest inflated primitive extent keeps max_span = 2);
   n_pairs (device) = total count, status (device) = MFEM_BP_* bits.  The
   reference has no live broad phase (its SAP and hash table are unwired,
   src/mfem/collision.py:840-895, refinement/geometry_hash.py:8-90). */
int mfem_broadphase_pt(const float* x, const int32_t* points, int n_points, const int32_t* tris, int n_tris,
                       float d_hat, float cell, int max_span, int32_t* pairs, int max_pairs, int32_t* n_pairs,
                       int32_t* status, void* workspace, size_t workspace_bytes, void* stream);
int mfem_broadphase_ee(const float* x, const int32_t* edges, int n_edges, float d_hat, float cell, int max_span,
                       int32_t* pairs, int max_pairs, int32_t* n_pairs, int32_t* status, void* workspace,
                       size_t workspace_bytes, void* stream);
size_t mfem_broadphase_workspace_bytes(int n_insert, int n_query, int max_span);

/* The broad phase on swept primitives (kind 0: point-triangle, prims =
   points, tris = triangles; kind 1: edge-edge, prims = edges, tris unused):
   each vertex's box holds it at x and at x + dx (dx FP64, may be null for
   the static boxes of mfem_broadphase_pt / _ee), so that the pairs cover
   every step size in [0, 1] of the Newton step dx. */
int mfem_broadphase_swept(int kind, const float* x, const double* dx, const int32_t* prims, int n_prims,
                          const int32_t* tris, int n_tris, float d_hat, float cell, int max_span, int32_t* pairs,
                          int max_pairs, int32_t* n_pairs, int32_t* status, void* workspace, size_t workspace_bytes,
                          void* stream);

/* Self-contact stages of mfem_step (mfem_selfcontact.hip; the fields
   self_contact .. sc_ws_bytes of mfem_solver).  No counterpart in the
   reference's live solve (the hook at src/mfem/solver.py:443-447 is
   commented out); the distance primitives and barrier follow
   src/mfem/collision.py:220-322, 529-546.
   _candidates: point-triangle and edge-edge candidate lists from the
   broad phase, on static boxes (swept = 0) or on the boxes swept by dx;
   _eval: gradient and diagonal Hessian blocks into sc_g / sc_h;
   _ccd: sc_alpha = the largest first line-search step (<= 1) that keeps
   every candidate pair at a positive distance (additive CCD). */
size_t mfem_self_contact_workspace_bytes(int n_points, int n_tris, int n_edges, int max_span);
int mfem_self_contact_candidates(const mfem_solver* S, int swept, void* stream);
int mfem_self_contact_eval(const mfem_solver* S, void* stream);
int mfem_self_contact_ccd(const mfem_solver* S, void* stream);
const char* mfem_self_contact_last_error(void);

/* eval_soft_rigid_distance (src/mfem/collision.py:380-476): for rigid/soft
   candidate triangle pairs, per soft particle the minimum of the 15 feature
   distances below d_hat (distance pre-filled with a sentinel >= d_hat). */
int mfem_soft_rigid_distance(const float* points, const int32_t* tris, int rigid_points_count, int rigid_tri_count,
                             const int32_t* pairs, const int32_t* pair_count, int max_pairs, float d_hat,
                             float* distance, void* stream);

/* IPC barrier kappa * b(d), b = -(d - d_hat)^2 ln(d / d_hat)
   (src/mfem/collision.py:529-546), on point-triangle (kind 0, pairs =
   (vertex, tri)) or edge-edge (kind 1, pairs = (edge, edge)) candidates:
   per pair the distance, the energy (FP64) and its gradient with respect to
   the pair's four vertices (grad[k][4][3], vertex ids in vids[k][4]). */
int mfem_pair_barrier(int kind, const float* x, const int32_t* pairs, const int32_t* n_pairs, int max_pairs,
                      const int32_t* tris, const int32_t* edges, float kappa, float d_hat, float* dist,
                      double* energy, float* grad, int32_t* vids, void* stream);
const char* mfem_collide_last_error(void);

/* ---------------------------------------------------------------------
   Adaptive refinement, one pass (mfem_refine.hip): replaces refine() of
   src/mfem/refinement/refinement.py:914-1322 ("legacy" scoring).  The
   reference's pass is race-dependent; this one is deterministic: greedy
   selection of split edges by (score, edge key), at most one split edge per
   tet, computed as iterated local maxima.  Call sequence:
     mfem_refine_select(phase 0)            scores, candidate table
     mfem_refine_select(phase 1) repeatedly  until *changed == 0
     mfem_refine_count                      totals[0..2] = tets, new vertices, triangles after the pass
     (host: drop the pass when the totals exceed the capacities)
     mfem_refine_apply                      new vertices appended in place, tets / triangles / per-tet data to O
   --------------------------------------------------------------------- */
#define MFEM_REFINE_TABLE_FULL 1

typedef struct mfem_refine_args {
  const int32_t* tets;         /* [n_tets][4] */
  const int32_t* tris;         /* [n_tris][3] */
  const float* x;              /* [n_verts][3] positions (scores, midpoint test) */
  const float* inv_mass;       /* [n_verts] 0 = pinned; edges between pinned vertices are never split */
  const float* tet_score;      /* [n_tets] (the solver passes the elastic energy) */
  const float* vertex_score;   /* [n_verts] (the solver passes the contact barrier energy) */
  const int32_t* shape_type;   /* [n_shapes] midpoint penetration test */
  const float* shape_xf;       /* [n_shapes][7] */
  const float* shape_scale;    /* [n_shapes][3] */
  const float* stretch;        /* [n_tets][6] */
  const float* lam;            /* [n_tets][6] */
  const float* materials;      /* [n_tets][3] */
  const float* dminv;          /* [n_tets][9] */
  int32_t n_verts, n_tets, n_tris, n_shapes;
  int32_t max_tets, max_tris, max_verts;
  float threshold;             /* min_refine_score */
  float w_tet, w_vertex;       /* refine_tet_score_weight, refine_vertex_score_weight */
  float penetrating_score;     /* score of an edge whose midpoint is inside a shape */
  float split_t;               /* refine_split_position */
  /* scoring 1 = "geometric" (refinement.py:123-317): per-edge maximum over
     incident tets and surface triangles of
     (L / h_min)(L / L_max)(w_e s_elastic + w_c s_contact), edges shorter
     than 2 h_min never split; threshold is then refine_geometric_threshold.
     s_elastic = max(0, rho / mean(rho) - 1) with rho = max(psi, 0) / (vol mu)
     (tet_score = psi), s_contact = max(0, (d1 - d) / d1) over endpoints whose
     nearest shape is a capsule (vertex_dist / vertex_sid), the triangle term
     w_t (d1 - d_tri) / d1 (1 - bary_k) on the edge opposite corner k
     (tri_dist / tri_bary, null = no triangle term).  0 = "legacy". */
  int32_t scoring;
  float min_edge_length, contact_d1, w_elastic, w_vertex_contact, w_tri_contact;
  const float* vertex_dist;    /* [n_verts] signed distance to the nearest shape */
  const int32_t* vertex_sid;   /* [n_verts] that shape, -1 none */
  const float* tri_dist;       /* [n_tris] triangle - capsule distance (or null) */
  const float* tri_bary;       /* [n_tris][3] barycentric weights of the closest point */
} mfem_refine_args;

typedef struct mfem_refine_out {
  float* x;           /* [max_verts][3] new vertices appended at n_verts.. (may alias args.x) */
  float* v;           /* [max_verts][3] */
  float* rest;        /* [max_verts][3] rest positions */
  int32_t* tets;      /* [max_tets][4] */
  float* stretch;     /* [max_tets][6] */
  float* lam;         /* [max_tets][6] */
  float* materials;   /* [max_tets][3] */
  float* dminv;       /* [max_tets][9] */
  int32_t* tris;      /* [max_tris][3] */
} mfem_refine_out;

size_t mfem_refine_workspace_bytes(int max_tets, int max_tris);
int mfem_refine_select(const mfem_refine_args* args, void* workspace, size_t workspace_bytes, int32_t* status,
                       int32_t* changed, int phase, void* stream);
int mfem_refine_count(const mfem_refine_args* args, void* workspace, size_t workspace_bytes, int32_t* totals,
                      void* stream);
int mfem_refine_apply(const mfem_refine_args* args, void* workspace, size_t workspace_bytes,
                      const mfem_refine_out* out, void* stream);
const char* mfem_refine_last_error(void);

/* Persistent-PCG variants for an autotuner, no counterpart in the reference:
   info[0..4] = chunks per wave, resident row slots (0 = streamed rows),
   mode (0 standard, 1 pipelined, 2 standard re-reading the preconditioner
   and right-hand side to run four waves per SIMD, 3 pipelined with the
   workgroup-local coarse correction), workgroups for S, applicable to S
   (0/1: co-resident, rows fit the slots, a batch's bodies aligned to its
   workgroups, mode 3 exactly when pcg_flags has MFEM_PCG_COARSE).  Reads
   only n_chunks, max_width, n_bodies, body_wg_align, local_sweeps and
   pcg_flags of S (a host-only struct may ask before any buffer exists).
   Returns MFEM_EINVAL when k >= mfem_pcg_variant_count(). */
/* the PCG variant index the calling thread's last mfem_global_solve (or
   mfem_step) launched, -1 before any (profiling: per-dispatch counters) */
int mfem_last_pcg_variant(void);
int mfem_pcg_variant_count(void);
/* Co-resident workgroups of the persistent line-search kernel (the bound
   for ls_blocks of a batch). */
int mfem_line_search_max_blocks(void);
int mfem_pcg_variant_info(const mfem_solver* S, int k, int32_t* info);

/* System-matrix pattern and assembly gather lists on the GPU, after a
   topology change (replaces the warp.sparse topology rebuild,
   src/mfem/refinement/solver.py:685-715): the same arrays, in the same
   order, as a sequential construction (all lists come from sorted 64-bit
   keys).  Two calls with one workspace: mfem_pattern_count writes rowptr
   [n_verts + 1], chunk_ptr [ceil(n_verts / 64) + 1] and counts = {nnz,
   n_slots}; the host sizes the outputs from counts, then mfem_pattern_fill
   writes cols [nnz], sell_col [n_slots * 64], diag_lane [n_verts],
   contrib_ptr [n_slots * 64 + 1], contrib_src [16 n_tets], vinc_ptr
   [n_verts + 1], vinc_src [4 n_tets], tcon_ptr [n_slots * 64 + 1],
   tcon_src [9 n_tris], vtri_ptr [n_verts + 1], vtri_src [3 n_tris];
   status[0] = 1 when a triangle edge is not a tet edge. */
size_t mfem_pattern_workspace_bytes(int max_verts, int max_tets, int max_tris);
int mfem_pattern_count(const int32_t* tets, int n_tets, int n_verts, int max_verts, int max_tets, int max_tris,
                       void* workspace, size_t workspace_bytes, int32_t* rowptr, int32_t* chunk_ptr, int32_t* counts,
                       void* stream);
int mfem_pattern_fill(const int32_t* tets, int n_tets, const int32_t* tris, int n_tris, int n_verts, int max_verts,
                      int max_tets, int max_tris, void* workspace, size_t workspace_bytes, const int32_t* rowptr,
                      const int32_t* chunk_ptr, int n_slots, int32_t* cols, int32_t* sell_col, int32_t* diag_lane,
                      int32_t* contrib_ptr, int32_t* contrib_src, int32_t* vinc_ptr, int32_t* vinc_src,
                      int32_t* tcon_ptr, int32_t* tcon_src, int32_t* vtri_ptr, int32_t* vtri_src, int32_t* status,
                      void* stream);
const char* mfem_pattern_last_error(void);

/* Contact.compute_distance (src/mfem/refinement/contact.py:98-164, 625-681):
   per vertex the signed distance to the nearest shape and its id (vdist,
   vsid: [n_verts]); with triangle contact on, the triangle - capsule
   distance, closest-point barycentric weights, normal and capsule id into
   S->tri_dist / tri_bary / tri_nrm / tri_sid.  At the bound state x. */
int mfem_contact_distance(const mfem_solver* S, float* vdist, int32_t* vsid, void* stream);

/* World poses of the rigid shapes for this step (contact.py:74-82,
   query_min_signed_distance): shape_xf[s] = body_q[shape_body[s]] *
   shape_local[s] (p(3), q(4) xyzw; rotation of the local pose applied
   first) for a shape attached to body shape_body[s] < n_body, and
   shape_local[s] for a static shape (shape_body[s] < 0).  All device
   pointers. */
int mfem_pose_shapes(const float* body_q, int n_body, const int32_t* shape_body, const float* shape_local,
                     int n_shapes, float* shape_xf, void* stream);

/* y = H x with the assembled SELL matrix (warp.sparse.bsr_mv). */
int mfem_spmv(const mfem_solver* S, const float* x, float* y, void* stream);

/* Line-search merit f(x + alpha dx, s + alpha ds) with lambda fixed
   (solver.py:803-869) into out[0] (device double). */
int mfem_merit(const mfem_solver* S, float dt, float alpha, double* out, void* stream);

/* ------------------------------------------------------------------------
 * Point-to-surface queries of the surface-tracking / correspondence losses
 * (src/mfem/refinement/surface_loss.py:202-245, 456-502), FP64.
 * ------------------------------------------------------------------------ */
/* The k (1..8) nearest of the n_cand vertices verts[vmap[j]] (vmap null:
   verts[0..n_cand)) to every query point, ascending by squared distance,
   ties to the lower index: out_idx / out_d2 [n_query][k] (index -1 and
   DBL_MAX when fewer than k candidates).  Replaces the cKDTree queries at
   surface_loss.py:214-220 and :486-490. */
int mfem_surface_knn(const double* query, int n_query, const double* verts, const int32_t* vmap, int n_cand, int k,
                     int32_t* out_idx, double* out_d2, void* stream);
/* Squared distance of every query point to the triangle mesh (verts, faces)
   over the triangles incident to its k nearest vertices (vert_faces:
   [n_verts][deg] padded with -1; face_ok: [n_faces] 0/1 or null), the
   smallest in neighbour-then-incidence order; without a usable triangle,
   the nearest vertex's squared distance and out_face = -1.  out_face and
   out_cp ([n_query][3] closest points) may be null.  Replaces the face
   projection at surface_loss.py:222-245 and :491-502. */
int mfem_surface_project(const double* query, int n_query, const double* verts, const int32_t* faces,
                         const int32_t* vert_faces, int deg, const uint8_t* face_ok, const int32_t* knn_idx,
                         const double* knn_d2, int k, double* out_sq, int32_t* out_face, double* out_cp,
                         void* stream);
const char* mfem_surface_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MFEM_AMD_H */
