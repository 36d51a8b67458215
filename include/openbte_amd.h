/*
 * libopenbte_amd — C ABI of the MI355X-native steady-state BTE path.
 *
 * Every function returns 0 on success and a negative OBTE_E* code on
 * failure; obte_last_error() returns the calling thread's last message.
 * Device pointers are plain HIP device addresses (e.g. torch
 * tensor.data_ptr()); every launch takes the caller's hipStream_t.  No entry
 * point allocates device memory: workspaces come in through the plan.
 * All floating-point data is FP64, all indices int32 unless noted.
 *
 * The reference path is pure Python; its "FFI" for this path is the pair of
 * Python entries BTE_RTA (openbte/bte.py:19-27) and Fourier
 * (openbte/fourier.py:15-24).  Each entry below replaces one numerical stage
 * inside those functions; the ctypes binding lives in openbte_amd/_lib.py.
 */
#ifndef OPENBTE_AMD_H
#define OPENBTE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* obte_stream_t; /* hipStream_t */

enum {
  OBTE_OK = 0,
  OBTE_EINVAL = -1,
  OBTE_ENOMEM = -2,
  OBTE_EHIP = -3,
  OBTE_EINTERNAL = -4,
  OBTE_ELIMIT = -5
};

const char* obte_last_error(void);

/* ---- host scheduling (replaces the implicit ordering inside SuperLU,
 *      openbte/bte.py:183-185) ------------------------------------------- */

/* Topological sweep order of every direction's upwind graph with the
 * periodic inflow faces (cut_candidate[f] != 0) cut.  i_mat/j_mat/k_mat are
 * the directed internal faces of openbte/bte.py:48-54 (k = n A / V_row).
 * max_width > 0 caps the cells of one level (0: longest-path levels). */
int obte_sched_create(int32_t n, int64_t n_faces, int32_t dim, const int32_t* i_mat,
                      const int32_t* j_mat, const double* k_mat, const uint8_t* cut_candidate,
                      int32_t n_dirs, const double* dirs, int32_t max_width, void** handle);
/* sizes[3*l + {0,1,2}] = levels, inflow entries, cut faces of direction l. */
int obte_sched_sizes(void* handle, int64_t* sizes);
int obte_sched_export(void* handle, int32_t* order, int32_t* pos, int32_t* lvl_ptr, int32_t* in_ptr,
                      int32_t* in_src, int32_t* in_face, int32_t* cut_exit, int32_t* cut_face);
/* Fixed-width inflow slots nb_src / nb_face ([L][kin][n] each, the layout of
 * obte_bte_plan_t) from the same schedule; kin >= the most inflow faces of
 * one cell, else OBTE_EINVAL. */
int obte_sched_export_slots(void* handle, int32_t kin, int32_t* nb_src, int32_t* nb_face);
int obte_sched_destroy(void* handle);
/* Threads of the host-side parallel loops (scheduling) that the calling
 * thread starts from now on (n >= 1).  No counterpart in the reference (its
 * preparation is serial): for callers that prepare several problems in
 * concurrent threads, which would otherwise each start one host thread per
 * core. */
int obte_set_host_threads(int32_t n);

/* ---- Fourier multigrid setup (host) ------------------------------------ */
/* Smoothed-aggregation hierarchy of the symmetric Fourier operator (the rows
 * of obte_fourier_plan_t: diag, off-diagonal weights w over row_ptr / col, in
 * a cell order where every `agg` consecutive rows form an aggregate, e.g. a
 * Morton order of the centroids).  P = (I - omega D^-1 A) Pt with Pt
 * piecewise constant on levels with at most smooth_max_row off-diagonal
 * entries per row on average (P = Pt on denser levels), coarse operator
 * P^T A P, levels until one has at most coarse_max rows.  Replaces the SuperLU factorisation of openbte/fourier.py:
 * 109-118 as the solver of the correction loop (no counterpart in the
 * reference).  Host memory only; the handle is freed by obte_amg_destroy. */
int obte_amg_create(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* w, const double* diag,
                    int32_t agg, int32_t coarse_max, double omega, int32_t smooth_max_row, void** handle);
/* *n_levels = levels including the coarsest; sizes[4*l + {0,1,2,3}] = rows,
 * off-diagonal entries, entries of P, entries of R of level l (sizes may be
 * NULL; P and R are empty on the coarsest level). */
int obte_amg_sizes(void* handle, int32_t* n_levels, int64_t* sizes);
/* Copy level `level` out (any pointer may be NULL to skip that array). */
int obte_amg_export(void* handle, int32_t level, int32_t* a_ptr, int32_t* a_col, double* a_w, double* a_diag,
                    int32_t* p_ptr, int32_t* p_col, double* p_val, int32_t* r_ptr, int32_t* r_col, double* r_val);
/* The coarsest level as a dense row-major matrix (rows x rows). */
int obte_amg_export_dense(void* handle, double* dense);
int obte_amg_destroy(void* handle);

/* ---- device plan of one BTE solve ----------------------------------- */

typedef struct obte_bte_plan {
  /* sizes */
  int32_t n;          /* cells */
  int32_t n_dirs;     /* L */
  int32_t n_mfp;      /* M */
  int32_t dim;        /* 2 or 3 */
  int32_t mc;         /* MFP bins per chunk: 4, 8, 16, 32 or 64 */
  int32_t n_chunks;   /* ceil(M / mc) */
  int32_t block;      /* threads per sweep workgroup: 64, 128 or 256 */
  int32_t n_cut;      /* total cut faces over all directions */
  int32_t n_wall;     /* diffuse wall sides */
  int32_t n_contact;  /* sides of the kappa contact */
  int32_t n_faces;    /* directed internal faces */
  int32_t contact_kind; /* 0 none, 1 periodic (e1,e2,alpha), 2 mixed (e1, alpha) */
  double normalization; /* alpha-hat of EffectiveThermalConductivity */
  double contact_jump;  /* alpha of openbte/bte.py:151,156 */

  /* schedule (per direction, concatenated; see obte_sched_export) */
  const int32_t* order;    /* [L][n] cell at position */
  const int32_t* pos;      /* [L][n] position of cell */
  const int32_t* lvl_off;  /* [L+1] offsets into lvl_ptr */
  const int32_t* lvl_ptr;  /* level start positions */
  int32_t kin;             /* inflow slots per position (nv - 1) */
  int32_t ncut_max;        /* max cut faces of one direction */
  int32_t lvl_pair_max;    /* max over directions and k of |level k| + |level k+1| (0: unknown) */
  int32_t sweep_window;    /* 1: sweeps may keep recent levels in an LDS window (needs
                              lvl_pair_max), 0: every upwind value is read from t */
  const int32_t* nb_src;   /* [L][kin][n] upstream row: position (< n), interface row n + cut,
                              or the position itself with a zero coefficient for unused slots */
  const int32_t* nb_face;  /* [L][kin][n] directed face of the slot, -1 if unused */
  const int32_t* cut_off;  /* [L+1] cut offsets */
  const int32_t* cut_exit; /* position of the upstream (exit) cell of each cut */

  /* geometry / material */
  const double* dirs;      /* [L][dim] s_l */
  const double* k_face;    /* [n_faces][dim] n A / V_row */
  const double* jump_face; /* [n_faces] periodic jump with the sign of openbte/bte.py:106,110 */
  const int32_t* cf_ptr;   /* [n+1] directed faces whose row is cell c */
  const int32_t* cf_face;
  const int32_t* wall_ptr; /* [n+1] wall entries of cell c (diffuse and mixed) */
  const int32_t* wall_side;/* diffuse index (>= 0), -1 for a mixed wall, or -2 - k for the
                              reflecting side k of a partially transmitting interface
                              (diffuse index k, no outflow into D_l: the interface face
                              already counts it; wall_temp then holds 1 - t) */
  const double* wall_db;   /* [n_wallent][dim] n A / V */
  const double* wall_temp; /* [n_wallent] T_wall of mixed walls, 1 - t of interface
                              sides, 0 for diffuse walls */
  const int32_t* wall_cell;/* [n_wall] cell of each diffuse side */
  const double* wall_rinv; /* [n_wall] 1 / sum_ml min(sigma.db, 0) (0 if the sum is 0) */
  const double* wall_dbs;  /* [n_wall][dim] db of each diffuse side */
  const int32_t* ce1;      /* [n_contact] */
  const int32_t* ce2;      /* [n_contact] (periodic contact only) */
  const double* c_na;      /* [n_contact][dim] normal areas of the contact */
  const double* c_t;       /* [n_contact] transmission of each contact side (NULL: 1) */
  const double* t_face;    /* [n_faces] transmission of each directed face (NULL: 1);
                              inflow coefficients are t * -min(s . k, 0); jump_face is
                              expected pre-multiplied by t */
  const double* mfps;      /* [M] Lambda_m (nm) */
  const double* t_coeff;   /* [M][L] */
  const double* sigma;     /* [M][L][dim] */
  const double* heat;      /* [n] heat_source * heat_source_coeff */
  double nano;             /* 1 for the nanowire perturbation -s_z (3D), else 0 */

  /* assembled operator (written by obte_bte_assemble) */
  double* pert;            /* [L][n] perturbation in position order */
  double* rec;             /* [L][n][8] sweep record of each position (64 B):
                              [0] D_l, [1] b0 = DeltaT + heat, [2] q = pert + diffuse-wall source,
                              [3..5] -min(s . k, 0) of the inflow slots (0 if unused),
                              [6] int32 pair (src0, src1), [7] int32 pair (src2, keep): keep = 1
                              when a later position reads the value from t (outside the LDS
                              window of the smallest window a sweep may use), it is a cut
                              exit, or its cell is a diffuse-wall or kappa-contact cell;
                              windowed sweeps store only those (final sweeps all of them
                              unless part_chunk is set and keep_t is 0);
                              [1] and [2] are rewritten by obte_bte_source every outer iteration */

  /* state */
  const double* delta_t;   /* [n] lattice temperature of the previous iteration */
  const double* tb;        /* [n_wall] wall temperatures of the previous iteration */
  double* t;               /* [L][n_chunks][n + ncut_max][mc] solution workspace; rows >= n
                              hold the interface inflows of the running sweep */
  double* part_tj;         /* [L][ps] per-direction sums in sweep order, ps = n (nf + 1)
                              rounded up to even, nf = 1 with sig_a, else dim: lattice
                              temperature and flux sums of position p at p (nf + 1) ...
                              p (nf + 1) + nf */
  double* part_tb;         /* [L][n_chunks][n_wall] */
  double* kappa;           /* [M][L] */
  double* supp;            /* [M][L] */
  double* scratch;         /* >= 4096 doubles of reduction scratch */

  /* final-sweep mode sums (windowed sweeps) */
  double* part_chunk;      /* [L][n_chunks][ps] per-chunk sums sum_m t_ml T_ml and
                              sum_m sigma_ml T_ml (with sig_a: sum_m sig_a_ml T_ml) of each
                              position, written by windowed final
                              sweeps, which then store in t only the values read later (keep
                              flags); NULL: final sweeps store every value and
                              obte_bte_reduce forms the sums from t */
  int32_t keep_t;          /* 1: windowed final sweeps still store every value in t (for
                              callers that read t afterwards) */

  /* separable group velocities (optional).  When every direction's velocities
     are parallel, sigma_ml = sig_a[m][l] sig_dir[l], the flux of a position
     needs one mode sum, sum_m sig_a_ml T_ml, instead of dim (the direction
     multiplies it in the sum over directions): part_tj and part_chunk then
     hold 2 doubles per position instead of dim + 1.  NULL: sigma's dim
     components are summed. */
  const double* sig_a;     /* [M][L] or NULL */
  const double* sig_dir;   /* [L][dim] (with sig_a) */
} obte_bte_plan_t;

/* Input / output buffers of one sweep launch for one problem. */
typedef struct obte_sweep_io {
  const double* y;            /* [n_chunks][n_cut][mc] cut-face inflows */
  double* x;                  /* [n_chunks][n_cut][mc] exit values (may be NULL) */
  const int32_t* active;      /* [L][n_chunks] task flags (NULL = all) */
  const int32_t* lane_active; /* [L][n_chunks][mc] per-system flags (mode 1; NULL = all) */
} obte_sweep_io_t;

/* Outputs of the mode-sum reduction of one problem. */
typedef struct obte_reduce_io {
  double* dt_new;  /* [n] */
  double* j_new;   /* [n][dim] */
  double* tb_new;  /* [n_wall] */
  double* norm2;   /* [1] ||dt_new||^2 */
} obte_reduce_io_t;

/* Face-flux assembly of D_l, the inflow coefficients and the perturbation
 * (openbte/bte.py:47-124) into pert and the static part of rec. */
int obte_bte_assemble(const obte_bte_plan_t* plan, obte_stream_t stream);

/* Per-iteration sources in position order, into rec[1] and rec[2]:
 * q = pert - TB Gbm2 (diffuse walls) and b0 = DeltaT + heat
 * (openbte/bte.py:199-200,211-214). */
int obte_bte_source(const obte_bte_plan_t* plan, obte_stream_t stream);

/* Level-scheduled upwind sweeps of all (l, m) systems (openbte/bte.py:217-218).
 * mode 0: source + interface inflow y -> exit values x
 * mode 1: zero source, interface inflow y -> exit values x (transfer operator)
 * mode 2: as 0, plus the wall temperatures (openbte/bte.py:227) into
 *         part_tb and kappa / supp (openbte/bte.py:232-243); the solutions
 *         stay in plan->t for obte_bte_reduce, or, for a windowed sweep of a
 *         plan with part_chunk, the per-chunk mode sums go to part_chunk.
 * y, x: [n_chunks][n_cut][mc].  active: [L][n_chunks] (or NULL = all);
 * lane_active: [L][n_chunks][mc] per-system flags used by mode 1 to skip
 * converged MFP bins (or NULL). */
int obte_bte_sweep(const obte_bte_plan_t* plan, int32_t mode, const double* y, double* x,
                   const int32_t* active, const int32_t* lane_active, obte_stream_t stream);

/* Positions of the LDS window that obte_bte_sweep / obte_batch_sweep use for
 * a plan with these fields (a power of two >= lvl_pair_max, at most 48 KB of
 * LDS), or 0 when sweeps read every upwind value from t.  Host-only query
 * (no counterpart in the reference: a launch-shape detail). */
int32_t obte_sweep_window(int32_t sweep_window, int32_t lvl_pair_max, int32_t mc);

/* Interface Krylov layer: GCR with a recycled basis on the cut-face
 * equations y = R S(B, y).  The transfer operator does not depend on the
 * outer iteration, so the basis (u_k, c_k = (I - T_op) u_k, c orthonormal)
 * is kept across iterations. */
typedef struct obte_gcr {
  int32_t kmax;        /* basis capacity */
  double rtol;         /* convergence: ||r||_2 <= rtol * max(1, ||y||_inf) * sqrt(q) */
  double* y;           /* [n_chunks][n_cut][mc] interface values */
  double* r;           /* residual */
  double* x;           /* exit values of the last sweep */
  double* u;           /* [kmax][n_chunks][n_cut][mc] */
  double* c;           /* [kmax][n_chunks][n_cut][mc] */
  int32_t* k;          /* [L][n_chunks] basis size */
  int32_t* active;     /* [L][n_chunks] 1 while some lane is unconverged */
  int32_t* lane_active;/* [L][n_chunks][mc] per-system flag */
  double* w;           /* [n_chunks][n_cut][mc] inflow of the extension sweep */
  int32_t* n_active;   /* [1] number of active (l, chunk) groups */
  double* rnorm;       /* [L][M] residual norms (diagnostics) */
} obte_gcr_t;

/* r = x - y, project r on the recycled basis, update y; set active flags. */
int obte_gcr_project(const obte_bte_plan_t* plan, const obte_gcr_t* g, obte_stream_t stream);
/* Copy r of active groups into basis slot k (the next sweep's inflow). */
int obte_gcr_begin_extend(const obte_bte_plan_t* plan, const obte_gcr_t* g, obte_stream_t stream);
/* c = u - x, orthogonalise against the basis, normalise, update y and r. */
int obte_gcr_end_extend(const obte_bte_plan_t* plan, const obte_gcr_t* g, obte_stream_t stream);
/* The whole extension loop in one launch: every active (direction, chunk)
 * task repeats begin_extend, the transfer sweep (obte_bte_sweep mode 1 with
 * y = g->w, x = g->x, active = g->active, lane_active = g->lane_active) and
 * end_extend until it converges or has made max_extend extensions, without
 * waiting for the other tasks.  Each task takes the steps of the host loop
 * "while any task is active and fewer than max_extend rounds: begin, sweep,
 * end" (openbte/bte.py:217-220 solves the same systems directly), so the
 * results are those of that loop.  On return *g->n_active (device) counts the
 * tasks still active and *ext_max (device int) holds the largest number of
 * extensions a task made.  No counterpart in the reference (its systems are
 * factorised, not iterated). */
int obte_gcr_extend(const obte_bte_plan_t* plan, const obte_gcr_t* g, int32_t max_extend, int32_t* ext_max,
                    obte_stream_t stream);
/* Clear the basis (k = 0); y = the plan's current delta_t at each cut's exit
 * cell (use_delta_t != 0) or 0. */
int obte_gcr_reset(const obte_bte_plan_t* plan, const obte_gcr_t* g, int32_t use_delta_t, obte_stream_t stream);

/* After a mode-2 sweep: mode sums DeltaT = sum t T and J = sum sigma (T - DeltaT_old)
 * (openbte/bte.py:222-230) from plan->t, then the fixed-order reduction over
 * directions into DeltaT, J, TB (openbte/bte.py:245-249 -> openbte/objects.py:86-101);
 * norm2 = ||DeltaT||^2.  plan->delta_t must still be the sweep's DeltaT_old. */
int obte_bte_reduce(const obte_bte_plan_t* plan, double* delta_t_new, double* j_new, double* tb_new,
                    double* norm2, obte_stream_t stream);

/* ---- batch of problems (many geometries / requests in one launch) -------
 *
 * All arrays below are device arrays.  Every problem keeps its own plan and
 * Krylov state; the sweep runs as one persistent launch whose workgroups drain
 * 8 XCD-local task queues.  A problem with live[p] == 0 is skipped by every
 * kernel, so its state stays that of its last iteration. */
typedef struct obte_bte_batch {
  int32_t n_problems;
  int32_t n_tasks;       /* sum over problems of n_dirs * n_chunks */
  int32_t mc;            /* shared launch shape (equal to every plan's) */
  int32_t block;
  int32_t kin;           /* max inflow slots over the problems (a plan with fewer
                            slots has zero coefficients in the rest of rec) */
  int32_t groups;        /* persistent workgroups of a sweep launch */
  int32_t max_elems;     /* max n_dirs * n over problems */
  int32_t max_dirs;      /* max n_dirs */
  int32_t max_cells;     /* max n */
  int32_t max_walls;     /* max n_wall */
  int32_t lvl_pair_max;  /* max of the plans' lvl_pair_max (0: unknown) */
  int32_t sweep_window;  /* as in the plan */
  int32_t part_sums;     /* 1: every plan has part_chunk (windowed final sweeps write
                            per-chunk mode sums) */
  const obte_bte_plan_t* plans;  /* [n_problems] */
  const obte_gcr_t* gcrs;        /* [n_problems] */
  const obte_sweep_io_t* io;     /* [3][n_problems]: sweep buffers of modes 0, 1, 2 */
  const obte_reduce_io_t* rio;   /* [n_problems] */
  const int32_t* live;           /* [n_problems] */
  const int32_t* task_pid;       /* [n_tasks] problem of each task, grouped by queue */
  const int32_t* task_local;     /* [n_tasks] l * n_chunks + chunk */
  const int32_t* queue_off;      /* [9] task range of each XCD queue */
  int32_t* queue_cnt;            /* [8] queue counters (cleared by each sweep launch) */
  /* Tasks run by a single wave (problems whose adjacent levels are narrow);
   * the waves of a workgroup take them once the workgroup queues are drained.
   * NULL wave_off: none.  Their window holds min_window(wave_pair_max)
   * positions per wave (block / 64 windows within the 48 KiB budget). */
  const int32_t* wave_off;       /* [9] task range of each XCD wave queue, or NULL */
  int32_t* wave_cnt;             /* [8] wave queue counters (cleared by each sweep launch) */
  int32_t wave_pair_max;         /* max lvl_pair_max of the wave tasks' plans */
} obte_bte_batch_t;

int obte_batch_assemble(const obte_bte_batch_t* b, obte_stream_t stream);
int obte_batch_source(const obte_bte_batch_t* b, obte_stream_t stream);
int obte_batch_sweep(const obte_bte_batch_t* b, int32_t mode, obte_stream_t stream);
/* op 0 project, 1 begin_extend, 2 end_extend (as the obte_gcr_* calls) */
int obte_batch_gcr(const obte_bte_batch_t* b, int32_t op, obte_stream_t stream);
int obte_batch_gcr_reset(const obte_bte_batch_t* b, int32_t use_delta_t, obte_stream_t stream);
/* obte_gcr_extend for every live problem of the batch in one persistent
 * launch: workgroups drain the task queues (workgroup tasks, then wave tasks,
 * each run by a whole workgroup here) and loop begin_extend, transfer sweep
 * and end_extend per task until it converges or has made max_extend
 * extensions, without waiting for other tasks (openbte/bte.py:217-220 solves
 * the same systems directly).  Clears and then sets *g.n_active of every
 * problem (tasks still active) and *ext_max (device int: the largest number
 * of extensions a task made).  The results are those of the round-by-round
 * obte_batch_gcr / obte_batch_sweep loop. */
int obte_batch_extend(const obte_bte_batch_t* b, int32_t max_extend, int32_t* ext_max, obte_stream_t stream);
/* *out (device int) = number of active GCR tasks over all problems; used to
 * record the rounds of a graph-captured iteration. */
int obte_batch_log_active(const obte_bte_batch_t* b, int32_t* out, obte_stream_t stream);
/* reduce into rio, then DeltaT, TB <- rio.dt_new, rio.tb_new for live problems */
int obte_batch_reduce(const obte_bte_batch_t* b, obte_stream_t stream);

/* ---- Fourier (openbte/fourier.py:15-185) ------------------------------ */

/* One level of the multigrid hierarchy on the device.  The operator rows are
 * (A x)_i = a_diag_i x_i - sum_k a_w[k] x[a_col[k]] over a_ptr[i] .. a_ptr[i+1];
 * P (n x n_next) prolongs from the next level, R = P^T restricts to it.  b, x,
 * res and z are work vectors of n entries (b: right-hand side of the level,
 * x: pre-smoothed iterate, res: its residual, z: the level's result).
 * a_entries / r_entries: off-diagonal entries of the level's operator and
 * entries of R (they set how many lanes share a row). */
typedef struct obte_amg_level {
  int32_t n;
  int64_t a_entries;
  int64_t r_entries;
  const int32_t* a_ptr;
  const int32_t* a_col;
  const double* a_w;
  const double* a_diag;
  const int32_t* p_ptr;
  const int32_t* p_col;
  const double* p_val;
  const int32_t* r_ptr;
  const int32_t* r_col;
  const double* r_val;
  double* b;
  double* x;
  double* res;
  double* z;
} obte_amg_level_t;

typedef struct obte_fourier_plan {
  int32_t n;          /* cells */
  int32_t dim;
  int32_t nv;         /* sides per element */
  int32_t n_int;      /* internal sides */
  int32_t pin;        /* pinned cell of openbte/fourier.py:43,115-118, or -1 */
  /* symmetric operator A = V F (+ boundary / modified terms):
   * (A x)_i = diag_i x_i - sum_k w_k x_col[k] over row i (no diagonal in CSR) */
  const int32_t* row_ptr;
  const int32_t* col;
  const double* w;
  const double* diag;
  const double* vol;       /* [n] element volumes */
  /* LSQ gradient (openbte/objects.py:322-405) */
  const int32_t* slot_cell;  /* [n][nv] neighbour cell or -1 */
  const int32_t* slot_mode;  /* [n][nv] 0: zero, 1: T[nb] - T + val, 2: val - T */
  const double* slot_val;    /* [n][nv] periodic jump (mode 1) or wall value (mode 2) */
  const double* gweights;    /* [n][dim][nv] pinv of the distance rows */
  /* non-orthogonal correction (openbte/fourier.py:161-165) */
  const int32_t* side_e;     /* [n_int][2] */
  const double* side_w;      /* [n_int] interpolation weights */
  const double* side_vn;     /* [n_int][dim] v_nonorth */
  const int32_t* slot_side;  /* [n][nv] internal side of the slot or -1 */
  const double* slot_sgn;    /* [n][nv] +1/V (cell is side_e[.][0]) or -1/V */
  /* diffusive kappa through the contact (openbte/objects.py:139-172) */
  int32_t n_contact;
  int32_t contact_kind;      /* 1 periodic, 2 mixed */
  const int32_t* ce1;
  const int32_t* ce2;
  const double* c_coef;      /* v_orth (periodic) or conductance-limited v_orth (mixed) */
  double c_value;            /* periodic jump or wall temperature */
  double normalization;
  /* multigrid preconditioner of the multi-launch CG (optional): one V-cycle of
   * the smoothed-aggregation hierarchy of obte_amg_create per CG iteration.
   * levels[0] is the plan's operator (its A fields are not read); the last
   * level restricts to the dense coarsest level, solved by coarse_inv. */
  int32_t n_levels;          /* levels above the coarsest, or 0: Jacobi only */
  const obte_amg_level_t* levels;  /* host array of n_levels descriptors */
  int32_t n_coarse;          /* rows of the coarsest level */
  const double* coarse_inv;  /* [n_coarse][n_coarse] inverse of the coarsest operator (plus a constant when
                                it is singular) */
  double* coarse_b;          /* [n_coarse] work vectors of the coarsest level */
  double* coarse_x;
} obte_fourier_plan_t;


/* Solve F T = B + C (openbte/fourier.py:152-156) as the symmetric system
 * A T = V (B + C) by preconditioned CG (Jacobi, or one multigrid V-cycle per
 * iteration when n_levels > 0), T holding the initial guess;
 * with a pinned cell the pinned row is replaced by the compatibility
 * condition and T is centred by (max + min) / 2.  ws: 6n + 64 doubles; scal: 16
 * doubles.  iters_out / relres_out are host pointers (may be NULL). */
int obte_fourier_solve(const obte_fourier_plan_t* plan, const double* b, const double* c, double* t, double* ws,
                       double* scal, double rtol, int32_t maxit, int32_t check_every, int32_t* iters_out,
                       double* relres_out, obte_stream_t stream);

/* obte_fourier_solve without the ws / scal arguments, allocating nothing:
 * its workspace is a file-scope device array of this library (6 * 262144 +
 * 80 doubles, 12 MiB, held on every device the library runs on until the
 * process ends).  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  Plans with more than OBTE_FOURIER_NOALLOC_MAX_CELLS cells fail
 * with OBTE_ELIMIT.  All calls on one device share that device's array and
 * leave their contents in it: calls that overlap on one device (two streams
 * or two threads) race and must be serialised by the caller. */
#define OBTE_FOURIER_NOALLOC_MAX_CELLS 262144
int obte_fourier_solve_noalloc(const obte_fourier_plan_t* plan, const double* b, const double* c, double* t,
                               double rtol, int32_t maxit, int32_t check_every, int32_t* iters_out,
                               double* relres_out, obte_stream_t stream);

/* LSQ cell gradient of T: grad [n][dim]. */
int obte_fourier_gradient(const obte_fourier_plan_t* plan, const double* t, double* grad, obte_stream_t stream);

/* Correction C [n] from the gradient; side_tmp: n_int doubles of workspace;
 * norm2 (device, 1 double) receives ||C||^2; ws: >= n/256+1 doubles. */
int obte_fourier_correction(const obte_fourier_plan_t* plan, const double* grad, double* c, double* side_tmp,
                            double* ws, double* norm2, obte_stream_t stream);

/* Diffusive kappa of T through the contact into out (device, 1 double). */
int obte_fourier_kappa(const obte_fourier_plan_t* plan, const double* t, double* out, obte_stream_t stream);

/* ---- differentiation (interfacial-transmission model) -------------------
 *
 * Replaces the reverse-mode gradient of the reference's JAX inverse solver
 * (openbte/inverse/bte.py:35-76 compute_kappa_and_gradient, via the custom
 * vjp of openbte/inverse/matinverse.py:647-662).  For every directed face f
 * (row cell i_face[f], upwind cell j_face[f], reflecting side side_of_face[f]
 * or -1) writes grad[f] = d kappa_eff / d t_f without the contact term:
 *   -V_i sum_l min(s_l . k_f, 0) sum_m t_ml Lambda_m mu'[lbar(l)][m][i] (T[l][m][j] - jump0_f - tb_k)
 * t_fwd / t_adj: solution workspaces (layout of plan->t) of the converged
 * forward solve and of the adjoint solve; lbar: reversed direction of each l. */
int obte_tim_face_grad(const obte_bte_plan_t* plan, const double* t_fwd, const double* t_adj, const int32_t* lbar,
                       const int32_t* i_face, const int32_t* j_face, const int32_t* side_of_face, const double* tb,
                       const double* jump0, const double* vol, int32_t n_faces, double* grad,
                       obte_stream_t stream);
/* The same sum restricted to directions l0 <= l < l1, whose forward
 * solutions t_fwd_dirs hold only those directions ([l1 - l0][n_chunks][n +
 * ncut_max][mc]); accumulate != 0 adds into grad.  Lets a caller recompute the
 * forward solutions a block of directions at a time instead of storing all of
 * them (same reference counterpart as obte_tim_face_grad). */
int obte_tim_face_grad_dirs(const obte_bte_plan_t* plan, const double* t_fwd_dirs, int32_t l0, int32_t l1,
                            int32_t accumulate, const double* t_adj, const int32_t* lbar, const int32_t* i_face,
                            const int32_t* j_face, const int32_t* side_of_face, const double* tb,
                            const double* jump0, const double* vol, int32_t n_faces, double* grad,
                            obte_stream_t stream);

/* ---- material interpolation ------------------------------------------------
 *
 * Splits n_modes phonon modes onto the (MFP bin, direction) grid and adds
 * them into w_out [n_mfp][n_ang] and sigma_out [n_mfp][n_ang][dim]
 * (zero-initialised by the caller; n_ang = n_phi in 2D, n_theta * n_phi in
 * 3D, direction index theta * n_phi + phi): replaces the np.add.at scatter of
 * openbte/material.py:59-77 (RTA3D) and 164-173 (RTA2DSym) with the index
 * rule of openbte/utils.py:281-320.  Per mode: weight W = C / tau, sigma = C v
 * and the vectorial MFP v tau ([n_modes][3], z ignored in 2D); grids
 * increasing, phi periodic in [0, 2 pi). */
int obte_rta_scatter(int32_t n_modes, int32_t dim, const double* weight, const double* sigma, const double* mfp,
                     int32_t n_mfp, const double* mfp_grid, int32_t n_phi, const double* phi_grid, int32_t n_theta,
                     const double* theta_grid, double* w_out, double* sigma_out, obte_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* OPENBTE_AMD_H */
