/*
 * cppe_hip.h -- C ABI of the MI355X (gfx950) polarizable-embedding field / induced-dipole path.
 *
 * Plain pointers and sizes only. Host-side vectors use the reference's layouts: "AoS3" is an
 * FP64 array of length 3*n indexed 3*i+c, in the caller's site order. Every entry point returns
 * CPPE_HIP_OK (0) or an error code; cppe_hip_last_error() then holds a thread-local message.
 *
 * A cppe_hip_system owns one set of sites resident in HBM (positions, polarizabilities, damping
 * parameters, exclusions, the octree and its interaction lists, and the solver work vectors).
 * It replaces, for the reference (maxscheurer/cppe), the per-call pair
 *   build_shared_tree<m,3>(...)            src/core/fmm/tree.cc:134-237
 *   Tree::compute_field_fmm / _exact(F)    src/core/fmm/tree.cc:359-394
 * and the objects built around it (BMatrix src/core/bmatrix.hh:19-36, InducedMoments
 * src/core/electric_fields.hh:51-80). The reference rebuilds the tree on every B apply
 * (src/core/bmatrix.cc:28-29); here it is built once per system.
 */
#ifndef CPPE_HIP_H
#define CPPE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define CPPE_HIP_OK 0
#define CPPE_HIP_ERR_ARG 1        /* invalid argument (reference: std::invalid_argument) */
#define CPPE_HIP_ERR_DEVICE 2     /* HIP runtime error or no device */
#define CPPE_HIP_ERR_NOCONV 3     /* "Failed to converge induced dipole moments." */
#define CPPE_HIP_ERR_TREE 4       /* octree could not be built (e.g. coincident sites) */
#define CPPE_HIP_ERR_UNSUPPORTED 5

#define CPPE_HIP_SUM_DIRECT 0
#define CPPE_HIP_SUM_FMM 1

#define CPPE_HIP_PREC_FP64 0      /* plain form: FP64 throughout */
#define CPPE_HIP_PREC_MIXED 1     /* FP32 field applies inside FP64 iterative refinement */

#define CPPE_HIP_PRECOND_JACOBI 0 /* z = alpha r, the reference's preconditioner (default) */
#define CPPE_HIP_PRECOND_LOCAL 1  /* z = alpha r + weight alpha T_loc alpha r (close pairs) */

typedef struct cppe_hip_options {
  int summation;    /* PeOptions::summation_induced_fields: 0 "direct", 1 "fmm" (pe_options.hh:28-29) */
  int ncrit;        /* PeOptions::tree_ncrit (pe_options.hh:30) */
  int order;        /* PeOptions::tree_expansion_order (pe_options.hh:31) */
  double theta;     /* PeOptions::theta (pe_options.hh:32) */
  double damping;   /* Thole factor used by this system's fields; 0 = undamped
                       (bmatrix.cc:18-21, electric_fields.cc:116-120) */
  int precision;    /* CPPE_HIP_PREC_* used by cppe_hip_solve */
} cppe_hip_options;

typedef struct cppe_hip_system cppe_hip_system;

/* Thread-local message of the last failing call on this thread ("" if none). */
const char* cppe_hip_last_error(void);

/* Number of visible HIP devices (0 on a host without a GPU; never fails). */
int cppe_hip_device_count(int* count);

/*
 * Create a device-resident system of n sites.
 *   pos         AoS3 positions (bohr)
 *   alpha6      n*6 packed polarizabilities (xx,xy,xz,yy,yz,zz) or NULL (no response operator)
 *   alpha_iso   n isotropic polarizabilities used for Thole damping (0 = site not damped), or NULL
 *   excl_off    n+1 CSR offsets, excl_idx values: exclusions compared with the *position* of the
 *               other site in this system, as the tree-side test does (fmm/calculate.cc:38,
 *               fmm/tree.hh:23-26). Values outside [0,n) never match. May both be NULL.
 * Replaces build_shared_tree(potentials, S, ncrit, order, theta, damping) (tree.cc:134-237)
 * together with the BMatrix constructor (bmatrix.hh:19-36).
 */
int cppe_hip_system_create(cppe_hip_system** out, int n, const double* pos, const double* alpha6,
                           const double* alpha_iso, const int* excl_off, const int* excl_idx,
                           const cppe_hip_options* opts);
int cppe_hip_system_destroy(cppe_hip_system* sys);

/*
 * Field at every site from order-m sources S (n*ncomp(m) values per site, the reference's
 * packing: q | mu_xyz | half-diagonal quadrupole), overwriting F (AoS3). Convention of the
 * reference's P2P<m,3> (fmm/field_m{0,1,2}.cc): m=1 returns the *negative* dipole field.
 * Replaces Tree<m,3>::compute_field_exact / compute_field_fmm (tree.cc:359-394).
 */
int cppe_hip_field(cppe_hip_system* sys, int m, const double* S, double* F);

/*
 * Static multipole field F(q) - F(mu) + F(theta') at every site, one fused device pass.
 * q, mu (AoS3), theta6 (n*6, full quadrupole with its trace removed; the half-diagonal packing
 * of electric_fields.cc:144-149 is applied inside). Any of mu/theta6 may be NULL.
 * Replaces MultipoleFields::compute (electric_fields.cc:106-189).
 */
int cppe_hip_multipole_fields(cppe_hip_system* sys, const double* q, const double* mu,
                              const double* theta6, double* F);

/* out = B x = field_m1(x) + alpha^-1 x  (BMatrix::apply, bmatrix.cc:9-40) */
int cppe_hip_apply_b(cppe_hip_system* sys, const double* x, double* out);
/* out_k = B x_k for nvec vectors (nvec consecutive 3n AoS blocks), applied together as in
 * cppe_hip_solve_multi: one near-field launch per group of three. One-snapshot systems only.
 * No counterpart in the reference (several BMatrix::apply calls). */
int cppe_hip_apply_b_multi(cppe_hip_system* sys, int nvec, const double* x, double* out);
/* out = alpha x (BMatrix::apply_diagonal_inverse, bmatrix.cc:42-51) */
int cppe_hip_apply_alpha(cppe_hip_system* sys, const double* x, double* out);
/* out = alpha^-1 x (BMatrix::apply_diagonal, bmatrix.cc:53-62) */
int cppe_hip_apply_alpha_inv(cppe_hip_system* sys, const double* x, double* out);

/*
 * Dense B (3n x 3n row-major): diagonal alpha^-1, off-diagonal -T2 (Thole-damped when the
 * system's damping > 0) with pairs zeroed where excl of i contains index[j] (the potfile index,
 * as bmatrix.cc:73 compares). BMatrix::to_dense_matrix (bmatrix.cc:64-92).
 */
int cppe_hip_dense_b(cppe_hip_system* sys, const int* index, double* out);

/* Inverse of the dense B by Gauss-Jordan elimination with partial pivoting on the device
 * (BMatrix::direct_inverse, bmatrix.cc:94-97). Test-size systems. */
int cppe_hip_dense_b_inverse(cppe_hip_system* sys, const int* index, double* out);

/*
 * Jacobi-preconditioned CG on B mu = rhs (InducedMoments::compute, electric_fields.cc:191-239).
 * make_guess != 0: x0 = alpha rhs; otherwise x0 = guess. Stops when ||r||_2 < thresh (absolute,
 * unpreconditioned). norms (length >= maxiter, may be NULL) receives the residual norm of each
 * iteration; *iters the number of iterations run. Returns CPPE_HIP_ERR_NOCONV after maxiter.
 * With opts.precision == CPPE_HIP_PREC_MIXED the field applies run in FP32 inside FP64
 * iterative refinement; the stopping test is the same FP64 residual norm.
 */
int cppe_hip_solve(cppe_hip_system* sys, const double* rhs, const double* guess, int make_guess,
                   double thresh, int maxiter, double* mu, int* iters, double* norms);

/*
 * Several right-hand sides solved together: nrhs independent CG recurrences (each with its own step
 * lengths and stopping test, exactly the single-solve recurrence) that share every B apply, so
 * the near-field pair geometry is evaluated once for all right-hand sides.
 * CppeState::induced_moments_eef (cppe_state.cc:111-128) solves its three unit fields this way
 * instead of with three InducedMoments::compute calls. rhs, guess and mu are nrhs consecutive
 * 3n AoS blocks (caller order); iters receives nrhs counts, norms (optional) nrhs x maxiter
 * residual norms. One-snapshot systems only.
 */
int cppe_hip_solve_multi(cppe_hip_system* sys, int nrhs, const double* rhs, const double* guess, int make_guess,
                         double thresh, int maxiter, double* mu, int* iters, double* norms);

/*
 * Device-resident variants (SCF-loop residency): rhs/guess are uploaded once and mu stays in
 * HBM between calls; cppe_hip_solve_resident times only device work (inputs already resident).
 * guess_mode: 0 = alpha rhs, 1 = previous mu held by the system, 2 = zero.
 * *device_ms (may be NULL) receives the solve's device time measured with HIP events.
 */
int cppe_hip_upload_rhs(cppe_hip_system* sys, const double* rhs);
int cppe_hip_solve_resident(cppe_hip_system* sys, int guess_mode, double thresh, int maxiter,
                            int* iters, double* norms, double* device_ms);
int cppe_hip_download_mu(cppe_hip_system* sys, double* mu);

/*
 * Snapshot batches (no counterpart in the reference, whose InducedMoments::compute,
 * electric_fields.cc:191-239, solves one environment per call). nseg independent systems, site
 * ranges [seg_off[g], seg_off[g+1]) of the concatenated caller arrays, are held by one handle:
 * one tree per snapshot (own root; seg_ncrit[g] = ncrit, or 0 for a single leaf = exact direct
 * sum), every field apply shared by all snapshots, and a CG recurrence per snapshot (own step
 * lengths, own FP64 stopping test ||F_g - B_g mu_g|| < thresh, frozen once converged).
 * Exclusion indices are global (caller positions in the concatenation) and must stay inside
 * their snapshot. The summation is always the tree path (opts.summation is ignored).
 * upload_rhs / solve_resident / download_mu / multipole_fields / apply_b act on all snapshots at
 * once; solve_resident's *iters is the shared loop's count (max over snapshots).
 * Not available on a batch: dense B, fields of order 0/2 through cppe_hip_field, descriptors.
 */
int cppe_hip_batch_create(cppe_hip_system** out, int nseg, const int* seg_off, const int* seg_ncrit,
                          const double* pos, const double* alpha6, const double* alpha_iso, const int* excl_off,
                          const int* excl_idx, const cppe_hip_options* opts);
/*
 * Dipole field (the P2P<1,3> convention of cppe_hip_field, m = 1: Thole-damped with the system's
 * factor, exclusions as in the system) at nt target sites (caller indices) by FP64 direct
 * summation over all sites of the target's snapshot, for mu (AoS3, caller order) or, with
 * mu == NULL, the system's resident moments. A measuring aid: it samples the tree path's error
 * against direct summation (Tree::compute_field_exact, tree.cc:387-394, restricted to targets).
 */
int cppe_hip_field_targets(cppe_hip_system* sys, const double* mu, int nt, const int* targets, double* F);

/* Per-snapshot outcome of the last solve (arrays of nseg; any may be NULL). */
int cppe_hip_segment_results(cppe_hip_system* sys, int* iters, double* rnorm, int* converged);
/* CG iterations replayed as captured HIP graphs (default 1) or launched kernel by kernel (0). */
int cppe_hip_set_graphs(cppe_hip_system* sys, int enable);
/*
 * Most CG steps launched as one sequence without the host in between (1 to 8; default 8).
 * After every decision the device publishes how many steps each snapshot can take before one of
 * them needs the host (an FP64 true residual, or none left active), predicted from their residual
 * reduction rates; the host launches that many steps (at most the limit) and reads the
 * status once. A step that starts after such an event changes nothing, so results do not depend
 * on this setting. No counterpart in the reference (a host loop, electric_fields.cc:213-234).
 */
int cppe_hip_set_graph_steps(cppe_hip_system* sys, int steps);
/* Last solve: launches[0] step sequences launched, [1] steps in them (those that started after an
 * event took none), [2] status reads of the host loop. */
int cppe_hip_last_solve_launches(cppe_hip_system* sys, int* launches);
/* Mixed form: factor by which a residual must fall before the FP64 true residual replaces the
 * recursive one (default 1e-5; convergence is declared on an FP64 true residual or on its
 * FP32-corrected form, cppe_hip_set_confirmation). */
int cppe_hip_set_reliable_drop(cppe_hip_system* sys, double drop);
/*
 * Mixed form: how a recursive residual below the threshold is confirmed (no counterpart in the
 * reference, whose FP64 CG stops on its recursive residual, electric_fields.cc:213-229). eta
 * (default 1e-5) bounds an FP32 apply's relative error (measured at or below 1e-6); forms, a bit
 * set (default 3):
 *  CPPE_HIP_CONFIRM_GAP (1): after an FP64 true residual, the gap between true and recursive
 *    residual is at most eta sum_k |alpha_k| ||B32 p_k|| over the steps since; a recursive
 *    residual r with ||r|| + that bound below the threshold has converged without further work;
 *  CPPE_HIP_CONFIRM_CORRECTED (2): otherwise, from the last FP64 true residual r_t (at moments
 *    x_t) and one FP32 apply of the correction since, r_c = r_t - B32 (x - x_t), accepted when
 *    ||r_c|| + eta ||B32 (x - x_t)|| < thresh (3 more vectors of 3 n doubles per right-hand side).
 * Both keep the FP64 residual ||rhs - B x|| below the threshold. What neither accepts, and
 * everything with eta = 0 or forms = 0, is decided by an FP64 true residual.
 */
#define CPPE_HIP_CONFIRM_GAP 1
#define CPPE_HIP_CONFIRM_CORRECTED 2
int cppe_hip_set_confirmation(cppe_hip_system* sys, double eta, int forms);
/* Last solve: FP32-corrected confirmations run (one FP32 apply each; FP64 true residuals are
 * counted by cppe_hip_last_solve_info). */
int cppe_hip_last_solve_confirmations(cppe_hip_system* sys, int* count);
/*
 * Mixed form, all knobs of the refinement: `drop` as above; `stall`: a recursive residual that
 * has not halved for this many steps also asks for an FP64 true residual (default 10);
 * `min_progress`: an FP64 true residual above this fraction of the previous one (after at least
 * one step) means the FP32 refinement stagnates, and the solve continues from the current moments
 * in the plain FP64 form (default 0.5). "Failed to converge" is returned only when that, too,
 * does not converge within maxiter. No counterpart in the reference (whose solve is FP64 only,
 * electric_fields.cc:191-239).
 */
int cppe_hip_set_refinement(cppe_hip_system* sys, double drop, int stall, double min_progress);
/*
 * CG preconditioner of this system's solves (no counterpart in the reference, whose
 * InducedMoments::compute preconditions with alpha r, electric_fields.cc:191-239).
 * kind CPPE_HIP_PRECOND_JACOBI: z = alpha r (default). kind CPPE_HIP_PRECOND_LOCAL: the first term
 * of the Neumann series of B^-1 = (alpha^-1 - T)^-1 beyond alpha, kept over close pairs:
 * z = alpha r + weight alpha T_loc alpha r, T_loc_ij the system's (Thole-damped) dipole tensor
 * -B_ij for pairs closer than `radius` (bohr) in partner leaves of the octree's near lists,
 * exclusions and the site itself left out; T_loc is symmetric, so CG keeps its recurrence and its
 * unpreconditioned stopping test. radius > 0 (5.669 = 3 A: about six neighbours of a water
 * site), weight >= 0 (0.8 measured best on water boxes).
 */
int cppe_hip_set_preconditioner(cppe_hip_system* sys, int kind, double radius, double weight);
/* Last solve: info[0] iterations of the shared loop, info[1] FP64 true-residual applies,
 * info[2] iteration at which a stagnating mixed solve continued in FP64 (-1: it did not). */
int cppe_hip_last_solve_info(cppe_hip_system* sys, int* info);

/*
 * Autotuner of the tree path's B-apply launch shapes (no counterpart in the reference). For the
 * FP64 and the FP32 apply in turn, times on this system's own tree (HIP events, mean of reps
 * after a warm-up) every near-field candidate, then every M2L group width (16/32/64 lanes per
 * target cell), and keeps the fastest of each. Near-field candidates (variant, block):
 *   0: one thread per target site, block 64/128/256;
 *   1: one wavefront per 64-target leaf chunk, uniform (scalar) source walk, block 64/128/256;
 *   2: one wavefront per leaf, sources staged through LDS, block 64 (only when every leaf has at
 *      most 1024 sites: its time grows with the square of the leaf size; otherwise not timed, 0 ms);
 *   3: one wavefront per leaf chunk, lanes = (target, source slot) over the leaf's expanded near
 *      list (built on first use: one int per near source site of each leaf), block 64/128/256;
 *   4: one wavefront per target cluster (whole consecutive leaves, <= 64 sites) walking the union
 *      of their near lists with per-source-leaf lane masks, block 64/128/256;
 *   5, 6, 7, 8: one wavefront per leaf chunk of at most 64, 32, 16, 8 targets, several per block, the
 *      expanded near list staged 64 sources at a time through wavefront-private LDS, lanes =
 *      (target, source slot), block 64/128/256;
 *   9, 10, 11: symmetric: each unordered near pair evaluated once for the field at both sites
 *      (the tensor is even in the displacement); one wavefront per chunk of at most 16 / 12 / 8
 *      (FP32) or 8 / 6 / 4 (FP64) targets of a leaf walking its own sites and those of its near
 *      leaves with a higher index, lanes = sources; the field at those sources goes to per-leaf
 *      reverse blocks that a gather pass adds in a fixed order (with the L2P term); block
 *      64/128/256. Single-right-hand-side B applies only: the system's other near-field launches
 *      then use variant 8.
 * Then, for a staged or symmetric near variant (5-11), the whole apply is timed with the far field (P2M, M2M,
 * M2L, L2L) in series, on a second stream beside the near field (joined before an L2P pass; enqueued
 * after the near field or ahead of it), and with only M2L / L2L beside the near field after the
 * upward pass (P2M, M2M) has run, each overlapped form with the near field in 1, 2 and 4 launches;
 * the fastest is kept.
 * choice[8]: (near variant, near block, m2l group, overlap) for FP64 then FP32; overlap = form +
 *            4 log2(near launches): form 0 = series, 1 = whole far field beside the near field,
 *            2 = upward pass first, 3 = as 1 with the far field enqueued ahead of the near field;
 *            with forms 1-3 the near field may be issued as 1, 2 or 4 launches of consecutive work
 *            items (so 0-3, 5-7, 9-11 are valid).
 * ms[78]: per precision, the 34 near timings (order above), the 3 M2L timings (ms/launch), then
 *         the apply in series and overlapped (the faster overlapped form; 0 when not timed).
 * Launch shapes never change results beyond FP summation order.
 */
int cppe_hip_autotune(cppe_hip_system* sys, int reps, int* choice, double* ms);
/* Applies a launch-shape choice (e.g. a cached autotune result). */
int cppe_hip_set_tune(cppe_hip_system* sys, const int* choice);

/*
 * Times `reps` B applies of a fixed pseudo-random vector on the system's stream, with HIP
 * events around each phase. phase_ms (6 values, mean per apply): [0] upward pass (P2M+M2M),
 * [1] M2L, [2] L2L, [3] near field (near-field kernel with fused L2P, plus the exclusion pass),
 * [4] whole apply, [5] the staged near-field kernel alone (-1 when another variant is tuned).
 * No counterpart in the reference (measurement support).
 */
int cppe_hip_profile_apply(cppe_hip_system* sys, int fp32, int reps, double* phase_ms);

/* NuclearFields::compute (electric_fields.cc:58-75): F_i = sum_a Z_a (r_i - R_a)/|r_i - R_a|^3 */
int cppe_hip_nuclear_fields(int n, const double* pos, int natoms, const double* charges,
                            const double* atom_pos, double* F);

/*
 * MultipoleExpansion::interaction_energy (multipole_expansion.cc:10-33): kmax[i] highest order
 * present at site i (-1: none), mul = 10 values per site (q, mu, full traceless quadrupole).
 */
int cppe_hip_nuclear_multipole_energy(int n, const double* pos, const int* kmax, const double* mul,
                                      int natoms, const double* charges, const double* atom_pos,
                                      double* energy);

/*
 * Allocation-free variant of cppe_hip_nuclear_multipole_energy. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph capture, loops that
 * re-evaluate the energy as the nuclei move). Every array argument and `energy` are DEVICE
 * pointers; the kernels are enqueued on `stream` (a hipStream_t, NULL = the null stream) and
 * *energy is written there, without synchronisation. The partial sums live in a fixed device
 * array of 16384 doubles (128 KiB per GPU, kept until the process ends): at most 4,194,304 sites,
 * larger n returns CPPE_HIP_ERR_ARG. Calls that overlap on one GPU race on that array.
 * Same kernels and summation order as cppe_hip_nuclear_multipole_energy (same result bits).
 */
int cppe_hip_nuclear_multipole_energy_noalloc(int n, const double* pos, const int* kmax, const double* mul,
                                              int natoms, const double* charges, const double* atom_pos,
                                              double* energy, void* stream);

/*
 * NuclearFields::nuclear_gradient (electric_fields.cc:77-104): derivative of the nuclear fields at
 * the n polarizable sites with respect to the atom positions, row-major (3 natoms) x (3 n):
 * grad[(3a + p) * 3n + 3i + q] = Z_a T2(r_i - R_a)_pq.
 */
int cppe_hip_nuclear_field_gradient(int n, const double* pos, int natoms, const double* charges,
                                    const double* atom_pos, double* grad);

/*
 * MultipoleExpansion::nuclear_gradient (multipole_expansion.cc:35-62): derivative of the
 * nuclear-multipole energy with respect to each atom position, natoms x 3 row-major. Same site
 * arguments as cppe_hip_nuclear_multipole_energy.
 */
int cppe_hip_nuclear_multipole_gradient(int n, const double* pos, const int* kmax, const double* mul,
                                        int natoms, const double* charges, const double* atom_pos,
                                        double* grad);

/* out = sum_i a_i b_i in FP64 on the device (polarization energies, cppe_state.cc:97-108) */
int cppe_hip_dot(long long n, const double* a, const double* b, double* out);

/*
 * Octree introspection for parity tests.
 * stats: [ncells, nleaves, n_m2l, n_p2p_cell_pairs, max_level, p2p_site_pairs]
 * geo (ncells*5: x,y,z,r,rmax), meta (ncells*4: parent, level, child mask, count),
 * leaf_start (ncells: first site in device order), perm (n: device order -> caller index),
 * m2l (n_m2l*2: source, target), p2p (n_p2p*2: target, source). Cells are in level order.
 */
int cppe_hip_tree_stats(cppe_hip_system* sys, long long* stats);
int cppe_hip_tree_export(cppe_hip_system* sys, double* geo, long long* meta, long long* cell_start,
                         int* perm, long long* m2l, long long* p2p);

/* ---------------------------------------------------------------------------------------------
 * Learned CG warm start (no counterpart in the reference). A per-site MLP 64 -> 128 -> 128 -> 3
 * (SiLU, BF16 MFMA with FP32 accumulation) maps 64 invariant descriptors of a site's
 * neighbourhood (6 A) and permanent field to coefficients c of three equivariant vectors
 * (v1 = alpha F, v2 and v3 = local damped Neumann terms); the guess is sum_a c_a v_a.
 * CG from that guess is unchanged and still stops on the FP64 residual; a zero or alpha F guess
 * remains available (guess_mode 0 / 2 of cppe_hip_solve_resident).
 * ------------------------------------------------------------------------------------------- */
#define CPPE_HIP_MLP_IN 64
#define CPPE_HIP_MLP_PARAMS 25219 /* W1[64][128] b1[128] W2[128][128] b2[128] W3[128][3] b3[3] */

typedef struct cppe_hip_mlp cppe_hip_mlp;

int cppe_hip_mlp_create(cppe_hip_mlp** out);
int cppe_hip_mlp_destroy(cppe_hip_mlp* m);
/* FP32 parameters (CPPE_HIP_MLP_PARAMS) and the feature scaling (64 means, 64 inverse std). */
int cppe_hip_mlp_set(cppe_hip_mlp* m, const float* params, const float* mean, const float* inv_std);
int cppe_hip_mlp_get(cppe_hip_mlp* m, float* params, float* mean, float* inv_std);
/* Descriptors X (n*64 FP32) and bases V (n*9: v1, v2, v3 as xyz triples) of the system's
 * resident rhs (cppe_hip_upload_rhs), caller order. */
int cppe_hip_ml_features(cppe_hip_system* sys, float* X, double* V);
/* y (ns*3) = MLP of raw features X (ns*64), scaled with the model's feature scaling. */
int cppe_hip_mlp_forward(cppe_hip_mlp* m, const float* X, long long ns, float* y);
/* Mean loss sum_s w_s |sum_a y_sa V_sa - T_s|^2 / ns and its gradient (CPPE_HIP_MLP_PARAMS). */
int cppe_hip_mlp_loss_grad(cppe_hip_mlp* m, const float* X, const float* V, const float* T, const float* w,
                           long long ns, float* grad, double* loss);
/* Training set in HBM: reserve room for ns samples; append every site of a system (features of its
 * resident rhs, bases, and its resident moments as targets); fix the feature scaling. */
int cppe_hip_mlp_reserve(cppe_hip_mlp* m, long long ns);
int cppe_hip_mlp_collect(cppe_hip_mlp* m, cppe_hip_system* sys);
int cppe_hip_mlp_prepare(cppe_hip_mlp* m);
/* Training objective, set before collecting: 0 (default) the squared dipole error relative to alpha F,
 * 1 the error in the alpha^-1 norm (e^T alpha^-1 e per site: the diagonal blocks of the B-norm that
 * CG minimises) relative to alpha F in that norm -- the collected bases and targets are mapped by
 * L^T, L L^T = alpha^-1 (Cholesky), so the same loss kernel trains either objective. */
int cppe_hip_mlp_set_objective(cppe_hip_mlp* m, int objective);
/* Collected samples s0 .. s0 + count - 1 as the loss sees them: V (9 floats), T (3), weight (1). */
int cppe_hip_mlp_samples(cppe_hip_mlp* m, long long s0, long long count, float* V, float* T, float* w);
/* Adam epochs over the collected set (minibatches of `batch` samples, tiles shuffled with `seed`);
 * losses receives the mean training loss of each epoch. */
int cppe_hip_mlp_train(cppe_hip_mlp* m, int epochs, int batch, float lr, unsigned seed, double* losses);
/* Learned guess into the system's resident moments (use guess_mode 1 of cppe_hip_solve_resident);
 * *device_ms receives the time of features + network + guess, waiting for them. device_ms NULL:
 * the guess is only enqueued on the system's stream, ahead of the solve that follows. */
int cppe_hip_ml_guess(cppe_hip_system* sys, cppe_hip_mlp* m, double* device_ms);

#ifdef __cplusplus
}
#endif
#endif /* CPPE_HIP_H */
