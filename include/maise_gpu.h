/*
 * maise_gpu.h -- C ABI of the MI355X-native MAISE neural-network potential engine (libmaise_gpu.so).
 *
 * Plain C types only: host pointers and sizes, opaque handles.  Every entry point returns 0 on success and
 * a negative code on failure; mg_last_error() gives the message of the last failure on the calling thread.
 *
 * Reference interfaces replaced (MAISE v3.0.00, paths relative to its source tree):
 *   CALL_MAISE ........................ declared in mlib.h (inc/cmod.h:30, src/cmod.c:1164-1224)
 *   mg_model_load ..................... READ_MODEL + Build_ANN + READ_ANN + Build_PRS
 *                                       (src/util.c:1233-1335, src/nutl.c:305-357,502-633, src/nprs.c:75-257)
 *   mg_set_structure .................. CAST_CELL (src/cmod.c:1084-1135)
 *   mg_eval ........................... CELL_FRC / CELL_ENE (src/cmod.c:128-191): JAR + LIST + PRS_BP +
 *                                       FRC_ANN/ENE_ANN + CELL_UREP (src/cell.c:82-185,490-508,
 *                                       src/nprs.c:445-773, src/nmlp.c:367-488, src/cmod.c:61-126)
 *   mg_get_neighbors .................. the LIST arrays Nn/Ni/S (src/cell.c:96-156), as a CSR export
 *   mg_get_descriptors ................ L->Cn after PRS_BP (src/nprs.c:504-566)
 */
#ifndef MAISE_GPU_H
#define MAISE_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mg_model mg_model;
typedef struct mg_ctx mg_ctx;

/* error codes */
#define MG_OK 0
#define MG_ERR_ARG -1
#define MG_ERR_MODEL -2
#define MG_ERR_HIP -3
#define MG_ERR_NEIGHBOR -4   /* neighbour capacity exceeded / inconsistent list */
#define MG_ERR_UNSUPPORTED -5
#define MG_ERR_IO -6          /* a file could not be opened, read or written */
#define MG_ERR_ALLOC -7       /* host memory */

/* mg_eval flags */
#define MG_EVAL_FORCES 1     /* compute forces and the virial as well as energies */
#define MG_EVAL_KEEP_G 2     /* keep descriptors G[N][D] for mg_get_descriptors */

/* MLP precision of mg_ctx_set_precision */
#define MG_PREC_FP32 0       /* descriptors, MLP and forces in FP32; energy/virial sums in FP64 */
#define MG_PREC_BF16 1       /* hidden-layer GEMMs (W1 h1, W1^T delta2) in one-pass BF16 on MFMA, FP32 accumulation; FP32 input layer and descriptors */
#define MG_PREC_BF16X3 2     /* MLP GEMMs in compensated BF16: split operands hi + lo, three MFMA passes, FP32 accumulation */

const char *mg_last_error(void);

/* --- model --- */
/* Load a MAISE nn_* model file (architecture D-H1-H2-1, G2/G4 basis appended). */
int mg_model_load(const char *path, mg_model **out);
void mg_model_free(mg_model *m);
/* info[0..15] = nspc, D, H1, H2, nsym, ng2, ng4, nw, spcz[0..3], G4 kernel variant (0 generic, 1 the 43-function
 * table, 2 the 22-function table), shared-eta flag, MFMA MLP flag, 0; *rc = neighbour-list cutoff (A). */
int mg_model_info(const mg_model *m, int *info, double *rc);
/* The model file's unique ID ("|  model unique ID |", read as READ_ANN does, src/nutl.c:530), as OUTCAR headers
 * print it; empty when the file has none. */
int mg_model_id(const mg_model *m, char *buf, int len);

/* --- evaluation context (one device stream, one structure or a batch of structures) --- */
int mg_ctx_create(const mg_model *m, int device, mg_ctx **out);
void mg_ctx_free(mg_ctx *c);
/* UREP short-range repulsion prefactor in eV (src/cmod.c:61-126); default 0 for the native API. */
int mg_ctx_set_urep(mg_ctx *c, double urep);
int mg_ctx_set_precision(mg_ctx *c, int prec);

/* Upload one structure (CAST_CELL semantics): N atoms, nd = 0 (cluster) or 3 (periodic), xt = 0 fractional
 * or 1 Cartesian (A), spc[N] species indices into the model's species list, lat[9] lattice rows (A),
 * x[3N] positions. */
int mg_set_structure(mg_ctx *c, int N, int nd, int xt, const int *spc, const double *lat, const double *x);

/* Upload a batch of S structures in CSR form: atom_off[S+1] offsets into spc/x, nd[S], lat[9S]; x Cartesian. */
int mg_set_batch(mg_ctx *c, int S, const int *atom_off, const int *nd, const double *lat, const int *spc,
                 const double *x);

/* Evaluate on the device.  All inputs already resident; results stay on the device until mg_get*.  Once the
 * capacities are known (after the first evaluation of a structure or batch) the call only enqueues the work on the
 * context's stream and returns; the next call that reads results (mg_get*, mg_features, dynamics) waits for it,
 * checks it, and evaluates again with larger buffers if the neighbour list outgrew them, so back-to-back
 * evaluations run without host round trips.  Errors of an enqueued evaluation are returned by that next call. */
int mg_eval(mg_ctx *c, int flags);
/* Copy results of structure s (0 for a single structure): E (eV), EA[natoms], F[3*natoms] (eV/A),
 * U[6] (eV, xx yy zz xy yz zx, U = -dE/dstrain), X[3*natoms] (wrapped positions).  NULL skips a field. */
int mg_get(mg_ctx *c, int s, double *E, double *EA, double *F, double *U, double *X);
/* Batch results: E[S], U[6S], F[3*Ntot] (NULL skips). */
int mg_get_batch(mg_ctx *c, double *E, double *U, double *F);
/* Batch positions: X[3*Ntot], the wrapped positions of every structure in batch order, in one copy (mg_get per
 * structure takes a copy per field and structure).  No counterpart in the reference, whose callers hold one
 * structure: it is for batched relaxation drivers (the evolutionary search reads its relaxed offspring with it). */
int mg_get_batch_x(mg_ctx *c, double *X);

/* One FP32 energy + forces + virial evaluation that allocates nothing: no counterpart in the reference; it is for
 * callers that cannot allocate device memory (HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same inputs and outputs as mg_set_structure + mg_eval(MG_EVAL_FORCES) + mg_get (host arrays:
 * spc[N], lat[9], x[3N] with xt = 0 fractional / 1 Cartesian, nd 0 or 3; E, EA[N], F[3N], U[6], NULL skips), on
 * the current device and its null stream, synchronously.  Every intermediate buffer is a slice of one fixed
 * 128 MiB device array per device, which the library holds until the process ends; the model tables are copied in
 * by each call.  Limits: N <= 32768 atoms, neighbour rows <= 256, and everything must fit the array (about 20,000
 * atoms at 80 neighbours and a 51-wide descriptor); beyond them the call fails with MG_ERR_ARG or MG_ERR_NEIGHBOR
 * and a message.  Calls share the array and a host work area: calls that overlap (threads, or devices driven from
 * several threads) race. */
int mg_eval_static(const mg_model *m, int N, int nd, int xt, const int *spc, const double *lat, const double *x,
                   double urep, double *E, double *EA, double *F, double *U);

/* Neighbour list of the last evaluation: row_off[N+1]; j[M], img[3M] (image indices), M = row_off[N].
 * Call with j == NULL to get M via *nedges only. */
int mg_get_neighbors(mg_ctx *c, int *nedges, int *row_off, int *j, int *img);
/* Descriptors G[N][D] of the last mg_eval with MG_EVAL_KEEP_G. */
int mg_get_descriptors(mg_ctx *c, float *G);
/* Stage times in milliseconds (device events on the context stream), averaged over the evaluations since timing
 * was enabled or last read (reading resets the average): t[0] neighbour build, t[1] descriptor kernel, t[2] MLP,
 * t[3] force/virial kernel, t[4] force gather + FP64 reductions, t[5] total.  t must hold 6 floats. */
int mg_get_timing(mg_ctx *c, float *t);
/* Issue rate of one vector ALU operation on the device, in operations (lanes) per second: op 0 FP32 FMA (non-packed),
 * 1 exp2, 2 cos, 3 sqrt (one add issued per sqrt), 4 rcp.  A microbenchmark with no counterpart in the reference:
 * it measures the rates that bound the descriptor and force kernels, which price their roofline. */
int mg_calibrate_rate(int device, int op, double *ops_per_s);
/* Enable (1) / disable (0) per-stage event timing (small overhead; evaluations stay asynchronous). */
int mg_ctx_set_timing(mg_ctx *c, int on);
/* Launch shape of the wave-per-atom descriptor and force kernels: atoms (64-lane waves) per workgroup, 1, 2 or 4;
 * 0 (the default) picks the largest whose LDS tile fits 64 KiB.  No counterpart in the reference (CPU only). */
int mg_ctx_set_wpb(mg_ctx *c, int wpb_desc, int wpb_force);
/* Relaxation and MD replay one captured iteration (evaluation chain + integrator step) as a HIP graph, with the
 * neighbour-edge buffers sized from the last synchronised evaluation plus `headroom` (fraction; default 0.125
 * and 8 extra row slots, 0 = exact).  A step that outgrows them is detected on the device, rolled back to the
 * last checkpoint (every 8 iterations) and repeated with larger buffers, so results do not depend on this
 * setting.  enable = 0 runs every iteration synchronised.  No counterpart in the reference (CPU only). */
int mg_ctx_set_graph(mg_ctx *c, int enable, double headroom);
/* The descriptor kernel stores the neighbour pairs (j, k) with r_jk < Rc that it enumerates for the G4 terms, and
 * the force kernel reads them back instead of enumerating again (default on; the list takes up to
 * rows x (rows - 1) / 2 x 4 bytes per atom and is skipped when that exceeds 1/16 of the device memory).  Results do
 * not depend on this setting.  No counterpart in the reference (PRS_BP enumerates the triplets once per call,
 * src/nprs.c:515-551, and derives forces in the same pass). */
int mg_ctx_set_pair_store(mg_ctx *c, int enable);
/* Autotuner (no counterpart in the reference, which runs one CPU code path): on the structure/batch currently
 * set, reps timed evaluations per candidate, it picks 1, 2 or 4 waves per workgroup for the neighbour fill, the
 * descriptor and the force kernel (each its fastest), then stored pair lists or enumeration by the descriptor +
 * force time.  No choice changes a result bit.  wpb_out[2] (may be NULL) receives (descriptor, force). */
int mg_ctx_autotune(mg_ctx *c, int reps, int *wpb_out);
/* The same; shape[4] receives (neighbour fill wpb, descriptor wpb, force wpb, pair store). */
int mg_ctx_autotune_shape(mg_ctx *c, int reps, int *shape);
/* Current launch shape: shape[4] = (neighbour fill wpb, descriptor wpb, force wpb, pair store); 0 = automatic. */
int mg_ctx_get_shape(mg_ctx *c, int *shape);
/* Waves per workgroup of the neighbour fill: 1, 2, 4, or 0 (4, fewer when the LDS rows of long lists need it). */
int mg_ctx_set_wpb_nbr(mg_ctx *c, int wpb);
/* Skin neighbour lists for fixed-cell runs (mg_relax, mg_md_run without a barostat) in graph mode: candidates
 * within rc + skin (A, 0 = off, the default; at most 4), rebuilt when an atom has moved more than skin/2 since their
 * build; every step applies LIST's own FP64 test (src/cell.c:109-123) to the candidates, so the neighbour rows,
 * and with them every result, are bit-identical to building them from scratch each step.  No counterpart in the
 * reference, which rebuilds its lists every call. */
int mg_ctx_set_skin(mg_ctx *c, double skin);
/* Skin-candidate builds over the context's life (fixed-cell runs with skin lists). */
int mg_ctx_skin_builds(mg_ctx *c, int *builds);
/* on != 0: mg_relax, mg_relax_cell and mg_md_run run the autotuner (3 reps) before their first captured iteration,
 * once per structure/batch set (evaluations at unchanged positions, so the trajectory is the same). */
int mg_ctx_set_autotune(mg_ctx *c, int on);

/* --- screening surrogate (no counterpart in the reference; its training derivative follows DE_ANN_PARA,
 * src/nmlp.c:292-354, and its training loop replaces the GSL minimiser of src/nmin.c:361-432 with Adam) ---
 * A nin-64-64-1 tanh network trained on the GPU with mean-squared error and Adam (b1 0.9, b2 0.999, eps 1e-8).
 * Parameters (floats, mg_sur_nparams(nin) of them): W0[64][nin] b0[64] W1[64][64] b1[64] w2[64] b2.
 * Inputs are standardised with (mu, sig) and the target with (ymu, ysig) set by mg_sur_set_norm. */
typedef struct mg_sur mg_sur;
int mg_sur_nparams(int nin);
/* nin <= 1024; weights ~ N(0, 1/fan_in) from the seed, biases 0, identity normalisation */
int mg_sur_create(int device, int nin, unsigned long long seed, mg_sur **out);
void mg_sur_free(mg_sur *s);
int mg_sur_get_params(mg_sur *s, float *p);
/* replaces the parameters and resets the Adam state */
int mg_sur_set_params(mg_sur *s, const float *p);
int mg_sur_set_norm(mg_sur *s, const float *mu, const float *sig, float ymu, float ysig);
/* epochs of seeded-shuffle mini-batch Adam over X[n][nin], y[n]; loss = final mean squared error (normalised) */
int mg_sur_train(mg_sur *s, const float *X, const float *y, int n, int epochs, int batch, float lr,
                 unsigned long long seed, float *loss);
/* The same run split in two: _begin copies the samples and enqueues every step on the surrogate's stream and returns
 * (the caller's arrays are free at once), _end waits for it and gives the final loss.  Any other call on the
 * surrogate waits for a run in flight first.  No counterpart in the reference. */
int mg_sur_train_begin(mg_sur *s, const float *X, const float *y, int n, int epochs, int batch, float lr,
                       unsigned long long seed);
int mg_sur_train_end(mg_sur *s, float *loss);
/* gradient (mg_sur_nparams floats) and value of the mean squared error over all n samples, parameters unchanged */
int mg_sur_grad(mg_sur *s, const float *X, const float *y, int n, float *grad, float *loss);
int mg_sur_predict(mg_sur *s, const float *X, int n, float *y);
/* Features of every structure of the context's last evaluation: per species q, the mean over that structure's
 * atoms of species q of their descriptor rows (D values), zero when q is absent.  feat holds S * nspc * D floats. */
int mg_features(mg_ctx *c, float *feat);

/* --- evolutionary operators (host C; no GPU work) ---
 * Cut-and-splice crossover after BULK_MATE (src/efnc.c:2021-2310) with MATE_LT's lattice choice
 * (src/eutl.c:139-172), then a Gaussian displacement of every atom (sigma, A) and a periodic minimum-distance
 * test (>= dmin), after at most N/4 steps of ADJT_CL's soft repulsion (src/eutl.c:97-135) when it fails.  Parents: npar structures of N species-grouped atoms (spc[npar*N], lat[9*npar] rows,
 * x[3*npar*N] Cartesian) drawn with the given weights; count children go to out_* with the same layout.  A slot
 * whose operators fail max_tries times is flagged in failed[count].  Returns the number of failed slots. */
int mg_evo_offspring(int npar, int N, int nspc, const int *spc, const double *lat, const double *x,
                     const double *weights, int count, double sigma, double dmin, unsigned long long seed,
                     int max_tries, int *out_spc, double *out_lat, double *out_x, int *failed);
/* smallest periodic interatomic distance of one structure (lattice rows L[9], Cartesian x[3n]) */
double mg_evo_min_distance(const double *L, const double *x, int n);
/* Mutation after BULK_MUTE with SHKE_CL (src/efnc.c:2316-2386, src/eutl.c:176-218): a parent drawn by weight, its
 * lattice strained (diagonal 1 + cl g, symmetric off-diagonal 0.5 cl g, g Gaussian; fractional coordinates kept),
 * every coordinate moved by a Gaussian of ca A, the rows reduced (SHRT_LV), the volume within 50% of the parent's,
 * then each atom's position swapped with probability cs with that of an atom of another species, and all periodic
 * distances >= dmin (after ADJT_CL's repair, as in mg_evo_offspring).  Layout and return value as mg_evo_offspring. */
int mg_evo_mutate(int npar, int N, int nspc, const int *spc, const double *lat, const double *x, const double *weights,
                  int count, double cl, double ca, double cs, double dmin, unsigned long long seed, int max_tries,
                  int *out_spc, double *out_lat, double *out_x, int *failed);
/* Structure fingerprints after RDF(C, 1) (src/cell.c:369-415) for S periodic structures (atom_off[S+1], lat[9S],
 * Cartesian x): Gaussian-broadened (sigma) radial distributions of every species pair m <= n on nrdf bins up to rmax,
 * tapered by a cosine from rmin to rmax, same-species pairs halved, scaled to unit length; out[S][npair * nrdf] floats.
 * The dot product of two fingerprints is CxC (src/cell.c:417-427), the structural part of COMP_CL. */
int mg_evo_fingerprint(int S, const int *atom_off, int nspc, const int *spc, const double *lat, const double *x,
                       int nrdf, double rmax, double rmin, double sigma, float *out);
/* The duplicate walk of the population update (RANK_TR's test with COMP_CL, src/emod.c:351-366,
 * src/eutl.c:394-413): the n structures in increasing energy E (ties by index; non-finite skipped) are kept unless
 * 0.8 CxC + 0.1 exp(-dE^2/(2 0.05^2)) + 0.1 exp(-dV^2/(2 0.01^2)) > cut against one already kept, where
 * dV = (V_a - V_b)/(2 (V_a + V_b)); pairs outside E_k - E_kept <= de_win, |dV| <= dv_win are not scored.  The first
 * n_unique structures are never compared with each other.  G[(k - n_unique) n + j] = CxC of structure k >= n_unique
 * with j (may be NULL when cut >= 1).  Up to limit indices go to keep; returns their number. */
int mg_evo_unique(int n, const double *E, const double *V, const float *G, int n_unique, double cut, double de_win,
                  double dv_win, int limit, int *keep);

/* --- device-resident dynamics (state stays on the device between steps) --- */
/* NVT molecular dynamics, replacing Dynamics/Maxwell (src/cmod.c:224-549): velocity Verlet with a Nose-Hoover
 * thermostat xi' = (2K/(3N kB T) - 1)/tau^2 (the reference's equation of motion, src/cmod.c:403-407).
 * mg_md_init draws seeded Maxwell-Boltzmann velocities at T (K), removes the centre-of-mass drift and rescales
 * to exactly T.  mg_md_run advances nsteps of dt (fs) at thermostat temperature T with time constant tau (fs;
 * tau <= 0 gives NVE); epot[nsteps] and ekin[nsteps] (eV, NULL skips) receive the energies after each step. */
int mg_md_init(mg_ctx *c, double T, unsigned long long seed);
int mg_md_set_velocities(mg_ctx *c, const double *v);
int mg_md_get_velocities(mg_ctx *c, double *v);
int mg_md_run(mg_ctx *c, int nsteps, double dt, double T, double tau, double *epot, double *ekin);
/* NPT (tau > 0) or isobaric NVE (tau = 0) steps: as mg_md_run plus the reference's Berendsen barostat
 * (MDTP 3 / 4, src/cmod.c:371, 408-411, 457-468; CELL_PRS src/cmod.c:552-561).  After every force evaluation the
 * lattice scales by lambda = (1 - (ktt dt / 3)(p - p_int))^(1/3), p_int = (U_xx + U_yy + U_zz)/(3V), and the
 * positions by the same factor at the next drift (all atoms; the reference's loop rescales only the first three).
 * p_gpa: target pressure (GPa); ktt = ICMP / CPLP (1/(GPa fs), the reference's default 0.005 / 50).  Optional
 * traces per step: epot, ekin (eV), vol (A^3), pint_gpa (GPa).  Periodic single structure only. */
int mg_md_run_npt(mg_ctx *c, int nsteps, double dt, double T, double tau, double p_gpa, double ktt, double *epot,
                  double *ekin, double *vol, double *pint_gpa);
/* Batched FIRE relaxation of every structure of the context (positions only), replacing the CELL_RELX /
 * CELL_MIN caller (src/cmod.c:901-977, src/cmin.c:177-301; the reference minimiser is GSL BFGS2).  A structure
 * stops when max_i |F_i| < fmax (eV/A); dt0/dtmax are FIRE's initial and maximum step, maxstep (A) caps the
 * per-structure displacement norm of one step.  steps[S] and done[S] (NULL skips) report iterations and
 * convergence; the final positions/energies/forces are read with mg_get / mg_get_batch. */
int mg_relax(mg_ctx *c, int max_steps, double fmax, double dt0, double dtmax, double maxstep, int *steps, int *done);
/* Batched L-BFGS relaxation of every structure of the context (positions only), replacing the same caller with the
 * quasi-Newton method the reference minimises with (CELL_MIN's default MINT 0, GSL vector_bfgs2, src/cmin.c:203-245),
 * in its limited-memory form without a line search: per structure the two-loop recursion over the last `memory`
 * (1..32) step / force-change pairs (pairs without positive curvature are skipped), gamma = s.y / y.y scaling, h0
 * (A^2/eV) as the initial inverse Hessian, steepest descent with the memory dropped when the direction is not a
 * descent direction, and every step scaled so no atom moves more than maxstep (A).  Stops a structure when
 * max_i |F_i| < fmax (eV/A).  Same batching, compaction, graph replay and skin lists as mg_relax; steps/done as
 * there. */
int mg_relax_lbfgs(mg_ctx *c, int max_steps, double fmax, double maxstep, int memory, double h0, int *steps,
                   int *done);
/* Wall time of the last relaxation by segment between its host checks: up to max records of 4 doubles
 * (iterations done at the segment's end, active structures, active atoms, seconds) go to rec; returns the number
 * of records there are.  No counterpart in the reference. */
int mg_relax_log(mg_ctx *c, int max, double *rec);
/* Full-cell relaxation (the reference's RLXT 3, CELL_RELX/CELL_MIN src/cmod.c:901-977, src/cmin.c:90-166: enthalpy
 * H = E + pV over positions and lattice, there with GSL BFGS2): batched FIRE over positions and the cell strain
 * (generalized force (U - pV 1)/N), to every |F| < fmax (eV/A) and every |U - pV 1|/V < smax (GPa), at pressure
 * p (GPa).  Lattices are updated on the device; mg_get_lattice returns them. */
int mg_relax_cell(mg_ctx *c, double p_gpa, int max_steps, double fmax, double smax_gpa, double dt0, double dtmax,
                  double maxstep, int *steps, int *done);
/* Lattice (rows, A) of structure s as last set or relaxed. */
int mg_get_lattice(mg_ctx *c, int s, double *lat);

/* ---- structure and output files (host only; no device needed) -------------------------------------------------
 * POSCAR/CONTCAR read (READ_CELL, src/util.c:820-953): title (first line, up to 199 characters; buffer >= 200),
 * lattice rows lat[9] scaled by the file's factor (A), species (symbols required, as in the reference; spcz[10]
 * atomic numbers, spcn[10] counts), natoms; with cap >= natoms also Cartesian positions x[3 natoms] (direct
 * coordinates converted with the lattice rows; Cartesian ones taken as written, unscaled, as the reference does;
 * clusters, nd = 0, centred on the origin as CENTER(C, 0.0)), selective-dynamics flags ff[3 natoms] (1 = free; may
 * be NULL) and velocities v[3 natoms] when every atom has one (*has_v; v may be NULL).  cap < natoms reads the
 * header and the count only. */
int mg_poscar_read(const char *path, int nd, int cap, char *title, double *lat, int *nspc, int *spcz, int *spcn,
                   int *natoms, double *x, int *ff, double *v, int *has_v);
/* CONTCAR write in the reference's format (SAVE_CELL, src/util.c:955-1030): lattice divided by scale, species symbols,
 * counts, "Selective dynamics" when sd, direct coordinates (clusters, nd = 0, centred in the box as CENTER(C, 0.5))
 * with T/F columns when sd or any flag in ff is 0, then velocities when v is given.  Atoms in species order. */
int mg_poscar_write(const char *path, const char *title, double scale, const double *lat, int nd, int nspc,
                    const int *spcz, const int *spcn, int natoms, const double *x, const int *ff, int sd,
                    const double *v);
/* One optimisation/dynamics iteration appended to dir/OUTCAR and dir/OSZICAR (CELL_OUT, src/cmod.c:618-687): the
 * run header when it == 0 (version, model id, optimizer and relaxation types, pressure, dimensionality, species);
 * the lattice with the volume per atom (CELL_VOL; clusters: NP_VOL, the inertia-tensor volume); with detail,
 * positions, forces and atom energies and the stress U/V (eV/A^3 and kB); the enthalpy/energy line; OSZICAR gets H. */
int mg_outcar_write(const char *dir, int it, const char *version, const char *model_id, int mint, int rlxt,
                    double p_gpa, int nd, int nspc, const int *spcz, const int *spcn, int natoms, const double *lat,
                    const double *x, const double *F, const double *EA, const double *U, double E, double H,
                    int detail);
/* Element symbol of atomic number z (1..100; "--" otherwise) and the atomic number of a symbol (0 if unknown). */
const char *mg_element_symbol(int z);
int mg_element_z(const char *symbol);

#ifdef __cplusplus
}
#endif
#endif
