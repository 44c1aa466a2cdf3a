/*
 * topodda.h -- C ABI of the gfx950 DDA solve path.
 *
 * Plain pointers and sizes only.  Complex numbers are interleaved (re, im)
 * doubles.  Every function returns 0 on success or a TDDA_ERR_* code;
 * tdda_last_error() gives the message of the last failure on this thread.
 *
 * The model entry points replace the reference's pybind11 DDAModel
 * (python_bindings/PythonBindings.cpp:16-70); each declaration cites the
 * reference member it stands for.
 */
#ifndef TOPODDA_H_
#define TOPODDA_H_

#ifdef __cplusplus
extern "C" {
#endif

#define TDDA_OK 0
#define TDDA_ERR_INVALID 1  /* reference: `throw 1` (pybind RuntimeError) */
#define TDDA_ERR_METHOD 2   /* reference: std::invalid_argument (pybind ValueError) */
#define TDDA_ERR_HIP 3      /* HIP runtime failure */
#define TDDA_ERR_MEMORY 4   /* device memory exhausted / over budget */
#define TDDA_ERR_STATE 5    /* call order violated */

#define TDDA_PREC_FP64 0          /* complex128 transforms and Krylov vectors */
#define TDDA_PREC_FP32_REFINED 1  /* complex64 transforms/Krylov, FP64 reliable updates */

#define TDDA_PRECOND_NONE 0    /* the reference's plain BiCGSTAB (DDAModel.cpp:899-1009) */
#define TDDA_PRECOND_JACOBI 1  /* right Jacobi: BiCGSTAB on A diag(1/alpha)^-1, same residual */

#define TDDA_GRAD_REFERENCE 0  /* one-sided FD partials as DDAModel.cpp:471-578 */
#define TDDA_GRAD_EXACT 1      /* closed-form partials, same adjoint solve */

#define TDDA_STATUS_ACTIVE 0
#define TDDA_STATUS_CONVERGED 1
#define TDDA_STATUS_BREAKDOWN 2
#define TDDA_STATUS_MAXIT 3

typedef struct tdda_model tdda_model;

/* Constructor arguments of DDAModel, in the order and meaning of
 * python_bindings/PythonBindings.cpp:19-59 (src/DDAModel.cpp:79). */
typedef struct {
    double betaMin, betaMax, ita;
    const char* betaType;
    const int* filterIterations;
    long nFilterIterations;
    const double* filterRadii;
    long nFilterRadii;
    int useFilter;
    const char* symmetryType; /* "None" | "4fold" */
    const double* symmetryAxis;
    long nSymmetryAxis;
    int isPeriodic;
    const char* objectiveName;
    const double* objPara;
    long nObjPara;
    const int* geometry; /* 3*num_total */
    long nGeometry;
    const double* inputDielectrics; /* 3*num_total */
    long nInputDielectrics;
    int num_x, num_y, num_z, num_total;
    const double* n_K; /* 3 */
    double E0;
    const double* n_E0; /* 3 */
    double lam;
    const double* material; /* nMaterial complex, interleaved */
    long nMaterial;
    double nback;
    int MAXm, MAXn;
    double Lm, Ln;                  /* image periods in cells */
    const char* matrixAMethodName;  /* "FCD" | "LDR" */
    double d;
    const double* sineIntegralValues;
    long nSine;
    const double* cosineIntegralValues;
    long nCosine;
    double integralDelta;
    int verbose;
} tdda_model_args;

typedef struct {
    int status;      /* TDDA_STATUS_* */
    int iterations;  /* the reference's ITERATION */
    double error;    /* final ||r|| / ||rhs|| (recursive residual in FP64 mode, true residual in refined mode) */
    int reliable_updates;
    double ms;       /* wall time of the solve */
    double err_est;  /* refined mode: estimated relative error of E = al P 3/(eps+2) (0: none) */
} tdda_solve_info;

/* last failure message of this thread ("" if none) */
const char* tdda_last_error(void);

/* DDAModel::DDAModel (src/DDAModel.cpp:79-333), PythonBindings.cpp:19-59 */
int tdda_model_create(const tdda_model_args* args, tdda_model** out);
/* DDAModel::~DDAModel (src/DDAModel.cpp:335-338) */
int tdda_model_destroy(tdda_model* m);
/* DDAModel::solveElectricField (src/DDAModel.cpp:874-885), PythonBindings.cpp:64;
 * P0: n complex (n = 3*num_total) */
int tdda_model_solve_electric_field(tdda_model* m, const double* P0, long n, int maxit, double tol);
/* DDAModel::calculateObjective (src/DDAModel.cpp:870-872), PythonBindings.cpp:61 */
int tdda_model_calculate_objective(tdda_model* m, double* value);
/* DDAModel::calculateGradients (src/DDAModel.cpp:398-469), PythonBindings.cpp:62;
 * grad: n = NFpara doubles */
int tdda_model_calculate_gradients(tdda_model* m, double epsilon_partial, double original_obj, int maxit,
                                   double tol, double* grad, long n);
/* DDAModel::UpdateParameters (src/DDAModel.cpp:747-772), PythonBindings.cpp:66 */
int tdda_model_set_parameters(tdda_model* m, const double* values, long n);
/* DDAModel::get_parameter_copy (src/DDAModel.cpp:1341-1343), PythonBindings.cpp:67 */
int tdda_model_get_parameters(tdda_model* m, double* out, long n);
/* DDAModel::get_dielectrics_copy (src/DDAModel.cpp:1346-1348), PythonBindings.cpp:68; n = 3*num_total */
int tdda_model_get_dielectrics(tdda_model* m, double* out, long n);
/* DDAModel::get_EResult (src/DDAModel.cpp:1254-1256), PythonBindings.cpp:69; n = 3*num_total complex */
int tdda_model_get_electric_field(tdda_model* m, double* out, long n);

/* Sizes for the getters above (NFpara, num_total). */
int tdda_model_sizes(tdda_model* m, long* nfpara, long* ndipoles);
/* Solver precision for later solves (TDDA_PREC_*). No reference counterpart: the reference is FP64 only. */
int tdda_model_set_precision(tdda_model* m, int precision);
/* Gradient partials mode (TDDA_GRAD_*). No reference counterpart. */
int tdda_model_set_gradient_mode(tdda_model* m, int mode);
/* Iteration count / residual of the last forward (which=0) or adjoint (which=1) solve;
 * the reference keeps ITERATION and Error unbound (src/DDAModel.h:64-66). */
int tdda_model_solve_info(tdda_model* m, int which, tdda_solve_info* out);
/* Current polarisation P (n = 3*num_total complex) and 1/alpha per dipole (n = num_total complex). */
int tdda_model_get_polarization(tdda_model* m, double* out, long n);
int tdda_model_get_inverse_alpha(tdda_model* m, double* out, long n);
/* y = (diag(al) + A) x on the device in FP64 (DDAModel::Aproductwithalb, src/DDAModel.cpp:1194-1203). */
int tdda_model_apply(tdda_model* m, const double* x, double* y, long n);
/* FFT grid chosen for the model (Lx, Ly, Lz) and Green's setup time in ms. */
int tdda_model_grid(tdda_model* m, int* lxyz, double* setup_ms);

/* Re-solve the forward system from P = 0 with the right-hand side and 1/alpha
 * already on the device and no copy back (throughput timing of
 * DDAModel::bicgstab, src/DDAModel.cpp:899-1009). */
int tdda_model_solve_resident(tdda_model* m, int maxit, double tol, tdda_solve_info* out);
/* Average ms per launch of the five matvec passes (X fwd, Y fwd, Z, Y inv,
 * X inv) at the current precision, timed with HIP events on the library stream;
 * the matvec is AProductCore::Aproduct + al*b (src/AProductCore.cpp:461-515). */
int tdda_model_time_passes(tdda_model* m, int reps, double* ms5);

/* Batched solves record one BiCGSTAB iteration as a HIP graph and replay it
 * (on = 1, the default); 0 issues the same launches one by one.  Results are
 * the same either way.  No reference counterpart. */
int tdda_set_graph_capture(int on);
/* c64 transforms of boxes whose y and z lengths allow it can run the y, z
 * and inverse y passes as one kernel per x frequency.  mode: TDDA_FUSED_OFF
 * never, TDDA_FUSED_SCHEDULE where the kernel schedule enables it for the
 * box's (Ly, Lz) (the default), TDDA_FUSED_ALL always.  Applies to batches
 * built afterwards; results agree to c64 rounding.  No reference counterpart. */
#define TDDA_FUSED_OFF 0
#define TDDA_FUSED_SCHEDULE 1
#define TDDA_FUSED_ALL 2
int tdda_set_fused_yz(int mode);
/* c64 transforms of boxes with Lx = Ly in {32, 48, 64, 96, 128} can run the
 * x and y transforms of each z-plane as one kernel per direction (the plane
 * stays in LDS between them; on by default).  on = 0: the separate X and Y
 * passes.  Applies to batches built afterwards; results agree to c64
 * rounding.  No reference counterpart. */
int tdda_set_plane_passes(int on);

/* Kernel schedule: the split variant of each length-specialised pass per
 * transform length and precision, the fused y/z choice per (Ly, Lz) and the
 * plane-pass choice per plane length L (plane passes take precedence over the
 * fused y/z pass).  Text, one entry per line: "var <c64|c128>
 * <xinv|y|z|xfwd> <L> <variant>", "yz <Ly> <Lz> <0|1>", "xy <L> <0|1>";
 * absent entries mean variant 0, no fused y/z pass and plane passes on.  It
 * is read when systems are made (and batches built), and loaded at start-up
 * from the file named by the environment variable TDDA_SCHEDULE.  Text
 * outputs are NUL-terminated and cut to cap bytes; *len receives the full
 * length.  No reference counterpart (the reference has one cuFFT plan per
 * model). */
int tdda_schedule_text(char* buf, long cap, long* len);
int tdda_schedule_parse(const char* text);   /* replaces every entry; a malformed line leaves it unchanged */
int tdda_schedule_load(const char* path);
int tdda_schedule_save(const char* path);
int tdda_schedule_clear(void);
/* Autotuner: for count LDR problems of each of the n boxes (nx, ny, nz
 * triples), times every split variant of every pass of the c64 and c128
 * matvecs (reps batched launches each), the c64 matvec with the fused y/z
 * pass against three passes, and per plane length the c64 matvec with the
 * plane passes against those choices; records the faster choices in the
 * schedule, and writes one report line per decision ("pass <prec> <pass> <L>
 * ms <v0> <v1|-> chosen <v>", "yz <Ly> <Lz> ms <three> <fused> chosen <0|1>",
 * "xy <L> ms <off> <on> chosen <0|1>").  A choice
 * leaves the default only when the alternative is at least 3 % faster (15 %
 * for a Z split whose kernel class differs from the default's).  Run it while
 * no other work is being set up: the variants it tries apply to every system
 * made meanwhile. */
int tdda_tune(const int* boxes, int n, int count, int reps, char* report, long cap, long* len);
/* Bytes of device memory currently held by the library. */
long tdda_device_bytes(void);
/* Drop cached Green's spectra nobody uses and return cached device blocks to
 * HIP (no reference counterpart: the reference frees on model destruction). */
int tdda_trim_caches(void);

/* ------------------------------------------------------------------------
 * Device-resident problems, batched solves and the request server.
 * No reference counterpart: the reference solves one model at a time.
 * A problem is a full Nx*Ny*Nz box (x-fastest dipoles, _generate_geometry of
 * python_bindings/dda_model/__init__.py:8-20) with one density per voxel.
 * ------------------------------------------------------------------------ */
typedef struct tdda_problem tdda_problem;
typedef struct tdda_server tdda_server;

typedef struct {
    int nx, ny, nz;
    double d;          /* nm */
    double lam;        /* nm */
    double nback;
    int method;        /* 0 FCD, 1 LDR */
    int maxm, maxn;    /* image lattice range (0: non-periodic) */
    double lm, ln;     /* image periods in cells */
    double n_k[3];
    double n_e0[3];
    double e0;
    double material[4]; /* background, foreground permittivity (re, im) */
} tdda_problem_desc;

/* precision: TDDA_PREC_*; tables may be null for LDR */
int tdda_problem_create(const tdda_problem_desc* d, int precision, const double* si, const double* ci, long nsici,
                        double delta, tdda_problem** out);
/* n problems at once: the Green's set-ups of their new operators (the
 * AProductCore constructor's lattice sum and spectrum, src/AProductCore.cpp:
 * 12-298, once per distinct operator) run as one launch per step for all of
 * them instead of five launches per operator; out: n handles */
int tdda_problems_create(const tdda_problem_desc* d, int n, int precision, const double* si, const double* ci,
                         long nsici, double delta, tdda_problem** out);
int tdda_problem_destroy(tdda_problem* p);
/* rho: nx*ny*nz densities (clipped to [0,1]); on_device: rho is a device pointer */
int tdda_problem_set_density(tdda_problem* p, const double* rho, int on_device);
/* solve all problems together (one batched launch per pass); infos: n entries */
int tdda_problems_solve(tdda_problem** ps, int n, int maxit, double tol, tdda_solve_info* infos);
/* Preconditioner of the problem's later solves (TDDA_PRECOND_*); no reference
 * counterpart: the reference's BiCGSTAB is unpreconditioned.  Right
 * preconditioning keeps the residual, hence the stopping rule, unchanged. */
int tdda_problem_set_preconditioner(tdda_problem* p, int precond);
/* FFT grid (Lx, Ly, Lz), Green's set-up time of the problem's operator in ms
 * (lattice sum, DFTs, spectrum layout; AProductCore::AProductCore,
 * src/AProductCore.cpp:12-298) and whether it came from the operator cache. */
int tdda_problem_grid(tdda_problem* p, int* lxyz, double* setup_ms, int* cached);
/* E = al P 3/(eps+2): 3N complex, to host or device memory */
int tdda_problem_field(tdda_problem* p, double* E, int on_device);
/* polarisation P: 3N complex, to host or device memory */
int tdda_problem_polarization(tdda_problem* p, double* P, int on_device);
#define TDDA_OBJ_INTEGRATED 0 /* log sum_i sum_c |al_i P_ic|^power (ObjDDAModel.cpp:130-212) */
#define TDDA_OBJ_TARGETS 1    /* mean |E|^2 over external points (DDAModel::solve_E, DDAModel.cpp:1153-1183) */
/* Objective of the current solution.  targets: ntargets (x,y,z) in nm for
 * TDDA_OBJ_TARGETS; target_fields (nullable): 3*ntargets complex E. */
int tdda_problem_objective(tdda_problem* p, int kind, int power, const double* targets, int ntargets, double* value,
                           double* target_fields);
/* d objective / d rho for every voxel by one adjoint solve (DDAModel::calculateGradients,
 * src/DDAModel.cpp:398-469, with closed-form partials); grad: nx*ny*nz doubles,
 * host or device.  value (nullable): the objective; adj (nullable): adjoint solve record. */
int tdda_problem_gradient(tdda_problem* p, int kind, int power, const double* targets, int ntargets, int maxit,
                          double tol, double* grad, int on_device, double* value, tdda_solve_info* adj);
/* y = (diag(al) + A) x for all n problems in one batched launch per pass
 * (AProductCore::Aproduct + DDAModel::Aproductwithalb, src/AProductCore.cpp:461-515,
 * src/DDAModel.cpp:1194-1203), with c64 transforms (c64 != 0; problems of
 * precision TDDA_PREC_FP32_REFINED) or FP64 ones.  x, y: host arrays of the
 * problems' vectors back to back, 3*nx*ny*nz complex each. */
int tdda_problems_apply(tdda_problem** ps, int n, int c64, const double* x, double* y);
/* y = (diag(al) + A) x with FP64 transforms for one problem, allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  x, y: DEVICE arrays of 3*nx*ny*nz
 * complex (interleaved doubles) that the caller owns.  The job descriptor and
 * pass maps go to a fixed 256 KiB device array per GPU (a problem needs 1 KiB
 * plus 8 bytes per pass tile; larger ones are refused with TDDA_ERR_MEMORY).
 * The work is enqueued on the library stream and not waited for; calls that
 * overlap on one GPU share that array and race. */
int tdda_problem_apply_noalloc(tdda_problem* p, const double* x, double* y);
/* Average ms per launch of the five matvec passes over all n problems in one
 * batched launch per pass, c64 (precision TDDA_PREC_FP32_REFINED problems) or
 * c128 transforms (c64 != 0 selects c64); HIP events on the library stream.
 * Jobs on the plane passes count their x+y forward kernel in slot 0 and their
 * inverse y+x kernel in slot 4; a slot with no launch reads 0. */
int tdda_problems_time_passes(tdda_problem** ps, int n, int c64, int reps, double* ms5);
/* Average ms per launch of the ten passes of one refined BiCGSTAB step over
 * all n problems (TDDA_PREC_FP32_REFINED): X fwd with the p update, Y, Z, Y,
 * X inv with <r0, v>; X fwd with t, Y, Z, Y, X inv with the step's dots.
 * Starts from P = 0 and overwrites the problems' solver state. */
int tdda_problems_time_iteration(tdda_problem** ps, int n, int reps, double* ms10);

typedef struct {
    int precision;       /* TDDA_PREC_* */
    int maxit;
    double tol;
    double ru_delta;     /* reliable-update trigger (refined precision); 0 = 0.01 */
    double hbm_budget_gb;
    int max_slots;       /* requests in flight */
    int chunk;           /* iterations between host checks */
    int keep_fields;     /* keep E of finished requests for tdda_server_take_field */
    double err_tol;      /* refined: relative field-error target; 0 = tol, < 0 = true residual only */
    int precond;         /* TDDA_PRECOND_* for every solve */
} tdda_server_opts;

typedef struct {
    long id;
    int status;          /* TDDA_STATUS_* */
    int iterations;
    int reliable_updates;
    double error;
    double t_submit, t_admit, t_done;  /* seconds since the server started */
    int operator_cached;
    double err_est;      /* refined: estimated relative field error */
} tdda_result;

typedef struct {
    long submitted, completed, chunks, rebuilds, launches;
    int max_active;
    long peak_bytes;
    /* host time of the loop thread (ms): admissions, waits for a chunk,
     * retirements, batch builds (maps and joins) */
    double host_admit_ms, host_wait_ms, host_retire_ms, host_build_ms;
    /* loop thread: enqueueing chunks (ms), and chunks replayed from a graph */
    double host_enqueue_ms;
    long graph_chunks;
} tdda_server_stats;

int tdda_server_create(const tdda_server_opts* o, const double* si, const double* ci, long nsici, double delta,
                       tdda_server** out);
int tdda_server_destroy(tdda_server* s);
int tdda_server_submit(tdda_server* s, long id, const tdda_problem_desc* d, const double* rho, long n);
int tdda_server_poll(tdda_server* s, tdda_result* out, int max, int* n_out);
/* Moves the field of a finished request out (3*nx*ny*nz complex); a buffer of
 * the wrong size is refused and the field stays. */
int tdda_server_take_field(tdda_server* s, long id, double* E, long n);
int tdda_server_idle(tdda_server* s, int* idle);
int tdda_server_get_stats(tdda_server* s, tdda_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* TOPODDA_H_ */
