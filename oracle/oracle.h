/*
 * oracle.h -- FP64 CPU restatement of UQCreator's IPM moment-solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check results against
 * (or the timed CPU baseline).  The product (uqcreator_amd/) never links it.
 *
 * Every routine names the reference file:line it restates (paths relative to
 * the UQCreator repository root).  Parity status of each routine is recorded
 * in the header comment of oracle.cpp.
 *
 * Layouts (identical to the reference's host layout):
 *   per-cell moments / duals : Matrix(S, Mmax) column-major, state fastest:
 *                              element (s, i) at [i*S + s]; cells contiguous.
 *   per-cell quadrature vals : Matrix(S, Q) column-major: (s, k) at [k*S + s].
 *   basis tables phi, phiw   : row k contiguous: (k, i) at [k*M + i].
 */
#ifndef UQK_ORACLE_H
#define UQK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_BB          = 0, /* bounded barrier, S states independent      */
    ORC_EULER1D     = 1, /* Euler 1D entropy closure, S = 3             */
    ORC_EULER2D     = 2, /* Euler 2D entropy closure, S = 4             */
    ORC_SW2D        = 3, /* shallow water 2D quadratic entropy, S = 3   */
    ORC_SW2D_BARRIER = 4 /* shallow water 2D with log barrier on h       */
};

typedef struct {
    int kind;
    int S;
    double gamma;       /* Euler adiabatic exponent (positive)            */
    double u_minus;     /* bounded barrier lower bound                    */
    double u_plus;      /* bounded barrier upper bound                    */
    double g;           /* gravity (shallow water)                        */
    double barrier;     /* log-barrier strength on h (SW2D_BARRIER)       */
    double alpha;       /* Newton damping                                 */
    double eps;         /* convergence threshold on ||grad L||_2          */
    int max_iter;       /* Newton iterations                              */
    int max_halvings;   /* line-search halvings before failure (1000)     */
    int du_exact;       /* 0: Euler1D DU as shipped, 1: exact DU          */
    int safe;           /* 1: SolveClosureSafe line search (reject NaN)   */
} orc_closure;

/* status codes per cell */
enum { ORC_OK = 0, ORC_MAXIT = 1, ORC_LINESEARCH = 2, ORC_NOT_SPD = 3, ORC_NONFINITE = 4 };

/* ---- quadrature, basis, index set ------------------------------------ */
int orc_gauss_legendre(int n, double* x, double* w);
double orc_legendre(int n, double x);
int orc_index_set(int d, int nlev, const int* degrees, int max_degree, int cap, int* idx, int* ntotal_per_level);
int orc_tables(int d, int q1, int ntotal, const int* idx, double* xi, double* wq, double* phi, double* phiw);

/* ---- entropy closures --------------------------------------------------- */
void orc_U(const orc_closure* c, const double* lam, double* out);
void orc_U_mat(const orc_closure* c, const double* lam, double* out);
void orc_DU(const orc_closure* c, const double* lam, double* y);
void orc_DS(const orc_closure* c, const double* u, double* ds);

/* ---- closure pieces ------------------------------------------------------ */
void orc_gradient(const orc_closure* c, int M, int Q, int nt, const double* phi, const double* phiw,
                  const double* lam, const double* u, double* g);
void orc_hessian(const orc_closure* c, int M, int Q, int nt, const double* phi, const double* phiw,
                 const double* lam, double* H);
int orc_posv_upper(int n, double* H, double* b);
int orc_solve_closure(const orc_closure* c, int M, int Q, const double* phi, const double* phiw, long ncell,
                      const int* nt_cell, const double* u, double* lam, int* status, int* iters, int nthreads);

/* ---- sweeps ---------------------------------------------------------------- */
void orc_eval_uq(const orc_closure* c, int M, int Q, const double* phi, long ncell, const int* nt_cell,
                 const double* lam, double* uq, int nthreads);
void orc_reconstruct(int S, int M, int Q, const double* phiw, long ncell, const int* nt_cell,
                     const double* uq, double* u, int nthreads);
void orc_mean_var(const orc_closure* c, int M, int Q, const double* phi, const double* phiw, long ncell,
                  const int* nt_cell, const double* lam, double* mv, int nthreads);

/* ---- problems, mesh, time loop ----------------------------------------- */
enum { ORC_P_BURGERS = 0, ORC_P_EULER2D = 1, ORC_P_SW2D = 2 };
enum { ORC_FLUX_REF = 0, ORC_FLUX_LF = 1 };
enum { ORC_BC_NONE = 0, ORC_BC_DIRICHLET = 1, ORC_BC_NOSLIP = 2, ORC_BC_SWWALL = 3 };

typedef struct {
    int problem;
    int flux;
    double cfl;
    double gamma;
    double g;
    double friction;     /* SW2D bed friction base coefficient (0: off)  */
    int friction_dim;    /* xi dimension carrying friction uncertainty   */
} orc_problem;

typedef struct {
    long ncell;
    int nedge;
    int dim;
    const int* nbr;             /* ncell*nedge; ncell marks a boundary   */
    const double* unit_normal;  /* ncell*nedge*dim                       */
    const double* normal;       /* ncell*nedge*dim (scaled by length)    */
    const double* area;         /* ncell                                 */
    const double* dx_param;     /* ncell: maxEdge/area                   */
    const int* bc;              /* ncell                                 */
    const double* bnormal;      /* ncell*dim boundary unit normal        */
} orc_mesh;

void orc_flux_point(const orc_problem* p, int S, const double* u, const double* v, const double* nu,
                    const double* n, int dim, double* out);
double orc_compute_dt(const orc_problem* p, int S, int Q, const double* uq, double dx);
int orc_advance(const orc_problem* p, const orc_mesh* mesh, int S, int M, int Q, int d, const double* phiw,
                const double* xi, const int* nt_cell, double dt, const double* u, const double* uq, double* unew,
                int nthreads);
int orc_run(const orc_closure* c, const orc_problem* p, const orc_mesh* mesh, int M, int Q, int d,
            const double* phi, const double* phiw, const double* xi, const int* nt_cell, double* u, double* lam,
            double t_end, long max_steps, double* t_out, long* steps_out, double* residual_out,
            long* newton_iters_out, int nthreads);

/* ---- error norm ------------------------------------------------------------ */
void orc_rel_error(int nf, long ncell, const double* sol, const double* ref, const double* area,
                   const unsigned char* mask, int lnorm, double* err);

#ifdef __cplusplus
}
#endif
#endif
