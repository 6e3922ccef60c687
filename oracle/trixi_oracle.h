/*
 * trixi_oracle.h -- CPU oracle for the DGSEM rhs! on a periodic TreeMesh.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a C++ (C-ABI), FP64 restatement of the
 * algorithm Trixi.jl runs on the CPU for the semidiscretization that
 * TrixiCUDA.jl offloads (reference: src/solvers/dg_{1,2,3}d.jl and
 * src/solvers/dg_{1,2,3}d_kernel.jl).  It exists to check the HIP library;
 * it is never linked into, called by, or shipped as part of the product path.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
 *
 * Array conventions follow Trixi.jl (Julia, column-major, 1-based) mapped to
 * 0-based C offsets:
 *   u[v + nv*(i + N*(j + N*(k + N*e)))]          ("[v,i,j,k,e]")
 *   node_coordinates[x + d*(i + N*(j + N*(k + N*e)))]
 *   neighbor_ids[s + 2*iface]  s=0 left, s=1 right (0-based element ids)
 *   orientations[iface] in {1,..,ndims}
 *   matrices row-major: M[r*N + c] == Julia M[r+1, c+1]
 */
#ifndef TRIXI_ORACLE_H
#define TRIXI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_EQ_LINADV = 0, OR_EQ_EULER = 1, OR_EQ_GLM_MHD = 2, OR_EQ_BURGERS = 3, OR_EQ_HYPDIFF = 4,
       OR_EQ_EULER_MULTI = 5 /* (rho v, rho e, rho_1..rho_nc) */,
       OR_EQ_EULER_QUASI1D = 6 /* (a rho, a rho v, a e, a) */,
       OR_EQ_SHALLOW_WATER = 7 /* 2D: (h, h v1, h v2, b) */ };
enum { OR_VI_WEAK = 0, OR_VI_FD = 1, OR_VI_HG = 2 };
enum {
    OR_FLUX_CENTRAL = 0,
    OR_FLUX_LLF = 1,          /* FluxLaxFriedrichs(max_abs_speed)            */
    OR_FLUX_LLF_NAIVE = 2,    /* FluxLaxFriedrichs(max_abs_speed_naive)      */
    OR_FLUX_RANOCHA = 3,
    OR_FLUX_SHIMA_ETAL = 4,
    OR_FLUX_HINDENLANG_GASSNER = 5,
    OR_FLUX_EC = 6,           /* Burgers flux_ec                             */
    OR_FLUX_HLL = 7,          /* FluxHLL(min_max_speed_davis), Euler         */
    OR_FLUX_CHAN_ETAL = 8,    /* quasi-1D: flux_chan_etal + flux_nonconservative_chan_etal */
    OR_FLUX_HLLE = 9,         /* GLM-MHD: FluxHLL(min_max_speed_einfeldt) */
    OR_FLUX_WINTERMEYER_ETAL = 10, /* shallow water (volume noncons g h_l b_r) */
    OR_FLUX_FJORDHOLM_ETAL = 11,   /* shallow water (surface noncons: Fjordholm) */
    OR_FLUX_HYDROSTATIC_LLF = 12   /* shallow water: LLF of Audusse's hydrostatic reconstruction */
};
enum {
    OR_IC_CONSTANT = 0,
    OR_IC_CONVERGENCE_TEST = 1, /* LinAdv: 1+0.5 sin(pi (x-a t)); MHD: Alfven wave */
    OR_IC_WEAK_BLAST_WAVE = 2,
    OR_IC_SEEDED_RANDOM = 3,    /* Euler: rho=1+0.1U, v=0.1N(0,1), p=1+0.05U     */
    OR_IC_SMOOTH_PERIODIC = 4,  /* Euler: seeded sums of periodic sines           */
    OR_IC_WEAK_BLAST_WAVE_SEEDED = 5, /* Euler: weak blast + U(-0.01,0.01) on rho,v,p */
    OR_IC_SHOCK = 6,          /* Burgers: 1.5 for x < 0.5, else 0.5 */
    OR_IC_RAREFACTION = 7,    /* Burgers: 0.5 for x < 0.5, else 1.5 */
    OR_IC_EOC_GRAVITY = 8,    /* Euler: initial_condition_eoc_test_coupled_euler_gravity */
    OR_IC_POISSON_NONPERIODIC = 9,  /* hyperbolic diffusion: guess at t = 0, else steady */
    OR_IC_HARMONIC_NONPERIODIC = 10 /* hyperbolic diffusion 1D harmonic state */
};
enum { OR_BC_PERIODIC = 0, OR_BC_DIRICHLET = 1, OR_BC_SLIP_WALL = 2 };
enum { OR_SOURCE_NONE = 0, OR_SOURCE_CONVERGENCE_TEST = 1, OR_SOURCE_EOC_TEST_EULER = 2,
       OR_SOURCE_POISSON_NONPERIODIC = 3, OR_SOURCE_HARMONIC = 4 };

typedef struct oracle_cfg {
    int32_t ndims;            /* 1, 2, 3 */
    int32_t equation;         /* OR_EQ_* */
    int32_t polydeg;          /* N = polydeg + 1 nodes per direction */
    int32_t volume_integral;  /* OR_VI_* */
    int32_t volume_flux;      /* OR_FLUX_* (DG volume flux for FD and HG) */
    int32_t surface_flux;     /* OR_FLUX_* */
    int32_t fv_flux;          /* OR_FLUX_* (HG subcell flux) */
    int32_t alpha_smooth;     /* HG smoothing on/off */
    double gamma;
    double c_h;               /* GLM cleaning speed */
    double advection_velocity[3];
    double alpha_max, alpha_min;
    double ic_length;         /* domain period used by the smooth periodic IC */
    int32_t boundary_condition; /* OR_BC_*: faces without an interface          */
    int32_t boundary_ic;      /* Dirichlet state: OR_IC_*                        */
    uint64_t boundary_seed;
    int32_t source_terms;     /* OR_SOURCE_*                                     */
    int32_t indicator_variable; /* HG: 0 density*pressure, 1 pressure, 2 density */
    const double* node_coordinates; /* [E][nodes][d], for boundary states and sources */
    double time;              /* rhs time; the start of the step for the RK step JVP */
    int64_t nmortars;         /* L2 mortars (Trixi init_mortars! container):          */
    const int32_t* mortar_neighbor_ids; /* [nmortars][2^(d-1)+1] small..., large   */
    const int32_t* mortar_orientations; /* [nmortars] 1..d                          */
    const int32_t* mortar_large_sides;  /* [nmortars] 1: large on the left, 2: right */
    int32_t boundary_outflow_faces; /* bit f (-x,+x,-y,+y,-z,+z): boundary flux f(u_inner) */
    int32_t reserved;
    double nu, Lr;            /* hyperbolic diffusion: T_r = Lr^2 / nu */
    int32_t ncomponents;      /* multicomponent Euler */
    int32_t reserved2;
    double gammas[4], gas_constants[4];
    double gravity, H0;       /* shallow water */
} oracle_cfg;

int oracle_nvars(const oracle_cfg* cfg);

/* ---- LGL basis (Trixi basis_lobatto_legendre.jl algorithms) -------------- */
int oracle_basis(int nnodes, double* nodes, double* weights, double* D,
                 double* Dsplit, double* Dhat, double* lhat_minus,
                 double* lhat_plus, double* Vinv);
int oracle_interpolation_matrix(int nin, const double* nodes_in, int nout,
                                const double* nodes_out, double* Vdm);

/* ---- ln_mean / inv_ln_mean (Trixi math.jl) -------------------------------- */
double oracle_ln_mean(double x, double y);
double oracle_inv_ln_mean(double x, double y);

/* ---- point fluxes ---------------------------------------------------------- */
void oracle_flux(const oracle_cfg* cfg, const double* u, int orientation, double* f);
void oracle_two_point_flux(const oracle_cfg* cfg, int kind, const double* ull,
                           const double* urr, int orientation, double* f);
void oracle_noncons_flux(const oracle_cfg* cfg, const double* ull,
                         const double* urr, int orientation, double* f);
void oracle_max_abs_speeds(const oracle_cfg* cfg, const double* u, double* lam);
double oracle_entropy(const oracle_cfg* cfg, const double* u);
void oracle_cons2entropy(const oracle_cfg* cfg, const double* u, double* w);
void oracle_entropy_potential(const oracle_cfg* cfg, const double* u,
                              int orientation, double* psi);

/* ---- uniform periodic TreeMesh (Trixi TreeMesh + init_elements/interfaces) */
int64_t oracle_mesh_nelements(int ndims, int level);
int oracle_mesh_uniform(int ndims, int level, double coord_min, double coord_max,
                        int polydeg, int32_t* neighbor_ids, int32_t* orientations,
                        double* inverse_jacobian, double* node_coordinates,
                        int32_t* cell_index /* [ndims*E] integer cell coords */);
/* The same mesh, periodic in direction d iff bit d of `periodic` is set:
 * no interfaces across the domain boundary in the other directions (those
 * faces are boundary faces). */
int64_t oracle_mesh_ninterfaces(int ndims, int level, int periodic);
int oracle_mesh(int ndims, int level, double coord_min, double coord_max, int polydeg, int periodic,
                int32_t* neighbor_ids, int32_t* orientations, double* inverse_jacobian,
                double* node_coordinates, int32_t* cell_index);
/* L2 mortar operators (row-major N x N): forward = interpolation from the
 * large face to the lower / upper half, reverse = L2 projection back
 * (Gauss quadrature). */
int oracle_mortar_matrices(int nnodes, double* forward_lower, double* forward_upper,
                           double* reverse_lower, double* reverse_upper);
/* source_terms_convergence_test at x, t (compressible Euler) */
void oracle_source_convergence_test(const oracle_cfg* cfg, const double* x, double t, double* s);

/* ---- initial conditions ------------------------------------------------------ */
void oracle_initial_condition(const oracle_cfg* cfg, int ic, const double* x,
                              double t, uint64_t seed, uint64_t node_index,
                              double* u);
void oracle_compute_coefficients(const oracle_cfg* cfg, int ic, int64_t nelements,
                                 const double* node_coordinates, double t,
                                 uint64_t seed, double* u);

/* ---- semidiscretization ------------------------------------------------------- */
void oracle_rhs(const oracle_cfg* cfg, int64_t nelements, int64_t ninterfaces,
                const int32_t* neighbor_ids, const int32_t* orientations,
                const double* inverse_jacobian, const double* u, double* du,
                double* alpha_out /* nullable, HG blending factors [E] */);
void oracle_indicator_hg(const oracle_cfg* cfg, int64_t nelements,
                         int64_t ninterfaces, const int32_t* neighbor_ids,
                         const double* u, double* alpha);

/* ---- callbacks ----------------------------------------------------------------- */
double oracle_max_dt(const oracle_cfg* cfg, int64_t nelements,
                     const double* inverse_jacobian, const double* u);
void oracle_error_norms(const oracle_cfg* cfg, int ic, int64_t nelements,
                        const double* inverse_jacobian,
                        const double* node_coordinates, const double* u, double t,
                        double total_volume, uint64_t seed, double* l2,
                        double* linf);
void oracle_integrate(const oracle_cfg* cfg, int64_t nelements,
                      const double* inverse_jacobian, const double* u,
                      double* integral /* [nvars] */);

/* ---- time integration: CarpenterKennedy2N54 ------------------------------------- */
void oracle_ck2n54_coefficients(double* A, double* B, double* c);
/* One step of size dt; k is the 2N register (overwritten). */
void oracle_ck2n54_step(const oracle_cfg* cfg, int64_t nelements,
                        int64_t ninterfaces, const int32_t* neighbor_ids,
                        const int32_t* orientations,
                        const double* inverse_jacobian, double* u, double* k,
                        double* du_scratch, double t, double dt);

/* ---- forward-mode JVP of rhs! by dual numbers (checks for the adjoint) ---------- */
void oracle_rhs_jvp(const oracle_cfg* cfg, int64_t nelements, int64_t ninterfaces,
                    const int32_t* neighbor_ids, const int32_t* orientations,
                    const double* inverse_jacobian, const double* u,
                    const double* udot, double* du, double* dudot);

void oracle_ck2n54_step_jvp(const oracle_cfg* cfg, int64_t nelements, int64_t ninterfaces,
                            const int32_t* neighbor_ids, const int32_t* orientations,
                            const double* inverse_jacobian, const double* u,
                            const double* udot, double dt, double* u_next,
                            double* u_next_dot);

/* ---- OpenMP control for the timed CPU baseline ---------------------------------- */
int oracle_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
