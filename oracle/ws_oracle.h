/*
 * ws_oracle.h -- CPU FP64 oracle for the inviscid-RHS path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py load this code, and only as a checker; the
 * shipped library never links it.
 *
 * Layouts (all dense, double precision):
 *   ghosted field   f[(i+4) + NXG*((j+4) + NYG*(k+4))], NXG = nx+8, NYG = ny+8,
 *                   i in [-4, nx+4): the reference's cudaField with LAP = 4
 *                   (include/windstencil/cuda_utility.h:21, get_Field_LAP);
 *   interior SoA    f[i + nx*(j + ny*(k + nz*m))], m = 0..4: the reference's
 *                   cudaSoA (cuda_utility.h:30, get_SoA).
 */
#ifndef WS_ORACLE_H
#define WS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int nx, ny, nz;          /* interior extents */
    double gamma;            /* ratio of specific heats (reference Gamma) */
    double ama;              /* reference Mach number (reference Ama) */
    double epsl_sw;          /* Steger-Warming smoothing epsilon (EPSL_SW) */
    double hx, hy, hz;       /* grid spacings, 1/(N_global-1) in the reference */
    int bc_planes;           /* 1: re-derive conserved vars on the 6 boundary planes after each stage (OCFD_bc) */
    int nthreads;            /* OpenMP threads, 0 = runtime default */
    int viscous;             /* 0: inviscid RHS only; 6 or 8: add the viscous terms with CD6 / CD8 */
    double re, pr, ref_t;    /* Re, Pr, Ref_T of the viscous terms */
    int periodic[3];         /* 1: ghost fill copies periodic images in that direction; 0: the
                                caller's ghost values stay (Iperiodic) */
    int d0_bound[6];         /* 1: one-sided closure of the inviscid scheme at that domain face
                                (x-, x+, y-, y+, z-, z+; the reference's D0_bound) */
    int non_ref[6];          /* 1: non-reflecting flag of the P (even index) / M (odd) pass of a
                                direction (the reference's Non_ref) */
} orc_params;

/* scalar kernels (for closed-form tests) */
double orc_weno7_symbo_p(const double *s8);
double orc_weno7_symbo_m(const double *s8);
void   orc_split_sw(double d, double u, double v, double w, double cc,
                    double A1, double A2, double A3,
                    double gamma, double epsl_sw, double fp[5], double fm[5]);

/* field kernels */
void orc_cons_to_pri(const orc_params *p, const double *f,
                     double *d, double *u, double *v, double *w, double *T, double *P);
void orc_pri_to_cons_box(const orc_params *p, double *f,
                         const double *d, const double *u, const double *v, const double *w, const double *T,
                         const int start[3], const int end[3]);
/* Periodic face ghosts of the directions with p->periodic[d] = 1 (others untouched). */
void orc_ghost_fill_periodic(const orc_params *p, double *fld);
void orc_sound_speed(const orc_params *p, const double *T, double *cc);

/* Inviscid RHS from ghosted primitives.  metrics[0..8] = Akx,Aky,Akz,Aix,Aiy,Aiz,Asx,Asy,Asz
 * (ghosted fields), ajac a ghosted field.  du (interior SoA) is ACCUMULATED into, as the
 * reference's decomposed kernels do; zero it first for a plain RHS. */
void orc_inviscid_rhs(const orc_params *p,
                      const double *d, const double *u, const double *v, const double *w, const double *cc,
                      const double *const metrics[9], const double *ajac, double *du);

/* Full RHS of one RK stage from the current primitives (du_comput):
 * zero du, periodic ghost fill of d,u,v,w,T, sound speed, inviscid RHS, then the
 * viscous terms when p->viscous is 6 or 8 (src/solver/OCFD_Stream.cpp:27-52). */
void orc_du_comput(const orc_params *p, double *d, double *u, double *v, double *w, double *T, double *cc,
                   const double *const metrics[9], const double *ajac, double *du);

/* SSP-RK3 stage update, src/solver/OCFD_time.cpp:15-65 (f_lap not produced). */
void orc_time_advance(const orc_params *p, int krk, double *f, const double *fn, const double *du, double dt);

/* OCFD_bc: conserved vars on the 6 boundary planes from primitives. */
void orc_bc_planes(const orc_params *p, double *f,
                   const double *d, const double *u, const double *v, const double *w, const double *T);

/* CFL time step: dt = cfl*min(hx,hy,hz)/max_cells max_dir(|u_dir| + c), c = sqrt(T)/Ma. */
double orc_cfl_dt(const orc_params *p, double cfl,
                  const double *u, const double *v, const double *w, const double *T);

/* N SSP-RK3 steps with Cartesian constant metrics (Akx=Aiy=Asz=Ajac=1, others 0).
 * State in/out: f (interior SoA) and the ghosted primitives d,u,v,w,T (the primitives must
 * correspond to f on entry, as after init or after a previous call).  dt > 0: fixed step;
 * dt <= 0: CFL step with cfl = -dt, recomputed at the start of every step. */
void orc_rk3_run(const orc_params *p, int nsteps, double dt,
                 double *f, double *d, double *u, double *v, double *w, double *T);

/* ---- viscous terms (du_vis_Jacobian3d, src/solver/OCFD_Stream.cpp:229-276) ---- */

/* Central first derivative of one stencil, src/kernels/OCFD_Schemes.cpp:99-105 (CD6,
 * s[0..6] centred on s[3]) and :130-138 (CD8, s[0..8] centred on s[4]); the caller
 * divides by h (put_d0_kernel, :80-97). */
double orc_cd6(const double *s7);
double orc_cd8(const double *s9);
/* Closure forms of the viscous derivatives at closed faces: s[0..2r] = f(c-r .. c+r) (r = 3 or
 * 4), c the index along the direction, n the extent, lo/hi the D0_bound flags of the two faces.
 * Returns 1 and the undivided derivative in *tmp when a form covers cell c. */
int orc_cd_closure(const double *s, int r, long c, long n, int lo, int hi, double *tmp);

/* Sutherland viscosity on the interior (get_Amu_kernal, src/kernels/commen_kernel.cpp:123-141):
 * amu = amu_C0 sqrt(T^3)/(Tsb + T), amu_C0 = (1 + 110.4/Ref_T)/Re, Tsb = 110.4/Ref_T
 * (opencfd_para_init, src/runtime/OCFD_init.cpp:379-384).  amu: interior array. */
void orc_amu(const orc_params *p, double re, double ref_t, const double *T, double *amu);

/* Viscous RHS accumulated into du components 1..4 (momentum x, y, z, energy), from ghosted
 * u, v, w, T whose periodic ghosts are filled: gradients by central differences of order
 * 6 or 8 (du_viscous_Jacobian3d_init, OCFD_NS_Jacobian3d.cpp:455-474), the stress and heat
 * flux per direction (vis_flux_ker, :477-620, with Cp/Pr = vis_flux_init_c), the periodic
 * exchange of those fluxes in the direction, their central derivative and du += Ajac * that
 * (du_viscous_Jacobian3d_[xyz]_final, :635-789; YF_Pe_XF, commen_kernel.cpp:159-170),
 * directions x, y, z in turn. */
void orc_viscous_rhs(const orc_params *p, double re, double pr, double ref_t, int order,
                     const double *u, const double *v, const double *w, const double *T,
                     const double *const metrics[9], const double *ajac, double *du);

/* Face value of the one-sided closures (D0_bound) on a line of n cells: s points at cell 0
 * (LAP ghost cells before it and after cell n-1); returns 1 and the P / M face value of
 * face f (between cells f-1 and f) when the closure covers that face, else 0. */
int orc_closure_p(const double *s, long n, long f, int lo, int hi, double *out);
int orc_closure_m(const double *s, long n, long f, int lo, int hi, double *out);

/* Internals of one P reconstruction (for first-principles tests): smoothness
 * indicators after the downwind max, the unnormalised product-form weights
 * d_k*prod_{j!=k}(eps+S_j) and the five candidate values q_k (before the
 * division by 12). */
void orc_weno7_symbo_p_parts(const double *s8, double S[5], double wts[5], double q[5]);

/* Metric terms Akx..Asz (out[0..8]) and Ajac (out[9]) from ghosted mesh coordinates
 * (comput_Jac3d, src/solver/OCFD_Comput_Jacobian3d.cpp:215-429); jacbound[d] = 1 selects
 * the one-sided derivative forms at the faces of direction d (the reference's Jacbound). */
void orc_comput_jacobian(const orc_params *p, const double *const *xyz, const int jacbound[3], double *const *out);

#ifdef __cplusplus
}
#endif
#endif
