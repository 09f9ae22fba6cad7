/* mhd2d_oracle.h -- C API of the CPU oracle (test infrastructure only). */
#ifndef MHD2D_ORACLE_H
#define MHD2D_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_BC_PERIODIC = 0, ORACLE_BC_SYMMETRIC = 1 };
enum { ORACLE_SCHEME_SINGLE = 0, ORACLE_SCHEME_MULTI_SHARE = 1 };

struct oracle_sim;

/* nx, ny include the 3 ghost layers per side (reference convention). */
struct oracle_sim *oracle_create(int nx, int ny, double dx, double dy, double gamma, double cfl,
                                 double eps, double eta, int bc_x, int bc_y);
void oracle_destroy(struct oracle_sim *s);
void oracle_set_state(struct oracle_sim *s, const double *aos);
void oracle_get_state(const struct oracle_sim *s, double *aos);
void oracle_get_stage1_face_b(const struct oracle_sim *s, double *bx, double *by);
void oracle_face_regions(struct oracle_sim *s, unsigned char *rx, unsigned char *ry);
double oracle_ct_contact_emf(const double *NF5, const double *NG4, const double *F5, const double *G4,
                             const double *NF0, const double *NG0);
int oracle_set_eta_field(struct oracle_sim *s, const double *eta);
int oracle_set_eta_staggered(struct oracle_sim *s, const double *xface, const double *yface, const double *cell);
void oracle_apply_boundary(struct oracle_sim *s);
void oracle_apply_boundary_axis(struct oracle_sim *s, int axis, int kind);
double oracle_calculate_dt(struct oracle_sim *s);
void oracle_step(struct oracle_sim *s, int nsteps);
double oracle_get_dt(const struct oracle_sim *s);
/* dt > 0: every later step uses dt instead of the CFL value (0: back to CFL). */
void oracle_set_fixed_dt(struct oracle_sim *s, double dt);
double oracle_get_time(const struct oracle_sim *s);
int oracle_check_crashed(const struct oracle_sim *s);
void oracle_get_stage1_fluxes(const struct oracle_sim *s, double *F, double *G);
void oracle_compute_fluxes(struct oracle_sim *s, double *aos_state, double *F, double *G);
double oracle_minmod(double x, double y, double eps);
void oracle_hlld_flux(const double *left8, const double *right8, double gamma, double eps, double *flux8);
double oracle_cell_dt(const double *cell8, double gamma, double dx, double dy, double eps);
int oracle_num_threads(void);
int oracle_set_scheme(struct oracle_sim *s, int scheme);
int oracle_check_crashed_multi(const struct oracle_sim *s);
void oracle_apply_halo(struct oracle_sim *s);

#ifdef __cplusplus
}
#endif
#endif
