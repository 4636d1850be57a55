/*
 * so_oracle.h — CPU f64 restatement of the split_operator crate's propagation
 * path. TEST INFRASTRUCTURE ONLY: this library is the checker that the HIP
 * product path is compared against. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the crate root). Complex arithmetic follows num-complex 0.4 formulas
 * (plain (ac-bd, ad+bc) products, division by Complex via norm_sqr), compiled
 * with -ffp-contract=off so rounding matches a non-FMA Rust build.
 */
#ifndef SO_ORACLE_H
#define SO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } orc_c;

/* threads the lane-parallel loops use (OpenMP) */
int orc_num_threads(void);

/* ---- Grid (src/grid.rs:13-89) ---------------------------------------- */
typedef struct orc_grid {
    char name[64];
    int dimension_no;
    int64_t nodes_no;
    double *nodes;
    double *weights;
} orc_grid;

orc_grid *orc_grid_linear_continuos(const char *name, double start, double end, int64_t n, int dim);
orc_grid *orc_grid_linear_countable(const char *name, double start, double end, int64_t n, int dim);
orc_grid *orc_grid_custom(const char *name, int64_t n, const double *nodes, const double *weights, int dim);
orc_grid *orc_grid_clone(const orc_grid *g);
void orc_grid_free(orc_grid *g);
const double *orc_grid_nodes(const orc_grid *g);
const double *orc_grid_weights(const orc_grid *g);
int64_t orc_grid_len(const orc_grid *g);
const char *orc_grid_name(const orc_grid *g);

/* ---- WaveFunction (src/wave_function.rs:10-154) ----------------------- */
typedef struct orc_wf orc_wf;
/* data: interleaved complex f64, C order; grids are cloned */
orc_wf *orc_wf_new(int ndim, const int64_t *shape, const double *data, orc_grid *const *grids);
orc_wf *orc_wf_clone(const orc_wf *wf);
void orc_wf_free(orc_wf *wf);
double *orc_wf_data(orc_wf *wf);
int64_t orc_wf_size(const orc_wf *wf);
int orc_wf_shape(const orc_wf *wf, int64_t *shape);
const orc_grid *orc_wf_grid(const orc_wf *wf, int axis);
double orc_wf_norm(orc_wf *wf);
void orc_wf_normalize(orc_wf *wf, double new_norm);
void orc_wf_dot(orc_wf *a, orc_wf *b, double *out_re_im);
void orc_wf_density(orc_wf *wf, double *out);
void orc_wf_state_density(orc_wf *wf, int axis, double *out);
void orc_gaussian_distribution(double x, double x0, double sigma, double momentum, double *out_re_im);

/* ---- TimeGrid (src/time_grid.rs:22-33) -------------------------------- */
void orc_select_step(int full, double step, int im_time, double *out_re_im);

/* ---- hamiltonian factory ---------------------------------------------- */
void orc_exp_propagator(const double *h, int64_t n, double dt_re, double dt_im, double *out);
void orc_exp_propagator_complex(const double *h, int64_t n, double dt_re, double dt_im, double *out);
void orc_momentum_grid(const orc_grid *g, double *momenta, double *weights);
void orc_kinetic_hamiltonian(const orc_grid *g, double red_mass, double *out);
void orc_legendre_polynomials(int64_t j, double x, double *out);
void orc_legendre_matrices(const orc_grid *polar, double *t, double *tinv);
void orc_rotational_hamiltonian(const orc_grid *radial, const orc_grid *polar, double red_mass,
                                double rot_const, double *out);
double orc_dispersion(double r, int n, double cn);
double orc_lennard_jones(double r, double d6, double r6);
double orc_harmonic(double r, double r0, double mass, double omega);
void orc_dumping_end(double mask_width, double mask_end, const orc_grid *g, double *out);
void orc_gauss_legendre(int64_t n, double *x, double *w);
double orc_double_factorial(int64_t n);
void orc_associated_legendre_polynomials(int64_t j, int m, double x, double *out);
double orc_legendre_normalization(int64_t l, int m);
void orc_associated_legendre_matrices(const orc_grid *polar, int omega, double *t, double *tinv);

/* ---- operations ------------------------------------------------------- */
typedef struct orc_op orc_op;
orc_op *orc_one_dim_propagator(int64_t n, int dim, const double *op);
orc_op *orc_n_dim_propagator(int ndim, const int64_t *shape, const double *op);
orc_op *orc_fft_transformation(const orc_grid *g, const char *transformed_name);
const orc_grid *orc_op_held_grid(const orc_op *op);
orc_op *orc_matrix_transformation(const orc_grid *g, const orc_grid *transformed,
                                  const double *t, const double *tinv);
orc_op *orc_non_diag_propagator(int dim, int64_t lanes, int64_t n, const double *ops);
orc_op *orc_state_matrix_transformation(int dep, const orc_grid *g, const orc_grid *transformed, int64_t count,
                                        const double *ts, const double *tinvs);
orc_op *orc_border_dumping(const double *mask, const orc_grid *g);
orc_op *orc_leak_control(void);
orc_op *orc_state_saver(const char *name, double step, int64_t step_no, const orc_grid *state_grid,
                        int64_t frames_no);
orc_op *orc_wave_function_saver(const char *name, double step, int64_t step_no, const orc_grid *x,
                                const orc_grid *y, int64_t frames_no);
/* frames_no == 0: checker without a LossSaver */
void orc_op_set_loss_checker(orc_op *op, const char *name, int64_t frames_no, double step,
                             int64_t step_no);
double orc_op_loss(orc_op *op);
int64_t orc_op_loss_frames(const orc_op *op, double *times, double *losses);
int64_t orc_saver_frames(const orc_op *op, double *data, double *times);
void orc_op_free(orc_op *op);

/* direct application (used by transformation unit tests) */
void orc_op_apply(orc_op *op, orc_wf *wf);
void orc_op_transform(orc_op *op, orc_wf *wf);
void orc_op_inverse_transform(orc_op *op, orc_wf *wf);

/* ---- Propagation (src/propagation.rs:11-292) --------------------------- */
typedef struct orc_prop orc_prop;
orc_prop *orc_prop_new(void);
void orc_prop_free(orc_prop *p); /* frees ops it owns */
void orc_prop_set_wave_function(orc_prop *p, const orc_wf *wf);
void orc_prop_set_time_grid(orc_prop *p, double step, int64_t step_no, int im_time);
void orc_prop_add_propagator(orc_prop *p, orc_op *op);
void orc_prop_add_transformation(orc_prop *p, orc_op *op, int inverse_first);
int orc_prop_add_saver(orc_prop *p, orc_op *op, int apply);
void orc_prop_add_control(orc_prop *p, orc_op *op, int apply);
void orc_prop_step(orc_prop *p);
void orc_prop_steps(orc_prop *p, int64_t n);
void orc_prop_propagate(orc_prop *p);
double orc_prop_mean_energy(orc_prop *p, int *err);
int64_t orc_prop_get_losses(orc_prop *p, double *out);
void orc_prop_reset_losses(orc_prop *p);
void orc_prop_reset_savers_state(orc_prop *p);
orc_wf *orc_prop_wave_function(orc_prop *p);

/* raw lane FFT for unit tests: unnormalised, sign -1 forward, +1 inverse */
void orc_fft_raw(const double *in, double *out, int64_t n, int sign);

#ifdef __cplusplus
}
#endif
#endif
