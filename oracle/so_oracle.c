/*
 * so_oracle.c — CPU f64 restatement of the split_operator crate (Rust) path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker). The product path never links this.
 * Each function names the reference file:line it follows. Lanes of the FFT,
 * matrix and one-dimensional operators run under OpenMP, as the crate runs
 * them under rayon; every reduction (norm, dot, state density) is a single
 * sequential fold in C order, as the crate's Zip::fold is.
 */
#include "so_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXDIM 8

static void orc_die(const char *msg) {
    fprintf(stderr, "so_oracle: %s\n", msg);
    abort();
}

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) orc_die("out of memory");
    return p;
}

/* num-complex 0.4 arithmetic, written out */
static inline orc_c c_mul(orc_c a, orc_c b) {
    orc_c r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline orc_c c_add(orc_c a, orc_c b) {
    orc_c r = {a.re + b.re, a.im + b.im};
    return r;
}
static inline orc_c c_sub(orc_c a, orc_c b) {
    orc_c r = {a.re - b.re, a.im - b.im};
    return r;
}
/* Complex / Complex: num-complex Div (norm_sqr denominator) */
static inline orc_c c_div(orc_c a, orc_c b) {
    double ns = b.re * b.re + b.im * b.im;
    double re = a.re * b.re + a.im * b.im;
    double im = a.im * b.re - a.re * b.im;
    orc_c r = {re / ns, im / ns};
    return r;
}
static inline orc_c c_div_real(orc_c a, double s) {
    orc_c r = {a.re / s, a.im / s};
    return r;
}
static inline orc_c c_scale(orc_c a, double s) {
    orc_c r = {a.re * s, a.im * s};
    return r;
}
static inline double c_norm_sqr(orc_c a) { return a.re * a.re + a.im * a.im; }

/* num-complex Complex::exp with its corner cases (+-inf, NaN real part) */
static orc_c c_exp(orc_c z) {
    double re = z.re, im = z.im;
    if (isinf(re)) {
        if (re < 0) {
            if (!isfinite(im)) {
                orc_c r = {0.0, 0.0};
                return r;
            }
        } else if (im == 0.0 || !isfinite(im)) {
            if (isinf(im)) im = NAN;
            orc_c r = {re, im};
            return r;
        }
    } else if (isnan(re) && im == 0.0) {
        return z;
    }
    double e = exp(re);
    orc_c r = {e * cos(im), e * sin(im)};
    return r;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ======================================================================
 * Grid — src/grid.rs:24-89
 * ====================================================================== */
static orc_grid *grid_alloc(const char *name, int64_t n, int dim) {
    orc_grid *g = (orc_grid *)xmalloc(sizeof(orc_grid));
    memset(g, 0, sizeof(*g));
    snprintf(g->name, sizeof(g->name), "%s", name);
    g->dimension_no = dim;
    g->nodes_no = n;
    g->nodes = (double *)xmalloc(sizeof(double) * (size_t)n);
    g->weights = (double *)xmalloc(sizeof(double) * (size_t)n);
    return g;
}

/* grid.rs:24-48 — trapezoid weights, half step at both ends */
orc_grid *orc_grid_linear_continuos(const char *name, double start, double end, int64_t n, int dim) {
    orc_grid *g = grid_alloc(name, n, dim);
    double step = (end - start) / ((double)n - 1.0);
    for (int64_t i = 0; i < n; i++) {
        g->nodes[i] = start + step * (double)i;
        g->weights[i] = 1.0 * step;
    }
    g->weights[0] = 0.5 * step;
    g->weights[n - 1] = 0.5 * step;
    return g;
}

/* grid.rs:51-73 — uniform weights */
orc_grid *orc_grid_linear_countable(const char *name, double start, double end, int64_t n, int dim) {
    orc_grid *g = grid_alloc(name, n, dim);
    double step = (end - start) / ((double)n - 1.0);
    for (int64_t i = 0; i < n; i++) {
        g->nodes[i] = start + step * (double)i;
        g->weights[i] = 1.0 * step;
    }
    return g;
}

/* grid.rs:76-84 */
orc_grid *orc_grid_custom(const char *name, int64_t n, const double *nodes, const double *weights, int dim) {
    orc_grid *g = grid_alloc(name, n, dim);
    memcpy(g->nodes, nodes, sizeof(double) * (size_t)n);
    memcpy(g->weights, weights, sizeof(double) * (size_t)n);
    return g;
}

orc_grid *orc_grid_clone(const orc_grid *s) {
    return orc_grid_custom(s->name, s->nodes_no, s->nodes, s->weights, s->dimension_no);
}

void orc_grid_free(orc_grid *g) {
    if (!g) return;
    free(g->nodes);
    free(g->weights);
    free(g);
}
const double *orc_grid_nodes(const orc_grid *g) { return g->nodes; }
const double *orc_grid_weights(const orc_grid *g) { return g->weights; }
int64_t orc_grid_len(const orc_grid *g) { return g->nodes_no; }
const char *orc_grid_name(const orc_grid *g) { return g->name; }

/* ======================================================================
 * WaveFunction + ChangeObserver — src/wave_function.rs:10-154,
 * src/change_observer.rs:6-47
 * ====================================================================== */
struct orc_wf {
    int ndim;
    int64_t shape[ORC_MAXDIM];
    int64_t size;
    orc_c *a;
    orc_grid *grids[ORC_MAXDIM];
    double *wamp; /* real part of weight_amplitude_array (imaginary part is 0) */
    /* ChangeObserver */
    char last_names[ORC_MAXDIM][64];
    double last_norm;
    int possible_norm_change;
};

/* wave_function.rs:50-65 — Π_axis sqrt(w_axis), multiplied in axis order */
static void wf_update_weight_amplitude(orc_wf *wf) {
    for (int64_t i = 0; i < wf->size; i++) wf->wamp[i] = 1.0;
    for (int ax = 0; ax < wf->ndim; ax++) {
        int64_t inner = 1;
        for (int b = ax + 1; b < wf->ndim; b++) inner *= wf->shape[b];
        int64_t n = wf->shape[ax];
        if (wf->grids[ax]->nodes_no != n) orc_die("grid weights do not match the lane length (shape mismatch in the crate)");
        const double *w = wf->grids[ax]->weights;
        for (int64_t i = 0; i < wf->size; i++) {
            int64_t k = (i / inner) % n;
            wf->wamp[i] = wf->wamp[i] * sqrt(w[k]);
        }
    }
}

static void obs_observe_grid(orc_wf *wf) {
    for (int ax = 0; ax < wf->ndim; ax++) snprintf(wf->last_names[ax], 64, "%s", wf->grids[ax]->name);
}
/* change_observer.rs:41-46 — names only */
static int obs_grid_changed(const orc_wf *wf) {
    for (int ax = 0; ax < wf->ndim; ax++)
        if (strcmp(wf->last_names[ax], wf->grids[ax]->name) != 0) return 1;
    return 0;
}
static void obs_observe_norm(orc_wf *wf, double n) {
    wf->last_norm = n;
    wf->possible_norm_change = 0;
}

/* wave_function.rs:23-47 + change_observer.rs:16-22 */
orc_wf *orc_wf_new(int ndim, const int64_t *shape, const double *data, orc_grid *const *grids) {
    if (ndim < 1 || ndim > ORC_MAXDIM) orc_die("bad ndim");
    orc_wf *wf = (orc_wf *)xmalloc(sizeof(orc_wf));
    memset(wf, 0, sizeof(*wf));
    wf->ndim = ndim;
    wf->size = 1;
    for (int i = 0; i < ndim; i++) {
        wf->shape[i] = shape[i];
        wf->size *= shape[i];
    }
    wf->a = (orc_c *)xmalloc(sizeof(orc_c) * (size_t)wf->size);
    memcpy(wf->a, data, sizeof(orc_c) * (size_t)wf->size);
    for (int i = 0; i < ndim; i++) {
        if (grids[i]->nodes_no != shape[i]) orc_die("grid length does not match the wave function shape");
        wf->grids[i] = orc_grid_clone(grids[i]);
    }
    wf->wamp = (double *)xmalloc(sizeof(double) * (size_t)wf->size);
    wf_update_weight_amplitude(wf);
    obs_observe_grid(wf);
    wf->last_norm = 1.0;
    wf->possible_norm_change = 1;
    return wf;
}

orc_wf *orc_wf_clone(const orc_wf *s) {
    orc_wf *wf = (orc_wf *)xmalloc(sizeof(orc_wf));
    memcpy(wf, s, sizeof(*wf));
    wf->a = (orc_c *)xmalloc(sizeof(orc_c) * (size_t)s->size);
    memcpy(wf->a, s->a, sizeof(orc_c) * (size_t)s->size);
    wf->wamp = (double *)xmalloc(sizeof(double) * (size_t)s->size);
    memcpy(wf->wamp, s->wamp, sizeof(double) * (size_t)s->size);
    for (int i = 0; i < s->ndim; i++) wf->grids[i] = orc_grid_clone(s->grids[i]);
    return wf;
}

void orc_wf_free(orc_wf *wf) {
    if (!wf) return;
    for (int i = 0; i < wf->ndim; i++) orc_grid_free(wf->grids[i]);
    free(wf->a);
    free(wf->wamp);
    free(wf);
}
double *orc_wf_data(orc_wf *wf) { return (double *)wf->a; }
int64_t orc_wf_size(const orc_wf *wf) { return wf->size; }
int orc_wf_shape(const orc_wf *wf, int64_t *shape) {
    for (int i = 0; i < wf->ndim; i++) shape[i] = wf->shape[i];
    return wf->ndim;
}
const orc_grid *orc_wf_grid(const orc_wf *wf, int axis) { return wf->grids[axis]; }

/* wave_function.rs:68-85 */
double orc_wf_norm(orc_wf *wf) {
    if (!wf->possible_norm_change) return wf->last_norm;
    if (obs_grid_changed(wf)) {
        wf_update_weight_amplitude(wf);
        obs_observe_grid(wf);
    }
    double acc = 0.0;
    for (int64_t i = 0; i < wf->size; i++) {
        double w = wf->wamp[i];
        acc = acc + c_norm_sqr(wf->a[i]) * (w * w + 0.0 * 0.0);
    }
    obs_observe_norm(wf, acc);
    return acc;
}

/* wave_function.rs:87-101 */
void orc_wf_dot(orc_wf *x, orc_wf *y, double *out) {
    double n1 = orc_wf_norm(x);
    double n2 = orc_wf_norm(y);
    if (x->size != y->size || memcmp(x->wamp, y->wamp, sizeof(double) * (size_t)x->size) != 0)
        orc_die("dot: weight arrays differ");
    orc_c acc = {0.0, 0.0};
    for (int64_t i = 0; i < x->size; i++) {
        orc_c yc = {y->a[i].re, -y->a[i].im};
        double w = x->wamp[i];
        acc = c_add(acc, c_scale(c_mul(x->a[i], yc), w * w + 0.0 * 0.0));
    }
    orc_c r = c_div_real(acc, sqrt(n1 * n2));
    out[0] = r.re;
    out[1] = r.im;
}

/* wave_function.rs:104-109 */
void orc_wf_normalize(orc_wf *wf, double new_norm) {
    double norm = orc_wf_norm(wf);
    orc_c s = {sqrt(new_norm / norm), 0.0};
    for (int64_t i = 0; i < wf->size; i++) wf->a[i] = c_mul(wf->a[i], s);
    obs_observe_norm(wf, new_norm);
}

/* wave_function.rs:112-118 */
void orc_wf_density(orc_wf *wf, double *out) {
    for (int64_t i = 0; i < wf->size; i++) out[i] = c_norm_sqr(wf->a[i]);
}

/* wave_function.rs:121-148 — unweighted for 1-D, weighted marginal for N-D */
void orc_wf_state_density(orc_wf *wf, int axis, double *out) {
    if (obs_grid_changed(wf)) {
        wf_update_weight_amplitude(wf);
        obs_observe_grid(wf);
    }
    if (wf->ndim == 1) {
        orc_wf_density(wf, out);
        return;
    }
    int64_t n = wf->shape[axis];
    int64_t inner = 1, outer = 1;
    for (int b = axis + 1; b < wf->ndim; b++) inner *= wf->shape[b];
    for (int b = 0; b < axis; b++) outer *= wf->shape[b];
    for (int64_t k = 0; k < n; k++) {
        double acc = 0.0;
        for (int64_t o = 0; o < outer; o++)
            for (int64_t i = 0; i < inner; i++) {
                int64_t idx = (o * n + k) * inner + i;
                double w = wf->wamp[idx];
                acc = acc + c_norm_sqr(wf->a[idx]) * (w * w + 0.0 * 0.0);
            }
        out[k] = acc;
    }
}

/* wave_function.rs:152-154 */
void orc_gaussian_distribution(double x, double x0, double sigma, double momentum, double *out) {
    double q = (x - x0) / (2.0 * sigma);
    /* -(q^2) - i*(x-x0)*k : Complex::i() * (x-x0) = (0*(x-x0), x-x0); times k */
    orc_c ix = {0.0 * (x - x0), 1.0 * (x - x0)};
    orc_c ixk = {ix.re * momentum, ix.im * momentum};
    orc_c z = {-(q * q) - ixk.re, 0.0 - ixk.im};
    orc_c r = c_exp(z);
    out[0] = r.re;
    out[1] = r.im;
}

/* ======================================================================
 * TimeGrid — src/time_grid.rs:22-33
 * ====================================================================== */
void orc_select_step(int full, double step, int im_time, double *out) {
    double ts = full ? step : step / 2.0;
    if (im_time) {
        /* -time_step * Complex::i() */
        out[0] = -ts * 0.0;
        out[1] = -ts * 1.0;
    } else {
        out[0] = ts;
        out[1] = 0.0;
    }
}

/* ======================================================================
 * Hamiltonian factory
 * ====================================================================== */
/* propagator_factory.rs:10-36 — exp(-i * h * dt) element-wise for real h */
void orc_exp_propagator(const double *h, int64_t n, double dt_re, double dt_im, double *out) {
    orc_c dt = {dt_re, dt_im};
    for (int64_t i = 0; i < n; i++) {
        double x = h[i];
        orc_c mi = {-0.0, -1.0};                /* -Complex::i() */
        orc_c a = {mi.re * x, mi.im * x};       /* * x (f64) */
        orc_c z = c_mul(a, dt);                 /* * dt */
        orc_c r = c_exp(z);
        out[2 * i] = r.re;
        out[2 * i + 1] = r.im;
    }
}

/* propagator_factory.rs:39-50 — complex hamiltonian */
void orc_exp_propagator_complex(const double *h, int64_t n, double dt_re, double dt_im, double *out) {
    orc_c dt = {dt_re, dt_im};
    for (int64_t i = 0; i < n; i++) {
        orc_c x = {h[2 * i], h[2 * i + 1]};
        orc_c mi = {-0.0, -1.0};
        orc_c z = c_mul(c_mul(mi, x), dt);
        orc_c r = c_exp(z);
        out[2 * i] = r.re;
        out[2 * i + 1] = r.im;
    }
}

/* fft_transformation.rs:28-37 — momentum nodes 0..N/2, -N/2..-1, weights dk
 * halved at N/2-1 and N/2 */
void orc_momentum_grid(const orc_grid *g, double *momenta, double *weights) {
    int64_t n = g->nodes_no;
    double dk = 2.0 * M_PI / (g->nodes[n - 1] - g->nodes[0]) * (1. - 1. / (double)n);
    int64_t len = n, c = 0;
    for (int64_t x = 0; x < len / 2; x++) momenta[c++] = (double)x * dk;
    for (int64_t x = -len / 2; x < 0; x++) momenta[c++] = (double)x * dk;
    /* odd N: the chain yields N-1 momenta (kinetic_operator.rs:12-15); the
     * caller gets a grid one shorter than the lane */
    for (int64_t i = 0; i < c; i++) weights[i] = dk;
    weights[len / 2 - 1] *= 0.5;
    weights[len / 2] *= 0.5;
}

/* kinetic_operator.rs:8-21 */
void orc_kinetic_hamiltonian(const orc_grid *g, double red_mass, double *out) {
    int64_t n = g->nodes_no;
    double dk = 2.0 * M_PI / (g->nodes[n - 1] - g->nodes[0]) * (1. - 1. / (double)n);
    int64_t c = 0;
    for (int64_t x = 0; x < n / 2; x++) {
        double k = (double)x * dk;
        out[c++] = k * k / (2.0 * red_mass);
    }
    for (int64_t x = -n / 2; x < 0; x++) {
        double k = (double)x * dk;
        out[c++] = k * k / (2.0 * red_mass);
    }
}

/* special_functions.rs:3-14 — Bonnet recursion (j >= 1; j == 0 panics in the crate) */
void orc_legendre_polynomials(int64_t j, double x, double *p) {
    if (j < 1) orc_die("legendre_polynomials: j must be >= 1 (index out of bounds in the crate)");
    p[0] = 1.0;
    p[1] = x;
    for (int64_t i = 2; i <= j; i++)
        p[i] = ((double)(2 * i - 1) / (double)i) * x * p[i - 1] - ((double)(i - 1) / (double)i) * p[i - 2];
}

/* legendre_diagonalization.rs:13-45 — T[i,j] = sqrt(l_i+1/2) P_{l_i}(cos th_j) w_j,
 * Tinv[j,i] = sqrt(l_i+1/2) P_{l_i}(cos th_j); both complex interleaved, row-major */
void orc_legendre_matrices(const orc_grid *polar, double *t, double *tinv) {
    int64_t n = polar->nodes_no;
    int64_t lmax = n - 1;
    double *pl = (double *)xmalloc(sizeof(double) * (size_t)(lmax + 1));
    for (int64_t j = 0; j < n; j++) {
        orc_legendre_polynomials(lmax, cos(polar->nodes[j]), pl);
        for (int64_t i = 0; i < n; i++) {
            double v = sqrt((double)i + 0.5) * pl[i];
            tinv[2 * (j * n + i)] = v;
            tinv[2 * (j * n + i) + 1] = 0.0;
            /* transformation[[i,j]] *= weights[j] : Complex * f64 */
            t[2 * (i * n + j)] = v * polar->weights[j];
            t[2 * (i * n + j) + 1] = 0.0 * polar->weights[j];
        }
    }
    free(pl);
}

/* special_functions.rs:43-55 — n!! as the crate forms it (n, n-2, ... down to 3 or 2) */
double orc_double_factorial(int64_t n) {
    if (n == 0) return 1.0;
    double v = (double)n;
    for (int64_t k = n - 2; k >= 2; k -= 2) v *= (double)k;
    return v;
}

/* special_functions.rs:57-80 */
static double negate_m(int64_t l, int m) {
    double v = (m % 2 == 0) ? 1.0 : -1.0;
    if (m > 0) {
        for (int64_t k = l - m; k < l + m; k++) v /= (double)k + 1.0;
    } else if (m < 0) {
        for (int64_t k = l + m; k < l - m; k++) v *= (double)k + 1.0;
    } else {
        v = 1.0;
    }
    return v;
}

/* special_functions.rs:17-41 — P_l^m(x), l = 0..j (entries below |m| are 0) */
void orc_associated_legendre_polynomials(int64_t j, int m, double x, double *p) {
    int64_t mu = m < 0 ? -m : m;
    if (m == 0) {
        orc_legendre_polynomials(j, x, p);
        return;
    }
    if (mu + 1 > j) orc_die("associated_legendre_polynomials: index out of bounds (the crate panics)");
    for (int64_t i = 0; i <= j; i++) p[i] = 0.0;
    p[mu] = ((m % 2 == 0) ? 1.0 : -1.0) * orc_double_factorial(2 * mu - 1) * pow(1.0 - x * x, (double)mu / 2.0);
    if (m < 0) {
        p[mu] *= negate_m(mu, (int)mu);
        p[mu + 1] = x * p[mu];
    } else {
        p[mu + 1] = x * (2.0 * (double)mu + 1.0) * p[mu];
    }
    for (int64_t i = mu + 2; i <= j; i++) {
        int64_t l = i - 1;
        p[i] = ((double)(2 * l + 1) * x * p[i - 1] - ((double)l + (double)m) * p[i - 2]) / ((double)l - (double)m + 1.0);
    }
}

/* special_functions.rs:82-116 — sqrt((l + 1/2) (l - m)! / (l + m)!) */
double orc_legendre_normalization(int64_t l, int m) {
    double v = 1.0;
    if (m > 0) {
        for (int64_t k = l - m; k < l + m; k++) v /= (double)k + 1.0;
    } else if (m < 0) {
        for (int64_t k = l + m; k < l - m; k++) v *= (double)k + 1.0;
    }
    return sqrt(v * ((double)l + 0.5));
}

/* legendre_diagonalization.rs:48-80 — transformation[i][j] = N_l P_l^omega(cos theta_j) w_j for i >= |omega|,
 * inverse = the unweighted matrix transposed */
void orc_associated_legendre_matrices(const orc_grid *polar, int omega, double *t, double *tinv) {
    int64_t n = polar->nodes_no, lmax = n - 1, om = omega < 0 ? -omega : omega;
    double *pl = (double *)xmalloc(sizeof(double) * (size_t)(lmax + 2));
    memset(t, 0, sizeof(double) * (size_t)(2 * n * n));
    memset(tinv, 0, sizeof(double) * (size_t)(2 * n * n));
    for (int64_t j = 0; j < n; j++) {
        orc_associated_legendre_polynomials(lmax, omega, cos(polar->nodes[j]), pl);
        for (int64_t i = om; i < n; i++) {
            double v = orc_legendre_normalization(i, omega) * pl[i];
            tinv[2 * (j * n + i)] = v;
            t[2 * (i * n + j)] = v * polar->weights[j];
        }
    }
    free(pl);
}

/* rotational_operator.rs:7-26 — (radial x polar) l(l+1)(B + 1/(2 mu r^2)) */
void orc_rotational_hamiltonian(const orc_grid *radial, const orc_grid *polar, double red_mass,
                                double rot_const, double *out) {
    int64_t nr = radial->nodes_no, np = polar->nodes_no;
    for (int64_t i = 0; i < nr; i++) {
        double r = radial->nodes[i];
        for (int64_t j = 0; j < np; j++)
            out[i * np + j] = ((double)(j * (j + 1))) * (rot_const + 1.0 / (2.0 * red_mass * r * r));
    }
}

/* analytic_potentials.rs:2-10 */
double orc_dispersion(double r, int n, double cn) {
    double v = cn * __builtin_powi(r, n);
    return isnan(v) ? INFINITY : v;
}
/* analytic_potentials.rs:13-22 */
double orc_lennard_jones(double r, double d6, double r6) {
    double rr = __builtin_powi(r6 / r, 6);
    double v = d6 * rr * (rr - 2.0);
    return isnan(v) ? INFINITY : v;
}
/* analytic_potentials.rs:24-32 */
double orc_harmonic(double r, double r0, double mass, double omega) {
    double v = 0.5 * mass * __builtin_powi(omega, 2) * __builtin_powi(r - r0, 2);
    return isnan(v) ? INFINITY : v;
}

/* border_dumping.rs:14-33 — mask as complex interleaved */
void orc_dumping_end(double mask_width, double mask_end, const orc_grid *g, double *out) {
    double rmax = g->nodes[g->nodes_no - 1];
    for (int64_t i = 0; i < g->nodes_no; i++) {
        double x = g->nodes[i], v;
        if (x < rmax - mask_width)
            v = 1.0;
        else if (x > rmax - mask_end)
            v = 0.0;
        else
            v = sin(M_PI / 2.0 * (rmax - x) / mask_width);
        out[2 * i] = v;
        out[2 * i + 1] = 0.0;
    }
}

/* Gauss-Legendre nodes (ascending) and weights: Newton iteration on P_n.
 * Not in the crate (its callers supply the polar grid); pinned against
 * numpy.polynomial.legendre.leggauss and exact quadrature of polynomials. */
void orc_gauss_legendre(int64_t n, double *x, double *w) {
    for (int64_t i = 0; i < n; i++) {
        /* root i counted from x = +1 downwards */
        double z = cos(M_PI * ((double)i + 0.75) / ((double)n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int64_t k = 2; k <= n; k++) {
                double p2 = ((double)(2 * k - 1) * z * p1 - (double)(k - 1) * p0) / (double)k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) {
                p1 = z;
                p0 = 1.0;
            }
            dp = (double)n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        {
            double p0 = 1.0, p1 = z;
            for (int64_t k = 2; k <= n; k++) {
                double p2 = ((double)(2 * k - 1) * z * p1 - (double)(k - 1) * p0) / (double)k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) p0 = 1.0;
            dp = (double)n * (z * p1 - p0) / (z * z - 1.0);
        }
        x[n - 1 - i] = z;
        w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

/* ======================================================================
 * FFT — the arithmetic of rustfft 6.2 (fft_transformation.rs:25-26,62,81):
 * an unnormalised DFT, X_k = sum_n x_n exp(-+2 pi i k n / N). Restated as a
 * recursive mixed-radix decimation in time with twiddles reduced to the first
 * octant, so quarter-turn twiddles are exact (rustfft's small butterflies are
 * exact there too; fft_tests.rs:49-77 asserts exact results at N = 4).
 * ====================================================================== */
static void twiddle_exact(int64_t x, int64_t n, double *c, double *s) {
    /* cos and sin of 2 pi x / n for 0 <= x < n, by symmetry */
    int neg_s = 0;
    if (2 * x > n) { /* theta -> 2pi - theta */
        x = n - x;
        neg_s = 1;
    }
    int neg_c = 0;
    if (4 * x > n) { /* theta -> pi - theta: (x', n') = (n - 2x, 2n) */
        int64_t xx = n - 2 * x;
        n = 2 * n;
        x = xx;
        neg_c = 1;
    }
    double cc, ss;
    if (8 * x > n) { /* theta -> pi/2 - theta: (n - 4x, 4n) */
        double th = 2.0 * M_PI * (double)(n - 4 * x) / (double)(4 * n);
        cc = sin(th);
        ss = cos(th);
    } else {
        double th = 2.0 * M_PI * (double)x / (double)n;
        cc = cos(th);
        ss = sin(th);
    }
    *c = neg_c ? -cc : cc;
    *s = neg_s ? -ss : ss;
}

typedef struct {
    int64_t n;
    orc_c *tw_fwd; /* exp(-2 pi i e / n) */
    orc_c *tw_inv;
} fft_plan;

static void plan_init(fft_plan *p, int64_t n) {
    p->n = n;
    p->tw_fwd = (orc_c *)xmalloc(sizeof(orc_c) * (size_t)n);
    p->tw_inv = (orc_c *)xmalloc(sizeof(orc_c) * (size_t)n);
    for (int64_t e = 0; e < n; e++) {
        double c, s;
        twiddle_exact(e, n, &c, &s);
        p->tw_fwd[e].re = c;
        p->tw_fwd[e].im = -s;
        p->tw_inv[e].re = c;
        p->tw_inv[e].im = s;
    }
}
static void plan_free(fft_plan *p) {
    free(p->tw_fwd);
    free(p->tw_inv);
}

static int64_t smallest_factor(int64_t n) {
    if (n % 2 == 0) return 2;
    for (int64_t f = 3; f * f <= n; f += 2)
        if (n % f == 0) return f;
    return n;
}

/* out[k] = sum_j in[j*stride] W^(j k), W = tw[twstride], length n */
static void fft_rec(const orc_c *in, int64_t stride, orc_c *out, int64_t n, const orc_c *tw, int64_t twstride) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int64_t p = smallest_factor(n);
    int64_t m = n / p;
    for (int64_t r = 0; r < p; r++) fft_rec(in + r * stride, stride * p, out + r * m, m, tw, twstride * p);
    if (p == 2) {
        for (int64_t k = 0; k < m; k++) {
            orc_c a0 = out[k];
            orc_c t = c_mul(tw[k * twstride], out[m + k]);
            out[k] = c_add(a0, t);
            out[k + m] = c_sub(a0, t);
        }
        return;
    }
    orc_c buf[64];
    orc_c *a = p <= 64 ? buf : (orc_c *)xmalloc(sizeof(orc_c) * (size_t)p);
    for (int64_t k = 0; k < m; k++) {
        for (int64_t r = 0; r < p; r++) a[r] = out[r * m + k];
        for (int64_t q = 0; q < p; q++) {
            int64_t kk = k + q * m;
            orc_c acc = a[0];
            for (int64_t r = 1; r < p; r++) {
                int64_t e = (r * kk) % n;
                acc = c_add(acc, c_mul(tw[e * twstride], a[r]));
            }
            out[kk] = acc;
        }
    }
    if (a != buf) free(a);
}

void orc_fft_raw(const double *in, double *out, int64_t n, int sign) {
    fft_plan p;
    plan_init(&p, n);
    fft_rec((const orc_c *)in, 1, (orc_c *)out, n, sign < 0 ? p.tw_fwd : p.tw_inv, 1);
    plan_free(&p);
}

/* ======================================================================
 * Operations
 * ====================================================================== */
enum {
    K_ONE_DIM = 1,
    K_N_DIM,
    K_FFT,
    K_MATRIX,
    K_NON_DIAG,
    K_BORDER,
    K_LEAK,
    K_STATE_SAVER,
    K_WF_SAVER,
    K_STATE_MATRIX
};

typedef struct {
    int present;
    char name[128];
    double loss;
    double current_norm;
    /* LossSaver (loss_saver.rs:5-57) */
    int has_saver;
    int64_t current_frame, frames_no, total_frames;
    double step;
    double *losses, *times;
    int64_t nrec;
} loss_checker;

struct orc_op {
    int kind;
    int dim;
    int64_t n;
    /* diagonal operators */
    orc_c *op;
    int nd_ndim;
    int64_t nd_shape[ORC_MAXDIM];
    int64_t op_size;
    /* transformations */
    orc_grid *held; /* grid_transformation */
    fft_plan plan;
    orc_c *t, *tinv;
    /* non-diagonal; state matrix: number of matrices and dependent axis */
    int64_t lanes;
    int dep;
    /* controls */
    double leak_norm;
    loss_checker chk;
    /* savers */
    char sname[256];
    double step;
    int64_t step_no;
    orc_grid *sgrid, *ygrid;
    int64_t frames_no, current_frame;
    double *data, *times;
    int64_t ntimes;
};

static orc_op *op_alloc(int kind) {
    orc_op *o = (orc_op *)xmalloc(sizeof(orc_op));
    memset(o, 0, sizeof(*o));
    o->kind = kind;
    return o;
}

/* LossChecker — loss_checker.rs:6-69, LossSaver — loss_saver.rs:29-39 */
static void chk_before(loss_checker *c, orc_wf *wf) { c->current_norm = orc_wf_norm(wf); }
static void chk_after(loss_checker *c, orc_wf *wf) {
    double nn = orc_wf_norm(wf);
    c->loss += c->current_norm - nn;
    if (c->has_saver) {
        int64_t freq = 2 * c->total_frames / c->frames_no;
        if (freq == 0) orc_die("LossSaver: frequency is zero (division by zero in the crate)");
        if (c->current_frame % freq == 0 && c->current_frame / freq < c->frames_no) {
            c->losses[c->nrec] = c->loss;
            c->times[c->nrec] = ((double)c->current_frame + 1.0) / 2. * c->step;
            c->nrec++;
        }
        c->current_frame++;
    }
    c->current_norm = nn;
}

void orc_op_set_loss_checker(orc_op *op, const char *name, int64_t frames_no, double step, int64_t step_no) {
    loss_checker *c = &op->chk;
    free(c->losses);
    free(c->times);
    memset(c, 0, sizeof(*c));
    c->present = 1;
    snprintf(c->name, sizeof(c->name), "%s", name);
    c->loss = 0.0;
    c->current_norm = 1.0;
    if (frames_no > 0) {
        c->has_saver = 1;
        c->frames_no = frames_no;
        c->total_frames = step_no;
        c->step = step;
        c->losses = (double *)xmalloc(sizeof(double) * (size_t)frames_no);
        c->times = (double *)xmalloc(sizeof(double) * (size_t)frames_no);
    }
}
double orc_op_loss(orc_op *op) { return op->chk.loss; }
int64_t orc_op_loss_frames(const orc_op *op, double *times, double *losses) {
    const loss_checker *c = &op->chk;
    for (int64_t i = 0; i < c->nrec; i++) {
        if (times) times[i] = c->times[i];
        if (losses) losses[i] = c->losses[i];
    }
    return c->nrec;
}

/* one_dim_propagator.rs:17-29 */
orc_op *orc_one_dim_propagator(int64_t n, int dim, const double *op) {
    orc_op *o = op_alloc(K_ONE_DIM);
    o->dim = dim;
    o->n = n;
    o->op = (orc_c *)xmalloc(sizeof(orc_c) * (size_t)n);
    memcpy(o->op, op, sizeof(orc_c) * (size_t)n);
    return o;
}

/* n_dim_propagator.rs:15-24 (op == NULL keeps the zeros([1]) default of :17) */
orc_op *orc_n_dim_propagator(int ndim, const int64_t *shape, const double *op) {
    orc_op *o = op_alloc(K_N_DIM);
    if (!op) {
        o->nd_ndim = 1;
        o->nd_shape[0] = 1;
        o->op_size = 1;
        o->op = (orc_c *)xmalloc(sizeof(orc_c));
        o->op[0].re = 0.0;
        o->op[0].im = 0.0;
        return o;
    }
    o->nd_ndim = ndim;
    o->op_size = 1;
    for (int i = 0; i < ndim; i++) {
        o->nd_shape[i] = shape[i];
        o->op_size *= shape[i];
    }
    o->op = (orc_c *)xmalloc(sizeof(orc_c) * (size_t)o->op_size);
    memcpy(o->op, op, sizeof(orc_c) * (size_t)o->op_size);
    return o;
}

/* fft_transformation.rs:24-48 */
orc_op *orc_fft_transformation(const orc_grid *g, const char *transformed_name) {
    orc_op *o = op_alloc(K_FFT);
    o->dim = g->dimension_no;
    o->n = g->nodes_no;
    int64_t n = g->nodes_no;
    double *mom = (double *)xmalloc(sizeof(double) * (size_t)n);
    double *w = (double *)xmalloc(sizeof(double) * (size_t)n);
    orc_momentum_grid(g, mom, w);
    int64_t len = 2 * (n / 2);
    o->held = orc_grid_custom(transformed_name, len, mom, w, g->dimension_no);
    free(mom);
    free(w);
    plan_init(&o->plan, n);
    return o;
}
const orc_grid *orc_op_held_grid(const orc_op *op) { return op->held; }

/* matrix_transformation.rs:21-51 */
orc_op *orc_matrix_transformation(const orc_grid *g, const orc_grid *transformed, const double *t,
                                  const double *tinv) {
    orc_op *o = op_alloc(K_MATRIX);
    o->dim = g->dimension_no;
    o->n = g->nodes_no;
    o->held = orc_grid_clone(transformed);
    size_t nn = (size_t)(o->n * o->n);
    o->t = (orc_c *)xmalloc(sizeof(orc_c) * nn);
    o->tinv = (orc_c *)xmalloc(sizeof(orc_c) * nn);
    memcpy(o->t, t, sizeof(orc_c) * nn);
    memcpy(o->tinv, tinv, sizeof(orc_c) * nn);
    return o;
}

/* state_matrix_transformation.rs:19-44 — count matrices n x n for each direction */
orc_op *orc_state_matrix_transformation(int dep, const orc_grid *g, const orc_grid *transformed, int64_t count,
                                        const double *ts, const double *tinvs) {
    if (dep <= g->dimension_no) orc_die("Dependent dimension should have higher dimension index");
    orc_op *o = op_alloc(K_STATE_MATRIX);
    o->dim = g->dimension_no;
    o->dep = dep;
    o->n = g->nodes_no;
    o->lanes = count;
    o->held = orc_grid_clone(transformed);
    size_t sz = (size_t)(count * o->n * o->n);
    o->t = (orc_c *)xmalloc(sizeof(orc_c) * (sz ? sz : 1));
    o->tinv = (orc_c *)xmalloc(sizeof(orc_c) * (sz ? sz : 1));
    if (sz) {
        memcpy(o->t, ts, sizeof(orc_c) * sz);
        memcpy(o->tinv, tinvs, sizeof(orc_c) * sz);
    }
    return o;
}

/* non_diagonal_propagator.rs:16-27 — ops: lanes x n x n complex */
orc_op *orc_non_diag_propagator(int dim, int64_t lanes, int64_t n, const double *ops) {
    orc_op *o = op_alloc(K_NON_DIAG);
    o->dim = dim;
    o->n = n;
    o->lanes = lanes;
    size_t sz = (size_t)(lanes * n * n);
    o->t = (orc_c *)xmalloc(sizeof(orc_c) * sz);
    memcpy(o->t, ops, sizeof(orc_c) * sz);
    return o;
}

/* border_dumping.rs:42-50 */
orc_op *orc_border_dumping(const double *mask, const orc_grid *g) {
    orc_op *o = orc_one_dim_propagator(g->nodes_no, g->dimension_no, mask);
    o->kind = K_BORDER;
    return o;
}

/* leak_control.rs:14-19 */
orc_op *orc_leak_control(void) {
    orc_op *o = op_alloc(K_LEAK);
    o->leak_norm = 0.0;
    return o;
}

/* wave_function_saver.rs:118-133 */
orc_op *orc_state_saver(const char *name, double step, int64_t step_no, const orc_grid *state_grid,
                        int64_t frames_no) {
    orc_op *o = op_alloc(K_STATE_SAVER);
    snprintf(o->sname, sizeof(o->sname), "%s", name);
    o->step = step;
    o->step_no = step_no;
    o->sgrid = orc_grid_clone(state_grid);
    o->frames_no = frames_no;
    o->data = (double *)calloc((size_t)(state_grid->nodes_no * frames_no), sizeof(double));
    o->times = (double *)xmalloc(sizeof(double) * (size_t)(frames_no + 1));
    return o;
}

/* wave_function_saver.rs:21-38 */
orc_op *orc_wave_function_saver(const char *name, double step, int64_t step_no, const orc_grid *x,
                                const orc_grid *y, int64_t frames_no) {
    orc_op *o = op_alloc(K_WF_SAVER);
    snprintf(o->sname, sizeof(o->sname), "%s", name);
    o->step = step;
    o->step_no = step_no;
    o->sgrid = orc_grid_clone(x);
    o->ygrid = orc_grid_clone(y);
    o->frames_no = frames_no;
    o->data = (double *)calloc((size_t)(x->nodes_no * y->nodes_no * frames_no), sizeof(double));
    o->times = (double *)xmalloc(sizeof(double) * (size_t)(frames_no + 1));
    return o;
}

int64_t orc_saver_frames(const orc_op *op, double *data, double *times) {
    int64_t per = op->sgrid->nodes_no * (op->ygrid ? op->ygrid->nodes_no : 1);
    if (data) memcpy(data, op->data, sizeof(double) * (size_t)(per * op->frames_no));
    if (times) memcpy(times, op->times, sizeof(double) * (size_t)op->ntimes);
    return op->ntimes;
}

void orc_op_free(orc_op *o) {
    if (!o) return;
    free(o->op);
    orc_grid_free(o->held);
    if (o->kind == K_FFT) plan_free(&o->plan);
    free(o->t);
    free(o->tinv);
    free(o->chk.losses);
    free(o->chk.times);
    orc_grid_free(o->sgrid);
    orc_grid_free(o->ygrid);
    free(o->data);
    free(o->times);
    free(o);
}

/* lane geometry along axis d */
static void lane_geom(const orc_wf *wf, int d, int64_t *outer, int64_t *inner) {
    int64_t o = 1, i = 1;
    for (int b = 0; b < d; b++) o *= wf->shape[b];
    for (int b = d + 1; b < wf->ndim; b++) i *= wf->shape[b];
    *outer = o;
    *inner = i;
}

/* one_dim_propagator.rs:37-46 — lane *= op */
static void apply_one_dim(orc_op *o, orc_wf *wf) {
    wf->possible_norm_change = 1;
    int d = o->dim;
    if (d >= wf->ndim || wf->shape[d] != o->n) orc_die("OneDimPropagator: operator length does not match the lane");
    int64_t outer, inner, n = o->n;
    lane_geom(wf, d, &outer, &inner);
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t a = 0; a < outer; a++)
        for (int64_t i = 0; i < inner; i++)
            for (int64_t k = 0; k < n; k++) {
                int64_t idx = (a * n + k) * inner + i;
                wf->a[idx] = c_mul(wf->a[idx], o->op[k]);
            }
}

/* n_dim_propagator.rs:32-36 — array *= operator (same shape, or a single element broadcast) */
static void apply_n_dim(orc_op *o, orc_wf *wf) {
    wf->possible_norm_change = 1;
    if (o->op_size == 1) {
        for (int64_t i = 0; i < wf->size; i++) wf->a[i] = c_mul(wf->a[i], o->op[0]);
        return;
    }
    if (o->op_size != wf->size) orc_die("NDimPropagator: operator shape does not broadcast to the wave function");
    for (int64_t i = 0; i < wf->size; i++) wf->a[i] = c_mul(wf->a[i], o->op[i]);
}

/* fft_transformation.rs:53-88 — per lane: copy, FFT, write back / sqrt(N) */
static void fft_lanes(orc_op *o, orc_wf *wf, int inverse) {
    int d = o->dim;
    orc_grid *tmp = wf->grids[d];
    wf->grids[d] = o->held;
    o->held = tmp;
    wf->possible_norm_change = 1;
    int64_t outer, inner, n = o->n;
    if (wf->shape[d] != n) orc_die("FFTTransformation: lane length mismatch");
    lane_geom(wf, d, &outer, &inner);
    double sq = sqrt((double)n);
    orc_c sqc = {sq, 0.0};
    const orc_c *tw = inverse ? o->plan.tw_inv : o->plan.tw_fwd;
#pragma omp parallel
    {
        orc_c *lane = (orc_c *)xmalloc(sizeof(orc_c) * (size_t)n);
        orc_c *res = (orc_c *)xmalloc(sizeof(orc_c) * (size_t)n);
#pragma omp for collapse(2) schedule(static)
        for (int64_t a = 0; a < outer; a++)
            for (int64_t i = 0; i < inner; i++) {
                for (int64_t k = 0; k < n; k++) lane[k] = wf->a[(a * n + k) * inner + i];
                fft_rec(lane, 1, res, n, tw, 1);
                for (int64_t k = 0; k < n; k++)
                    wf->a[(a * n + k) * inner + i] = inverse ? c_div(res[k], sqc) : c_div_real(res[k], sq);
            }
        free(lane);
        free(res);
    }
}

/* matrix_transformation.rs:58-75 — lane = M . lane */
static void matrix_lanes(orc_op *o, orc_wf *wf, const orc_c *m) {
    int d = o->dim;
    orc_grid *tmp = wf->grids[d];
    wf->grids[d] = o->held;
    o->held = tmp;
    wf->possible_norm_change = 1;
    int64_t outer, inner, n = o->n;
    if (wf->shape[d] != n) orc_die("MatrixTransformation: lane length mismatch");
    lane_geom(wf, d, &outer, &inner);
#pragma omp parallel
    {
        orc_c *lane = (orc_c *)xmalloc(sizeof(orc_c) * (size_t)n);
#pragma omp for collapse(2) schedule(static)
        for (int64_t a = 0; a < outer; a++)
            for (int64_t i = 0; i < inner; i++) {
                for (int64_t k = 0; k < n; k++) lane[k] = wf->a[(a * n + k) * inner + i];
                for (int64_t r = 0; r < n; r++) {
                    orc_c acc = {0.0, 0.0};
                    const orc_c *row = m + r * n;
                    for (int64_t k = 0; k < n; k++) acc = c_add(acc, c_mul(row[k], lane[k]));
                    wf->a[(a * n + r) * inner + i] = acc;
                }
            }
        free(lane);
    }
}

/* non_diagonal_propagator.rs:29-37 — operator k applied to lane k (lanes in C order) */
static void apply_non_diag(orc_op *o, orc_wf *wf) {
    wf->possible_norm_change = 1;
    int d = o->dim;
    int64_t outer, inner, n = o->n;
    if (wf->shape[d] != n) orc_die("NonDiagPropagator: lane length mismatch");
    lane_geom(wf, d, &outer, &inner);
    int64_t nl = outer * inner < o->lanes ? outer * inner : o->lanes;
#pragma omp parallel
    {
        orc_c *lane = (orc_c *)xmalloc(sizeof(orc_c) * (size_t)n);
#pragma omp for schedule(static)
        for (int64_t l = 0; l < nl; l++) {
            int64_t a = l / inner, i = l % inner;
            for (int64_t k = 0; k < n; k++) lane[k] = wf->a[(a * n + k) * inner + i];
            const orc_c *m = o->t + l * n * n;
            for (int64_t r = 0; r < n; r++) {
                orc_c acc = {0.0, 0.0};
                for (int64_t k = 0; k < n; k++) acc = c_add(acc, c_mul(m[r * n + k], lane[k]));
                wf->a[(a * n + r) * inner + i] = acc;
            }
        }
        free(lane);
    }
}

/* Propagator::apply with the optional LossChecker around it
 * (one_dim_propagator.rs:55-65, n_dim_propagator.rs:44-54) */
void orc_op_apply(orc_op *o, orc_wf *wf) {
    if (o->chk.present) chk_before(&o->chk, wf);
    switch (o->kind) {
    case K_ONE_DIM: apply_one_dim(o, wf); break;
    case K_N_DIM: apply_n_dim(o, wf); break;
    case K_NON_DIAG: apply_non_diag(o, wf); break;
    default: orc_die("apply on a non-propagator");
    }
    if (o->chk.present) chk_after(&o->chk, wf);
}

/* state_matrix_transformation.rs:52-74 — swap the grid; slice j of the dependent axis has every lane along the
 * axis replaced by mats[j] . lane (slices past the matrices, or matrices past the slices, are skipped by zip) */
static void state_matrix_lanes(orc_op *o, orc_wf *wf, const orc_c *mats) {
    int d = o->dim;
    orc_grid *tmp = wf->grids[d];
    wf->grids[d] = o->held;
    o->held = tmp;
    wf->possible_norm_change = 1;
    int64_t outer, inner, n = o->n;
    if (wf->shape[d] != n) orc_die("StateMatrixTransformation: lane length mismatch");
    lane_geom(wf, d, &outer, &inner);
    int64_t dep_stride = 1;
    for (int a = o->dep + 1; a < wf->ndim; a++) dep_stride *= wf->shape[a];
    int64_t dep_n = wf->shape[o->dep];
#pragma omp parallel
    {
        orc_c *lane = (orc_c *)xmalloc(sizeof(orc_c) * (size_t)n);
#pragma omp for collapse(2) schedule(static)
        for (int64_t a = 0; a < outer; a++)
            for (int64_t i = 0; i < inner; i++) {
                int64_t jj = (i / dep_stride) % dep_n;
                if (jj >= o->lanes) continue;
                const orc_c *m = mats + jj * n * n;
                for (int64_t k = 0; k < n; k++) lane[k] = wf->a[(a * n + k) * inner + i];
                for (int64_t r = 0; r < n; r++) {
                    orc_c acc = {0.0, 0.0};
                    for (int64_t k = 0; k < n; k++) acc = c_add(acc, c_mul(m[r * n + k], lane[k]));
                    wf->a[(a * n + r) * inner + i] = acc;
                }
            }
        free(lane);
    }
}

void orc_op_transform(orc_op *o, orc_wf *wf) {
    if (o->kind == K_FFT)
        fft_lanes(o, wf, 0);
    else if (o->kind == K_MATRIX)
        matrix_lanes(o, wf, o->t);
    else if (o->kind == K_STATE_MATRIX)
        state_matrix_lanes(o, wf, o->t);
    else
        orc_die("transform on a non-transformation");
}
void orc_op_inverse_transform(orc_op *o, orc_wf *wf) {
    if (o->kind == K_FFT)
        fft_lanes(o, wf, 1);
    else if (o->kind == K_MATRIX)
        matrix_lanes(o, wf, o->tinv);
    else if (o->kind == K_STATE_MATRIX)
        state_matrix_lanes(o, wf, o->tinv);
    else
        orc_die("inverse_transform on a non-transformation");
}

/* border_dumping.rs:62-84 — the raw mask in each enabled half, checked */
static void border_half(orc_op *o, orc_wf *wf) {
    if (o->chk.present) chk_before(&o->chk, wf);
    apply_one_dim(o, wf);
    if (o->chk.present) chk_after(&o->chk, wf);
}

/* leak_control.rs:31-45 */
static void leak_first(orc_op *o, orc_wf *wf) {
    o->leak_norm = orc_wf_norm(wf);
    if (o->chk.present) chk_before(&o->chk, wf);
}
static void leak_second(orc_op *o, orc_wf *wf) {
    if (o->chk.present) chk_after(&o->chk, wf);
    orc_wf_normalize(wf, o->leak_norm);
}

/* wave_function_saver.rs:41-64 (2-D density) and :137-151 (state density) */
static void saver_monitor(orc_op *o, orc_wf *wf) {
    if (o->frames_no == 0) orc_die("saver: frames_no is zero");
    int64_t freq = o->step_no / o->frames_no;
    if (freq == 0) orc_die("saver: frequency is zero (division by zero in the crate)");
    if (o->current_frame % freq == 0 && o->current_frame / freq < o->frames_no) {
        int64_t col = o->current_frame / freq;
        if (o->kind == K_WF_SAVER) {
            if (wf->ndim != 2) orc_die("Wave function must be 2d for now");
            int64_t nx = o->sgrid->nodes_no, ny = o->ygrid->nodes_no;
            for (int64_t i = 0; i < nx * ny; i++)
                o->data[i * o->frames_no + col] = c_norm_sqr(wf->a[i]);
        } else {
            int64_t n = o->sgrid->nodes_no;
            double *st = (double *)xmalloc(sizeof(double) * (size_t)wf->size);
            orc_wf_state_density(wf, o->sgrid->dimension_no, st);
            for (int64_t i = 0; i < n; i++) o->data[i * o->frames_no + col] = st[i];
            free(st);
        }
        o->times[o->ntimes++] = o->step * ((double)o->current_frame + 1.);
    }
    o->current_frame++;
}

/* ======================================================================
 * Propagation — src/propagation.rs:67-292
 * ====================================================================== */
enum { T_PROPAGATOR = 0, T_TRANSFORMATION = 1, T_SAVER = 2, T_CONTROL = 3 };
typedef struct {
    int type;
    orc_op *op;
    int flag; /* Order (0 Normal, 1 InverseFirst) or Apply bits */
} stack_item;

struct orc_prop {
    orc_wf *wf;
    double step;
    int64_t step_no;
    int im_time;
    stack_item *items;
    int nitems, cap;
};

orc_prop *orc_prop_new(void) {
    orc_prop *p = (orc_prop *)xmalloc(sizeof(orc_prop));
    memset(p, 0, sizeof(*p));
    return p;
}
void orc_prop_free(orc_prop *p) {
    if (!p) return;
    for (int i = 0; i < p->nitems; i++) orc_op_free(p->items[i].op);
    free(p->items);
    orc_wf_free(p->wf);
    free(p);
}
void orc_prop_set_wave_function(orc_prop *p, const orc_wf *wf) {
    orc_wf_free(p->wf);
    p->wf = orc_wf_clone(wf);
}
void orc_prop_set_time_grid(orc_prop *p, double step, int64_t step_no, int im_time) {
    p->step = step;
    p->step_no = step_no;
    p->im_time = im_time;
}
static void push_item(orc_prop *p, int type, orc_op *op, int flag) {
    if (p->nitems == p->cap) {
        p->cap = p->cap ? 2 * p->cap : 8;
        p->items = (stack_item *)realloc(p->items, sizeof(stack_item) * (size_t)p->cap);
    }
    p->items[p->nitems].type = type;
    p->items[p->nitems].op = op;
    p->items[p->nitems].flag = flag;
    p->nitems++;
}
void orc_prop_add_propagator(orc_prop *p, orc_op *op) { push_item(p, T_PROPAGATOR, op, 0); }
void orc_prop_add_transformation(orc_prop *p, orc_op *op, int inverse_first) {
    push_item(p, T_TRANSFORMATION, op, inverse_first);
}
/* propagation.rs:55-59 — apply must not be Apply::None */
int orc_prop_add_saver(orc_prop *p, orc_op *op, int apply) {
    if (apply == 0) return -1;
    push_item(p, T_SAVER, op, apply);
    return 0;
}
void orc_prop_add_control(orc_prop *p, orc_op *op, int apply) { push_item(p, T_CONTROL, op, apply); }

static void control_half(orc_op *o, orc_wf *wf, int second) {
    if (o->kind == K_LEAK)
        second ? leak_second(o, wf) : leak_first(o, wf);
    else if (o->kind == K_BORDER)
        border_half(o, wf);
    else
        orc_die("unknown control");
}

/* propagation.rs:134-182 */
void orc_prop_step(orc_prop *p) {
    orc_wf *wf = p->wf;
    for (int i = 0; i < p->nitems; i++) {
        stack_item *it = &p->items[i];
        switch (it->type) {
        case T_PROPAGATOR: orc_op_apply(it->op, wf); break;
        case T_TRANSFORMATION:
            if (it->flag == 0) orc_op_transform(it->op, wf);
            else orc_op_inverse_transform(it->op, wf);
            break;
        case T_SAVER:
            if (it->flag & 1) saver_monitor(it->op, wf);
            break;
        case T_CONTROL:
            if (it->flag & 1) control_half(it->op, wf, 0);
            break;
        }
    }
    for (int i = p->nitems - 2; i >= 0; i--) {
        stack_item *it = &p->items[i];
        switch (it->type) {
        case T_PROPAGATOR: orc_op_apply(it->op, wf); break;
        case T_TRANSFORMATION:
            if (it->flag == 0) orc_op_inverse_transform(it->op, wf);
            else orc_op_transform(it->op, wf);
            break;
        case T_SAVER:
            if (it->flag & 2) saver_monitor(it->op, wf);
            break;
        case T_CONTROL:
            if (it->flag & 2) control_half(it->op, wf, 1);
            break;
        }
    }
}

void orc_prop_steps(orc_prop *p, int64_t n) {
    for (int64_t i = 0; i < n; i++) orc_prop_step(p);
}

/* propagation.rs:185-190 (without the per-step println) */
void orc_prop_propagate(orc_prop *p) { orc_prop_steps(p, p->step_no); }

/* propagation.rs:245-291; *err = 1 when the imaginary-time precondition fails */
double orc_prop_mean_energy(orc_prop *p, int *err) {
    if (err) *err = 0;
    if (p->im_time) {
        if (p->nitems == 0 || p->items[0].type != T_CONTROL || p->items[0].op->kind != K_LEAK ||
            !p->items[0].op->chk.present) {
            if (err) *err = 1;
            return NAN;
        }
        p->items[0].op->chk.loss = 0.0;
        orc_prop_step(p);
        double decay = p->items[0].op->chk.loss;
        return -log(orc_wf_norm(p->wf) - decay) / p->step / 2.;
    }
    orc_wf *before = orc_wf_clone(p->wf);
    orc_prop_step(p);
    orc_wf *after = orc_wf_clone(p->wf);
    double d[2];
    orc_wf_dot(after, before, d);
    orc_wf_free(before);
    orc_wf_free(after);
    return -atan2(d[1], d[0]) / p->step;
}

/* propagation.rs:216-231 — Propagator losses only */
int64_t orc_prop_get_losses(orc_prop *p, double *out) {
    int64_t c = 0;
    for (int i = 0; i < p->nitems; i++)
        if (p->items[i].type == T_PROPAGATOR && p->items[i].op->chk.present) {
            if (out) out[c] = p->items[i].op->chk.loss;
            c++;
        }
    return c;
}
/* propagation.rs:124-131 */
void orc_prop_reset_losses(orc_prop *p) {
    for (int i = 0; i < p->nitems; i++)
        if (p->items[i].type == T_PROPAGATOR && p->items[i].op->chk.present) p->items[i].op->chk.loss = 0.0;
}
/* propagation.rs:115-122, wave_function_saver.rs:99-101,180-182 (the frame
 * counter and times are not reset, as in the crate) */
void orc_prop_reset_savers_state(orc_prop *p) {
    for (int i = 0; i < p->nitems; i++)
        if (p->items[i].type == T_SAVER) {
            orc_op *o = p->items[i].op;
            int64_t per = o->sgrid->nodes_no * (o->ygrid ? o->ygrid->nodes_no : 1);
            memset(o->data, 0, sizeof(double) * (size_t)(per * o->frames_no));
        }
}
orc_wf *orc_prop_wave_function(orc_prop *p) { return p->wf; }
