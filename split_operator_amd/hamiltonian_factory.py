"""Host-side (f64, numpy) operator construction, mirroring the crate's setup code.

These run once before propagation and hand arrays to the device path:
src/time_grid.rs:22-33, src/propagator/propagator_factory.rs:10-50,
src/hamiltonian_factory/*.rs, src/special_functions.rs:3-14,
src/border_dumping.rs:14-33, src/wave_function.rs:152-154.
"""
import math

import numpy as np


def select_step(full, step, im_time):
    """time_grid.rs:22-33: dt (real time) or -i dt (imaginary), half or full."""
    ts = step if full else step / 2.0
    return complex(-ts * 0.0, -ts) if im_time else complex(ts, 0.0)


def exp_propagator(hamiltonian, dt):
    """propagator_factory.rs:10-36: exp(-i H dt) element-wise (H real)."""
    h = np.asarray(hamiltonian, dtype=np.float64)
    with np.errstate(invalid="ignore", over="ignore"):
        z = (-1j * h) * complex(dt)
        out = np.exp(z)
    # exp(-inf + i*0) = 0 (an infinite barrier in imaginary time)
    out = np.where(np.isneginf(z.real) & np.isfinite(z.imag), 0.0, out)
    return out


def exp_propagator_complex(hamiltonian, dt):
    """propagator_factory.rs:39-50: exp(-i H dt) for complex H."""
    h = np.asarray(hamiltonian, dtype=np.complex128)
    return np.exp((-1j * h) * complex(dt))


def momentum_grid(grid):
    """fft_transformation.rs:28-37: k = dk (0..N/2, -N/2..-1), weights dk halved at N/2-1 and N/2."""
    n = grid.nodes_no
    nodes = grid.nodes
    dk = 2.0 * math.pi / (nodes[-1] - nodes[0]) * (1.0 - 1.0 / n)
    idx = np.concatenate([np.arange(0, n // 2), np.arange(-(n // 2), 0)]).astype(np.float64)
    k = idx * dk
    w = np.full(k.shape, dk)
    w[n // 2 - 1] *= 0.5
    w[n // 2] *= 0.5
    return k, w


def kinetic_hamiltonian(grid, red_mass):
    """kinetic_operator.rs:8-21: k^2 / (2 mu) in FFT order."""
    k, _ = momentum_grid(grid)
    return k * k / (2.0 * red_mass)


def legendre_polynomials(j, x):
    """special_functions.rs:3-14 (Bonnet recursion)."""
    if j < 1:
        raise IndexError("legendre_polynomials needs j >= 1 (special_functions.rs:7)")
    p = np.zeros(j + 1)
    p[0] = 1.0
    p[1] = x
    for i in range(2, j + 1):
        p[i] = ((2 * i - 1) / i) * x * p[i - 1] - ((i - 1) / i) * p[i - 2]
    return p


def legendre_matrices(polar_grid):
    """legendre_diagonalization.rs:13-45: T[i,j] = sqrt(l_i+1/2) P_l(cos th_j) w_j and
    Tinv[j,i] = sqrt(l_i+1/2) P_l(cos th_j) (complex arrays with zero imaginary part)."""
    n = polar_grid.nodes_no
    tinv = np.zeros((n, n), dtype=np.complex128)
    for j in range(n):
        pl = legendre_polynomials(n - 1, math.cos(polar_grid.nodes[j]))
        tinv[j, :] = np.sqrt(np.arange(n) + 0.5) * pl
    t = tinv.T * polar_grid.weights[None, :]
    return np.ascontiguousarray(t), tinv


def double_factorial(n):
    """special_functions.rs:43-55: n (n-2) (n-4) ... down to 3 or 2; 1 for n = 0."""
    if n >= 100:
        raise ValueError("double_factorial needs n < 100 (special_functions.rs:44)")
    v = float(n) if n > 0 else 1.0
    for k in range(n - 2, 1, -2):
        v *= k
    return v


def _negate_m(l, m):
    """special_functions.rs:57-80."""
    v = -1.0 if m % 2 else 1.0
    if m > 0:
        for k in range(l - m, l + m):
            v /= k + 1.0
    elif m < 0:
        for k in range(l + m, l - m):
            v *= k + 1.0
    else:
        v = 1.0
    return v


def associated_legendre_polynomials(j, m, x):
    """special_functions.rs:17-41: P_l^m(x) for l = 0..j (Condon-Shortley phase), upward in l; 0 below |m|."""
    mu = abs(m)
    if m == 0:
        return legendre_polynomials(j, x)
    if mu + 1 > j:
        raise IndexError("associated_legendre_polynomials needs j > |m| (index out of bounds in the crate)")
    p = np.zeros(j + 1)
    p[mu] = (-1.0 if m % 2 else 1.0) * double_factorial(2 * mu - 1) * (1.0 - x * x) ** (mu / 2.0)
    if m < 0:
        p[mu] *= _negate_m(mu, mu)
        p[mu + 1] = x * p[mu]
    else:
        p[mu + 1] = x * (2.0 * mu + 1.0) * p[mu]
    for i in range(mu + 2, j + 1):
        l = i - 1
        p[i] = ((2 * l + 1) * x * p[i - 1] - (l + m) * p[i - 2]) / (l - m + 1.0)
    return p


def normalization(l, m):
    """special_functions.rs:82-116: sqrt((l + 1/2) (l - m)! / (l + m)!)."""
    v = 1.0
    if m > 0:
        for k in range(l - m, l + m):
            v /= k + 1.0
    elif m < 0:
        for k in range(l + m, l - m):
            v *= k + 1.0
    return math.sqrt(v * (l + 0.5))


def associated_legendre_matrices(polar_grid, omega):
    """legendre_diagonalization.rs:48-80: T[i,j] = N_l P_l^omega(cos th_j) w_j for l_i >= |omega| (0 above),
    Tinv = the unweighted matrix transposed."""
    n = polar_grid.nodes_no
    om = abs(int(omega))
    tr = np.zeros((n, n), dtype=np.complex128)
    for j in range(n):
        pl = associated_legendre_polynomials(n - 1, int(omega), math.cos(polar_grid.nodes[j]))
        for i in range(om, n):
            tr[i, j] = normalization(i, int(omega)) * pl[i]
    tinv = np.ascontiguousarray(tr.T)
    t = tr * polar_grid.weights[None, :]
    return np.ascontiguousarray(t), tinv


def rotational_hamiltonian(radial_grid, polar_grid, red_mass, rot_const):
    """rotational_operator.rs:7-26: H[r, l] = l(l+1)(B + 1/(2 mu r^2))."""
    l = np.arange(polar_grid.nodes_no, dtype=np.float64)
    r = radial_grid.nodes
    return (l * (l + 1))[None, :] * (rot_const + 1.0 / (2.0 * red_mass * r * r))[:, None]


def dispersion(r, n, cn):
    """analytic_potentials.rs:2-10."""
    v = cn * r ** n
    return math.inf if v != v else v


def lennard_jones(r, d6, r6):
    """analytic_potentials.rs:13-22."""
    rr = (r6 / r) ** 6
    v = d6 * rr * (rr - 2.0)
    return math.inf if v != v else v


def harmonic(r, r0, mass, omega):
    """analytic_potentials.rs:24-32."""
    v = 0.5 * mass * omega ** 2 * (r - r0) ** 2
    return math.inf if v != v else v


def dumping_end(mask_width, mask_end, grid):
    """border_dumping.rs:14-33: 1 below r_max - width, 0 beyond r_max - end, sine ramp between."""
    x = grid.nodes
    rmax = x[-1]
    m = np.where(x < rmax - mask_width, 1.0,
                 np.where(x > rmax - mask_end, 0.0, np.sin(math.pi / 2.0 * (rmax - x) / mask_width)))
    return m.astype(np.complex128)


def gaussian_distribution(x, x0, sigma, momentum):
    """wave_function.rs:152-154: exp(-((x-x0)/(2 sigma))^2 - i (x-x0) k)."""
    q = (x - x0) / (2.0 * sigma)
    return complex(np.exp(complex(-(q * q), -(x - x0) * momentum)))


def two_dim_into_n_dim_operator(shape, hamiltonian, dim1, dim2):
    """hamiltonian_broadcasting.rs:6-23: broadcast a (dim1, dim2) operator to `shape`."""
    h = np.asarray(hamiltonian)
    ndim = len(shape)
    idx = [None] * ndim
    idx[dim1] = slice(None)
    idx[dim2] = slice(None)
    hh = h if dim1 < dim2 else h.T
    return np.broadcast_to(hh[tuple(idx)], shape).copy()
