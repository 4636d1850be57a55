"""ctypes wrapper over the C f64 oracle (oracle/so_oracle.c).

TEST INFRASTRUCTURE ONLY: this is the checker. It may be imported by tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py, never by
the product package. The class and method names mirror the split_operator
crate (src/grid.rs, src/wave_function.rs, src/propagation.rs, ...) so that a
scenario reads the same whether it is built on the oracle or on the HIP path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libso_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -O2, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        sig = {
            "orc_grid_linear_continuos": (vp, [C.c_char_p, C.c_double, C.c_double, C.c_int64, C.c_int]),
            "orc_grid_linear_countable": (vp, [C.c_char_p, C.c_double, C.c_double, C.c_int64, C.c_int]),
            "orc_grid_custom": (vp, [C.c_char_p, C.c_int64, _dp, _dp, C.c_int]),
            "orc_grid_free": (None, [vp]),
            "orc_grid_nodes": (_dp, [vp]),
            "orc_grid_weights": (_dp, [vp]),
            "orc_grid_len": (C.c_int64, [vp]),
            "orc_grid_name": (C.c_char_p, [vp]),
            "orc_wf_new": (vp, [C.c_int, _i64p, _dp, C.POINTER(vp)]),
            "orc_wf_clone": (vp, [vp]),
            "orc_wf_free": (None, [vp]),
            "orc_wf_data": (_dp, [vp]),
            "orc_wf_size": (C.c_int64, [vp]),
            "orc_wf_shape": (C.c_int, [vp, _i64p]),
            "orc_wf_grid": (vp, [vp, C.c_int]),
            "orc_wf_norm": (C.c_double, [vp]),
            "orc_wf_normalize": (None, [vp, C.c_double]),
            "orc_wf_dot": (None, [vp, vp, _dp]),
            "orc_wf_density": (None, [vp, _dp]),
            "orc_wf_state_density": (None, [vp, C.c_int, _dp]),
            "orc_gaussian_distribution": (None, [C.c_double] * 4 + [_dp]),
            "orc_select_step": (None, [C.c_int, C.c_double, C.c_int, _dp]),
            "orc_exp_propagator": (None, [_dp, C.c_int64, C.c_double, C.c_double, _dp]),
            "orc_exp_propagator_complex": (None, [_dp, C.c_int64, C.c_double, C.c_double, _dp]),
            "orc_momentum_grid": (None, [vp, _dp, _dp]),
            "orc_kinetic_hamiltonian": (None, [vp, C.c_double, _dp]),
            "orc_legendre_polynomials": (None, [C.c_int64, C.c_double, _dp]),
            "orc_legendre_matrices": (None, [vp, _dp, _dp]),
            "orc_double_factorial": (C.c_double, [C.c_int64]),
            "orc_associated_legendre_polynomials": (None, [C.c_int64, C.c_int, C.c_double, _dp]),
            "orc_legendre_normalization": (C.c_double, [C.c_int64, C.c_int]),
            "orc_associated_legendre_matrices": (None, [vp, C.c_int, _dp, _dp]),
            "orc_state_matrix_transformation": (vp, [C.c_int, vp, vp, C.c_int64, _dp, _dp]),
            "orc_rotational_hamiltonian": (None, [vp, vp, C.c_double, C.c_double, _dp]),
            "orc_dispersion": (C.c_double, [C.c_double, C.c_int, C.c_double]),
            "orc_lennard_jones": (C.c_double, [C.c_double] * 3),
            "orc_harmonic": (C.c_double, [C.c_double] * 4),
            "orc_dumping_end": (None, [C.c_double, C.c_double, vp, _dp]),
            "orc_gauss_legendre": (None, [C.c_int64, _dp, _dp]),
            "orc_one_dim_propagator": (vp, [C.c_int64, C.c_int, _dp]),
            "orc_n_dim_propagator": (vp, [C.c_int, _i64p, _dp]),
            "orc_fft_transformation": (vp, [vp, C.c_char_p]),
            "orc_op_held_grid": (vp, [vp]),
            "orc_matrix_transformation": (vp, [vp, vp, _dp, _dp]),
            "orc_non_diag_propagator": (vp, [C.c_int, C.c_int64, C.c_int64, _dp]),
            "orc_border_dumping": (vp, [_dp, vp]),
            "orc_leak_control": (vp, []),
            "orc_state_saver": (vp, [C.c_char_p, C.c_double, C.c_int64, vp, C.c_int64]),
            "orc_wave_function_saver": (vp, [C.c_char_p, C.c_double, C.c_int64, vp, vp, C.c_int64]),
            "orc_op_set_loss_checker": (None, [vp, C.c_char_p, C.c_int64, C.c_double, C.c_int64]),
            "orc_op_loss": (C.c_double, [vp]),
            "orc_op_loss_frames": (C.c_int64, [vp, _dp, _dp]),
            "orc_saver_frames": (C.c_int64, [vp, _dp, _dp]),
            "orc_op_free": (None, [vp]),
            "orc_op_apply": (None, [vp, vp]),
            "orc_op_transform": (None, [vp, vp]),
            "orc_op_inverse_transform": (None, [vp, vp]),
            "orc_prop_new": (vp, []),
            "orc_prop_free": (None, [vp]),
            "orc_prop_set_wave_function": (None, [vp, vp]),
            "orc_prop_set_time_grid": (None, [vp, C.c_double, C.c_int64, C.c_int]),
            "orc_prop_add_propagator": (None, [vp, vp]),
            "orc_prop_add_transformation": (None, [vp, vp, C.c_int]),
            "orc_prop_add_saver": (C.c_int, [vp, vp, C.c_int]),
            "orc_prop_add_control": (None, [vp, vp, C.c_int]),
            "orc_prop_step": (None, [vp]),
            "orc_prop_steps": (None, [vp, C.c_int64]),
            "orc_prop_propagate": (None, [vp]),
            "orc_prop_mean_energy": (C.c_double, [vp, C.POINTER(C.c_int)]),
            "orc_prop_get_losses": (C.c_int64, [vp, _dp]),
            "orc_prop_reset_losses": (None, [vp]),
            "orc_prop_reset_savers_state": (None, [vp]),
            "orc_prop_wave_function": (vp, [vp]),
            "orc_fft_raw": (None, [_dp, _dp, C.c_int64, C.c_int]),
            "orc_num_threads": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def num_threads():
    return lib().orc_num_threads()


def _d(a):
    return a.ctypes.data_as(_dp)


def _c128(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


# ---------------------------------------------------------------- Grid (src/grid.rs)
class Grid:
    def __init__(self, handle, owned=True):
        self._h = handle
        self._owned = owned

    @staticmethod
    def new_linear_continuos(name, start, end, n, dim):
        return Grid(lib().orc_grid_linear_continuos(name.encode(), start, end, n, dim))

    @staticmethod
    def new_linear_countable(name, start, end, n, dim):
        return Grid(lib().orc_grid_linear_countable(name.encode(), start, end, n, dim))

    @staticmethod
    def new_custom(name, nodes, weights, dim):
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        weights = np.ascontiguousarray(weights, dtype=np.float64)
        return Grid(lib().orc_grid_custom(name.encode(), len(nodes), _d(nodes), _d(weights), dim))

    @property
    def nodes(self):
        n = lib().orc_grid_len(self._h)
        return np.ctypeslib.as_array(lib().orc_grid_nodes(self._h), shape=(n,)).copy()

    @property
    def weights(self):
        n = lib().orc_grid_len(self._h)
        return np.ctypeslib.as_array(lib().orc_grid_weights(self._h), shape=(n,)).copy()

    @property
    def name(self):
        return lib().orc_grid_name(self._h).decode()

    @property
    def nodes_no(self):
        return lib().orc_grid_len(self._h)

    def __del__(self):
        if getattr(self, "_owned", False) and self._h and _lib is not None:
            _lib.orc_grid_free(self._h)
            self._h = None


# ---------------------------------------------------------------- WaveFunction
class WaveFunction:
    def __init__(self, array, grids, _handle=None, _owned=True):
        if _handle is not None:
            self._h = _handle
            self._owned = _owned
            return
        a = _c128(array)
        self.shape = a.shape
        shape = np.array(a.shape, dtype=np.int64)
        hs = (C.c_void_p * len(grids))(*[g._h for g in grids])
        self._h = lib().orc_wf_new(a.ndim, shape.ctypes.data_as(_i64p), a.view(np.float64).ctypes.data_as(_dp), hs)
        self._owned = True

    @property
    def array(self):
        n = lib().orc_wf_size(self._h)
        shape = np.zeros(8, dtype=np.int64)
        nd = lib().orc_wf_shape(self._h, shape.ctypes.data_as(_i64p))
        a = np.ctypeslib.as_array(lib().orc_wf_data(self._h), shape=(2 * n,)).copy().view(np.complex128)
        return a.reshape(tuple(shape[:nd]))

    def grid(self, axis):
        return Grid(lib().orc_wf_grid(self._h, axis), owned=False)

    def norm(self):
        return lib().orc_wf_norm(self._h)

    def normalize(self, new_norm):
        lib().orc_wf_normalize(self._h, new_norm)

    def dot(self, other):
        out = np.zeros(2)
        lib().orc_wf_dot(self._h, other._h, _d(out))
        return complex(out[0], out[1])

    def density(self):
        n = lib().orc_wf_size(self._h)
        out = np.zeros(n)
        lib().orc_wf_density(self._h, _d(out))
        return out

    def state_density(self, axis, n_axis):
        n = lib().orc_wf_size(self._h)
        out = np.zeros(max(n, n_axis))
        lib().orc_wf_state_density(self._h, axis, _d(out))
        return out[:n_axis] if n != n_axis else out

    def clone(self):
        return WaveFunction(None, None, _handle=lib().orc_wf_clone(self._h))

    def __del__(self):
        if getattr(self, "_owned", False) and self._h and _lib is not None:
            _lib.orc_wf_free(self._h)
            self._h = None


def gaussian_distribution(x, x0, sigma, momentum):
    out = np.zeros(2)
    lib().orc_gaussian_distribution(x, x0, sigma, momentum, _d(out))
    return complex(out[0], out[1])


def select_step(full, step, im_time):
    out = np.zeros(2)
    lib().orc_select_step(1 if full else 0, step, 1 if im_time else 0, _d(out))
    return complex(out[0], out[1])


# ---------------------------------------------------------------- factories
def exp_propagator(h, dt):
    h = np.ascontiguousarray(h, dtype=np.float64)
    out = np.zeros(h.size, dtype=np.complex128)
    lib().orc_exp_propagator(_d(h.ravel()), h.size, dt.real, dt.imag, _d(out.view(np.float64)))
    return out.reshape(h.shape)


def exp_propagator_complex(h, dt):
    h = _c128(h)
    out = np.zeros(h.size, dtype=np.complex128)
    lib().orc_exp_propagator_complex(_d(h.view(np.float64).ravel()), h.size, dt.real, dt.imag,
                                     _d(out.view(np.float64)))
    return out.reshape(h.shape)


def momentum_grid(grid):
    n = grid.nodes_no
    k = np.zeros(n)
    w = np.zeros(n)
    lib().orc_momentum_grid(grid._h, _d(k), _d(w))
    m = 2 * (n // 2)
    return k[:m], w[:m]


def kinetic_hamiltonian(grid, red_mass):
    n = grid.nodes_no
    out = np.zeros(n)
    lib().orc_kinetic_hamiltonian(grid._h, red_mass, _d(out))
    return out[: 2 * (n // 2)]


def legendre_polynomials(j, x):
    out = np.zeros(j + 1)
    lib().orc_legendre_polynomials(j, x, _d(out))
    return out


def legendre_matrices(polar):
    n = polar.nodes_no
    t = np.zeros((n, n), dtype=np.complex128)
    ti = np.zeros((n, n), dtype=np.complex128)
    lib().orc_legendre_matrices(polar._h, _d(t.view(np.float64)), _d(ti.view(np.float64)))
    return t, ti


def double_factorial(n):
    return lib().orc_double_factorial(n)


def associated_legendre_polynomials(j, m, x):
    out = np.zeros(max(j, abs(m) + 1) + 1)
    lib().orc_associated_legendre_polynomials(j, m, x, _d(out))
    return out[: j + 1]


def normalization(l, m):
    return lib().orc_legendre_normalization(l, m)


def associated_legendre_matrices(polar, omega):
    n = polar.nodes_no
    t = np.zeros((n, n), dtype=np.complex128)
    ti = np.zeros((n, n), dtype=np.complex128)
    lib().orc_associated_legendre_matrices(polar._h, int(omega), _d(t.view(np.float64)), _d(ti.view(np.float64)))
    return t, ti


def rotational_hamiltonian(radial, polar, red_mass, rot_const):
    out = np.zeros((radial.nodes_no, polar.nodes_no))
    lib().orc_rotational_hamiltonian(radial._h, polar._h, red_mass, rot_const, _d(out))
    return out


def dispersion(r, n, cn):
    return lib().orc_dispersion(r, n, cn)


def lennard_jones(r, d6, r6):
    return lib().orc_lennard_jones(r, d6, r6)


def harmonic(r, r0, mass, omega):
    return lib().orc_harmonic(r, r0, mass, omega)


def dumping_end(mask_width, mask_end, grid):
    out = np.zeros(grid.nodes_no, dtype=np.complex128)
    lib().orc_dumping_end(mask_width, mask_end, grid._h, _d(out.view(np.float64)))
    return out


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _d(x), _d(w))
    return x, w


def fft_raw(x, sign=-1):
    x = _c128(x)
    out = np.zeros_like(x)
    lib().orc_fft_raw(_d(x.view(np.float64)), _d(out.view(np.float64)), x.size, sign)
    return out


# ---------------------------------------------------------------- operations
class Op:
    """An operation handle. Ownership moves to the Propagation when added."""

    def __init__(self, handle):
        self._h = handle
        self._owned = True

    def set_loss_checked(self, name, frames_no=0, step=0.0, step_no=0):
        lib().orc_op_set_loss_checker(self._h, name.encode(), frames_no, step, step_no)

    add_loss_checker = set_loss_checked

    def loss(self):
        return lib().orc_op_loss(self._h)

    def loss_frames(self):
        n = lib().orc_op_loss_frames(self._h, None, None)
        t = np.zeros(max(n, 1))
        l = np.zeros(max(n, 1))
        lib().orc_op_loss_frames(self._h, _d(t), _d(l))
        return t[:n], l[:n]

    def loss_tail(self, count):
        """The last `count` cumulative losses the LossSaver recorded (oldest first)."""
        return self.loss_frames()[1][-count:]

    def held_grid(self):
        return Grid(lib().orc_op_held_grid(self._h), owned=False)

    def apply(self, wf):
        lib().orc_op_apply(self._h, wf._h)

    def transform(self, wf):
        lib().orc_op_transform(self._h, wf._h)

    def inverse_transform(self, wf):
        lib().orc_op_inverse_transform(self._h, wf._h)

    def __del__(self):
        if getattr(self, "_owned", False) and self._h and _lib is not None:
            _lib.orc_op_free(self._h)
            self._h = None


def one_dim_propagator(op, dim):
    op = _c128(op)
    return Op(lib().orc_one_dim_propagator(op.size, dim, _d(op.view(np.float64))))


def n_dim_propagator(op=None):
    if op is None:
        return Op(lib().orc_n_dim_propagator(0, None, None))
    op = _c128(op)
    shape = np.array(op.shape, dtype=np.int64)
    return Op(lib().orc_n_dim_propagator(op.ndim, shape.ctypes.data_as(_i64p), _d(op.view(np.float64))))


def fft_transformation(grid, name):
    return Op(lib().orc_fft_transformation(grid._h, name.encode()))


def matrix_transformation(grid, transformed, t, tinv):
    t = _c128(t)
    ti = _c128(tinv)
    return Op(lib().orc_matrix_transformation(grid._h, transformed._h, _d(t.view(np.float64)),
                                              _d(ti.view(np.float64))))


def state_matrix_transformation(dimension_no_dependent, grid, grid_transformation, transformations,
                                inverse_transformations):
    ts = _c128(transformations)
    tis = _c128(inverse_transformations)
    return Op(lib().orc_state_matrix_transformation(dimension_no_dependent, grid._h, grid_transformation._h,
                                                    ts.shape[0], _d(ts.view(np.float64)), _d(tis.view(np.float64))))


def non_diag_propagator(dim, ops):
    ops = _c128(ops)
    return Op(lib().orc_non_diag_propagator(dim, ops.shape[0], ops.shape[1], _d(ops.view(np.float64))))


def border_dumping(mask, grid):
    m = _c128(mask)
    return Op(lib().orc_border_dumping(_d(m.view(np.float64)), grid._h))


def leak_control():
    return Op(lib().orc_leak_control())


class Saver(Op):
    def __init__(self, handle, per_frame):
        super().__init__(handle)
        self.per_frame = per_frame

    def frames(self, frames_no):
        data = np.zeros(self.per_frame * frames_no)
        times = np.zeros(frames_no + 1)
        n = lib().orc_saver_frames(self._h, _d(data), _d(times))
        return data.reshape(-1, frames_no), times[:n]


def state_saver(name, step, step_no, state_grid, frames_no):
    return Saver(lib().orc_state_saver(name.encode(), step, step_no, state_grid._h, frames_no), state_grid.nodes_no)


def wave_function_saver(name, step, step_no, x, y, frames_no):
    return Saver(lib().orc_wave_function_saver(name.encode(), step, step_no, x._h, y._h, frames_no),
                 x.nodes_no * y.nodes_no)


APPLY_FIRST = 1
APPLY_SECOND = 2


class Propagation:
    def __init__(self):
        self._h = lib().orc_prop_new()
        self._ops = []

    def set_wave_function(self, wf):
        lib().orc_prop_set_wave_function(self._h, wf._h)

    def set_time_grid(self, step, step_no, im_time):
        self.step = step
        lib().orc_prop_set_time_grid(self._h, step, step_no, 1 if im_time else 0)

    def _take(self, op):
        op._owned = False
        self._ops.append(op)

    def add_propagator(self, op):
        self._take(op)
        lib().orc_prop_add_propagator(self._h, op._h)

    def add_transformation(self, op, inverse_first=False):
        self._take(op)
        lib().orc_prop_add_transformation(self._h, op._h, 1 if inverse_first else 0)

    def add_saver(self, op, apply):
        if lib().orc_prop_add_saver(self._h, op._h, apply) != 0:
            raise ValueError("saver apply must not be Apply::None")
        self._take(op)

    def add_control(self, op, apply):
        self._take(op)
        lib().orc_prop_add_control(self._h, op._h, apply)

    def step_once(self):
        lib().orc_prop_step(self._h)

    def steps(self, n):
        lib().orc_prop_steps(self._h, n)

    def propagate(self):
        lib().orc_prop_propagate(self._h)

    def mean_energy(self):
        err = C.c_int(0)
        e = lib().orc_prop_mean_energy(self._h, C.byref(err))
        if err.value:
            raise RuntimeError("For imaginary time propagation first control must be LeakControl with a loss checker")
        return e

    def get_losses(self):
        n = lib().orc_prop_get_losses(self._h, None)
        out = np.zeros(max(n, 1))
        lib().orc_prop_get_losses(self._h, _d(out))
        return list(out[:n])

    def reset_losses(self):
        lib().orc_prop_reset_losses(self._h)

    def reset_savers_state(self):
        lib().orc_prop_reset_savers_state(self._h)

    def wave_function(self):
        return WaveFunction(None, None, _handle=lib().orc_prop_wave_function(self._h), _owned=False)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_prop_free(self._h)
            self._h = None
