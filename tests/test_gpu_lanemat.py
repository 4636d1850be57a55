"""NonDiagPropagator and StateMatrixTransformation on the GPU against the oracle and numpy.

Both apply a matrix per lane: NonDiag one per lane in the crate's lanes_mut order
(non_diagonal_propagator.rs:29-37), StateMatrix one per index of the dependent axis to every lane of that
slice (state_matrix_transformation.rs:52-74). `zip` leaves lanes (slices) beyond the matrices untouched;
the cases below have fewer matrices than lanes so that is checked too.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

so = pytest.importorskip("split_operator_amd")
from oracle import oracle as O  # noqa: E402


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


def _grids(B, shape):
    return [B.Grid.new_linear_continuos(f"g{d}", -1.0, 1.0, n, d) for d, n in enumerate(shape)]


@pytest.mark.parametrize("dtype,tol", [("f64", 1e-12), ("f32", 3e-6), ("f32x", 3e-7)])
@pytest.mark.parametrize("shape,axis,count", [((16, 6), 1, 10), ((6, 16), 0, 16), ((4, 5, 3), 1, 12)])
def test_non_diag_propagator(dtype, tol, shape, axis, count):
    rng = np.random.default_rng(sum(shape) + axis)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    n = shape[axis]
    ops = (rng.standard_normal((count, n, n)) + 1j * rng.standard_normal((count, n, n))) / np.sqrt(n)
    out = []
    for B in (so, O):
        g = _grids(B, shape)
        wf = B.WaveFunction(a, g, dtype=dtype) if B is so else B.WaveFunction(a, g)
        op = B.non_diag_propagator(axis, ops)
        op.apply(wf)
        out.append(wf.array)
    # numpy: lanes in C order of the other axes
    moved = np.moveaxis(a, axis, -1).reshape(-1, n).copy()
    for k in range(min(count, moved.shape[0])):
        moved[k] = ops[k] @ moved[k]
    ref = np.moveaxis(moved.reshape(np.moveaxis(a, axis, -1).shape), -1, axis)
    assert rel(out[1], ref) < 1e-13
    assert rel(out[0], ref) < tol


def test_non_diag_in_propagation_with_loss_checker():
    """A unitary-ish coupled-channel step inside a stack, its loss checked (the crate's check_before /
    check_after around apply), against the oracle."""
    shape = (64, 4)
    rng = np.random.default_rng(7)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    q, _ = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))
    ops = np.stack([q * (0.999 if k % 2 else 1.0) for k in range(64)])
    res = []
    for B in (so, O):
        g = _grids(B, shape)
        wf = B.WaveFunction(a, g)
        wf.normalize(1.0)
        p = B.Propagation()
        p.set_wave_function(wf)
        p.set_time_grid(0.01, 30, False)
        nd = B.non_diag_propagator(1, ops)
        nd.add_loss_checker("coupling")
        p.add_propagator(nd)
        fft = B.fft_transformation(g[0], "k")
        p.add_transformation(fft)
        p.add_propagator(B.one_dim_propagator(B.exp_propagator(B.kinetic_hamiltonian(g[0], 1.0),
                                                               B.select_step(True, 0.01, False)), 0))
        p.propagate()
        res.append((p.wave_function().array, p.get_losses()))
    assert rel(res[0][0], res[1][0]) < 1e-11
    assert len(res[0][1]) == len(res[1][1]) == 1
    assert abs(res[0][1][0] - res[1][1][0]) < 1e-12


@pytest.mark.parametrize("dtype,tol", [("f64", 1e-12), ("f32", 3e-6)])
@pytest.mark.parametrize("shape,dim,dep,count", [((12, 8, 5), 1, 2, 4), ((6, 8, 3, 2), 1, 3, 2), ((8, 7, 6), 0, 1, 6)])
def test_state_matrix_transformation(dtype, tol, shape, dim, dep, count):
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    n = shape[dim]
    ts = rng.standard_normal((count, n, n)) + 1j * rng.standard_normal((count, n, n))
    tis = np.linalg.inv(ts)
    outs = []
    for B in (so, O):
        g = _grids(B, shape)
        gl = B.Grid.new_linear_countable("l", 0.0, float(n - 1), n, dim)
        wf = B.WaveFunction(a, g, dtype=dtype) if B is so else B.WaveFunction(a, g)
        op = B.state_matrix_transformation(dep, g[dim], gl, ts, tis)
        op.transform(wf)
        mid = wf.array
        assert wf.grid(dim).name == "l"
        op.inverse_transform(wf)
        outs.append((mid, wf.array))
    ref = a.copy()
    am = np.moveaxis(a, (dim, dep), (-1, -2))
    rm = np.moveaxis(ref, (dim, dep), (-1, -2))
    for j in range(min(count, shape[dep])):
        rm[..., j, :] = am[..., j, :] @ ts[j].T
    assert rel(outs[1][0], ref) < 1e-12
    assert rel(outs[0][0], ref) < tol
    assert rel(outs[0][1], a) < tol * 10
    assert rel(outs[1][1], a) < 1e-10


def test_associated_legendre_stack_vs_oracle():
    """r x theta x Omega propagation with the associated Legendre operator (one basis per Omega,
    legendre_diagonalization.rs:83-109) against the oracle."""
    n_r, n_t, n_o = 64, 12, 3
    dt = 0.002
    res = []
    for B in (so, O):
        radial = B.Grid.new_linear_continuos("radial", 0.5, 10.0, n_r, 0)
        x, w = np.polynomial.legendre.leggauss(n_t)
        theta = np.arccos(x)[::-1].copy()
        polar = B.Grid.new_custom("polar", theta, w[::-1].copy(), 1)
        omega = B.Grid.new_linear_countable("omega", 0.0, float(n_o - 1), n_o, 2)
        r = radial.nodes
        psi0 = (np.exp(-((r - 4.0) ** 2))[:, None, None] * np.ones((1, n_t, n_o))).astype(np.complex128)
        psi0 = psi0 * np.sin(theta)[None, :, None] ** np.arange(n_o)[None, None, :]
        wf = B.WaveFunction(psi0, [radial, polar, omega])
        wf.normalize(1.0)
        p = B.Propagation()
        p.set_wave_function(wf)
        p.set_time_grid(dt, 20, False)
        V = (0.5 * (r - 4.0) ** 2)[:, None, None] * (1 + 0.1 * np.cos(theta))[None, :, None] * np.ones((1, 1, n_o))
        p.add_propagator(B.n_dim_propagator(B.exp_propagator(V, B.select_step(False, dt, False))))
        if B is so:
            op = so.associated_legendre_operator(polar, omega)
        else:
            mats = [O.associated_legendre_matrices(polar, int(om)) for om in omega.nodes]
            lg = O.Grid.new_linear_countable("angular_momentum", 0.0, float(n_t - 1), n_t, 1)
            op = O.state_matrix_transformation(2, polar, lg, np.stack([m[0] for m in mats]),
                                               np.stack([m[1] for m in mats]))
        p.add_transformation(op)
        p.add_transformation(B.fft_transformation(radial, "k"))
        p.add_propagator(B.one_dim_propagator(B.exp_propagator(B.kinetic_hamiltonian(radial, 1.0),
                                                               B.select_step(True, dt, False)), 0))
        p.propagate()
        res.append(p.wave_function().array)
    assert rel(res[0], res[1]) < 1e-10
