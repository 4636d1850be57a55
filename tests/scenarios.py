"""Scenario builders shared by the oracle tests and the GPU parity tests.

`B` is a backend module exposing the crate's API names: the oracle wrapper
(oracle/oracle.py) or the product package (split_operator_amd). Each builder
restates one of the crate's scenario tests (file:line given) or one of the
benchmark workloads, with the same operation stack.
"""
import numpy as np

# Li6+Li7 reduced mass in atomic units, recovered from the crate's fixture
# tests/test_data/harmonic_potential.npy as 2 V(x) / (omega^2 x^2)
MU_LI6_LI7 = 5903.538543342382
# Kelvin -> hartree (CODATA 2018 k_B / E_h); used by tests/ground_state_tests.rs:345-348
# and tests/free_wave.rs:451 through the quantum crate
KELVIN_TO_AU = 3.166811563455546e-06
# Dalton -> electron masses (CODATA 2018)
DALTON_TO_AU = 1822.888486209


def harmonic_trap(B, im_time, name="harmonic", steps=1000, frames=50):
    """tests/harmonic_trap.rs:34-111 (Li6+Li7 in a harmonic trap, 256 points)."""
    g = B.Grid.new_linear_continuos("space", -4.0, 4.0, 256, 0)
    x = g.nodes
    arr = np.array([B.gaussian_distribution(xi, 2.0, 0.2, 0.0) for xi in x])
    wf = B.WaveFunction(arr, [g])
    wf.normalize(1.0)
    p = B.Propagation()
    p.set_wave_function(wf)
    p.set_time_grid(3.0, steps, im_time)
    V = np.array([B.harmonic(xi, 0.0, MU_LI6_LI7, 0.001) for xi in x])
    vp = B.one_dim_propagator(B.exp_propagator(V, B.select_step(False, 3.0, im_time)), 0)
    fft = B.fft_transformation(g, "momentum")
    T = B.kinetic_hamiltonian(g, MU_LI6_LI7)
    tp = B.one_dim_propagator(B.exp_propagator(T, B.select_step(True, 3.0, im_time)), 0)
    sv = B.state_saver(name, 3.0, steps, g, frames)
    p.add_saver(sv, B.APPLY_FIRST)
    p.add_control(B.leak_control(), B.APPLY_FIRST | B.APPLY_SECOND)
    p.add_propagator(vp)
    p.add_transformation(fft)
    p.add_propagator(tp)
    return p, sv, V


def lj_ground(B, steps=1000, frames=120):
    """tests/ground_state_tests.rs:51-139 (Lennard-Jones, imaginary time, 512 points)."""
    g = B.Grid.new_linear_continuos("space", 8.0, 20.0, 512, 0)
    x = g.nodes
    k = np.sqrt(2.0 * MU_LI6_LI7 * 1e-7 * KELVIN_TO_AU)
    arr = np.array([B.gaussian_distribution(xi, 14.0, 2.0, k) for xi in x])
    wf = B.WaveFunction(arr, [g])
    wf.normalize(1.0)
    p = B.Propagation()
    p.set_wave_function(wf)
    p.set_time_grid(50.0, steps, True)
    V = np.array([B.lennard_jones(xi, 0.0003, 9.7) for xi in x])
    vp = B.one_dim_propagator(B.exp_propagator(V, B.select_step(False, 50.0, True)), 0)
    T = B.kinetic_hamiltonian(g, MU_LI6_LI7)
    tp = B.one_dim_propagator(B.exp_propagator(T, B.select_step(True, 50.0, True)), 0)
    fft = B.fft_transformation(g, "momentum")
    sv = B.state_saver("lj_ground_space", 50.0, steps, g, frames)
    lc = B.leak_control()
    lc.add_loss_checker("leak control")
    p.add_control(lc, B.APPLY_FIRST | B.APPLY_SECOND)
    p.add_saver(sv, B.APPLY_FIRST)
    p.add_propagator(vp)
    p.add_transformation(fft)
    p.add_propagator(tp)
    return p, sv, V


def free_wave(B, momentum_saver=False, steps=500, frames=50):
    """tests/free_wave.rs:10-84 (free packet, 12 Da + 12 Da at 3000 K, 512 points).

    With momentum_saver the StateSaver sits between the FFT and the kinetic
    propagator, which reproduces tests/test_data/free_momentum*.npy (the crate
    holds no test that writes them)."""
    mu = 6.0 * DALTON_TO_AU
    k = np.sqrt(2.0 * mu * 3000.0 * KELVIN_TO_AU)
    g = B.Grid.new_linear_continuos("space", 0.0, 30.0, 512, 0)
    x = g.nodes
    arr = np.array([B.gaussian_distribution(xi, 15.0, 0.5, k) for xi in x])
    wf = B.WaveFunction(arr, [g])
    wf.normalize(1.0)
    p = B.Propagation()
    p.set_wave_function(wf)
    p.set_time_grid(50.0, steps, False)
    V = np.zeros(512)
    vp = B.one_dim_propagator(B.exp_propagator(V, B.select_step(False, 50.0, False)), 0)
    fft = B.fft_transformation(g, "momentum")
    T = B.kinetic_hamiltonian(g, mu)
    tp = B.one_dim_propagator(B.exp_propagator(T, B.select_step(True, 50.0, False)), 0)
    if momentum_saver:
        kg = fft.held_grid()
        sv = B.state_saver("free_momentum", 50.0, steps, kg, frames)
        p.add_control(B.leak_control(), B.APPLY_FIRST | B.APPLY_SECOND)
        p.add_propagator(vp)
        p.add_transformation(fft)
        p.add_saver(sv, B.APPLY_FIRST)
        p.add_propagator(tp)
    else:
        sv = B.state_saver("free_wave", 50.0, steps, g, frames)
        p.add_saver(sv, B.APPLY_FIRST)
        p.add_control(B.leak_control(), B.APPLY_FIRST | B.APPLY_SECOND)
        p.add_propagator(vp)
        p.add_transformation(fft)
        p.add_propagator(tp)
    return p, sv
