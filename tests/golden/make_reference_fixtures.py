"""Writes tests/golden/reference_fixtures.npz from the crate's own test data.

The crate's scenario tests (tests/harmonic_tests.rs, tests/ground_state_tests.rs,
tests/free_wave.rs) saved StateSaver frames, grids, potentials and frame times
as .npy files under tests/test_data/. This script keeps the grids, potentials
and times whole and a subset of the density frames (columns listed in FRAMES),
so the fixtures stay small. It reads the reference tree, so it runs only where
that tree exists; its output is committed and is what the tests load.

The 4x4 FFT known answer of tests/fft_tests.rs:15-74 (input 1..16, expected
transform along axis 0) is stored as data too.
"""
import os

import numpy as np

SRC = "/root/reference/tests/test_data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.npz")
FRAMES = {
    "harmonic": [0, 1, 2, 10, 25, 49],
    "harmonic_imaginary": [0, 1, 2, 10, 25, 49],
    "lj_ground_space": [0, 1, 5, 30, 60, 119],
    "free_wave": [0, 1, 10, 49],
    "free_momentum": [0, 1, 10, 49],
}
WHOLE = [
    "harmonic_potential", "harmonic_space_grid", "harmonic_time",
    "harmonic_imaginary_potential", "harmonic_imaginary_space_grid", "harmonic_imaginary_time",
    "lj_ground_space_potential", "lj_ground_space_space_grid", "lj_ground_space_time",
    "free_wave_space_grid", "free_wave_time", "free_momentum_momentum_grid", "free_momentum_time",
]


def main():
    out = {}
    for name in WHOLE:
        out[name] = np.load(os.path.join(SRC, name + ".npy"), allow_pickle=False)
    for name, cols in FRAMES.items():
        a = np.load(os.path.join(SRC, name + ".npy"), allow_pickle=False)
        out[name + "__cols"] = np.array(cols, dtype=np.int64)
        out[name + "__frames"] = a[:, cols]
        out[name + "__nframes"] = np.array(a.shape[1], dtype=np.int64)
    # tests/fft_tests.rs:15-40 (input) and :49-74 (expected transform along axis 0)
    out["fft4x4_input"] = np.arange(1, 17, dtype=np.float64).reshape(4, 4).astype(np.complex128)
    out["fft4x4_expected"] = np.array(
        [[14, 16, 18, 20], [-4 + 4j] * 4, [-4] * 4, [-4 - 4j] * 4], dtype=np.complex128)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
