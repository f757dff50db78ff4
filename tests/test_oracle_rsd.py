"""Pins the float32 restatement of pos_redshift_space (redshift_space_library.pyx:29-43)."""
import numpy as np

from oracle import rsd


def test_factor_is_float32_of_double_ratio():
    f = rsd.factor(np.float32(67.3), np.float32(0.5))
    assert f.dtype == np.float32
    assert f == np.float32(1.5 / float(np.float32(67.3)))


def test_exact_shifts_and_wraps():
    L = 100.0
    pos = np.array([[10.0, 20.0, 30.0], [10.0, 99.0, 30.0], [10.0, 1.0, 30.0], [10.0, 50.0, 30.0],
                    [10.0, 75.0, 30.0], [10.0, 0.0, 30.0]], np.float32)
    # H = 64, z = 0: factor 1/64 exactly, so every shift below is exact
    vel = np.zeros_like(pos)
    vel[:, 1] = np.array([128.0, 128.0, -192.0, 6400.0, -9600.0, 0.0], np.float32)
    out = rsd.pos_redshift_space(pos, vel, L, 64.0, 0.0, 1)
    assert np.array_equal(out[:, [0, 2]], pos[:, [0, 2]])             # other axes untouched
    assert out[0, 1] == np.float32(22.0)                               # plain shift
    assert out[1, 1] == np.float32(1.0)                                # 101 -> wraps to 1
    assert out[2, 1] == np.float32(98.0)                               # -2 -> 98
    assert out[3, 1] == np.float32(50.0)                               # 150 -> fmod(250, 100) = 50
    assert out[4, 1] == np.float32(25.0)                               # 75 - 150 = -75 -> fmod(25, 100)
    assert out[5, 1] == np.float32(0.0)


def test_box_edge_not_wrapped_and_multiple_crossings():
    L = np.float32(64.0)
    pos = np.array([[0.0, 0.0, 32.0], [0.0, 0.0, 10.0]], np.float32)
    vel = np.array([[0.0, 0.0, 2048.0], [0.0, 0.0, 10240.0]], np.float32)   # factor 1/64: +32, +160
    out = rsd.pos_redshift_space(pos, vel, L, 64.0, 0.0, 2)
    assert out[0, 2] == L                        # exactly L: kept (the deposit wraps cell N to 0)
    assert out[1, 2] == np.float32(42.0)         # 170 -> fmod(234, 64) = 42
