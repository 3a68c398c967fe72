"""Known-answer data taken from the reference's own tests (values only).

tests.cpp:20-88 (small products), :581-597 (FFT vector), :623-642 (FWHT
vector), :939-986 / :1147-1194 (16x16 cases), test.py:23-48, :53-64,
:122-133 (boundary test tables).
"""
import numpy as np

from oracle import oracle as orc


def _aligned(a, alignment=64):
    a = np.asarray(a, np.float64)
    buf = np.empty(a.nbytes + alignment, np.uint8)
    off = (-buf.ctypes.data) % alignment
    out = buf[off:off + a.nbytes].view(np.float64).reshape(a.shape)
    out[...] = a
    return out


def aligned_empty(shape, dtype=np.float64, alignment=64):
    nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
    buf = np.empty(nbytes + alignment, np.uint8)
    off = (-buf.ctypes.data) % alignment
    return buf[off:off + nbytes].view(dtype).reshape(shape)


AT_4x4 = _aligned(np.arange(1, 17, dtype=float).reshape(4, 4).T)
B_4x4 = _aligned(np.arange(17, 33, dtype=float).reshape(4, 4))
AT_2x4 = _aligned(np.arange(1, 9, dtype=float).reshape(4, 2).T)
B_2x4 = _aligned(np.arange(9, 17, dtype=float).reshape(2, 4))
AT_4x2 = _aligned(np.arange(1, 9, dtype=float).reshape(2, 4).T)
B_4x2 = _aligned(np.arange(9, 17, dtype=float).reshape(4, 2))
B_4x8 = _aligned(np.arange(1, 33, dtype=float).reshape(4, 8))

C_4x4x4 = np.array([[250, 260, 270, 280], [618, 644, 670, 696],
                    [986, 1028, 1070, 1112], [1354, 1412, 1470, 1528]], float)
C_4x2x4 = np.array([[35, 38, 41, 44], [79, 86, 93, 100],
                    [123, 134, 145, 156], [167, 182, 197, 212]], float)
C_2x4x2 = np.array([[130, 140], [322, 348]], float)
C_2x4x8 = np.array([[170, 180, 190, 200, 210, 220, 230, 240],
                    [378, 404, 430, 456, 482, 508, 534, 560]], float)

# tests.cpp uses B_4x8 = 1..32 with AT_4x2; test.py uses 9..40 for its own
# 2x4x8 case (test.py:47-48).
B_4x8_PY = _aligned(np.arange(9, 41, dtype=float).reshape(4, 8))
C_2x4x8_PY = AT_4x2.T @ B_4x8_PY

SMALL_PRODUCTS = [(AT_4x4, B_4x4, C_4x4x4), (AT_2x4, B_2x4, C_4x2x4),
                  (AT_4x2, B_4x2, C_2x4x2), (AT_4x2, B_4x8, C_2x4x8)]
PY_PRODUCTS = [(AT_4x4, B_4x4, C_4x4x4), (AT_4x2, B_4x2, C_2x4x2),
               (AT_2x4, B_2x4, C_4x2x4), (AT_4x2, B_4x8_PY, C_2x4x8_PY)]

FFT_X = np.array([
    0.91527326, 0.07064576, 0.53040597, 0.70546434, 0.36414312, 0.14137265,
    0.50579731, 0.55227775, 0.37729946, 0.95291381, 0.09620142, 0.72017424,
    0.75809317, 0.90013873, 0.46557539, 0.61504573, 0.58568483, 0.25007278,
    0.79962251, 0.22426499, 0.07811993, 0.42820229, 0.72256178, 0.99788856,
    0.1011867, 0.93158399, 0.6872351, 0.95747717, 0.14232377, 0.04623507,
    0.1833904, 0.45529912])
FFT_X_CORRECT = np.array([
    16.2619711 + 0.j, -1.03093929 - 0.05840623j,
    -1.03078571 + 1.11588142j, 0.10465414 - 2.23170224j,
    1.13343466 - 0.7142413j, 0.89874511 - 0.48200699j,
    0.60648699 - 0.68331135j, 1.02030708 + 2.99294096j,
    -0.66866564 + 1.50672682j, -0.09505786 - 0.3173026j,
    1.48484497 + 1.82538006j, 1.3403471 - 0.56288428j,
    0.14009386 - 0.24196106j, -0.0927295 - 0.86637594j,
    3.02934147 + 1.79195727j, 0.49138066 + 0.28645587j,
    -1.63614286 + 0.j])

FWHT_X = np.array([
    0.77956498, 0.24374313, 0.66121831, 0.89375154, 0.95073372, 0.11784131,
    0.50483831, 0.44331508, 0.41599355, 0.5194395, 0.26875662, 0.87481817,
    0.9575786, 0.33804587, 0.49530665, 0.73303024, 0.42873154, 0.87784849,
    0.19554847, 0.18266582, 0.78896532, 0.05496766, 0.40847289, 0.30497385,
    0.67434276, 0.70394152, 0.45398693, 0.02708591, 0.01455404, 0.69647227,
    0.69386635, 0.45950925])
FWHT_X_CORRECT = np.array([
    1.61639087e+01, 1.22100944e+00, 9.61619873e-01, 1.69532008e+00,
    2.38965842e-01, -2.11131125e+00, 1.20992754e+00, -9.90377533e-01,
    -4.89547794e-01, 1.97692390e+00, 3.33603927e-01, 1.72112646e+00,
    1.13896243e+00, -1.61840476e+00, -9.18335302e-01, -1.20064715e+00,
    2.23204248e+00, 5.19002383e-01, -2.06580837e+00, 4.10387227e+00,
    -5.77377406e-03, -1.25357604e+00, -3.06495295e+00, 2.75067875e-01,
    4.73622156e-01, 1.07388134e+00, -8.74375765e-01, -1.36142179e+00,
    8.74043671e-01, 2.61878400e+00, 1.65027151e-01, 1.90390037e+00])

# (family, d, b, seed) of tests.cpp:977-981 (FFT) and :1185-1189 (FWHT)
CASES_16x16 = {
    "fft": [(0, 5, 1024, 15955), (1, 5, 1024, 15956), (2, 11, 2048, 15957),
            (3, 11, 1024, 15958)],
    "fwht": [(0, 5, 2048, 15955), (1, 5, 2048, 15956), (2, 11, 2048, 15957),
             (3, 11, 1024, 15958)],
}


def drng_16x16():
    """tests.cpp:959-963: AT[i], B[i] drawn alternately from drng(15954)."""
    s = orc.drng_stream(15954, 512)
    return _aligned(s[0::2].reshape(16, 16)), _aligned(s[1::2].reshape(16, 16))


HASHES = [None, "multiply shift", "multiply add shift", "tabulation8", "tabulation16"]

TEST_PY_PARAMS = {
    ("fft", None): (1, 256), ("fft", "multiply shift"): (1, 128),
    ("fft", "multiply add shift"): (1, 256), ("fft", "tabulation8"): (1, 128),
    ("fft", "tabulation16"): (1, 64),
    ("fwht", None): (1, 32), ("fwht", "multiply shift"): (1, 64),
    ("fwht", "multiply add shift"): (1, 32), ("fwht", "tabulation8"): (1, 64),
    ("fwht", "tabulation16"): (1, 64),
}
TEST_PY_PARAMS_16 = {
    ("fft", None): (7, 1024), ("fft", "multiply shift"): (5, 1024),
    ("fft", "multiply add shift"): (7, 1024), ("fft", "tabulation8"): (9, 2048),
    ("fft", "tabulation16"): (9, 1024),
    ("fwht", None): (11, 1024), ("fwht", "multiply shift"): (9, 1024),
    ("fwht", "multiply add shift"): (11, 1024), ("fwht", "tabulation8"): (9, 1024),
    ("fwht", "tabulation16"): (11, 2048),
}


def int_16x16():
    """test.py:135-138: integers in [-10,10] from default_rng(27840)."""
    rng = np.random.default_rng(27840)
    A = _aligned(rng.integers(-10, 11, (16, 16)).astype(np.float64))
    B = _aligned(rng.integers(-10, 11, (16, 16)).astype(np.float64))
    return _aligned(np.ascontiguousarray(A.T)), B, A @ B
