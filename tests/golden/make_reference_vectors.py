"""Writes reference_known_answers.json: the known-answer vectors that heFFTe's own
unit tests hold for the single-rank path, transcribed as data.

Sources (paths relative to the heFFTe 2.4.1 tree):
  * 1D c2c / r2c on the 2x3x4 box with inputs 1..24:
      test/test_units_nompi.cpp:92-97 (input), :136-202 (make_fft0/1/2, *_r2c),
      :209-262 and :336-367 (how they are applied: one executor per axis)
  * 1D DCT-II / DST-II with batch replication: test/test_units_nompi.cpp:288-333
  * 3D DCT-II / DST-II / DCT-I of iota(1..24) on the 2x3x4 world: test/test_cos.cpp:36-54
  * in-node transposes of the 2x3x4 box: test/test_units_nompi.cpp:577-637

Complex numbers are stored as [re, im] pairs.  Run from any directory:
    python tests/golden/make_reference_vectors.py
"""
import json
import math
import os

SQ3 = 1.73205080756888  # the reference writes sqrt(3) with 15 digits


def c(re, im=0.0):
    return [float(re), float(im)]


def make_fft0():
    # 1D transforms of length 2 along axis 0: pairs (2i+1, 2i+2) -> (4i+3, -1)
    r = [None] * 24
    for i in range(0, 24, 2):
        r[i] = c(3 + 2 * i)
        r[i + 1] = c(-1.0)
    return r


def make_fft1():
    r = [None] * 24
    for j in range(4):
        for i in range(2):
            r[6 * j + i] = c((2 * j + i + 1) * 9.0 - i * 6.0)
            r[6 * j + i + 2] = c(-3.0, SQ3)
            r[6 * j + i + 4] = c(-3.0, -SQ3)
    return r


def make_fft1_r2c():
    r = [None] * 16
    for j in range(4):
        for i in range(2):
            r[4 * j + i] = c((2 * j + i + 1) * 9.0 - i * 6.0)
            r[4 * j + i + 2] = c(-3.0, SQ3)
    return r


def make_fft2():
    r = [None] * 24
    for i in range(6):
        r[i] = c(40.0 + 4 * i)
        r[i + 6] = c(-12.0, 12.0)
        r[i + 12] = c(-12.0)
        r[i + 18] = c(-12.0, -12.0)
    return r


def make_fft2_r2c():
    r = [None] * 18
    for i in range(6):
        r[i] = c(40.0 + 4 * i)
        r[i + 6] = c(-12.0, 12.0)
        r[i + 12] = c(-12.0)
    return r


data = {
    "source": "heFFTe 2.4.1 test/test_units_nompi.cpp and test/test_cos.cpp (transcribed known answers)",
    "box_2x3x4": {"low": [0, 0, 0], "high": [1, 2, 3], "input": list(range(1, 25))},
    "c2c_1d": {
        "axis0": make_fft0(),  # test/test_units_nompi.cpp:136-145
        "axis1": make_fft1(),  # :151-161
        "axis2": make_fft2(),  # :180-189
    },
    "r2c_1d": {
        "axis0": make_fft0(),       # :343 (make_fft0 serves both)
        "axis1": make_fft1_r2c(),   # :165-175
        "axis2": make_fft2_r2c(),   # :193-202
        "backward_divisor": [2, 3, 4],  # :364 r /= (2.0 + i)
    },
    "r2r_1d": {
        # box {0,0,0}-{3,1,2}: length 4 lines, 6 of them; box5 {0,0,0}-{4,0,2}: length 5, 3 lines
        "cos_len4": {"input": [1.0, 2.0, 3.0, 4.0], "batch": 6,
                     "output": [2.0000000000000000e+01, -6.3086440597978992e+00, 0.0, -4.4834152916796510e-01]},
        "cos_len5": {"input": [1.0, 2.0, 3.0, 4.0, 5.0], "batch": 3,
                     "output": [30.0, -9.9595931395311208, 0.0, -8.9805595315917053e-01, 0.0]},
        "sin_len4": {"input": [1.0, 2.0, 3.0, 4.0], "batch": 6,
                     "output": [1.3065629648763766e+01, -5.6568542494923806e+00, 5.4119610014619699e+00, -4.0e+00]},
        "sin_len5": {"input": [1.0, 2.0, 3.0, 4.0, 5.0], "batch": 3,
                     "output": [1.9416407864998735e+01, -8.5065080835203979e+00, 7.4164078649987371e+00,
                                -5.2573111211913348e+00, 6.0e+00]},
        # round-trip divisor of the GPU/stock convention: 4 N (test/test_units_nompi.cpp:299-304)
        "roundtrip_factor_per_length": 4,
    },
    "r2r_3d": {
        # test/test_cos.cpp:36-54, world {0,0,0}-{1,2,3}, input iota(1..24)
        "cos": [2.4e+03, -6.7882250993908571e+01, -2.2170250336881628e+02, 0.0, 0.0, 0.0, -9.0844474461089760e+02,
                0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -6.4561180200187039e+01, 0.0, 0.0, 0.0, 0.0,
                0.0],
        "sin": [7.3910362600902943e+02, -4.1810014876044050e+01, -1.0241320258448191e+02, 0.0,
                3.6955181300451477e+02, -2.0905007438022025e+01, -3.8400000000000006e+02, 0.0, 0.0, 0.0,
                -1.9200000000000003e+02, 0.0, 3.0614674589207186e+02, -1.7318275204678301e+01,
                -4.2420937476555700e+01, 0.0, 1.5307337294603599e+02, -8.6591376023391504e+00,
                -2.7152900397563417e+02, 0.0, 0.0, 0.0, -1.3576450198781720e+02, 0.0],
        # test/test_cos.cpp:48-50 (rocfft_cos1 / stock_cos1: DCT-I), same world and input
        "cos1": [600.0, -24.0, -48.0, 0.0, 0.0, 0.0, -192.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
                 0.0, -48.0, 0.0, 0.0, 0.0, 0.0, 0.0],
    },
    "transpose": {
        # test/test_units_nompi.cpp:595-636; source box order (0,1,2), payload 1..24
        "order_120": [1.0, 3.0, 5.0, 7.0, 9.0, 11.0, 13.0, 15.0, 17.0, 19.0, 21.0, 23.0,
                      2.0, 4.0, 6.0, 8.0, 10.0, 12.0, 14.0, 16.0, 18.0, 20.0, 22.0, 24.0],
        "order_210": [1.0, 7.0, 13.0, 19.0, 3.0, 9.0, 15.0, 21.0, 5.0, 11.0, 17.0, 23.0,
                      2.0, 8.0, 14.0, 20.0, 4.0, 10.0, 16.0, 22.0, 6.0, 12.0, 18.0, 24.0],
        "order_021": [1.0, 2.0, 7.0, 8.0, 13.0, 14.0, 19.0, 20.0,
                      3.0, 4.0, 9.0, 10.0, 15.0, 16.0, 21.0, 22.0,
                      5.0, 6.0, 11.0, 12.0, 17.0, 18.0, 23.0, 24.0],
    },
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_known_answers.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1)
    print(out)
