"""The device's single-precision sine/cosine (csrc/hip/libm_sincosf.h) against
the host C library, which the reference's std::sin/std::cos(float) calls reach
(model/matrixFunctions.h:51-60).  CPU only: the header compiled for the host
(tests/cpp/trig_host.cpp); the device build is checked in test_gpu_parity.py.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "cpp", "_bin", "libtrig_host.so")


def _lib():
    L = ctypes.CDLL(LIB)
    L.trig_host_mismatches.restype = ctypes.c_uint64
    L.trig_host_mismatches.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
    L.trig_host_eval.restype = None
    L.trig_host_eval.argtypes = [ctypes.c_void_p, ctypes.c_size_t] + [ctypes.c_void_p] * 4
    return L


def test_every_finite_float_matches_the_c_library():
    """All 2 x 2^31 - 2^24 finite floats (both signs), bit for bit, against
    sincosf, sinf and cosf."""
    L = _lib()
    assert L.trig_host_mismatches(0, 0x7F800000, 1, 8) == 0


def test_special_values():
    L = _lib()
    x = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, 2.0 ** -12, 0.75, np.pi / 4,
                  np.pi / 2, 119.99, 120.0, 3e38], np.float32)
    out = [np.zeros_like(x) for _ in range(4)]
    L.trig_host_eval(x.ctypes.data, x.size, *[o.ctypes.data for o in out])
    s_lib, c_lib, s_own, c_own = out
    assert np.array_equal(s_lib.view(np.uint32)[~np.isnan(s_lib)], s_own.view(np.uint32)[~np.isnan(s_lib)])
    assert np.array_equal(np.isnan(s_lib), np.isnan(s_own))
    assert np.array_equal(np.isnan(c_lib), np.isnan(c_own))
    assert np.array_equal(c_lib[~np.isnan(c_lib)], c_own[~np.isnan(c_lib)])
    assert np.signbit(s_own[1])  # sin(-0) = -0


def _pi_bits(nbits):
    """floor(pi * 2^nbits) from Machin's formula in integer arithmetic."""
    guard = 64
    one = 1 << (nbits + guard)

    def arctan_inv(k):
        s, term, n, sign = 0, one // k, 1, 1
        while term:
            s += sign * (term // n)
            term //= k * k
            n += 2
            sign = -sign
        return s
    return (16 * arctan_inv(5) - 4 * arctan_inv(239)) >> guard


def test_two_over_pi_words_are_the_bits_of_two_over_pi():
    """Word k of the Payne-Hanek table is floor(2/pi * 2^(8k+8)) mod 2^32."""
    src = open(os.path.join(ROOT, "gpuattract_2.0_amd", "csrc", "hip", "libm_sincosf.h")).read()
    body = src[src.index("kTwoOverPiWords[24] = {") + len("kTwoOverPiWords[24] = {"):]
    words = [int(w, 16) for w in body[:body.index("}")].replace("\n", " ").split(",")]
    assert len(words) == 24
    P = 400
    pi = _pi_bits(P)
    for k, w in enumerate(words):
        e = 8 * k + 8
        # floor(2/pi * 2^e) = floor(2^(e+1+P) / (pi * 2^P)); exact unless the
        # quotient is within one unit of an integer (checked by both bounds)
        lo = (1 << (e + 1 + P)) // (pi + 1)
        hi = (1 << (e + 1 + P)) // pi
        assert lo == hi
        assert w == hi % (1 << 32), k
