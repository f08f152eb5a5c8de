"""Real solid-harmonic and concomitant-field basis functions (oracle).

Restates ``computeHarmonics/basisFuncHarm.m:24-96`` (indices 1..36, 0th to 5th
order) and ``computeHarmonics/basisFuncConc.m:13-24`` (indices 1..4).  Index
``n`` is 1-based like the reference.

Pinned without retyping the formulas (``tests/test_oracle_highorder.py``):
every harmonic basis function is a homogeneous polynomial of the stated order
(``b(l*r) = l^order b(r)``), every one of order >= 2 is harmonic (its
Laplacian vanishes; checked by finite differences), the 16 (and 36) functions
are linearly independent, and the known-answer DFT test of
``@sampHighOrder/tests.m:265-283`` pins indices 2 and 3 to ``x`` and ``y``.
"""
import numpy as np

# polynomial order of each harmonic index (n = 1..36)
HARM_ORDER = np.array([0] + [1] * 3 + [2] * 5 + [3] * 7 + [4] * 9 + [5] * 11)


def basis_harm(x, y, z, n):
    """``basisFuncHarm(x, y, z, n)`` for a scalar 1-based index ``n``."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    r2 = x ** 2 + y ** 2 + z ** 2
    if n == 1:
        return np.ones_like(x)
    if n == 2:
        return x
    if n == 3:
        return y
    if n == 4:
        return z
    if n == 5:
        return x * y
    if n == 6:
        return z * y
    if n == 7:
        return 3 * z ** 2 - r2
    if n == 8:
        return x * z
    if n == 9:
        return x ** 2 - y ** 2
    if n == 10:
        return 3 * y * x ** 2 - y ** 3
    if n == 11:
        return x * z * y
    if n == 12:
        return (5 * z ** 2 - r2) * y
    if n == 13:
        return 5 * z ** 3 - 3 * z * r2
    if n == 14:
        return (5 * z ** 2 - r2) * x
    if n == 15:
        return x ** 2 * z - y ** 2 * z
    if n == 16:
        return x ** 3 - 3 * x * y ** 2
    if n == 17:
        return (x * y) * (x ** 2 - y ** 2)
    if n == 18:
        return (y * z) * (3 * x ** 2 - y ** 2)
    if n == 19:
        return (x * y) * (7 * z ** 2 - r2)
    if n == 20:
        return y * (7 * z ** 3 - 3 * z * r2)
    if n == 21:
        return 35 * z ** 4 - 30 * z ** 2 * r2 + 3 * r2 ** 2
    if n == 22:
        return x * (7 * z ** 3 - 3 * z * r2)
    if n == 23:
        return (x ** 2 - y ** 2) * (7 * z ** 2 - r2)
    if n == 24:
        return x * (x ** 2 - 3 * y ** 2) * z
    if n == 25:
        return x ** 2 * (x ** 2 - 3 * y ** 2) - y ** 2 * (3 * x ** 2 - y ** 2)
    if n == 26:
        return y * (5 * x ** 4 - 10 * x ** 2 * y ** 2 + y ** 4)
    if n == 27:
        return x * y * z * (x ** 2 - y ** 2)
    if n == 28:
        return y * (r2 - 9 * z ** 2) * (y ** 2 - 3 * x ** 2)
    if n == 29:
        return x * y * z * (3 * z ** 2 - r2)
    if n == 30:
        return y * (r2 ** 2 - 14 * r2 * z ** 2 + 21 * z ** 4)
    if n == 31:
        return 63 * z ** 5 - 70 * z ** 3 * r2 + 15 * z * r2 ** 2
    if n == 32:
        return x * (r2 ** 2 - 14 * r2 * z ** 2 + 21 * z ** 4)
    if n == 33:
        return z * (3 * z ** 2 - r2) * (x ** 2 - y ** 2)
    if n == 34:
        return x * (r2 - 9 * z ** 2) * (3 * y ** 2 - x ** 2)
    if n == 35:
        return z * (x ** 4 - 6 * x ** 2 * y ** 2 + y ** 4)
    if n == 36:
        return x ** 5 - 10 * x ** 3 * y ** 2 + 5 * x * y ** 4
    raise ValueError("Basis function undefined.")


def basis_conc(x, y, z, n):
    """``basisFuncConc(x, y, z, n)`` for a scalar 1-based index ``n``."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    if n == 1:
        return z * z
    if n == 2:
        return x * x + y * y
    if n == 3:
        return x * z
    if n == 4:
        return y * z
    raise ValueError("Basis function undefined.")
