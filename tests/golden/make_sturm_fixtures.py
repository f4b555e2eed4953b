"""Write golden Sturm-solver fixtures from the reference's own root finder.

Runs in the build container only (needs /root/reference): compiles the
reference's std-only header inc/maths/sturm_polynomial_solver.h into
oracle/_ref/libref_sturm.so (`make -C oracle ref`) and records, for each
degree-10 polynomial, the roots bisect_sturm<10> returns.

Polynomials (ascending coefficients, 11 each):
  real_*    -- products of (z - r_i) with 10 real roots (some clustered);
  mixed_*   -- real roots times quadratic factors with complex roots;
  random_*  -- i.i.d. normal coefficients (any root structure);
  scaled_*  -- mixed polynomials scaled by 1e-8 .. 1e8.
Output: tests/golden/sturm_fixtures.npz (coeffs [K, 11], roots [K, 10],
counts [K]).
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    L = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_sturm.so"))
    L.ref_bisect_sturm10.restype = ctypes.c_int
    rng = np.random.default_rng(20261202)
    polys = []
    for _ in range(40):
        r = rng.uniform(-5, 5, 10)
        if rng.uniform() < 0.3:
            r[:3] = r[0] + rng.normal(0, 1e-3, 3)
        polys.append(np.poly(r)[::-1] * rng.uniform(0.5, 2.0))
    for _ in range(40):
        k = int(rng.integers(0, 5))  # quadratic factors with complex roots
        real = rng.uniform(-3, 3, 10 - 2 * k)
        p = np.poly(real)
        for _ in range(k):
            a, b = rng.uniform(-2, 2), rng.uniform(0.1, 2)
            p = np.polymul(p, [1.0, -2 * a, a * a + b * b])
        polys.append(p[::-1] * rng.normal())
    for _ in range(60):
        polys.append(rng.normal(size=11))
    for s in (1e-8, 1e-4, 1e4, 1e8):
        for _ in range(5):
            polys.append(np.poly(rng.uniform(-2, 2, 10))[::-1] * s)
    coeffs = np.ascontiguousarray(np.array(polys), np.float64)
    roots = np.zeros((len(polys), 10))
    counts = np.zeros(len(polys), np.int32)
    for i, c in enumerate(coeffs):
        out = np.zeros(10)
        counts[i] = L.ref_bisect_sturm10(c.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
        roots[i] = out
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "sturm_fixtures.npz"), coeffs=coeffs, roots=roots,
                        counts=counts)
    print("wrote %d polynomials, %d roots" % (len(polys), counts.sum()))


if __name__ == "__main__":
    main()
