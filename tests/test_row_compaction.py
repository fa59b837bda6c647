"""The compact PGS row record stores only the angular parts of M^-1 J^T; a row
step rebuilds each linear component as 0 + (1/m) j.  That equals, bit for bit,
the full 6x6 product the reference forms (NMat.h:91-103: a running sum from
+0 over all six columns, ConstraintSolver.cpp:240-272), because the other five
terms of a linear row are (+0) * x = signed zeros.  Checked here in IEEE
double arithmetic (Python floats, no fused multiply-add) over random values
and the signed-zero and static-body (1/m = 0) cases, for the normal row
(J_a linear = -n, J_b = n) and both friction rows (u, w)."""
import math
import random
import struct

import pytest


def bits(x):
    return struct.pack("<d", x)


def full_linear(im, j6, r):
    """Row r < 3 of M^-1 J^T with M^-1 = diag(im, im, im, I): sum from +0 in
    column order, the zero entries multiplying as +0 * j."""
    s = 0.0
    for q in range(6):
        m = (1.0 if q == r else 0.0) * im if q < 3 else 0.0
        s += m * j6[q]
    return s


def compact_linear(im, j6, r):
    return 0.0 + im * j6[r]


def cases():
    rng = random.Random(7)
    specials = [0.0, -0.0, 1.0, -1.0, 1e-300, -1e-300, 5e-324, -5e-324]
    for _ in range(4000):
        n = [rng.uniform(-1, 1) for _ in range(3)]
        ang = [rng.uniform(-3, 3) for _ in range(3)]
        im = rng.choice([0.0, rng.uniform(1e-3, 10.0), 1.0 / rng.uniform(0.1, 50.0)])
        yield im, [-n[0], -n[1], -n[2]] + ang
        yield im, n + ang
    for a in specials:
        for b in specials:
            for im in (0.0, 0.5, 3.0):
                yield im, [a, b, -a, b, -b, a]


@pytest.mark.parametrize("r", [0, 1, 2])
def test_linear_part_rebuild_is_bit_exact(r):
    for im, j6 in cases():
        f, c = full_linear(im, j6, r), compact_linear(im, j6, r)
        assert bits(f) == bits(c), (im, j6, r, f, c)


def test_a_dropped_zero_prefix_would_differ():
    # the +0 start matters: im * j alone keeps -0 where the product gives +0
    im, j6 = 0.0, [-1.0, 0.0, 0.0, 1.0, 1.0, 1.0]
    assert math.copysign(1.0, full_linear(im, j6, 0)) == 1.0
    assert math.copysign(1.0, im * j6[0]) == -1.0
    assert bits(compact_linear(im, j6, 0)) == bits(full_linear(im, j6, 0))
