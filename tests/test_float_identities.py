"""Exact float identities the collision solve relies on to skip work
(gravity_amd/csrc/collide.hip, accumulate_contact): for equal materials the
mass share m / (m + m) and the pair friction sqrt(mu * mu) are computed
without the division and the square root. Both identities are checked here
over every float32 of the ranges the kernel admits, in IEEE round-to-nearest
(numpy's float32 division and square root are correctly rounded)."""
import numpy as np


def _floats(first_exponent, last_exponent):
    """Every positive float32 with a biased exponent in [first, last], one
    exponent (2^23 values) at a time."""
    mant = np.arange(1 << 23, dtype=np.uint32)
    for e in range(first_exponent, last_exponent + 1):
        yield ((np.uint32(e) << np.uint32(23)) | mant).view(np.float32)


def test_half_share_for_equal_masses():
    # every normal m below 2^127 (biased exponents 1..253)
    for m in _floats(1, 253):
        assert np.all(m / (m + m) == np.float32(0.5))


def test_square_root_of_a_square_is_the_friction():
    # mu in [2^-63, 2^63): biased exponents 64..189
    for mu in _floats(64, 189):
        assert np.all(np.sqrt(mu * mu) == mu)
    zero = np.float32(0.0)
    assert np.sqrt(zero * zero) == zero
