"""The WENO5BND kernels share a divisor's refined reciprocal among several
numerators (hg_div.h) where the operands lie in the ranges in which the
hardware's division steps neither rescale nor fix up; elsewhere they divide
in full.  On the device, over generated operands that cover every exponent
(zeros, denormals and values outside those ranges included), both sides of
the ranges' edges and structured mantissas, the shortcut must give the bits
of the hardware's own x / d: for three numerators of one divisor, for 1 / d,
and for four signed numerators of 12 (zero quotients compared without their
sign, which the reconstruction never carries)."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [1, 2])
def test_shared_reciprocal_division_is_the_hardware_division(seed):
    from harm_finalwham_amd import harmgpu
    n = 1 << 24
    print("case division self-check seed", seed, "n", n, flush=True)
    shared, recip, const12, fast = harmgpu.check_division(n, seed)
    print("mismatches", shared, recip, const12, "operand sets on the shortcut", fast, flush=True)
    assert (shared, recip, const12) == (0, 0, 0)
    # both paths are exercised: a good share of the sets inside the ranges,
    # a good share outside
    assert n // 10 < fast < n - n // 10
