"""Kernel handles (CPU; the library loads without a GPU): level bounds and
memory planning as the reference's tests/test_kernel_management.py:20-127
check them."""
import pytest

from gpucsl_amd.pc.kernel_management import CompactKernel, DiscreteKernel, GaussianKernel, Kernels


def test_handles_resolve_prebuilt_entry_points():
    g, d, c = GaussianKernel(6, 1, 1), DiscreteKernel(1, 1, 1, 1, memory_restriction=1 << 30), CompactKernel(1, 1)
    for fn in (g.get_function_for_level(0), g.get_function_for_level(1), c.get_function_for_level(0),
               d.get_function_for_level(0), d.get_function_for_level(1)):
        assert fn is not None


def test_cannot_execute_greater_ci_kernel_than_max_level():
    with pytest.raises(AssertionError):
        GaussianKernel(6, 1, 1).get_function_for_level(2)
    with pytest.raises(AssertionError):
        DiscreteKernel(1, 1, 1, 1, memory_restriction=1 << 30).get_function_for_level(2)
    with pytest.raises(AssertionError):
        GaussianKernel(6, 1, 1)(2)


@pytest.mark.parametrize("memory_restriction", [100_000_000, 500_000_000])
def test_discrete_kernel_configuration_is_valid_for_memory_restriction(memory_restriction):
    k = DiscreteKernel(4, 10, 10, 100, memory_restriction=memory_restriction)
    assert k.max_memory_size < memory_restriction


def test_discrete_kernel_fails_if_not_enough_memory():
    # one contingency table alone would hold 100^100 cells
    with pytest.raises(RuntimeError):
        DiscreteKernel(100, 100, 100, 100, memory_restriction=1 << 40)
    with pytest.raises(RuntimeError):
        DiscreteKernel(4, 10, 10, 100, memory_restriction=1000)


def test_kernels_factories():
    k = Kernels.for_gaussian_ci(20, 1, 3)
    assert k.ci_test.max_level == 3 and k.compact.variable_count == 20
    k = Kernels.for_discrete_ci(2, 20, 4, 1000, memory_restriction=1 << 30)
    assert k.ci_test.max_memory_size > 0
