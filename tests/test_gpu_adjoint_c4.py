"""The adjoint on the 64x64x32 BuenoOrovio workload over its whole 5000-step (50 ms)
horizon against the oracle's tangent-linear run: the loss and the derivative along six
random conductivity-scale directions and tau_fi, tau_si (0.65 G cell-steps of tangent
arithmetic on the host, about a minute)."""
import pytest

from tests.test_gpu_adjoint import _say, _tangent_check

pytestmark = pytest.mark.gpu


def test_adjoint_c4_full_horizon_vs_tangent_linear(gpu, orc):
    _say("adjoint 64x64x32, 5000 steps, vs tangent-linear oracle")
    lrel, drel = _tangent_check(gpu, orc, (64, 64, 32), 2e-4, 5000, 256, [1.5e-3] * 3, 6,
                                20240604, 100)
    assert lrel <= 1e-10
    assert drel <= 1e-7
