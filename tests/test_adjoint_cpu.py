"""Host-side argument checks of cardiax_amd.adjoint.rollout (no GPU needed:
the shape check runs before the device is touched)."""
import numpy as np
import pytest
import torch

from cardiax_amd import adjoint
from cardiax_amd.params import PARAMSET_3 as prm
from oracle import fk_oracle as o


@pytest.mark.parametrize("D_shape", [(20, 16), (15, 20), (16, 21)])
def test_rollout_refuses_a_D_of_another_shape(D_shape):
    st = tuple(np.asarray(x, np.float32) for x in o.init((16, 20)))
    D = torch.full(D_shape, 1e-3)
    with pytest.raises(ValueError, match="shape"):
        adjoint.rollout(D, prm.tau_d, prm.tau_r, prm.tau_si, st, prm, [], 0, 10, 5, 0.01, 0.01)


def test_rollout_refuses_a_transposed_D():
    st = tuple(np.asarray(x, np.float32) for x in o.init((16, 20)))
    D = torch.full((16, 20), 1e-3).t()  # (20, 16): the transposed field
    with pytest.raises(ValueError, match="shape"):
        adjoint.rollout(D, prm.tau_d, prm.tau_r, prm.tau_si, st, prm, [], 0, 10, 5, 0.01, 0.01)
