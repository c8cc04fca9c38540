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


def test_rollout_batch_needs_the_gpu():
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    st = tuple(np.asarray(x, np.float32) for x in o.init((16, 20)))
    with pytest.raises(RuntimeError, match="GPU"):
        adjoint.RolloutBatch([(st, prm, [], 0)], 10, 5, 0.01, 0.01)


def test_checkpoint_interval():
    assert adjoint.checkpoint_interval(2000) == 45  # ceil(sqrt(2000))
    assert adjoint.checkpoint_interval(1000) == 32
    assert adjoint.checkpoint_interval(1) == 1
    assert adjoint.checkpoint_interval(1000, 1000) == 1000


def test_batched_plan_needs_one_schedule():
    """All instances of a batched plan share n_steps and frame_every."""
    from cardiax_amd import inverse

    def inst(n_steps, fe):
        return inverse.Instance(torch.zeros((8, 8)), torch.zeros((1, 8, 8)), None, prm, [],
                                n_steps, fe)

    with pytest.raises(ValueError, match="n_steps and frame_every"):
        inverse._plan([inst(100, 50), inst(200, 50)], 0.01, 0.01)
    with pytest.raises(ValueError, match="n_steps and frame_every"):
        inverse._plan([inst(100, 50), inst(100, 25)], 0.01, 0.01)
