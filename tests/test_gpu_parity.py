"""GPU parity of the HIP forward path against the FP64 oracle.

Every call goes through the C-ABI (cardiax_amd.solve -> libcardiax_amd.so).
Tolerances: fields within 1e-4 max-abs of the FP64 oracle over 1e3 steps
(scaled by the field magnitude where a case overdrives u far above 1), and
activation times within 1 step; derivatives within 2e-5 relative to their
largest magnitude (fp32 arithmetic against FP64).
"""
import numpy as np
import pytest
import torch

from cardiax_amd import params as P
from cardiax_amd import solve, workloads as wl
from cardiax_amd.stimulus import Protocol
from oracle import c_oracle, fk_oracle as o

pytestmark = pytest.mark.gpu

KERNEL_VARIANTS = ["plain", "tile16", "tile32", "stream8", "stream16", "stream32",
                   "tb64x2", "tb64x4", "tb32x2", "tb32x4", "streamc8", "streamc16", "stream4",
                   "streamc4", "stream2", "streamc2", "pair2", "pair4", "streamt8", "streamt12",
                   "streamt16", "streamt24"]


@pytest.fixture(params=KERNEL_VARIANTS)
def variant(request, gpu):
    """Run the test with one forward kernel variant, then restore the default."""
    from cardiax_amd import _lib

    L = _lib.lib()
    prev = L.fk_set_forward_variant(_lib.VARIANTS[request.param])
    yield request.param
    L.fk_set_forward_variant(prev)

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def npy(t):
    return t.detach().cpu().numpy().astype(np.float64)


def rand_state(shape, seed, excited=True):
    rng = np.random.default_rng(seed)
    if excited:
        return o.State(rng.uniform(0.2, 1.0, shape), rng.uniform(0.2, 1.0, shape),
                       rng.uniform(-0.05, 1.05, shape))
    return o.init(shape)


def f32(st):
    return tuple(np.asarray(x, np.float32) for x in st)


def test_init(gpu):
    st = solve.init((37, 70))
    assert all(x.dtype == torch.float32 and x.is_cuda for x in st)
    assert torch.all(st.v == 1) and torch.all(st.w == 1) and torch.all(st.u == 0)


@pytest.mark.parametrize("shape", [(5, 9), (9, 5), (33, 70), (130, 257)])
@pytest.mark.parametrize("axis", [0, 1])
def test_gradient(gpu, shape, axis):
    if shape[axis] < 5:
        pytest.skip("axis too short")
    a = np.random.default_rng(7).normal(size=shape).astype(np.float32)
    g = npy(solve.gradient(a, axis))
    ref = o.gradient(a.astype(np.float64), axis)
    np.testing.assert_allclose(g, ref, atol=2e-6 * np.abs(a).max() * 10)


def test_gradient_rejects_short_axis(gpu):
    with pytest.raises(ValueError):
        solve.gradient(np.ones((4, 9), np.float32), 0)


def test_stimulate(gpu):
    shape = (20, 30)
    X = np.random.default_rng(1).normal(size=shape).astype(np.float32)
    f1 = np.zeros(shape, np.float32); f1[:, :5] = 2.0
    f2 = np.zeros(shape, np.float32); f2[5:9, 3:20] = -1.5
    stims = [(Protocol(0, 2, 50), f1), (Protocol(1, 2, 1e9), f2)]
    ostims = [o.Stimulus(o.Protocol(*p), f) for p, f in stims]
    for t in [0, 1, 2, 3, 50, 51, 52]:
        got = npy(solve.stimulate(t, X, stims))
        np.testing.assert_array_equal(got, o.stimulate(t, X.astype(np.float64), ostims))


# includes stages that end exactly on the grid edge: H = (ty+1) T + 4K for
# 64-tiles with K = 2 (72, 136) and K = 4 (80), 32-tiles with K = 2 / 4 (40, 48)
SHAPES = [(3, 3), (5, 7), (16, 16), (33, 65), (64, 64), (100, 37), (130, 200), (37, 129),
          (72, 80), (136, 48), (40, 200)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("pname", ["3", "1A", "7"])
def test_step_derivatives(variant, shape, pname):
    prm = P.ALL_PARAMSETS[pname]
    st = f32(rand_state(shape, 11))
    rng = np.random.default_rng(12)
    D = rng.uniform(5e-4, 1e-3, shape).astype(np.float32)
    f = np.zeros(shape, np.float32); f[: max(1, shape[0] // 3)] = 20.0
    stims = [(Protocol(2, 2, 1e9), f)]
    for t in (0.0, 2.0):
        got = solve.step(st, t, prm, D, stims, 0.01)
        ref = o.step(o.State(*[x.astype(np.float64) for x in st]), t, prm,
                     D.astype(np.float64), [o.Stimulus(o.Protocol(2, 2, 1e9), f)], 0.01)
        for name, g, r in zip("vwu", got, ref):
            scale = max(1.0, np.abs(r).max())
            np.testing.assert_allclose(npy(g), r, atol=2e-5 * scale, err_msg=name)


@pytest.mark.parametrize("shape", [(3, 4), (45, 70), (129, 64)])
def test_step_euler(variant, shape):
    st = f32(rand_state(shape, 3))
    D = np.full(shape, 1e-3, np.float32)
    got = solve.step_euler(st, 5, P.PARAMSET_3, D, [], 0.01, 0.01)
    ref = o.step_euler(o.State(*[x.astype(np.float64) for x in st]), 5, P.PARAMSET_3,
                       D.astype(np.float64), [], 0.01, 0.01)
    for g, r in zip(got, ref):
        np.testing.assert_allclose(npy(g), r, atol=2e-6)


@pytest.mark.parametrize("name", ["fk_16_p3_west", "fk_32_p1a_overlap", "fk_48x40_p5_edges",
                                  "fk_64_p7_s1s2", "fk_5x7_p3_tiny"])
def test_golden_fixtures(variant, name):
    z = np.load("%s/%s.npz" % (GOLDEN, name))
    prm = P.Params(*z["params"])
    stims = [(Protocol(*p), f) for p, f in zip(z["protocols"], z["fields"])]
    st = tuple(z["state0"])
    t = 0
    for c, exp in zip(z["checkpoints"], z["expected"]):
        st = solve._forward_euler(st, t, int(c), prm, z["D"], stims, float(z["dt"]),
                                  float(z["dx"]))
        for k, (g, r) in enumerate(zip(st, exp)):
            scale = max(1.0, np.abs(r).max())
            err = np.abs(npy(g) - r).max()
            assert err <= 1e-4 * scale, (name, int(c), "vwu"[k], err)
        t = int(c)


@pytest.mark.parametrize("shape", [(72, 80), (136, 48), (40, 200), (97, 131), (5, 64)])
def test_forward_matches_oracle_odd_shapes(variant, shape):
    """60 Euler steps from a perturbed state with two stimuli on grids whose
    tiles / stages end on or near the edges."""
    H, W = shape
    ii, jj = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    u0 = 0.5 + 0.45 * np.sin(2 * np.pi * ii / max(H, 8) * 1.3) * np.cos(2 * np.pi * jj / W * 0.7)
    st = f32(o.State(0.8 + 0.1 * np.cos(jj / 7.0), 0.9 - 0.05 * np.sin(ii / 5.0), u0))
    rng = np.random.default_rng(22)
    D = rng.uniform(6e-4, 1e-3, shape).astype(np.float32)
    f1 = np.zeros(shape, np.float32); f1[:, :3] = 20.0
    f2 = np.zeros(shape, np.float32); f2[-4:, :] = -2.0
    stims = [(Protocol(0, 2, 1e9), f1), (Protocol(5, 3, 20), f2)]
    got = solve._forward_euler(st, 0, 60, P.PARAMSET_3, D, stims, 0.01, 0.01)
    ref = o.forward_euler(o.State(*[x.astype(np.float64) for x in st]), 0, 60, P.PARAMSET_3,
                          D.astype(np.float64),
                          [o.Stimulus(o.Protocol(*p), f) for p, f in stims], 0.01, 0.01)
    for name, g, r in zip("vwu", got, ref):
        assert np.abs(npy(g) - r).max() <= 1e-4 * max(1.0, np.abs(r).max()), name


def test_plain_and_tile_kernels_agree(gpu):
    shape = (100, 75)
    sim = wl.c1_inputs(3, shape=shape, steps=300)
    stims = [(s.protocol._replace(start=50 if k else 0), s.field)
             for k, s in enumerate(sim.stimuli)]
    st0 = solve.init(shape)
    a = solve._forward_euler(st0, 0, 300, sim.params, sim.D, stims, 0.01, 0.01)
    b = solve._forward_euler_plain(st0, 0, 300, sim.params, sim.D, stims, 0.01, 0.01)
    for x, y in zip(a, b):
        assert float((x - y).abs().max()) <= 1e-5


def _activation(us, theta=0.5):
    """First step (1-based) after which u >= theta; inf where never."""
    act = np.full(us[0].shape, np.inf)
    for k, u in enumerate(us, start=1):
        newly = np.isinf(act) & (u >= theta)
        act[newly] = k
    return act


def test_c1_parity_1000_steps(variant):
    """128^2, PARAMSET_3, S1 + S2, 1000 steps: u within 1e-4 of the FP64
    oracle every 100 steps; activation times within 1 step."""
    sim = wl.c1_inputs(0)
    stims = [(s.protocol, s.field) for s in sim.stimuli]
    ostims = [o.Stimulus(o.Protocol(*s.protocol), s.field) for s in sim.stimuli]
    g = solve.init(sim.shape)
    r = o.init(sim.shape)
    gu, ru = [], []
    worst = 0.0
    for t in range(0, 1000, 10):
        g = solve._forward_euler(g, t, t + 10, sim.params, sim.D, stims, 0.01, 0.01)
        r = c_oracle.forward_euler(r, t, t + 10, sim.params, sim.D, ostims, 0.01, 0.01)
        gu.append(npy(g.u))
        ru.append(r.u)
        if (t + 10) % 100 == 0:
            for name, a, b in zip("vwu", g, r):
                err = float(np.abs(npy(a) - b).max())
                worst = max(worst, err) if name == "u" else worst
                assert err <= 1e-4, (t + 10, name, err)
    ga, ra = _activation(gu), _activation(ru)
    assert np.array_equal(np.isinf(ga), np.isinf(ra))
    fin = np.isfinite(ga)
    assert fin.sum() > 1000  # both waves did propagate
    # sampled every 10 steps: activation times agree to the sample
    assert np.abs(ga[fin] - ra[fin]).max() <= 1
    print("C1 worst u max-abs vs FP64 over 1000 steps: %.3e" % worst)


def test_c1_activation_every_step(gpu):
    """Activation maps at single-step resolution over all 1000 steps of C1:
    the S1 wave and the S2 wave (S2 fires at step 640), first crossings of
    u >= 0.5 within one step of the FP64 oracle's."""
    sim = wl.c1_inputs(0, steps=1000)
    s2 = sim.meta["s2_start"]
    assert s2 < 1000
    stims = [(s.protocol, s.field) for s in sim.stimuli]
    ostims = [o.Stimulus(o.Protocol(*s.protocol), s.field) for s in sim.stimuli]
    g, r = solve.init(sim.shape), o.init(sim.shape)
    ga = np.full(sim.shape, np.inf)
    ra = np.full(sim.shape, np.inf)
    for t in range(1000):
        g = solve._forward_euler(g, t, t + 1, sim.params, sim.D, stims, 0.01, 0.01)
        r = c_oracle.forward_euler(r, t, t + 1, sim.params, sim.D, ostims, 0.01, 0.01)
        for act, u in ((ga, npy(g.u)), (ra, r.u)):
            act[np.isinf(act) & (u >= 0.5)] = t + 1
    assert np.array_equal(np.isinf(ga), np.isinf(ra))
    fin = np.isfinite(ga)
    assert fin.sum() > 500
    assert (ga[fin] > s2).sum() > 500  # cells first activated by the S2 wave
    assert np.abs(ga[fin] - ra[fin]).max() <= 1
    print("C1 per-step activation: %d cells, %d after S2 (step %d), max |diff| %d steps"
          % (fin.sum(), (ga[fin] > s2).sum(), s2, np.abs(ga[fin] - ra[fin]).max()))


def test_forward_checkpoints_and_composition(gpu):
    shape = (40, 50)
    sim = wl.c1_inputs(1, shape=shape)
    stims = [(s.protocol._replace(start=20 if k else 0), s.field)
             for k, s in enumerate(sim.stimuli)]
    st = solve.init(shape)
    states = solve.forward(st, [0, 30, 70, 100], sim.params, sim.D, stims, 0.01, 0.01)
    assert len(states) == 3
    whole = solve._forward_euler(st, 0, 100, sim.params, sim.D, stims, 0.01, 0.01)
    for a, b in zip(states[-1], whole):
        assert torch.equal(a, b)  # deterministic: same kernels, same order
    seg = solve._forward_euler(st, 0, 30, sim.params, sim.D, stims, 0.01, 0.01)
    for a, b in zip(states[0], seg):
        assert torch.equal(a, b)
    # a custom integrator plugs in like the reference's
    calls = []

    def integ(*args):
        calls.append(args[1:3])
        return solve._forward_euler(*args)

    solve.forward(st, [0, 5, 9], sim.params, sim.D, stims, 0.01, 0.01, integrator=integ)
    assert calls == [(0.0, 5.0), (5.0, 9.0)]


def test_inputs_are_not_modified_and_zero_steps(gpu):
    shape = (20, 20)
    st = tuple(torch.from_numpy(x.astype(np.float32)).cuda() for x in rand_state(shape, 5))
    keep = tuple(x.clone() for x in st)
    D = torch.full(shape, 1e-3, device="cuda")
    out = solve._forward_euler(st, 0, 7, P.PARAMSET_3, D, [], 0.01, 0.01)
    for a, b in zip(st, keep):
        assert torch.equal(a, b)
    same = solve._forward_euler(st, 4, 4, P.PARAMSET_3, D, [], 0.01, 0.01)
    for a, b in zip(same, st):
        assert torch.equal(a, b) and a.data_ptr() != b.data_ptr()
    one = solve._forward_euler(st, 0, 1, P.PARAMSET_3, D, [], 0.01, 0.01)
    ref = solve.step_euler(st, 0, P.PARAMSET_3, D, [], 0.01, 0.01)
    for a, b in zip(one, ref):
        assert torch.equal(a, b)
    assert not torch.equal(out.u, st[2])


def test_argument_errors(gpu):
    st = solve.init((10, 12))
    with pytest.raises(ValueError):
        solve._forward_euler(st, 0, 5, P.PARAMSET_3, np.ones((10, 11), np.float32), [],
                             0.01, 0.01)
    with pytest.raises(ValueError):
        solve._forward_euler(st, 0, 5, P.PARAMSET_3, np.ones((10, 12), np.float32),
                             [(Protocol(0, 2, 1e9), np.ones((3, 3), np.float32))], 0.01, 0.01)
    many = [(Protocol(0, 2, 1e9), np.ones((10, 12), np.float32))] * 33
    with pytest.raises(NotImplementedError):
        solve._forward_euler(st, 0, 5, P.PARAMSET_3, np.ones((10, 12), np.float32), many,
                             0.01, 0.01)
    small = solve.init((2, 5))  # like jnp.ones, init accepts any shape ...
    with pytest.raises(ValueError, match="too small"):  # ... the solver needs >= 3 x 3
        solve._forward_euler(small, 0, 1, P.PARAMSET_3, np.ones((2, 5), np.float32), [],
                             0.01, 0.01)


def test_full_size_mirror_symmetry_1024(variant):
    """1024^2 scarred C2 inputs, 500 steps: solving the left-right mirrored
    problem gives the mirrored solution (size-independent property), to the
    parity bar: fp32 rounding of the two runs differs at the 1e-7 level (the
    one-sided edge rules and their mirror images round differently), and
    500 steps of excitable fronts amplify that to ~1e-5; an index or
    orientation mistake gives O(1) differences."""
    sim = wl.c2_inputs(0, steps=500)
    stims = [(s.protocol._replace(start=0 if k == 0 else 100), s.field)
             for k, s in enumerate(sim.stimuli)]
    mir = [(p, np.ascontiguousarray(f[:, ::-1])) for p, f in stims]
    st = solve.init(sim.shape)
    a = solve._forward_euler(st, 0, 500, sim.params, sim.D, stims, 0.01, 0.01)
    b = solve._forward_euler(st, 0, 500, sim.params, np.ascontiguousarray(sim.D[:, ::-1]),
                             mir, 0.01, 0.01)
    for x, y in zip(a, b):
        assert float((x - torch.flip(y, dims=[1])).abs().max()) <= 1e-4
    assert float(a.u.max()) > 0.5  # the S1 wave is under way


def test_full_size_planar_wave_row_invariance_1024(gpu):
    """Uniform D with a WEST stripe: every row of the 1024^2 solution is the
    same 1-D solution. The one-sided rules applied to a constant column round
    to ~1e-7 instead of 0, so the 3-4 edge rows drift first and the excitable
    front carries that inwards over time; until it arrives the bulk rows are
    bitwise identical, which rules out any row-dependent arithmetic."""
    shape = (1024, 1024)
    f = np.zeros(shape, np.float32); f[:, :50] = 20.0
    D = np.full(shape, 1e-3, np.float32)
    early = solve._forward_euler(solve.init(shape), 0, 10, P.PARAMSET_3, D,
                                 [(Protocol(0, 2, 1e9), f)], 0.01, 0.01)
    for x in early:
        assert torch.equal(x[8:1016], x[512:513].expand(1008, -1))
    out = solve._forward_euler(early, 10, 1500, P.PARAMSET_3, D,
                               [(Protocol(0, 2, 1e9), f)], 0.01, 0.01)
    for x in out:
        assert float((x - x[512:513]).abs().max()) <= 1e-4
    assert float(out.u.max()) > 0.5
