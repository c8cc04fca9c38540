"""Unit helpers on the forward path.

Mirrors every helper of cardiax/convert.py: ``realsize_to_shape``
(cardiax/convert.py:29-30) and ``ms_to_units`` (:45-46), both truncating with
``int()``, their inverses, the linear diffusivity rescale used by the D-map
generators, the voltage maps, and the dict-based unit conversions of
parameters and stimuli (which, as in the reference, change their argument in
place).
"""


def params_to_units(params, dx, dt):
    """cardiax/convert.py:5-16 -- a dict of dimensional parameters to grid
    units, in place: Cm divided by dx^2 / dt, every tau divided by dt.
    Returns the same dict."""
    params["Cm"] /= (dx ** 2) / dt
    for k in ("tau_d", "tau_v1_minus", "tau_v2_minus", "tau_v_plus", "tau_0", "tau_r",
              "tau_si", "tau_w_minus", "tau_w_plus"):
        params[k] /= dt
    return params


def diffusivity_to_units(d, dt):
    """cardiax/convert.py:19-20."""
    return d / dt


def realsize_to_shape(field, dx):
    """cardiax/convert.py:29-30 -- tissue size (cm) to grid shape, truncated."""
    return (int(field[0] / dx), int(field[1] / dx))


def shape_to_realsize(field, dx):
    """cardiax/convert.py:33-34."""
    return (int(field[0] * dx), int(field[1] * dx))


def cm_to_units(value, dx):
    """cardiax/convert.py:37-38."""
    return int(value / dx)


def units_to_cm(value, dx):
    """cardiax/convert.py:41-42."""
    return value * dx


def ms_to_units(value, dt):
    """cardiax/convert.py:45-46 -- milliseconds to steps, truncated."""
    return int(value / dt)


def units_to_ms(value, dt):
    """cardiax/convert.py:49-50."""
    return value * dt


def diffusivity_rescale(c, domain):
    """cardiax/convert.py:23-26 -- affine map of [min, max] onto ``domain``."""
    a, b = c.min(), c.max()
    y, z = domain[0], domain[1]
    return (c - a) * (z - y) / (b - a) + y


def stimuli_to_units(stimuli, dx, dt):
    """cardiax/convert.py:53-59 -- a list of dict stimuli with start,
    duration and period in ms to steps (truncated), in place."""
    stimuli = list(stimuli)
    for i in range(len(stimuli)):
        for k in ("start", "duration", "period"):
            stimuli[i][k] = ms_to_units(stimuli[i][k], dt)
    return stimuli


def u_to_V(u, V0=-85, Vfi=15):
    """cardiax/convert.py:62-63 -- dimensionless u to millivolts."""
    return ((Vfi - V0) * u) + V0


def V_to_u(V, V0=-85, Vfi=15):
    """cardiax/convert.py:66-67."""
    return (V - V0) / (Vfi - V0)
