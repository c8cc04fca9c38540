"""Unit helpers on the forward path.

Mirrors the helpers of cardiax/convert.py that callers of the solver use:
``realsize_to_shape`` (cardiax/convert.py:29-30) and ``ms_to_units``
(cardiax/convert.py:45-46), both truncating with ``int()``, plus their inverses
and the linear diffusivity rescale used by the D-map generators.
"""


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


def u_to_V(u, V0=-85, Vfi=15):
    """cardiax/convert.py:62-63 -- dimensionless u to millivolts."""
    return ((Vfi - V0) * u) + V0


def V_to_u(V, V0=-85, Vfi=15):
    """cardiax/convert.py:66-67."""
    return (V - V0) / (Vfi - V0)
