"""Split time integrators (vlapy/core/vlasov_poisson.py).

Compositions of the pusher and field closures, in the reference's order and
with its coefficients (the coefficients are the published method constants
and must be identical).  The closures may be called with numpy arrays or with
device tensors; the composition does not care.
"""


def get_full_leapfrog_step(vdfdx, edfdv, field_solve, dt, driver_function):
    """Strang split kick(e, dt/2), drift(dt), field(t + dt), kick(E, dt/2)
    (vlasov_poisson.py:24-60; the code order is v-x-v)."""

    def full_leapfrog_ps_step(e, f, t):
        f = edfdv(f=f, e=e, dt=0.5 * dt)
        f = vdfdx(f=f, dt=dt)
        e = field_solve(driver_field=driver_function(t + dt), f=f)
        f = edfdv(f=f, e=e, dt=0.5 * dt)
        return e, f

    return full_leapfrog_ps_step


# Omelyan, Mryglod & Folk (2002) position-extended Forest-Ruth-like constants
_XSI = 0.1786178958448091
_LAMBDA = -0.2123418310626054
_CHI = -0.6626458266981849e-1


def get_full_pefrl_step(vdfdx, edfdv, field_solve, dt, driver_function):
    """4th-order PEFRL: five drifts, each followed by a field solve at the
    cumulative drift time, and four kicks (vlasov_poisson.py:63-154)."""
    drift_w = (_XSI, _CHI, 1.0 - 2.0 * (_CHI + _XSI), _CHI, _XSI)
    kick_w = (0.5 * (1.0 - 2.0 * _LAMBDA), _LAMBDA, _LAMBDA, 0.5 * (1.0 - 2.0 * _LAMBDA))

    def full_pefrl_ps_step(e, f, t):
        tc = t
        for i in range(5):
            ddt = drift_w[i] * dt
            f = vdfdx(f, ddt)
            tc = tc + ddt
            e = field_solve(driver_function(tc), f=f)
            if i < 4:
                f = edfdv(f, e, kick_w[i] * dt)
        return e, f

    return full_pefrl_ps_step


# Casas, Crouseilles, Faou & Mehrenberger (2017) 6th-order constants
_H6 = (0.168735950563437422448196, 0.377851589220928303880766, -0.093175079568731452657924,
       0.049086460976116245491441, 0.264177609888976700200146, 0.186735929134907054308413,
       -0.000069728715055305084099, -0.000625704827430047189169, -0.002213085124045325561636,
       -2.916600457689847816445691e-6, 3.048480261700038788680723e-5, 4.985549387875068121593988e-7)


def get_6th_order_integrator(vdfdx, edfdv, field_solve, dt, driver_function):
    """6th-order splitting: kicks D1 D2 D3 D3 D2 D1 around drifts a1 a2 a3 a2 a1,
    the driver taken at t + a_k dt (vlasov_poisson.py:157-232)."""
    a1, a2, a3, b1, b2, b3, c1, c2, c3, d2, d3, e3 = _H6

    def sixth_order_step(e, f, t):
        k1 = b1 + 2.0 * c1 * dt ** 2.0
        k2 = b2 + 2.0 * c2 * dt ** 2.0 + 4.0 * d2 * dt ** 4.0
        k3 = b3 + 2.0 * c3 * dt ** 2.0 + 4.0 * d3 * dt ** 4.0 - 8.0 * e3 * dt ** 6.0
        kicks = (k1, k2, k3, k3, k2, k1)
        drifts = (a1, a2, a3, a2, a1)
        f = edfdv(f=f, e=e, dt=kicks[0] * dt)
        for i in range(5):
            f = vdfdx(f=f, dt=drifts[i] * dt)
            e = field_solve(driver_field=driver_function(t + drifts[i] * dt), f=f)
            f = edfdv(f=f, e=e, dt=kicks[i + 1] * dt)
        return e, f

    return sixth_order_step


_INTEGRATORS = {
    "leapfrog": get_full_leapfrog_step,
    "pefrl": get_full_pefrl_step,
    "h-sixth": get_6th_order_integrator,
}


def get_time_integrator(time_integrator_name, vdfdx, edfdv, field_solver, stuff_for_time_loop):
    """Name dispatch (vlasov_poisson.py:235-280)."""
    maker = _INTEGRATORS.get(time_integrator_name)
    if maker is None:
        raise NotImplementedError("df/dt : <" + str(time_integrator_name) + "> has not yet been implemented")
    return maker(vdfdx=vdfdx, edfdv=edfdv, field_solve=field_solver, dt=stuff_for_time_loop["dt"],
                 driver_function=stuff_for_time_loop["driver_function"])
