"""Shared set-ups of the reference's example problems (used by oracle and GPU tests).

Each returns (interior q, delta, bc_lower, bc_upper, problem kwargs) following the
reference example script named in the docstring.
"""
import numpy as np

from oracle import oracle as O
from oracle.driver import cell_centers


def grid(lower, upper, n):
    xs = [cell_centers(l, h, m) for l, h, m in zip(lower, upper, n)]
    return np.meshgrid(*[c for c, _ in xs], indexing="ij"), [d for _, d in xs]


def advection_2d():
    """examples/advection_2d/advection_2d.py: qinit 19-26, setup 28-75."""
    (X, Y), d = grid([0, 0], [1, 1], [50, 50])
    q = (0.9 * (0.1 < X) * (X < 0.6) * (0.1 < Y) * (Y < 0.6) + 0.1)[None]
    return q, d


def advection_reaction_2d():
    """examples/advection_reaction_2d/advection_reaction.py: source_step 43-47,
    psi 50-52, set_velocities 54-63, setup 66-107 (100^2 on [0,1]^2, t_period 4,
    epsilon 0.1).  Returns (q, aux, delta, nodes, set_velocities(t, dt),
    source(q, dt)): the edge-velocity field for a time and the reaction step on
    a host q (meqn, mx, my)."""
    t_period, epsilon = 4.0, 0.1
    (Xc, Yc), d = grid([0, 0], [1, 1], [100, 100])
    xe = np.linspace(0.0, 1.0, 101)
    Xe, Ye = np.meshgrid(xe, xe, indexing="ij")
    dx, dy = d

    def psi(x, y):
        return (np.sin(np.pi * x)) ** 2 * (np.sin(np.pi * y)) ** 2 / np.pi

    def velocities(v_t=1.0):
        aux = np.empty((2, 100, 100), order="F")
        aux[0] = -(psi(Xe[:-1, 1:], Ye[:-1, 1:]) - psi(Xe[:-1, :-1], Ye[:-1, :-1])) / dy
        aux[1] = (psi(Xe[1:, :-1], Ye[1:, :-1]) - psi(Xe[:-1, :-1], Ye[:-1, :-1])) / dx
        aux[:] = v_t * aux[:]
        return aux

    def set_velocities(t, dt):
        return velocities(np.cos(2 * np.pi * (t + dt / 2) / t_period))

    def source(q, dt):
        qq = q.copy()
        q[0, :, :] = q[0, :, :] + epsilon * dt * qq[1, :, :]
        q[1, :, :] = q[1, :, :] + epsilon * dt * qq[0, :, :]

    q = np.array([(Xc <= 0.5) * 1.0, (Yc <= 0.5) * 1.0])
    return q, velocities(), d, set_velocities, source


def advection_annulus():
    """examples/advection_2d_annulus/advection_annulus.py: the polar map
    mapc2p_annulus 22-38, two Gaussian pulses qinit 41-61, ghost edge
    velocities 64-95, edge velocities and area ratio from the stream function
    pi*(x^2 + y^2) 98-147, setup 150-213 (r in [0.2, 1] x 40 cells, theta in
    [0, 2 pi] x 120, index_capa 2, vanleer, trans_cor, cfl 0.4/0.5, tfinal 1).

    Returns a dict: q (1, 40, 120), aux (3, 40, 120), auxbc (aux with its ghosts
    filled as _apply_bcs orders it: custom r sides, then periodic theta), delta,
    and the example's callbacks for the PyClaw API (mapc2p, edge_aux,
    ghost_lower / ghost_upper with the reference's custom-BC signature)."""
    lo, hi, n = (0.2, 0.0), (1.0, np.pi * 2.0), (40, 120)
    d = [(h - l) / float(m) for l, h, m in zip(lo, hi, n)]
    ng = 2

    def nodes(k, ghosts=0):
        # Dimension.nodes / nodes_with_ghost (geometry.py:538-545, 591-597)
        inner = np.array([lo[k] + i * d[k] for i in range(n[k] + 1)])
        if ghosts == 0:
            return inner
        below = np.linspace(lo[k] - ghosts * d[k], lo[k] - d[k], ghosts)
        above = np.linspace(hi[k] + d[k], hi[k] + ghosts * d[k], ghosts)
        return np.concatenate((below, inner, above))

    def mapc2p(xc, yc):
        return [xc * np.cos(yc), xc * np.sin(yc)]

    def stream(x, y):
        return np.pi * (x ** 2 + y ** 2)

    def edge_aux(X, Y, dx, dy):
        """aux of the cells spanned by corner arrays X, Y (nx+1, ny+1): u at the
        left edge, v at the bottom edge (stream-function differences) and the
        physical area over dx*dy (shoelace over the four corners)."""
        x0, y0 = X[:-1, :-1], Y[:-1, :-1]    # lower left
        x1, y1 = X[:-1, 1:], Y[:-1, 1:]      # upper left
        x2, y2 = X[1:, 1:], Y[1:, 1:]        # upper right
        x3, y3 = X[1:, :-1], Y[1:, :-1]      # lower right
        a = np.empty((3,) + x0.shape, order="F")
        a[0] = (stream(x1, y1) - stream(x0, y0)) / dy
        a[1] = -(stream(x3, y3) - stream(x0, y0)) / dx
        area = 0.5 * ((y0 + y1) * (x1 - x0) + (y1 + y2) * (x2 - x1) +
                      (y2 + y3) * (x3 - x2) + (y3 + y0) * (x0 - x3))
        a[2] = area / (dx * dy)
        return a

    def ghost_lower(state, dim, t, qbc, auxbc, num_ghost):
        grid = state.grid
        if dim != grid.dimensions[0]:
            raise Exception("custom aux BC only on the inner radius")
        dx, dy = grid.delta
        X, Y = grid.p_nodes_with_ghost(num_ghost=2)
        auxbc[:, 0:num_ghost, :] = edge_aux(X[0:num_ghost + 1, :], Y[0:num_ghost + 1, :], dx, dy)

    def ghost_upper(state, dim, t, qbc, auxbc, num_ghost):
        grid = state.grid
        if dim != grid.dimensions[0]:
            raise Exception("custom aux BC only on the outer radius")
        dx, dy = grid.delta
        X, Y = grid.p_nodes_with_ghost(num_ghost=2)
        auxbc[:, -num_ghost:, :] = edge_aux(X[-num_ghost - 1:, :], Y[-num_ghost - 1:, :], dx, dy)

    Xn, Yn = mapc2p(*np.meshgrid(nodes(0), nodes(1), indexing="ij"))
    aux = edge_aux(Xn, Yn, d[0], d[1])
    Xg, Yg = mapc2p(*np.meshgrid(nodes(0, ng), nodes(1, ng), indexing="ij"))
    auxbc = np.zeros((3, n[0] + 2 * ng, n[1] + 2 * ng), order="F")
    auxbc[:, ng:-ng, ng:-ng] = aux
    auxbc[:, 0:ng, :] = edge_aux(Xg[0:ng + 1, :], Yg[0:ng + 1, :], d[0], d[1])
    auxbc[:, -ng:, :] = edge_aux(Xg[-ng - 1:, :], Yg[-ng - 1:, :], d[0], d[1])
    auxbc[:, :, 0:ng] = auxbc[:, :, -2 * ng:-ng]       # theta periodic (solver.py:461-463)
    auxbc[:, :, -ng:] = auxbc[:, :, ng:2 * ng]

    centers = [np.array([lo[k] + (i + 0.5) * d[k] for i in range(n[k])]) for k in range(2)]
    X, Y = mapc2p(*np.meshgrid(*centers, indexing="ij"))
    q = (1.0 * np.exp(-40.0 * (np.square(X + 0.5) + np.square(Y - 0.0))) +
         -1.0 * np.exp(-40.0 * (np.square(X - 0.5) + np.square(Y - 0.0))))[None]
    return dict(q=np.asfortranarray(q), aux=aux, auxbc=auxbc, delta=d, lower=lo, upper=hi,
                num_cells=n, mapc2p=mapc2p, edge_aux=edge_aux, ghost_lower=ghost_lower,
                ghost_upper=ghost_upper)


def acoustics_2d_variable(num_cells=(50, 50)):
    """examples/acoustics_2d_variable/acoustics_2d_interface.py:22-96: [-1, 1]^2,
    density 4 | 1 and bulk modulus 4 | 4 left | right of x = 0 (aux = rho, c),
    a ring pulse of pressure centred at (-0.5, 0); walls at the lower sides,
    extrapolation at the upper ones (q and aux); tfinal 0.6, 20 outputs.
    Returns (q, aux, delta)."""
    (X, Y), d = grid([-1.0, -1.0], [1.0, 1.0], list(num_cells))
    rho_l, rho_r, bulk_l, bulk_r = 4.0, 1.0, 4.0, 4.0
    c_l, c_r = np.sqrt(bulk_l / rho_l), np.sqrt(bulk_r / rho_r)
    aux = np.empty((2,) + X.shape, order="F")
    aux[0] = rho_l * (X < 0.) + rho_r * (X >= 0.)
    aux[1] = c_l * (X < 0.) + c_r * (X >= 0.)
    r = np.sqrt((X + 0.5) ** 2 + (Y - 0.) ** 2)
    width, rad = 0.1, 0.25
    q = np.zeros((3,) + X.shape, order="F")
    q[0] = (np.abs(r - rad) <= width) * (1. + np.cos(np.pi * (r - rad) / width))
    return q, aux, d


def euler_quadrants():
    """examples/euler_2d/quadrants.py:80-110."""
    (X, Y), d = grid([0, 0], [1, 1], [100, 100])
    l, r, b, t = X < 0.8, X >= 0.8, Y < 0.8, Y >= 0.8
    rho = 1.5 * r * t + 0.532258064516129 * l * t + 0.137992831541219 * l * b + 0.532258064516129 * r * b
    u = 0.0 * r * t + 1.206045378311055 * l * t + 1.206045378311055 * l * b + 0.0 * r * b
    v = 0.0 * r * t + 0.0 * l * t + 1.206045378311055 * l * b + 1.206045378311055 * r * b
    p = 1.5 * r * t + 0.3 * l * t + 0.029032258064516 * l * b + 0.3 * r * b
    q = np.array([rho, rho * u, rho * v, 0.5 * rho * (u ** 2 + v ** 2) + p / (1.4 - 1.0)])
    return q, d


def shock_bubble(mx=160, my=40, gamma=1.4):
    """examples/euler_2d/shock_bubble_interaction.py: a planar shock (p = 5 behind
    it) entering at x = 0 and hitting a low-density bubble (rho 0.1, radius 0.2,
    centre (0.5, 0)) in r-z coordinates on [0,2]x[0,0.5].  qinit 41-87 (bubble
    edge cells averaged with scipy's quad), auxinit 90-96 (aux[0] = r), the
    incoming-shock ghost state 99-115 and the two-stage geometric source
    step_Euler_radial 118-153.  Returns (q, aux, delta, shock_state,
    step_source(q, aux, dt))."""
    from scipy import integrate
    g1 = gamma - 1.
    x0, y0, r0 = 0.5, 0., 0.2
    rhoin, pinf, rhoout, pout, pin, xshock = 0.1, 5., 1., 1., 1., 0.2
    (X, Y), d = grid([0., 0.], [2., 0.5], [mx, my])
    dx, dy = d
    rinf = (g1 + pinf * (gamma + 1.)) / ((gamma + 1.) + g1 * pinf)
    vinf = 1. / np.sqrt(gamma) * (pinf - 1.) / np.sqrt(0.5 * ((gamma + 1.) / gamma) * pinf
                                                       + 0.5 * g1 / gamma)
    einf = 0.5 * rinf * vinf ** 2 + pinf / g1
    r = np.sqrt((X - x0) ** 2 + (Y - y0) ** 2)
    q = np.zeros((5, mx, my), order="F")
    q[0] = rinf * (X < xshock) + rhoin * (r <= r0) + rhoout * (r > r0) * (X >= xshock)
    q[1] = rinf * vinf * (X < xshock)
    q[3] = einf * (X < xshock) + (pin * (r <= r0) + pout * (r > r0) * (X >= xshock)) / g1
    q[4] = 1. * (r <= r0)

    def height_inside(x, ylo, yhi):
        # length of [ylo, yhi] under the bubble's upper boundary at x
        if (x - x0) ** 2 < r0 ** 2:
            return max(min(y0 + np.sqrt(r0 ** 2 - (x - x0) ** 2), yhi) - ylo, 0.)
        return 0

    half_diag = np.linalg.norm([dx, dy]) / 2.
    for i in range(mx):
        for j in range(my):
            if abs(r[i, j] - r0) < half_diag:
                frac, _ = integrate.quad(height_inside, X[i, j] - dx / 2., X[i, j] + dx / 2.,
                                         args=(Y[i, j] - dy / 2., Y[i, j] + dy / 2.),
                                         epsabs=1.e-8, epsrel=1.e-5)
                frac = frac / (dx * dy)
                q[0, i, j] = rhoin * frac + rhoout * (1. - frac)
                q[3, i, j] = (pin * frac + pout * (1. - frac)) / g1
                q[4, i, j] = 1. * frac
    aux = np.asfortranarray(Y[None].copy())
    shock_state = np.array([rinf, rinf * vinf, 0., einf, 0.])

    def step_source(q, aux, dt):
        rad = aux[0]

        def rates(qq):
            rho = qq[0]
            u, v = qq[1] / rho, qq[2] / rho
            press = g1 * (qq[3] - 0.5 * rho * (u ** 2 + v ** 2))
            return rho, u, v, press

        rho, u, v, press = rates(q)
        h = dt / 2.
        qs = np.empty(q.shape)
        qs[0] = q[0] - h / rad * q[2]
        qs[1] = q[1] - h / rad * rho * u * v
        qs[2] = q[2] - h / rad * rho * v * v
        qs[3] = q[3] - h / rad * v * (q[3] + press)
        rho, u, v, press = rates(qs)
        q[0] = q[0] - dt / rad * qs[2]
        q[1] = q[1] - dt / rad * rho * u * v
        q[2] = q[2] - dt / rad * rho * v * v
        q[3] = q[3] - dt / rad * v * (qs[3] + press)

    return q, aux, d, shock_state, step_source


def shock_bubble_dq_src(gamma=1.4):
    """examples/euler_2d/shock_bubble_interaction.py dq_Euler_radial 156-178: the
    SharpClaw form of the radial source, dt times its value.  Returns
    dq_src(q, aux, dt)."""
    g1 = gamma - 1.

    def dq_src(q, aux, dt):
        rad = aux[0]
        rho = q[0]
        u, v = q[1] / rho, q[2] / rho
        press = g1 * (q[3] - 0.5 * rho * (u ** 2 + v ** 2))
        dq = np.empty(q.shape)
        dq[0] = -dt / rad * q[2]
        dq[1] = -dt / rad * rho * u * v
        dq[2] = -dt / rad * rho * v * v
        dq[3] = -dt / rad * v * (q[3] + press)
        dq[4] = 0
        return dq

    return dq_src


def shock_bubble_3d(num_cells=(64, 16, 16), gamma=1.4):
    """examples/euler_3d/shock_bubble.py: a planar shock (p = 5 behind it,
    40-43) entering at x = 0 (incoming_shock 63-72) and hitting a low-density
    sphere (rho 0.1, radius 0.2, centre (0.5, 0, 0)) on [0,2]x[0,0.5]x[0,0.5],
    walls at y = 0 and z = 0 (setup 75-150).  Cells within max(delta) of the
    sphere's surface take the volume fraction inside it from scipy's dblquad
    over x and y of the sphere's z-extent clipped to the cell (bubble 46-61,
    epsabs 1e-3, epsrel 1e-2).  Returns (q, delta, incoming_state)."""
    from scipy import integrate
    g1 = gamma - 1.
    x0, y0, z0, rb = 0.5, 0., 0., 0.2
    rhoout, pout, rhoin, pin, xshock = 1.0, 1.0, 0.1, 1.0, 0.2
    p_shock = 5.0
    rho_shock = (g1 + p_shock * (gamma + 1.)) / ((gamma + 1.) + g1 * p_shock)
    v_shock = (p_shock - 1.) / np.sqrt(0.5 * ((gamma + 1.) * p_shock + g1))
    e_shock = 0.5 * rho_shock * v_shock ** 2 + p_shock / g1
    (X, Y, Z), d = grid([0., 0., 0.], [2., 0.5, 0.5], list(num_cells))
    r0 = np.sqrt((X - x0) ** 2 + (Y - y0) ** 2 + (Z - z0) ** 2)
    q = np.zeros((5,) + tuple(num_cells), order="F")
    q[0] = rho_shock * (X < xshock) + rhoout * (X >= xshock)
    q[1] = rho_shock * v_shock * (X < xshock)
    q[4] = e_shock * (X < xshock) + pout / g1 * (X >= xshock)

    def extent(yy, xx, zdown, zup):
        z2 = rb ** 2 - (xx - x0) ** 2 - (yy - y0) ** 2
        top = 0 if z2 < 0 else z0 + np.sqrt(z2)
        bot = 0 if z2 < 0 else z0 - np.sqrt(z2)
        top = min(top, zup)
        return top - min(top, max(bot, zdown))

    dx, dy, dz = d
    dmax = max(d)
    for i, j, k in zip(*np.nonzero(r0 - dmax <= rb)):
        frac, _ = integrate.dblquad(extent, X[i, j, k] - dx / 2., X[i, j, k] + dx / 2.,
                                    lambda x, y=Y[i, j, k]: y - dy / 2.,
                                    lambda x, y=Y[i, j, k]: y + dy / 2.,
                                    args=(Z[i, j, k] - dz / 2., Z[i, j, k] + dz / 2.),
                                    epsabs=1.e-3, epsrel=1.e-2)
        frac = frac / (dx * dy * dz)
        q[0, i, j, k] = rhoout * (1. - frac) + rhoin * frac
        q[4, i, j, k] = (pout * (1. - frac) + pin * frac) / g1
    return q, d, np.array([rho_shock, rho_shock * v_shock, 0., 0., e_shock])


def sill():
    """examples/shallow_2d/sill.py: bathymetry 22-24 (a Gaussian hump
    b = 0.8 exp(-10((x-1)^2 + (y-0.5)^2))), the wave maker 26-36 (h mirrored into
    the lower-x ghost cells; in ghost cell num_ghost-1 the momentum 2 vwall - hu
    of the first interior cell, vwall = 0.1 sin(pi t / 1.5) until t = 5), setup
    41-82 (80 x 20 cells on [0,4]x[0,1], h = 1 - b, g = 1, dry tolerance 1e-3).
    Returns (q, aux, delta, wave_maker(t, qbc, ng)); wave_maker leaves the other
    ghost values of qbc as they are, as the example's function does."""
    (X, Y), d = grid([0., 0.], [4., 1.], [80, 20])
    aux = (0.8 * np.exp(-10 * ((X - 1.) ** 2 + (Y - 0.5) ** 2)))[None]
    q = np.zeros((3, 80, 20), order="F")
    q[0] = 1. - aux[0]

    def wave_maker(t, qbc, ng):
        qbc[0, :ng, :] = qbc[0, ng, :]
        vwall = 0.1 * np.sin(t * np.pi / 1.5) if t <= 5. else 0.
        for ibc in range(ng - 1):
            qbc[1, ng - ibc - 1, :] = 2 * vwall - qbc[1, ng + ibc, :]

    return q, np.asfortranarray(aux), d, wave_maker


def radial_dam_break():
    """examples/shallow_2d/radial_dam_break.py: qinit 22-30, setup 33-80."""
    (X, Y), d = grid([-2.5, -2.5], [2.5, 2.5], [150, 150])
    r = np.sqrt(X ** 2 + Y ** 2)
    q = np.array([2. * (r <= 0.5) + 1. * (r > 0.5), 0 * X, 0 * X])
    return q, d


def acoustics_2d():
    """examples/acoustics_2d_homogeneous/acoustics_2d.py: setup 58-90, qinit 92-98."""
    (X, Y), d = grid([-1, -1], [1, 1], [100, 100])
    r = np.sqrt(X ** 2 + Y ** 2)
    w = 0.2
    q = np.array([(np.abs(r - 0.5) <= w) * (1. + np.cos(np.pi * (r - 0.5) / w)), 0 * X, 0 * X])
    return q, d


def shocktube_3d():
    """examples/euler_3d/shocktube.py:30-52 (mx=10, my=10, mz=128)."""
    (X, Y, Z), d = grid([-1, -1, -1], [1, 1, 1], [10, 10, 128])
    p = 3. * (Z <= 0) + 1. * (Z > 0)
    q = np.array([3. * (Z <= 0) + 1. * (Z > 0), 0 * X, 0 * X, 0 * X, p / (1.4 - 1.)])
    return q, d


def seeded_state(rng, kind, shape):
    """Seeded random admissible states (positive density/depth and pressure)."""
    meqn = O.MEQN[kind]
    q = np.zeros((meqn,) + shape)
    if kind == O.ACOUSTICS:
        q[:] = rng.uniform(-1, 1, q.shape)
    elif kind in (O.EULER, O.EULER3, O.EULER5):
        nd = 3 if kind == O.EULER3 else 2
        rho = rng.uniform(0.5, 1.5, shape)
        vel = [rng.uniform(-0.5, 0.5, shape) for _ in range(nd)]
        p = rng.uniform(0.1, 1.0, shape)
        q[0] = rho
        for d in range(nd):
            q[1 + d] = rho * vel[d]
        q[nd + 1] = p / 0.4 + 0.5 * rho * sum(v * v for v in vel)
        if kind == O.EULER5:
            q[4] = rng.uniform(0, 1, shape)
    elif kind in (O.SHALLOW, O.SHALLOW_HLLE):
        h = rng.uniform(0.5, 1.5, shape)
        q[0] = h
        q[1] = h * rng.uniform(-0.5, 0.5, shape)
        q[2] = h * rng.uniform(-0.5, 0.5, shape)
    elif kind == O.ADVECTION:
        q[:] = rng.uniform(0, 1, q.shape)
    return np.asfortranarray(q)
