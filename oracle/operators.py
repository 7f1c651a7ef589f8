"""CPU oracle: PDE operator callbacks `operator(mod, ctx)` of the reference.

TEST INFRASTRUCTURE ONLY (see `oracle/tape.py`). Each factory returns a
callable with the reference's callback signature; evaluated through
`oracle.tape` it reproduces the reference's residuals, tape Jacobians and loss
gradients in float64.
"""

import numpy as np
import torch


def extrap_quadh(u0, u1, u1p):
    """Quadratic extrapolation through points 0, 1 and the half point 1.5 to
    point 2 (`util_op.py:1017-1023`)."""
    return (u0 - 6 * u1 + 8 * u1p) / 3


def extrap_quad(u0, u1, u2):
    """Quadratic extrapolation from points 0, 1, 2 to 3
    (`util_op.py:1026-1029`)."""
    return u0 - 3 * u1 + 3 * u2


def extrap_linear(u0, u1):
    """Linear extrapolation from points 0, 1 to 2 (`util_op.py:1032-1035`)."""
    return 2 * u1 - u0


# ---------------------------------------------------------------------------
# 1D wave equation on a (t, x) space-time grid, `wave.py:67-122`.


def wave_exact(t, x):
    """Exact solution of the reference wave case and its time derivative
    (`wave.py:27-39`): mean of 10 travelling cosines with k = i*pi."""
    t = np.asarray(t, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    u = np.zeros(np.broadcast(t, x).shape)
    ut = np.zeros_like(u)
    for i in range(1, 6):
        k = i * np.pi
        a = (x - t + 0.5) * k
        b = (x + t - 0.5) * k
        u = u + np.cos(a) + np.cos(b)
        ut = ut + k * np.sin(a) - k * np.sin(b)
    return u / 10, ut / 10


def make_wave_operator(init_u, init_ut, left_u, right_u):
    """Space-time residual of u_tt = u_xx (`wave.py:67-122`).

    init_u, init_ut: arrays over x at t=0; left_u, right_u: boundary values
    over t at x=lower/upper. Row it=0 imposes u = init_u + dt/2*init_ut; row
    it=1 uses init_ut for the backward time difference; x-halos at t-1 come
    from quadratic extrapolation through the boundary value at t-1.
    """
    init_u = np.asarray(init_u, dtype=np.float64)
    init_ut = np.asarray(init_ut, dtype=np.float64)
    left_u = np.asarray(left_u, dtype=np.float64)
    right_u = np.asarray(right_u, dtype=np.float64)

    def operator(mod, ctx):
        dt = ctx.step('t')
        dx = ctx.step('x')
        it = ctx.cell_index('t')
        ix = ctx.cell_index('x')
        nx = ctx.size('x')
        u = ctx.field('u')
        utm = ctx.field('u', -1, 0)
        utmm = ctx.field('u', -2, 0)
        uxm = ctx.field('u', -1, -1)
        uxp = ctx.field('u', -1, 1)
        # Boundary values at the previous time row (periodic roll in t).
        lprev = torch.as_tensor(np.roll(left_u, 1))[:, None]
        rprev = torch.as_tensor(np.roll(right_u, 1))[:, None]
        uxm = mod.where(ix == 0, extrap_quadh(uxp, utm, lprev), uxm)
        uxp = mod.where(ix == nx - 1, extrap_quadh(uxm, utm, rprev), uxp)
        d_now = (u - utm) / dt
        d_prev = (utm - utmm) / dt
        d_prev = mod.where(it == 1, torch.as_tensor(init_ut)[None, :],
                           d_prev)
        u_tt = (d_now - d_prev) / dt
        u_xx = (uxm - 2 * utm + uxp) / dx**2
        fu = u_tt - u_xx
        u0 = torch.as_tensor(init_u + 0.5 * dt * init_ut)[None, :]
        fu = mod.where(it == 0, u - u0, fu)
        return [fu]

    return operator


def wave_reference_setup(nt, nx):
    """Domain and boundary data of the reference wave case
    (`wave.py:433-464`): t in [0,1], x in [-1,1], exact cosine solution."""
    from oracle.tape import Domain
    domain = Domain(ndim=2, shape=(nt, nx), lower=(0, -1), upper=(1, 1),
                    varnames=('t', 'x'), fieldnames=['u'])
    t1 = domain.cell_center_1d(0)
    x1 = domain.cell_center_1d(1)
    left_u, _ = wave_exact(t1, t1 * 0 + domain.lower[1])
    right_u, _ = wave_exact(t1, t1 * 0 + domain.upper[1])
    init_u, init_ut = wave_exact(x1 * 0 + domain.lower[0], x1)
    t, x = domain.cell_center_all()
    exact_u, _ = wave_exact(t, x)
    op = make_wave_operator(init_u, init_ut, left_u, right_u)
    return domain, op, exact_u


def wave_gaussian_setup(nt, nx, seed):
    """Wave case with a seeded sum of 3 Gaussians as u(x,0), u_t(x,0)=0 and
    Dirichlet zero at x=+-1. Centres U(-0.5,0.5), widths U(0.05,0.15),
    amplitudes U(0.5,1) from `numpy.random.default_rng(seed)`."""
    from oracle.tape import Domain
    domain = Domain(ndim=2, shape=(nt, nx), lower=(0, -1), upper=(1, 1),
                    varnames=('t', 'x'), fieldnames=['u'])
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.5, 0.5, 3)
    w = rng.uniform(0.05, 0.15, 3)
    a = rng.uniform(0.5, 1.0, 3)
    x1 = domain.cell_center_1d(1)
    init_u = np.zeros(nx)
    for k in range(3):
        init_u += a[k] * np.exp(-(x1 - c[k])**2 / (2 * w[k]**2))
    init_ut = np.zeros(nx)
    zeros_t = np.zeros(nt)
    op = make_wave_operator(init_u, init_ut, zeros_t, zeros_t)
    return domain, op, (init_u, init_ut, zeros_t, zeros_t)


# ---------------------------------------------------------------------------
# Steady lid-driven cavity, collocated grid with Rhie-Chow correction,
# `ns_steady.py:142-394` (defaults bc=0, rhiechow=1, no imposed points).


def make_cavity_operator(mu, bc=0, rhiechow=1, imposed=None, alpha=1.,
                         kreg=0.):
    """Residuals (fu, fv, fp) of the steady Navier-Stokes cavity.

    Velocity derivatives use deferred correction: second-order upwind of the
    frozen field plus first-order upwind of the live field minus first-order
    upwind of the frozen field (`ns_steady.py:171-204,282-283,320-323`).
    The lid (u=1) is at y=1 (`:224`).

    Inverse-problem terms appended as in `ns_steady.py:396-411`: with
    `imposed` = (mask, u_imp, v_imp), mask (u - u_imp) coeff and
    mask (v - v_imp) coeff, coeff = alpha N / n_imp (n_imp the nonzeros of
    mask); with kreg != 0, kreg times u_xx, u_yy, v_xx, v_yy of the
    halo-corrected stencils (`laplace_split`, `:213-217`).
    """

    def operator(mod, ctx):
        dx = ctx.step('x')
        dy = ctx.step('y')
        ix = ctx.cell_index('x')
        iy = ctx.cell_index('y')
        nx = ctx.size('x')
        ny = ctx.size('y')
        zeros = ctx.field('zeros')

        def five(key, freeze=False):
            return [
                ctx.field(key, freeze=freeze),
                ctx.field(key, -1, 0, freeze=freeze),
                ctx.field(key, 1, 0, freeze=freeze),
                ctx.field(key, 0, -1, freeze=freeze),
                ctx.field(key, 0, 1, freeze=freeze),
            ]

        def bc_wall(st, lid):
            # Halo from quadratic extrapolation through the wall value.
            st[1] = mod.where(ix == 0, extrap_quadh(st[2], st[0], 0), st[1])
            st[2] = mod.where(ix == nx - 1, extrap_quadh(st[1], st[0], 0),
                              st[2])
            st[3] = mod.where(iy == 0, extrap_quadh(st[4], st[0], 0), st[3])
            st[4] = mod.where(iy == ny - 1, extrap_quadh(st[3], st[0], lid),
                              st[4])

        def bc_lin(st):
            st[1] = mod.where(ix == 0, extrap_linear(st[2], st[0]), st[1])
            st[2] = mod.where(ix == nx - 1, extrap_linear(st[1], st[0]),
                              st[2])
            st[3] = mod.where(iy == 0, extrap_linear(st[4], st[0]), st[3])
            st[4] = mod.where(iy == ny - 1, extrap_linear(st[3], st[0]),
                              st[4])

        def bc_npen(st, normal_x):
            # No-penetration: wall value 0 across the normal faces,
            # extrapolation along the tangential ones (`:243-257`).
            if normal_x:
                st[1] = mod.where(ix == 0, extrap_quadh(st[2], st[0], 0),
                                  st[1])
                st[2] = mod.where(ix == nx - 1,
                                  extrap_quadh(st[1], st[0], 0), st[2])
                st[3] = mod.where(iy == 0, extrap_linear(st[4], st[0]),
                                  st[3])
                st[4] = mod.where(iy == ny - 1, extrap_linear(st[3], st[0]),
                                  st[4])
            else:
                st[1] = mod.where(ix == 0, extrap_linear(st[2], st[0]),
                                  st[1])
                st[2] = mod.where(ix == nx - 1, extrap_linear(st[1], st[0]),
                                  st[2])
                st[3] = mod.where(iy == 0, extrap_quadh(st[4], st[0], 0),
                                  st[3])
                st[4] = mod.where(iy == ny - 1,
                                  extrap_quadh(st[3], st[0], 0), st[4])

        def apply_bc(st, which):
            if bc == 0:
                bc_wall(st, 1 if which == 'u' else 0)
            elif bc == 1:
                bc_npen(st, which == 'u')
            elif bc == 2:
                bc_lin(st)
            else:
                raise ValueError("Unknown bc={:}".format(bc))

        def second_neighbours(st):
            # Neighbours at distance 2 built by rolling the halo-corrected
            # distance-1 neighbours; at the wall they repeat the halo.
            xmm = mod.where(ix == 0, st[1], mod.roll(st[1], 1, 0))
            xpp = mod.where(ix == nx - 1, st[2], mod.roll(st[2], -1, 0))
            ymm = mod.where(iy == 0, st[3], mod.roll(st[3], 1, 1))
            ypp = mod.where(iy == ny - 1, st[4], mod.roll(st[4], -1, 1))
            return xmm, xpp, ymm, ypp

        def upwind2(st, st5, a, b):
            q, qxm, qxp, qym, qyp = st
            qxmm, qxpp, qymm, qypp = st5
            dxm = (qxmm - 4 * qxm + 3 * q) / (2 * dx)
            dxp = (-3 * q + 4 * qxp - qxpp) / (2 * dx)
            dym = (qymm - 4 * qym + 3 * q) / (2 * dy)
            dyp = (-3 * q + 4 * qyp - qypp) / (2 * dy)
            dxm = mod.where(ix == 0, (q - qxm) / dx, dxm)
            dxp = mod.where(ix == nx - 1, (qxp - q) / dx, dxp)
            dym = mod.where(iy == 0, (q - qym) / dy, dym)
            dyp = mod.where(iy == ny - 1, (qyp - q) / dy, dyp)
            return mod.where(a > 0, dxm, dxp), mod.where(b > 0, dym, dyp)

        def upwind1(st, a, b):
            q, qxm, qxp, qym, qyp = st
            gx = mod.where(a > 0, (q - qxm) / dx, (qxp - q) / dx)
            gy = mod.where(b > 0, (q - qym) / dy, (qyp - q) / dy)
            return gx, gy

        def lap(st):
            q, qxm, qxp, qym, qyp = st
            return (qxp - 2 * q + qxm) / dx**2 + (qyp - 2 * q + qym) / dy**2

        u_st = five('u')
        uf_st = five('u', freeze=True)
        apply_bc(u_st, 'u')
        apply_bc(uf_st, 'u')
        uf = uf_st[0]
        v_st = five('v')
        vf_st = five('v', freeze=True)
        apply_bc(v_st, 'v')
        apply_bc(vf_st, 'v')
        vf = vf_st[0]
        p_st = five('p')
        bc_lin(p_st)

        hx, hy = upwind2(uf_st, second_neighbours(uf_st), uf, vf)
        lx, ly = upwind1(u_st, uf, vf)
        fx, fy = upwind1(uf_st, uf, vf)
        u_x, u_y = hx + lx - fx, hy + ly - fy
        hx, hy = upwind2(vf_st, second_neighbours(vf_st), uf, vf)
        lx, ly = upwind1(v_st, uf, vf)
        fx, fy = upwind1(vf_st, uf, vf)
        v_x, v_y = hx + lx - fx, hy + ly - fy
        p_x = (p_st[2] - p_st[1]) / (2 * dx)
        p_y = (p_st[4] - p_st[3]) / (2 * dy)

        fu = uf * u_x + vf * u_y - mu * lap(u_st) + p_x
        fv = uf * v_x + vf * v_y - mu * lap(v_st) + p_y

        if rhiechow:
            g = mod.abs(uf) / dx + mod.abs(vf) / dy + mu * (2 / dx**2 +
                                                            2 / dy**2)
            gxm = mod.roll(g, 1, 0)
            gxp = mod.roll(g, -1, 0)
            gym = mod.roll(g, 1, 1)
            gyp = mod.roll(g, -1, 1)
            u, uxm, uxp = u_st[0], u_st[1], u_st[2]
            v, vym, vyp = v_st[0], v_st[3], v_st[4]
            pf_st = five('p', freeze=True)
            bc_lin(pf_st)
            pf, pfxm, pfxp, pfym, pfyp = pf_st
            pfxmm, pfxpp, pfymm, pfypp = second_neighbours(pf_st)
            p, pxm, pxp, pym, pyp = p_st
            # Face velocities with pressure-gradient correction.
            qxm = ((u + uxm) * 0.5 + (1 / gxm * (pf - pfxmm) /
                                      (2 * dx)) * 0.5 +
                   (1 / g * (pfxp - pfxm) / (2 * dx)) * 0.5 -
                   (1 / gxm + 1 / g) * 0.5 * (p - pxm) / dx)
            qxp = ((u + uxp) * 0.5 + (1 / gxp * (pfxpp - pf) /
                                      (2 * dx)) * 0.5 +
                   (1 / g * (pfxp - pfxm) / (2 * dx)) * 0.5 -
                   (1 / gxp + 1 / g) * 0.5 * (pxp - p) / dx)
            qym = ((v + vym) * 0.5 + (1 / gym * (pf - pfymm) /
                                      (2 * dy)) * 0.5 +
                   (1 / g * (pfyp - pfym) / (2 * dy)) * 0.5 -
                   (1 / gym + 1 / g) * 0.5 * (p - pym) / dy)
            qyp = ((v + vyp) * 0.5 + (1 / gyp * (pfypp - pf) /
                                      (2 * dy)) * 0.5 +
                   (1 / g * (pfyp - pfym) / (2 * dy)) * 0.5 -
                   (1 / gyp + 1 / g) * 0.5 * (pyp - p) / dy)
            if bc == 2:
                qxm = mod.where(ix == 0, (u + uxm) * 0.5, qxm)
                qxp = mod.where(ix == nx - 1, (u + uxp) * 0.5, qxp)
                qym = mod.where(iy == 0, (v + vym) * 0.5, qym)
                qyp = mod.where(iy == ny - 1, (v + vyp) * 0.5, qyp)
            else:
                qxm = mod.where(ix == 0, zeros, qxm)
                qxp = mod.where(ix == nx - 1, zeros, qxp)
                qym = mod.where(iy == 0, zeros, qym)
                qyp = mod.where(iy == ny - 1, zeros, qyp)
            fp = (qxp - qxm) / dx + (qyp - qym) / dy
        else:
            u_xc = (u_st[2] - u_st[1]) / (2 * dx)
            v_yc = (v_st[4] - v_st[3]) / (2 * dy)
            fp = u_xc + v_yc
        res = [fu, fv, fp]
        if imposed is not None:
            mask, uimp, vimp = (torch.as_tensor(np.asarray(a, np.float64))
                                for a in imposed)
            n_imp = int(np.count_nonzero(np.asarray(imposed[0])))
            coeff = alpha * nx * ny / n_imp
            res += [mask * (u_st[0] - uimp) * coeff,
                    mask * (v_st[0] - vimp) * coeff]
        if kreg:
            def laplace_split(st):
                q, qxm, qxp, qym, qyp = st
                return ((qxp - 2 * q + qxm) / dx**2,
                        (qyp - 2 * q + qym) / dy**2)
            u_xx, u_yy = laplace_split(u_st)
            v_xx, v_yy = laplace_split(v_st)
            res += [u_xx * kreg, u_yy * kreg, v_xx * kreg, v_yy * kreg]
        return res

    return operator


def cavity_domain(n, ny=None):
    from oracle.tape import Domain
    return Domain(ndim=2, shape=(n, ny or n), varnames=('x', 'y'),
                  fieldnames=['u', 'v', 'p'])


# ---------------------------------------------------------------------------
# Linear advection on a (t, x, y) space-time grid, `advection_2d.py:103-223`
# with problem='direct', periodic=1, central=0, implicit=0. The velocity is a
# given constant (ax, ay); the unknown is the advected field `p`.


def make_advection_operator(ax, ay, p_init):
    """fp = (p - p(t-1))/dt + a . grad p, first-order upwind on sign(a),
    periodic in x and y. At it=0 the previous time row is `p_init`
    (`advection_2d.py:200-202,209,217,220`)."""
    p_init = np.asarray(p_init, dtype=np.float64)

    def operator(mod, ctx):
        dt = ctx.step('t')
        dx = ctx.step('x')
        dy = ctx.step('y')
        it = ctx.cell_index('t')
        p = ctx.field('p')
        pxm = ctx.field('p', 0, -1, 0)
        pxp = ctx.field('p', 0, 1, 0)
        pym = ctx.field('p', 0, 0, -1)
        pyp = ctx.field('p', 0, 0, 1)
        ptm = ctx.field('p', -1, 0, 0)
        ptm = mod.where(it == 0, torch.as_tensor(p_init)[None], ptm)
        p_x = (p - pxm) / dx if ax > 0 else (pxp - p) / dx
        p_y = (p - pym) / dy if ay > 0 else (pyp - p) / dy
        fp = (p - ptm) / dt + ax * p_x + ay * p_y
        return [fp]

    return operator


def make_advection_inverse_operator(p_exact, alpha=1., anneal_half=0.,
                                    onlyfinal=False, kreg=1e-4, ktreg=0.,
                                    divfree=0., periodic=True):
    """`operator` of advection_2d.py:103-250 with problem='inverse',
    central=0, implicit=0: unknowns u, v, p on (t, x, y), equations
    [fe, fp] + [fsu, fsv] (kreg) + [ftregu, ftregv] (ktreg) + [fdiv]
    (divfree). The upwind direction of p_x, p_y comes from the frozen
    velocity (`uf`, `vf`), the advecting velocity is live; `fe` is annealed
    by 0.5^(epoch / anneal_half) (`get_anneal_factor`, :25-26). Non-periodic
    grids zero the x/y halo of every stencil (`apply_bc_zero`); at it = 0
    the previous time level of p is p_exact (:200-202)."""
    pex = torch.as_tensor(np.asarray(p_exact, dtype=np.float64))

    def operator(mod, ctx):
        dt = ctx.step('t')
        dx = ctx.step('x')
        dy = ctx.step('y')
        it = ctx.cell_index('t')
        ix = ctx.cell_index('x')
        iy = ctx.cell_index('y')
        nt = ctx.size('t')
        nx = ctx.size('x')
        ny = ctx.size('y')
        zeros = ctx.field('zeros')

        def st(key, freeze=False, shift_t=0):
            return [ctx.field(key, shift_t, 0, 0, freeze=freeze),
                    ctx.field(key, shift_t, -1, 0, freeze=freeze),
                    ctx.field(key, shift_t, 1, 0, freeze=freeze),
                    ctx.field(key, shift_t, 0, -1, freeze=freeze),
                    ctx.field(key, shift_t, 0, 1, freeze=freeze)]

        def bc_zero(s):
            if periodic:
                return s
            return [s[0], mod.where(ix == 0, zeros, s[1]),
                    mod.where(ix == nx - 1, zeros, s[2]),
                    mod.where(iy == 0, zeros, s[3]),
                    mod.where(iy == ny - 1, zeros, s[4])]

        u_st = bc_zero(st('u'))
        v_st = bc_zero(st('v'))
        uf = ctx.field('u', freeze=True)
        vf = ctx.field('v', freeze=True)
        p_st = bc_zero(st('p'))
        ptm = ctx.field('p', -1, 0, 0)
        ptm = mod.where(it == 0, pex, ptm)
        p, pxm, pxp, pym, pyp = p_st
        p_x = mod.where(uf > 0, (p - pxm) / dx, (pxp - p) / dx)
        p_y = mod.where(vf > 0, (p - pym) / dy, (pyp - p) / dy)
        u, v = u_st[0], v_st[0]
        fp = (p - ptm) / dt + u * p_x + v * p_y
        anneal = 0.5**(ctx.epoch / anneal_half) if anneal_half else 1.
        fe = (p - pex) * alpha * anneal
        if onlyfinal:
            fe = mod.where((it == 0) | (it == nt - 1), fe, zeros)
        res = [fe, fp]

        def lap(s):
            q, qxm, qxp, qym, qyp = s
            return (qxp - 2 * q + qxm) / dx**2 + (qyp - 2 * q + qym) / dy**2

        if kreg:
            res += [lap(u_st) * kreg, lap(v_st) * kreg]
        if ktreg:
            fu = (u - ctx.field('u', -1, 0, 0)) / dt * ktreg
            fv = (v - ctx.field('v', -1, 0, 0)) / dt * ktreg
            res += [mod.where(it == 0, zeros, fu),
                    mod.where(it == 0, zeros, fv)]
        if divfree:
            u_xc = (u_st[2] - u_st[1]) / (2 * dx)
            v_yc = (v_st[4] - v_st[3]) / (2 * dy)
            res.append((u_xc + v_yc) * divfree)
        return res

    return operator


def advection_inverse_domain(nt, nx, ny, tmax=1.0):
    from oracle.tape import Domain
    return Domain(ndim=3, shape=(nt, nx, ny), upper=(tmax, 1., 1.),
                  varnames=('t', 'x', 'y'), fieldnames=['u', 'v', 'p'])


def advection_setup(nt, nx, ny, seed, tmax=1.0):
    """Seeded advection case: a ~ U(-1,1)^2, Gaussian blob p_init with centre
    U(0.3,0.7)^2 and sigma 0.1 on x,y in [0,1]."""
    from oracle.tape import Domain
    rng = np.random.default_rng(seed)
    ax, ay = rng.uniform(-1, 1, 2)
    cx, cy = rng.uniform(0.3, 0.7, 2)
    domain = Domain(ndim=3, shape=(nt, nx, ny), upper=(tmax, 1., 1.),
                    varnames=('t', 'x', 'y'), fieldnames=['p'])
    x = domain.cell_center_1d(1)
    y = domain.cell_center_1d(2)
    xx, yy = np.meshgrid(x, y, indexing='ij')
    p_init = np.exp(-((xx - cx)**2 + (yy - cy)**2) / (2 * 0.1**2))
    return domain, make_advection_operator(ax, ay, p_init), (ax, ay, p_init)


# ---------------------------------------------------------------------------
# Laplace and Poisson operators.


def operator_laplace(mod, ctx):
    """Two decoupled Laplacians with linear-extrapolation halos plus a zero
    equation (`util_op.py:1038-1079`)."""
    dx = ctx.step('x')
    dy = ctx.step('y')
    ix = ctx.cell_index('x')
    iy = ctx.cell_index('y')
    nx = ctx.size('x')
    ny = ctx.size('y')
    zeros = ctx.field('zeros')

    def lap_of(key):
        q = ctx.field(key)
        qxm = ctx.field(key, -1, 0)
        qxp = ctx.field(key, 1, 0)
        qym = ctx.field(key, 0, -1)
        qyp = ctx.field(key, 0, 1)
        qxm = mod.where(ix == 0, extrap_linear(qxp, q), qxm)
        qxp = mod.where(ix == nx - 1, extrap_linear(qxm, q), qxp)
        qym = mod.where(iy == 0, extrap_linear(qyp, q), qym)
        qyp = mod.where(iy == ny - 1, extrap_linear(qym, q), qyp)
        return (qxp - 2 * q + qxm) / dx**2 + (qyp - 2 * q + qym) / dy**2

    return [lap_of('u'), lap_of('v'), zeros]


def make_poisson_operator(src):
    """f = lap(u) - src with Dirichlet u=0 on the walls, halos by quadratic
    extrapolation through the wall value (the no-slip rule of
    `ns_steady.py:219-225` applied to a scalar)."""
    src = np.asarray(src, dtype=np.float64)

    def operator(mod, ctx):
        dx = ctx.step('x')
        dy = ctx.step('y')
        ix = ctx.cell_index('x')
        iy = ctx.cell_index('y')
        nx = ctx.size('x')
        ny = ctx.size('y')
        q = ctx.field('u')
        qxm = ctx.field('u', -1, 0)
        qxp = ctx.field('u', 1, 0)
        qym = ctx.field('u', 0, -1)
        qyp = ctx.field('u', 0, 1)
        qxm = mod.where(ix == 0, extrap_quadh(qxp, q, 0), qxm)
        qxp = mod.where(ix == nx - 1, extrap_quadh(qxm, q, 0), qxp)
        qym = mod.where(iy == 0, extrap_quadh(qyp, q, 0), qym)
        qyp = mod.where(iy == ny - 1, extrap_quadh(qym, q, 0), qyp)
        lap = (qxp - 2 * q + qxm) / dx**2 + (qyp - 2 * q + qym) / dy**2
        return [lap - torch.as_tensor(src)]

    return operator


def gaussian_mixture_source(n, seed, nblobs=3):
    """Seeded source on the unit square: sum of `nblobs` Gaussians with centres
    U(0.2,0.8)^2, widths U(0.05,0.15), amplitudes U(-1,1)."""
    rng = np.random.default_rng(seed)
    x = (np.arange(n) + 0.5) / n
    xx, yy = np.meshgrid(x, x, indexing='ij')
    s = np.zeros((n, n))
    for _ in range(nblobs):
        cx, cy = rng.uniform(0.2, 0.8, 2)
        w = rng.uniform(0.05, 0.15)
        a = rng.uniform(-1, 1)
        s += a * np.exp(-((xx - cx)**2 + (yy - cy)**2) / (2 * w**2))
    return s
