"""CPU oracle for the inverse wave problem: recover the wave speed c(x) from
sensor traces. TEST INFRASTRUCTURE ONLY (imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline only, as the checker).

The reference has no inverse wave case; the problem is built from its
pieces:
* the space-time wave residual of `wave.py:67-122` with the spatial term
  scaled by c(x)^2 (c = 1 gives the reference operator exactly);
* data misfit on sensor columns in the form of the reference's imposed
  points, `(u - u_imp) * k_imp` as an extra equation (`ns_steady.py:
  396-411`);
* Tikhonov regularisation k_reg^2 * mean((dc/dx)^2) on the parameter.
Loss = mean(f_wave^2) + mean(f_data^2) + reg, the `Problem._eval_loss` sum
(util_op.py:628-654); gradients with respect to u and c by torch autograd
in float64.
"""

import numpy as np
import torch

from oracle.operators import extrap_quadh


def wave_inverse_setup(nt, nx, nsensor=64, seed=0):
    """Seeded case on t in [0, 1], x in [-1, 1]: initial u(x, 0) a sum of
    3 Gaussians (centres U(-0.5, 0.5), widths U(0.05, 0.15), amplitudes
    U(0.5, 1)), u_t(x, 0) = 0, walls u = 0; true speed
    c(x) = 1 + sum_4 s_k tanh((x - x_k) / 0.05), s_k ~ U(-0.25, 0.25),
    x_k ~ U(-0.8, 0.8), clipped to [0.5, 1.5]; `nsensor` evenly spaced
    sensor columns; data = u of the forward march with the true speed."""
    rng = np.random.default_rng(seed)
    dt, dx = 1. / nt, 2. / nx
    x1 = -1 + dx * (np.arange(nx) + 0.5)
    cc = rng.uniform(-0.5, 0.5, 3)
    ww = rng.uniform(0.05, 0.15, 3)
    aa = rng.uniform(0.5, 1.0, 3)
    init_u = sum(aa[k] * np.exp(-(x1 - cc[k])**2 / (2 * ww[k]**2))
                 for k in range(3))
    s = rng.uniform(-0.25, 0.25, 4)
    xk = rng.uniform(-0.8, 0.8, 4)
    c_true = np.clip(1 + sum(s[k] * np.tanh((x1 - xk[k]) / 0.05)
                             for k in range(4)), 0.5, 1.5)
    if c_true.max() * dt / dx > 1:
        raise ValueError('explicit march unstable: c dt / dx = {:.3f} > 1 '
                         '(use nt >= nx)'.format(c_true.max() * dt / dx))
    sensors = np.unique(np.round(np.linspace(0, nx - 1, nsensor)).astype(int))
    case = dict(nt=nt, nx=nx, dt=dt, dx=dx, x=x1, init_u=init_u,
                init_ut=np.zeros(nx), left_u=np.zeros(nt),
                right_u=np.zeros(nt), c_true=c_true, sensors=sensors)
    u_true = wave_march(c_true, case)
    case['u_true'] = u_true
    case['data'] = u_true[:, sensors]
    return case


def _uxx_row(row, lb, rb, dx):
    """u_xx of one time row with the quadratic half-cell ghosts through the
    wall values lb, rb (wave.py:93-102)."""
    um = np.empty_like(row)
    up = np.empty_like(row)
    um[1:] = row[:-1]
    up[:-1] = row[1:]
    um[0] = extrap_quadh(row[1], row[0], lb)
    up[-1] = extrap_quadh(row[-2], row[-1], rb)
    return (um - 2 * row + up) / dx**2


def wave_march(c, case):
    """The space-time solution of f_wave = 0 row by row (explicit):
    u_0 = u0 + dt/2 ut0; u_1 = u_0 + dt ut0 + dt^2 c^2 u_xx(u_0);
    u_t = 2 u_{t-1} - u_{t-2} + dt^2 c^2 u_xx(u_{t-1})."""
    nt, nx, dt, dx = case['nt'], case['nx'], case['dt'], case['dx']
    c2 = np.asarray(c, dtype=np.float64)**2
    u = np.zeros((nt, nx))
    u[0] = case['init_u'] + 0.5 * dt * case['init_ut']
    lap = _uxx_row(u[0], case['left_u'][0], case['right_u'][0], dx)
    u[1] = u[0] + dt * case['init_ut'] + dt**2 * c2 * lap
    for t in range(2, nt):
        lap = _uxx_row(u[t - 1], case['left_u'][t - 1],
                       case['right_u'][t - 1], dx)
        u[t] = 2 * u[t - 1] - u[t - 2] + dt**2 * c2 * lap
    return u


def wave_c_residual(u, c, case):
    """f_wave(u, c) as a torch expression (float64): the reference wave
    residual with u_xx scaled by c(x)^2."""
    nt, nx, dt, dx = case['nt'], case['nx'], case['dt'], case['dx']
    it = torch.arange(nt)[:, None]
    ix = torch.arange(nx)[None, :]
    utm = torch.roll(u, 1, 0)
    utmm = torch.roll(u, 2, 0)
    uxm = torch.roll(utm, 1, 1)
    uxp = torch.roll(utm, -1, 1)
    lprev = torch.as_tensor(np.roll(case['left_u'], 1))[:, None]
    rprev = torch.as_tensor(np.roll(case['right_u'], 1))[:, None]
    uxm = torch.where(ix == 0, extrap_quadh(uxp, utm, lprev), uxm)
    uxp = torch.where(ix == nx - 1, extrap_quadh(uxm, utm, rprev), uxp)
    d_now = (u - utm) / dt
    d_prev = torch.where(it == 1,
                         torch.as_tensor(case['init_ut'])[None, :],
                         (utm - utmm) / dt)
    u_tt = (d_now - d_prev) / dt
    u_xx = (uxm - 2 * utm + uxp) / dx**2
    f = u_tt - c[None, :]**2 * u_xx
    u0 = torch.as_tensor(case['init_u'] + 0.5 * dt * case['init_ut'])
    return torch.where(it == 0, u - u0[None, :], f)


def inverse_loss_grad(u, c, case, k_data=1.0, k_reg=1e-2):
    """(loss, split [wave, data, reg], grad_u, grad_c) in float64."""
    u = torch.tensor(np.asarray(u, dtype=np.float64), requires_grad=True)
    c = torch.tensor(np.asarray(c, dtype=np.float64), requires_grad=True)
    fw = wave_c_residual(u, c, case)
    mask = torch.zeros(case['nx'], dtype=torch.float64)
    mask[case['sensors']] = 1.
    d = torch.zeros((case['nt'], case['nx']), dtype=torch.float64)
    d[:, case['sensors']] = torch.as_tensor(case['data'])
    fd = k_data * mask[None, :] * (u - d)
    dc = (c[1:] - c[:-1]) / case['dx']
    split = [torch.mean(fw**2), torch.mean(fd**2),
             k_reg**2 * torch.mean(dc**2)]
    loss = sum(split)
    gu, gc = torch.autograd.grad(loss, [u, c])
    return (loss.item(), [s.item() for s in split], gu.numpy(), gc.numpy())
