"""Pins of the oracle functions that the FP64 oracle's other tests leave open:
lagged friction, rigid-shape signed distances, the triangle-vs-capsule
barrier, the local (lambda, s) solve and the GMW81 correction.

Each pin is a known answer from the reference's own tests (file:line), a
closed form, a finite difference or a mathematical property of the converged
mixed-FEM solution, chosen so that a plausible mistake (a sign, a dropped
factor, a transposed operand, eps_v instead of eps_v dt) fails one of them.
The planted-mistake tests at the end restate a function with one such mistake
and show that the pin catches it.
"""
import numpy as np
import pytest
import scipy.linalg
from scipy.spatial.transform import Rotation

from oracle import Kernels, Oracle
from mfem_amd import scene

K = Kernels(64)


def _fd_grad(f, x, h):
    x = np.asarray(x, float)
    g = np.zeros_like(x)
    for k in range(x.size):
        e = np.zeros_like(x)
        e.flat[k] = h
        g.flat[k] = (f(x + e) - f(x - e)) / (2 * h)
    return g


def _unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


# --------------------------------------------------------------------------
# Friction (src/mfem/refinement/contact.py:34-59, 304-387)
# --------------------------------------------------------------------------
def _slip_point(rng, n, y):
    """x_start, x with tangential slip |u_T| = y and a random normal part."""
    xs = rng.normal(size=3)
    t = _unit(np.cross(n, rng.normal(size=3)))
    return xs, xs + y * t + rng.normal() * 0.3 * y * n


@pytest.mark.parametrize("regime", ["static", "sliding"])
@pytest.mark.parametrize("seed", range(4))
def test_friction_gradient_is_energy_derivative(regime, seed):
    rng = np.random.default_rng(seed)
    n = _unit(rng.normal(size=3))
    ev, lamn, mu = 1e-2, 3.7, 0.4
    y = (0.35 if regime == "static" else 2.5) * ev
    xs, x = _slip_point(rng, n, y)
    E, g, H = K.friction(x, xs, n, lamn, mu, ev)
    fd = _fd_grad(lambda z: K.friction(z, xs, n, lamn, mu, ev)[0], x, 1e-6 * ev)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max())
    # no force along the (lagged) normal
    assert abs(g @ n) < 1e-12 * np.linalg.norm(g)


@pytest.mark.parametrize("seed", range(3))
def test_friction_closed_forms_and_limits(seed):
    rng = np.random.default_rng(10 + seed)
    n = _unit(rng.normal(size=3))
    ev, lamn, mu = 2e-3, 1.3, 0.3
    sc = mu * lamn
    # sliding: Coulomb, E = mu lamn |u_T|, |grad| = mu lamn
    xs, x = _slip_point(rng, n, 4 * ev)
    E, g, _ = K.friction(x, xs, n, lamn, mu, ev)
    assert E == pytest.approx(sc * 4 * ev, rel=1e-12)
    assert np.linalg.norm(g) == pytest.approx(sc, rel=1e-12)
    # at rest: f0(0) = ev / 3
    E0, g0, H0 = K.friction(xs, xs, n, lamn, mu, ev)
    assert E0 == pytest.approx(sc * ev / 3, rel=1e-12)
    # f1(y) / y -> 2 / ev as y -> 0: H = mu lamn (2 / ev) (I - n n^T)
    np.testing.assert_allclose(H0, sc * 2 / ev * (np.eye(3) - np.outer(n, n)), rtol=1e-12, atol=1e-12)
    xs, x = _slip_point(rng, n, 1e-9 * ev)
    _, g, _ = K.friction(x, xs, n, lamn, mu, ev)
    ut = (np.eye(3) - np.outer(n, n)) @ (x - xs)
    np.testing.assert_allclose(g, sc * 2 / ev * ut, rtol=1e-6)
    # C1 at y = ev: value and slope continuous from both sides
    t = _unit(np.cross(n, rng.normal(size=3)))
    vals = [K.friction(xs + y * t, xs, n, lamn, mu, ev) for y in (ev * (1 - 1e-9), ev * (1 + 1e-9))]
    assert vals[0][0] == pytest.approx(vals[1][0], rel=1e-8)
    np.testing.assert_allclose(vals[0][1], vals[1][1], rtol=1e-7)
    assert vals[0][0] == pytest.approx(sc * ev, rel=1e-8)


@pytest.mark.parametrize("seed", range(4))
def test_friction_hessian_is_lagged_projection(seed):
    # The reference's Hessian mu lamn (f1/y)(I - n n^T) is PSD and equals the
    # exact Hessian on tangent directions orthogonal to the slip; in the
    # static regime the exact Hessian differs from it by exactly
    # -mu lamn / (ev^2 y) u_T u_T^T (f1'(y) - f1(y)/y = -y / ev^2).
    rng = np.random.default_rng(20 + seed)
    n = _unit(rng.normal(size=3))
    ev, lamn, mu = 1e-2, 2.0, 0.5
    y = 0.5 * ev
    xs, x = _slip_point(rng, n, y)
    _, _, H = K.friction(x, xs, n, lamn, mu, ev)
    Hfd = np.stack([_fd_grad(lambda z: K.friction(z, xs, n, lamn, mu, ev)[1][k], x, 1e-6 * ev) for k in range(3)])
    ut = (np.eye(3) - np.outer(n, n)) @ (x - xs)
    assert np.linalg.norm(ut) == pytest.approx(y, rel=1e-12)
    np.testing.assert_allclose(Hfd, H - mu * lamn / (ev * ev * y) * np.outer(ut, ut), rtol=1e-5,
                               atol=1e-6 * np.abs(H).max())
    w = _unit(np.cross(n, ut))
    assert w @ H @ w == pytest.approx(w @ Hfd @ w, rel=1e-5)
    np.testing.assert_allclose(H @ n, 0, atol=1e-12 * np.abs(H).max())
    assert np.linalg.eigvalsh(H).min() > -1e-12 * np.abs(H).max()


def test_friction_in_the_solver_uses_eps_v_dt_and_the_lagged_normal_force():
    # One tet resting on a tilted plane: the normal force and normal are
    # frozen at the start of the step, lamn = kappa max(-b'(d), 0) of the
    # start distance, and the slip potential uses ev = eps_v * dt.
    x = np.array([[0.0, 0.0, 0.0], [0.05, 0.0, 0.0], [0.0, 0.05, 0.0], [0.0, 0.0, 0.05]])
    rot = Rotation.from_euler("xy", [0.3, -0.2])
    shape = scene.Shape(scene.SHAPE_PLANE, pos=np.array([0.01, -0.02, -0.004]), quat=rot.as_quat(), mu=0.25)
    nrm = rot.apply([0.0, 0.0, 1.0])
    x = x - (nrm @ (x[0] - shape.pos) - 4e-4) * nrm  # vertex 0 at 0.4 mm from the plane
    m = scene.build_model(x, np.array([[0, 1, 2, 3]]), np.array([[1e4, 1e4, 0.0]]), 1000.0, shapes=[shape])
    d0, d1, kappa, eps_v, dt = 1e-4, 1e-3, 7.0, 3e-3, 0.02
    o = Oracle(m, 64, params=dict(contact_d0=d0, contact_d1=d1, contact_stiffness=kappa, friction_eps_v=eps_v,
                                  friction_mu=[0.25], energy="arap"))
    dist = (x - shape.pos) @ nrm
    assert dist[0] == pytest.approx(4e-4, rel=1e-9)
    o.begin_step(dt)
    for i, di in enumerate(dist):
        want = kappa * max(-K.barrier(di, d0, d1)[1], 0.0) if di < d1 else 0.0
        assert o.buf("lag_fn")[i] == pytest.approx(want, rel=1e-10, abs=1e-300)
        np.testing.assert_allclose(o.buf("lag_n")[3 * i:3 * i + 3], nrm if di < d1 else 0.0, atol=1e-12)
    lamn = o.buf("lag_fn")[0]
    assert lamn > 0
    # slide vertex 0 tangentially by y < ev and read the friction energy
    t = _unit(np.cross(nrm, [1.0, 0.0, 0.0]))
    ev = eps_v * dt
    y = 0.4 * ev
    xn = o.x.copy()
    xn[0] += y * t + 1e-5 * nrm
    o.buf("x")[:] = xn.ravel()
    o.update_system(dt)
    f0 = -y ** 3 / (3 * ev ** 2) + y ** 2 / ev + ev / 3
    assert o.buf("eD")[0] == pytest.approx(0.25 * lamn * f0, rel=1e-9)
    np.testing.assert_allclose(o.buf("gD")[:3], 0.25 * lamn * (-y / ev ** 2 + 2 / ev) * y * t, rtol=1e-8,
                               atol=1e-12 * lamn)


# --------------------------------------------------------------------------
# Shape signed distances (src/mfem/ipc/distance.py:4-57, contact.py:62-164)
# --------------------------------------------------------------------------
def _pose(rng):
    return np.concatenate([rng.normal(size=3) * 0.2, Rotation.random(random_state=rng).as_quat()])


@pytest.mark.parametrize("seed", range(6))
def test_plane_sdf_posed(seed):
    rng = np.random.default_rng(30 + seed)
    xf = _pose(rng)
    n = Rotation.from_quat(xf[3:]).apply([0.0, 0.0, 1.0])
    x = rng.normal(size=3)
    d, g, H = K.shape_sdf(scene.SHAPE_PLANE, xf, np.zeros(3), x)
    assert d == pytest.approx(n @ (x - xf[:3]), rel=1e-12, abs=1e-14)
    np.testing.assert_allclose(g, n, atol=1e-12)
    np.testing.assert_allclose(H, 0, atol=1e-14)


def _segment_distance(x, p0, p1):
    ax = p1 - p0
    t = np.clip((x - p0) @ ax / (ax @ ax), 0.0, 1.0)
    return np.linalg.norm(x - (p0 + t * ax))


@pytest.mark.parametrize("where", ["cylinder", "cap"])
@pytest.mark.parametrize("seed", range(5))
def test_capsule_sdf_posed(where, seed):
    rng = np.random.default_rng(40 + seed)
    xf = _pose(rng)
    rad, hh = 0.05, 0.2
    R = Rotation.from_quat(xf[3:])
    axis = R.apply([0.0, 0.0, 1.0])
    p0, p1 = xf[:3] - hh * axis, xf[:3] + hh * axis
    side = _unit(np.cross(axis, rng.normal(size=3)))
    if where == "cylinder":
        x = xf[:3] + rng.uniform(-0.8, 0.8) * hh * axis + rng.uniform(0.01, 0.2) * side
    else:
        x = xf[:3] + (hh + rng.uniform(0.02, 0.1)) * axis * rng.choice([-1, 1]) + rng.uniform(0.0, 0.1) * side
    d, g, H = K.shape_sdf(scene.SHAPE_CAPSULE, xf, [rad, hh, 0.0], x)
    assert d == pytest.approx(_segment_distance(x, p0, p1) - rad, rel=1e-12)
    f = lambda z: K.shape_sdf(scene.SHAPE_CAPSULE, xf, [rad, hh, 0.0], z)[0]
    np.testing.assert_allclose(g, _fd_grad(f, x, 1e-6), atol=1e-8)
    assert np.linalg.norm(g) == pytest.approx(1.0, rel=1e-12)
    Hfd = np.stack([_fd_grad(lambda z: K.shape_sdf(scene.SHAPE_CAPSULE, xf, [rad, hh, 0.0], z)[1][k], x, 1e-6)
                    for k in range(3)])
    np.testing.assert_allclose(H, Hfd, atol=1e-5 * max(1.0, np.abs(H).max()))


def test_capsule_sdf_known_answers():
    # local frame: axis along z, radius 0.1, half height 0.5
    xf = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0])
    sc = [0.1, 0.5, 0.0]
    d, g, _ = K.shape_sdf(scene.SHAPE_CAPSULE, xf, sc, [0.3, 0.0, 0.2])
    assert d == pytest.approx(0.2) and np.allclose(g, [1, 0, 0])
    d, g, _ = K.shape_sdf(scene.SHAPE_CAPSULE, xf, sc, [0.0, 0.0, -0.9])
    assert d == pytest.approx(0.3) and np.allclose(g, [0, 0, -1])
    # a capsule lying along world x (90 degrees about y), 1 unit up
    q = Rotation.from_euler("y", 90, degrees=True).as_quat()
    d, g, _ = K.shape_sdf(scene.SHAPE_CAPSULE, np.concatenate([[0.0, 0.0, 1.0], q]), sc, [0.4, 0.0, 0.6])
    assert d == pytest.approx(0.3) and np.allclose(g, [0, 0, -1], atol=1e-12)
    d, _, _ = K.shape_sdf(scene.SHAPE_CAPSULE, np.concatenate([[0.0, 0.0, 1.0], q]), sc, [0.8, 0.0, 1.0])
    assert d == pytest.approx(0.2)


def test_solver_distance_is_minimum_over_shapes():
    # contact.py:62-95: minimum over shapes; the gradient of the winner
    x = np.array([[0.0, 0.0, 0.0], [0.05, 0.0, 0.0], [0.0, 0.05, 0.0], [0.0, 0.0, 0.05]]) + [0.0, 0.0, 0.003]
    cap = scene.Shape(scene.SHAPE_CAPSULE, pos=np.array([0.0, 0.0, 0.05 + 0.02 + 0.004]),
                      quat=Rotation.from_euler("x", 90, degrees=True).as_quat(), scale=np.array([0.02, 0.1, 0.0]))
    m = scene.build_model(x, np.array([[0, 1, 2, 3]]), np.array([[1e4, 1e4, 0.0]]), 1000.0,
                          shapes=[scene.Shape(scene.SHAPE_PLANE), cap])
    o = Oracle(m, 64, params=dict(contact_d0=1e-4, contact_d1=1e-2, contact_stiffness=1.0, energy="arap"))
    o.update_system(0.01)
    dist = o.buf("dist")
    # vertices 0-2 are 3 mm above the plane; vertex 3 (z = 53 mm) is 1 mm
    # below the capsule's surface (axis at z = 74 mm, radius 20 mm)
    np.testing.assert_allclose(dist[:3], [0.003, 0.003, 0.003], rtol=1e-9)
    assert dist[3] == pytest.approx(0.001, rel=1e-9)


# --------------------------------------------------------------------------
# Triangle vs capsule barrier (tests/test_tri_contact.py:154-277 of the
# reference, same constants and configurations)
# --------------------------------------------------------------------------
D0, D1, STIFF, RADIUS, HALF_HEIGHT = 0.05, 0.2, 10.0, 0.1, 0.5
BIG_TRI = np.array([(-3.0, -3.0, 0.0), (3.0, -3.0, 0.0), (0.0, 3.0, 0.0)])
GAP = 0.08


def _capsule(pos, rot=(0.0, 0.0, 0.0, 1.0)):
    return (scene.SHAPE_CAPSULE, np.concatenate([pos, rot]), np.array([RADIUS, HALF_HEIGHT, 0.0]))


def _tri(verts, pos, rot=(0.0, 0.0, 0.0, 1.0)):
    return K.tri_barrier(np.asarray(verts, float), [_capsule(np.asarray(pos, float), rot)], D0, D1, STIFF)


def test_tri_barrier_catches_what_vertices_miss():
    pos = np.array([0.0, -1.0, GAP + RADIUS + HALF_HEIGHT])
    for v in BIG_TRI:
        assert K.shape_sdf(scene.SHAPE_CAPSULE, np.concatenate([pos, [0, 0, 0, 1]]), [RADIUS, HALF_HEIGHT, 0], v)[0] > D1
    e, g, H, d, bary, n = _tri(BIG_TRI, pos)
    assert d == pytest.approx(GAP, abs=1e-12)
    np.testing.assert_allclose(n, [0.0, 0.0, -1.0], atol=1e-12)
    assert e > 0 and e == pytest.approx(STIFF * K.barrier(GAP, D0, D1)[0], rel=1e-12)
    assert bary.sum() == pytest.approx(1.0, abs=1e-12)
    # moving the face away from the capsule (-z) lowers the energy
    assert (g[:, 2] > 0).all()
    tot = g.sum(axis=0)
    assert np.abs(tot[:2]).max() < 1e-12 * abs(tot[2])


def test_tri_barrier_axis_through_face_is_deepest():
    e_touch, *_ = _tri(BIG_TRI, (0.0, -1.0, HALF_HEIGHT + RADIUS))
    e_pierce, _, _, d, _, _ = _tri(BIG_TRI, (0.0, -1.0, 0.0))
    assert d == pytest.approx(-RADIUS, abs=1e-12)
    assert e_pierce > e_touch > 0.0


def _quat_axis_angle(axis, degrees):
    return Rotation.from_rotvec(np.radians(degrees) * _unit(axis)).as_quat()


@pytest.mark.parametrize("pos,rot", [
    ((0.3, -0.8, GAP + RADIUS + HALF_HEIGHT), _unit([0.1, 0.05, 0.0, 0.9937])),
    ((0.4, -3.3, 0.1), (Rotation.from_quat(_quat_axis_angle([0, 0, 1], 15.0))
                        * Rotation.from_quat(_quat_axis_angle([0, 1, 0], 90.0))).as_quat()),
], ids=["face", "edge"])
def test_tri_barrier_gradient_fd_and_psd_hessian(pos, rot):
    e, g, H, d, bary, _ = _tri(BIG_TRI, pos, tuple(rot))
    assert 0.0 < d < D1
    assert bary.max() < 1.0 - 1e-4
    f = lambda v: _tri(v.reshape(3, 3), pos, tuple(rot))[0]
    fd = _fd_grad(f, BIG_TRI.ravel(), 1e-3).reshape(3, 3)
    # the reference's own bound (h = 1e-3, 2e-2 of the largest entry) and a
    # tight one at FP64 with a small step
    assert np.abs(g - fd).max() < 2e-2 * np.abs(fd).max()
    fd6 = _fd_grad(f, BIG_TRI.ravel(), 1e-6).reshape(3, 3)
    np.testing.assert_allclose(g, fd6, rtol=1e-6, atol=1e-7 * np.abs(fd6).max())
    H9 = H.transpose(0, 2, 1, 3).reshape(9, 9)
    np.testing.assert_allclose(H9, H9.T, atol=1e-12 * np.abs(H9).max())
    assert np.linalg.eigvalsh(H9).min() > -1e-10 * np.abs(H9).max()
    # Gauss-Newton form: w_a w_b kappa max(b'', 0) n n^T
    b2 = max(K.barrier(d, D0, D1)[2], 0.0)
    _, _, _, _, w, n = _tri(BIG_TRI, pos, tuple(rot))
    np.testing.assert_allclose(H9, STIFF * b2 * np.kron(np.outer(w, w), np.outer(n, n)), rtol=1e-12, atol=1e-14)


def test_tri_barrier_in_the_solver_matches_the_single_triangle_function():
    # a one-tet body whose bottom face is poked by a capsule: the oracle's
    # solver folds the triangle barrier into eB / gB exactly as the single
    # triangle function scatters it
    x = np.array([[0.0, 0.0, 0.0], [0.3, 0.0, 0.0], [0.0, 0.3, 0.0], [0.0, 0.0, 0.3]]) + [0.0, 0.0, 1.0]
    pos = np.array([0.08, 0.07, 1.0 - 0.04 - 0.02])
    cap = scene.Shape(scene.SHAPE_CAPSULE, pos=pos, quat=Rotation.from_euler("y", 90, degrees=True).as_quat(),
                      scale=np.array([0.02, 0.05, 0.0]))
    m = scene.build_model(x, np.array([[0, 1, 2, 3]]), np.array([[1e4, 1e4, 0.0]]), 1000.0, shapes=[cap])
    o = Oracle(m, 64, params=dict(contact_d0=1e-3, contact_d1=0.05, contact_stiffness=3.0, energy="arap",
                                  tri_contact=1))
    o.update_system(0.01)
    tris = m.tri_indices
    eB = np.zeros(4)
    gB = np.zeros((4, 3))
    for t in tris:
        e, g, _, d, w, _ = K.tri_barrier(x[t], [(cap.type, np.concatenate([cap.pos, cap.quat]), cap.scale)], 1e-3,
                                         0.05, 3.0)
        eB[t] += w * e
        gB[t] += g
    assert eB.sum() > 0
    np.testing.assert_allclose(o.buf("eB"), eB, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(o.buf("gB").reshape(-1, 3), gB, rtol=1e-12, atol=1e-15)


# --------------------------------------------------------------------------
# Local solve: the converged mixed solution satisfies the stationarity
# conditions of the Lagrangian, and equals the displacement-form equilibrium
# (src/mfem/refinement/kernels.py:252-283, solver.py:537)
# --------------------------------------------------------------------------
def _converged_beam(energy="neohookean", iters=30, line_search=True):
    m = scene.beam_scene(cells=(3, 2, 2), energy=energy)
    o = Oracle(m, 64, params=dict(cg_tolerance=1e-13, cg_max_iterations=5000, line_search_max_iterations=40,
                                  line_search=line_search))
    dt = 1.0 / 60.0
    o.begin_step(dt)
    for _ in range(iters):
        o.update_system(dt)
        o.assemble(dt)
        o.global_solve()
        o.local_solve()
        if line_search:
            o.line_search(dt)
        else:
            o.buf("x")[:] += o.buf("dx")
            o.buf("s")[:] += o.buf("ds")
    o.update_system(dt)
    o.assemble(dt)
    o.global_solve()
    o.local_solve()
    return m, o, dt


# ARAP: the reference's energy uses a different I2 than its gradient, so its
# merit is not the potential of the iteration; it runs without line search.
@pytest.mark.parametrize("energy,ls", [("neohookean", True), ("arap", False)])
def test_local_solve_multipliers_are_stationary_at_convergence(energy, ls):
    m, o, dt = _converged_beam(energy, line_search=ls)
    nt, nv = o.nt, o.nv
    sym = np.array([1, 1, 1, 2, 2, 2], float)
    c = o.buf("c").reshape(nt, 6)
    s = o.buf("s").reshape(nt, 6)
    lam = o.buf("lam").reshape(nt, 6)
    gs = o.buf("gs").reshape(nt, 6)
    G = o.buf("G").reshape(nt, 6, 12)
    # converged: the Newton step and the stretch update vanish
    assert np.abs(o.buf("dx")).max() < 1e-9
    assert np.abs(o.buf("ds")).max() < 1e-9
    # constraint satisfied: s = sym6(S(x))
    assert np.abs(c).max() < 1e-9 * np.abs(s).max()
    # stationarity in s: dPsi/ds + SYM lambda = 0
    assert np.abs(gs + sym * lam).max() < 1e-7 * np.abs(gs).max()
    # stationarity in x (momentum balance with the multipliers as element
    # forces): g_x + grad B + grad D + sum_e G_e^T lambda_e = 0, where
    # G = -dc/dx, so that -lambda . c contributes +G^T lambda
    fx = (o.buf("gx") + o.buf("gB") + o.buf("gD")).reshape(nv, 3).copy()
    for e, t in enumerate(m.tet_indices):
        fx[t] += (G[e].T @ lam[e]).reshape(4, 3)
    scale = np.abs(o.buf("gx")).max()
    assert np.abs(fx).max() < 1e-6 * scale
    if ls:
        # and the merit's x-gradient (finite differences) vanishes
        x = o.x.copy().ravel()
        f = lambda z: o.merit(z.reshape(-1, 3), s, dt)
        for k in np.random.default_rng(0).choice(x.size, 12, replace=False):
            e = np.zeros(x.size)
            e[k] = 1e-6
            gk = (f(x + e) - f(x - e)) / 2e-6
            assert abs(gk) < 1e-5 * scale, (k, gk, scale)


def test_converged_mixed_solution_is_displacement_equilibrium():
    # With s = S(x) the mixed stationary point is the minimiser of the
    # displacement-form energy sum Psi(S(x)) + kinetic + attachment terms, so
    # its gradient there (finite differences through scipy's polar
    # decomposition, not the oracle's kinematics) vanishes.
    m, o, dt = _converged_beam("neohookean")
    x = o.x.copy()
    xt = o.buf("x_tilde").reshape(-1, 3).copy()
    pins = np.nonzero(m.particle_inv_mass == 0)[0]
    mu, lam = m.tet_materials[0, 0], m.tet_materials[0, 1]
    vols = 1.0 / (6.0 * np.linalg.det(m.tet_poses))

    def total(y):
        y = y.reshape(-1, 3)
        E = 0.5 * np.sum(m.particle_mass[:, None] * (y - xt) ** 2) / dt ** 2
        E += 0.5 * 10.0 * np.sum((y[pins] - m.particle_q[pins]) ** 2) / dt ** 2
        for e, t in enumerate(m.tet_indices):
            F = np.stack([y[t[1]] - y[t[0]], y[t[2]] - y[t[0]], y[t[3]] - y[t[0]]], 1) @ m.tet_poses[e]
            _, P = scipy.linalg.polar(F)
            J = np.linalg.det(P)
            E += vols[e] * (mu / 2 * (np.sum(P * P) - 3) - mu * (J - 1) + lam / 2 * (J - 1) ** 2)
        return E

    scale = np.abs(m.particle_mass).max() * 9.81
    idx = np.random.default_rng(1).choice(x.size, 10, replace=False)
    for k in idx:
        e = np.zeros(x.size)
        e[k] = 1e-7
        gk = (total(x.ravel() + e) - total(x.ravel() - e)) / 2e-7
        assert abs(gk) < 1e-4 * scale, (k, gk)


# --------------------------------------------------------------------------
# GMW81 (src/mfem/numerical/gmw81.py:18-86): an independent LDL^T
# restatement of Gill, Murray & Wright (Practical Optimization, 1981,
# sec. 4.4.2.2) and a hand-worked 2 x 2 case
# --------------------------------------------------------------------------
def gmw81_ldl(A, eps=1.19e-7, xi_term=True):
    """Modified LDL^T: d_j = max(|c_jj|, theta_j^2 / beta^2, delta) with
    beta^2 = max(gamma, xi / sqrt(n^2 - 1), eps), delta = sqrt(eps) (the
    reference's constants).  Returns L~ = L diag(sqrt(d)) and E = L~L~^T - A."""
    A = np.asarray(A, float)
    n = A.shape[0]
    gamma = np.abs(np.diag(A)).max()
    xi = np.abs(A[np.triu_indices(n, 1)]).max() if n > 1 else 0.0
    beta2 = max(gamma, xi / np.sqrt(n * n - 1) if xi_term else 0.0, eps)
    delta = np.sqrt(eps)
    L = np.eye(n)
    d = np.zeros(n)
    for j in range(n):
        c = np.array([A[i, j] - sum(d[s] * L[i, s] * L[j, s] for s in range(j)) for i in range(n)])
        theta = np.abs(c[j + 1:]).max() if j < n - 1 else 0.0
        d[j] = max(abs(c[j]), theta ** 2 / beta2, delta)
        L[j + 1:, j] = c[j + 1:] / d[j]
    Lt = L * np.sqrt(d)[None, :]
    return Lt, Lt @ Lt.T - A


def test_gmw81_hand_worked_2x2():
    # A = [[1, 2], [2, 1]] (eigenvalues 3, -1): gamma = 1, xi = 2,
    # beta^2 = 2 / sqrt(3); d0 = theta0^2 / beta^2 = 2 sqrt(3), l10 = 1 / sqrt(3),
    # c11 = 1 - 2 sqrt(3) / 3 < 0, d1 = |c11|  =>  E = diag(2 sqrt(3) - 1, 4 sqrt(3) / 3 - 2)
    A = np.array([[1.0, 2.0], [2.0, 1.0]])
    L = K.gmw81(A)
    r3 = np.sqrt(3.0)
    np.testing.assert_allclose(L, [[np.sqrt(2 * r3), 0.0], [2.0 / np.sqrt(2 * r3), np.sqrt(2 * r3 / 3 - 1)]],
                               rtol=1e-12)
    fixed, inv = K.psd_fix(A)
    np.testing.assert_allclose(fixed - A, np.diag([2 * r3 - 1, 4 * r3 / 3 - 2]), atol=1e-12)
    np.testing.assert_allclose(fixed @ inv, np.eye(2), atol=1e-12)


@pytest.mark.parametrize("which", ["A3_non_pd", "A6_non_pd", "random"])
def test_gmw81_matches_textbook_ldl(which):
    from test_oracle import A3_non_pd, A6_non_pd
    if which == "A3_non_pd":
        mats = [A3_non_pd]
    elif which == "A6_non_pd":
        mats = [np.asarray(A6_non_pd, float)]
    else:
        rng = np.random.default_rng(7)
        mats = []
        for _ in range(20):
            Q = np.linalg.qr(rng.normal(size=(6, 6)))[0]
            mats.append(Q @ np.diag(rng.uniform(-3, 5, size=6)) @ Q.T)
    for A in mats:
        Lt, E = gmw81_ldl(A)
        np.testing.assert_allclose(K.gmw81(A), Lt, rtol=1e-10, atol=1e-12)
        fixed, _ = K.psd_fix(A)
        np.testing.assert_allclose(fixed - A, E, atol=1e-10 * max(1.0, np.abs(A).max()))
        assert np.abs(E - np.diag(np.diag(E))).max() < 1e-10 * np.abs(A).max()
        assert np.all(np.diag(E) >= -1e-12)
        if np.linalg.eigvalsh(A).min() < 0:
            assert np.diag(E).max() > 0.1  # indefinite input: a real correction


# --------------------------------------------------------------------------
# Planted mistakes: a restatement with one error each fails its pin
# --------------------------------------------------------------------------
def _friction_planted(x, xs, n, lamn, mu, ev, mistake):
    u = x - xs
    ut = u - (u @ n) * n if mistake != "no_projection" else u
    y = np.linalg.norm(ut)
    if mistake == "eps_v_not_dt":
        ev = ev * 100.0
    f0 = y if y >= ev else -y ** 3 / (3 * ev ** 2) + y ** 2 / ev + ev / 3
    f1y = 1 / y if y >= ev else -y / ev ** 2 + 2 / ev
    sc = mu * lamn if mistake != "no_mu" else lamn
    g = sc * f1y * ut * (-1.0 if mistake == "sign" else 1.0)
    return sc * f0, g


@pytest.mark.parametrize("mistake", ["sign", "no_projection"])
def test_planted_friction_mistakes_fail_the_fd_pin(mistake):
    rng = np.random.default_rng(3)
    n = _unit(rng.normal(size=3))
    ev, lamn, mu = 1e-2, 3.7, 0.4
    xs, x = _slip_point(rng, n, 0.35 * ev)
    _, g = _friction_planted(x, xs, n, lamn, mu, ev, mistake)
    fd = _fd_grad(lambda z: K.friction(z, xs, n, lamn, mu, ev)[0], x, 1e-6 * ev)
    assert not np.allclose(g, fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max())


@pytest.mark.parametrize("mistake", ["eps_v_not_dt", "no_mu"])
def test_planted_friction_mistakes_fail_the_closed_forms(mistake):
    n = np.array([0.0, 0.0, 1.0])
    ev, lamn, mu = 2e-3, 1.3, 0.3
    E0, _ = _friction_planted(np.zeros(3), np.zeros(3), n, lamn, mu, ev, mistake)
    assert E0 != pytest.approx(mu * lamn * ev / 3, rel=1e-6)


def test_planted_gmw81_mistake_fails_the_textbook_comparison():
    # beta^2 without the xi / sqrt(n^2 - 1) term
    A = np.array([[1.0, 2.0], [2.0, 1.0]])
    Lt, _ = gmw81_ldl(A, xi_term=False)
    assert not np.allclose(K.gmw81(A), Lt, rtol=1e-6)


def test_planted_multiplier_sign_fails_the_momentum_balance():
    m, o, dt = _converged_beam("neohookean")
    nt, nv = o.nt, o.nv
    G = o.buf("G").reshape(nt, 6, 12)
    lam = o.buf("lam").reshape(nt, 6)
    fx = (o.buf("gx") + o.buf("gB") + o.buf("gD")).reshape(nv, 3).copy()
    for e, t in enumerate(m.tet_indices):
        fx[t] -= (G[e].T @ lam[e]).reshape(4, 3)
    assert np.abs(fx).max() > 1e-2 * np.abs(o.buf("gx")).max()
