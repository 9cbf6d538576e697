"""The sweep half of the FP64 CPU oracle pinned by what the mathematics fixes:
physical fluxes against independently written F(u).n, the HLL upwind branches,
the time step against its closed form, the update against conservation,
free-stream preservation, the divergence of a smooth field and the wall ghost's
zero normal mass flux, the friction source against its projection in closed
form, the variance against a high-order integral of the closed form, and the
error norm on hand numbers.

Every physical formula in this file is written from the equations (Euler:
mass rho v.n, momentum rho v (v.n) + p n, energy (E + p) v.n; shallow water:
h v.n, h v (v.n) + g h^2/2 n), not taken from the oracle."""
import numpy as np
import pytest

GAMMA = 1.4
G = 9.81


# ---- independent physics --------------------------------------------------

def euler_cons(rho, vx, vy, p):
    return np.array([rho, rho * vx, rho * vy, p / (GAMMA - 1) + 0.5 * rho * (vx * vx + vy * vy)])


def euler_prim(u):
    rho = u[0]
    vx, vy = u[1] / rho, u[2] / rho
    p = (GAMMA - 1) * (u[3] - 0.5 * rho * (vx * vx + vy * vy))
    return rho, vx, vy, p


def euler_Fn(u, n):
    rho, vx, vy, p = euler_prim(u)
    vn = vx * n[0] + vy * n[1]
    return np.array([rho * vn, rho * vx * vn + p * n[0], rho * vy * vn + p * n[1], (u[3] + p) * vn])


def sw_cons(h, vx, vy):
    return np.array([h, h * vx, h * vy])


def sw_Fn(u, n):
    h, vx, vy = u[0], u[1] / u[0], u[2] / u[0]
    vn = vx * n[0] + vy * n[1]
    return np.array([h * vn, h * vx * vn + 0.5 * G * h * h * n[0], h * vy * vn + 0.5 * G * h * h * n[1]])


def sound(u, sw):
    if sw:
        return np.sqrt(G * u[0])
    rho, _, _, p = euler_prim(u)
    return np.sqrt(GAMMA * p / rho)


def normal_speed(u, nu):
    return (u[1] * nu[0] + u[2] * nu[1]) / u[0]


# ---- physical flux and HLL branches ----------------------------------------

def test_rest_state_flux_is_pressure_only(oracle):
    # rho = 1, v = 0, p = 1: F.n = (0, p n_x, p n_y, 0); Rusanov and HLL both reduce to it
    u = euler_cons(1.0, 0.0, 0.0, 1.0)
    nu = np.array([0.6, 0.8])
    n = 0.5 * nu
    for flux in (0, 1):
        got = oracle.flux_point(oracle.problem(1, flux=flux, gamma=GAMMA), u, u, nu, n)
        np.testing.assert_allclose(got, [0.0, 0.3, 0.4, 0.0], atol=1e-15)
    h = sw_cons(2.0, 0.0, 0.0)
    for flux in (0, 1):
        got = oracle.flux_point(oracle.problem(2, flux=flux, g=G), h, h, nu, n)
        np.testing.assert_allclose(got, [0.0, 0.5 * G * 4 * 0.3, 0.5 * G * 4 * 0.4], rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("sw", [False, True])
def test_physical_flux_on_random_states(oracle, sw):
    # consistency G(u, u) = F(u).n against the independent F, both flux flavours
    rng = np.random.default_rng(11)
    kind = 2 if sw else 1
    for _ in range(25):
        if sw:
            u = sw_cons(rng.uniform(0.3, 3), rng.uniform(-2, 2), rng.uniform(-2, 2))
        else:
            u = euler_cons(rng.uniform(0.3, 3), rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.3, 3))
        ang = rng.uniform(0, 2 * np.pi)
        nu = np.array([np.cos(ang), np.sin(ang)])
        n = rng.uniform(0.1, 2) * nu
        want = sw_Fn(u, n) if sw else euler_Fn(u, n)
        for flux in (0, 1):
            p = oracle.problem(kind, flux=flux, gamma=GAMMA, g=G)
            np.testing.assert_allclose(oracle.flux_point(p, u, u, nu, n), want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("sw", [False, True])
def test_hll_supersonic_branches_upwind(oracle, sw):
    """lambda_min = u_L.n - a_L >= 0 takes F(u_L).n, lambda_max = u_R.n + a_R <= 0
    takes F(u_R).n (euler2d.cpp:39-49, shallowwater2d.cpp:35-40), exactly, for
    any other state."""
    rng = np.random.default_rng(5)
    kind = 2 if sw else 1
    p = oracle.problem(kind, flux=0, gamma=GAMMA, g=G)
    hits = [0, 0]
    for _ in range(40):
        ang = rng.uniform(0, 2 * np.pi)
        nu = np.array([np.cos(ang), np.sin(ang)])
        t = np.array([-nu[1], nu[0]])
        n = 0.7 * nu
        def state(vn, vt):
            v = vn * nu + vt * t
            if sw:
                return sw_cons(rng.uniform(0.2, 1.0), v[0], v[1])
            return euler_cons(rng.uniform(0.5, 2), v[0], v[1], rng.uniform(0.2, 1.0))
        Fn = sw_Fn if sw else euler_Fn
        # supersonic out of the cell: normal speed far above the sound speed
        uL = state(rng.uniform(6, 9), rng.uniform(-1, 1))
        vR = state(rng.uniform(-3, 3), rng.uniform(-1, 1))
        assert normal_speed(uL, nu) - sound(uL, sw) >= 0
        np.testing.assert_allclose(oracle.flux_point(p, uL, vR, nu, n), Fn(uL, n), rtol=1e-14, atol=1e-14)
        hits[0] += 1
        # supersonic into the cell: u_R.n + a_R <= 0
        uL = state(rng.uniform(-3, 0), rng.uniform(-1, 1))
        vR = state(-rng.uniform(6, 9), rng.uniform(-1, 1))
        assert normal_speed(vR, nu) + sound(vR, sw) <= 0 and normal_speed(uL, nu) - sound(uL, sw) < 0
        np.testing.assert_allclose(oracle.flux_point(p, uL, vR, nu, n), Fn(vR, n), rtol=1e-14, atol=1e-14)
        hits[1] += 1
    assert hits == [40, 40]


@pytest.mark.parametrize("sw", [False, True])
def test_hll_and_rusanov_subsonic_closed_form(oracle, sw):
    """Subsonic HLL with the Davis speeds lambda_min = u_L.n - a_L, lambda_max =
    u_R.n + a_R (euler2d.cpp:36-47): (l+ F_L.n - l- F_R.n + l+ l- (v - u)|n|)/(l+ - l-);
    Rusanov: (F_L + F_R).n/2 - s_max (v - u)|n|/2."""
    rng = np.random.default_rng(7)
    kind = 2 if sw else 1
    for _ in range(25):
        if sw:
            u = sw_cons(rng.uniform(0.5, 2), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5))
            v = sw_cons(rng.uniform(0.5, 2), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5))
        else:
            u = euler_cons(rng.uniform(0.5, 2), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(0.5, 2))
            v = euler_cons(rng.uniform(0.5, 2), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(0.5, 2))
        ang = rng.uniform(0, 2 * np.pi)
        nu = np.array([np.cos(ang), np.sin(ang)])
        L = rng.uniform(0.1, 3)
        n = L * nu
        Fn = sw_Fn if sw else euler_Fn
        lm = normal_speed(u, nu) - sound(u, sw)
        lp = normal_speed(v, nu) + sound(v, sw)
        assert lm < 0 < lp
        hll = (lp * Fn(u, n) - lm * Fn(v, n) + lp * lm * (v - u) * L) / (lp - lm)
        got = oracle.flux_point(oracle.problem(kind, flux=0, gamma=GAMMA, g=G), u, v, nu, n)
        np.testing.assert_allclose(got, hll, rtol=1e-12, atol=1e-13)
        smax = max(abs(normal_speed(u, nu)) + sound(u, sw), abs(normal_speed(v, nu)) + sound(v, sw))
        rus = 0.5 * (Fn(u, n) + Fn(v, n)) - 0.5 * smax * (v - u) * L
        got = oracle.flux_point(oracle.problem(kind, flux=1, gamma=GAMMA, g=G), u, v, nu, n)
        np.testing.assert_allclose(got, rus, rtol=1e-12, atol=1e-13)


def test_burgers_fluxes_closed_form(oracle):
    # upwind (burgers.cpp:21-28): F(u) n if u nUnit > 0 else F(v) n
    p = oracle.problem(0, flux=0)
    assert oracle.flux_point(p, [3.0], [5.0], [1.0], [2.0])[0] == 0.5 * 9.0 * 2.0
    assert oracle.flux_point(p, [-3.0], [5.0], [1.0], [2.0])[0] == 0.5 * 25.0 * 2.0
    assert oracle.flux_point(p, [3.0], [5.0], [-1.0], [-2.0])[0] == 0.5 * 25.0 * -2.0
    # Lax-Friedrichs: (f(u) + f(v)) n / 2 - max(|u|, |v|) (v - u) |n| / 2
    p = oracle.problem(0, flux=1)
    got = oracle.flux_point(p, [3.0], [-5.0], [-1.0], [-0.5])[0]
    assert abs(got - (0.5 * (4.5 + 12.5) * -0.5 - 0.5 * 5.0 * (-8.0) * 0.5)) < 1e-14


# ---- time step ---------------------------------------------------------------

def test_compute_dt_closed_forms(oracle):
    # Burgers (burgers.cpp:114-122): min_k cfl dx / |u_k|
    p = oracle.problem(0, cfl=0.5)
    uq = np.array([[3.0], [-12.0], [7.0]])
    assert abs(oracle.compute_dt(p, uq, 0.01) - 0.5 * 0.01 / 12.0) < 1e-18
    # Euler 2D (euler2d.cpp:114-134): (cfl / dx) min over points of
    # 1 / max(|vx| + a, |vy| + a) with dx = maxEdge / area
    p = oracle.problem(1, cfl=0.4, gamma=GAMMA)
    states = [(1.0, 0.3, -0.2, 1.0), (0.5, -0.9, 0.1, 2.0), (2.0, 0.0, 1.5, 0.7)]
    uq = np.array([euler_cons(*s) for s in states])
    want = min(0.4 / 8.0 / (max(abs(vx), abs(vy)) + np.sqrt(GAMMA * pr / rho)) for rho, vx, vy, pr in states)
    assert abs(oracle.compute_dt(p, uq, 8.0) - want) < 1e-15 * want
    # shallow water: the same with a = sqrt(g h)
    p = oracle.problem(2, cfl=0.45, g=G)
    states = [(2.0, 0.4, -0.1), (0.5, -1.2, 0.3)]
    uq = np.array([sw_cons(*s) for s in states])
    want = min(0.45 / 3.0 / (max(abs(vx), abs(vy)) + np.sqrt(G * h)) for h, vx, vy in states)
    assert abs(oracle.compute_dt(p, uq, 3.0) - want) < 1e-15 * want


# ---- update ------------------------------------------------------------------

def periodic_grid(nx, ny, Lx=1.0, Ly=1.0, skew=0.0):
    """Doubly periodic quadrilateral grid (a torus): cell (i, j) = j*nx + i,
    edges j-, i+, j+, i- with outward normals.  skew shears the grid
    (x += skew*y) so that normals are not axis aligned; every cell's scaled
    normals still sum to zero."""
    N = nx * ny
    dx, dy = Lx / nx, Ly / ny
    # sheared parallelogram cells: edge vectors a = (dx, 0), b = (skew*dy, dy)
    a = np.array([dx, 0.0])
    b = np.array([skew * dy, dy])
    area = abs(a[0] * b[1] - a[1] * b[0])
    # outward normals (scaled by edge length) of edges j- (along a), i+ (along b), j+ (-a), i- (-b)
    def rot(e):  # outward for counter-clockwise traversal
        return np.array([e[1], -e[0]])
    sn = np.array([rot(a), rot(b), rot(-a), rot(-b)])
    length = np.hypot(sn[:, 0], sn[:, 1])
    un = sn / length[:, None]
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    ii, jj = ii.ravel(), jj.ravel()
    cid = lambda i, j: (j % ny) * nx + (i % nx)
    nbr = np.stack([cid(ii, jj - 1), cid(ii + 1, jj), cid(ii, jj + 1), cid(ii - 1, jj)], 1)
    centers = np.stack([(ii + 0.5) * dx + skew * (jj + 0.5) * dy, (jj + 0.5) * dy], 1)
    return dict(nbr=nbr, unit_normal=np.broadcast_to(un, (N, 4, 2)).copy(),
                normal=np.broadcast_to(sn, (N, 4, 2)).copy(), area=np.full(N, area),
                dx_param=np.full(N, length.max() / area), bc=np.zeros(N, np.int32), bnormal=np.zeros((N, 2)),
                center=centers)


def mesh_of(oracle, m):
    return oracle.MeshArrays(m["nbr"], m["unit_normal"], m["normal"], m["area"], m["dx_param"], m["bc"], m["bnormal"])


def tables_2d(oracle, N=2, q=3):
    idx, _ = oracle.index_set(2, [N])
    xi, wq, phi, phiw = oracle.tables(2, q, idx)
    return xi, phi, phiw


@pytest.mark.parametrize("kind,flux", [(1, 0), (1, 1), (2, 0), (2, 1)])
def test_uniform_free_stream_stays_uniform(oracle, kind, flux):
    """Each closed cell's scaled normals sum to zero and G(u, u) = F(u).n, so a
    uniform state (constant in x, varying in xi) is a steady solution on a
    periodic sheared grid."""
    m = periodic_grid(6, 5, skew=0.4)
    mesh = mesh_of(oracle, m)
    xi, phi, phiw = tables_2d(oracle)
    Q, M = phi.shape
    N = m["area"].size
    S = 4 if kind == 1 else 3
    # xi-dependent but x-uniform point values
    uq1 = np.array([euler_cons(1.0 + 0.2 * x0, 0.5 + 0.1 * x1, -0.3, 1.0 + 0.1 * x0 * x1) if S == 4
                    else sw_cons(1.5 + 0.2 * x0, 0.4, -0.2 + 0.1 * x1) for x0, x1 in xi])
    uq = np.broadcast_to(uq1, (N, Q, S)).copy()
    u = np.einsum("nks,ki->nis", uq, phiw)
    p = oracle.problem(kind, flux=flux, gamma=GAMMA, g=G)
    unew = oracle.advance(p, mesh, phiw, xi, 1e-3, u, uq)
    np.testing.assert_allclose(unew, u, rtol=0, atol=1e-14 * np.abs(u).max())


def test_hll_colliding_supersonic_states_take_the_left_flux(oracle):
    """The reference's branch order (euler2d.cpp:39-44) tests lambda_min >= 0
    first, so when u moves supersonically towards v and v supersonically
    towards u (lambda_min >= 0 and lambda_max <= 0) G(u, v, n) = F(u).n while
    -G(v, u, -n) = F(v).n: the two sides of the edge disagree.  The oracle
    keeps this as the reference has it."""
    p = oracle.problem(1, flux=0, gamma=GAMMA)
    nu = np.array([1.0, 0.0])
    n = 2.0 * nu
    u = euler_cons(1.0, 3.0, 0.0, 0.5)
    v = euler_cons(0.8, -3.0, 0.2, 0.6)
    assert normal_speed(u, nu) - sound(u, False) >= 0 and normal_speed(v, nu) + sound(v, False) <= 0
    np.testing.assert_allclose(oracle.flux_point(p, u, v, nu, n), euler_Fn(u, n), rtol=1e-15)
    np.testing.assert_allclose(-oracle.flux_point(p, v, u, -nu, -n), euler_Fn(v, n), rtol=1e-15)


@pytest.mark.parametrize("kind,flux", [(1, 0), (1, 1), (2, 0), (2, 1)])
def test_update_conserves_on_closed_mesh(oracle, kind, flux):
    """sum_j A_j u_j changes only by boundary fluxes: on a periodic mesh it is
    constant for every state and moment (G(u, v, n) = -G(v, u, -n)), while the
    cells themselves do change.  The states are subsonic: the reference's HLL
    is not conservative for colliding supersonic states (see the next test)."""
    m = periodic_grid(7, 6, skew=0.3)
    mesh = mesh_of(oracle, m)
    xi, phi, phiw = tables_2d(oracle)
    Q, M = phi.shape
    N = m["area"].size
    rng = np.random.default_rng(2)
    S = 4 if kind == 1 else 3
    uq = np.empty((N, Q, S))
    for j in range(N):
        for k in range(Q):
            uq[j, k] = (euler_cons(rng.uniform(0.5, 2), rng.uniform(-.3, .3), rng.uniform(-.3, .3), rng.uniform(0.5, 2))
                        if S == 4 else sw_cons(rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1)))
    u = np.einsum("nks,ki->nis", uq, phiw)
    p = oracle.problem(kind, flux=flux, gamma=GAMMA, g=G)
    unew = oracle.advance(p, mesh, phiw, xi, 1e-2, u, uq)
    before = np.einsum("j,jis->is", m["area"], u)
    after = np.einsum("j,jis->is", m["area"], unew)
    np.testing.assert_allclose(after, before, rtol=0, atol=1e-13 * np.abs(before).max())
    assert np.abs(unew - u).max() > 1e-4


def test_update_approximates_the_divergence_euler2d(oracle):
    """A smooth density wave on a periodic grid (v, p constant): (u - unew)/dt
    converges to div F(u) = (v . grad rho) (1, vx, vy, |v|^2/2) at first order
    for the Rusanov flux, which fixes the update's sign and the dt / A scaling."""
    vx, vy, pr = 0.7, -0.4, 1.0
    errs = []
    for nx in (32, 64, 128):
        m = periodic_grid(nx, nx, skew=0.0)
        mesh = mesh_of(oracle, m)
        N = m["area"].size
        xi = np.zeros((1, 1))
        phiw = np.ones((1, 1))
        c = m["center"]
        rho = 1.0 + 0.2 * np.sin(2 * np.pi * c[:, 0]) * np.cos(2 * np.pi * c[:, 1])
        uq = euler_cons(rho, vx, vy, pr).T[:, None, :].copy()
        u = uq.copy()
        dt = 1e-4
        unew = oracle.advance(oracle.problem(1, flux=1, gamma=GAMMA), mesh, phiw, xi, dt, u, uq)
        drx = 0.4 * np.pi * np.cos(2 * np.pi * c[:, 0]) * np.cos(2 * np.pi * c[:, 1])
        dry = -0.4 * np.pi * np.sin(2 * np.pi * c[:, 0]) * np.sin(2 * np.pi * c[:, 1])
        adv = vx * drx + vy * dry
        want = adv[:, None] * np.array([1.0, vx, vy, 0.5 * (vx * vx + vy * vy)])[None, :]
        got = (u - unew)[:, 0, :] / dt
        errs.append(np.abs(got - want).max() / np.abs(want).max())
    # Rusanov's dissipation, s_max dx |laplacian rho| / 2, is the leading error
    assert errs[-1] < 0.15
    assert errs[0] / errs[1] > 1.7 and errs[1] / errs[2] > 1.7


def test_update_approximates_the_divergence_burgers(oracle):
    """1D Lax-Friedrichs: (u - unew)/dt -> d/dx(u^2/2) = u u' at first order."""
    errs = []
    for N in (50, 100, 200):
        x = (np.arange(N) + 0.5) / N
        dx = 1.0 / N
        nbr = np.stack([(np.arange(N) - 1) % N, (np.arange(N) + 1) % N], 1)
        un = np.zeros((N, 2, 1))
        un[:, 0, 0], un[:, 1, 0] = -1.0, 1.0
        mesh = oracle.MeshArrays(nbr, un, un.copy(), np.full(N, dx), np.full(N, dx), np.zeros(N, np.int32),
                                 np.zeros((N, 1)))
        uu = 5.0 + np.sin(2 * np.pi * x)
        uq = uu[:, None, None]
        dt = 1e-5
        unew = oracle.advance(oracle.problem(0, flux=1), mesh, np.ones((1, 1)), np.zeros((1, 1)), dt, uq.copy(), uq)
        got = (uq - unew)[:, 0, 0] / dt
        want = uu * 2 * np.pi * np.cos(2 * np.pi * x)
        errs.append(np.abs(got - want).max() / np.abs(want).max())
    assert errs[-1] < 0.05
    assert errs[0] / errs[1] > 1.7 and errs[1] / errs[2] > 1.7


def one_cell_wall_mesh(oracle, bc, bn_angle=0.3):
    """One square cell of side 1 rotated by bn_angle whose edge 0 is a wall
    (neighbour id 1 = n_cells) and whose other edges see the cell itself."""
    c, s = np.cos(bn_angle), np.sin(bn_angle)
    R = np.array([[c, -s], [s, c]])
    sn = np.array([[0.0, -1.0], [1.0, 0.0], [0.0, 1.0], [-1.0, 0.0]]) @ R.T
    nbr = np.array([[1, 0, 0, 0]])
    return oracle.MeshArrays(nbr, sn[None], sn[None].copy(), np.ones(1), np.ones(1), np.array([bc], np.int32),
                             sn[0][None].copy()), sn


@pytest.mark.parametrize("flux", [0, 1])
def test_noslip_wall_ghost(oracle, flux):
    """expliciteuler.cpp:25-43: the ghost state mirrors out the normal velocity,
    keeps the tangential one and the pressure, so the wall edge carries no mass
    or energy flux and exactly the pressure p n in momentum."""
    mesh, sn = one_cell_wall_mesh(oracle, oracle.BC_NOSLIP)
    u0 = euler_cons(1.3, 0.8, -0.5, 0.9)
    uq = u0[None, None, :]
    dt = 0.01
    unew = oracle.advance(oracle.problem(1, flux=flux, gamma=GAMMA), mesh, np.ones((1, 1)), np.zeros((1, 2)), dt,
                          uq.copy(), uq)
    rhs = (uq - unew)[0, 0] / dt
    wall = rhs - euler_Fn(u0, sn[1] + sn[2] + sn[3])
    pr = euler_prim(u0)[3]
    np.testing.assert_allclose(wall, [0.0, pr * sn[0][0], pr * sn[0][1], 0.0], atol=1e-14)


@pytest.mark.parametrize("flux", [0, 1])
def test_swwall_ghost(oracle, flux):
    """expliciteuler.cpp:44-58: zero normal mass flux, momentum flux g h^2/2 n."""
    mesh, sn = one_cell_wall_mesh(oracle, oracle.BC_SWWALL, bn_angle=-0.7)
    u0 = sw_cons(1.7, -0.6, 0.9)
    uq = u0[None, None, :]
    dt = 0.01
    unew = oracle.advance(oracle.problem(2, flux=flux, g=G), mesh, np.ones((1, 1)), np.zeros((1, 2)), dt,
                          uq.copy(), uq)
    rhs = (uq - unew)[0, 0] / dt
    wall = rhs - sw_Fn(u0, sn[1] + sn[2] + sn[3])
    ph = 0.5 * G * 1.7 ** 2
    np.testing.assert_allclose(wall, [0.0, ph * sn[0][0], ph * sn[0][1]], atol=1e-13)


def test_dirichlet_cells_are_frozen(oracle):
    m = periodic_grid(4, 4)
    m["bc"][[0, 5, 15]] = oracle.BC_DIRICHLET
    mesh = mesh_of(oracle, m)
    xi, phi, phiw = tables_2d(oracle)
    rng = np.random.default_rng(4)
    N, (Q, M) = m["area"].size, phi.shape
    uq = np.array([[euler_cons(rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 2))
                    for _ in range(Q)] for _ in range(N)])
    u = np.einsum("nks,ki->nis", uq, phiw)
    unew = oracle.advance(oracle.problem(1, flux=0, gamma=GAMMA), mesh, phiw, xi, 0.01, u, uq)
    np.testing.assert_array_equal(unew[[0, 5, 15]], u[[0, 5, 15]])
    assert np.abs(unew[1] - u[1]).max() > 1e-5


def test_friction_source_projection(oracle):
    """Bed friction of the dam-break config, S = -c_f (1 + xi_d) |v| (0, hu, hv)/h
    per point, projected on the basis.  With a state constant in xi and a
    periodic mesh (fluxes cancel) the moments change by dt times
    -c_f |v| h v <(1 + xi_d) phi_i>: the mean for i = 0 and
    <xi_d sqrt(3) xi_d> = 1/sqrt(3) for the degree-1 moment of dimension d."""
    m = periodic_grid(3, 3)
    mesh = mesh_of(oracle, m)
    idx, _ = oracle.index_set(3, [2])
    xi, wq, phi, phiw = oracle.tables(3, 3, idx)
    Q, M = phi.shape
    N = m["area"].size
    h, vx, vy, cf, dim = 1.4, 0.6, -0.8, 0.01, 2
    uq = np.broadcast_to(sw_cons(h, vx, vy), (N, Q, 3)).copy()
    u = np.einsum("nks,ki->nis", uq, phiw)
    dt = 0.05
    p = oracle.problem(2, flux=0, g=G, friction=cf, friction_dim=dim)
    unew = oracle.advance(p, mesh, phiw, xi, dt, u, uq)
    du = (unew - u)[0]
    speed = np.hypot(vx, vy)
    i_lin = [i for i in range(M) if list(idx[i]) == [0, 0, 1]][0]
    for s, v in ((1, vx), (2, vy)):
        assert abs(du[0, s] - dt * -cf * speed * v) < 1e-15
        assert abs(du[i_lin, s] - dt * -cf * speed * v / np.sqrt(3)) < 1e-15
        others = [i for i in range(M) if i not in (0, i_lin)]
        assert np.abs(du[others, s]).max() < 1e-16
    assert np.abs(du[:, 0]).max() < 1e-16


# ---- mean, variance and the error norm ------------------------------------

def test_mean_and_variance_against_high_order_integral(oracle):
    """Bounded barrier, one uncertain dimension: U(xi) = u- + (u+ - u-)
    / (1 + exp(-lambda . phi(xi))) with phi_i = sqrt(2i+1) P_i; E = 1/2 int U,
    Var = 1/2 int (U - E)^2, integrated with numpy's 200-point Gauss rule."""
    from numpy.polynomial import legendre as npl

    um, up = 2.9, 12.1
    c = oracle.closure(0, u_minus=um, u_plus=up)
    idx, _ = oracle.index_set(1, [3])
    xi, wq, phi, phiw = oracle.tables(1, 40, idx)
    lam = np.array([[[0.4], [1.3], [-0.6], [0.25]], [[-1.0], [0.2], [0.9], [0.0]]])
    mv = oracle.mean_var(c, phi, phiw, lam)
    X, W = npl.leggauss(200)
    basis = np.stack([np.sqrt(2 * i + 1) * npl.legval(X, np.eye(4)[i]) for i in range(4)], 1)
    for j in range(2):
        U = um + (up - um) / (1 + np.exp(-(basis @ lam[j, :, 0])))
        E = 0.5 * (W * U).sum()
        V = 0.5 * (W * (U - E) ** 2).sum()
        assert abs(mv[j, 0] - E) < 1e-12 * E
        assert abs(mv[j, 1] - V) < 1e-10 * V


def test_mean_and_variance_euler2d_cell_of_known_moments(oracle):
    """For a dual with only constant terms U is the same at every point:
    E = U, Var = 0; with a linear xi term the variance is positive."""
    c = oracle.closure(2)
    idx, _ = oracle.index_set(2, [3])
    xi, wq, phi, phiw = oracle.tables(2, 5, idx)
    M = phi.shape[1]
    lam = np.zeros((1, M, 4))
    lam[0, 0] = [-3.0, 0.4, 0.2, -1.3]
    mv = oracle.mean_var(c, phi, phiw, lam)
    np.testing.assert_allclose(mv[0, :4], oracle.U(c, lam[0, 0]), rtol=1e-13)
    assert np.abs(mv[0, 4:]).max() < 1e-25
    lam[0, 1, 0] = 0.1
    mv = oracle.mean_var(c, phi, phiw, lam)
    assert mv[0, 4] > 0


def test_rel_error_hand_numbers(oracle):
    """momentsolver.cpp:687-749: (sum |e|^p A / sum |r|^p A)^(1/max(p,1)), Linf as
    max |e| A / max |r| A, optionally over a box of cells (mask)."""
    sol = np.array([[1.0, 3.0], [2.0, 1.0], [4.0, 0.0]])
    ref = np.array([[1.0, 1.0], [1.0, 1.0], [2.0, 1.0]])
    area = np.array([1.0, 2.0, 0.5])
    e2 = oracle.rel_error(sol, ref, area, 2)
    assert abs(e2[0] - np.sqrt((0 + 2 + 2) / (1 + 2 + 2))) < 1e-15
    assert abs(e2[1] - np.sqrt((4 + 0 + 0.5) / (1 + 2 + 0.5))) < 1e-15
    e1 = oracle.rel_error(sol, ref, area, 1)
    assert abs(e1[0] - (0 + 2 + 1) / (1 + 2 + 1)) < 1e-15
    assert abs(e1[1] - (2 + 0 + 0.5) / (1 + 2 + 0.5)) < 1e-15
    e0 = oracle.rel_error(sol, ref, area, 0)
    assert abs(e0[0] - 2.0 / 2.0) < 1e-15 and abs(e0[1] - 2.0 / 2.0) < 1e-15
    em = oracle.rel_error(sol, ref, area, 2, mask=np.array([1, 0, 1]))
    assert abs(em[0] - np.sqrt(2 / 3)) < 1e-15
