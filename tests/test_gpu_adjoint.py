"""Far-field objective and discrete adjoint gradient (aem_objective_*,
aem_adjoint_*) on the device.

* The transposed step passes the dot-product test <lam, A x> = <A^T lam, x>
  over the full state (bulk, CPML with material coefficients inside the
  slabs), for one and several steps.
* The device objective F = r E_pol(theta, phi, f) equals a numpy restatement
  written from the surface-equivalence formulas (M = E x n, J = n x H at the
  face-cell centres) on the oracle's fields, and its magnitude equals the
  spectrum of the reference's time-domain far field (the e_theta / e_phi
  series of the same run) at that frequency.
* d|F|^2/d eps_r from the adjoint equals a central finite difference of FP64
  oracle runs to 1e-3 relative.
"""
import math
import os

import numpy as np
import pytest

from alpineem_amd import decks
from alpineem_amd.solver import Plan, Simulation
from tests.oracle_ffi import Oracle

pytestmark = pytest.mark.gpu

C0 = 299792458.0
MU0 = float(np.float32(1.25663706e-6))
EP0 = float(np.float32(8.85418782e-12))
ETA = math.sqrt(MU0 / EP0)
NAMES = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]


# ---------------------------------------------------------------------------
# numpy restatement of the objective
# ---------------------------------------------------------------------------
class FarFieldRef:
    """F = r E_pol of the frequency-domain near-to-far-field transform over the
    far-field box of ``plan``: every field component is averaged to the
    centres of the face cells, M = E x n and J = n x H with n the outward
    normal, L = sum M dS e^{jk u.r'} e^{-jw n dt}, N the same with J at
    (n - 1/2) dt, r' = ((i - ic) dx, (j - jc) dy, (k - kc) dz)."""

    def __init__(self, plan, theta_deg, phi_deg, freq, pol="theta"):
        v = plan.view
        assert not any(v.pbc[a] for a in range(3))
        self.lo = [int(v.ff_lo[a]) for a in range(3)]
        self.hi = [int(v.ff_hi[a]) for a in range(3)]
        self.c = (int(v.ic), int(v.jc), int(v.kc))
        self.d = (float(v.dx), float(v.dy), float(v.dz))
        self.dt = float(v.dt)
        self.w = 2 * math.pi * freq
        self.k = self.w / C0
        th = float(np.float32(theta_deg)) * math.pi / 180
        ph = float(np.float32(phi_deg)) * math.pi / 180
        self.u = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
        self.th_v = np.array([math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph), -math.sin(th)])
        self.ph_v = np.array([-math.sin(ph), math.cos(ph), 0.0])
        self.pol = pol
        self.L = np.zeros(3, complex)
        self.N = np.zeros(3, complex)

    @staticmethod
    def _centre(F, r):
        """Six components averaged to the centres of cells r = ((i0,i1),(j0,j1),(k0,k1))."""
        (i0, i1), (j0, j1), (k0, k1) = r

        def s(a, di, dj, dk):
            return a[k0 - 1 + dk:k1 + dk, j0 - 1 + dj:j1 + dj, i0 - 1 + di:i1 + di]

        Ex, Ey, Ez, Hx, Hy, Hz = (F[n] for n in NAMES)
        E = np.stack([(s(Ex, 0, 0, 0) + s(Ex, 0, 1, 0) + s(Ex, 0, 0, 1) + s(Ex, 0, 1, 1)) / 4,
                      (s(Ey, 0, 0, 0) + s(Ey, 1, 0, 0) + s(Ey, 0, 0, 1) + s(Ey, 1, 0, 1)) / 4,
                      (s(Ez, 0, 0, 0) + s(Ez, 1, 0, 0) + s(Ez, 0, 1, 0) + s(Ez, 1, 1, 0)) / 4])
        H = np.stack([(s(Hx, 0, 0, 0) + s(Hx, 1, 0, 0)) / 2,
                      (s(Hy, 0, 0, 0) + s(Hy, 0, 1, 0)) / 2,
                      (s(Hz, 0, 0, 0) + s(Hz, 0, 0, 1)) / 2])
        return E, H

    def add(self, F, n):
        """Accumulate the fields after step n (dict name -> (NZ, NY, PX) array)."""
        eM = np.exp(-1j * self.w * n * self.dt)
        eJ = np.exp(-1j * self.w * (n - 0.5) * self.dt)
        for a in range(3):
            b, c = [x for x in range(3) if x != a]
            dS = self.d[b] * self.d[c]
            for low in (True, False):
                pos = self.lo[a] if low else self.hi[a]
                r = [(self.lo[x], self.hi[x]) for x in range(3)]
                r[a] = (pos, pos)
                E, H = self._centre(F, r)
                nrm = np.zeros(3)
                nrm[a] = -1.0 if low else 1.0
                M = np.stack([E[1] * nrm[2] - E[2] * nrm[1], E[2] * nrm[0] - E[0] * nrm[2],
                              E[0] * nrm[1] - E[1] * nrm[0]])
                J = np.stack([nrm[1] * H[2] - nrm[2] * H[1], nrm[2] * H[0] - nrm[0] * H[2],
                              nrm[0] * H[1] - nrm[1] * H[0]])
                kk, jj, ii = np.meshgrid(np.arange(r[2][0], r[2][1] + 1), np.arange(r[1][0], r[1][1] + 1),
                                         np.arange(r[0][0], r[0][1] + 1), indexing="ij")
                ph = self.k * (self.u[0] * (ii - self.c[0]) * self.d[0] + self.u[1] * (jj - self.c[1]) * self.d[1] +
                               self.u[2] * (kk - self.c[2]) * self.d[2])
                e = np.exp(1j * ph) * dS
                self.L += (M * e).reshape(3, -1).sum(axis=1) * eM
                self.N += (J * e).reshape(3, -1).sum(axis=1) * eJ

    def value(self) -> complex:
        if self.pol == "theta":
            return -1j * self.k / (4 * math.pi) * (self.L @ self.ph_v + ETA * (self.N @ self.th_v))
        return 1j * self.k / (4 * math.pi) * (self.L @ self.th_v - ETA * (self.N @ self.ph_v))


def oracle_objective(plan, fp64, *args, **kw) -> complex:
    """Run the oracle over the whole deck and accumulate the objective."""
    orc = Oracle(plan, fp64=fp64)
    ref = FarFieldRef(plan, *args, **kw)
    get = orc.field64 if fp64 else orc.field
    for n in range(1, plan.time_steps + 1):
        orc.step(1)
        ref.add({name: np.asarray(get(name), np.float64) for name in NAMES}, n)
    return ref.value()


# ---------------------------------------------------------------------------
# decks
# ---------------------------------------------------------------------------
def dot_deck():
    """No source; a lossy dielectric block reaching into the x and z CPML."""
    return decks.Deck(cells=(32, 32, 32), cell_size=(1e-3, 1e-3, 1e-3), time_steps=8, excitation="none",
                      materials=[decks.Material(2, (4.0, 3.0, 5.0), (0.5, 0.2, 0.8))],
                      objects=[decks.Block(2, (4, 10, 3), (12, 12, 16))])


DESIGN_LO, DESIGN_HI = (19, 19, 20), (20, 20, 20)
THETA, PHI, FREQ = 45.0, 90.0, 10e9


def scatter_deck(eps=(4.0, 4.5, 5.0, 3.5)):
    """Oblique TF/SF plane wave on a lossy block with four single-cell design
    voxels (materials 1..4) on top of it; one far-field angle."""
    mats = [decks.Material(11 + q, (e, e, e), (0.5, 0.5, 0.5)) for q, e in enumerate(eps)]
    mats.append(decks.Material(20, (3.0, 3.0, 3.0), (0.1, 0.1, 0.1)))
    objs = [decks.Block(20, (17, 17, 17), (6, 6, 3))]
    q = 0
    for j in (19, 20):
        for i in (19, 20):
            objs.append(decks.Block(11 + q, (i, j, 20), (1, 1, 1)))
            q += 1
    return decks.Deck(cells=(40, 40, 40), cell_size=(1e-3, 1e-3, 1e-3), time_steps=240, factor=1.0,
                      f_center=10e9, pulse_type=2, excitation="plane wave", amplitude=1.0,
                      angles=(60.0, 30.0, 45.0), phase_centre=(20, 20, 20), far_field=[(THETA, PHI)],
                      materials=mats, objects=objs)


def masks(plan):
    """Components the update changes: H on [1, N-1]^3; E tangential to the
    outer PEC faces stays fixed (Ex needs j, k >= 2, Ey i, k >= 2, Ez i, j >= 2)."""
    v = plan.view
    NZ, NY, PX = plan.shape
    k = np.arange(1, NZ + 1)[:, None, None]
    j = np.arange(1, NY + 1)[None, :, None]
    i = np.arange(1, PX + 1)[None, None, :]
    inb = (i <= v.NX - 1) & (j <= v.NY - 1) & (k <= v.NZ - 1)
    return {"Ex": inb & (j >= 2) & (k >= 2), "Ey": inb & (i >= 2) & (k >= 2), "Ez": inb & (i >= 2) & (j >= 2),
            "Hx": inb, "Hy": inb, "Hz": inb}


# ---------------------------------------------------------------------------
# tests
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("steps", [1, 3])
def test_transposed_step_dot_product(steps):
    plan = Plan.from_text(dot_deck().text())
    m = masks(plan)
    rng = np.random.default_rng(17 + steps)
    x = {n: np.where(m[n], rng.standard_normal(plan.shape), 0.0).astype(np.float32) for n in NAMES}
    lam = {n: np.where(m[n], rng.standard_normal(plan.shape), 0.0).astype(np.float32) for n in NAMES}
    # H and E differ in scale by ~eta in a physical state; random unit
    # states exercise every coefficient equally
    sim = Simulation(plan)
    for n in NAMES:
        sim.set_field(n, x[n])
    sim.run(steps, graph=False)
    y = {n: sim.field(n).astype(np.float64) for n in NAMES}
    for n in NAMES:
        sim.set_adjoint_field(n, lam[n])
    sim.adjoint_step(steps)
    mu = {n: sim.adjoint_field(n).astype(np.float64) for n in NAMES}
    terms_l = np.concatenate([(lam[n] * y[n]).ravel() for n in NAMES])
    terms_r = np.concatenate([(np.where(m[n], mu[n], 0.0) * x[n]).ravel() for n in NAMES])
    lhs, rhs = terms_l.sum(), terms_r.sum()
    scale = math.sqrt((terms_l ** 2).sum() + (terms_r ** 2).sum())
    print(f"steps={steps} <lam,Ax>={lhs:.9e} <A^T lam,x>={rhs:.9e} rel={(lhs - rhs) / scale:.3e}")
    assert abs(lhs - rhs) <= 1e-5 * scale
    sim.close()


@pytest.fixture(scope="module")
def scatter_run():
    plan = Plan.from_text(scatter_deck().text())
    sim = Simulation(plan)
    sim.set_objective(THETA, PHI, FREQ, "theta")
    sim.record_design(DESIGN_LO, DESIGN_HI)
    sim.run()
    return plan, sim


def test_objective_matches_restatement(scatter_run):
    plan, sim = scatter_run
    got = sim.objective()
    ref = oracle_objective(plan, False, THETA, PHI, FREQ, "theta")
    print(f"device F={got:.9e} numpy F={ref:.9e}")
    assert abs(got - ref) <= 1e-5 * abs(ref)


def x_dipole():
    """The dipole workload turned along x, on non-cubic cells."""
    d = decks.dipole(n=48, steps=900, angles=0, arm=10)
    c = 24
    d.cell_size = (1e-3, 1.1e-3, 0.9e-3)
    d.far_field = [(90.0, 90.0)]
    d.objects = [decks.Block(4, (c - 10, c, c), (10, 1, 1)), decks.Block(4, (c + 1, c, c), (10, 1, 1))]
    d.ports = [decks.Port("x", 50.0, (c, c, c), (1, 0, 0))]
    return d


@pytest.mark.parametrize("case", ["z_dipole_theta", "x_dipole_phi"])
def test_objective_matches_time_domain_far_field(case):
    """F against the spectrum of the reference's time-domain far field (the
    e_theta / e_phi series) of the same run, broadside to a dipole.  The
    series starts at the far-field time t0 (aem_get_ff_time) less one step,
    so its spectrum is F e^{-jw (t0 - dt)}; the two transforms differ by the
    time-derivative and retarded-time interpolation error, O((w dt)^2) ~ 2e-4
    at 3 GHz."""
    if case == "z_dipole_theta":
        d = decks.dipole(n=48, steps=900, angles=0, arm=10)
        d.far_field = [(90.0, 0.0)]
        pol, kind = "theta", "e_theta"
    else:
        d, pol, kind = x_dipole(), "phi", "e_phi"
    plan = Plan.from_text(d.text())
    f = 3e9
    sim = Simulation(plan)
    sim.set_objective(90.0, d.far_field[0][1], f, pol)
    sim.run()
    sim.finalize()
    F = sim.objective()
    spec = sim.dft(kind, [f], 0)[0]
    w = 2 * math.pi * f
    pred = F * np.exp(-1j * w * (float(sim.ff_time()[0]) - float(plan.view.dt)))
    print(f"{case}: F={F:.6e} spectrum={spec:.6e} rel={abs(spec - pred) / abs(pred):.3e}")
    assert abs(spec - pred) <= 1e-3 * abs(pred)


def test_eps_gradient_matches_finite_difference(scatter_run):
    plan, sim = scatter_run
    g_adj = sim.eps_gradient(4)
    eps0 = [4.0, 4.5, 5.0, 3.5]
    h = 0.05
    g_fd = np.zeros(4)
    for m in range(4):
        J = []
        e32 = []
        for sgn in (1, -1):
            eps = list(eps0)
            eps[m] += sgn * h
            e32.append(float(np.float32(eps[m])))
            p = Plan.from_text(scatter_deck(tuple(eps)).text())
            J.append(abs(oracle_objective(p, True, THETA, PHI, FREQ, "theta")) ** 2)
        g_fd[m] = (J[0] - J[1]) / (e32[0] - e32[1])
    print("adjoint", g_adj, "\nfinite difference", g_fd)
    assert np.linalg.norm(g_adj - g_fd) <= 1e-3 * np.linalg.norm(g_fd)
    assert np.all(np.abs(g_adj - g_fd) <= 2e-3 * np.abs(g_fd) + 1e-4 * np.abs(g_fd).max())


# ---------------------------------------------------------------------------
# per-cell design region
# ---------------------------------------------------------------------------
BOX_LO, BOX_HI = (18, 18, 19), (20, 20, 20)      # 3 x 3 x 2 design cells


# the 10 x 10 x 5 box of the sixteen-voxel finite-difference fixture
# (tests/golden/make_adjoint_fd.py): inside the TF/SF box (14..27), over the
# lossy block and the four single-cell voxels
BOX16_LO, BOX16_HI = (16, 16, 18), (25, 25, 22)


def design_eps(seed=5, lo=BOX_LO, hi=BOX_HI):
    """Per-cell permittivity of the design box, U[1, 4], shape (nz, ny, nx)."""
    rng = np.random.default_rng(seed)
    shape = tuple(hi[a] - lo[a] + 1 for a in (2, 1, 0))
    return rng.uniform(1.0, 4.0, shape).astype(np.float32)


def design_deck(eps_cells=None, lo=BOX_LO, hi=BOX_HI):
    """The plane-wave scatterer with a sigma = 0.5 block (eps 2) filling the
    design box; with ``eps_cells`` every box cell is its own material
    (100, 101, ...) of that permittivity -- the per-voxel geometry the
    reference reaches through one material per voxel."""
    d = scatter_deck()
    d.materials.append(decks.Material(30, (2.0, 2.0, 2.0), (0.5, 0.5, 0.5)))
    size = tuple(hi[a] - lo[a] + 1 for a in range(3))
    d.objects.append(decks.Block(30, lo, size))
    if eps_cells is not None:
        q = 0
        for k in range(size[2]):
            for j in range(size[1]):
                for i in range(size[0]):
                    e = float(eps_cells[k, j, i])
                    d.materials.append(decks.Material(100 + q, (e, e, e), (0.5, 0.5, 0.5)))
                    d.objects.append(decks.Block(100 + q, (lo[0] + i, lo[1] + j, lo[2] + k), (1, 1, 1)))
                    q += 1
    return d


def test_design_defaults_reproduce_the_deck():
    """A design box left at the deck's permittivity changes no bit."""
    plan = Plan.from_text(design_deck().text())
    a = Simulation(plan)
    b = Simulation(plan)
    b.design_box(BOX_LO, BOX_HI)
    for s in (a, b):
        s.run()
        s.finalize()
    for n in NAMES:
        assert np.array_equal(a.field(n), b.field(n)), n
    assert np.array_equal(a.series("e_theta", 0), b.series("e_theta", 0))


def test_design_eps_matches_per_voxel_materials():
    """Per-cell permittivity through the design box is bit-identical to the
    same permittivity declared as one material per voxel."""
    eps = design_eps()
    a = Simulation(Plan.from_text(design_deck(eps).text()))
    b = Simulation(Plan.from_text(design_deck().text()))
    b.design_box(BOX_LO, BOX_HI)
    b.set_design_eps(eps)
    for s in (a, b):
        s.run()
        s.finalize()
    for n in NAMES:
        assert np.array_equal(a.field(n), b.field(n)), n
    assert np.array_equal(a.series("e_theta", 0), b.series("e_theta", 0))


def test_design_transposed_step_dot_product():
    """Dot-product test of the transposed step with a design box (per-edge
    coefficients applied after the neutral bulk update)."""
    plan = Plan.from_text(dot_deck().text())
    m = masks(plan)
    rng = np.random.default_rng(29)
    x = {n: np.where(m[n], rng.standard_normal(plan.shape), 0.0).astype(np.float32) for n in NAMES}
    lam = {n: np.where(m[n], rng.standard_normal(plan.shape), 0.0).astype(np.float32) for n in NAMES}
    sim = Simulation(plan)
    sim.design_box((12, 13, 12), (16, 18, 15))
    sim.set_design_eps(rng.uniform(1.0, 6.0, sim.design_shape))
    for n in NAMES:
        sim.set_field(n, x[n])
    sim.run(3, graph=False)
    y = {n: sim.field(n).astype(np.float64) for n in NAMES}
    for n in NAMES:
        sim.set_adjoint_field(n, lam[n])
    sim.adjoint_step(3)
    mu = {n: sim.adjoint_field(n).astype(np.float64) for n in NAMES}
    terms_l = np.concatenate([(lam[n] * y[n]).ravel() for n in NAMES])
    terms_r = np.concatenate([(np.where(m[n], mu[n], 0.0) * x[n]).ravel() for n in NAMES])
    lhs, rhs = terms_l.sum(), terms_r.sum()
    scale = math.sqrt((terms_l ** 2).sum() + (terms_r ** 2).sum())
    print(f"design: <lam,Ax>={lhs:.9e} <A^T lam,x>={rhs:.9e} rel={(lhs - rhs) / scale:.3e}")
    assert abs(lhs - rhs) <= 1e-5 * scale


def test_design_gradient_matches_material_gradient(scatter_run):
    """On the four single-cell materials of the scatterer, the per-cell
    gradient of a design box over them equals the per-material gradient."""
    plan, sim = scatter_run
    g_mat = sim.eps_gradient(4)
    d = Simulation(plan)
    d.set_objective(THETA, PHI, FREQ, "theta")
    d.design_box(DESIGN_LO, DESIGN_HI)
    d.run()
    assert abs(d.objective() - sim.objective()) <= 1e-12 * abs(sim.objective())
    g = d.design_gradient()           # (1, 2, 2): cells (19..20, 19..20, 20)
    g_cells = g.reshape(-1)           # material q sits at cell q (x fastest)
    print("material", g_mat, "\ncells", g_cells)
    assert np.allclose(g_cells, g_mat, rtol=1e-5, atol=0)


def test_design_gradient_matches_finite_difference():
    eps = design_eps()
    plan = Plan.from_text(design_deck().text())
    sim = Simulation(plan)
    sim.set_objective(THETA, PHI, FREQ, "theta")
    sim.design_box(BOX_LO, BOX_HI)
    sim.set_design_eps(eps)
    sim.run()
    g_adj = sim.design_gradient()
    rng = np.random.default_rng(16)
    cells = [tuple(int(v) for v in np.unravel_index(q, eps.shape)) for q in rng.choice(eps.size, 4, replace=False)]
    h = 0.05
    for c in cells:
        J = []
        e32 = []
        for sgn in (1, -1):
            e = eps.copy()
            e[c] = np.float32(e[c] + sgn * h)
            e32.append(float(e[c]))
            J.append(abs(oracle_objective(Plan.from_text(design_deck(e).text()), True, THETA, PHI, FREQ)) ** 2)
        g_fd = (J[0] - J[1]) / (e32[0] - e32[1])
        print(f"cell {c}: adjoint {g_adj[c]:.9e} finite difference {g_fd:.9e}")
        assert abs(g_adj[c] - g_fd) <= 1e-3 * abs(g_fd)


def test_design_gradient_matches_sixteen_voxel_finite_differences(golden_dir):
    """d|F|^2/d eps at sixteen random voxels of a 10 x 10 x 5 design box (500
    cells, eps U[1, 4]) against central differences of FP64 oracle runs with
    one material per voxel (tests/golden/adjoint_fd16.npz, written by
    tests/golden/make_adjoint_fd.py; h = 2e-2, whose differences agree with
    h = 1e-2 to 4e-4) to 1e-3 relative; the device objective against the
    FP64 oracle's to 1e-5."""
    fx = np.load(os.path.join(golden_dir, "adjoint_fd16.npz"))
    lo, hi = tuple(int(x) for x in fx["box_lo"]), tuple(int(x) for x in fx["box_hi"])
    assert (lo, hi) == (BOX16_LO, BOX16_HI) and float(fx["h"][2]) == 2e-2
    eps = fx["eps"]
    np.testing.assert_array_equal(eps, design_eps(5, lo, hi))
    sim = Simulation(Plan.from_text(design_deck(None, lo, hi).text()))
    sim.set_objective(THETA, PHI, FREQ, "theta")
    sim.design_box(lo, hi)
    sim.set_design_eps(eps)
    sim.run()
    J = abs(sim.objective()) ** 2
    print(f"device J={J:.9e} FP64 oracle J={float(fx['J0']):.9e}")
    assert abs(J - float(fx["J0"])) <= 2e-5 * float(fx["J0"])
    g_adj = sim.design_gradient()
    g_fd = fx["g_fd"][2]
    worst = 0.0
    for q, c in enumerate(fx["cells"]):
        c = tuple(int(x) for x in c)
        rel = abs(g_adj[c] - g_fd[q]) / abs(g_fd[q])
        worst = max(worst, rel)
        print(f"cell {c}: adjoint {g_adj[c]:.9e} finite difference {g_fd[q]:.9e} rel {rel:.2e}")
    print(f"worst relative difference over 16 voxels: {worst:.2e}")
    assert worst <= 1e-3


def test_autograd_wrapper():
    """torch.autograd: J and dJ/d eps through far_field_power equal the direct
    API, and a small step along the gradient changes J as the gradient
    predicts to first order."""
    torch = pytest.importorskip("torch")
    from alpineem_amd.autograd import DesignProblem, far_field_power

    plan = Plan.from_text(design_deck().text())
    prob = DesignProblem(plan, BOX_LO, BOX_HI, THETA, PHI, FREQ)
    base = np.maximum(design_eps(), 1.1).astype(np.float32)
    eps = torch.tensor(base, dtype=torch.float64, requires_grad=True)
    J = far_field_power(eps, prob)
    J.backward()
    sim = prob.simulate(base)
    g = sim.design_gradient()
    assert J.item() == pytest.approx(abs(sim.objective()) ** 2, rel=1e-12)
    assert np.allclose(eps.grad.numpy(), g, rtol=1e-6, atol=0)
    step = (0.01 * g / np.abs(g).max()).astype(np.float32)
    with torch.no_grad():
        J2 = far_field_power(torch.tensor(base + step, dtype=torch.float64), prob).item()
    first = float((g * step.astype(np.float64)).sum())
    print(f"J={J.item():.9e} J(eps+s)={J2:.9e} predicted change {first:.3e} actual {J2 - J.item():.3e}")
    assert abs((J2 - J.item()) - first) <= 0.02 * abs(first)


# ---------------------------------------------------------------------------
# the reference's own gradient formula, on this solver's fields
# ---------------------------------------------------------------------------
def _edge_dft(plan, freq, lo, hi, with_objective=False):
    """FP32 oracle run of the whole deck: DFT at ``freq`` of Ex, Ey, Ez on
    the edges of the cell box [lo, hi] (array (3, nz+1, ny+1, nx+1)), and the
    objective when asked."""
    v = plan.view
    orc = Oracle(plan)
    w = 2 * math.pi * freq
    dt = float(v.dt)
    acc = 0.0
    ref = FarFieldRef(plan, THETA, PHI, freq, "theta") if with_objective else None
    for n in range(1, plan.time_steps + 1):
        orc.step(1)
        comps = [orc.field(c).astype(np.float64)[lo[2] - 1:hi[2] + 1, lo[1] - 1:hi[1] + 1, lo[0] - 1:hi[0] + 1]
                 for c in ("Ex", "Ey", "Ez")]
        acc = acc + np.stack(comps) * np.exp(-1j * w * n * dt)
        if ref is not None:
            ref.add({c: orc.field(c).astype(np.float64) for c in NAMES}, n)
    return acc, (ref.value() if ref is not None else None)


def test_design_gradient_matches_reference_workflow_formula():
    """The reference's adjoint workflow (optimization_scripts/
    adjoint_optimization/adjoint_optimization.py:76-134) computes
    dJ/d eps = Re(sum E_fwd E_adj) dV eps0 w^2 x (adjoint source scaling) from
    a second simulation.  Restated for J = |F|^2 with F = r E_theta: by
    reciprocity the adjoint simulation is a plane wave arriving from the
    observation direction with theta polarisation, and
        dF/d eps_v = k^2 dV / (4 pi E0) * sum over the 12 edges of voxel v of
                     (1/4) E_fwd . E_adj        (each edge averages 4 cells)
        dJ/d eps_v = 2 Re(conj(F) dF/d eps_v),
    E0 being the adjoint wave's amplitude at the phase reference (measured on
    an empty grid).  This continuous-adjoint formula approximates the exact
    discrete gradient to the grid's discretisation error, O((k dx)^2) = 4.4% at
    30 cells per wavelength; the test bounds the gap."""
    T = 800
    eps = design_eps()
    fwd = design_deck(eps)
    fwd.time_steps = T
    # exact discrete gradient on the device
    base = design_deck()
    base.time_steps = T
    plan = Plan.from_text(base.text())
    sim = Simulation(plan)
    sim.set_objective(THETA, PHI, FREQ, "theta")
    sim.design_box(BOX_LO, BOX_HI)
    sim.set_design_eps(eps)
    sim.run()
    g_adj = sim.design_gradient()
    # forward, adjoint and empty-grid fields from the oracle
    p_fwd = Plan.from_text(fwd.text())
    E1, F = _edge_dft(p_fwd, FREQ, BOX_LO, BOX_HI, with_objective=True)
    adj = design_deck(eps)
    adj.time_steps = T
    adj.angles = (180.0 - THETA, (PHI + 180.0) % 360.0, 270.0)   # arriving from (THETA, PHI), E along theta^
    E2, _ = _edge_dft(Plan.from_text(adj.text()), FREQ, BOX_LO, BOX_HI)
    adj.objects = []
    p_emp = Plan.from_text(adj.text())
    E2e, _ = _edge_dft(p_emp, FREQ, BOX_LO, BOX_HI)
    v = p_fwd.view
    k = 2 * math.pi * FREQ / C0
    th, ph = math.radians(THETA), math.radians(PHI)
    rh = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    thv = np.array([math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph), -math.sin(th)])
    d3 = np.array([float(v.dx), float(v.dy), float(v.dz)])
    p0 = (np.array([v.ic, v.jc, v.kc]) + 0.5) * d3          # phase reference of F (centre of the phase-centre cell)
    K, J, I = np.meshgrid(*(np.arange(BOX_LO[a], BOX_HI[a] + 2) for a in (2, 1, 0)), indexing="ij")
    est = []
    for c in range(3):
        if abs(thv[c]) < 1e-9:
            continue
        half = np.zeros(3)
        half[c] = 0.5                                          # Yee position of component c
        pos = np.stack([(I + half[0]) * d3[0], (J + half[1]) * d3[1], (K + half[2]) * d3[2]])
        phase = np.exp(1j * k * np.tensordot(rh, pos - p0[:, None, None, None], axes=1))
        est.append((E2e[c] / (thv[c] * phase)).ravel())
    est = np.concatenate(est)
    E0 = est.mean()
    assert np.abs(est - E0).max() < 1e-3 * abs(E0)            # the empty-grid wave is the assumed plane wave
    P = E1 * E2
    g_rec = np.zeros_like(g_adj)
    dV = float(d3.prod())
    for kk in range(g_adj.shape[0]):
        for jj in range(g_adj.shape[1]):
            for ii in range(g_adj.shape[2]):
                s = (P[0, kk:kk + 2, jj:jj + 2, ii].sum() + P[1, kk:kk + 2, jj, ii:ii + 2].sum() +
                     P[2, kk, jj:jj + 2, ii:ii + 2].sum())
                dF = k * k * dV / (4 * math.pi) * 0.25 * s / E0
                g_rec[kk, jj, ii] = 2 * np.real(np.conj(F) * dF)
    rel = np.abs(g_rec - g_adj) / np.abs(g_adj)
    cos = float((g_rec * g_adj).sum() / (np.linalg.norm(g_rec) * np.linalg.norm(g_adj)))
    print("discrete adjoint", g_adj.ravel(), "\nreference formula", g_rec.ravel(),
          f"\nmax rel gap {rel.max():.3f}, cosine {cos:.5f}")
    # measured: cosine 1.00000, a uniform 5.4% scale -- the O((k dx)^2) = 4.4% order of the gap
    assert cos > 0.999
    assert rel.max() < 0.08


# ---------------------------------------------------------------------------
# running-DFT far field at many directions and frequencies
# ---------------------------------------------------------------------------
def test_farfield_dft_matches_objective_and_restatement():
    """aem_ffdft_*: at the objective's (theta, phi, f) it equals the
    objective F; at every (direction, frequency) it equals the numpy
    restatement on the oracle's fields, for both polarisations."""
    plan = Plan.from_text(scatter_deck().text())
    angles = [(THETA, PHI), (60.0, 30.0), (150.0, 300.0)]
    f0, df, nf = 6e9, 2e9, 5                       # 6, 8, 10, 12, 14 GHz
    sim = Simulation(plan)
    sim.set_objective(THETA, PHI, FREQ, "theta")
    sim.set_farfield_dft(angles, f0, df, nf)
    sim.run()
    ff = sim.farfield_dft()
    F = sim.objective()
    assert abs(ff[0, 2, 0] - F) <= 1e-9 * abs(F)
    orc = Oracle(plan)
    refs = {(a, q, pol): FarFieldRef(plan, th, ph, f0 + q * df, pol)
            for a, (th, ph) in enumerate(angles) for q in range(nf) for pol in ("theta", "phi")}
    for n in range(1, plan.time_steps + 1):
        orc.step(1)
        fields = {c: orc.field(c).astype(np.float64) for c in NAMES}
        for r in refs.values():
            r.add(fields, n)
    worst = 0.0
    for (a, q, pol), r in refs.items():
        want = r.value()
        got = ff[a, q, 0 if pol == "theta" else 1]
        worst = max(worst, abs(got - want) / np.abs(ff[a, q]).max())
    print(f"running-DFT far field vs restatement: worst {worst:.2e}")
    assert worst <= 1e-5


def test_design_gradient_single_pass():
    """A deck that qualifies for the single-pass step (CPML, no TF/SF; here
    source-free with an initial field, as the adjoint takes no ports) runs
    its recorded forward as the single pass, the design box's update
    following it in k_e_design: fields and objective equal the split
    launches bit for bit, and so does the design gradient (the backward
    sweep is deterministic: test_gradient_is_deterministic)."""
    from tests.test_gpu_parity import single_pass_deck
    deck = single_pass_deck(steps=240)
    deck.ports = []
    deck.excitation = "none"
    text = deck.text()
    lo, hi = (25, 15, 22), (30, 19, 25)   # clear of the wire and the sphere
    shape = (hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
    eps = np.random.default_rng(7).uniform(1.0, 3.0, shape).astype(np.float32)
    out = []
    for sp in (True, False):
        sim = Simulation(Plan.from_text(text), ntff="off", single_pass=sp)
        ez = np.zeros(sim.plan.shape, np.float32)
        ez[17:21, 20:25, 33:38] = np.random.default_rng(3).standard_normal((4, 5, 5)).astype(np.float32)
        sim.set_field("Ez", ez)
        sim.set_objective(60.0, 0.0, 4e9, "theta")
        sim.design_box(lo, hi)
        sim.set_design_eps(eps)
        assert sim.single_pass == sp
        sim.run()
        fields = [sim.field(n) for n in NAMES]
        out.append((fields, sim.objective(), sim.design_gradient()))
        sim.close()
    (fa, Fa, ga), (fb, Fb, gb) = out
    for n, x, y in zip(NAMES, fa, fb):
        assert np.array_equal(x, y), n
    assert Fa == Fb and abs(Fa) > 0
    assert np.array_equal(ga, gb) and np.abs(ga).max() > 0


def test_gradient_is_deterministic():
    """The backward sweep sums in a fixed order (no atomics): two runs of the
    same design problem give bit-identical gradients, per voxel and per
    material."""
    eps = design_eps()
    g = []
    for _ in range(2):
        sim = Simulation(Plan.from_text(design_deck().text()), ntff="off")
        sim.set_objective(THETA, PHI, FREQ, "theta")
        sim.design_box(BOX_LO, BOX_HI)
        sim.set_design_eps(eps)
        sim.run()
        g.append(sim.design_gradient())
        sim.close()
    assert np.array_equal(g[0], g[1]) and np.abs(g[0]).max() > 0


def test_general_objective_gradient():
    """aem_adjoint with a caller's dJ/dF: for J = |F|^2 (dJ/dF = 2F) it is the
    design gradient bit for bit; for a goal mismatch J = |F - F_goal|^2 it
    matches FP64 central differences of that J on two cells."""
    eps = design_eps()
    plan = Plan.from_text(design_deck().text())
    sim = Simulation(plan)
    sim.set_objective(THETA, PHI, FREQ, "theta")
    sim.design_box(BOX_LO, BOX_HI)
    sim.set_design_eps(eps)
    sim.run()
    F = sim.objective()
    np.testing.assert_array_equal(sim.gradient(2 * F), sim.design_gradient())
    goal = 0.5 * F * np.exp(0.7j)
    g = sim.gradient(2 * (F - goal))
    rng = np.random.default_rng(21)
    cells = [tuple(int(v) for v in np.unravel_index(q, eps.shape)) for q in rng.choice(eps.size, 2, replace=False)]
    h = 0.05
    for c in cells:
        J, e32 = [], []
        for sgn in (1, -1):
            e = eps.copy()
            e[c] = np.float32(e[c] + sgn * h)
            e32.append(float(e[c]))
            Fo = oracle_objective(Plan.from_text(design_deck(e).text()), True, THETA, PHI, FREQ)
            J.append(abs(Fo - goal) ** 2)
        g_fd = (J[0] - J[1]) / (e32[0] - e32[1])
        print(f"cell {c}: adjoint {g[c]:.9e} finite difference {g_fd:.9e}")
        assert abs(g[c] - g_fd) <= 1e-3 * abs(g_fd)


# ---------------------------------------------------------------------------
# Drude media in the transposed step (fdtd_solver.f90:5782-5854, 6812-6826)
# ---------------------------------------------------------------------------
DRUDE_POLES = [((3e10, 3e10, 3e10), (1e10, 1e10, 1e10)), ((2e10, 2.2e10, 1.8e10), (5e9, 6e9, 4e9))]


def drude_dot_deck():
    """The dot-product deck with a two-pole Drude block clear of the CPML."""
    d = dot_deck()
    d.materials.append(decks.Material(7, (1.5, 2.0, 1.2), (0.01, 0.0, 0.02), DRUDE_POLES))
    d.objects.append(decks.Block(7, (15, 14, 15), (3, 4, 3)))
    return d


def drude_scatter_deck(eps=(4.0, 4.5, 5.0, 3.5)):
    """The plane-wave scatterer with a Drude slab above the design voxels
    (cells 17..22 x 17..22 x 23..24, clear of the record box's edges)."""
    d = scatter_deck(eps)
    d.materials.append(decks.Material(21, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), DRUDE_POLES))
    d.objects.append(decks.Block(21, (17, 17, 23), (6, 6, 2)))
    return d


@pytest.mark.parametrize("steps", [1, 4])
def test_drude_transposed_step_dot_product(steps):
    """<lam, A x> = <A^T lam, x> with the Drude current and E_old as state:
    over several steps the current's recursion is transposed too."""
    plan = Plan.from_text(drude_dot_deck().text())
    assert plan.view.has_drude
    m = masks(plan)
    rng = np.random.default_rng(41 + steps)
    x = {n: np.where(m[n], rng.standard_normal(plan.shape), 0.0).astype(np.float32) for n in NAMES}
    lam = {n: np.where(m[n], rng.standard_normal(plan.shape), 0.0).astype(np.float32) for n in NAMES}
    sim = Simulation(plan)
    for n in NAMES:
        sim.set_field(n, x[n])
    sim.run(steps, graph=False)
    y = {n: sim.field(n).astype(np.float64) for n in NAMES}
    for n in NAMES:
        sim.set_adjoint_field(n, lam[n])
    sim.adjoint_step(steps)
    mu = {n: sim.adjoint_field(n).astype(np.float64) for n in NAMES}
    terms_l = np.concatenate([(lam[n] * y[n]).ravel() for n in NAMES])
    terms_r = np.concatenate([(np.where(m[n], mu[n], 0.0) * x[n]).ravel() for n in NAMES])
    lhs, rhs = terms_l.sum(), terms_r.sum()
    scale = math.sqrt((terms_l ** 2).sum() + (terms_r ** 2).sum())
    print(f"drude steps={steps} <lam,Ax>={lhs:.9e} <A^T lam,x>={rhs:.9e} rel={(lhs - rhs) / scale:.3e}")
    assert abs(lhs - rhs) <= 1e-5 * scale
    sim.close()


def test_drude_eps_gradient_matches_finite_difference():
    """d|F|^2/d eps_r of the four design materials with a Drude slab in the
    scattered field: the adjoint equals FP64 central differences of the
    oracle to 1e-3, and the slab changes F (the transposed Drude step is on
    the gradient's path)."""
    plan = Plan.from_text(drude_scatter_deck().text())
    sim = Simulation(plan)
    sim.set_objective(THETA, PHI, FREQ, "theta")
    sim.record_design(DESIGN_LO, DESIGN_HI)
    sim.run()
    F = sim.objective()
    g_adj = sim.eps_gradient(4)
    sim.close()
    plain = Simulation(Plan.from_text(scatter_deck().text()))
    plain.set_objective(THETA, PHI, FREQ, "theta")
    plain.run()
    F0 = plain.objective()
    plain.close()
    print(f"F with the slab {F:.6e}, without {F0:.6e}")
    assert abs(F - F0) > 0.05 * abs(F0)
    eps0 = [4.0, 4.5, 5.0, 3.5]
    h = 0.05
    g_fd = np.zeros(4)
    for q in range(4):
        J, e32 = [], []
        for sgn in (1, -1):
            eps = list(eps0)
            eps[q] += sgn * h
            e32.append(float(np.float32(eps[q])))
            p = Plan.from_text(drude_scatter_deck(tuple(eps)).text())
            J.append(abs(oracle_objective(p, True, THETA, PHI, FREQ, "theta")) ** 2)
        g_fd[q] = (J[0] - J[1]) / (e32[0] - e32[1])
    print("adjoint", g_adj, "\nfinite difference", g_fd)
    assert np.linalg.norm(g_adj - g_fd) <= 1e-3 * np.linalg.norm(g_fd)
    assert np.all(np.abs(g_adj - g_fd) <= 2e-3 * np.abs(g_fd) + 1e-4 * np.abs(g_fd).max())


def test_drude_design_box_refusals():
    """A design box meeting the Drude media's bounding box, and Drude media
    at the CPML, are refused with AEM_ERR_UNSUPPORTED."""
    from alpineem_amd._lib import AemError
    sim = Simulation(Plan.from_text(drude_scatter_deck().text()))
    with pytest.raises(AemError, match="Drude"):
        sim.design_box((18, 18, 21), (20, 20, 22))
    sim.close()
    d = drude_dot_deck()
    d.objects[-1] = decks.Block(7, (4, 14, 15), (3, 4, 3))
    sim = Simulation(Plan.from_text(d.text()))
    with pytest.raises(AemError, match="Drude"):
        sim.design_box((15, 15, 15), (16, 16, 16))
    sim.close()


# ---------------------------------------------------------------------------
# lumped ports in the transposed step, and the port-series objective
# (fdtd_solver.f90:4800-5013, 6120-6223; adjoint_optimization.py:76-80)
# ---------------------------------------------------------------------------
def port_dot_deck():
    """The dot-product deck with two lumped ports (z and x gaps) clear of
    the CPML; the lines start at rest."""
    d = dot_deck()
    d.ports = [decks.Port("z", 50.0, (16, 16, 16), (0, 0, 1)), decks.Port("x", 75.0, (14, 18, 17), (2, 0, 0))]
    return d


@pytest.mark.parametrize("steps", [1, 5])
def test_port_transposed_step_dot_product(steps):
    """<lam, A x> = <A^T lam, x> with the ports' lines, loop currents and
    gap fields in the step: from one step on, the gap field couples the
    field state to the line and the line back to H."""
    plan = Plan.from_text(port_dot_deck().text())
    assert plan.view.n_ports == 2
    m = masks(plan)
    rng = np.random.default_rng(53 + steps)
    x = {n: np.where(m[n], rng.standard_normal(plan.shape), 0.0).astype(np.float32) for n in NAMES}
    lam = {n: np.where(m[n], rng.standard_normal(plan.shape), 0.0).astype(np.float32) for n in NAMES}
    sim = Simulation(plan)
    for n in NAMES:
        sim.set_field(n, x[n])
    sim.run(steps, graph=False)
    y = {n: sim.field(n).astype(np.float64) for n in NAMES}
    for n in NAMES:
        sim.set_adjoint_field(n, lam[n])
    sim.adjoint_step(steps)
    mu = {n: sim.adjoint_field(n).astype(np.float64) for n in NAMES}
    terms_l = np.concatenate([(lam[n] * y[n]).ravel() for n in NAMES])
    terms_r = np.concatenate([(np.where(m[n], mu[n], 0.0) * x[n]).ravel() for n in NAMES])
    lhs, rhs = terms_l.sum(), terms_r.sum()
    scale = math.sqrt((terms_l ** 2).sum() + (terms_r ** 2).sum())
    print(f"ports steps={steps} <lam,Ax>={lhs:.9e} <A^T lam,x>={rhs:.9e} rel={(lhs - rhs) / scale:.3e}")
    assert abs(lhs - rhs) <= 1e-5 * scale
    sim.close()


PORT_LO, PORT_HI = (26, 23, 22), (28, 25, 26)     # 3 x 3 x 5 design cells beside the dipole's gap
PORT_F0 = 3e9
S11_GOAL = 0.6 * np.exp(0.4j)


def port_deck(eps_cells=None):
    """The 48^3 dipole (50 ohm port, arms of 10 cells) with a lossy
    dielectric block (eps 2, sigma 0.05) filling the design box; with
    ``eps_cells`` every box cell is its own material of that permittivity."""
    d = decks.dipole(n=48, steps=900, angles=0, arm=10)
    size = tuple(PORT_HI[a] - PORT_LO[a] + 1 for a in range(3))
    d.materials.append(decks.Material(30, (2.0, 2.0, 2.0), (0.05, 0.05, 0.05)))
    d.objects.append(decks.Block(30, PORT_LO, size))
    if eps_cells is not None:
        q = 0
        for k in range(size[2]):
            for j in range(size[1]):
                for i in range(size[0]):
                    e = float(eps_cells[k, j, i])
                    d.materials.append(decks.Material(100 + q, (e, e, e), (0.05, 0.05, 0.05)))
                    d.objects.append(decks.Block(100 + q, (PORT_LO[0] + i, PORT_LO[1] + j, PORT_LO[2] + k),
                                                 (1, 1, 1)))
                    q += 1
    return d


def port_eps(seed=8):
    shape = tuple(PORT_HI[a] - PORT_LO[a] + 1 for a in (2, 1, 0))
    return np.random.default_rng(seed).uniform(2.0, 8.0, shape).astype(np.float32)


def spectrum(x, f, dt):
    n = np.arange(len(x), dtype=np.float64)
    p = f * dt * n
    return complex(np.sum(x * np.exp(-2j * np.pi * (p - np.floor(p)))))


def oracle_port_objectives(plan):
    """FP64 oracle run: J_s11 = |goal - S11|^2 at f0 (S11 = V/Vinc - 1) and
    J_i = |I(f0)|^2 of the port current, with V, I, Vinc in double."""
    orc = Oracle(plan, fp64=True)
    orc.step(plan.time_steps)
    orc.finalize()
    dt = float(plan.view.dt)
    V = spectrum(orc.series64("port_v", 0), PORT_F0, dt)
    I = spectrum(orc.series64("port_i", 0), PORT_F0, dt)
    vinc = spectrum(orc.series64("incident", 0), PORT_F0, dt)
    s11 = V / vinc - 1
    return abs(S11_GOAL - s11) ** 2, abs(I) ** 2


def test_port_s11_gradient_matches_finite_difference():
    """The reference workflow's objective J = |S11_goal - S11|^2 at 3 GHz
    (adjoint_optimization.py:76-80), and J = |I(f0)|^2 of the port current,
    differentiated per design cell through the transposed port lines: the
    adjoint equals FP64 central differences of the oracle to 1e-3 on four
    cells."""
    eps = port_eps()
    plan = Plan.from_text(port_deck().text())
    sim = Simulation(plan, ntff="off")
    sim.design_box(PORT_LO, PORT_HI)
    sim.set_design_eps(eps)
    sim.run()
    sim.finalize()
    V = complex(sim.dft("port_v", [PORT_F0], 0)[0])
    I = complex(sim.dft("port_i", [PORT_F0], 0)[0])
    vinc = complex(sim.dft("incident", [PORT_F0], 0)[0])
    s11 = V / vinc - 1
    print(f"S11(3 GHz) = {s11:.6f}")
    g_s11 = sim.port_gradient(0, PORT_F0, 2 * (s11 - S11_GOAL) / np.conj(vinc))
    # J = |I|^2 through the current series: dJ/dI(step n+1) = Re(conj(2 I) e^{-i w n dt})
    T = plan.time_steps
    ph = PORT_F0 * float(plan.view.dt) * np.arange(T, dtype=np.float64)
    wI = np.real(np.conj(2 * I) * np.exp(-2j * np.pi * (ph - np.floor(ph))))[None, :]
    g_i = sim.series_gradient(dJ_dI=wI)
    sim.close()
    rng = np.random.default_rng(31)
    cells = [tuple(int(v) for v in np.unravel_index(q, eps.shape)) for q in rng.choice(eps.size, 4, replace=False)]
    h = 0.05
    for c in cells:
        Js, Ji, e32 = [], [], []
        for sgn in (1, -1):
            e = eps.copy()
            e[c] = np.float32(e[c] + sgn * h)
            e32.append(float(e[c]))
            js, ji = oracle_port_objectives(Plan.from_text(port_deck(e).text()))
            Js.append(js)
            Ji.append(ji)
        fd_s = (Js[0] - Js[1]) / (e32[0] - e32[1])
        fd_i = (Ji[0] - Ji[1]) / (e32[0] - e32[1])
        print(f"cell {c}: S11 adjoint {g_s11[c]:.9e} fd {fd_s:.9e} | I adjoint {g_i[c]:.9e} fd {fd_i:.9e}")
        assert abs(g_s11[c] - fd_s) <= 1e-3 * abs(fd_s)
        assert abs(g_i[c] - fd_i) <= 1e-3 * abs(fd_i)


def test_port_s11_autograd_wrapper():
    """torch.autograd: (goal - port_s11(eps)).abs()**2 and its gradient equal
    the direct API's."""
    torch = pytest.importorskip("torch")
    from alpineem_amd.autograd import PortProblem, port_s11

    eps0 = port_eps()
    plan = Plan.from_text(port_deck().text())
    prob = PortProblem(plan, PORT_LO, PORT_HI, PORT_F0, port=0)
    eps = torch.tensor(eps0, dtype=torch.float64, requires_grad=True)
    s11 = port_s11(eps, prob)
    J = (torch.tensor(S11_GOAL, dtype=torch.complex128) - s11).abs() ** 2
    J.backward()
    sim = prob.simulate(eps0)
    V = complex(sim.dft("port_v", [PORT_F0], 0)[0])
    vinc = complex(sim.dft("incident", [PORT_F0], 0)[0])
    s = V / vinc - 1
    g = sim.port_gradient(0, PORT_F0, 2 * (s - S11_GOAL) / np.conj(vinc))
    sim.close()
    assert complex(s11.item()) == pytest.approx(s, rel=1e-12)
    assert J.item() == pytest.approx(abs(S11_GOAL - s) ** 2, rel=1e-12)
    assert np.allclose(eps.grad.numpy(), g, rtol=1e-9, atol=0)


def test_port_adjoint_refusals():
    """A design box meeting a port's gap or loop, and a port at the CPML,
    are refused with AEM_ERR_UNSUPPORTED; SPICE decks stay refused."""
    from alpineem_amd._lib import AemError
    sim = Simulation(Plan.from_text(port_deck().text()))
    with pytest.raises(AemError, match="port"):
        sim.design_box((24, 23, 22), (26, 25, 26))
    sim.close()
    d = port_dot_deck()
    d.ports = [decks.Port("z", 50.0, (16, 16, 5), (0, 0, 1))]
    sim = Simulation(Plan.from_text(d.text()))
    with pytest.raises(AemError, match="port"):
        sim.design_box((16, 16, 16), (17, 17, 17))
    sim.close()


# ---------------------------------------------------------------------------
# checkpoints: a record of `every` steps and recomputation between snapshots
# ---------------------------------------------------------------------------
def _grad_run(deck_text, box, every, objective, eps=None):
    sim = Simulation(Plan.from_text(deck_text), ntff="off")
    if every:
        sim.checkpoint(every)
    if objective == "farfield":
        sim.set_objective(THETA, PHI, FREQ, "theta")
    sim.design_box(*box)
    if eps is not None:
        sim.set_design_eps(eps)
    sim.run()
    fields = [sim.field(n) for n in NAMES]
    if objective == "farfield":
        g = sim.design_gradient()
    else:
        sim.finalize()
        V = complex(sim.dft("port_v", [PORT_F0], 0)[0])
        vinc = complex(sim.dft("incident", [PORT_F0], 0)[0])
        g = sim.port_gradient(0, PORT_F0, 2 * (V / vinc - 1 - S11_GOAL) / np.conj(vinc))
    after = [sim.field(n) for n in NAMES]
    g2 = sim.design_gradient() if objective == "farfield" else None
    sim.close()
    return fields, after, g, g2


@pytest.mark.parametrize("case", ["farfield", "port", "drude"])
def test_checkpointed_gradient_equals_full_record(case):
    """With checkpoints every K steps (K not dividing the step count) the
    gradient equals the full record's bit for bit, the fields after the
    gradient call are the run's last fields again, and a second gradient
    call gives the same result."""
    if case == "farfield":
        text, box, K, obj, eps = design_deck().text(), (BOX_LO, BOX_HI), 70, "farfield", design_eps()
    elif case == "port":
        text, box, K, obj, eps = port_deck().text(), (PORT_LO, PORT_HI), 130, "port", port_eps()
    else:
        d = drude_scatter_deck()
        text, box, K, obj, eps = d.text(), (DESIGN_LO, DESIGN_HI), 45, "farfield", None
    f_full, a_full, g_full, _ = _grad_run(text, box, 0, obj, eps)
    f_ck, a_ck, g_ck, g_ck2 = _grad_run(text, box, K, obj, eps)
    for n, x, y in zip(NAMES, f_full, f_ck):
        assert np.array_equal(x, y), n
    for n, x, y in zip(NAMES, f_ck, a_ck):
        assert np.array_equal(x, y), n
    print(f"{case}: |g| {np.linalg.norm(g_full):.6e}, checkpointed {np.linalg.norm(g_ck):.6e}")
    assert np.abs(g_full).max() > 0
    np.testing.assert_array_equal(g_ck, g_full)
    if g_ck2 is not None:
        np.testing.assert_array_equal(g_ck2, g_full)
