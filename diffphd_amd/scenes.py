"""Synthetic scenes of the PD benchmark configurations.

Every builder returns a plain dict `spec` (mesh arrays, energies, boundary
conditions, initial state, options); `apply(spec, body)` replays it on any
object with the reference Deformable's method names (the product's
HexDeformable/TetDeformable or a test oracle). Seeds are fixed per scene.
"""
import numpy as np

from .hex_mesh import hex_grid, tet_grid


def lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def default_options(**kw):
    opt = {"max_pd_iter": 1000, "abs_tol": 1e-12, "rel_tol": 1e-9, "verbose": 0, "thread_ct": 8}
    opt.update(kw)
    return opt


def apply(spec, body):
    body.InitializeFromArrays(spec["V"], spec["E"], spec["density"])
    if spec.get("gravity") is not None:
        body.AddStateForce("gravity", list(spec["gravity"]))
    for kind, params in spec.get("state_forces", []):
        body.AddStateForce(kind, list(params))
    for kind, params, idx in spec["energies"]:
        body.AddPdEnergy(kind, list(params), list(idx))
    for stiffness, fiber, idx in spec.get("muscles", []):
        body.AddActuation(float(stiffness), list(fiber), list(idx))
    for dof, val in spec.get("dirichlet", []):
        body.SetDirichletBoundaryCondition(int(dof), float(val))
    b = spec.get("boundary")
    if b is not None:
        body.SetFrictionalBoundary(b[0], list(b[1]), list(b[2]))
    return body


def cantilever(cells=(16, 4, 4), dx=0.01, E=1e5, nu=0.45, rho=1000.0, per_element_energies=False, seed=0):
    """Cantilever beam (16x4x4 hex, x=0 face fixed, gravity): plain FP64 PD, no
    contact. Loss = dot(q_T, g), g ~ N(0, 1) with the given seed.

    per_element_energies: one 'corotated' energy per element (the reference's
    layout for per-element mu gradients) instead of one per-element-parameter energy."""
    V, Eidx, _, _ = hex_grid(np.ones(cells), dx, (0.0, 0.0, 0.0))
    ne = Eidx.shape[1]
    mu, la = lame(E, nu)
    if per_element_energies:
        energies = [("corotated", [2 * mu], [e]) for e in range(ne)] + [("volume", [la], [])]
    else:
        energies = [("corotated", [2 * mu], []), ("volume", [la], [])]
    dirichlet = []
    for i in range(V.shape[1]):
        if V[0, i] < 1e-12:
            dirichlet += [(3 * i + k, V[k, i]) for k in range(3)]
    q0 = V.T.ravel().copy()
    rng = np.random.default_rng(seed)
    return {
        "name": "cantilever", "edim": 8, "V": V, "E": Eidx, "density": rho, "gravity": (0.0, 0.0, -9.81),
        "energies": energies, "dirichlet": dirichlet, "boundary": None, "q0": q0, "v0": np.zeros_like(q0),
        "dt": 1e-2, "frames": 100, "loss_g": rng.standard_normal(q0.size), "mu": mu, "la": la,
        "options": default_options(max_pd_iter=5000),
    }


def actuated_cantilever(cells=(8, 2, 2), dx=0.01, E=1e5, nu=0.45, rho=1000.0, stiffness=(2e5, 1e5), seed=0):
    """Cantilever with muscle fibers (CantileverEnv3d,
    python/py_diff_pd/env/cantilever_env_3d.py:86-96): the bottom row of
    elements (j = 0, k = 0) actuated along x, plus a second group, the top row
    (j = ny - 1, k = nz - 1), along the diagonal (1, 0, 1) to exercise a fiber
    off the grid axes. `act` holds one seeded level in [0.7, 1.0] per muscle
    element (1 = the rest length, below 1 contracts)."""
    spec = cantilever(cells, dx, E, nu, rho, seed=seed)
    nx, ny, nz = cells
    bottom = [i * ny * nz for i in range(nx)]
    top = [i * ny * nz + (ny - 1) * nz + (nz - 1) for i in range(nx)]
    spec["name"] = "actuated_cantilever"
    spec["muscles"] = [(stiffness[0], [1.0, 0.0, 0.0], bottom), (stiffness[1], [1.0, 0.0, 1.0], top)]
    spec["act"] = np.random.default_rng(seed + 100).uniform(0.7, 1.0, len(bottom) + len(top))
    return spec


def penalty_drop(cells=(4, 2, 2), dx=0.01, E=1e5, nu=0.45, rho=1000.0, kn=5.0, kf=0.005, mu=0.4, seed=0):
    """A free block on a penalty floor (PlanarContactStateForce, the state force
    of AddStateForce("planar_contact", [n, offset, p, kn, kf, mu]),
    planar_contact_state_force.cpp): the bottom starts 0.2 dx below z = 0 with a
    downward and a sideways velocity. The force is explicit (evaluated at the
    frame start), so kn and kf stay below m / h^2 and m / h (1e-3 kg vertices,
    dt = 1e-2); with the defaults the friction is clamped at kf (viscous), with
    kf large it is Coulomb, mu f_n."""
    V, Eidx, _, _ = hex_grid(np.ones(cells), dx, (0.0, 0.0, -0.2 * dx))
    mu_l, la = lame(E, nu)
    q0 = V.T.ravel().copy()
    rng = np.random.default_rng(seed)
    v0 = np.tile([0.3, 0.1, -0.2], V.shape[1]) + 0.02 * rng.standard_normal(q0.size)
    return {
        "name": "penalty_drop", "edim": 8, "V": V, "E": Eidx, "density": rho, "gravity": (0.0, 0.0, -9.81),
        "state_forces": [("planar_contact", [0.0, 0.0, 1.0, 0.0, 2, kn, kf, mu])],
        "energies": [("corotated", [2 * mu_l], []), ("volume", [la], [])], "dirichlet": [], "boundary": None,
        "q0": q0, "v0": v0, "dt": 1e-2, "frames": 10, "loss_g": rng.standard_normal(q0.size),
        "options": default_options(max_pd_iter=5000),
    }


def arc_drop(kn=5.0, kf=0.005, mu=0.4, seed=0):
    """The penalty_drop block in a cylindrical trough (ArcContactStateForce,
    AddStateForce("arc_contact", [center, dir, start, radius, angle, p, kn, kf,
    mu]), arc_contact_state_force.cpp): axis along x through (0, 0.01, 0.05),
    the arc spanning 90 degrees centred on straight down (start (0, -1, -1),
    toward aux = dir x start = (0, 1, -1) / sqrt 2), radius 0.0515, so the bottom
    vertices (0.052-0.053 from the axis) start just outside it."""
    spec = penalty_drop(kn=kn, kf=kf, mu=mu, seed=seed)
    s = 1.0 / np.sqrt(2.0)
    spec["name"] = "arc_drop"
    spec["state_forces"] = [("arc_contact", [0.0, 0.01, 0.05, 1.0, 0.0, 0.0, 0.0, -s, -s, 0.0515, np.pi / 2, 2,
                                             kn, kf, mu])]
    return spec


def mesh_stack(kn=5.0, kf=0.005, mu=0.4, radius=0.003, dx=0.01, E=1e5, nu=0.45, rho=1000.0, seed=0):
    """Two 3x3x2-cell blocks in one body, the upper one 0.2 dx above the lower,
    meeting through MeshContactStateForce (AddStateForce("mesh_contact",
    [radius, kn_left, kn_right, kf, mu, #faces A, #faces B, faces A, faces B]),
    mesh_contact_state_force.cpp): A = the upper block's bottom face, B = the
    lower block's top face (two triangles per quad). The lower block's bottom
    layer is fixed; the upper block moves sideways and down."""
    vox = np.ones((3, 3, 5), dtype=bool)
    vox[:, :, 2] = False
    V, Eidx, _, vf = hex_grid(vox, dx, (0.0, 0.0, 0.0))
    upper = np.unique(vf[:, :, 3:][vf[:, :, 3:] >= 0])
    V[2, upper] -= 0.8 * dx

    def face(k, flip):
        tris = []
        for i in range(3):
            for j in range(3):
                a, b, c, d = vf[i, j, k], vf[i + 1, j, k], vf[i + 1, j + 1, k], vf[i, j + 1, k]
                tris += [(a, c, b), (a, d, c)] if flip else [(a, b, c), (a, c, d)]
        return tris

    fa, fb = face(3, True), face(2, False)
    prm = [radius, kn, kn, kf, mu, len(fa), len(fb)] + [float(x) for t in fa + fb for x in t]
    mu_l, la = lame(E, nu)
    q0 = V.T.ravel().copy()
    rng = np.random.default_rng(seed)
    v0 = np.zeros_like(q0)
    v0.reshape(-1, 3)[upper] = [0.2, 0.1, -0.1]
    v0 += 0.01 * rng.standard_normal(q0.size) * np.repeat(np.isin(np.arange(V.shape[1]), upper), 3)
    bottom = vf[:, :, 0].ravel()
    return {
        "name": "mesh_stack", "edim": 8, "V": V, "E": Eidx, "density": rho, "gravity": (0.0, 0.0, -9.81),
        "state_forces": [("mesh_contact", prm)], "faces": (fa, fb), "upper": upper,
        "energies": [("corotated", [2 * mu_l], []), ("volume", [la], [])],
        "dirichlet": [(3 * int(i) + d, float(q0[3 * int(i) + d])) for i in bottom for d in range(3)], "boundary": None,
        "q0": q0, "v0": v0, "dt": 1e-2, "frames": 10, "loss_g": rng.standard_normal(q0.size),
        "options": default_options(max_pd_iter=5000),
    }


def billiard(balls=8, dx=0.01, radius=1.05, E=1e5, nu=0.45, rho=1000.0, seed=0, k=None, mu=None):
    """A row of one-cell cubes 2 dx apart along x as balls of
    BilliardBallStateForce (AddStateForce("billiard_ball", [radius, vertices per
    ball, stiffness per ball, friction per ball]), billiard_ball_state_force.cpp):
    each cube's 8 vertices are one ball (8 balls, so the reference's backward is
    defined), ball radius `radius` dx, so neighbours overlap by 2 (radius - 1) dx;
    random sideways velocities set the friction directions; no gravity."""
    vox = np.zeros((2 * balls - 1, 1, 1), dtype=bool)
    vox[::2] = True
    V, Eidx, _, _ = hex_grid(vox, dx, (0.0, 0.0, 0.0))
    rng = np.random.default_rng(seed)
    k = list(rng.uniform(3.0, 6.0, balls)) if k is None else list(k)
    mu = list(rng.uniform(0.2, 0.5, balls)) if mu is None else list(mu)
    mu_l, la = lame(E, nu)
    q0 = V.T.ravel().copy()
    vb = rng.standard_normal((balls, 3)) * [0.05, 0.2, 0.05]
    v0 = np.repeat(vb, 8, axis=0).ravel() + 0.005 * rng.standard_normal(q0.size)
    return {
        "name": "billiard", "edim": 8, "V": V, "E": Eidx, "density": rho, "gravity": None,
        "state_forces": [("billiard_ball", [radius * dx, 8] + k + mu)],
        "energies": [("corotated", [2 * mu_l], []), ("volume", [la], [])], "dirichlet": [], "boundary": None,
        "q0": q0, "v0": v0, "dt": 1e-2, "frames": 10, "loss_g": rng.standard_normal(q0.size),
        "options": default_options(max_pd_iter=5000),
    }


HYDRO_CD = [0.0, 0.05, 0.4, 0.05, 0.7, 1.85, 1.0, 2.05]
HYDRO_CT = [-1.0, -0.8, -0.3, -0.5, 0.3, 0.1, 1.0, 2.5]


def swim(kind="hex", v_water=(-0.6, 0.1, 0.05), max_force=1e3, faces=None, dx=0.01, E=1e5, nu=0.45, rho=1000.0,
         rho_water=1000.0, seed=0):
    """A free block in moving water (HydrodynamicsStateForce,
    AddStateForce("hydrodynamics", [rho, v_water, Cd points, Ct points,
    max_force, faces]), hydrodynamics_state_force.cpp): a 4x2x2-cell hex block
    with its boundary quads, or a 2x2x2-cube tet block (5 tets per cube) with its
    boundary triangles, every face ordered so that its normal points outward.
    `faces` restricts the surface to the given face indices. The tet block's
    vertices weigh a fifth of the hex block's (the reference's per-vertex mass is
    density times the mean element volume), so a rollout of it wants a lighter
    fluid (rho_water)."""
    if kind == "hex":
        V, Eidx, _, vf = hex_grid(np.ones((4, 2, 2)), dx, (0.0, 0.0, 0.0))
        n = np.array(vf.shape) - 1
        quads = []
        for ax in range(3):
            for side in (0, n[ax]):
                o1, o2 = [a for a in range(3) if a != ax]
                for u in range(n[o1]):
                    for w in range(n[o2]):
                        idx = []
                        for du, dw in ((0, 0), (1, 0), (1, 1), (0, 1)):
                            c = [0, 0, 0]
                            c[ax], c[o1], c[o2] = side, u + du, w + dw
                            idx.append(int(vf[tuple(c)]))
                        quads.append(idx)
        surf = quads
        edim = 8
    else:
        V, Eidx = tet_grid(2, dx, (0.0, 0.0, 0.0))
        count = {}
        for e in Eidx.T:
            for drop in range(4):
                f = tuple(int(x) for x in np.delete(e, drop))
                key = tuple(sorted(f))
                count.setdefault(key, []).append((f, int(e[drop])))
        surf = [lst[0][0] + (lst[0][1],) for lst in count.values() if len(lst) == 1]
        edim = 4
    X = V.T
    center = X.mean(axis=0)
    oriented = []
    for f in surf:
        f = list(f[:4] if edim == 8 else f[:3])
        P = X[f]
        if edim == 8:
            un = np.cross(P[2] - P[0], P[3] - P[1])
        else:
            un = np.cross(P[1] - P[0], P[2] - P[1])
        if un @ (P.mean(axis=0) - center) < 0:
            f = f[::-1]
        oriented.append(f)
    if faces is not None:
        oriented = [oriented[i] for i in faces]
    mu_l, la = lame(E, nu)
    q0 = V.T.ravel().copy()
    rng = np.random.default_rng(seed)
    prm = [rho_water] + list(v_water) + HYDRO_CD + HYDRO_CT + [max_force] + [float(i) for f in oriented for i in f]
    return {
        "name": "swim_" + kind, "edim": edim, "V": V, "E": Eidx, "density": rho, "gravity": None,
        "state_forces": [("hydrodynamics", prm)], "surface": oriented,
        "energies": [("corotated", [2 * mu_l], []), ("volume", [la], [])], "dirichlet": [], "boundary": None,
        "q0": q0, "v0": 0.02 * rng.standard_normal(q0.size), "dt": 1e-2, "frames": 10,
        "loss_g": rng.standard_normal(q0.size), "options": default_options(max_pd_iter=5000),
    }


def napkin(cells=(10, 10, 1), contact_ratio=0.4, E=4e5, nu=0.45, seed=42):
    """Heterogeneous napkin dropped on a sphere (python/example/napkin_3d_test.py:32-45,
    python/py_diff_pd/env/napkin_env_3d.py:19-79): stiff half / 0.1x soft half
    by element index, spherical frictional boundary under the bottom nodes
    within `contact_ratio * 0.2` of the axis, q0 jittered with the legacy
    numpy RandomState seeded with `seed`."""
    node_nums = (cells[0] + 1, cells[1] + 1, cells[2] + 1)
    dx = 0.4 / cells[0]
    origin = (-0.2, -0.2, 0.52)
    np.random.seed(seed)
    V, Eidx, _, _ = hex_grid(np.ones(cells), dx, origin)
    mu, la = lame(E, nu)
    n = cells[0] * cells[1] * cells[2]
    stiff, soft = list(range(n // 2)), list(range(n // 2, n))
    energies = [("corotated", [2 * mu], stiff), ("volume", [la], stiff),
                ("corotated", [2 * mu * 0.1], soft), ("volume", [la * 0.1], soft)]
    radius = 0.2 * contact_ratio
    fidx = []
    for i in range(node_nums[0]):
        for j in range(node_nums[1]):
            k = i * node_nums[1] * node_nums[2] + j * node_nums[2]
            px, py, _ = V[:, k]
            if px ** 2 + py ** 2 <= radius ** 2:
                fidx.append(k)
    dofs = 3 * V.shape[1]
    q0 = V.T.ravel() + np.random.uniform(low=-2 * dx * 0.01, high=2 * dx * 0.01, size=dofs)
    return {
        "name": "napkin", "edim": 8, "V": V, "E": Eidx, "density": 1e3, "gravity": (0.0, 0.0, -9.81),
        "energies": energies, "dirichlet": [], "boundary": ("spherical", [0.0, 0.0, 0.0, 0.5], fidx),
        "q0": q0, "v0": np.zeros(dofs), "dt": 2e-3, "frames": 125,
        "options": {"max_pd_iter": 5000, "max_ls_iter": 10, "abs_tol": 1e-9, "rel_tol": 1e-4, "verbose": 0,
                    "thread_ct": 8, "use_bfgs": 1, "use_acc": 1, "use_sparse": 0, "project_newton": 0, "use_abs": 0},
    }


def _smoothed_field(shape, rng, sigma_cells):
    """Gaussian random field smoothed by a separable Gaussian (FFT), zero mean."""
    white = rng.standard_normal(shape)
    f = np.fft.fftn(white)
    grids = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    k2 = sum(g ** 2 for g in grids)
    field = np.real(np.fft.ifftn(f * np.exp(-2 * (np.pi ** 2) * (sigma_cells ** 2) * k2)))
    return field - field.mean()


def press(cells=(64, 64, 32), dx=0.01, E0=1e5, decades=1.0, nu=0.45, rho=1000.0, fmu=0.3, frames=100, seed=0,
          bump=0.05):
    """Contact-rich press: hex block, corotated + volume with per-element E from
    a seeded smoothed Gaussian random field in log space spanning `decades`
    decades, pressed by a rigid plane lowered by 20% of the block height over
    the rollout into a bumpy heightfield ground (sum of 16 seeded sinusoids,
    total amplitude `bump` (0.05) of the block size; 0 gives a flat ground).
    Loss = mean squared distance of the final vertex positions to a seeded
    target."""
    rng = np.random.default_rng(seed)
    V, Eidx, _, _ = hex_grid(np.ones(cells), dx, (0.0, 0.0, 0.0))
    ne = Eidx.shape[1]
    size = max(cells) * dx
    height = cells[2] * dx
    # ground heightfield
    nw = 16
    amp = bump * size
    u = rng.uniform(0.5, 1.0, nw)
    a = amp * u / u.sum()
    kx = 2 * np.pi * rng.integers(1, 5, nw) / (cells[0] * dx) * rng.choice([-1, 1], nw)
    ky = 2 * np.pi * rng.integers(1, 5, nw) / (cells[1] * dx) * rng.choice([-1, 1], nw)
    ph = rng.uniform(0, 2 * np.pi, nw)
    waves = np.stack([a, kx, ky, ph], 1).ravel()
    # place the block just above the highest bump of its footprint: a gap of
    # dx/50 keeps every vertex off d == 0, where the NCP weight
    # omega = 1 - d / max(|d|, 1e-14) jumps between 0 and 2 on round-off
    xs = np.unique(V[0]); ys = np.unique(V[1])
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    H = np.zeros_like(X)
    for i in range(nw):
        H += a[i] * np.sin(kx[i] * X + ky[i] * Y + ph[i])
    z0 = -H.max() - 0.02 * dx
    plane0 = height
    # per-element E: log-space smoothed GRF spanning `decades`
    field = _smoothed_field(cells, rng, sigma_cells=3.0).ravel()
    field = (field - field.min()) / max(field.max() - field.min(), 1e-30)
    Ee = E0 * 10.0 ** (decades * field)
    mu_e, la_e = lame(Ee, nu)
    energies = [("corotated", list(2 * mu_e), []), ("volume", list(la_e), [])]
    bottom = np.flatnonzero(np.abs(V[2]) < 1e-12)
    top = np.flatnonzero(np.abs(V[2] - height) < 1e-12)
    fidx = sorted(list(bottom) + list(top))
    params = [z0, nw] + list(waves) + [plane0]
    q0 = V.T.ravel().copy()
    target = q0 + 0.01 * rng.standard_normal(q0.size)
    return {
        "name": "press", "edim": 8, "V": V, "E": Eidx, "density": rho, "gravity": (0.0, 0.0, -9.81),
        "energies": energies, "dirichlet": [], "boundary": ("press", params, fidx),
        "q0": q0, "v0": np.zeros_like(q0), "dt": 1e-2, "frames": frames, "target": target,
        "plane_schedule": [plane0 - 0.2 * height * (f + 1) / frames for f in range(frames)],
        "options": default_options(friction_mu=fmu, max_pd_iter=500, rel_tol=1e-9),
        "E_elem": Ee,
    }


def beam(cells=(16, 16, 8), dx=0.01, E=1e5, nu=0.45, rho=1000.0, fmu=0.5, v_std=0.1, seed=0):
    """One beam of the batched workload: a free hex block (corotated + volume)
    resting a gap of dx/50 above a planar frictional ground z = 0 (bottom-face
    vertices are the candidates), gravity, initial velocity ~ N(0, v_std) per
    DOF. Loss = dot(q_T, g), g ~ N(0, 1)."""
    rng = np.random.default_rng(seed)
    V, Eidx, _, _ = hex_grid(np.ones(cells), dx, (0.0, 0.0, 0.02 * dx))
    mu, la = lame(E, nu)
    energies = [("corotated", [2 * mu], []), ("volume", [la], [])]
    bottom = sorted(np.flatnonzero(np.abs(V[2] - 0.02 * dx) < 1e-12).tolist())
    q0 = V.T.ravel().copy()
    return {
        "name": "beam", "edim": 8, "V": V, "E": Eidx, "density": rho, "gravity": (0.0, 0.0, -9.81),
        "energies": energies, "dirichlet": [], "boundary": ("planar", [0.0, 0.0, 1.0, 0.0], bottom),
        "q0": q0, "v0": v_std * rng.standard_normal(q0.size), "dt": 1e-2, "frames": 50,
        "loss_g": rng.standard_normal(q0.size),
        "options": default_options(friction_mu=fmu, max_pd_iter=500, rel_tol=1e-9),
    }


def beam_batch(n=256, seed=0, dims=((8, 32), (8, 32), (8, 16))):
    """The batched-rollout workload: `n` beams with cell counts uniform in
    `dims`, E log-uniform in [1e4, 1e6], nu uniform in [0.3, 0.49], initial
    velocities ~ N(0, 0.1 m/s), ground friction 0.5, dt 1e-2, 50 frames."""
    rng = np.random.default_rng(seed)
    specs = []
    for b in range(n):
        cells = tuple(int(rng.integers(lo, hi + 1)) for lo, hi in dims)
        E = 10.0 ** rng.uniform(4.0, 6.0)
        nu = rng.uniform(0.3, 0.49)
        specs.append(beam(cells=cells, E=E, nu=nu, seed=int(rng.integers(1 << 30))))
    return specs


def tet_block(n=40, dx=None, seed=0, nu=0.4, inclusions=8, fmu=0.5):
    """Heterogeneous Neo-Hookean tet block on a ground plane: n^3 cubes split
    into 5 tets, two-phase per-element mu (seeded spherical inclusions at 100x
    contrast, log-uniform within each phase over [1e4, 1e6] Pa), lambda = 4 mu."""
    rng = np.random.default_rng(seed)
    dx = dx if dx is not None else 0.4 / n
    V, Eidx = tet_grid(n, dx, (0.0, 0.0, 0.02))
    ne = Eidx.shape[1]
    cent = V[:, Eidx].mean(axis=1)
    L = n * dx
    inc_c = rng.uniform(0.2 * L, 0.8 * L, (inclusions, 3)) + np.array([0, 0, 0.02])
    inc_r = rng.uniform(0.08 * L, 0.15 * L, inclusions)
    inside = np.zeros(ne, dtype=bool)
    for c, r in zip(inc_c, inc_r):
        inside |= np.linalg.norm(cent.T - c, axis=1) < r
    base = 10 ** rng.uniform(4, 5, ne)          # soft phase
    mu = np.where(inside, base * 100, base)      # inclusions 100x stiffer, within [1e4, 1e7]
    mu = np.clip(mu, 1e4, 1e6 * 100)
    la = 4 * mu
    params = np.stack([mu, la], 1).ravel()
    bottom = np.flatnonzero(np.abs(V[2] - 0.02) < 1e-12)
    q0 = V.T.ravel().copy()
    return {
        "name": "tet_block", "edim": 4, "V": V, "E": Eidx, "density": 1000.0, "gravity": (0.0, 0.0, -9.81),
        "energies": [("pd_neohookean", list(params), [])], "dirichlet": [],
        "boundary": ("planar", [0.0, 0.0, 1.0, 0.0], list(bottom)), "q0": q0, "v0": np.zeros_like(q0),
        "dt": 5e-3, "frames": 200, "options": default_options(friction_mu=fmu, max_pd_iter=1000),
    }


def stack(top=(4, 4, 2), bottom=(8, 8, 2), dx=0.01, E=1e5, nu=0.45, rho=1000.0, fmu=0.5, gap=5e-4, radius=2e-4,
          fixed_bottom="all", v_down=0.3, both_sides=True):
    """Two hex blocks in one body in contact through the mesh frictional
    boundary (MeshFrictionalBoundary, mesh_frictional_boundary.cpp): the upper
    block (vertices 0 .. split - 1, its bottom face the left triangle set)
    falls with speed v_down onto the lower block (its top face the right
    set), starting `gap` above it; contact at distance `radius`. Upper
    vertices query the right set, lower vertices the left set.
    fixed_bottom: "all" (every lower vertex fixed: a rigid slab) or "base"
    (only the lower block's bottom face). both_sides: the lower block's top
    vertices are candidates too (else only the upper block's bottom ones)."""
    Vb, Eb, _, fb_flag = hex_grid(np.ones(bottom), dx, (0.0, 0.0, -bottom[2] * dx))
    ox, oy = (bottom[0] - top[0]) * dx / 2, (bottom[1] - top[1]) * dx / 2
    Vt, Et, _, ft_flag = hex_grid(np.ones(top), dx, (ox, oy, gap))
    nt = Vt.shape[1]
    V = np.concatenate([Vt, Vb], axis=1)
    Eidx = np.concatenate([Et, Eb + nt], axis=1)

    def face_tris(flag, k, up, off):
        tris = []
        for i in range(flag.shape[0] - 1):
            for j in range(flag.shape[1] - 1):
                a, b, c, d = flag[i, j, k], flag[i + 1, j, k], flag[i + 1, j + 1, k], flag[i, j + 1, k]
                if up:  # counter-clockwise seen from +z: outward normal +z
                    tris += [(a, b, c), (a, c, d)]
                else:
                    tris += [(a, c, b), (a, d, c)]
        return (np.asarray(tris, dtype=np.int64) + off).ravel()

    left = face_tris(ft_flag, 0, False, 0)                       # upper block's bottom face
    right = face_tris(fb_flag, bottom[2], True, nt)              # lower block's top face
    cand_up = sorted(int(v) for v in ft_flag[:, :, 0].ravel())
    cand_low = sorted(int(v) + nt for v in fb_flag[:, :, bottom[2]].ravel())
    fbv = cand_up + (cand_low if both_sides else [])
    params = [radius, nt, left.size // 3, right.size // 3] + [float(x) for x in left] + [float(x) for x in right]
    mu, la = lame(E, nu)
    if fixed_bottom == "all":
        fixed = range(nt, V.shape[1])
    else:
        fixed = [int(v) + nt for v in fb_flag[:, :, 0].ravel()]
    dirichlet = [(3 * i + k, float(V[k, i])) for i in fixed for k in range(3)]
    q0 = V.T.ravel().copy()
    v0 = np.zeros_like(q0)
    v0[2:3 * nt:3] = -v_down
    return {
        "name": "stack", "edim": 8, "V": V, "E": Eidx, "density": rho, "gravity": (0.0, 0.0, -9.81),
        "energies": [("corotated", [2 * mu], []), ("volume", [la], [])], "dirichlet": dirichlet,
        "boundary": ("mesh", params, fbv), "q0": q0, "v0": v0, "dt": 1e-2, "frames": 20, "split": nt,
        "mu": mu, "la": la, "options": default_options(friction_mu=fmu, max_pd_iter=500, rel_tol=1e-9),
    }
