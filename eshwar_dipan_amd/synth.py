"""Synthetic microstructures and Input.in text for the grand-potential step.

Everything here is seeded and deterministic (SplitMix64), so a workload is named
by its parameters alone.  Fields are interior-only arrays phi[P][MX][MY],
mu[C][MX][MY] (y fastest), the layout of the reference's VTK blocks.
"""
import numpy as np

_M64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): state += golden gamma, then mix."""

    def __init__(self, seed):
        self.s = seed & _M64

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self):
        """double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo, hi):
        """integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))

    def normal(self):
        """standard normal by Box-Muller (first of the pair)."""
        u1 = self.uniform()
        while u1 <= 0.0:
            u1 = self.uniform()
        u2 = self.uniform()
        return float(np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2))


def splitmix_uniforms(seed, n, skip=0):
    """The next n uniforms of SplitMix64(seed) after skipping `skip` draws,
    computed in bulk (state after i draws is seed + i * gamma mod 2**64)."""
    with np.errstate(over="ignore"):
        i = np.arange(skip + 1, skip + n + 1, dtype=np.uint64)
        z = np.uint64(seed & _M64) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def splitmix_normals(seed, n):
    """n standard normals by Box-Muller from consecutive uniform pairs."""
    u = splitmix_uniforms(seed, 2 * n)
    u1 = np.where(u[0::2] <= 0.0, 2.0 ** -53, u[0::2])
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u[1::2])


def base_params(P=2, C=1, mesh=(64, 64), **kw):
    """Input.in keys (as a dict of key -> value string) for a P-phase, (C+1)-component
    F_01 run with the liquid as the last phase.  Keyword overrides replace keys."""
    MX, MY = mesh
    d = {
        "DIMENSION": "2", "MESH_X": str(MX), "MESH_Y": str(MY), "MESH_Z": "1",
        "DELTA_X": "1.0", "DELTA_Y": "1.0", "DELTA_Z": "1.0", "DELTA_t": "0.05",
        "NUMPHASES": str(P), "NUMCOMPONENTS": str(C + 1),
        "NTIMESTEPS": "100", "NSMOOTH": "1", "SAVET": "100", "STARTTIME": "0", "RESTART": "0",
        "COMPONENTS": "{" + ", ".join(["Al", "Cu", "Ni", "Zn", "Mg"][: C + 1]) + "}",
        "PHASES": "{" + ", ".join([f"s{a}" for a in range(P - 1)] + ["liquid"]) + "}",
        "GAMMA": "{" + ",".join(["1.0"] * (P * (P - 1) // 2)) + "}",
        "R": "1.0", "V": "1.0",
        "ISOTHERMAL": "1", "T": "0.95", "Equilibrium_temperature": "1.0", "Filling_temperature": "1.0",
        "WRITEFORMAT": "BINARY", "TRACK_PROGRESS": "100",
        "epsilon": "4.0", "tau": "1.31",
        "Tau": "{" + ",".join(["0.28"] * (P * (P - 1) // 2)) + "}",
        "Function_anisotropy": "0", "Anisotropy_type": "4",
        "dab": "{" + ",".join(["0.0"] * (P * (P - 1) // 2)) + "}",
        "Function_W": "1",
        "Gamma_abc": "{" + ",".join(["0.0"] * (P * (P - 1) * (P - 2) // 6)) + "}",
        "Shift": "0", "Shiftj": "30", "Writecomposition": "0",
        "Noise_phasefield": "0", "Amp_Noise_Phase": "0.0",
        "Tempgrady": "{0.96, 0.06, 800.0, 0, 0.016}",
        "Function_F": "1",
    }
    d.update({k: str(v) for k, v in kw.items()})
    return d


def thermo_lines(P, C, A_diag=1.0, A_off=0.0, c_sol=0.2, c_liq=0.4, slope=-0.45,
                 D_sol=0.0, D_liq=1.0, solver="BINARY"):
    """Multi-valued Input.in lines: A, ceq, cfill, slopes, DIFFUSIVITY and the
    solver flag.  Solid a prefers component a % C (solute-poor elsewhere)."""
    L = P - 1
    out = []
    if solver == "BINARY":
        out.append("BINARY = 1;")
    elif solver == "TERNARY":
        out.append("TERNARY = 1;")
    elif solver == "DILUTE":
        out.append("DILUTE = 1;")
    for a in range(P):
        offs = [f"{A_off}"] * (C * (C - 1) // 2)
        out.append("A = {" + ", ".join([str(a)] + [f"{A_diag}"] * C + offs) + "};")
    for a in range(P):
        diag = [f"{D_sol if a < L else D_liq}"] * C
        out.append("DIFFUSIVITY = {" + ", ".join(["1", str(a)] + diag) + "};")

    def comp(a, b):
        if a == L and b == L:
            return [c_liq] * C
        if a == b:
            v = [c_sol] * C
            v[a % C] = c_sol + 0.05 * (a % 3)
            return v
        return [c_liq] * C

    for a in range(P):
        for b in (a, L):
            for key in ("ceq", "cfill"):
                out.append(f"{key} = {{{a}, {b}, " + ", ".join(f"{v:.6g}" for v in comp(a, b)) + "};")
            out.append(f"slopes = {{{a}, {b}, " + ", ".join([f"{slope}"] * C) + "};")
        if a != L:
            for key in ("ceq", "cfill"):
                out.append(f"{key} = {{{L}, {a}, " + ", ".join(f"{v:.6g}" for v in comp(a, L)) + "};")
            out.append(f"slopes = {{{L}, {a}, " + ", ".join([f"{slope}"] * C) + "};")
    return out


def bc_lines(phi=(1, 1, 1, 1), mu=(1, 1, 1, 1)):
    """BOUNDARY lines, tuples in the file's order {X+, X-, Y+, Y-}."""
    out = []
    for name, t in (("phi", phi), ("mu", mu)):
        out.append("BOUNDARY = {%s, %d, %d, %d, %d, 0, 0};" % ((name,) + tuple(t)))
    for name in ("phi", "mu"):
        out.append("BOUNDARY_VALUE = {%s, 0, 0, 0, 0, 0, 0};" % name)
    return out


def rotation_lines(P, angles=None):
    """Rotation_matrix lines for every pair; angles[(a,b)] = (tx, ty, tz) degrees."""
    out = []
    for a in range(P):
        for b in range(a + 1, P):
            t = (angles or {}).get((a, b), (0, 0, 0))
            out.append(f"Rotation_matrix = {{{a}, {b}, {t[0]}, {t[1]}, {t[2]}}};")
    return out


def infile_text(keys, extra_lines=()):
    lines = [f"{k} = {v};" for k, v in keys.items()]
    return "\n".join(lines + list(extra_lines)) + "\n"


# --------------------------------------------------------------------------
# initial fields
# --------------------------------------------------------------------------

def seeds(MX, MY, P, nseeds, rmin, rmax, seed, diffuse=0.0, periodic=False):
    """Circular solid seeds with uniform-random centres, radii in [rmin, rmax] and
    uniform-random solid labels in {0..P-2}; the liquid (P-1) fills the rest.
    diffuse > 0 gives tanh profiles of that half-width instead of sharp disks.
    Overlapping seeds are resolved by the later seed (sharp) or by rescaling the
    solids to sum <= 1 (diffuse)."""
    rng = SplitMix64(seed)
    phi = np.zeros((P, MX, MY))
    X, Y = np.meshgrid(np.arange(MX, dtype=float), np.arange(MY, dtype=float), indexing="ij")
    for _ in range(nseeds):
        cx, cy = rng.uniform() * MX, rng.uniform() * MY
        r = rmin + (rmax - rmin) * rng.uniform()
        lab = rng.randint(0, P - 2)
        dxv, dyv = X - cx, Y - cy
        if periodic:
            dxv = (dxv + MX / 2) % MX - MX / 2
            dyv = (dyv + MY / 2) % MY - MY / 2
        d = np.sqrt(dxv * dxv + dyv * dyv) - r
        if diffuse > 0:
            v = 0.5 * (1.0 - np.tanh(d / diffuse))
            phi[lab] = np.maximum(phi[lab], v)
        else:
            m = d <= 0
            phi[:, m] = 0.0
            phi[lab, m] = 1.0
    s = phi[: P - 1].sum(0)
    over = s > 1.0
    if over.any():
        phi[: P - 1, over] /= s[over]
        s = phi[: P - 1].sum(0)
    phi[P - 1] = 1.0 - s
    phi[P - 1][phi[P - 1] < 1e-12] = 0.0
    return phi


def voronoi(MX, MY, P, npoints, seed, liquid_fraction=0.0):
    """Sharp Voronoi grains: npoints seeded sites with labels in {0..P-2}; each cell
    takes its nearest site's label; a seeded fraction of sites is liquid."""
    rng = SplitMix64(seed)
    pts = []
    for _ in range(npoints):
        cx, cy = rng.uniform() * MX, rng.uniform() * MY
        lab = P - 1 if rng.uniform() < liquid_fraction else rng.randint(0, P - 2)
        pts.append((cx, cy, lab))
    X, Y = np.meshgrid(np.arange(MX, dtype=float), np.arange(MY, dtype=float), indexing="ij")
    best = np.full((MX, MY), np.inf)
    lab = np.zeros((MX, MY), dtype=np.int64)
    for cx, cy, l in pts:
        d = (X - cx) ** 2 + (Y - cy) ** 2
        m = d < best
        best[m] = d[m]
        lab[m] = l
    phi = np.zeros((P, MX, MY))
    for a in range(P):
        phi[a][lab == a] = 1.0
    return phi


def mu_fill(MX, MY, c_liq, Mu, noise=0.0, seed=0):
    """mu = Mu_liq(c_liq, Teq) everywhere plus N(0, noise) per cell and component.
    Mu(c, k) must return the liquid chemical potential of composition c."""
    C = len(c_liq)
    mu = np.empty((C, MX, MY))
    for k in range(C):
        mu[k] = Mu(c_liq, k)
    if noise > 0:
        mu += noise * splitmix_normals(seed, C * MX * MY).reshape(C, MX, MY)
    return mu
