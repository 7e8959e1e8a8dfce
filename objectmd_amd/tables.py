"""Potential tables in the reference's table-file format, and unit constants.

Table files use the ``#$<NAME> key value`` header lines and the
``#$<NAME> --`` data marker that TableReader reads
(src/omd/treader.cpp:83-136; tables/eam.aluminum is an example).  An EAM file
holds the PAIR, EDENS and EMBED tables (src/gadget/team.cpp:47-66, 206-222);
a pair file holds one unnamed table (include/potential/tpair.h:55-60).
"""
import numpy as np

# include/omd/conversion.h
NORMAL_AMU = 9.64853397731809e3      # amu -> eV ps^2 / A^2 divisor
MD_BOLTZMANN = 8.61734279e-05        # eV / K
MD_GIGAPASCAL = 160.22               # GPa per eV/A^3

MASS_AL = 26.98154                   # tables/def.aluminum
MASS_AR = 39.948
MASS_SI = 28.0855


def mass_sim(amu):
    """Mass in eV ps^2 / A^2 (container.cpp:94)."""
    return amu / NORMAL_AMU


class Table:
    """Equispaced samples y of one table: r_i = r0 + i dr."""

    def __init__(self, y, dr, r0, rmult=False):
        self.y = np.asarray(y, dtype=np.float64)
        self.dr = float(dr)
        self.r0 = float(r0)
        self.rmult = bool(rmult)

    @property
    def n(self):
        return len(self.y)

    @property
    def max_range(self):
        return self.r0 + (self.n - 1) * self.dr


def write_table_file(path, tables, comment=""):
    """tables: dict NAME -> Table ('' for an unnamed pair table)."""
    with open(path, "w") as fl:
        for line in comment.splitlines():
            fl.write("# " + line + "\n")
        for name, t in tables.items():
            tag = "#$" + name
            fl.write(f"{tag} NumberOfData {t.n}\n")
            fl.write(f"{tag} Spacing {t.dr:.17e}\n")
            fl.write(f"{tag} Offset {t.r0:.17e}\n")
            fl.write(f"{tag} Format {'RMULT' if t.rmult else 'PLAIN'}\n")
            fl.write(f"{tag} --\n")
            for v in t.y:
                fl.write(f"{v:.17e}\n")


# ---------------------------------------------------------------------------
# Johnson-type analytic EAM for aluminium
# ---------------------------------------------------------------------------
def _taper(g, dg, r, rc):
    """Shifted-force cut: g(r) - g(rc) - (r - rc) g'(rc), zero value and slope
    at rc, so pairs crossing the cutoff do not jump in energy or force."""
    return g(r) - g(rc) - (r - rc) * dg(rc)


class JohnsonEAM:
    """f(r) = exp(-beta (r/re - 1)); Morse pair with its minimum at rm,
    phi(r) = D [exp(-2 alpha (r - rm)) - 2 exp(-alpha (r - rm))];
    F(rho) = -A sqrt(rho).  D and A are fitted so the fcc crystal at a0 has
    cohesive energy Ec and zero pressure; alpha is chosen for the bulk
    modulus B.  With rm beyond the nearest-neighbour distance the pair term
    is repulsive at a0 and the embedding term binds (D > 0, A > 0).  Density
    and pair term are tapered to zero value and slope at rc."""

    def __init__(self, a0=4.05, re=2.86, beta=6.0, ec=3.36, rc=5.5,
                 bulk_gpa=76.0, rm=3.3):
        self.a0, self.re, self.beta, self.ec, self.rc = a0, re, beta, ec, rc
        self.rm = rm
        self.alpha = self._fit_alpha(bulk_gpa)
        self.D, self.A = self._fit_DA(self.alpha)

    # raw functions
    def f_raw(self, r):
        return np.exp(-self.beta * (r / self.re - 1.0))

    def df_raw(self, r):
        return -self.beta / self.re * self.f_raw(r)

    def morse(self, r, alpha):
        e = np.exp(-alpha * (r - self.rm))
        return e * e - 2.0 * e

    def dmorse(self, r, alpha):
        e = np.exp(-alpha * (r - self.rm))
        return -2.0 * alpha * e * e + 2.0 * alpha * e

    # tapered
    def f(self, r):
        r = np.asarray(r, dtype=np.float64)
        return np.where(r < self.rc, _taper(self.f_raw, self.df_raw, r, self.rc), 0.0)

    def phi_unit(self, r, alpha=None):
        a = self.alpha if alpha is None else alpha
        r = np.asarray(r, dtype=np.float64)
        g = lambda x: self.morse(x, a)
        dg = lambda x: self.dmorse(x, a)
        return np.where(r < self.rc, _taper(g, dg, r, self.rc), 0.0)

    def phi(self, r):
        return self.D * self.phi_unit(r)

    def F(self, rho):
        return -self.A * np.sqrt(np.maximum(rho, 0.0))

    # fcc lattice sums
    def _shells(self, a):
        m = int(np.ceil(self.rc / a)) + 1
        rng = np.arange(-2 * m, 2 * m + 1)
        i, j, k = np.meshgrid(rng, rng, rng, indexing="ij")
        sel = ((i + j + k) % 2 == 0)
        pts = np.stack([i[sel], j[sel], k[sel]], 1) * (a / 2.0)
        r = np.sqrt((pts ** 2).sum(1))
        return r[(r > 0) & (r < self.rc)]

    def _energy_terms(self, a, alpha):
        r = self._shells(a)
        return 0.5 * self.phi_unit(r, alpha).sum(), self.f(r).sum()

    def _fit_DA(self, alpha):
        h = 1e-4
        s0, rho0 = self._energy_terms(self.a0, alpha)
        sp, rhop = self._energy_terms(self.a0 + h, alpha)
        sm, rhom = self._energy_terms(self.a0 - h, alpha)
        ds = (sp - sm) / (2 * h)
        drho = (rhop - rhom) / (2 * h)
        # D s0 - A sqrt(rho0) = -Ec ;  D ds - A drho / (2 sqrt(rho0)) = 0
        M = np.array([[s0, -np.sqrt(rho0)], [ds, -drho / (2 * np.sqrt(rho0))]])
        D, A = np.linalg.solve(M, np.array([-self.ec, 0.0]))
        return float(D), float(A)

    def energy_per_atom(self, a, alpha=None, D=None, A=None):
        alpha = self.alpha if alpha is None else alpha
        D = self.D if D is None else D
        A = self.A if A is None else A
        s, rho = self._energy_terms(a, alpha)
        return D * s - A * np.sqrt(rho)

    def bulk_modulus(self, alpha=None):
        alpha = self.alpha if alpha is None else alpha
        D, A = self._fit_DA(alpha)
        h = 1e-3
        e = [self.energy_per_atom(self.a0 + k * h, alpha, D, A) for k in (-1, 0, 1)]
        d2 = (e[0] - 2 * e[1] + e[2]) / h ** 2
        v0 = self.a0 ** 3 / 4.0
        # B = V d2E/dV2 with V = a^3/4:  dV/da = 3a^2/4
        return d2 / (9.0 * v0 / self.a0 ** 2) * MD_GIGAPASCAL

    def _fit_alpha(self, bulk_gpa):
        grid = np.linspace(0.6, 1.6, 41)
        vals = np.array([self.bulk_modulus(a) for a in grid])
        k = int(np.argmin(np.abs(vals - bulk_gpa)))
        lo, hi = grid[max(0, k - 1)], grid[min(len(grid) - 1, k + 1)]
        for _ in range(40):
            mid = 0.5 * (lo + hi)
            if (self.bulk_modulus(lo) - bulk_gpa) * (self.bulk_modulus(mid) - bulk_gpa) <= 0:
                hi = mid
            else:
                lo = mid
        return float(0.5 * (lo + hi))

    def rho_lattice(self, a=None):
        return float(self.f(self._shells(self.a0 if a is None else a)).sum())

    def tables(self, n=5000, r_start=1.0, rho_max_factor=2.5):
        """PAIR, EDENS, EMBED tables with n knots each; PAIR and EDENS share
        the grid [r_start, rc], EMBED covers [0, rho_max_factor rho0]."""
        dr = (self.rc - r_start) / (n - 1)
        r = r_start + dr * np.arange(n)
        rho_max = rho_max_factor * self.rho_lattice()
        drho = rho_max / (n - 1)
        rho = drho * np.arange(n)
        return {
            "PAIR": Table(self.phi(r), dr, r_start),
            "EDENS": Table(self.f(r), dr, r_start),
            "EMBED": Table(self.F(rho), drho, 0.0),
        }

    def write(self, path, n=5000):
        write_table_file(
            path, self.tables(n),
            comment=(f"Johnson-type EAM aluminium: re={self.re} beta={self.beta} "
                     f"Ec={self.ec} rc={self.rc}\nMorse alpha={self.alpha:.10f} "
                     f"rm={self.rm} D={self.D:.10f}; F=-A sqrt(rho), A={self.A:.10f}"))


# ---------------------------------------------------------------------------
# Lennard-Jones
# ---------------------------------------------------------------------------
def lj_table(eps=0.0104, sigma=3.40, rc=None, n=5000, r_start=2.0, rmult=True):
    """Truncated 12-6 Lennard-Jones table on [r_start, rc] (rc = 2.5 sigma
    by default).  RMULT tables store r V(r), as tables/pair.argon.lj does."""
    rc = 2.5 * sigma if rc is None else rc
    dr = (rc - r_start) / (n - 1)
    r = r_start + dr * np.arange(n)
    sr6 = (sigma / r) ** 6
    v = 4.0 * eps * (sr6 * sr6 - sr6)
    return Table(r * v if rmult else v, dr, r_start, rmult)


def lj_energy(r, eps=0.0104, sigma=3.40):
    sr6 = (sigma / r) ** 6
    return 4.0 * eps * (sr6 * sr6 - sr6)


def lj_force_over_r(r, eps=0.0104, sigma=3.40):
    """-V'(r)/r"""
    sr6 = (sigma / r) ** 6
    return 24.0 * eps * (2.0 * sr6 * sr6 - sr6) / (r * r)
